// Open-time arithmetic gate (the role of device/selfcheck.rs in the reference): the device must
// give the host's IEEE single precision results bit for bit on the operations the kernels rely
// on -- correctly rounded division, no fused multiply-add, preserved denormals, signed zeros, the
// truncation rules and the double precision rounding step -- or the backend refuses to open.

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "twx_common.hpp"

namespace twx {

namespace {

constexpr int NOPS = 9;

__host__ __device__ inline uint32_t fbits(float f) {
#ifdef __HIP_DEVICE_COMPILE__
    return __float_as_uint(f);
#else
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
#endif
}

__host__ __device__ inline int32_t trunc_x86(float x) {
    if (x != x || x >= 2147483648.0f || x < -2147483648.0f) return INT_MIN;
    return (int32_t)x;
}

__host__ __device__ inline int64_t trunc_x86_d(double x) {
    if (x != x || x >= 9.223372036854775808e18 || x < -9.223372036854775808e18) return LLONG_MIN;
    return (int64_t)x;
}

// The checked operations; the same source is compiled for the host and for gfx950.
__host__ __device__ inline uint32_t op(int k, float a, float b, float c) {
    switch (k) {
        case 0: return fbits(a / b);                          // correctly rounded division
        case 1: return fbits(a * b + c);                      // two roundings, never an FMA
        case 2: return fbits((b - a) * c + a);                // the lerp chain of the samplers
        case 3: return fbits(a * b);                          // products into the denormal range
        case 4: return (uint32_t)trunc_x86(a * b);            // truncation with INT_MIN for NaN / range
        case 5: return (uint32_t)(trunc_x86_d((double)(a * c) + 0.5) & 0xffff);  // q() of the premultiply
        case 6: return fbits((a - b) + 0.0f);                 // -0 + 0 = +0
        case 7: return fbits((float)((double)a / (double)b)); // double division narrowed
        case 8: return fbits((a * a - b * c) / (a + 40.0f));  // the new_disp shape
    }
    return 0;
}

__global__ void k_selfcheck(const float* __restrict__ in, uint32_t* __restrict__ out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = in[i * 3], b = in[i * 3 + 1], c = in[i * 3 + 2];
    for (int k = 0; k < NOPS; k++) out[(long)i * NOPS + k] = op(k, a, b, c);
}

float from_bits(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}

}  // namespace

int run_selfcheck(hipStream_t s, std::string& err) {
    std::vector<float> in;
    auto add = [&](float a, float b, float c) { in.push_back(a); in.push_back(b); in.push_back(c); };
    // adversarial corpus: denormals, signed zeros, specials, .5 ties, cast bounds, FMA-sensitive products
    const uint32_t specials[] = {0x00000000u, 0x80000000u, 0x00000001u, 0x007fffffu, 0x00800000u, 0x3f800000u,
                                 0x7f7fffffu, 0x7f800000u, 0xff800000u, 0x7fc00000u, 0x4f000000u, 0xcf000000u,
                                 0x3effffffu, 0x3f000000u, 0x437f0000u, 0x3b808081u};
    for (uint32_t x : specials)
        for (uint32_t y : specials) add(from_bits(x), from_bits(y), 0.5f);
    add(1.0f + 1.1920929e-7f, 1.0f - 5.9604645e-8f, -1.0f);  // a*b+c differs from fma(a,b,c)
    add(0.1f, 10.0f, -1.0f);
    add(1e-20f, 1e-20f, 0.0f);
    add(3.0e9f, 1.0f, 0.0f);
    add(-3.0e9f, 1.0f, 0.0f);
    add(254.49998f, 1.0f, 1.0f);
    add(0.49999997f, 1.0f, 1.0f);
    std::mt19937 rng(1234);
    std::uniform_real_distribution<float> u(-300.0f, 300.0f);
    std::uniform_int_distribution<uint32_t> bits(0, 0xffffffffu);
    for (int i = 0; i < 4096; i++) add(u(rng), u(rng), u(rng) * 1e-3f);
    for (int i = 0; i < 2048; i++) add(from_bits(bits(rng)), from_bits(bits(rng)), from_bits(bits(rng)));
    int n = (int)(in.size() / 3);
    float* d_in = nullptr;
    uint32_t* d_out = nullptr;
    std::vector<uint32_t> got((size_t)n * NOPS);
    if (hipMalloc(&d_in, in.size() * 4) != hipSuccess) { err = "hipMalloc failed"; return TWX_E_DEVICE; }
    if (hipMalloc(&d_out, got.size() * 4) != hipSuccess) { (void)hipFree(d_in); err = "hipMalloc failed"; return TWX_E_DEVICE; }
    hipError_t e = hipMemcpyAsync(d_in, in.data(), in.size() * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        k_selfcheck<<<dim3((n + 255) / 256), dim3(256), 0, s>>>(d_in, d_out, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(got.data(), d_out, got.size() * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    if (e != hipSuccess) { err = std::string("self-check launch: ") + hipGetErrorString(e); return TWX_E_DEVICE; }
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < NOPS; k++) {
            uint32_t want = op(k, in[i * 3], in[i * 3 + 1], in[i * 3 + 2]);
            uint32_t have = got[(size_t)i * NOPS + k];
            // NaN payloads are outside the contract (no NaN inputs on the path): compare NaN-ness only
            bool wn = (k != 4 && k != 5) && std::isnan(from_bits(want));
            bool hn = (k != 4 && k != 5) && std::isnan(from_bits(have));
            if (wn || hn) {
                if (wn != hn) { err = "self-check: NaN mismatch in op " + std::to_string(k); return TWX_E_UNSUPPORTED; }
                continue;
            }
            if (want != have) {
                char buf[160];
                snprintf(buf, sizeof buf, "self-check: op %d on (%a, %a, %a): device %08x, host %08x", k, in[i * 3],
                         in[i * 3 + 1], in[i * 3 + 2], have, want);
                err = buf;
                return TWX_E_UNSUPPORTED;
            }
        }
    }
    return TWX_OK;
}

}  // namespace twx
