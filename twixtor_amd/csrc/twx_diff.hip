// The differentiable warp+blend: out = blend(W(A, F_BA, 1 - wa), W(B, F_AB, wa), 1 - wa) on float
// images of C channels (interleaved, rows top first), with W the inverse warp with hold edges of
// warp.rs:477-512 (flow sampled bilinearly at the working size, the sample point clipped to the
// image along its segment, clip_to_rect warp.rs:178-215, the image sampled bilinearly, bilerp4
// warp.rs:306-331) and the blend of blend.rs:55-114 on floats, ((tB - tA) * w + tA) + 0.
//
// The backward pass recomputes every sample point instead of storing it: per output pixel the
// flow sample, the clip (with the Jacobian of the clipped point with respect to the flow value
// carried through each clip step) and the image sample, then
//   dL/dframes: each output pixel's upstream gradient scattered to the four texels it sampled,
//               summed in 64-bit fixed point (a power-of-two scale from the largest upstream
//               magnitude): integer sums do not depend on the order of the atomic adds, so the
//               gradients are the same bits on every run;
//   dL/dflow:   the gradient with respect to each pixel's flow value (two floats per pixel), then
//               gathered per flow texel over the output pixels whose flow sample uses it, in a
//               fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "twx_common.hpp"

namespace twx {

namespace {

__device__ __forceinline__ int32_t cvtt_d(float x) {
    if (isnan(x) || x >= 2147483648.0f || x < -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

// The flow sample of output pixel (x, y): corner (jx, jy), fractions (gx, gy), value (dx, dy).
struct FlowSample {
    int jx, jy;
    float gx, gy;
    float d[2];
};

__device__ __forceinline__ void flow_sample(const float* __restrict__ F, int fw, int fh, float rx, float ry, int x, int y,
                                            FlowSample& s) {
    const float qx = (float)x * rx, qy = (float)y * ry;
    int32_t jx = cvtt_d(qx), jy = cvtt_d(qy);
    float gx = qx - (float)jx, gy = qy - (float)jy;
    if (jx >= fw - 1) { jx = fw - 2; gx = 1.0f; }
    if (jy >= fh - 1) { jy = fh - 2; gy = 1.0f; }
    const long f00 = ((long)jy * fw + jx) * 2, f01 = f00 + (long)fw * 2;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float p00 = F[f00 + k], p10 = F[f00 + 2 + k], p01 = F[f01 + k], p11 = F[f01 + 2 + k];
        const float top = (p10 - p00) * gx + p00;
        const float bot = (p11 - p01) * gx + p01;
        s.d[k] = (bot - top) * gy + top;
    }
    s.jx = jx;
    s.jy = jy;
    s.gx = gx;
    s.gy = gy;
}

// clip_to_rect (warp.rs:178-215) with J = d(ix, iy) / d(dx, dy) carried through each step.
__device__ __forceinline__ bool clip_jac(float x, float y, float& ix, float& iy, float w1, float h1, float J[2][2]) {
    for (int n = 0; n < 4096; n++) {
        if (ix < 0.0f || ix > w1) {
            const float bound = ix < 0.0f ? 0.0f : w1;
            const float dy = y - iy, dxs = x - ix, num = x - bound;
            const float t = num / dxs;
            // iy' = y - t * dy, t = num / (x - ix): dt = t / dxs * d ix, d dy = -d iy
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float dt = t / dxs * J[0][c];
                J[1][c] = -(dt * dy) + t * J[1][c];
                J[0][c] = 0.0f;
            }
            ix = bound;
            iy = y - t * dy;
            continue;
        }
        if (iy < 0.0f || iy > h1) {
            const float bound = iy < 0.0f ? 0.0f : h1;
            const float dyy = y - iy, num = y - bound, dx = x - ix;
            const float t = num / dyy;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const float dt = t / dyy * J[1][c];
                J[0][c] = -(dt * dx) + t * J[0][c];
                J[1][c] = 0.0f;
            }
            ix = x - t * dx;
            iy = bound;
            continue;
        }
        return true;
    }
    return false;
}

// The sample point of one warp at output pixel (x, y) and its Jacobian; false for a NaN point.
__device__ __forceinline__ bool warp_point_jac(const float* __restrict__ F, int fw, int fh, int W, int H, float amount, int x,
                                               int y, FlowSample& fs, float& ix, float& iy, float J[2][2]) {
    const float rx = (float)(fw - 1) / (float)(W - 1), ry = (float)(fh - 1) / (float)(H - 1);
    const float sx = amount / rx, sy = amount / ry;
    flow_sample(F, fw, fh, rx, ry, x, y, fs);
    ix = fs.d[0] * sx + (float)x;
    iy = fs.d[1] * sy + (float)y;
    J[0][0] = sx;
    J[0][1] = 0.0f;
    J[1][0] = 0.0f;
    J[1][1] = sy;
    if (!clip_jac((float)x, (float)y, ix, iy, (float)(W - 1), (float)(H - 1), J) || isnan(ix) || isnan(iy)) return false;
    return true;
}

// The image sample at (ix, iy) (bilerp4): corner, fractions and whether each fraction moves with
// its coordinate (not at the clamped last cell).
struct ImgSample {
    int kx, ky;
    float ax, ay;
    float dax, day;
};

__device__ __forceinline__ void img_cell(float ix, float iy, int W, int H, ImgSample& s) {
    int32_t kx = cvtt_d(ix), ky = cvtt_d(iy);
    float ax = ix - (float)kx, ay = iy - (float)ky;
    s.dax = 1.0f;
    s.day = 1.0f;
    if (kx >= W - 1) { kx = W - 2; ax = 1.0f; s.dax = 0.0f; }
    if (ky >= H - 1) { ky = H - 2; ay = 1.0f; s.day = 0.0f; }
    s.kx = kx;
    s.ky = ky;
    s.ax = ax;
    s.ay = ay;
}

struct DiffArgs {
    const float* a;
    const float* b;
    const float* f_ab;
    const float* f_ba;
    const float* grad_out;
    float* out;
    float* grad_a;
    float* grad_b;
    float* gp_ba;   // per output pixel dL/d(flow value) of the warp of A
    float* gp_ab;   // of the warp of B
    float* grad_f_ab;
    float* grad_f_ba;
    int* bad;
    int C, W, H, fw, fh;
    float wa, frac;
    int copy_a, copy_b;  // the |amount| early-out of each warp: a plain copy
    int mode;            // the blend's branch (blend.rs:55-72): 0 copy A, 1 copy B, 3 per-sample mix
};

// One warp at (x, y): the sampled value (C channels), and with `g` the backward: the scatter into
// the image gradient and the flow-value gradient `gflow`.
template <bool BWD>
__device__ __forceinline__ bool warp_one(const DiffArgs& P, const float* __restrict__ src, const float* __restrict__ F, float amount,
                                         int copy, int x, int y, float* v, const float* g, float* gsrc, float gflow[2]) {
    const int C = P.C, W = P.W, H = P.H;
    const long pi = (long)y * W + x;
    if (copy) {
        for (int c = 0; c < C; c++) v[c] = src[pi * C + c];
        if (BWD) {
            if (gsrc)
                for (int c = 0; c < C; c++) atomicAdd(&gsrc[pi * C + c], g[c]);
            gflow[0] = gflow[1] = 0.0f;
        }
        return true;
    }
    FlowSample fs;
    float ix, iy, J[2][2];
    if (!warp_point_jac(F, P.fw, P.fh, W, H, amount, x, y, fs, ix, iy, J)) {
        for (int c = 0; c < C; c++) v[c] = 0.0f;
        if (BWD) gflow[0] = gflow[1] = 0.0f;
        return false;
    }
    ImgSample s;
    img_cell(ix, iy, W, H, s);
    const long i00 = (long)s.ky * W + s.kx, i01 = i00 + W;
    float dlx = 0.0f, dly = 0.0f;
    for (int c = 0; c < C; c++) {
        const float p00 = src[i00 * C + c], p10 = src[(i00 + 1) * C + c];
        const float p01 = src[i01 * C + c], p11 = src[(i01 + 1) * C + c];
        const float top = p00 + (p10 - p00) * s.ax;
        const float bot = p01 + (p11 - p01) * s.ax;
        v[c] = top + (bot - top) * s.ay;
        if (BWD) {
            const float dvx = (p10 - p00) * (1.0f - s.ay) + (p11 - p01) * s.ay;
            const float dvy = bot - top;
            dlx += g[c] * dvx;
            dly += g[c] * dvy;
            if (gsrc) {
                atomicAdd(&gsrc[i00 * C + c], g[c] * ((1.0f - s.ax) * (1.0f - s.ay)));
                atomicAdd(&gsrc[(i00 + 1) * C + c], g[c] * (s.ax * (1.0f - s.ay)));
                atomicAdd(&gsrc[i01 * C + c], g[c] * ((1.0f - s.ax) * s.ay));
                atomicAdd(&gsrc[(i01 + 1) * C + c], g[c] * (s.ax * s.ay));
            }
        }
    }
    if (BWD) {
        dlx *= s.dax;
        dly *= s.day;
        gflow[0] = J[0][0] * dlx + J[1][0] * dly;
        gflow[1] = J[0][1] * dlx + J[1][1] * dly;
    }
    return true;
}

// Pixel (x, y) of block b of a row-major grid of 256-pixel row segments, the blocks taken in XCD
// order: the hardware deals block b to XCD b % 8, and segment (b % 8) * ceil(n / 8) + b / 8 gives
// each XCD a contiguous band of rows, whose source rows its own L2 then serves to the rows below
// (with 5 segments per 1280-pixel row the plain order spread each source row over every XCD).
__device__ __forceinline__ bool wb_pixel(const DiffArgs& P, int& x, int& y) {
    const int per_row = (P.W + 255) / 256, n = per_row * P.H;
    const int per = (n + 7) / 8, b = blockIdx.x;
    const int seg = (b & 7) * per + (b >> 3);
    if (seg >= n) return false;
    y = seg / per_row;
    x = (seg - y * per_row) * 256 + threadIdx.x;
    return x < P.W;
}

__global__ void k_wb_forward(DiffArgs P) {
    int x, y;
    if (!wb_pixel(P, x, y)) return;
    float ta[4], tb[4];
    bool ok = warp_one<false>(P, P.a, P.f_ba, P.frac, P.copy_a, x, y, ta, nullptr, nullptr, nullptr);
    ok = warp_one<false>(P, P.b, P.f_ab, P.wa, P.copy_b, x, y, tb, nullptr, nullptr, nullptr) && ok;
    if (!ok) atomicOr(P.bad, 1);
    const long pi = (long)y * P.W + x;
    for (int c = 0; c < P.C; c++)
        P.out[pi * P.C + c] = P.mode == 0 ? ta[c] : (P.mode == 1 ? tb[c] : ((tb[c] - ta[c]) * P.frac + ta[c]) + 0.0f);
}

__global__ void k_wb_backward(DiffArgs P) {
    int x, y;
    if (!wb_pixel(P, x, y)) return;
    const long pi = (long)y * P.W + x;
    float g[4], ga[4], gb[4], v[4];
    for (int c = 0; c < P.C; c++) {
        g[c] = P.grad_out[pi * P.C + c];
        // out = tA (1 - w) + tB w, or one of the two
        ga[c] = P.mode == 0 ? g[c] : (P.mode == 1 ? 0.0f : g[c] * (1.0f - P.frac));
        gb[c] = P.mode == 0 ? 0.0f : (P.mode == 1 ? g[c] : g[c] * P.frac);
    }
    float gfa[2], gfb[2];
    bool ok = warp_one<true>(P, P.a, P.f_ba, P.frac, P.copy_a, x, y, v, ga, nullptr, gfa);
    ok = warp_one<true>(P, P.b, P.f_ab, P.wa, P.copy_b, x, y, v, gb, nullptr, gfb) && ok;
    if (!ok) atomicOr(P.bad, 1);
    P.gp_ba[pi * 2] = gfa[0];
    P.gp_ba[pi * 2 + 1] = gfa[1];
    P.gp_ab[pi * 2] = gfb[0];
    P.gp_ab[pi * 2 + 1] = gfb[1];
}

// dL/dflow per flow texel (i, j): the sum over the output pixels whose flow sample has (i, j) as
// a corner of the corner's bilinear weight times the pixel's flow-value gradient, in a fixed order
// (rows, then columns). blockIdx.z: 0 the flow of the warp of A (F_BA), 1 of B (F_AB); a null
// output is skipped. The candidate window is one pixel wider on each side than the cells' exact
// reach (the float product x * rx may round across a cell boundary); each pixel is then tested.
__global__ void k_wb_flow_gather(DiffArgs P) {
    const int which = blockIdx.z;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const int fw = P.fw, fh = P.fh, W = P.W, H = P.H;
    if (i >= fw || j >= fh) return;
    const float* gp = which ? P.gp_ab : P.gp_ba;
    float* out = which ? P.grad_f_ab : P.grad_f_ba;
    if (!out) return;
    const float rx = (float)(fw - 1) / (float)(W - 1), ry = (float)(fh - 1) / (float)(H - 1);
    // output columns whose cell index is i - 1 or i (with the last-cell clamp), and rows likewise
    const int x0 = max(0, (int)floorf((float)(i - 1) / rx) - 1), x1 = min(W - 1, (int)ceilf((float)(i + 1) / rx) + 1);
    const int y0 = max(0, (int)floorf((float)(j - 1) / ry) - 1), y1 = min(H - 1, (int)ceilf((float)(j + 1) / ry) + 1);
    float s0 = 0.0f, s1 = 0.0f;
    for (int y = y0; y <= y1; y++) {
        const float qy = (float)y * ry;
        int32_t jy = cvtt_d(qy);
        float gy = qy - (float)jy;
        if (jy >= fh - 1) { jy = fh - 2; gy = 1.0f; }
        float wy;
        if (jy == j) wy = 1.0f - gy;
        else if (jy + 1 == j) wy = gy;
        else continue;
        for (int x = x0; x <= x1; x++) {
            const float qx = (float)x * rx;
            int32_t jx = cvtt_d(qx);
            float gx = qx - (float)jx;
            if (jx >= fw - 1) { jx = fw - 2; gx = 1.0f; }
            float wx;
            if (jx == i) wx = 1.0f - gx;
            else if (jx + 1 == i) wx = gx;
            else continue;
            const float wgt = wx * wy;
            const long pi = (long)y * W + x;
            s0 += gp[pi * 2] * wgt;
            s1 += gp[pi * 2 + 1] * wgt;
        }
    }
    out[((long)j * fw + i) * 2] = s0;
    out[((long)j * fw + i) * 2 + 1] = s1;
}

// ---- image gradients in fixed point ----
//
// flags[0] is the NaN-point flag, flags[1] the largest |upstream gradient| as float bits (an
// atomic max on the bits orders non-negative floats). A contribution v is summed as
// rint(v * 2^k) in 64 bits with k = 33 - exponent(max): each term is below 2^34 in magnitude,
// so even 2^28 of them (every pixel of the largest frame on one texel) stay below 2^62.

__device__ __forceinline__ double fixed_scale(const int* flags) {
    const unsigned bits = (unsigned)flags[1];
    if (bits == 0u || bits >= 0x7f800000u) return 1.0;  // all zero, or not finite (refused by the host)
    return ldexp(1.0, 33 - ilogbf(__uint_as_float(bits)));
}

__global__ void k_wb_absmax(const float* __restrict__ g, size_t n, int* flags) {
    unsigned m = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        m = max(m, __float_as_uint(g[i]) & 0x7fffffffu);
    for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
    if ((threadIdx.x & 63) == 0 && m) atomicMax((unsigned*)&flags[1], m);
}

__global__ void k_wb_acc_clear(unsigned long long* __restrict__ acc, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) acc[i] = 0ull;
}

// Channels c0 .. c0 + nch - 1 (nch <= 2) of one warp's image gradient (which: 0 the warp of A,
// 1 of B): each output pixel's upstream gradient through the blend, times the bilinear weight of
// each texel it sampled (the float products of the reference scatter), added in fixed point into
// acc[texel * 2 + channel - c0].
__global__ void k_wb_img_scatter(DiffArgs P, int which, int c0, int nch, unsigned long long* __restrict__ acc) {
    int x, y;
    if (!wb_pixel(P, x, y)) return;
    const int C = P.C, W = P.W, H = P.H;
    const long pi = (long)y * W + x;
    const double scale = fixed_scale(P.bad);
    float gc[2];
    for (int k = 0; k < nch; k++) {
        const float g = P.grad_out[pi * C + c0 + k];
        gc[k] = which == 0 ? (P.mode == 0 ? g : (P.mode == 1 ? 0.0f : g * (1.0f - P.frac)))
                           : (P.mode == 0 ? 0.0f : (P.mode == 1 ? g : g * P.frac));
    }
    auto add = [&](long texel, int k, float v) {
        atomicAdd(&acc[texel * 2 + k], (unsigned long long)__double2ll_rn((double)v * scale));
    };
    if (which == 0 ? P.copy_a : P.copy_b) {
        for (int k = 0; k < nch; k++) add(pi, k, gc[k]);
        return;
    }
    FlowSample fs;
    float ix, iy, J[2][2];
    if (!warp_point_jac(which ? P.f_ab : P.f_ba, P.fw, P.fh, W, H, which ? P.wa : P.frac, x, y, fs, ix, iy, J)) return;
    ImgSample s;
    img_cell(ix, iy, W, H, s);
    const long i00 = (long)s.ky * W + s.kx, i01 = i00 + W;
    for (int k = 0; k < nch; k++) {
        add(i00, k, gc[k] * ((1.0f - s.ax) * (1.0f - s.ay)));
        add(i00 + 1, k, gc[k] * (s.ax * (1.0f - s.ay)));
        add(i01, k, gc[k] * ((1.0f - s.ax) * s.ay));
        add(i01 + 1, k, gc[k] * (s.ax * s.ay));
    }
}

// The fixed-point sums of channels c0 .. c0 + nch - 1 into the float gradient, and the
// accumulator cleared for the next pass.
__global__ void k_wb_img_finish(unsigned long long* __restrict__ acc, float* __restrict__ grad, int c0, int nch, int C,
                                size_t n_px, const int* flags) {
    const double inv = 1.0 / fixed_scale(flags);  // a power of two: exact
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_px; i += stride)
        for (int k = 0; k < nch; k++) {
            grad[i * C + c0 + k] = (float)((double)(long long)acc[i * 2 + k] * inv);
            acc[i * 2 + k] = 0ull;
        }
}

// ---- photometric refinement steps ----

// One step's forward, L2 loss and flow-value gradients: per output pixel the blend of both
// warps, d = out - truth, the loss term sum_c d^2 (block sums in a fixed order into
// loss_part[block]), and with `grad` the upstream gradient (2 d) / N through both warps into the
// per-pixel flow-value gradients (gp_ba, gp_ab), as k_wb_backward computes them.
__global__ __launch_bounds__(256) void k_rf_step(DiffArgs P, const float* __restrict__ truth, float inv_n,
                                                 float* __restrict__ loss_part, int grad) {
    __shared__ float red[256];
    int x, y;
    float sq = 0.0f;
    if (wb_pixel(P, x, y)) {
        const long pi = (long)y * P.W + x;
        float ta[4], tb[4];
        bool ok = warp_one<false>(P, P.a, P.f_ba, P.frac, P.copy_a, x, y, ta, nullptr, nullptr, nullptr);
        ok = warp_one<false>(P, P.b, P.f_ab, P.wa, P.copy_b, x, y, tb, nullptr, nullptr, nullptr) && ok;
        if (!ok) atomicOr(P.bad, 1);
        float ga[4], gb[4];
        for (int c = 0; c < P.C; c++) {
            const float o = P.mode == 0 ? ta[c] : (P.mode == 1 ? tb[c] : ((tb[c] - ta[c]) * P.frac + ta[c]) + 0.0f);
            const float d = o - truth[pi * P.C + c];
            sq = sq + d * d;
            const float g = (2.0f * d) * inv_n;
            ga[c] = P.mode == 0 ? g : (P.mode == 1 ? 0.0f : g * (1.0f - P.frac));
            gb[c] = P.mode == 0 ? 0.0f : (P.mode == 1 ? g : g * P.frac);
        }
        if (grad) {
            float v[4], gfa[2], gfb[2];
            warp_one<true>(P, P.a, P.f_ba, P.frac, P.copy_a, x, y, v, ga, nullptr, gfa);
            warp_one<true>(P, P.b, P.f_ab, P.wa, P.copy_b, x, y, v, gb, nullptr, gfb);
            P.gp_ba[pi * 2] = gfa[0];
            P.gp_ba[pi * 2 + 1] = gfa[1];
            P.gp_ab[pi * 2] = gfb[0];
            P.gp_ab[pi * 2 + 1] = gfb[1];
        }
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// The mean of the step's loss terms: the block sums in a fixed order, in double.
__global__ __launch_bounds__(256) void k_rf_loss(const float* __restrict__ loss_part, int n, double inv_n,
                                                 float* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += (double)loss_part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = (float)(red[0] * inv_n);
}

// One Adam update of both flows (blockIdx.y: 0 F_AB, 1 F_BA), in the operation order of
// torch.optim.Adam's default (foreach) step: m.lerp_(g, 1 - b1); v.mul_(b2).addcmul_(g, g, 1 - b2);
// f.addcdiv_(m, v.sqrt() / sqrt(1 - b2^t) + eps, -lr / (1 - b1^t)).
struct AdamArgs {
    float* f[2];
    const float* g[2];
    float* m[2];
    float* v[2];
    int n;
    float w1, b2, w2, bc2_sqrt, eps, neg_step;
};
__global__ void k_rf_adam(AdamArgs A) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= A.n) return;
    const float g = A.g[k][i];
    float m = A.m[k][i];
    m = m + A.w1 * (g - m);
    float v = A.v[k][i] * A.b2;
    v = v + (A.w2 * g) * g;
    A.m[k][i] = m;
    A.v[k][i] = v;
    const float denom = sqrtf(v) / A.bc2_sqrt + A.eps;
    A.f[k][i] = A.f[k][i] + A.neg_step * (m / denom);
}

}  // namespace

// The |amount| early-out of inv_warp (warp.rs:749-764).
static bool diff_warp_is_copy(float amount, int fw, int W) {
    float a = amount < 0.0f ? -amount : amount;
    if (0.005 > (double)a) return true;
    float r = (float)(fw - 1) / (float)(W - 1);
    float q = (float)((double)amount / (double)r);
    if (q < 0.0f) q = -q;
    return 0.0001 > (double)q;
}

// The arguments shared by the forward, the backward and the refinement; TWX_E_* on a shape or
// weight the kernels do not serve.
static int diff_setup(int C, int W, int H, const float* a, const float* b, const float* f_ab, const float* f_ba, int fw,
                      int fh, float wa, DiffArgs& P, std::string& err) {
    if (C < 1 || C > 4 || W < 2 || H < 2 || fw < 2 || fh < 2) {
        err = "warp+blend: 1 to 4 channels, images and flows of at least 2x2";
        return TWX_E_MISMATCH;
    }
    if (int st = check_frame_pixels(W, H, err)) return st;
    P = DiffArgs{};
    P.a = a;
    P.b = b;
    P.f_ab = f_ab;
    P.f_ba = f_ba;
    P.C = C;
    P.W = W;
    P.H = H;
    P.fw = fw;
    P.fh = fh;
    P.wa = wa;
    P.frac = 1.0f - wa;
    P.copy_a = diff_warp_is_copy(P.frac, fw, W) ? 1 : 0;
    P.copy_b = diff_warp_is_copy(wa, fw, W) ? 1 : 0;
    const float w = P.frac;
    if (0.005 > (double)std::fabs(w)) P.mode = 0;
    else if (0.005 > std::fabs((double)w - 1.0)) P.mode = 1;
    else if (w < 0.0f || w > 1.0f) {
        err = "warp+blend: a blend weight 1 - wa outside [0, 1] (the extrapolating lerp) has no backward here";
        return TWX_E_UNSUPPORTED;
    } else P.mode = 3;
    return TWX_OK;
}

static dim3 wb_grid(int W, int H) {
    const long segs = (long)((W + 255) / 256) * H;
    return dim3((unsigned)(8 * ((segs + 7) / 8)));  // wb_pixel's XCD order
}

int warp_blend_run(hipStream_t s, bool backward, int C, int W, int H, const float* a, const float* b, const float* f_ab,
                   const float* f_ba, int fw, int fh, float wa, float* out, const float* grad_out, float* grad_a,
                   float* grad_b, float* grad_f_ab, float* grad_f_ba, float* gp_scratch, int* flags, std::string& err) {
    DiffArgs P;
    if (int st = diff_setup(C, W, H, a, b, f_ab, f_ba, fw, fh, wa, P, err)) return st;
    P.out = out;
    P.grad_out = grad_out;
    P.grad_a = grad_a;
    P.grad_b = grad_b;
    P.grad_f_ab = grad_f_ab;
    P.grad_f_ba = grad_f_ba;
    P.gp_ba = gp_scratch;
    P.gp_ab = gp_scratch + (size_t)W * H * 2;
    P.bad = flags;
    dim3 blk(256), grd = wb_grid(W, H);
    const double img = 4.0 * C * (double)W * H, fl = 8.0 * (double)fw * fh;
    if (!backward) {
        TWX_LAUNCH(s, F_DIFF_FWD, 3.0 * img + 2.0 * fl, k_wb_forward<<<grd, blk, 0, s>>>(P));
        return TWX_OK;
    }
    // reads A, B, both flows and the upstream gradient, writes the two per-pixel flow-value
    // gradients
    TWX_LAUNCH(s, F_DIFF_BWD, 3.0 * img + 2.0 * fl + 16.0 * (double)W * H, k_wb_backward<<<grd, blk, 0, s>>>(P));
    dim3 fgrd((fw + 255) / 256, fh, 2);
    if (grad_f_ba || grad_f_ab)
        TWX_LAUNCH(s, F_DIFF_GATHER, (grad_f_ba && grad_f_ab ? 2.0 : 1.0) * (8.0 * (double)W * H + fl),
                   k_wb_flow_gather<<<fgrd, blk, 0, s>>>(P));
    if (!grad_a && !grad_b) return TWX_OK;
    // the image gradients, two channels a pass, in fixed point in the scratch the flow-value
    // gradients held (16 bytes a pixel: two 64-bit sums)
    const size_t n_px = (size_t)W * H, n = n_px * C;
    const unsigned cg = (unsigned)std::min<size_t>((n_px + 255) / 256, 2048);
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(gp_scratch);
    TWX_LAUNCH(s, F_DIFF_IMG, 4.0 * n, (k_wb_absmax<<<(unsigned)std::min<size_t>((n + 255) / 256, 2048), blk, 0, s>>>(grad_out, n, flags)));
    TWX_LAUNCH(s, F_DIFF_IMG, 16.0 * n_px, (k_wb_acc_clear<<<cg, blk, 0, s>>>(acc, n_px * 2)));
    for (int which = 0; which < 2; which++) {
        float* g = which ? grad_b : grad_a;
        if (!g) continue;
        for (int c0 = 0; c0 < C; c0 += 2) {
            const int nch = std::min(2, C - c0);
            TWX_LAUNCH(s, F_DIFF_IMG, 4.0 * nch * (double)n_px + fl + 4.0 * 8.0 * nch * (double)n_px,
                       (k_wb_img_scatter<<<grd, blk, 0, s>>>(P, which, c0, nch, acc)));
            TWX_LAUNCH(s, F_DIFF_IMG, 12.0 * nch * (double)n_px,
                       (k_wb_img_finish<<<cg, blk, 0, s>>>(acc, g, c0, nch, C, n_px, flags)));
        }
    }
    return TWX_OK;
}

int refine_loss_blocks(int W, int H) { return (int)wb_grid(W, H).x; }

int refine_flows_run(hipStream_t s, int C, int W, int H, const float* a, const float* b, const float* truth, float* f_ab,
                     float* f_ba, int fw, int fh, float wa, int steps, float lr, const RefineBufs& B, std::string& err) {
    DiffArgs P;
    if (int st = diff_setup(C, W, H, a, b, f_ab, f_ba, fw, fh, wa, P, err)) return st;
    P.grad_f_ab = B.g_ab;
    P.grad_f_ba = B.g_ba;
    P.gp_ba = B.gp;
    P.gp_ab = B.gp + (size_t)W * H * 2;
    P.bad = B.flags;
    dim3 blk(256), grd = wb_grid(W, H), fgrd((fw + 255) / 256, fh, 2);
    const double N = (double)C * W * H, img = 4.0 * N, fl = 8.0 * (double)fw * fh;
    const float inv_n = (float)(1.0 / N);
    const int nb = (int)grd.x;
    AdamArgs A{};
    A.f[0] = f_ab;
    A.f[1] = f_ba;
    A.g[0] = B.g_ab;
    A.g[1] = B.g_ba;
    A.m[0] = B.m_ab;
    A.m[1] = B.m_ba;
    A.v[0] = B.v_ab;
    A.v[1] = B.v_ba;
    A.n = fw * fh * 2;
    const double b1 = 0.9, b2 = 0.999;
    A.w1 = (float)(1.0 - b1);
    A.b2 = (float)b2;
    A.w2 = (float)(1.0 - b2);
    A.eps = 1e-8f;
    for (int t = 0; t <= steps; t++) {
        const int grad = t < steps;
        // reads A, B, the truth and both flows, writes the per-pixel flow-value gradients
        TWX_LAUNCH(s, F_DIFF_STEP, 3.0 * img + 2.0 * fl + (grad ? 16.0 * (double)W * H : 0.0),
                   (k_rf_step<<<grd, blk, 0, s>>>(P, truth, inv_n, B.loss_part, grad)));
        TWX_LAUNCH(s, F_DIFF_STEP, 4.0 * nb, (k_rf_loss<<<1, blk, 0, s>>>(B.loss_part, nb, 1.0 / N, B.losses + t)));
        if (!grad) break;
        TWX_LAUNCH(s, F_DIFF_GATHER, 2.0 * (8.0 * (double)W * H + fl), k_wb_flow_gather<<<fgrd, blk, 0, s>>>(P));
        const double bc1 = 1.0 - std::pow(b1, (double)(t + 1)), bc2 = 1.0 - std::pow(b2, (double)(t + 1));
        A.bc2_sqrt = (float)std::sqrt(bc2);
        A.neg_step = (float)((lr / bc1) * -1.0);
        // reads the flow, its gradient and both moments, writes the flow and the moments
        TWX_LAUNCH(s, F_DIFF_ADAM, 2.0 * 28.0 * (double)A.n,
                   (k_rf_adam<<<dim3((unsigned)((A.n + 255) / 256), 2), blk, 0, s>>>(A)));
    }
    return TWX_OK;
}

// ---- the backward without device allocation ---------------------------------------------------
//
// The backward's scratch (the per-pixel flow-value gradients of both warps, 16 bytes a pixel, and
// the NaN flag) lives in a fixed device array of this library instead of memory taken per call:
// 256 bytes of flags, then the gradients of up to NOALLOC_MAX_PIXELS pixels. Each device holds its
// own copy until the process ends. In host code `scratch` only names the symbol: its device
// address on each device comes from hipGetSymbolAddress, taken once per device with that device
// current (at twx_open), and only that address is ever handed to kernels.

constexpr size_t NOALLOC_FLAG_BYTES = 256;
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_FLAG_BYTES + (size_t)TWX_NOALLOC_MAX_PIXELS * 16;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Zeroes the NaN flag and the largest-gradient bits: a kernel, not hipMemsetAsync, so a caller
// may capture the call's launches into a graph.
__global__ void k_wb_clear(int* flags) {
    flags[0] = 0;
    flags[1] = 0;
}

}  // namespace

int noalloc_scratch_setup(int device, std::string& err) {
    if (device < 0 || device >= MAX_DEVICES) {
        err = "device index beyond the allocation-free scratch table";
        return TWX_E_DEVICE;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return TWX_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return TWX_E_DEVICE;
    }
    g_scratch_ptr[device] = p;
    return TWX_OK;
}

int warp_blend_backward_noalloc(hipStream_t s, int device, int C, int W, int H, const float* a, const float* b,
                                const float* f_ab, const float* f_ba, int fw, int fh, float wa, const float* grad_out,
                                float* grad_a, float* grad_b, float* grad_f_ab, float* grad_f_ba, std::string& err) {
    if (C < 1 || C > 4 || W < 2 || H < 2 || fw < 2 || fh < 2) {
        err = "warp+blend: 1 to 4 channels, images and flows of at least 2x2";
        return TWX_E_MISMATCH;
    }
    if ((uint64_t)W * (uint64_t)H > (uint64_t)TWX_NOALLOC_MAX_PIXELS) {
        err = "warp+blend backward without allocation: " + std::to_string(W) + "x" + std::to_string(H) +
              " is over its fixed scratch of " + std::to_string(TWX_NOALLOC_MAX_PIXELS) + " pixels";
        return TWX_E_UNSUPPORTED;
    }
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        if (device >= 0 && device < MAX_DEVICES) base = g_scratch_ptr[device];
    }
    if (!base) {
        err = "the allocation-free scratch of this device was not set up (twx_open)";
        return TWX_E_DEVICE;
    }
    int* bad = (int*)base;
    float* gp = (float*)((unsigned char*)base + NOALLOC_FLAG_BYTES);
    k_wb_clear<<<1, 1, 0, s>>>(bad);
    if (hipGetLastError() != hipSuccess) {
        err = "warp+blend backward: the clear launch failed";
        return TWX_E_DEVICE;
    }
    int st = warp_blend_run(s, true, C, W, H, a, b, f_ab, f_ba, fw, fh, wa, nullptr, grad_out, grad_a, grad_b, grad_f_ab,
                            grad_f_ba, gp, bad, err);
    if (st) return st;
    int hb[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(hb, bad, sizeof(hb), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        err = std::string("warp+blend backward: ") + hipGetErrorString(e);
        return TWX_E_DEVICE;
    }
    return diff_flags_status(hb, err);
}

int diff_flags_status(const int flags[2], std::string& err) {
    if (flags[0]) {
        err = "warp+blend: a sample point is NaN or never converges";
        return TWX_E_UNSUPPORTED;
    }
    if ((unsigned)flags[1] >= 0x7f800000u) {
        err = "warp+blend: the upstream gradient holds an infinity or a NaN";
        return TWX_E_UNSUPPORTED;
    }
    return TWX_OK;
}

}  // namespace twx
