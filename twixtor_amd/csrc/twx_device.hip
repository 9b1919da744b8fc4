// HIP kernels of the frame-synthesis path for gfx950 (CDNA4), bit-exact with the reference CPU
// path. Every kernel follows the CPU's operation order (plugins/retime-engine/src/flow.rs,
// flow/conv.rs, flow/sample.rs, prep.rs, warp.rs, blend.rs, image.rs); the library is compiled
// with -ffp-contract=off, IEEE denormals and correctly rounded division, so every float operation
// is the one the CPU performs.
//
// Batching: every kernel takes a table of jobs and maps blockIdx.z onto it, so one launch serves
// one pair, every pair of a clip, or the pairs of many clips of different sizes (blocks outside a
// job's extent exit at once).

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include <climits>
#include <cmath>

#include "twx_common.hpp"
#include "twx_pixel.hpp"

namespace twx {

// -------------------------------------------------------------------------------------------
// Arithmetic helpers with the reference's semantics

// the hardware's float -> int32 truncation (v_cvt_i32_f32: NaN gives 0, out of range saturates),
// for callers that only use the result where cvtt gives the same value
__device__ __forceinline__ int32_t trunc_hw(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
// global-memory views of device pointers (global loads rather than flat loads)
typedef __attribute__((address_space(1))) const float gfloat;
typedef __attribute__((address_space(1))) float gfloat_rw;
typedef __attribute__((address_space(1))) const char gchar;

// element i of a plane whose base is wave-uniform: a 32-bit byte offset from that base (planes
// hold fewer than 2^30 elements), so the address needs no 64-bit arithmetic per lane
__device__ __forceinline__ float ldg(const gfloat* base, uint32_t i) { return *(const gfloat*)((gchar*)base + i * 4u); }

// N consecutive elements starting at element i (one wide load where the compiler can)
template <int N>
__device__ __forceinline__ void ldrun(const gfloat* base, uint32_t i, float out[N]) {
    gchar* q = (gchar*)base + i * 4u;
#pragma unroll
    for (int j = 0; j < N; j++) out[j] = *(const gfloat*)(q + 4 * j);
}
// the N channels of pixel i of an N-channel plane
template <int N>
__device__ __forceinline__ void ldgn(const gfloat* base, uint32_t i, float out[N]) {
    ldrun<N>(base, i * N, out);
}




// -------------------------------------------------------------------------------------------
// Wire -> engine image (image.rs:173-203, frame.rs:79-111, image.rs:264-298)

// The engine pixel of wire pixel si (n: the frame's pixels, for the f32 planes).
__device__ __forceinline__ uchar4 engine_px_u8(const uint8_t* __restrict__ wire, long si) {
    uchar4 c = reinterpret_cast<const uchar4*>(wire)[si];
    float a = (float)c.w * bits_f(0x3b808081u);
    uchar4 o;
    o.x = c.w;
    o.y = (uint8_t)qround((float)c.x * a);
    o.z = (uint8_t)qround((float)c.y * a);
    o.w = (uint8_t)qround((float)c.z * a);
    return o;
}
__device__ __forceinline__ ushort4 engine_px_u16(const uint8_t* __restrict__ wire, long si) {
    ushort4 c = reinterpret_cast<const ushort4*>(wire)[si];
    auto q15 = [](uint32_t v) -> uint32_t { return (v * 32768u + 32767u) / 65535u; };
    uint32_t r = q15(c.x), g = q15(c.y), b = q15(c.z), al = q15(c.w);
    float a = (float)al * bits_f(0x387fff00u);
    ushort4 o;
    o.x = (uint16_t)qround((float)al * bits_f(0x3fffff00u));
    o.y = (uint16_t)qround((float)r * a);
    o.z = (uint16_t)qround((float)g * a);
    o.w = (uint16_t)qround((float)b * a);
    return o;
}
__device__ __forceinline__ float4 engine_px_f32(const uint8_t* __restrict__ wire, long si, long n) {
    const float* pl = reinterpret_cast<const float*>(wire);  // planes G, B, R, A
    float g = pl[si], b = pl[n + si], r = pl[2 * n + si], a = pl[3 * n + si];
    float4 o = make_float4(a, a * r, a * g, a * b);
    if (isnan(o.x) || isnan(o.y) || isnan(o.z) || isnan(o.w))  // image.rs:291-300, the NaN rule
        o = make_float4(a, nr_mul(a, r), nr_mul(a, g), nr_mul(a, b));
    return o;
}

template <int FMT>
__device__ __forceinline__ void ingest_pixel(const uint8_t* __restrict__ wire, uint8_t* __restrict__ engine, int W, int H,
                                             int x, int y, int reverse) {
    long n = (long)W * H;
    long si = (long)y * W + x;
    int yo = reverse ? H - 1 - y : y;
    long di = (long)yo * W + x;
    if (FMT == TWX_FMT_U8) reinterpret_cast<uchar4*>(engine)[di] = engine_px_u8(wire, si);
    else if (FMT == TWX_FMT_U16) reinterpret_cast<ushort4*>(engine)[di] = engine_px_u16(wire, si);
    else reinterpret_cast<float4*>(engine)[di] = engine_px_f32(wire, si, n);
}

template <int FMT>
__global__ void k_ingest(const uint8_t* __restrict__ wire, uint8_t* __restrict__ engine, int W, int H, int reverse) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    if (x >= W) return;
    ingest_pixel<FMT>(wire, engine, W, H, x, y, reverse);
}

// Every source frame of a batch in one launch per format (blockIdx.z = op); a transparent frame
// (outside the clip) is written as zeros.
template <int FMT>
__global__ void k_ingest_batch(const IngestOp* __restrict__ ops) {
    const IngestOp& op = ops[blockIdx.z];
    if (op.fmt != FMT) return;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= op.W || y >= op.H) return;
    if (op.zero) {
        const long i = (long)y * op.W + x;
        if (FMT == TWX_FMT_U8) reinterpret_cast<uchar4*>(op.engine)[i] = make_uchar4(0, 0, 0, 0);
        else if (FMT == TWX_FMT_U16) reinterpret_cast<ushort4*>(op.engine)[i] = make_ushort4(0, 0, 0, 0);
        else reinterpret_cast<float4*>(op.engine)[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    ingest_pixel<FMT>(op.wire, op.engine, op.W, op.H, x, y, op.reverse);
}

std::vector<std::pair<int, int>> size_buckets(const std::vector<double>& areas) {
    std::vector<std::pair<int, int>> out;
    int n = (int)areas.size();
    for (int b0 = 0; b0 < n;) {
        int b1 = b0 + 1;
        while (b1 < n && areas[b1] * 2.0 >= areas[b0]) b1++;
        out.push_back({b0, b1});
        b0 = b1;
    }
    return out;
}

void launch_ingest_batch(hipStream_t s, const IngestOp* d_ops, const std::vector<IngestOp>& ops) {
    std::vector<double> area;
    for (auto& o : ops) area.push_back((double)o.W * o.H);
    for (auto [b0, b1] : size_buckets(area)) {
        int mw = 0, mh = 0;
        bool has[3] = {false, false, false};
        double bytes = 0;
        for (int i = b0; i < b1; i++) {
            mw = std::max(mw, ops[i].W);
            mh = std::max(mh, ops[i].H);
            has[ops[i].fmt] = true;
            bytes += (ops[i].zero ? 1.0 : 2.0) * image_bytes(ops[i].W, ops[i].H, ops[i].fmt);
        }
        dim3 blk(256), grd((mw + 255) / 256, mh, b1 - b0);
        if (has[0]) TWX_LAUNCH(s, F_INGEST, bytes, k_ingest_batch<0><<<grd, blk, 0, s>>>(d_ops + b0));
        if (has[1]) TWX_LAUNCH(s, F_INGEST, bytes, k_ingest_batch<1><<<grd, blk, 0, s>>>(d_ops + b0));
        if (has[2]) TWX_LAUNCH(s, F_INGEST, bytes, k_ingest_batch<2><<<grd, blk, 0, s>>>(d_ops + b0));
    }
}

void launch_ingest(hipStream_t s, const void* wire, void* engine, int W, int H, int fmt, bool reverse_rows) {
    dim3 blk(256), grd((W + 255) / 256, H);
    double bytes = 2.0 * image_bytes(W, H, fmt);
    if (fmt == TWX_FMT_U8)
        TWX_LAUNCH(s, F_INGEST, bytes, k_ingest<0><<<grd, blk, 0, s>>>((const uint8_t*)wire, (uint8_t*)engine, W, H, reverse_rows));
    else if (fmt == TWX_FMT_U16)
        TWX_LAUNCH(s, F_INGEST, bytes, k_ingest<1><<<grd, blk, 0, s>>>((const uint8_t*)wire, (uint8_t*)engine, W, H, reverse_rows));
    else
        TWX_LAUNCH(s, F_INGEST, bytes, k_ingest<2><<<grd, blk, 0, s>>>((const uint8_t*)wire, (uint8_t*)engine, W, H, reverse_rows));
}

// ---- profiler (host) ----

const char* const PROF_NAMES[F_NFAM] = {"ingest", "pyr_luma_blur5_band_resample",
                                        "tensor_gauss", "bm_transpose", "solve_init", "calc_b_gauss", "new_disp_gauss",
                                        "bm_rows", "bm_cols", "range", "upsample", "negate", "frames",
                                        "wb_mask", "wb_sharpen_blend", "fill_pull", "fill_push", "fill_io",
                                        "warp_blend_fwd", "warp_blend_bwd", "warp_blend_flow_gather",
                                        "warp_blend_image_grad", "refine_step", "refine_adam",
                                        "fw_mesh", "fw_raster", "fw_scan", "fw_fold", "vector_coding", "coarse_solve"};
thread_local Prof* g_prof = nullptr;

hipEvent_t Prof::take() {
    if (!spare.empty()) {
        hipEvent_t e = spare.back();
        spare.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void Prof::flush() {
    for (auto& r : pending) {
        float t = 0.0f;
        if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
            ms[r.fam] += t;
            bytes[r.fam] += r.bytes;
            launches[r.fam] += 1;
        }
        spare.push_back(r.a);
        spare.push_back(r.b);
    }
    pending.clear();
}

void Prof::reset() {
    for (int i = 0; i < F_NFAM; i++) ms[i] = bytes[i] = 0, launches[i] = 0;
}

Prof::~Prof() {
    for (auto& r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : spare) (void)hipEventDestroy(e);
}

hipEvent_t prof_begin(hipStream_t s) {
    if (!g_prof || !g_prof->on) return nullptr;
    hipEvent_t a = g_prof->take();
    (void)hipEventRecord(a, s);
    return a;
}

void prof_end(hipStream_t s, int fam, double bytes, hipEvent_t a) {
    if (!a || !g_prof) return;
    hipEvent_t b = g_prof->take();
    (void)hipEventRecord(b, s);
    g_prof->pending.push_back(Prof::Rec{fam, a, b, bytes});
}

// -------------------------------------------------------------------------------------------
// Luma (prep.rs:168-235, retime path: no alpha, no video range)

__device__ __forceinline__ float luma_px(uchar4 p) {
    float r = (float)p.y, g = (float)p.z, b = (float)p.w;
    return (r * 0.3f + g * 0.59f) + b * 0.11f;
}
__device__ __forceinline__ float luma_px(ushort4 p) {
    double r = p.y, g = p.z, b = p.w;
    return (float)(((g * 0.59 + r * 0.3) + b * 0.11) / 257.0);
}
__device__ __forceinline__ float luma_px(float4 p) {
    double r = p.y, g = p.z, b = p.w;
    return (float)(((g * 0.59 + r * 0.3) + b * 0.11) * 255.0);
}
__device__ __forceinline__ float luma_at(const void* img, int fmt, long i) {
    if (fmt == TWX_FMT_U8) return luma_px(reinterpret_cast<const uchar4*>(img)[i]);
    if (fmt == TWX_FMT_U16) return luma_px(reinterpret_cast<const ushort4*>(img)[i]);
    return luma_px(reinterpret_cast<const float4*>(img)[i]);
}

// The luma of engine pixel (x, y) of a frame read from its wire bytes (the engine pixel as the
// ingest makes it, rows reversed with `reverse`), stored into the engine image `eng` when that is
// not null.
__device__ __forceinline__ float wire_luma(const uint8_t* __restrict__ wire, uint8_t* __restrict__ eng, int fmt, int W,
                                           int H, int reverse, int x, int y) {
    const long si = (long)(reverse ? H - 1 - y : y) * W + x, di = (long)y * W + x;
    if (fmt == TWX_FMT_U8) {
        const uchar4 p = engine_px_u8(wire, si);
        if (eng) reinterpret_cast<uchar4*>(eng)[di] = p;
        return luma_px(p);
    }
    if (fmt == TWX_FMT_U16) {
        const ushort4 p = engine_px_u16(wire, si);
        if (eng) reinterpret_cast<ushort4*>(eng)[di] = p;
        return luma_px(p);
    }
    const float4 p = engine_px_f32(wire, si, (long)W * H);
    if (eng) reinterpret_cast<float4*>(eng)[di] = p;
    return luma_px(p);
}

// -------------------------------------------------------------------------------------------
// Separable convolution (flow/conv.rs). The element arithmetic is shared by the pyramid planes
// and the solver's buffers.

// conv.rs:99-141 border element; get(t) gives the sample or false outside the line.
template <int C, bool LEFT_FIRST>
__device__ __forceinline__ float border_elem(const float* __restrict__ line, long stride, int len, int i, const Taps& t) {
    // `line` points at element 0 of the line for this channel; `stride` steps one pixel.
    float num = t.k[0] * line[(long)i * stride];
    float den = t.k[0];
#pragma unroll
    for (int j = MAX_TAPS - 1; j >= 1; j--) {
        if (j >= t.n) continue;
        float kj = t.k[j];
#pragma unroll
        for (int side = 0; side < 2; side++) {
            bool right = (side == 0) != LEFT_FIRST;
            int p = right ? i + j : i - j;
            bool in = p >= 0 && p < len;
            float xv, w;
            if (C == 1) {
                if (in) { xv = line[(long)p * stride]; w = kj; }
                else { xv = 0.0f; w = 0.0f; }
            } else {
                float v = in ? line[(long)p * stride] : -1.0f;
                float mask = (v != -1.0f) ? 1.0f : 0.0f;
                xv = v;
                w = mask * kj;
            }
            den = den + w;
            num = num + xv * w;
        }
    }
    return num / den;
}

// One output element of the horizontal pass (conv.rs:144-199) for pixel x, channel c.
template <int C>
__device__ __forceinline__ float hpass_elem(const float* __restrict__ row, int w, int x, int c, const Taps& t) {
    int n = t.n;
    int head = min(n, w);
    int interior = w > 2 * n ? w - 2 * n : 0;
    if (x >= head && x < head + interior) {
        const float* p = row + (long)x * C + c;
        float s = p[0] * t.k[0];
#pragma unroll
        for (int j = MAX_TAPS - 1; j >= 1; j--) {
            if (j >= n) continue;
            s = s + (p[j * C] + p[-j * C]) * t.k[j];
        }
        return s;
    }
    return border_elem<C, false>(row + c, C, w, x, t);
}

// One output element of the vertical pass (conv.rs:272-311, 584-606): border rows above first,
// or below first (the engine's default order) with ABOVE false.
template <int C, bool ABOVE = true>
__device__ __forceinline__ float vpass_elem(const float* __restrict__ src, int w, int h, int x, int y, int c, const Taps& t) {
    int n = t.n;
    int head = min(n, h);
    int interior = h > 2 * n ? h - 2 * n : 0;
    long row = (long)w * C;
    const float* col = src + (long)x * C + c;
    if (y >= head && y < head + interior) {
        const float* p = col + (long)y * row;
        float s = p[0] * t.k[0];
#pragma unroll
        for (int j = MAX_TAPS - 1; j >= 1; j--) {
            if (j >= n) continue;
            s = s + (p[j * row] + p[-j * row]) * t.k[j];
        }
        return s;
    }
    return border_elem<C, ABOVE>(col, row, h, y, t);
}

// -------------------------------------------------------------------------------------------
// Samplers (flow/sample.rs:31-112)

template <bool FINITE = false>
__device__ __forceinline__ void cell(float x, int len, int& io, float& fo) {
    // FINITE: the caller guarantees x in [0, len - 1] or NaN, and len >= 2; there the hardware
    // truncation gives cvtt's io and fo (for NaN both end at io = 0, fo NaN)
    int32_t i = FINITE ? trunc_hw(x) : cvtt(x);
    float f = x - (float)i;
    int32_t last = len - 1;
    if (i >= last) { i = last - 1; f = 1.0f; }
    if (FINITE) {  // i is in 0 .. len - 2 already: the clamps below leave it
        io = i;
        fo = f;
        return;
    }
    int ii = i < 0 ? 0 : i;
    int cap = len >= 2 ? len - 2 : 0;
    io = min(ii, cap);
    fo = f;
}

__device__ __forceinline__ float bilinear1(const float* __restrict__ img, int w, int h, float x, float y) {
    int ix, iy;
    float fx, fy;
    cell(x, w, ix, fx);
    cell(y, h, iy, fy);
    long i = (long)iy * w + ix;
    float p00 = img[i], p10 = img[i + 1], p01 = img[i + w], p11 = img[i + w + 1];
    float top = (p10 - p00) * fx + p00;
    float bot = (p11 - p01) * fx + p01;
    return (bot - top) * fy + top;
}

// -------------------------------------------------------------------------------------------
// Solver kernels (flow.rs:281-442), one launch per step and level over all jobs.
//
// Flow buffers of a job: a level starts in flow[0]; the three Gaussian iterations go 0 -> 1 ->
// 0 -> 1 (no kernel reads and writes one buffer, so tiles may overlap their halos freely); the
// upsample writes the next level into flow[0]. flow[1] is the caller's output, so level 0 ends
// in place.

// One element of a separable pass over a line held with a stride, `origin` being the line index
// of base[0]; the interior chain or the renormalised border (conv.rs:99-199, 272-311).
template <int R, bool LEFT_FIRST, bool MULTI>
__device__ __forceinline__ float line_elem(const float* __restrict__ base, int stride, int origin, int len, int i,
                                           const Taps& t) {
    constexpr int n = R + 1;
    const int head = min(n, len);
    const int interior = len > 2 * n ? len - 2 * n : 0;
    const float* p = base + (i - origin) * stride;
    if (i >= head && i < head + interior) {
        float s = p[0] * t.k[0];
#pragma unroll
        for (int j = R; j >= 1; j--) s = s + (p[j * stride] + p[-j * stride]) * t.k[j];
        return s;
    }
    float num = t.k[0] * p[0];
    float den = t.k[0];
#pragma unroll
    for (int j = R; j >= 1; j--) {
        const float kj = t.k[j];
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const bool right = (side == 0) != LEFT_FIRST;
            const int q = right ? i + j : i - j;
            const bool in = q >= 0 && q < len;
            float xv, wt;
            if (!MULTI) {
                xv = in ? p[(right ? j : -j) * stride] : 0.0f;
                wt = in ? kj : 0.0f;
            } else {
                float v = in ? p[(right ? j : -j) * stride] : -1.0f;
                float mask = (v != -1.0f) ? 1.0f : 0.0f;
                xv = v;
                wt = mask * kj;
            }
            den = den + wt;
            num = num + xv * wt;
        }
    }
    return num / den;
}

enum Produce { P_TENSOR = 0, P_CALCB = 1, P_NEWDISP = 2 };

// The value of the field the Gaussian runs on, at PU pixels (px[u], py[u]) of a lane at once:
//  P_TENSOR: calc_grad + calc_a (flow.rs:283-328); the tile's own pixels also store the gradient
//  P_CALCB:  calc_b (flow.rs:334-357), I1 = X, I0 = Y, with the flow in buffer `src`
//  P_NEWDISP: new_disp (flow.rs:361-371) applied to the flow in buffer `src`
// The pixels' loads are issued together, phase by phase (calc_b's samples of Y depend on its flow
// loads), through global-address-space pointers and 32-bit offsets; every load is unconditional
// (in-bounds addresses, results selected). The arithmetic of each pixel is the reference's.
template <int MODE, int C, int PU, bool INNER>
__device__ __forceinline__ void produce(const SolveJob& J, int level, int src, int w, int h, const int px[PU],
                                        const int py[PU], const bool own[PU], float v[PU][C]) {
    uint32_t e[PU];
#pragma unroll
    for (int u = 0; u < PU; u++) e[u] = (uint32_t)py[u] * w + px[u];
    if (MODE == P_TENSOR) {
        const gfloat* img = (const gfloat*)J.X[level];
        float c[PU], vl[PU], vu[PU], vr[PU], vd[PU];
#pragma unroll
        for (int u = 0; u < PU; u++) {
            const int x = px[u], y = py[u];
            c[u] = ldg(img, e[u]);
            vl[u] = ldg(img, e[u] - (INNER || x >= 1 ? 1 : 0));
            vu[u] = ldg(img, e[u] - (INNER || y >= 1 ? w : 0));
            vr[u] = ldg(img, e[u] + (INNER || x + 1 <= w - 1 ? 1 : 0));
            vd[u] = ldg(img, e[u] + (INNER || y + 1 <= h - 1 ? w : 0));
        }
#pragma unroll
        for (int u = 0; u < PU; u++) {
            const int x = px[u], y = py[u];
            float nx = 0.0f, dx = 0.0f, ny = 0.0f, dy = 0.0f;
            if (INNER || x >= 1) { nx = (c[u] - vl[u]) + 0.0f; dx = 1.0f; }
            if (INNER || y >= 1) { ny = (c[u] - vu[u]) + 0.0f; dy = 1.0f; }
            if (INNER || x + 1 <= w - 1) { nx = nx + (vr[u] - c[u]); dx = dx + 1.0f; }
            if (INNER || y + 1 <= h - 1) { ny = ny + (vd[u] - c[u]); dy = dy + 1.0f; }
            const float gx = nx / dx, gy = ny / dy;
            if (own[u]) {
                *(gfloat_rw*)((gchar*)J.grad + 8u * e[u]) = gx;
                *(gfloat_rw*)((gchar*)J.grad + 8u * e[u] + 4u) = gy;
            }
            v[u][0] = gx * gx;
            v[u][1] = gx * gy;
            v[u][C - 1] = gy * gy;
        }
    } else if (MODE == P_CALCB) {
        const gfloat* fl = (const gfloat*)J.flow[src];
        const gfloat* X = (const gfloat*)J.X[level];
        const gfloat* Y = (const gfloat*)J.Y[level];
        const gfloat* G = (const gfloat*)J.grad;
        float fu[PU], fv[PU], xv[PU], gx[PU], gy[PU];
#pragma unroll
        for (int u = 0; u < PU; u++) {
            float f2[2], g2[2];
            ldgn<2>(fl, e[u], f2);
            ldgn<2>(G, e[u], g2);
            fu[u] = f2[0];
            fv[u] = f2[1];
            xv[u] = ldg(X, e[u]);
            gx[u] = g2[0];
            gy[u] = g2[1];
        }
        bool in[PU];
        float fx[PU], fy[PU], p00[PU], p10[PU], p01[PU], p11[PU];
#pragma unroll
        for (int u = 0; u < PU; u++) {
            const float sx = (float)px[u] - fu[u];
            const float sy = (float)py[u] - fv[u];
            in[u] = !(sx < 0.0f || (float)w - 1.0f < sx || sy < 0.0f || (float)h - 1.0f < sy);
            // bilinear1 at (sx, sy), or at (0, 0) outside
            int ix, iy;
            cell<true>(in[u] ? sx : 0.0f, w, ix, fx[u]);
            cell<true>(in[u] ? sy : 0.0f, h, iy, fy[u]);
            const uint32_t i = (uint32_t)iy * w + ix;
            p00[u] = ldg(Y, i);
            p10[u] = ldg(Y, i + 1);
            p01[u] = ldg(Y, i + w);
            p11[u] = ldg(Y, i + w + 1);
        }
#pragma unroll
        for (int u = 0; u < PU; u++) {
            const float top = (p10[u] - p00[u]) * fx[u] + p00[u];
            const float bot = (p11[u] - p01[u]) * fx[u] + p01[u];
            const float s = (bot - top) * fy[u] + top;
            const float r = s - xv[u];
            v[u][0] = in[u] ? r * gx[u] : 0.0f;
            v[u][C - 1] = in[u] ? r * gy[u] : 0.0f;
        }
    } else {
        const gfloat* fl = (const gfloat*)J.flow[src];
        const gfloat* ta = (const gfloat*)J.ta;
        const gfloat* tb = (const gfloat*)J.tb;
        float fu[PU], fv[PU], aa[PU], ab[PU], ad[PU], bx[PU], by[PU];
#pragma unroll
        for (int u = 0; u < PU; u++) {
            float f2[2], a3[3], b2[2];
            ldgn<2>(fl, e[u], f2);
            ldgn<3>(ta, e[u], a3);
            ldgn<2>(tb, e[u], b2);
            fu[u] = f2[0];
            fv[u] = f2[1];
            aa[u] = a3[0];
            ab[u] = a3[1];
            ad[u] = a3[2];
            bx[u] = b2[0];
            by[u] = b2[1];
        }
#pragma unroll
        for (int u = 0; u < PU; u++) {
            float uu = fu[u], vv = fv[u];
            const float t = ab[u] * ab[u] - ad[u] * aa[u];
            if (fabsf(t) > 40.0f) {
                uu = (ab[u] * by[u] - ad[u] * bx[u]) / t + uu;
                vv = (ab[u] * bx[u] - aa[u] * by[u]) / t + vv;
            }
            v[u][0] = uu;
            v[u][C - 1] = vv;
        }
    }
}

constexpr int TILE_X = 64, TILE_Y = 32, TILE_THREADS = 256, TILE_RUN = 8;

// The interior formula (every tap inside the line) for TILE_RUN consecutive outputs of one lane
// from a register window: win[m + R + d] is the sample at offset d of output m. The operations are
// those of line_elem's interior branch, in the same order.
template <int R>
__device__ __forceinline__ void run_interior(const float* win, const Taps& t, float* out) {
#pragma unroll
    for (int m = 0; m < TILE_RUN; m++) {
        float s = win[m + R] * t.k[0];
#pragma unroll
        for (int j = R; j >= 1; j--) s = s + (win[m + R + j] + win[m + R - j]) * t.k[j];
        out[m] = s;
    }
}

// The border elements of a tile's horizontal pass: the tile's columns within nh = min(R + 1, w) of
// either end of the line, whose taps leave it and which take the renormalised formula. They form at
// most two runs per row, [x0, min(x0 + TILE_X, nh, w)) and [max(x0, nh + hint), min(x0 + TILE_X, w)),
// so at most 2 (R + 1) columns; the lanes of the block share them out element by element, while
// every run of the tile takes the interior formula (whose results the write-back drops for these
// columns). A run holding one border element otherwise sent its whole wave through the border
// formula for all TILE_RUN elements of every lane.
template <int R, int C, int RH, int RW, int NT>
struct HBorder {
    static constexpr int PER = (RH * 2 * (R + 1) + NT - 1) / NT;
    float v[PER][C];
    int x0, bl, brs, nbc;
    __device__ __forceinline__ HBorder(int tx0, int w) : x0(tx0) {
        const int nh = min(R + 1, w);
        const int hint = w > 2 * (R + 1) ? w - 2 * (R + 1) : 0;
        const int xe = min(x0 + TILE_X, w);
        bl = max(0, min(xe, nh) - x0);
        brs = max(x0, nh + hint);
        nbc = bl + max(0, xe - brs);
    }
    // element e of the list: region row e / nbc, the k-th border column of the tile
    __device__ __forceinline__ bool at(int e, int y0, int h, int& ry, int& x) const {
        if (e >= RH * nbc) return false;
        ry = e / nbc;
        const int k = e - ry * nbc;
        x = k < bl ? x0 + k : brs + (k - bl);
        const int y = y0 - R + ry;
        return y >= 0 && y < h;
    }
    template <bool MULTI>
    __device__ __forceinline__ void compute(const float* s_in, int y0, int w, int h, const Taps& t, int tid) {
        if (nbc == 0) return;
#pragma unroll
        for (int b = 0; b < PER; b++) {
            int ry, x;
            if (!at(tid + b * NT, y0, h, ry, x)) continue;
            const float* line = s_in + ry * RW * C;
#pragma unroll
            for (int c = 0; c < C; c++) v[b][c] = line_elem<R, false, MULTI>(line + c, C, x0 - R, w, x, t);
        }
    }
    __device__ __forceinline__ void store(float* s_in, int y0, int h, int tid) const {
        if (nbc == 0) return;
#pragma unroll
        for (int b = 0; b < PER; b++) {
            int ry, x;
            if (!at(tid + b * NT, y0, h, ry, x)) continue;
#pragma unroll
            for (int c = 0; c < C; c++) s_in[(ry * RW + R + x - x0) * C + c] = v[b][c];
        }
    }
};

// Tiles of a level in an order for the 8 XCDs: the hardware deals workgroup b to XCD b % 8, so
// tile (b % 8) * ceil(n / 8) + b / 8 gives each XCD a contiguous band of tiles (row-major within a
// job), whose halos the XCD's own L2 then serves.
__device__ __forceinline__ int xcd_tile(int n) {
    const int per = (n + 7) / 8;
    const int b = blockIdx.x;
    return (b & 7) * per + (b >> 3);
}

// A producer step fused with the separable Gaussian that follows it (flow.rs:1277-1287): the
// tile's region (tile plus the kernel's reach on every side) is produced into LDS, the horizontal
// pass runs over the region's rows (results back into the region's middle columns, all reads
// before any write), the vertical pass (border rows above first) over the tile. Each lane
// computes runs of TILE_RUN outputs from a register window where every tap lies inside the image,
// and the border formula element by element elsewhere. Every element is computed exactly as by
// the separate passes, so the bits are the same.
template <int C, int R, int MODE, int TY, int PUV>
__global__ __launch_bounds__(TILE_X * TY / TILE_RUN) void k_tile_gauss(const SolveJob* __restrict__ jobs, int level, int src,
                                                             int dst, Taps t, int tiles_x, int tiles_y, int njobs,
                                                             int scan) {
    const int per_job = tiles_x * tiles_y;
    const int lt = xcd_tile(per_job * njobs);
    if (lt >= per_job * njobs) return;
    const int job = lt / per_job, rem = lt - job * per_job;
    const int tyi = rem / tiles_x, txi = rem - tyi * tiles_x;
    const SolveJob& J = jobs[job];
    if (level > J.top || (MODE == P_TENSOR && J.tensor_shared)) return;
    const int w = J.w[level], h = J.h[level];
    const int x0 = txi * TILE_X, y0 = tyi * TY;
    constexpr int NT = TILE_X * TY / TILE_RUN;  // one vertical run of TILE_RUN rows per lane
    if (x0 >= w || y0 >= h) return;
    constexpr int RW = TILE_X + 2 * R, RH = TY + 2 * R;
    constexpr int RUNS_X = TILE_X / TILE_RUN, HITEMS = RH * RUNS_X;
    constexpr int HPER = (HITEMS + NT - 1) / NT;
    constexpr bool MULTI = C > 1;
    __shared__ float s_in[RH * RW * C];
    const int tid = threadIdx.x;
    // the region, PU pixels per lane at a time with their loads in flight together; a pixel
    // outside the image is computed at a clamped position and not stored
    constexpr int NPROD = (RH * RW + NT - 1) / NT, PU = PUV;
    // region element k = tid + j * NT of the lane's j-th pixel, as (ry, rx), stepped without division.
    // INNER: the region and a one-pixel margin lie inside the image (most tiles), so no pixel needs a
    // clamp or a border test; the lanes past the region's end compute at a safe inner pixel and
    // store nothing.
    auto region = [&](auto inner) {
        constexpr bool INNER = decltype(inner)::value;
        int ry = tid / RW, rx = tid - (tid / RW) * RW;
#pragma unroll 1
        for (int k0 = 0; k0 < NPROD; k0 += PU) {
            float v[PU][C];
            int slot[PU], px[PU], py[PU];
            bool own[PU];
#pragma unroll
            for (int u = 0; u < PU; u++) {
                const int x = x0 - R + rx, y = y0 - R + ry;
                const bool keep = INNER ? ry < RH : ry < RH && x >= 0 && x < w && y >= 0 && y < h;
                own[u] = keep && rx >= R && rx < R + TILE_X && ry >= R && ry < R + TY;
                px[u] = INNER ? x : min(max(x, 0), w - 1);
                py[u] = INNER ? (keep ? y : y0) : min(max(y, 0), h - 1);
                slot[u] = keep ? ry * RW + rx : -1;
                rx += NT % RW;
                ry += NT / RW;
                if (rx >= RW) { rx -= RW; ry++; }
            }
            produce<MODE, C, PU, INNER>(J, level, src, w, h, px, py, own, v);
#pragma unroll
            for (int u = 0; u < PU; u++)
                if (slot[u] >= 0) {
#pragma unroll
                    for (int c = 0; c < C; c++) s_in[slot[u] * C + c] = v[u][c];
                }
        }
    };
    // a tile at the image's edge: only the part of its region inside the image, [xa, xb) x [ya, yb)
    // in region coordinates, is produced (at a small level a 64-wide region is mostly outside)
    auto region_clipped = [&]() {
        const int xa = max(0, R - x0), xb = min(RW, w - x0 + R);
        const int ya = max(0, R - y0), yb = min(RH, h - y0 + R);
        const int bw = xb - xa, cnt = bw * (yb - ya);
#pragma unroll 1
        for (int k0 = 0; k0 < cnt; k0 += PU * NT) {
            float v[PU][C];
            int slot[PU], px[PU], py[PU];
            bool own[PU];
#pragma unroll
            for (int u = 0; u < PU; u++) {
                const int k = k0 + u * NT + tid;
                const bool keep = k < cnt;
                const int kk = keep ? k : 0;
                const int ry = ya + kk / bw, rx = xa + kk - (kk / bw) * bw;
                px[u] = x0 - R + rx;
                py[u] = y0 - R + ry;
                own[u] = keep && rx >= R && rx < R + TILE_X && ry >= R && ry < R + TY;
                slot[u] = keep ? ry * RW + rx : -1;
            }
            produce<MODE, C, PU, false>(J, level, src, w, h, px, py, own, v);
#pragma unroll
            for (int u = 0; u < PU; u++)
                if (slot[u] >= 0) {
#pragma unroll
                    for (int c = 0; c < C; c++) s_in[slot[u] * C + c] = v[u][c];
                }
        }
    };
    if (x0 - R >= 1 && x0 + TILE_X + R <= w - 1 && y0 - R >= 1 && y0 + TY + R <= h - 1) region(std::true_type{});
    else region_clipped();
    __syncthreads();
    // horizontal pass: item = (region row, run of TILE_RUN columns) with the interior formula, the
    // border columns element by element (HBorder); results held until every lane has read its window
    const int nh = min(R + 1, w);
    const int hint = w > 2 * (R + 1) ? w - 2 * (R + 1) : 0;
    float hres[HPER][TILE_RUN][C];
#pragma unroll
    for (int q = 0; q < HPER; q++) {
        const int it = tid + q * NT;
        if (it >= HITEMS) continue;
        const int ry = it / RUNS_X, rx0 = (it - ry * RUNS_X) * TILE_RUN;
        const int y = y0 - R + ry, xs = x0 + rx0;
        if (y < 0 || y >= h || xs >= w) continue;
        const float* line = s_in + ry * RW * C;
#pragma unroll
        for (int c = 0; c < C; c++) {
            float win[TILE_RUN + 2 * R];
#pragma unroll
            for (int d = 0; d < TILE_RUN + 2 * R; d++) win[d] = line[(rx0 + d) * C + c];
            float o[TILE_RUN];
            run_interior<R>(win, t, o);
#pragma unroll
            for (int m = 0; m < TILE_RUN; m++) hres[q][m][c] = o[m];
        }
    }
    HBorder<R, C, RH, RW, NT> hb(x0, w);
    hb.template compute<MULTI>(s_in, y0, w, h, t, tid);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < HPER; q++) {
        const int it = tid + q * NT;
        if (it >= HITEMS) continue;
        const int ry = it / RUNS_X, rx0 = (it - ry * RUNS_X) * TILE_RUN;
        const int y = y0 - R + ry, xs = x0 + rx0;
        if (y < 0 || y >= h || xs >= w) continue;
#pragma unroll
        for (int m = 0; m < TILE_RUN; m++) {
            if (xs + m < nh || xs + m >= nh + hint) continue;  // outside the line's interior (HBorder's)
#pragma unroll
            for (int c = 0; c < C; c++) s_in[(ry * RW + R + rx0 + m) * C + c] = hres[q][m][c];
        }
    }
    hb.store(s_in, y0, h, tid);
    __syncthreads();
    // vertical pass: item = (column, run of TILE_RUN rows), one per lane
    float* out = MODE == P_TENSOR ? J.ta : (MODE == P_CALCB ? J.tb : J.flow[dst]);
    const int hx = tid % TILE_X, ry0 = (tid / TILE_X) * TILE_RUN;
    const int x = x0 + hx, ys = y0 + ry0;
    if (x >= w || ys >= h) return;
    const int nv = min(R + 1, h);
    const int vint = h > 2 * (R + 1) ? h - 2 * (R + 1) : 0;
    const float* col = s_in + (R + hx) * C;
    if (ys >= nv && ys + TILE_RUN - 1 < nv + vint) {
        float o[C][TILE_RUN];
#pragma unroll
        for (int c = 0; c < C; c++) {
            float win[TILE_RUN + 2 * R];
#pragma unroll
            for (int d = 0; d < TILE_RUN + 2 * R; d++) win[d] = col[(ry0 + d) * RW * C + c];
            run_interior<R>(win, t, o[c]);
        }
        bool nan = false;
#pragma unroll
        for (int m = 0; m < TILE_RUN; m++) {
            float* p = out + ((long)(ys + m) * w + x) * C;
#pragma unroll
            for (int c = 0; c < C; c++) {
                p[c] = o[c][m];
                nan = nan || isnan(o[c][m]);
            }
        }
        // scan: the last pass of a solve whose flow is not negated afterwards also takes over the
        // negation pass's NaN scan (device/solver/mod.rs:575-582)
        if (MODE == P_NEWDISP && scan && nan) atomicOr(&J.state[2], 1);
        return;
    }
    bool nan = false;
#pragma unroll
    for (int m = 0; m < TILE_RUN; m++) {
        if (ys + m >= h) break;
#pragma unroll
        for (int c = 0; c < C; c++) {
            const float v = line_elem<R, true, MULTI>(col + c, RW * C, y0 - R, h, ys + m, t);
            out[((long)(ys + m) * w + x) * C + c] = v;
            nan = nan || isnan(v);
        }
    }
    if (MODE == P_NEWDISP && scan && nan) atomicOr(&J.state[2], 1);
}

// The separable pass of one TILE_X x TILE_Y tile of a one-channel plane (conv.rs:144-199,
// 272-311): the tile and R pixels around it into LDS (s_in, (TILE_Y + 2R) x (TILE_X + 2R)
// floats; load(i) gives the plane's element i), the row pass with the border columns right first, then the column pass with the border
// rows above first (ABOVE) or below first. Lane (column, run of TILE_RUN rows) receives its run's
// outputs in o (rows inside the plane only) at column x from row ys; the return is false for a
// lane with no output. Every element is computed as by hpass_elem and vpass_elem, so the bits are
// the same. Every lane of the block calls it (it holds barriers).
template <int R, bool ABOVE, class Load>
__device__ __forceinline__ bool sep_tile(const Load& load, int w, int h, int x0, int y0, const Taps& t, float* s_in,
                                         float o[TILE_RUN], int& x, int& ys) {
    constexpr int RW = TILE_X + 2 * R, RH = TILE_Y + 2 * R;
    constexpr int RUNS_X = TILE_X / TILE_RUN, HITEMS = RH * RUNS_X;
    constexpr int HPER = (HITEMS + TILE_THREADS - 1) / TILE_THREADS;
    const int tid = threadIdx.x;
    for (int k = tid; k < RH * RW; k += TILE_THREADS) {
        const int ry = k / RW, rx = k - ry * RW;
        const int xx = x0 - R + rx, yy = y0 - R + ry;
        if (xx < 0 || xx >= w || yy < 0 || yy >= h) continue;
        s_in[k] = load(xx, yy);
    }
    __syncthreads();
    const int nh = min(R + 1, w);
    const int hint = w > 2 * (R + 1) ? w - 2 * (R + 1) : 0;
    float hres[HPER][TILE_RUN];
#pragma unroll
    for (int q = 0; q < HPER; q++) {
        const int it = tid + q * TILE_THREADS;
        if (it >= HITEMS) continue;
        const int ry = it / RUNS_X, rx0 = (it - ry * RUNS_X) * TILE_RUN;
        const int yy = y0 - R + ry, xs = x0 + rx0;
        if (yy < 0 || yy >= h || xs >= w) continue;
        const float* line = s_in + ry * RW;
        float win[TILE_RUN + 2 * R];
#pragma unroll
        for (int d = 0; d < TILE_RUN + 2 * R; d++) win[d] = line[rx0 + d];
        run_interior<R>(win, t, hres[q]);
    }
    HBorder<R, 1, RH, RW, TILE_THREADS> hb(x0, w);
    hb.template compute<false>(s_in, y0, w, h, t, tid);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < HPER; q++) {
        const int it = tid + q * TILE_THREADS;
        if (it >= HITEMS) continue;
        const int ry = it / RUNS_X, rx0 = (it - ry * RUNS_X) * TILE_RUN;
        const int yy = y0 - R + ry, xs = x0 + rx0;
        if (yy < 0 || yy >= h || xs >= w) continue;
#pragma unroll
        for (int m = 0; m < TILE_RUN; m++)
            if (xs + m >= nh && xs + m < nh + hint) s_in[ry * RW + R + rx0 + m] = hres[q][m];
    }
    hb.store(s_in, y0, h, tid);
    __syncthreads();
    const int hx = tid % TILE_X, ry0 = (tid / TILE_X) * TILE_RUN;
    x = x0 + hx;
    ys = y0 + ry0;
    if (x >= w || ys >= h) return false;
    const int nv = min(R + 1, h);
    const int vint = h > 2 * (R + 1) ? h - 2 * (R + 1) : 0;
    const float* col = s_in + R + hx;
    if (ys >= nv && ys + TILE_RUN - 1 < nv + vint) {
        float win[TILE_RUN + 2 * R];
#pragma unroll
        for (int d = 0; d < TILE_RUN + 2 * R; d++) win[d] = col[(ry0 + d) * RW];
        run_interior<R>(win, t, o);
    } else {
#pragma unroll
        for (int m = 0; m < TILE_RUN; m++)
            if (ys + m < h) o[m] = line_elem<R, ABOVE, false>(col, RW, y0 - R, h, ys + m, t);
    }
    return true;
}

// blur5 with its band and the resample to the next level in one pass (conv.rs:598-629,
// flow.rs:228-275; the resample is the bilinear of flow/sample.rs:31-112): tiles at a pitch of
// PYR_PX x PYR_PY blur a TILE_X x TILE_Y region, so the 2 x 2 footprint of every next-level
// sample whose cell lies in the tile's pitch is inside the region, and the blurred plane stays in
// LDS. op.aux (the band; may be null) takes src - blur over the tile's pitch, op.dst (the next
// level, op.dw x op.dh; null for the last level) the samples whose cells the tile holds. The blur
// is computed as by the separate row and column passes (sep_tile) and each sample as by
// bilinear1, so the bits are the same. The ratios op.rx, op.ry are at least 2 (every level
// halves), which bounds a tile's samples to PYR_CX x PYR_CY candidates from
// (floor((x0 - 2) / rx), floor((y0 - 2) / ry)).
constexpr int PYR_PX = TILE_X - 1, PYR_PY = TILE_Y - 1, PYR_CX = 36, PYR_CY = 20;

__global__ __launch_bounds__(TILE_THREADS) void k_blur5_resample(const PlaneOp* __restrict__ ops, Taps t, int tiles_x,
                                                                 int tiles_y, int nops) {
    constexpr int R = 2;
    const int per_op = tiles_x * tiles_y;
    const int lt = xcd_tile(per_op * nops);
    if (lt >= per_op * nops) return;
    const int job = lt / per_op, rem = lt - job * per_op;
    const int tyi = rem / tiles_x, txi = rem - tyi * tiles_x;
    const PlaneOp& op = ops[job];
    const int w = op.sw, h = op.sh;
    const int x0 = txi * PYR_PX, y0 = tyi * PYR_PY;
    if (x0 >= w || y0 >= h) return;
    __shared__ float s_in[(TILE_Y + 2 * R) * (TILE_X + 2 * R)];
    __shared__ float s_blur[TILE_Y * TILE_X];
    // the plane: a float plane, or (op.luma) the luma of an engine image of format op.fmt, or
    // (op.wire) of the engine image of a wire frame, whose pixels in the tile's pitch the tile then
    // stores into op.eng (when not null) as the ingest would
    const float* __restrict__ src = (const float*)op.src;
    auto plain = [&](int xx, int yy) { return src[(long)yy * w + xx]; };
    auto luma = [&](int xx, int yy) { return luma_at(op.src, op.fmt, (long)yy * w + xx); };
    auto wire = [&](int xx, int yy) {
        const bool own = xx >= x0 && xx < x0 + PYR_PX && yy >= y0 && yy < y0 + PYR_PY;
        return wire_luma(op.wire, own ? op.eng : nullptr, op.fmt, w, h, op.reverse, xx, yy);
    };
    float o[TILE_RUN];
    int x, ys;
    const bool any = op.wire ? sep_tile<R, true>(wire, w, h, x0, y0, t, s_in, o, x, ys)
                   : op.luma ? sep_tile<R, true>(luma, w, h, x0, y0, t, s_in, o, x, ys)
                             : sep_tile<R, true>(plain, w, h, x0, y0, t, s_in, o, x, ys);
    if (any) {
#pragma unroll
        for (int m = 0; m < TILE_RUN; m++) {
            if (ys + m >= h) break;
            s_blur[(ys + m - y0) * TILE_X + (x - x0)] = o[m];
            if (op.aux && x < x0 + PYR_PX && ys + m < y0 + PYR_PY) {
                const long i = (long)(ys + m) * w + x;
                const float v = op.wire ? wire_luma(op.wire, nullptr, op.fmt, w, h, op.reverse, x, ys + m)
                                        : (op.luma ? luma(x, ys + m) : src[i]);
                op.aux[i] = v - o[m];
            }
        }
    }
    if (!op.dst) return;  // the same for the whole block
    __syncthreads();
    const float sx = op.rx, sy = op.ry;
    const int xg = max(0, (int)floorf((float)(x0 - 2) / sx));
    const int yg = max(0, (int)floorf((float)(y0 - 2) / sy));
    for (int c = threadIdx.x; c < PYR_CX * PYR_CY; c += TILE_THREADS) {
        const int oy = yg + c / PYR_CX, ox = xg + c % PYR_CX;
        if (ox >= op.dw || oy >= op.dh) continue;
        int ix, iy;
        float fx, fy;
        cell((float)ox * sx, w, ix, fx);
        cell((float)oy * sy, h, iy, fy);
        if (ix < x0 || ix >= x0 + PYR_PX || iy < y0 || iy >= y0 + PYR_PY) continue;
        const float* bp = s_blur + (iy - y0) * TILE_X + (ix - x0);
        const float p00 = bp[0], p10 = bp[1], p01 = bp[TILE_X], p11 = bp[TILE_X + 1];
        const float top = (p10 - p00) * fx + p00;
        const float bot = (p11 - p01) * fx + p01;
        op.dst[(long)oy * op.dw + ox] = (bot - top) * fy + top;
    }
}

void launch_blur5_resample(hipStream_t s, const PlaneOp* d_ops, const std::vector<PlaneOp>& ops, int maxw, int maxh,
                           const Taps& t) {
    const int tx = (maxw + PYR_PX - 1) / PYR_PX, ty = (maxh + PYR_PY - 1) / PYR_PY;
    const long n = (long)tx * ty * (long)ops.size();
    double bytes = 0;
    for (auto& o : ops)
        bytes += (double)o.sw * o.sh * ((o.luma || o.wire ? 4.0 * sample_size(o.fmt) : 4.0) + (o.aux ? 4.0 : 0.0) +
                                        (o.eng ? 4.0 * sample_size(o.fmt) : 0.0)) +
                 (o.dst ? 4.0 * o.dw * o.dh : 0.0);
    TWX_LAUNCH(s, F_PYR_BLUR, bytes,
               k_blur5_resample<<<dim3((unsigned)(8 * ((n + 7) / 8))), dim3(TILE_THREADS), 0, s>>>(d_ops, t, tx, ty,
                                                                                                   (int)ops.size()));
}

// The range check of a level (all_flow_out_of_range while no coarser level found a pixel in
// range, flow.rs:1291-1297) and the damping at the start level (flow.rs:1298-1300) in one launch:
// one 1024-lane block per job, which reduces "any pixel in range" over the level itself, zeroes
// the flow when there is none (or ends the check when there is one), then damps. A job with the
// check over and no damping due returns at once, so a launch past the start level costs one wave
// per job.
__global__ __launch_bounds__(1024) void k_range(const SolveJob* __restrict__ jobs, int level, int buf) {
    const SolveJob& J = jobs[blockIdx.x];
    if (level > J.top) return;
    const bool active = J.state[0] != 0;
    const bool damp = level == J.start && J.scale_damp;
    if (!active && !damp) return;  // the flow stays as it is (uniform over the block)
    const int w = J.w[level], h = J.h[level];
    float* flow = J.flow[buf];
    bool zero = false;
    if (active) {
        int any = 0;
        for (int i = threadIdx.x; i < w * h && !any; i += blockDim.x) {
            const int y = i / w, x = i - y * w;
            const float px = (float)x + flow[(long)i * 2];
            const float py = (float)y + flow[(long)i * 2 + 1];
            any = 0.0f <= px && px <= (float)(w - 1) && 0.0f <= py && py <= (float)(h - 1);
        }
        any = __syncthreads_or(any);
        zero = !any;
        if (threadIdx.x == 0) {
            if (any) J.state[0] = 0;
            J.state[1] = 0;
            // a zeroed level 0 holds no NaN: clear what a scan in its last pass found
            if (zero && level == 0) J.state[2] = 0;
        }
    }
    if (!zero && !damp) return;
    for (int i = threadIdx.x; i < w * h; i += blockDim.x) {
        float u = flow[(long)i * 2], v = flow[(long)i * 2 + 1];
        if (zero) { u = 0.0f; v = 0.0f; }
        if (damp) { u = J.damp * u; v = J.damp * v; }
        flow[(long)i * 2] = u;
        flow[(long)i * 2 + 1] = v;
    }
}

// upsample_flow (flow.rs:376-391) from level l (buffer 1) into level l-1 (buffer 0)
// UP_ROWS destination rows per thread: the cell column and its weight are computed once, and a
// launch has a quarter of the waves (one row per 256-lane block ran at 2.5 TB/s, bound by wave
// launches rather than by memory)
constexpr int UP_ROWS = 4;
__global__ void k_upsample(const SolveJob* __restrict__ jobs, int level) {
    const SolveJob& J = jobs[blockIdx.z];
    if (level > J.top || level == 0) return;
    const int sw = J.w[level], sh = J.h[level];
    const int dw = J.w[level - 1], dh = J.h[level - 1];
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y0 = blockIdx.y * UP_ROWS;
    if (x >= dw || y0 >= dh) return;
    const gfloat* src = (const gfloat*)J.flow[1];
    float* dst = J.flow[0];
    const float sx = J.up[level][0], sy = J.up[level][1], rx = J.up[level][2], ry = J.up[level][3];
    int ix;
    float ax;
    cell((float)x * rx, sw, ix, ax);
    // the cell's two flow pixels of its top row, then of its bottom row (a level's flow is under
    // 2^32 bytes: working sizes are at most 16384 x 16384)
    float t0[UP_ROWS][4], t1[UP_ROWS][4], ay[UP_ROWS];
#pragma unroll
    for (int r = 0; r < UP_ROWS; r++) {
        const int y = min(y0 + r, dh - 1);
        int iy;
        cell((float)y * ry, sh, iy, ay[r]);
        const uint32_t q = (uint32_t)iy * (uint32_t)sw + (uint32_t)ix;
        ldrun<4>(src, 2 * q, t0[r]);
        ldrun<4>(src, 2 * (q + (uint32_t)sw), t1[r]);
    }
#pragma unroll
    for (int r = 0; r < UP_ROWS; r++) {
        const int y = y0 + r;
        if (y >= dh) break;
        float out[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            float p00 = t0[r][c], p10 = t0[r][2 + c], p01 = t1[r][c], p11 = t1[r][2 + c];
            float top = (p10 - p00) * ax + p00;
            float bot = (p11 - p01) * ax + p01;
            out[c] = (bot - top) * ay[r] + top;
        }
        reinterpret_cast<float2*>(dst)[(long)y * dw + x] = make_float2(out[0] * sx, out[1] * sy);
    }
}

void solve_job_ratios(SolveJob& J) {
    for (int l = 1; l < J.levels && l < MAXL; l++) {
        const int sw = J.w[l], sh = J.h[l], dw = J.w[l - 1], dh = J.h[l - 1];
        const float sx = ((float)dw - 1.0f) / (float)(sw - 1);
        const float sy = ((float)dh - 1.0f) / (float)(sh - 1);
        J.up[l][0] = sx;
        J.up[l][1] = sy;
        J.up[l][2] = (float)(1.0 / (double)sx);
        J.up[l][3] = (float)(1.0 / (double)sy);
    }
}

// final negation (flow.rs:1305) of level 0's flow (buffer 1, the output) and the NaN scan the
// reference device path applies to the downloaded flow (device/solver/mod.rs:575-582); a
// grid-stride loop over the job's floats, four at a time where the buffer is 16-byte aligned
__global__ void k_negate(const SolveJob* __restrict__ jobs) {
    const SolveJob& J = jobs[blockIdx.z];
    const long n = (long)J.w[0] * J.h[0] * 2;
    float* flow = J.flow[1];
    const long stride = (long)gridDim.x * blockDim.x;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    long done = 0;
    if (((uintptr_t)flow & 15) == 0) {
        float4* f4 = reinterpret_cast<float4*>(flow);
        done = n / 4 * 4;
        for (long i = t; i < n / 4; i += stride) {
            float4 v = f4[i];
            v.x = -1.0f * v.x;
            v.y = -1.0f * v.y;
            v.z = -1.0f * v.z;
            v.w = -1.0f * v.w;
            bad |= isnan(v.x) || isnan(v.y) || isnan(v.z) || isnan(v.w);
            f4[i] = v;
        }
    }
    for (long i = done + t; i < n; i += stride) {
        const float v = -1.0f * flow[i];
        bad |= isnan(v);
        flow[i] = v;
    }
    if (bad) atomicOr(&J.state[2], 1);
}

// -------------------------------------------------------------------------------------------
// Block matching (flow.rs:446-622)

// A sweep's coordinates: `a` along its chains, `c` across them (rows: a = x, c = y; columns:
// a = y, c = x). In walk layout element (a, c) of a plane sits at a * m + c (m: the extent across),
// so the lanes of a wave, one chain each, read consecutive addresses. The column sweeps run on the
// plain planes, which are in walk layout for them. The row sweeps read transposed copies of X and
// Y; the plain flow is in cross layout for them (c * n + a, n: the extent along): sweep 1 reads it
// there and writes its result in walk layout into the job's scratch, and sweep 2 reads that and
// writes its result back into the plain flow, so the flow needs no transposing pass.

// The B plane of a sweep (walk layout) behind a buffer resource: the four runs of a cost's 4 x 4
// texels are one per-lane offset plus four wave-uniform ones.
struct BmPlane {
    __amdgpu_buffer_rsrc_t rsrc;
    int w, h;          // the level's extent
    uint32_t rs;       // elements per run (m)
    uint32_t off_max;  // the largest first-texel offset whose 4 x 4 block stays in the plane
    float wm2, hm2;
    __device__ BmPlane(const float* base, int w_, int h_, int n_al, int m)
        : rsrc(__builtin_amdgcn_make_buffer_rsrc((void*)base, 0, w_ * h_ * 4, 0x00020000)), w(w_), h(h_),
          rs((uint32_t)m), off_max((uint32_t)((n_al - 4) * m + m - 4)), wm2((float)(w_ - 2)), hm2((float)(h_ - 2)) {}
    __device__ float word(uint32_t i) const {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)(i * 4u), 0, 0));
    }
};

// Block::cost (flow.rs:472-524) without branches: `ok` false where the reference returns None (the
// value is then meaningless). The operations and their order are those of the reference.
template <bool T, typename Between>
__device__ __forceinline__ float bm_cost(const BmPlane& B, const float p[9], float x2, float y2, bool& ok,
                                         Between&& between) {
    ok = !(x2 > B.wm2 || y2 > B.hm2);
    // cvtt(x2) for x2 < w - 2: the hardware truncation equals cvtt wherever the cost exists and
    // gives ix < 1 wherever cvtt does (NaN gives 0, out of range saturates); at x2 == w - 2 the
    // reference takes ix = w - 3 and fx = 1, which the clamp gives (x2 - (w - 3) is exactly 1)
    const int32_t ix = min(trunc_hw(x2), B.w - 3), iy = min(trunc_hw(y2), B.h - 3);
    ok = ok && !(ix < 1 || iy < 1);
    const float fx = x2 - (float)ix, fy = y2 - (float)iy;
    const float w11 = fy * fx;
    const float w10 = fx - w11;
    const float w01 = fy - w11;
    const float w00 = ((1.0f - fx) - fy) + w11;
    // texels (ix - 1 .. ix + 2, iy - 1 .. iy + 2) as four runs of four, the first at element
    // (r - 1) * rs + (e - 1) (r, e: run and element of texel (ix, iy)). Sides are below 2^24, so
    // the 24-bit product is exact wherever the cost exists; any other offset is clamped into the
    // plane, and its texels are not used.
    const uint32_t r = (uint32_t)(T ? ix : iy), e = (uint32_t)(T ? iy : ix);
    const uint32_t off = min(__umul24(r, B.rs) + e - (B.rs + 1u), B.off_max) * 4u;
    float t[4][4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(B.rsrc, (int)off, (int)((uint32_t)k * B.rs * 4u), 0);
#pragma unroll
        for (int i = 0; i < 4; i++) t[k][i] = __uint_as_float(v[i]);
    }
    between();  // the caller's work between the texel loads and their use
    // texel (ix - 1 + dx, iy - 1 + dy)
    auto tx = [&](int dx, int dy) { return T ? t[dx][dy] : t[dy][dx]; };
    float sq[9];
#pragma unroll
    for (int bb = 0; bb < 3; bb++) {
#pragma unroll
        for (int aa = 0; aa < 3; aa++) {
            float q = ((w00 * tx(aa, bb) + w10 * tx(aa + 1, bb)) + w01 * tx(aa, bb + 1)) + w11 * tx(aa + 1, bb + 1);
            float d = p[bb * 3 + aa] - q;
            sq[bb * 3 + aa] = d * d;
        }
    }
    float s = sq[0] + sq[1];
#pragma unroll
    for (int k = 2; k < 8; k++) s = s + sq[k];
    return sq[8] + s;
}

// A flow plane as a sweep reads or writes it: pixel (a, c) at byte base + a * step + lane, where
// `lane` is the chain's own part (c * 8 in walk layout, c * n * 8 in cross layout) and a * step
// is wave-uniform.
struct BmFlow {
    gchar* base;
    uint32_t step, lane;
    __device__ BmFlow(const float* p, bool walk, int n_al, int m, int c)
        : base((gchar*)p), step(walk ? 8u * (uint32_t)m : 8u), lane(walk ? 8u * (uint32_t)c : 8u * (uint32_t)c * (uint32_t)n_al) {}
    __device__ float2 at(int a) const {
        const gfloat* q = (const gfloat*)(base + ((uint32_t)a * step + lane));
        return make_float2(q[0], q[1]);
    }
    __device__ void put(int a, float u, float v) const {
        gfloat_rw* q = (gfloat_rw*)(base + ((uint32_t)a * step + lane));
        q[0] = u;
        q[1] = v;
    }
};

// A cost plane as a sweep reads or writes it: the same addressing as BmFlow with one float per pixel.
struct BmCost {
    gchar* base;
    uint32_t step, lane;
    __device__ BmCost(float* p, bool walk, int n_al, int m, int c)
        : base((gchar*)p), step(walk ? 4u * (uint32_t)m : 4u), lane(walk ? 4u * (uint32_t)c : 4u * (uint32_t)c * (uint32_t)n_al) {}
    __device__ float at(int a) const { return *(const gfloat*)(base + ((uint32_t)a * step + lane)); }
    __device__ void put(int a, float v) const { *(gfloat_rw*)(base + ((uint32_t)a * step + lane)) = v; }
};

// The transposed copy of a frame's level for the row sweeps (element (x, y) at x * h + y): 32 x 32
// tiles through LDS. One plane per distinct frame of the launch (blockIdx.z), which every job with
// the frame as X or Y reads.
__global__ void k_plane_transpose(const PlaneT* __restrict__ ops, int level) {
    const PlaneT& P = ops[blockIdx.z];
    if (level > P.top) return;
    const int w = P.w[level], h = P.h[level];
    if (w < 4 || h < 4) return;  // no sweep runs on such a level
    __shared__ float t[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    if (bx >= w || by >= h) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
    const float* src = P.lv[level];
    for (int r = ty; r < 32; r += 8) {
        const int x = bx + tx, y = by + r;
        if (x < w && y < h) t[r][tx] = src[(long)y * w + x];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int x = bx + r, y = by + tx;
        if (x < w && y < h) P.dst[(long)x * h + y] = t[tx][r];
    }
}

// The flow into its transposed scratch copy for the row sweeps (OUT false) and back (OUT true),
// for levels whose row sweeps run on the transposed flow; a flow pixel (u, v) moves as one 8-byte
// access (the scratch slice starts 8-byte aligned).
// Back out, the carried costs move from the walk layout into the plain one as well.
template <bool OUT>
__global__ void k_bm_transpose(const SolveJob* __restrict__ jobs, int level, int buf) {
    const SolveJob& J = jobs[blockIdx.z];
    if (level > J.top) return;
    const int w = J.w[level], h = J.h[level];
    if (w < 4 || h < 4) return;
    __shared__ float t[OUT ? 3 : 2][32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    if (bx >= w || by >= h) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 256 threads: 32 x 8
    float2* flow = reinterpret_cast<float2*>(J.flow[buf]);
    float2* ft = reinterpret_cast<float2*>(J.tmp);
    if (!OUT) {
        for (int r = ty; r < 32; r += 8) {
            const int x = bx + tx, y = by + r;
            if (x < w && y < h) {
                const float2 f = flow[(long)y * w + x];
                t[0][r][tx] = f.x;
                t[1][r][tx] = f.y;
            }
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            const int x = bx + r, y = by + tx;
            if (x < w && y < h) ft[(long)x * h + y] = make_float2(t[0][tx][r], t[1][tx][r]);
        }
    } else {
        for (int r = ty; r < 32; r += 8) {
            const int x = bx + r, y = by + tx;
            if (x < w && y < h) {
                const float2 f = ft[(long)x * h + y];
                t[0][tx][r] = f.x;
                t[1][tx][r] = f.y;
                t[OUT ? 2 : 0][tx][r] = J.cost[0][(long)x * h + y];
            }
        }
        __syncthreads();
        for (int r = ty; r < 32; r += 8) {
            const int x = bx + tx, y = by + r;
            if (x < w && y < h) {
                flow[(long)y * w + x] = make_float2(t[0][r][tx], t[1][r][tx]);
                J.cost[1][(long)y * w + x] = t[OUT ? 2 : 0][r][tx];
            }
        }
    }
}

// steps per unrolled chunk of a block-matching chain (a multiple of the patch's three runs), and
// how many steps ahead a chain prefetches its lines
constexpr int BM_CHUNK = 6;
// (measured: the row sweeps gain from prefetching four steps ahead, the column sweeps lose at
// every distance tried and do not prefetch)
constexpr int BM_PREFETCH_ROWS = 4, BM_PREFETCH_COLS = 0;
// The row sweeps read and write the plain flow directly (cross layout: one line per chain) on
// levels of up to this many pixels over the launch's jobs, and run on a transposed copy above it
// (measured on 148 solves of 960 x 540: cross layout at the 960 x 540 level costs 1.5 ms more
// than the copies, at 480 x 270 and below it saves their time)
constexpr double BM_CROSS_MAX_PIXELS = 32e6;

// The state of one lane of a block-matching chain and its steps (Block::step, flow.rs:528-544),
// run CH at a time with the loop unrolled: the patch runs sit in ring slots that rotate with the
// step (slot of patch run r at step s: (s + r) mod 3 for DIR = +1, (r - s) mod 3 for DIR = -1;
// CH is a multiple of 3, so every chunk starts from the same assignment), and the decision
// updates the incoming vector by selects, so values move between registers only at chunk ends.
//
// The decision needs no range test of its own: the reference's test that the pixel's own target
// lies in [0, w - 1] x [0, h - 1] holds wherever the own cost exists (that cost needs the target
// in [1, w - 2] x [1, h - 2]). A cost that does not exist is replaced by -inf for the own vector
// and +inf for the incoming one, so "both exist and incoming < own" is one comparison.
//
// Every sweep visits the same pixels (the interior), and a pixel's vector after a sweep is the one
// whose cost that sweep computed last for it: the incoming vector's when the pixel took it, its own
// otherwise. The cost is a function of the pixel and the vector alone, so that value is the own
// cost the next sweep would compute. The second row sweep (a lane pair: role 0 the own vector,
// role 1 the incoming one) stores it for every pixel (COST_OUT), and the column sweeps (CARRIED:
// one lane per chain) read it instead of computing it, cost only the incoming vector, and the
// first of them stores the new value where the pixel changes. (The row sweeps keep the lane pair:
// their chains are fewer and longer, and with half the waves a carried row sweep ran slower.)
template <bool T, int DIR, int PD, bool CROSS, bool CARRIED, bool COST_OUT>
struct BmChain {
    static constexpr bool ALWAYS = T && CROSS;  // the sweep writes every pixel (it moves the flow between layouts)
    // the costs are stored for every pixel: the first sweep, and a sweep that moves them between layouts
    static constexpr bool COST_ALL = COST_OUT && (!CARRIED || ALWAYS);
    const BmPlane& B;
    gchar* A;        // X in walk layout
    uint32_t a_lane;  // byte offset of the chain's patch run, across (ce - 1) * 4
    uint32_t a_step;  // bytes per run (m * 4)
    const BmFlow& fr;
    const BmFlow& fw;
    const BmCost& ccr;  // the carried costs (CARRIED)
    const BmCost& ccw;  // where the costs go (COST_OUT)
    float pf[3];  // the prefetched words of the last step
    int m, n_al, n, ce;
    int role;
    float sent;   // the cost where it does not exist: -inf for role 0, +inf for role 1
    float cf;     // (float)ce
    int a;
    float cu, cv, nu, nv;
    float oc;     // CARRIED: the own cost of the current pixel (-inf where it does not exist)
    bool wr;      // the lane stores (role 0 of a live chain)
    // COST_ALL: the costs of the last chunk, stored during the next chunk's first step after its
    // loads (so that the in-order memory counter makes none of them wait for the stores) or after
    // the chain's last chunk. (Deferring the flow's stores the same way made the sweeps slower.)
    float pc[BM_CHUNK];
    int pa = 0, pn = 0;

    __device__ gchar* arun(int r) const { return A + ((uint32_t)r * a_step + a_lane); }

    template <int CH, bool CHECK>
    __device__ __forceinline__ void chunk(int k0, float (&ring)[3][3]) {
        static_assert(CH % 3 == 0, "the ring's rotation repeats every three steps");
        static_assert(CH == BM_CHUNK, "the pending costs hold one chunk");
        const int a_first = a;
        float cb[CH];
#pragma unroll
        for (int st = 0; st < CH; st++) {
            const int k = k0 + st;
            if (CHECK && k >= n) break;
            const int an = k + 1 < n ? a + DIR : a;
            const float2 cn = fr.at(an);
            const float ocn = CARRIED ? ccr.at(an) : 0.0f;
            // the run entering the patch at the next step (a + 2 DIR), clamped at the chain's end
            const int ar = min(max(an + DIR, 0), n_al - 1);
            float nr[3];
            gchar* q = arun(ar);
#pragma unroll
            for (int i = 0; i < 3; i++) nr[i] = *(const gfloat*)(q + 4 * i);
            const int s0 = DIR > 0 ? st % 3 : (3 - st % 3) % 3;        // slot of patch run 0
            const int s1 = DIR > 0 ? (st + 1) % 3 : (4 - st % 3) % 3;  // run 1
            const int s2 = DIR > 0 ? (st + 2) % 3 : (5 - st % 3) % 3;  // run 2
            const int sl[3] = {s0, s1, s2};
            float p[9];
#pragma unroll
            for (int bb = 0; bb < 3; bb++)
#pragma unroll
                for (int aa = 0; aa < 3; aa++) p[bb * 3 + aa] = T ? ring[sl[aa]][bb] : ring[sl[bb]][aa];
            const float af = (float)a;
            const float xf = T ? af : cf, yf = T ? cf : af;
            // role 0 costs the pixel's own vector, role 1 (and every lane of a CARRIED sweep) the
            // incoming one
            const float vu = CARRIED || role ? nu : cu, vv = CARRIED || role ? nv : cv;
            // Every step touches new lines of A, B and the flow; a load that misses L2 costs the
            // step its full latency, since the step waits on its dependent B texels and the memory
            // counter retires loads in order. So each step asks, right after its own B loads, for
            // one word of each line the chain will touch PD steps on (the leading B run around the
            // lane's vector of now, the A run, the flow). The words are consumed one step later,
            // after that step's own wait has covered them, so they cost the chain no wait of their
            // own, and by their step the lines sit in L2.
            auto prefetch = [&]() {
                if (COST_ALL && st == 0) flush();
                if (PD == 0) return;
                asm volatile("" ::"v"(pf[0]), "v"(pf[1]), "v"(pf[2]));
                const int ap = a + PD * DIR;
                const float al = (float)ap - (T ? vu : vv), ac = cf - (T ? vv : vu);
                const int rb = min(max(trunc_hw(fminf(fmaxf(al, -8.0f), (float)n_al + 8.0f)) + (DIR > 0 ? 2 : -1), 0), n_al - 1);
                const int eb = min(max(trunc_hw(fminf(fmaxf(ac, -8.0f), (float)m + 8.0f)) - 1, 0), m - 1);
                const int ra = min(max(ap + 2 * DIR, 0), n_al - 1);
                const int rf = min(max(ap + DIR, 0), n_al - 1);
                pf[0] = B.word(__umul24((uint32_t)rb, (uint32_t)m) + (uint32_t)eb);
                pf[1] = *(const gfloat*)arun(ra);
                pf[2] = fr.at(rf).x;
            };
            bool ok;
            const float c = bm_cost<T>(B, p, xf - vu, yf - vv, ok, prefetch);
            const float cs = ok ? c : sent;
            float inc, own;
            if (CARRIED) {
                inc = cs;
                own = oc;
            } else {
                // the partner lane's cost by a quad permutation [1, 0, 3, 2] (every lane active)
                const float co = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(cs), 0xB1, 0xF, 0xF, false));
                // incoming < own: role 0 holds own, role 1 incoming
                inc = role ? cs : co;
                own = role ? co : cs;
            }
            const bool take = inc < own;
            nu = take ? nu : cu;
            nv = take ? nv : cv;
            if (wr && (ALWAYS || take)) fw.put(a, nu, nv);
            if (COST_ALL) cb[st] = take ? inc : own;
            else if (COST_OUT && wr && take) ccw.put(a, inc);
            // the run leaving the patch (run 0 for DIR = +1, run 2 for DIR = -1) makes room for the
            // entering one
#pragma unroll
            for (int i = 0; i < 3; i++) ring[DIR > 0 ? s0 : s2][i] = nr[i];
            a += DIR;
            cu = cn.x;
            cv = cn.y;
            if (CARRIED) oc = ocn;
        }
        if (COST_ALL) {
#pragma unroll
            for (int st = 0; st < CH; st++) pc[st] = cb[st];
            pa = a_first;
            pn = CHECK ? n - k0 : CH;
        }
    }
    // the pending costs out (after the chain's last chunk: called once more)
    __device__ __forceinline__ void flush() {
        if (!COST_ALL || !wr) return;
#pragma unroll
        for (int i = 0; i < BM_CHUNK; i++)
            if (i < pn) ccw.put(pa + i * DIR, pc[i]);
        pn = 0;
    }
};

// The four sweeps of block matching (flow.rs:551-614) as one kernel: T = true runs sweeps 1
// (DIR = +1, left to right) and 2 (DIR = -1) with one serial chain per interior row, T = false
// sweeps 3 (top to bottom) and 4 with one chain per interior column (adjacent chains are adjacent
// lanes). Whether a pixel keeps its own vector decides the next pixel's input, and in practice
// more than half the pixels take the incoming vector, in long runs: a chain does not split
// usefully. A lane pair runs each chain: role 0 computes the cost of the pixel's own vector, role
// 1 the incoming vector's, in one instruction stream, and both take the decision (Block::step,
// flow.rs:528-544). The A patch slides along the chain, one new run per step; the loop runs in
// unrolled chunks of CH steps so that the values loaded one step ahead need fewer register copies.
// CARRIED (sweeps 3 and 4): one lane per chain, the own vectors' costs read from the previous
// sweep (BmChain); COST_OUT: the sweep stores the costs of its pixels' vectors for the next one.
template <bool T, int DIR, int CH, bool CROSS, bool CARRIED, bool COST_OUT>
__global__ __launch_bounds__(64) void k_bm_sweep(const SolveJob* __restrict__ jobs, int level, int buf) {
    const SolveJob& J = jobs[blockIdx.z];
    if (level > J.top) return;
    const int w = J.w[level], h = J.h[level];
    // with a side below 4 no cost exists (it needs 1 <= ix <= w - 3), so no pixel changes
    if (w < 4 || h < 4) return;
    const int n_al = T ? w : h, m = T ? h : w;
    constexpr int CPB = CARRIED ? 64 : 32;  // chains per 64-lane block
    const int c0 = 1 + blockIdx.x * CPB;
    if (c0 > m - 2) return;  // the whole workgroup is past the chains (uniform)
    const int n = n_al - 2;
    const int a0 = DIR > 0 ? 1 : n_al - 2;
    const int lane = threadIdx.x;
    const int role = CARRIED ? 1 : lane & 1;
    const int cc = c0 + (CARRIED ? lane : lane >> 1);
    const bool live = cc <= m - 2;
    const int ce = live ? cc : m - 2;
    const BmPlane B(T ? J.yt : J.Y[level], w, h, n_al, m);
    gchar* A = (gchar*)(T ? J.xt : J.X[level]);
    // flow planes: the plain flow (walk layout for the columns, cross for the rows) and the rows'
    // scratch (walk layout). CROSS rows: sweep 1 reads plain and writes scratch, sweep 2 the other
    // way round; otherwise the rows run on a transposed copy of the flow in the scratch.
    const BmFlow plain(J.flow[buf], !T, n_al, m, ce), scratch(J.tmp, true, n_al, m, ce);
    const BmFlow& fr = T && (DIR < 0 || !CROSS) ? scratch : plain;
    const BmFlow& fw = T && (DIR > 0 || !CROSS) ? scratch : plain;
    const BmFlow& fi = T && !CROSS ? scratch : plain;
    // cost planes: [0] the rows' walk layout, [1] plain (walk layout for the columns). On levels
    // whose rows run on the transposed flow, sweep 2 writes [0], the transposes back move it into
    // [1], and sweeps 3 and 4 read [1] (3 updates it in place).
    const BmCost cwalk(J.cost[0], true, n_al, m, ce), cplain(J.cost[1], !T, n_al, m, ce);
    const BmCost& ccr = cplain;
    const BmCost& ccw = T && !CROSS ? cwalk : cplain;
    // the incoming vector of the first pixel: the chain's end pixel, which no sweep changes
    const float2 f0 = fi.at(a0 - DIR), f1 = fr.at(a0);
    float nu = f0.x, nv = f0.y, cu = f1.x, cv = f1.y;
    // the A patch (Block::patch, flow.rs:459-470) as three runs of three: run r is A at along
    // a - 1 + r, across ce - 1 .. ce + 1; the runs live in a ring whose slots rotate with the step
    const uint32_t a_lane = 4u * (uint32_t)(ce - 1), a_step = 4u * (uint32_t)m;
    float ring[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        gchar* q = A + ((uint32_t)(a0 - 1 + r) * a_step + a_lane);
#pragma unroll
        for (int i = 0; i < 3; i++) ring[r][i] = *(const gfloat*)(q + 4 * i);
    }
    asm volatile("" ::"v"(nu), "v"(nv), "v"(cu), "v"(cv), "v"(ring[0][0]), "v"(ring[0][1]), "v"(ring[0][2]),
                 "v"(ring[1][0]), "v"(ring[1][1]), "v"(ring[1][2]), "v"(ring[2][0]), "v"(ring[2][1]), "v"(ring[2][2]));
    const float oc = CARRIED ? ccr.at(a0) : 0.0f;
    BmChain<T, DIR, T ? BM_PREFETCH_ROWS : BM_PREFETCH_COLS, CROSS, CARRIED, COST_OUT> ch{
        B, A, a_lane, a_step, fr, fw, ccr, ccw, {0.0f, 0.0f, 0.0f}, m, n_al, n, ce, role,
        role ? INFINITY : -INFINITY, (float)ce, a0, cu, cv, nu, nv, oc, live && (CARRIED || role == 0)};
    const int nfull = n / CH;
    for (int j = 0; j < nfull; j++) ch.template chunk<CH, false>(j * CH, ring);
    if (nfull * CH < n) ch.template chunk<CH, true>(nfull * CH, ring);
    ch.flush();
    asm volatile("" ::"v"(ch.pf[0]), "v"(ch.pf[1]), "v"(ch.pf[2]));
}

// -------------------------------------------------------------------------------------------
// The coarse levels of a solve in one persistent kernel (flow.rs:1258-1304 from the start level
// down to the last level whose planes fit in LDS): a fixed grid of workgroups takes jobs from a
// device work queue, and each runs its job's coarse levels entirely in LDS (X, Y, the flow, the
// gradient, the blurred tensor and two three-channel work planes: 15 floats per pixel), with
// block barriers between the steps. It hands the first level left to the per-level kernels its
// upsampled flow in flow[0] (or, when every level fits, the finished level 0 in flow[1]). Every
// element is computed with the operations of the per-level kernels, so the bits do not depend on
// which path ran a level. One launch replaces about fifteen per level, whose time at these sizes
// is latency, not work.

constexpr int COARSE_THREADS = 512;

static std::atomic<bool> g_coarse_on{true};
void coarse_solver_enable(bool on) { g_coarse_on.store(on); }
bool coarse_solver_enabled() { return g_coarse_on.load(); }

// i / w for the pixel indices of a coarse level (i < 2^16, w < 2^16): the high word of i times
// ceil(2^32 / w), exact in that range (its error, below i / 2^32 <= 2^-16 <= 1 / w, never carries
// past the next multiple of 1 / w); one multiply instead of a division.
struct RowDiv {
    uint32_t m;
    int w;
    __device__ explicit RowDiv(int w_) : m(w_ > 1 ? (uint32_t)(0xFFFFFFFFu / (uint32_t)w_) + 1u : 0u), w(w_) {}
    __device__ int operator()(int i) const { return w > 1 ? (int)__umulhi((uint32_t)i, m) : i; }
};

// A C-channel plane's Gaussian (conv.rs:144-199, 272-311) in LDS: the row pass (border columns
// right first) src -> tmp, then the column pass (border rows above first) tmp -> dst.
template <int R, int C>
__device__ __forceinline__ void lds_gauss(const float* src, float* tmp, float* dst, int w, int h, const Taps& t,
                                          const RowDiv& row) {
    constexpr bool MULTI = C > 1;
    const int n = w * h * C;
    for (int e = threadIdx.x; e < n; e += COARSE_THREADS) {
        const int i = e / C, c = e - i * C;
        const int y = row(i), x = i - y * w;
        tmp[e] = line_elem<R, false, MULTI>(src + (y * w) * C + c, C, 0, w, x, t);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += COARSE_THREADS) {
        const int i = e / C, c = e - i * C;
        const int y = row(i), x = i - y * w;
        dst[e] = line_elem<R, true, MULTI>(tmp + x * C + c, w * C, 0, h, y, t);
    }
    __syncthreads();
}

// Block::cost (flow.rs:472-524) on a plane in LDS; false where the reference returns None.
__device__ __forceinline__ bool lds_cost(const float* B, int w, int h, const float p[9], float x2, float y2, float& cost) {
    const float wm2 = (float)(w - 2), hm2 = (float)(h - 2);
    if (x2 > wm2 || y2 > hm2) return false;
    int32_t ix, iy;
    float fx, fy;
    if (x2 == wm2) { ix = cvtt(wm2 - 1.0f); fx = 1.0f; } else { ix = cvtt(x2); fx = x2 - (float)ix; }
    if (y2 == hm2) { iy = cvtt(hm2 - 1.0f); fy = 1.0f; } else { iy = cvtt(y2); fy = y2 - (float)iy; }
    if (ix < 1 || iy < 1) return false;
    const float w11 = fy * fx;
    const float w10 = fx - w11;
    const float w01 = fy - w11;
    const float w00 = ((1.0f - fx) - fy) + w11;
    const float* t = B + ((iy - 1) * w + ix - 1);
    float sq[9];
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float q = ((w00 * t[b * w + a] + w10 * t[b * w + a + 1]) + w01 * t[(b + 1) * w + a]) +
                            w11 * t[(b + 1) * w + a + 1];
            const float d = p[b * 3 + a] - q;
            sq[b * 3 + a] = d * d;
        }
    float s2 = sq[0] + sq[1];
#pragma unroll
    for (int k = 2; k < 8; k++) s2 = s2 + sq[k];
    cost = sq[8] + s2;
    return true;
}

// One sweep of block matching (flow.rs:551-614) on the flow F in LDS: the own vectors' costs of
// every interior pixel first (a pixel holds its own vector until its chain reaches it; -inf where
// the cost does not exist) into OC, then one lane per chain costing the incoming vector (+inf
// where it does not exist), which replaces the pixel's vector when it is lower (Block::step,
// flow.rs:528-544). ROWS: chains are rows (sweeps 1, 2); DIR: the direction along them.
template <bool ROWS, int DIR>
__device__ void lds_sweep(const float* X, const float* Y, float* F, float* OC, int w, int h, const RowDiv& row) {
    for (int e = threadIdx.x; e < w * h; e += COARSE_THREADS) {
        const int y = row(e), x = e - y * w;
        if (x < 1 || x > w - 2 || y < 1 || y > h - 2) continue;
        float p[9], c;
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int a = 0; a < 3; a++) p[b * 3 + a] = X[(y - 1 + b) * w + x - 1 + a];
        OC[e] = lds_cost(Y, w, h, p, (float)x - F[2 * e], (float)y - F[2 * e + 1], c) ? c : -INFINITY;
    }
    __syncthreads();
    const int chains = ROWS ? h - 2 : w - 2, len = ROWS ? w - 2 : h - 2;
    for (int k = threadIdx.x; k < chains; k += COARSE_THREADS) {
        // the chain's end pixel (which no sweep changes) gives the first incoming vector
        int x = ROWS ? (DIR > 0 ? 0 : w - 1) : 1 + k;
        int y = ROWS ? 1 + k : (DIR > 0 ? 0 : h - 1);
        int e = y * w + x;
        const int step = ROWS ? DIR : DIR * w;
        float nu = F[2 * e], nv = F[2 * e + 1];
        // the patch around the first pixel, then one new run of three per step
        float p[9];
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int px = ROWS ? x + DIR * a : x - 1 + a, py = ROWS ? y - 1 + b : y + DIR * b;
                const int pa = ROWS ? (DIR > 0 ? a : 2 - a) : a, pb = ROWS ? b : (DIR > 0 ? b : 2 - b);
                p[pb * 3 + pa] = X[py * w + px];
            }
        for (int j = 0; j < len; j++) {
            x += ROWS ? DIR : 0;
            y += ROWS ? 0 : DIR;
            e += step;
            if (j > 0) {
                // slide the patch one pixel along the chain
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    if (ROWS) {
                        const float v = X[(y - 1 + r) * w + x + DIR];
                        if (DIR > 0) { p[r * 3] = p[r * 3 + 1]; p[r * 3 + 1] = p[r * 3 + 2]; p[r * 3 + 2] = v; }
                        else { p[r * 3 + 2] = p[r * 3 + 1]; p[r * 3 + 1] = p[r * 3]; p[r * 3] = v; }
                    } else {
                        const float v = X[(y + DIR) * w + x - 1 + r];
                        if (DIR > 0) { p[r] = p[3 + r]; p[3 + r] = p[6 + r]; p[6 + r] = v; }
                        else { p[6 + r] = p[3 + r]; p[3 + r] = p[r]; p[r] = v; }
                    }
                }
            }
            const float cu = F[2 * e], cv = F[2 * e + 1], own = OC[e];
            float c;
            const float inc = lds_cost(Y, w, h, p, (float)x - nu, (float)y - nv, c) ? c : INFINITY;
            if (inc < own) {
                F[2 * e] = nu;
                F[2 * e + 1] = nv;
            } else {
                nu = cu;
                nv = cv;
            }
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(COARSE_THREADS) void k_coarse_solve(const SolveJob* __restrict__ jobs, int njobs,
                                                                  int* __restrict__ queue, Taps g7, Taps g5,
                                                                  int skip_refine) {
    __shared__ float smem[15 * COARSE_MAX_PIXELS];
    __shared__ int s_job;
    for (;;) {
        if (threadIdx.x == 0) s_job = atomicAdd(queue, 1);
        __syncthreads();
        const int job = s_job;
        __syncthreads();  // every lane has read s_job before lane 0 may take the next one
        if (job >= njobs) return;
        const SolveJob& J = jobs[job];
        if (J.top >= J.start) continue;  // no coarse part
        float* X = smem;
        float* Y = X + COARSE_MAX_PIXELS;
        float* F = Y + COARSE_MAX_PIXELS;
        float* G = F + 2 * COARSE_MAX_PIXELS;
        float* TA = G + 2 * COARSE_MAX_PIXELS;
        float* S1 = TA + 3 * COARSE_MAX_PIXELS;
        float* S2 = S1 + 3 * COARSE_MAX_PIXELS;
        bool active = true;  // the range check (k_solve_init started it)
        {
            const int n = J.w[J.start] * J.h[J.start] * 2;
            for (int i = threadIdx.x; i < n; i += COARSE_THREADS) F[i] = 0.0f;
        }
        for (int l = J.start; l > J.top; l--) {
            const int w = J.w[l], h = J.h[l], n = w * h;
            const RowDiv row(w);
            for (int i = threadIdx.x; i < n; i += COARSE_THREADS) {
                X[i] = J.X[l][i];
                Y[i] = J.Y[l][i];
            }
            __syncthreads();
            // calc_grad + calc_a (flow.rs:283-328), the tensor blurred r = 7
            for (int i = threadIdx.x; i < n; i += COARSE_THREADS) {
                const int y = row(i), x = i - y * w;
                const float c = X[i];
                float nx = 0.0f, dx = 0.0f, ny = 0.0f, dy = 0.0f;
                if (x >= 1) { nx = (c - X[i - 1]) + 0.0f; dx = 1.0f; }
                if (y >= 1) { ny = (c - X[i - w]) + 0.0f; dy = 1.0f; }
                if (x + 1 <= w - 1) { nx = nx + (X[i + 1] - c); dx = dx + 1.0f; }
                if (y + 1 <= h - 1) { ny = ny + (X[i + w] - c); dy = dy + 1.0f; }
                const float gx = nx / dx, gy = ny / dy;
                G[2 * i] = gx;
                G[2 * i + 1] = gy;
                S1[3 * i] = gx * gx;
                S1[3 * i + 1] = gx * gy;
                S1[3 * i + 2] = gy * gy;
            }
            __syncthreads();
            lds_gauss<7, 3>(S1, S2, TA, w, h, g7, row);
            for (int it = 0; it < 3; it++) {
                // calc_b (flow.rs:334-357), blurred r = 7
                for (int i = threadIdx.x; i < n; i += COARSE_THREADS) {
                    const int y = row(i), x = i - y * w;
                    const float sx = (float)x - F[2 * i], sy = (float)y - F[2 * i + 1];
                    float bx = 0.0f, by = 0.0f;
                    if (!(sx < 0.0f || (float)w - 1.0f < sx || sy < 0.0f || (float)h - 1.0f < sy)) {
                        const float r = bilinear1(Y, w, h, sx, sy) - X[i];
                        bx = r * G[2 * i];
                        by = r * G[2 * i + 1];
                    }
                    S1[2 * i] = bx;
                    S1[2 * i + 1] = by;
                }
                __syncthreads();
                lds_gauss<7, 2>(S1, S2, S1, w, h, g7, row);
                // new_disp (flow.rs:361-371), the flow blurred r = 5
                for (int i = threadIdx.x; i < n; i += COARSE_THREADS) {
                    const float aa = TA[3 * i], ab = TA[3 * i + 1], ad = TA[3 * i + 2];
                    const float bx = S1[2 * i], by = S1[2 * i + 1];
                    float u = F[2 * i], v = F[2 * i + 1];
                    const float t = ab * ab - ad * aa;
                    if (fabsf(t) > 40.0f) {
                        u = (ab * by - ad * bx) / t + u;
                        v = (ab * bx - aa * by) / t + v;
                    }
                    S2[2 * i] = u;
                    S2[2 * i + 1] = v;
                }
                __syncthreads();
                lds_gauss<5, 2>(S2, S1, F, w, h, g5, row);
                if (it == 1 && l == 0 && !skip_refine && w >= 4 && h >= 4) {
                    lds_sweep<true, 1>(X, Y, F, S1, w, h, row);
                    lds_sweep<true, -1>(X, Y, F, S1, w, h, row);
                    lds_sweep<false, 1>(X, Y, F, S1, w, h, row);
                    lds_sweep<false, -1>(X, Y, F, S1, w, h, row);
                }
            }
            if (l != 0 && !skip_refine && w >= 4 && h >= 4) {
                lds_sweep<true, 1>(X, Y, F, S1, w, h, row);
                lds_sweep<true, -1>(X, Y, F, S1, w, h, row);
                lds_sweep<false, 1>(X, Y, F, S1, w, h, row);
                lds_sweep<false, -1>(X, Y, F, S1, w, h, row);
            }
            // the range check while no level has found a pixel in range (flow.rs:1291-1297), the
            // damping at the start level (flow.rs:1298-1300)
            bool zero = false;
            if (active) {
                int any = 0;
                for (int i = threadIdx.x; i < n && !any; i += COARSE_THREADS) {
                    const int y = row(i), x = i - y * w;
                    const float px = (float)x + F[2 * i], py = (float)y + F[2 * i + 1];
                    any = 0.0f <= px && px <= (float)(w - 1) && 0.0f <= py && py <= (float)(h - 1);
                }
                any = __syncthreads_or(any);
                zero = !any;
                active = !any;
            }
            const bool damp = l == J.start && J.scale_damp;
            if (zero || damp) {
                for (int i = threadIdx.x; i < 2 * n; i += COARSE_THREADS) {
                    float u = F[i];
                    if (zero) u = 0.0f;
                    if (damp) u = J.damp * u;
                    F[i] = u;
                }
                __syncthreads();
            }
            if (l > 0) {
                // upsample_flow (flow.rs:376-391) into the next level: S1, or flow[0] at the handoff
                const int dw = J.w[l - 1], dh = J.h[l - 1];
                const float sx = J.up[l][0], sy = J.up[l][1], rx = J.up[l][2], ry = J.up[l][3];
                float* dst = l - 1 > J.top ? S1 : J.flow[0];
                const RowDiv drow(dw);
                for (int i = threadIdx.x; i < dw * dh; i += COARSE_THREADS) {
                    const int y = drow(i), x = i - y * dw;
                    int ix, iy;
                    float ax, ay;
                    cell((float)x * rx, w, ix, ax);
                    cell((float)y * ry, h, iy, ay);
                    const float* t0 = F + 2 * (iy * w + ix);
                    const float* t1 = t0 + 2 * w;
                    float out[2];
#pragma unroll
                    for (int c = 0; c < 2; c++) {
                        const float p00 = t0[c], p10 = t0[2 + c], p01 = t1[c], p11 = t1[2 + c];
                        const float top = (p10 - p00) * ax + p00;
                        const float bot = (p11 - p01) * ax + p01;
                        out[c] = (bot - top) * ay + top;
                    }
                    dst[2 * i] = out[0] * sx;
                    dst[2 * i + 1] = out[1] * sy;
                }
                __syncthreads();
                if (l - 1 > J.top) {
                    for (int i = threadIdx.x; i < 2 * dw * dh; i += COARSE_THREADS) F[i] = S1[i];
                    __syncthreads();
                }
            } else {
                // the whole solve ran here: level 0 into the output, scanned for NaNs (the negation
                // pass scans it again when it runs)
                int nan = 0;
                for (int i = threadIdx.x; i < 2 * n; i += COARSE_THREADS) {
                    const float u = F[i];
                    J.flow[1][i] = u;
                    nan |= isnan(u) ? 1 : 0;
                }
                nan = __syncthreads_or(nan);
                if (threadIdx.x == 0 && nan) J.state[2] = 1;
            }
        }
        if (threadIdx.x == 0) {
            J.state[0] = active ? 1 : 0;
            J.state[1] = 0;
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------------------------------------
// The solve driver: per level, the kernel sequence of flow.rs:1271-1304 over every job.

// the flow starts at zero at the start level, in buffer 0; the range check starts active
// (flow.rs:1268-1270)
__global__ void k_solve_init(const SolveJob* __restrict__ jobs, int* __restrict__ queue) {
    const SolveJob& J = jobs[blockIdx.z];
    if (queue && blockIdx.x == 0 && blockIdx.z == 0 && threadIdx.x == 0) *queue = 0;  // the coarse solver's work queue
    long n = (long)J.w[J.start] * J.h[J.start] * 2;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) J.flow[0][i] = 0.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        J.state[0] = 1;
        J.state[1] = 0;
        J.state[2] = 0;
    }
}

// The tile shapes of the fused producer + Gaussian kernels, picked by measurement: a tile of
// 64 x TY outputs (TY 32, 16 or 64: 256, 128 or 512 lanes) whose region producer keeps PU pixels'
// loads in flight per lane (4 or 1). Taller tiles recompute less of the region's halo and need
// more LDS per workgroup. Every variant computes every output element with the same operations,
// so the choice changes time, not bits. The first launch of a (kernel, level size, job count)
// class times each variant twice on the live data (the kernels are idempotent: their inputs and
// outputs are distinct buffers) and keeps the fastest for the process; launches while a graph is
// being captured use the default shape until a plain launch has tuned their class; a shape
// forced through tile_force applies to every launch.
static std::mutex& tune_mu() {
    static std::mutex m;
    return m;
}
static std::map<std::tuple<int, int, int>, int>& tune_table() {
    static std::map<std::tuple<int, int, int>, int> t;
    return t;
}
static std::atomic<int> g_tile_force{-1};
constexpr int TILE_NVAR = 6;

int tile_force(int v) {
    if (v < -1 || v >= TILE_NVAR) return -1;
    g_tile_force.store(v);
    return 0;
}

template <int C, int R, int MODE, int TY, int PU>
static void tile_launch_v(hipStream_t s, const SolveJob* d_jobs, int l, int src, int dst, const Taps& t, int mw, int mh, int nj,
                          int scan) {
    const int tx = (mw + TILE_X - 1) / TILE_X, ty = (mh + TY - 1) / TY;
    const long ntile = (long)tx * ty * nj;
    k_tile_gauss<C, R, MODE, TY, PU><<<dim3((unsigned)(8 * ((ntile + 7) / 8))), dim3(TILE_X * TY / TILE_RUN), 0, s>>>(
        d_jobs, l, src, dst, t, tx, ty, nj, scan);
}

template <int C, int R, int MODE>
static void tile_launch_idx(int v, hipStream_t s, const SolveJob* d_jobs, int l, int src, int dst, const Taps& t, int mw,
                            int mh, int nj, int scan) {
    switch (v) {
        case 1: tile_launch_v<C, R, MODE, 32, 1>(s, d_jobs, l, src, dst, t, mw, mh, nj, scan); break;
        case 2: tile_launch_v<C, R, MODE, 16, 4>(s, d_jobs, l, src, dst, t, mw, mh, nj, scan); break;
        case 3: tile_launch_v<C, R, MODE, 16, 1>(s, d_jobs, l, src, dst, t, mw, mh, nj, scan); break;
        case 4: tile_launch_v<C, R, MODE, 64, 4>(s, d_jobs, l, src, dst, t, mw, mh, nj, scan); break;
        case 5: tile_launch_v<C, R, MODE, 64, 1>(s, d_jobs, l, src, dst, t, mw, mh, nj, scan); break;
        default: tile_launch_v<C, R, MODE, 32, 4>(s, d_jobs, l, src, dst, t, mw, mh, nj, scan); break;
    }
}

static int ceil_log2(long v) {
    int k = 0;
    while ((1L << k) < v) k++;
    return k;
}

template <int C, int R, int MODE>
static void tile_launch(hipStream_t s, const SolveJob* d_jobs, int l, int src, int dst, const Taps& t, int mw, int mh,
                        int nj, int scan = 0) {
    const int forced = g_tile_force.load();
    if (forced >= 0) {
        tile_launch_idx<C, R, MODE>(forced, s, d_jobs, l, src, dst, t, mw, mh, nj, scan);
        return;
    }
    const auto key = std::make_tuple(MODE, ceil_log2((long)mw * mh), ceil_log2(nj));
    int choice = -1;
    {
        std::lock_guard<std::mutex> g(tune_mu());
        auto it = tune_table().find(key);
        if (it != tune_table().end()) choice = it->second;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(s, &cs);
    if (choice < 0 && cs == hipStreamCaptureStatusNone) {
        // two rounds over the shapes; each shape's time is the faster of its two launches
        constexpr int NL = 2 * TILE_NVAR;
        hipEvent_t ev[NL + 1];
        for (auto& e : ev) (void)hipEventCreate(&e);
        (void)hipEventRecord(ev[0], s);
        for (int i = 0; i < NL; i++) {
            tile_launch_idx<C, R, MODE>(i % TILE_NVAR, s, d_jobs, l, src, dst, t, mw, mh, nj, scan);
            (void)hipEventRecord(ev[i + 1], s);
        }
        (void)hipEventSynchronize(ev[NL]);
        float best = 0.0f;
        for (int v = 0; v < TILE_NVAR; v++) {
            float a = 0.0f, b = 0.0f;
            (void)hipEventElapsedTime(&a, ev[v], ev[v + 1]);
            (void)hipEventElapsedTime(&b, ev[v + TILE_NVAR], ev[v + TILE_NVAR + 1]);
            const float ms = a < b ? a : b;
            if (choice < 0 || ms < best) { best = ms; choice = v; }
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        std::lock_guard<std::mutex> g(tune_mu());
        tune_table()[key] = choice;
    }
    // with MODE new_disp the default keeps one pixel's loads in flight (its producer is heavier)
    if (choice < 0) choice = MODE == P_NEWDISP ? 1 : 0;
    tile_launch_idx<C, R, MODE>(choice, s, d_jobs, l, src, dst, t, mw, mh, nj, scan);
}

int tile_tuning_report(int* out, int cap) {
    std::lock_guard<std::mutex> g(tune_mu());
    int n = 0;
    for (auto& kv : tune_table()) {
        if (n * 4 + 3 < cap) {
            out[n * 4] = std::get<0>(kv.first);
            out[n * 4 + 1] = std::get<1>(kv.first);
            out[n * 4 + 2] = std::get<2>(kv.first);
            out[n * 4 + 3] = kv.second;
        }
        n++;
    }
    return n;
}



static void launch_solve_range(hipStream_t s, const SolveJob* d_jobs, const std::vector<SolveJob>& jobs,
                               const PlaneT* d_planes, const std::vector<PlaneT>& planes, int* queue, bool skip_refine,
                               bool raw) {
    int nj = (int)jobs.size();
    if (nj == 0) return;
    int max_top = -1, coarse = 0;
    long max_init = 0;
    double Pc = 0;  // pixels the coarse solver reads and writes
    for (auto& j : jobs) {
        max_top = std::max(max_top, j.top);
        max_init = std::max(max_init, (long)j.w[j.start] * j.h[j.start] * 2);
        if (j.top < j.start) {
            coarse++;
            for (int l = j.start; l > j.top; l--) Pc += (double)j.w[l] * j.h[l];
        }
    }
    TWX_LAUNCH(s, F_SOLVE_INIT, 0.0,
               k_solve_init<<<dim3((unsigned)((max_init + 255) / 256), 1, nj), dim3(256), 0, s>>>(d_jobs, coarse ? queue : nullptr));
    const Taps g7 = gaussian_taps(7.0f), g5 = gaussian_taps(5.0f);
    if (coarse) {
        // X and Y 4 + 4 in per pixel of each coarse level, the flow handed over 8 out
        TWX_LAUNCH(s, F_COARSE, 16.0 * Pc,
                   k_coarse_solve<<<dim3((unsigned)std::min(coarse, 256)), dim3(COARSE_THREADS), 0, s>>>(
                       d_jobs, nj, queue, g7, g5, skip_refine ? 1 : 0));
    }
    // block matching on the flow in buffer `buf`
    auto bm = [&](int l, int mw, int mh, double P, int buf) {
        dim3 tb(256), tg((mw + 31) / 32, (mh + 31) / 32, nj);
        // a lane pair per chain, 32 chains per 64-lane block; column sweeps with carried own costs:
        // a lane per chain, 64 per block
        dim3 cb(64), rg((mh + 31) / 32, 1, nj), cg((mw + 31) / 32, 1, nj), cg1((mw + 63) / 64, 1, nj);
        // the transposed X and Y of the level, once per distinct frame: 4 + 4 B per frame pixel
        int fw = 0, fh = 0;
        double Pf = 0;
        for (auto& f : planes)
            if (l <= f.top) {
                fw = std::max(fw, f.w[l]);
                fh = std::max(fh, f.h[l]);
                Pf += (double)f.w[l] * f.h[l];
            }
        if (Pf > 0)
            TWX_LAUNCH(s, F_BM_TRANSPOSE, 8.0 * Pf,
                       (k_plane_transpose<<<dim3((fw + 31) / 32, (fh + 31) / 32, (unsigned)planes.size()), tb, 0, s>>>(
                           d_planes, l)));
        // sweeps: A 4 + B 4 + flow read and write 16 B/px each (the carried costs, 4 B/px written
        // by sweeps 2 and 3 and read by sweeps 3 and 4, are layout work the CPU path does not have)
        if (P <= BM_CROSS_MAX_PIXELS) {
            // the row sweeps move the flow between layouts; no costs are carried (storing them in
            // the plain layout from a row sweep cost more than the column sweeps saved)
            TWX_LAUNCH(s, F_BM_ROWS, 24.0 * P,
                       (k_bm_sweep<true, 1, BM_CHUNK, true, false, false><<<rg, cb, 0, s>>>(d_jobs, l, buf)));
            TWX_LAUNCH(s, F_BM_ROWS, 24.0 * P,
                       (k_bm_sweep<true, -1, BM_CHUNK, true, false, false><<<rg, cb, 0, s>>>(d_jobs, l, buf)));
            TWX_LAUNCH(s, F_BM_COLS, 24.0 * P,
                       (k_bm_sweep<false, 1, BM_CHUNK, false, false, false><<<cg, cb, 0, s>>>(d_jobs, l, buf)));
            TWX_LAUNCH(s, F_BM_COLS, 24.0 * P,
                       (k_bm_sweep<false, -1, BM_CHUNK, false, false, false><<<cg, cb, 0, s>>>(d_jobs, l, buf)));
        } else {
            // the rows on the transposed flow (8 + 8 B/px there, and back 8 + 8 B/px with the costs
            // 4 + 4); sweep 2 stores the costs, which the column sweeps carry
            TWX_LAUNCH(s, F_BM_TRANSPOSE, 16.0 * P, (k_bm_transpose<false><<<tg, tb, 0, s>>>(d_jobs, l, buf)));
            TWX_LAUNCH(s, F_BM_ROWS, 24.0 * P,
                       (k_bm_sweep<true, 1, BM_CHUNK, false, false, false><<<rg, cb, 0, s>>>(d_jobs, l, buf)));
            TWX_LAUNCH(s, F_BM_ROWS, 24.0 * P,
                       (k_bm_sweep<true, -1, BM_CHUNK, false, false, true><<<rg, cb, 0, s>>>(d_jobs, l, buf)));
            TWX_LAUNCH(s, F_BM_TRANSPOSE, 16.0 * P, (k_bm_transpose<true><<<tg, tb, 0, s>>>(d_jobs, l, buf)));
            TWX_LAUNCH(s, F_BM_COLS, 24.0 * P,
                       (k_bm_sweep<false, 1, BM_CHUNK, false, true, true><<<cg1, cb, 0, s>>>(d_jobs, l, buf)));
            TWX_LAUNCH(s, F_BM_COLS, 24.0 * P,
                       (k_bm_sweep<false, -1, BM_CHUNK, false, true, false><<<cg1, cb, 0, s>>>(d_jobs, l, buf)));
        }
    };
    for (int l = max_top; l >= 0; l--) {
        int mw = 0, mh = 0;
        double P = 0, Pn = 0;  // pixels of this level and of the next finer one, over the jobs
        double Ps = 0;         // pixels of the jobs that start at this level
        for (auto& j : jobs)
            if (l <= j.top) {
                mw = std::max(mw, j.w[l]);
                mh = std::max(mh, j.h[l]);
                P += (double)j.w[l] * j.h[l];
                if (l == j.start) Ps += (double)j.w[l] * j.h[l];
                if (l > 0) Pn += (double)j.w[l - 1] * j.h[l - 1];
            }
        dim3 blk(256), grd((mw + 255) / 256, mh, nj);
        // grad + tensor, blurred r = 7, once per distinct X: X 4 B in, grad 8 + tensor 12 B out. The
        // jobs that compute them lead the range; a job sharing another's planes is skipped
        int nt = 0, tw = 0, th = 0;
        double Pt = 0;
        for (int k = 0; k < nj; k++)
            if (!jobs[k].tensor_shared && l <= jobs[k].top) {
                nt = k + 1;
                tw = std::max(tw, jobs[k].w[l]);
                th = std::max(th, jobs[k].h[l]);
                Pt += (double)jobs[k].w[l] * jobs[k].h[l];
            }
        if (nt > 0)
            TWX_LAUNCH(s, F_GRAD_TENSOR, 24.0 * Pt, tile_launch<3, 7, P_TENSOR>(s, d_jobs, l, 0, 0, g7, tw, th, nt));
        for (int it = 0; it < 3; it++) {
            const int src = it & 1, dst = src ^ 1;
            // calc_b + B blur: flow 8 + X 4 + Y 4 + grad 8 in, B 8 out
            TWX_LAUNCH(s, F_CALC_B, 32.0 * P, tile_launch<2, 7, P_CALCB>(s, d_jobs, l, src, 0, g7, mw, mh, nj));
            // new_disp + flow blur: tensor 12 + B 8 + flow 8 in, flow 8 out; a raw solve's last pass scans
            // for NaNs
            const int scan = raw && l == 0 && it == 2 ? 1 : 0;
            TWX_LAUNCH(s, F_NEW_DISP, 36.0 * P, tile_launch<2, 5, P_NEWDISP>(s, d_jobs, l, src, dst, g5, mw, mh, nj, scan));
            if (it == 1 && l == 0 && !skip_refine) bm(l, mw, mh, P, 0);
        }
        if (l != 0 && !skip_refine) bm(l, mw, mh, P, 1);
        // the check reads the flow while no level has found a pixel in range, and the flow is only
        // rewritten when it is zeroed or damped: in practice both at the start level alone
        TWX_LAUNCH(s, F_RANGE, 24.0 * Ps, k_range<<<dim3((unsigned)nj), dim3(1024), 0, s>>>(d_jobs, l, 1));
        if (l > 0) {
            int uw = 0, uh = 0;
            for (auto& j : jobs)
                if (l <= j.top) { uw = std::max(uw, j.w[l - 1]); uh = std::max(uh, j.h[l - 1]); }
            TWX_LAUNCH(s, F_UPSAMPLE, 8.0 * P + 8.0 * Pn,
                       k_upsample<<<dim3((uw + 255) / 256, (uh + UP_ROWS - 1) / UP_ROWS, nj), blk, 0, s>>>(d_jobs, l));
        }
    }
    int w0 = 0, h0 = 0;
    double P0 = 0;
    for (auto& j : jobs) { w0 = std::max(w0, j.w[0]); h0 = std::max(h0, j.h[0]); P0 += (double)j.w[0] * j.h[0]; }
    if (raw) return;
    // about 8 float4 per lane over the largest job
    const long nb = std::max<long>(1, ((long)w0 * h0 / 2 + 256 * 8 - 1) / (256 * 8));
    TWX_LAUNCH(s, F_NEGATE, 16.0 * P0, k_negate<<<dim3((unsigned)nb, 1, nj), dim3(256), 0, s>>>(d_jobs));
}

void launch_solve(hipStream_t s, const SolveJob* d_jobs, const std::vector<SolveJob>& jobs, const PlaneT* d_planes,
                  const std::vector<PlaneT>& planes, const std::vector<SolveBucket>& buckets, bool skip_refine,
                  bool raw) {
    for (const SolveBucket& b : buckets)
        launch_solve_range(s, d_jobs + b.j0, std::vector<SolveJob>(jobs.begin() + b.j0, jobs.begin() + b.j1),
                           d_planes + b.f0, std::vector<PlaneT>(planes.begin() + b.f0, planes.begin() + b.f1), b.queue,
                           skip_refine, raw);
}

__global__ void k_nan_scan(const float* __restrict__ d, size_t n, int* flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && isnan(d[i])) atomicOr(flag, 1);
}
void launch_nan_scan(hipStream_t s, const float* data, size_t n, int* flag) {
    if (n == 0) return;
    k_nan_scan<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(data, n, flag);
}

// -------------------------------------------------------------------------------------------
// Inverse warp (warp.rs:99-786, hold edges), blend (blend.rs:55-257), finish (image.rs:205-245)

// warp.rs:178-215
__device__ __forceinline__ bool clip_to_rect(float x, float y, float& ix, float& iy, float w1, float h1) {
    for (int n = 0; n < 4096; n++) {
        if (ix < 0.0f) {
            float dy = y - iy, dxs = x - ix, num = x - 0.0f;
            ix = 0.0f;
            iy = y - (num / dxs) * dy;
            continue;
        }
        if (ix > w1) {
            float dy = y - iy, dxs = x - ix, num = x - w1;
            ix = w1;
            iy = y - (num / dxs) * dy;
            continue;
        }
        if (iy < 0.0f) {
            float dyy = y - iy, num = y - 0.0f, dx = x - ix;
            ix = x - (num / dyy) * dx;
            iy = 0.0f;
            continue;
        }
        if (iy > h1) {
            float dyy = y - iy, num = y - h1, dx = x - ix;
            ix = x - (num / dyy) * dx;
            iy = h1;
            continue;
        }
        return true;
    }
    return false;
}

template <int FMT> struct Px;
template <> struct Px<0> { using T = uint8_t; };
template <> struct Px<1> { using T = uint16_t; };
template <> struct Px<2> { using T = float; };

// Pixel i of an engine image (images hold fewer than 2^28 pixels: a 32-bit byte offset from the
// image's base).
template <int FMT>
__device__ __forceinline__ void load4(const void* img, uint32_t i, float o[4]) {
    typedef unsigned char u8x4 __attribute__((ext_vector_type(4)));
    typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    gchar* q = (gchar*)img + i * (uint32_t)(4 * (FMT == 0 ? 1 : (FMT == 1 ? 2 : 4)));
    if (FMT == 0) {
        const u8x4 p = *(const __attribute__((address_space(1))) u8x4*)q;
        o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = p.w;
    } else if (FMT == 1) {
        const u16x4 p = *(const __attribute__((address_space(1))) u16x4*)q;
        o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = p.w;
    } else {
        const f32x4 p = *(const __attribute__((address_space(1))) f32x4*)q;
        o[0] = p.x; o[1] = p.y; o[2] = p.z; o[3] = p.w;
    }
}


// The sample point of output pixel (x, y) of W(src, flow, amount) (warp.rs:477-486): the flow
// sampled bilinearly at (x * rx, y * ry) (Flow::bilerp, warp.rs:415-442), scaled by amount / r.
// rx = (fw - 1) / (W - 1), ry = (fh - 1) / (H - 1) and sx = amount / rx, sy = amount / ry are the
// same for every pixel of a frame: the host computes them once (frame_job_ratios).
__device__ __forceinline__ void flow_sample(const float* __restrict__ flow, int fw, int fh, float rx, float ry, int x, int y,
                                            float d[2]) {
    float xf = (float)x, yf = (float)y;
    float qx = xf * rx, qy = yf * ry;
    int fw1 = fw - 1, fh1 = fh - 1;
    // qx, qy lie in [0, fw - 1] x [0, fh - 1] (the vector path needs frames of 2 x 2 or more, so
    // rx and ry are finite): plain truncation is cvtt there
    int32_t jx = (int32_t)qx, jy = (int32_t)qy;
    float gx = qx - (float)jx, gy = qy - (float)jy;
    if (jx >= fw1) { jx = fw1 - 1; gx = 1.0f; }
    if (jy >= fh1) { jy = fh1 - 1; gy = 1.0f; }
    const uint32_t q00 = (uint32_t)jy * fw + jx;
    const gfloat* gf = (const gfloat*)flow;
    float t0[4], t1[4];  // the two flow pixels of the cell's top row, then of its bottom row
    ldrun<4>(gf, 2 * q00, t0);
    ldrun<4>(gf, 2 * (q00 + fw), t1);
#pragma unroll
    for (int k = 0; k < 2; k++) {
        float p00 = t0[k], p10 = t0[2 + k], p01 = t1[k], p11 = t1[2 + k];
        float top = (p10 - p00) * gx + p00;
        float bot = (p11 - p01) * gx + p01;
        d[k] = (bot - top) * gy + top;
    }
}
// The sample point of output (x, y) from the flow sample d, scaled by (sx, sy). raw: the flow
// holds v where the reference reads -1 * v. Negation is exact, so the reference's sample of the
// negated flow is -d except that an exactly zero result may carry the other sign, which vanishes in
// the sum with the pixel position; xf - d * sx is xf + (-d) * sx bit for bit.
__device__ __forceinline__ void warp_point_at(const float d[2], float sx, float sy, int x, int y, float& ix, float& iy,
                                              bool raw) {
    const float xf = (float)x, yf = (float)y;
    ix = raw ? xf - d[0] * sx : d[0] * sx + xf;
    iy = raw ? yf - d[1] * sy : d[1] * sy + yf;
}
__device__ __forceinline__ void warp_point(const float* __restrict__ flow, int fw, int fh, float rx, float ry, float sx,
                                           float sy, int x, int y, float& ix, float& iy, bool raw) {
    float d[2];
    flow_sample(flow, fw, fh, rx, ry, x, y, d);
    warp_point_at(d, sx, sy, x, y, ix, iy, raw);
}

// bilerp4 (warp.rs:306-331) at a point inside [0, W-1] x [0, H-1], before the store rule, in two
// parts: the cell's four texels (t[0] = (kx, ky), t[1] = (kx + 1, ky), t[2] = (kx, ky + 1),
// t[3] = (kx + 1, ky + 1)) and the fractions, then the lerps. A caller with several samples
// fetches them all before mixing any, so their loads are in flight together.
template <int FMT>
__device__ __forceinline__ void bilerp_fetch(const void* src, int W, int H, float ix, float iy, float t[4][4], float& ax,
                                             float& ay) {
    int w1i = W - 1, h1i = H - 1;
    int32_t kx = (int32_t)ix, ky = (int32_t)iy;  // the point is inside the image: cvtt's values
    ax = ix - (float)kx;
    ay = iy - (float)ky;
    if (kx >= w1i) { kx = w1i - 1; ax = 1.0f; }
    if (ky >= h1i) { ky = h1i - 1; ay = 1.0f; }
    uint32_t i00 = (uint32_t)ky * W + kx, i01 = i00 + W;
    load4<FMT>(src, i00, t[0]);
    load4<FMT>(src, i00 + 1, t[1]);
    load4<FMT>(src, i01, t[2]);
    load4<FMT>(src, i01 + 1, t[3]);
}
template <int FMT>
__device__ __forceinline__ void bilerp_mix(const float t[4][4], float ax, float ay, float o[4]) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float top = t[0][k] + (t[1][k] - t[0][k]) * ax;
        float bot = t[2][k] + (t[3][k] - t[2][k]) * ax;
        o[k] = top + (bot - top) * ay;
    }
    if (FMT == 2 && (isnan(o[0]) || isnan(o[1]) || isnan(o[2]) || isnan(o[3]))) {
        // warp.rs:334-343: the same lerps again with the NaN rule
#pragma unroll
        for (int k = 0; k < 4; k++)
            o[k] = nr_lerp_bf(nr_lerp_bf(t[0][k], t[1][k], ax), nr_lerp_bf(t[2][k], t[3][k], ax), ay);
    }
}
template <int FMT>
__device__ __forceinline__ void bilerp_src(const void* src, int W, int H, float ix, float iy, float o[4]) {
    float t[4][4], ax, ay;
    bilerp_fetch<FMT>(src, W, H, ix, iy, t, ax, ay);
    bilerp_mix<FMT>(t, ax, ay, o);
}

// One warped pixel W(src, flow, amount) at output (x, y) with hold edges (warp.rs:500-512), as the
// stored sample values (integer formats hold exact integers in the floats). Returns false for a
// NaN coordinate.
template <int FMT>
__device__ __forceinline__ bool warp_pixel(const void* src, int W, int H, const float* __restrict__ flow, int fw, int fh,
                                           const FrameJob& F, int wi, int x, int y, float o[4]) {
    float ix, iy;
    warp_point(flow, fw, fh, F.rx, F.ry, F.sx[wi], F.sy[wi], x, y, ix, iy, F.flow_raw != 0);
    if (!clip_to_rect((float)x, (float)y, ix, iy, (float)(W - 1), (float)(H - 1)) || isnan(ix) || isnan(iy)) {
        o[0] = o[1] = o[2] = o[3] = 0.0f;
        return false;
    }
    float s[4];
    bilerp_src<FMT>(src, W, H, ix, iy, s);
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = store_rule<FMT>(s[k]);
    return true;
}

// Both warped pixels of a blend, W(A, F_BA, frac) into pa and W(B, F_AB, wa) into pb, with hold
// edges: the values of two warp_pixel calls, with the two flow samples' loads issued together and
// then the eight texels' loads together (a point that clip_to_rect refuses or that is NaN samples
// the image's first cell and gives zeros), so a pixel waits for two memory round trips, not four.
// Returns false when either point is refused.
template <int FMT>
__device__ __forceinline__ bool warp_pixel_pair_at(const FrameJob& F, const float dba[2], const float dab[2], int x, int y,
                                                   float pa[4], float pb[4]) {
    float ixa, iya, ixb, iyb;
    warp_point_at(dba, F.sx[0], F.sy[0], x, y, ixa, iya, F.flow_raw != 0);
    warp_point_at(dab, F.sx[1], F.sy[1], x, y, ixb, iyb, F.flow_raw != 0);
    const float w1 = (float)(F.W - 1), h1 = (float)(F.H - 1);
    const bool oka = clip_to_rect((float)x, (float)y, ixa, iya, w1, h1) && !isnan(ixa) && !isnan(iya);
    const bool okb = clip_to_rect((float)x, (float)y, ixb, iyb, w1, h1) && !isnan(ixb) && !isnan(iyb);
    ixa = oka ? ixa : 0.0f;
    iya = oka ? iya : 0.0f;
    ixb = okb ? ixb : 0.0f;
    iyb = okb ? iyb : 0.0f;
    float ta[4][4], tb[4][4], axa, aya, axb, ayb;
    bilerp_fetch<FMT>(F.a, F.W, F.H, ixa, iya, ta, axa, aya);
    bilerp_fetch<FMT>(F.b, F.W, F.H, ixb, iyb, tb, axb, ayb);
    float sa[4], sb[4];
    bilerp_mix<FMT>(ta, axa, aya, sa);
    bilerp_mix<FMT>(tb, axb, ayb, sb);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        pa[k] = oka ? store_rule<FMT>(sa[k]) : 0.0f;
        pb[k] = okb ? store_rule<FMT>(sb[k]) : 0.0f;
    }
    return oka && okb;
}
template <int FMT>
__device__ __forceinline__ bool warp_pixel_pair(const FrameJob& F, int x, int y, float pa[4], float pb[4]) {
    float dba[2], dab[2];
    flow_sample(F.f_ba, F.fw, F.fh, F.rx, F.ry, x, y, dba);
    flow_sample(F.f_ab, F.fw, F.fh, F.rx, F.ry, x, y, dab);
    return warp_pixel_pair_at<FMT>(F, dba, dab, x, y, pa, pb);
}

// The store of a faded edge sample (warp.rs:237-295): the stored sample times f,
// trunc((double)(b * f) + 0.5), low byte or word; float v * f.
template <int FMT>
__device__ __forceinline__ float store_faded(float v, float f) {
    if (FMT == 2) return v * f;
    float b = FMT == 0 ? sat_u8(v) : sat_u16(v);
    float s = b * f;
    return (float)(cvtt64d((double)s + 0.5) & (FMT == 0 ? 0xff : 0xffff));
}

// One pixel of the warp without hold edges (warp.rs:519-561): transparent one pixel or more
// outside, faded in the border band, plain inside; `cov` gets the coverage byte. Returns false
// only where the clip of a border-band point fails (never for finite data).
template <int FMT>
__device__ __forceinline__ bool warp_fade_pixel(const void* src, int W, int H, const float* __restrict__ flow, int fw,
                                                int fh, const FrameJob& F, int wi, int x, int y, float o[4], int& cov) {
    float ix, iy;
    warp_point(flow, fw, fh, F.rx, F.ry, F.sx[wi], F.sy[wi], x, y, ix, iy, F.flow_raw != 0);
    const float wf = (float)W, hf = (float)H, w1 = (float)(W - 1), h1 = (float)(H - 1);
    o[0] = o[1] = o[2] = o[3] = 0.0f;
    if (!(ix > -1.0f && ix < wf && iy > -1.0f && iy < hf)) {
        cov = 0;
        return true;
    }
    float s[4];
    if (ix >= 0.0f && ix <= w1 && iy >= 0.0f && iy <= h1) {
        bilerp_src<FMT>(src, W, H, ix, iy, s);
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = store_rule<FMT>(s[k]);
        cov = 255;
        return true;
    }
    float f = ix < 0.0f ? ix + 1.0f : (ix <= w1 ? 1.0f : wf - ix);
    if (iy < 0.0f) f *= iy + 1.0f;
    else if (iy > h1) f *= hf - iy;
    if (!clip_to_rect((float)x, (float)y, ix, iy, w1, h1) || isnan(ix) || isnan(iy)) {
        cov = 0;
        return false;
    }
    bilerp_src<FMT>(src, W, H, ix, iy, s);
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = store_faded<FMT>(s[k], f);
    float cv = f * 255.0f + 0.5f;
    cov = cv <= 255.0f ? (int)(cvtt64d((double)cv) & 0xff) : 255;
    return true;
}

// `a` over `b` (smartblend.rs:167-178), premultiplied, alpha in units of white
__device__ __forceinline__ void over4(const float a[4], const float b[4], float white, float o[4]) {
    if (a[0] == 0.0f) { for (int k = 0; k < 4; k++) o[k] = b[k]; return; }
    if (a[0] == white) { for (int k = 0; k < 4; k++) o[k] = a[k]; return; }
    float ia = 1.0f - a[0] / white;
    o[0] = (white - b[0]) * (a[0] / white) + b[0];
    for (int k = 1; k < 4; k++) o[k] = b[k] * ia + a[k];
}

// over4 with the NaN rule's operations
__device__ __forceinline__ void over4_nr(const float a[4], const float b[4], float white, float o[4]) {
    if (a[0] == 0.0f) { for (int k = 0; k < 4; k++) o[k] = b[k]; return; }
    if (a[0] == white) { for (int k = 0; k < 4; k++) o[k] = a[k]; return; }
    float ia = nr_sub(1.0f, nr_div(a[0], white));
    o[0] = nr_add(nr_mul(nr_sub(white, b[0]), nr_div(a[0], white)), b[0]);
    for (int k = 1; k < 4; k++) o[k] = nr_add(nr_mul(b[k], ia), a[k]);
}

// One pixel of the Smart Blend composite (smartblend.rs:193-212) with its stores
// (smartblend.rs:84-150): the main path `+ 0.5` and saturation (integers) or `+ 0.0` (float), the
// "alpha is 0" fallback the plain truncation (integers) or the value (float).
template <int FMT>
__device__ __forceinline__ void smart_px(const float p1[4], const float p2[4], float w, float o[4]) {
    const float white = FMT == 0 ? 255.0f : (FMT == 1 ? 65535.0f : 1.0f);
    const int64_t m = FMT == 0 ? 0xff : 0xffff;
    float c1[4], c2[4], v[4];
    over4(p1, p2, white, c1);
    over4(p2, p1, white, c2);
    float amax = p1[0] > p2[0] ? p1[0] : p2[0];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = (c2[k] - c1[k]) * w + c1[k];
    // a NaN in the composite (float frames only): the arithmetic again with the NaN rule's
    // operations, under which the references were rendered (nanrule.rs:14-21)
    const bool nr = FMT == 2 && (isnan(v[0]) || isnan(v[1]) || isnan(v[2]) || isnan(v[3]));
    if (nr) {
        over4_nr(p1, p2, white, c1);
        over4_nr(p2, p1, white, c2);
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = nr_lerp(c1[k], c2[k], w);
    }
    bool raw = false;
    if (v[0] != amax) {
        if (v[0] == 0.0f) {
            raw = true;
            for (int k = 0; k < 4; k++) v[k] = c1[0] > c2[0] ? c1[k] : c2[k];
        } else {
            float alpha = v[0];
            for (int k = 1; k < 4; k++) v[k] = nr ? nr_div(nr_mul(v[k], amax), alpha) : (v[k] * amax) / alpha;
            v[0] = amax;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (FMT == 2) { o[k] = raw ? v[k] : v[k] + 0.0f; continue; }
        if (raw) { o[k] = (float)(cvtt64f(v[k]) & m); continue; }
        float y = v[k] + 0.5f;
        o[k] = 0.0f > y ? 0.0f : (y > white ? white : (float)(cvtt64f(y) & m));
    }
}


// Image::to_frame of one pixel into the wire frame (image.rs:300-351, frame.rs:113-141): the
// values of finish_values, stored in the wire's channel order
template <int FMT>
__device__ __forceinline__ void finish_pixel(void* out, long n, long i, const float e[4]) {
    uint32_t v[4];
    finish_values<FMT>(e, v);
    if (FMT == 0) {
        reinterpret_cast<uchar4*>(out)[i] = make_uchar4((uint8_t)v[1], (uint8_t)v[2], (uint8_t)v[3], (uint8_t)v[0]);
    } else if (FMT == 1) {
        reinterpret_cast<ushort4*>(out)[i] = make_ushort4((uint16_t)v[1], (uint16_t)v[2], (uint16_t)v[3], (uint16_t)v[0]);
    } else {
        float* pl = reinterpret_cast<float*>(out);  // planes G, B, R, A
        pl[i] = bits_f(v[2]);
        pl[n + i] = bits_f(v[3]);
        pl[2 * n + i] = bits_f(v[1]);
        pl[3 * n + i] = bits_f(v[0]);
    }
}

template <int FMT>
__global__ void k_frames(const FrameJob* __restrict__ jobs) {
    // one block per 256 pixels of a row: the hardware deals consecutive blocks to the 8 XCDs in
    // turn, so each XCD keeps a vertical strip of the frame and the source rows that vertically
    // neighbouring pixels share stay in its L2 (HBM fetch 1.03x the algorithmic bytes; 64x4
    // blocks ran 2.6 % faster at 1.45x, and 64x4 blocks in per-XCD bands 7 % slower at 0.93x)
    const FrameJob& F = jobs[blockIdx.z];
    if (F.fmt != FMT) return;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;  // motion row
    if (x >= F.W || y >= F.H) return;
    long n = (long)F.W * F.H;
    long mi = (long)y * F.W + x;
    float pa[4], pb[4], e[4];
    bool ok = true;
    if (F.kind == 0) {
        load4<FMT>(F.a, mi, pa);
        load4<FMT>(F.b, mi, pb);
        if (FMT == 2 && F.plus_zero) {
            // blend.rs:280-300 takes another route for non-finite or huge samples (its NaN rule),
            // which is not served here: flag the frame (status 3)
            bool odd = false;
            for (int k = 0; k < 4; k++) {
                odd |= !isfinite(pa[k]) || !isfinite(pb[k]) || fabsf(pa[k]) > 1e30f || fabsf(pb[k]) > 1e30f;
                pa[k] = pa[k] + 0.0f;
                pb[k] = pb[k] + 0.0f;
            }
            if (odd) atomicOr(F.bad, 2);
        }
        blend_pixel<FMT>(pa, pb, F.blend_mode, F.weight, F.lerp_w, e);
    } else if (F.kind == 1 && F.pre_a) {
        load4<FMT>(F.pre_a, mi, e);
    } else if (F.kind == 1) {
        if (F.nearest_from_a) {
            if (F.warp_a_copy) load4<FMT>(F.a, mi, e);
            else ok = warp_pixel<FMT>(F.a, F.W, F.H, F.f_ba, F.fw, F.fh, F, 0, x, y, e);
        } else {
            if (F.warp_b_copy) load4<FMT>(F.b, mi, e);
            else ok = warp_pixel<FMT>(F.b, F.W, F.H, F.f_ab, F.fw, F.fh, F, 1, x, y, e);
        }
    } else if (F.kind >= 4) {
        // Smart Blend, first stage (smartblend.rs:323-419): the warps without hold edges and their
        // coverage; Nearest keeps its one warp, Blend the composite, Weighted Blend both warps and
        // their difference masks; the holes mask is 255 - max(cov1, cov2) (smart_finish)
        using T = typename Px<FMT>::T;
        int c1 = 255, c2 = 255;
        if (F.pre_a) {
            // pushed frames and their coverages (warping code 2)
            load4<FMT>(F.pre_a, mi, F.kind == 4 ? e : pa);
            c1 = F.pre_cov_a[mi];
            if (F.kind != 4) {
                load4<FMT>(F.pre_b, mi, pb);
                c2 = F.pre_cov_b[mi];
                if (F.kind == 5) smart_px<FMT>(pa, pb, F.frac, e);
            }
        } else if (F.kind == 4) {
            if (F.nearest_from_a) {
                if (F.warp_a_copy) load4<FMT>(F.a, mi, e);
                else ok = warp_fade_pixel<FMT>(F.a, F.W, F.H, F.f_ba, F.fw, F.fh, F, 0, x, y, e, c1);
            } else {
                if (F.warp_b_copy) load4<FMT>(F.b, mi, e);
                else ok = warp_fade_pixel<FMT>(F.b, F.W, F.H, F.f_ab, F.fw, F.fh, F, 1, x, y, e, c1);
            }
        } else {
            if (F.warp_a_copy) load4<FMT>(F.a, mi, pa);
            else ok = warp_fade_pixel<FMT>(F.a, F.W, F.H, F.f_ba, F.fw, F.fh, F, 0, x, y, pa, c1);
            if (F.warp_b_copy) load4<FMT>(F.b, mi, pb);
            else ok = warp_fade_pixel<FMT>(F.b, F.W, F.H, F.f_ab, F.fw, F.fh, F, 1, x, y, pb, c2) && ok;
            if (F.kind == 5) smart_px<FMT>(pa, pb, F.frac, e);
        }
        if (F.kind == 6) {
            T* oa = reinterpret_cast<T*>(F.ta) + mi * 4;
            T* ob = reinterpret_cast<T*>(F.tb) + mi * 4;
            float a0[4], b0[4];
            load4<FMT>(F.a, mi, a0);
            load4<FMT>(F.b, mi, b0);
            for (int k = 0; k < 4; k++) { oa[k] = (T)pa[k]; ob[k] = (T)pb[k]; }
            float da = (fabsf(pa[1] - a0[1]) + fabsf(pa[2] - a0[2])) + fabsf(pa[3] - a0[3]);
            float db = (fabsf(pb[1] - b0[1]) + fabsf(pb[2] - b0[2])) + fabsf(pb[3] - b0[3]);
            F.mask[0][0][mi] = da >= F.limit ? 255 : 0;
            F.mask[1][0][mi] = db >= F.limit ? 255 : 0;
        } else {
            T* oc = reinterpret_cast<T*>(F.comp) + mi * 4;
            for (int k = 0; k < 4; k++) oc[k] = (T)e[k];
        }
        // Nearest inverts its one coverage (smartblend.rs:336), the others the combined one
        const int hole = 255 - (F.kind == 4 ? c1 : max(c1, c2));
        F.holes[mi] = (uint8_t)hole;
        if (hole != 0) *F.has_hole = 1;  // the same value from every writer
        if (!ok) atomicOr(F.bad, 1);
        return;
    } else {
        if (F.pre_a) {
            load4<FMT>(F.pre_a, mi, pa);
            load4<FMT>(F.pre_b, mi, pb);
        } else if (!F.warp_a_copy && !F.warp_b_copy) {
            ok = warp_pixel_pair<FMT>(F, x, y, pa, pb);
        } else {
            if (F.warp_a_copy) load4<FMT>(F.a, mi, pa);
            else ok = warp_pixel<FMT>(F.a, F.W, F.H, F.f_ba, F.fw, F.fh, F, 0, x, y, pa);
            if (F.warp_b_copy) load4<FMT>(F.b, mi, pb);
            else ok = warp_pixel<FMT>(F.b, F.W, F.H, F.f_ab, F.fw, F.fh, F, 1, x, y, pb) && ok;
        }
        if (F.kind == 3) {
            // Weighted Blend, first stage: keep both warped frames and their difference masks
            // against the unwarped frames (filters.rs:295-335: (|dR| + |dG|) + |dB| >= limit)
            using T = typename Px<FMT>::T;
            T* oa = reinterpret_cast<T*>(F.ta) + mi * 4;
            T* ob = reinterpret_cast<T*>(F.tb) + mi * 4;
            float a0[4], b0[4];
            load4<FMT>(F.a, mi, a0);
            load4<FMT>(F.b, mi, b0);
            for (int k = 0; k < 4; k++) { oa[k] = (T)pa[k]; ob[k] = (T)pb[k]; }
            float da = (fabsf(pa[1] - a0[1]) + fabsf(pa[2] - a0[2])) + fabsf(pa[3] - a0[3]);
            float db = (fabsf(pb[1] - b0[1]) + fabsf(pb[2] - b0[2])) + fabsf(pb[3] - b0[3]);
            F.mask[0][0][mi] = da >= F.limit ? 255 : 0;
            F.mask[1][0][mi] = db >= F.limit ? 255 : 0;
            if (!ok) atomicOr(F.bad, 1);
            return;
        }
        blend_pixel<FMT>(pa, pb, F.blend_mode, F.weight, F.lerp_w, e);
    }
    if (!ok) atomicOr(F.bad, 1);
    if (!F.finish) {
        typename Px<FMT>::T* o = reinterpret_cast<typename Px<FMT>::T*>(F.out) + mi * 4;
        for (int k = 0; k < 4; k++) o[k] = (typename Px<FMT>::T)e[k];
        return;
    }
    // the composed image is in motion rows: row y is the output's row H-1-y (finish 2: rows as given)
    long oi = (long)(F.finish == 2 ? y : F.H - 1 - y) * F.W + x;
    finish_pixel<FMT>(F.out, n, oi, e);
}

// The common vector-path frame (kind 2: Blend of two inverse warps, neither an early-out copy,
// no pushed frames) in a kernel of its own, a fraction of k_frames' size (whose other kinds and
// their NaN-rule paths make a large body), with the same functions as k_frames' branch for it.
// The frames of one source pair go together (same_pair: the same frames, flows, size and format;
// the weights differ): a workgroup takes the pixel run of every frame of its group, the flows
// sampled once for all of them, and the later frames' texels come from the cache lines the first
// one brought in.
constexpr int FG_MAX = 384;
struct FrameGroups {
    int n;
    int first[FG_MAX];
    int count[FG_MAX];
};

template <int FMT>
__global__ void k_frames_blend_group(const FrameJob* __restrict__ jobs, FrameGroups G) {
    const int j0 = G.first[blockIdx.z], nj = G.count[blockIdx.z];
    const FrameJob& F0 = jobs[j0];
    if (F0.fmt != FMT) return;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;  // motion row
    if (x >= F0.W || y >= F0.H) return;
    const long n = (long)F0.W * F0.H;
    float dba[2], dab[2];
    flow_sample(F0.f_ba, F0.fw, F0.fh, F0.rx, F0.ry, x, y, dba);
    flow_sample(F0.f_ab, F0.fw, F0.fh, F0.rx, F0.ry, x, y, dab);
    for (int j = 0; j < nj; j++) {
        const FrameJob& F = jobs[j0 + j];
        float pa[4], pb[4], e[4];
        const bool ok = warp_pixel_pair_at<FMT>(F, dba, dab, x, y, pa, pb);
        blend_pixel<FMT>(pa, pb, F.blend_mode, F.weight, F.lerp_w, e);
        if (!ok) atomicOr(F.bad, 1);
        if (!F.finish) {
            typename Px<FMT>::T* o = reinterpret_cast<typename Px<FMT>::T*>(F.out) + ((long)y * F.W + x) * 4;
            for (int k = 0; k < 4; k++) o[k] = (typename Px<FMT>::T)e[k];
        } else {
            finish_pixel<FMT>(F.out, n, (long)(F.finish == 2 ? y : F.H - 1 - y) * F.W + x, e);
        }
    }
}

// Jobs k and k + 1 blend the same pair the same way but for their weights: the same source frames,
// flows, sizes, format, flow convention and finish.
static bool same_pair(const FrameJob& a, const FrameJob& b) {
    return a.a == b.a && a.b == b.b && a.f_ab == b.f_ab && a.f_ba == b.f_ba && a.W == b.W && a.H == b.H &&
           a.fw == b.fw && a.fh == b.fh && a.fmt == b.fmt && a.flow_raw == b.flow_raw && a.finish == b.finish &&
           a.rx == b.rx && a.ry == b.ry;
}

void frame_job_ratios(FrameJob& F) {
    F.rx = (float)(F.fw - 1) / (float)(F.W - 1);
    F.ry = (float)(F.fh - 1) / (float)(F.H - 1);
    F.sx[0] = F.frac / F.rx;
    F.sy[0] = F.frac / F.ry;
    F.sx[1] = F.wa / F.rx;
    F.sy[1] = F.wa / F.ry;
}

static std::vector<std::pair<int, int>> frame_buckets(const std::vector<FrameJob>& jobs) {
    std::vector<double> area;
    for (auto& f : jobs) area.push_back((double)f.W * f.H);
    return size_buckets(area);
}

static void launch_frames_range(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs);

void launch_frames(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs) {
    for (auto [b0, b1] : frame_buckets(jobs))
        launch_frames_range(s, d_jobs + b0, std::vector<FrameJob>(jobs.begin() + b0, jobs.begin() + b1));
}

static void launch_frames_run(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs, bool pair);

static void launch_frames_range(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs) {
    // runs of consecutive jobs: the blend pairs through k_frames_blend_group, the others through k_frames
    for (size_t r0 = 0; r0 < jobs.size();) {
        const bool pair = frame_is_blend_pair(jobs[r0]);
        size_t r1 = r0 + 1;
        while (r1 < jobs.size() && frame_is_blend_pair(jobs[r1]) == pair) r1++;
        launch_frames_run(s, d_jobs + r0, std::vector<FrameJob>(jobs.begin() + r0, jobs.begin() + r1), pair);
        r0 = r1;
    }
}

static void launch_frames_run(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs, bool pair) {
    if (jobs.empty()) return;
    int mw = 0, mh = 0;
    bool has[3] = {false, false, false};
    for (auto& f : jobs) { mw = std::max(mw, f.W); mh = std::max(mh, f.H); has[f.fmt] = true; }
    dim3 blk(256), grd((mw + 255) / 256, mh, (unsigned)jobs.size());
    // per output pixel: the two source pixels and the output (Bpp each), plus the flows read at
    // working size (8 B per flow and working pixel)
    double bytes = 0;
    for (auto& f : jobs) {
        double px = (double)f.W * f.H, bpp = 4.0 * sample_size(f.fmt);
        double flows = f.kind == 0 ? 0.0 : ((f.kind == 1 || f.kind == 4) ? 8.0 : 16.0) * f.fw * f.fh;
        // the frames read and written (Bpp each) and the u8 masks written
        double imgs = 3.0, masks = 0.0;
        if (f.kind == 1 || f.kind == 4) imgs = 2.0;
        if (f.kind == 3 || f.kind == 6) { imgs = 4.0; masks = 2.0; }
        if (f.kind >= 4) masks += 1.0;
        bytes += imgs * px * bpp + masks * px + flows;
    }
    if (pair) {
        // consecutive jobs of one pair as a group, up to FG_MAX groups a launch (the table travels in
        // the kernel's arguments)
        FrameGroups G{};
        double gbytes = 0.0;
        auto flush = [&]() {
            if (G.n == 0) return;
            const dim3 gg((mw + 255) / 256, mh, (unsigned)G.n);
            if (has[0]) TWX_LAUNCH(s, F_FRAMES, gbytes, k_frames_blend_group<0><<<gg, blk, 0, s>>>(d_jobs, G));
            if (has[1]) TWX_LAUNCH(s, F_FRAMES, gbytes, k_frames_blend_group<1><<<gg, blk, 0, s>>>(d_jobs, G));
            if (has[2]) TWX_LAUNCH(s, F_FRAMES, gbytes, k_frames_blend_group<2><<<gg, blk, 0, s>>>(d_jobs, G));
            G.n = 0;
            gbytes = 0.0;
        };
        for (size_t k = 0; k < jobs.size();) {
            size_t e = k + 1;
            while (e < jobs.size() && same_pair(jobs[k], jobs[e])) e++;
            if (G.n == FG_MAX) flush();
            G.first[G.n] = (int)k;
            G.count[G.n] = (int)(e - k);
            G.n++;
            for (size_t i = k; i < e; i++)
                gbytes += 3.0 * (double)jobs[i].W * jobs[i].H * 4.0 * sample_size(jobs[i].fmt) + 16.0 * jobs[i].fw * jobs[i].fh;
            k = e;
        }
        flush();
        (void)bytes;
        return;
    }
    if (has[0]) TWX_LAUNCH(s, F_FRAMES, bytes, k_frames<0><<<grd, blk, 0, s>>>(d_jobs));
    if (has[1]) TWX_LAUNCH(s, F_FRAMES, bytes, k_frames<1><<<grd, blk, 0, s>>>(d_jobs));
    if (has[2]) TWX_LAUNCH(s, F_FRAMES, bytes, k_frames<2><<<grd, blk, 0, s>>>(d_jobs));
}

// -------------------------------------------------------------------------------------------
// Weighted Blend on the vector path, after the first stage of k_frames (kind 3): the clean-up of
// both change masks (filters.rs:337-508), then one tiled kernel that runs the sharpening step of
// both warped frames (blur r = 3, unsharp lerp, masked composite; filters.rs:240-283, 532-656),
// the blend and the finish (compose/vector.rs:268-274).

enum MaskOp { M_PRUNE, M_DILATE, M_FILL };

// One clean-up pass over a snapshot: buffer `src` in, the other buffer out. With `clamp` the
// snapshot is read through the border replication that the clean-up runs after each fill: on a
// mask of at least 3x3 pixels mask_replicate_border (filters.rs:423-484) makes
// m(x, y) = m(clamp(x, 1, w - 2), clamp(y, 1, h - 2)).
template <int OP>
__global__ void k_mask_pass(const FrameJob* __restrict__ jobs, int src, int limit, int value, int clamp) {
    const FrameJob& F = jobs[blockIdx.z >> 1];
    if ((F.kind != 3 && F.kind != 6) || F.mask_small) return;
    const int half = blockIdx.z & 1;
    const int W = F.W, H = F.H;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W || y >= H) return;
    const uint8_t* __restrict__ t = F.mask[half][src];
    uint8_t* __restrict__ d = F.mask[half][src ^ 1];
    auto at = [&](int xx, int yy) -> int {
        if (clamp) { xx = min(max(xx, 1), W - 2); yy = min(max(yy, 1), H - 2); }
        return (int)t[(long)yy * W + xx];
    };
    int v = at(x, y);
    if (OP == M_PRUNE) {
        // mask_prune (filters.rs:352-375): the rows y - 2, y and y + 2 around a set pixel
        if (H >= 5 && x >= 1 && x <= W - 2 && y >= 2 && y <= H - 3 && v != 0) {
            int sum = at(x, y - 2) + at(x - 1, y - 2) + at(x + 1, y - 2) + at(x - 1, y) + at(x + 1, y) +
                      at(x - 1, y + 2) + at(x, y + 2) + at(x + 1, y + 2);
            if (sum < limit) v = 0;
        }
    } else if (OP == M_DILATE) {
        // mask_dilate_rows (filters.rs:403-417)
        if (y >= 1 && y <= H - 2 && (at(x, y - 1) == 255 || at(x, y + 1) == 255)) v = 255;
    } else {
        // mask_fill (filters.rs:384-399): the full ring around an empty pixel
        if (x >= 1 && x <= W - 2 && y >= 1 && y <= H - 2 && v == 0) {
            int sum = at(x, y - 1) + at(x - 1, y - 1) + at(x + 1, y - 1) + at(x - 1, y) + at(x + 1, y) +
                      at(x - 1, y + 1) + at(x, y + 1) + at(x + 1, y + 1);
            if (limit < sum) v = value;
        }
    }
    d[(long)y * W + x] = (uint8_t)v;
}

// mask_replicate_border (filters.rs:423-484) as written, rectangle tests in float, for masks with
// a side under 3 pixels where it is not the clamp of the inner ring.
__device__ void mask_border_seq(uint8_t* m, long wl, long hl) {
    int64_t w = wl, h = hl;
    float p2 = 1.0f, p3 = 1.0f, p4 = (float)(w - 2), p5 = (float)(h - 2);
    int64_t iv14 = (int64_t)(int32_t)p5;
    float fy0 = p3, fy1 = p5;
    if (p5 < p3) { fy0 = p5; fy1 = p3; }
    float fx0 = p2, fx1 = p4;
    if (p4 < p2) { fx0 = p4; fx1 = p2; }
    auto ok = [&](int64_t x, int64_t y) { return x >= 0 && x < w && y >= 0 && y < h; };
    auto outside = [&](int64_t x, int64_t y) {
        float xf = (float)x, yf = (float)y;
        return yf < fy0 || fy1 < yf || fx1 < xf || xf < fx0;
    };
    int64_t sy = (int64_t)(int32_t)p3;
    for (int64_t y = 0; (float)y < fy0; y++)
        for (int64_t x = 0; x < w; x++)
            if (outside(x, y) && ok(x, y) && ok(x, sy)) m[y * w + x] = m[sy * w + x];
    for (int64_t y = (int64_t)(int32_t)fy1; y < h; y++)
        for (int64_t x = 0; x < w; x++)
            if (outside(x, y) && ok(x, y) && ok(x, iv14)) m[y * w + x] = m[iv14 * w + x];
    int64_t sx = (int64_t)(int32_t)p2, first = (int64_t)(int32_t)fx1, base = (int64_t)(int32_t)p4;
    for (int64_t y = 0; y < h; y++) {
        for (int64_t x = 0; (float)x < fx0; x++)
            if (ok(x, y) && ok(sx, y)) m[y * w + x] = m[y * w + sx];
        for (int64_t j = 0; first + j < w; j++)
            if (ok(base + j, y) && ok(base, y)) m[y * w + base + j] = m[y * w + base];
    }
}

// The whole clean-up of a mask with a side under 3 pixels, one thread per mask: prune and fill
// need 3 columns and 3 rows and do nothing; the row growth (3 rows) and the four border
// replications run as written. The result goes to buffer 1.
__global__ void k_mask_small(const FrameJob* __restrict__ jobs) {
    const FrameJob& F = jobs[blockIdx.x >> 1];
    if ((F.kind != 3 && F.kind != 6) || !F.mask_small || threadIdx.x != 0) return;
    const int half = blockIdx.x & 1;
    const long W = F.W, H = F.H, n = W * H;
    uint8_t* m = F.mask[half][0];
    uint8_t* t = F.mask[half][1];
    if (H >= 3) {
        for (long i = 0; i < n; i++) t[i] = m[i];
        for (long y = 1; y < H - 1; y++)
            for (long x = 0; x < W; x++)
                if (t[(y - 1) * W + x] == 255 || t[(y + 1) * W + x] == 255) m[y * W + x] = 255;
    }
    for (int k = 0; k < 4; k++) mask_border_seq(m, W, H);
    for (long i = 0; i < n; i++) t[i] = m[i];
}

constexpr int WB_TW = 64, WB_TH = 16, WB_R = 3;  // tile, and the reach of the r = 3 blur (4 taps)
constexpr int WB_RW = WB_TW + 2 * WB_R, WB_RH = WB_TH + 2 * WB_R;

struct BlurTaps {
    float k[4];
};

// The store of the blur passes (filters.rs:112-119): trunc(v + 0.5), clamped, trunc again.
template <int FMT>
__device__ __forceinline__ float wb_store(float v) {
    if (FMT == 2) return v;
    const float mx = FMT == 0 ? 255.0f : 65535.0f;
    float f = (float)cvtt(v + 0.5f);
    f = f > 0.0f ? f : 0.0f;
    f = f < mx ? f : mx;
    return (float)cvtt(f + 0.5f);
}

// One output pixel of the 1D blur (filters.rs:190-236) on four lanes: `p` is the pixel in LDS,
// `stride` the float4 distance to the next pixel of the line, i its index in a line of n. Taps
// outside the image hold the sentinel -1.0; in the edge formula a lane equal to it carries no
// weight, as in the reference.
__device__ __forceinline__ float4 wb_conv(const float4* p, int stride, int i, int n, const BlurTaps& bt, bool left_first) {
    const float S = -1.0f;
    const int left = min(n, 4);
    const int mid = n > 8 ? n - 8 : 0;
    float4 c = p[0];
    float acc[4] = {c.x * bt.k[0], c.y * bt.k[0], c.z * bt.k[0], c.w * bt.k[0]};
    if (i >= left && i < left + mid) {
#pragma unroll
        for (int j = 3; j >= 1; j--) {
            float4 r = p[j * stride], l = p[-j * stride];
            acc[0] = acc[0] + (r.x + l.x) * bt.k[j];
            acc[1] = acc[1] + (r.y + l.y) * bt.k[j];
            acc[2] = acc[2] + (r.z + l.z) * bt.k[j];
            acc[3] = acc[3] + (r.w + l.w) * bt.k[j];
        }
        return make_float4(acc[0], acc[1], acc[2], acc[3]);
    }
    float wsum[4] = {bt.k[0], bt.k[0], bt.k[0], bt.k[0]};
#pragma unroll
    for (int j = 3; j >= 1; j--) {
        float4 r4 = i + j < n ? p[j * stride] : make_float4(S, S, S, S);
        float4 l4 = i >= j ? p[-j * stride] : make_float4(S, S, S, S);
        if (left_first) { float4 t = r4; r4 = l4; l4 = t; }
        const float r[4] = {r4.x, r4.y, r4.z, r4.w}, l[4] = {l4.x, l4.y, l4.z, l4.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float wr = (r[k] != S ? 1.0f : 0.0f) * bt.k[j];
            wsum[k] = wsum[k] + wr;
            float rt = wr * r[k];
            float wl = (l[k] != S ? 1.0f : 0.0f) * bt.k[j];
            float lt = l[k] * wl;
            wsum[k] = wsum[k] + wl;
            acc[k] = (acc[k] + rt) + lt;
        }
    }
    return make_float4(acc[0] / wsum[0], acc[1] / wsum[1], acc[2] / wsum[2], acc[3] / wsum[3]);
}

// A 64x16 output tile: for each warped frame, its region with a 3 pixel apron into LDS, the row
// pass over the region rows, the column pass, the unsharp lerp against the frame and the masked
// composite (inverted mask as the weight of the frame); then blend and finish.
template <int FMT>
__global__ __launch_bounds__(256) void k_weighted(const FrameJob* __restrict__ jobs, BlurTaps bt) {
    const FrameJob& F = jobs[blockIdx.z];
    if (F.fmt != FMT || (F.kind != 3 && F.kind != 6)) return;
    const int W = F.W, H = F.H;
    const int x0 = blockIdx.x * WB_TW, y0 = blockIdx.y * WB_TH;
    if (x0 >= W || y0 >= H) return;
    __shared__ float4 reg[WB_RH][WB_RW];
    __shared__ float4 hb[WB_RH][WB_TW];
    const int tid = threadIdx.x, tx = tid & 63, ty = tid >> 6;
    const int64_t imask = FMT == 0 ? 0xff : 0xffff;
    float res[2][4][4];
#pragma unroll
    for (int half = 0; half < 2; half++) {
        const void* img = half ? F.tb : F.ta;
        const float c = half ? F.cb : F.ca;
        const uint8_t* __restrict__ cm = F.mask[half][1];
        __syncthreads();
        for (int i = tid; i < WB_RH * WB_RW; i += 256) {
            int r = i / WB_RW, cc = i - r * WB_RW;
            int gy = y0 - WB_R + r, gx = x0 - WB_R + cc;
            float v[4] = {-1.0f, -1.0f, -1.0f, -1.0f};
            if (gx >= 0 && gx < W && gy >= 0 && gy < H) load4<FMT>(img, (long)gy * W + gx, v);
            reg[r][cc] = make_float4(v[0], v[1], v[2], v[3]);
        }
        __syncthreads();
        // lerp with weight 0 copies the frame (the sharpened copy is not needed)
        if (c != 0.0f) {
            for (int i = tid; i < WB_RH * WB_TW; i += 256) {
                int r = i >> 6, j = i & 63;
                int gy = y0 - WB_R + r, gx = x0 + j;
                float4 o = make_float4(-1.0f, -1.0f, -1.0f, -1.0f);
                if (gy >= 0 && gy < H && gx < W) {
                    float4 v = wb_conv(&reg[r][j + WB_R], 1, gx, W, bt, false);
                    o = make_float4(wb_store<FMT>(v.x), wb_store<FMT>(v.y), wb_store<FMT>(v.z), wb_store<FMT>(v.w));
                }
                hb[r][j] = o;
            }
            __syncthreads();
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int ly = ty + 4 * q, gx = x0 + tx, gy = y0 + ly;
            if (gx >= W || gy >= H) continue;
            float4 f4 = reg[ly + WB_R][tx + WB_R];
            float f[4] = {f4.x, f4.y, f4.z, f4.w};
            float sh[4] = {f[0], f[1], f[2], f[3]};
            if (c != 0.0f) {
                float4 v = wb_conv(&hb[ly + WB_R][tx], WB_TW, gy, H, bt, F.above_first != 0);
                sh[0] = wb_store<FMT>(v.x);
                sh[1] = wb_store<FMT>(v.y);
                sh[2] = wb_store<FMT>(v.z);
                sh[3] = wb_store<FMT>(v.w);
                if (c != 1.0f) lerp_pixel<FMT>(sh, f, c);
            }
            int mx = gx, my = gy;
            if (!F.mask_small) { mx = min(max(mx, 1), W - 2); my = min(max(my, 1), H - 2); }
            const int m = 255 - (int)cm[(long)my * W + mx];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (m == 0) { res[half][q][k] = sh[k]; continue; }
                float fw = (float)m / 255.0f;
                float g = (float)(1.0 - (double)fw);
                if (FMT == 2) res[half][q][k] = (g * sh[k] + fw * f[k]) + 0.0f;
                else res[half][q][k] = (float)(cvtt64f((sh[k] * g + f[k] * fw) + 0.5f) & imask);
            }
        }
    }
    const long n = (long)W * H;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int gx = x0 + tx, gy = y0 + ty + 4 * q;
        if (gx >= W || gy >= H) continue;
        float e[4];
        long mi = (long)gy * W + gx;
        if (F.kind == 6) {
            // Smart Blend of the sharpened frames into the image the hole fill works on
            smart_px<FMT>(res[0][q], res[1][q], F.frac, e);
            typename Px<FMT>::T* o = reinterpret_cast<typename Px<FMT>::T*>(F.comp) + mi * 4;
            for (int k = 0; k < 4; k++) o[k] = (typename Px<FMT>::T)e[k];
            continue;
        }
        blend_pixel<FMT>(res[0][q], res[1][q], F.blend_mode, F.weight, F.lerp_w, e);
        if (!F.finish) {
            typename Px<FMT>::T* o = reinterpret_cast<typename Px<FMT>::T*>(F.out) + mi * 4;
            for (int k = 0; k < 4; k++) o[k] = (typename Px<FMT>::T)e[k];
        } else {
            finish_pixel<FMT>(F.out, n, (long)(F.finish == 2 ? gy : H - 1 - gy) * W + gx, e);
        }
    }
}

static void launch_weighted_range(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs);

void launch_weighted(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs) {
    for (auto [b0, b1] : frame_buckets(jobs))
        launch_weighted_range(s, d_jobs + b0, std::vector<FrameJob>(jobs.begin() + b0, jobs.begin() + b1));
}

static void launch_weighted_range(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs) {
    int mw = 0, mh = 0, nw = 0;
    bool has[3] = {false, false, false}, small = false, big = false;
    double px = 0, img = 0;
    for (auto& f : jobs) {
        if (f.kind != 3 && f.kind != 6) continue;
        nw++;
        mw = std::max(mw, f.W);
        mh = std::max(mh, f.H);
        has[f.fmt] = true;
        (f.mask_small ? small : big) = true;
        px += (double)f.W * f.H;
        img += (double)f.W * f.H * 4 * sample_size(f.fmt);
    }
    if (!nw) return;
    const unsigned nz = (unsigned)jobs.size();
    if (big) {
        dim3 blk(256), grd((mw + 255) / 256, mh, 2 * nz);
        double b = 4.0 * px;  // both masks read and written once per pass
        TWX_LAUNCH(s, F_WB_MASK, b, k_mask_pass<M_PRUNE><<<grd, blk, 0, s>>>(d_jobs, 0, 254, 0, 0));
        TWX_LAUNCH(s, F_WB_MASK, b, k_mask_pass<M_PRUNE><<<grd, blk, 0, s>>>(d_jobs, 1, 509, 0, 0));
        TWX_LAUNCH(s, F_WB_MASK, b, k_mask_pass<M_DILATE><<<grd, blk, 0, s>>>(d_jobs, 0, 0, 0, 0));
        TWX_LAUNCH(s, F_WB_MASK, b, k_mask_pass<M_FILL><<<grd, blk, 0, s>>>(d_jobs, 1, 254, 255, 0));
        TWX_LAUNCH(s, F_WB_MASK, b, k_mask_pass<M_FILL><<<grd, blk, 0, s>>>(d_jobs, 0, 254, 255, 1));
        TWX_LAUNCH(s, F_WB_MASK, b, k_mask_pass<M_FILL><<<grd, blk, 0, s>>>(d_jobs, 1, 254, 128, 1));
        TWX_LAUNCH(s, F_WB_MASK, b, k_mask_pass<M_FILL><<<grd, blk, 0, s>>>(d_jobs, 0, 127, 64, 1));
    }
    if (small) TWX_LAUNCH(s, F_WB_MASK, 4.0 * px, k_mask_small<<<dim3(2 * nz), dim3(64), 0, s>>>(d_jobs));
    Taps g = gaussian_taps(3.0f);  // the same formula as the blur's taps (filters.rs:71-95)
    BlurTaps bt{{g.k[0], g.k[1], g.k[2], g.k[3]}};
    dim3 blk(256), grd((mw + WB_TW - 1) / WB_TW, (mh + WB_TH - 1) / WB_TH, nz);
    double b = 3.0 * img + 2.0 * px;  // both warped frames and masks read, the frame written
    if (has[0]) TWX_LAUNCH(s, F_WB_SHARPEN, b, k_weighted<0><<<grd, blk, 0, s>>>(d_jobs, bt));
    if (has[1]) TWX_LAUNCH(s, F_WB_SHARPEN, b, k_weighted<1><<<grd, blk, 0, s>>>(d_jobs, bt));
    if (has[2]) TWX_LAUNCH(s, F_WB_SHARPEN, b, k_weighted<2><<<grd, blk, 0, s>>>(d_jobs, bt));
}

// -------------------------------------------------------------------------------------------
// The push-pull hole fill (holefill.rs:214-336), batched over planes (frame x channel): the
// frame's weight plane and its four colour planes are pulled alike (blur5, resample), the colour
// planes pushed (combine, Gaussian r = 5, resample). The combine of a level is fused into the
// resample that produces it (and into the top-level init and the final store). The levels from
// a plane's `lo` up run in k_fill_coarse, in LDS; the per-level kernels handle the levels below.
// Each kernel returns at once for a frame without holes (the fill is the identity then).

// Load (holefill.rs:314-321): the channels in native units, holes zeroed; weights 0 at holes,
// else 1. One thread per pixel of an image writes level 0 of all its planes (the colour planes
// follow each other from the image's channel 0 plane, then its weight plane); the blocks of the
// other planes return at once.
__global__ void k_fill_load(const FillPlane* __restrict__ planes) {
    const FillPlane& P = planes[blockIdx.z];
    if (!*P.gate || P.chan != 0) return;
    const int W = P.w[0], H = P.h[0];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const long i = (long)y * W + x;
    const bool hole = P.holes[i] != 0;  // !(eps > m) for a u8 mask
    const int nc = P.fmt == FILL_FLOW2 ? 2 : 4;
    float v[4];
    if (P.fmt == FILL_FLOW2) {
        const float2 f = reinterpret_cast<const float2*>(P.img)[i];
        v[0] = f.x;
        v[1] = f.y;
        v[2] = v[3] = 0.0f;
    } else if (P.fmt == TWX_FMT_U8) {
        const uchar4 q = reinterpret_cast<const uchar4*>(P.img)[i];
        v[0] = (float)q.x;
        v[1] = (float)q.y;
        v[2] = (float)q.z;
        v[3] = (float)q.w;
    } else if (P.fmt == TWX_FMT_U16) {
        const ushort4 q = reinterpret_cast<const ushort4*>(P.img)[i];
        v[0] = (float)q.x;
        v[1] = (float)q.y;
        v[2] = (float)q.z;
        v[3] = (float)q.w;
    } else {
        const float4 q = reinterpret_cast<const float4*>(P.img)[i];
        v[0] = q.x;
        v[1] = q.y;
        v[2] = q.z;
        v[3] = q.w;
    }
#pragma unroll
    for (int c = 0; c < 4; c++)
        if (c < nc) planes[blockIdx.z + c].val[0][i] = hole ? 0.0f : v[c];
    planes[blockIdx.z + nc].val[0][i] = hole ? 0.0f : 1.0f;
}

// The separable pass of a fill level as LDS tiles: the pull blur5 of level `level - 1`'s values
// into t2 (push false, R = 2), or the push Gaussian (r = 5, R = 5) of level `level`'s solution
// into t1 (push true); the border rows above or below first as the plane says. Levels above the
// plane's `lo` run in k_fill_coarse.
template <int R>
__global__ __launch_bounds__(TILE_THREADS) void k_fill_tile(const FillPlane* __restrict__ planes, int level, int push, Taps t,
                                                            int tiles_x, int tiles_y, int np) {
    const int per = tiles_x * tiles_y;
    const int lt = xcd_tile(per * np);
    if (lt >= per * np) return;
    const int q = lt / per, rem = lt - q * per;
    const int tyi = rem / tiles_x, txi = rem - tyi * tiles_x;
    const FillPlane& P = planes[q];
    if (!*P.gate || level > P.lo || (push && P.chan == 4)) return;
    const int l = push ? level : level - 1;
    const int w = P.w[l], h = P.h[l];
    const int x0 = txi * TILE_X, y0 = tyi * TILE_Y;
    if (x0 >= w || y0 >= h) return;
    __shared__ float s_in[(TILE_Y + 2 * R) * (TILE_X + 2 * R)];
    const float* src = push ? P.sol[l] : P.val[l];
    float* dst = push ? P.t1 : P.t2;
    float o[TILE_RUN];
    int x, ys;
    auto load = [&](int xx, int yy) { return src[(long)yy * w + xx]; };
    const bool any = P.above_first ? sep_tile<R, true>(load, w, h, x0, y0, t, s_in, o, x, ys)
                                   : sep_tile<R, false>(load, w, h, x0, y0, t, s_in, o, x, ys);
    if (!any) return;
#pragma unroll
    for (int m = 0; m < TILE_RUN; m++) {
        if (ys + m >= h) break;
        dst[(long)(ys + m) * w + x] = o[m];
    }
}

// Pull: t2 (level - 1 blurred) resampled into val[level]. Push: t1 (level's solution after its
// Gaussian) resampled into level - 1, then the combine of that level (sol = value / weight where
// the weight is not 0).
__global__ void k_fill_resample(const FillPlane* __restrict__ planes, int level, int push) {
    const FillPlane& P = planes[blockIdx.z];
    if (!*P.gate || level > P.top || level < 1 || (push && P.chan == 4)) return;
    if (level > P.lo) return;  // inside the coarse levels: k_fill_coarse
    const int sl = push ? level : level - 1, dl = push ? level - 1 : level;
    const int sw = P.w[sl], sh = P.h[sl], dw = P.w[dl], dh = P.h[dl];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= dw || y >= dh) return;
    const float sx = ((float)sw - 1.0f) / (float)(dw - 1);
    const float sy = ((float)sh - 1.0f) / (float)(dh - 1);
    const long i = (long)y * dw + x;
    float v = bilinear1(push ? P.t1 : P.t2, sw, sh, (float)x * sx, (float)y * sy);
    if (!push) { P.val[dl][i] = v; return; }
    const float wt = P.wts[dl][i];
    if (wt != 0.0f) v = P.val[dl][i] / wt;
    P.sol[dl][i] = v;
}

// The coarse levels lo .. top of a colour plane in one workgroup, every plane in LDS: the pull of
// the values and of the frame's weights from level lo up (blur5, resample), the top level's
// solution (the combine over zeros, holefill.rs:322-326), and the push down to level lo (Gaussian r = 5, resample, combine), whose solution
// (before its Gaussian) goes back to memory; the same operations, in the same order, as the
// per-level kernels.
__global__ __launch_bounds__(1024) void k_fill_coarse(const FillPlane* __restrict__ planes, Taps t5, Taps tg) {
    const FillPlane& P = planes[blockIdx.x];
    if (!*P.gate || P.chan == 4) return;  // the same for the whole workgroup
    __shared__ float lds[FILL_LDS_FLOATS];
    const int lo = P.lo, top = P.top, tid = threadIdx.x, nt = blockDim.x;
    // level l's value, weight and solution planes follow each other from lds + base(l); the
    // offsets are computed, not kept in runtime-indexed arrays (those would go to scratch memory)
    auto base = [&](int l) {
        int b = 0;
        for (int k = lo; k < l; k++) b += 3 * P.w[k] * P.h[k];
        return b;
    };
    auto val = [&](int l) { return lds + base(l); };
    auto wts = [&](int l) { return lds + base(l) + P.w[l] * P.h[l]; };
    auto sol = [&](int l) { return lds + base(l) + 2 * P.w[l] * P.h[l]; };
    float* t1 = lds + base(top + 1);
    float* t2 = t1 + P.w[lo] * P.h[lo];
    {
        const int a = P.w[lo] * P.h[lo];
        float* v = val(lo);
        float* wt = wts(lo);
        for (int i = tid; i < a; i += nt) {
            v[i] = P.val[lo][i];
            wt[i] = P.wts[lo][i];
        }
    }
    __syncthreads();
    // pull: level k from level k - 1, for the values and the weights alike
    for (int k = lo + 1; k <= top; k++) {
        const int W = P.w[k - 1], H = P.h[k - 1], dw = P.w[k], dh = P.h[k];
        const float sx = ((float)W - 1.0f) / (float)(dw - 1);
        const float sy = ((float)H - 1.0f) / (float)(dh - 1);
        for (int pl = 0; pl < 2; pl++) {
            const float* src = pl ? wts(k - 1) : val(k - 1);
            float* dst = pl ? wts(k) : val(k);
            for (int i = tid; i < W * H; i += nt) {
                const int y = i / W, x = i - y * W;
                t1[i] = hpass_elem<1>(src + y * W, W, x, 0, t5);
            }
            __syncthreads();
            for (int i = tid; i < W * H; i += nt) {
                const int y = i / W, x = i - y * W;
                t2[i] = P.above_first ? vpass_elem<1, true>(t1, W, H, x, y, 0, t5) : vpass_elem<1, false>(t1, W, H, x, y, 0, t5);
            }
            __syncthreads();
            for (int i = tid; i < dw * dh; i += nt) {
                const int y = i / dw, x = i - y * dw;
                dst[i] = bilinear1(t2, W, H, (float)x * sx, (float)y * sy);
            }
            __syncthreads();
        }
    }
    // the top level's solution: the combine over zeros
    {
        const int a = P.w[top] * P.h[top];
        const float* v = val(top);
        const float* wt = wts(top);
        float* so = sol(top);
        for (int i = tid; i < a; i += nt) so[i] = wt[i] != 0.0f ? v[i] / wt[i] : 0.0f;
    }
    __syncthreads();
    // push: the Gaussian of level l, then its resample into level l - 1 with that level's
    // combine, down to level lo (whose Gaussian k_fill_tile runs)
    for (int l = top; l > lo; l--) {
        const int W = P.w[l], H = P.h[l];
        float* so = sol(l);
        for (int i = tid; i < W * H; i += nt) {
            const int y = i / W, x = i - y * W;
            t1[i] = hpass_elem<1>(so + y * W, W, x, 0, tg);
        }
        __syncthreads();
        for (int i = tid; i < W * H; i += nt) {
            const int y = i / W, x = i - y * W;
            so[i] = P.above_first ? vpass_elem<1, true>(t1, W, H, x, y, 0, tg) : vpass_elem<1, false>(t1, W, H, x, y, 0, tg);
        }
        __syncthreads();
        const int dw = P.w[l - 1], dh = P.h[l - 1];
        const float sx = ((float)W - 1.0f) / (float)(dw - 1);
        const float sy = ((float)H - 1.0f) / (float)(dh - 1);
        const float* dv = val(l - 1);
        const float* dwt = wts(l - 1);
        float* dso = sol(l - 1);
        for (int i = tid; i < dw * dh; i += nt) {
            const int y = i / dw, x = i - y * dw;
            float v = bilinear1(so, W, H, (float)x * sx, (float)y * sy);
            const float wt = dwt[i];
            if (wt != 0.0f) v = dv[i] / wt;
            dso[i] = v;
        }
        __syncthreads();
    }
    const int a = P.w[lo] * P.h[lo];
    const float* so = sol(lo);
    for (int i = tid; i < a; i += nt) P.sol[lo][i] = so[i];
}

// The last combine of level 0 and the write back (holefill.rs:330-333, 366-386), one thread per
// pixel of an image for all its colour planes (the whole pixel is written); the blocks of the
// other planes return at once.
__global__ void k_fill_store(const FillPlane* __restrict__ planes) {
    const FillPlane& P = planes[blockIdx.z];
    if (!*P.gate || P.chan != 0) return;
    const int W = P.w[0], H = P.h[0];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W || y >= H) return;
    const long i = (long)y * W + x;
    const int nc = P.fmt == FILL_FLOW2 ? 2 : 4;
    const float wt = P.wts[0][i];  // the image's weight plane, shared by its colour planes
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (c >= nc) {
            v[c] = 0.0f;
            continue;
        }
        const FillPlane& Q = planes[blockIdx.z + c];
        v[c] = wt != 0.0f ? Q.val[0][i] / wt : Q.t1[i];  // t1: level 0's solution after its Gaussian
    }
    if (P.fmt == FILL_FLOW2) {
        reinterpret_cast<float2*>(P.img)[i] = make_float2(v[0], v[1]);
        return;
    }
    if (P.fmt == TWX_FMT_F32) {
        reinterpret_cast<float4*>(P.img)[i] = make_float4(v[0], v[1], v[2], v[3]);
        return;
    }
    const float white = P.fmt == TWX_FMT_U8 ? 255.0f : 65535.0f;
    const int64_t wi = P.fmt == TWX_FMT_U8 ? 255 : 65535;
    uint32_t o[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        // clamped to [0, white]; a NaN sample truncates to INT64_MIN, whose low bits are 0
        const float yv = v[c] + 0.5f;
        const int64_t q = 0.0f > yv ? 0 : (yv > white ? wi : cvtt64f(yv));
        o[c] = (uint32_t)(q & wi);
    }
    if (P.fmt == TWX_FMT_U8) reinterpret_cast<uchar4*>(P.img)[i] = make_uchar4(o[0], o[1], o[2], o[3]);
    else reinterpret_cast<ushort4*>(P.img)[i] = make_ushort4(o[0], o[1], o[2], o[3]);
}

static void launch_fill_range(hipStream_t s, const FillPlane* d_planes, const std::vector<FillPlane>& planes);

void launch_fill(hipStream_t s, const FillPlane* d_planes, const std::vector<FillPlane>& planes) {
    std::vector<double> area;
    for (auto& p : planes) area.push_back((double)p.w[0] * p.h[0]);
    for (auto [b0, b1] : size_buckets(area))
        launch_fill_range(s, d_planes + b0, std::vector<FillPlane>(planes.begin() + b0, planes.begin() + b1));
}

static void launch_fill_range(hipStream_t s, const FillPlane* d_planes, const std::vector<FillPlane>& planes) {
    if (planes.empty()) return;
    const unsigned np = (unsigned)planes.size();
    int maxtop = 0;
    for (auto& p : planes) maxtop = std::max(maxtop, p.top);
    // the largest level sizes over the planes, and the pixels of each level over the planes
    std::vector<int> mw(maxtop + 1, 0), mh(maxtop + 1, 0);
    std::vector<double> px(maxtop + 1, 0.0);
    for (auto& p : planes)
        for (int l = 0; l <= p.top; l++) {
            mw[l] = std::max(mw[l], p.w[l]);
            mh[l] = std::max(mh[l], p.h[l]);
            px[l] += (double)p.w[l] * p.h[l];
        }
    auto grid = [&](int l) { return dim3((mw[l] + 255) / 256, mh[l], np); };
    const dim3 blk(256);
    const Taps t5 = binomial_taps(), tg = gaussian_taps(5.0f);
    // levels above `hi` run only inside k_fill_coarse (every plane has 1 <= lo <= top: its 2x2
    // top level always fits in LDS); the per-level kernels stop there
    int hi = 0;
    double lo_px = 0.0;
    for (auto& p : planes) {
        hi = std::max(hi, p.lo);
        if (p.chan != 4) lo_px += (double)p.w[p.lo] * p.h[p.lo];
    }
    // the tiles of level l over the planes, in XCD order
    auto tiles = [&](int l, int& tx, int& ty) {
        tx = (mw[l] + TILE_X - 1) / TILE_X;
        ty = (mh[l] + TILE_Y - 1) / TILE_Y;
        const long n = (long)tx * ty * np;
        return dim3((unsigned)(8 * ((n + 7) / 8)));
    };
    int tx, ty;
    TWX_LAUNCH(s, F_FILL_IO, 8.0 * px[0], k_fill_load<<<grid(0), blk, 0, s>>>(d_planes));
    for (int k = 1; k <= hi; k++) {
        const dim3 tg = tiles(k - 1, tx, ty);
        TWX_LAUNCH(s, F_FILL_PULL, 8.0 * px[k - 1], k_fill_tile<2><<<tg, TILE_THREADS, 0, s>>>(d_planes, k, 0, t5, tx, ty, (int)np));
        TWX_LAUNCH(s, F_FILL_PULL, 4.0 * (px[k - 1] + px[k]), k_fill_resample<<<grid(k), blk, 0, s>>>(d_planes, k, 0));
    }
    // the coarse kernel reads level lo's values and weights and writes its solution
    TWX_LAUNCH(s, F_FILL_PUSH, 12.0 * lo_px, k_fill_coarse<<<dim3(np), dim3(1024), 0, s>>>(d_planes, t5, tg));
    for (int l = hi; l >= 0; l--) {
        const dim3 gt = tiles(l, tx, ty);
        TWX_LAUNCH(s, F_FILL_PUSH, 8.0 * px[l], k_fill_tile<5><<<gt, TILE_THREADS, 0, s>>>(d_planes, l, 1, tg, tx, ty, (int)np));
        if (l > 0)
            TWX_LAUNCH(s, F_FILL_PUSH, 4.0 * px[l] + 16.0 * px[l - 1],
                       k_fill_resample<<<grid(l - 1), blk, 0, s>>>(d_planes, l, 1));
    }
    TWX_LAUNCH(s, F_FILL_IO, 16.0 * px[0], k_fill_store<<<grid(0), blk, 0, s>>>(d_planes));
}

// The Smart Blend frames after the fill: the filled engine image into the wire frame, rows turned
// back (image.rs:205-245).
template <int FMT>
__global__ void k_finish(const FrameJob* __restrict__ jobs) {
    const FrameJob& F = jobs[blockIdx.z];
    if (F.fmt != FMT || F.kind < 4 || !F.finish) return;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= F.W || y >= F.H) return;
    const long mi = (long)y * F.W + x;
    float e[4];
    load4<FMT>(F.comp, mi, e);
    finish_pixel<FMT>(F.out, (long)F.W * F.H, (long)(F.finish == 2 ? y : F.H - 1 - y) * F.W + x, e);
}

static void launch_finish_range(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs);

void launch_finish(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs) {
    for (auto [b0, b1] : frame_buckets(jobs))
        launch_finish_range(s, d_jobs + b0, std::vector<FrameJob>(jobs.begin() + b0, jobs.begin() + b1));
}

static void launch_finish_range(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs) {
    int mw = 0, mh = 0;
    bool has[3] = {false, false, false};
    double bytes = 0;
    for (auto& f : jobs) {
        if (f.kind < 4 || !f.finish) continue;
        mw = std::max(mw, f.W);
        mh = std::max(mh, f.H);
        has[f.fmt] = true;
        bytes += 2.0 * f.W * f.H * 4 * sample_size(f.fmt);
    }
    if (!mw) return;
    dim3 blk(256), grd((mw + 255) / 256, mh, (unsigned)jobs.size());
    if (has[0]) TWX_LAUNCH(s, F_FILL_IO, bytes, k_finish<0><<<grd, blk, 0, s>>>(d_jobs));
    if (has[1]) TWX_LAUNCH(s, F_FILL_IO, bytes, k_finish<1><<<grd, blk, 0, s>>>(d_jobs));
    if (has[2]) TWX_LAUNCH(s, F_FILL_IO, bytes, k_finish<2><<<grd, blk, 0, s>>>(d_jobs));
}

// -------------------------------------------------------------------------------------------
// Vector image coding (vectors.rs:57-130, :474-549) and the flow steps of Vectors In
// (compose/vector.rs:335-397)

// Flow::bilerp (warp.rs:415-442)
__device__ __forceinline__ void flow_bilerp(const float* __restrict__ flow, int fw, int fh, float x, float y, float d[2]) {
    int32_t jx = cvtt(x), jy = cvtt(y);
    float gx = x - (float)jx, gy = y - (float)jy;
    if (jx >= fw - 1) { jx = fw - 2; gx = 1.0f; }
    if (jy >= fh - 1) { jy = fh - 2; gy = 1.0f; }
    const long f00 = ((long)jy * fw + jx) * 2, f01 = f00 + (long)fw * 2;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float p00 = flow[f00 + k], p10 = flow[f00 + 2 + k], p01 = flow[f01 + k], p11 = flow[f01 + 2 + k];
        const float top = (p10 - p00) * gx + p00;
        const float bot = (p11 - p01) * gx + p01;
        d[k] = (bot - top) * gy + top;
    }
}

__device__ __forceinline__ void clamp_unit(float& v) {  // vectors.rs:132-141 (a NaN falls through)
    if (v > 1.0f) v = 1.0f;
    else if (-1.0f > v) v = -1.0f;
}

template <int FMT>
__global__ void k_vec_encode(const VecJob* __restrict__ jobs) {
    const VecJob& J = jobs[blockIdx.z];
    if (J.fmt != FMT) return;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= J.W || y >= J.H) return;
    float d[2];
    flow_bilerp(J.flow, J.fw, J.fh, (float)x * J.sx, (float)y * J.sy, d);
    float a = d[0] * J.kx, b = d[1] * J.ky;
    float e[4];
    if (FMT == 2) {
        e[0] = 1.0f;
        e[1] = (a + 1.0f) * 0.5f;
        e[2] = (b + 1.0f) * 0.5f;
        e[3] = 0.0f;
    } else {
        clamp_unit(a);
        clamp_unit(b);
        a = a + 1.0f;
        b = b + 1.0f;
        const double q = FMT == 0 ? 127.0 : 32767.0;
        const uint32_t mask = FMT == 0 ? 0xffu : 0xffffu;  // the low byte or word of the 64-bit truncation
        e[0] = FMT == 0 ? 255.0f : 65535.0f;
        e[1] = (float)((uint32_t)cvtt64d((double)a * q + 0.5) & mask);
        e[2] = (float)((uint32_t)cvtt64d((double)b * q + 0.5) & mask);
        e[3] = 0.0f;
    }
    const long n = (long)J.W * J.H;
    if (J.finish) {
        finish_pixel<FMT>(J.out, n, (long)(J.H - 1 - y) * J.W + x, e);
        return;
    }
    typename Px<FMT>::T* o = reinterpret_cast<typename Px<FMT>::T*>(J.out) + ((long)y * J.W + x) * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = (typename Px<FMT>::T)e[k];
}

void vec_job_scales(VecJob& J, float scale) {
    const float dw1 = (float)(J.W - 1), dh1 = (float)(J.H - 1), fw1 = (float)(J.fw - 1), fh1 = (float)(J.fh - 1);
    J.sx = fw1 / dw1;
    J.sy = fh1 / dh1;
    J.kx = (dw1 / fw1) / scale;
    J.ky = (dh1 / fh1) / scale;
}

void launch_vec_encode(hipStream_t s, const VecJob* d_jobs, const std::vector<VecJob>& jobs) {
    int mw = 0, mh = 0;
    bool has[3] = {false, false, false};
    double bytes = 0;
    for (auto& j : jobs) {
        mw = std::max(mw, j.W);
        mh = std::max(mh, j.H);
        has[j.fmt] = true;
        bytes += (double)j.fw * j.fh * 8 + (double)image_bytes(j.W, j.H, j.fmt);
    }
    if (!mw) return;
    dim3 blk(256), grd((mw + 255) / 256, mh, (unsigned)jobs.size());
    if (has[0]) TWX_LAUNCH(s, F_VEC, bytes, k_vec_encode<0><<<grd, blk, 0, s>>>(d_jobs));
    if (has[1]) TWX_LAUNCH(s, F_VEC, bytes, k_vec_encode<1><<<grd, blk, 0, s>>>(d_jobs));
    if (has[2]) TWX_LAUNCH(s, F_VEC, bytes, k_vec_encode<2><<<grd, blk, 0, s>>>(d_jobs));
}

template <int FMT>
__global__ void k_vec_decode(const void* __restrict__ vec, long n, float md, float* __restrict__ flow,
                             uint8_t* __restrict__ holes, int* gate) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // the hole flag of the inverted alpha: u8 any non-zero, u16 and f32 at least 3.1e-5 or unordered
    const double wide_eps = __longlong_as_double(0x3f0040bfe3b03e21ll);
    float o0, o1;
    bool hole;
    if (FMT == 2) {
        const float* p = reinterpret_cast<const float*>(vec) + i * 4;
        if (p[0] == 0.0f) { o0 = o1 = 0.0f; }
        else {
            o0 = (p[1] * 2.0f - 1.0f) * md;
            o1 = (p[2] * 2.0f - 1.0f) * md;
        }
        hole = !(wide_eps > (double)(1.0f - p[0]));
    } else {
        using T = typename Px<FMT>::T;
        const T* p = reinterpret_cast<const T*>(vec) + i * 4;
        const double den = FMT == 0 ? 254.0 : 65534.0;
        const double mdd = (double)md;
        if (p[0] == 0) { o0 = o1 = 0.0f; }
        else {
            o0 = (float)((((double)p[1] * 2.0) / den - 1.0) * mdd);
            o1 = (float)((((double)p[2] * 2.0) / den - 1.0) * mdd);
        }
        hole = FMT == 0 ? p[0] != 0xff : !(wide_eps > (double)(uint16_t)(0xffff - p[0]));
    }
    flow[i * 2] = o0;
    flow[i * 2 + 1] = o1;
    holes[i] = hole ? 1 : 0;
    if (hole) *gate = 1;  // the same value from every writer
}

void launch_vec_decode(hipStream_t s, const void* vec, int W, int H, int fmt, float max_displace, float* flow,
                       uint8_t* holes, int* gate) {
    const long n = (long)W * H;
    const unsigned g = (unsigned)((n + 255) / 256);
    const double bytes = (double)image_bytes(W, H, fmt) + 9.0 * n;
    if (fmt == 0) TWX_LAUNCH(s, F_VEC, bytes, k_vec_decode<0><<<g, 256, 0, s>>>(vec, n, max_displace, flow, holes, gate));
    else if (fmt == 1) TWX_LAUNCH(s, F_VEC, bytes, k_vec_decode<1><<<g, 256, 0, s>>>(vec, n, max_displace, flow, holes, gate));
    else TWX_LAUNCH(s, F_VEC, bytes, k_vec_decode<2><<<g, 256, 0, s>>>(vec, n, max_displace, flow, holes, gate));
}

__global__ void k_vec_scale(float* __restrict__ flow, long n, float sx, float sy) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flow[i * 2] = sx * flow[i * 2];
    flow[i * 2 + 1] = sy * flow[i * 2 + 1];
}

void launch_vec_scale(hipStream_t s, float* flow, long n, float sx, float sy) {
    TWX_LAUNCH(s, F_VEC, 16.0 * n, k_vec_scale<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(flow, n, sx, sy));
}

__global__ void k_invert_prep(const float* __restrict__ flow, float* __restrict__ t, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    reinterpret_cast<float4*>(t)[i] = make_float4(-1.0f * 1.0f, -1.0f * flow[i * 2], -1.0f * flow[i * 2 + 1], -1.0f * 1.0f);
}

__global__ void k_invert_post(const float* __restrict__ pushed, const uint8_t* __restrict__ cov, float* __restrict__ flow,
                              uint8_t* __restrict__ holes, int* gate, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = reinterpret_cast<const float4*>(pushed)[i];
    flow[i * 2] = p.y;
    flow[i * 2 + 1] = p.z;
    const uint8_t hole = (uint8_t)(255 - cov[i]);
    holes[i] = hole;
    if (hole) *gate = 1;
}

void launch_invert_prep(hipStream_t s, const float* flow, float* t, long n) {
    TWX_LAUNCH(s, F_VEC, 24.0 * n, k_invert_prep<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(flow, t, n));
}

void launch_invert_post(hipStream_t s, const float* pushed, const uint8_t* cov, float* flow, uint8_t* holes, int* gate,
                        long n) {
    TWX_LAUNCH(s, F_VEC, 26.0 * n, k_invert_post<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pushed, cov, flow, holes, gate, n));
}

__global__ void k_zero_ints(int* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void launch_zero(hipStream_t s, int* p, size_t n) {
    if (n) k_zero_ints<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(p, n);
}

// -------------------------------------------------------------------------------------------
// Taps (conv.rs:44-86), on the host

Taps gaussian_taps(float r) {
    Taps t{};
    int n = std::max((int32_t)r + 1, 0);
    float s3 = r / 3.0f;
    float d = s3 * s3;
    float sum = 0.0f;
    for (int i = 0; i < n; i++) {
        float fi = (float)i;
        float e = -((fi * fi) / d);
        float v = (float)std::pow(2.0, (double)e);
        t.k[i] = v;
        sum += v;
    }
    t.n = n;
    if (n == 0) return t;
    float den = sum * 2.0f - t.k[0];
    for (int i = 0; i < n; i++) t.k[i] /= den;
    return t;
}

Taps binomial_taps() {
    Taps t{};
    t.k[0] = 0.375f;
    t.k[1] = 0.25f;
    t.k[2] = 0.0625f;
    t.n = 3;
    return t;
}

}  // namespace twx
