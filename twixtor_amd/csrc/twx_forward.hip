// The Gouraud forward warp (plugins/retime-engine/src/forward.rs:229-373, forward/mesh.rs:11-52,
// forward/raster.rs:59-300): every source pixel is a vertex of a mesh displaced by the flow, the two
// triangles of each cell are scan converted in cell order, and every covered pixel folds the
// interpolated vertex colours into a running mean in triangle order.
//
// The running mean depends on the order of the hits of a pixel, so the device keeps that order
// explicitly instead of relying on scheduling:
//   mesh   one thread per vertex: the displaced positions (with the tie breaking epsilon);
//   count  one thread per triangle: the scan converter run for positions only, hits counted per
//          pixel (atomic increments: the counts do not depend on the order);
//   scan   exclusive prefix sum of the counts (64-bit offsets);
//   emit   one thread per triangle: the scan converter with the colour payload, each hit stored
//          as (triangle index, colour) in its pixel's segment (slot order within a segment is
//          arbitrary);
//   fold   one thread per pixel: its hits taken in ascending triangle index, the running mean,
//          the conversion to the engine format and the 255/0 coverage.
// Everything follows the CPU's single precision operation order (no contraction; the double
// precision epsilon, the span and last-row thresholds and the running mean's reciprocal in double).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>

#include "twx_common.hpp"
#include "twx_pixel.hpp"

namespace twx {

namespace {

__device__ __forceinline__ int32_t fw_cvtt(float x) {
    if (isnan(x) || x >= 2147483648.0f || x < -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}

__device__ __forceinline__ int64_t fw_cvtt64(float x) {
    if (isnan(x) || x >= 9223372036854775808.0f || x < -9223372036854775808.0f) return INT64_MIN;
    return (int64_t)x;
}

// A vertex: its displaced position and its pixel, whose colour is read where an edge or the
// last-row pass needs it (keeping the colours of all three vertices live took registers enough
// to halve the emit pass's occupancy).
struct FwV {
    float x, y;
    uint32_t idx;
};

struct FwE {
    float ay, by, x, dx;
    float p[4], dp[4];
};

// raster.rs:59-86: activates the edge a -> b for a.y <= y < b.y, state set for scanline y - 1,
// into e1 when an edge is active already, else into e0. The edge pair lives in two named
// structs (no runtime-indexed private arrays, which would go to scratch memory).
template <bool PAY, class Tex>
__device__ __forceinline__ void fw_edge_init(const FwV& a, const FwV& b, float y, FwE& e0, FwE& e1, uint32_t& count,
                                             const Tex& tex) {
    if (a.y <= y && y < b.y) {
        const bool second = count != 0;
        count++;
        const float dx = b.x - a.x;
        const float inv = 1.0f / (b.y - a.y);
        const float t = ((y - 1.0f) - a.y) * inv;
        FwE E = second ? e1 : e0;
        E.ay = a.y;
        E.by = b.y;
        E.x = dx * t + a.x;
        E.dx = dx * inv;
        if (PAY) {
            float pa[4], pb[4];
            tex(a.idx, pa);
            tex(b.idx, pb);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float dk = pb[k] - pa[k];
                E.p[k] = dk * t + pa[k];
                E.dp[k] = dk * inv;
            }
        }
        if (second) e1 = E;
        else e0 = E;
    }
}

// one of three vertices by a runtime index, field by field (selects, no indexed private array)
__device__ __forceinline__ FwV fw_pick(const FwV& v0, const FwV& v1, const FwV& v2, int a) {
    FwV o;
    o.x = a == 0 ? v0.x : (a == 1 ? v1.x : v2.x);
    o.y = a == 0 ? v0.y : (a == 1 ? v1.y : v2.y);
    o.idx = a == 0 ? v0.idx : (a == 1 ? v1.idx : v2.idx);
    return o;
}

// raster.rs:88-112: ties keep the given order except v2.y == v0.y < v1.y -> v2, v0, v1
__device__ __forceinline__ void fw_sort_by_y(const FwV v[3], FwV o[3]) {
    const float y0 = v[0].y, y1 = v[1].y, y2 = v[2].y;
    int a, b, c;
    if (y1 < y0) {
        if (y2 < y0) {
            if (y2 < y1) { a = 2; b = 1; c = 0; } else { a = 1; b = 2; c = 0; }
        } else { a = 1; b = 0; c = 2; }
    } else if (y2 < y1) {
        if (y0 < y2) { a = 0; b = 2; c = 1; } else { a = 2; b = 0; c = 1; }
    } else { a = 0; b = 1; c = 2; }
    o[0] = fw_pick(v[0], v[1], v[2], a);
    o[1] = fw_pick(v[0], v[1], v[2], b);
    o[2] = fw_pick(v[0], v[1], v[2], c);
}

// raster.rs:153-214: on reaching vertex K, the edge pair pruned of the edges that ended, then
// K's two edges (0: 1, 2), (1: 2, 0), (2: 0, 1) activated
template <bool PAY, int K, class Tex>
__device__ __forceinline__ void fw_vertex_edges(const FwV vs[3], float y, FwE& e0, FwE& e1, uint32_t& count, const Tex& tex) {
    if (count == 1) {
        if (e0.ay <= y && e0.by <= y) count = 0;
    } else if (count == 2) {
        if (y < e1.ay || y < e1.by) {
            if (!(y < e0.ay || y < e0.by)) {
                e0 = e1;
                count = 1;
            }
        } else if (e0.ay <= y && e0.by <= y) {
            count = 0;
        } else {
            count = 1;
        }
    }
    constexpr int N0 = K == 0 ? 1 : (K == 1 ? 2 : 0);
    constexpr int N1 = K == 0 ? 2 : (K == 1 ? 0 : 1);
    fw_edge_init<PAY>(vs[K], vs[N0], y, e0, e1, count, tex);
    fw_edge_init<PAY>(vs[K], vs[N1], y, e0, e1, count, tex);
}

// raster.rs:114-151: one span of the main pass. `hit(x, y, p)` returns false to stop the triangle.
template <bool PAY, class Hit>
__device__ __forceinline__ bool fw_span(const FwE& l, const FwE& r, float y, int32_t w1, Hit& hit) {
    const int32_t x0 = max(fw_cvtt(ceilf(l.x)), 0);
    const int32_t x1 = min(fw_cvtt(floorf(r.x)), w1);
    const float wid = r.x - l.x;
    if (x0 > x1) return true;
    if (!((double)wid > 1e-08)) return true;
    const int32_t yi = fw_cvtt(y);
    float p[4] = {0.0f, 0.0f, 0.0f, 0.0f}, dp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (x0 == x1) {
        const float t = ((float)x0 - l.x) * (1.0f / wid);
        if (PAY) {
#pragma unroll
            for (int k = 0; k < 4; k++) p[k] = (r.p[k] - l.p[k]) * t + l.p[k];
        }
        return hit(x0, yi, p);
    }
    const float inv = 1.0f / wid;
    const float t = ((float)x0 - l.x) * inv;
    if (PAY) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            dp[k] = r.p[k] - l.p[k];
            p[k] = dp[k] * t + l.p[k];
        }
    }
    int32_t x = x0;
    while (x < x1) {
        if (!hit(x, yi, p)) return false;
        if (PAY) {
#pragma unroll
            for (int k = 0; k < 4; k++) p[k] = p[k] + dp[k] * inv;
        }
        x++;
    }
    return hit(x, yi, p);
}

// raster.rs:153-300. Returns 0, or 1 when the last row pass would write outside the image, or 2
// when `hit` stopped the triangle.
template <bool PAY, class Hit, class Tex>
__device__ int fw_triangle(const FwV tri[3], int32_t w1, int32_t h1, Hit& hit, const Tex& tex) {
    const float w1f = (float)w1, h1f = (float)h1;
    const bool ok = (0.0f <= tri[0].x || 0.0f <= tri[1].x || 0.0f <= tri[2].x) &&
                    (0.0f <= tri[0].y || 0.0f <= tri[1].y || 0.0f <= tri[2].y) &&
                    (tri[0].x <= w1f || tri[1].x <= w1f || tri[2].x <= w1f) &&
                    (tri[0].y <= h1f || tri[1].y <= h1f || tri[2].y <= h1f);
    if (!ok) return 0;
    FwV vs[3];
    fw_sort_by_y(tri, vs);
    float y = ceilf(vs[0].y), y_end = floorf(vs[2].y);
    if (y < 0.0f) y = 0.0f;
    if (h1f < y_end) y_end = h1f;
    if (y <= y_end) {
        int vi = 0;  // the next vertex to reach; vertices are reached in order, several per row
        uint32_t count = 0;
        FwE e0, e1;
        e0.ay = e0.by = e0.x = e0.dx = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; k++) e0.p[k] = e0.dp[k] = 0.0f;
        e1 = e0;
        for (;;) {
            if (vi == 0 && !(y < vs[0].y)) {
                fw_vertex_edges<PAY, 0>(vs, y, e0, e1, count, tex);
                vi = 1;
            }
            if (vi == 1 && !(y < vs[1].y)) {
                fw_vertex_edges<PAY, 1>(vs, y, e0, e1, count, tex);
                vi = 2;
            }
            if (vi == 2 && !(y < vs[2].y)) {
                fw_vertex_edges<PAY, 2>(vs, y, e0, e1, count, tex);
                vi = 3;
            }
            e0.x += e0.dx;
            e1.x += e1.dx;
            if (PAY) {
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    e0.p[k] = e0.p[k] + e0.dp[k];
                    e1.p[k] = e1.p[k] + e1.dp[k];
                }
            }
            if (count == 2) {
                // the left and right edge by value (a select per field)
                const bool swap = e1.x < e0.x;
                FwE l, r;
                l.x = swap ? e1.x : e0.x;
                r.x = swap ? e0.x : e1.x;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    l.p[k] = swap ? e1.p[k] : e0.p[k];
                    r.p[k] = swap ? e0.p[k] : e1.p[k];
                }
                if (!fw_span<PAY>(l, r, y, w1, hit)) return 2;
            }
            y += 1.0f;
            if (!(y <= y_end)) break;
        }
    }
    // the degenerate pass: both lower vertices exactly on the last row
    const FwV& mid = vs[1];
    const FwV& bot = vs[2];
    int status = 0;
    if (mid.y == h1f && mid.y == bot.y) {
        const bool bl = bot.x <= mid.x;
        const FwV l = fw_pick(bot, mid, mid, bl ? 0 : 1);
        const FwV r = fw_pick(mid, bot, bot, bl ? 0 : 1);
        const float yy = l.y, yt = (float)fw_cvtt(yy);
        if (0.0f <= yt && yt <= h1f && !((double)(yy - yt) > 0.01)) {
            const int32_t x0 = max(fw_cvtt(ceilf(l.x)), 0);
            const int32_t x1 = min(fw_cvtt(floorf(r.x)), w1);
            const float wid = r.x - l.x;
            if ((double)wid > 1e-05) {
                const float inv = 1.0f / wid;
                const int32_t yi = fw_cvtt(yt);
                float p[4] = {0.0f, 0.0f, 0.0f, 0.0f}, dp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                float lp[4], rp[4];
                if (PAY) {
                    tex(l.idx, lp);
                    tex(r.idx, rp);
                }
                const float tt = ((float)x0 - l.x) * inv;
                if (x0 == x1) {
                    if (PAY) {
#pragma unroll
                        for (int k = 0; k < 4; k++) p[k] = (rp[k] - lp[k]) * tt + lp[k];
                    }
                    if (x0 < 0 || x0 > w1 || yi < 0 || yi > h1) status = 1;
                    else if (!hit(x0, yi, p)) return 2;
                } else {
                    if (PAY) {
#pragma unroll
                        for (int k = 0; k < 4; k++) {
                            dp[k] = rp[k] - lp[k];
                            p[k] = dp[k] * tt + lp[k];
                        }
                    }
                    int32_t x = x0;
                    for (;;) {
                        // `while x < x1` then one more emit (a sub-pixel span still writes x0)
                        if (x < 0 || x > w1 || yi < 0 || yi > h1) status = 1;
                        else if (!hit(x, yi, p)) return 2;
                        if (!(x < x1)) break;
                        if (PAY) {
#pragma unroll
                            for (int k = 0; k < 4; k++) p[k] = p[k] + dp[k] * inv;
                        }
                        x++;
                    }
                }
            }
        }
    }
    return status;
}

template <int FMT> struct FwPx;
template <> struct FwPx<0> { using T = uint8_t; };
template <> struct FwPx<1> { using T = uint16_t; };
template <> struct FwPx<2> { using T = float; };

template <int FMT>
__device__ __forceinline__ void fw_texel(const void* img, long i, float o[4]) {
    const typename FwPx<FMT>::T* p = reinterpret_cast<const typename FwPx<FMT>::T*>(img) + i * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = (float)p[k];
}

// The vertices of triangle t: cell (r, c) = t / 2, T1 = (v00, v01, v10), T2 = (v11, v01, v10).
__device__ __forceinline__ void fw_vertices(const FwArgs& A, uint32_t t, FwV tri[3]) {
    const uint32_t cell = t >> 1, W1 = (uint32_t)A.W - 1;
    const uint32_t r = cell / W1, c = cell - r * W1;
    const uint32_t i00 = r * (uint32_t)A.W + c, i01 = i00 + 1, i10 = i00 + (uint32_t)A.W, i11 = i10 + 1;
    const uint32_t idx[3] = {(t & 1) ? i11 : i00, i01, i10};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const float2 p = A.pos[idx[q]];
        tri[q].x = p.x;
        tri[q].y = p.y;
        tri[q].idx = idx[q];
    }
}

// Flow::bilerp (warp.rs:415-442) of the warp's flow at (qx, qy) inside [0, fw - 1] x [0, fh - 1]
__device__ __forceinline__ void fw_flow_at(const FwArgs& A, float qx, float qy, float d[2]) {
    int32_t jx = fw_cvtt(qx), jy = fw_cvtt(qy);
    float gx = qx - (float)jx, gy = qy - (float)jy;
    if (jx >= A.fw - 1) { jx = A.fw - 2; gx = 1.0f; }
    if (jy >= A.fh - 1) { jy = A.fh - 2; gy = 1.0f; }
    const long f00 = ((long)jy * A.fw + jx) * 2, f01 = f00 + (long)A.fw * 2;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float p00 = A.flow[f00 + k], p10 = A.flow[f00 + 2 + k], p01 = A.flow[f01 + k], p11 = A.flow[f01 + 2 + k];
        const float top = (p10 - p00) * gx + p00;
        const float bot = (p11 - p01) * gx + p01;
        d[k] = (bot - top) * gy + top;
    }
}

// mesh.rs:33-52: x' = fx * sx + c, y' = fy * sy + r, the flow sampled bilinearly at (c * rx,
// r * ry); rows 1.. and columns 1.. add 0.0001 in double
__global__ void k_fw_mesh(FwArgs A) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c >= A.W || r >= A.H) return;
    const float cf = (float)c, rf = (float)r;
    float d[2];
    fw_flow_at(A, cf * A.rx, rf * A.ry, d);
    float px = d[0] * A.sx + cf, py = d[1] * A.sy + rf;
    if (r != 0 && c >= 1) {
        px = (float)((double)px + 0.0001);
        py = (float)((double)py + 0.0001);
    }
    const long i = (long)r * A.W + c;
    A.pos[i] = make_float2(px, py);
    if (A.fmt == TWX_FMT_F32 && !A.textured) {
        // a NaN colour at any vertex (every pixel is a vertex of some triangle when W, H >= 2); the
        // textured variant checks the colours it samples instead
        float col[4];
        fw_texel<2>(A.src, i, col);
        if (isnan(col[0]) || isnan(col[1]) || isnan(col[2]) || isnan(col[3])) atomicOr(A.flags, 4);
    }
}

// mesh.rs:54-88 up to the logarithm: the argument m of vertex (c, r)'s mip level, from the
// distances of its 8 displaced neighbours (indices mirrored at the borders) to it: the diagonals
// summed in single precision, widened and scaled by 1/sqrt(2), the axis neighbours added in double,
// narrowed, times 0.125. The host takes d = -(logf(m) * (1 / logf(2))) + 0 with its C runtime's
// logf, whose last bits a device logarithm would not reproduce.
__global__ void k_fw_mip_arg(FwArgs A, float* __restrict__ m) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c >= A.W || r >= A.H) return;
    const int32_t w = A.W, h = A.H;
    const int32_t up = r - 1 < 0 ? 1 : r - 1, down = h - 1 < r + 1 ? h - 2 : r + 1;
    const int32_t left = c - 1 < 0 ? 1 : c - 1, right = w - 1 < c + 1 ? w - 2 : c + 1;
    float f[2];
    fw_flow_at(A, (float)c * A.rx, (float)r * A.ry, f);
    const float xc = f[0] * A.sx + (float)c, yc = f[1] * A.sy + (float)r;
    auto len = [&](int32_t ci, int32_t ri) {
        float n[2];
        fw_flow_at(A, (float)ci * A.rx, (float)ri * A.ry, n);
        const float dx = (n[0] * A.sx + (float)ci) - xc, dy = (n[1] * A.sy + (float)ri) - yc;
        return sqrtf(dy * dy + dx * dx);
    };
    const float ru = len(right, up), lu = len(left, up), ld = len(left, down), rd = len(right, down);
    const float u = len(c, up), dn = len(c, down), l = len(left, r), rr = len(right, r);
    const float diag = ((ru + lu) + ld) + rd;
    double t = (double)diag * 0.7071067811865476;
    t += (double)u;
    t += (double)dn;
    t += (double)l;
    t += (double)rr;
    m[(long)r * w + c] = (float)((double)(float)t * 0.125);
}

// Pixel i of a four-channel image of format FMT as floats.
template <int FMT>
__device__ __forceinline__ float4 fw_ld4(const void* img, long i) {
    float o[4];
    fw_texel<FMT>(img, i, o);
    return make_float4(o[0], o[1], o[2], o[3]);
}

// The store of the pyramid's blur passes (filters.rs:112-119): trunc(v + 0.5), clamped, trunc
// again for the integer formats; floats as they are.
template <int FMT>
__device__ __forceinline__ float fwp_store(float v) {
    if (FMT == 2) return v;
    const float mx = FMT == 0 ? 255.0f : 65535.0f;
    float f = (float)fw_cvtt(v + 0.5f);
    f = f > 0.0f ? f : 0.0f;
    f = f < mx ? f : mx;
    return (float)fw_cvtt(f + 0.5f);
}

// One output pixel of the pyramid's blur (pyramid.rs:126-177, the 1D filter of filters.rs with the
// 5-tap binomial half kernel [0.375, 0.25, 0.0625]): `q[j]` for j = 0..4 is the line's pixel
// i - 2 + j (only those inside the line are read); edge pixels renormalise over the taps inside
// the line, a lane equal to the sentinel -1.0 carrying no weight.
__device__ __forceinline__ float4 fwp_conv(const float4 q[5], int i, int n) {
    const float S = -1.0f, k[3] = {0.375f, 0.25f, 0.0625f};
    const int left = min(n, 3), mid = n > 6 ? n - 6 : 0;
    const float4 c = q[2];
    float acc[4] = {c.x * k[0], c.y * k[0], c.z * k[0], c.w * k[0]};
    if (i >= left && i < left + mid) {
#pragma unroll
        for (int j = 2; j >= 1; j--) {
            const float4 r = q[2 + j], l = q[2 - j];
            acc[0] = acc[0] + (r.x + l.x) * k[j];
            acc[1] = acc[1] + (r.y + l.y) * k[j];
            acc[2] = acc[2] + (r.z + l.z) * k[j];
            acc[3] = acc[3] + (r.w + l.w) * k[j];
        }
        return make_float4(acc[0], acc[1], acc[2], acc[3]);
    }
    float wsum[4] = {k[0], k[0], k[0], k[0]};
#pragma unroll
    for (int j = 2; j >= 1; j--) {
        const float4 r4 = i + j < n ? q[2 + j] : make_float4(S, S, S, S);
        const float4 l4 = i >= j ? q[2 - j] : make_float4(S, S, S, S);
        const float r[4] = {r4.x, r4.y, r4.z, r4.w}, l[4] = {l4.x, l4.y, l4.z, l4.w};
#pragma unroll
        for (int a = 0; a < 4; a++) {
            const float wr = (r[a] != S ? 1.0f : 0.0f) * k[j];
            wsum[a] = wsum[a] + wr;
            const float rt = wr * r[a];
            const float wl = (l[a] != S ? 1.0f : 0.0f) * k[j];
            const float lt = l[a] * wl;
            wsum[a] = wsum[a] + wl;
            acc[a] = (acc[a] + rt) + lt;
        }
    }
    return make_float4(acc[0] / wsum[0], acc[1] / wsum[1], acc[2] / wsum[2], acc[3] / wsum[3]);
}

// The row pass of a level's blur: the previous level (format FMT) into `out` (the stored values
// as floats).
template <int FMT>
__global__ void k_fwp_rows(const void* __restrict__ src, float4* __restrict__ out, int w, int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    float4 q[5];
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const int xx = min(max(x - 2 + j, 0), w - 1);  // outside the line: never read by fwp_conv
        q[j] = fw_ld4<FMT>(src, (long)y * w + xx);
    }
    const float4 v = fwp_conv(q, x, w);
    out[(long)y * w + x] = make_float4(fwp_store<FMT>(v.x), fwp_store<FMT>(v.y), fwp_store<FMT>(v.z), fwp_store<FMT>(v.w));
}

// The column pass, on the row pass's result as a whole (pyramid.rs:185-194).
template <int FMT>
__global__ void k_fwp_cols(const float4* __restrict__ src, float4* __restrict__ out, int w, int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w || y >= h) return;
    float4 q[5];
#pragma unroll
    for (int j = 0; j < 5; j++) {
        const int yy = min(max(y - 2 + j, 0), h - 1);
        q[j] = src[(long)yy * w + x];
    }
    const float4 v = fwp_conv(q, y, h);
    out[(long)y * w + x] = make_float4(fwp_store<FMT>(v.x), fwp_store<FMT>(v.y), fwp_store<FMT>(v.z), fwp_store<FMT>(v.w));
}

// bilerp4 (warp.rs:306-331) of four floats per pixel: truncating conversion, the upper clamp
// (ix >= w1 -> w1 - 1, fx = 1), the three lerps, and with a NaN among the results the same lerps
// by the NaN rule. `bad` is set where the CPU would index below the image (a coordinate below 0
// that does not truncate to 0, or NaN); the result is then zeros.
template <class Load>
__device__ __forceinline__ void fw_bilerp4(const Load& ld, int w, int w1, int h1, float x, float y, float o[4], bool& bad) {
    int32_t ix = fw_cvtt(x), iy = fw_cvtt(y);
    float fx = x - (float)ix, fy = y - (float)iy;
    if (ix >= w1) { ix = w1 - 1; fx = 1.0f; }
    if (iy >= h1) { iy = h1 - 1; fy = 1.0f; }
    if (ix < 0 || iy < 0) {
        bad = true;
        o[0] = o[1] = o[2] = o[3] = 0.0f;
        return;
    }
    const long i00 = (long)iy * w + ix, i01 = i00 + w;
    const float4 a = ld(i00), b = ld(i00 + 1), c = ld(i01), d = ld(i01 + 1);
    const float pa[4] = {a.x, a.y, a.z, a.w}, pb[4] = {b.x, b.y, b.z, b.w};
    const float pc[4] = {c.x, c.y, c.z, c.w}, pd[4] = {d.x, d.y, d.z, d.w};
    bool nan = false;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float top = (pb[k] - pa[k]) * fx + pa[k];
        const float bot = (pd[k] - pc[k]) * fx + pc[k];
        o[k] = (bot - top) * fy + top;
        nan = nan || isnan(o[k]);
    }
    if (nan) {
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = nr_lerp_bf(nr_lerp_bf(pa[k], pb[k], fx), nr_lerp_bf(pc[k], pd[k], fx), fy);
    }
}

// pyramid.rs resample_plane: level (dw, dh) from the blurred previous level, corner aligned,
// through bilerp4 and the warp's store (Pix::store)
template <int FMT>
__global__ void k_fwp_resample(const float4* __restrict__ src, int sw, int sh, void* __restrict__ dst, int dw, int dh,
                               float sx, float sy) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= dw || y >= dh) return;
    float v[4];
    bool bad = false;
    fw_bilerp4([&](long i) { return src[i]; }, sw, sw - 1, sh - 1, (float)x * sx, (float)y * sy, v, bad);
    typename FwPx<FMT>::T* o = reinterpret_cast<typename FwPx<FMT>::T*>(dst) + ((long)y * dw + x) * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = (typename FwPx<FMT>::T)store_rule<FMT>(v[k]);
}

// pyramid.rs:97-110: one bilinear tap of level i at (w1 * u, h1 * v)
template <int FMT>
__device__ __forceinline__ void fw_tap(const FwArgs& A, int i, float u, float v, float o[4], bool& bad) {
    const int w1 = A.lw[i] - 1, h1 = A.lh[i] - 1;
    const void* img = A.lev[i];
    fw_bilerp4([&](long j) { return fw_ld4<FMT>(img, j); }, A.lw[i], w1, h1, (float)w1 * u, (float)h1 * v, o, bad);
}

// pyramid.rs:68-92 and forward.rs:130-141: the colour of a textured hit at (u, v) of mip level d
template <int FMT>
__device__ __forceinline__ void fw_sample(const FwArgs& A, float u, float v, float d, float o[4], bool& bad) {
    if (0.0f < d) {
        if (d < (float)(A.nlev - 1)) {
            const int32_t i = fw_cvtt(d);
            const float frac = d - (float)i;
            float a[4], b[4];
            fw_tap<FMT>(A, i, u, v, a, bad);
            fw_tap<FMT>(A, i + 1, u, v, b, bad);
#pragma unroll
            for (int k = 0; k < 4; k++) o[k] = (b[k] - a[k]) * frac + a[k];
            return;
        }
        fw_tap<FMT>(A, A.nlev - 1, u, v, o, bad);
        return;
    }
    fw_tap<FMT>(A, 0, u, v, o, bad);
}

// Per-triangle work cap: a triangle that would cover more pixels stops and the warp is refused.
constexpr uint32_t FW_TRI_CAP = 1u << 16;

// flags: 1 the last row pass would write outside, 2 a triangle over FW_TRI_CAP hits, 4 a NaN
// colour (set by k_fw_mesh), 8 two hits of one triangle on one pixel, 16 the emit pass disagreed
// with the count pass
__global__ void k_fw_count(FwArgs A) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.ntri) return;
    FwV tri[3];
    fw_vertices(A, (uint32_t)t, tri);
    uint32_t n = 0;
    auto hit = [&](int32_t x, int32_t y, const float*) -> bool {
        if (++n > FW_TRI_CAP) return false;
        atomicAdd(&A.counts[(long)y * A.W + x], 1u);
        return true;
    };
    auto tex = [](uint32_t, float*) {};
    const int st = fw_triangle<false>(tri, A.W - 1, A.H - 1, hit, tex);
    if (st == 1) atomicOr(A.flags, 1);
    if (st == 2) atomicOr(A.flags, 2);
}

// TEX: the textured variant, whose vertices carry (u, v, d) = (c / (w - 1), r / (h - 1), the
// vertex's mip level) (forward.rs:271-289) and whose hits sample the pyramid there; else the
// Gouraud variant, whose vertices carry their source pixel's colour.
template <int FMT, bool TEX>
__global__ void k_fw_emit(FwArgs A) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.ntri) return;
    FwV tri[3];
    fw_vertices(A, (uint32_t)t, tri);
    const float w1f = (float)(A.W - 1), h1f = (float)(A.H - 1);
    auto tex = [&](uint32_t i, float o[4]) {
        if (TEX) {
            const uint32_t r = i / (uint32_t)A.W, c = i - r * (uint32_t)A.W;
            o[0] = (float)(int32_t)c / w1f;
            o[1] = (float)(int32_t)r / h1f;
            o[2] = A.mips ? A.mips[i] : 0.0f;
            o[3] = 0.0f;
        } else {
            fw_texel<FMT>(A.src, i, o);
        }
    };
    bool bad = false, nanc = false;
    auto hit = [&](int32_t x, int32_t y, const float* p) -> bool {
        const long i = (long)y * A.W + x;
        const uint32_t k = atomicAdd(&A.fill[i], 1u);
        if (k >= A.counts[i]) {  // the count pass saw fewer hits: never write past the segment
            atomicOr(A.flags, 16);
            return false;
        }
        const uint64_t slot = A.offsets[i] + k;
        A.rec_t[slot] = (uint32_t)t;
        if (TEX) {
            float c[4];
            fw_sample<FMT>(A, p[0], p[1], p[2], c, bad);
            nanc = nanc || isnan(c[0]) || isnan(c[1]) || isnan(c[2]) || isnan(c[3]);
            A.rec_c[slot] = make_float4(c[0], c[1], c[2], c[3]);
        } else {
            A.rec_c[slot] = make_float4(p[0], p[1], p[2], p[3]);
        }
        return true;
    };
    fw_triangle<true>(tri, A.W - 1, A.H - 1, hit, tex);
    if (TEX && (bad || nanc)) atomicOr(A.flags, (bad ? 32 : 0) | (nanc ? 64 : 0));
}

// forward.rs:104-128
__device__ __forceinline__ void fw_fold(float acc[4], uint32_t& n, const float4 c) {
    const float cc[4] = {c.x, c.y, c.z, c.w};
    if (n == 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] = cc[k] + 0.0f;
        n = 1;
    } else if (n == 0xff) {
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] = (acc[k] * 255.0f + cc[k]) * 0.00390625f + 0.0f;
    } else {
        const float nf = (float)n;
        const float r = (float)(1.0 / ((double)nf + 1.0));
#pragma unroll
        for (int k = 0; k < 4; k++) acc[k] = (nf * acc[k] + cc[k]) * r + 0.0f;
        n++;
    }
}

// In-place heap sort of a pixel's records by triangle index (long segments).
__device__ void fw_heap_sort(uint32_t* t, float4* c, uint64_t n) {
    auto sift = [&](uint64_t root, uint64_t end) {
        for (;;) {
            uint64_t child = 2 * root + 1;
            if (child >= end) return;
            if (child + 1 < end && t[child] < t[child + 1]) child++;
            if (!(t[root] < t[child])) return;
            const uint32_t tt = t[root];
            t[root] = t[child];
            t[child] = tt;
            const float4 cc = c[root];
            c[root] = c[child];
            c[child] = cc;
            root = child;
        }
    };
    for (uint64_t s = n / 2; s-- > 0;) sift(s, n);
    for (uint64_t e = n; e-- > 1;) {
        const uint32_t tt = t[0];
        t[0] = t[e];
        t[e] = tt;
        const float4 cc = c[0];
        c[0] = c[e];
        c[e] = cc;
        sift(0, e);
    }
}

// forward.rs:139-172 conversion, and the coverage
template <int FMT>
__global__ void k_fw_fold(FwArgs A) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= A.W || y >= A.H) return;
    const long i = (long)y * A.W + x;
    const uint32_t n = A.counts[i];
    const uint64_t base = A.offsets[i];
    if (A.fill[i] != n) atomicOr(A.flags, 16);
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t cnt = 0;
    if (n <= 64) {
        // short segments: the hits in ascending triangle index, each found by a scan of the
        // segment (a triangle hits a pixel at most once; a repeated index is reported)
        bool dup = false;
        for (uint32_t j = 1; j < n; j++)
            for (uint32_t q = 0; q < j; q++) dup |= A.rec_t[base + q] == A.rec_t[base + j];
        if (dup) {
            atomicOr(A.flags, 8);
        } else {
            uint32_t last = 0;
            for (uint32_t k = 0; k < n; k++) {
                uint32_t best = 0xffffffffu, at = 0;
                for (uint32_t j = 0; j < n; j++) {
                    const uint32_t tj = A.rec_t[base + j];
                    if ((k == 0 || tj > last) && tj < best) {
                        best = tj;
                        at = j;
                    }
                }
                fw_fold(acc, cnt, A.rec_c[base + at]);
                last = best;
            }
        }
    } else {
        fw_heap_sort(A.rec_t + base, A.rec_c + base, n);
        for (uint32_t k = 0; k < n; k++) {
            if (k > 0 && A.rec_t[base + k] == A.rec_t[base + k - 1]) atomicOr(A.flags, 8);
            fw_fold(acc, cnt, A.rec_c[base + k]);
        }
    }
    typename FwPx<FMT>::T* o = reinterpret_cast<typename FwPx<FMT>::T*>(A.out) + i * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if constexpr (FMT == 2) o[k] = acc[k] + 0.0f;
        else o[k] = (typename FwPx<FMT>::T)fw_cvtt64(acc[k] + 0.5f);  // the low byte or word, no saturation
    }
    if (A.cov) A.cov[i] = n != 0 ? 0xff : 0;
}

// Exclusive prefix sum of the per-pixel counts into 64-bit offsets: blocks of FW_SCAN_ITEMS
// counts, the block totals scanned by one block, then added back.
constexpr int FW_SCAN_THREADS = 256;
constexpr int FW_SCAN_PER_THREAD = 8;
constexpr int FW_SCAN_ITEMS = FW_SCAN_THREADS * FW_SCAN_PER_THREAD;

__device__ uint64_t fw_block_exclusive(uint64_t v, uint64_t* sh, uint64_t& total) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < FW_SCAN_THREADS; off <<= 1) {
        const uint64_t add = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    total = sh[FW_SCAN_THREADS - 1];
    const uint64_t incl = sh[tid];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(FW_SCAN_THREADS) k_fw_scan_blocks(const uint32_t* __restrict__ counts, uint64_t* __restrict__ offsets,
                                                                    uint64_t* __restrict__ block_totals, long n) {
    __shared__ uint64_t sh[FW_SCAN_THREADS];
    const long base = (long)blockIdx.x * FW_SCAN_ITEMS + (long)threadIdx.x * FW_SCAN_PER_THREAD;
    // a lane's eight counts as two 16-byte loads and its eight offsets as four 16-byte stores where
    // the run lies inside the array (both arrays start 256-byte aligned), element by element at the end
    const bool whole = base + FW_SCAN_PER_THREAD <= n;
    uint64_t v[FW_SCAN_PER_THREAD], s = 0;
    if (whole) {
        const uint4 c0 = *reinterpret_cast<const uint4*>(counts + base);
        const uint4 c1 = *reinterpret_cast<const uint4*>(counts + base + 4);
        v[0] = c0.x; v[1] = c0.y; v[2] = c0.z; v[3] = c0.w;
        v[4] = c1.x; v[5] = c1.y; v[6] = c1.z; v[7] = c1.w;
    } else {
#pragma unroll
        for (int k = 0; k < FW_SCAN_PER_THREAD; k++) v[k] = base + k < n ? counts[base + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < FW_SCAN_PER_THREAD; k++) s += v[k];
    uint64_t total;
    uint64_t run = fw_block_exclusive(s, sh, total);
    uint64_t o[FW_SCAN_PER_THREAD];
#pragma unroll
    for (int k = 0; k < FW_SCAN_PER_THREAD; k++) {
        o[k] = run;
        run += v[k];
    }
    if (whole) {
#pragma unroll
        for (int k = 0; k < FW_SCAN_PER_THREAD; k += 2)
            *reinterpret_cast<ulonglong2*>(offsets + base + k) = make_ulonglong2(o[k], o[k + 1]);
    } else {
#pragma unroll
        for (int k = 0; k < FW_SCAN_PER_THREAD; k++)
            if (base + k < n) offsets[base + k] = o[k];
    }
    if (threadIdx.x == 0) block_totals[blockIdx.x] = total;
}

// One block: the exclusive scan of the block totals in place (nb <= FW_SCAN_THREADS * 64), and
// the grand total into total[0].
__global__ void __launch_bounds__(FW_SCAN_THREADS) k_fw_scan_totals(uint64_t* __restrict__ block_totals, int nb,
                                                                    uint64_t* __restrict__ total_out) {
    __shared__ uint64_t sh[FW_SCAN_THREADS];
    const int per = (nb + FW_SCAN_THREADS - 1) / FW_SCAN_THREADS;
    const int base = threadIdx.x * per;
    uint64_t s = 0;
    for (int k = 0; k < per; k++)
        if (base + k < nb) s += block_totals[base + k];
    uint64_t total;
    uint64_t run = fw_block_exclusive(s, sh, total);
    for (int k = 0; k < per; k++) {
        if (base + k < nb) {
            const uint64_t v = block_totals[base + k];
            block_totals[base + k] = run;
            run += v;
        }
    }
    if (threadIdx.x == 0) total_out[0] = total;
}

__global__ void k_fw_scan_add(uint64_t* __restrict__ offsets, const uint64_t* __restrict__ block_totals, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) offsets[i] += block_totals[i / FW_SCAN_ITEMS];
}

}  // namespace

namespace {

// forward.rs:183-195 on a device flow: sets *flag when a component moves by more than 1e-4 (a
// NaN does not: it fails the comparison, as on the CPU)
// Four 16-byte loads in flight per lane, the last n % 4 components one by one; one atomic per
// block that saw a moving component (a few hundred blocks: every wave's atomic on the one word cost
// about 12 ns each, serialised).
__device__ __forceinline__ bool fw_moves(float v) {
    const float a = v < 0.0f ? -v : v;
    return 0.0001 < (double)a;
}
__global__ void k_fw_flow_moves(const float* __restrict__ f, long n, int* flag) {
    bool moves = false;
    // a flow that is not 16-byte aligned takes the element loop below for all of it
    const long n4 = (reinterpret_cast<uintptr_t>(f) & 15) ? 0 : n / 4, stride = (long)gridDim.x * blockDim.x;
    const float4* f4 = reinterpret_cast<const float4*>(f);
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += 4 * stride) {
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = i + u * stride < n4 ? f4[i + u * stride] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int u = 0; u < 4; u++) moves |= fw_moves(v[u].x) || fw_moves(v[u].y) || fw_moves(v[u].z) || fw_moves(v[u].w);
    }
    for (long j = n4 * 4 + (long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) moves |= fw_moves(f[j]);
    if (__syncthreads_or(moves) && threadIdx.x == 0) atomicOr(flag, 1);
}

}  // namespace

void launch_fw_flow_moves(hipStream_t s, const float* flow, long n, int* flag) {
    const unsigned blocks = (unsigned)std::max<long>(1, std::min<long>((n / 4 + 4095) / 4096, 256));
    TWX_LAUNCH(s, F_FW_MESH, 4.0 * (double)n, k_fw_flow_moves<<<blocks, 1024, 0, s>>>(flow, n, flag));
}

size_t fw_scan_blocks(long n) { return (size_t)((n + FW_SCAN_ITEMS - 1) / FW_SCAN_ITEMS); }

bool fw_flow_is_zero(const float* flow, long n) {
    for (long i = 0; i < n; i++) {
        const float v = flow[i];
        const float a = v < 0.0f ? -v : v;
        if (0.0001 < (double)a) return false;
    }
    return true;
}

void fw_grid(FwArgs& A, float amount) {
    A.rx = (float)(A.fw - 1) / (float)(A.W - 1);
    A.ry = (float)(A.fh - 1) / (float)(A.H - 1);
    A.sx = amount / A.rx;
    A.sy = amount / A.ry;
    A.ntri = 2ull * (uint64_t)(A.W - 1) * (uint64_t)(A.H - 1);
}

void launch_fw_count(hipStream_t s, const FwArgs& A) {
    const double px = (double)A.W * A.H, ss = (double)sample_size(A.fmt);
    dim3 blk(256);
    const double nan_scan = A.fmt == TWX_FMT_F32 ? 4.0 * ss * px : 0.0;
    TWX_LAUNCH(s, F_FW_MESH, 8.0 * A.fw * A.fh + 8.0 * px + nan_scan, k_fw_mesh<<<dim3((A.W + 255) / 256, A.H), blk, 0, s>>>(A));
    const dim3 tg((unsigned)((A.ntri + 255) / 256));
    TWX_LAUNCH(s, F_FW_RASTER, 8.0 * px + 4.0 * px, k_fw_count<<<tg, blk, 0, s>>>(A));
    const long n = (long)A.W * A.H;
    const int nb = (int)fw_scan_blocks(n);
    TWX_LAUNCH(s, F_FW_SCAN, 12.0 * px, k_fw_scan_blocks<<<nb, FW_SCAN_THREADS, 0, s>>>(A.counts, A.offsets, A.block_totals, n));
    TWX_LAUNCH(s, F_FW_SCAN, 16.0 * nb, k_fw_scan_totals<<<1, FW_SCAN_THREADS, 0, s>>>(A.block_totals, nb, A.total));
    TWX_LAUNCH(s, F_FW_SCAN, 16.0 * px, k_fw_scan_add<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(A.offsets, A.block_totals, n));
}

void launch_fw_emit_fold(hipStream_t s, const FwArgs& A, uint64_t records) {
    const double px = (double)A.W * A.H, ss = (double)sample_size(A.fmt);
    dim3 blk(256);
    const dim3 tg((unsigned)((A.ntri + 255) / 256));
    // algorithmic bytes of the warp: the source read, the positions, the records written and read
    // back, the output and coverage written
    const double emit_bytes = 8.0 * px + 4.0 * ss * px + 20.0 * (double)records;
    const double fold_bytes = 12.0 * px + 20.0 * (double)records + 4.0 * ss * px + (A.cov ? px : 0.0);
    const bool tx = A.textured != 0;
    if (A.fmt == TWX_FMT_U8) {
        if (tx) TWX_LAUNCH(s, F_FW_RASTER, emit_bytes, (k_fw_emit<0, true><<<tg, blk, 0, s>>>(A)));
        else TWX_LAUNCH(s, F_FW_RASTER, emit_bytes, (k_fw_emit<0, false><<<tg, blk, 0, s>>>(A)));
        TWX_LAUNCH(s, F_FW_FOLD, fold_bytes, k_fw_fold<0><<<dim3((A.W + 255) / 256, A.H), blk, 0, s>>>(A));
    } else if (A.fmt == TWX_FMT_U16) {
        if (tx) TWX_LAUNCH(s, F_FW_RASTER, emit_bytes, (k_fw_emit<1, true><<<tg, blk, 0, s>>>(A)));
        else TWX_LAUNCH(s, F_FW_RASTER, emit_bytes, (k_fw_emit<1, false><<<tg, blk, 0, s>>>(A)));
        TWX_LAUNCH(s, F_FW_FOLD, fold_bytes, k_fw_fold<1><<<dim3((A.W + 255) / 256, A.H), blk, 0, s>>>(A));
    } else {
        if (tx) TWX_LAUNCH(s, F_FW_RASTER, emit_bytes, (k_fw_emit<2, true><<<tg, blk, 0, s>>>(A)));
        else TWX_LAUNCH(s, F_FW_RASTER, emit_bytes, (k_fw_emit<2, false><<<tg, blk, 0, s>>>(A)));
        TWX_LAUNCH(s, F_FW_FOLD, fold_bytes, k_fw_fold<2><<<dim3((A.W + 255) / 256, A.H), blk, 0, s>>>(A));
    }
}

int fw_level_count(int w, int h) {
    int n = 1;
    int32_t a = std::max(w, 2), b = std::max(h, 2);
    for (;;) {
        a /= 2;
        if (a < 2) break;
        b /= 2;
        if (b < 2) break;
        n++;
    }
    return n;
}

void launch_fw_mip_arg(hipStream_t s, const FwArgs& A, float* m) {
    const double px = (double)A.W * A.H;
    TWX_LAUNCH(s, F_FW_MESH, 8.0 * A.fw * A.fh + 4.0 * px, k_fw_mip_arg<<<dim3((A.W + 255) / 256, A.H), dim3(256), 0, s>>>(A, m));
}

// pyramid.rs:37-63: level i from level i - 1, blurred (rows, then columns on the row result) and
// resampled to the level's size; `scratch` holds two float4 planes of the source's size
void launch_fw_pyramid(hipStream_t s, FwArgs& A, float* scratch) {
    float4* t0 = reinterpret_cast<float4*>(scratch);
    float4* t1 = t0 + (size_t)A.lw[0] * A.lh[0];
    const dim3 blk(256);
    for (int i = 1; i < A.nlev; i++) {
        const int pw = A.lw[i - 1], ph = A.lh[i - 1], dw = A.lw[i], dh = A.lh[i];
        const dim3 g((pw + 255) / 256, ph), gd((dw + 255) / 256, dh);
        const float sx = ((float)pw - 1.0f) / (float)(dw - 1), sy = ((float)ph - 1.0f) / (float)(dh - 1);
        const double bytes = (double)pw * ph * (4.0 * sample_size(A.fmt) + 16.0 * 3.0) + (double)dw * dh * 4.0 * sample_size(A.fmt);
        if (A.fmt == TWX_FMT_U8) {
            TWX_LAUNCH(s, F_FW_MESH, bytes, (k_fwp_rows<0><<<g, blk, 0, s>>>(A.lev[i - 1], t0, pw, ph)));
            TWX_LAUNCH(s, F_FW_MESH, 0.0, (k_fwp_cols<0><<<g, blk, 0, s>>>(t0, t1, pw, ph)));
            TWX_LAUNCH(s, F_FW_MESH, 0.0, (k_fwp_resample<0><<<gd, blk, 0, s>>>(t1, pw, ph, (void*)A.lev[i], dw, dh, sx, sy)));
        } else if (A.fmt == TWX_FMT_U16) {
            TWX_LAUNCH(s, F_FW_MESH, bytes, (k_fwp_rows<1><<<g, blk, 0, s>>>(A.lev[i - 1], t0, pw, ph)));
            TWX_LAUNCH(s, F_FW_MESH, 0.0, (k_fwp_cols<1><<<g, blk, 0, s>>>(t0, t1, pw, ph)));
            TWX_LAUNCH(s, F_FW_MESH, 0.0, (k_fwp_resample<1><<<gd, blk, 0, s>>>(t1, pw, ph, (void*)A.lev[i], dw, dh, sx, sy)));
        } else {
            TWX_LAUNCH(s, F_FW_MESH, bytes, (k_fwp_rows<2><<<g, blk, 0, s>>>(A.lev[i - 1], t0, pw, ph)));
            TWX_LAUNCH(s, F_FW_MESH, 0.0, (k_fwp_cols<2><<<g, blk, 0, s>>>(t0, t1, pw, ph)));
            TWX_LAUNCH(s, F_FW_MESH, 0.0, (k_fwp_resample<2><<<gd, blk, 0, s>>>(t1, pw, ph, (void*)A.lev[i], dw, dh, sx, sy)));
        }
    }
}

}  // namespace twx
