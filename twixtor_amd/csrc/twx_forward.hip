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

// mesh.rs:33-52: x' = fx * sx + c, y' = fy * sy + r, the flow sampled bilinearly at (c * rx,
// r * ry) (Flow::bilerp, warp.rs:415-442); rows 1.. and columns 1.. add 0.0001 in double
__global__ void k_fw_mesh(FwArgs A) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y;
    if (c >= A.W || r >= A.H) return;
    const float cf = (float)c, rf = (float)r;
    const float qx = cf * A.rx, qy = rf * A.ry;
    int32_t jx = fw_cvtt(qx), jy = fw_cvtt(qy);
    float gx = qx - (float)jx, gy = qy - (float)jy;
    if (jx >= A.fw - 1) { jx = A.fw - 2; gx = 1.0f; }
    if (jy >= A.fh - 1) { jy = A.fh - 2; gy = 1.0f; }
    const long f00 = ((long)jy * A.fw + jx) * 2, f01 = f00 + (long)A.fw * 2;
    float d[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float p00 = A.flow[f00 + k], p10 = A.flow[f00 + 2 + k], p01 = A.flow[f01 + k], p11 = A.flow[f01 + 2 + k];
        const float top = (p10 - p00) * gx + p00;
        const float bot = (p11 - p01) * gx + p01;
        d[k] = (bot - top) * gy + top;
    }
    float px = d[0] * A.sx + cf, py = d[1] * A.sy + rf;
    if (r != 0 && c >= 1) {
        px = (float)((double)px + 0.0001);
        py = (float)((double)py + 0.0001);
    }
    const long i = (long)r * A.W + c;
    A.pos[i] = make_float2(px, py);
    if (A.fmt == TWX_FMT_F32) {
        // a NaN colour at any vertex (every pixel is a vertex of some triangle when W, H >= 2)
        float col[4];
        fw_texel<2>(A.src, i, col);
        if (isnan(col[0]) || isnan(col[1]) || isnan(col[2]) || isnan(col[3])) atomicOr(A.flags, 4);
    }
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

template <int FMT>
__global__ void k_fw_emit(FwArgs A) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.ntri) return;
    FwV tri[3];
    fw_vertices(A, (uint32_t)t, tri);
    auto tex = [&](uint32_t i, float o[4]) { fw_texel<FMT>(A.src, i, o); };
    auto hit = [&](int32_t x, int32_t y, const float* p) -> bool {
        const long i = (long)y * A.W + x;
        const uint32_t k = atomicAdd(&A.fill[i], 1u);
        if (k >= A.counts[i]) {  // the count pass saw fewer hits: never write past the segment
            atomicOr(A.flags, 16);
            return false;
        }
        const uint64_t slot = A.offsets[i] + k;
        A.rec_t[slot] = (uint32_t)t;
        A.rec_c[slot] = make_float4(p[0], p[1], p[2], p[3]);
        return true;
    };
    fw_triangle<true>(tri, A.W - 1, A.H - 1, hit, tex);
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
    uint64_t v[FW_SCAN_PER_THREAD], s = 0;
#pragma unroll
    for (int k = 0; k < FW_SCAN_PER_THREAD; k++) {
        v[k] = base + k < n ? counts[base + k] : 0;
        s += v[k];
    }
    uint64_t total;
    uint64_t run = fw_block_exclusive(s, sh, total);
#pragma unroll
    for (int k = 0; k < FW_SCAN_PER_THREAD; k++) {
        if (base + k < n) offsets[base + k] = run;
        run += v[k];
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
__global__ void k_fw_flow_moves(const float* __restrict__ f, long n, int* flag) {
    bool moves = false;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float v = f[i];
        const float a = v < 0.0f ? -v : v;
        moves |= 0.0001 < (double)a;
    }
    if (__any(moves) && (threadIdx.x & 63) == 0) atomicOr(flag, 1);
}

}  // namespace

void launch_fw_flow_moves(hipStream_t s, const float* flow, long n, int* flag) {
    const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 2048);
    TWX_LAUNCH(s, F_FW_MESH, 4.0 * (double)n, k_fw_flow_moves<<<blocks, 256, 0, s>>>(flow, n, flag));
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
    if (A.fmt == TWX_FMT_U8) {
        TWX_LAUNCH(s, F_FW_RASTER, emit_bytes, k_fw_emit<0><<<tg, blk, 0, s>>>(A));
        TWX_LAUNCH(s, F_FW_FOLD, fold_bytes, k_fw_fold<0><<<dim3((A.W + 255) / 256, A.H), blk, 0, s>>>(A));
    } else if (A.fmt == TWX_FMT_U16) {
        TWX_LAUNCH(s, F_FW_RASTER, emit_bytes, k_fw_emit<1><<<tg, blk, 0, s>>>(A));
        TWX_LAUNCH(s, F_FW_FOLD, fold_bytes, k_fw_fold<1><<<dim3((A.W + 255) / 256, A.H), blk, 0, s>>>(A));
    } else {
        TWX_LAUNCH(s, F_FW_RASTER, emit_bytes, k_fw_emit<2><<<tg, blk, 0, s>>>(A));
        TWX_LAUNCH(s, F_FW_FOLD, fold_bytes, k_fw_fold<2><<<dim3((A.W + 255) / 256, A.H), blk, 0, s>>>(A));
    }
}

}  // namespace twx
