// cpuref -- CPU restatement of the reference retime engine's CPU path (the ORACLE).
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the HIP backend is compared with; only the
// tests, the smoke check and the CPU-baseline leg of the benchmark load it. The product library
// (twixtor_amd/libtwx.so) never links, loads or calls it.
//
// It restates, operation for operation, the Rust CPU path of baairon/twixtor:
//   plugins/retime-engine/src/flow/conv.rs     separable convolution, border renormalisation
//   plugins/retime-engine/src/flow/sample.rs   cvtt, cell, bilinear samplers, resample
//   plugins/retime-engine/src/prep.rs          geometry, luma, reduce
//   plugins/retime-engine/src/flow.rs          pyramid, solver kernels, block matching, driver
//   plugins/retime-engine/src/warp.rs          inverse warp (hold edges), store rules
//   plugins/retime-engine/src/blend.rs         blend, lerp, copy
//   plugins/retime-engine/src/image.rs         premultiply / unpremultiply, wire conversions
//   crates/fx-core/src/frame.rs                wire byte layouts, 16 bpc range mapping
//   plugins/retime-engine/src/timeline.rs      rates, Speed time map, selection
//   crates/fx-core/src/curves.rs, fixed.rs     keyframe sampling, 16.16 fixed point
//   plugins/retime-engine/src/compose.rs, compose/vector.rs   fast path and vector path order
//
// Numerics: single precision in the order written, every operation rounded on its own
// (compiled with -ffp-contract=off, no fast-math), the double precision steps where the
// reference widens, and x86 truncation semantics (NaN and out-of-range give INT_MIN).
// Rows run in parallel with OpenMP; no float sum crosses rows, so the bytes do not depend on
// the thread count.
//
// Parity status per function is recorded in the test files that pin it (tests/test_oracle_*.py).

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../include/twx.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace cref {

// ------------------------------------------------------------------------------------------
// Truncations (flow/sample.rs:18-25, warp.rs:104-121, image.rs:249-262, timeline.rs:43-58)

static inline int32_t cvtt(float x) {
    if (std::isnan(x) || x >= 2147483648.0f || x < -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
static inline int64_t cvtt64f(float x) {  // blend.rs:18-25
    const float two63 = 9223372036854775808.0f;
    if (std::isnan(x) || x >= two63 || x < -two63) return INT64_MIN;
    return (int64_t)x;
}
static inline int64_t cvtt64d(double x) {  // image.rs:249-255, warp.rs:114-121
    if (std::isnan(x) || x >= 9.223372036854775808e18 || x < -9.223372036854775808e18) return INT64_MIN;
    return (int64_t)x;
}
static inline int32_t cvtt_i32d(double x) {  // timeline.rs:43-49
    if (x > -2147483649.0 && x < 2147483648.0) return (int32_t)x;
    return INT32_MIN;
}
static inline int32_t cvtt_i32f(float x) {  // timeline.rs:52-58
    if (x > -2147483904.0f && x < 2147483648.0f) return (int32_t)x;
    return INT32_MIN;
}

// ------------------------------------------------------------------------------------------
// Convolution (flow/conv.rs)

static const float BINOMIAL5[3] = {0.375f, 0.25f, 0.0625f};  // conv.rs:49
static const float PAD_MULTI = -1.0f;                         // conv.rs:89

// conv.rs:60-86: taps exp2(-(i*i)/d) as double pow narrowed, normalised by sum*2 - k0.
static std::vector<float> gaussian_kernel(float r) {
    int n = std::max((int32_t)r + 1, 0);
    float s3 = r / 3.0f;
    float d = s3 * s3;
    std::vector<float> k;
    float sum = 0.0f;
    for (int i = 0; i < n; i++) {
        float fi = (float)i;
        float t = fi * fi;
        float e = -(t / d);
        float v = (float)std::pow(2.0, (double)e);
        k.push_back(v);
        sum += v;
    }
    if (n == 0) return k;
    float den = sum * 2.0f - k[0];
    for (auto& v : k) v /= den;
    return k;
}

// conv.rs:99-141: border element with renormalisation. `get(t, &v)` returns false outside.
// LEFT_FIRST takes the tap at i-j before the tap at i+j.
template <int C, bool LEFT_FIRST, class Get>
static inline float border_value(const Get& get, long i, const float* k, int n) {
    float centre;
    if (!get(i, centre)) centre = 0.0f;
    float num = k[0] * centre;
    float den = k[0];
    for (int j = n - 1; j >= 1; j--) {
        float kj = k[j];
        for (int side = 0; side < 2; side++) {
            bool right = (side == 0) != LEFT_FIRST;
            long t = right ? i + j : i - j;
            float xv, w;
            float v;
            if (C == 1) {
                if (get(t, v)) { xv = v; w = kj; } else { xv = 0.0f; w = 0.0f; }
            } else {
                if (!get(t, v)) v = PAD_MULTI;
                float mask = (v != PAD_MULTI) ? 1.0f : 0.0f;
                xv = v;
                w = mask * kj;
            }
            den += w;
            num = num + xv * w;
        }
    }
    return num / den;
}

// conv.rs:144-199: one row of the horizontal pass.
template <int C>
static void hrow(const float* x, float* out, long len, const float* k, int n) {
    long head = std::min((long)n, len);
    long interior = len > 2L * n ? len - 2L * n : 0;
    auto border = [&](long i) {
        for (int c = 0; c < C; c++) {
            auto get = [&](long t, float& v) -> bool {
                if (t < 0 || t >= len) return false;
                v = x[t * C + c];
                return true;
            };
            out[i * C + c] = border_value<C, false>(get, i, k, n);
        }
    };
    for (long i = 0; i < head; i++) border(i);
    if (interior > 0) {
        long a = head * C, b = (head + interior) * C;
        float k0 = k[0];
        for (long t = a; t < b; t++) out[t] = x[t] * k0;
        for (int j = n - 1; j >= 1; j--) {
            long off = (long)j * C;
            float kj = k[j];
            for (long t = a; t < b; t++) out[t] = out[t] + (x[t + off] + x[t - off]) * kj;
        }
    }
    for (long i = head + interior; i < len; i++) border(i);
}

// conv.rs:272-311: vertical pass output row y; ABOVE_FIRST takes row y-j before y+j in borders.
template <int C, bool ABOVE_FIRST>
static void vrow(const float* src, float* out, long y, long w, long h, const float* k, int n) {
    long row = w * C;
    long head = std::min((long)n, h);
    long interior = h > 2L * n ? h - 2L * n : 0;
    if (y >= head && y < head + interior) {
        float k0 = k[0];
        const float* s0 = src + y * row;
        for (long t = 0; t < row; t++) out[t] = s0[t] * k0;
        for (int j = n - 1; j >= 1; j--) {
            const float* below = src + (y + j) * row;
            const float* above = src + (y - j) * row;
            float kj = k[j];
            for (long t = 0; t < row; t++) out[t] = out[t] + (below[t] + above[t]) * kj;
        }
    } else {
        for (long xc = 0; xc < w; xc++) {
            for (int c = 0; c < C; c++) {
                auto get = [&](long t, float& v) -> bool {
                    if (t < 0 || t >= h) return false;
                    v = src[(t * w + xc) * C + c];
                    return true;
                };
                out[xc * C + c] = border_value<C, ABOVE_FIRST>(get, y, k, n);
            }
        }
    }
}

template <int C, bool ABOVE_FIRST>
static void convolve_c(float* data, long w, long h, const float* k, int n) {
    long len = w * h * C;
    if (len == 0 || n == 0) return;
    std::vector<float> tmp(len);
    long row = w * C;
#pragma omp parallel for schedule(static)
    for (long y = 0; y < h; y++) hrow<C>(data + y * row, tmp.data() + y * row, w, k, n);
#pragma omp parallel for schedule(static)
    for (long y = 0; y < h; y++) vrow<C, ABOVE_FIRST>(tmp.data(), data + y * row, y, w, h, k, n);
}

// conv.rs:584-606 convolve_above_first (the solver's and the pyramid's), and conv.rs:417-428
// convolve (below first) for completeness.
static void convolve(float* data, long w, long h, int c, const float* k, int n, bool above_first) {
    if (above_first) {
        if (c == 1) convolve_c<1, true>(data, w, h, k, n);
        else if (c == 2) convolve_c<2, true>(data, w, h, k, n);
        else convolve_c<3, true>(data, w, h, k, n);
    } else {
        if (c == 1) convolve_c<1, false>(data, w, h, k, n);
        else if (c == 2) convolve_c<2, false>(data, w, h, k, n);
        else convolve_c<3, false>(data, w, h, k, n);
    }
}

// conv.rs:610-616
static void blur_gaussian_above_first(float* data, long w, long h, int c, float r) {
    if ((double)r < 0.001) return;
    std::vector<float> k = gaussian_kernel(r);
    convolve(data, w, h, c, k.data(), (int)k.size(), true);
}

// ------------------------------------------------------------------------------------------
// Samplers (flow/sample.rs)

// sample.rs:31-41
static inline void cell(float x, long len, long& io, float& fo) {
    int32_t i = cvtt(x);
    float f = x - (float)i;
    int32_t last = (int32_t)len - 1;
    if (i >= last) { i = last - 1; f = 1.0f; }
    long ii = i < 0 ? 0 : (long)i;
    long cap = len >= 2 ? len - 2 : 0;
    io = std::min(ii, cap);
    fo = f;
}

// sample.rs:46-64; lerp_g of nanrule.rs:227: (b - a) * f + a
static inline float bilinear1(const float* img, long w, long h, float x, float y) {
    long ix, iy;
    float fx, fy;
    cell(x, w, ix, fx);
    cell(y, h, iy, fy);
    long i = iy * w + ix;
    float p00 = img[i], p10 = img[i + 1], p01 = img[i + w], p11 = img[i + w + 1];
    float top = (p10 - p00) * fx + p00;
    float bot = (p11 - p01) * fx + p01;
    return (bot - top) * fy + top;
}

// sample.rs:69-87
static inline void bilinear2(const float* img, long w, long h, float x, float y, float out[2]) {
    long ix, iy;
    float fx, fy;
    cell(x, w, ix, fx);
    cell(y, h, iy, fy);
    long i = (iy * w + ix) * 2;
    long r = w * 2;
    for (int c = 0; c < 2; c++) {
        float p00 = img[i + c], p10 = img[i + 2 + c], p01 = img[i + r + c], p11 = img[i + r + 2 + c];
        float top = (p10 - p00) * fx + p00;
        float bot = (p11 - p01) * fx + p01;
        out[c] = (bot - top) * fy + top;
    }
}

// sample.rs:93-112
static std::vector<float> resample1(const float* src, long sw, long sh, long dw, long dh) {
    float sx = ((float)sw - 1.0f) / (float)((int32_t)dw - 1);
    float sy = ((float)sh - 1.0f) / (float)((int32_t)dh - 1);
    std::vector<float> dst(dw * dh, 0.0f);
    if (dw == 0 || dh == 0) return dst;
#pragma omp parallel for schedule(static)
    for (long y = 0; y < dh; y++) {
        float fy = (float)(int32_t)y * sy;
        for (long x = 0; x < dw; x++) dst[y * dw + x] = bilinear1(src, sw, sh, (float)(int32_t)x * sx, fy);
    }
    return dst;
}

// ------------------------------------------------------------------------------------------
// Geometry and luma (prep.rs)

// prep.rs:63-89
static void flow_resolution(int q, bool hiq, long w, long h, int& r, long& w0, long& h0) {
    r = (q <= 2) ? 3 : (q == 3 ? 2 : (q == 4 ? 1 : 0));
    if (!hiq) {
        if (r == 0) r = 1;
        else if (r == 1) r = 2;
        else if (r == 2) r = 3;
    }
    while (r > 0 && ((w >> r) < 2 || (h >> r) < 2)) r--;
    w0 = w;
    h0 = h;
    for (int i = 0; i < r; i++) { w0 /= 2; h0 /= 2; }
}

// prep.rs:116-134
static int level_count(long w0, long h0, long min_size) {
    int levels = 1;
    long w = w0, h = h0;
    for (;;) {
        w /= 2;
        if (w < min_size) break;
        h /= 2;
        if (h < min_size) break;
        if (w > 2 && h > 2 && !(w == 2 && h == 2)) levels++;
        else break;
    }
    return levels;
}

// prep.rs:137-146
static std::vector<std::pair<long, long>> level_sizes(long w0, long h0, int levels) {
    std::vector<std::pair<long, long>> v;
    long w = w0, h = h0;
    for (int i = 0; i < levels; i++) {
        v.push_back({w, h});
        w = std::max(w / 2, 2L);
        h = std::max(h / 2, 2L);
    }
    return v;
}

// prep.rs:168-235 (track_alpha and video_range off: the retime path)
static std::vector<float> luma(const void* img, long w, long h, int fmt) {
    std::vector<float> out(w * h);
    if (fmt == TWX_FMT_U8) {
        const uint8_t* s = (const uint8_t*)img;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < w * h; i++) {
            const uint8_t* px = s + i * 4;
            float r = (float)px[1], g = (float)px[2], b = (float)px[3];
            out[i] = (r * 0.3f + g * 0.59f) + b * 0.11f;
        }
    } else if (fmt == TWX_FMT_U16) {
        const uint16_t* s = (const uint16_t*)img;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < w * h; i++) {
            const uint16_t* px = s + i * 4;
            double r = px[1], g = px[2], b = px[3];
            double sum = (g * 0.59 + r * 0.3) + b * 0.11;
            out[i] = (float)(sum / 257.0);
        }
    } else {
        const float* s = (const float*)img;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < w * h; i++) {
            const float* px = s + i * 4;
            double r = px[1], g = px[2], b = px[3];
            double sum = (g * 0.59 + r * 0.3) + b * 0.11;
            out[i] = (float)(sum * 255.0);
        }
    }
    return out;
}

// ------------------------------------------------------------------------------------------
// Flow geometry, start level (flow.rs:118-197)

struct Geometry {
    int reduction;
    long width, height;
    int levels;
    std::vector<std::pair<long, long>> sizes;
};

// flow.rs:118-137. Returns a status.
static int geometry(const twx_flow_settings& fs, long out_w, long out_h, Geometry& g, std::string& err) {
    if (fs.quality == 0 || fs.quality > 5) { err = "flow quality outside 1..=5"; return TWX_E_MISMATCH; }
    if (!(fs.sensitivity >= 0.0f) || !std::isfinite(fs.sensitivity)) { err = "bad sensitivity"; return TWX_E_MISMATCH; }
    flow_resolution(fs.quality, fs.hiq != 0, out_w, out_h, g.reduction, g.width, g.height);
    if (g.width < 2 || g.height < 2) { err = "optical flow: working size below 2x2"; return TWX_E_UNSUPPORTED; }
    g.levels = level_count(g.width, g.height, 2);
    g.sizes = level_sizes(g.width, g.height, g.levels);
    return TWX_OK;
}

// flow.rs:179-197
static void start_level(float sensitivity, int levels, int& start, float& damp) {
    float p = (float)std::pow((double)sensitivity, 2.0);
    float l1 = (float)std::log((double)p);
    float l2 = (float)std::log(2.0);
    float lv = (float)levels;
    float x = l1 / l2 + lv;
    if (0.0f <= x || std::isnan(x)) {
        if (x < lv || std::isnan(x)) {
            float f = std::floor(x);
            int32_t s = cvtt(f);
            long su = s < 0 ? 0 : (long)s;
            start = (int)std::min(su, (long)levels - 1);
            damp = x - (float)su;
        } else {
            start = levels - 1;
            damp = 1.0f;
        }
    } else {
        start = 0;
        damp = 0.0f;
    }
}

// ------------------------------------------------------------------------------------------
// Pyramid (prep.rs:246-259, flow.rs:242-275)

struct Pyramid {
    Geometry geo;
    std::vector<std::vector<float>> levels;
};

static void blur5(std::vector<float>& p, long w, long h) { convolve(p.data(), w, h, 1, BINOMIAL5, 3, true); }

static int build_pyramid(const void* img, long w, long h, int fmt, const twx_flow_settings& fs, Pyramid& out,
                         std::string& err) {
    int st = geometry(fs, w, h, out.geo, err);
    if (st) return st;
    if (fs.prep_contrast_edge || fs.prep_delinearize || fs.wrap || fs.track_alpha || fs.video_range_luma) {
        err = "oracle: prep, wrap, track alpha and video range luma are not restated";
        return TWX_E_UNSUPPORTED;
    }
    std::vector<float> t = luma(img, w, h, fmt);
    long tw = w, th = h;
    for (int i = 0; i < out.geo.reduction; i++) {  // prep.rs:252-259 reduce
        blur5(t, tw, th);
        long nw = tw / 2, nh = th / 2;
        t = resample1(t.data(), tw, th, nw, nh);
        tw = nw;
        th = nh;
    }
    int n = out.geo.levels;
    std::vector<float> orig = t;
    out.levels.clear();
    for (int i = 0; i < n; i++) {
        long lw = out.geo.sizes[i].first, lh = out.geo.sizes[i].second;
        std::vector<float> tmp = orig;
        blur5(tmp, lw, lh);
        std::vector<float> next;
        bool has_next = i + 1 < n;
        if (has_next) next = resample1(tmp.data(), lw, lh, out.geo.sizes[i + 1].first, out.geo.sizes[i + 1].second);
        std::vector<float> band = orig;
        for (size_t q = 0; q < band.size(); q++) band[q] = band[q] - tmp[q];
        out.levels.push_back(std::move(band));
        if (!has_next) break;
        orig = std::move(next);
    }
    return TWX_OK;
}

// ------------------------------------------------------------------------------------------
// Solver kernels (flow.rs:281-442)

static std::vector<float> calc_grad(const float* img, long w, long h) {
    std::vector<float> g(w * h * 2, 0.0f);
#pragma omp parallel for schedule(static)
    for (long y = 0; y < h; y++) {
        for (long x = 0; x < w; x++) {
            float c = img[y * w + x];
            float nx = 0.0f, dx = 0.0f, ny = 0.0f, dy = 0.0f;
            if (x >= 1) { nx = (c - img[y * w + x - 1]) + 0.0f; dx = 1.0f; }
            if (y >= 1) { ny = (c - img[(y - 1) * w + x]) + 0.0f; dy = 1.0f; }
            if (x + 1 <= w - 1) { nx = nx + (img[y * w + x + 1] - c); dx = dx + 1.0f; }
            if (y + 1 <= h - 1) { ny = ny + (img[(y + 1) * w + x] - c); dy = dy + 1.0f; }
            g[(y * w + x) * 2] = nx / dx;
            g[(y * w + x) * 2 + 1] = ny / dy;
        }
    }
    return g;
}

static std::vector<float> calc_a(const std::vector<float>& grad, long w, long h) {
    std::vector<float> a(w * h * 3);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < w * h; i++) {
        float gx = grad[i * 2], gy = grad[i * 2 + 1];
        a[i * 3] = gx * gx;
        a[i * 3 + 1] = gx * gy;
        a[i * 3 + 2] = gy * gy;
    }
    return a;
}

static std::vector<float> calc_b(const float* i1, const float* i0, const float* grad, const float* flow, long w,
                                 long h) {
    std::vector<float> b(w * h * 2, 0.0f);
    float fw = (float)w, fh = (float)h;
#pragma omp parallel for schedule(static)
    for (long y = 0; y < h; y++) {
        float fy = (float)y;
        for (long x = 0; x < w; x++) {
            float fx = (float)x;
            long i = y * w + x;
            float sx = fx - flow[i * 2];
            float sy = fy - flow[i * 2 + 1];
            if (sx < 0.0f || fw - 1.0f < sx || sy < 0.0f || fh - 1.0f < sy) {
                b[i * 2] = 0.0f;
                b[i * 2 + 1] = 0.0f;
            } else {
                float s = bilinear1(i0, w, h, sx, sy);
                float r = s - i1[i];
                b[i * 2] = r * grad[i * 2];
                b[i * 2 + 1] = r * grad[i * 2 + 1];
            }
        }
    }
    return b;
}

static void new_disp(float* flow, const float* a, const float* b, long n) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        float aa = a[i * 3], ab = a[i * 3 + 1], ad = a[i * 3 + 2];
        float bx = b[i * 2], by = b[i * 2 + 1];
        float t = ab * ab - ad * aa;
        if (std::fabs(t) > 40.0f) {
            flow[i * 2] = (ab * by - ad * bx) / t + flow[i * 2];
            flow[i * 2 + 1] = (ab * bx - aa * by) / t + flow[i * 2 + 1];
        }
    }
}

static std::vector<float> upsample_flow(const float* src, long sw, long sh, long dw, long dh) {
    float sx = ((float)dw - 1.0f) / (float)((int32_t)sw - 1);
    float sy = ((float)dh - 1.0f) / (float)((int32_t)sh - 1);
    float rx = (float)(1.0 / (double)sx);
    float ry = (float)(1.0 / (double)sy);
    std::vector<float> dst(dw * dh * 2, 0.0f);
#pragma omp parallel for schedule(static)
    for (long y = 0; y < dh; y++) {
        float fy = (float)(int32_t)y * ry;
        for (long x = 0; x < dw; x++) {
            float v[2];
            bilinear2(src, sw, sh, (float)(int32_t)x * rx, fy, v);
            dst[(y * dw + x) * 2] = v[0] * sx;
            dst[(y * dw + x) * 2 + 1] = v[1] * sy;
        }
    }
    return dst;
}

static void scale_values(std::vector<float>& d, float f) {
    if (f == 1.0f) return;
    for (auto& v : d) v = f * v;
}

static bool all_flow_out_of_range(const float* flow, long w, long h) {
    float fw = (float)(w - 1), fh = (float)(h - 1);
    for (long y = 0; y < h; y++) {
        float fy = (float)y;
        for (long x = 0; x < w; x++) {
            float px = (float)(int32_t)x + flow[(y * w + x) * 2];
            float py = fy + flow[(y * w + x) * 2 + 1];
            if (0.0f <= px && px <= fw && 0.0f <= py && py <= fh) return false;
        }
    }
    return true;
}

// ------------------------------------------------------------------------------------------
// Block matching (flow.rs:446-622)

struct Block {
    const float* a;
    const float* b;
    long w, h;

    inline void patch(long x, long y, float p[9]) const {
        long i = (y - 1) * w + x - 1;
        p[0] = a[i];
        p[1] = a[i + 1];
        p[2] = a[i + 2];
        p[3] = a[i + w];
        p[4] = a[i + w + 1];
        p[5] = a[i + w + 2];
        p[6] = a[i + 2 * w];
        p[7] = a[i + 2 * w + 1];
        p[8] = a[i + 2 * w + 2];
    }

    // flow.rs:472-524; returns false for "None".
    inline bool cost(const float p[9], float x2, float y2, float& out) const {
        float wm2 = (float)((int32_t)w - 2);
        float hm2 = (float)((int32_t)h - 2);
        if (x2 > wm2 || y2 > hm2) return false;
        int32_t ix, iy;
        float fx, fy;
        if (x2 == wm2) { ix = cvtt(wm2 - 1.0f); fx = 1.0f; }
        else { ix = cvtt(x2); fx = x2 - (float)ix; }
        if (y2 == hm2) { iy = cvtt(hm2 - 1.0f); fy = 1.0f; }
        else { iy = cvtt(y2); fy = y2 - (float)iy; }
        if (ix < 1 || iy < 1) return false;
        float w11 = fy * fx;
        float w10 = fx - w11;
        float w01 = fy - w11;
        float w00 = ((1.0f - fx) - fy) + w11;
        long base = ((long)iy - 1) * w + (long)ix - 1;
        float sq[9];
        int n = 0;
        for (int bb = 0; bb < 3; bb++) {
            for (int aa = 0; aa < 3; aa++) {
                const float* t0 = b + base + (long)bb * w + aa;
                const float* t1 = t0 + w;
                float q = ((w00 * t0[0] + w10 * t0[1]) + w01 * t1[0]) + w11 * t1[1];
                float d = p[n] - q;
                sq[n] = d * d;
                n++;
            }
        }
        float s = sq[0] + sq[1];
        for (int k = 2; k < 8; k++) s = s + sq[k];
        out = sq[8] + s;
        return true;
    }

    // flow.rs:528-544; returns true when the neighbour's vector replaces the pixel's.
    inline bool step(long x, long y, const float cur[2], const float nb[2]) const {
        float tx = (float)x - cur[0];
        float ty = (float)y - cur[1];
        float w1 = (float)(w - 1), h1 = (float)(h - 1);
        if (0.0f <= tx && tx <= w1 && 0.0f <= ty && ty <= h1) {
            float p[9];
            patch(x, y, p);
            float c0, c1;
            if (cost(p, tx, ty, c0)) {
                if (cost(p, (float)x - nb[0], (float)y - nb[1], c1)) {
                    if (c1 < c0) return true;
                }
            }
        }
        return false;
    }
};

static void block_match(const float* a, const float* b, float* flow, long w, long h) {
    if (w < 3 || h < 3) return;
    Block bm{a, b, w, h};
    // sweep 1: left to right
#pragma omp parallel for schedule(static)
    for (long y = 1; y <= h - 2; y++) {
        float* row = flow + y * w * 2;
        for (long x = 1; x <= w - 2; x++) {
            float cur[2] = {row[x * 2], row[x * 2 + 1]};
            float nb[2] = {row[(x - 1) * 2], row[(x - 1) * 2 + 1]};
            if (bm.step(x, y, cur, nb)) { row[x * 2] = nb[0]; row[x * 2 + 1] = nb[1]; }
        }
    }
    // sweep 2: right to left
#pragma omp parallel for schedule(static)
    for (long y = 1; y <= h - 2; y++) {
        float* row = flow + y * w * 2;
        for (long x = w - 2; x >= 1; x--) {
            float cur[2] = {row[x * 2], row[x * 2 + 1]};
            float nb[2] = {row[(x + 1) * 2], row[(x + 1) * 2 + 1]};
            if (bm.step(x, y, cur, nb)) { row[x * 2] = nb[0]; row[x * 2 + 1] = nb[1]; }
        }
    }
    // sweeps 3 and 4: columns (the reference transposes for contiguity; the values are the same)
#pragma omp parallel for schedule(static)
    for (long x = 1; x <= w - 2; x++) {
        for (long y = 1; y <= h - 2; y++) {
            float* c = flow + (y * w + x) * 2;
            float* n = flow + ((y - 1) * w + x) * 2;
            float cur[2] = {c[0], c[1]};
            float nb[2] = {n[0], n[1]};
            if (bm.step(x, y, cur, nb)) { c[0] = nb[0]; c[1] = nb[1]; }
        }
    }
#pragma omp parallel for schedule(static)
    for (long x = 1; x <= w - 2; x++) {
        for (long y = h - 2; y >= 1; y--) {
            float* c = flow + (y * w + x) * 2;
            float* n = flow + ((y + 1) * w + x) * 2;
            float cur[2] = {c[0], c[1]};
            float nb[2] = {n[0], n[1]};
            if (bm.step(x, y, cur, nb)) { c[0] = nb[0]; c[1] = nb[1]; }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Driver (flow.rs:1258-1353)

static int solve_pyramids(const Pyramid& X, const Pyramid& Y, const twx_flow_settings& fs, std::vector<float>& out,
                          std::string& err) {
    if (X.geo.sizes != Y.geo.sizes) { err = "pyramids of different geometry"; return TWX_E_MISMATCH; }
    const Geometry& geo = X.geo;
    int start;
    float damp;
    start_level(fs.sensitivity, geo.levels, start, damp);
    std::vector<float> flow(geo.sizes[start].first * geo.sizes[start].second * 2, 0.0f);
    bool check = true;
    for (int l = start; l >= 0; l--) {
        long w = geo.sizes[l].first, h = geo.sizes[l].second;
        const float* img_a = X.levels[l].data();
        const float* img_b = Y.levels[l].data();
        std::vector<float> grad = calc_grad(img_a, w, h);
        std::vector<float> a = calc_a(grad, w, h);
        blur_gaussian_above_first(a.data(), w, h, 3, 7.0f);
        for (int it = 0; it < 3; it++) {
            std::vector<float> b = calc_b(img_a, img_b, grad.data(), flow.data(), w, h);
            blur_gaussian_above_first(b.data(), w, h, 2, 7.0f);
            new_disp(flow.data(), a.data(), b.data(), w * h);
            blur_gaussian_above_first(flow.data(), w, h, 2, 5.0f);
            if (it == 1 && l == 0 && !fs.skip_refine) block_match(img_a, img_b, flow.data(), w, h);
        }
        if (l != 0 && !fs.skip_refine) block_match(img_a, img_b, flow.data(), w, h);
        if (check) {
            if (all_flow_out_of_range(flow.data(), w, h)) std::fill(flow.begin(), flow.end(), 0.0f);
            else check = false;
        }
        if (l == start && std::fabs((double)damp - 1.0) >= 1e-6) scale_values(flow, damp);
        if (l > 0) flow = upsample_flow(flow.data(), w, h, geo.sizes[l - 1].first, geo.sizes[l - 1].second);
    }
    scale_values(flow, -1.0f);
    out = std::move(flow);
    return TWX_OK;
}

// ------------------------------------------------------------------------------------------
// Engine images (image.rs:249-351, frame.rs:79-141)

static inline int64_t q(float x) { return cvtt64d((double)x + 0.5); }  // image.rs:259-261
static const float INV_255 = [] { uint32_t b = 0x3b808081u; float f; memcpy(&f, &b, 4); return f; }();
static const float PREMUL_16_COLOR = [] { uint32_t b = 0x387fff00u; float f; memcpy(&f, &b, 4); return f; }();
static const float PREMUL_16_ALPHA = [] { uint32_t b = 0x3fffff00u; float f; memcpy(&f, &b, 4); return f; }();
static const float UNPREMUL_16_ALPHA = [] { uint32_t b = 0x3f000080u; float f; memcpy(&f, &b, 4); return f; }();

static inline uint16_t u16_to_q15(uint16_t v) { return (uint16_t)(((uint32_t)v * 32768u + 32767u) / 65535u); }
static inline uint16_t q15_to_u16(uint16_t v) {
    uint32_t m = std::min<uint32_t>(v, 32768u);
    return (uint16_t)((m * 65535u + 16384u) / 32768u);
}

static inline void premultiply_u8(const uint8_t p[4], uint8_t o[4]) {
    float a = (float)p[0] * INV_255;
    o[0] = p[0];
    for (int c = 1; c < 4; c++) o[c] = (uint8_t)q((float)p[c] * a);
}
static inline void premultiply_u16(const uint16_t p[4], uint16_t o[4]) {
    float a = (float)p[0] * PREMUL_16_COLOR;
    o[0] = (uint16_t)q((float)p[0] * PREMUL_16_ALPHA);
    for (int c = 1; c < 4; c++) o[c] = (uint16_t)q((float)p[c] * a);
}
// ------------------------------------------------------------------------------------------
// The NaN rule (nanrule.rs:1-120, the emulator rule the references were rendered under): numbers
// never depend on it. A NaN result of an f32 sum, difference, product or quotient takes: the one
// NaN operand, quieted; of two NaN operands the one with the larger fraction (equal fractions: the
// positive one; a quiet NaN before a signalling one), quieted; for an invalid operation (inf - inf,
// 0 * inf, 0 / 0) the x86 default NaN 0xffc00000. The call sites compute plainly and repeat a
// result with these operations only when it holds a NaN, as the reference does.
static inline uint32_t nr_u(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }
static inline float nr_f(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }
static inline float nr_quiet(float x) { return nr_f(nr_u(x) | 0x00400000u); }
static inline float nr_pick(float a, float b) {  // nanrule.rs:44-73
    const uint32_t ba = nr_u(a), bb = nr_u(b);
    const bool sa = !(ba & 0x00400000u), sb = !(bb & 0x00400000u);
    const uint32_t fa = ba & 0x007fffffu, fb = bb & 0x007fffffu;
    const bool a_larger = fa > fb || (fa == fb && (ba >> 31) < (bb >> 31));
    const bool take_b = (sa == sb) ? !a_larger : sa;
    return nr_quiet(take_b ? b : a);
}
static inline float nr_fix(float r, float a, float b) {  // nanrule.rs:76-93
    if (!std::isnan(r)) return r;
    if (std::isnan(a)) return std::isnan(b) ? nr_pick(a, b) : nr_quiet(a);
    if (std::isnan(b)) return nr_quiet(b);
    return nr_f(0xffc00000u);
}
static inline float nr_add(float a, float b) { return nr_fix(a + b, a, b); }
static inline float nr_sub(float a, float b) { return nr_fix(a - b, a, b); }
static inline float nr_mul(float a, float b) { return nr_fix(a * b, a, b); }
static inline float nr_div(float a, float b) { return nr_fix(a / b, a, b); }
// lerp (nanrule.rs:240): (b - a) * f + a; lerp_base_first (:247): a + (b - a) * f
static inline float nr_lerp(float a, float b, float f) { return nr_add(nr_mul(nr_sub(b, a), f), a); }
static inline float nr_lerp_base_first(float a, float b, float f) { return nr_add(a, nr_mul(nr_sub(b, a), f)); }

// image.rs:291-300: colour times alpha, the products again with the NaN rule when one is a NaN
static inline void premultiply_f32(const float p[4], float o[4]) {
    o[0] = p[0];
    bool nan = std::isnan(p[0]);
    for (int c = 1; c < 4; c++) {
        o[c] = p[0] * p[c];
        nan |= std::isnan(o[c]);
    }
    if (nan)
        for (int c = 1; c < 4; c++) o[c] = nr_mul(p[0], p[c]);
}
static inline void unpremultiply_u8(const uint8_t p[4], uint8_t o[4]) {
    if (p[0] == 0) { memcpy(o, p, 4); return; }
    float s = 255.0f / (float)p[0];
    o[0] = p[0];
    for (int c = 1; c < 4; c++) {
        float v = (float)p[c] * s;
        if (v > 255.0f) v = 255.0f;
        o[c] = (uint8_t)q(v);
    }
}
static inline void unpremultiply_u16(const uint16_t p[4], uint16_t o[4]) {
    if (p[0] == 0) { memcpy(o, p, 8); return; }
    float a = (float)p[0];
    float s = 32768.0f / a;
    o[0] = (uint16_t)q(a * UNPREMUL_16_ALPHA);
    for (int c = 1; c < 4; c++) o[c] = (uint16_t)q((float)p[c] * s);
}
static inline void unpremultiply_f32(const float p[4], float o[4]) {
    if ((double)p[0] < 1e-12) { o[0] = o[1] = o[2] = o[3] = 0.0f; return; }
    float inv = 1.0f / p[0];
    o[0] = p[0];
    bool nan = std::isnan(p[0]);
    for (int c = 1; c < 4; c++) {
        o[c] = p[c] * inv;
        nan |= std::isnan(o[c]);
    }
    if (nan)  // image.rs:344-348: the colour is the first operand of the product
        for (int c = 1; c < 4; c++) o[c] = nr_mul(p[c], inv);
}

static inline size_t sample_size(int fmt) { return fmt == TWX_FMT_U8 ? 1 : (fmt == TWX_FMT_U16 ? 2 : 4); }
static inline size_t image_bytes(long w, long h, int fmt) { return (size_t)w * h * 4 * sample_size(fmt); }

// Image::from_frame of one wire frame (image.rs:173-203, frame.rs:79-111).
static void from_frame(const uint8_t* wire, long w, long h, int fmt, void* engine) {
    long n = w * h;
    if (fmt == TWX_FMT_U8) {
        uint8_t* o = (uint8_t*)engine;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            const uint8_t* c = wire + i * 4;
            uint8_t p[4] = {c[3], c[0], c[1], c[2]};
            premultiply_u8(p, o + i * 4);
        }
    } else if (fmt == TWX_FMT_U16) {
        uint16_t* o = (uint16_t*)engine;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            const uint8_t* c = wire + i * 8;
            uint16_t v[4];
            for (int k = 0; k < 4; k++) v[k] = u16_to_q15((uint16_t)(c[2 * k] | (c[2 * k + 1] << 8)));
            uint16_t p[4] = {v[3], v[0], v[1], v[2]};
            premultiply_u16(p, o + i * 4);
        }
    } else {
        float* o = (float*)engine;
        const float* pl = (const float*)wire;  // planes G, B, R, A
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            float p[4] = {pl[3 * n + i], pl[2 * n + i], pl[0 * n + i], pl[1 * n + i]};
            premultiply_f32(p, o + i * 4);
        }
    }
}

// Image::to_frame (image.rs:205-245, frame.rs:113-132)
static void to_frame(const void* engine, long w, long h, int fmt, uint8_t* wire) {
    long n = w * h;
    if (fmt == TWX_FMT_U8) {
        const uint8_t* s = (const uint8_t*)engine;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            uint8_t o[4];
            unpremultiply_u8(s + i * 4, o);
            uint8_t* c = wire + i * 4;
            c[0] = o[1]; c[1] = o[2]; c[2] = o[3]; c[3] = o[0];
        }
    } else if (fmt == TWX_FMT_U16) {
        const uint16_t* s = (const uint16_t*)engine;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            uint16_t o[4];
            unpremultiply_u16(s + i * 4, o);
            uint16_t v[4] = {q15_to_u16(o[1]), q15_to_u16(o[2]), q15_to_u16(o[3]), q15_to_u16(o[0])};
            uint8_t* c = wire + i * 8;
            for (int k = 0; k < 4; k++) { c[2 * k] = (uint8_t)(v[k] & 0xff); c[2 * k + 1] = (uint8_t)(v[k] >> 8); }
        }
    } else {
        const float* s = (const float*)engine;
        float* pl = (float*)wire;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            float o[4];
            unpremultiply_f32(s + i * 4, o);
            // o = [A, R, G, B]; planes G, B, R, A
            pl[0 * n + i] = o[2];
            pl[1 * n + i] = o[3];
            pl[2 * n + i] = o[1];
            pl[3 * n + i] = o[0];
        }
    }
}

// Image::rows_reversed (image.rs:96-114)
static void rows_reversed(const void* src, long w, long h, int fmt, void* dst) {
    size_t row = (size_t)w * 4 * sample_size(fmt);
    for (long y = 0; y < h; y++) memcpy((uint8_t*)dst + y * row, (const uint8_t*)src + (h - 1 - y) * row, row);
}

// ------------------------------------------------------------------------------------------
// Inverse warp, hold edges (warp.rs:99-786)

static inline uint8_t sat_u8(float v) {
    int32_t i = cvtt(v + 0.5f);
    int32_t t = (int32_t)((uint32_t)i + (uint32_t)-128);
    int32_t w = std::min(std::max(t, -32768), 32767);
    int32_t b = std::min(std::max(w, -128), 127);
    return (uint8_t)(b + 128);
}
static inline uint16_t sat_u16(float v) {
    int32_t i = cvtt(v + 0.5f);
    int32_t t = (int32_t)((uint32_t)i + (uint32_t)-32768);
    int32_t w = std::min(std::max(t, -32768), 32767);
    return (uint16_t)(w + 32768);
}

// warp.rs:178-215
static bool clip_to_rect(float x, float y, float& ix, float& iy, float w1, float h1) {
    for (uint32_t n = 0; n < 4096; n++) {
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

template <class P> static inline float to_f(P v) { return (float)v; }
template <class P> static inline P store(float v);
template <> inline uint8_t store<uint8_t>(float v) { return sat_u8(v); }
template <> inline uint16_t store<uint16_t>(float v) { return sat_u16(v); }
template <> inline float store<float>(float v) { return v; }

// warp.rs:306-360 bilerp4 + mix4 (lerp_base_first: a + (b - a) * f, the same value as (b - a) * f + a)
template <class P>
static inline void bilerp4(const P* src, long w, int32_t w1, int32_t h1, float x, float y, float out[4]) {
    int32_t ix = cvtt(x), iy = cvtt(y);
    float fx = x - (float)ix, fy = y - (float)iy;
    if (ix >= w1) { ix = w1 - 1; fx = 1.0f; }
    if (iy >= h1) { iy = h1 - 1; fy = 1.0f; }
    long i00 = ((long)iy * w + ix) * 4;
    long i01 = i00 + w * 4;
    bool nan = false;
    for (int k = 0; k < 4; k++) {
        float p00 = to_f(src[i00 + k]), p10 = to_f(src[i00 + 4 + k]);
        float p01 = to_f(src[i01 + k]), p11 = to_f(src[i01 + 4 + k]);
        float top = p00 + (p10 - p00) * fx;
        float bot = p01 + (p11 - p01) * fx;
        out[k] = top + (bot - top) * fy;
        nan |= std::isnan(out[k]);
    }
    if (nan) {  // warp.rs:334-343: the same lerps again with the NaN rule
        for (int k = 0; k < 4; k++) {
            float p00 = to_f(src[i00 + k]), p10 = to_f(src[i00 + 4 + k]);
            float p01 = to_f(src[i01 + k]), p11 = to_f(src[i01 + 4 + k]);
            out[k] = nr_lerp_base_first(nr_lerp_base_first(p00, p10, fx), nr_lerp_base_first(p01, p11, fx), fy);
        }
    }
}

// warp.rs:415-442 Flow::bilerp
static inline void flow_bilerp(const float* d, long fw, int32_t fw1, int32_t fh1, float x, float y, float out[2]) {
    int32_t ix = cvtt(x), iy = cvtt(y);
    float fx = x - (float)ix, fy = y - (float)iy;
    if (ix >= fw1) { ix = fw1 - 1; fx = 1.0f; }
    if (iy >= fh1) { iy = fh1 - 1; fy = 1.0f; }
    long i00 = ((long)iy * fw + ix) * 2;
    long i01 = i00 + fw * 2;
    for (int k = 0; k < 2; k++) {
        float p00 = d[i00 + k], p10 = d[i00 + 2 + k], p01 = d[i01 + k], p11 = d[i01 + 2 + k];
        float top = (p10 - p00) * fx + p00;
        float bot = (p11 - p01) * fx + p01;
        out[k] = (bot - top) * fy + top;
    }
}

// warp.rs:237-295 store of a faded edge sample: the stored sample scaled, trunc((double)(b * f) + 0.5)
template <class P> static inline P store_faded(float v, float f);
template <> inline uint8_t store_faded<uint8_t>(float v, float f) {
    float s = (float)sat_u8(v) * f;
    return (uint8_t)cvtt64d((double)s + 0.5);
}
template <> inline uint16_t store_faded<uint16_t>(float v, float f) {
    float s = (float)sat_u16(v) * f;
    return (uint16_t)cvtt64d((double)s + 0.5);
}
template <> inline float store_faded<float>(float v, float f) { return v * f; }

// warp.rs:477-580, the row kernel with `hold` (clip, then sample) or without it (transparent
// beyond one pixel, faded in the border band, with a coverage byte per pixel into `cov`).
template <class P>
static bool warp_run(P* out, const P* src, const float* flow, long fw, long fh, float amount, long w, long h, bool hold,
                     uint8_t* cov) {
    int32_t w1i = (int32_t)w - 1, h1i = (int32_t)h - 1;
    float rx = (float)((int32_t)fw - 1) / (float)w1i;
    float ry = (float)((int32_t)fh - 1) / (float)h1i;
    float sx = amount / rx, sy = amount / ry;
    float w1 = (float)w1i, h1 = (float)h1i;
    float wf = (float)w, hf = (float)h;
    int32_t fw1 = (int32_t)fw - 1, fh1 = (int32_t)fh - 1;
    bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
    for (long y = 0; y < h; y++) {
        float yf = (float)y;
        float fy = yf * ry;
        for (long x = 0; x < w; x++) {
            float xf = (float)x;
            float d[2];
            flow_bilerp(flow, fw, fw1, fh1, xf * rx, fy, d);
            float ix = d[0] * sx + xf;
            float iy = d[1] * sy + yf;
            P* px = out + (y * w + x) * 4;
            float s[4];
            if (hold) {
                if (!clip_to_rect(xf, yf, ix, iy, w1, h1) || std::isnan(ix) || std::isnan(iy)) {
                    bad = true;
                    for (int k = 0; k < 4; k++) px[k] = (P)0;
                    continue;
                }
                bilerp4(src, w, w1i, h1i, ix, iy, s);
                for (int k = 0; k < 4; k++) px[k] = store<P>(s[k]);
                continue;
            }
            // pixel_fade (warp.rs:519-561)
            uint8_t c;
            if (!(ix > -1.0f && ix < wf && iy > -1.0f && iy < hf)) {
                for (int k = 0; k < 4; k++) px[k] = (P)0;
                c = 0;
            } else if (ix >= 0.0f && ix <= w1 && iy >= 0.0f && iy <= h1) {
                bilerp4(src, w, w1i, h1i, ix, iy, s);
                for (int k = 0; k < 4; k++) px[k] = store<P>(s[k]);
                c = 255;
            } else {
                float f = ix < 0.0f ? ix + 1.0f : (ix <= w1 ? 1.0f : wf - ix);
                if (iy < 0.0f) f *= iy + 1.0f;
                else if (iy > h1) f *= hf - iy;
                if (!clip_to_rect(xf, yf, ix, iy, w1, h1) || std::isnan(ix) || std::isnan(iy)) {
                    bad = true;
                    for (int k = 0; k < 4; k++) px[k] = (P)0;
                    c = 0;
                } else {
                    bilerp4(src, w, w1i, h1i, ix, iy, s);
                    for (int k = 0; k < 4; k++) px[k] = store_faded<P>(s[k], f);
                    float cv = f * 255.0f + 0.5f;
                    c = cv <= 255.0f ? (uint8_t)cvtt64d((double)cv) : 255;
                }
            }
            if (cov) cov[y * w + x] = c;
        }
    }
    return !bad;
}

// warp.rs:693-786 (mode 1, is360 = false): `hold` set, or the fading warp whose coverage goes to
// `cov` (pre-filled with 255 by the caller: the copy early-outs leave it so)
static int inv_warp(void* out, const void* src, long w, long h, int fmt, const float* flow, long fw, long fh,
                    float amount, std::string& err, bool hold = true, uint8_t* cov = nullptr) {
    if (w < 2 || h < 2 || fw < 2 || fh < 2) { err = "inverse warp needs images and a flow of at least 2x2"; return TWX_E_MISMATCH; }
    float a = amount < 0.0f ? -amount : amount;
    bool copy_only = 0.005 > (double)a;
    if (!copy_only) {
        float r = (float)((int32_t)fw - 1) / (float)((int32_t)w - 1);
        float qv = (float)((double)amount / (double)r);
        if (qv < 0.0f) qv = -qv;
        copy_only = 0.0001 > (double)qv;
    }
    if (copy_only) { memcpy(out, src, image_bytes(w, h, fmt)); return TWX_OK; }
    bool ok;
    if (fmt == TWX_FMT_U8) ok = warp_run<uint8_t>((uint8_t*)out, (const uint8_t*)src, flow, fw, fh, amount, w, h, hold, cov);
    else if (fmt == TWX_FMT_U16) ok = warp_run<uint16_t>((uint16_t*)out, (const uint16_t*)src, flow, fw, fh, amount, w, h, hold, cov);
    else ok = warp_run<float>((float*)out, (const float*)src, flow, fw, fh, amount, w, h, hold, cov);
    if (!ok) { err = "inverse warp: a sample coordinate is NaN or never converges, which is not supported"; return TWX_E_UNSUPPORTED; }
    return TWX_OK;
}

// ------------------------------------------------------------------------------------------
// Blend, lerp, copy (blend.rs:55-257)

template <class T>
static void lerp_int(T* f, const T* s, float w, float scale) {
    auto diff = [&](int i) { return (float)((int32_t)f[i] - (int32_t)s[i]); };
    auto sf = [&](int i) { return (float)(int32_t)s[i]; };
    float a = diff(0) * w + sf(0);
    float n = a / scale;
    if (!(n > 0.0f)) { for (int k = 0; k < 4; k++) f[k] = (T)cvtt64f(0.0f); return; }
    float r = diff(1) * w + sf(1);
    float g = diff(2) * w + sf(2);
    float b = diff(3) * w + sf(3);
    if (n > 1.0f) { r = r / n; g = g / n; b = b / n; a = scale; }
    auto clamp = [&](float c) { return c > a ? a : (0.0f > c ? 0.0f : c); };
    r = clamp(r); g = clamp(g); b = clamp(b);
    f[1] = (T)cvtt64f(r + 0.5f);
    f[2] = (T)cvtt64f(g + 0.5f);
    f[3] = (T)cvtt64f(b + 0.5f);
    f[0] = (T)cvtt64f(a + 0.5f);
}
static void lerp_f32(float* f, const float* s, float w) {
    bool nan = false;
    for (int k = 0; k < 4; k++) nan |= std::isnan(f[k]) || std::isnan(s[k]);
    if (nan) {  // blend.rs:181-216: a NaN among the inputs, the arithmetic with the NaN rule
        float a = nr_lerp(s[0], f[0], w);
        if (!(a > 0.0f)) { f[0] = f[1] = f[2] = f[3] = 0.0f; return; }
        float r = nr_lerp(s[1], f[1], w);
        float g = nr_lerp(s[2], f[2], w);
        float b = nr_lerp(s[3], f[3], w);
        if (a > 1.0f) { r = nr_div(r, a); g = nr_div(g, a); b = nr_div(b, a); a = 1.0f; }
        f[1] = nr_add(r, 0.0f); f[2] = nr_add(g, 0.0f); f[3] = nr_add(b, 0.0f); f[0] = nr_add(a, 0.0f);
        return;
    }
    float a = (f[0] - s[0]) * w + s[0];
    if (!(a > 0.0f)) { f[0] = f[1] = f[2] = f[3] = 0.0f; return; }
    float r = (f[1] - s[1]) * w + s[1];
    float g = (f[2] - s[2]) * w + s[2];
    float b = (f[3] - s[3]) * w + s[3];
    if (a > 1.0f) { r = r / a; g = g / a; b = b / a; a = 1.0f; }
    f[1] = r + 0.0f; f[2] = g + 0.0f; f[3] = b + 0.0f; f[0] = a + 0.0f;
}

// blend.rs:145-178
static void lerp(void* first, const void* second, long w, long h, int fmt, float wt) {
    if (wt == 1.0f) return;
    if (wt == 0.0f) { memcpy(first, second, image_bytes(w, h, fmt)); return; }
    long n = w * h;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        if (fmt == TWX_FMT_U8) lerp_int<uint8_t>((uint8_t*)first + i * 4, (const uint8_t*)second + i * 4, wt, 255.0f);
        else if (fmt == TWX_FMT_U16) lerp_int<uint16_t>((uint16_t*)first + i * 4, (const uint16_t*)second + i * 4, wt, 65535.0f);
        else lerp_f32((float*)first + i * 4, (const float*)second + i * 4, wt);
    }
}

// blend.rs:55-114
static void blend(void* out, const void* a, const void* b, long w, long h, int fmt, float wt) {
    size_t bytes = image_bytes(w, h, fmt);
    if (0.005 > (double)std::fabs(wt)) { memcpy(out, a, bytes); return; }
    if (0.005 > std::fabs((double)wt - 1.0)) { memcpy(out, b, bytes); return; }
    if (wt < 0.0f || wt > 1.0f) {
        memcpy(out, a, bytes);
        lerp(out, b, w, h, fmt, (float)(1.0 - (double)wt));
        return;
    }
    long n = w * h * 4;
    if (fmt == TWX_FMT_U8) {
        uint8_t* o = (uint8_t*)out;
        const uint8_t *pa = (const uint8_t*)a, *pb = (const uint8_t*)b;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            float af = pa[i], bf = pb[i];
            o[i] = (uint8_t)cvtt64f(((bf - af) * wt + af) + 0.5f);
        }
    } else if (fmt == TWX_FMT_U16) {
        uint16_t* o = (uint16_t*)out;
        const uint16_t *pa = (const uint16_t*)a, *pb = (const uint16_t*)b;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            float af = pa[i], bf = pb[i];
            o[i] = (uint16_t)cvtt64f(((bf - af) * wt + af) + 0.5f);
        }
    } else {
        float* o = (float*)out;
        const float *pa = (const float*)a, *pb = (const float*)b;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; i++) {
            o[i] = ((pb[i] - pa[i]) * wt + pa[i]) + 0.0f;
            // blend.rs:105-110: a NaN result again with the NaN rule
            if (std::isnan(o[i])) o[i] = nr_add(nr_lerp(pa[i], pb[i], wt), 0.0f);
        }
    }
}

// ------------------------------------------------------------------------------------------
// The filters of the Weighted Blend step (filters.rs:54-712): the separable blur, the difference
// mask and its clean-up, the masked composite and the sharpening step. Finite samples only (the
// NaN rule of nanrule.rs is not restated).

static const float BLUR_RADIUS = 3.0f;   // filters.rs:55
static const float MASK_THRESHOLD = 0.4f;  // filters.rs:285

// filters.rs:71-95: w[i] = powf(2, -(i*i)/t^2) for i <= (int)radius, over sum * 2 - w[0]
static std::vector<float> blur_taps(float radius) {
    float t = radius / 3.0f;
    float t2 = t * t;
    int count = (std::isnan(radius) || radius < 0.0f) ? 1 : (int)radius + 1;  // Rust `as i32` saturates
    std::vector<float> w;
    float sum = 0.0f;
    for (int i = 0; i < count; i++) {
        float fi = (float)i;
        float x = fi * fi;
        x = x / t2;
        x = -x;
        float p = powf(2.0f, x);
        w.push_back(p);
        sum += p;
    }
    float denom = sum * 2.0f - w[0];
    for (auto& p : w) p /= denom;
    return w;
}

// filters.rs:190-236, one output pixel of the 1D filter. `get(j)` gives the four lanes of pixel j
// of a line of n pixels; the sentinel -1.0 stands for a tap outside the line, and in the edge
// formula a lane that equals it carries no weight.
template <class Get>
static inline void fconv_pixel(long i, long n, const float* taps, int count, const Get& get, bool left_first,
                               float out[4]) {
    const float S = -1.0f;
    const float sent[4] = {S, S, S, S};
    long left = std::min(n, (long)count);
    long mid = n > 2L * count ? n - 2L * count : 0;
    const float* c = get(i);
    float w0 = taps[0];
    float acc[4];
    for (int l = 0; l < 4; l++) acc[l] = c[l] * w0;
    if (i >= left && i < left + mid) {
        for (int k = count - 1; k >= 1; k--) {
            const float* r = get(i + k);
            const float* lf = get(i - k);
            for (int l = 0; l < 4; l++) acc[l] = acc[l] + (r[l] + lf[l]) * taps[k];
        }
        for (int l = 0; l < 4; l++) out[l] = acc[l];
        return;
    }
    float wsum[4] = {w0, w0, w0, w0};
    for (int k = count - 1; k >= 1; k--) {
        const float* r = i + k < n ? get(i + k) : sent;
        const float* lf = i >= k ? get(i - k) : sent;
        if (left_first) std::swap(r, lf);
        float wk = taps[k];
        for (int l = 0; l < 4; l++) {
            float wr = (r[l] != S ? 1.0f : 0.0f) * wk;
            wsum[l] = wsum[l] + wr;
            float rt = wr * r[l];
            float wl = (lf[l] != S ? 1.0f : 0.0f) * wk;
            float lt = lf[l] * wl;
            wsum[l] = wsum[l] + wl;
            acc[l] = (acc[l] + rt) + lt;
        }
    }
    for (int l = 0; l < 4; l++) out[l] = acc[l] / wsum[l];
}

// filters.rs:112-119: trunc(v + 0.5), clamped to 0..max, trunc(x + 0.5) again
static inline int32_t store_int(float v, float max) {
    int32_t t = cvtt(v + 0.5f);
    float f = (float)t;
    f = f > 0.0f ? f : 0.0f;
    f = f < max ? f : max;
    return cvtt(f + 0.5f);
}
template <class T> static inline T fstore(float v);
template <> inline uint8_t fstore<uint8_t>(float v) { return (uint8_t)store_int(v, 255.0f); }
template <> inline uint16_t fstore<uint16_t>(float v) { return (uint16_t)store_int(v, 65535.0f); }
template <> inline float fstore<float>(float v) { return v; }

// filters.rs:240-266: rows, then columns on a snapshot of the row result
template <class T>
static void blur_plane(T* d, long w, long h, const std::vector<float>& taps, bool above_first) {
    if (w == 0 || h == 0) return;
    int count = (int)taps.size();
#pragma omp parallel for schedule(static)
    for (long y = 0; y < h; y++) {
        std::vector<float> line(w * 4);
        for (long i = 0; i < w * 4; i++) line[i] = (float)d[y * w * 4 + i];
        auto get = [&](long j) { return &line[j * 4]; };
        for (long x = 0; x < w; x++) {
            float v[4];
            fconv_pixel(x, w, taps.data(), count, get, false, v);
            for (int l = 0; l < 4; l++) d[(y * w + x) * 4 + l] = fstore<T>(v[l]);
        }
    }
    std::vector<T> snap(d, d + w * h * 4);
#pragma omp parallel for schedule(static)
    for (long x = 0; x < w; x++) {
        std::vector<float> col(h * 4);
        for (long y = 0; y < h; y++)
            for (int l = 0; l < 4; l++) col[y * 4 + l] = (float)snap[(y * w + x) * 4 + l];
        auto get = [&](long j) { return &col[j * 4]; };
        for (long y = 0; y < h; y++) {
            float v[4];
            fconv_pixel(y, h, taps.data(), count, get, above_first, v);
            for (int l = 0; l < 4; l++) d[(y * w + x) * 4 + l] = fstore<T>(v[l]);
        }
    }
}

// filters.rs:272-283 (blur_with): a negative radius does nothing
static void blur_with(void* img, long w, long h, int fmt, float radius, bool above_first) {
    if (0.0 > (double)radius) return;
    std::vector<float> taps = blur_taps(radius);
    if (fmt == TWX_FMT_U8) blur_plane((uint8_t*)img, w, h, taps, above_first);
    else if (fmt == TWX_FMT_U16) blur_plane((uint16_t*)img, w, h, taps, above_first);
    else blur_plane((float*)img, w, h, taps, above_first);
}

// filters.rs:287-293: ((threshold * white) * threshold) * threshold * 3
static float mask_limit(int fmt, float threshold) {
    float white = fmt == TWX_FMT_U8 ? 255.0f : (fmt == TWX_FMT_U16 ? 65535.0f : 1.0f);
    return threshold * white * threshold * threshold * 3.0f;
}

// filters.rs:295-335: 255 where (|dR| + |dG|) + |dB| >= limit, differences in float
template <class T>
static void difference_rows(const T* a, const T* b, long n, float limit, uint8_t* m) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        auto d = [&](int c) { return std::fabs((float)a[i * 4 + c] - (float)b[i * 4 + c]); };
        float sum = (d(1) + d(2)) + d(3);
        m[i] = sum >= limit ? 0xff : 0;
    }
}
static void difference_mask(const void* warped, const void* original, long w, long h, int fmt, float threshold,
                            uint8_t* m) {
    float limit = mask_limit(fmt, threshold);
    long n = w * h;
    if (fmt == TWX_FMT_U8) difference_rows((const uint8_t*)warped, (const uint8_t*)original, n, limit, m);
    else if (fmt == TWX_FMT_U16) difference_rows((const uint16_t*)warped, (const uint16_t*)original, n, limit, m);
    else difference_rows((const float*)warped, (const float*)original, n, limit, m);
}

// filters.rs:345-350: the ring of eight around (x + 1, r1) from rows r0, r1, r2
static inline int ring_sum(const uint8_t* t, long w, long x, long r0, long r1, long r2) {
    auto at = [&](long y, long dx) { return (int)t[y * w + x + dx]; };
    return at(r0, 1) + at(r0, 0) + at(r0, 2) + at(r1, 0) + at(r1, 2) + at(r2, 0) + at(r2, 1) + at(r2, 2);
}

// filters.rs:352-375
static void mask_prune(uint8_t* m, long w, long h, int limit) {
    if (h < 5 || w < 3) return;
    std::vector<uint8_t> t(m, m + w * h);
    for (long y0 = 0; y0 < h - 4; y0++)
        for (long x = 0; x < w - 2; x++) {
            int sum = ring_sum(t.data(), w, x, y0, y0 + 2, y0 + 4);
            long dst = (y0 + 2) * w + x + 1;
            if (sum < limit && m[dst] != 0) m[dst] = 0;
        }
}

// filters.rs:384-399
static void mask_fill(uint8_t* m, long w, long h, int limit, uint8_t value) {
    if (h < 3 || w < 3) return;
    std::vector<uint8_t> t(m, m + w * h);
    for (long y0 = 0; y0 < h - 2; y0++)
        for (long x = 0; x < w - 2; x++) {
            int sum = ring_sum(t.data(), w, x, y0, y0 + 1, y0 + 2);
            long dst = (y0 + 1) * w + x + 1;
            if (limit < sum && m[dst] == 0) m[dst] = value;
        }
}

// filters.rs:403-417
static void mask_dilate_rows(uint8_t* m, long w, long h) {
    if (h < 3) return;
    std::vector<uint8_t> t(m, m + w * h);
    for (long y = 1; y < h - 1; y++)
        for (long x = 0; x < w; x++)
            if (t[(y - 1) * w + x] == 255 || t[(y + 1) * w + x] == 255) m[y * w + x] = 255;
}

// filters.rs:423-484: the rectangle tests in float, as the plugin runs them for any size
static void mask_replicate_border(uint8_t* m, long wl, long hl) {
    int64_t w = wl, h = hl;
    if (w == 0 || h == 0) return;
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
    for (int64_t y = 0; (float)y < fy0; y++)
        for (int64_t x = 0; x < w; x++) {
            int64_t sy = (int64_t)(int32_t)p3;
            if (outside(x, y) && ok(x, y) && ok(x, sy)) m[y * w + x] = m[sy * w + x];
        }
    for (int64_t y = (int64_t)(int32_t)fy1; y < h; y++)
        for (int64_t x = 0; x < w; x++)
            if (outside(x, y) && ok(x, y) && ok(x, iv14)) m[y * w + x] = m[iv14 * w + x];
    for (int64_t y = 0; y < h; y++) {
        int64_t sx = (int64_t)(int32_t)p2;
        for (int64_t x = 0; (float)x < fx0; x++)
            if (ok(x, y) && ok(sx, y)) m[y * w + x] = m[y * w + sx];
        int64_t first = (int64_t)(int32_t)fx1;
        int64_t base = (int64_t)(int32_t)p4;
        for (int64_t j = 0; first + j < w; j++)
            if (ok(base + j, y) && ok(base, y)) m[y * w + base + j] = m[y * w + base];
    }
}

// filters.rs:495-508
static void clean_mask(uint8_t* m, long w, long h) {
    mask_prune(m, w, h, 254);
    mask_prune(m, w, h, 509);
    mask_dilate_rows(m, w, h);
    for (int i = 0; i < 2; i++) {
        mask_fill(m, w, h, 254, 255);
        mask_replicate_border(m, w, h);
    }
    int limit = 254, value = 128;
    for (int i = 0; i < 2; i++) {
        mask_fill(m, w, h, limit, (uint8_t)value);
        mask_replicate_border(m, w, h);
        limit = (int)((double)limit * 0.5 + 0.5);
        value = (int)((double)value * 0.5 + 0.5);
    }
}

// filters.rs:532-607 and 611-616: mask 0 takes src, else f = m / 255, g = 1 - f (in double)
template <class T>
static void composite_rows(T* d, const T* s, long n, const uint8_t* m) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        uint8_t mv = m[i];
        for (int c = 0; c < 4; c++) {
            long k = i * 4 + c;
            if (mv == 0) { d[k] = s[k]; continue; }
            float f = (float)mv / 255.0f;
            float g = (float)(1.0 - (double)f);
            d[k] = (T)cvtt64f(((float)s[k] * g + (float)d[k] * f) + 0.5f);
        }
    }
}
static void composite_rows_f32(float* d, const float* s, long n, const uint8_t* m) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        uint8_t mv = m[i];
        for (int c = 0; c < 4; c++) {
            long k = i * 4 + c;
            if (mv == 0) { d[k] = s[k]; continue; }
            float f = (float)mv / 255.0f;
            float g = (float)(1.0 - (double)f);
            d[k] = (g * s[k] + f * d[k]) + 0.0f;
        }
    }
}
static void composite_masked(void* dest, const void* src, long w, long h, int fmt, const uint8_t* m) {
    long n = w * h;
    if (fmt == TWX_FMT_U8) composite_rows((uint8_t*)dest, (const uint8_t*)src, n, m);
    else if (fmt == TWX_FMT_U16) composite_rows((uint16_t*)dest, (const uint16_t*)src, n, m);
    else composite_rows_f32((float*)dest, (const float*)src, n, m);
}

// filters.rs:628-633: -(1 - |0.5 - t| * 2) * 0.5, the inner value narrowed and negated in float
static float sharpen_weight(float t) {
    double x = 1.0 - std::fabs(0.5 - (double)t) * 2.0;
    float neg = -(float)x;
    return (float)((double)neg * 0.5);
}

// filters.rs:649-656: mask of the change, blurred copy, unsharp lerp, inverted mask, composite
static void sharpen_step_with(void* frame, const void* original, long w, long h, int fmt, float t, bool above_first) {
    size_t bytes = image_bytes(w, h, fmt);
    std::vector<uint8_t> mask(w * h);
    difference_mask(frame, original, w, h, fmt, MASK_THRESHOLD, mask.data());
    clean_mask(mask.data(), w, h);
    std::vector<uint8_t> sharp((uint8_t*)frame, (uint8_t*)frame + bytes);
    blur_with(sharp.data(), w, h, fmt, BLUR_RADIUS, above_first);
    std::vector<uint8_t> base((uint8_t*)frame, (uint8_t*)frame + bytes);
    lerp(sharp.data(), base.data(), w, h, fmt, sharpen_weight(t));
    for (auto& v : mask) v = (uint8_t)(255 - v);
    composite_masked(frame, sharp.data(), w, h, fmt, mask.data());
}

// ------------------------------------------------------------------------------------------
// Smart Blend (smartblend.rs:72-330) and the push-pull hole fill (holefill.rs:145-491).
// Finite samples only (the NaN rule of nanrule.rs is not restated).

// smartblend.rs:84-150: the stores of the main path and of the "alpha is 0" fallback
template <class P> struct SmartPx;
template <> struct SmartPx<uint8_t> {
    static constexpr float WHITE = 255.0f;
    static uint8_t plain(float x) {
        float y = x + 0.5f;
        return 0.0f > y ? 0 : (y > 255.0f ? 255 : (uint8_t)cvtt64f(y));
    }
    static uint8_t raw(float x) { return (uint8_t)cvtt64f(x); }
};
template <> struct SmartPx<uint16_t> {
    static constexpr float WHITE = 65535.0f;
    static uint16_t plain(float x) {
        float y = x + 0.5f;
        return 0.0f > y ? 0 : (y > 65535.0f ? 65535 : (uint16_t)cvtt64f(y));
    }
    static uint16_t raw(float x) { return (uint16_t)cvtt64f(x); }
};
template <> struct SmartPx<float> {
    static constexpr float WHITE = 1.0f;
    static float plain(float x) { return x + 0.0f; }
    static float raw(float x) { return x; }
};

// smartblend.rs:167-178: `a` over `b`, premultiplied, alpha in units of white
static inline void over(const float a[4], const float b[4], float white, float o[4]) {
    if (a[0] == 0.0f) { for (int k = 0; k < 4; k++) o[k] = b[k]; return; }
    if (a[0] == white) { for (int k = 0; k < 4; k++) o[k] = a[k]; return; }
    float ia = 1.0f - a[0] / white;
    o[0] = (white - b[0]) * (a[0] / white) + b[0];
    for (int k = 1; k < 4; k++) o[k] = b[k] * ia + a[k];
}

// over() with the NaN rule's operations
static inline void over_nr(const float a[4], const float b[4], float white, float o[4]) {
    if (a[0] == 0.0f) { for (int k = 0; k < 4; k++) o[k] = b[k]; return; }
    if (a[0] == white) { for (int k = 0; k < 4; k++) o[k] = a[k]; return; }
    float ia = nr_sub(1.0f, nr_div(a[0], white));
    o[0] = nr_add(nr_mul(nr_sub(white, b[0]), nr_div(a[0], white)), b[0]);
    for (int k = 1; k < 4; k++) o[k] = nr_add(nr_mul(b[k], ia), a[k]);
}

// smartblend.rs:193-212: returns true for the raw store of the fallback colour. A NaN in the
// composite: the same arithmetic again with the NaN rule's operations (the references were
// rendered under it, nanrule.rs:14-21).
static inline bool smart_pixel(const float p1[4], const float p2[4], float w, float white, float v[4]) {
    float c1[4], c2[4];
    over(p1, p2, white, c1);
    over(p2, p1, white, c2);
    float amax = p1[0] > p2[0] ? p1[0] : p2[0];
    bool nan = false;
    for (int k = 0; k < 4; k++) {
        v[k] = (c2[k] - c1[k]) * w + c1[k];
        nan |= std::isnan(v[k]);
    }
    if (nan) {
        over_nr(p1, p2, white, c1);
        over_nr(p2, p1, white, c2);
        for (int k = 0; k < 4; k++) v[k] = nr_lerp(c1[k], c2[k], w);
        if (v[0] == amax) return false;
        if (v[0] == 0.0f) {
            const float* c = c1[0] > c2[0] ? c1 : c2;
            for (int k = 0; k < 4; k++) v[k] = c[k];
            return true;
        }
        float alpha = v[0];
        for (int k = 1; k < 4; k++) v[k] = nr_div(nr_mul(v[k], amax), alpha);
        v[0] = amax;
        return false;
    }
    if (v[0] == amax) return false;
    if (v[0] == 0.0f) {
        const float* c = c1[0] > c2[0] ? c1 : c2;
        for (int k = 0; k < 4; k++) v[k] = c[k];
        return true;
    }
    float alpha = v[0];
    for (int k = 1; k < 4; k++) v[k] = (v[k] * amax) / alpha;
    v[0] = amax;
    return false;
}

template <class P>
static void smart_rows(P* out, const P* p1, const P* p2, long n, float w) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) {
        float a[4], b[4], v[4];
        for (int k = 0; k < 4; k++) { a[k] = (float)p1[i * 4 + k]; b[k] = (float)p2[i * 4 + k]; }
        bool raw = smart_pixel(a, b, w, SmartPx<P>::WHITE, v);
        for (int k = 0; k < 4; k++) out[i * 4 + k] = raw ? SmartPx<P>::raw(v[k]) : SmartPx<P>::plain(v[k]);
    }
}

// smartblend.rs:290-293: out = smart_blend(p1, p2, w)
static void smart_blend(void* out, const void* p1, const void* p2, long w, long h, int fmt, float wt) {
    long n = w * h;
    if (fmt == TWX_FMT_U8) smart_rows((uint8_t*)out, (const uint8_t*)p1, (const uint8_t*)p2, n, wt);
    else if (fmt == TWX_FMT_U16) smart_rows((uint16_t*)out, (const uint16_t*)p1, (const uint16_t*)p2, n, wt);
    else smart_rows((float*)out, (const float*)p1, (const float*)p2, n, wt);
}

// holefill.rs:145-176: level sizes of the fill pyramid; the push starts at `top`
struct HoleGeometry {
    std::vector<std::pair<long, long>> dims;
    int top;
};
static HoleGeometry hole_geometry(long w, long h) {
    int64_t a = w, b = h;
    int count = 1;
    for (;;) {
        a /= 2;
        b /= 2;
        if (a < 2) a = 2;
        if (b < 2) b = 2;
        count++;
        if (!(a > 1 && b > 1 && (a != 2 || b != 2))) break;
    }
    HoleGeometry g;
    long cw = w, ch = h;
    for (int i = 0; i <= count; i++) {
        g.dims.push_back({cw, ch});
        cw = std::max(cw / 2, 2L);
        ch = std::max(ch / 2, 2L);
    }
    g.top = count - 1;
    return g;
}

static const float GAUSS_FILL_RADIUS = 5.0f;  // holefill.rs:344

// holefill.rs:347-363: sol = value / weight where the weight is not 0
static void hole_combine(std::vector<float>& sol, const std::vector<float>& v, const std::vector<float>& wt) {
    for (size_t i = 0; i < sol.size(); i++)
        if (wt[i] != 0.0f) sol[i] = v[i] / wt[i];
}

// holefill.rs:214-336: the weight pyramid of the mask, then the pull and push of one plane
struct HoleFiller {
    HoleGeometry geo;
    std::vector<uint8_t> hole;
    std::vector<std::vector<float>> wts;
    bool above_first;
    std::vector<float> k5;

    void blur5(std::vector<float>& d, long w, long h) const { convolve(d.data(), w, h, 1, BINOMIAL5, 3, above_first); }
    void gauss(std::vector<float>& d, long w, long h) const {
        convolve(d.data(), w, h, 1, k5.data(), (int)k5.size(), above_first);
    }

    void fill(float* plane) const {
        long w = geo.dims[0].first, h = geo.dims[0].second;
        std::vector<std::vector<float>> vals;
        std::vector<float> v0(plane, plane + w * h);
        for (long i = 0; i < w * h; i++)
            if (hole[i]) v0[i] = 0.0f;
        vals.push_back(v0);
        for (int k = 1; k <= geo.top; k++) {
            auto [pw, ph] = geo.dims[k - 1];
            auto [cw, ch] = geo.dims[k];
            std::vector<float> tmp = vals[k - 1];
            blur5(tmp, pw, ph);
            vals.push_back(resample1(tmp.data(), pw, ph, cw, ch));
        }
        auto [tw, th] = geo.dims[geo.top];
        std::vector<float> sol(tw * th, 0.0f);
        for (int level = geo.top; level >= 0; level--) {
            auto [lw, lh] = geo.dims[level];
            hole_combine(sol, vals[level], wts[level]);
            gauss(sol, lw, lh);
            if (level == 0) {
                hole_combine(sol, vals[0], wts[0]);
            } else {
                auto [nw, nh] = geo.dims[level - 1];
                sol = resample1(sol.data(), lw, lh, nw, nh);
            }
        }
        memcpy(plane, sol.data(), sizeof(float) * w * h);
    }
};

// holefill.rs:90-91 and 128-134: a u8 mask value m is a hole when !(eps > (double)m)
static const double HOLE_U8_EPS = [] { uint64_t b = 0x20202020202020fcull; double d; memcpy(&d, &b, 8); return d; }();

// holefill.rs:262-296. Returns 0 with `has` false when the mask has no hole (the fill is then the
// identity), an error for a hole in an image with a side of 1.
static int hole_filler(long w, long h, const uint8_t* mask, bool above_first, HoleFiller& f, bool& has,
                       std::string& err) {
    f.hole.resize(w * h);
    has = false;
    for (long i = 0; i < w * h; i++) {
        f.hole[i] = !(HOLE_U8_EPS > (double)mask[i]);
        has = has || f.hole[i];
    }
    if (!has) return TWX_OK;
    if (w < 2 || h < 2) { err = "hole fill of an image with a side of 1"; return TWX_E_UNSUPPORTED; }
    f.above_first = above_first;
    f.k5 = gaussian_kernel(GAUSS_FILL_RADIUS);
    f.geo = hole_geometry(w, h);
    std::vector<float> w0(w * h);
    for (long i = 0; i < w * h; i++) w0[i] = f.hole[i] ? 0.0f : 1.0f;
    f.wts.push_back(w0);
    for (int k = 1; k <= f.geo.top; k++) {
        auto [pw, ph] = f.geo.dims[k - 1];
        auto [cw, ch] = f.geo.dims[k];
        std::vector<float> tmp = f.wts[k - 1];
        f.blur5(tmp, pw, ph);
        f.wts.push_back(resample1(tmp.data(), pw, ph, cw, ch));
    }
    return TWX_OK;
}

// holefill.rs:366-386: write back of the integer depths
template <class T> static inline T hole_store(float x, float white) {
    float y = x + 0.5f;
    return 0.0f > y ? (T)0 : (y > white ? (T)white : (T)cvtt64f(y));
}

// holefill.rs:444-481: every channel of the engine image, alpha first
static int fill_holes(void* img, long w, long h, int fmt, const uint8_t* holes, bool above_first, std::string& err) {
    HoleFiller f;
    bool has;
    int st = hole_filler(w, h, holes, above_first, f, has, err);
    if (st || !has) return st;
    long n = w * h;
    std::vector<float> plane(n);
#pragma omp parallel for schedule(static) firstprivate(plane)
    for (int c = 0; c < 4; c++) {
        for (long i = 0; i < n; i++) {
            if (fmt == TWX_FMT_U8) plane[i] = (float)((uint8_t*)img)[i * 4 + c];
            else if (fmt == TWX_FMT_U16) plane[i] = (float)((uint16_t*)img)[i * 4 + c];
            else plane[i] = ((float*)img)[i * 4 + c];
        }
        f.fill(plane.data());
        for (long i = 0; i < n; i++) {
            if (fmt == TWX_FMT_U8) ((uint8_t*)img)[i * 4 + c] = hole_store<uint8_t>(plane[i], 255.0f);
            else if (fmt == TWX_FMT_U16) ((uint16_t*)img)[i * 4 + c] = hole_store<uint16_t>(plane[i], 65535.0f);
            else ((float*)img)[i * 4 + c] = plane[i];
        }
    }
    return TWX_OK;
}

// smartblend.rs:304-316: coverage max, invert, hole fill
static int smart_finish(void* out, long w, long h, int fmt, std::vector<uint8_t>& cov1, const std::vector<uint8_t>& cov2,
                        bool above_first, std::string& err) {
    for (size_t i = 0; i < cov1.size(); i++) cov1[i] = (uint8_t)(255 - std::max(cov1[i], cov2[i]));
    return fill_holes(out, w, h, fmt, cov1.data(), above_first, err);
}

// The compositions of warp code 1 (smartblend.rs:323-419, compose/vector.rs:222-247): the fading
// warps with coverage, the composite (Nearest: the one warp), the Weighted Blend sharpening, the
// coverage finish and the hole fill. `flow` is the one flow of Nearest; interp 1 and 2 read both.
static int compose_smart(int interp, float wa, const void* a, const void* b, long w, long h, int fmt, const float* f_ab,
                         const float* f_ba, long fw, long fh, bool above_first, void* out, std::string& err) {
    size_t bytes = image_bytes(w, h, fmt);
    float frac = 1.0f - wa;
    int st;
    if (interp == 0) {
        std::vector<uint8_t> cov(w * h, 255);
        if (!(0.5f > wa)) st = inv_warp(out, a, w, h, fmt, f_ba, fw, fh, 1.0f - wa, err, false, cov.data());
        else st = inv_warp(out, b, w, h, fmt, f_ab, fw, fh, wa, err, false, cov.data());
        if (st) return st;
        for (auto& c : cov) c = (uint8_t)(255 - c);
        return fill_holes(out, w, h, fmt, cov.data(), above_first, err);
    }
    std::vector<uint8_t> ta(bytes, 0), tb(bytes, 0);
    std::vector<uint8_t> cov1(w * h, 255), cov2(w * h, 255);
    if ((st = inv_warp(ta.data(), a, w, h, fmt, f_ba, fw, fh, frac, err, false, cov1.data()))) return st;
    if ((st = inv_warp(tb.data(), b, w, h, fmt, f_ab, fw, fh, wa, err, false, cov2.data()))) return st;
    if (interp == 2) {
        sharpen_step_with(ta.data(), a, w, h, fmt, frac, above_first);
        sharpen_step_with(tb.data(), b, w, h, fmt, wa, above_first);
    }
    smart_blend(out, ta.data(), tb.data(), w, h, fmt, frac);
    return smart_finish(out, w, h, fmt, cov1, cov2, above_first, err);
}

// ------------------------------------------------------------------------------------------
// Forward warp (forward.rs, forward/mesh.rs, forward/raster.rs, forward/pyramid.rs): the displaced
// vertex grid, the triangle scan converter, the running mean per pixel in triangle order, the
// textured (mip pyramid) and the Gouraud (vertex colour) variants. Finite colours only: the NaN
// rule of the colour arithmetic (nanrule.rs) is not restated.

// mesh.rs:11-31
struct FwGrid {
    long w, h;
    float rx, ry, sx, sy;
};
static FwGrid fw_grid(long w, long h, long fw, long fh, float amount) {
    FwGrid g;
    g.w = w;
    g.h = h;
    g.rx = (float)((int32_t)fw - 1) / (float)((int32_t)w - 1);
    g.ry = (float)((int32_t)fh - 1) / (float)((int32_t)h - 1);
    g.sx = amount / g.rx;
    g.sy = amount / g.ry;
    return g;
}

// mesh.rs:33-52: x' = fx * sx + c, y' = fy * sy + r; rows 1.. and columns 1.. get the tie
// breaking epsilon as a double addition, narrowed
static std::vector<float> fw_positions(const FwGrid& g, const float* flow, long fw, long fh) {
    std::vector<float> pos((size_t)g.w * g.h * 2);
    for (long r = 0; r < g.h; r++) {
        float fy = (float)(int32_t)r * g.ry, rf = (float)(int32_t)r;
        for (long c = 0; c < g.w; c++) {
            float cf = (float)(int32_t)c, f[2];
            flow_bilerp(flow, fw, (int32_t)fw - 1, (int32_t)fh - 1, cf * g.rx, fy, f);
            float px = f[0] * g.sx + cf, py = f[1] * g.sy + rf;
            if (r != 0 && c >= 1) {
                px = (float)((double)px + 0.0001);
                py = (float)((double)py + 0.0001);
            }
            pos[(r * g.w + c) * 2] = px;
            pos[(r * g.w + c) * 2 + 1] = py;
        }
    }
    return pos;
}

// mesh.rs:54-88: the mip level of vertex (c, r) from the distances to its 8 displaced neighbours
static float fw_mip_level(const FwGrid& g, const float* flow, long fw, long fh, int32_t c, int32_t r) {
    int32_t w = (int32_t)g.w, h = (int32_t)g.h;
    int32_t up = r - 1 < 0 ? 1 : r - 1, down = h - 1 < r + 1 ? h - 2 : r + 1;
    int32_t left = c - 1 < 0 ? 1 : c - 1, right = w - 1 < c + 1 ? w - 2 : c + 1;
    float f[2];
    flow_bilerp(flow, fw, (int32_t)fw - 1, (int32_t)fh - 1, (float)c * g.rx, (float)r * g.ry, f);
    float xc = f[0] * g.sx + (float)c, yc = f[1] * g.sy + (float)r;
    auto len = [&](int32_t ci, int32_t ri) {
        float n[2];
        flow_bilerp(flow, fw, (int32_t)fw - 1, (int32_t)fh - 1, (float)ci * g.rx, (float)ri * g.ry, n);
        float dx = (n[0] * g.sx + (float)ci) - xc, dy = (n[1] * g.sy + (float)ri) - yc;
        return sqrtf(dy * dy + dx * dx);
    };
    float ru = len(right, up), lu = len(left, up), ld = len(left, down), rd = len(right, down);
    float u = len(c, up), dn = len(c, down), l = len(left, r), rr = len(right, r);
    float diag = ((ru + lu) + ld) + rd;
    double t = (double)diag * 0.7071067811865476;
    t += (double)u;
    t += (double)dn;
    t += (double)l;
    t += (double)rr;
    float m = (float)((double)(float)t * 0.125);
    float inv_ln2 = 1.0f / logf(2.0f);
    return -(logf(m) * inv_ln2) + 0.0f;
}

// pyramid.rs:20-35
static int fw_level_count(long w, long h) {
    int n = 1;
    int32_t a = (int32_t)std::max(w, 2L), b = (int32_t)std::max(h, 2L);
    for (;;) {
        a /= 2;
        if (a < 2) break;
        b /= 2;
        if (b < 2) break;
        n++;
    }
    return n;
}

// pyramid.rs:37-95: levels 1.. (blur5 of the previous level with the filters.rs edge rule, rows
// then columns, right tap first; then the corner aligned resample with the warp stores)
struct FwPyramid {
    int count = 1;
    std::vector<std::vector<uint8_t>> levels;  // levels 1..count-1
    std::vector<std::pair<long, long>> dims;   // dims[0] = the source
};
template <class P>
static void fw_resample(P* dst, long dw, long dh, const P* src, long sw, long sh) {
    float sx = ((float)(int32_t)sw - 1.0f) / (float)((int32_t)dw - 1);
    float sy = ((float)(int32_t)sh - 1.0f) / (float)((int32_t)dh - 1);
    for (long y = 0; y < dh; y++) {
        float fy = (float)(int32_t)y * sy;
        for (long x = 0; x < dw; x++) {
            float v[4];
            bilerp4(src, sw, (int32_t)sw - 1, (int32_t)sh - 1, (float)(int32_t)x * sx, fy, v);
            for (int k = 0; k < 4; k++) dst[(y * dw + x) * 4 + k] = store<P>(v[k]);
        }
    }
}
template <class P>
static FwPyramid fw_pyramid(const P* src, long w, long h) {
    FwPyramid p;
    p.count = fw_level_count(w, h);
    p.dims.push_back({w, h});
    long cw = std::max(w, 2L), ch = std::max(h, 2L);
    const std::vector<float> taps(BINOMIAL5, BINOMIAL5 + 3);
    for (int i = 1; i < p.count; i++) {
        const P* prev = i == 1 ? src : (const P*)p.levels[i - 2].data();
        long pw = p.dims[i - 1].first, ph = p.dims[i - 1].second;
        std::vector<P> tmp(prev, prev + pw * ph * 4);
        blur_plane(tmp.data(), pw, ph, taps, false);
        cw = std::max(cw / 2, 2L);
        ch = std::max(ch / 2, 2L);
        std::vector<uint8_t> lvl((size_t)cw * ch * 4 * sizeof(P));
        fw_resample((P*)lvl.data(), cw, ch, tmp.data(), pw, ph);
        p.levels.push_back(std::move(lvl));
        p.dims.push_back({cw, ch});
    }
    return p;
}
// pyramid.rs:97-110: one bilinear tap at (w1 * u, h1 * v)
template <class P>
static void fw_tap(const P* img, long w, long h, float u, float v, float out[4]) {
    int32_t w1 = (int32_t)w - 1, h1 = (int32_t)h - 1;
    bilerp4(img, w, w1, h1, (float)w1 * u, (float)h1 * v, out);
}
// pyramid.rs:68-92: the blend of the two levels around d
template <class P>
static void fw_sample(const FwPyramid& p, const P* src, float u, float v, float d, float out[4]) {
    auto level = [&](int i) { return i == 0 ? src : (const P*)p.levels[i - 1].data(); };
    if (0.0f < d) {
        if (d < (float)(p.count - 1)) {
            int32_t i = cvtt(d);
            float frac = d - (float)i;
            float a[4], b[4];
            fw_tap(level(i), p.dims[i].first, p.dims[i].second, u, v, a);
            fw_tap(level(i + 1), p.dims[i + 1].first, p.dims[i + 1].second, u, v, b);
            for (int k = 0; k < 4; k++) out[k] = (b[k] - a[k]) * frac + a[k];
            return;
        }
        int last = p.count - 1;
        fw_tap(level(last), p.dims[last].first, p.dims[last].second, u, v, out);
        return;
    }
    fw_tap(src, p.dims[0].first, p.dims[0].second, u, v, out);
}

// raster.rs: a vertex with its position and payload (u, v, d of the textured variant; the four
// colours of the Gouraud variant, whose two passes over the same pixels are carried together)
struct FwVtx {
    float x, y;
    float p[4];
};
struct FwEdge {
    float ay = 0, by = 0, x = 0, dx = 0;
    float p[4] = {0, 0, 0, 0}, dp[4] = {0, 0, 0, 0};
};

// raster.rs:59-86
static inline void fw_edge_init(const FwVtx& a, const FwVtx& b, float y, FwEdge e[2], uint32_t& count) {
    if (a.y <= y && y < b.y) {
        int slot = count != 0 ? 1 : 0;
        count++;
        float dx = b.x - a.x;
        float inv = 1.0f / (b.y - a.y);
        float t = ((y - 1.0f) - a.y) * inv;
        FwEdge& E = e[slot];
        E.ay = a.y;
        E.by = b.y;
        E.x = dx * t + a.x;
        E.dx = dx * inv;
        for (int k = 0; k < 4; k++) {
            float dk = b.p[k] - a.p[k];
            E.p[k] = dk * t + a.p[k];
            E.dp[k] = dk * inv;
        }
    }
}

// raster.rs:88-112: ties keep the given order except v2.y == v0.y < v1.y -> v2, v0, v1
static inline void fw_sort_by_y(const FwVtx v[3], FwVtx o[3]) {
    float y0 = v[0].y, y1 = v[1].y, y2 = v[2].y;
    int a, b, c;
    if (y1 < y0) {
        if (y2 < y0) {
            if (y2 < y1) { a = 2; b = 1; c = 0; } else { a = 1; b = 2; c = 0; }
        } else { a = 1; b = 0; c = 2; }
    } else if (y2 < y1) {
        if (y0 < y2) { a = 0; b = 2; c = 1; } else { a = 2; b = 0; c = 1; }
    } else { a = 0; b = 1; c = 2; }
    o[0] = v[a];
    o[1] = v[b];
    o[2] = v[c];
}

// raster.rs:114-151: one span of the main pass
template <class Hit>
static inline void fw_span(const FwEdge& l, const FwEdge& r, float y, int32_t w1, const Hit& hit) {
    int32_t x0 = std::max(cvtt(ceilf(l.x)), 0);
    int32_t x1 = std::min(cvtt(floorf(r.x)), w1);
    float wid = r.x - l.x;
    if (x0 > x1) return;
    if (!((double)wid > 1e-08)) return;
    int32_t yi = cvtt(y);
    float p[4];
    if (x0 == x1) {
        float t = ((float)x0 - l.x) * (1.0f / wid);
        for (int k = 0; k < 4; k++) p[k] = (r.p[k] - l.p[k]) * t + l.p[k];
        hit(x0, yi, p);
        return;
    }
    float inv = 1.0f / wid, dp[4];
    float t = ((float)x0 - l.x) * inv;
    for (int k = 0; k < 4; k++) {
        dp[k] = r.p[k] - l.p[k];
        p[k] = dp[k] * t + l.p[k];
    }
    int32_t x = x0;
    while (x < x1) {
        hit(x, yi, p);
        for (int k = 0; k < 4; k++) p[k] = p[k] + dp[k] * inv;
        x++;
    }
    hit(x, yi, p);
}

// raster.rs:153-300: rows y_start..=y_end with closed spans, then the degenerate last row pass.
// Returns false when that pass would write outside the image (nothing is written for it).
template <class Hit>
static bool fw_triangle(const FwVtx tri[3], int32_t w1, int32_t h1, const Hit& hit) {
    float w1f = (float)w1, h1f = (float)h1;
    const FwVtx* t = tri;
    bool ok = (0.0f <= t[0].x || 0.0f <= t[1].x || 0.0f <= t[2].x) && (0.0f <= t[0].y || 0.0f <= t[1].y || 0.0f <= t[2].y) &&
              (t[0].x <= w1f || t[1].x <= w1f || t[2].x <= w1f) && (t[0].y <= h1f || t[1].y <= h1f || t[2].y <= h1f);
    if (!ok) return true;
    FwVtx vs[3];
    fw_sort_by_y(tri, vs);
    float y = ceilf(vs[0].y), y_end = floorf(vs[2].y);
    if (y < 0.0f) y = 0.0f;
    if (h1f < y_end) y_end = h1f;
    if (y <= y_end) {
        int vi = 0;
        uint32_t count = 0;
        FwEdge e[2];
        for (;;) {
            if (vi < 3) {
                for (;;) {
                    if (y < vs[vi].y) break;
                    if (count == 1) {
                        if (e[0].ay <= y && e[0].by <= y) count = 0;
                    } else if (count == 2) {
                        if (y < e[1].ay || y < e[1].by) {
                            if (!(y < e[0].ay || y < e[0].by)) {
                                e[0] = e[1];
                                count = 1;
                            }
                        } else if (e[0].ay <= y && e[0].by <= y) {
                            count = 0;
                        } else {
                            count = 1;
                        }
                    }
                    static const int NEXT[3][2] = {{1, 2}, {2, 0}, {0, 1}};
                    fw_edge_init(vs[vi], vs[NEXT[vi][0]], y, e, count);
                    fw_edge_init(vs[vi], vs[NEXT[vi][1]], y, e, count);
                    vi++;
                    if (vi >= 3) break;
                }
            }
            for (int k = 0; k < 2; k++) {
                e[k].x += e[k].dx;
                for (int q = 0; q < 4; q++) e[k].p[q] = e[k].p[q] + e[k].dp[q];
            }
            if (count == 2) {
                if (e[1].x < e[0].x) fw_span(e[1], e[0], y, w1, hit);
                else fw_span(e[0], e[1], y, w1, hit);
            }
            y += 1.0f;
            if (!(y <= y_end)) break;
        }
    }
    const FwVtx& mid = vs[1];
    const FwVtx& bot = vs[2];
    bool inside = true;
    if (mid.y == h1f && mid.y == bot.y) {
        const FwVtx& l = bot.x <= mid.x ? bot : mid;
        const FwVtx& r = bot.x <= mid.x ? mid : bot;
        float yy = l.y, yt = (float)cvtt(yy);
        if (0.0f <= yt && yt <= h1f && !((double)(yy - yt) > 0.01)) {
            int32_t x0 = std::max(cvtt(ceilf(l.x)), 0);
            int32_t x1 = std::min(cvtt(floorf(r.x)), w1);
            float wid = r.x - l.x;
            if ((double)wid > 1e-05) {
                float inv = 1.0f / wid;
                int32_t yi = cvtt(yt);
                auto emit = [&](int32_t x, const float* p) {
                    if (x < 0 || x > w1 || yi < 0 || yi > h1) inside = false;
                    else hit(x, yi, p);
                };
                float p[4], dp[4];
                if (x0 == x1) {
                    float tt = ((float)x0 - l.x) * inv;
                    for (int k = 0; k < 4; k++) p[k] = (r.p[k] - l.p[k]) * tt + l.p[k];
                    emit(x0, p);
                } else {
                    float tt = ((float)x0 - l.x) * inv;
                    for (int k = 0; k < 4; k++) {
                        dp[k] = r.p[k] - l.p[k];
                        p[k] = dp[k] * tt + l.p[k];
                    }
                    int32_t x = x0;
                    while (x < x1) {
                        emit(x, p);
                        for (int k = 0; k < 4; k++) p[k] = p[k] + dp[k] * inv;
                        x++;
                    }
                    emit(x, p);
                }
            }
        }
    }
    return inside;
}

// forward.rs:104-128: the running mean with the saturating u8 hit counter
static inline void fw_fold(float* acc, uint8_t& n, const float c[4]) {
    if (n == 0) {
        for (int k = 0; k < 4; k++) acc[k] = c[k] + 0.0f;
        n = 1;
    } else if (n == 0xff) {
        for (int k = 0; k < 4; k++) acc[k] = (acc[k] * 255.0f + c[k]) * 0.00390625f + 0.0f;
    } else {
        float nf = (float)n;
        float r = (float)(1.0 / ((double)nf + 1.0));
        for (int k = 0; k < 4; k++) acc[k] = (nf * acc[k] + c[k]) * r + 0.0f;
        n = (uint8_t)(n + 1);
    }
}

// forward.rs:183-195
static bool fw_flow_is_zero(const float* flow, long n) {
    for (long i = 0; i < n; i++) {
        float v = flow[i];
        float a = v < 0.0f ? -v : v;
        if (0.0001 < (double)a) return false;
    }
    return true;
}

template <class P> static inline P fw_convert(float a);
template <> inline uint8_t fw_convert<uint8_t>(float a) { return (uint8_t)cvtt64f(a + 0.5f); }
template <> inline uint16_t fw_convert<uint16_t>(float a) { return (uint16_t)cvtt64f(a + 0.5f); }
template <> inline float fw_convert<float>(float a) { return a + 0.0f; }

// forward.rs:229-320 (textured, pyramid on) and :327-373 (Gouraud): cells row-major, T1 = (v00,
// v01, v10) then T2 = (v11, v01, v10); the running mean, the conversion and the 255/0 coverage
template <class P>
static int fw_run(P* out, const P* src, long w, long h, const float* flow, long fw, long fh, float amount, bool gouraud,
                  uint8_t* cov, std::string& err) {
    FwGrid g = fw_grid(w, h, fw, fh, amount);
    std::vector<float> pos = fw_positions(g, flow, fw, fh);
    int32_t w1 = (int32_t)w - 1, h1 = (int32_t)h - 1;
    float w1f = (float)w1, h1f = (float)h1;
    std::vector<float> acc((size_t)w * h * 4, 0.0f);
    std::vector<uint8_t> cnt((size_t)w * h, 0);
    FwPyramid pyr;
    std::vector<float> mips;
    if (!gouraud) {
        pyr = fw_pyramid(src, w, h);
        if (pyr.count > 1) {
            mips.resize((size_t)w * h);
            for (long r = 0; r < h; r++)
                for (long c = 0; c < w; c++) mips[r * w + c] = fw_mip_level(g, flow, fw, fh, (int32_t)c, (int32_t)r);
        }
    }
    auto vtx = [&](long r, long c) {
        FwVtx v;
        v.x = pos[(r * w + c) * 2];
        v.y = pos[(r * w + c) * 2 + 1];
        if (gouraud) {
            for (int k = 0; k < 4; k++) v.p[k] = to_f(src[(r * w + c) * 4 + k]);
        } else {
            v.p[0] = (float)(int32_t)c / w1f;
            v.p[1] = (float)(int32_t)r / h1f;
            v.p[2] = mips.empty() ? 0.0f : mips[r * w + c];
            v.p[3] = 0.0f;
        }
        return v;
    };
    auto hit = [&](int32_t x, int32_t y, const float* p) {
        long i = (long)y * w + x;
        float c[4];
        if (gouraud) {
            for (int k = 0; k < 4; k++) c[k] = p[k];
        } else if (0.0f < p[2]) {
            fw_sample(pyr, src, p[0], p[1], p[2], c);
        } else {
            fw_tap(src, w, h, p[0], p[1], c);
        }
        fw_fold(&acc[i * 4], cnt[i], c);
    };
    bool inside = true;
    for (long r = 0; r < h - 1; r++) {
        for (long c = 0; c < w - 1; c++) {
            FwVtx v00 = vtx(r, c), v01 = vtx(r, c + 1), v10 = vtx(r + 1, c), v11 = vtx(r + 1, c + 1);
            FwVtx t1[3] = {v00, v01, v10}, t2[3] = {v11, v01, v10};
            inside = fw_triangle(t1, w1, h1, hit) && inside;
            inside = fw_triangle(t2, w1, h1, hit) && inside;
        }
    }
    if (!inside) {
        err = "forward warp: the degenerate last row pass of the rasteriser would write outside the image, which is not "
              "supported";
        return TWX_E_UNSUPPORTED;
    }
    for (long i = 0; i < w * h * 4; i++) out[i] = fw_convert<P>(acc[i]);
    if (cov)
        for (long i = 0; i < w * h; i++) cov[i] = cnt[i] != 0 ? 0xff : 0;
    return TWX_OK;
}

static int forward_warp(void* out, const void* src, long w, long h, int fmt, const float* flow, long fw, long fh,
                        float amount, bool gouraud, uint8_t* cov, std::string& err) {
    if (w < 2 || h < 2 || fw < 2 || fh < 2) {
        err = "forward warp needs images and a flow of at least 2x2";
        return TWX_E_MISMATCH;
    }
    float a = 0.0f > amount ? -amount : amount;
    if (1e-05 > (double)a || fw_flow_is_zero(flow, fw * fh * 2)) {
        memcpy(out, src, image_bytes(w, h, fmt));
        if (cov) memset(cov, 0xff, (size_t)w * h);
        return TWX_OK;
    }
    if (fmt == TWX_FMT_U8) return fw_run((uint8_t*)out, (const uint8_t*)src, w, h, flow, fw, fh, amount, gouraud, cov, err);
    if (fmt == TWX_FMT_U16) return fw_run((uint16_t*)out, (const uint16_t*)src, w, h, flow, fw, fh, amount, gouraud, cov, err);
    return fw_run((float*)out, (const float*)src, w, h, flow, fw, fh, amount, gouraud, cov, err);
}

// The compositions of warping codes 2 and 3 (compose/vector.rs:117-147, forward.rs:391-572):
// each frame pushed with its own flow (A with F_AB, B with F_BA; Nearest the one flow), code 3
// without coverage (holes transparent), code 2 with coverage, the Smart Blend composite and the
// hole fill. `gouraud` selects the forward warp variant of the stage set (stages::GouraudForward).
static int compose_forward(int code, int interp, float wa, const void* a, const void* b, long w, long h, int fmt,
                           const float* f_ab, const float* f_ba, long fw, long fh, bool above_first, bool gouraud,
                           void* out, std::string& err) {
    size_t bytes = image_bytes(w, h, fmt);
    float frac = 1.0f - wa;
    bool fill = code == 2;
    int st;
    if (interp == 0) {
        const float* flow = (0.5f <= wa) ? f_ab : f_ba;
        std::vector<uint8_t> cov(fill ? w * h : 0);
        if (!(0.5f > wa)) st = forward_warp(out, a, w, h, fmt, flow, fw, fh, 1.0f - wa, gouraud, fill ? cov.data() : nullptr, err);
        else st = forward_warp(out, b, w, h, fmt, flow, fw, fh, wa, gouraud, fill ? cov.data() : nullptr, err);
        if (st || !fill) return st;
        for (auto& c : cov) c = (uint8_t)(255 - c);
        return fill_holes(out, w, h, fmt, cov.data(), above_first, err);
    }
    std::vector<uint8_t> ta(bytes, 0), tb(bytes, 0);
    std::vector<uint8_t> cov1(fill ? w * h : 0), cov2(fill ? w * h : 0);
    if ((st = forward_warp(ta.data(), a, w, h, fmt, f_ab, fw, fh, frac, gouraud, fill ? cov1.data() : nullptr, err))) return st;
    if ((st = forward_warp(tb.data(), b, w, h, fmt, f_ba, fw, fh, wa, gouraud, fill ? cov2.data() : nullptr, err))) return st;
    if (interp == 2) {
        sharpen_step_with(ta.data(), a, w, h, fmt, frac, above_first);
        sharpen_step_with(tb.data(), b, w, h, fmt, wa, above_first);
    }
    if (!fill) {
        blend(out, ta.data(), tb.data(), w, h, fmt, frac);
        return TWX_OK;
    }
    smart_blend(out, ta.data(), tb.data(), w, h, fmt, frac);
    return smart_finish(out, w, h, fmt, cov1, cov2, above_first, err);
}

// ------------------------------------------------------------------------------------------
// Vector image coding (vectors.rs:57-130, :474-583) and the flow steps of Vectors In
// (compose/vector.rs:335-397): decode, the hole fill of the flow, Vec Scale, and the flow
// inversion through the forward warp.

static inline void vec_clamp_unit(float& v) {  // vectors.rs:132-141 (a NaN falls through)
    if (v > 1.0f) v = 1.0f;
    else if (-1.0f > v) v = -1.0f;
}

// vectors.rs:57-130: the flow sampled at every output pixel, scaled to output pixels over `scale`
static int encode_vectors(void* dst, long w, long h, int fmt, const float* flow, long fw, long fh, float scale,
                          std::string& err) {
    if (w < 2 || h < 2 || fw < 2 || fh < 2) {
        err = "vector encode needs images and a flow of at least 2x2";
        return TWX_E_MISMATCH;
    }
    float dw1 = (float)((int32_t)w - 1), dh1 = (float)((int32_t)h - 1);
    float fw1 = (float)((int32_t)fw - 1), fh1 = (float)((int32_t)fh - 1);
    float sx = fw1 / dw1, sy = fh1 / dh1;
    float kx = (dw1 / fw1) / scale, ky = (dh1 / fh1) / scale;
    for (long y = 0; y < h; y++)
        for (long x = 0; x < w; x++) {
            float f[2];
            flow_bilerp(flow, fw, (int32_t)fw - 1, (int32_t)fh - 1, (float)x * sx, (float)y * sy, f);
            float a = f[0] * kx, b = f[1] * ky;
            long i = (y * w + x) * 4;
            if (fmt == TWX_FMT_F32) {
                float* o = (float*)dst;
                o[i] = 1.0f;
                o[i + 1] = (a + 1.0f) * 0.5f;
                o[i + 2] = (b + 1.0f) * 0.5f;
                o[i + 3] = 0.0f;
                continue;
            }
            vec_clamp_unit(a);
            vec_clamp_unit(b);
            a = a + 1.0f;
            b = b + 1.0f;
            if (fmt == TWX_FMT_U8) {
                uint8_t* o = (uint8_t*)dst;
                o[i] = 0xff;
                o[i + 1] = (uint8_t)cvtt64d((double)a * 127.0 + 0.5);
                o[i + 2] = (uint8_t)cvtt64d((double)b * 127.0 + 0.5);
                o[i + 3] = 0;
            } else {
                uint16_t* o = (uint16_t*)dst;
                o[i] = 0xffff;
                o[i + 1] = (uint16_t)cvtt64d((double)a * 32767.0 + 0.5);
                o[i + 2] = (uint16_t)cvtt64d((double)b * 32767.0 + 0.5);
                o[i + 3] = 0;
            }
        }
    return TWX_OK;
}

// vectors.rs:474-521: alpha 0 is the zero vector; integers ((c * 2) / den - 1) * scale in double
static void decode_vectors(const void* vec, long w, long h, int fmt, float max_displace, float* flow) {
    double md = (double)max_displace;
    for (long i = 0; i < w * h; i++) {
        float* o = flow + i * 2;
        if (fmt == TWX_FMT_F32) {
            const float* p = (const float*)vec + i * 4;
            if (p[0] == 0.0f) { o[0] = o[1] = 0.0f; continue; }
            o[0] = (p[1] * 2.0f - 1.0f) * max_displace;
            o[1] = (p[2] * 2.0f - 1.0f) * max_displace;
        } else if (fmt == TWX_FMT_U16) {
            const uint16_t* p = (const uint16_t*)vec + i * 4;
            if (p[0] == 0) { o[0] = o[1] = 0.0f; continue; }
            o[0] = (float)((((double)p[1] * 2.0) / 65534.0 - 1.0) * md);
            o[1] = (float)((((double)p[2] * 2.0) / 65534.0 - 1.0) * md);
        } else {
            const uint8_t* p = (const uint8_t*)vec + i * 4;
            if (p[0] == 0) { o[0] = o[1] = 0.0f; continue; }
            o[0] = (float)((((double)p[1] * 2.0) / 254.0 - 1.0) * md);
            o[1] = (float)((((double)p[2] * 2.0) / 254.0 - 1.0) * md);
        }
    }
}

// vectors.rs:543-549 with holefill.rs:124-133: the inverted alpha of the vector image as a hole
// flag (u8: any non-zero; u16 and f32: at least 3.1e-5 or unordered)
static std::vector<uint8_t> vector_holes(const void* vec, long w, long h, int fmt) {
    std::vector<uint8_t> holes((size_t)w * h);
    const double wide_eps = [] { uint64_t b = 0x3f0040bfe3b03e21ull; double d; memcpy(&d, &b, 8); return d; }();
    for (long i = 0; i < w * h; i++) {
        bool hole;
        if (fmt == TWX_FMT_U8) hole = (uint8_t)(0xff - ((const uint8_t*)vec)[i * 4]) != 0;
        else if (fmt == TWX_FMT_U16) hole = !(wide_eps > (double)(uint16_t)(0xffff - ((const uint16_t*)vec)[i * 4]));
        else hole = !(wide_eps > (double)(1.0f - ((const float*)vec)[i * 4]));
        holes[i] = hole ? 1 : 0;
    }
    return holes;
}

// holefill.rs:506-521: the fill of a two channel float flow
static int fill_flow_holes(float* flow, long w, long h, const uint8_t* holes, bool above_first, std::string& err) {
    HoleFiller f;
    bool has;
    int st = hole_filler(w, h, holes, above_first, f, has, err);
    if (st || !has) return st;
    std::vector<float> plane((size_t)w * h);
    for (int c = 0; c < 2; c++) {
        for (long i = 0; i < w * h; i++) plane[i] = flow[i * 2 + c];
        f.fill(plane.data());
        for (long i = 0; i < w * h; i++) flow[i * 2 + c] = plane[i];
    }
    return TWX_OK;
}

// compose/vector.rs:335-345 with vectors.rs:566-571: decode, fill, Vec Scale
static int flow_from_vector_image(const void* vec, long w, long h, int fmt, float max_displace, float scale_x,
                                  float scale_y, bool above_first, float* flow, std::string& err) {
    decode_vectors(vec, w, h, fmt, max_displace, flow);
    std::vector<uint8_t> holes = vector_holes(vec, w, h, fmt);
    int st = fill_flow_holes(flow, w, h, holes.data(), above_first, err);
    if (st) return st;
    for (long i = 0; i < w * h; i++) {
        flow[i * 2] = scale_x * flow[i * 2];
        flow[i * 2 + 1] = scale_y * flow[i * 2 + 1];
    }
    return TWX_OK;
}

// compose/vector.rs:370-397: T = (-1, -fx, -fy, -1) pushed along the flow itself (amount 1, with
// coverage), the uncovered pixels filled; `gouraud` is the stage set's forward variant
static int invert_flow(const float* flow, long w, long h, bool gouraud, bool above_first, float* out, std::string& err) {
    std::vector<float> t((size_t)w * h * 4), pushed((size_t)w * h * 4);
    for (long i = 0; i < w * h; i++) {
        t[i * 4] = -1.0f * 1.0f;
        t[i * 4 + 1] = -1.0f * flow[i * 2];
        t[i * 4 + 2] = -1.0f * flow[i * 2 + 1];
        t[i * 4 + 3] = -1.0f * 1.0f;
    }
    std::vector<uint8_t> cov((size_t)w * h);
    int st = forward_warp(pushed.data(), t.data(), w, h, TWX_FMT_F32, flow, w, h, 1.0f, gouraud, cov.data(), err);
    if (st) return st;
    for (auto& c : cov) c = (uint8_t)(255 - c);
    for (long i = 0; i < w * h; i++) {
        out[i * 2] = pushed[i * 4 + 1];
        out[i * 2 + 1] = pushed[i * 4 + 2];
    }
    return fill_flow_holes(out, w, h, cov.data(), above_first, err);
}

// ------------------------------------------------------------------------------------------
// Time model (timeline.rs, curves.rs, fixed.rs, frontend.rs)

static inline int32_t float_to_fixed(double f) {  // fixed.rs:8-10 (Rust `as i32` saturates, NaN 0)
    double v = f * 65536.0 + (f < 0.0 ? -0.5 : 0.5);
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int32_t)v;
}
static inline float fixed_reader(int32_t fixed, float scale) {  // frontend.rs:130-132
    return (float)((double)fixed * (1.0 / 65536.0)) * scale;
}

// curves.rs:537-544
static inline double eased_progress(double u, bool es, bool ee) {
    if (!es && !ee) return u;
    if (es && ee) return u * u * (3.0 - 2.0 * u);
    if (es) return u * u * (2.0 - u);
    return u * (1.0 + u * (1.0 - u));
}

// curves.rs:344-370 for a float parameter (no Bezier handles)
static double speed_at(const twx_retime_params& p, int64_t time) {
    if (p.n_speed_keys <= 0) return p.speed;
    const twx_key* k = p.speed_keys;
    int last = p.n_speed_keys - 1;
    if (time <= k[0].time) return k[0].value;
    if (time >= k[last].time) return k[last].value;
    int i = 0;
    while (i + 1 <= last && k[i + 1].time <= time) i++;
    const twx_key& a = k[i];
    const twx_key& b = k[i + 1];
    if (a.hold) return a.value;
    double f = eased_progress((double)(time - a.time) / (double)(b.time - a.time), a.ease != 0, b.ease != 0);
    return a.value + (b.value - a.value) * f;
}

static inline int32_t speed_fixed(const twx_retime_params& p, int32_t time) { return float_to_fixed(speed_at(p, time)); }

static double snap(double rate) {  // timeline.rs:69-76
    const double t[3] = {29.97, 59.94, 23.976};
    for (double c : t)
        if ((double)std::fabs((float)(rate - c)) < 2e-5) rate = c;
    return rate;
}

struct Rates {
    double in_fps, out_fps;
    float dur;
    int32_t lts;
};

static int make_rates(const twx_retime_params& p, Rates& r, std::string& err) {  // timeline.rs:108-133
    double in_fps = (double)((float)p.input_frame_rate * 1.0f);
    double out_fps = p.fps;
    if (p.in_fps_is_out_fps) { in_fps = 30.0; out_fps = 30.0; }
    in_fps = snap(in_fps);
    out_fps = snap(out_fps);
    if (0.01 > out_fps || 0.01 >= in_fps) { err = "error 5 (bad parameter): frame rates"; return TWX_E_MISMATCH; }
    r.in_fps = in_fps;
    r.out_fps = out_fps;
    r.lts = p.time_step;
    r.dur = (float)(((double)p.time_step * out_fps) / in_fps);
    return TWX_OK;
}

// timeline.rs:280-331 (Speed mode, src_time_offset 0 so base stays 0)
static int src_pos(const twx_retime_params& p, const Rates& r, float t, float& out_pos, std::string& err) {
    float lts_f = (float)r.lts;
    float out = 0.0f;
    int32_t n2 = cvtt_i32f(t * 2.0f);
    if (n2 >= (1 << 24)) { err = "output position beyond 2^23 frames is not supported"; return TWX_E_UNSUPPORTED; }
    float k = 1.0f;
    while ((float)n2 >= k) {
        int32_t time = cvtt_i32d((double)((lts_f * k) / 2.0f) + 0.05);
        float sp = (float)((double)speed_fixed(p, time) / 6553600.0);
        out = (sp / 2.0f) + out;
        k += 1.0f;
    }
    float tt = t * 2.0f;
    float rem = tt - (k + -1.0f);
    if ((double)rem > 0.0001) {
        int32_t time = cvtt_i32d((double)((tt * lts_f) / 2.0f) + 0.05);
        float sp = (float)((double)speed_fixed(p, time) / 6553600.0);
        out = ((rem * sp) / 2.0f) + out;
    }
    out_pos = (float)((r.in_fps / r.out_fps) * (double)out);
    return TWX_OK;
}

// timeline.rs:342-350 frame_time with base 0
static int32_t frame_time(const Rates& r, float f) {
    float d = f - 0.0f;
    if (std::fabs(r.in_fps - r.out_fps) < 0.05) return cvtt_i32d((double)d * (double)r.dur + 0.5);
    return cvtt_i32d((double)(d * r.dur) + (double)r.dur * 0.05);
}

// timeline.rs:397-409
static bool exact_position(float s0, float s1, float blur) {
    bool exact = true;
    float frac = s0 - std::floor(s0);
    if ((double)std::fabs(frac) > 0.005 && (double)std::fabs((float)((double)frac - 1.0)) > 0.005) exact = false;
    if ((double)std::fabs(s0 - s1) > 1e-6 && (double)blur > 0.005) exact = false;
    return exact;
}

static int select_frame(const twx_retime_params& p, int32_t n, twx_selection& sel, Rates& rates, std::string& err) {
    int st = make_rates(p, rates, err);
    if (st) return st;
    int32_t cur = n * p.time_step;
    // timeline.rs:137-142 frame_positions
    float curf = (float)cur, ltsf = (float)p.time_step;
    float t0 = curf / ltsf;
    float t1 = ((float)p.time_step + curf) / ltsf;
    float s0, s1;
    if ((st = src_pos(p, rates, t0, s0, err))) return st;
    if ((st = src_pos(p, rates, t1, s1, err))) return st;
    // frontend.rs:215-218 and prep.rs:106-112: popups minus one, the tiny frames rule
    int vq = p.vectors - 1;
    if (!(p.width >= 2 && p.height >= 2)) vq = 0;
    // the segment marks are at their default (popup 1 = Cut A): marker 0 everywhere, no override
    int32_t a = cvtt_i32f(s0);
    float frac = s0 - (float)a;
    sel.s0 = s0;
    sel.s1 = s1;
    sel.a = a;
    sel.b = (int32_t)((uint32_t)a + 1u);
    sel.wa = 1.0f - frac;
    sel.exact = exact_position(s0, s1, 0.0f) ? 1 : 0;
    sel.vectors_quality = vq;
    sel.interp = p.frame_interp - 1;
    return TWX_OK;
}

// ------------------------------------------------------------------------------------------
// The per-frame render (compose.rs:93-246, compose/vector.rs:163-321, frontend.rs:290-312)

struct Clip {
    const uint8_t* wire;
    int32_t n;
    long w, h;
    int fmt;
};

// host.rs main_frame: the clip frame at frame_time(position), transparent outside the clip
// (clip.rs:170-172 frame_at gives None; host.rs:123-127 transparent image).
static std::vector<uint8_t> main_frame(const Clip& c, const Rates& r, int32_t lts, int32_t pos) {
    std::vector<uint8_t> img(image_bytes(c.w, c.h, c.fmt), 0);
    int32_t time = frame_time(r, (float)pos);
    int64_t idx = (int64_t)std::floor((double)time / (double)lts);  // TimeBase::frame_of, div_euclid
    if (idx >= 0 && idx < c.n) {
        size_t fb = image_bytes(c.w, c.h, c.fmt);
        from_frame(c.wire + (size_t)idx * fb, c.w, c.h, c.fmt, img.data());
    }
    return img;
}

static int render_frame(const twx_retime_params& p, const Clip& c, int32_t n, uint8_t* out_wire, std::string& err) {
    twx_selection sel;
    Rates rates;
    int st = select_frame(p, n, sel, rates, err);
    if (st) return st;
    long w = c.w, h = c.h;
    int fmt = c.fmt;
    size_t bytes = image_bytes(w, h, fmt);
    std::vector<uint8_t> out(bytes, 0);
    if (sel.vectors_quality != 0 && !sel.exact) {
        // vector path (compose/vector.rs:289-314)
        if (!(sel.interp >= 0 && sel.interp <= 2)) { err = "frame_interp outside 1..3"; return TWX_E_MISMATCH; }
        std::vector<uint8_t> a0 = main_frame(c, rates, p.time_step, sel.a);
        std::vector<uint8_t> b0 = main_frame(c, rates, p.time_step, sel.b);
        std::vector<uint8_t> a(bytes), b(bytes);
        rows_reversed(a0.data(), w, h, fmt, a.data());
        rows_reversed(b0.data(), w, h, fmt, b.data());
        twx_flow_settings fs{};
        fs.quality = sel.vectors_quality;
        fs.hiq = p.quality_high;
        fs.sensitivity = fixed_reader(float_to_fixed(p.sensitivity), 1.0f) / 100.0f;
        Pyramid pa, pb;
        if ((st = build_pyramid(a.data(), w, h, fmt, fs, pa, err))) return st;
        if ((st = build_pyramid(b.data(), w, h, fmt, fs, pb, err))) return st;
        long fw = pa.geo.width, fh = pa.geo.height;
        float wa = sel.wa;
        float frac = 1.0f - wa;
        std::vector<uint8_t> composed(bytes, 0);
        if (p.smart_blend) {
            // warping code 1 (frontend.rs:228: warp_mode = smart_blend), the fills and the
            // sharpening blur with the row above first (frontend.rs:238-241)
            std::vector<float> f_ba, f_ab;
            if (sel.interp == 0) {
                if (wa < 0.5f) st = solve_pyramids(pa, pb, fs, f_ab, err);  // nearest_uses_f_ab, code 1
                else st = solve_pyramids(pb, pa, fs, f_ba, err);
                if (st) return st;
            } else {
                if ((st = solve_pyramids(pb, pa, fs, f_ba, err))) return st;
                if ((st = solve_pyramids(pa, pb, fs, f_ab, err))) return st;
            }
            if ((st = compose_smart(sel.interp, wa, a.data(), b.data(), w, h, fmt, f_ab.data(), f_ba.data(), fw, fh, true,
                                    composed.data(), err)))
                return st;
        } else if (sel.interp == 0) {
            bool use_ab = wa < 0.5f;  // nearest_uses_f_ab with warping code 0
            std::vector<float> flow;
            if (use_ab) { if ((st = solve_pyramids(pa, pb, fs, flow, err))) return st; }
            else { if ((st = solve_pyramids(pb, pa, fs, flow, err))) return st; }
            if (0.5f <= wa) st = inv_warp(composed.data(), a.data(), w, h, fmt, flow.data(), fw, fh, frac, err);
            else st = inv_warp(composed.data(), b.data(), w, h, fmt, flow.data(), fw, fh, wa, err);
            if (st) return st;
        } else {
            std::vector<float> f_ba, f_ab;
            if ((st = solve_pyramids(pb, pa, fs, f_ba, err))) return st;
            if ((st = solve_pyramids(pa, pb, fs, f_ab, err))) return st;
            std::vector<uint8_t> ta(bytes, 0), tb(bytes, 0);
            if ((st = inv_warp(ta.data(), a.data(), w, h, fmt, f_ba.data(), fw, fh, frac, err))) return st;
            if ((st = inv_warp(tb.data(), b.data(), w, h, fmt, f_ab.data(), fw, fh, wa, err))) return st;
            if (sel.interp == 2) {
                // Weighted Blend (compose/vector.rs:268-273): the sharpening step of each warped
                // frame against its unwarped frame, strength position frac for A and wa for B; the
                // retime front-end's stages take the row above first (frontend.rs:238-241)
                sharpen_step_with(ta.data(), a.data(), w, h, fmt, frac, true);
                sharpen_step_with(tb.data(), b.data(), w, h, fmt, wa, true);
            }
            blend(composed.data(), ta.data(), tb.data(), w, h, fmt, frac);
        }
        rows_reversed(composed.data(), w, h, fmt, out.data());
    } else {
        // fast path (compose.rs:214-246)
        std::vector<uint8_t> a = main_frame(c, rates, p.time_step, sel.a);
        std::vector<uint8_t> b = main_frame(c, rates, p.time_step, sel.b);
        if (sel.interp == 0) {
            memcpy(out.data(), sel.wa >= 0.5f ? a.data() : b.data(), bytes);
        } else if (sel.interp == 1) {
            blend(out.data(), a.data(), b.data(), w, h, fmt, (float)(1.0 - (double)sel.wa));
        } else if (sel.interp == 2) {
            // weighted_blend_noflow (blend.rs:280-300): the plain Blend with 1 - wa in f32, and for
            // float images the inputs plus 0.0 (finite inputs of magnitude up to 1e30)
            float t = 1.0f - sel.wa;
            if (fmt == TWX_FMT_F32) {
                float* fa = (float*)a.data();
                float* fb = (float*)b.data();
                for (long i = 0; i < w * h * 4; i++) {
                    if (!std::isfinite(fa[i]) || !std::isfinite(fb[i]) || std::fabs(fa[i]) > 1e30f || std::fabs(fb[i]) > 1e30f) {
                        err = "oracle: non-finite float input to the weighted blend is not restated";
                        return TWX_E_UNSUPPORTED;
                    }
                    fa[i] = fa[i] + 0.0f;
                    fb[i] = fb[i] + 0.0f;
                }
            }
            blend(out.data(), a.data(), b.data(), w, h, fmt, t);
        } else {
            err = "frame_interp outside 1..3";
            return TWX_E_MISMATCH;
        }
    }
    to_frame(out.data(), w, h, fmt, out_wire);
    return TWX_OK;
}

// ------------------------------------------------------------------------------------------
// The Vectors In render (plugins/vectors-in/src/glue.rs:380-495, vectorin.rs:1-230,
// timemap.rs:1-70): the time map of that effect, the flows from its two vector layers, the
// core of compose.rs:122-130 with the forward warps of stages::GouraudForward and the taps above
// first (glue.rs:253). Every frame a render reads is turned upside down before the render sees it
// and the result is turned back (glue.rs:398-410).

extern "C" int cref_compose_vectors(int interp, float wa, const void* a, const void* b, int w, int h, int fmt,
                                    const float* f_ab, const float* f_ba, int fw, int fh, int warp_mode, int above_first,
                                    int gouraud, void* out, char* err, size_t errlen);

// `vfx apply vectors-in` parameters (plugins/vectors-in/src/lib.rs:82-130), the struct of the C ABI
using cref_vin_params = twx_vin_params;

// One layer of a render: wire frames, frame i showing at ticks [i * time_step, (i + 1) * time_step)
// (crates/fx-cli/src/temporal.rs:21-23); `wire == nullptr` is a layer that is not set.
struct VinLayer {
    const uint8_t* wire;
    int32_t n;
};

// The stage set and the row order of a render: the effect's glue installs the Gouraud forward
// warp and the taps above first and turns every frame upside down (glue.rs:57-75, :253, :398-410):
// {true, true, true}. The engine's plain stage set on rows as given, {false, false, false}, is the
// configuration the reference CPU path recorded its grid hashes with (results/base.grid.tsv).
struct VinStages {
    bool gouraud, above_first, flip;
};

// The layer's frame at `time`, premultiplied (rows reversed with `flip`); transparent outside the
// clip (clip.rs:170-172, host.rs:123-127).
static std::vector<uint8_t> vin_frame(const VinLayer& L, long w, long h, int fmt, int32_t lts, int64_t time, bool flip) {
    const size_t fb = image_bytes(w, h, fmt);
    std::vector<uint8_t> img(fb, 0), up(fb, 0);
    const int64_t idx = (int64_t)std::floor((double)time / (double)lts);  // TimeBase::frame_of
    if (L.wire && idx >= 0 && idx < L.n) from_frame(L.wire + (size_t)idx * fb, w, h, fmt, img.data());
    if (!flip) return img;
    rows_reversed(img.data(), w, h, fmt, up.data());
    return up;
}

static int vin_render_frame(const cref_vin_params& p, const VinLayer& input, const VinLayer& source, const VinLayer& prev,
                            const VinLayer& next, const VinStages& vs, int32_t n, uint8_t* out_wire, std::string& err) {
    const long w = p.width, h = p.height;
    const int fmt = p.fmt;
    const size_t bytes = image_bytes(w, h, fmt);
    const VinLayer& layer = source.wire ? source : input;  // the retimed layer (glue.rs:86-89 IDS)
    // the rates (glue.rs:200-209): input 30 fps with "In FPS is Out FPS", else the slider; output the
    // composition's; Rates::new without its equal-rates switch (timeline.rs:108-133)
    double in_fps = p.in_fps_is_out_fps ? 30.0 : (double)((float)p.input_frame_rate * 1.0f);
    double out_fps = p.fps;
    in_fps = snap(in_fps);
    out_fps = snap(out_fps);
    const bool rate_error = 0.01 > out_fps || 0.01 >= in_fps;
    // the probes (glue.rs:268-300): neither vector layer is the size mismatch fill, as a rate error
    if (rate_error || (!prev.wire && !next.wire)) {
        const size_t ss = sample_size(fmt);
        for (long i = 0; i < w * h; i++) {  // image.rs:391-425: red at half alpha, R, G, B, A order
            uint8_t* q = out_wire;
            if (fmt == TWX_FMT_U8) {
                q[i * 4 + 0] = 0xff; q[i * 4 + 1] = 0; q[i * 4 + 2] = 0; q[i * 4 + 3] = 0x80;
            } else if (fmt == TWX_FMT_U16) {
                uint16_t v[4] = {q15_to_u16(0x8000), 0, 0, q15_to_u16(0x4000)};
                memcpy(q + i * 8, v, 8);
            } else {
                float* pl = (float*)q;  // planes G, B, R, A
                pl[i] = 0.0f; pl[w * h + i] = 0.0f; pl[2 * w * h + i] = 1.0f; pl[3 * w * h + i] = 0.5f;
            }
            (void)ss;
        }
        return TWX_OK;
    }
    Rates r{in_fps, out_fps, (float)(((double)p.time_step * out_fps) / in_fps), p.time_step};
    const int32_t cur = n * p.time_step;
    if (p.display == 2) {  // the retimed layer at current_time through the prepare and the finish
        std::vector<uint8_t> img = vin_frame(layer, w, h, fmt, p.time_step, cur, false);
        to_frame(img.data(), w, h, fmt, out_wire);
        return TWX_OK;
    }
    if (p.blur != 0.0) { err = "oracle: the blur compensation tail is not restated"; return TWX_E_UNSUPPORTED; }
    // the time map (timemap.rs:56-70): src_pos of RemapTimeMap without the source time offset,
    // T(k) = trunc((f64)dur * (f64)k + 0.5)
    twx_retime_params rp{};
    rp.time_step = p.time_step;
    rp.speed = p.speed;
    rp.n_speed_keys = p.n_speed_keys;
    rp.speed_keys = p.speed_keys;
    auto src_pos_at = [&](float t, float& s) -> int {
        if (p.time_remap_mode == 2) {  // timeline.rs:274-279, a constant Frame Num
            const float v = (float)p.frame_num;
            s = (float)((r.in_fps / r.out_fps) * (double)v);
            return TWX_OK;
        }
        return src_pos(rp, r, t, s, err);
    };
    auto T = [&](float k) -> int64_t { return cvtt_i32d((double)r.dur * (double)k + 0.5); };
    const float curf = (float)cur, ltsf = (float)p.time_step;
    const float t0 = curf / ltsf, t1 = ((float)p.time_step + curf) / ltsf;
    float s0, s1;
    int st;
    if ((st = src_pos_at(t0, s0))) return st;
    if ((st = src_pos_at(t1, s1))) return st;
    // the constant VECTORS_QUALITY 4 (prerender.rs:50), 0 for a frame with a side of 1
    int vq = (w >= 2 && h >= 2) ? 4 : 0;
    int interp = p.frame_interp - 1;
    // the segment override (timeline.rs:418-440) with a constant Mark Segments: both markers equal
    const int mark = (p.mark_segments >= 1 && p.mark_segments <= 5) ? p.mark_segments - 1 : 0;
    if (mark == 3) { vq = 0; interp = 0; }
    if (mark == 4) { vq = 0; interp = 1; }
    const int32_t a = cvtt_i32f(s0);
    const int32_t b = (int32_t)((uint32_t)a + 1u);
    const float wa = 1.0f - (s0 - (float)a);
    const bool exact = exact_position(s0, s1, (float)p.blur);
    std::vector<uint8_t> out(bytes, 0);
    if (vq != 0 && !exact) {
        if (interp < 0 || interp > 2) { err = "frame_interp outside 1..3"; return TWX_E_MISMATCH; }
        const int warp_mode = p.warping == 2 ? 2 : 0;  // glue.rs:96-103
        // the loaders (vectorin.rs:200-225): Max Displace read once, Vec Scale at the load's time
        const float md = (1.0f / 1.0f) * fixed_reader(float_to_fixed(p.max_displace), 1.0f);
        const float sx = fixed_reader(float_to_fixed(p.vec_scale_x), 1.0f);
        const float sy = fixed_reader(float_to_fixed(p.vec_scale_y), 1.0f);
        const size_t fn = (size_t)w * h * 2;
        std::vector<float> f_ab, f_ba;
        if (next.wire) {
            std::vector<uint8_t> vec = vin_frame(next, w, h, fmt, p.time_step, T((float)a), vs.flip);
            f_ab.assign(fn, 0.0f);
            if ((st = flow_from_vector_image(vec.data(), w, h, fmt, md, sx, sy, vs.above_first, f_ab.data(), err))) return st;
        }
        if (prev.wire) {
            std::vector<uint8_t> vec = vin_frame(prev, w, h, fmt, p.time_step, T((float)b), vs.flip);
            f_ba.assign(fn, 0.0f);
            if ((st = flow_from_vector_image(vec.data(), w, h, fmt, md, sx, sy, vs.above_first, f_ba.data(), err))) return st;
        }
        // complete_flows (compose/vector.rs:430-456): the missing flow is the other's inversion, made
        // only when the composition reads it
        bool need_ab = true, need_ba = true;
        if (interp == 0) {
            const bool ab = (wa < 0.5f && warp_mode < 2) || (0.5f <= wa && warp_mode > 1);  // nearest_uses_f_ab
            need_ab = ab;
            need_ba = !ab;
        }
        if (f_ab.empty() && need_ab) {
            f_ab.assign(fn, 0.0f);
            if ((st = invert_flow(f_ba.data(), w, h, vs.gouraud, vs.above_first, f_ab.data(), err))) return st;
        }
        if (f_ba.empty() && need_ba) {
            f_ba.assign(fn, 0.0f);
            if ((st = invert_flow(f_ab.data(), w, h, vs.gouraud, vs.above_first, f_ba.data(), err))) return st;
        }
        if (f_ab.empty()) f_ab.assign(fn, 0.0f);  // not read by this composition
        if (f_ba.empty()) f_ba.assign(fn, 0.0f);
        std::vector<uint8_t> ia = vin_frame(layer, w, h, fmt, p.time_step, T((float)a), vs.flip);
        std::vector<uint8_t> ib = vin_frame(layer, w, h, fmt, p.time_step, T((float)b), vs.flip);
        char e[512] = {0};
        st = cref_compose_vectors(interp, wa, ia.data(), ib.data(), (int)w, (int)h, fmt, f_ab.data(), f_ba.data(), (int)w,
                                  (int)h, warp_mode, vs.above_first ? 1 : 0, vs.gouraud ? 1 : 0, out.data(), e, sizeof e);
        if (st) { err = e; return st; }
    } else {
        std::vector<uint8_t> ia = vin_frame(layer, w, h, fmt, p.time_step, T((float)a), vs.flip);
        std::vector<uint8_t> ib = vin_frame(layer, w, h, fmt, p.time_step, T((float)b), vs.flip);
        if (interp == 0) {
            memcpy(out.data(), wa >= 0.5f ? ia.data() : ib.data(), bytes);
        } else if (interp == 1 || interp == 2) {
            // Blend, and Weighted Blend without flows (blend.rs:280-300: the plain Blend with
            // 1 - wa in f32, the float inputs plus 0.0; non-finite or huge samples not restated)
            const float t = interp == 1 ? (float)(1.0 - (double)wa) : 1.0f - wa;
            if (interp == 2 && fmt == TWX_FMT_F32) {
                float* fa = (float*)ia.data();
                float* fb = (float*)ib.data();
                for (long i = 0; i < w * h * 4; i++) {
                    if (!std::isfinite(fa[i]) || !std::isfinite(fb[i]) || std::fabs(fa[i]) > 1e30f || std::fabs(fb[i]) > 1e30f) {
                        err = "oracle: non-finite float input to the weighted blend is not restated";
                        return TWX_E_UNSUPPORTED;
                    }
                    fa[i] = fa[i] + 0.0f;
                    fb[i] = fb[i] + 0.0f;
                }
            }
            blend(out.data(), ia.data(), ib.data(), w, h, fmt, t);
        } else {
            err = "frame_interp outside 1..3";
            return TWX_E_MISMATCH;
        }
    }
    std::vector<uint8_t> back(bytes);
    if (vs.flip) rows_reversed(out.data(), w, h, fmt, back.data());
    else back = out;
    to_frame(back.data(), w, h, fmt, out_wire);
    return TWX_OK;
}

}  // namespace cref

// ==========================================================================================
// C ABI for the tests and the CPU-baseline leg of the benchmark (ctypes)

using namespace cref;

static void set_err(char* err, size_t errlen, const std::string& m) {
    if (err && errlen) {
        size_t n = std::min(errlen - 1, m.size());
        memcpy(err, m.data(), n);
        err[n] = 0;
    }
}

extern "C" {

int cref_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int cref_gaussian_kernel(float r, float* out, int cap) {
    std::vector<float> k = gaussian_kernel(r);
    for (int i = 0; i < (int)k.size() && i < cap; i++) out[i] = k[i];
    return (int)k.size();
}

void cref_convolve(float* data, int w, int h, int c, const float* k, int n, int above_first) {
    convolve(data, w, h, c, k, n, above_first != 0);
}

void cref_resample1(const float* src, int sw, int sh, int dw, int dh, float* dst) {
    std::vector<float> d = resample1(src, sw, sh, dw, dh);
    memcpy(dst, d.data(), d.size() * 4);
}

float cref_bilinear1(const float* img, int w, int h, float x, float y) { return bilinear1(img, w, h, x, y); }

void cref_luma(const void* img, int w, int h, int fmt, float* out) {
    std::vector<float> l = luma(img, w, h, fmt);
    memcpy(out, l.data(), l.size() * 4);
}

int cref_geometry(const twx_flow_settings* fs, int w, int h, int* fw, int* fh, int* levels, int* sizes_wh) {
    Geometry g;
    std::string e;
    int st = geometry(*fs, w, h, g, e);
    if (st) return st;
    *fw = (int)g.width;
    *fh = (int)g.height;
    *levels = g.levels;
    if (sizes_wh)
        for (int i = 0; i < g.levels; i++) { sizes_wh[2 * i] = (int)g.sizes[i].first; sizes_wh[2 * i + 1] = (int)g.sizes[i].second; }
    return TWX_OK;
}

void cref_start_level(float sensitivity, int levels, int* start, float* damp) { start_level(sensitivity, levels, *start, *damp); }

// The band-pass levels of one engine image, concatenated finest first.
int cref_build_pyramid(const void* img, int w, int h, int fmt, const twx_flow_settings* fs, float* out, char* err, size_t errlen) {
    Pyramid p;
    std::string e;
    int st = build_pyramid(img, w, h, fmt, *fs, p, e);
    if (st) { set_err(err, errlen, e); return st; }
    size_t off = 0;
    for (auto& l : p.levels) { memcpy(out + off, l.data(), l.size() * 4); off += l.size(); }
    return TWX_OK;
}

void cref_calc_grad(const float* img, int w, int h, float* out) {
    std::vector<float> g = calc_grad(img, w, h);
    memcpy(out, g.data(), g.size() * 4);
}
void cref_calc_b(const float* i1, const float* i0, const float* grad, const float* flow, int w, int h, float* out) {
    std::vector<float> b = calc_b(i1, i0, grad, flow, w, h);
    memcpy(out, b.data(), b.size() * 4);
}
void cref_new_disp(float* flow, const float* a, const float* b, int n) { new_disp(flow, a, b, n); }
void cref_upsample_flow(const float* src, int sw, int sh, int dw, int dh, float* dst) {
    std::vector<float> d = upsample_flow(src, sw, sh, dw, dh);
    memcpy(dst, d.data(), d.size() * 4);
}
void cref_block_match(const float* a, const float* b, float* flow, int w, int h) { block_match(a, b, flow, w, h); }
int cref_block_cost(const float* a, const float* b, int w, int h, int x, int y, float x2, float y2, float* out) {
    Block bm{a, b, w, h};
    float p[9];
    bm.patch(x, y, p);
    return bm.cost(p, x2, y2, *out) ? 1 : 0;
}
int cref_all_flow_out_of_range(const float* flow, int w, int h) { return all_flow_out_of_range(flow, w, h) ? 1 : 0; }

// flow::solve(X, Y) and flow::solve_pair(A, B) on engine images.
int cref_solve(const void* x, const void* y, int w, int h, int fmt, const twx_flow_settings* fs, float* out, char* err, size_t errlen) {
    Pyramid px, py;
    std::string e;
    int st = build_pyramid(x, w, h, fmt, *fs, px, e);
    if (!st) st = build_pyramid(y, w, h, fmt, *fs, py, e);
    std::vector<float> f;
    if (!st) st = solve_pyramids(px, py, *fs, f, e);
    if (st) { set_err(err, errlen, e); return st; }
    memcpy(out, f.data(), f.size() * 4);
    return TWX_OK;
}

int cref_solve_pair(const void* a, const void* b, int w, int h, int fmt, const twx_flow_settings* fs, float* f_ab, float* f_ba,
                    char* err, size_t errlen) {
    Pyramid pa, pb;
    std::string e;
    int st = build_pyramid(a, w, h, fmt, *fs, pa, e);
    if (!st) st = build_pyramid(b, w, h, fmt, *fs, pb, e);
    std::vector<float> ba, ab;
    if (!st) st = solve_pyramids(pb, pa, *fs, ba, e);
    if (!st) st = solve_pyramids(pa, pb, *fs, ab, e);
    if (st) { set_err(err, errlen, e); return st; }
    memcpy(f_ab, ab.data(), ab.size() * 4);
    memcpy(f_ba, ba.data(), ba.size() * 4);
    return TWX_OK;
}

int cref_inv_warp(void* out, const void* src, int w, int h, int fmt, const float* flow, int fw, int fh, float amount, char* err,
                  size_t errlen) {
    std::string e;
    int st = inv_warp(out, src, w, h, fmt, flow, fw, fh, amount, e);
    if (st) set_err(err, errlen, e);
    return st;
}

// The forward warp (forward.rs:229-373): `pyramid` 1 the textured variant, 0 the Gouraud variant;
// `cov` (w * h bytes) may be null.
int cref_forward_warp(void* out, const void* src, int w, int h, int fmt, const float* flow, int fw, int fh, float amount,
                      int pyramid, uint8_t* cov, char* err, size_t errlen) {
    std::string e;
    int st = forward_warp(out, src, w, h, fmt, flow, fw, fh, amount, pyramid == 0, cov, e);
    if (st) set_err(err, errlen, e);
    return st;
}

// One triangle through the scan converter (raster.rs:153-300): the hits (x, y, payload[4]) into
// `hits` (at most `cap`), their count returned; *outside set when the last row pass would write
// outside the image.
int cref_fw_triangle(const float* vtx /* 3 x (x, y, p0..p3) */, int w1, int h1, float* hits, int cap, int* outside) {
    FwVtx t[3];
    for (int i = 0; i < 3; i++) {
        t[i].x = vtx[i * 6];
        t[i].y = vtx[i * 6 + 1];
        for (int k = 0; k < 4; k++) t[i].p[k] = vtx[i * 6 + 2 + k];
    }
    int n = 0;
    bool inside = fw_triangle(t, w1, h1, [&](int32_t x, int32_t y, const float* p) {
        if (n < cap) {
            hits[n * 6] = (float)x;
            hits[n * 6 + 1] = (float)y;
            for (int k = 0; k < 4; k++) hits[n * 6 + 2 + k] = p[k];
        }
        n++;
    });
    *outside = inside ? 0 : 1;
    return n;
}

// The vertex positions (mesh.rs:33-52) and mip levels (mesh.rs:54-88) of a w x h grid.
void cref_fw_mesh(const float* flow, int fw, int fh, int w, int h, float amount, float* pos, float* mips) {
    FwGrid g = fw_grid(w, h, fw, fh, amount);
    std::vector<float> p = fw_positions(g, flow, fw, fh);
    memcpy(pos, p.data(), p.size() * sizeof(float));
    if (mips)
        for (int r = 0; r < h; r++)
            for (int c = 0; c < w; c++) mips[r * w + c] = fw_mip_level(g, flow, fw, fh, c, r);
}

// The running mean step (forward.rs:104-128) on one pixel.
void cref_fw_fold(float* acc, uint8_t* n, const float* c) { fw_fold(acc, *n, c); }

// Vector image coding and the Vectors In flow steps.
int cref_encode_vectors(void* dst, int w, int h, int fmt, const float* flow, int fw, int fh, float scale, char* err,
                        size_t errlen) {
    std::string e;
    int st = encode_vectors(dst, w, h, fmt, flow, fw, fh, scale, e);
    if (st) set_err(err, errlen, e);
    return st;
}
void cref_decode_vectors(const void* vec, int w, int h, int fmt, float max_displace, float* flow) {
    decode_vectors(vec, w, h, fmt, max_displace, flow);
}
void cref_vector_holes(const void* vec, int w, int h, int fmt, uint8_t* holes) {
    std::vector<uint8_t> v = vector_holes(vec, w, h, fmt);
    memcpy(holes, v.data(), v.size());
}
int cref_fill_flow_holes(float* flow, int w, int h, const uint8_t* holes, int above_first, char* err, size_t errlen) {
    std::string e;
    int st = fill_flow_holes(flow, w, h, holes, above_first != 0, e);
    if (st) set_err(err, errlen, e);
    return st;
}
int cref_flow_from_vector_image(const void* vec, int w, int h, int fmt, float max_displace, float sx, float sy,
                                int above_first, float* flow, char* err, size_t errlen) {
    std::string e;
    int st = flow_from_vector_image(vec, w, h, fmt, max_displace, sx, sy, above_first != 0, flow, e);
    if (st) set_err(err, errlen, e);
    return st;
}
int cref_invert_flow(const float* flow, int w, int h, int gouraud, int above_first, float* out, char* err, size_t errlen) {
    std::string e;
    int st = invert_flow(flow, w, h, gouraud != 0, above_first != 0, out, e);
    if (st) set_err(err, errlen, e);
    return st;
}

// The fading warp (hold off) with its coverage image.
int cref_inv_warp_cov(void* out, const void* src, int w, int h, int fmt, const float* flow, int fw, int fh, float amount,
                      uint8_t* cov, char* err, size_t errlen) {
    std::string e;
    memset(cov, 255, (size_t)w * h);
    int st = inv_warp(out, src, w, h, fmt, flow, fw, fh, amount, e, false, cov);
    if (st) set_err(err, errlen, e);
    return st;
}

// Smart Blend and the hole fill, for the tests.
void cref_smart_blend(void* out, const void* p1, const void* p2, int w, int h, int fmt, float wt) {
    smart_blend(out, p1, p2, w, h, fmt, wt);
}
int cref_hole_geometry(int w, int h, int* dims, int cap) {
    HoleGeometry g = hole_geometry(w, h);
    for (int i = 0; i < (int)g.dims.size() && 2 * i + 1 < cap; i++) {
        dims[2 * i] = (int)g.dims[i].first;
        dims[2 * i + 1] = (int)g.dims[i].second;
    }
    return g.top * 1000 + (int)g.dims.size();
}
int cref_fill_holes(void* img, int w, int h, int fmt, const uint8_t* holes, int above_first, char* err, size_t errlen) {
    std::string e;
    int st = fill_holes(img, w, h, fmt, holes, above_first != 0, e);
    if (st) set_err(err, errlen, e);
    return st;
}
int cref_fill_plane(float* plane, int w, int h, const uint8_t* holes, int above_first, char* err, size_t errlen) {
    std::string e;
    HoleFiller f;
    bool has;
    int st = hole_filler(w, h, holes, above_first != 0, f, has, e);
    if (!st && has) f.fill(plane);
    if (st) set_err(err, errlen, e);
    return st;
}

void cref_blend(void* out, const void* a, const void* b, int w, int h, int fmt, float wt) { blend(out, a, b, w, h, fmt, wt); }
void cref_lerp(void* first, const void* second, int w, int h, int fmt, float wt) { lerp(first, second, w, h, fmt, wt); }
void cref_from_frame(const uint8_t* wire, int w, int h, int fmt, void* engine) { from_frame(wire, w, h, fmt, engine); }
void cref_to_frame(const void* engine, int w, int h, int fmt, uint8_t* wire) { to_frame(engine, w, h, fmt, wire); }
uint32_t cref_sat_u8(float v) { return sat_u8(v); }
uint32_t cref_sat_u16(float v) { return sat_u16(v); }

// compose_vectors, warping code 0 (compose/vector.rs:249-276), on engine images in motion rows.
// `above_first` is the vertical tap order of the Weighted Blend blur (the stage set's
// vertical_taps).
int cref_compose_vectors(int interp, float wa, const void* a, const void* b, int w, int h, int fmt, const float* f_ab,
                         const float* f_ba, int fw, int fh, int warp_mode, int above_first, int gouraud, void* out,
                         char* err, size_t errlen) {
    std::string e;
    float frac = 1.0f - wa;
    size_t bytes = image_bytes(w, h, fmt);
    int st;
    if (warp_mode == 1 && interp >= 0 && interp <= 2) {
        st = compose_smart(interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above_first != 0, out, e);
    } else if ((warp_mode == 2 || warp_mode == 3) && interp >= 0 && interp <= 2) {
        st = compose_forward(warp_mode, interp, wa, a, b, w, h, fmt, f_ab, f_ba, fw, fh, above_first != 0, gouraud != 0, out, e);
    } else if (warp_mode != 0) {
        e = "warping code outside 0..3";
        st = TWX_E_MISMATCH;
    } else if (interp == 0) {
        if (0.5f <= wa) st = inv_warp(out, a, w, h, fmt, f_ba, fw, fh, frac, e);
        else st = inv_warp(out, b, w, h, fmt, f_ab, fw, fh, wa, e);
    } else if (interp == 1 || interp == 2) {
        std::vector<uint8_t> ta(bytes), tb(bytes);
        st = inv_warp(ta.data(), a, w, h, fmt, f_ba, fw, fh, frac, e);
        if (!st) st = inv_warp(tb.data(), b, w, h, fmt, f_ab, fw, fh, wa, e);
        if (!st && interp == 2) {
            sharpen_step_with(ta.data(), a, w, h, fmt, frac, above_first != 0);
            sharpen_step_with(tb.data(), b, w, h, fmt, wa, above_first != 0);
        }
        if (!st) blend(out, ta.data(), tb.data(), w, h, fmt, frac);
    } else {
        e = "interpolation mode outside 0..2";
        st = TWX_E_MISMATCH;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

// The filters of the Weighted Blend step, for the tests.
int cref_blur_taps(float radius, float* out, int cap) {
    std::vector<float> t = blur_taps(radius);
    for (int i = 0; i < (int)t.size() && i < cap; i++) out[i] = t[i];
    return (int)t.size();
}
void cref_blur(void* img, int w, int h, int fmt, float radius, int above_first) {
    blur_with(img, w, h, fmt, radius, above_first != 0);
}
float cref_mask_limit(int fmt, float threshold) { return mask_limit(fmt, threshold); }
void cref_difference_mask(const void* warped, const void* original, int w, int h, int fmt, float threshold, uint8_t* m) {
    difference_mask(warped, original, w, h, fmt, threshold, m);
}
// op: 0 prune(limit), 1 fill(limit, value), 2 dilate rows, 3 replicate border, 4 clean, 5 invert
void cref_mask_op(uint8_t* m, int w, int h, int op, int limit, int value) {
    switch (op) {
        case 0: mask_prune(m, w, h, limit); break;
        case 1: mask_fill(m, w, h, limit, (uint8_t)value); break;
        case 2: mask_dilate_rows(m, w, h); break;
        case 3: mask_replicate_border(m, w, h); break;
        case 4: clean_mask(m, w, h); break;
        default: for (long i = 0; i < (long)w * h; i++) m[i] = (uint8_t)(255 - m[i]); break;
    }
}
void cref_composite_masked(void* dest, const void* src, int w, int h, int fmt, const uint8_t* m) {
    composite_masked(dest, src, w, h, fmt, m);
}
float cref_sharpen_weight(float t) { return sharpen_weight(t); }
void cref_sharpen_step(void* frame, const void* original, int w, int h, int fmt, float t, int above_first) {
    sharpen_step_with(frame, original, w, h, fmt, t, above_first != 0);
}

int cref_select(const twx_retime_params* p, int n, twx_selection* out, char* err, size_t errlen) {
    std::string e;
    Rates r;
    int st = select_frame(*p, n, *out, r, e);
    if (st) set_err(err, errlen, e);
    return st;
}

// Source clip index of frame position `pos` (main_frame), or -1 outside the clip.
int cref_frame_index(const twx_retime_params* p, int pos, int n_in) {
    Rates r;
    std::string e;
    if (make_rates(*p, r, e)) return -2;
    int32_t time = frame_time(r, (float)pos);
    int64_t idx = (int64_t)std::floor((double)time / (double)p->time_step);
    return (idx >= 0 && idx < n_in) ? (int)idx : -1;
}

// Output frames [first, first + count) of `vfx apply retime` over a wire clip of n_in frames.
int cref_retime_frames(const twx_retime_params* p, const uint8_t* clip, int n_in, int first, int count, uint8_t* out, char* err,
                       size_t errlen) {
    Clip c{clip, n_in, p->width, p->height, p->fmt};
    size_t fb = image_bytes(p->width, p->height, p->fmt);
    for (int i = 0; i < count; i++) {
        std::string e;
        int st = render_frame(*p, c, first + i, out + (size_t)i * fb, e);
        if (st) { set_err(err, errlen, "frame " + std::to_string(first + i) + ": " + e); return st; }
    }
    return TWX_OK;
}

// Output frames [first, first + count) of `vfx apply vectors-in`: `input` the footage, `source`
// the Color Source layer, `prev` and `next` the vector layers (a null pointer: the layer is not
// set), each with its frame count; gouraud, above_first, flip: the VinStages of the render (the
// effect's own: 1, 1, 1).
int cref_vectors_in_frames(const cref_vin_params* p, const uint8_t* input, int n_input, const uint8_t* source, int n_source,
                           const uint8_t* prev, int n_prev, const uint8_t* next, int n_next, int gouraud, int above_first,
                           int flip, int first, int count, uint8_t* out, char* err, size_t errlen) {
    const size_t fb = image_bytes(p->width, p->height, p->fmt);
    VinLayer li{input, n_input}, ls{source, n_source}, lp{prev, n_prev}, ln{next, n_next};
    const VinStages vs{gouraud != 0, above_first != 0, flip != 0};
    for (int i = 0; i < count; i++) {
        std::string e;
        int st = vin_render_frame(*p, li, ls, lp, ln, vs, first + i, out + (size_t)i * fb, e);
        if (st) { set_err(err, errlen, "frame " + std::to_string(first + i) + ": " + e); return st; }
    }
    return TWX_OK;
}

}  // extern "C"
