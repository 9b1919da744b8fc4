// The per-pixel arithmetic of the frame kernels with the reference's semantics, compiled for the
// device (the kernels) and for the host (the open-time gate, which runs these functions on the
// device and compares every result with the host's, bit for bit). Nothing here may depend on the
// side it is compiled for.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>

namespace twx {

#define TWX_HD __host__ __device__ __forceinline__

TWX_HD uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
TWX_HD float u2f(uint32_t b) { return __builtin_bit_cast(float, b); }
TWX_HD float bits_f(uint32_t b) { return u2f(b); }
TWX_HD bool is_nan(float x) { return x != x; }
TWX_HD bool is_nan(double x) { return x != x; }
TWX_HD int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
TWX_HD int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

// flow/sample.rs:18-25: truncation, NaN and out of range give INT_MIN (the x86 cvtt value)
TWX_HD int32_t cvtt(float x) {
    if (is_nan(x) || x >= 2147483648.0f || x < -2147483648.0f) return INT_MIN;
    return (int32_t)x;
}
// blend.rs:18-25 (64 bit truncation of a float; the callers keep the low byte or word)
TWX_HD int64_t cvtt64f(float x) {
    const float two63 = 9223372036854775808.0f;
    if (is_nan(x) || x >= two63 || x < -two63) return LLONG_MIN;
    return (int64_t)x;
}
// image.rs:249-255 (64 bit truncation of a double)
TWX_HD int64_t cvtt64d(double x) {
    if (is_nan(x) || x >= 9.223372036854775808e18 || x < -9.223372036854775808e18) return LLONG_MIN;
    return (int64_t)x;
}
// image.rs:259-261: q(x) = trunc((double)x + 0.5)
TWX_HD int64_t qround(float x) { return cvtt64d((double)x + 0.5); }

// -------------------------------------------------------------------------------------------
// The reference's NaN rule (nanrule.rs:1-120): numbers never depend on it; a NaN result of an f32
// sum, difference, product or quotient takes the one NaN operand, quieted; of two NaN operands the
// one with the larger fraction (equal fractions: the positive one; a quiet NaN before a
// signalling one), quieted; for an invalid operation (inf - inf, 0 * inf, 0 / 0) the x86 default
// NaN 0xffc00000. The f32 kernels compute plainly and repeat a result with these operations only
// where it holds a NaN (a branch no finite pixel takes), as the reference's call sites do.

TWX_HD float nr_quiet(float x) { return u2f(f2u(x) | 0x00400000u); }
TWX_HD float nr_pick(float a, float b) {
    const uint32_t ba = f2u(a), bb = f2u(b);
    const bool sa = !(ba & 0x00400000u), sb = !(bb & 0x00400000u);
    const uint32_t fa = ba & 0x007fffffu, fb = bb & 0x007fffffu;
    const bool a_larger = fa > fb || (fa == fb && (ba >> 31) < (bb >> 31));
    const bool take_b = (sa == sb) ? !a_larger : sa;
    return nr_quiet(take_b ? b : a);
}
TWX_HD float nr_fix(float r, float a, float b) {
    if (!is_nan(r)) return r;
    if (is_nan(a)) return is_nan(b) ? nr_pick(a, b) : nr_quiet(a);
    if (is_nan(b)) return nr_quiet(b);
    return u2f(0xffc00000u);
}
TWX_HD float nr_add(float a, float b) { return nr_fix(a + b, a, b); }
TWX_HD float nr_sub(float a, float b) { return nr_fix(a - b, a, b); }
TWX_HD float nr_mul(float a, float b) { return nr_fix(a * b, a, b); }
TWX_HD float nr_div(float a, float b) { return nr_fix(a / b, a, b); }
// lerp (nanrule.rs:240): (b - a) * f + a; lerp_base_first (:247): a + (b - a) * f
TWX_HD float nr_lerp(float a, float b, float f) { return nr_add(nr_mul(nr_sub(b, a), f), a); }
TWX_HD float nr_lerp_bf(float a, float b, float f) { return nr_add(a, nr_mul(nr_sub(b, a), f)); }

// warp.rs:129-146 store rules
TWX_HD float sat_u8(float v) {
    int32_t i = cvtt(v + 0.5f);
    int32_t t = (int32_t)((uint32_t)i + (uint32_t)-128);
    int32_t w = imin(imax(t, -32768), 32767);
    int32_t b = imin(imax(w, -128), 127);
    return (float)(b + 128);
}
TWX_HD float sat_u16(float v) {
    int32_t i = cvtt(v + 0.5f);
    int32_t t = (int32_t)((uint32_t)i + (uint32_t)-32768);
    int32_t w = imin(imax(t, -32768), 32767);
    return (float)(w + 32768);
}
template <int FMT>
TWX_HD float store_rule(float v) {
    return FMT == 0 ? sat_u8(v) : (FMT == 1 ? sat_u16(v) : v);
}

// lerp of one pixel, weight on the first (blend.rs:180-257)
template <int FMT>
TWX_HD void lerp_pixel(float f[4], const float s[4], float w) {
    if (FMT == 2 && (is_nan(f[0]) || is_nan(f[1]) || is_nan(f[2]) || is_nan(f[3]) || is_nan(s[0]) || is_nan(s[1]) ||
                     is_nan(s[2]) || is_nan(s[3]))) {
        // blend.rs:181-216: a NaN among the inputs, the arithmetic with the NaN rule
        float a = nr_lerp(s[0], f[0], w);
        if (!(a > 0.0f)) { f[0] = f[1] = f[2] = f[3] = 0.0f; return; }
        float r = nr_lerp(s[1], f[1], w);
        float g = nr_lerp(s[2], f[2], w);
        float b = nr_lerp(s[3], f[3], w);
        if (a > 1.0f) { r = nr_div(r, a); g = nr_div(g, a); b = nr_div(b, a); a = 1.0f; }
        f[1] = nr_add(r, 0.0f); f[2] = nr_add(g, 0.0f); f[3] = nr_add(b, 0.0f); f[0] = nr_add(a, 0.0f);
    } else if (FMT == 2) {
        float a = (f[0] - s[0]) * w + s[0];
        if (!(a > 0.0f)) { f[0] = f[1] = f[2] = f[3] = 0.0f; return; }
        float r = (f[1] - s[1]) * w + s[1];
        float g = (f[2] - s[2]) * w + s[2];
        float b = (f[3] - s[3]) * w + s[3];
        if (a > 1.0f) { r = r / a; g = g / a; b = b / a; a = 1.0f; }
        f[1] = r + 0.0f; f[2] = g + 0.0f; f[3] = b + 0.0f; f[0] = a + 0.0f;
    } else {
        const float scale = FMT == 0 ? 255.0f : 65535.0f;
        const int64_t mask = FMT == 0 ? 0xff : 0xffff;
        float a = (float)((int32_t)f[0] - (int32_t)s[0]) * w + s[0];
        float n = a / scale;
        if (!(n > 0.0f)) { f[0] = f[1] = f[2] = f[3] = (float)(cvtt64f(0.0f) & mask); return; }
        float r = (float)((int32_t)f[1] - (int32_t)s[1]) * w + s[1];
        float g = (float)((int32_t)f[2] - (int32_t)s[2]) * w + s[2];
        float b = (float)((int32_t)f[3] - (int32_t)s[3]) * w + s[3];
        if (n > 1.0f) { r = r / n; g = g / n; b = b / n; a = scale; }
        auto clamp = [&](float c) { return c > a ? a : (0.0f > c ? 0.0f : c); };
        r = clamp(r); g = clamp(g); b = clamp(b);
        f[1] = (float)(cvtt64f(r + 0.5f) & mask);
        f[2] = (float)(cvtt64f(g + 0.5f) & mask);
        f[3] = (float)(cvtt64f(b + 0.5f) & mask);
        f[0] = (float)(cvtt64f(a + 0.5f) & mask);
    }
}

// blend(out, first, second, w) of one pixel (blend.rs:55-114); mode decided on the host
template <int FMT>
TWX_HD void blend_pixel(const float p[4], const float q[4], int mode, float w, float lw, float o[4]) {
    if (mode == 0) { for (int k = 0; k < 4; k++) o[k] = p[k]; return; }
    if (mode == 1) { for (int k = 0; k < 4; k++) o[k] = q[k]; return; }
    if (mode == 2) {
        for (int k = 0; k < 4; k++) o[k] = p[k];
        if (lw == 1.0f) return;
        if (lw == 0.0f) { for (int k = 0; k < 4; k++) o[k] = q[k]; return; }
        lerp_pixel<FMT>(o, q, lw);
        return;
    }
    const int64_t mask = FMT == 0 ? 0xff : 0xffff;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (FMT == 2) {
            o[k] = ((q[k] - p[k]) * w + p[k]) + 0.0f;
            if (is_nan(o[k])) o[k] = nr_add(nr_lerp(p[k], q[k], w), 0.0f);  // blend.rs:105-110
        } else {
            o[k] = (float)(cvtt64f(((q[k] - p[k]) * w + p[k]) + 0.5f) & mask);
        }
    }
}

// Image::to_frame of one engine pixel (A, R, G, B) into its wire values, in the engine's channel
// order (image.rs:300-351, frame.rs:113-141): u8 and u16 as integers (u16 already the wire's
// 0..65535), f32 as float bits.
template <int FMT>
TWX_HD void finish_values(const float e[4], uint32_t v[4]) {
    if (FMT == 0) {
        uint8_t p[4] = {(uint8_t)e[0], (uint8_t)e[1], (uint8_t)e[2], (uint8_t)e[3]};
        uint8_t o[4] = {p[0], p[1], p[2], p[3]};
        if (p[0] != 0) {
            float s = 255.0f / (float)p[0];
            for (int c = 1; c < 4; c++) {
                float x = (float)p[c] * s;
                if (x > 255.0f) x = 255.0f;
                o[c] = (uint8_t)qround(x);
            }
        }
        for (int c = 0; c < 4; c++) v[c] = o[c];
    } else if (FMT == 1) {
        uint16_t p[4] = {(uint16_t)e[0], (uint16_t)e[1], (uint16_t)e[2], (uint16_t)e[3]};
        uint16_t o[4] = {p[0], p[1], p[2], p[3]};
        if (p[0] != 0) {
            float a = (float)p[0];
            float s = 32768.0f / a;
            o[0] = (uint16_t)qround(a * bits_f(0x3f000080u));
            for (int c = 1; c < 4; c++) o[c] = (uint16_t)qround((float)p[c] * s);
        }
        for (int c = 0; c < 4; c++) {
            uint32_t m = o[c] < 32768u ? o[c] : 32768u;
            v[c] = (m * 65535u + 16384u) / 32768u;
        }
    } else {
        float o[4];
        if ((double)e[0] < 1e-12) {
            o[0] = o[1] = o[2] = o[3] = 0.0f;
        } else {
            // 1 / NaN: the NaN quieted, as the CPU's division gives it, whatever the hardware's
            // payload rule
            float inv = 1.0f / e[0];
            if (is_nan(inv)) inv = nr_div(1.0f, e[0]);
            o[0] = e[0];
            for (int c = 1; c < 4; c++) o[c] = e[c] * inv;
            if (is_nan(e[0]) || is_nan(o[1]) || is_nan(o[2]) || is_nan(o[3]))  // image.rs:344-348
                for (int c = 1; c < 4; c++) o[c] = nr_mul(e[c], inv);
        }
        for (int c = 0; c < 4; c++) v[c] = f2u(o[c]);
    }
}

#undef TWX_HD

}  // namespace twx
