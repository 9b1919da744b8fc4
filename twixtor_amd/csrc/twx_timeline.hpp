// Host-side time model of the retime effect: which source frames and weights each output frame
// uses. Scalar per-frame work that stays on the host; it plans the batched device launches.
//
// Follows plugins/retime-engine/src/timeline.rs (rates :108-133, frame_positions :137-142, the
// Speed time map :280-331, frame_time :342-350, exact_position :397-409, select :447-464),
// crates/fx-core/src/curves.rs:344-370 and :537-544 (keyframe sampling), fixed.rs:8-10 (16.16),
// frontend.rs:130-132 and :184-243 (the parameter readers) and clip.rs:170-172 (frame lookup).
// Precision is part of that contract: f32 where the reference narrows, f64 where it widens.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/twx.h"
#include "twx_common.hpp"

namespace twx {
namespace timeline {

// x86 truncations (timeline.rs:43-58)
inline int32_t trunc_i32_f64(double x) { return (x > -2147483649.0 && x < 2147483648.0) ? (int32_t)x : INT32_MIN; }
inline int32_t trunc_i32_f32(float x) { return (x > -2147483904.0f && x < 2147483648.0f) ? (int32_t)x : INT32_MIN; }

// fixed.rs:8-10 (`as i32` saturates, NaN gives 0)
inline int32_t to_fixed(double f) {
    double v = f * 65536.0 + (f < 0.0 ? -0.5 : 0.5);
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int32_t)v;
}

inline float read_fixed(int32_t fixed, float scale) { return (float)((double)fixed * (1.0 / 65536.0)) * scale; }

inline double progress(double u, bool ease_start, bool ease_end) {
    if (ease_start && ease_end) return u * u * (3.0 - 2.0 * u);
    if (ease_start) return u * u * (2.0 - u);
    if (ease_end) return u * (1.0 + u * (1.0 - u));
    return u;
}

// The "Speed %" curve at a tick.
inline double speed_value(const twx_retime_params& p, int64_t tick) {
    int n = p.n_speed_keys;
    if (n <= 0) return p.speed;
    const twx_key* k = p.speed_keys;
    if (tick <= k[0].time) return k[0].value;
    if (tick >= k[n - 1].time) return k[n - 1].value;
    int lo = 0, hi = n - 1;  // last key with time <= tick
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (k[mid].time <= tick) lo = mid; else hi = mid;
    }
    const twx_key& a = k[lo];
    const twx_key& b = k[lo + 1];
    if (a.hold) return a.value;
    double f = progress((double)(tick - a.time) / (double)(b.time - a.time), a.ease != 0, b.ease != 0);
    return a.value + (b.value - a.value) * f;
}

struct Rates {
    double in_fps = 0, out_fps = 0;
    float dur = 0;
    int32_t lts = 0;
};

inline double snap_rate(double rate) {
    static const double targets[3] = {29.97, 59.94, 23.976};
    for (double t : targets)
        if ((double)std::fabs((float)(rate - t)) < 2e-5) rate = t;
    return rate;
}

inline int rates(const twx_retime_params& p, Rates& r, std::string& err) {
    double in = (double)((float)p.input_frame_rate * 1.0f);
    double out = p.fps;
    if (p.in_fps_is_out_fps) in = out = 30.0;
    in = snap_rate(in);
    out = snap_rate(out);
    if (0.01 > out || 0.01 >= in) {
        err = "error 5 (bad parameter): input and output rates must be above 0.01";
        return TWX_E_MISMATCH;
    }
    r.in_fps = in;
    r.out_fps = out;
    r.lts = p.time_step;
    r.dur = (float)(((double)p.time_step * out) / in);
    return TWX_OK;
}

// The Speed integral with the running sums cached (RemapTimeMap::with_speed_cache): the sum after
// k half steps depends only on the samples, so it is computed once per run.
class SpeedMap {
  public:
    SpeedMap(const twx_retime_params& p, const Rates& r) : p_(p), r_(r) { sums_.push_back(0.0f); }

    int src_pos(float t, float& pos, std::string& err) {
        const float lts_f = (float)r_.lts;
        int32_t n2 = trunc_i32_f32(t * 2.0f);
        if (n2 >= (1 << 24)) {
            err = "output position beyond 2^23 frames is not supported";
            return TWX_E_UNSUPPORTED;
        }
        float out = 0.0f, k = 1.0f;
        while ((int64_t)sums_.size() <= (int64_t)n2) {
            float step = (float)sums_.size();
            int32_t tick = trunc_i32_f64((double)((lts_f * step) / 2.0f) + 0.05);
            float sp = (float)((double)to_fixed(speed_value(p_, tick)) / 6553600.0);
            sums_.push_back((sp / 2.0f) + sums_.back());
        }
        if (n2 >= 1) {
            out = sums_[n2];
            k = (float)n2 + 1.0f;
        }
        float tt = t * 2.0f;
        float rem = tt - (k + -1.0f);
        if ((double)rem > 0.0001) {
            int32_t tick = trunc_i32_f64((double)((tt * lts_f) / 2.0f) + 0.05);
            float sp = (float)((double)to_fixed(speed_value(p_, tick)) / 6553600.0);
            out = ((rem * sp) / 2.0f) + out;
        }
        pos = (float)((r_.in_fps / r_.out_fps) * (double)out);
        return TWX_OK;
    }

  private:
    const twx_retime_params& p_;
    Rates r_;
    std::vector<float> sums_;
};

// Layer time of source frame `f` (base 0: no source offset).
inline int32_t frame_time(const Rates& r, float f) {
    float d = f - 0.0f;
    if (std::fabs(r.in_fps - r.out_fps) < 0.05) return trunc_i32_f64((double)d * (double)r.dur + 0.5);
    return trunc_i32_f64((double)(d * r.dur) + (double)r.dur * 0.05);
}

// Clip index shown at source frame position `pos`, or -1 outside the clip (transparent).
inline int64_t clip_index(const Rates& r, int32_t pos, int64_t n_frames) {
    int64_t t = frame_time(r, (float)pos);
    int64_t step = r.lts;
    int64_t q = t / step, m = t % step;
    if (m < 0) q -= 1;  // div_euclid
    return (q >= 0 && q < n_frames) ? q : -1;
}

inline bool is_exact(float s0, float s1, float blur) {
    bool exact = true;
    float frac = s0 - std::floor(s0);
    if ((double)std::fabs(frac) > 0.005 && (double)std::fabs((float)((double)frac - 1.0)) > 0.005) exact = false;
    if ((double)std::fabs(s0 - s1) > 1e-6 && (double)blur > 0.005) exact = false;
    return exact;
}

// Frame selection of output frame n.
inline int select(const twx_retime_params& p, SpeedMap& map, int32_t n, twx_selection& sel, std::string& err) {
    int32_t cur = n * p.time_step;
    float curf = (float)cur, ltsf = (float)p.time_step;
    float t0 = curf / ltsf;
    float t1 = ((float)p.time_step + curf) / ltsf;
    float s0, s1;
    int st = map.src_pos(t0, s0, err);
    if (!st) st = map.src_pos(t1, s1, err);
    if (st) return st;
    int vq = p.vectors - 1;
    if (!(p.width >= 2 && p.height >= 2)) vq = 0;  // prep.rs:106-112, frames with a side of 1
    int32_t a = trunc_i32_f32(s0);
    sel.s0 = s0;
    sel.s1 = s1;
    sel.a = a;
    sel.b = (int32_t)((uint32_t)a + 1u);
    sel.wa = 1.0f - (s0 - (float)a);
    sel.exact = is_exact(s0, s1, 0.0f) ? 1 : 0;
    sel.vectors_quality = vq;
    sel.interp = p.frame_interp - 1;
    return TWX_OK;
}

// ---- the Vectors In effect's time model (plugins/vectors-in/src/timemap.rs, glue.rs:190-253) ----

// Its rates: the input is 30 fps with "In FPS is Out FPS" checked, the slider otherwise; the
// output is always the composition's (Rates::new without its equal-rates switch). Returns true
// for the rate error (out < 0.01 or in <= 0.01), which renders the error fill.
inline bool vin_rates(const twx_vin_params& p, Rates& r) {
    double in = p.in_fps_is_out_fps ? 30.0 : (double)((float)p.input_frame_rate * 1.0f);
    double out = p.fps;
    in = snap_rate(in);
    out = snap_rate(out);
    if (0.01 > out || 0.01 >= in) return true;
    r.in_fps = in;
    r.out_fps = out;
    r.lts = p.time_step;
    r.dur = (float)(((double)p.time_step * out) / in);
    return false;
}

// The Speed parameters of a Vectors In run in the form SpeedMap reads.
inline twx_retime_params vin_speed_params(const twx_vin_params& p) {
    twx_retime_params rp{};
    rp.width = p.width;
    rp.height = p.height;
    rp.fmt = p.fmt;
    rp.time_step = p.time_step;
    rp.fps = p.fps;
    rp.speed = p.speed;
    rp.n_speed_keys = p.n_speed_keys;
    rp.speed_keys = p.speed_keys;
    return rp;
}

// T(k) = trunc((f64)dur * (f64)k + 0.5) (timemap.rs:56-59): the layer time of source frame k.
inline int32_t vin_time(const Rates& r, float k) { return trunc_i32_f64((double)r.dur * (double)k + 0.5); }

// The layer clip index shown at layer time `t` (TimeBase::frame_of), -1 outside the clip.
inline int64_t vin_clip_index(int64_t t, int32_t lts, int64_t n_frames) {
    int64_t q = t / lts, m = t % lts;
    if (m < 0) q -= 1;
    return (q >= 0 && q < n_frames) ? q : -1;
}

// The frame selection of output frame n (timeline.rs:447-464 with the Vectors In time map and a
// constant Mark Segments, whose two markers are equal): the constant vectors quality 4
// (prerender.rs:50), 0 for a frame with a side of 1.
inline int vin_select(const twx_vin_params& p, SpeedMap& map, const Rates& r, int32_t n, twx_selection& sel,
                      std::string& err) {
    const int32_t cur = n * p.time_step;
    const float curf = (float)cur, ltsf = (float)p.time_step;
    const float t0 = curf / ltsf, t1 = ((float)p.time_step + curf) / ltsf;
    float s0, s1;
    if (p.time_remap_mode == 2) {  // Frame Number (timeline.rs:274-279), a constant Frame Num
        s0 = s1 = (float)((r.in_fps / r.out_fps) * (double)(float)p.frame_num);
    } else {
        int st = map.src_pos(t0, s0, err);
        if (!st) st = map.src_pos(t1, s1, err);
        if (st) return st;
    }
    int vq = (p.width >= 2 && p.height >= 2) ? 4 : 0;
    int ip = p.frame_interp - 1;
    const int mark = (p.mark_segments >= 1 && p.mark_segments <= 5) ? p.mark_segments - 1 : 0;
    if (mark == 3) { vq = 0; ip = 0; }
    if (mark == 4) { vq = 0; ip = 1; }
    const int32_t a = trunc_i32_f32(s0);
    sel.s0 = s0;
    sel.s1 = s1;
    sel.a = a;
    sel.b = (int32_t)((uint32_t)a + 1u);
    sel.wa = 1.0f - (s0 - (float)a);
    sel.exact = is_exact(s0, s1, (float)p.blur) ? 1 : 0;
    sel.vectors_quality = vq;
    sel.interp = ip;
    return TWX_OK;
}

inline int check_params(const twx_retime_params& p, std::string& err) {
    if (p.width <= 0 || p.height <= 0 || p.fmt < 0 || p.fmt > 2 || p.time_step <= 0) {
        err = "retime parameters: bad frame size, format or time step";
        return TWX_E_MISMATCH;
    }
    if (p.vectors < 1 || p.vectors > 5 || p.frame_interp < 1 || p.frame_interp > 3) {
        err = "retime parameters: popup out of range";
        return TWX_E_MISMATCH;
    }
    return check_frame_pixels(p.width, p.height, err);
}

}  // namespace timeline
}  // namespace twx
