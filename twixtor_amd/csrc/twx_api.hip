// The C ABI of include/twx.h: device open and gate, the Backend-trait entry points
// (plugins/retime-engine/src/device/exec/backend.rs:27-46) and the batched per-clip render of
// `vfx apply retime` (crates/fx-cli/src/temporal.rs:183-250).
//
// A call plans on the host (frame selection, unique source frames, unique frame pairs), then
// issues batched launches on the handle's stream: ingest of every needed source frame, one
// pyramid build over all of them, one solve over every flow job, one compose-and-finish over
// every output frame. Identical pairs are solved once (the reference solves every in-between
// frame's pair again, compose/vector.rs:163-206; the flows are the same, so are the bytes).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <tuple>
#include <vector>

#include "twx_common.hpp"
#include "twx_timeline.hpp"

using namespace twx;

// Grow-only memory for the calls on a handle: device blocks (work buffers, pyramids, flows) and
// pinned host blocks (the job tables uploaded to the device). A call bump-allocates from the
// start; the same call pattern reuses the same blocks, so a steady state allocates nothing.
struct BlockArena {
    struct Block {
        char* p;
        size_t cap;
    };
    std::vector<Block> blocks;
    size_t cur = 0, off = 0;
    bool pinned = false;
    size_t min_block = 256u << 20;

    void reset() { cur = 0; off = 0; }
    char* take(size_t bytes);
    void release();
};

// A batched render captured into a graph, with its own memory (the arenas the captured launches
// and table copies point into), replayed when a call repeats the same arguments.
struct GraphPlan {
    std::string key;
    hipGraphExec_t exec = nullptr;
    BlockArena dev, host;
    int* bad = nullptr;     // the render's device flags (see BatchFlags)
    int* states = nullptr;
    size_t nsolves = 0;
    twx_stats delta{};      // the counters one render adds
    GraphPlan() { host.pinned = true; }
    ~GraphPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        dev.release();
        host.release();
    }
};

// The resident pyramids behind twx_solve / twx_solve_pair (the reference's frame cache,
// device/exec/cache.rs:1-140, used by backend.rs:191-208): an entry is the device pyramid of one
// host engine image, found again only by the image's exact bytes (its size, format and every
// sample's bits; a digest of a few samples filters first) and the working geometry the settings
// give, so a hit is always a pyramid made from these very samples. Least recently used first out,
// within a quarter of the handle's device budget (the reference's share, backend.rs:147) and at
// most MAX_ENTRIES entries (the keys are host copies of the images).
struct PyrEntry {
    std::vector<uint8_t> key;  // the host image's bytes
    int W = 0, H = 0, fmt = 0, reduction = 0, levels = 0;
    uint64_t digest = 0;
    float* dev = nullptr;      // the pyramid, finest level first
    size_t bytes = 0;          // device bytes of `dev`
    bool ready = false;        // the pyramid's build has completed
};

struct PyramidCache {
    static constexpr size_t MAX_ENTRIES = 64;
    std::vector<std::unique_ptr<PyrEntry>> entries;  // most recently used last
    uint64_t budget = 0, used = 0;
    uint64_t hits = 0, misses = 0, evictions = 0;
    void drop(size_t i) {
        (void)hipFree(entries[i]->dev);
        used -= entries[i]->bytes;
        entries.erase(entries.begin() + (long)i);
    }
    void clear() {
        while (!entries.empty()) drop(entries.size() - 1);
    }
};

struct twx_handle {
    int device = 0;
    hipStream_t stream = nullptr;  // non-blocking: ordered after the caller's work by `caller_ev`
    hipEvent_t caller_ev = nullptr;
    uint64_t budget = 0;
    std::mutex mu;
    std::thread::id owner;
    bool busy = false;
    twx_stats stats{};
    Prof prof;
    BlockArena dev;
    BlockArena host;
    bool graphs = true;                // capture batched renders into graphs and replay them
    std::unique_ptr<GraphPlan> graph;  // the last captured render
    uint64_t graph_replays = 0;
    PyramidCache pyramids;             // twx_solve / twx_solve_pair
};

namespace {

void set_err(char* err, size_t errlen, const std::string& m) {
    if (err && errlen) {
        size_t n = std::min(errlen - 1, m.size());
        memcpy(err, m.data(), n);
        err[n] = 0;
    }
}

struct HipError {
    std::string msg;
};

#define HIPCHECK(expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

}  // namespace

char* BlockArena::take(size_t bytes) {
    bytes = std::max<size_t>((bytes + 255) & ~(size_t)255, 256);
    while (cur < blocks.size()) {
        if (off + bytes <= blocks[cur].cap) {
            char* r = blocks[cur].p + off;
            off += bytes;
            return r;
        }
        cur++;
        off = 0;
    }
    size_t cap = std::max(bytes, min_block);
    void* p = nullptr;
    if (pinned) HIPCHECK(hipHostMalloc(&p, cap, hipHostMallocDefault));
    else HIPCHECK(hipMalloc(&p, cap));
    blocks.push_back(Block{(char*)p, cap});
    cur = blocks.size() - 1;
    off = bytes;
    return (char*)p;
}

void BlockArena::release() {
    for (auto& b : blocks) {
        if (pinned) (void)hipHostFree(b.p);
        else (void)hipFree(b.p);
    }
    blocks.clear();
    reset();
}

namespace {

// The buffers of one call: bump allocations from the handle's arenas (reset when the call
// starts), table uploads through pinned staging. Every call synchronises its stream before it
// returns, so the next call may reuse the memory.
class DevPool {
  public:
    DevPool(twx_handle* h) : dev_(h->dev), host_(h->host), s_(h->stream) { reset(); }
    // over other arenas (a graph plan's own memory)
    DevPool(BlockArena& dev, BlockArena& host, hipStream_t s) : dev_(dev), host_(host), s_(s) { reset(); }
    void reset() {
        dev_.reset();
        host_.reset();
    }
    template <class T>
    T* alloc(size_t count) {
        return (T*)dev_.take(count * sizeof(T));
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* d = alloc<T>(v.size());
        if (!v.empty()) {
            size_t bytes = v.size() * sizeof(T);
            char* st = host_.take(bytes);
            memcpy(st, v.data(), bytes);
            HIPCHECK(hipMemcpyAsync(d, st, bytes, hipMemcpyHostToDevice, s_));
        }
        return d;
    }
    template <class T>
    T* pinned(size_t count) {
        return (T*)host_.take(count * sizeof(T));
    }

  private:
    BlockArena& dev_;
    BlockArena& host_;
    hipStream_t s_;
};

// Serialises the calls on a handle and refuses a nested call from the same thread
// (exec/backend.rs:151-159).
class Guard {
  public:
    Guard(twx_handle* h, int& st, std::string& err) : h_(h) {
        {
            std::lock_guard<std::mutex> g(meta_mu());
            if (h->busy && h->owner == std::this_thread::get_id()) {
                st = TWX_E_UNSUPPORTED;
                err = "nested device call from the same thread";
                return;
            }
        }
        h->mu.lock();
        {
            std::lock_guard<std::mutex> g(meta_mu());
            h->busy = true;
            h->owner = std::this_thread::get_id();
        }
        locked_ = true;
        (void)hipSetDevice(h->device);
        g_prof = &h->prof;
    }
    ~Guard() {
        if (!locked_) return;
        if (h_->prof.on) {
            (void)hipStreamSynchronize(h_->stream);
            h_->prof.flush();
        }
        g_prof = nullptr;
        {
            std::lock_guard<std::mutex> g(meta_mu());
            h_->busy = false;
        }
        h_->mu.unlock();
    }
    bool ok() const { return locked_; }

  private:
    static std::mutex& meta_mu() {
        static std::mutex m;
        return m;
    }
    twx_handle* h_;
    bool locked_ = false;
};

// The handle's stream does not wait for the legacy default stream by itself (it is created
// non-blocking). An entry point that reads or writes the caller's device buffers first makes it
// wait for everything queued on the default stream so far, so device work the caller queued there
// (a fill of the output, the upload of the clip) is complete before the call's kernels touch the
// buffers. Work on other streams of the caller is the caller's to finish before the call.
void after_caller(twx_handle* h) {
    HIPCHECK(hipEventRecord(h->caller_ev, nullptr));
    HIPCHECK(hipStreamWaitEvent(h->stream, h->caller_ev, 0));
}

// ---- geometry (prep.rs:63-146, flow.rs:118-197) ------------------------------------------------

struct Geo {
    int reduction = 0;
    int w0 = 0, h0 = 0;
    int levels = 0;
    int w[MAXL], h[MAXL];
};

int geometry(const twx_flow_settings& fs, int W, int H, Geo& g, std::string& err) {
    if (fs.quality == 0 || fs.quality > 5) { err = "flow quality is outside 1..=5"; return TWX_E_MISMATCH; }
    if (!(fs.sensitivity >= 0.0f) || !std::isfinite(fs.sensitivity)) { err = "sensitivity must be finite and not negative"; return TWX_E_MISMATCH; }
    if (fs.prep_contrast_edge || fs.prep_delinearize || fs.track_alpha || fs.video_range_luma || fs.wrap) {
        err = "prep, track alpha, video range luma and wrap-around solves are not served by this backend";
        return TWX_E_UNSUPPORTED;
    }
    if (int st = check_frame_pixels(W, H, err)) return st;
    int q = fs.quality;
    int r = (q <= 2) ? 3 : (q == 3 ? 2 : (q == 4 ? 1 : 0));
    if (!fs.hiq) r = (r == 0) ? 1 : (r == 1 ? 2 : (r == 2 ? 3 : r));
    while (r > 0 && ((W >> r) < 2 || (H >> r) < 2)) r--;
    int w0 = W, h0 = H;
    for (int i = 0; i < r; i++) { w0 /= 2; h0 /= 2; }
    if (w0 < 2 || h0 < 2) { err = "optical flow: working size below 2x2 is not supported"; return TWX_E_UNSUPPORTED; }
    int levels = 1;
    {
        int w = w0, h = h0;
        for (;;) {
            w /= 2;
            if (w < 2) break;
            h /= 2;
            if (h < 2) break;
            if (w > 2 && h > 2 && !(w == 2 && h == 2)) levels++;
            else break;
        }
    }
    if (levels > MAXL) { err = "pyramid deeper than this backend's level table"; return TWX_E_UNSUPPORTED; }
    if (w0 > 16384 || h0 > 16384) { err = "working size above 16384"; return TWX_E_UNSUPPORTED; }
    g.reduction = r;
    g.w0 = w0;
    g.h0 = h0;
    g.levels = levels;
    int w = w0, h = h0;
    for (int i = 0; i < levels; i++) {
        g.w[i] = w;
        g.h[i] = h;
        w = std::max(w / 2, 2);
        h = std::max(h / 2, 2);
    }
    return TWX_OK;
}

void start_level(float sensitivity, int levels, int& start, float& damp) {
    float p = (float)std::pow((double)sensitivity, 2.0);
    float l1 = (float)std::log((double)p);
    float l2 = (float)std::log(2.0);
    float lv = (float)levels;
    float x = l1 / l2 + lv;
    if (0.0f <= x || std::isnan(x)) {
        if (x < lv || std::isnan(x)) {
            float f = std::floor(x);
            int32_t s = (std::isnan(f) || f >= 2147483648.0f || f < -2147483648.0f) ? INT32_MIN : (int32_t)f;
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

size_t pyramid_floats(const Geo& g) {
    size_t n = 0;
    for (int i = 0; i < g.levels; i++) n += (size_t)g.w[i] * g.h[i];
    return n;
}

// ---- batched pyramid build (flow.rs:242-275, prep.rs:246-259) ----------------------------------

struct PyrIn {
    const void* img;  // engine image on the device
    int W, H, fmt;
    Geo geo;
    float* levels;    // pyramid_floats(geo) floats
    // or the frame's wire bytes (rows reversed into the engine image's order with `reverse`),
    // read by the first step in place of img; `eng` (may be null) then receives the engine image
    const uint8_t* wire = nullptr;
    uint8_t* eng = nullptr;
    int reverse = 0;
};

void build_pyramids(hipStream_t s, DevPool& pool, const std::vector<PyrIn>& in) {
    if (in.empty()) return;
    const Taps bin = binomial_taps();
    const size_t chunk = 128;  // pyramids per launch sequence (the small levels' launches are latency)
    // per frame scratch: the planes of steps 1, 3, ... in B (half size) and of steps 2, 4, ... in A
    // (quarter size); one set for the largest chunk, reused by every chunk in stream order (step 0
    // reads the engine image, and the blurred planes stay in LDS)
    auto r64 = [](size_t v) { return (v + 63) & ~(size_t)63; };
    auto quarter = [&](const PyrIn& p) { return r64((size_t)(p.W / 4 + 1) * (p.H / 4 + 1)); };
    auto frame_scratch = [&](const PyrIn& p) { return quarter(p) + r64((size_t)(p.W / 2 + 1) * (p.H / 2 + 1)); };
    size_t max_floats = 0;
    for (size_t c0 = 0; c0 < in.size(); c0 += chunk) {
        size_t f = 0;
        for (size_t i = c0; i < std::min(in.size(), c0 + chunk); i++) f += frame_scratch(in[i]);
        max_floats = std::max(max_floats, f);
    }
    float* sbase = pool.alloc<float>(max_floats);
    for (size_t c0 = 0; c0 < in.size(); c0 += chunk) {
        size_t c1 = std::min(in.size(), c0 + chunk);
        int n = (int)(c1 - c0);
        std::vector<float*> bufA(n), bufB(n);
        int maxR = 0, maxL = 0;
        float* sp = sbase;
        for (int i = 0; i < n; i++) {
            const PyrIn& p = in[c0 + i];
            bufA[i] = sp;
            sp += quarter(p);
            bufB[i] = sp;
            sp += r64((size_t)(p.W / 2 + 1) * (p.H / 2 + 1));
            maxR = std::max(maxR, p.geo.reduction);
            maxL = std::max(maxL, p.geo.levels);
        }
        // every step is one launch over the chunk; a frame without the step gets an empty op
        std::vector<std::vector<PlaneOp>> steps;
        std::vector<int> kinds;  // 2: blur5 (+ band) with the resample of the blurred plane
        auto level_ptr = [&](int i, int l) {
            const PyrIn& p = in[c0 + i];
            float* base = p.levels;
            for (int k = 0; k < l; k++) base += (size_t)p.geo.w[k] * p.geo.h[k];
            return base;
        };
        // the buffer holding the plane of step k (reductions, then levels): A, B, A, ...
        auto cur = [&](int i, int k) { return (k % 2 == 0) ? bufA[i] : bufB[i]; };
        const PlaneOp empty{};
        // a frame's first step reads its engine image, or its wire frame, and takes the luma there
        // (prep.rs:168-235)
        auto source = [&](int i, int k, PlaneOp& op) {
            if (k == 0) {
                const PyrIn& p = in[c0 + i];
                op.src = p.img;
                op.fmt = p.fmt;
                op.luma = 1;
                op.wire = p.wire;
                op.eng = p.eng;
                op.reverse = p.reverse;
            }
        };
        // reduction: R x (blur5, resample to the halved size)
        for (int r = 0; r < maxR; r++) {
            std::vector<PlaneOp> op(n, empty);
            for (int i = 0; i < n; i++) {
                const PyrIn& p = in[c0 + i];
                if (r >= p.geo.reduction) continue;
                int cw = p.W, ch = p.H;
                for (int k = 0; k < r; k++) { cw /= 2; ch /= 2; }
                op[i] = PlaneOp{cur(i, r), cur(i, r + 1), cw, ch, cw / 2, ch / 2, 0, 0, nullptr};
                source(i, r, op[i]);
            }
            steps.push_back(op); kinds.push_back(2);
        }
        // levels: blur = blur5(orig), band = orig - blur into the pyramid, next = resample(blur)
        // (none after the last level)
        for (int l = 0; l < maxL; l++) {
            std::vector<PlaneOp> op(n, empty);
            for (int i = 0; i < n; i++) {
                const PyrIn& p = in[c0 + i];
                if (l >= p.geo.levels) continue;
                const int k = p.geo.reduction + l;
                int w = p.geo.w[l], h = p.geo.h[l];
                const bool next = l + 1 < p.geo.levels;
                op[i] = PlaneOp{cur(i, k), next ? cur(i, k + 1) : nullptr, w, h, next ? p.geo.w[l + 1] : 0,
                                next ? p.geo.h[l + 1] : 0, 0, 0, level_ptr(i, l)};
                source(i, k, op[i]);
            }
            steps.push_back(op); kinds.push_back(2);
        }
        // the resample ratios: every target is 2 x 2 or larger and halves its source (reductions
        // stop at 2 pixels, pyramid levels are 3 or more), so both ratios are at least 2, which
        // the fused kernel's tiling relies on
        for (size_t k = 0; k < steps.size(); k++)
            for (auto& op : steps[k])
                if ((op.src || op.wire) && op.dst && kinds[k] == 2) {
                    op.rx = ((float)op.sw - 1.0f) / (float)(op.dw - 1);
                    op.ry = ((float)op.sh - 1.0f) / (float)(op.dh - 1);
                    if (!(op.rx >= 2.0f && op.ry >= 2.0f)) throw HipError{"pyramid: a level that does not halve"};
                }
        std::vector<PlaneOp> flat;
        for (auto& st : steps) flat.insert(flat.end(), st.begin(), st.end());
        PlaneOp* d = pool.upload(flat);
        for (size_t k = 0; k < steps.size(); k++) {
            const PlaneOp* dk = d + k * n;
            int mw = 0, mh = 0;
            for (auto& op : steps[k]) {
                mw = std::max({mw, op.sw, op.dw});
                mh = std::max({mh, op.sh, op.dh});
            }
            if (mw == 0 || mh == 0) continue;
            switch (kinds[k]) {
                case 2: launch_blur5_resample(s, dk, steps[k], mw, mh, bin); break;
            }
        }
    }
}

// ---- batched solves ---------------------------------------------------------------------------

struct SolveIn {
    const float* X;  // pyramid levels (concatenated) of X and Y
    const float* Y;
    Geo geo;
    float* out;      // w0 * h0 * 2 floats, the flow X(p) ~ Y(p + out(p))
};

// The coarsest level the per-level kernels run for a solve of these level sizes: the levels above
// it, from the start level down while a level fits in LDS, run in the coarse solver kernel.
static int coarse_top(const int* w, const int* h, int start) {
    if (!coarse_solver_enabled()) return start;
    int l = start;
    while (l >= 0 && (long)w[l] * h[l] <= COARSE_MAX_PIXELS) l--;
    return l;
}

// Enqueues every solve, in chunks that fit the scratch budget, with no host synchronisation:
// solve k keeps its state in states[4k .. 4k + 3] (device; [2] is set for a NaN in its flow).
int enqueue_solves(hipStream_t s, DevPool& pool, const std::vector<SolveIn>& in, const twx_flow_settings& fs,
                   uint64_t scratch_budget, int* states, std::string& err, bool raw) {
    if (in.empty()) return TWX_OK;
    // chunks by the scratch bytes of each job: flow[0] 2 + grad 2 + tmp 3 + tb 2 + ta 3 + block-matching
    // costs 2 floats/pixel
    std::vector<std::pair<size_t, size_t>> chunks;
    size_t max_px = 0;
    for (size_t i0 = 0; i0 < in.size();) {
        size_t i1 = i0, px_sum = 0;
        while (i1 < in.size()) {
            size_t px = (size_t)in[i1].geo.w0 * in[i1].geo.h0;
            if (i1 > i0 && (px_sum + px) * 14 * sizeof(float) > scratch_budget) break;
            px_sum += px;
            i1++;
        }
        chunks.push_back({i0, i1});
        max_px = std::max(max_px, px_sum);
        i0 = i1;
    }
    // one scratch set, reused by every chunk in stream order
    float* scratch = pool.alloc<float>(max_px * 14);
    for (auto& ch : chunks) {
        int n = (int)(ch.second - ch.first);
        // Job order: area descending, in size buckets (one launch sequence each); within a bucket
        // first one job per distinct X (it computes X's gradient and blurred tensor), then the jobs
        // whose X an earlier job of the bucket has: they read that job's grad and ta planes, since
        // both depend on X alone (flow.rs:1274-1277). Each bucket also lists its distinct frames,
        // whose transposed levels the row sweeps of all its jobs read.
        std::vector<int> ord(n);
        for (int k = 0; k < n; k++) ord[k] = k;
        auto area = [&](int k) { return (long)in[ch.first + k].geo.w0 * in[ch.first + k].geo.h0; };
        std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return area(a) > area(b); });
        std::vector<double> areas;
        for (int k : ord) areas.push_back((double)area(k));
        std::vector<int> order, owner(n, -1);  // owner[k]: position in `order` of job k's tensor job
        std::vector<SolveBucket> buckets;
        std::vector<const float*> frames;      // the bucket's distinct frames, bucket after bucket
        std::vector<int> frame_geo;            // a job (index into `in` of the chunk) with the frame
        for (auto [b0, b1] : size_buckets(areas)) {
            SolveBucket B{(int)order.size(), 0, (int)frames.size(), 0, nullptr};
            std::map<const float*, int> first;
            std::set<const float*> seen;
            std::vector<int> shared;
            for (int i = b0; i < b1; i++) {
                const int k = ord[i];
                const SolveIn& si = in[ch.first + k];
                auto it = first.find(si.X);
                if (it == first.end()) {
                    first[si.X] = (int)order.size();
                    order.push_back(k);
                } else {
                    owner[k] = it->second;
                    shared.push_back(k);
                }
                for (const float* f : {si.X, si.Y})
                    if (seen.insert(f).second) {
                        frames.push_back(f);
                        frame_geo.push_back(k);
                    }
            }
            order.insert(order.end(), shared.begin(), shared.end());
            B.j1 = (int)order.size();
            B.f1 = (int)frames.size();
            buckets.push_back(B);
        }
        // the frames' transposed planes, one level at a time (the largest, level 0, sized)
        std::vector<PlaneT> planes(frames.size());
        std::map<const float*, const float*> tplane;
        for (size_t f = 0; f < frames.size(); f++) {
            const SolveIn& si = in[ch.first + frame_geo[f]];
            PlaneT& P = planes[f];
            memset(&P, 0, sizeof(P));
            const float* lv = frames[f];
            for (int l = 0; l < si.geo.levels; l++) {
                P.lv[l] = lv;
                P.w[l] = si.geo.w[l];
                P.h[l] = si.geo.h[l];
                lv += (size_t)si.geo.w[l] * si.geo.h[l];
            }
            float damp;
            int start;
            start_level(fs.sensitivity, si.geo.levels, start, damp);
            P.top = coarse_top(P.w, P.h, start);
            P.dst = pool.alloc<float>((size_t)si.geo.w0 * si.geo.h0);
            tplane[frames[f]] = P.dst;
        }
        std::vector<SolveJob> jobs(n);
        float* sp = scratch;
        for (int q = 0; q < n; q++) {
            const int k = order[q];
            const SolveIn& si = in[ch.first + k];
            SolveJob& J = jobs[q];
            memset(&J, 0, sizeof(J));
            const float* xl = si.X;
            const float* yl = si.Y;
            for (int l = 0; l < si.geo.levels; l++) {
                J.X[l] = xl;
                J.Y[l] = yl;
                J.w[l] = si.geo.w[l];
                J.h[l] = si.geo.h[l];
                xl += (size_t)si.geo.w[l] * si.geo.h[l];
                yl += (size_t)si.geo.w[l] * si.geo.h[l];
            }
            J.levels = si.geo.levels;
            start_level(fs.sensitivity, si.geo.levels, J.start, J.damp);
            J.top = coarse_top(J.w, J.h, J.start);
            J.scale_damp = (std::fabs((double)J.damp - 1.0) >= 1e-6 && J.damp != 1.0f) ? 1 : 0;
            size_t px = (size_t)si.geo.w0 * si.geo.h0;
            J.flow[0] = sp;      // where each level starts (the upsample target)
            J.flow[1] = si.out;  // where each level ends: level 0 ends in the output
            J.grad = sp + px * 2;
            J.tmp = sp + px * 4;  // 8-byte aligned (flow pixels are moved as float2 through it)
            J.tb = sp + px * 7;
            J.ta = sp + px * 9;
            J.cost[0] = sp + px * 12;
            J.cost[1] = sp + px * 13;
            if (owner[k] >= 0) {
                J.grad = jobs[owner[k]].grad;
                J.ta = jobs[owner[k]].ta;
                J.tensor_shared = 1;
            }
            sp += px * 14;
            J.xt = tplane[si.X];
            J.yt = tplane[si.Y];
            J.state = states + (ch.first + k) * 4;
            J.max_w = si.geo.w0;
            J.max_h = si.geo.h0;
            solve_job_ratios(J);
        }
        int* queues = pool.alloc<int>(buckets.size());  // zeroed by each bucket's init kernel
        for (size_t b = 0; b < buckets.size(); b++) buckets[b].queue = queues + b;
        SolveJob* d = pool.upload(jobs);
        PlaneT* dp = pool.upload(planes);
        launch_solve(s, d, jobs, dp, planes, buckets, fs.skip_refine != 0, raw);  // starts with the per-job init kernel
    }
    (void)err;
    return TWX_OK;
}

// After the stream's work: TWX_E_NAN_OUTPUT when a solved flow holds a NaN (the reference's
// download_flow refusal, device/solver/mod.rs:575-582).
int check_solve_states(hipStream_t s, DevPool& pool, const int* states, size_t n, std::string& err) {
    if (!n) return TWX_OK;
    int* hs = pool.pinned<int>(n * 4);
    HIPCHECK(hipMemcpyAsync(hs, states, n * 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    for (size_t k = 0; k < n; k++)
        if (hs[k * 4 + 2]) { err = "the solved flow holds a NaN"; return TWX_E_NAN_OUTPUT; }
    return TWX_OK;
}

// Runs every solve and checks the flows. Returns a status.
int run_solves(hipStream_t s, DevPool& pool, const std::vector<SolveIn>& in, const twx_flow_settings& fs,
               uint64_t scratch_budget, std::string& err) {
    int* states = pool.alloc<int>(std::max<size_t>(in.size(), 1) * 4);
    int st = enqueue_solves(s, pool, in, fs, scratch_budget, states, err, false);
    if (st) return st;
    return check_solve_states(s, pool, states, in.size(), err);
}

// ---- frame composition tables -----------------------------------------------------------------

// The host-side branch decisions of blend (blend.rs:55-72) for weight w.
void blend_branch(float w, FrameJob& F) {
    F.weight = w;
    if (0.005 > (double)std::fabs(w)) F.blend_mode = 0;
    else if (0.005 > std::fabs((double)w - 1.0)) F.blend_mode = 1;
    else if (w < 0.0f || w > 1.0f) { F.blend_mode = 2; F.lerp_w = (float)(1.0 - (double)w); }
    else F.blend_mode = 3;
}

// The limit of the difference mask (filters.rs:287-293): ((0.4 * white) * 0.4) * 0.4 * 3.
float mask_limit(int fmt) {
    const float t = 0.4f;
    float white = fmt == TWX_FMT_U8 ? 255.0f : (fmt == TWX_FMT_U16 ? 65535.0f : 1.0f);
    return t * white * t * t * 3.0f;
}

// The strength of the unsharp step at position t (filters.rs:628-633): -min(t, 1 - t), with the
// inner value narrowed to float and negated there.
float sharpen_weight(float t) {
    double x = 1.0 - std::fabs(0.5 - (double)t) * 2.0;
    float neg = -(float)x;
    return (float)((double)neg * 0.5);
}

// A Weighted Blend frame of the vector path (kind 3): buffers of the warped frames and of the two
// change masks, the limit, the sharpening weights (frac for A, wa for B) and the blend branch.
void setup_weighted(DevPool& pool, FrameJob& F, bool above_first) {
    size_t fb = image_bytes(F.W, F.H, F.fmt);
    size_t np = (size_t)F.W * F.H;
    F.kind = 3;
    F.ta = pool.alloc<uint8_t>(fb);
    F.tb = pool.alloc<uint8_t>(fb);
    for (int hh = 0; hh < 2; hh++)
        for (int b = 0; b < 2; b++) F.mask[hh][b] = pool.alloc<uint8_t>(np);
    F.limit = mask_limit(F.fmt);
    F.ca = sharpen_weight(F.frac);
    F.cb = sharpen_weight(F.wa);
    F.above_first = above_first ? 1 : 0;
    F.mask_small = (F.W < 3 || F.H < 3) ? 1 : 0;
    blend_branch(F.frac, F);
}

// A Smart Blend frame of the vector path (warping code 1): kind 4 Nearest, 5 Blend, 6 Weighted
// Blend (the Weighted Blend buffers as well). The composite goes to its own image when the frame
// is finished into wire bytes, else straight into `out`; the hole fill then works in place.
// `has_hole` must be cleared before the first kernel.
void setup_smart(DevPool& pool, FrameJob& F, int interp, bool above_first, int* has_hole) {
    if (interp == 2) setup_weighted(pool, F, above_first);
    F.kind = 4 + interp;
    F.above_first = above_first ? 1 : 0;
    F.comp = F.finish ? (void*)pool.alloc<uint8_t>(image_bytes(F.W, F.H, F.fmt)) : F.out;
    F.holes = pool.alloc<uint8_t>((size_t)F.W * F.H);
    F.has_hole = has_hole;
    F.nearest_from_a = !(0.5f > F.wa) ? 1 : 0;  // nearest_smart (smartblend.rs:330-337)
}

// The level sizes of the hole-fill pyramid (holefill.rs:145-176): halve (at least 2) until both
// sides are 2; the push starts at `top`.
void hole_geometry(int w, int h, std::vector<std::pair<int, int>>& dims, int& top) {
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
    dims.clear();
    int cw = w, ch = h;
    for (int i = 0; i <= count; i++) {
        dims.push_back({cw, ch});
        cw = std::max(cw / 2, 2);
        ch = std::max(ch / 2, 2);
    }
    top = count - 1;
}

// One image of a batched hole fill: `chans` planes of `img` (4 of an engine image in its format,
// or 2 of a flow with fmt FILL_FLOW2), the holes mask and the has-a-hole flag that gates it.
struct FillItem {
    int W, H;
    void* img;
    int fmt, chans;
    const uint8_t* holes;
    const int* gate;
    int above_first;
};

// The push-pull hole fill of a set of images in chunks whose buffers fit a few GB; an image
// without holes passes through every kernel at once.
int fill_items(hipStream_t s, DevPool& pool, const std::vector<FillItem>& items, std::string& err) {
    if (items.empty()) return TWX_OK;
    // per image: its pull pyramids (values of each plane and the weights), push pyramids, two
    // level-0 scratch planes per pull plane
    size_t per_item = 0;
    for (const FillItem& it : items) {
        std::vector<std::pair<int, int>> dims;
        int top;
        hole_geometry(it.W, it.H, dims, top);
        if (top + 1 > MAXFL) { err = "hole fill: an image too large for the fill pyramid"; return TWX_E_UNSUPPORTED; }
        size_t pyr = 0;
        for (int l = 0; l <= top; l++) pyr += (size_t)dims[l].first * dims[l].second;
        const size_t np = (size_t)it.chans + 1;
        // every buffer rounded up to 64 floats
        per_item = std::max(per_item, 4 * ((2 * np - 1) * pyr + 2 * np * (size_t)it.W * it.H +
                                           64 * ((2 * np - 1) * (size_t)(top + 1) + 2 * np)));
    }
    const size_t chunk = std::max<size_t>(1, std::min<size_t>(items.size(), ((size_t)4 << 30) / per_item));
    char* scratch = pool.alloc<char>(chunk * per_item);
    for (size_t c0 = 0; c0 < items.size(); c0 += chunk) {
        std::vector<FillPlane> planes;
        for (size_t k = c0; k < std::min(items.size(), c0 + chunk); k++) {
            const FillItem& it = items[k];
            std::vector<std::pair<int, int>> dims;
            int top;
            hole_geometry(it.W, it.H, dims, top);
            float* base = reinterpret_cast<float*>(scratch + (k - c0) * per_item);
            auto take = [&](size_t n) { float* p = base; base += (n + 63) & ~(size_t)63; return p; };
            FillPlane fp[5];
            memset(fp, 0, sizeof(fp));
            // the colour planes 0..chans-1 and the weight plane (chan 4)
            int chans[5], nplanes = 0;
            for (int c = 0; c < it.chans; c++) chans[nplanes++] = c;
            chans[nplanes++] = 4;
            for (int q = 0; q < nplanes; q++) {
                FillPlane& P = fp[q];
                P.top = top;
                P.chan = chans[q];
                P.gate = it.gate;
                P.img = it.img;
                P.holes = it.holes;
                P.fmt = it.fmt;
                P.above_first = it.above_first;
                for (int l = 0; l <= top; l++) {
                    P.w[l] = dims[l].first;
                    P.h[l] = dims[l].second;
                    P.val[l] = take((size_t)dims[l].first * dims[l].second);
                    if (P.chan < 4) P.sol[l] = take((size_t)dims[l].first * dims[l].second);
                }
                P.t1 = take((size_t)it.W * it.H);
                P.t2 = take((size_t)it.W * it.H);
            }
            // the coarse levels that fit in one workgroup's LDS (from level 1 on; the 2x2 top level
            // always does)
            int lo = top + 1;
            for (int l = top; l >= 1; l--) {
                size_t pyr_l = 0;
                for (int k = l; k <= top; k++) pyr_l += (size_t)dims[k].first * dims[k].second;
                if (3 * pyr_l + 2 * (size_t)dims[l].first * dims[l].second > (size_t)FILL_LDS_FLOATS) break;
                lo = l;
            }
            if (lo > top) {
                err = "hole fill: the top pyramid level does not fit in LDS";
                return TWX_E_DEVICE;
            }
            for (int q = 0; q < nplanes; q++) {
                fp[q].lo = lo;
                for (int l = 0; l <= top; l++) fp[q].wts[l] = fp[nplanes - 1].val[l];
            }
            for (int q = 0; q < nplanes; q++) planes.push_back(fp[q]);
        }
        FillPlane* dp = pool.upload(planes);
        launch_fill(s, dp, planes);
    }
    return TWX_OK;
}

// The hole fill of every Smart Blend frame of `jobs` (kinds 4-6).
int run_fills(hipStream_t s, DevPool& pool, const std::vector<FrameJob>& jobs, std::string& err) {
    std::vector<FillItem> items;
    for (const FrameJob& F : jobs)
        if (F.kind >= 4) items.push_back(FillItem{F.W, F.H, F.comp, F.fmt, 4, F.holes, F.has_hole, F.above_first});
    return fill_items(s, pool, items, err);
}

// The |amount| early-out of inv_warp (warp.rs:749-764): true means a plain copy.
bool warp_is_copy(float amount, int fw, int W) {
    float a = amount < 0.0f ? -amount : amount;
    if (0.005 > (double)a) return true;
    float r = (float)(fw - 1) / (float)(W - 1);
    float q = (float)((double)amount / (double)r);
    if (q < 0.0f) q = -q;
    return 0.0001 > (double)q;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

twx_flow_settings settings_for(const twx_retime_params& p, int vq) {
    twx_flow_settings fs{};
    fs.quality = vq;
    fs.hiq = p.quality_high;
    fs.sensitivity = timeline::read_fixed(timeline::to_fixed(p.sensitivity), 1.0f) / 100.0f;
    return fs;
}

// One clip of a batched render: output frames [first, first + count) of a retime run over a clip
// of n_in wire frames, clip and output on the device.
struct ClipRun {
    twx_retime_params p;
    const uint8_t* clip;
    int n_in, first, count;
    uint8_t* out;
};

// What a batched render leaves on the device for the host to check after the stream's work: the
// per-frame flags (bad[0]: 1 a NaN warp coordinate, 2 non-finite samples in the Weighted Blend
// without vectors) and the solves' states.
struct BatchFlags {
    int* bad = nullptr;
    int* states = nullptr;
    size_t nsolves = 0;
};

// The batched render of every output frame of every run: one plan on the host, then one ingest of
// every source frame, one pyramid build, one solve per flow-settings group, one compose-and-finish
// launch set over every output frame. Frame pairs are keyed per run, so identical pairs of a clip
// are solved once. Everything is enqueued on the handle's stream with no host synchronisation and
// no memset, so the same call can be captured into a graph; `delta` receives the counters.
int retime_batch_enqueue(twx_handle* h, DevPool& pool, const std::vector<ClipRun>& runs, BatchFlags& flags,
                         twx_stats& delta, std::string& err) {
    hipStream_t s = h->stream;
    int st;
    struct Plan {
        twx_selection sel;
        int vector;   // 0 fast path, 1 vector path
        int sa, sb;   // slots of frames a and b
    };
    struct Slot {
        int run;
        long idx;     // clip index, -1 transparent
        uint8_t* eng;
    };
    std::vector<std::vector<Plan>> plans(runs.size());
    std::vector<Slot> slots;
    std::map<std::pair<int, long>, int> slot_of;
    long total_frames = 0;
    for (size_t r = 0; r < runs.size(); r++) {
        const ClipRun& R = runs[r];
        if ((st = timeline::check_params(R.p, err))) return st;
        if (R.count < 0 || R.n_in < 0) { err = "negative frame count"; return TWX_E_MISMATCH; }
        timeline::Rates rates;
        if ((st = timeline::rates(R.p, rates, err))) return st;
        timeline::SpeedMap map(R.p, rates);
        const size_t fb = image_bytes(R.p.width, R.p.height, R.p.fmt);
        auto slot = [&](long idx) {
            auto key = std::make_pair((int)r, idx);
            auto it = slot_of.find(key);
            if (it != slot_of.end()) return it->second;
            int k = (int)slots.size();
            slots.push_back(Slot{(int)r, idx, pool.alloc<uint8_t>(fb)});
            slot_of[key] = k;
            return k;
        };
        plans[r].resize(R.count);
        for (int i = 0; i < R.count; i++) {
            Plan& pl = plans[r][i];
            if ((st = timeline::select(R.p, map, R.first + i, pl.sel, err))) return st;
            pl.vector = (pl.sel.vectors_quality != 0 && !pl.sel.exact) ? 1 : 0;
            if (pl.sel.interp < 0 || pl.sel.interp > 2) { err = "frame_interp outside 1..3"; return TWX_E_MISMATCH; }
            pl.sa = slot(timeline::clip_index(rates, pl.sel.a, R.n_in));
            pl.sb = slot(timeline::clip_index(rates, pl.sel.b, R.n_in));
        }
        total_frames += R.count;
    }

    // pyramids of the frames the vector path solves on, and the unique solves; flow settings come
    // from the run (quality flag, sensitivity) and the frame's vectors quality
    std::map<std::pair<int, int>, int> pyr_of;  // (slot, vq) -> pyramid
    std::vector<PyrIn> pyrs;
    struct SolveKey {
        int x, y, vq;
        bool operator<(const SolveKey& o) const { return std::tie(x, y, vq) < std::tie(o.x, o.y, o.vq); }
    };
    std::map<SolveKey, int> solve_of;
    std::vector<SolveIn> solves;
    std::vector<std::tuple<int, int, uint32_t>> solve_group;  // (vq, hiq, sensitivity bits)
    std::map<std::pair<int, int>, Geo> geo_of;                 // (run, vq)
    auto geo_for = [&](int r, int vq, Geo& g) -> int {
        auto it = geo_of.find({r, vq});
        if (it != geo_of.end()) { g = it->second; return TWX_OK; }
        twx_flow_settings fs = settings_for(runs[r].p, vq);
        int e = geometry(fs, runs[r].p.width, runs[r].p.height, g, err);
        if (!e) geo_of[{r, vq}] = g;
        return e;
    };
    auto pyramid = [&](int sl, int vq) -> int {
        auto it = pyr_of.find({sl, vq});
        if (it != pyr_of.end()) return it->second;
        const ClipRun& R = runs[slots[sl].run];
        Geo g;
        (void)geo_for(slots[sl].run, vq, g);
        PyrIn pi{slots[sl].eng, R.p.width, R.p.height, R.p.fmt, g, pool.alloc<float>(pyramid_floats(g))};
        int k = (int)pyrs.size();
        pyrs.push_back(pi);
        pyr_of[{sl, vq}] = k;
        return k;
    };
    auto solve = [&](int sx, int sy, int vq) -> int {
        SolveKey key{sx, sy, vq};
        auto it = solve_of.find(key);
        if (it != solve_of.end()) { delta.pairs_reused++; return it->second; }
        int px = pyramid(sx, vq), py = pyramid(sy, vq);
        Geo g = pyrs[px].geo;
        SolveIn si{pyrs[px].levels, pyrs[py].levels, g, pool.alloc<float>((size_t)g.w0 * g.h0 * 2)};
        int k = (int)solves.size();
        solves.push_back(si);
        twx_flow_settings fs = settings_for(runs[slots[sx].run].p, vq);
        uint32_t sb;
        memcpy(&sb, &fs.sensitivity, 4);
        solve_group.push_back(std::make_tuple(vq, fs.hiq, sb));
        solve_of[key] = k;
        return k;
    };
    struct FrameSolves {
        int f_ab = -1, f_ba = -1;
    };
    std::vector<std::vector<FrameSolves>> fsv(runs.size());
    for (size_t r = 0; r < runs.size(); r++) {
        fsv[r].resize(plans[r].size());
        for (size_t i = 0; i < plans[r].size(); i++) {
            const Plan& pl = plans[r][i];
            if (!pl.vector) continue;
            Geo g;
            if ((st = geo_for((int)r, pl.sel.vectors_quality, g))) return st;
            int vq = pl.sel.vectors_quality;
            if (pl.sel.interp == 0) {
                // Nearest: one solve, the one nearest_uses_f_ab names (compose/vector.rs:170-176)
                if (pl.sel.wa < 0.5f) fsv[r][i].f_ab = solve(pl.sa, pl.sb, vq);
                else fsv[r][i].f_ba = solve(pl.sb, pl.sa, vq);
            } else {
                fsv[r][i].f_ba = solve(pl.sb, pl.sa, vq);
                fsv[r][i].f_ab = solve(pl.sa, pl.sb, vq);
            }
        }
    }
    // A frame with a pyramid is read from the clip by its pyramids' first step, and the first of
    // them writes its engine image (the ingest's bytes) on the way; every other source frame goes
    // through the ingest (rows reversed for the motion path), largest first.
    std::vector<char> fused(slots.size(), 0);
    for (auto& [key, k] : pyr_of) {
        const int sl = key.first;
        const Slot& S = slots[sl];
        if (S.idx < 0) continue;  // a transparent frame: the ingest writes its zeros
        const ClipRun& R = runs[S.run];
        pyrs[k].wire = R.clip + (size_t)S.idx * image_bytes(R.p.width, R.p.height, R.p.fmt);
        pyrs[k].reverse = 1;
        if (!fused[sl]) pyrs[k].eng = S.eng;
        fused[sl] = 1;
    }
    {
        std::vector<IngestOp> ops;
        for (size_t k = 0; k < slots.size(); k++) {
            const Slot& sl = slots[k];
            if (fused[k]) continue;
            const ClipRun& R = runs[sl.run];
            const size_t fb = image_bytes(R.p.width, R.p.height, R.p.fmt);
            IngestOp op{};
            op.wire = sl.idx >= 0 ? R.clip + (size_t)sl.idx * fb : nullptr;
            op.engine = sl.eng;
            op.W = R.p.width;
            op.H = R.p.height;
            op.fmt = R.p.fmt;
            op.zero = sl.idx < 0 ? 1 : 0;
            op.reverse = 1;
            ops.push_back(op);
        }
        std::stable_sort(ops.begin(), ops.end(), [](const IngestOp& a, const IngestOp& b) {
            return (long)a.W * a.H > (long)b.W * b.H;
        });
        if (!ops.empty()) launch_ingest_batch(s, pool.upload(ops), ops);
    }
    {
        std::vector<PyrIn> sorted = pyrs;
        std::stable_sort(sorted.begin(), sorted.end(),
                         [](const PyrIn& a, const PyrIn& b) { return (long)a.W * a.H > (long)b.W * b.H; });
        build_pyramids(s, pool, sorted);
    }
    delta.pyramid_builds += pyrs.size();
    delta.pair_solves += solves.size();
    flags.nsolves = solves.size();
    flags.states = pool.alloc<int>(std::max<size_t>(solves.size(), 1) * 4);
    {
        std::map<std::tuple<int, int, uint32_t>, std::vector<SolveIn>> groups;
        for (size_t k = 0; k < solves.size(); k++) groups[solve_group[k]].push_back(solves[k]);
        size_t state_off = 0;
        for (auto& kv : groups) {
            twx_flow_settings fs{};
            fs.quality = std::get<0>(kv.first);
            fs.hiq = std::get<1>(kv.first);
            uint32_t sb = std::get<2>(kv.first);
            memcpy(&fs.sensitivity, &sb, 4);
            std::stable_sort(kv.second.begin(), kv.second.end(), [](const SolveIn& a, const SolveIn& b) {
                return (long)a.geo.w0 * a.geo.h0 > (long)b.geo.w0 * b.geo.h0;
            });
            // the flows' only readers are this render's inverse warps (flow_raw below)
            if ((st = enqueue_solves(s, pool, kv.second, fs, h->budget / 2, flags.states + state_off * 4, err, true)))
                return st;
            state_off += kv.second.size();
        }
    }

    // compose and finish every output frame; the NaN flag and one has-a-hole flag per output frame
    int* bad = pool.alloc<int>(1 + (size_t)total_frames);
    launch_zero(s, bad, 1 + (size_t)total_frames);
    flags.bad = bad;
    std::vector<FrameJob> jobs;
    jobs.reserve(total_frames);
    for (size_t r = 0; r < runs.size(); r++) {
        const ClipRun& R = runs[r];
        const int W = R.p.width, H = R.p.height;
        const size_t fb = image_bytes(W, H, R.p.fmt);
        for (size_t i = 0; i < plans[r].size(); i++) {
            const Plan& pl = plans[r][i];
            FrameJob F;
            memset(&F, 0, sizeof(F));
            F.a = slots[pl.sa].eng;
            F.b = slots[pl.sb].eng;
            F.out = R.out + i * fb;
            F.bad = bad;
            F.fmt = R.p.fmt;
            F.W = W;
            F.H = H;
            F.finish = 1;
            F.wa = pl.sel.wa;
            F.frac = 1.0f - pl.sel.wa;
            if (!pl.vector) {
                F.kind = 0;
                delta.fast_frames++;
                if (pl.sel.interp == 0) {
                    F.blend_mode = pl.sel.wa >= 0.5f ? 0 : 1;
                } else if (pl.sel.interp == 1) {
                    blend_branch((float)(1.0 - (double)pl.sel.wa), F);
                } else {
                    // weighted_blend_noflow (blend.rs:280-300): Blend with 1 - wa in f32, float inputs + 0.0
                    blend_branch(1.0f - pl.sel.wa, F);
                    F.plus_zero = 1;
                }
            } else {
                if (W < 2 || H < 2) { err = "inverse warp needs images of at least 2x2"; return TWX_E_MISMATCH; }
                delta.composes++;
                Geo g = geo_of[{(int)r, pl.sel.vectors_quality}];
                F.fw = g.w0;
                F.fh = g.h0;
                if (fsv[r][i].f_ab >= 0) F.f_ab = solves[fsv[r][i].f_ab].out;
                if (fsv[r][i].f_ba >= 0) F.f_ba = solves[fsv[r][i].f_ba].out;
                F.flow_raw = 1;
                F.warp_a_copy = warp_is_copy(F.frac, g.w0, W) ? 1 : 0;
                F.warp_b_copy = warp_is_copy(F.wa, g.w0, W) ? 1 : 0;
                if (R.p.smart_blend) {
                    // warping code 1 (frontend.rs:228); the retime front-end's stages take the row
                    // above first in the hole fill and the sharpening blur (frontend.rs:238-241)
                    setup_smart(pool, F, pl.sel.interp, true, bad + 1 + jobs.size());
                } else if (pl.sel.interp == 0) {
                    F.kind = 1;
                    F.nearest_from_a = 0.5f <= pl.sel.wa ? 1 : 0;
                } else if (pl.sel.interp == 1) {
                    F.kind = 2;
                    blend_branch(F.frac, F);
                } else {
                    // Weighted Blend: the retime front-end's stages take the row above first in
                    // the blur of the sharpening step (frontend.rs:238-241)
                    setup_weighted(pool, F, true);
                }
            }
            jobs.push_back(F);
        }
    }
    // largest first; within a size, the blend pairs together (they run in a kernel of their own)
    std::stable_sort(jobs.begin(), jobs.end(), [](const FrameJob& a, const FrameJob& b) {
        const long aa = (long)a.W * a.H, ab = (long)b.W * b.H;
        return aa != ab ? aa > ab : frame_is_blend_pair(a) > frame_is_blend_pair(b);
    });
    if (!jobs.empty()) {
        for (auto& F : jobs) frame_job_ratios(F);
        FrameJob* dj = pool.upload(jobs);
        launch_frames(s, dj, jobs);
        launch_weighted(s, dj, jobs);
        if ((st = run_fills(s, pool, jobs, err))) return st;
        launch_finish(s, dj, jobs);
    }
    return TWX_OK;
}

// After the stream's work of a batched render: the NaN flags of the solves and the frame flags.
// `pool` only provides the pinned read-back buffers.
int retime_batch_check(twx_handle* h, DevPool& pool, const BatchFlags& flags, std::string& err) {
    hipStream_t s = h->stream;
    int* hbad = pool.pinned<int>(1);
    HIPCHECK(hipMemcpyAsync(hbad, flags.bad, sizeof(int), hipMemcpyDeviceToHost, s));
    int st = check_solve_states(s, pool, flags.states, flags.nsolves, err);  // synchronises the stream
    if (!flags.nsolves) HIPCHECK(hipStreamSynchronize(s));
    HIPCHECK(hipGetLastError());
    if (st) return st;
    if (*hbad & 2) {
        err = "weighted blend without vectors: non-finite or huge float samples (the NaN rule of blend.rs:280-300) are not "
              "served";
        return TWX_E_UNSUPPORTED;
    }
    if (*hbad) { err = "inverse warp: a sample coordinate is NaN or never converges, which is not supported"; return TWX_E_UNSUPPORTED; }
    return TWX_OK;
}

static void add_stats(twx_stats& a, const twx_stats& d) {
    a.pair_solves += d.pair_solves;
    a.pairs_reused += d.pairs_reused;
    a.composes += d.composes;
    a.fast_frames += d.fast_frames;
    a.pyramid_builds += d.pyramid_builds;
}

// The batched render, enqueued and checked in one go (no graph).
int retime_batch_dev(twx_handle* h, DevPool& pool, const std::vector<ClipRun>& runs, std::string& err) {
    double t0 = now_ms();
    BatchFlags flags;
    twx_stats delta{};
    int st = retime_batch_enqueue(h, pool, runs, flags, delta, err);
    if (!st) st = retime_batch_check(h, pool, flags, err);
    add_stats(h->stats, delta);
    h->stats.device_ms += now_ms() - t0;
    return st;
}

// The batched render of output frames [first, first + count) of one clip on the device.
int retime_frames_dev(twx_handle* h, DevPool& pool, const twx_retime_params& p, const uint8_t* clip, int n_in, int first,
                      int count, uint8_t* out, std::string& err) {
    return retime_batch_dev(h, pool, {ClipRun{p, clip, n_in, first, count, out}}, err);
}

// The arguments that decide a batched render's launches: the parameter blocks (keyframes by value),
// the device pointers and the frame ranges.
static std::string batch_key(const std::vector<ClipRun>& runs) {
    std::string k;
    auto put = [&](const void* p, size_t n) { k.append((const char*)p, n); };
    for (const ClipRun& r : runs) {
        twx_retime_params p = r.p;
        p.speed_keys = nullptr;
        put(&p, sizeof(p));
        if (r.p.n_speed_keys > 0 && r.p.speed_keys) put(r.p.speed_keys, sizeof(twx_key) * (size_t)r.p.n_speed_keys);
        put(&r.clip, sizeof(r.clip));
        put(&r.n_in, sizeof(r.n_in));
        put(&r.first, sizeof(r.first));
        put(&r.count, sizeof(r.count));
        put(&r.out, sizeof(r.out));
    }
    return k;
}

// A batched render through the graph cache: a call that repeats the arguments of the last captured
// render replays its graph; otherwise the render runs on a new plan's own memory (this call's
// output), then the same launch sequence is captured over the same memory (the same allocation
// sequence gives the same addresses; the table uploads stay in the plan's pinned staging) and kept
// for the next call. With profiling on or graphs off, the plain path runs on the shared arenas.
int retime_batch_run(twx_handle* h, const std::vector<ClipRun>& runs, std::string& err) {
    if (!h->graphs || h->prof.on) {
        DevPool pool(h);
        return retime_batch_dev(h, pool, runs, err);
    }
    hipStream_t s = h->stream;
    const double t0 = now_ms();
    std::string key = batch_key(runs);
    if (h->graph && h->graph->exec && h->graph->key == key) {
        GraphPlan& g = *h->graph;
        HIPCHECK(hipGraphLaunch(g.exec, s));
        DevPool readback(h);
        BatchFlags f;
        f.bad = g.bad;
        f.states = g.states;
        f.nsolves = g.nsolves;
        int st = retime_batch_check(h, readback, f, err);
        add_stats(h->stats, g.delta);
        h->stats.device_ms += now_ms() - t0;
        h->graph_replays++;
        return st;
    }
    h->graph.reset();
    auto g = std::make_unique<GraphPlan>();
    g->key = key;
    {
        DevPool pool(g->dev, g->host, s);
        BatchFlags f;
        twx_stats delta{};
        int st = retime_batch_enqueue(h, pool, runs, f, delta, err);
        if (!st) {
            DevPool readback(h);
            st = retime_batch_check(h, readback, f, err);
        }
        add_stats(h->stats, delta);
        h->stats.device_ms += now_ms() - t0;
        if (st) return st;
        g->delta = delta;
    }
    DevPool pool(g->dev, g->host, s);
    BatchFlags f;
    twx_stats delta{};
    hipGraph_t graph = nullptr;
    HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    std::string e2;
    int st = TWX_OK;
    try {
        st = retime_batch_enqueue(h, pool, runs, f, delta, e2);
    } catch (const HipError&) {
        (void)hipStreamEndCapture(s, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        return TWX_OK;  // this call's output is complete; no graph is kept
    }
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (st || ec != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return TWX_OK;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        (void)hipGetLastError();
        return TWX_OK;
    }
    g->exec = exec;
    g->bad = f.bad;
    g->states = f.states;
    g->nsolves = f.nsolves;
    h->graph = std::move(g);
    return TWX_OK;
}

// A digest of an engine image for the pyramid cache's first filter: its size, format and 4096
// evenly spaced 8-byte words (FNV-1a); equal images always have equal digests. Matches are then
// confirmed byte for byte.
uint64_t image_digest(const uint8_t* p, size_t bytes, int W, int H, int fmt) {
    uint64_t hsh = 0xcbf29ce484222325ull;
    auto eat = [&](uint64_t v) {
        for (int i = 0; i < 8; i++) {
            hsh ^= (v >> (8 * i)) & 0xff;
            hsh *= 0x100000001b3ull;
        }
    };
    eat((uint64_t)W);
    eat((uint64_t)H);
    eat((uint64_t)fmt);
    const size_t words = bytes / 8, step = std::max<size_t>(words / 4096, 1);
    for (size_t i = 0; i < words; i += step) {
        uint64_t v;
        memcpy(&v, p + i * 8, 8);
        eat(v);
    }
    for (size_t i = words * 8; i < bytes; i++) eat(p[i]);
    return hsh;
}

// The Backend::solve / solve_pair path on host engine images (rows as given).
int solve_host(twx_handle* h, const void* x, const void* y, int W, int H, int fmt, const twx_flow_settings& fs,
               float* f_xy, float* f_yx, std::string& err) {
    Geo g;
    int st = geometry(fs, W, H, g, err);
    if (st) return st;
    if (fmt < 0 || fmt > 2) { err = "bad format code"; return TWX_E_MISMATCH; }
    hipStream_t s = h->stream;
    double t0 = now_ms();
    DevPool pool(h);
    const size_t fb = image_bytes(W, H, fmt);
    const size_t pyr_bytes = pyramid_floats(g) * sizeof(float);
    PyramidCache& C = h->pyramids;
    for (size_t i = C.entries.size(); i-- > 0;)  // builds an earlier call did not complete
        if (!C.entries[i]->ready) C.drop(i);
    // the pyramid of each image: found in the cache, or built (into a new cache entry when it fits
    // the cache, else into this call's memory)
    const float* lv[2] = {nullptr, nullptr};
    PyrEntry* used_now[2] = {nullptr, nullptr};
    std::vector<PyrIn> build;
    const void* imgs[2] = {x, y};
    std::vector<PyrEntry*> made;  // entries whose pyramids this call builds
    for (int k = 0; k < 2; k++) {
        const uint8_t* src = (const uint8_t*)imgs[k];
        if (k == 1 && memcmp(x, y, fb) == 0) {  // solve(X, X): one pyramid
            lv[1] = lv[0];
            C.hits++;
            h->stats.pyramid_hits++;
            break;
        }
        const uint64_t d = image_digest(src, fb, W, H, fmt);
        PyrEntry* hit = nullptr;
        for (size_t i = 0; i < C.entries.size(); i++) {
            PyrEntry& E = *C.entries[i];
            if (E.ready && E.digest == d && E.W == W && E.H == H && E.fmt == fmt && E.reduction == g.reduction &&
                E.levels == g.levels && memcmp(E.key.data(), src, fb) == 0) {
                std::unique_ptr<PyrEntry> keep = std::move(C.entries[i]);
                C.entries.erase(C.entries.begin() + (long)i);
                C.entries.push_back(std::move(keep));
                hit = C.entries.back().get();
                break;
            }
        }
        if (hit) {
            C.hits++;
            h->stats.pyramid_hits++;
            lv[k] = hit->dev;
            used_now[k] = hit;
            continue;
        }
        C.misses++;
        h->stats.pyramid_builds++;
        uint8_t* dimg = pool.alloc<uint8_t>(fb);
        HIPCHECK(hipMemcpyAsync(dimg, src, fb, hipMemcpyHostToDevice, s));
        float* dst = nullptr;
        if (pyr_bytes <= C.budget) {
            // room for the new entry: the least recently used entries this call does not read go
            for (size_t i = 0; i < C.entries.size() &&
                               (C.used + pyr_bytes > C.budget || C.entries.size() >= PyramidCache::MAX_ENTRIES);) {
                if (C.entries[i].get() == used_now[0]) { i++; continue; }
                C.drop(i);
                C.evictions++;
            }
            if (C.used + pyr_bytes <= C.budget && C.entries.size() < PyramidCache::MAX_ENTRIES) {
                auto E = std::make_unique<PyrEntry>();
                HIPCHECK(hipMalloc(&E->dev, pyr_bytes));
                E->key.assign(src, src + fb);
                E->W = W;
                E->H = H;
                E->fmt = fmt;
                E->reduction = g.reduction;
                E->levels = g.levels;
                E->digest = d;
                E->bytes = pyr_bytes;
                C.used += pyr_bytes;
                dst = E->dev;
                used_now[k] = E.get();
                made.push_back(E.get());
                C.entries.push_back(std::move(E));
            }
        }
        if (!dst) dst = pool.alloc<float>(pyramid_floats(g));
        lv[k] = dst;
        build.push_back(PyrIn{dimg, W, H, fmt, g, dst});
    }
    build_pyramids(s, pool, build);
    // an entry becomes findable once its pyramid is complete; one whose build did not complete
    // (an exception leaves this function) stays unready and is dropped by the next eviction
    HIPCHECK(hipStreamSynchronize(s));
    for (PyrEntry* E : made) E->ready = true;
    size_t fn = (size_t)g.w0 * g.h0 * 2;
    std::vector<SolveIn> solves;
    // solve_pair order: F_BA (X = B) then F_AB (X = A); both in one batched launch
    if (f_yx) solves.push_back(SolveIn{lv[1], lv[0], g, pool.alloc<float>(fn)});
    solves.push_back(SolveIn{lv[0], lv[1], g, pool.alloc<float>(fn)});
    if ((st = run_solves(s, pool, solves, fs, h->budget / 2, err))) return st;
    if (f_yx) {
        HIPCHECK(hipMemcpyAsync(f_yx, solves[0].out, fn * 4, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(f_xy, solves[1].out, fn * 4, hipMemcpyDeviceToHost, s));
    } else {
        HIPCHECK(hipMemcpyAsync(f_xy, solves[0].out, fn * 4, hipMemcpyDeviceToHost, s));
    }
    HIPCHECK(hipStreamSynchronize(s));
    h->stats.pair_solves += solves.size();
    h->stats.solve_calls++;
    h->stats.device_ms += now_ms() - t0;
    return TWX_OK;
}

}  // namespace

// ================================================================================================

extern "C" {

int twx_open(int index, twx_handle** out, char* err, size_t errlen) {
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_err(err, errlen, "no HIP device");
        return TWX_E_DEVICE;
    }
    int dev = index < 0 ? 0 : index;
    if (dev >= n) {
        set_err(err, errlen, "device index out of range");
        return TWX_E_DEVICE;
    }
    twx_handle* h = new twx_handle();
    h->device = dev;
    try {
        HIPCHECK(hipSetDevice(dev));
        HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        HIPCHECK(hipEventCreateWithFlags(&h->caller_ev, hipEventDisableTiming));
        size_t free_b = 0, total_b = 0;
        HIPCHECK(hipMemGetInfo(&free_b, &total_b));
        h->budget = (uint64_t)(total_b * 0.75);
        h->pyramids.budget = h->budget / 4;
        h->host.pinned = true;
        h->host.min_block = 8u << 20;
        // keep freed stream-ordered memory in the pool between calls
        hipMemPool_t pool;
        HIPCHECK(hipDeviceGetDefaultMemPool(&pool, dev));
        uint64_t thresh = UINT64_MAX;
        HIPCHECK(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thresh));
        std::string e;
        // the device address of the allocation-free backward's fixed scratch, with this device current
        int st = noalloc_scratch_setup(dev, e);
        if (st) {
            set_err(err, errlen, e);
            (void)hipEventDestroy(h->caller_ev);
            (void)hipStreamDestroy(h->stream);
            delete h;
            return st;
        }
        st = run_selfcheck(h->stream, e);
        if (st) {
            set_err(err, errlen, "device self-check refused the device: " + e);
            (void)hipEventDestroy(h->caller_ev);
            (void)hipStreamDestroy(h->stream);
            delete h;
            return st;
        }
    } catch (const HipError& e) {
        set_err(err, errlen, e.msg);
        delete h;
        return TWX_E_DEVICE;
    }
    *out = h;
    return TWX_OK;
}

void twx_close(twx_handle* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
    }
    if (h->caller_ev) (void)hipEventDestroy(h->caller_ev);
    h->pyramids.clear();
    h->dev.release();
    h->host.release();
    delete h;
}

int twx_selfcheck(twx_handle* h, char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (!g.ok()) { set_err(err, errlen, e); return st; }
        st = run_selfcheck(h->stream, e);
    } catch (const HipError& x) {
        set_err(err, errlen, x.msg);
        return TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

uint64_t twx_memory_budget(twx_handle* h) { return h ? h->budget : 0; }

void* twx_stream(twx_handle* h) { return h ? (void*)h->stream : nullptr; }

int twx_flow_geometry(const twx_flow_settings* fs, int32_t width, int32_t height, int32_t* flow_w, int32_t* flow_h,
                      int32_t* levels) {
    Geo g;
    std::string e;
    int st = geometry(*fs, width, height, g, e);
    if (st) return st;
    if (flow_w) *flow_w = g.w0;
    if (flow_h) *flow_h = g.h0;
    if (levels) *levels = g.levels;
    return TWX_OK;
}

int twx_solve(twx_handle* h, const void* x, const void* y, int32_t width, int32_t height, int32_t fmt,
              const twx_flow_settings* fs, float* flow_out, char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) st = solve_host(h, x, y, width, height, fmt, *fs, flow_out, nullptr, e);
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

int twx_solve_pair(twx_handle* h, const void* a, const void* b, int32_t width, int32_t height, int32_t fmt,
                   const twx_flow_settings* fs, float* f_ab_out, float* f_ba_out, char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) st = solve_host(h, a, b, width, height, fmt, *fs, f_ab_out, f_ba_out, e);
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

// The mip levels of a textured forward warp's vertices (mesh.rs:54-88) from their arguments m:
// -(logf(m) * (1 / logf(2))) + 0, in place, with the host C runtime's logf (the reference's device
// path takes the logarithm on the host too; a device logarithm would differ in the last bits).
static void fw_mip_levels_host(float* m, size_t n) {
    const float inv_ln2 = 1.0f / logf(2.0f);
    auto run = [=](size_t a, size_t b) {
        for (size_t i = a; i < b; i++) m[i] = -(logf(m[i]) * inv_ln2) + 0.0f;
    };
    const size_t nt = n < (1u << 18) ? 1 : std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency()));
    if (nt == 1) {
        run(0, n);
        return;
    }
    std::vector<std::thread> th;
    const size_t step = (n + nt - 1) / nt;
    for (size_t a = 0; a < n; a += step) th.emplace_back(run, a, std::min(n, a + step));
    for (auto& t : th) t.join();
}

// The forward warp of a device image into `out` (and `cov`, may be null): the Gouraud variant
// (vertex colours), or with `textured` the pyramid-sampled variant (ForwardOpts::pyramid,
// forward.rs:229-320); the flow on the device, with its host copy for the early-out test
// (forward.rs:245-256). `dflags` is a device int the caller cleared; its bits 8, 16, 32 and 64 are
// read by the caller after its final synchronisation (check_fw_flags).
static int forward_warp_dev(twx_handle* h, DevPool& pool, const void* src, int W, int H, int fmt, const float* d_flow,
                            const float* h_flow, int fw, int fh, float amount, bool textured, void* out, uint8_t* cov,
                            int* dflags, std::string& err) {
    hipStream_t s = h->stream;
    const float a = 0.0f > amount ? -amount : amount;
    bool still = 1e-05 > (double)a;
    auto copy_still = [&]() {
        HIPCHECK(hipMemcpyAsync(out, src, image_bytes(W, H, fmt), hipMemcpyDeviceToDevice, s));
        if (cov) HIPCHECK(hipMemsetAsync(cov, 0xff, (size_t)W * H, s));
        return TWX_OK;
    };
    // no host copy of the flow: the same test on the device, read at the warp's first synchronisation
    // (the passes enqueued before it write only this warp's own buffers)
    int* hm = nullptr;
    if (!still && h_flow) {
        still = fw_flow_is_zero(h_flow, (long)fw * fh * 2);
    } else if (!still) {
        int* moves = pool.alloc<int>(1);
        hm = pool.pinned<int>(1);
        launch_zero(s, moves, 1);
        launch_fw_flow_moves(s, d_flow, (long)fw * fh * 2, moves);
        HIPCHECK(hipMemcpyAsync(hm, moves, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    if (still) return copy_still();
    if ((long)W * H > (long)UINT32_MAX) {  // the rasteriser's vertices carry 32-bit pixel indices
        err = "forward warp: an image of more than 2^32 - 1 pixels";
        return TWX_E_UNSUPPORTED;
    }
    FwArgs A{};
    A.src = src;
    A.out = out;
    A.cov = cov;
    A.flow = d_flow;
    A.W = W;
    A.H = H;
    A.fw = fw;
    A.fh = fh;
    A.fmt = fmt;
    A.flags = dflags;
    A.textured = textured ? 1 : 0;
    fw_grid(A, amount);
    const long n = (long)W * H;
    if (textured) {
        // pyramid.rs:20-63: the levels (halved, at least 2, from max(w, 2) x max(h, 2)) below the
        // source, and the mip level of every vertex when there is more than one level
        A.nlev = fw_level_count(W, H);
        if (A.nlev > FW_MAX_LEVELS) {
            err = "forward warp: a texture pyramid deeper than this backend's level table";
            return TWX_E_UNSUPPORTED;
        }
        A.lev[0] = src;
        A.lw[0] = W;
        A.lh[0] = H;
        int cw = std::max(W, 2), ch = std::max(H, 2);
        for (int i = 1; i < A.nlev; i++) {
            cw = std::max(cw / 2, 2);
            ch = std::max(ch / 2, 2);
            A.lw[i] = cw;
            A.lh[i] = ch;
            A.lev[i] = pool.alloc<uint8_t>(image_bytes(cw, ch, fmt));
        }
        if (A.nlev > 1) {
            launch_fw_pyramid(s, A, pool.alloc<float>((size_t)n * 8));
            float* m = pool.alloc<float>(n);
            float* hmip = pool.pinned<float>(n);
            launch_fw_mip_arg(s, A, m);
            HIPCHECK(hipMemcpyAsync(hmip, m, n * sizeof(float), hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
            if (hm && *hm == 0) return copy_still();
            fw_mip_levels_host(hmip, n);
            HIPCHECK(hipMemcpyAsync(m, hmip, n * sizeof(float), hipMemcpyHostToDevice, s));
            A.mips = m;
        }
    }
    A.pos = pool.alloc<float2>(n);
    A.counts = pool.alloc<uint32_t>(n);
    A.fill = pool.alloc<uint32_t>(n);
    A.offsets = pool.alloc<uint64_t>(n);
    A.block_totals = pool.alloc<uint64_t>(fw_scan_blocks(n));
    A.total = pool.alloc<uint64_t>(1);
    HIPCHECK(hipMemsetAsync(A.counts, 0, n * sizeof(uint32_t), s));
    HIPCHECK(hipMemsetAsync(A.fill, 0, n * sizeof(uint32_t), s));
    // the count pass's flags (1, 2, 4) in a word of this warp's own, so that a warp that turns out
    // still leaves nothing in the caller's flags; the emit and fold passes report into dflags
    int* cflags = pool.alloc<int>(1);
    launch_zero(s, cflags, 1);
    A.flags = cflags;
    launch_fw_count(s, A);
    uint64_t* ht = pool.pinned<uint64_t>(1);
    int* hf = pool.pinned<int>(1);
    HIPCHECK(hipMemcpyAsync(ht, A.total, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipMemcpyAsync(hf, cflags, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (hm && *hm == 0) return copy_still();
    A.flags = dflags;
    if (*hf & 1) {
        err = "forward warp: the degenerate last row pass of the rasteriser would write outside the image, which is not "
              "supported";
        return TWX_E_UNSUPPORTED;
    }
    if (*hf & 2) {
        err = "forward warp: a triangle of the displaced mesh covers more than 65536 pixels (no device kernel for that)";
        return TWX_E_UNSUPPORTED;
    }
    if (*hf & 4) {
        err = "forward warp: a NaN colour (the NaN rule of the colour arithmetic has no device kernel)";
        return TWX_E_UNSUPPORTED;
    }
    const uint64_t total = *ht;
    if (total * 20 > h->budget / 2) {
        err = "forward warp: the mesh covers too many pixels for the device memory budget";
        return TWX_E_UNSUPPORTED;
    }
    A.rec_t = pool.alloc<uint32_t>(std::max<uint64_t>(total, 1));
    A.rec_c = pool.alloc<float4>(std::max<uint64_t>(total, 1));
    launch_fw_emit_fold(s, A, total);
    return TWX_OK;
}

static int check_fw_flags(int f, std::string& err) {
    if (f & 24) {
        err = "forward warp: the device hit lists were inconsistent";
        return TWX_E_DEVICE;
    }
    if (f & 32) {
        err = "forward warp: a textured sample below the image (a pyramid tap the CPU path would index out of bounds)";
        return TWX_E_UNSUPPORTED;
    }
    if (f & 64) {
        err = "forward warp: a NaN colour (the NaN rule of the colour arithmetic has no device kernel)";
        return TWX_E_UNSUPPORTED;
    }
    return TWX_OK;
}

int twx_forward_warp(twx_handle* h, const void* src, int32_t width, int32_t height, int32_t fmt, const float* flow,
                     int32_t flow_w, int32_t flow_h, float amount, int32_t pyramid, void* out, uint8_t* coverage,
                     char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) {
            if (width < 2 || height < 2 || flow_w < 2 || flow_h < 2 || fmt < 0 || fmt > 2) {
                e = "forward warp needs images and a flow of at least 2x2";
                st = TWX_E_MISMATCH;
            } else if ((st = check_frame_pixels(width, height, e))) {
            } else {
                hipStream_t s = h->stream;
                DevPool pool(h);
                const size_t fb = image_bytes(width, height, fmt), fn = (size_t)flow_w * flow_h * 2;
                uint8_t* img = pool.alloc<uint8_t>(fb * 2);
                float* fl = pool.alloc<float>(fn);
                uint8_t* cov = coverage ? pool.alloc<uint8_t>((size_t)width * height) : nullptr;
                int* flags = pool.alloc<int>(1);
                HIPCHECK(hipMemcpyAsync(img, src, fb, hipMemcpyHostToDevice, s));
                HIPCHECK(hipMemcpyAsync(fl, flow, fn * 4, hipMemcpyHostToDevice, s));
                HIPCHECK(hipMemsetAsync(flags, 0, sizeof(int), s));
                st = forward_warp_dev(h, pool, img, width, height, fmt, fl, flow, flow_w, flow_h, amount, pyramid != 0,
                                      img + fb, cov, flags, e);
                if (!st) {
                    int* hf = pool.pinned<int>(1);
                    HIPCHECK(hipMemcpyAsync(out, img + fb, fb, hipMemcpyDeviceToHost, s));
                    if (cov) HIPCHECK(hipMemcpyAsync(coverage, cov, (size_t)width * height, hipMemcpyDeviceToHost, s));
                    HIPCHECK(hipMemcpyAsync(hf, flags, sizeof(int), hipMemcpyDeviceToHost, s));
                    HIPCHECK(hipStreamSynchronize(s));
                    st = check_fw_flags(*hf, e);
                }
            }
        }
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

// ---- vector image coding and the Vectors In flow steps ----------------------------------------

}  // extern "C"

// The common frame of an entry point: the handle's lock and device, HIP errors as status 1, the
// message into `err`.
template <class F>
static int entry(twx_handle* h, char* err, size_t errlen, F&& body) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) st = body(e);
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

extern "C" {

int twx_encode_vectors(twx_handle* h, const float* flow, int32_t flow_w, int32_t flow_h, int32_t width, int32_t height,
                       int32_t fmt, float scale, void* out, char* err, size_t errlen) {
    return entry(h, err, errlen, [&](std::string& e) -> int {
        if (width < 2 || height < 2 || flow_w < 2 || flow_h < 2 || fmt < 0 || fmt > 2) {
            e = "vector encode needs images and a flow of at least 2x2";
            return TWX_E_MISMATCH;
        }
        if (int st = check_frame_pixels(width, height, e)) return st;
        hipStream_t s = h->stream;
        DevPool pool(h);
        const size_t fb = image_bytes(width, height, fmt), fn = (size_t)flow_w * flow_h * 2;
        float* fl = pool.alloc<float>(fn);
        uint8_t* img = pool.alloc<uint8_t>(fb);
        HIPCHECK(hipMemcpyAsync(fl, flow, fn * 4, hipMemcpyHostToDevice, s));
        VecJob J{};
        J.flow = fl;
        J.out = img;
        J.fw = flow_w;
        J.fh = flow_h;
        J.W = width;
        J.H = height;
        J.fmt = fmt;
        J.finish = 0;
        vec_job_scales(J, scale);
        std::vector<VecJob> jobs{J};
        launch_vec_encode(s, pool.upload(jobs), jobs);
        HIPCHECK(hipMemcpyAsync(out, img, fb, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        return TWX_OK;
    });
}

int twx_decode_vectors(twx_handle* h, const void* vec, int32_t width, int32_t height, int32_t fmt, float max_displace,
                       float* flow_out, uint8_t* holes_out, char* err, size_t errlen) {
    return entry(h, err, errlen, [&](std::string& e) -> int {
        if (width < 1 || height < 1 || fmt < 0 || fmt > 2) {
            e = "vector decode: bad size or format";
            return TWX_E_MISMATCH;
        }
        if (int st = check_frame_pixels(width, height, e)) return st;
        hipStream_t s = h->stream;
        DevPool pool(h);
        const size_t fb = image_bytes(width, height, fmt), n = (size_t)width * height;
        uint8_t* img = pool.alloc<uint8_t>(fb);
        float* fl = pool.alloc<float>(n * 2);
        uint8_t* holes = pool.alloc<uint8_t>(n);
        int* gate = pool.alloc<int>(1);
        HIPCHECK(hipMemcpyAsync(img, vec, fb, hipMemcpyHostToDevice, s));
        launch_vec_decode(s, img, width, height, fmt, max_displace, fl, holes, gate);
        HIPCHECK(hipMemcpyAsync(flow_out, fl, n * 8, hipMemcpyDeviceToHost, s));
        if (holes_out) HIPCHECK(hipMemcpyAsync(holes_out, holes, n, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        return TWX_OK;
    });
}

// The flow of a Vectors In layer on the device: decode, the hole fill of the alpha-0 holes,
// Vec Scale (compose/vector.rs:335-345). `flow` (width * height * 2) is written.
static int flow_from_vectors_dev(hipStream_t s, DevPool& pool, const void* d_vec, int W, int H, int fmt, float md,
                                 float sx, float sy, bool above_first, float* d_flow, std::string& err) {
    const size_t n = (size_t)W * H;
    uint8_t* holes = pool.alloc<uint8_t>(n);
    int* gate = pool.alloc<int>(1);
    HIPCHECK(hipMemsetAsync(gate, 0, sizeof(int), s));
    launch_vec_decode(s, d_vec, W, H, fmt, md, d_flow, holes, gate);
    if (W >= 2 && H >= 2) {
        std::vector<FillItem> items{FillItem{W, H, d_flow, FILL_FLOW2, 2, holes, gate, above_first ? 1 : 0}};
        int st = fill_items(s, pool, items, err);
        if (st) return st;
    } else {
        int* hg = pool.pinned<int>(1);
        HIPCHECK(hipMemcpyAsync(hg, gate, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        if (*hg) {
            err = "hole fill of an image with a side of 1";
            return TWX_E_UNSUPPORTED;
        }
    }
    launch_vec_scale(s, d_flow, (long)n, sx, sy);
    return TWX_OK;
}

// invert_flow_on (compose/vector.rs:370-397) of a device flow into `out` (both width x height): T =
// (-1, -fx, -fy, -1) pushed along the flow by the forward warp with coverage (Gouraud, or the
// textured variant), the read-out, the push-pull fill of the uncovered pixels. `h_flow` is the
// flow's host copy or null; the forward warp's flags go to `fw_flags` (a device int the caller
// cleared).
static int invert_flow_dev(twx_handle* h, DevPool& pool, const float* d_flow, const float* h_flow, int W, int H,
                           bool above_first, bool textured, float* out, int* fw_flags, std::string& err) {
    hipStream_t s = h->stream;
    const size_t n = (size_t)W * H;
    float* t = pool.alloc<float>(n * 4);
    float* pushed = pool.alloc<float>(n * 4);
    uint8_t* cov = pool.alloc<uint8_t>(n);
    uint8_t* holes = pool.alloc<uint8_t>(n);
    int* gate = pool.alloc<int>(1);
    launch_zero(s, gate, 1);
    launch_invert_prep(s, d_flow, t, (long)n);
    int st = forward_warp_dev(h, pool, t, W, H, TWX_FMT_F32, d_flow, h_flow, W, H, 1.0f, textured, pushed, cov, fw_flags,
                              err);
    if (st) return st;
    launch_invert_post(s, pushed, cov, out, holes, gate, (long)n);
    std::vector<FillItem> items{FillItem{W, H, out, FILL_FLOW2, 2, holes, gate, above_first ? 1 : 0}};
    return fill_items(s, pool, items, err);
}

int twx_flow_from_vectors(twx_handle* h, const void* vec, int32_t width, int32_t height, int32_t fmt, float max_displace,
                          float scale_x, float scale_y, int32_t taps_above_first, float* flow_out, char* err,
                          size_t errlen) {
    return entry(h, err, errlen, [&](std::string& e) -> int {
        if (width < 1 || height < 1 || fmt < 0 || fmt > 2) {
            e = "vector decode: bad size or format";
            return TWX_E_MISMATCH;
        }
        if (int st = check_frame_pixels(width, height, e)) return st;
        hipStream_t s = h->stream;
        DevPool pool(h);
        const size_t fb = image_bytes(width, height, fmt), n = (size_t)width * height;
        uint8_t* img = pool.alloc<uint8_t>(fb);
        float* fl = pool.alloc<float>(n * 2);
        HIPCHECK(hipMemcpyAsync(img, vec, fb, hipMemcpyHostToDevice, s));
        int st = flow_from_vectors_dev(s, pool, img, width, height, fmt, max_displace, scale_x, scale_y,
                                       taps_above_first != 0, fl, e);
        if (st) return st;
        HIPCHECK(hipMemcpyAsync(flow_out, fl, n * 8, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        return TWX_OK;
    });
}

int twx_invert_flow(twx_handle* h, const float* flow, int32_t width, int32_t height, int32_t forward_gouraud,
                    int32_t taps_above_first, float* out, char* err, size_t errlen) {
    return entry(h, err, errlen, [&](std::string& e) -> int {
        if (width < 2 || height < 2) {
            e = "forward warp needs images and a flow of at least 2x2";
            return TWX_E_UNSUPPORTED;
        }
        if (int st = check_frame_pixels(width, height, e)) return st;
        hipStream_t s = h->stream;
        DevPool pool(h);
        const size_t n = (size_t)width * height;
        float* fl = pool.alloc<float>(n * 2);
        float* res = pool.alloc<float>(n * 2);
        int* flags = pool.alloc<int>(1);
        HIPCHECK(hipMemcpyAsync(fl, flow, n * 8, hipMemcpyHostToDevice, s));
        launch_zero(s, flags, 1);
        int st = invert_flow_dev(h, pool, fl, flow, width, height, taps_above_first != 0, forward_gouraud == 0, res, flags,
                                 e);
        if (st) return st;
        int* hf = pool.pinned<int>(1);
        HIPCHECK(hipMemcpyAsync(out, res, n * 8, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(hf, flags, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        return check_fw_flags(*hf, e);
    });
}

int twx_create_vectors_dev(twx_handle* h, const twx_flow_settings* fs, int32_t width, int32_t height, int32_t fmt,
                           const void* clip, int32_t n_in, int32_t count, const int32_t* frame_x, const int32_t* frame_y,
                           float max_displace, void* out, char* err, size_t errlen) {
    return entry(h, err, errlen, [&](std::string& e) -> int {
        if (width < 2 || height < 2) {
            e = "create vectors: frames with a side of 1 take the zero vector image on the CPU";
            return TWX_E_UNSUPPORTED;
        }
        if (fmt < 0 || fmt > 2 || count < 0 || n_in < 0 || !fs || (count > 0 && (!frame_x || !frame_y))) {
            e = "create vectors: bad arguments";
            return TWX_E_MISMATCH;
        }
        for (int k = 0; k < count; k++)
            if (frame_x[k] < 0 || frame_x[k] >= n_in || frame_y[k] < 0 || frame_y[k] >= n_in) {
                e = "create vectors: a frame index outside the clip";
                return TWX_E_MISMATCH;
            }
        Geo g;
        int st = geometry(*fs, width, height, g, e);
        if (st) return st;
        after_caller(h);
        hipStream_t s = h->stream;
        double t0 = now_ms();
        DevPool pool(h);
        const size_t fb = image_bytes(width, height, fmt);
        // the pyramids of the frames used, once each, read from the clip (engine rows flipped, as
        // the ingest would store them; nothing else needs the engine images)
        std::map<int, int> slot_of;
        std::vector<PyrIn> pyrs;
        auto slot = [&](int idx) {
            auto it = slot_of.find(idx);
            if (it != slot_of.end()) return it->second;
            PyrIn pi{nullptr, width, height, fmt, g, pool.alloc<float>(pyramid_floats(g))};
            pi.wire = (const uint8_t*)clip + (size_t)idx * fb;
            pi.reverse = 1;
            pyrs.push_back(pi);
            int k = (int)pyrs.size() - 1;
            slot_of[idx] = k;
            return k;
        };
        std::map<std::pair<int, int>, int> solve_of;
        std::vector<SolveIn> solves;
        std::vector<int> solve_for(count);
        for (int k = 0; k < count; k++) {
            auto key = std::make_pair(frame_x[k], frame_y[k]);
            auto it = solve_of.find(key);
            if (it != solve_of.end()) {
                solve_for[k] = it->second;
                h->stats.pairs_reused++;
                continue;
            }
            int px = slot(frame_x[k]), py = slot(frame_y[k]);
            solves.push_back(SolveIn{pyrs[px].levels, pyrs[py].levels, g, pool.alloc<float>((size_t)g.w0 * g.h0 * 2)});
            solve_for[k] = (int)solves.size() - 1;
            solve_of[key] = solve_for[k];
        }
        build_pyramids(s, pool, pyrs);
        h->stats.pyramid_builds += pyrs.size();
        h->stats.pair_solves += solves.size();
        if ((st = run_solves(s, pool, solves, *fs, h->budget / 2, e))) return st;
        std::vector<VecJob> jobs;
        for (int k = 0; k < count; k++) {
            VecJob J{};
            J.flow = solves[solve_for[k]].out;
            J.out = (uint8_t*)out + (size_t)k * fb;
            J.fw = g.w0;
            J.fh = g.h0;
            J.W = width;
            J.H = height;
            J.fmt = fmt;
            J.finish = 1;
            vec_job_scales(J, max_displace);
            jobs.push_back(J);
        }
        if (!jobs.empty()) launch_vec_encode(s, pool.upload(jobs), jobs);
        HIPCHECK(hipStreamSynchronize(s));
        HIPCHECK(hipGetLastError());
        h->stats.device_ms += now_ms() - t0;
        return TWX_OK;
    });
}

// ---- many solve_pair calls in one batch (device/exec/backend.rs:191-208 per pair) ----

int twx_solve_pairs_batched(twx_handle* h, int32_t n, const twx_pair* pairs, int32_t width, int32_t height, int32_t fmt,
                            const twx_flow_settings* fs, char* err, size_t errlen) {
    return entry(h, err, errlen, [&](std::string& e) -> int {
        if (n < 0 || (n > 0 && !pairs) || !fs || fmt < 0 || fmt > 2) {
            e = "solve pairs: bad arguments";
            return TWX_E_MISMATCH;
        }
        for (int k = 0; k < n; k++)
            if (!pairs[k].a || !pairs[k].b) {
                e = "solve pairs: pair " + std::to_string(k) + " has a null image";
                return TWX_E_MISMATCH;
            }
        Geo g;
        int st = geometry(*fs, width, height, g, e);
        if (st) return st;
        after_caller(h);
        hipStream_t s = h->stream;
        double t0 = now_ms();
        DevPool pool(h);
        // one pyramid per distinct image of the call
        std::map<const void*, int> slot_of;
        std::vector<PyrIn> pyrs;
        auto levels = [&](const void* img) -> const float* {
            auto it = slot_of.find(img);
            if (it != slot_of.end()) {
                h->stats.pyramid_hits++;
                return pyrs[it->second].levels;
            }
            pyrs.push_back(PyrIn{img, width, height, fmt, g, pool.alloc<float>(pyramid_floats(g))});
            slot_of[img] = (int)pyrs.size() - 1;
            return pyrs.back().levels;
        };
        std::vector<SolveIn> solves;
        for (int k = 0; k < n; k++) {
            const twx_pair& P = pairs[k];
            if (!P.f_ab && !P.f_ba) continue;
            const float* la = levels(P.a);
            const float* lb = levels(P.b);
            // solve_pair order: F_BA (X = B), then F_AB (X = A); the flows end in the caller's buffers
            if (P.f_ba) solves.push_back(SolveIn{lb, la, g, P.f_ba});
            if (P.f_ab) solves.push_back(SolveIn{la, lb, g, P.f_ab});
        }
        build_pyramids(s, pool, pyrs);
        h->stats.pyramid_builds += pyrs.size();
        h->stats.pair_solves += solves.size();
        h->stats.solve_calls++;
        if ((st = run_solves(s, pool, solves, *fs, h->budget / 2, e))) return st;
        HIPCHECK(hipStreamSynchronize(s));
        HIPCHECK(hipGetLastError());
        h->stats.device_ms += now_ms() - t0;
        return TWX_OK;
    });
}

// ---- the Vectors In render (plugins/vectors-in/src/glue.rs:380-495, vectorin.rs, timemap.rs) ----

// One layer of a Vectors In render: wire frames on the device, frame i showing at ticks
// [i * time_step, (i + 1) * time_step); a null `wire` is a layer that is not set.
struct VinClip {
    const uint8_t* wire;
    int32_t n;
};

// Output frames [first, first + count) into `out` (wire frames on the device). Every frame the
// render reads is ingested upside down (glue.rs:398-410), once per (layer, clip frame); every
// vector frame's flow is decoded, filled and scaled once (vectorin.rs:112-121); a single vector
// layer's flow is inverted once where a composition reads the missing one (compose/vector.rs:
// 430-456); the compositions are the engine's with the effect's stages, the Gouraud forward warp
// and the taps above first (glue.rs:253), batched over the output frames. The caller
// synchronises the stream and checks `bad` ([0] a NaN warp coordinate, [1 + i] output frame i's
// has-a-hole flag, [1 + count] the forward warps' flags).
static int vectors_in_enqueue(twx_handle* h, DevPool& pool, const twx_vin_params& p, const VinClip& input,
                              const VinClip& source, const VinClip& prev, const VinClip& next, int first, int count,
                              uint8_t* out, int* bad, std::string& err) {
    hipStream_t s = h->stream;
    const int W = p.width, H = p.height, fmt = p.fmt;
    const size_t fb = image_bytes(W, H, fmt), npx = (size_t)W * H;
    timeline::Rates r;
    if (timeline::vin_rates(p, r) || (!prev.wire && !next.wire)) {
        // the rate error or no vector layer: the size mismatch fill (image.rs:391-425), red at half
        // alpha in the wire's R, G, B, A order, without the finish
        std::vector<uint8_t> fill(fb);
        for (size_t i = 0; i < npx; i++) {
            if (fmt == TWX_FMT_U8) {
                const uint8_t v[4] = {0xff, 0, 0, 0x80};
                memcpy(&fill[i * 4], v, 4);
            } else if (fmt == TWX_FMT_U16) {
                auto q = [](uint32_t v) { return (uint16_t)(((v < 32768u ? v : 32768u) * 65535u + 16384u) / 32768u); };
                const uint16_t v[4] = {q(0x8000), 0, 0, q(0x4000)};
                memcpy(&fill[i * 8], v, 8);
            } else {
                float* pl = (float*)fill.data();  // planes G, B, R, A
                pl[i] = 0.0f;
                pl[npx + i] = 0.0f;
                pl[2 * npx + i] = 1.0f;
                pl[3 * npx + i] = 0.5f;
            }
        }
        uint8_t* d_fill = pool.upload(fill);
        for (int i = 0; i < count; i++)
            HIPCHECK(hipMemcpyAsync(out + (size_t)i * fb, d_fill, fb, hipMemcpyDeviceToDevice, s));
        return TWX_OK;
    }
    const VinClip& layer = source.wire ? source : input;  // the retimed layer (glue.rs:86-89)
    // the stage set: the effect's (glue.rs:57-75: Gouraud forward warp, taps above first, every
    // frame turned upside down and back) or the engine's plain set on rows as given
    const bool plain = p.plain_stages != 0;
    const bool gouraud = !plain, above = !plain, flip = !plain;
    const twx_retime_params rp = timeline::vin_speed_params(p);
    timeline::SpeedMap map(rp, r);
    // the frames each output frame reads: (layer 0 retimed / 1 Vectors: Prev / 2 Vectors: Next, clip
    // frame or -1 outside) -> an engine image slot
    std::map<std::pair<int, int64_t>, int> slot_of;
    std::vector<IngestOp> ops;
    const VinClip* clips[3] = {&layer, &prev, &next};
    auto slot = [&](int lid, int64_t time) {
        const int64_t idx = timeline::vin_clip_index(time, p.time_step, clips[lid]->n);
        auto it = slot_of.find({lid, idx});
        if (it != slot_of.end()) return it->second;
        IngestOp op{};
        op.wire = idx >= 0 ? clips[lid]->wire + (size_t)idx * fb : nullptr;
        op.engine = pool.alloc<uint8_t>(fb);
        op.W = W;
        op.H = H;
        op.fmt = fmt;
        op.zero = idx < 0 ? 1 : 0;
        op.reverse = flip ? 1 : 0;
        ops.push_back(op);
        const int k = (int)ops.size() - 1;
        slot_of[{lid, idx}] = k;
        return k;
    };
    struct VPlan {
        int kind = 0;  // 0 the source display, 1 the fast path, 2 the vector path
        twx_selection sel{};
        int sa = -1, sb = -1, vab = -1, vba = -1;  // slots of A, B, the Next image at A, the Prev image at B
        bool need_ab = false, need_ba = false;
    };
    std::vector<VPlan> plans(count);
    const int warp_mode = p.warping == 2 ? 2 : 0;  // glue.rs:96-103
    for (int i = 0; i < count; i++) {
        VPlan& pl = plans[i];
        const int32_t n = first + i;
        if (p.display == 2) {  // compose.rs:62-71: the retimed layer at current_time
            pl.kind = 0;
            pl.sa = slot(0, (int64_t)n * p.time_step);
            continue;
        }
        int st = timeline::vin_select(p, map, r, n, pl.sel, err);
        if (st) return st;
        if (pl.sel.interp < 0 || pl.sel.interp > 2) { err = "frame_interp outside 1..3"; return TWX_E_MISMATCH; }
        const int64_t ta = timeline::vin_time(r, (float)pl.sel.a), tb = timeline::vin_time(r, (float)pl.sel.b);
        pl.sa = slot(0, ta);
        pl.sb = slot(0, tb);
        if (pl.sel.vectors_quality != 0 && !pl.sel.exact) {
            pl.kind = 2;
            if (next.wire) pl.vab = slot(2, ta);
            if (prev.wire) pl.vba = slot(1, tb);
            pl.need_ab = pl.need_ba = true;
            if (pl.sel.interp == 0) {  // nearest_uses_f_ab (compose/vector.rs:61-64)
                const bool ab = (pl.sel.wa < 0.5f && warp_mode < 2) || (0.5f <= pl.sel.wa && warp_mode > 1);
                pl.need_ab = ab;
                pl.need_ba = !ab;
            }
        } else {
            pl.kind = 1;
        }
    }
    if (!ops.empty()) launch_ingest_batch(s, pool.upload(ops), ops);
    // the layers' flows (vectorin.rs:200-225): Max Displace read once, Vec Scale constant
    const float md = (1.0f / 1.0f) * timeline::read_fixed(timeline::to_fixed(p.max_displace), 1.0f);
    const float vsx = timeline::read_fixed(timeline::to_fixed(p.vec_scale_x), 1.0f);
    const float vsy = timeline::read_fixed(timeline::to_fixed(p.vec_scale_y), 1.0f);
    std::map<int, float*> flow_of;  // vector slot -> its flow
    auto flow = [&](int vslot) -> float* {
        auto it = flow_of.find(vslot);
        if (it != flow_of.end()) return it->second;
        float* f = pool.alloc<float>(npx * 2);
        int st = flow_from_vectors_dev(s, pool, ops[vslot].engine, W, H, fmt, md, vsx, vsy, above, f, err);
        if (st) throw st;
        flow_of[vslot] = f;
        return f;
    };
    std::map<int, float*> inverse_of;  // vector slot -> the inversion of its flow
    int* fw_flags = bad + 1 + count;
    auto inverse = [&](int vslot) -> float* {
        auto it = inverse_of.find(vslot);
        if (it != inverse_of.end()) return it->second;
        float* f = pool.alloc<float>(npx * 2);
        int st = invert_flow_dev(h, pool, flow(vslot), nullptr, W, H, above, !gouraud, f, fw_flags, err);
        if (st) throw st;
        inverse_of[vslot] = f;
        return f;
    };
    std::vector<FrameJob> jobs;
    try {
        for (int i = 0; i < count; i++) {
            const VPlan& pl = plans[i];
            FrameJob F;
            memset(&F, 0, sizeof(F));
            F.a = ops[pl.sa].engine;
            F.b = pl.sb >= 0 ? ops[pl.sb].engine : F.a;
            F.out = out + (size_t)i * fb;
            F.bad = bad;
            F.fmt = fmt;
            F.W = W;
            F.H = H;
            F.finish = flip ? 1 : 2;
            if (pl.kind == 0) {  // the frame through the prepare and the finish
                F.kind = 0;
                F.blend_mode = 0;
                h->stats.fast_frames++;
                jobs.push_back(F);
                continue;
            }
            F.wa = pl.sel.wa;
            F.frac = 1.0f - pl.sel.wa;
            const int interp = pl.sel.interp;
            if (pl.kind == 1) {  // compose.rs:214-246
                F.kind = 0;
                h->stats.fast_frames++;
                if (interp == 0) {
                    F.blend_mode = pl.sel.wa >= 0.5f ? 0 : 1;
                } else if (interp == 1) {
                    blend_branch((float)(1.0 - (double)pl.sel.wa), F);
                } else {
                    blend_branch(1.0f - pl.sel.wa, F);
                    F.plus_zero = 1;
                }
                jobs.push_back(F);
                continue;
            }
            if (W < 2 || H < 2) { err = "the warps need images of at least 2x2"; return TWX_E_MISMATCH; }
            h->stats.composes++;
            // the flows: as loaded, the missing one inverted where this composition reads it
            const float* f_ab = pl.vab >= 0 ? flow(pl.vab) : nullptr;
            const float* f_ba = pl.vba >= 0 ? flow(pl.vba) : nullptr;
            if (!f_ab && pl.need_ab) f_ab = inverse(pl.vba);
            if (!f_ba && pl.need_ba) f_ba = inverse(pl.vab);
            F.f_ab = f_ab;
            F.f_ba = f_ba;
            F.fw = W;
            F.fh = H;
            F.warp_a_copy = warp_is_copy(F.frac, W, W) ? 1 : 0;
            F.warp_b_copy = warp_is_copy(F.wa, W, W) ? 1 : 0;
            if (warp_mode == 2) {
                // forward.rs:391-572 with coverage and fill: A pushed with F_AB at 1 - wa, B with F_BA
                // at wa; Nearest the one frame nearest_uses_f_ab names
                int st;
                if (interp == 0) {
                    const bool from_a = !(0.5f > pl.sel.wa);
                    void* dst = pool.alloc<uint8_t>(fb);
                    uint8_t* cov = pool.alloc<uint8_t>(npx);
                    st = forward_warp_dev(h, pool, from_a ? F.a : F.b, W, H, fmt, from_a ? f_ab : f_ba, nullptr, W, H,
                                          from_a ? F.frac : F.wa, !gouraud, dst, cov, fw_flags, err);
                    F.pre_a = dst;
                    F.pre_cov_a = cov;
                } else {
                    void* ta = pool.alloc<uint8_t>(fb);
                    void* tb = pool.alloc<uint8_t>(fb);
                    uint8_t* c1 = pool.alloc<uint8_t>(npx);
                    uint8_t* c2 = pool.alloc<uint8_t>(npx);
                    st = forward_warp_dev(h, pool, F.a, W, H, fmt, f_ab, nullptr, W, H, F.frac, !gouraud, ta, c1,
                                          fw_flags, err);
                    if (!st)
                        st = forward_warp_dev(h, pool, F.b, W, H, fmt, f_ba, nullptr, W, H, F.wa, !gouraud, tb, c2,
                                              fw_flags, err);
                    F.pre_a = ta;
                    F.pre_b = tb;
                    F.pre_cov_a = c1;
                    F.pre_cov_b = c2;
                }
                if (st) return st;
                setup_smart(pool, F, interp, above, bad + 1 + i);
            } else if (interp == 0) {
                F.kind = 1;
                F.nearest_from_a = 0.5f <= pl.sel.wa ? 1 : 0;
            } else if (interp == 1) {
                F.kind = 2;
                blend_branch(F.frac, F);
            } else {
                setup_weighted(pool, F, above);
            }
            jobs.push_back(F);
        }
    } catch (int st) {
        return st;
    }
    if (!jobs.empty()) {
        for (auto& F : jobs) frame_job_ratios(F);
        FrameJob* dj = pool.upload(jobs);
        launch_frames(s, dj, jobs);
        launch_weighted(s, dj, jobs);
        int st = run_fills(s, pool, jobs, err);
        if (st) return st;
        launch_finish(s, dj, jobs);
    }
    return TWX_OK;
}

// Vectors In over output frames [first, first + count): checks, then chunks of output frames that
// fit the memory budget, each enqueued, synchronised and checked.
static int vectors_in_run(twx_handle* h, const twx_vin_params& p, const VinClip& input, const VinClip& source,
                          const VinClip& prev, const VinClip& next, int first, int count, uint8_t* out, std::string& err) {
    if (p.width <= 0 || p.height <= 0 || p.fmt < 0 || p.fmt > 2 || p.time_step <= 0 || count < 0) {
        err = "vectors-in: bad frame size, format, time step or frame count";
        return TWX_E_MISMATCH;
    }
    if (int st = check_frame_pixels(p.width, p.height, err)) return st;
    if (p.frame_interp < 1 || p.frame_interp > 3 || p.display < 1 || p.display > 6) {
        err = "vectors-in: popup out of range";
        return TWX_E_MISMATCH;
    }
    if (p.blur != 0.0) {
        err = "vectors-in: the blur compensation tail is not served by this backend";
        return TWX_E_UNSUPPORTED;
    }
    if (!input.wire || input.n < 0 || (source.wire && source.n < 0) || (prev.wire && prev.n < 0) || (next.wire && next.n < 0)) {
        err = "vectors-in: missing input clip";
        return TWX_E_MISMATCH;
    }
    const double t0 = now_ms();
    const size_t fb = image_bytes(p.width, p.height, p.fmt), npx = (size_t)p.width * p.height;
    // a frame reads at most four images and holds two flows, an inversion, two pushed frames and
    // their coverages, the fill's pyramids: about 10 frames and 64 bytes a pixel
    const size_t per_frame = 10 * fb + 64 * npx;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(64, (h->budget / 4) / std::max<size_t>(per_frame, 1)));
    for (int c0 = 0; c0 < count; c0 += chunk) {
        const int n = std::min(chunk, count - c0);
        DevPool pool(h);
        int* bad = pool.alloc<int>(2 + (size_t)n);
        launch_zero(h->stream, bad, 2 + (size_t)n);
        int st = vectors_in_enqueue(h, pool, p, input, source, prev, next, first + c0, n, out + (size_t)c0 * fb, bad, err);
        if (st) return st;
        int* hb = pool.pinned<int>(2 + (size_t)n);
        HIPCHECK(hipMemcpyAsync(hb, bad, (2 + (size_t)n) * sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHECK(hipStreamSynchronize(h->stream));
        if (hb[0]) {
            err = "inverse warp: a sample coordinate is NaN or never converges, which is not supported";
            return TWX_E_UNSUPPORTED;
        }
        if ((st = check_fw_flags(hb[1 + n], err))) return st;
    }
    h->stats.device_ms += now_ms() - t0;
    return TWX_OK;
}

int twx_vectors_in_frames_dev(twx_handle* h, const twx_vin_params* p, const void* input, int32_t n_input,
                              const void* source, int32_t n_source, const void* prev, int32_t n_prev, const void* next,
                              int32_t n_next, int32_t first, int32_t count, void* out, char* err, size_t errlen) {
    return entry(h, err, errlen, [&](std::string& e) -> int {
        if (!p) { e = "vectors-in: no parameters"; return TWX_E_MISMATCH; }
        after_caller(h);
        return vectors_in_run(h, *p, VinClip{(const uint8_t*)input, n_input}, VinClip{(const uint8_t*)source, n_source},
                              VinClip{(const uint8_t*)prev, n_prev}, VinClip{(const uint8_t*)next, n_next}, first, count,
                              (uint8_t*)out, e);
    });
}

int twx_vectors_in_frames(twx_handle* h, const twx_vin_params* p, const void* input, int32_t n_input, const void* source,
                          int32_t n_source, const void* prev, int32_t n_prev, const void* next, int32_t n_next,
                          int32_t first, int32_t count, void* out, char* err, size_t errlen) {
    return entry(h, err, errlen, [&](std::string& e) -> int {
        if (!p || p->fmt < 0 || p->fmt > 2 || p->width <= 0 || p->height <= 0 || count < 0) {
            e = "vectors-in: bad parameters";
            return TWX_E_MISMATCH;
        }
        if (int st = check_frame_pixels(p->width, p->height, e)) return st;
        hipStream_t s = h->stream;
        const size_t fb = image_bytes(p->width, p->height, p->fmt);
        // the clips in memory of their own (the render's chunks reset the handle's arena)
        std::vector<void*> owned;
        auto up = [&](const void* src, int32_t n) -> const uint8_t* {
            if (!src || n <= 0) return (const uint8_t*)src;
            void* d = nullptr;
            HIPCHECK(hipMalloc(&d, fb * n));
            owned.push_back(d);
            HIPCHECK(hipMemcpyAsync(d, src, fb * n, hipMemcpyHostToDevice, s));
            return (const uint8_t*)d;
        };
        int st = TWX_OK;
        void* dout = nullptr;
        try {
            const uint8_t *di = up(input, n_input), *ds = up(source, n_source), *dp = up(prev, n_prev), *dn = up(next, n_next);
            if (count > 0) {
                HIPCHECK(hipMalloc(&dout, fb * count));
                owned.push_back(dout);
            }
            st = vectors_in_run(h, *p, VinClip{di, n_input}, VinClip{ds, n_source}, VinClip{dp, n_prev}, VinClip{dn, n_next},
                                first, count, (uint8_t*)dout, e);
            if (!st && count > 0) {
                HIPCHECK(hipMemcpyAsync(out, dout, fb * count, hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
            }
        } catch (...) {
            (void)hipStreamSynchronize(s);
            for (void* d : owned) (void)hipFree(d);
            throw;
        }
        (void)hipStreamSynchronize(s);
        for (void* d : owned) (void)hipFree(d);
        return st;
    });
}

int twx_compose_vectors(twx_handle* h, int32_t interp, float wa, const void* a, const void* b, int32_t width,
                        int32_t height, int32_t fmt, const float* f_ab, const float* f_ba, int32_t flow_w,
                        int32_t flow_h, int32_t warp_mode, int32_t taps_above_first, int32_t forward_gouraud,
                        void* out, char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (!g.ok()) { set_err(err, errlen, e); return st; }
        if (interp < 0 || interp > 2) {
            set_err(err, errlen, "interpolation mode outside 0 (Nearest), 1 (Blend), 2 (Weighted Blend)");
            return TWX_E_MISMATCH;
        }
        if (warp_mode < 0 || warp_mode > 3) {
            set_err(err, errlen, "warping code outside 0..3");
            return TWX_E_MISMATCH;
        }
        const bool forward = warp_mode >= 2, textured = forward_gouraud == 0;
        if (width < 2 || height < 2 || flow_w < 2 || flow_h < 2 || fmt < 0 || fmt > 2) {
            set_err(err, errlen, "the warps need images and a flow of at least 2x2");
            return TWX_E_MISMATCH;
        }
        if ((st = check_frame_pixels(width, height, e))) { set_err(err, errlen, e); return st; }
        hipStream_t s = h->stream;
        double t0 = now_ms();
        DevPool pool(h);
        size_t fb = image_bytes(width, height, fmt);
        size_t fn = (size_t)flow_w * flow_h * 2;
        uint8_t* img = pool.alloc<uint8_t>(fb * 3);
        float* fl = pool.alloc<float>(fn * 2);
        HIPCHECK(hipMemcpyAsync(img, a, fb, hipMemcpyHostToDevice, s));
        HIPCHECK(hipMemcpyAsync(img + fb, b, fb, hipMemcpyHostToDevice, s));
        FrameJob F;
        memset(&F, 0, sizeof(F));
        F.a = img;
        F.b = img + fb;
        F.out = img + 2 * fb;
        F.fmt = fmt;
        F.W = width;
        F.H = height;
        F.fw = flow_w;
        F.fh = flow_h;
        F.wa = wa;
        F.frac = 1.0f - wa;
        F.finish = 0;
        F.warp_a_copy = warp_is_copy(F.frac, flow_w, width) ? 1 : 0;
        F.warp_b_copy = warp_is_copy(wa, flow_w, width) ? 1 : 0;
        // compose/vector.rs:61-64 (nearest_uses_f_ab): the one flow of Nearest
        const bool nearest_ab = (wa < 0.5f && !forward) || (0.5f <= wa && forward);
        bool need_ab = interp != 0 || nearest_ab;
        bool need_ba = interp != 0 || !nearest_ab;
        if (need_ab) {
            if (!f_ab) { set_err(err, errlen, "the composition needs the flow F_AB"); return TWX_E_MISMATCH; }
            HIPCHECK(hipMemcpyAsync(fl, f_ab, fn * 4, hipMemcpyHostToDevice, s));
            F.f_ab = fl;
        }
        if (need_ba) {
            if (!f_ba) { set_err(err, errlen, "the composition needs the flow F_BA"); return TWX_E_MISMATCH; }
            HIPCHECK(hipMemcpyAsync(fl + fn, f_ba, fn * 4, hipMemcpyHostToDevice, s));
            F.f_ba = fl + fn;
        }
        int* bad = pool.alloc<int>(3);
        HIPCHECK(hipMemsetAsync(bad, 0, 3 * sizeof(int), s));
        bool direct = false;  // warping code 3 Nearest: the pushed frame is the output
        if (forward) {
            // forward.rs:391-572: each frame pushed with its own flow (A with F_AB at 1 - wa, B with
            // F_BA at wa; Nearest the one flow), with coverage for code 2
            const bool fill = warp_mode == 2;
            const size_t np = (size_t)width * height;
            if (interp == 0) {
                const bool from_a = !(0.5f > wa);
                void* dst = fill ? (void*)pool.alloc<uint8_t>(fb) : F.out;
                uint8_t* cov = fill ? pool.alloc<uint8_t>(np) : nullptr;
                st = forward_warp_dev(h, pool, from_a ? F.a : F.b, width, height, fmt, from_a ? F.f_ab : F.f_ba,
                                      from_a ? f_ab : f_ba, flow_w, flow_h, from_a ? 1.0f - wa : wa, textured, dst, cov,
                                      bad + 2, e);
                F.pre_a = dst;
                F.pre_cov_a = cov;
                direct = !fill;
            } else {
                void* ta = pool.alloc<uint8_t>(fb);
                void* tb = pool.alloc<uint8_t>(fb);
                uint8_t* c1 = fill ? pool.alloc<uint8_t>(np) : nullptr;
                uint8_t* c2 = fill ? pool.alloc<uint8_t>(np) : nullptr;
                st = forward_warp_dev(h, pool, F.a, width, height, fmt, F.f_ab, f_ab, flow_w, flow_h, F.frac, textured,
                                      ta, c1, bad + 2, e);
                if (!st)
                    st = forward_warp_dev(h, pool, F.b, width, height, fmt, F.f_ba, f_ba, flow_w, flow_h, wa, textured, tb,
                                          c2, bad + 2, e);
                F.pre_a = ta;
                F.pre_b = tb;
                F.pre_cov_a = c1;
                F.pre_cov_b = c2;
            }
            if (st) { set_err(err, errlen, e); return st; }
            if (fill) {
                setup_smart(pool, F, interp, taps_above_first != 0, bad + 1);
            } else if (interp == 0) {
                F.kind = 1;
            } else if (interp == 1) {
                F.kind = 2;
                blend_branch(F.frac, F);
            } else {
                setup_weighted(pool, F, taps_above_first != 0);
            }
        } else if (warp_mode == 1) {
            setup_smart(pool, F, interp, taps_above_first != 0, bad + 1);
        } else if (interp == 0) {
            F.kind = 1;
            F.nearest_from_a = 0.5f <= wa ? 1 : 0;
        } else if (interp == 1) {
            F.kind = 2;
            blend_branch(F.frac, F);
        } else {
            setup_weighted(pool, F, taps_above_first != 0);
        }
        F.bad = bad;
        if (!direct) {
            std::vector<FrameJob> jobs{F};
            frame_job_ratios(jobs[0]);
            FrameJob* dj = pool.upload(jobs);
            launch_frames(s, dj, jobs);
            launch_weighted(s, dj, jobs);
            if ((st = run_fills(s, pool, jobs, e))) { set_err(err, errlen, e); return st; }
            launch_finish(s, dj, jobs);
        }
        int* hbad = pool.pinned<int>(3);
        HIPCHECK(hipMemcpyAsync(out, img + 2 * fb, fb, hipMemcpyDeviceToHost, s));
        HIPCHECK(hipMemcpyAsync(hbad, bad, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHECK(hipStreamSynchronize(s));
        h->stats.composes++;
        h->stats.device_ms += now_ms() - t0;
        if (hbad[0]) st = TWX_E_UNSUPPORTED, e = "inverse warp: a sample coordinate is NaN or never converges, which is not supported";
        else st = check_fw_flags(hbad[2], e);
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

int twx_select(const twx_retime_params* p, int32_t n, twx_selection* out, char* err, size_t errlen) {
    std::string e;
    int st = timeline::check_params(*p, e);
    timeline::Rates r;
    if (!st) st = timeline::rates(*p, r, e);
    if (!st) {
        timeline::SpeedMap map(*p, r);
        st = timeline::select(*p, map, n, *out, e);
    }
    if (st) set_err(err, errlen, e);
    return st;
}

int twx_retime_frames_dev(twx_handle* h, const twx_retime_params* p, const void* clip, int32_t n_in, int32_t first,
                          int32_t count, void* out, char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) {
            after_caller(h);
            st = retime_batch_run(h, {ClipRun{*p, (const uint8_t*)clip, n_in, first, count, (uint8_t*)out}}, e);
        }
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

int twx_retime_batch(twx_handle* h, int32_t n_clips, const twx_retime_params* params, const void* const* clips,
                     const int32_t* n_in, const int32_t* first, const int32_t* count, void* const* outs, char* err,
                     size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) {
            if (n_clips < 0 || (n_clips > 0 && (!params || !clips || !n_in || !first || !count || !outs))) {
                e = "batch: missing arrays";
                st = TWX_E_MISMATCH;
            } else {
                std::vector<ClipRun> runs;
                for (int i = 0; i < n_clips; i++)
                    runs.push_back(ClipRun{params[i], (const uint8_t*)clips[i], n_in[i], first[i], count[i], (uint8_t*)outs[i]});
                after_caller(h);
                st = retime_batch_run(h, runs, e);
            }
        }
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

int twx_retime_frames(twx_handle* h, const twx_retime_params* p, const void* clip, int32_t n_in, int32_t first,
                      int32_t count, void* out, char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (!g.ok()) { set_err(err, errlen, e); return st; }
        if ((st = timeline::check_params(*p, e))) { set_err(err, errlen, e); return st; }
        hipStream_t s = h->stream;
        size_t fb = image_bytes(p->width, p->height, p->fmt);
        DevPool pool(h);
        uint8_t* dclip = pool.alloc<uint8_t>(fb * std::max(n_in, 1));
        uint8_t* dout = pool.alloc<uint8_t>(fb * std::max(count, 1));
        if (n_in > 0) HIPCHECK(hipMemcpyAsync(dclip, clip, fb * n_in, hipMemcpyHostToDevice, s));
        st = retime_frames_dev(h, pool, *p, dclip, n_in, first, count, dout, e);
        if (!st) {
            HIPCHECK(hipMemcpyAsync(out, dout, fb * count, hipMemcpyDeviceToHost, s));
            HIPCHECK(hipStreamSynchronize(s));
        }
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

static int warp_blend_entry(twx_handle* h, bool backward, int32_t channels, int32_t width, int32_t height, const float* a,
                            const float* b, const float* f_ab, const float* f_ba, int32_t flow_w, int32_t flow_h, float wa,
                            float* out, const float* grad_out, float* grad_a, float* grad_b, float* grad_f_ab,
                            float* grad_f_ba, char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) {
            after_caller(h);
            hipStream_t s = h->stream;
            DevPool pool(h);
            const size_t px = (size_t)std::max(width, 0) * std::max(height, 0);
            int* flags = pool.alloc<int>(2);
            HIPCHECK(hipMemsetAsync(flags, 0, 2 * sizeof(int), s));
            float* gp = backward ? pool.alloc<float>(px * 4) : nullptr;
            st = warp_blend_run(s, backward, channels, width, height, a, b, f_ab, f_ba, flow_w, flow_h, wa, out, grad_out,
                                grad_a, grad_b, grad_f_ab, grad_f_ba, gp, flags, e);
            if (!st) {
                int* hb = pool.pinned<int>(2);
                HIPCHECK(hipMemcpyAsync(hb, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
                HIPCHECK(hipStreamSynchronize(s));
                st = diff_flags_status(hb, e);
            }
        }
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

int twx_warp_blend(twx_handle* h, int32_t channels, int32_t width, int32_t height, const float* a, const float* b,
                   const float* f_ab, const float* f_ba, int32_t flow_w, int32_t flow_h, float wa, float* out, char* err,
                   size_t errlen) {
    return warp_blend_entry(h, false, channels, width, height, a, b, f_ab, f_ba, flow_w, flow_h, wa, out, nullptr, nullptr,
                            nullptr, nullptr, nullptr, err, errlen);
}

int twx_warp_blend_backward(twx_handle* h, int32_t channels, int32_t width, int32_t height, const float* a, const float* b,
                            const float* f_ab, const float* f_ba, int32_t flow_w, int32_t flow_h, float wa,
                            const float* grad_out, float* grad_a, float* grad_b, float* grad_f_ab, float* grad_f_ba,
                            char* err, size_t errlen) {
    return warp_blend_entry(h, true, channels, width, height, a, b, f_ab, f_ba, flow_w, flow_h, wa, nullptr, grad_out,
                            grad_a, grad_b, grad_f_ab, grad_f_ba, err, errlen);
}

int twx_warp_blend_backward_noalloc(twx_handle* h, int32_t channels, int32_t width, int32_t height, const float* a,
                                    const float* b, const float* f_ab, const float* f_ba, int32_t flow_w, int32_t flow_h,
                                    float wa, const float* grad_out, float* grad_a, float* grad_b, float* grad_f_ab,
                                    float* grad_f_ba, char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) {
            g_prof = nullptr;  // no profiling events: they are taken from a growing pool
            after_caller(h);
            st = warp_blend_backward_noalloc(h->stream, h->device, channels, width, height, a, b, f_ab, f_ba, flow_w,
                                             flow_h, wa, grad_out, grad_a, grad_b, grad_f_ab, grad_f_ba, e);
        }
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

int twx_refine_flows(twx_handle* h, int32_t channels, int32_t width, int32_t height, const float* a, const float* b,
                     const float* truth, float* f_ab, float* f_ba, int32_t flow_w, int32_t flow_h, float wa, int32_t steps,
                     float lr, float* losses, char* err, size_t errlen) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) {
            if (steps < 0 || !a || !b || !truth || !f_ab || !f_ba || !losses || !(lr >= 0.0f)) {
                e = "refine flows: bad arguments";
                st = TWX_E_MISMATCH;
            } else {
                after_caller(h);
                hipStream_t s = h->stream;
                double t0 = now_ms();
                DevPool pool(h);
                const size_t px = (size_t)std::max(width, 0) * std::max(height, 0);
                const size_t fn = (size_t)std::max(flow_w, 0) * std::max(flow_h, 0) * 2;
                RefineBufs B{};
                B.gp = pool.alloc<float>(std::max<size_t>(px * 4, 1));
                B.g_ab = pool.alloc<float>(std::max<size_t>(fn, 1));
                B.g_ba = pool.alloc<float>(std::max<size_t>(fn, 1));
                float* mv = pool.alloc<float>(std::max<size_t>(fn * 4, 1));
                HIPCHECK(hipMemsetAsync(mv, 0, std::max<size_t>(fn * 4, 1) * sizeof(float), s));
                B.m_ab = mv;
                B.v_ab = mv + fn;
                B.m_ba = mv + 2 * fn;
                B.v_ba = mv + 3 * fn;
                B.loss_part = pool.alloc<float>((size_t)std::max(refine_loss_blocks(std::max(width, 1), std::max(height, 1)), 1));
                B.losses = pool.alloc<float>((size_t)steps + 1);
                B.flags = pool.alloc<int>(2);
                HIPCHECK(hipMemsetAsync(B.flags, 0, 2 * sizeof(int), s));
                st = refine_flows_run(s, channels, width, height, a, b, truth, f_ab, f_ba, flow_w, flow_h, wa, steps, lr, B, e);
                if (!st) {
                    int* hb = pool.pinned<int>(2);
                    HIPCHECK(hipMemcpyAsync(hb, B.flags, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
                    HIPCHECK(hipMemcpyAsync(losses, B.losses, ((size_t)steps + 1) * sizeof(float), hipMemcpyDeviceToHost, s));
                    HIPCHECK(hipStreamSynchronize(s));
                    st = diff_flags_status(hb, e);
                }
                h->stats.device_ms += now_ms() - t0;
            }
        }
    } catch (const HipError& x) {
        e = x.msg;
        st = TWX_E_DEVICE;
    }
    if (st) set_err(err, errlen, e);
    return st;
}

int twx_set_tile_shape(int32_t variant) { return tile_force(variant) == 0 ? TWX_OK : TWX_E_MISMATCH; }

int twx_tile_choices(int32_t* out, int32_t cap) { return tile_tuning_report(out, out ? cap : 0); }

int twx_set_coarse_solver(int32_t on) {
    coarse_solver_enable(on != 0);
    return TWX_OK;
}

int twx_set_graphs(twx_handle* h, int32_t on) {
    std::string e;
    int st = TWX_OK;
    try {
        Guard g(h, st, e);
        if (g.ok()) {
            h->graphs = on != 0;
            if (!h->graphs) h->graph.reset();
        }
    } catch (const HipError&) {
        st = TWX_E_DEVICE;
    }
    return st;
}

int twx_cache_stats(twx_handle* h, uint64_t* hits, uint64_t* misses, uint64_t* evictions, uint64_t* bytes_held) {
    if (!h) return TWX_E_MISMATCH;
    std::lock_guard<std::mutex> g(h->mu);
    if (hits) *hits = h->pyramids.hits;
    if (misses) *misses = h->pyramids.misses;
    if (evictions) *evictions = h->pyramids.evictions;
    if (bytes_held) *bytes_held = h->pyramids.used;
    return TWX_OK;
}

int twx_get_stats(twx_handle* h, twx_stats* out) {
    if (!h || !out) return TWX_E_MISMATCH;
    *out = h->stats;
    out->graph_replays = h->graph_replays;
    return TWX_OK;
}

int twx_profile_enable(twx_handle* h, int on) {
    if (!h) return TWX_E_MISMATCH;
    std::lock_guard<std::mutex> g(h->mu);
    h->prof.on = on != 0;
    return TWX_OK;
}

int twx_profile_reset(twx_handle* h) {
    if (!h) return TWX_E_MISMATCH;
    std::lock_guard<std::mutex> g(h->mu);
    h->prof.reset();
    return TWX_OK;
}

int twx_profile_read(twx_handle* h, int32_t family, const char** name, double* ms, double* bytes, uint64_t* launches) {
    if (!h || family < 0 || family >= F_NFAM) return TWX_E_MISMATCH;
    std::lock_guard<std::mutex> g(h->mu);
    if (name) *name = PROF_NAMES[family];
    if (ms) *ms = h->prof.ms[family];
    if (bytes) *bytes = h->prof.bytes[family];
    if (launches) *launches = h->prof.launches[family];
    return TWX_OK;
}

}  // extern "C"
