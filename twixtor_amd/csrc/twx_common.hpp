// Shared definitions of the HIP backend: device job tables and the host-side launch API.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/twx.h"

namespace twx {

constexpr int MAXL = 16;     // pyramid levels (a side of 65536 would give 16)
constexpr int MAX_TAPS = 8;  // half kernels used: r=7 -> 8, r=5 -> 6, binomial -> 3

// The frame kernels address a pixel of a frame by a 32-bit byte offset from the frame's base, and
// the widest pixel is 16 bytes (f32 A,R,G,B engine images, 4-channel float images): a frame must
// hold fewer than 2^28 pixels. Larger frames are refused (TWX_E_UNSUPPORTED, the CPU's cue), never
// rendered with wrapped offsets. Block matching forms texel offsets with 24-bit multiplies, so a
// side of 2^24 pixels or more is refused the same way.
constexpr uint64_t MAX_FRAME_PIXELS = 1ull << 28;
constexpr long long MAX_FRAME_SIDE = 1ll << 24;
inline int check_frame_pixels(long long w, long long h, std::string& err) {
    if (w > 0 && h > 0 && (uint64_t)w * (uint64_t)h >= MAX_FRAME_PIXELS) {
        err = "a frame of 2^28 pixels or more (" + std::to_string(w) + "x" + std::to_string(h) +
              ") is not served by this backend";
        return TWX_E_UNSUPPORTED;
    }
    if (w >= MAX_FRAME_SIDE || h >= MAX_FRAME_SIDE) {
        err = "a frame side of 2^24 pixels or more (" + std::to_string(w) + "x" + std::to_string(h) +
              ") is not served by this backend";
        return TWX_E_UNSUPPORTED;
    }
    return TWX_OK;
}

struct Taps {
    float k[MAX_TAPS];
    int n;
};

// A plane operation of the pyramid build, one per frame of a batched launch (blockIdx.z).
struct PlaneOp {
    const void* src;
    float* dst;
    int sw, sh, dw, dh;
    int fmt;   // with luma: the engine image format
    int luma;  // k_blur5_resample: src is an engine image whose luma is the plane
    float* aux;  // k_blur5_resample: where the band src - blur goes (null: none)
    float rx, ry;  // k_blur5_resample: (sw - 1) / (dw - 1) and (sh - 1) / (dh - 1), computed on the host
    // k_blur5_resample, a frame's first step: the plane is the luma of the engine image of this
    // wire frame (format fmt, sw x sh, rows reversed with `reverse`) instead of src; eng (may be
    // null) receives that engine image, written by the tiles as the ingest would write it
    const uint8_t* wire;
    uint8_t* eng;
    int reverse;
};

// One flow solve X -> Y over a pyramid pair (flow.rs:1258-1307). Device resident table entry;
// every per-level kernel reads its job from `jobs[blockIdx.z]` and skips levels above `start`.
struct SolveJob {
    const float* X[MAXL];  // band-pass levels of X (imgA of the solver), finest first
    const float* Y[MAXL];  // band-pass levels of Y (imgB)
    int w[MAXL], h[MAXL];
    int levels;
    int start;             // start level (flow.rs:179-197)
    // the coarsest level the per-level kernels run; the levels above it (top + 1 .. start, whose
    // planes fit in LDS) run in the coarse solver kernel, which hands level top its flow in flow[0].
    // top = start: no coarse part; top = -1: the whole solve runs there
    int top;
    float damp;
    int scale_damp;        // |damp - 1| >= 1e-6 and damp != 1
    float* flow[2];        // flow, double buffered across the upsample: level l lives in flow[l & 1]
    float* grad;           // 2 channels
    float* ta;             // structure tensor, 3 channels
    float* tb;             // right-hand side B, 2 channels
    float* tmp;            // block-matching scratch: the transposed flow (2 channels)
    // block-matching costs carried between sweeps (one float per pixel each): the cost of every
    // interior pixel's vector after a sweep, [0] in the row sweeps' walk layout, [1] plain
    float* cost[2];
    int* state;            // [0] range check still active, [1] any pixel in range, [2] NaN seen
    // 1: grad and ta are those of an earlier job of the same launch with the same X, which
    // computes them (they depend on X alone, flow.rs:1274-1277); the tensor kernel skips this job
    int tensor_shared;
    // the transposed X and Y of the level whose rows are being swept: one plane per distinct frame
    // of the launch, shared by its jobs (written by the plane transposes, PlaneT)
    const float* xt;
    const float* yt;
    int max_w, max_h;
    // upsample_flow from level l to l - 1 (flow.rs:376-391), the same for every pixel: the scales
    // sx = (dw - 1) / (sw - 1), sy and their reciprocals (through double), at index l. Filled by
    // solve_job_ratios before upload.
    float up[MAXL][4];
};

// Fills a job's per-level upsample ratios (IEEE division on the host, as on the device).
void solve_job_ratios(SolveJob& J);

// One output frame of a retime run (compose.rs:93-246 fast path, compose/vector.rs:249-276 vector
// path code 0), including the finish into wire bytes (image.rs:205-245).
struct FrameJob {
    const void* a;         // engine images in motion-row order (bottom row first)
    const void* b;
    const float* f_ab;     // flows at the working size (vector path)
    const float* f_ba;
    void* out;             // wire frame
    int* bad;              // set when a warp coordinate is NaN (StageError::Unsupported)
    int kind;              // 0 fast path, 1 vector Nearest, 2 vector Blend
    int fmt;
    int W, H, fw, fh;
    float wa, frac;        // Selection::wa and 1 - wa (the amounts of the two warps)
    float weight;          // blend weight
    int blend_mode;        // 0 copy first, 1 copy second, 2 copy first then lerp, 3 per-sample mix
    float lerp_w;          // lerp weight of the extrapolation branch
    int warp_a_copy;       // the |amount| early-out of the warp of A: a plain copy
    int warp_b_copy;
    int nearest_from_a;    // Nearest: warp A with F_BA (wa >= 0.5), else warp B with F_AB
    int plus_zero;         // weighted blend without flows on float images: inputs + 0.0 first
    int finish;            // 1: unpremultiply into the wire frame, rows turned back; 2: the same on rows as given; 0: engine pixel as is
    // 1: f_ab and f_ba are the solver's flows before its final negation (flow.rs:1305), which the
    // warps take into account by subtracting the scaled flow sample from the pixel position
    int flow_raw;
    // Weighted Blend on the vector path (kind 3, compose/vector.rs:268-273, filters.rs:649-656):
    // the warped frames, two buffers per change mask (ping-pong of the clean-up passes), the
    // difference-mask limit, the sharpening weights of A and B and the blur's vertical tap order.
    void* ta;
    void* tb;
    uint8_t* mask[2][2];   // [half A/B][buffer]
    float limit;
    float ca, cb;
    int above_first;
    int mask_small;        // a side under 3 pixels: the clean-up runs as one sequential pass
    // Smart Blend (warping code 1, kinds 4 Nearest, 5 Blend, 6 Weighted Blend; smartblend.rs:
    // 323-419): the composited engine image before the hole fill, the holes mask (255 - the
    // combined coverage) and the frame's has-a-hole flag that gates the fill.
    void* comp;
    uint8_t* holes;
    int* has_hole;
    int pad2;
    // Forward warps (warping codes 2 and 3, forward.rs:391-572): the pushed frames computed before
    // this job (engine images) and their coverages; when pre_a is set the job reads them in place
    // of its inverse warps (Nearest: pre_a is the one pushed frame).
    const void* pre_a;
    const void* pre_b;
    const uint8_t* pre_cov_a;
    const uint8_t* pre_cov_b;
    // The warps' per-frame ratios (warp.rs:500-512), the same for every pixel: flow step per pixel
    // (fw - 1) / (W - 1) and (fh - 1) / (H - 1), and each warp's amount over them ([0] the warp of
    // A by frac, [1] the warp of B by wa). Filled by frame_job_ratios before upload.
    float rx, ry;
    float sx[2], sy[2];
};

// Fills the per-frame warp ratios of a job (IEEE division on the host, as on the device).
void frame_job_ratios(FrameJob& F);
// A frame launch_frames sends through its kernel for the common vector-path frame (kind 2, both
// warps computed, no pushed frames).
__host__ __device__ inline bool frame_is_blend_pair(const FrameJob& F) {
    return F.kind == 2 && !F.pre_a && !F.warp_a_copy && !F.warp_b_copy;
}

constexpr int MAXFL = 24;  // levels of the hole-fill pyramid (holefill.rs:145-176)
constexpr int FILL_FLOW2 = 3;  // FillPlane.fmt of a two-channel float flow (holefill.rs:506-521)

// One plane of a batched hole fill (holefill.rs:214-336): a colour channel of a frame (chan 0..3)
// or the frame's weight plane (chan 4). Pull: val[k] = resample(blur5(val[k - 1])); push from
// sol[top] down to sol[0]. Every kernel of the fill returns at once when *gate is 0 (no hole).
struct FillPlane {
    float* val[MAXFL];        // pull pyramid (values, or the weights for chan 4)
    float* sol[MAXFL];        // push solution per level (colour planes)
    const float* wts[MAXFL];  // the frame's weight pyramid
    int w[MAXFL], h[MAXFL];
    int top;
    int chan;
    float* t1;                // scratch of the level-0 size
    float* t2;
    const int* gate;
    void* img;                // the frame's engine image (load and store)
    const uint8_t* holes;
    int fmt;
    int above_first;
    int lo;                   // levels lo .. top (1 <= lo <= top) run in one workgroup in LDS
};

// LDS floats of the fill's coarse levels: the value, weight and solution pyramids of levels
// lo .. top plus two planes of level lo's size
constexpr int FILL_LDS_FLOATS = 32768;

// Per-kernel-family timing with HIP events on the launching stream, and the algorithmic bytes of
// each launch (the step-boundary model: distinct inputs read once, outputs written once).
enum ProfFam {
    F_INGEST, F_PYR_BLUR, F_GRAD_TENSOR, F_BM_TRANSPOSE,
    F_SOLVE_INIT, F_CALC_B, F_NEW_DISP, F_BM_ROWS, F_BM_COLS, F_RANGE, F_UPSAMPLE, F_NEGATE, F_FRAMES,
    F_WB_MASK, F_WB_SHARPEN, F_FILL_PULL, F_FILL_PUSH, F_FILL_IO, F_DIFF_FWD, F_DIFF_BWD, F_DIFF_GATHER,
    F_DIFF_IMG, F_DIFF_STEP, F_DIFF_ADAM,
    F_FW_MESH, F_FW_RASTER, F_FW_SCAN, F_FW_FOLD, F_VEC, F_COARSE, F_NFAM
};
extern const char* const PROF_NAMES[F_NFAM];

struct Prof {
    bool on = false;
    struct Rec {
        int fam;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> spare;
    double ms[F_NFAM] = {};
    double bytes[F_NFAM] = {};
    uint64_t launches[F_NFAM] = {};
    hipEvent_t take();
    void flush();  // after a stream synchronisation: collect the elapsed times
    void reset();
    ~Prof();
};
// The profiler of the call in progress on this thread (null when off).
extern thread_local Prof* g_prof;
hipEvent_t prof_begin(hipStream_t s);
void prof_end(hipStream_t s, int fam, double bytes, hipEvent_t a);

#define TWX_LAUNCH(stream, fam, bytes, ...)                 \
    do {                                                    \
        hipEvent_t e0_ = prof_begin(stream);                \
        __VA_ARGS__;                                        \
        prof_end(stream, fam, (double)(bytes), e0_);        \
    } while (0)

// One source frame of a batch: wire frame -> engine image (rows reversed), or zeros.
struct IngestOp {
    const uint8_t* wire;
    uint8_t* engine;
    int W, H, fmt;
    int zero;     // a frame outside the clip: transparent
    int reverse;  // the motion path's row order
    int pad;
};

// Consecutive ranges of a table sorted by area, largest first, in which the largest area is at
// most twice the smallest: a mixed-size batch launches once per range, one size gives one range.
std::vector<std::pair<int, int>> size_buckets(const std::vector<double>& areas);

// Host side launchers (twx_device.hip). Tables with several frame sizes are sorted by area,
// largest first.
void launch_ingest(hipStream_t s, const void* wire, void* engine, int W, int H, int fmt, bool reverse_rows);
void launch_ingest_batch(hipStream_t s, const IngestOp* d_ops, const std::vector<IngestOp>& ops);
void launch_blur5_resample(hipStream_t s, const PlaneOp* d_ops, const std::vector<PlaneOp>& ops, int maxw, int maxh,
                           const Taps& t);
// A distinct frame of a solve launch: its band-pass levels and the plane that receives the
// transposed level for the row sweeps of its jobs (levels up to start).
struct PlaneT {
    const float* lv[MAXL];
    float* dst;
    int w[MAXL], h[MAXL];
    int top;  // the coarsest level its jobs' per-level kernels run (SolveJob::top)
};
// A size bucket of a solve launch: jobs [j0, j1) and the frames [f0, f1) they read. The jobs that
// compute a gradient and tensor (tensor_shared 0) lead the bucket's jobs. queue: a device counter
// for the coarse solver's work queue.
struct SolveBucket {
    int j0, j1, f0, f1;
    int* queue;
};
// The largest level (in pixels) the coarse solver kernel holds in LDS: 15 floats per pixel.
constexpr int COARSE_MAX_PIXELS = 2560;
// the coarse solver on or off for the solves enqueued from now on (on by default)
void coarse_solver_enable(bool on);
bool coarse_solver_enabled();
// raw: the flows stay as the last level leaves them, without the final negation (their only
// readers are inverse warps that know it); the NaN scan then runs in the last Gaussian pass.
void launch_solve(hipStream_t s, const SolveJob* d_jobs, const std::vector<SolveJob>& jobs, const PlaneT* d_planes,
                  const std::vector<PlaneT>& planes, const std::vector<SolveBucket>& buckets, bool skip_refine,
                  bool raw);
void launch_frames(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs);
// The Weighted Blend jobs (kinds 3 and 6) after launch_frames: mask clean-up, sharpening, then
// blend and finish (kind 3) or the Smart Blend composite into the frame's image (kind 6).
void launch_weighted(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs);
// The hole fill of a set of planes (every level's buffers set up by the caller); `maxw`, `maxh`
// per level over the planes.
void launch_fill(hipStream_t s, const FillPlane* d_planes, const std::vector<FillPlane>& planes);
// The Smart Blend frames (kinds 4-6) with finish set: the filled image into the wire frame.
void launch_finish(hipStream_t s, const FrameJob* d_jobs, const std::vector<FrameJob>& jobs);
void launch_nan_scan(hipStream_t s, const float* data, size_t n, int* flag);
// Clears n ints with a kernel (a memset node would not replay correctly inside a captured graph).
void launch_zero(hipStream_t s, int* p, size_t n);
// Tile shapes of the fused producer + Gaussian kernels: variant v in 0..3 forced for every
// launch (-1 returns to measured choice); the measured choices as (kernel, log2 level area,
// log2 job count, variant) quadruples, their number returned.
int tile_force(int v);
int tile_tuning_report(int* out, int cap);
int run_selfcheck(hipStream_t s, std::string& err);
// The differentiable warp+blend (twx_diff.hip): the forward into `out`, or the backward into the
// gradients that are not null; gp_scratch holds 4 floats per pixel; flags (2 ints, zeroed by the
// caller) receive the NaN-point flag and the largest upstream magnitude, read back into
// diff_flags_status after the stream's work.
int warp_blend_run(hipStream_t s, bool backward, int C, int W, int H, const float* a, const float* b, const float* f_ab,
                   const float* f_ba, int fw, int fh, float wa, float* out, const float* grad_out, float* grad_a,
                   float* grad_b, float* grad_f_ab, float* grad_f_ba, float* gp_scratch, int* flags, std::string& err);
int diff_flags_status(const int flags[2], std::string& err);
// Photometric refinement on the device: `steps` Adam steps of the flows (in place) on the mean
// squared error of the warp+blend against `truth`; losses[t] (device) the loss before step t,
// losses[steps] after the last. The buffers: gp 4 floats per pixel, g/m/v fw * fh * 2 floats each
// (m and v zeroed), loss_part refine_loss_blocks(W, H) floats, losses steps + 1, flags 2 ints zeroed.
struct RefineBufs {
    float* gp;
    float* g_ab;
    float* g_ba;
    float* m_ab;
    float* v_ab;
    float* m_ba;
    float* v_ba;
    float* loss_part;
    float* losses;
    int* flags;
};
int refine_loss_blocks(int W, int H);
int refine_flows_run(hipStream_t s, int C, int W, int H, const float* a, const float* b, const float* truth, float* f_ab,
                     float* f_ba, int fw, int fh, float wa, int steps, float lr, const RefineBufs& B, std::string& err);
// The backward of the warp+blend on the library's fixed per-device scratch (no allocation): the
// scratch address of `device` is taken once by noalloc_scratch_setup with that device current.
int noalloc_scratch_setup(int device, std::string& err);
int warp_blend_backward_noalloc(hipStream_t s, int device, int C, int W, int H, const float* a, const float* b,
                                const float* f_ab, const float* f_ba, int fw, int fh, float wa, const float* grad_out,
                                float* grad_a, float* grad_b, float* grad_f_ab, float* grad_f_ba, std::string& err);

// The forward warp (twx_forward.hip): device buffers of one warp. The caller allocates pos (W*H),
// counts and fill (W*H, zeroed), offsets (W*H), block_totals (fw_scan_blocks(W*H)), total (1) and
// flags (1, zeroed); after launch_fw_count and a read of *total, rec_t and rec_c (total records
// each), then launch_fw_emit_fold. flags: 1 the last row pass would write outside the image, 2 a
// triangle over the per-triangle hit cap, 4 a NaN vertex colour (Gouraud), 8 and 16 a hit list
// inconsistent with the count pass, 32 a textured tap below the image, 64 a NaN textured sample.
constexpr int FW_MAX_LEVELS = 16;  // mip levels of the textured forward warp (a side of 65536 gives 16)
struct FwArgs {
    const void* src;
    void* out;
    uint8_t* cov;  // may be null
    const float* flow;
    float2* pos;
    uint32_t* counts;
    uint32_t* fill;
    uint64_t* offsets;
    uint64_t* block_totals;
    uint64_t* total;
    uint32_t* rec_t;
    float4* rec_c;
    int* flags;
    int W, H, fw, fh, fmt;
    float rx, ry, sx, sy;
    uint64_t ntri;
    // the textured variant (forward.rs:229-320): the vertices carry (u, v, d) and every hit samples
    // the source's mip pyramid (forward/pyramid.rs) at them; mips: the per-vertex mip level d (null
    // when the pyramid has one level); lev[0] is the source, lev[1..nlev) its blurred levels
    int textured;
    int nlev;
    const float* mips;
    const void* lev[FW_MAX_LEVELS];
    int lw[FW_MAX_LEVELS], lh[FW_MAX_LEVELS];
};
size_t fw_scan_blocks(long n);
bool fw_flow_is_zero(const float* flow, long n);  // forward.rs:183-195 on a host flow
// the same test on a device flow: sets *flag (cleared by the caller) when some component moves
void launch_fw_flow_moves(hipStream_t s, const float* flow, long n, int* flag);
void fw_grid(FwArgs& A, float amount);             // mesh.rs:11-31
void launch_fw_count(hipStream_t s, const FwArgs& A);
void launch_fw_emit_fold(hipStream_t s, const FwArgs& A, uint64_t records);
// The textured variant: the mip pyramid's level count (pyramid.rs:20-35), the argument m of each
// vertex's mip level d = -(logf(m) / ln 2) (mesh.rs:54-88; the logarithm is the host C runtime's,
// as in the reference), and the pyramid levels 1.. of the source into lev (pyramid.rs:37-63).
int fw_level_count(int w, int h);
void launch_fw_mip_arg(hipStream_t s, const FwArgs& A, float* m);
void launch_fw_pyramid(hipStream_t s, FwArgs& A, float* scratch);

// Vector image coding (vectors.rs:57-130): one output image per job (blockIdx.z). The flow at
// the working size is sampled at every output pixel; `finish` 1 writes the wire frame with the
// rows turned back (Create Vectors runs on flipped frames), 0 the engine image as is.
struct VecJob {
    const float* flow;
    void* out;
    int fw, fh, W, H, fmt, finish;
    float sx, sy, kx, ky;
};
void vec_job_scales(VecJob& J, float scale);  // vectors.rs:70-76
void launch_vec_encode(hipStream_t s, const VecJob* d_jobs, const std::vector<VecJob>& jobs);
// vectors.rs:474-521 and the hole flags of the inverted alpha (vectors.rs:543-549, holefill.rs:
// 124-133); *gate is set when a hole exists
void launch_vec_decode(hipStream_t s, const void* vec, int W, int H, int fmt, float max_displace, float* flow,
                       uint8_t* holes, int* gate);
void launch_vec_scale(hipStream_t s, float* flow, long n, float sx, float sy);  // vectors.rs:566-571
// compose/vector.rs:370-397: T = (-1, -fx, -fy, -1); after the forward warp, the holes of the
// inverted coverage and the two flow channels of the pushed image
void launch_invert_prep(hipStream_t s, const float* flow, float* t, long n);
void launch_invert_post(hipStream_t s, const float* pushed, const uint8_t* cov, float* flow, uint8_t* holes, int* gate,
                        long n);

// Taps of the solver's Gaussians and of the pyramid filter (conv.rs:44-86).
Taps gaussian_taps(float r);
Taps binomial_taps();

inline size_t sample_size(int fmt) { return fmt == TWX_FMT_U8 ? 1 : (fmt == TWX_FMT_U16 ? 2 : 4); }
inline size_t image_bytes(long w, long h, int fmt) { return (size_t)w * h * 4 * sample_size(fmt); }

}  // namespace twx
