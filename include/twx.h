/*
 * twx.h -- C ABI of the MI355X (gfx950) frame-synthesis backend of the retime engine.
 *
 * The entry points mirror the engine's device backend trait
 * (plugins/retime-engine/src/device/exec/backend.rs:27-46, `trait Backend`) and the per-frame
 * render loop that drives it (crates/fx-cli/src/temporal.rs:183-250). Plain C types only: host
 * or device pointers, sizes, POD structs. No torch, no C++ types.
 *
 * Conventions shared by every entry point
 *  - Engine images are interleaved A,R,G,B, premultiplied, rows top first
 *    (plugins/retime-engine/src/image.rs:1-14); `fmt` is the format code of image.rs:42-48
 *    (0 = 8 bpc u8, 1 = 16 bpc u16 in 0..65535, 2 = 32 bpc f32).
 *  - Wire frames are what `vfx` reads and writes (crates/fx-core/src/frame.rs:79-141): packed
 *    RGBA u8, packed RGBA u16 little endian 0..65535, or planar float planes G,B,R,A.
 *  - Flows are interleaved (dx, dy) f32 at the solver's working size (flow.rs:37-45).
 *  - Every call blocks until its results are in place (the reference wrappers block,
 *    device/cl.rs:20-24). One in-order HIP stream per handle; the caller serialises calls on a
 *    handle. A nested call from a stage callback is refused with TWX_E_UNSUPPORTED
 *    (exec/backend.rs:151-159).
 *  - Ordering of DEVICE pointer arguments: the handle's stream is non-blocking. An entry point
 *    that takes device buffers first makes its stream wait (on the device) for all work queued on
 *    the legacy default stream before the call; work the caller queued on any other stream that
 *    writes the inputs or touches the outputs must be complete before the call. When the call
 *    returns, its outputs are complete.
 *  - A frame of 2^28 pixels or more, or with a side of 2^24 pixels or more, is refused with
 *    TWX_E_UNSUPPORTED (the kernels address frames by 32-bit byte offsets, 16 bytes a pixel at
 *    most, and block matching forms texel offsets with 24-bit multiplies).
 *  - Status codes map onto StageError (device/image.rs:14-23): any non-zero status makes the
 *    engine rerun the stage on the CPU. `err`/`errlen` receive a message (may be NULL/0).
 */
#ifndef TWX_H
#define TWX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TWX_ABI_VERSION 1

enum twx_status {
    TWX_OK = 0,
    TWX_E_DEVICE = 1,      /* StageError::Cl (driver / runtime failure) */
    TWX_E_NAN_OUTPUT = 2,  /* StageError::NanOutput (a NaN in a float flow) */
    TWX_E_UNSUPPORTED = 3, /* StageError::Unsupported (settings or sizes this backend refuses) */
    TWX_E_MISMATCH = 4     /* FxError::Mismatch (shapes or arguments that do not fit) */
};

enum twx_fmt { TWX_FMT_U8 = 0, TWX_FMT_U16 = 1, TWX_FMT_F32 = 2 };

/* FlowSettings (plugins/retime-engine/src/flow.rs:83-110). */
typedef struct twx_flow_settings {
    int32_t quality;     /* vectors quality 1..5 (4 = Best, half resolution) */
    int32_t hiq;         /* high quality render */
    float sensitivity;   /* slider / 100 (0.7 by default) */
    int32_t prep_contrast_edge, prep_delinearize, track_alpha, video_range_luma, skip_refine, wrap;
} twx_flow_settings;

/* One keyframe of a float parameter (crates/fx-core/src/curves.rs `Keyframe`): `time` in ticks. */
typedef struct twx_key {
    int64_t time;
    double value;
    int32_t hold;
    int32_t ease;
} twx_key;

/* The parameters of one `vfx apply retime` run that the frame-synthesis path reads
 * (plugins/retime/src/lib.rs:59-95 for ids and defaults; frontend.rs:184-243 for the derivation).
 * Parameters not listed keep their defaults (display Output, no 360, no prep, Speed mode, no
 * source offset, no blur compensation, no segment marks). */
typedef struct twx_retime_params {
    int32_t width, height, fmt;  /* frame size and wire depth code */
    int32_t time_step;           /* ticks per output frame (also the local time step) */
    int32_t time_scale;          /* ticks per second */
    double fps;                  /* composition rate (cx.fps as f64) */
    int32_t vectors;             /* "Motion Vectors" popup 1..5, default 5 (Best) */
    int32_t frame_interp;        /* "Frame Interp" popup 1..3, default 2 (Blend) */
    int32_t smart_blend;         /* checkbox, default 0 */
    int32_t quality_high;        /* 1 = --quality high (default), 0 = draft */
    double sensitivity;          /* "Motion Sensitivity" percent, default 70 */
    int32_t in_fps_is_out_fps;   /* checkbox, default 1 */
    double input_frame_rate;     /* slider, used when in_fps_is_out_fps == 0 */
    double speed;                /* "Speed %" constant, used when n_speed_keys == 0 */
    int32_t n_speed_keys;        /* keyframes of "Speed %", sorted by time, or 0 */
    const twx_key* speed_keys;
} twx_retime_params;

/* The parameters of one `vfx apply vectors-in` run (plugins/vectors-in/src/lib.rs:82-130 for ids
 * and defaults; glue.rs:190-253 for the derivation). Constant values; keyframes only for Speed %. */
typedef struct twx_vin_params {
    int32_t width, height, fmt;  /* frame size and wire depth code */
    int32_t time_step;           /* ticks per output frame (also the layers' frame duration) */
    double fps;                  /* composition rate */
    int32_t frame_interp;        /* "Frame Interp" 1..3, default 2 (Blend) */
    int32_t warping;             /* "Warping" 1 Inverse, 2 Forward (default) */
    int32_t display;             /* "Display" 1..6, default 1; 2 shows the retimed layer */
    int32_t mark_segments;       /* "Src: Mark Segments" 1..5, default 1 */
    int32_t time_remap_mode;     /* 1 Speed (default), 2 Frame Number */
    double speed;                /* "Speed %" constant, used when n_speed_keys == 0 */
    int32_t n_speed_keys;
    const twx_key* speed_keys;
    double frame_num;            /* "Frame Num" */
    int32_t in_fps_is_out_fps;   /* checkbox, default 1 (input 30 fps) */
    double input_frame_rate;     /* slider, used when in_fps_is_out_fps == 0 */
    double max_displace;         /* "Vectors: Max Displace", default 2048 */
    double vec_scale_x, vec_scale_y;
    double blur;                 /* "Motion Blur Compensation"; non-zero is refused */
    int32_t quality_high;
    /* The engine stage set of the render: 0 the effect's own (the Gouraud forward warp, hole-fill
     * taps above first, frames turned upside down: plugins/vectors-in/src/glue.rs:57-75, :253,
     * :398-410); 1 the engine's plain set on rows as given (the textured forward warp, taps below
     * first: stages.rs), the configuration the reference's recorded Vectors In grid outputs
     * (parity/vectors-in results/base.grid.tsv) were made with. */
    int32_t plain_stages;
} twx_vin_params;

/* The frame selection of one output frame (plugins/retime-engine/src/timeline.rs:359-380). */
typedef struct twx_selection {
    float s0, s1;
    int32_t a, b;
    float wa;
    int32_t exact;
    int32_t vectors_quality; /* popup - 1, after the searchable test and the segment override */
    int32_t interp;          /* popup - 1 */
} twx_selection;

typedef struct twx_handle twx_handle;

/* ---- device open / close / gate ---------------------------------------------------------- */

/* Opens device `index` (-1 = the first device) and runs the self-check gate
 * (device/exec/open.rs:45-78: self-check, then the program build). */
int twx_open(int index, twx_handle** out, char* err, size_t errlen);
void twx_close(twx_handle* h);
/* The adversarial arithmetic self-check (device/selfcheck.rs): float ops, conversions,
 * rounding, denormals and the order rules the kernels rely on, compared with host results. */
int twx_selfcheck(twx_handle* h, char* err, size_t errlen);
/* Bytes of device memory the handle may use (75 % of global memory, docs/engine.md:124). */
uint64_t twx_memory_budget(twx_handle* h);

/* ---- Backend::solve / solve_pair (exec/backend.rs:27-46, flow.rs:1328-1353) --------------- */

/* Working geometry of a solve for a frame of width x height (flow.rs:118-137, prep.rs:63-146).
 * Returns TWX_E_UNSUPPORTED for working sizes below 2x2. */
int twx_flow_geometry(const twx_flow_settings* fs, int32_t width, int32_t height,
                      int32_t* flow_w, int32_t* flow_h, int32_t* levels);

/* flow::solve(X, Y): X(p) ~ Y(p + out(p)). Host engine images in, host flow out
 * (flow_w * flow_h * 2 floats, from twx_flow_geometry). */
int twx_solve(twx_handle* h, const void* x, const void* y, int32_t width, int32_t height,
              int32_t fmt, const twx_flow_settings* fs, float* flow_out, char* err, size_t errlen);

/* flow::solve_pair(A, B): F_AB = solve(A, B), F_BA = solve(B, A), pyramids built once. */
int twx_solve_pair(twx_handle* h, const void* a, const void* b, int32_t width, int32_t height,
                   int32_t fmt, const twx_flow_settings* fs, float* f_ab_out, float* f_ba_out,
                   char* err, size_t errlen);

/* Many solve_pair calls (Backend::solve_pair, device/exec/backend.rs:191-208) in one batched
 * launch sequence, for a caller that wants flows rather than frames (the batched solve of a
 * clip's pairs): pair k writes F_AB = solve(A, B) into f_ab and F_BA = solve(B, A) into f_ba;
 * either may be NULL to skip that direction. `a` and `b` are DEVICE engine images (rows as given,
 * like twx_solve's host images), all of one size and format; f_ab and f_ba are DEVICE buffers of
 * flow_w * flow_h * 2 floats (twx_flow_geometry), distinct from each other and from the images.
 * Each distinct image pointer has its pyramid built once per call. The flows are those of n
 * twx_solve_pair calls, byte for byte; a NaN in any flow gives TWX_E_NAN_OUTPUT (the outputs
 * are then unspecified). */
typedef struct {
    const void* a;
    const void* b;
    float* f_ab;
    float* f_ba;
} twx_pair;
int twx_solve_pairs_batched(twx_handle* h, int32_t n, const twx_pair* pairs, int32_t width,
                            int32_t height, int32_t fmt, const twx_flow_settings* fs, char* err,
                            size_t errlen);

/* ---- Backend::compose_vectors (compose/vector.rs:209-284) -------------------------------- */

/* The composition of one output frame from its two frames and their flows, mode 1, no 360.
 * `warp_mode` is EngineSettings::warp_mode: 0 the inverse warps with hold edges, 1 the inverse
 * warps without hold edges and with coverage, the Smart Blend composite and the push-pull hole
 * fill of the uncovered pixels (smartblend.rs:323-419), 3 the forward warps (each frame pushed
 * with its own flow, holes transparent), 2 the forward warps with coverage, the Smart Blend
 * composite and the hole fill (compose/vector.rs:117-147, forward.rs:391-572).
 * `interp` 0 Nearest, 1 Blend, 2 Weighted Blend (each warped frame through the sharpening step
 * against its unwarped frame, filters.rs:649-656, before the blend or composite); wa is the
 * weight of A (Selection::wa). Nearest reads only the flow nearest_uses_f_ab names
 * (compose/vector.rs:61-64; the other may be NULL). `taps_above_first` is the vertical tap order
 * of the Weighted Blend blur and of the hole fill (the stage set's vertical_taps,
 * stages.rs:60-62): 1 for the retime and vectors-in front-ends (frontend.rs:238-241,
 * vectors-in/src/glue.rs:242), 0 for the plain CPU stages. `forward_gouraud` is the forward warp
 * variant of the stage set: 1 the Gouraud variant (stages::GouraudForward, stages.rs:101-130,
 * which vectors-in and retime-fgbg install), 0 the textured variant (the engine's plain stage
 * set: pyramid-sampled vertices, forward.rs:229-320). Host engine images. */
int twx_compose_vectors(twx_handle* h, int32_t interp, float wa, const void* a, const void* b,
                        int32_t width, int32_t height, int32_t fmt, const float* f_ab,
                        const float* f_ba, int32_t flow_w, int32_t flow_h, int32_t warp_mode,
                        int32_t taps_above_first, int32_t forward_gouraud, void* out, char* err,
                        size_t errlen);

/* ---- Backend::forward_warp_with (device/exec/backend.rs:34-36, forward.rs:229-373) ------------ */

/* The forward warp of `src` by `flow` times `amount`: the source pixels as the vertices of a
 * displaced mesh, its triangles scan converted in cell order, the running mean of the colours per
 * covered pixel in triangle order. `pyramid` is ForwardOpts::pyramid: 0 the Gouraud variant
 * (vertex colours), 1 the textured variant (each vertex carries its texture coordinates and mip
 * level, mesh.rs:54-88, and each covered pixel samples the source's mip pyramid there,
 * pyramid.rs). `coverage` (width * height bytes: 255 covered, 0 hole) may be NULL. The early outs
 * copy `src` with full coverage. TWX_E_UNSUPPORTED as the CPU path for the degenerate last row
 * pass that would write outside the image, and for what this device does not serve: a triangle
 * over 65536 covered pixels, NaN colours (a NaN vertex colour of the Gouraud variant, a NaN sample
 * of the textured one), a textured tap below the image. Host buffers. */
int twx_forward_warp(twx_handle* h, const void* src, int32_t width, int32_t height, int32_t fmt,
                     const float* flow, int32_t flow_w, int32_t flow_h, float amount,
                     int32_t pyramid, void* out, uint8_t* coverage, char* err, size_t errlen);

/* ---- vector image coding (vectors.rs) and the Vectors In flow steps (compose/vector.rs) ---- */

/* encode_vectors (vectors.rs:57-130): `flow` (flow_w x flow_h, any size of at least 2x2) sampled
 * at every pixel of a width x height engine image of `fmt`, scaled to pixels of that size over
 * `scale` (Max Displace): A white, R and G the centred components, B 0. Host buffers. No
 * counterpart in the reference's device backend (the reference encodes on the CPU). */
int twx_encode_vectors(twx_handle* h, const float* flow, int32_t flow_w, int32_t flow_h,
                       int32_t width, int32_t height, int32_t fmt, float scale, void* out,
                       char* err, size_t errlen);

/* decode_vectors (vectors.rs:474-521): the flow (width x height x 2) of an encoded engine image
 * with `max_displace`; alpha 0 is the zero vector. `holes_out` (may be NULL) receives the hole
 * flags of the inverted alpha (vectors.rs:543-549 as the hole fill reads them,
 * holefill.rs:124-133). Host buffers. */
int twx_decode_vectors(twx_handle* h, const void* vec, int32_t width, int32_t height, int32_t fmt,
                       float max_displace, float* flow_out, uint8_t* holes_out, char* err,
                       size_t errlen);

/* flow_from_vector_image_with (compose/vector.rs:335-345): decode, the push-pull fill of the
 * holes (holefill.rs:506-521, the vertical tap order `taps_above_first`), Vec Scale
 * (vectors.rs:566-571). Host buffers. */
int twx_flow_from_vectors(twx_handle* h, const void* vec, int32_t width, int32_t height,
                          int32_t fmt, float max_displace, float scale_x, float scale_y,
                          int32_t taps_above_first, float* flow_out, char* err, size_t errlen);

/* invert_flow_on (compose/vector.rs:370-397): T = (-1, -fx, -fy, -1) pushed along the flow by
 * the forward warp with coverage (amount 1), the uncovered pixels filled. `forward_gouraud` is
 * the stage set's forward variant (1 for Vectors In; 0 the textured one of the plain stage set).
 * Host buffers. */
int twx_invert_flow(twx_handle* h, const float* flow, int32_t width, int32_t height,
                    int32_t forward_gouraud, int32_t taps_above_first, float* out, char* err,
                    size_t errlen);

/* The Create Vectors pass over a clip (create_vectors.rs:162-179 with the row flips of
 * create-vectors/src/glue.rs:517-566): output k is solve(X = frame frame_x[k], Y = frame
 * frame_y[k]) on the flipped engine images, encoded at the frame size with `max_displace`,
 * flipped back and written as a wire frame (image.rs:205-245). Identical (X, Y) pairs are solved
 * once. `clip` (n_in wire frames) and `out` (count wire frames) are DEVICE pointers; frame_x and
 * frame_y are host arrays of count indices into the clip. */
int twx_create_vectors_dev(twx_handle* h, const twx_flow_settings* fs, int32_t width,
                           int32_t height, int32_t fmt, const void* clip, int32_t n_in,
                           int32_t count, const int32_t* frame_x, const int32_t* frame_y,
                           float max_displace, void* out, char* err, size_t errlen);

/* The per-frame render of `vfx apply vectors-in` (plugins/vectors-in/src/glue.rs:380-495): output
 * frames [first, first + count) of a run over the footage `input` (n_input wire frames), the optional
 * Color Source layer `source` and the vector layers `prev` (Vectors: Prev) and `next` (Vectors:
 * Next); a NULL layer is a layer that is not set. Every frame is read upside down and the result
 * turned back (glue.rs:398-410); the flows come from the vector images (decode, push-pull fill,
 * Vec Scale; vectorin.rs:200-225), a single layer's flow is inverted through the Gouraud forward
 * warp where a composition reads the missing one; the compositions are the engine's with the
 * effect's stages (the Gouraud forward warp, taps above first, glue.rs:253). Without a vector layer,
 * or with a rate below 0.01, the frames are the size-mismatch error fill. The blur compensation
 * tail and the FG matte layers are refused (TWX_E_UNSUPPORTED). DEVICE pointers. */
int twx_vectors_in_frames_dev(twx_handle* h, const twx_vin_params* p, const void* input, int32_t n_input,
                              const void* source, int32_t n_source, const void* prev, int32_t n_prev,
                              const void* next, int32_t n_next, int32_t first, int32_t count, void* out,
                              char* err, size_t errlen);
/* The same with host pointers (uploads the clips, downloads the frames). */
int twx_vectors_in_frames(twx_handle* h, const twx_vin_params* p, const void* input, int32_t n_input,
                          const void* source, int32_t n_source, const void* prev, int32_t n_prev,
                          const void* next, int32_t n_next, int32_t first, int32_t count, void* out,
                          char* err, size_t errlen);

/* ---- the per-frame render loop of `vfx apply retime` ------------------------------------- */

/* Frame selection of output frame `n` (compose.rs:76-99 and timeline.rs:447-464). */
int twx_select(const twx_retime_params* p, int32_t n, twx_selection* out, char* err, size_t errlen);

/* Renders output frames [first, first + count) of a retime run over a clip of `n_in` wire
 * frames, as `vfx apply retime` writes them (temporal.rs:183-250: select, fetch and
 * premultiply, rows reversed, solve_pair, compose, finish). `clip` and `out` are DEVICE
 * pointers (the clip resident in HBM, count wire frames out). Identical frame pairs are solved
 * once. */
int twx_retime_frames_dev(twx_handle* h, const twx_retime_params* p, const void* clip,
                          int32_t n_in, int32_t first, int32_t count, void* out,
                          char* err, size_t errlen);

/* Many clips, of any sizes, depths and settings, in one batched render: clip i renders its
 * output frames [first[i], first[i] + count[i]) from clips[i] (n_in[i] wire frames) into outs[i].
 * DEVICE pointers. Every frame pair of every clip goes through the same ingest, pyramid, solve and
 * compose launches (the frame tables sorted by size, one launch per size class); the bytes of each
 * clip equal a twx_retime_frames_dev call on it alone. No counterpart in the reference, whose
 * render loop takes one output frame at a time (temporal.rs:183-250). */
int twx_retime_batch(twx_handle* h, int32_t n_clips, const twx_retime_params* params,
                     const void* const* clips, const int32_t* n_in, const int32_t* first,
                     const int32_t* count, void* const* outs, char* err, size_t errlen);

/* The same with host pointers (uploads the clip, downloads the frames). */
int twx_retime_frames(twx_handle* h, const twx_retime_params* p, const void* clip, int32_t n_in,
                      int32_t first, int32_t count, void* out, char* err, size_t errlen);

/* ---- the differentiable warp+blend (no counterpart in the reference) ------------------------ */

/* out = blend(W(A, F_BA, 1 - wa), W(B, F_AB, wa), 1 - wa) on float images of `channels` (1..4)
 * interleaved channels, width x height, rows top first; W is the inverse warp with hold edges
 * of warp.rs:477-512 (flows interleaved (dx, dy) at flow_w x flow_h), the blend the float blend
 * of blend.rs:55-114 without its snap branches, ((tB - tA) * w + tA) + 0. DEVICE pointers. */
int twx_warp_blend(twx_handle* h, int32_t channels, int32_t width, int32_t height, const float* a,
                   const float* b, const float* f_ab, const float* f_ba, int32_t flow_w,
                   int32_t flow_h, float wa, float* out, char* err, size_t errlen);

/* The backward pass of twx_warp_blend for the upstream gradient grad_out (same shape as out):
 * dL/dA, dL/dB (images) and dL/dF_AB, dL/dF_BA (flows), each written when its pointer is not
 * NULL. Sample points are recomputed, not stored. The flow gradients are summed per flow texel
 * in a fixed order; the image gradients are scattered to the sampled texels and summed in 64-bit
 * fixed point (each term rounded at 2^-34 of the largest |grad_out|, power-of-two scaled), so
 * every gradient is the same bits on every run. An upstream gradient holding an infinity or a NaN
 * is refused with TWX_E_UNSUPPORTED when an image gradient is asked for. DEVICE pointers. */
int twx_warp_blend_backward(twx_handle* h, int32_t channels, int32_t width, int32_t height,
                            const float* a, const float* b, const float* f_ab, const float* f_ba,
                            int32_t flow_w, int32_t flow_h, float wa, const float* grad_out,
                            float* grad_a, float* grad_b, float* grad_f_ab, float* grad_f_ba,
                            char* err, size_t errlen);

/* twx_warp_blend_backward without any device or pinned-host allocation. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * refinement loops, harnesses that forbid hipMalloc). The same arguments, kernels and results as
 * twx_warp_blend_backward; its scratch (the per-pixel flow-value gradients, 16 bytes a pixel, and
 * a flag) is a fixed array of the library that each device holds from twx_open until the process
 * ends, so width * height must be at most TWX_NOALLOC_MAX_PIXELS (TWX_E_UNSUPPORTED above). Every
 * call on one device shares that array: calls that overlap on one device (two handles, or two
 * threads) race. Its launches are not timed by twx_profile_*. DEVICE pointers. */
#define TWX_NOALLOC_MAX_PIXELS 2097152u /* 2^21: 1920 x 1080 fits; 32 MiB of scratch a device */
int twx_warp_blend_backward_noalloc(twx_handle* h, int32_t channels, int32_t width, int32_t height,
                                    const float* a, const float* b, const float* f_ab,
                                    const float* f_ba, int32_t flow_w, int32_t flow_h, float wa,
                                    const float* grad_out, float* grad_a, float* grad_b,
                                    float* grad_f_ab, float* grad_f_ba, char* err, size_t errlen);

/* Photometric refinement of a flow pair on the device, in one call: `steps` Adam steps (betas 0.9
 * and 0.999, eps 1e-8, torch.optim.Adam's default step order) of F_AB and F_BA, updated in place,
 * on the mean squared error of twx_warp_blend(a, b, F_AB, F_BA, wa) against `truth` (H x W x C).
 * Each step is the fused forward, loss and flow-value gradient (2 (out - truth) / N through both
 * warps), the fixed-order flow gathers and the Adam update; nothing returns to the host between
 * steps, and every sum has a fixed order, so the result is the same bits on every run.
 * losses (HOST, steps + 1 floats) receives the loss before each step and after the last. No
 * counterpart in the reference (the differentiable extension of the warp+blend). DEVICE
 * pointers for the images and flows. */
int twx_refine_flows(twx_handle* h, int32_t channels, int32_t width, int32_t height, const float* a,
                     const float* b, const float* truth, float* f_ab, float* f_ba, int32_t flow_w,
                     int32_t flow_h, float wa, int32_t steps, float lr, float* losses, char* err,
                     size_t errlen);

/* ---- counters (Exec::summary, device/exec/mod.rs:292-322) -------------------------------- */

typedef struct twx_stats {
    uint64_t solve_calls, pair_solves, pairs_reused, composes, fast_frames;
    uint64_t pyramid_builds, pyramid_hits;
    double device_ms;
    uint64_t graph_replays;  /* batched renders replayed from a captured graph */
} twx_stats;
int twx_get_stats(twx_handle* h, twx_stats* out);

/* Backend::cache_stats (device/exec/backend.rs:42-45): the counters of the resident pyramid cache
 * behind twx_solve and twx_solve_pair (hits, misses, evictions, device bytes held; any pointer may
 * be NULL). A pyramid is found again only for an image with the same size, format and sample bits
 * and the same working geometry (device/exec/cache.rs:1-27); least recently used entries go first,
 * within a quarter of twx_memory_budget. pyramid_builds / pyramid_hits of twx_stats count the
 * same misses and hits. */
int twx_cache_stats(twx_handle* h, uint64_t* hits, uint64_t* misses, uint64_t* evictions,
                    uint64_t* bytes_held);

/* Graph capture of batched renders (on by default; no counterpart in the reference, whose render
 * loop launches per frame): a twx_retime_frames_dev or twx_retime_batch call runs on memory of its
 * own and captures its launches (ingest, pyramids, every solve, compose, finish) into one hipGraph;
 * a later call with the same parameters, pointers and frame ranges replays it in one launch. The
 * captured plan holds its own device and pinned memory until another plan replaces it or graphs are
 * turned off. With profiling on, calls take the plain path (per-launch events). */
int twx_set_graphs(twx_handle* h, int32_t on);

/* Tile shapes of the fused gradient/tensor, calc_b and new_disp kernels of the flow solve (no
 * counterpart in the reference, whose passes run per row on the CPU): 64 x 32, 64 x 16 or 64 x 64
 * outputs per workgroup, with 4 or 1 producer pixels in flight per lane (variants 0: 32/4, 1: 32/1,
 * 2: 16/4, 3: 16/1, 4: 64/4, 5: 64/1). By default the first plain launch of each (kernel, level
 * area class, job count class) times the six (twice each) on the live data and keeps the fastest for the
 * process; every variant gives the same bits. twx_set_tile_shape forces one variant for every
 * launch (-1: measured choice again; other values TWX_E_MISMATCH). twx_tile_choices writes up to
 * cap/4 quadruples (kernel 0 tensor, 1 calc_b, 2 new_disp; ceil log2 of level area; ceil log2 of
 * job count; variant) and returns how many choices exist. Process-wide. */
int twx_set_tile_shape(int32_t variant);
int twx_tile_choices(int32_t* out, int32_t cap);

/* The coarse solver (no counterpart in the reference, whose solve runs level by level on the CPU):
 * the levels of a solve from its start level down while a level holds at most 2560 pixels run in
 * one persistent kernel per launch, whose workgroups take the jobs from a device work queue and
 * keep each job's planes in LDS, instead of one launch per step and level. On by default;
 * twx_set_coarse_solver(0) runs every level through the per-level kernels, with the same bits.
 * Process-wide; applies to solves enqueued afterwards (a captured graph replays the launches it
 * captured). */
int twx_set_coarse_solver(int32_t on);

/* Device pointer of the handle's HIP stream (hipStream_t), so a caller can order its own work. */
void* twx_stream(twx_handle* h);

/* Per-kernel-family timing (the reference's VFX_DEVICE_REPORT counters, device/exec/mod.rs:
 * 292-322, extended with roofline data): when enabled, every launch is bracketed by HIP events
 * on the handle's stream and its algorithmic bytes (each step's distinct inputs read once and
 * outputs written once) are accumulated per family. Families are numbered 0 .. n-1; reading
 * family n returns TWX_E_MISMATCH. */
int twx_profile_enable(twx_handle* h, int on);
int twx_profile_reset(twx_handle* h);
int twx_profile_read(twx_handle* h, int32_t family, const char** name, double* ms, double* bytes,
                     uint64_t* launches);

#ifdef __cplusplus
}
#endif
#endif /* TWX_H */
