/*
 * bt_abi.h -- C ABI of blacktracer_amd, the MI355X (gfx950) implementation
 * of BlackTracer's data-parallel core: the Kerr null-geodesic grid build
 * (Path A) and the per-pixel render pass (Path B).
 *
 * Every entry point returns an int status (BT_OK == 0), never prints,
 * never exits and never resets the device.  Host buffers are owned by the
 * caller and borrowed for the duration of the call; device memory is owned
 * by a bt_context (one per device; calls on one context are serialised by
 * the context's mutex, calls on different contexts run concurrently).
 * Spin lives in the pose, not in process globals.
 *
 * Reference interfaces replaced (annemiekie/BlackTracer,
 * BlackTracer2/BlackTracer/):
 *   bt_camera_init      Camera::Camera(...) + BlackHole::setA   Camera.h:27-61, BlackHole.h:12-15
 *   bt_trace_rays       Grid::integration_wrapper               Grid.h:514-556
 *   bt_grid_build       Grid::Grid(maxLevel, startLevel, ...)   Grid.h:605-624
 *   bt_grid_export_psh  Grid::saveAsGpuHash / PSHOffsetTable    Grid.h:642-654, PSHOffsetTable.h:350-398
 *   bt_make_image       makeImage(...)                          MakeImageKernel.cu:1206-1236, kernel.cuh:28-33
 * Entry points with no counterpart in the reference (declared with their
 * reason): bt_trace_rays_device, bt_grid_build_batch, the resident render
 * objects (bt_scene, bt_view, bt_render_frame) and the server (bt_server_*).
 */
#ifndef BT_ABI_H
#define BT_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BT_ABI_VERSION 2

/* ---- status codes ---- */
#define BT_OK              0
#define BT_ERR_INVALID    -1  /* bad argument (null pointer, size, level range) */
#define BT_ERR_DEVICE     -2  /* no HIP device / device index out of range */
#define BT_ERR_HIP        -3  /* a HIP runtime call failed */
#define BT_ERR_NOMEM      -4  /* device or host allocation failed */
#define BT_ERR_BUDGET     -5  /* request does not fit the HBM admission budget */
#define BT_ERR_NOTFOUND   -6  /* handle or cache entry not found */
#define BT_ERR_STATE      -7  /* call not valid in the object's state */

/* Message of the last failing call on this thread ("" if none). */
const char* bt_last_error(void);
int         bt_abi_version(void);

/* ---- context (device memory pools, stream, timers) ---- */
typedef struct bt_context bt_context;
int  bt_context_create(int device, bt_context** out);
void bt_context_destroy(bt_context* ctx);
/* Blocks until all work queued on the context's stream is done. */
int  bt_context_sync(bt_context* ctx);
/* Per-kernel timing with HIP events recorded on the context's stream.
 * enable != 0 switches it on; bt_timer_get returns launches and total
 * milliseconds for a kernel name (e.g. "rk_f64"). */
int  bt_timing_enable(bt_context* ctx, int enable);
int  bt_timer_get(bt_context* ctx, const char* kernel, int64_t* launches, double* total_ms);
int  bt_timer_reset(bt_context* ctx);
/* Device buffers owned by the caller (frame outputs of bt_render_frame,
 * inputs of bt_trace_rays_device); freed with bt_device_free. */
int  bt_device_alloc(bt_context* ctx, size_t bytes, void** out);
void bt_device_free(bt_context* ctx, void* ptr);
int  bt_memcpy(bt_context* ctx, void* dst, const void* src, size_t bytes, int to_device);

/* ---- pose: Camera(theta, phi, r, br, btheta, bphi) + BlackHole(a) ---- */
typedef struct {
    double a;                  /* spin a/M (BlackHole.h:12) */
    double theta, phi, r;      /* camera Boyer-Lindquist position */
    double br, btheta, bphi;   /* direction of motion in the FIDO frame */
    double speed;              /* used only when user_speed != 0 */
    int    user_speed;         /* 0: speed = calcSpeed(r, theta) (circular orbit) */
} bt_pose;

/* Derived camera quantities, Camera::initforms (Camera.h:54-61):
 * out6 = {speed, alpha, w, wbar, Delta, ro}.  cam7 (nullable) receives
 * getParamArray() (Camera.h:67-78): {speed, alpha, w, wbar, br, btheta, bphi}. */
int bt_camera_init(const bt_pose* pose, double* out6, float* cam7);

/* ---- integrator parameters ---- */
#define BT_PREC_FP64  0   /* plain FP64 path: the drop-in */
#define BT_PREC_MIXED 1   /* FP32 Cash-Karp bulk, FP64 set-up/fallback/accumulation */

typedef struct {
    double eps;               /* Metric.h:312 (1e-5) */
    double h1;                /* Metric.h:313 (0.01) */
    double z_end;             /* Metric.h:311 (-1e7) */
    int    max_steps;         /* Metric.h:31 (1000) */
    int    apriori_capture;   /* 1: checkCelest (Metric.h:169) decides capture */
    double r_escape;          /* >0: stop as escaped once r > r_escape */
    double r_capture_factor;  /* >0: stop as captured once r < f * r_horizon */
    int    newton_cap;        /* cap on checkB_Q's Newton loop (reference: none) */
    int    precision;         /* BT_PREC_FP64 / BT_PREC_MIXED */
    double crit_eps;          /* mixed: FP64 for rays whose (b, q) lie within crit_eps of the
                                 critical curve (distance in the (b, q) plane; 0.03) */
    double hmin_fp64;         /* mixed: switch a ray to FP64 once |h| < hmin_fp64 */
    double noise_budget;      /* mixed: the exact FP64 path for a ray once the summed rounding-noise
                                 bound of its FP32 error estimates exceeds this (0.13; 0: off) */
    double decision_margin;   /* mixed: the exact FP64 path for a ray whose FP32 accept/reject
                                 decision lies within this many noise bounds of the threshold
                                 (4; 0: off) */
    double drift_budget;      /* mixed: the exact FP64 path for a ray once the FP32 rounding its
                                 accepted theta and phi increments carry sums past this many
                                 radians (1.2e-6; 0: off) */
    double polar_b;           /* mixed: the exact FP64 path for rays with |b| below this: they pass
                                 within about |b| / sqrt(q) of a pole, where the FP32 error
                                 estimate turns hypersensitive to the step size (1e-3; 0: off) */
} bt_integ_params;

void bt_integ_params_default(bt_integ_params* p);

/* Per-call work counters (algorithmic-work tally and fallback fraction). */
typedef struct {
    int64_t rays;          /* rays handled */
    int64_t integrated;    /* rays that entered the Cash-Karp loop */
    int64_t attempts;      /* Cash-Karp attempts, all precisions */
    int64_t attempts_fp64; /* attempts done in FP64 */
    int64_t newton;        /* checkB_Q Newton iterations */
    int64_t golden;        /* findMinGoldSec iterations */
    int64_t fallback_rays; /* mixed: rays finished in FP64 */
} bt_trace_stats;

/* Grid::integration_wrapper (Grid.h:514-556): camera-sky (theta, phi) in,
 * celestial (theta', phi') out in place, (-1,-1) for captured rays, plus
 * the step index of the last Cash-Karp attempt.  Host buffers of n. */
int bt_trace_rays(bt_context* ctx, const bt_pose* pose, const bt_integ_params* prm,
                  int64_t n, double* theta, double* phi, int32_t* step,
                  bt_trace_stats* stats);
/* Same on device buffers, enqueued on the context's stream (asynchronous;
 * stats, when non-null, forces a synchronisation). */
int bt_trace_rays_device(bt_context* ctx, const bt_pose* pose, const bt_integ_params* prm,
                         int64_t n, double* d_theta, double* d_phi, int32_t* d_step,
                         bt_trace_stats* stats);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): bt_trace_rays_device on the FP64 path (prm null or
 * precision BT_PREC_FP64; BT_ERR_INVALID otherwise) whose work queue lives in
 * a fixed device array of the library (about 36 MiB on each GPU it runs
 * on, held until the process ends) instead of the context's buffers, so the
 * call allocates nothing.  n <= BT_NOALLOC_MAX_RAYS, else BT_ERR_INVALID
 * and nothing runs.  No statistics.  Calls that overlap on one GPU (through
 * different contexts) share the array and race; calls through one context
 * are serialised by it. */
#define BT_NOALLOC_MAX_RAYS (1 << 20)
int bt_trace_rays_device_noalloc(bt_context* ctx, const bt_pose* pose, const bt_integ_params* prm,
                                 int64_t n, double* d_theta, double* d_phi, int32_t* d_step);

/* Self-check (no reference counterpart): the FP64 integrator divides by a
 * shared reciprocal where several quotients have one denominator (the
 * division sequence's own steps, the reciprocal refined once).  This runs
 * that form against the device's IEEE division on n generated operand pairs
 * (exponents within 2^+-200, with all-ones significands, zero numerators
 * and power-of-two denominators among them) and stores in *mismatches the
 * pairs whose quotients differ in any bit.  n at most 2^30. */
int bt_selfcheck_division(bt_context* ctx, int64_t n, uint64_t seed, int64_t* mismatches);

/* ---- launch-shape tuning of the persistent integrators (no reference
 * counterpart: the reference launches one thread per ray with a fixed block,
 * integration_wrapper Grid.h:514-556) ---- */
typedef struct {
    int chunk;             /* queue positions a wave reserves per atomic (>= 64) */
    int block;             /* threads per workgroup: 64, 128 or 256 */
    int waves;             /* resident waves per SIMD, 0 = as many as registers allow */
    double rays_per_lane;  /* bt_autotune only: calibration rays per lane of the widest launch */
} bt_tune;

int bt_tune_get(bt_context* ctx, bt_tune* out);
/* BT_ERR_INVALID for a block other than 64/128/256, chunk outside
 * [64, 2^20] or waves outside 0..8.  Results never depend on the shape. */
int bt_tune_set(bt_context* ctx, const bt_tune* t);
/* Times about n_rays (<= 0: 2^20; at most 2^24) directions of the pose's
 * equal-angle camera sky (the lattice grid builds trace) under prm for
 * candidate shapes (coordinate descent over chunk {64..1024}, block
 * {64,128,256}, waves {0..4}) and keeps the fastest in the context; out and
 * best_ms (nullable) receive it and its time.  The best shape depends on
 * the batch size: calibrate at the size that will be traced. */
int bt_autotune(bt_context* ctx, const bt_pose* pose, const bt_integ_params* prm, int64_t n_rays,
                bt_tune* out, double* best_ms);

/* ---- adaptive grid: Grid(maxLevel, startLevel, angle, cam, bh) ---- */
typedef struct bt_grid bt_grid;

typedef struct {
    double refine_thresh;   /* PRECCELEST, Grid.h:24 (0.015) */
    int    fix_tvertices;   /* 1: Grid.h:620-622, pinned order level asc, key asc */
} bt_grid_params;

void bt_grid_params_default(bt_grid_params* p);

int  bt_grid_build(bt_context* ctx, const bt_pose* pose, const bt_integ_params* prm,
                   int max_level, int start_level, int angle, const bt_grid_params* gp,
                   bt_grid** out);
/* Several grids at once (no counterpart in the reference, which builds one
 * Grid at a time, Grid.h:605): n poses with their own max/start levels and
 * shared integration parameters.  The grids advance level by level in
 * lockstep and the rays of each level of every grid are traced together in
 * one set of persistent launches over a shared device work queue.  out
 * receives n grids (each freed with bt_grid_free); stats (nullable) the
 * batch's counters; bt_grid_stats of a batched grid holds only its ray
 * count.  1 <= n <= 128. */
int  bt_grid_build_batch(bt_context* ctx, int n, const bt_pose* poses, const bt_integ_params* prm,
                         const int* max_levels, const int* start_levels, int angle,
                         const bt_grid_params* gp, bt_grid** out, bt_trace_stats* stats);
int  bt_grid_info(const bt_grid* g, int* N, int* M, int* max_level, int* start_level,
                  int* equafactor, int64_t* rays_traced);
int  bt_grid_stats(const bt_grid* g, bt_trace_stats* stats);
/* Host copies of the dense grid, row-major (i*M + j):
 * cel N*M*2 doubles ((-2,-2) = not traced, (-1,-1) = captured),
 * steps N*M int32, levels N*M int8 (-1 = not a block key). Any may be null. */
int  bt_grid_export(const bt_grid* g, double* cel, int32_t* steps, int8_t* levels);
/* Perfect spatial hash of the traced points (PSHOffsetTable layout):
 * call with null buffers to get the widths, then with buffers of
 * 2*hw*hw floats / 2*ow*ow ints / 2*hw*hw ints. */
int  bt_grid_export_psh(const bt_grid* g, int* hash_width, int* offset_width,
                        float* hash_table, int32_t* offset_table, int32_t* hash_pos_tag);
void bt_grid_free(bt_grid* g);

/* ---- render: makeImage ---- */
typedef struct {
    int   star;        /* star pass (MakeImageKernel.cu:955) */
    int   map;         /* sky-map pass (:956) */
    int   lr;          /* lensing + redshift (:957) */
    float hor, ver;    /* view offsets (:959-960) */
    float offset;      /* celestial phi offset of the frame (:1413) */
} bt_render_flags;

void bt_render_flags_default(bt_render_flags* f);

/* makeImage (MakeImageKernel.cu:1206) with cv::Mat csImage replaced by a
 * BGR8 buffer and the PNG written by the reference replaced by out_bgra
 * (N*M*4 bytes, B,G,R,255 per pixel).  Arrays of int2/float2 are passed as
 * interleaved pairs.  G must be 1: the reference renders its G > 1 frames
 * only in loops that do not run from makeImage (movielength = 0,
 * MakeImageKernel.cu:1410); bt_render_frame_blend renders them. */
int bt_make_image(bt_context* ctx, const float* stars, const int32_t* star_tree,
                  int star_size, const float* cam_param, const float* magnitude,
                  int tree_level, int M, int N, int step, const uint8_t* sky_bgr,
                  int sky_rows, int sky_cols, int G, float grid_start, float grid_step,
                  const float* hit, const float* viewthing, float view_angle, int GM,
                  int GN, int gridlvl, const int32_t* offset_tables,
                  const float* hash_tables, const int32_t* hash_pos_tag,
                  const int32_t* table_sizes, int otsize, int htsize,
                  const bt_render_flags* flags, uint8_t* out_bgra);

/* ---- render inputs computed on the host (as the reference does) ---- */

/* Viewer(viewangle, offUp, offRight, w, h, sphere).viewMatrix
 * (Viewer.h:20-133): (height+1)*(width+1) float2 pixel-corner directions. */
int bt_viewer_matrix(double view_angle, double off_up, double off_right, int width, int height,
                     int sphere, float* out);
/* StarProcessor::makeTree (StarProcessor.h:172-212): stars4 is
 * star_count x (theta, phi, magnitude, B-V); tree has (1 << (L+1)) - 1
 * ints; star_pos / star_mag 2*star_count floats in leaf order. */
int bt_star_tree(const float* stars4, int star_count, int tree_level, int32_t* tree,
                 float* star_pos, float* star_mag);
/* StarProcessor(...)'s split of a catalogue at the magnitude cut
 * (StarProcessor.h:36-52) with addLowLightStarsToImage
 * (StarProcessor.h:108-158): every star of stars4 fainter than
 * magnitude_cut is painted into the BGR8 sky image (rows x cols, in place)
 * as a 5 x 5 blackbody splat; the brighter ones belong in the star tree
 * (bt_star_catalog_parse with the same cut).  Call before building the
 * scene from the sky. */
int bt_sky_add_low_stars(uint8_t* sky_bgr, int rows, int cols, const float* stars4, int star_count,
                         float magnitude_cut);

/* ---- device-resident render objects (no counterpart in the reference,
 * which re-uploads everything per makeImage call; used by the serving
 * loop so that a frame touches only HBM-resident data) ---- */
typedef struct bt_scene bt_scene;   /* star tree + summed-area sky table */
typedef struct bt_view bt_view;     /* pixel-corner directions of one resolution */

int  bt_scene_create(bt_context* ctx, const float* stars4, int star_count, int tree_level,
                     const uint8_t* sky_bgr, int sky_rows, int sky_cols, bt_scene** out);
/* Same from a star tree already built (the arrays of a .star file,
 * StarProcessor.h:19-30): tree 2^(L+1)-1 ints, pos and mag 2 floats per star. */
int  bt_scene_create_tree(bt_context* ctx, const int32_t* tree, int tree_level, const float* pos,
                          const float* mag, int star_count, const uint8_t* sky_bgr, int sky_rows,
                          int sky_cols, bt_scene** out);
/* Host copy of the summed-area table, sky_rows*sky_cols*4 floats. */
int  bt_scene_export_table(const bt_scene* s, float* table);
void bt_scene_free(bt_scene* s);
int  bt_view_create(bt_context* ctx, double view_angle, double off_up, double off_right,
                    int width, int height, int sphere, bt_view** out);
void bt_view_free(bt_view* v);
/* One frame from a device-resident grid.  cam7 null: the grid's own pose
 * (getParamArray).  Output to a host buffer (width*height*4, BGRA) and/or a
 * device buffer; with only a device buffer the call is asynchronous. */
int  bt_render_frame(bt_context* ctx, const bt_grid* g, const bt_scene* s, const bt_view* v,
                     const float* cam7, const bt_render_flags* flags, int step,
                     uint8_t* out_bgra, void* d_out_bgra);
/* Multi-grid camera-motion blending: the G > 1 frame of the reference
 * (pixInterpolation with Gr > 1, camUpdate, findBhCenter, findBhBorders,
 * smoothBorder; MakeImageKernel.cu:386-510,1415-1451), which its display
 * loop renders between two cached grids (MakeImageKernel.cu:986-1013).  The
 * frame lies between the poses of g0 and g1 at fraction alpha in [0, 1]:
 * the camera parameters are blended linearly (cam14: the two poses' 7-float
 * arrays; null: the grids' own poses), and near the shadow both grids are
 * sampled at positions warped so that their shadow borders (1000
 * directions) meet.  The grids must have equal sizes and levels.  Optional
 * outputs: thphi (h+1)(w+1)*2, border (2 + 2*1000)*2 floats. */
int  bt_render_frame_blend(bt_context* ctx, const bt_grid* g0, const bt_grid* g1, float alpha,
                           const bt_scene* s, const bt_view* v, const float* cam14,
                           const bt_render_flags* flags, int step, uint8_t* out_bgra,
                           void* d_out_bgra, float* thphi, float* border);
/* Star trails and diffraction of a frame sequence: the effects the
 * reference's display loop adds to consecutive frames (clearArrays,
 * makeGradField, addTrails, addDiffraction; MakeImageKernel.cu:98-174,
 * 517-537,618-629,775-800, called at :1020-1031).  A bt_trails holds, per
 * star, the pixels of its images in the last two frames; each
 * bt_render_frame_trails call renders the next frame of the sequence
 * (frame 0 after create or reset) with the trails from the previous frame
 * drawn into the star light, and, when a diffraction sprite was given
 * ((width/16)^2 RGB8 pixels, the reference's sprite size), the sprite
 * scattered around every pixel brighter than 255^2.  Where the reference
 * depends on thread scheduling (which 30 images of a star are kept, which
 * of two crossing trails wins a pixel) this build keeps a star's images in
 * pixel order and the per-channel maximum of crossing trails.  Optional
 * outputs: star (w*h*3, summed star light with trails and diffraction),
 * trail (w*h*3, this frame's trails).  The frame size, scene and context
 * must be the ones the state was made for. */
typedef struct bt_trails bt_trails;
int  bt_trails_create(bt_context* ctx, const bt_scene* s, int width, int height,
                      const uint8_t* diffraction_rgb, bt_trails** out);
void bt_trails_reset(bt_trails* t);
/* frames rendered since create/reset; images of the last frame dropped
 * because a star had more than 32 in one frame */
int  bt_trails_stats(bt_trails* t, int64_t* frames, int64_t* dropped);
void bt_trails_free(bt_trails* t);
int  bt_render_frame_trails(bt_context* ctx, const bt_grid* g, const bt_scene* s, const bt_view* v,
                            const float* cam7, const bt_render_flags* flags, int step,
                            bt_trails* trails, uint8_t* out_bgra, void* d_out_bgra, float* star,
                            float* trail);
/* Device bytes a frame of this size works in (intermediates, queues and
 * output; not the grid, scene or view): what the server's admission
 * reserves per frame.  No reference counterpart. */
size_t bt_frame_bytes(int width, int height, int step);
/* Same, also copying the frame's intermediates to host buffers (any may be
 * null): thphi (h+1)(w+1)*2, gap (h+1)(w+1), bh w*h, smoothed area w*h,
 * summed star light w*h*3.  For parity tests. */
int  bt_render_frame_stages(bt_context* ctx, const bt_grid* g, const bt_scene* s,
                            const bt_view* v, const float* cam7, const bt_render_flags* flags,
                            int step, uint8_t* out_bgra, void* d_out_bgra, float* thphi,
                            int32_t* gap, uint8_t* bh, float* area, float* star);

/* ---- the reference's on-disk caches and catalogue (host only) ----
 * cereal binary archives as the reference writes them (BlackTracer.cpp:20-47):
 * 32-bit ints raw, vectors as a 64-bit count then the elements, little
 * endian; float2/int2 elements as x then y. */
/* readStars2 + the magnitude cut (StarProcessor.h:34-46,71-96): records of
 * five numbers (magnitude, B-V, ra, dec, ignored; radians) ->
 * stars4 = {theta = pi/2 - dec, ra (wrapped), magnitude, B-V} of every star
 * with magnitude <= cut, in file order.  capacity 0: count only. */
int  bt_star_catalog_parse(const char* text, size_t len, float magnitude_cut, float* stars4,
                           int capacity, int* count);
/* .star: StarProcessor::serialize (StarProcessor.h:19-23). */
int  bt_star_file_write(const char* path, const int32_t* tree, int tree_level, const float* pos,
                        const float* mag, int star_count);
/* Null arrays: star_count and tree_level only. */
int  bt_star_file_read(const char* path, int32_t* tree, float* pos, float* mag, int* star_count,
                       int* tree_level);
/* .grid: Grid::serialize (Grid.h:33-39) with its PSHOffsetTable
 * (PSHOffsetTable.h:44-48), built by bt_grid_export_psh; a uniform grid
 * (start level == max level) writes empty tables, as the reference does
 * not build one (Grid.h:623). */
int  bt_grid_file_write(const char* path, const bt_grid* g);
/* Null arrays: header only.  hash 2*hw*hw floats, offsets 2*ow*ow ints,
 * tags 2*hw*hw ints: the bt_make_image table arguments. */
int  bt_grid_file_read(const char* path, int* max_level, int* N, int* M, int* hash_width,
                       int* offset_width, float* hash, int32_t* offsets, int32_t* tags);

/* ---- serving: continuous batching of render requests (no counterpart in
 * the reference, whose frame loop renders one frame at a time,
 * MakeImageKernel.cu:975-1204) ----
 * A server renders requests of any resolution and pose from one scene.  A
 * builder thread builds and caches in HBM the grid of each new pose
 * (Grid(max_level, start_level) at the configured integrator settings) and
 * the view of each new resolution; a dispatcher admits pending requests in
 * arrival order whenever one of `slots` frame slots is free and the frame's
 * working set fits the HBM budget next to the cached grids (idle grids are
 * evicted least recently used first).  Each slot has its own stream, so
 * admitted frames overlap on the device.  A request's result does not
 * depend on what else is in flight. */
typedef struct {
    size_t hbm_budget;        /* device bytes the server may hold (16 GiB): cached grids,
                                 views, every slot's work buffers at their high-water mark,
                                 a running build (bt_server_footprint) */
    int    slots;             /* frames in flight at most, 1..16 (4) */
    int    max_level, start_level;   /* grids of new poses (10, 1) */
    bt_integ_params integ;    /* integrator settings of those grids (defaults) */
    int    step;              /* star splat step (1) */
    double view_angle, off_up, off_right;   /* Viewer of every resolution (pi/2, 0, pi/16) */
} bt_server_config;

typedef struct {
    double   a, r, theta;     /* pose: camera on the circular orbit at (r, theta) of spin a */
    float    offset;          /* orbit position: the frame's celestial phi offset (>= 0) */
    int      width, height;
    int      star, map, lr;   /* passes, as bt_render_flags */
    uint8_t* out_bgra;        /* width*height*4 bytes, written before completion is posted */
    uint64_t tag;             /* caller's id, returned with the completion */
} bt_request;

typedef struct {
    uint64_t tag;
    int      status;          /* BT_OK or the error of the request */
    int      grid_hit;        /* the pose's grid was already resident on arrival */
    double   latency_ms;      /* arrival (submit) -> frame in the caller's buffer */
    double   queue_ms;        /* arrival -> admission */
    double   frame_ms;        /* admission -> frame in the caller's buffer */
} bt_completion;

typedef struct {
    int64_t submitted, completed, failed;
    int64_t grid_builds, grid_hits, evictions;
    int     peak_inflight;
    size_t  peak_bytes;       /* peak of the device bytes counted against the budget */
    double  p50_ms, p99_ms, mean_ms;   /* latency of completed requests */
    double  frames_per_s;     /* completed / (last completion - first arrival) */
    double  wall_s;
} bt_server_stats;

typedef struct bt_server bt_server;

/* What a server counts against its budget, per item. */
typedef struct {
    size_t slot;    /* a slot's work buffers after a width x height frame */
    size_t view;    /* the view of one resolution */
    size_t grid;    /* a cached grid (its arrays and FP32 render grid) */
    size_t build;   /* a grid build while it runs (arrays, render grid, work buffers) */
} bt_server_bytes;

void bt_server_config_default(bt_server_config* cfg);
int  bt_server_footprint(const bt_server_config* cfg, int width, int height, bt_server_bytes* out);
/* The scene must outlive the server. */
int  bt_server_create(bt_context* ctx, const bt_scene* scene, const bt_server_config* cfg,
                      bt_server** out);
/* Non-blocking.  BT_ERR_BUDGET if the frame alone exceeds the budget. */
int  bt_server_submit(bt_server* s, const bt_request* req);
/* Takes up to max completions (oldest first); *n receives the count. */
int  bt_server_poll(bt_server* s, bt_completion* out, int max, int* n);
/* Blocks until every submitted request has completed (timeout_s <= 0: no limit). */
int  bt_server_drain(bt_server* s, double timeout_s);
int  bt_server_stats_get(bt_server* s, bt_server_stats* stats);
void bt_server_destroy(bt_server* s);

#ifdef __cplusplus
}
#endif

#endif /* BT_ABI_H */
