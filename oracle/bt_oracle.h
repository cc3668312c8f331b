/*
 * bt_oracle.h -- CPU oracle for the BlackTracer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference algorithm (annemiekie/BlackTracer, BlackTracer2/BlackTracer/).
 * It is the checker that the HIP product is compared against, and the
 * CPU baseline that bench.py times.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Path A (grid build, FP64) follows Metric.h, Camera.h, BlackHole.h and
 * Grid.h operation for operation, built with -ffp-contract=off.
 * Path B (render, FP32) follows kernel.cuh and MakeImageKernel.cu.
 *
 * Pinning status of every function is listed in the header of
 * bt_oracle.c and bt_oracle_render.c.
 */
#ifndef BT_ORACLE_H
#define BT_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pose: BlackHole(a) + Camera(theta, phi, r, br, btheta, bphi)
 * (BlackHole.h:12-15, Camera.h:40-61).  All derived quantities are
 * computed by bto_pose_init exactly as Camera::initforms does. */
typedef struct {
    double a, asq;
    double theta, phi, r;
    double speed, br, btheta, bphi;
    double alpha, w, wbar, Delta, ro;
} bto_pose;

/* Integration parameters.  Reference defaults (Metric.h:307-318,27-33):
 * eps 1e-5, h1 0.01, z_end -1e7, max_steps 1000, a-priori capture on,
 * no radius stops. */
typedef struct {
    double eps;
    double h1;
    double z_end;
    int    max_steps;
    int    apriori_capture;   /* 1: checkCelest decides capture (reference) */
    double r_escape;          /* >0: stop as escaped once r > r_escape */
    double r_capture_factor;  /* >0: stop as captured once r < f * r_horizon */
    int    newton_cap;        /* iteration cap on checkB_Q's Newton loop */
} bto_params;

/* Per-ray counters used for the algorithmic-work tally. */
typedef struct {
    int64_t attempts;   /* Cash-Karp attempts (derivs + rkqs) */
    int64_t newton;     /* checkB_Q Newton iterations */
    int64_t golden;     /* findMinGoldSec iterations */
    int64_t rays;       /* rays handled */
    int64_t integrated; /* rays that entered odeint1 */
} bto_stats;

enum { BTO_ESCAPED = 0, BTO_CAPTURED = 1, BTO_MAXSTEP = 2 };

void   bto_default_params(bto_params* p);
void   bto_pose_init(bto_pose* p, double a, double theta, double phi, double r,
                     double br, double btheta, double bphi);
void   bto_pose_init_speed(bto_pose* p, double a, double theta, double phi,
                           double r, double speed);

/* Metric.h closed forms (exposed for pinning tests). */
double bto_calc_speed(double a, double r, double theta);
double bto_b0(double a, double r0);
double bto_q0(double a, double r0);
int    bto_check_bq(double a, double b, double q, int newton_cap, int64_t* iters);
int    bto_check_rup(double a, double r, double theta, double b, double q, int64_t* iters);
int    bto_check_celest(double a, double pr, double r, double theta, double b,
                        double q, int newton_cap, int64_t* newton, int64_t* golden);
void   bto_derivs(double a, const double* var, double* out, double b, double q);
void   bto_rkck(double a, const double* var, const double* dvdz, double h,
                double* out, double* err, double b, double q);
void   bto_wrap_to_pi(double* theta, double* phi);
/* Test knob: scale derivs' sin/cos by (1 + k 2^-53) (0: off, the default). */
void   bto_set_trig_ulp(int k);
void   bto_set_pow_ulp(int k);

/* Initial conditions of one camera-sky direction (Grid.h:524-547):
 * out[0..3] = pR, pTheta, b, q. */
void   bto_initial_conditions(const bto_pose* pose, double theta, double phi,
                              double* out);

/* One ray: integration_wrapper's loop body (Grid.h:524-554).
 * Returns BTO_ESCAPED/BTO_CAPTURED/BTO_MAXSTEP; (theta, phi) in/out
 * follow the reference's in-place convention: (-1,-1) when captured. */
int    bto_trace_ray(const bto_pose* pose, const bto_params* prm,
                     double* theta, double* phi, int* step, bto_stats* st);

/* Batch: integration_wrapper (Grid.h:514-556), optional OpenMP. */
void   bto_trace_rays(const bto_pose* pose, const bto_params* prm, int n,
                      double* theta, double* phi, int* step, int nthreads,
                      bto_stats* st);

/* Adaptive grid: Grid(maxLevel, startLevel, angle, cam, bh) (Grid.h:605-624)
 * on dense arrays.  cel is N*M*2 doubles ((-2,-2) = not traced),
 * steps N*M ints, level N*M int8 (-1 = no block key here).
 * refine_thresh: PRECCELEST (0.015, Grid.h:24).
 * fix_tvertices: 1 to run fixTvertices in the pinned order
 * (level ascending, then key ascending). */
typedef struct {
    int N, M, max_level, start_level, equafactor;
    int64_t rays_traced;
    double* cel;
    int*    steps;
    int8_t* level;
    uint8_t* present;   /* 1 where CamToCel holds the key */
} bto_grid;

int    bto_grid_dims(int max_level, int angle, int* N, int* M);
int    bto_grid_build(const bto_pose* pose, const bto_params* prm, int max_level,
                      int start_level, int angle, double refine_thresh,
                      int fix_tvertices, int nthreads, bto_grid* g, bto_stats* st);
void   bto_grid_fix_tvertices(bto_grid* g);
void   bto_grid_free(bto_grid* g);

/* Catmull-Rom (tension 0, bias 0) Hermite of Grid.h:183-204 on (theta, phi). */
void   bto_hermite2(double v, const double* x0, const double* x1, const double* x2,
                    const double* x3, double* out);

/* ---------------- Path B (bt_oracle_render.c) ---------------- */

/* Viewer(viewangle, offUp, offRight, w, h, false) (Viewer.h:20-33,109-133):
 * (h+1)*(w+1) float2 camera-sky directions of the pixel corners. */
int    btor_camera_view(double viewangle, double off_up, double off_right, int w, int h,
                        float* view);
void   btor_sphere_view(int w, int h, float* view);
/* StarProcessor::makeTree (StarProcessor.h:172-212); stars4 = S x
 * (theta, phi, magnitude, B-V); tree (1 << (L+1)) - 1 ints; pos/mag 2*S. */
int    btor_make_tree(const float* stars4, int S, int tree_level, int32_t* tree,
                      float* pos, float* mag);
/* Summed-area sky table (MakeImageKernel.cu:1312-1331): rows*cols*4 floats. */
void   btor_sum_table(const uint8_t* sky_bgr, int rows, int cols, float* table);
/* addLowLightStarsToImage (StarProcessor.h:108-158): stars4 (theta, phi,
 * magnitude, B-V) of the stars below the catalogue cut painted into the
 * BGR8 sky in place. */
void   btor_sky_low_stars(uint8_t* sky_bgr, int rows, int cols, const float* stars4, int n);
/* makeGrid from a perfect spatial hash (MakeImageKernel.cu:369-384). */
void   btor_make_grid(const float* hash_table, const int32_t* offset_table,
                      const int32_t* hash_pos_tag, int hw, int ow, int GN, int GM, float* grid);
/* One G == 1 frame (MakeImageKernel.cu:1455-1529): out_bgra w*h*4.  The
 * o_* intermediates are optional (thphi (h+1)(w+1)*2, gap (h+1)(w+1),
 * bh w*h, area w*h (smoothed), star w*h*3 (summed star light)). */
int    btor_frame(const float* grid, int GN, int GM, int gridlvl, const float* view, int w,
                  int h, const float* cam7, const int32_t* tree, int tree_level,
                  const float* star_pos, const float* star_mag, int S, const float* sum_table,
                  int sky_rows, int sky_cols, int star, int map, int lr, float hor, float ver,
                  float offset, int step, uint8_t* out_bgra, float* o_thphi, int32_t* o_gap,
                  uint8_t* o_bh, float* o_area, float* o_star);
/* Multi-grid blending (the G > 1 frame, MakeImageKernel.cu:386-510):
 * border = (2 + 2*1000) float2 from grids A and B (GN x GM, GN the rows of
 * the whole-sky grid); the frame between them at alpha, cam14 = the two
 * poses' parameter arrays.  o_thphi (h+1)(w+1)*2 and o_border optional. */
void   btor_blend_border(const float* grid_a, const float* grid_b, int GN, int GM, float* border);
int    btor_frame_blend(const float* grid_a, const float* grid_b, int GN, int GM, int gridlvl,
                        float alpha, const float* view, int w, int h, const float* cam14,
                        const int32_t* tree, int tree_level, const float* star_pos,
                        const float* star_mag, int S, const float* sum_table, int sky_rows,
                        int sky_cols, int star, int map, int lr, float hor, float ver,
                        float offset, int step, uint8_t* out_bgra, float* o_thphi,
                        float* o_border);
/* A frame sequence with star trails and diffraction (MakeImageKernel.cu:
 * 98-174,517-537,618-629,775-800): the state holds, per star, the pixels
 * of its images in the last two frames; each btor_frame_trails call is the
 * next frame.  sprite: (w/16)^2 RGB8 or NULL.  o_star (w*h*3, light with
 * trails and diffraction) and o_trail (w*h*3) optional. */
typedef struct btor_trails btor_trails;
btor_trails* btor_trails_new(int S, int w, int h, const uint8_t* sprite);
void   btor_trails_free(btor_trails* t);
int    btor_frame_trails(btor_trails* t, const float* grid, int GN, int GM, int gridlvl,
                         const float* view, int w, int h, const float* cam7, const int32_t* tree,
                         int tree_level, const float* star_pos, const float* star_mag, int S,
                         const float* sum_table, int sky_rows, int sky_cols, int star, int map,
                         int lr, float hor, float ver, float offset, int step, uint8_t* out_bgra,
                         float* o_star, float* o_trail);
/* Helpers exposed for pinning tests. */
void   btor_rgb_to_hsp(float r, float g, float b, float* hsp);
void   btor_hsp_to_rgb(float h, float s, float p, float* rgb);
float  btor_calc_area4(const float* t, const float* p);
float  btor_calc_area3(const float* t, const float* p);
float  btor_redshift(float theta, float phi, const float* cam7);
int    btor_star_in_polygon(const float* t, const float* p, float st, float sp, int sgn);
int    btor_search_tree(const int32_t* tree, const float* mn, const float* mx, int tree_level,
                        int* out, int cap);

#ifdef __cplusplus
}
#endif

#endif
