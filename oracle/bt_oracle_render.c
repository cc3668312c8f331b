/*
 * bt_oracle_render.c -- CPU oracle, Path B (render pass, FP32).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP render kernels.
 *
 * Restates, operation for operation in float, the reference's G == 1 frame
 * pipeline of MakeImageKernel.cu (makeGrid, pixInterpolation,
 * findBlackPixels, findArea, smoothAreaH/V, distortStarMap, sumStarLight,
 * makePix, distortEnvironmentMap, addStarsAndBackground) and kernel.cuh's
 * device helpers, plus the host-side inputs Viewer::makeCameraView,
 * StarProcessor::makeTree and the summed-area sky table
 * (MakeImageKernel.cu:1312-1331).
 *
 * Transcendentals.  The reference used CUDA 10's device libm (sinf, cosf,
 * atanf, powf, expf, log10f, exp10f), which is not available here.  Every
 * such call is evaluated in double and rounded to float, i.e. correctly
 * rounded except in ~2^-28 of cases; the HIP kernels do the same, so the
 * oracle and the device agree bit-for-bit.  Against the reference itself
 * these values are "parity unpinned" (1/255 per channel is the only bound).
 *
 * Divergences from the reference, all where the reference has undefined
 * behaviour (out-of-bounds reads or a data race), in favour of the evident
 * intent:
 *   - grid reads outside rows [0, GN) in findPoint read "absent" (-2,-2);
 *   - the per-pixel star-leaf list (16 slots, MakeImageKernel.cu:1299) is
 *     unbounded instead of overflowing into the neighbour's slots;
 *   - sumStarLight is bounds-checked (no writes past the row when W % 24 != 0);
 *   - summed-table rows -1 (single-column branch) read as zero and rows
 *     beyond the last are clamped;
 *   - the horizontal area-smoothing window wraps with a non-negative modulo;
 *   - the blackbody index is clamped to the table's last entry (390); the
 *     reference clamps to 1170 and reads past its 391-entry table.
 * Pixel 0 of the sky pass stays black, as in the reference (ind > 0 test).
 * Trails and diffraction are not part of this path (both are identities on
 * a first frame with no diffraction sprite).
 *
 * Multi-grid blending (the G > 1 frame: camUpdate, findBhCenter,
 * findBhBorders, smoothBorder and pixInterpolation with Gr > 1,
 * MakeImageKernel.cu:386-510) is restated in btor_blend_border and
 * btor_frame_blend, with two more divergences on undefined behaviour: the
 * border walk stops at the grid's edge (the reference walks on, reading
 * outside its grid), and the direction slot is clamped to angleNum - 1
 * (the reference reads one entry past its border array when the angle
 * rounds to 2 pi).  findBhCenter's extremes start from (100, 0) on every
 * call; the reference's accumulate over the grid pairs of a session.
 *
 * Star trails and diffraction of a frame sequence (clearArrays,
 * makeGradField, addTrails, addDiffraction; MakeImageKernel.cu:98-174,
 * 517-537,618-629,775-800) are restated in btor_frame_trails with the
 * scheduling-independent rules of the device: a star's images are kept in
 * pixel order (the reference keeps the first 30 in atomic order), crossing
 * trails keep the per-channel maximum (the reference's last writer wins),
 * diffraction sources and thresholds are the light before diffraction, the
 * trail walk is bounded (the reference's never ends for some moves) and
 * stays in the frame, gradient reads past the corner array count as
 * invalid, and shadow corners get a zero gradient (unwritten there).
 *
 * Pinning (each function, the check that would fail on a plausible
 * mistake in it; tests/test_oracle_render_pins.py (P),
 * tests/test_oracle_render_closed_forms.py (C), test_oracle_blend_pins.py,
 * test_oracle_trail_pins.py, test_low_stars_cpu.py):
 *   camera / sphere views     closed-form pinhole and equirect directions (C)
 *   make_tree, search_tree    brute-force leaf membership and cover (P)
 *   calc_area3/4              Van Oosterom-Strackee; quads tile 4 pi (P)
 *   pi_check(_tot), wrap_to_pi_f   seam unwrapping; brute force (C)
 *   hermite_f, intersection   exact on quadratics; line crossing (C)
 *   interpolate_* / find_*    quadratic fields reproduced, through coarse
 *                             blocks and next to captured points (C, P)
 *   rgb_to_hsp / hsp_to_rgb   round trip (P)
 *   redshift                  static observer; E_inf / E_cam; flat-space
 *                             special-relativistic Doppler (P, C)
 *   gaussian, interpolate_star, the star taps and the x20 sum
 *                             normalised profile; affine sub-pixel position;
 *                             taps = 100 frac g 10^(-0.4 (m + 4 log10 z))
 *                             times the blackbody colour; sums over stars (C)
 *   make_pix, composite       saturation / hue preservation; root sum of
 *                             squares (C)
 *   smooth_area               constant fields kept, a delta spread as
 *                             1/(2r+1) per axis (C)
 *   the sky pass              constant sky; footprint mean of a known sky;
 *                             lensing and redshift in HSP (C)
 *   sum_table, make_grid      summed-area identity (P); hash round trip (C)
 *   blend_*, trails           border walk on a disk, self-blend, endpoints;
 *                             trails along the motion, sprites (blend/trail)
 *   btor_sky_low_stars        5 x 5 splat closed form, overwrite order,
 *                             seam and poles (low stars)
 */
#include "bt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* kernel.cuh:23-25: float constants of the render code */
#define FPI2 6.283185307179586476f
#define FPI 3.141592653589793238f
#define SQRT2PI 2.506628274631f
/* Const.h doubles, used by the host-side Viewer */
static const double DPI = 3.141592653589793238;
static const double DPI1_2 = 3.141592653589793238 / 2.;
static const double DPI2 = 3.141592653589793238 * 2.;

static const int kTempRGB[1173] = {
#include "blackbody_rgb.inc"
};

/* correctly rounded float transcendentals (see header) */
static float f_sin(float x) { return (float)sin((double)x); }
static float f_cos(float x) { return (float)cos((double)x); }
static float f_atan(float x) { return (float)atan((double)x); }
static float f_atan2(float y, float x) { return (float)atan2((double)y, (double)x); }
static float f_acos(float x) { return (float)acos((double)x); }
static float f_pow(float x, float y) { return (float)pow((double)x, (double)y); }
static float f_exp(float x) { return (float)exp((double)x); }
static float f_log10(float x) { return (float)log10((double)x); }
static float f_exp10(float x) { return (float)pow(10.0, (double)x); }

typedef struct { float x, y; } f2;

/* float -> int as CUDA's cvt.rzi.s32.f32 does it: truncate, saturate,
 * NaN -> 0 (a plain C cast of NaN or an out-of-range value is undefined). */
static int cvt_i32(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return -2147483647 - 1;
    return (int)x;
}

/* ------------------------------------------------------------ Viewer */

/* Viewer.h:20-33 + 109-133 */
int btor_camera_view(double viewangle, double offsetVer, double offsetHor, int pixw, int pixh,
                     float* view) {
    double viewAngleWide = viewangle;
    double viewAngleUp = 1. * pixh * viewAngleWide / (1. * pixw);
    memset(view, 0, sizeof(float) * 2 * (size_t)(pixh + 1) * (pixw + 1));
    if (fabs(offsetVer) > (DPI - viewAngleUp) / 2.) return -1;
    float yleft = (float)tan(viewAngleWide / 2.f);
    float step = (2.f * yleft) / pixw;
    float zup = yleft * pixh / pixw;
    for (int i = 0; i < pixh + 1; i++) {
        float zval = zup - i * step;
        float halfH = i * 1.f / (pixh / 2.f);
        float theta = halfH == 1.f ? (float)(DPI1_2 - offsetVer)
                                   : (float)(atan(1. / zval) + (halfH > 1.f) * DPI - offsetVer);
        for (int j = 0; j < pixw + 1; j++) {
            float yval = yleft - j * step;
            float halfV = j * 1.f / (pixw / 2.f);
            float phi = halfV == 1.f ? (float)(DPI - offsetHor)
                                     : (float)(atan(1. / yval) + (halfV > 1.f) * DPI + DPI1_2 - offsetHor);
            size_t o = (size_t)(pixw + 1) * i + pixw - j;
            view[2 * o] = theta;
            view[2 * o + 1] = phi;
        }
    }
    return 0;
}

/* Viewer.h:82-88 */
void btor_sphere_view(int pixw, int pixh, float* view) {
    for (int i = 0; i < pixh + 1; i++)
        for (int j = 0; j < pixw + 1; j++) {
            size_t o = (size_t)i * (pixw + 1) + j;
            view[2 * o] = (float)(1.f * i * DPI / pixh);
            view[2 * o + 1] = (float)(1.f * j * DPI2 / pixw);
        }
}

/* ------------------------------------------------------------ star tree */

typedef struct {
    int level_max;
    int32_t* tree;
    float* pos;
    float* mag;
} tree_ctx;

/* StarProcessor.h:172-212; stars are index lists into s4 (theta, phi, mag, bv) */
static void make_tree(tree_ctx* c, const float* s4, int* idx, int n, int level, float thphi[2],
                      float size[2], int writePos) {
    int searchPos = 0;
    if (writePos != 0 && ((writePos + 1) & writePos) != 0) searchPos = c->tree[writePos - 1];
    c->tree[writePos] = searchPos + n;
    if (level == c->level_max) {
        for (int i = 0; i < n; i++) {
            const float* s = s4 + 4 * (size_t)idx[i];
            c->pos[(i + searchPos) * 2] = s[0];
            c->pos[(i + searchPos) * 2 + 1] = s[1];
            c->mag[(i + searchPos) * 2] = s[2];
            c->mag[(i + searchPos) * 2 + 1] = s[3];
        }
        return;
    }
    level++;
    int nn = level % 2;
    size[nn] = size[nn] * .5f;
    int* lu = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    int* rd = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    int nl = 0, nr = 0;
    float check = thphi[nn] + size[nn];
    for (int i = 0; i < n; i++) {
        if (s4[4 * (size_t)idx[i] + nn] < check)
            lu[nl++] = idx[i];
        else
            rd[nr++] = idx[i];
    }
    float sizeNew[2] = {size[0], size[1]};
    float thphiNew[2] = {thphi[0], thphi[1]};
    make_tree(c, s4, lu, nl, level, thphiNew, sizeNew, writePos * 2 + 1);
    float thphiNew2[2] = {thphi[0], thphi[1]};
    float sizeNew2[2] = {size[0], size[1]};
    thphiNew2[nn] += size[nn];
    make_tree(c, s4, rd, nr, level, thphiNew2, sizeNew2, writePos * 2 + 2);
    free(lu);
    free(rd);
}

/* StarProcessor.h:190-198: tree of (1 << (L+1)) - 1 ints, pos/mag 2*S floats */
int btor_make_tree(const float* stars4, int S, int tree_level, int32_t* tree, float* pos, float* mag) {
    tree_ctx c = {tree_level, tree, pos, mag};
    int* idx = (int*)malloc(sizeof(int) * (S > 0 ? S : 1));
    for (int i = 0; i < S; i++) idx[i] = i;
    float thphi[2] = {0, 0};
    float size[2] = {(float)DPI, (float)DPI2};
    make_tree(&c, stars4, idx, S, 0, thphi, size, 0);
    free(idx);
    return 0;
}

/* ------------------------------------------------------------ math helpers */

/* kernel.cuh:586-610 */
static float calc_area4(const float t[4], const float p[4]) {
    float x[4], y[4], z[4];
    for (int q = 0; q < 4; q++) {
        float sint = f_sin(t[q]);
        x[q] = sint * f_cos(p[q]);
        y[q] = sint * f_sin(p[q]);
        z[q] = f_cos(t[q]);
    }
    float dotpr1 = 1.f;
    dotpr1 += x[0] * x[2] + y[0] * y[2] + z[0] * z[2];
    float dotpr2 = dotpr1;
    dotpr1 += x[2] * x[1] + y[2] * y[1] + z[2] * z[1];
    dotpr1 += x[0] * x[1] + y[0] * y[1] + z[0] * z[1];
    dotpr2 += x[0] * x[3] + y[0] * y[3] + z[0] * z[3];
    dotpr2 += x[2] * x[3] + y[2] * y[3] + z[2] * z[3];
    float triprod1 = fabsf(x[0] * (y[1] * z[2] - y[2] * z[1]) - y[0] * (x[1] * z[2] - x[2] * z[1]) +
                           z[0] * (x[1] * y[2] - x[2] * y[1]));
    float triprod2 = fabsf(x[0] * (y[2] * z[3] - y[3] * z[2]) - y[0] * (x[2] * z[3] - x[3] * z[2]) +
                           z[0] * (x[2] * y[3] - x[3] * y[2]));
    return 2.f * (f_atan(triprod1 / dotpr1) + f_atan(triprod2 / dotpr2));
}

/* kernel.cuh:612-649 */
static float calc_area3(const float t[3], const float p[3]) {
    float xi[3], yi[3], zi[3];
    for (int q = 0; q < 3; q++) {
        float sint = f_sin(t[q]);
        xi[q] = sint * f_cos(p[q]);
        yi[q] = sint * f_sin(p[q]);
        zi[q] = f_cos(t[q]);
    }
    float dot01 = xi[0] * xi[1] + yi[0] * yi[1] + zi[0] * zi[1];
    float dot02 = xi[0] * xi[2] + yi[0] * yi[2] + zi[0] * zi[2];
    float dot12 = xi[2] * xi[1] + yi[2] * yi[1] + zi[2] * zi[1];
    float x[3] = {xi[0], xi[1], xi[2]};
    float y[3] = {yi[0], yi[1], yi[2]};
    float z[3] = {zi[0], zi[1], zi[2]};
    if (dot01 < dot02 && dot01 < dot12) {
        x[0] = xi[2]; x[1] = xi[0]; x[2] = xi[1];
        y[0] = yi[2]; y[1] = yi[0]; y[2] = yi[1];
        z[0] = zi[2]; z[1] = zi[0]; z[2] = zi[1];
    } else if (dot02 < dot12 && dot02 < dot01) {
        x[0] = xi[1]; x[1] = xi[2]; x[2] = xi[0];
        y[0] = yi[1]; y[1] = yi[2]; y[2] = yi[0];
        z[0] = zi[1]; z[1] = zi[2]; z[2] = zi[0];
    }
    float dotpr1 = 1.f;
    dotpr1 += x[0] * x[2] + y[0] * y[2] + z[0] * z[2];
    dotpr1 += x[2] * x[1] + y[2] * y[1] + z[2] * z[1];
    dotpr1 += x[0] * x[1] + y[0] * y[1] + z[0] * z[1];
    float triprod1 = fabsf(x[0] * (y[1] * z[2] - y[2] * z[1]) - y[0] * (x[1] * z[2] - x[2] * z[1]) +
                           z[0] * (x[1] * y[2] - x[2] * y[1]));
    return 2.f * (f_atan(triprod1 / dotpr1));
}

/* kernel.cuh:274-295 */
static int pi_check(float* p, float factor) {
    float factor1 = FPI2 * (1.f - factor);
    int check = 0;
    for (int q = 0; q < 4; q++)
        if (p[q] > factor1) {
            check = 1;
            break;
        }
    if (!check) return 0;
    check = 0;
    float factor2 = FPI2 * factor;
    for (int q = 0; q < 4; q++)
        if (p[q] < factor2) {
            p[q] += FPI2;
            check = 1;
        }
    return check;
}

/* kernel.cuh:303-322 */
static int pi_check_tot(f2* tp, float factor, int size) {
    float factor1 = FPI2 * (1.f - factor);
    int check = 0;
    for (int q = 0; q < size; q++)
        if (tp[q].y > factor1) {
            check = 1;
            break;
        }
    if (!check) return 0;
    check = 0;
    float factor2 = FPI2 * factor;
    for (int q = 0; q < size; q++)
        if (tp[q].y < factor2) {
            tp[q].y += FPI2;
            check = 1;
        }
    return check;
}

/* kernel.cuh:694-703 */
static void wrap_to_pi_f(float* thetaW, float* phiW) {
    *thetaW = fmodf(*thetaW, FPI2);
    while (*thetaW < 0.f) *thetaW += FPI2;
    if (*thetaW > FPI) {
        *thetaW -= 2.f * (*thetaW - FPI);
        *phiW += FPI;
    }
    while (*phiW < 0.f) *phiW += FPI2;
    *phiW = fmodf(*phiW, FPI2);
}

/* kernel.cuh:790-815 with tension = bias = 0 */
static f2 hermite_f(float v, f2 x0, f2 x1, f2 x2, f2 x3) {
    float const v2 = v * v;
    float const v3 = v * v2;
    float const aa = (1.f + 0.f) * (1.f - 0.f) / 2.f;
    float const bb = (1.f - 0.f) * (1.f - 0.f) / 2.f;
    float const m0T = aa * (x1.x - x0.x) + bb * (x2.x - x1.x);
    float const m0P = aa * (x1.y - x0.y) + bb * (x2.y - x1.y);
    float const m1T = aa * (x2.x - x1.x) + bb * (x3.x - x2.x);
    float const m1P = aa * (x2.y - x1.y) + bb * (x3.y - x2.y);
    float const u0 = 2.f * v3 - 3.f * v2 + 1.f;
    float const u1 = v3 - 2.f * v2 + v;
    float const u2 = v3 - v2;
    float const u3 = -2.f * v3 + 3.f * v2;
    f2 r = {u0 * x1.x + u1 * m0T + u2 * m1T + u3 * x2.x, u0 * x1.y + u1 * m0P + u2 * m1P + u3 * x2.y};
    return r;
}

/* kernel.cuh:754-772 (double intersection of two lines) */
static f2 intersection(float ax, float ay, float bx, float by, float cx, float cy, float dx, float dy) {
    double a1 = by - ay;
    double b1 = ax - bx;
    double c1 = a1 * (ax) + b1 * (ay);
    double a2 = dy - cy;
    double b2 = cx - dx;
    double c2 = a2 * (cx) + b2 * (cy);
    double determinant = a1 * b2 - a2 * b1;
    f2 r = {-1, -1};
    if (determinant == 0) return r;
    double x = (b2 * c1 - b1 * c2) / determinant;
    double y = (a1 * c2 - a2 * c1) / determinant;
    r.x = (float)x;
    r.y = (float)y;
    return r;
}

/* ------------------------------------------------------------ grid interpolation */

typedef struct {
    const float* grid; /* GN x GM float2 */
    int GN, GM;
} gridv;

static f2 grid_at(const gridv* g, int i, int j) {
    f2 r = {-2.f, -2.f};
    if (i < 0 || i >= g->GN || j < 0 || j >= g->GM) return r;
    size_t o = (size_t)i * g->GM + j;
    r.x = g->grid[2 * o];
    r.y = g->grid[2 * o + 1];
    return r;
}

/* kernel.cuh:774-788 */
static f2 interpolate_linear(float percDown, float percRight, const f2* c) {
    float phi[4] = {c[0].y, c[1].y, c[2].y, c[3].y};
    float theta[4] = {c[0].x, c[1].x, c[2].x, c[3].x};
    pi_check(phi, 0.2f);
    float leftT = theta[0] + percDown * (theta[2] - theta[0]);
    float leftP = phi[0] + percDown * (phi[2] - phi[0]);
    float rightT = theta[1] + percDown * (theta[3] - theta[1]);
    float rightP = phi[1] + percDown * (phi[3] - phi[1]);
    float upT = theta[0] + percRight * (theta[1] - theta[0]);
    float upP = phi[0] + percRight * (phi[1] - phi[0]);
    float downT = theta[2] + percRight * (theta[3] - theta[2]);
    float downP = phi[2] + percRight * (phi[3] - phi[2]);
    return intersection(upT, upP, downT, downP, leftT, leftP, rightT, rightP);
}

static f2 interpolate_hermite(const gridv* g, int i, int j, int gap, float percDown, float percRight,
                              f2* cornersCel, int count);

/* kernel.cuh:817-889 */
static f2 find_point(const gridv* g, int i, int j, int offver, int offhor, int gap, int count) {
    const int GM = g->GM, GN = g->GN;
    f2 gridpt = grid_at(g, i, j);
    f2 m1 = {-1, -1};
    if (gridpt.x == -2 && gridpt.y == -2) {
        int j2 = (j + offhor * gap + GM) % GM;
        int i2 = i + offver * gap;
        f2 ij2 = grid_at(g, i2, j2);
        if (ij2.x == -1 && ij2.y == -1) return m1;
        else if (ij2.x != -2 && ij2.y != -2) {
            int j0 = (j - offhor * gap + GM) % GM;
            int i0 = (i - offver * gap);
            f2 ij0 = grid_at(g, i0, j0);
            if (ij0.x < 0) return m1;
            int jprev = (j - 3 * offhor * gap + GM) % GM;
            int jnext = (j + 3 * offhor * gap + GM) % GM;
            int iprev = i - offver * 3 * gap;
            int inext = i + offver * 3 * gap;
            if (offver != 0) {
                if (i2 == 0) {
                    jnext = (j0 + GM / 2) % GM;
                    inext = i0;
                } else if (i0 == GN - 1) {
                    jprev = (j0 + GM / 2) % GM;
                    iprev = i2;
                } else if (i2 == GN - 1) {
                    inext = i0;
                    jnext = (j0 + GM / 2) % GM;
                }
            }
            f2 ijprev = grid_at(g, iprev, jprev);
            f2 ijnext = grid_at(g, inext, jnext);
            if (ijprev.x > -2 && ijnext.x > -2) {
                f2 pt[4] = {ijprev, ij0, ij2, ijnext};
                if (pt[0].x != -1 && pt[3].x != -1) {
                    pi_check_tot(pt, 0.2f, 4);
                    return hermite_f(0.5f, pt[0], pt[1], pt[2], pt[3]);
                }
            }
            f2 pt[2] = {ij2, ij0};
            pi_check_tot(pt, 0.2f, 2);
            f2 r = {.5f * (pt[0].x + pt[1].x), .5f * (pt[0].y + pt[1].y)};
            return r;
        } else {
            if (count > 0) return m1;
            int j0 = (j + gap) % GM;
            int j1 = (j - gap + GM) % GM;
            if (i - gap < 0) return m1;
            f2 cc[12];
            cc[0] = grid_at(g, i + gap, j0);
            cc[1] = grid_at(g, i - gap, j0);
            cc[2] = grid_at(g, i - gap, j1);
            cc[3] = grid_at(g, i + gap, j1);
            for (int q = 0; q < 4; q++)
                if (cc[q].x == -1 || cc[q].x == -2) return m1;
            return interpolate_hermite(g, i - gap, j1, 2 * gap, .5f, .5f, cc, 1);
        }
    }
    return gridpt;
}

/* kernel.cuh:891-938 */
static f2 interpolate_hermite(const gridv* g, int i, int j, int gap, float percDown, float percRight,
                              f2* c, int count) {
    const int GM = g->GM, GN = g->GN;
    int k = i + gap;
    int l = (j + gap) % GM;
    int imin1 = i - gap;
    int kplus1 = k + gap;
    int jmin1 = (j - gap + GM) % GM;
    int lplus1 = (l + gap) % GM;
    int jx = j, jy = j, lx = l, ly = l;
    if (i == 0) {
        jx = (j + GM / 2) % GM;
        lx = (jx + gap) % GM;
        imin1 = k;
    } else if (k == GN - 1) {
        jy = (j + GM / 2) % GM;
        ly = (jy + gap) % GM;
        kplus1 = i;
    }
    c[4] = find_point(g, i, jmin1, 0, -1, gap, count);
    c[5] = find_point(g, i, lplus1, 0, 1, gap, count);
    c[6] = find_point(g, k, jmin1, 0, -1, gap, count);
    c[7] = find_point(g, k, lplus1, 0, 1, gap, count);
    c[8] = find_point(g, imin1, jx, -1, 0, gap, count);
    c[9] = find_point(g, imin1, lx, -1, 0, gap, count);
    c[10] = find_point(g, kplus1, jy, 1, 0, gap, count);
    c[11] = find_point(g, kplus1, ly, 1, 0, gap, count);
    for (int q = 4; q < 12; q++)
        if (c[q].x == -1) return interpolate_linear(percDown, percRight, c);
    pi_check_tot(c, 0.2f, 12);
    f2 up = hermite_f(percRight, c[4], c[0], c[1], c[5]);
    f2 down = hermite_f(percRight, c[6], c[2], c[3], c[7]);
    f2 upup = {c[8].x + (c[9].x - c[8].x) * percRight, c[8].y + (c[9].y - c[8].y) * percRight};
    f2 downdown = {c[10].x + (c[11].x - c[10].x) * percRight, c[10].y + (c[11].y - c[10].y) * percRight};
    return hermite_f(percDown, upup, up, down, downdown);
}

/* kernel.cuh:940-968 */
static f2 interpolate_spline(const gridv* g, int i, int j, int gap, float thetaCam, float phiCam,
                             f2* c, const float* cornersCam) {
    float thetaUp = cornersCam[0];
    float thetaDown = cornersCam[2];
    float phiLeft = cornersCam[1];
    float phiRight = cornersCam[3];
    if (thetaUp == thetaCam) {
        if (phiLeft == phiCam) return c[0];
        if (phiRight == phiCam) return c[1];
        if (i == 0.f) return c[0];
    } else if (thetaDown == thetaCam) {
        if (phiLeft == phiCam) return c[2];
        if (phiRight == phiCam) return c[3];
    }
    for (int q = 0; q < 4; q++)
        if (c[q].x == -1 && c[q].y == -1) {
            f2 m1 = {-1.f, -1.f};
            return m1;
        }
    float percDown = (thetaCam - thetaUp) / (thetaDown - thetaUp);
    float percRight = (phiCam - phiLeft) / (phiRight - phiLeft);
    return interpolate_hermite(g, i, j, gap, percDown, percRight, c, 0);
}

/* kernel.cuh:736-752 */
static void find_block(const gridv* g, float theta, float phi, int* i, int* j, int* gap, int level) {
    const int GM = g->GM;
    for (int s = 0; s < level + 1; s++) {
        int ngap = *gap / 2;
        int k = *i + ngap;
        int l = *j + ngap;
        if (*gap <= 1 || grid_at(g, k, l).x == -2.f) return;
        float thHalf = FPI2 * k / (1.f * GM);
        float phHalf = FPI2 * l / (1.f * GM);
        if (thHalf <= theta) *i = k;
        if (phHalf <= phi) *j = l;
        *gap = ngap;
    }
}

/* MakeImageKernel.cu:328-351 */
static f2 interpolate_pix(const gridv* g, float theta, float phi, int gridlvl, int* gapsave) {
    const int GM = g->GM;
    int half = (phi < FPI) ? 0 : 1;
    int a = 0;
    int b = half * GM / 2;
    int gap = GM / 2;
    find_block(g, theta, phi, &a, &b, &gap, gridlvl);
    *gapsave = gap;
    int k = a + gap;
    int l = b + gap;
    float factor = FPI2 / (1.f * GM);
    float cornersCam[4] = {factor * a, factor * b, factor * k, factor * l};
    l = l % GM;
    f2 nul = {-1, -1};
    f2 c[12] = {grid_at(g, a, b), grid_at(g, a, l), grid_at(g, k, b), grid_at(g, k, l),
                nul, nul, nul, nul, nul, nul, nul, nul};
    f2 r = interpolate_spline(g, a, b, gap, theta, phi, c, cornersCam);
    if (!(r.x == -1 && r.y == -1)) wrap_to_pi_f(&r.x, &r.y);
    return r;
}

/* ------------------------------------------------------------ colour helpers */

#define Pr .299f
#define Pg .587f
#define Pb .114f

/* kernel.cuh:388-427 */
static void rgb_to_hsp(float R, float G, float B, float* H, float* S, float* P) {
    *P = sqrtf(R * R * Pr + G * G * Pg + B * B * Pb);
    if (R == G && R == B) {
        *H = 0.f;
        *S = 0.f;
        return;
    }
    if (R >= G && R >= B) {
        if (B >= G) {
            *H = 6.f / 6.f - 1.f / 6.f * (B - G) / (R - G);
            *S = 1.f - G / R;
        } else {
            *H = 0.f / 6.f + 1.f / 6.f * (G - B) / (R - B);
            *S = 1.f - B / R;
        }
    } else if (G >= R && G >= B) {
        if (R >= B) {
            *H = 2.f / 6.f - 1.f / 6.f * (R - B) / (G - B);
            *S = 1.f - B / G;
        } else {
            *H = 2.f / 6.f + 1.f / 6.f * (B - R) / (G - R);
            *S = 1.f - R / G;
        }
    } else {
        if (G >= R) {
            *H = 4.f / 6.f - 1.f / 6.f * (G - R) / (B - R);
            *S = 1.f - R / B;
        } else {
            *H = 4.f / 6.f + 1.f / 6.f * (R - G) / (B - G);
            *S = 1.f - G / B;
        }
    }
}

/* kernel.cuh:440-524 */
static void hsp_to_rgb(float H, float S, float P, float* R, float* G, float* B) {
    float part, minOverMax = 1.f - S;
    if (minOverMax > 0.f) {
        if (H < 1.f / 6.f) {
            H = 6.f * H;
            part = 1.f + H * (1.f / minOverMax - 1.f);
            *B = P / sqrtf(Pr / minOverMax / minOverMax + Pg * part * part + Pb);
            *R = *B / minOverMax;
            *G = *B + H * (*R - *B);
        } else if (H < 2.f / 6.f) {
            H = 6.f * (-H + 2.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            *B = P / sqrtf(Pg / minOverMax / minOverMax + Pr * part * part + Pb);
            *G = *B / minOverMax;
            *R = *B + H * (*G - *B);
        } else if (H < 3.f / 6.f) {
            H = 6.f * (H - 2.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            *R = P / sqrtf(Pg / minOverMax / minOverMax + Pb * part * part + Pr);
            *G = *R / minOverMax;
            *B = *R + H * (*G - *R);
        } else if (H < 4.f / 6.f) {
            H = 6.f * (-H + 4.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            *R = P / sqrtf(Pb / minOverMax / minOverMax + Pg * part * part + Pr);
            *B = *R / minOverMax;
            *G = *R + H * (*B - *R);
        } else if (H < 5.f / 6.f) {
            H = 6.f * (H - 4.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            *G = P / sqrtf(Pb / minOverMax / minOverMax + Pr * part * part + Pg);
            *B = *G / minOverMax;
            *R = *G + H * (*B - *G);
        } else {
            H = 6.f * (-H + 6.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            *G = P / sqrtf(Pr / minOverMax / minOverMax + Pb * part * part + Pg);
            *R = *G / minOverMax;
            *B = *G + H * (*R - *G);
        }
    } else {
        if (H < 1.f / 6.f) {
            H = 6.f * (H);
            *R = sqrtf(P * P / (Pr + Pg * H * H));
            *G = *R * H;
            *B = 0.f;
        } else if (H < 2.f / 6.f) {
            H = 6.f * (-H + 2.f / 6.f);
            *G = sqrtf(P * P / (Pg + Pr * H * H));
            *R = *G * H;
            *B = 0.f;
        } else if (H < .5f) {
            H = 6.f * (H - 2.f / 6.f);
            *G = sqrtf(P * P / (Pg + Pb * H * H));
            *B = *G * H;
            *R = 0.f;
        } else if (H < 4.f / 6.f) {
            H = 6.f * (-H + 4.f / 6.f);
            *B = sqrtf(P * P / (Pb + Pg * H * H));
            *G = *B * H;
            *R = 0.f;
        } else if (H < 5.f / 6.f) {
            H = 6.f * (H - 4.f / 6.f);
            *B = sqrtf(P * P / (Pb + Pr * H * H));
            *R = *B * H;
            *G = 0.f;
        } else {
            H = 6.f * (-H + 1.f);
            *R = sqrtf(P * P / (Pr + Pb * H * H));
            *B = *R * H;
            *G = 0.f;
        }
    }
}

void btor_rgb_to_hsp(float r, float g, float b, float* hsp) { rgb_to_hsp(r, g, b, &hsp[0], &hsp[1], &hsp[2]); }
void btor_hsp_to_rgb(float h, float s, float p, float* rgb) { hsp_to_rgb(h, s, p, &rgb[0], &rgb[1], &rgb[2]); }
float btor_calc_area4(const float* t, const float* p) { return calc_area4(t, p); }
float btor_calc_area3(const float* t, const float* p) { return calc_area3(t, p); }

/* kernel.cuh:348-365; cam = getParamArray() {speed, alpha, w, wbar, br, btheta, bphi} */
static float redshift(float theta, float phi, const float* cam) {
    float xCam = f_sin(theta) * f_cos(phi);
    float zCam = f_cos(theta);
    float yCam = f_sin(theta) * f_sin(phi);
    float part = (1.f - cam[0] * yCam);
    float betaPart = sqrtf(1.f - cam[0] * cam[0]) / part;
    float xFido = -sqrtf(1.f - cam[0] * cam[0]) * xCam / part;
    float zFido = -sqrtf(1.f - cam[0] * cam[0]) * zCam / part;
    float yFido = (-yCam + cam[0]) / part;
    float k = sqrtf(1.f - cam[5] * cam[5]);
    float phiFido = -xFido * cam[4] / k + cam[6] * yFido + cam[6] * cam[5] / k * zFido;
    float eF = 1.f / (cam[1] + cam[2] * cam[3] * phiFido);
    float b = eF * cam[3] * phiFido;
    return 1.f / (betaPart * (1.f - b * cam[2]) / cam[1]);
}
float btor_redshift(float theta, float phi, const float* cam) { return redshift(theta, phi, cam); }

/* smoothAreaH then smoothAreaV (MakeImageKernel.cu:570-616): a box mean of
 * radius floor(max corner gap / 2) over the non-shadow pixels of the row
 * (wrapping in j), then of the column (clamped in i).  area (W x H) is
 * replaced by the result; tmp holds the horizontal pass. */
static void smooth_area(float* area, const uint8_t* bh, const int32_t* gap, int W, int H, float* tmp) {
    const int M = W, N = H, M1 = W + 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < N; i++)
        for (int j = 0; j < M; j++) {
            size_t ij = (size_t)i * M + j;
            tmp[ij] = 0.f;
            if (bh[ij] == 1) continue;
            int fs = gap[i * M1 + j];
            if (gap[i * M1 + M1 + j] > fs) fs = gap[i * M1 + M1 + j];
            if (gap[i * M1 + j + 1] > fs) fs = gap[i * M1 + j + 1];
            if (gap[i * M1 + M1 + j + 1] > fs) fs = gap[i * M1 + M1 + j + 1];
            fs = fs / 2;
            float sum = 0.f;
            int count = 0;
            for (int h = -fs; h <= fs; h++) {
                int jj = ((j + h) % M + M) % M;
                if (bh[(size_t)i * M + jj] == 0) {
                    sum += area[(size_t)i * M + jj];
                    count++;
                }
            }
            tmp[ij] = sum / (1.f * count);
        }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < N; i++)
        for (int j = 0; j < M; j++) {
            size_t ij = (size_t)i * M + j;
            if (bh[ij] == 1) continue;
            int fs = gap[i * M1 + j];
            if (gap[i * M1 + M1 + j] > fs) fs = gap[i * M1 + M1 + j];
            if (gap[i * M1 + j + 1] > fs) fs = gap[i * M1 + j + 1];
            if (gap[i * M1 + M1 + j + 1] > fs) fs = gap[i * M1 + M1 + j + 1];
            fs = fs / 2;
            float sum = 0.f;
            int minv = i - fs > 0 ? i - fs : 0;
            int maxv = i + fs < N - 1 ? i + fs : N - 1;
            int count = 0;
            for (int h = minv; h <= maxv; h++)
                if (bh[(size_t)h * M + j] == 0) {
                    sum += tmp[(size_t)h * M + j];
                    count++;
                }
            area[ij] = sum / (1.f * count);
        }
}
void btor_smooth_area(float* area, const uint8_t* bh, const int32_t* gap, int W, int H) {
    float* tmp = (float*)malloc(sizeof(float) * (size_t)W * H);
    smooth_area(area, bh, gap, W, H, tmp);
    free(tmp);
}

/* makePix (MakeImageKernel.cu:802-835): summed squared light -> BGRA8.
 * rgb = sqrt(light); if a channel exceeds 255 the colour is renormalised
 * through HSP with P = 1. */
static void make_pix(float br, float bg, float bb, uint8_t* o) {
    float rgb[3] = {sqrtf(br), sqrtf(bg), sqrtf(bb)};
    float mxc = fmaxf(fmaxf(rgb[0], rgb[1]), rgb[2]);
    if (mxc > 255.f) {
        float hh, ss, pp;
        rgb_to_hsp(rgb[0] / 255.f, rgb[1] / 255.f, rgb[2] / 255.f, &hh, &ss, &pp);
        pp = 1.0f;
        hsp_to_rgb(hh, ss, pp, &rgb[0], &rgb[1], &rgb[2]);
        rgb[0] *= 255;
        rgb[1] *= 255;
        rgb[2] *= 255;
    }
    int cb = cvt_i32(rgb[2] + 0.f), cg = cvt_i32(rgb[1] + 0.f), cr = cvt_i32(rgb[0] + 0.f);
    o[0] = (uint8_t)(cb < 255 ? cb : 255);
    o[1] = (uint8_t)(cg < 255 ? cg : 255);
    o[2] = (uint8_t)(cr < 255 ? cr : 255);
    o[3] = 255;
}
void btor_make_pix(float r, float g, float b, uint8_t* bgra) { make_pix(r, g, b, bgra); }

/* ------------------------------------------------------------ frame stages */

typedef struct {
    int W, H;
    const float* view;
    float* thphi;
} frame_ctx;

/* kernel.cuh:652-672; returns 0 (ind = -1) when a corner is negative */
static int retrieve_corners(const float* thphi, int W, int ind, float* t, float* p, int* picheck,
                            float offset) {
    const int M1 = W + 1;
    t[0] = thphi[2 * (ind + M1)];
    t[1] = thphi[2 * ind];
    t[2] = thphi[2 * (ind + 1)];
    t[3] = thphi[2 * (ind + M1 + 1)];
    p[0] = thphi[2 * (ind + M1) + 1];
    p[1] = thphi[2 * ind + 1];
    p[2] = thphi[2 * (ind + 1) + 1];
    p[3] = thphi[2 * (ind + M1 + 1) + 1];
    for (int q = 0; q < 4; q++) {
        if (p[q] < 0 || t[q] < 0) return 0;
        p[q] = fmodf(p[q] + offset, FPI2);
    }
    *picheck = pi_check(p, .2f);
    return 1;
}

/* kernel.cuh:674-692 */
static void lensing_redshift(const float* t, const float* p, int W, int ind, const float* cam,
                             const float* view, float* frac, float* redshft, float solidAngle) {
    const int M1 = W + 1;
    if (solidAngle == 0.f) {
        float th1[3] = {t[0], t[1], t[2]}, ph1[3] = {p[0], p[1], p[2]};
        float th2[3] = {t[0], t[2], t[3]}, ph2[3] = {p[0], p[2], p[3]};
        solidAngle = calc_area3(th1, ph1) + calc_area3(th2, ph2);
    }
    float ver4[4] = {view[2 * ind], view[2 * (ind + 1)], view[2 * (ind + M1 + 1)], view[2 * (ind + M1)]};
    float hor4[4] = {view[2 * ind + 1], view[2 * (ind + 1) + 1], view[2 * (ind + M1 + 1) + 1],
                     view[2 * (ind + M1) + 1]};
    float pixArea = calc_area4(ver4, hor4);
    *frac = pixArea / solidAngle;
    float thetaCam = (ver4[0] + ver4[1] + ver4[2] + ver4[3]) * .25f;
    float phiCam = (hor4[0] + hor4[1] + hor4[2] + hor4[3]) * .25f;
    *redshft = redshift(thetaCam, phiCam, cam);
}

/* kernel.cuh:245-266 */
static int cross_ok(float t_a, float t_b, float p_a, float p_b, float st, float sp, int sgn) {
    float c1t = (float)sgn * (t_a - t_b);
    float c1p = (float)sgn * (p_a - p_b);
    float c2t = sgn ? st - t_b : st - t_a;
    float c2p = sgn ? sp - p_b : sp - p_a;
    return (c1t * c2p - c2t * c1p) > 0;
}
static int star_in_polygon(const float* t, const float* p, float st, float sp, int sgn) {
    return cross_ok(t[0], t[1], p[0], p[1], st, sp, sgn) && cross_ok(t[1], t[2], p[1], p[2], st, sp, sgn) &&
           cross_ok(t[2], t[3], p[2], p[3], st, sp, sgn) && cross_ok(t[3], t[0], p[3], p[0], st, sp, sgn);
}
int btor_star_in_polygon(const float* t, const float* p, float st, float sp, int sgn) {
    return star_in_polygon(t, p, st, sp, sgn);
}

/* kernel.cuh:112-200: sub-pixel position by 10-step bisection */
static void interpolate_star(float t0, float t1, float t2, float t3, float p0, float p1, float p2, float p3,
                             float* start, float* starp, int sgn) {
    float error = 0.00001f;
    float midT = (t0 + t1 + t2 + t3) * .25f;
    float midP = (p0 + p1 + p2 + p3) * .25f;
    float starInPixY = 0.5f;
    float starInPixX = 0.5f;
    float perc = 0.5f;
    for (int q = 0; q < 10; q++) {
        if ((fabsf(*start - midT) < error) && (fabsf(*starp - midP) < error)) break;
        float half01T = (t0 + t1) * .5f;
        float half23T = (t2 + t3) * .5f;
        float half12T = (t2 + t1) * .5f;
        float half03T = (t0 + t3) * .5f;
        float half01P = (p0 + p1) * .5f;
        float half23P = (p2 + p3) * .5f;
        float half12P = (p2 + p1) * .5f;
        float half03P = (p0 + p3) * .5f;
        float line01to23T = half23T - half01T;
        float line03to12T = half12T - half03T;
        float line01to23P = half23P - half01P;
        float line03to12P = half12P - half03P;
        float line01toStarT = *start - half01T;
        float line03toStarT = *start - half03T;
        float line01toStarP = *starp - half01P;
        float line03toStarP = *starp - half03P;
        int a = (((line03to12T * line03toStarP) - (line03toStarT * line03to12P)) > 0.f) ? 1 : -1;
        int b = (((line01to23T * line01toStarP) - (line01toStarT * line01to23P)) > 0.f) ? 1 : -1;
        perc *= 0.5f;
        if (sgn * a > 0) {
            if (sgn * b > 0) {
                t2 = half12T; t0 = half01T; t3 = midT;
                p2 = half12P; p0 = half01P; p3 = midP;
                starInPixX -= perc; starInPixY -= perc;
            } else {
                t2 = midT; t1 = half01T; t3 = half03T;
                p2 = midP; p1 = half01P; p3 = half03P;
                starInPixX -= perc; starInPixY += perc;
            }
        } else {
            if (sgn * b > 0) {
                t1 = half12T; t3 = half23T; t0 = midT;
                p1 = half12P; p3 = half23P; p0 = midP;
                starInPixX += perc; starInPixY -= perc;
            } else {
                t0 = half03T; t1 = midT; t2 = half23T;
                p0 = half03P; p1 = midP; p2 = half23P;
                starInPixX += perc; starInPixY += perc;
            }
        }
        midT = (t0 + t1 + t2 + t3) * .25f;
        midP = (p0 + p1 + p2 + p3) * .25f;
    }
    *start = starInPixY;
    *starp = starInPixX;
}

/* kernel.cuh:335-338 */
static float gaussian(float dist) {
    float sigma = 1.f / 2.5f;
    return f_exp(-.5f * dist / (sigma * sigma)) * 1.f / (sigma * SQRT2PI);
}

float btor_gaussian(float dist) { return gaussian(dist); }
void btor_interpolate_star(const float* t, const float* p, float st, float sp, int sgn, float* yx) {
    interpolate_star(t[0], t[1], t[2], t[3], p[0], p[1], p[2], p[3], &st, &sp, sgn);
    yx[0] = st;
    yx[1] = sp;
}
int btor_pi_check(float* p, float factor) { return pi_check(p, factor); }
void btor_hermite(float v, const float* x4, float* out) {
    f2 x0 = {x4[0], x4[1]}, x1 = {x4[2], x4[3]}, x2 = {x4[4], x4[5]}, x3 = {x4[6], x4[7]};
    f2 r = hermite_f(v, x0, x1, x2, x3);
    out[0] = r.x;
    out[1] = r.y;
}
void btor_intersection(const float* abcd, float* out) {
    f2 r = intersection(abcd[0], abcd[1], abcd[2], abcd[3], abcd[4], abcd[5], abcd[6], abcd[7]);
    out[0] = r.x;
    out[1] = r.y;
}
void btor_wrap_to_pi(float* t, float* p) { wrap_to_pi_f(t, p); }

/* MakeImageKernel.cu:27-96, searching the implicit star tree for leaves
 * (or nearly covered inner nodes) that overlap the pixel's bounding box.
 * Nodes are appended to out (unbounded); returns their count. */
static int search_tree(const int32_t* tree, const float* mn, const float* mx, int treeLevel, int* out,
                       int cap, int picheck) {
    float nodeStart[2] = {0.f, 0.f + picheck * FPI};
    float nodeSize[2] = {FPI, FPI2};
    int node = 0;
    unsigned bitMask = (unsigned)powf(2, treeLevel);
    int level = 0;
    int lvl = 0;
    int pos = 0;
    while (bitMask != 0) {
        bitMask &= ~(1UL << (treeLevel - level));
        for (lvl = level + 1; lvl <= treeLevel; lvl++) {
            int star_n = tree[node];
            if (node != 0 && ((node + 1) & node) != 0) star_n -= tree[node - 1];
            int tp = lvl & 1;
            float x_overlap = fmaxf(0.f, fminf(mx[0], nodeStart[0] + nodeSize[0]) - fmaxf(mn[0], nodeStart[0]));
            float y_overlap = fmaxf(0.f, fminf(mx[1], nodeStart[1] + nodeSize[1]) - fmaxf(mn[1], nodeStart[1]));
            float overlapArea = x_overlap * y_overlap;
            int size = overlapArea / (nodeSize[0] * nodeSize[1]) > 0.8f;
            nodeSize[tp] = nodeSize[tp] * .5f;
            if (star_n == 0) {
                node = node * 2 + 1;
                break;
            }
            float check = nodeStart[tp] + nodeSize[tp];
            int lu = mn[tp] < check;
            int rd = mx[tp] >= check;
            if (lvl == 1 && picheck) {
                int tmp = lu;
                lu = rd;
                rd = tmp;
            }
            if (lvl == treeLevel || (rd && lu && size)) {
                if (rd) {
                    if (pos < cap) out[pos] = node * 2 + 2;
                    pos++;
                }
                if (lu) {
                    if (pos < cap) out[pos] = node * 2 + 1;
                    pos++;
                }
                node = node * 2 + 1;
                break;
            } else {
                node = node * 2 + 1;
                if (rd) bitMask |= 1UL << (treeLevel - lvl);
                if (!lu) break;
                else if (lvl == 1 && picheck) nodeStart[1] += nodeSize[1];
            }
        }
        level = treeLevel - __builtin_ffs((int)bitMask) + 1;
        if (level >= 0) {
            int diff = lvl - level;
            for (int i = 0; i < diff; i++) {
                int tp = (lvl - i) & 1;
                if (!(node & 1)) nodeStart[tp] -= nodeSize[tp];
                nodeSize[tp] = nodeSize[tp] * 2.f;
                node = (node - 1) / 2;
            }
            node++;
            int tp = level & 1;
            if (picheck && level == 1) nodeStart[tp] -= nodeSize[tp];
            else nodeStart[tp] += nodeSize[tp];
        }
    }
    return pos;
}
int btor_search_tree(const int32_t* tree, const float* mn, const float* mx, int treeLevel, int* out, int cap) {
    return search_tree(tree, mn, mx, treeLevel, out, cap, 0);
}

/* ------------------------------------------------------------ summed table */

/* MakeImageKernel.cu:1312-1331: table[(p * cols + q) * 4 + c] (BGR + area) */
/* StarProcessor::addLowLightStarsToImage (StarProcessor.h:108-158): the
 * stars fainter than the catalogue cut are painted into the sky image (BGR8,
 * rows x cols, in place), each as a 5 x 5 splat (taps within 2.5 pixels)
 * of brightness 400 * 10^(-0.4 (m - 2.5 log10 g(d))) with
 * g(d) = 0.318310 exp(-2 d) of the squared tap distance d, times the star's
 * squared blackbody colour; a tap overwrites what an earlier star left
 * there (catalogue order); then every pixel becomes
 * min(255, (int) sqrt(star + original^2)) per channel.  Divergences on
 * undefined behaviour: the splat buffer starts at zero (the reference's is
 * an uninitialised cv::Mat) and the colour index is clamped to 390 (the
 * reference clamps to 1172 and reads past its 1173-entry table). */
void btor_sky_low_stars(uint8_t* sky_bgr, int rows, int cols, const float* stars4, int n) {
    int* buf = (int*)calloc((size_t)rows * cols * 3, sizeof(int));
    const int step = 2;
    const float maxDistSq = (step + .5f) * (step + .5f);
    for (int s = 0; s < n; s++) {
        const float thetastar = stars4[4 * s], phistar = stars4[4 * s + 1];
        const float magnitude = stars4[4 * s + 2], bvcolor = stars4[4 * s + 3];
        /* the float angle over the double PI: evaluated in double, stored as float */
        const float pixPosTheta = (float)(thetastar / DPI * 1.f * rows);
        const float pixPosPhi = (float)(phistar / DPI2 * 1.f * cols);
        const int pixHor = (int)pixPosPhi, pixVert = (int)pixPosTheta;
        const float temp = 46.f * ((1.f / ((0.92f * bvcolor) + 1.7f)) + (1.f / ((0.92f * bvcolor) + 0.62f))) - 10.f;
        int index = (int)floorf(temp);
        index = index < 0 ? 0 : index > 390 ? 390 : index;
        const float rx = (float)(kTempRGB[3 * index] * kTempRGB[3 * index]);
        const float ry = (float)(kTempRGB[3 * index + 1] * kTempRGB[3 * index + 1]);
        const float rz = (float)(kTempRGB[3 * index + 2] * kTempRGB[3 * index + 2]);
        const int start = 0 > -(pixVert - step) ? 0 : -(pixVert - step);
        const int diff = pixVert + step - rows + 1;
        const int stop = (pixVert + step >= rows) ? 2 * step - diff : 2 * step;
        for (int u = start; u <= stop; u++) {
            for (int v = 0; v <= 2 * step; v++) {
                const float a = -step + u + .5f - (pixPosTheta - pixVert);
                const float b = -step + v + .5f - (pixPosPhi - pixHor);
                const float dist = a * a + b * b;
                if (dist > maxDistSq) continue;
                const float gauss = f_exp(-2.f * dist) * 0.318310f;
                const float appMag = -.4f * (magnitude - 2.5f * f_log10(gauss));
                const float brightness = 400.f * f_pow(10.f, appMag);
                const size_t o = 3 * ((size_t)(-step + u + pixVert) * cols + (size_t)((-step + v + pixHor + cols) % cols));
                buf[o] = cvt_i32(brightness * rz);
                buf[o + 1] = cvt_i32(brightness * ry);
                buf[o + 2] = cvt_i32(brightness * rx);
            }
        }
    }
    for (size_t q = 0; q < (size_t)rows * cols * 3; q++) {
        const int orig = sky_bgr[q];
        const int v = (int)sqrt((double)(buf[q] + orig * orig));
        sky_bgr[q] = (uint8_t)(v < 255 ? v : 255);
    }
    free(buf);
}

void btor_sum_table(const uint8_t* sky_bgr, int rows, int cols, float* table) {
    float pxsz = FPI2 / (1.f * cols);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int q = 0; q < cols; q++) {
        float prev[4] = {0.f, 0.f, 0.f, 0.f};
        float phi[4] = {pxsz * q, pxsz * q, pxsz * (q + 1), pxsz * (q + 1)};
        for (int p = 0; p < rows; p++) {
            float theta[4] = {pxsz * (p + 1), pxsz * p, pxsz * p, pxsz * (p + 1)};
            float area = calc_area4(theta, phi);
            const uint8_t* px = sky_bgr + 3 * ((size_t)p * cols + q);
            prev[0] += f_pow((float)px[0], 2.2f) * area;
            prev[1] += f_pow((float)px[1], 2.2f) * area;
            prev[2] += f_pow((float)px[2], 2.2f) * area;
            prev[3] += area;
            float* o = table + 4 * ((size_t)p * cols + q);
            o[0] = prev[0];
            o[1] = prev[1];
            o[2] = prev[2];
            o[3] = prev[3];
        }
    }
}

/* ------------------------------------------------------------ dense grid from the PSH */

/* MakeImageKernel.cu:369-384 with hashLookup (kernel.cuh:705-733), G == 1 */
void btor_make_grid(const float* hash_table, const int32_t* offset_table, const int32_t* hash_pos_tag,
                    int hw, int ow, int GN, int GM, float* grid /* GN1 * GM * 2 */) {
    int sym = (float)GM / (float)GN > 3 ? 1 : 0;
    int GN1 = sym ? 2 * GN - 1 : GN;
    for (int i = 0; i < GN; i++)
        for (int j = 0; j < GM; j++) {
            int ix = (i + ow) % ow, iy = (j + ow) % ow;
            int ax = (i + hw) % hw + offset_table[2 * (ix * ow + iy)];
            int ay = (j + hw) % hw + offset_table[2 * (ix * ow + iy) + 1];
            int hx = (ax + hw) % hw, hy = (ay + hw) % hw;
            int s = hx * hw + hy;
            float vx = -2.f, vy = -2.f;
            if (hash_pos_tag[2 * s] == i && hash_pos_tag[2 * s + 1] == j) {
                vx = hash_table[2 * s];
                vy = hash_table[2 * s + 1];
            }
            grid[2 * ((size_t)i * GM + j)] = vx;
            grid[2 * ((size_t)i * GM + j) + 1] = vy;
            if (sym == 1) {
                if (vx != -1 && vx != -2) vx = FPI - vx;
                int x = GN1 - i - 1;
                grid[2 * ((size_t)x * GM + j)] = vx;
                grid[2 * ((size_t)x * GM + j) + 1] = vy;
            }
        }
}

/* ------------------------------------------------------------ the frame */

/* ------------------------------------------------------------ multi-grid blending */

#define ANGLE_NUM 1000 /* MakeImageKernel.cu:1257 */

/* findBhCenter (MakeImageKernel.cu:391-404), findBhBorders (:406-431) and
 * smoothBorder x 4 (:444-456, called at :1439-1442) for grids A and B of
 * GN x GM (GN: rows of the whole-sky grid).  border: (2 + 2*ANGLE_NUM) float2. */
void btor_blend_border(const float* grid_a, const float* grid_b, int GN, int GM, float* border) {
    gridv ga = {grid_a, GN, GM}, gb = {grid_b, GN, GM};
    float b0x = 100.f, b0y = 0.f, b1x = 100.f, b1y = 0.f; /* :1308-1310 */
    for (int i = 0; i < GN; i++)
        for (int j = 0; j < GM; j++)
            if (grid_at(&ga, i, j).x == -1 || grid_at(&gb, i, j).x == -1) {
                float gridsize = FPI / (1.f * GN);
                float ti = gridsize * (float)i, tj = gridsize * (float)j;
                if (ti < b0x) b0x = ti;
                if (ti > b0y) b0y = ti;
                if (tj < b1x) b1x = tj;
                if (tj > b1y) b1y = tj;
            }
    float* cur = border;
    float* tmp = (float*)malloc(sizeof(float) * 2 * (2 + 2 * ANGLE_NUM));
    cur[0] = b0x;
    cur[1] = b0y;
    cur[2] = b1x;
    cur[3] = b1y;
    for (int i = 0; i < 2 * ANGLE_NUM; i++) {
        int ii = i / 2;
        float angle = FPI2 / (1.f * ANGLE_NUM) * 1.f * ii;
        float thetaChange = -f_sin(angle);
        float phiChange = f_cos(angle);
        float ptx = .5f * b0x + .5f * b0y, pty = .5f * b1x + .5f * b1y;
        ptx = ptx / FPI2 * GM;
        pty = pty / FPI2 * GM;
        const gridv* g = (i % 2) ? &gb : &ga;
        float gA = -2.f, gB = -2.f;
        while (!(gA > 0 && gB == -1)) {
            gB = gA;
            ptx += thetaChange;
            pty += phiChange;
            int gx = cvt_i32(ptx), gy = cvt_i32(pty);
            if (gx < 0 || gx >= GN || gy < 0 || gy >= GM) break; /* the walk left the grid */
            gA = grid_at(g, gx, gy).x;
        }
        cur[2 * (2 + i)] = (ptx - thetaChange) * FPI2 / (1.f * GM);
        cur[2 * (2 + i) + 1] = (pty - phiChange) * FPI2 / (1.f * GM);
    }
    for (int pass = 0; pass < 4; pass++) {
        const float* src = (pass % 2 == 0) ? border : tmp;
        float* dst = (pass % 2 == 0) ? tmp : border;
        dst[0] = src[0];
        dst[1] = src[1];
        dst[2] = src[2];
        dst[3] = src[3];
        for (int i = 0; i < 2 * ANGLE_NUM; i++) {
            int prev = (i - 2 + 2 * ANGLE_NUM) % (2 * ANGLE_NUM);
            int next = (i + 2) % (2 * ANGLE_NUM);
            for (int d = 0; d < 2; d++)
                dst[2 * (i + 2) + d] =
                    1.f / 3.f * (src[2 * (prev + 2) + d] + src[2 * (i + 2) + d] + src[2 * (next + 2) + d]);
        }
    }
    free(tmp);
}

/* pixInterpolation, Gr > 1 (MakeImageKernel.cu:466-505), one corner */
static f2 blend_corner(const gridv* ga, const gridv* gb, int gridlvl, const float* border, float alpha,
                       float theta, float phi, int* gs) {
    float cx = .5f * border[0] + .5f * border[1], cy = .5f * border[2] + .5f * border[3];
    float stretchRad = fmaxf(border[1] - border[0], border[2] - border[3]) * 0.75f;
    float centerdist = (theta - cx) * (theta - cx) + (phi - cy) * (phi - cy);
    f2 A, B, r;
    *gs = 0;
    if (centerdist < stretchRad * stretchRad) {
        float angle = f_atan2(cx - theta, phi - cy);
        angle = fmodf(angle + FPI2, FPI2);
        int slot = cvt_i32(angle / FPI2 * ANGLE_NUM);
        if (slot < 0) slot = 0;
        if (slot > ANGLE_NUM - 1) slot = ANGLE_NUM - 1;
        float ax = border[2 * (2 * slot + 2)], ay = border[2 * (2 * slot + 2) + 1];
        float bx = border[2 * (2 * slot + 3)], by = border[2 * (2 * slot + 3) + 1];
        float nx = (1.f - alpha) * ax + alpha * bx;
        float ny = (1.f - alpha) * ay + alpha * by;
        if (centerdist <= (nx - cx) * (nx - cx) + (ny - cy) * (ny - cy)) {
            r.x = -1.f;
            r.y = -1.f;
            return r;
        }
        float tStoB = (cx - stretchRad * f_sin(angle) - nx);
        float pStoB = (cy + stretchRad * f_cos(angle) - ny);
        float thetaPerc = fabsf(tStoB) < 1E-5 ? 0 : 1.f - (theta - nx) / tStoB;
        float phiPerc = fabsf(pStoB) < 1E-5 ? 0 : 1.f - (phi - ny) / pStoB;
        float thetaA = theta - thetaPerc * (nx - ax);
        float phiA = phi - phiPerc * (ny - ay);
        float thetaB = theta - thetaPerc * (nx - bx);
        float phiB = phi - phiPerc * (ny - by);
        A = interpolate_pix(ga, thetaA, phiA, gridlvl, gs);
        B = interpolate_pix(gb, thetaB, phiB, gridlvl, gs);
    } else {
        A = interpolate_pix(ga, theta, phi, gridlvl, gs);
        B = interpolate_pix(gb, theta, phi, gridlvl, gs);
    }
    if (A.x == -1 || B.x == -1) {
        r.x = -1.f;
        r.y = -1.f;
        return r;
    }
    if (A.y < .2f * FPI2 && B.y > .8f * FPI2) A.y += FPI2;
    if (B.y < .2f * FPI2 && A.y > .8f * FPI2) B.y += FPI2;
    r.x = (1.f - alpha) * A.x + alpha * B.x;
    r.y = fmodf((1.f - alpha) * A.y + alpha * B.y, FPI2);
    return r;
}

/* ------------------------------------------------------------ trails and diffraction */

#define TRAIL_NUM 30  /* MakeImageKernel.cu:1255 */
#define TRAIL_HITS 32 /* images recorded per star and frame (the device's bound) */

struct btor_trails {
    int S, W, H, fs;
    long frame;
    int* cache; /* [2][S][TRAIL_NUM][2] (row, column) */
    int* count; /* [2][S] */
    uint8_t* sprite;
};

btor_trails* btor_trails_new(int S, int W, int H, const uint8_t* sprite) {
    btor_trails* t = (btor_trails*)calloc(1, sizeof(btor_trails));
    t->S = S;
    t->W = W;
    t->H = H;
    t->fs = W / 16;
    t->cache = (int*)calloc((size_t)2 * (S > 0 ? S : 1) * TRAIL_NUM * 2, sizeof(int));
    t->count = (int*)calloc((size_t)2 * (S > 0 ? S : 1), sizeof(int));
    if (sprite) {
        t->sprite = (uint8_t*)malloc((size_t)3 * t->fs * t->fs);
        memcpy(t->sprite, sprite, (size_t)3 * t->fs * t->fs);
    }
    return t;
}

void btor_trails_free(btor_trails* t) {
    if (!t) return;
    free(t->cache);
    free(t->count);
    free(t->sprite);
    free(t);
}

typedef struct {
    int q;
    uint32_t pix;
    float part, frac;
    int index;
} trail_hit;

static int hit_cmp(const void* a, const void* b) {
    const trail_hit* x = (const trail_hit*)a;
    const trail_hit* y = (const trail_hit*)b;
    if (x->q != y->q) return x->q < y->q ? -1 : 1;
    return x->pix < y->pix ? -1 : (x->pix > y->pix ? 1 : 0);
}

static void trail_put(float* trail, size_t ij, float r2, float g2, float b2) {
    float v[3] = {r2, g2, b2};
    for (int c = 0; c < 3; c++)
        if (v[c] >= 0.f && v[c] > trail[3 * ij + c]) trail[3 * ij + c] = v[c];
}

/* makeGradField (MakeImageKernel.cu:517-537) */
static void trail_grad(const float* thphi, int W, int H, float* grad) {
    const int M1 = W + 1;
    const size_t NC = (size_t)M1 * (H + 1);
    for (int i = 0; i <= H; i++)
        for (int j = 0; j <= W; j++) {
#define TH(c) ((size_t)(c) < NC ? thphi[2 * (size_t)(c)] : -1.f)
            float up = TH((size_t)(i + 1 < H ? i + 1 : H) * M1 + j);
            float down = TH((size_t)(i - 1 > 0 ? i - 1 : 0) * M1 + j);
            float left = TH((size_t)i * M1 + (j - 1 > 0 ? j - 1 : 0));
            float right = TH((size_t)i * M1 + (j + 1 < M1 ? j + 1 : M1));
            float mid = TH((size_t)i * M1 + j);
#undef TH
            float gx = 0.f, gy = 0.f;
            if (mid > 0) {
                float xdir = 0;
                if (up > 0 && down > 0) xdir = .5f * (up - mid) - .5f * (down - mid);
                float ydir = 0;
                if (left > 0 && right > 0) ydir = .5f * (right - mid) - .5f * (left - mid);
                gx = -ydir;
                gy = xdir;
            }
            grad[2 * ((size_t)i * M1 + j)] = gx;
            grad[2 * ((size_t)i * M1 + j) + 1] = gy;
        }
}

/* addTrails (MakeImageKernel.cu:106-172) for one image */
static void trail_line(const btor_trails* T, int prev_half, const trail_hit* h, const float* grad, float* trail) {
    const int W = T->W, H = T->H, M = W, M1 = W + 1;
    const int i = (int)(h->pix / (uint32_t)W), j = (int)(h->pix % (uint32_t)W);
    const int pc = T->count[(size_t)prev_half * T->S + h->q];
    const int* pcache = T->cache + ((size_t)prev_half * T->S + h->q) * TRAIL_NUM * 2;
    float traildist = (float)M, angle = FPI2;
    int px = 0, py = 0, line = 0;
    for (int w = 0; w < pc; w++) {
        int prx = pcache[2 * w], pry = pcache[2 * w + 1];
        int dx = prx - i, dy = pry - j;
        int dxx = dx * dx, dyy = dy * dy;
        if (dxx <= 1 && dyy <= 1) {
            line = 0;
            break;
        }
        float grx = grad[2 * ((size_t)prx * M1 + pry)], gry = grad[2 * ((size_t)prx * M1 + pry) + 1];
        float dist = sqrtf((float)(dxx + dyy));
        if (dist > M / 25.f) continue;
        float div = (dist * sqrtf(grx * grx + gry * gry));
        float a1 = f_acos((1.f * dx * grx + 1.f * dy * gry) / div);
        float a2 = f_acos((1.f * dx * -grx + 1.f * dy * -gry) / div);
        float a = fminf(a1, a2);
        if (a > angle || a > FPI * .25f) continue;
        if (a < angle) {
            angle = a;
            traildist = dist;
            px = prx;
            py = pry;
            line = 1;
        }
    }
    if (!line) return;
    int deltax = i - px, deltay = j - py;
    int sgnX = deltax < 0 ? -1 : 1, sgnY = deltay < 0 ? -1 : 1;
    float deltaerr = deltay == 0 ? fabsf((float)deltax) : fabsf(deltax / (1.f * deltay));
    float error = 0.f;
    int y = py, x = px;
    float appMag = h->part - 2.5f * f_log10(h->frac);
    float bright0 = 100 * f_exp10(-.4f * appMag);
    float r0 = (float)kTempRGB[3 * h->index], g0 = (float)kTempRGB[3 * h->index + 1],
          b0 = (float)kTempRGB[3 * h->index + 2];
    float hh, ss, pp;
    rgb_to_hsp(r0 / 255.f, g0 / 255.f, b0 / 255.f, &hh, &ss, &pp);
    int maxit = 2 * (abs(deltax) + abs(deltay)) + 2;
    for (int it = 0; it < maxit && (y != j || x != i); it++) {
        if (error < 1.f) {
            y += sgnY;
            error += deltaerr;
        }
        if (error >= 0.5f) {
            x += sgnX;
            error -= 1.f;
        }
        if (x < 0 || x >= H || y < 0 || y >= W) break;
        float dist = (float)((x - i) * (x - i) + (y - j) * (y - j));
        float fade = (traildist - sqrtf(dist)) * (traildist - sqrtf(dist)) / (traildist * traildist);
        float brightness = bright0 * fade;
        float r, g, b;
        hsp_to_rgb(hh, ss, pp * brightness, &r, &g, &b);
        r *= 255.f;
        g *= 255.f;
        b *= 255.f;
        trail_put(trail, (size_t)x * M + y, r * r, g * g, b * b);
        if (dist <= 1.f) break;
    }
}

static int frame_impl(const float* grid, const float* grid_b, float alpha, const float* border, int GN, int GM,
                      int gridlvl, const float* view, int W, int H, const float* cam, const int32_t* tree,
                      int tree_level, const float* star_pos, const float* star_mag, int S,
                      const float* sum_table, int sky_rows, int sky_cols, int star, int map, int lr, float hor,
                      float ver, float offset, int step, uint8_t* out_bgra, float* o_thphi, int32_t* o_gap,
                      uint8_t* o_bh, float* o_area, float* o_star, btor_trails* tr, float* o_trail);

int btor_frame_trails(btor_trails* t, const float* grid, int GN, int GM, int gridlvl, const float* view,
                      int W, int H, const float* cam, const int32_t* tree, int tree_level,
                      const float* star_pos, const float* star_mag, int S, const float* sum_table, int sky_rows,
                      int sky_cols, int star, int map, int lr, float hor, float ver, float offset, int step,
                      uint8_t* out_bgra, float* o_star, float* o_trail) {
    if (!t || t->W != W || t->H != H || t->S != S) return -1;
    int rc = frame_impl(grid, NULL, 0.f, NULL, GN, GM, gridlvl, view, W, H, cam, tree, tree_level, star_pos,
                        star_mag, S, sum_table, sky_rows, sky_cols, star, map, lr, hor, ver, offset, step, out_bgra,
                        NULL, NULL, NULL, NULL, o_star, t, o_trail);
    t->frame++;
    return rc;
}

int btor_frame(const float* grid, int GN, int GM, int gridlvl, const float* view, int W, int H,
               const float* cam, const int32_t* tree, int tree_level, const float* star_pos,
               const float* star_mag, int S, const float* sum_table, int sky_rows, int sky_cols,
               int star, int map, int lr, float hor, float ver, float offset, int step,
               uint8_t* out_bgra, float* o_thphi, int32_t* o_gap, uint8_t* o_bh, float* o_area,
               float* o_star) {
    return frame_impl(grid, NULL, 0.f, NULL, GN, GM, gridlvl, view, W, H, cam, tree, tree_level, star_pos,
                      star_mag, S, sum_table, sky_rows, sky_cols, star, map, lr, hor, ver, offset, step, out_bgra,
                      o_thphi, o_gap, o_bh, o_area, o_star, NULL, NULL);
}

int btor_frame_blend(const float* grid_a, const float* grid_b, int GN, int GM, int gridlvl, float alpha,
                     const float* view, int W, int H, const float* cam14, const int32_t* tree, int tree_level,
                     const float* star_pos, const float* star_mag, int S, const float* sum_table, int sky_rows,
                     int sky_cols, int star, int map, int lr, float hor, float ver, float offset, int step,
                     uint8_t* out_bgra, float* o_thphi, float* o_border) {
    float* border = o_border ? o_border : (float*)malloc(sizeof(float) * 2 * (2 + 2 * ANGLE_NUM));
    btor_blend_border(grid_a, grid_b, GN, GM, border);
    float cam[7];
    for (int q = 0; q < 7; q++) cam[q] = (1.f - alpha) * cam14[q] + alpha * cam14[7 + q]; /* camUpdate */
    int rc = frame_impl(grid_a, grid_b, alpha, border, GN, GM, gridlvl, view, W, H, cam, tree, tree_level,
                        star_pos, star_mag, S, sum_table, sky_rows, sky_cols, star, map, lr, hor, ver, offset,
                        step, out_bgra, o_thphi, NULL, NULL, NULL, NULL, NULL, NULL);
    if (!o_border) free(border);
    return rc;
}

static int frame_impl(const float* grid, const float* grid_b, float alpha, const float* border, int GN, int GM,
                      int gridlvl, const float* view, int W, int H, const float* cam, const int32_t* tree,
                      int tree_level, const float* star_pos, const float* star_mag, int S,
                      const float* sum_table, int sky_rows, int sky_cols, int star, int map, int lr, float hor,
                      float ver, float offset, int step, uint8_t* out_bgra, float* o_thphi, int32_t* o_gap,
                      uint8_t* o_bh, float* o_area, float* o_star, btor_trails* tr, float* o_trail) {
    const int M = W, N = H, M1 = W + 1, N1 = H + 1;
    const size_t NP = (size_t)N * M, NC = (size_t)N1 * M1;
    gridv g = {grid, GN, GM};
    gridv g2 = {grid_b, GN, GM};
    float* thphi = o_thphi ? o_thphi : (float*)malloc(sizeof(float) * 2 * NC);
    int32_t* gap = o_gap ? o_gap : (int32_t*)malloc(sizeof(int32_t) * NC);
    uint8_t* bh = o_bh ? o_bh : (uint8_t*)malloc(NP);
    float* area = o_area ? o_area : (float*)malloc(sizeof(float) * NP);
    float* areaSmooth = (float*)malloc(sizeof(float) * NP);
    (void)S;

    /* pixInterpolation (MakeImageKernel.cu:458-515), G == 1 or (grid_b) Gr > 1 */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8)
#endif
    for (int i = 0; i < N1; i++)
        for (int j = 0; j < M1; j++) {
            size_t c = (size_t)i * M1 + j;
            float theta = view[2 * c] + ver;
            float phi = fmodf(view[2 * c + 1] + hor + FPI2, FPI2);
            int gs = 0;
            f2 r = grid_b ? blend_corner(&g, &g2, gridlvl, border, alpha, theta, phi, &gs)
                          : interpolate_pix(&g, theta, phi, gridlvl, &gs);
            thphi[2 * c] = r.x;
            thphi[2 * c + 1] = r.y;
            gap[c] = gs;
        }

    /* findBlackPixels + findArea (MakeImageKernel.cu:539-568) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < N; i++)
        for (int j = 0; j < M; j++) {
            float t[4], p[4];
            int pc = 0;
            int ok = retrieve_corners(thphi, W, i * M1 + j, t, p, &pc, 0.0f);
            bh[(size_t)i * M + j] = ok ? 0 : 1;
            float a = 0.f;
            if (ok) {
                float th1[3] = {t[0], t[1], t[2]}, ph1[3] = {p[0], p[1], p[2]};
                float th2[3] = {t[0], t[2], t[3]}, ph2[3] = {p[0], p[2], p[3]};
                a = calc_area3(th1, ph1) + calc_area3(th2, ph2);
            }
            area[(size_t)i * M + j] = a;
        }

    /* smoothAreaH / smoothAreaV (MakeImageKernel.cu:570-616) */
    smooth_area(area, bh, gap, W, H, areaSmooth);

    uint8_t* img_star = NULL;
    uint8_t* img_map = NULL;

    if (star) {
        const int fw = 2 * step + 1;
        float* temp = (float*)calloc(NP * fw * fw * 3, sizeof(float));
        float maxDistSq = (step + .5f) * (step + .5f);
        /* distortStarMap (MakeImageKernel.cu:206-326) */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
        for (int i = 0; i < N; i++) {
            int* nodes = (int*)malloc(sizeof(int) * (1 << (tree_level + 1)));
            for (int j = 0; j < M; j++) {
                size_t ij = (size_t)i * M + j;
                if (bh[ij] != 0) continue;
                float t[4], p[4];
                int picheck = 0;
                int ind = i * M1 + j;
                retrieve_corners(thphi, W, ind, t, p, &picheck, offset);
                float mxv[2] = {fmaxf(fmaxf(t[0], t[1]), fmaxf(t[2], t[3])),
                                fmaxf(fmaxf(p[0], p[1]), fmaxf(p[2], p[3]))};
                float mnv[2] = {fminf(fminf(t[0], t[1]), fminf(t[2], t[3])),
                                fminf(fminf(p[0], p[1]), fminf(p[2], p[3]))};
                int pos = search_tree(tree, mnv, mxv, tree_level, nodes, 1 << (tree_level + 1), 0);
                if (pos == 0) continue;
                float orient = (t[1] - t[0]) * (p[1] + p[0]) + (t[2] - t[1]) * (p[2] + p[1]) +
                               (t[3] - t[2]) * (p[3] + p[2]) + (t[0] - t[3]) * (p[0] + p[3]);
                int sgn = orient < 0 ? -1 : 1;
                float redshft = 1.f, frac = 1.f;
                if (lr) lensing_redshift(t, p, W, ind, cam, view, &frac, &redshft, area[ij]);
                float red = 4.f * f_log10(redshft);
                float* sl = temp + 3 * (size_t)fw * fw * ij;
                for (int s = 0; s < pos; s++) {
                    int node = nodes[s];
                    int startN = 0;
                    if (node != 0 && ((node + 1) & node) != 0) startN = tree[node - 1];
                    for (int q = startN; q < tree[node]; q++) {
                        float st = star_pos[2 * q];
                        float sp = star_pos[2 * q + 1];
                        int in = star_in_polygon(t, p, st, sp, sgn);
                        if (picheck && !in && sp < FPI2 * .2f) {
                            sp += FPI2;
                            in = star_in_polygon(t, p, st, sp, sgn);
                        }
                        if (!in) continue;
                        interpolate_star(t[0], t[1], t[2], t[3], p[0], p[1], p[2], p[3], &st, &sp, sgn);
                        float part = star_mag[2 * q] + red;
                        float bv = star_mag[2 * q + 1];
                        float temp_k = 46.f / redshft * ((1.f / ((0.92f * bv) + 1.7f)) +
                                                         (1.f / ((0.92f * bv) + 0.62f))) - 10.f;
                        /* the reference clamps to 1170 (= 3 * 390, an element index)
                         * and reads past its 391-entry table; clamp to entry 390 */
                        int index = cvt_i32(floorf(temp_k));
                        if (index > 390) index = 390;
                        if (index < 0) index = 0;
                        float r0 = (float)kTempRGB[3 * index], g0 = (float)kTempRGB[3 * index + 1],
                              b0 = (float)kTempRGB[3 * index + 2];
                        float hh, ss, pp;
                        rgb_to_hsp(r0 / 255.f, g0 / 255.f, b0 / 255.f, &hh, &ss, &pp);
                        for (int u = 0; u <= 2 * step; u++)
                            for (int v = 0; v <= 2 * step; v++) {
                                float dist = (-step + u + .5f - st) * (-step + u + .5f - st) +
                                             (-step + v + .5f - sp) * (-step + v + .5f - sp);
                                if (dist > maxDistSq) continue;
                                float appMag = part - 2.5f * f_log10(frac * gaussian(dist));
                                float brightness = 100 * f_exp10(-.4f * appMag);
                                float rr, gg, bb;
                                hsp_to_rgb(hh, ss, pp * brightness, &rr, &gg, &bb);
                                rr *= 255.f;
                                gg *= 255.f;
                                bb *= 255.f;
                                float* o = sl + 3 * (fw * u + v);
                                o[0] += rr * rr;
                                o[1] += gg * gg;
                                o[2] += bb * bb;
                            }
                    }
                }
            }
            free(nodes);
        }
        float* trail = NULL;
        if (tr) {
            /* clearArrays (MakeImageKernel.cu:618-629) */
            const int cur = (int)(tr->frame % 2);
            for (int q = 0; q < S; q++) {
                tr->count[(size_t)cur * S + q] = 0;
                if (tr->frame == 0) tr->count[(size_t)(1 - cur) * S + q] = 0;
            }
            float* grad = (float*)malloc(sizeof(float) * 2 * NC);
            trail_grad(thphi, W, H, grad);
            trail = (float*)calloc(NP * 3, sizeof(float));
            /* the frame's images (MakeImageKernel.cu:273-303) */
            size_t nh = 0, caph = 1024;
            trail_hit* hits = (trail_hit*)malloc(sizeof(trail_hit) * caph);
            int* nodes = (int*)malloc(sizeof(int) * (1 << (tree_level + 1)));
            for (int i = 0; i < N; i++)
                for (int j = 0; j < M; j++) {
                    size_t ij = (size_t)i * M + j;
                    if (bh[ij] != 0) continue;
                    float t[4], p[4];
                    int picheck = 0;
                    int ind = i * M1 + j;
                    retrieve_corners(thphi, W, ind, t, p, &picheck, offset);
                    float mxv[2] = {fmaxf(fmaxf(t[0], t[1]), fmaxf(t[2], t[3])),
                                    fmaxf(fmaxf(p[0], p[1]), fmaxf(p[2], p[3]))};
                    float mnv[2] = {fminf(fminf(t[0], t[1]), fminf(t[2], t[3])),
                                    fminf(fminf(p[0], p[1]), fminf(p[2], p[3]))};
                    int pos = search_tree(tree, mnv, mxv, tree_level, nodes, 1 << (tree_level + 1), 0);
                    int starsToCheck = 0;
                    for (int s2 = 0; s2 < pos; s2++) {
                        int node = nodes[s2];
                        int startN = 0;
                        if (node != 0 && ((node + 1) & node) != 0) startN = tree[node - 1];
                        starsToCheck += tree[node] - startN;
                    }
                    if (!(starsToCheck < S / 100)) continue;
                    float orient = (t[1] - t[0]) * (p[1] + p[0]) + (t[2] - t[1]) * (p[2] + p[1]) +
                                   (t[3] - t[2]) * (p[3] + p[2]) + (t[0] - t[3]) * (p[0] + p[3]);
                    int sgn = orient < 0 ? -1 : 1;
                    float redshft = 1.f, frac = 1.f;
                    if (lr) lensing_redshift(t, p, W, ind, cam, view, &frac, &redshft, area[ij]);
                    float red = 4.f * f_log10(redshft);
                    for (int s2 = 0; s2 < pos; s2++) {
                        int node = nodes[s2];
                        int startN = 0;
                        if (node != 0 && ((node + 1) & node) != 0) startN = tree[node - 1];
                        for (int q = startN; q < tree[node]; q++) {
                            float st = star_pos[2 * q], sp = star_pos[2 * q + 1];
                            int in = star_in_polygon(t, p, st, sp, sgn);
                            if (picheck && !in && sp < FPI2 * .2f) {
                                sp += FPI2;
                                in = star_in_polygon(t, p, st, sp, sgn);
                            }
                            if (!in) continue;
                            float bv = star_mag[2 * q + 1];
                            float temp_k = 46.f / redshft * ((1.f / ((0.92f * bv) + 1.7f)) +
                                                             (1.f / ((0.92f * bv) + 0.62f))) - 10.f;
                            int index = cvt_i32(floorf(temp_k));
                            if (index > 390) index = 390;
                            if (index < 0) index = 0;
                            if (nh == caph) {
                                caph *= 2;
                                hits = (trail_hit*)realloc(hits, sizeof(trail_hit) * caph);
                            }
                            trail_hit h = {q, (uint32_t)ij, star_mag[2 * q] + red, frac, index};
                            hits[nh++] = h;
                        }
                    }
                }
            free(nodes);
            qsort(hits, nh, sizeof(trail_hit), hit_cmp);
            /* the first TRAIL_HITS images of each star are recorded; the first
             * TRAIL_NUM become this frame's cache */
            size_t a = 0;
            while (a < nh) {
                size_t b = a;
                while (b < nh && hits[b].q == hits[a].q) b++;
                const int q = hits[a].q;
                const size_t n = b - a < TRAIL_HITS ? b - a : TRAIL_HITS;
                const size_t keep = n < TRAIL_NUM ? n : TRAIL_NUM;
                int* c = tr->cache + ((size_t)cur * S + q) * TRAIL_NUM * 2;
                for (size_t k = 0; k < keep; k++) {
                    c[2 * k] = (int)(hits[a + k].pix / (uint32_t)W);
                    c[2 * k + 1] = (int)(hits[a + k].pix % (uint32_t)W);
                }
                tr->count[(size_t)cur * S + q] = (int)keep;
                for (size_t k = 0; k < n; k++) trail_line(tr, 1 - cur, &hits[a + k], grad, trail);
                a = b;
            }
            free(hits);
            free(grad);
        }
        /* sumStarLight (MakeImageKernel.cu:753-773) and makePix (:802-835) */
        img_star = (uint8_t*)malloc(NP * 4);
        float* light = (tr && tr->sprite) ? (float*)malloc(sizeof(float) * 3 * NP) : NULL;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
        for (int i = 0; i < N; i++)
            for (int j = 0; j < M; j++) {
                float br = 0.f, bg = 0.f, bb = 0.f;
                int start = step - i > 0 ? step - i : 0;
                int stop = 2 * step < step + N - i - 1 ? 2 * step : step + N - i - 1;
                float factor = 20.f;
                for (int u = start; u <= stop; u++)
                    for (int v = 0; v <= 2 * step; v++) {
                        size_t src = (size_t)(i + u - step) * M + ((j + v - step + M) % M);
                        const float* s = temp + 3 * ((size_t)fw * fw * src + fw * fw - (fw * u + v + 1));
                        br += factor * s[0];
                        bg += factor * s[1];
                        bb += factor * s[2];
                    }
                if (trail) {
                    size_t ij = (size_t)i * M + j;
                    br += trail[3 * ij];
                    bg += trail[3 * ij + 1];
                    bb += trail[3 * ij + 2];
                }
                if (light) {
                    size_t ij = (size_t)i * M + j;
                    light[3 * ij] = br;
                    light[3 * ij + 1] = bg;
                    light[3 * ij + 2] = bb;
                    continue;
                }
                if (o_star) {
                    size_t ij = (size_t)i * M + j;
                    o_star[3 * ij] = br;
                    o_star[3 * ij + 1] = bg;
                    o_star[3 * ij + 2] = bb;
                }
                make_pix(br, bg, bb, img_star + 4 * ((size_t)i * M + j));
            }
        free(temp);
        if (light) {
            /* addDiffraction (MakeImageKernel.cu:775-800) from the light before it */
            float* acc = (float*)malloc(sizeof(float) * 3 * NP);
            memcpy(acc, light, sizeof(float) * 3 * NP);
            const int fs = tr->fs, fh = fs / 2;
            for (int i = 0; i < N; i++)
                for (int j = 0; j < M; j++) {
                    size_t ij = (size_t)i * M + j;
                    float lx = light[3 * ij], ly = light[3 * ij + 1], lz = light[3 * ij + 2];
                    float mx = fmaxf(fmaxf(lx, ly), lz);
                    if (!(mx > 65025.f)) continue;
                    int startx = 0, endx = fs;
                    if (i < fh) startx = fh - i;
                    if (i > (N - fh)) endx = N - i + fh;
                    float div = 10000 * f_pow(mx, 0.6f);
                    for (int qq = startx; qq < endx; qq++) {
                        int row = i - fh + qq;
                        if (row < 0 || row >= N) continue;
                        for (int pp = 0; pp < fs; pp++) {
                            const uint8_t* sp = tr->sprite + 3 * ((size_t)qq * fs + pp);
                            float dx = lx / div * (float)sp[0], dy = ly / div * (float)sp[1],
                                  dz = lz / div * (float)sp[2];
                            size_t o = (size_t)M * row + (size_t)(((j - fh + pp) % M + M) % M);
                            acc[3 * o] += fminf(65025.f, dx * dx);
                            acc[3 * o + 1] += fminf(65025.f, dy * dy);
                            acc[3 * o + 2] += fminf(65025.f, dz * dz);
                        }
                    }
                }
            for (size_t ij = 0; ij < NP; ij++) {
                if (o_star) memcpy(o_star + 3 * ij, acc + 3 * ij, sizeof(float) * 3);
                make_pix(acc[3 * ij], acc[3 * ij + 1], acc[3 * ij + 2], img_star + 4 * ij);
            }
            free(acc);
            free(light);
        }
        if (trail) {
            if (o_trail) memcpy(o_trail, trail, sizeof(float) * 3 * NP);
            free(trail);
        }
    }

    if (map) {
        img_map = (uint8_t*)malloc(NP * 4);
        const int rows = sky_rows, cols = sky_cols;
        /* distortEnvironmentMap (MakeImageKernel.cu:631-751) */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
        for (int i = 0; i < N; i++)
            for (int j = 0; j < M; j++) {
                size_t ij = (size_t)i * M + j;
                float color[4] = {0.f, 0.f, 0.f, 0.f};
                if (bh[ij] == 0) {
                    float t[4], p[4];
                    int ind = i * M1 + j;
                    int picheck = 0;
                    retrieve_corners(thphi, W, ind, t, p, &picheck, offset);
                    if (ind > 0) {
                        float pixSize = FPI / (float)rows;
                        float phMax = fmaxf(fmaxf(p[0], p[1]), fmaxf(p[2], p[3]));
                        float phMin = fminf(fminf(p[0], p[1]), fminf(p[2], p[3]));
                        int pixMax = cvt_i32(phMax / pixSize);
                        int pixMin = cvt_i32(phMin / pixSize);
                        int pixNum = pixMax - pixMin + 1;
                        if (pixNum > 1 && pixMax * pixSize > phMax) pixNum -= 1;
                        float thMax = fmaxf(fmaxf(t[0], t[1]), fmaxf(t[2], t[3]));
                        float thMin = fminf(fminf(t[0], t[1]), fminf(t[2], t[3]));
                        if (pixNum == 1) {
                            int rmax = cvt_i32(thMax / pixSize);
                            int rmin = cvt_i32(thMin / pixSize) - 1;
                            if (rmax > rows - 1) rmax = rows - 1;
                            if (rmax < 0) rmax = 0;
                            int col = pixMin < 0 ? 0 : (pixMin > cols - 1 ? cols - 1 : pixMin);
                            const float* mxc = sum_table + 4 * ((size_t)rmax * cols + col);
                            float mnc[4] = {0.f, 0.f, 0.f, 0.f};
                            if (rmin >= 0) {
                                if (rmin > rows - 1) rmin = rows - 1;
                                const float* s = sum_table + 4 * ((size_t)rmin * cols + col);
                                mnc[0] = s[0]; mnc[1] = s[1]; mnc[2] = s[2]; mnc[3] = s[3];
                            }
                            for (int c = 0; c < 4; c++) color[c] += mxc[c] - mnc[c];
                        } else {
                            float max1 = -100, min1 = cols + 1, max2 = -100, min2 = cols + 1;
                            if (pixNum < cols) {
                                for (int s = 0; s <= pixNum; s++) {
                                    float cp = (pixMin + s) * pixSize;
                                    for (int q = 0; q < 4; q++) {
                                        float ap = p[q], bp = p[(q + 1) % 4];
                                        float at = t[q], bt = t[(q + 1) % 4];
                                        if ((cp - ap) * (cp - bp) <= 0) {
                                            float ct = at + (bt - at) / (bp - ap) * (cp - ap);
                                            min2 = ct < min2 ? ct : min2;
                                            max2 = ct > max2 ? ct : max2;
                                        }
                                    }
                                    if (s > 0) {
                                        float max_ = fmaxf(max1, max2);
                                        float min_ = fminf(min1, min2);
                                        int rmax = cvt_i32(max_ / pixSize);
                                        if (rmax > rows - 1) rmax = rows - 1;
                                        if (rmax < 0) rmax = 0;
                                        int col = (pixMin + s) % cols;
                                        const float* mxc = sum_table + 4 * ((size_t)rmax * cols + col);
                                        int index2 = (cvt_i32(min_ / pixSize) - 1) * cols + col;
                                        float mnc[4] = {0.f, 0.f, 0.f, 0.f};
                                        if (index2 > 0) {
                                            if (index2 >= rows * cols) index2 = (rows - 1) * cols + col;
                                            const float* s2 = sum_table + 4 * (size_t)index2;
                                            mnc[0] = s2[0]; mnc[1] = s2[1]; mnc[2] = s2[2]; mnc[3] = s2[3];
                                        }
                                        for (int c = 0; c < 4; c++) color[c] += mxc[c] - mnc[c];
                                    }
                                    max1 = max2;
                                    min1 = min2;
                                }
                            } else {
                                int thMaxPix = cvt_i32(thMax / pixSize);
                                int thMinPix = cvt_i32(thMin / pixSize);
                                if (thMaxPix > rows - 1) thMaxPix = rows - 1;
                                if (thMinPix > rows - 1) thMinPix = rows - 1;
                                if (thMaxPix < 0) thMaxPix = 0;
                                if (thMinPix < 0) thMinPix = 0;
                                thMaxPix *= cols;
                                thMinPix *= cols;
                                for (int q = 0; q < pixNum; q++) {
                                    const float* mxc = sum_table + 4 * (size_t)(thMaxPix + (pixMin + q) % cols);
                                    const float* mnc = sum_table + 4 * (size_t)(thMinPix + (pixMin + q) % cols);
                                    for (int c = 0; c < 4; c++) color[c] += mxc[c] - mnc[c];
                                }
                            }
                        }
                        color[0] = fminf(255.f, f_pow(color[0] / color[3], 1.f / 2.2f));
                        color[1] = fminf(255.f, f_pow(color[1] / color[3], 1.f / 2.2f));
                        color[2] = fminf(255.f, f_pow(color[2] / color[3], 1.f / 2.2f));
                        float Hh, Ss, Pp;
                        rgb_to_hsp(color[2] / 255.f, color[1] / 255.f, color[0] / 255.f, &Hh, &Ss, &Pp);
                        if (lr) {
                            float redshft, frac;
                            lensing_redshift(t, p, W, ind, cam, view, &frac, &redshft, area[ij]);
                            Pp *= frac;
                            Pp = f_pow(Pp, redshft);
                        }
                        hsp_to_rgb(Hh, Ss, fminf(1.f, Pp), &color[2], &color[1], &color[0]);
                    }
                }
                uint8_t* o = img_map + 4 * ij;
                int cx = cvt_i32(color[0] * 255), cy = cvt_i32(color[1] * 255), cz = cvt_i32(color[2] * 255);
                o[0] = (uint8_t)(cx < 255 ? cx : 255);
                o[1] = (uint8_t)(cy < 255 ? cy : 255);
                o[2] = (uint8_t)(cz < 255 ? cz : 255);
                o[3] = 255;
            }
    }

    /* addStarsAndBackground (MakeImageKernel.cu:837-853) */
    for (size_t q = 0; q < NP; q++) {
        uint8_t* o = out_bgra + 4 * q;
        if (star && map) {
            const uint8_t* s = img_star + 4 * q;
            const uint8_t* b = img_map + 4 * q;
            for (int c = 0; c < 3; c++) {
                float sv = (float)s[c] * s[c];
                float bv = (float)b[c] * b[c];
                float p = 1.f;
                int v = cvt_i32(sqrtf(p * sv + (2.f - p) * bv));
                o[c] = (uint8_t)(v < 255 ? v : 255);
            }
            o[3] = 255;
        } else if (map) {
            memcpy(o, img_map + 4 * q, 4);
        } else if (star) {
            memcpy(o, img_star + 4 * q, 4);
        } else {
            o[0] = o[1] = o[2] = 0;
            o[3] = 255;
        }
    }
    free(img_star);
    free(img_map);
    free(areaSmooth);
    if (!o_thphi) free(thphi);
    if (!o_gap) free(gap);
    if (!o_bh) free(bh);
    if (!o_area) free(area);
    return 0;
}
