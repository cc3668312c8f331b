/*
 * bt_oracle.c -- CPU oracle, Path A (adaptive Kerr geodesic grid build).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the CPU
 * baseline timed by bench.py.  Nothing in blacktracer_amd/ links this.
 *
 * A plain-C restatement of the reference, operation for operation, built
 * with -O2 -ffp-contract=off (no FMA contraction, no fast-math) so that the
 * IEEE double results equal the reference compiled with the same libm.
 * Citations are to BlackTracer2/BlackTracer/ of annemiekie/BlackTracer.
 *
 * Pinning (what checks each function besides the HIP comparison):
 *   pose (calc_speed, alpha/w/wbar/...)     oracle/_ref bit-exact; ZAMO
 *                                            circular-orbit closed form
 *   b0/q0/check_bq/check_rup/check_celest   oracle/_ref bit-exact; Bardeen
 *                                            critical curve and photon-orbit
 *                                            radii closed forms
 *   derivs/rkck/rkqs/odeint1/wrap_to_pi     oracle/_ref bit-exact; Cash-Karp
 *                                            5th-order convergence; b,q and
 *                                            null-constraint invariants;
 *                                            Schwarzschild deflection limit
 *   initial_conditions (Grid.h:524-547)     Grid.h cannot be compiled here
 *                                            (needs cuda_runtime.h); pinned by
 *                                            the null constraint H=0 at the
 *                                            camera and by b,q closed forms
 *                                            for axis-aligned directions
 *   grid build / refinement (Grid.h:291-505) brute-force set semantics on
 *                                            small grids (refinement is an
 *                                            order-free set); level-count
 *                                            invariants
 *   fixTvertices (Grid.h:103-177)           pinned order (level asc, key
 *                                            asc); affine/quadratic fields
 *                                            are reproduced exactly.  The
 *                                            reference's own order comes from
 *                                            unordered_map iteration and is
 *                                            compiler-dependent, so post-fix
 *                                            values at order-sensitive
 *                                            T-vertices are "parity unpinned"
 *                                            against the reference itself.
 */
#include "bt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Const.h:2-3 */
static const double PI  = 3.141592653589793238;
static const double PI2 = 3.141592653589793238 * 2.;

/* Metric.h:27-33 */
#define SAFETY   0.9
#define PGROW   -0.2
#define PSHRNK  -0.25
#define ERRCON   1.89e-4
#define TINY     1.0e-30
#define ADAPTIVE 5.0

/* Cash-Karp tableau, Metric.h:41-49 (Numerical Recipes 16.2). */
static const double b21 = 0.2,
    b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.3, b42 = -0.9, b43 = 1.2,
    b51 = -11.0 / 54.0, b52 = 2.5, b53 = -70.0 / 27.0, b54 = 35.0 / 27.0,
    b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0, b63 = 575.0 / 13824.0,
    b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0, c1 = 37.0 / 378.0,
    c3 = 250.0 / 621.0, c4 = 125.0 / 594.0, c6 = 512.0 / 1771.0,
    dc5 = -277.00 / 14336.0;
static const double dc1 = 37.0 / 378.0 - 2825.0 / 27648.0,
    dc3 = 250.0 / 621.0 - 18575.0 / 48384.0,
    dc4 = 125.0 / 594.0 - 13525.0 / 55296.0, dc6 = 512.0 / 1771.0 - 0.25;

static double sq(double x) { return x * x; }
static double sq3(double x) { return x * x * x; }

/* Metric.h:64-102 */
static double k_delta(double a, double r) { return sq(r) - 2. * r + a * a; }
static double k_sigma(double a, double r, double th) {
    double asq = a * a;
    return sqrt(sq(sq(r) + asq) - asq * k_delta(a, r) * sq(sin(th)));
}
static double k_w(double a, double r, double th) { return 2. * a * r / sq(k_sigma(a, r, th)); }
static double k_ro(double a, double r, double th) { return sqrt(sq(r) + a * a * sq(cos(th))); }
static double k_wbar(double a, double r, double th) {
    return k_sigma(a, r, th) * sin(th) / k_ro(a, r, th);
}
static double k_alpha(double a, double r, double th) {
    return k_ro(a, r, th) * sqrt(k_delta(a, r)) / k_sigma(a, r, th);
}
static double k_P(double a, double r, double b) { return sq(r) + a * a - a * b; }
static double k_R(double a, double r, double b, double q) {
    return sq(k_P(a, r, b)) - k_delta(a, r) * (sq((b - a)) + q);
}

/* Metric.h:104-108 */
double bto_calc_speed(double a, double r, double theta) {
    double omega = 1. / (a + pow(r, 1.5));
    double sp = k_wbar(a, r, theta) / k_alpha(a, r, theta) * (omega - k_w(a, r, theta));
    return sp;
}

/* Camera.h:54-61 */
static void pose_forms(bto_pose* p) {
    p->alpha = k_alpha(p->a, p->r, p->theta);
    p->w = k_w(p->a, p->r, p->theta);
    p->wbar = k_wbar(p->a, p->r, p->theta);
    p->Delta = k_delta(p->a, p->r);
    p->ro = k_ro(p->a, p->r, p->theta);
}

/* BlackHole.h:12-15 + Camera.h:40-52 */
void bto_pose_init(bto_pose* p, double a, double theta, double phi, double r,
                   double br, double btheta, double bphi) {
    p->a = a;
    p->asq = a * a;
    p->theta = theta;
    p->phi = phi;
    p->r = r;
    p->bphi = bphi;
    p->btheta = btheta;
    p->br = br;
    p->speed = bto_calc_speed(a, r, theta);
    pose_forms(p);
}

/* Camera.h:27-38 (user-speed variant) */
void bto_pose_init_speed(bto_pose* p, double a, double theta, double phi,
                         double r, double speed) {
    p->a = a;
    p->asq = a * a;
    p->theta = theta;
    p->phi = phi;
    p->r = r;
    p->speed = speed;
    p->bphi = 1;
    p->btheta = 0;
    p->br = 0;
    pose_forms(p);
}

void bto_default_params(bto_params* p) {
    p->eps = 1e-5;          /* Metric.h:312 */
    p->h1 = 0.01;           /* Metric.h:313 */
    p->z_end = -10000000;   /* Metric.h:311 */
    p->max_steps = 1000;    /* Metric.h:31 */
    p->apriori_capture = 1; /* Grid.h:552 */
    p->r_escape = 0.0;
    p->r_capture_factor = 0.0;
    p->newton_cap = 100000;
}

/* Metric.h:110-125 */
static double find_min_gold_sec(double a, double theta, double bval, double qval,
                                double ax, double b, double tol, int64_t* iters) {
    (void)theta;
    double gr = (sqrt(5.0) + 1.) / 2.;
    double c = b - (b - ax) / gr;
    double d = ax + (b - ax) / gr;
    int64_t it = 0;
    while (fabs(c - d) > tol) {
        if (k_R(a, c, bval, qval) < k_R(a, d, bval, qval)) {
            b = d;
        } else {
            ax = c;
        }
        c = b - (b - ax) / gr;
        d = ax + (b - ax) / gr;
        it++;
    }
    if (iters) *iters += it;
    return (ax + b) / 2;
}

/* Metric.h:127-131 */
int bto_check_rup(double a, double r, double theta, double b, double q, int64_t* iters) {
    if (a == 0) return 0;
    double mn = find_min_gold_sec(a, theta, b, q, r, 4 * r, 0.00001, iters);
    return k_R(a, mn, b, q) >= 0;
}

/* Metric.h:133-143 */
double bto_b0(double a, double r0) {
    double asq = a * a;
    return -(sq3(r0) - 3. * sq(r0) + asq * r0 + asq) / (a * (r0 - 1.));
}
static double b0diff(double a, double r0) {
    double asq = a * a;
    return (asq + a - 2. * r0 * (sq(r0) - 3. * r0 + 3.)) / (a * sq(r0 - 1.));
}
double bto_q0(double a, double r0) {
    double asq = a * a;
    return -sq3(r0) * (sq3(r0) - 6. * sq(r0) + 9. * r0 - 4. * asq) / (asq * sq(r0 - 1.));
}

/* Metric.h:145-167.  The reference loop has no iteration cap; newton_cap
 * bounds it (after the cap the last iterate is used as-is). */
int bto_check_bq(double a, double bV, double qV, int newton_cap, int64_t* iters) {
    double r1 = 2. * (1. + cos(2. * acos(-a) / 3.));
    double r2 = 2. * (1. + cos(2. * acos(a) / 3.));
    double error = 0.0000001;
    double r0V = 2.0;
    double bcheck = 100;
    int64_t it = 0;
    while (fabs(bV - bcheck) > error) {
        if (it >= newton_cap) break;
        bcheck = bto_b0(a, r0V);
        double bdiffcheck = b0diff(a, r0V);
        double rnew = r0V - (bcheck - bV) / bdiffcheck;
        if (rnew < 1) {
            r0V = 1.0001;
        } else {
            r0V = rnew;
        }
        it++;
    }
    if (iters) *iters += it;
    double qb = bto_q0(a, r0V);
    double b1 = bto_b0(a, r2);
    double b2 = bto_b0(a, r1);
    return (b1 >= bV) || (b2 <= bV) || (qV >= qb);
}

/* Metric.h:169-174 */
int bto_check_celest(double a, double pRV, double rV, double thetaV, double bV,
                     double qV, int newton_cap, int64_t* newton, int64_t* golden) {
    int check1 = bto_check_bq(a, bV, qV, newton_cap, newton);
    int check2 = !check1 && (pRV < 0);
    int check3 = check1 && bto_check_rup(a, rV, thetaV, bV, qV, golden);
    return check2 || check3;
}

/* Metric.h:176-185 */
void bto_wrap_to_pi(double* thetaW, double* phiW) {
    *thetaW = fmod(*thetaW, PI2);
    while (*thetaW < 0) *thetaW += PI2;
    if (*thetaW > PI) {
        *thetaW -= 2 * (*thetaW - PI);
        *phiW += PI;
    }
    while (*phiW < 0) *phiW += PI2;
    *phiW = fmod(*phiW, PI2);
}

/* Test knobs (default off): move every sin/cos of derivs, or every pow of
 * the step-size control, one ulp up (k = 1) or down (k = -1), so that tests
 * can find the rays whose result a last-ulp difference of these functions
 * moves (the device's libm need not round as the host's does). */
static int g_trig_ulp = 0, g_pow_ulp = 0;
void bto_set_trig_ulp(int k) { g_trig_ulp = k; }
void bto_set_pow_ulp(int k) { g_pow_ulp = k; }
static double ulp_move(double v, int k) { return k ? nextafter(v, k > 0 ? HUGE_VAL : -HUGE_VAL) : v; }
static double d_sin(double x) { return ulp_move(sin(x), g_trig_ulp); }
static double d_cos(double x) { return ulp_move(cos(x), g_trig_ulp); }
static double d_pow(double x, double y) { return ulp_move(pow(x, y), g_pow_ulp); }

/* Metric.h:198-226.  var = (r, theta, phi, pR, pTheta). */
void bto_derivs(double a, const double* var, double* out, double b, double q) {
    const double asq = a * a;
    const double r = var[0], th = var[1], pr = var[3], pt = var[4];
    double cosv = d_cos(th);
    double sinv = d_sin(th);
    double cossq = sq(cosv);
    double sinsq = sq(sinv);
    double bsq = sq(b);
    double delta = sq(r) - 2. * r + asq;                 /* _Delta */
    double rosq = sq(r) + asq * sq(d_cos(th));            /* _rosq */
    double P = sq(r) + asq - a * b;                       /* _P */
    double prsq = sq(pr);
    double pthetasq = sq(pt);
    double R = k_R(a, r, b, q);                           /* _R */
    double partR = (q + sq(a - b));
    double btheta = q - sq(d_cos(th)) * (sq(b) / sq(d_sin(th)) - asq); /* _BigTheta */
    double rosqsq = sq(2 * rosq);
    double sqrosqdel = (sq(rosq) * delta);
    double asqcossin = asq * cosv * sinv;
    double rtwo = 2 * r - 2;

    out[0] = delta / rosq * pr;
    out[1] = 1. / rosq * pt;
    out[2] = (2 * a * P - (2 * a - 2 * b) * delta + (2 * b * cossq * delta) / sinsq) / (rosq * 2 * delta);
    out[3] = (rtwo * btheta - rtwo * partR + 4 * r * P) / (rosq * (2 * delta)) - (prsq * rtwo) / (2 * rosq)
        + (4 * pthetasq * r) / rosqsq - ((4 * r - 4) * (btheta * (delta) + R)) / (rosq * sq(2 * delta))
        + (4 * prsq * r * (delta)) / rosqsq - (r * (btheta * delta + R)) / sqrosqdel;
    out[4] = ((2 * cosv * sinv * (bsq / sinsq - asq) + (2 * bsq * sq3(cosv)) / sq3(sinv)) * delta) /
        (rosq * 2 * delta) - (4 * asqcossin * pthetasq) / rosqsq - (4 * asqcossin * prsq * delta) /
        rosqsq + (asqcossin * (btheta * delta + R)) / sqrosqdel;
}

/* Metric.h:228-251 */
void bto_rkck(double a, const double* var, const double* dvdz, double h,
              double* varOut, double* varErr, double b, double q) {
    double aks[25];
    double tmp[5];
    int i;
    const int n = 5;
    for (i = 0; i < n; i++) tmp[i] = var[i] + b21 * h * dvdz[i];
    bto_derivs(a, tmp, aks, b, q);
    for (i = 0; i < n; i++) tmp[i] = var[i] + h * (b31 * dvdz[i] + b32 * aks[i]);
    bto_derivs(a, tmp, aks + 5, b, q);
    for (i = 0; i < n; i++) tmp[i] = var[i] + h * (b41 * dvdz[i] + b42 * aks[i] + b43 * aks[i + 5]);
    bto_derivs(a, tmp, aks + 10, b, q);
    for (i = 0; i < n; i++)
        tmp[i] = var[i] + h * (b51 * dvdz[i] + b52 * aks[i] + b53 * aks[i + 5] + b54 * aks[i + 10]);
    bto_derivs(a, tmp, aks + 15, b, q);
    for (i = 0; i < n; i++)
        tmp[i] = var[i] + h * (b61 * dvdz[i] + b62 * aks[i] + b63 * aks[i + 5] + b64 * aks[i + 10] +
                               b65 * aks[i + 15]);
    bto_derivs(a, tmp, aks + 20, b, q);
    for (i = 0; i < n; i++)
        varOut[i] = var[i] + h * (c1 * dvdz[i] + c3 * aks[i + 5] + c4 * aks[i + 10] + c6 * aks[i + 20]);
    for (i = 0; i < n; i++)
        varErr[i] = h * (dc1 * dvdz[i] + dc3 * aks[i + 5] + dc4 * aks[i + 10] + dc5 * aks[i + 15] +
                         dc6 * aks[i + 20]);
}

/* Metric.h:253-270 */
static void rkqs(double a, double* var, double* dvdz, double* z, double* h, double eps,
                 const double* varScal, double b, double q) {
    double varErr[5], varTemp[5];
    bto_rkck(a, var, dvdz, *h, varTemp, varErr, b, q);
    double errmax = 0.0;
    for (int i = 0; i < 5; i++) errmax = fmax(errmax, fabs(varErr[i] / varScal[i]));
    errmax /= eps;
    if (errmax <= 1.0) {
        *z += *h;
        for (int i = 0; i < 5; i++) var[i] = varTemp[i];
        if (errmax > ERRCON) *h = SAFETY * *h * d_pow(errmax, PGROW);
        else *h = ADAPTIVE * *h;
    } else {
        *h = fmin(SAFETY * *h * d_pow(errmax, PSHRNK), 0.1 * *h);
    }
}

/* Metric.h:272-274: sign of a float argument. */
static int sgnf(float val) { return (0 < val) - (val < 0); }

/* Metric.h:276-303 (+ optional radius stops).  Returns BTO_ESCAPED when
 * z reached z_end (or r > r_escape), BTO_CAPTURED when r < r_cap, and
 * BTO_MAXSTEP when the attempt budget ran out (varStart left untouched). */
static int odeint1(double a, double* varStart, const bto_params* prm, double r_cap,
                   double b, double q, int* step, int64_t* attempts) {
    double varScal[5], var[5], dvdz[5];
    double z = 0.0;
    double h = prm->h1 * sgnf((float)prm->z_end);
    for (int i = 0; i < 5; i++) var[i] = varStart[i];
    for (int nstp = 0; nstp < prm->max_steps; nstp++) {
        bto_derivs(a, var, dvdz, b, q);
        for (int i = 0; i < 5; i++) varScal[i] = fabs(var[i]) + fabs(dvdz[i] * h) + TINY;
        rkqs(a, var, dvdz, &z, &h, prm->eps, varScal, b, q);
        *step = nstp;
        if (attempts) (*attempts)++;
        if (z <= prm->z_end || (prm->r_escape > 0 && var[0] > prm->r_escape)) {
            for (int i = 0; i < 5; i++) varStart[i] = var[i];
            return BTO_ESCAPED;
        }
        if (r_cap > 0 && var[0] < r_cap) {
            for (int i = 0; i < 5; i++) varStart[i] = var[i];
            return BTO_CAPTURED;
        }
    }
    return BTO_MAXSTEP;
}

/* Grid.h:524-547 */
void bto_initial_conditions(const bto_pose* cam, double theta, double phi, double* out) {
    double thetaS = cam->theta;
    double sp = cam->speed;
    double xCam = sin(theta) * cos(phi);
    double yCam = sin(theta) * sin(phi);
    double zCam = cos(theta);

    double yFido = (-yCam + sp) / (1 - sp * yCam);
    double xFido = -sqrtf(1 - sp * sp) * xCam / (1 - sp * yCam);
    double zFido = -sqrtf(1 - sp * sp) * zCam / (1 - sp * yCam);

    double k = sqrt(1 - cam->btheta * cam->btheta);
    double rFido = xFido * cam->bphi / k + cam->br * yFido + cam->br * cam->btheta / k * zFido;
    double thetaFido = cam->btheta * yFido - k * zFido;
    double phiFido = -xFido * cam->br / k + cam->bphi * yFido + cam->bphi * cam->btheta / k * zFido;

    double eF = 1. / (cam->alpha + cam->w * cam->wbar * phiFido);

    double pR = eF * cam->ro * rFido / sqrtf(cam->Delta);
    double pTheta = eF * cam->ro * thetaFido;
    double pPhi = eF * cam->wbar * phiFido;

    double b = pPhi;
    double q = pTheta * pTheta + cos(thetaS) * cos(thetaS) * (b * b / (sin(thetaS) * sin(thetaS)) - cam->asq);
    out[0] = pR;
    out[1] = pTheta;
    out[2] = b;
    out[3] = q;
}

/* Grid.h:524-554 loop body + Metric.h:307-324 */
int bto_trace_ray(const bto_pose* cam, const bto_params* prm, double* theta, double* phi,
                  int* step, bto_stats* st) {
    double ic[4];
    bto_initial_conditions(cam, *theta, *phi, ic);
    const double pR = ic[0], pTheta = ic[1], b = ic[2], q = ic[3];
    int64_t newton = 0, golden = 0, attempts = 0;
    int status;

    *theta = -1;
    *phi = -1;
    *step = 0;
    int escapes = 1;
    if (prm->apriori_capture)
        escapes = bto_check_celest(cam->a, pR, cam->r, cam->theta, b, q, prm->newton_cap,
                                   &newton, &golden);
    if (escapes) {
        double rh = 1. + sqrt(1. - cam->asq);
        double r_cap = prm->r_capture_factor > 0 ? prm->r_capture_factor * rh : 0.0;
        double varStart[5] = {cam->r, cam->theta, cam->phi, pR, pTheta};
        status = odeint1(cam->a, varStart, prm, r_cap, b, q, step, &attempts);
        if (status == BTO_CAPTURED) {
            *theta = -1;
            *phi = -1;
        } else {
            *theta = varStart[1];
            *phi = varStart[2];
            bto_wrap_to_pi(theta, phi);
        }
    } else {
        status = BTO_CAPTURED;
    }
    if (st) {
        st->attempts += attempts;
        st->newton += newton;
        st->golden += golden;
        st->rays += 1;
        st->integrated += escapes ? 1 : 0;
    }
    return status;
}

void bto_trace_rays(const bto_pose* cam, const bto_params* prm, int n, double* theta,
                    double* phi, int* step, int nthreads, bto_stats* st) {
    int64_t att = 0, nw = 0, gd = 0, integ = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(+ : att, nw, gd, integ)
#endif
    for (int i = 0; i < n; i++) {
        bto_stats s = {0, 0, 0, 0, 0};
        bto_trace_ray(cam, prm, &theta[i], &phi[i], &step[i], &s);
        att += s.attempts;
        nw += s.newton;
        gd += s.golden;
        integ += s.integrated;
    }
    (void)nthreads;
    if (st) {
        st->attempts += att;
        st->newton += nw;
        st->golden += gd;
        st->rays += n;
        st->integrated += integ;
    }
}

/* ------------------------------------------------------------------ */
/* Adaptive grid (Grid.h), dense layout: cell (i, j) at index i*M + j.  */
/* ------------------------------------------------------------------ */

typedef struct {
    const bto_pose* cam;
    const bto_params* prm;
    int nthreads;
    bto_stats* st;
    uint8_t* present;
    int64_t* blocks;
    int64_t nblocks, capblocks;
} build_ctx;

#define CEL_T(g, i, j) ((g)->cel[2 * ((int64_t)(i) * (g)->M + (j))])
#define CEL_P(g, i, j) ((g)->cel[2 * ((int64_t)(i) * (g)->M + (j)) + 1])
#define IDX(g, i, j) ((int64_t)(i) * (g)->M + (j))

int bto_grid_dims(int max_level, int angle, int* N, int* M) {
    int eq = angle ? 1 : 0;
    /* Grid.h:612-615 */
    *N = (int)(uint32_t)round(pow(2, max_level) / (2 - eq) + 1);
    *M = (2 - eq) * 2 * (*N - 1);
    return 0;
}

/* Grid.h:382-401 (callKernel) + Grid.h:347-355 (fillGridCam) */
static void call_kernel(bto_grid* g, build_ctx* bc, const int64_t* keys, int64_t n) {
    if (n == 0) return;
    double* th = (double*)malloc(sizeof(double) * n);
    double* ph = (double*)malloc(sizeof(double) * n);
    int* stp = (int*)malloc(sizeof(int) * n);
    for (int64_t q = 0; q < n; q++) {
        uint32_t i = (uint32_t)(keys[q] / g->M), j = (uint32_t)(keys[q] % g->M);
        th[q] = (double)i / (g->N - 1) * PI / (2 - g->equafactor);
        ph[q] = (double)j / g->M * PI2;
    }
    bto_trace_rays(bc->cam, bc->prm, (int)n, th, ph, stp, bc->nthreads, bc->st);
    for (int64_t q = 0; q < n; q++) {
        g->cel[2 * keys[q]] = th[q];
        g->cel[2 * keys[q] + 1] = ph[q];
        g->steps[keys[q]] = stp[q];
        bc->present[keys[q]] = 1;
    }
    g->rays_traced += n;
    free(th);
    free(ph);
    free(stp);
}

static void push(int64_t** arr, int64_t* n, int64_t* cap, int64_t v) {
    if (*n == *cap) {
        *cap = *cap ? *cap * 2 : 1024;
        *arr = (int64_t*)realloc(*arr, sizeof(int64_t) * (*cap));
    }
    (*arr)[(*n)++] = v;
}

/* Grid.h:410-445 */
static int refine_check(bto_grid* g, uint32_t i, uint32_t j, int gap, int level, double thresh) {
    uint32_t k = i + gap;
    uint32_t l = (j + gap) % g->M;
    double th1 = CEL_T(g, i, j), th2 = CEL_T(g, k, j), th3 = CEL_T(g, i, l), th4 = CEL_T(g, k, l);
    double ph1 = CEL_P(g, i, j), ph2 = CEL_P(g, k, j), ph3 = CEL_P(g, i, l), ph4 = CEL_P(g, k, l);
    double diag = (th1 - th4) * (th1 - th4) + (ph1 - ph4) * (ph1 - ph4);
    double diag2 = (th2 - th3) * (th2 - th3) + (ph2 - ph3) * (ph2 - ph3);
    double mx = (diag < diag2) ? diag2 : diag; /* std::max(diag, diag2) */
    if (level < 6 && mx > 1E-10) return 1;
    if (mx > thresh) return 1;
    g->level[IDX(g, i, j)] = (int8_t)level;
    return 0;
}

/* Grid.h:453-459 */
static void fill_vector(bto_grid* g, build_ctx* bc, int64_t** v, int64_t* n, int64_t* cap,
                        uint32_t i, uint32_t j) {
    int64_t id = IDX(g, i, j);
    if (!bc->present[id]) {
        push(v, n, cap, id);
        bc->present[id] = 1;
        g->cel[2 * id] = -10;
        g->cel[2 * id + 1] = -10;
    }
}

int bto_grid_build(const bto_pose* pose, const bto_params* prm, int max_level, int start_level,
                   int angle, double thresh, int fix_tv, int nthreads, bto_grid* g,
                   bto_stats* st) {
    memset(g, 0, sizeof(*g));
    g->max_level = max_level;
    g->start_level = start_level;
    g->equafactor = angle ? 1 : 0;
    bto_grid_dims(max_level, angle, &g->N, &g->M);
    const int N = g->N, M = g->M, eq = g->equafactor;
    int64_t NM = (int64_t)N * M;
    g->cel = (double*)malloc(sizeof(double) * 2 * NM);
    g->steps = (int*)calloc(NM, sizeof(int));
    g->level = (int8_t*)malloc(NM);
    for (int64_t q = 0; q < 2 * NM; q++) g->cel[q] = -2;
    memset(g->level, -1, NM);

    build_ctx bc;
    memset(&bc, 0, sizeof(bc));
    bc.cam = pose;
    bc.prm = prm;
    bc.nthreads = nthreads;
    bc.st = st;
    g->present = (uint8_t*)calloc(NM, 1);
    bc.present = g->present;

    int64_t *blocks = NULL, nb = 0, capb = 0;

    /* raytrace(), Grid.h:291-319 */
    int gap = (int)pow(2, max_level - start_level);
    int64_t start_keys[2] = {0, IDX(g, N - 1, 0)};
    call_kernel(g, &bc, start_keys, 1 + eq);
    for (uint32_t j = 0; j < (uint32_t)M; j += gap) {
        CEL_T(g, 0, j) = CEL_T(g, 0, 0);
        CEL_P(g, 0, j) = CEL_P(g, 0, 0);
        g->steps[IDX(g, 0, j)] = g->steps[0];
        bc.present[IDX(g, 0, j)] = 1;
        push(&blocks, &nb, &capb, IDX(g, 0, j));
        if (eq) {
            CEL_T(g, N - 1, j) = CEL_T(g, N - 1, 0);
            CEL_P(g, N - 1, j) = CEL_P(g, N - 1, 0);
            g->steps[IDX(g, N - 1, j)] = g->steps[0];
            bc.present[IDX(g, N - 1, j)] = 1;
        }
    }

    /* integrateFirst(gap), Grid.h:325-338 */
    {
        int64_t *toint = NULL, nt = 0, capt = 0;
        for (uint32_t i = gap; i < (uint32_t)(N - eq); i += gap) {
            for (uint32_t j = 0; j < (uint32_t)M; j += gap) {
                push(&toint, &nt, &capt, IDX(g, i, j));
                if (i == (uint32_t)(N - 1))
                    ;
                else if (max_level == start_level)
                    g->level[IDX(g, i, j)] = (int8_t)start_level;
                else
                    push(&blocks, &nb, &capb, IDX(g, i, j));
            }
        }
        call_kernel(g, &bc, toint, nt);
        free(toint);
    }

    /* adaptiveBlockIntegration(startLevel), Grid.h:465-505 */
    int level = start_level;
    int early = 0;
    while (level < max_level) {
        if (nb == 0) {
            early = 1;
            break;
        }
        int64_t *todo = NULL, ntodo = 0, captodo = 0;
        int64_t *toint = NULL, nt = 0, capt = 0;
        uint32_t bgap = (uint32_t)pow(2, max_level - level);
        for (int64_t q = 0; q < nb; q++) {
            uint32_t i = (uint32_t)(blocks[q] / M), j = (uint32_t)(blocks[q] % M);
            uint32_t k = i + bgap / 2;
            uint32_t l = j + bgap / 2;
            j = j % M;
            if (refine_check(g, i, j, bgap, level, thresh)) {
                fill_vector(g, &bc, &toint, &nt, &capt, k, j);
                fill_vector(g, &bc, &toint, &nt, &capt, k, l);
                fill_vector(g, &bc, &toint, &nt, &capt, i, l);
                fill_vector(g, &bc, &toint, &nt, &capt, i + bgap, l);
                fill_vector(g, &bc, &toint, &nt, &capt, k, (j + bgap) % M);
                push(&todo, &ntodo, &captodo, IDX(g, i, j));
                push(&todo, &ntodo, &captodo, IDX(g, k, j));
                push(&todo, &ntodo, &captodo, IDX(g, k, l));
                push(&todo, &ntodo, &captodo, IDX(g, i, l));
            }
        }
        call_kernel(g, &bc, toint, nt);
        free(toint);
        level++;
        free(blocks);
        blocks = todo;
        nb = ntodo;
        capb = captodo;
    }
    if (!early)
        for (int64_t q = 0; q < nb; q++) g->level[blocks[q]] = (int8_t)level;
    free(blocks);

    if (fix_tv) bto_grid_fix_tvertices(g);
    return 0;
}

/* ---------------- fixTvertices (Grid.h:103-204) ---------------- */

static int present_cell(const bto_grid* g, uint32_t i, uint32_t j) {
    if (i >= (uint32_t)g->N || j >= (uint32_t)g->M) return 0;
    return g->present[IDX(g, i, j)] != 0;
}

/* Grid.h:183-204 with aTension = aBias = 0 */
void bto_hermite2(double aValue, const double* aX0, const double* aX1, const double* aX2,
                  const double* aX3, double* out) {
    double const v = aValue;
    double const v2 = v * v;
    double const v3 = v * v2;
    double const aTension = 0., aBias = 0.;
    double const aa = (double)(1 + aBias) * (double)(1 - aTension) / (double)(2);
    double const bb = (double)(1 - aBias) * (double)(1 - aTension) / (double)(2);
    double const u0 = (double)(2) * v3 - (double)(3) * v2 + (double)(1);
    double const u1 = v3 - (double)(2) * v2 + v;
    double const u2 = v3 - v2;
    double const u3 = (double)(-2) * v3 + (double)(3) * v2;
    for (int c = 0; c < 2; c++) {
        double m0 = (aX1[c] - aX0[c]) * aa + (aX2[c] - aX1[c]) * bb;
        double m1 = (aX2[c] - aX1[c]) * aa + (aX3[c] - aX2[c]) * bb;
        out[c] = aX1[c] * u0 + m0 * u1 + m1 * u2 + aX2[c] * u3;
    }
}

/* Metric.h:333-339 / 355-364 with factor 5. on the phi components */
static void fix_2pi(double* phis, int n) {
    int cross = 0;
    for (int q = 0; q < n; q++)
        if (phis[q] > PI2 * (1. - 1. / (double)5.f)) cross = 1;
    if (!cross) return;
    for (int q = 0; q < n; q++)
        if (phis[q] < PI2 * (1. / (double)5.f)) phis[q] += PI2;
}

/* Grid.h:132-177 */
static void check_adjacent_block(bto_grid* g, uint32_t ii, uint32_t jj, uint32_t ii2, uint32_t jj2,
                                 int level, int udlr, int gap) {
    const uint32_t M = (uint32_t)g->M, N = (uint32_t)g->N;
    uint32_t i = ii + udlr * gap / 2;
    uint32_t j = jj + (1 - udlr) * gap / 2;
    if (!present_cell(g, i, j)) return;
    uint32_t jprev = (jj - (1 - udlr) * gap + M) % M;
    uint32_t jnext = (jj + (1 - udlr) * 2 * gap) % M;
    uint32_t iprev = ii - udlr * gap;
    uint32_t inext = ii + 2 * udlr * gap;
    int half = 0;
    if (ii == 0) {
        jprev = (jj + M / 2) % M;
        iprev = gap;
    } else if (inext > N - 1) {
        inext = ii;
        if (g->equafactor) jnext = (jj + M / 2) % M;
        else half = 1;
    }
    int succes = 0;
    if (present_cell(g, iprev, jprev) && present_cell(g, inext, jnext)) {
        double c0[2] = {CEL_T(g, iprev, jprev), CEL_P(g, iprev, jprev)};
        double c1[2] = {CEL_T(g, ii, jj), CEL_P(g, ii, jj)};
        double c2[2] = {CEL_T(g, ii2, jj2), CEL_P(g, ii2, jj2)};
        double c3[2] = {CEL_T(g, inext, jnext), CEL_P(g, inext, jnext)};
        if ((c0[0] != -1 || c0[1] != -1) && (c3[0] != -1 || c3[1] != -1)) {
            succes = 1;
            if (half) c3[0] = PI - c3[0];
            double phis[4] = {c0[1], c1[1], c2[1], c3[1]};
            fix_2pi(phis, 4);
            c0[1] = phis[0];
            c1[1] = phis[1];
            c2[1] = phis[2];
            c3[1] = phis[3];
            double out[2];
            bto_hermite2(0.5, c0, c1, c2, c3, out);
            CEL_T(g, i, j) = out[0];
            CEL_P(g, i, j) = out[1];
        }
    }
    if (!succes) {
        double c0[2] = {CEL_T(g, ii, jj), CEL_P(g, ii, jj)};
        double c1[2] = {CEL_T(g, ii2, jj2), CEL_P(g, ii2, jj2)};
        double phis[2] = {c0[1], c1[1]};
        fix_2pi(phis, 2);
        CEL_T(g, i, j) = 1. / 2. * (c1[0] + c0[0]);
        CEL_P(g, i, j) = 1. / 2. * (phis[1] + phis[0]);
    }
    if (level + 1 == g->max_level) return;
    check_adjacent_block(g, ii, jj, i, j, level + 1, udlr, gap / 2);
    check_adjacent_block(g, i, j, ii2, jj2, level + 1, udlr, gap / 2);
}

/* Grid.h:103-119 */
static void fix_tvertices_block(bto_grid* g, uint32_t i, uint32_t j, int level) {
    if (level == g->max_level) return;
    if (CEL_P(g, i, j) < 0) return;
    int gap = (int)pow(2, g->max_level - level);
    uint32_t k = i + gap;
    uint32_t l = (j + gap) % g->M;
    check_adjacent_block(g, i, j, k, j, level, 1, gap);
    check_adjacent_block(g, i, j, i, l, level, 0, gap);
    check_adjacent_block(g, i, l, k, l, level, 1, gap);
    check_adjacent_block(g, k, j, k, l, level, 0, gap);
}

/* Grid.h:620-622 in the pinned order: level ascending, then key
 * ((uint64)i << 32 | j, i.e. row-major index) ascending. */
void bto_grid_fix_tvertices(bto_grid* g) {
    int64_t NM = (int64_t)g->N * g->M;
    for (int L = g->start_level; L < g->max_level; L++)
        for (int64_t id = 0; id < NM; id++)
            if (g->level[id] == L) fix_tvertices_block(g, (uint32_t)(id / g->M), (uint32_t)(id % g->M), L);
}

void bto_grid_free(bto_grid* g) {
    free(g->cel);
    free(g->steps);
    free(g->level);
    free(g->present);
    memset(g, 0, sizeof(*g));
}
