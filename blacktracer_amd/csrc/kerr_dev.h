// kerr_dev.h -- device-side Kerr geodesic building blocks in FP64:
// initial conditions, a-priori capture test, Hamiltonian right-hand side
// and one Cash-Karp attempt with the reference's step control.
//
// Included only by translation units compiled with -ffp-contract=off, so
// that every expression rounds exactly as the reference's host code does
// (Metric.h is compiled with /fp:precise).  Expression order below is kept
// operation for operation; do not reassociate.
#pragma once

#include "bt_internal.h"

namespace bt {
namespace kerr {

constexpr double PI = 3.141592653589793238;   // Const.h:2
constexpr double PI2 = 3.141592653589793238 * 2.;

// Metric.h:27-33
constexpr double SAFETY = 0.9;
constexpr double PGROW = -0.2;
constexpr double PSHRNK = -0.25;
constexpr double ERRCON = 1.89e-4;
constexpr double TINY = 1.0e-30;
constexpr double ADAPTIVE = 5.0;

// Cash-Karp embedded RK4(5) tableau, Metric.h:41-49.
constexpr double b21 = 0.2, b31 = 3.0 / 40.0, b32 = 9.0 / 40.0, b41 = 0.3, b42 = -0.9,
                 b43 = 1.2, b51 = -11.0 / 54.0, b52 = 2.5, b53 = -70.0 / 27.0,
                 b54 = 35.0 / 27.0, b61 = 1631.0 / 55296.0, b62 = 175.0 / 512.0,
                 b63 = 575.0 / 13824.0, b64 = 44275.0 / 110592.0, b65 = 253.0 / 4096.0,
                 c1 = 37.0 / 378.0, c3 = 250.0 / 621.0, c4 = 125.0 / 594.0,
                 c6 = 512.0 / 1771.0, dc5 = -277.00 / 14336.0;
constexpr double dc1 = 37.0 / 378.0 - 2825.0 / 27648.0, dc3 = 250.0 / 621.0 - 18575.0 / 48384.0,
                 dc4 = 125.0 / 594.0 - 13525.0 / 55296.0, dc6 = 512.0 / 1771.0 - 0.25;

__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ double sq3(double x) { return x * x * x; }

// Division by a shared denominator d, bit-identical to the compiler's FP64
// division x / d when |x| (unless zero), |d| and |x / d| lie between about
// 2^-900 and 2^900: the same steps -- the hardware reciprocal estimate, two
// Newton-Raphson refinements, the quotient, its remainder by FMA and one
// correction (Markstein's scheme) -- with the reciprocal refined once per
// denominator instead of once per division.  In that range the division
// sequence's pre-scaling and final fix-up leave every value unchanged; they
// act only on zero, denormal, huge and non-finite operands.
struct SharedDiv {
    double d, y;
    __device__ __forceinline__ explicit SharedDiv(double den) : d(den) {
        double r = __builtin_amdgcn_rcp(den);
        double e = fma(-den, r, 1.0);
        r = fma(r, e, r);
        e = fma(-den, r, 1.0);
        y = fma(r, e, r);
    }
    // d > 0.  The remainder is taken as d q - x (the negated x - d q, exact),
    // which gives a zero x's quotient the sign x / d has.
    __device__ __forceinline__ double pos(double x) const {
        const double q = x * y;
        const double rem = fma(d, q, -x);
        return fma(-rem, y, q);
    }
    // Any sign of d: a zero quotient takes the sign of x * y, that of x / d
    // (a nonzero one has it already).
    __device__ __forceinline__ double any(double x) const {
        const double q = x * y;
        return copysign(fma(fma(-d, q, x), y, q), q);
    }
};

// R(r) of Metric.h:96-98 with P and Delta of Metric.h:64-66,92-94.
__device__ __forceinline__ double Rfun(double a, double asq, double r, double b, double q) {
    double P = sq(r) + asq - a * b;
    double D = sq(r) - 2. * r + asq;
    return sq(P) - D * (sq((b - a)) + q);
}

// Metric.h:133-143
__device__ __forceinline__ double b0(double a, double asq, double r0) {
    return -(sq3(r0) - 3. * sq(r0) + asq * r0 + asq) / (a * (r0 - 1.));
}
__device__ __forceinline__ double b0diff(double a, double asq, double r0) {
    return (asq + a - 2. * r0 * (sq(r0) - 3. * r0 + 3.)) / (a * sq(r0 - 1.));
}
__device__ __forceinline__ double q0(double asq, double r0) {
    return -sq3(r0) * (sq3(r0) - 6. * sq(r0) + 9. * r0 - 4. * asq) / (asq * sq(r0 - 1.));
}

// Grid.h:524-547: camera-sky direction -> (pR, pTheta, b, q).
__device__ __forceinline__ void initial_conditions(const PoseK& P, double theta, double phi,
                                                   double& pR, double& pTheta, double& b,
                                                   double& q) {
    double st, ct, sp, cp;
    sincos(theta, &st, &ct);
    sincos(phi, &sp, &cp);
    double xCam = st * cp;
    double yCam = st * sp;
    double zCam = ct;
    const SharedDiv by_dop(1 - P.speed * yCam);
    double yFido = by_dop.any(-yCam + P.speed);
    double xFido = by_dop.any(P.neg_s * xCam);
    double zFido = by_dop.any(P.neg_s * zCam);
    double rFido = xFido * P.bphi / P.k + P.br * yFido + P.brbt_k * zFido;
    double thetaFido = P.btheta * yFido - P.k * zFido;
    double phiFido = -xFido * P.br / P.k + P.bphi * yFido + P.bpbt_k * zFido;
    double eF = 1. / (P.alpha + P.ww * phiFido);
    pR = eF * P.ro * rFido / P.sqrtD;
    pTheta = eF * P.ro * thetaFido;
    double pPhi = eF * P.wbar * phiFido;
    b = pPhi;
    q = pTheta * pTheta + P.cos2S * (b * b / P.sin2S - P.asq);
}

// Metric.h:145-167 (Newton loop bounded by newton_cap).
// *r0 (nullable) receives the b-matched photon-orbit radius of the Newton solve.
__device__ __forceinline__ bool check_bq(const PoseK& P, double bV, double qV, int cap,
                                         unsigned& iters, double* r0 = nullptr) {
    const double error = 0.0000001;
    double r0V = 2.0;
    double bcheck = 100;
    int it = 0;
    while (fabs(bV - bcheck) > error) {
        if (it >= cap) break;
        bcheck = b0(P.a, P.asq, r0V);
        double bdiffcheck = b0diff(P.a, P.asq, r0V);
        double rnew = r0V - (bcheck - bV) / bdiffcheck;
        r0V = (rnew < 1) ? 1.0001 : rnew;
        it++;
    }
    iters += it;
    double qb = q0(P.asq, r0V);
    if (r0) *r0 = r0V;
    return (P.b1 >= bV) || (P.b2 <= bV) || (qV >= qb);
}

// dq0/dr0 of q0 above: with N = -r^3 (r^3 - 6r^2 + 9r - 4a^2),
// q0 = N / (a^2 (r-1)^2) and q0' = (N' (r-1) - 2N) / (a^2 (r-1)^3).
__device__ __forceinline__ double q0diff(double asq, double r) {
    double N = -sq3(r) * (sq3(r) - 6. * sq(r) + 9. * r - 4. * asq);
    double Nd = -(6. * sq(sq(r)) * r - 30. * sq(sq(r)) + 36. * sq3(r) - 12. * asq * sq(r));
    return (Nd * (r - 1.) - 2. * N) / (asq * sq3(r - 1.));
}

// Euclidean distance in the (b, q) plane from a ray's constants of motion
// to the critical curve of spherical photon orbits {(b0(r0), q0(r0)) :
// r0 in [r_ph1, r_ph2]} (Bardeen's curve, the shadow boundary; q0 = 0 at
// both ends).  Clamped Gauss-Newton projection from the b-matched r0; for
// a = 0 the curve is b^2 + q = 27.
__device__ __forceinline__ double critical_distance(const PoseK& P, double b, double q, double r0) {
    if (P.a_zero) {
        double t = fmin(fmax(b, -sqrt(27.)), sqrt(27.));
        for (int it = 0; it < 8; it++) {
            double dq = -2. * t;
            double num = (b - t) + (q - (27. - t * t)) * dq;
            t = fmin(fmax(t + num / (1. + dq * dq), -sqrt(27.)), sqrt(27.));
        }
        return hypot(b - t, q - (27. - t * t));
    }
    double lo = P.r_ph1, hi = P.r_ph2;
    double r = fmin(fmax(r0, lo), hi);
    for (int it = 0; it < 8; it++) {
        double db = b0diff(P.a, P.asq, r), dq = q0diff(P.asq, r);
        double num = (b - b0(P.a, P.asq, r)) * db + (q - q0(P.asq, r)) * dq;
        r = fmin(fmax(r + num / (db * db + dq * dq), lo), hi);
    }
    return hypot(b - b0(P.a, P.asq, r), q - q0(P.asq, r));
}

// True when (b, q) lies provably farther than eps from the critical curve,
// without the projection: outside the curve's b range by more than eps, or
// (with r0 the b-matched radius of the Newton solve, b0(r0) = b) at a
// vertical distance |q - q0(r0)| > eps sqrt(1 + S^2), S the curve's steepest
// slope over b (a point at vertical distance D from a graph of slope at most
// S is at least D / sqrt(1 + S^2) from it).
__device__ __forceinline__ bool far_from_critical(const PoseK& P, double b, double q, double r0, double eps) {
    const double blo = fmin(P.b1, P.b2), bhi = fmax(P.b1, P.b2);
    if (P.a_zero) {
        if (fabs(b) > sqrt(27.) + eps) return true;
        return fabs(q - (27. - b * b)) > eps * sqrt(1. + P.crit_slope * P.crit_slope);
    }
    if (b < blo - eps || b > bhi + eps) return true;
    if (!(P.crit_slope < 1e30) || !(r0 >= P.r_ph1 && r0 <= P.r_ph2)) return false;
    // the curve's height at b differs from q0(r0) by at most S |b0(r0) - b|
    const double slack = P.crit_slope * fabs(b0(P.a, P.asq, r0) - b);
    return fabs(q - q0(P.asq, r0)) - slack > eps * sqrt(1. + P.crit_slope * P.crit_slope);
}

// Metric.h:110-131
__device__ __forceinline__ bool check_rup(const PoseK& P, double rV, double bV, double qV,
                                          unsigned& iters) {
    if (P.a_zero) return false;
    double ax = rV, bb = 4 * rV;
    const double tol = 0.00001;
    const SharedDiv by_gr(P.gr);   // the golden ratio; (bb - ax) / gr below
    double c = bb - by_gr.pos(bb - ax);
    double d = ax + by_gr.pos(bb - ax);
    int it = 0;
    while (fabs(c - d) > tol) {
        if (Rfun(P.a, P.asq, c, bV, qV) < Rfun(P.a, P.asq, d, bV, qV))
            bb = d;
        else
            ax = c;
        c = bb - by_gr.pos(bb - ax);
        d = ax + by_gr.pos(bb - ax);
        it++;
    }
    iters += it;
    double mn = (ax + bb) / 2;
    return Rfun(P.a, P.asq, mn, bV, qV) >= 0;
}

// Metric.h:169-174: true when the ray escapes (is integrated).
__device__ __forceinline__ bool check_celest(const PoseK& P, double pR, double bV, double qV,
                                             int cap, unsigned& newton, unsigned& golden,
                                             double* r0 = nullptr) {
    bool check1 = check_bq(P, bV, qV, cap, newton, r0);
    bool check2 = !check1 && (pR < 0);
    bool check3 = check1 && check_rup(P, P.r, bV, qV, golden);
    return check2 || check3;
}

// Metric.h:198-226.  v = (r, theta, phi, pR, pTheta).  Every operation is
// the reference's, in its order; the sixteen divisions run through
// SharedDiv, fourteen of them sharing five reciprocals (rosq, sinsq,
// rosq 2 delta, rosqsq and sqrosqdel; the denominators rosq * 2 * delta and
// rosq * (2 * delta) are the same double, and x / (2 rosq) is half of
// x / rosq exactly).
__device__ __forceinline__ void derivs(double a, double asq, const double* v, double* o,
                                       double b, double q) {
    const double r = v[0], pr = v[3], pt = v[4];
    double sinv, cosv;
    sincos(v[1], &sinv, &cosv);
    double cossq = sq(cosv);
    double sinsq = sq(sinv);
    double bsq = sq(b);
    double delta = sq(r) - 2. * r + asq;
    double rosq = sq(r) + asq * cossq;
    double P = sq(r) + asq - a * b;
    double prsq = sq(pr);
    double pthetasq = sq(pt);
    double R = sq(P) - delta * (sq((b - a)) + q);
    double partR = (q + sq(a - b));
    const SharedDiv by_rosq(rosq), by_sinsq(sinsq);
    const double bsq_sinsq = by_sinsq.pos(bsq);   // bsq / sinsq
    double btheta = q - cossq * (bsq_sinsq - asq);
    double rosqsq = sq(2 * rosq);
    double sqrosqdel = (sq(rosq) * delta);
    double asqcossin = asq * cosv * sinv;
    double rtwo = 2 * r - 2;
    const SharedDiv by_rd2(rosq * 2 * delta), by_rosqsq(rosqsq), by_srd(sqrosqdel);
    const double bdr = btheta * delta + R;

    o[0] = by_rosq.pos(delta) * pr;
    o[1] = by_rosq.pos(1.) * pt;
    o[2] = by_rd2.any(2 * a * P - (2 * a - 2 * b) * delta + by_sinsq.pos(2 * b * cossq * delta));
    o[3] = by_rd2.any(rtwo * btheta - rtwo * partR + 4 * r * P) - by_rosq.pos(prsq * rtwo) * 0.5
        + by_rosqsq.pos(4 * pthetasq * r) - SharedDiv(rosq * sq(2 * delta)).pos((4 * r - 4) * bdr)
        + by_rosqsq.pos(4 * prsq * r * (delta)) - by_srd.any(r * bdr);
    o[4] = by_rd2.any((2 * cosv * sinv * (bsq_sinsq - asq) + SharedDiv(sq3(sinv)).any(2 * bsq * sq3(cosv))) *
                      delta)
        - by_rosqsq.pos(4 * asqcossin * pthetasq) - by_rosqsq.pos(4 * asqcossin * prsq * delta)
        + by_srd.any(asqcossin * bdr);
}

// odeint1's loop body (Metric.h:292-297): derivs, scale, rkqs (Metric.h:253-270)
// with rkck (Metric.h:228-251).  Updates var/z/h in place; returns errmax.
// Each stage sum of rkck is accumulated as its stages arrive, in rkck's
// left-to-right order (b41 k1 + b42 k2, then + b43 k3, ...): with no FMA
// contraction every partial sum rounds exactly as in the one-expression
// form, and the stage derivatives need not stay live to the end.
__device__ __forceinline__ double attempt(double a, double asq, double* var, double& z, double& h,
                                          double eps, double b, double q) {
    double dvdz[5], scal[5], k[5], t[5], s4[5], s5[5], s6[5], pc[5], pd[5];
    derivs(a, asq, var, dvdz, b, q);
#pragma unroll
    for (int i = 0; i < 5; i++) scal[i] = fabs(var[i]) + fabs(dvdz[i] * h) + TINY;
    const double hh = h;
#pragma unroll
    for (int i = 0; i < 5; i++) t[i] = var[i] + b21 * hh * dvdz[i];
    derivs(a, asq, t, k, b, q);   // k2
#pragma unroll
    for (int i = 0; i < 5; i++) {
        t[i] = var[i] + hh * (b31 * dvdz[i] + b32 * k[i]);
        s4[i] = b41 * dvdz[i] + b42 * k[i];
        s5[i] = b51 * dvdz[i] + b52 * k[i];
        s6[i] = b61 * dvdz[i] + b62 * k[i];
    }
    derivs(a, asq, t, k, b, q);   // k3
#pragma unroll
    for (int i = 0; i < 5; i++) {
        t[i] = var[i] + hh * (s4[i] + b43 * k[i]);
        s5[i] = s5[i] + b53 * k[i];
        s6[i] = s6[i] + b63 * k[i];
        pc[i] = c1 * dvdz[i] + c3 * k[i];
        pd[i] = dc1 * dvdz[i] + dc3 * k[i];
    }
    derivs(a, asq, t, k, b, q);   // k4
#pragma unroll
    for (int i = 0; i < 5; i++) {
        t[i] = var[i] + hh * (s5[i] + b54 * k[i]);
        s6[i] = s6[i] + b64 * k[i];
        pc[i] = pc[i] + c4 * k[i];
        pd[i] = pd[i] + dc4 * k[i];
    }
    derivs(a, asq, t, k, b, q);   // k5
#pragma unroll
    for (int i = 0; i < 5; i++) {
        t[i] = var[i] + hh * (s6[i] + b65 * k[i]);
        pd[i] = pd[i] + dc5 * k[i];
    }
    derivs(a, asq, t, k, b, q);   // k6
    double errmax = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        t[i] = var[i] + hh * (pc[i] + c6 * k[i]);
        double err = hh * (pd[i] + dc6 * k[i]);
        // scal >= TINY > 0.  An err below the division's range (2^-900)
        // gives a ratio far below ERRCON, which decides nothing: errmax's
        // value is used (pow) only above ERRCON.
        errmax = fmax(errmax, fabs(SharedDiv(scal[i]).pos(err)));
    }
    errmax = SharedDiv(eps).pos(errmax);   // eps > 0
    // One pow call serves both branches (the same arguments as the
    // reference's two calls), so a wave whose lanes split between accept
    // and reject evaluates it once rather than twice.
    const bool accept = errmax <= 1.0;
    const double pw = pow(errmax, accept ? PGROW : PSHRNK);
    if (accept) {
        z += h;
#pragma unroll
        for (int i = 0; i < 5; i++) var[i] = t[i];
        if (errmax > ERRCON)
            h = SAFETY * h * pw;
        else
            h = ADAPTIVE * h;
    } else {
        h = fmin(SAFETY * h * pw, 0.1 * h);
    }
    return errmax;
}

// ---- the mixed path's FP64 warm-up attempt ------------------------------
// The warm-up reproduces the FP64 step-size growth of a ray (its decisions
// errmax <= ERRCON hold with orders of magnitude to spare) and hands its
// state to the FP32 bulk, which parts from the FP64 path at the 1e-7 level
// anyway; it needs FP64 accuracy, not the reference's exact roundings.  So
// it evaluates the regrouped right-hand side of the FP32 kernel in FP64,
// with reciprocals from the hardware estimate refined by two Newton steps
// and sin/cos from their Taylor series after a two-constant range reduction,
// at about a third of the exact attempt's instructions.

// 1/x to about one ulp: the hardware estimate and two Newton-Raphson steps.
__device__ __forceinline__ double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// sin and cos of th: th = n pi/2 + v with pi/2 in two parts, |v| <= pi/4,
// Taylor series to v^17 / v^18 (truncation below 1e-19 on |v| <= pi/4).
__device__ __forceinline__ void sincos64(double th, double& s, double& c) {
    constexpr double PIO2_HI = 1.5707963267948966192;          // pi/2 rounded
    constexpr double PIO2_LO = 6.123233995736766036e-17;       // pi/2 - PIO2_HI
    constexpr double TWO_OVER_PI = 0.63661977236758134308;
    const double n = rint(th * TWO_OVER_PI);
    double v = fma(-n, PIO2_HI, th);
    v = fma(-n, PIO2_LO, v);
    const double v2 = v * v;
    double ps = 1.0 / 355687428096000.0;                        // 1/17!
    ps = fma(ps, v2, -1.0 / 1307674368000.0);                  // 1/15!
    ps = fma(ps, v2, 1.0 / 6227020800.0);                      // 1/13!
    ps = fma(ps, v2, -1.0 / 39916800.0);                       // 1/11!
    ps = fma(ps, v2, 1.0 / 362880.0);                          // 1/9!
    ps = fma(ps, v2, -1.0 / 5040.0);
    ps = fma(ps, v2, 1.0 / 120.0);
    ps = fma(ps, v2, -1.0 / 6.0);
    const double sv = fma(v * v2, ps, v);
    double pc = 1.0 / 6402373705728000.0;                       // 1/18!
    pc = fma(pc, v2, -1.0 / 20922789888000.0);                 // 1/16!
    pc = fma(pc, v2, 1.0 / 87178291200.0);                     // 1/14!
    pc = fma(pc, v2, -1.0 / 479001600.0);                      // 1/12!
    pc = fma(pc, v2, 1.0 / 3628800.0);                         // 1/10!
    pc = fma(pc, v2, -1.0 / 40320.0);
    pc = fma(pc, v2, 1.0 / 720.0);
    pc = fma(pc, v2, -1.0 / 24.0);
    pc = fma(pc, v2, 0.5);
    const double cv = fma(-v2, pc, 1.0);
    const int qd = (int)n & 3;
    const double s0 = (qd & 1) ? cv : sv;
    const double c0 = (qd & 1) ? sv : cv;
    s = (qd & 2) ? -s0 : s0;
    c = ((qd + 1) & 2) ? -c0 : c0;
}

// Per-ray constants of the warm-up right-hand side.
struct RayW {
    double a, asq, b, bsq, q, partR, e, Poff;   // e = 1 - a^2, Poff = a^2 - a b
    __device__ __forceinline__ void set(double a_, double asq_, double b_, double q_) {
        a = a_;
        asq = asq_;
        b = b_;
        q = q_;
        bsq = b_ * b_;
        partR = q_ + (a_ - b_) * (a_ - b_);
        e = 1.0 - asq_;
        Poff = asq_ - a_ * b_;
    }
};

// The right-hand side of Metric.h:198-226 regrouped (as the FP32 kernel's):
// iR = 1/rho^2, iD = 1/Delta, Ph = P/Delta and the null-constraint term X.
__device__ __forceinline__ void derivs_w(const RayW& C, const double* v, double* o) {
    const double r = v[0], pr = v[3], pt = v[4];
    double s, c;
    sincos64(v[1], s, c);
    const double rm1 = r - 1.0;
    const double D = fma(rm1, rm1, -C.e);
    const double P = fma(r, r, C.Poff);
    const double c2 = c * c;
    const double rosq = fma(C.asq, c2, r * r);
    const double iR = rcp64(rosq), iD = rcp64(D), is = rcp64(s);
    const double is2 = is * is;
    const double Ph = P * iD;
    const double Theta = C.q - c2 * (C.bsq * is2 - C.asq);
    const double prsq = pr * pr;
    const double X = pt * pt - Theta + C.partR + prsq * D - P * Ph;
    const double iR2X = iR * iR * X;
    o[0] = D * iR * pr;
    o[1] = iR * pt;
    o[2] = iR * (C.a * Ph - C.a + C.b * is2);
    o[3] = iR * (2.0 * r * Ph - rm1 * (Ph * Ph + prsq)) + r * iR2X;
    o[4] = iR * c * (C.bsq * is2 * is - C.asq * s) - C.asq * c * s * iR2X;
}

// attempt() with derivs_w; the same tableau and step control.
__device__ __forceinline__ double attempt_w(const RayW& C, double* var, double& z, double& h, double eps) {
    double dvdz[5], scal[5], k2[5], k3[5], k4[5], k5[5], k6[5], t[5];
    derivs_w(C, var, dvdz);
#pragma unroll
    for (int i = 0; i < 5; i++) scal[i] = fabs(var[i]) + fabs(dvdz[i] * h) + TINY;
    const double hh = h;
#pragma unroll
    for (int i = 0; i < 5; i++) t[i] = var[i] + b21 * hh * dvdz[i];
    derivs_w(C, t, k2);
#pragma unroll
    for (int i = 0; i < 5; i++) t[i] = var[i] + hh * (b31 * dvdz[i] + b32 * k2[i]);
    derivs_w(C, t, k3);
#pragma unroll
    for (int i = 0; i < 5; i++) t[i] = var[i] + hh * (b41 * dvdz[i] + b42 * k2[i] + b43 * k3[i]);
    derivs_w(C, t, k4);
#pragma unroll
    for (int i = 0; i < 5; i++)
        t[i] = var[i] + hh * (b51 * dvdz[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i]);
    derivs_w(C, t, k5);
#pragma unroll
    for (int i = 0; i < 5; i++)
        t[i] = var[i] + hh * (b61 * dvdz[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i]);
    derivs_w(C, t, k6);
    double errmax = 0.0;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        t[i] = var[i] + hh * (c1 * dvdz[i] + c3 * k3[i] + c4 * k4[i] + c6 * k6[i]);
        double err = hh * (dc1 * dvdz[i] + dc3 * k3[i] + dc4 * k4[i] + dc5 * k5[i] + dc6 * k6[i]);
        errmax = fmax(errmax, fabs(err) * rcp64(scal[i]));
    }
    errmax /= eps;
    const bool accept = errmax <= 1.0;
    const double pw = pow(errmax, accept ? PGROW : PSHRNK);   // one call for both branches
    if (accept) {
        z += h;
#pragma unroll
        for (int i = 0; i < 5; i++) var[i] = t[i];
        if (errmax > ERRCON)
            h = SAFETY * h * pw;
        else
            h = ADAPTIVE * h;
    } else {
        h = fmin(SAFETY * h * pw, 0.1 * h);
    }
    return errmax;
}

// Metric.h:176-185 (the while loops are kept: repeated addition of 2*PI
// rounds differently from a single fmod).
__device__ __forceinline__ void wrap_to_pi(double& thetaW, double& phiW) {
    thetaW = fmod(thetaW, PI2);
    while (thetaW < 0) thetaW += PI2;
    if (thetaW > PI) {
        thetaW -= 2 * (thetaW - PI);
        phiW += PI;
    }
    while (phiW < 0) phiW += PI2;
    phiW = fmod(phiW, PI2);
}

// Camera-sky direction of ray k (Grid.h:386-390 for grid keys).
__device__ __forceinline__ void ray_direction(const RayIO& io, int64_t k, double& th, double& ph) {
    if (io.keys) {
        uint32_t key = io.keys[k];
        uint32_t i = key / (uint32_t)io.M, j = key % (uint32_t)io.M;
        th = (double)i / (io.N - 1) * PI / (2 - io.eq);
        ph = (double)j / io.M * PI2;
    } else {
        th = io.in_th[k];
        ph = io.in_ph[k];
    }
}

__device__ __forceinline__ void ray_store(const RayIO& io, int64_t k, double th, double ph, int step) {
    int64_t o = io.keys ? (int64_t)io.keys[k] : k;
    io.out_th[o * io.ostride] = th;
    io.out_ph[o * io.ostride] = ph;
    io.out_step[o] = step;
}

// ---- wavefront helpers (wave64) ----
__device__ __forceinline__ unsigned lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Adds this thread's counts c[i] into *dst[i], summed over the block first:
// at most one atomic per counter and block, none for a zero sum.  A launch's
// statistics share one cache line, where an atomic from every wave would
// queue behind the others (one every few nanoseconds).  Every thread of the
// block (at most 256) calls it, once, at the end of the kernel.
template <int N>
__device__ __forceinline__ void block_add_stats(const unsigned long long (&c)[N],
                                                unsigned long long* const (&dst)[N]) {
    __shared__ unsigned long long s_part[4][N];
    const unsigned wave = (threadIdx.x >> 6) & 3;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const unsigned long long v = wave_sum(c[i]);
        if (lane_id() == 0) s_part[wave][i] = v;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; i++)
        if (threadIdx.x == (unsigned)i) {
            unsigned long long t = 0;
            for (unsigned w = 0; w < (blockDim.x + 63) / 64; w++) t += s_part[w][i];
            if (t) atomicAdd(dst[i], t);
        }
}

// Wave-cooperative claim of queue positions for the lanes in ballot mask m
// (every lane of the wave calls it with the same m).  A wave reserves
// `chunk` (>= 64) consecutive positions per atomic on *head and hands them
// out in order; next/end describe the unused part of its reservation and
// are wave-uniform.  Returns this lane's position (meaningful when its bit
// is set in m).  Positions >= the queue length mean "queue exhausted".
__device__ __forceinline__ unsigned claim_positions(unsigned long long m, unsigned lane,
                                                    unsigned* head, unsigned chunk,
                                                    unsigned& next, unsigned& end) {
    const unsigned cnt = (unsigned)__popcll(m);
    const unsigned rank = (unsigned)__popcll(m & ((1ull << lane) - 1));
    const unsigned avail = end - next;
    unsigned my;
    if (cnt <= avail) {
        my = next + rank;
        next += cnt;
    } else {
        int leader = __ffsll((long long)m) - 1;
        unsigned base = 0;
        if (lane == (unsigned)leader) base = atomicAdd(head, chunk);
        base = __shfl(base, leader);
        my = rank < avail ? next + rank : base + (rank - avail);
        next = base + (cnt - avail);
        end = base + chunk;
    }
    return my;
}

// Pose constants a lane needs while it traces one ray.
struct LanePose {
    double a, asq, r, theta, phi, r_cap;
    __device__ __forceinline__ void load(const PoseK& P) {
        a = P.a;
        asq = P.asq;
        r = P.r;
        theta = P.theta;
        phi = P.phi;
        r_cap = P.r_cap;
    }
};

// Launch shape of a persistent kernel: one resident set of workgroups
// (Context::tune: block size, resident waves per SIMD), never more than the
// rays need.
struct Launch {
    unsigned blocks, threads;
};

template <class K>
int persistent_launch(Context* c, K kernel, int64_t n_max, Launch* L) {
    const int threads = (int)c->tune.block;
    if (threads != 64 && threads != 128 && threads != 256)
        return fail(BT_ERR_INVALID, "tune: block size must be 64, 128 or 256");
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, threads, 0) != hipSuccess || occ < 1)
        occ = 1;
    const int wpb = threads / 64;
    int waves_cu = occ * wpb;
    if (c->tune.waves > 0 && c->tune.waves * 4 < waves_cu) waves_cu = c->tune.waves * 4;
    int bpc = waves_cu / wpb;
    if (bpc < 1) bpc = 1;
    int64_t blocks = (int64_t)c->num_cus * bpc;
    int64_t need = ((n_max + 63) / 64 + wpb - 1) / wpb;
    if (need < blocks) blocks = need;
    if (blocks < 1) blocks = 1;
    L->blocks = (unsigned)blocks;
    L->threads = (unsigned)threads;
    if (blocks * threads > c->launch_lanes) c->launch_lanes = blocks * threads;
    return BT_OK;
}

// Wave issue priority (s_setprio) of a persistent integrator wave: base,
// raised by one once a lane's ray passes kLongRay attempts and by two past
// kVeryLongRay.  The few rays near the photon orbit run up to the attempt
// cap while the bulk takes tens of attempts; a launch ends with its longest
// ray, so its wave gets first claim on the SIMD's issue slots and the short
// rays of other waves fill the rest.  Only the order of issue changes.
constexpr int kLongRay = 128;
constexpr int kVeryLongRay = 512;

__device__ __forceinline__ void set_wave_prio(int p) {
    p = __builtin_amdgcn_readfirstlane(p);   // uniform: a scalar branch
    if (p <= 0)
        __builtin_amdgcn_s_setprio(0);
    else if (p == 1)
        __builtin_amdgcn_s_setprio(1);
    else if (p == 2)
        __builtin_amdgcn_s_setprio(2);
    else
        __builtin_amdgcn_s_setprio(3);
}

// Updates the wave's priority from its lanes' attempt counts; cur is the
// priority last set (wave-uniform).
__device__ __forceinline__ void update_wave_prio(int base, bool live, int nstp, int& cur) {
    const int boost = __ballot(live && nstp >= kVeryLongRay) ? 2 : __ballot(live && nstp >= kLongRay) ? 1 : 0;
    int want = base + boost;
    if (want > 3) want = 3;
    if (want != cur) {
        set_wave_prio(want);
        cur = want;
    }
}

// Appends the lanes with `take` set to queue Q (one atomic per wave).  Every
// lane of the wave must call it.
__device__ __forceinline__ void queue_push(const QueueK& Q, bool take, uint32_t k, unsigned lane) {
    unsigned long long m = __ballot(take);
    if (!m) return;
    int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == (unsigned)leader) base = atomicAdd(&Q.counters[0], (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (take) Q.slots[base + (unsigned)__popcll(m & ((1ull << lane) - 1))] = k;
}

}  // namespace kerr
}  // namespace bt
