// kerr_mixed.hip -- mixed-precision ray path: the bulk of the Cash-Karp
// attempts run in FP32, everything that decides or accumulates runs in FP64.
//
//   FP64  initial conditions, constants of motion (b, q), the capture test
//         and the critical-curve test (ic_capture<true>, kerr_f64.hip);
//         the initial step-size growth of every ray, until its first
//         attempt with errmax > ERRCON (rk_f64's warm-up mode): there the
//         FP64 error estimate is far below ERRCON while an FP32 one is
//         rounding noise above it, and the two step sequences would part
//         from the start; the state (r, theta, phi, pR, pTheta) and the
//         affine parameter z, to which every accepted FP32 increment is
//         added; at each stage point Delta = (r-1)^2 - (1-a^2), P = r^2 +
//         a^2 - ab and the range reduction of theta, so that no FP32
//         cancellation happens near the horizon or the poles.  Every ray
//         that leaves the FP32 bulk runs the exact FP64 path (rk_f64's exact
//         mode) from its initial state and so ends exactly as there: the
//         rays near the critical curve (from the start, on the auxiliary
//         stream), the equatorial rays and those with b near 0 (which pass
//         close to a pole), the rays whose step collapses
//         (|h| < hmin_fp64) or which near the attempt cap, and the rays
//         whose FP32 step control may part from the FP64 one (an
//         accept/reject decision within the rounding noise of the
//         threshold, or too much noise summed over the ray's attempts; see
//         the noise accounting below).
//   FP32  the six right-hand-side evaluations, stage combinations, error
//         estimate and step-size control.
//
// The right-hand side is the reference's (Metric.h:198-226) regrouped with
// iR = 1/rho^2, iD = 1/Delta and Ph = P/Delta, using partR + R/Delta =
// P^2/Delta (R = P^2 - Delta*partR, Metric.h:96-98):
//   dr/dz      = Delta iR pR
//   dtheta/dz  = iR pTheta
//   dphi/dz    = iR (a Ph - a + b / sin^2)
//   dpR/dz     = iR (2 r Ph - (r-1)(Ph^2 + pR^2)) + r iR^2 X
//   dpTheta/dz = iR cos (b^2 / sin^3 - a^2 sin) - a^2 cos sin iR^2 X
// with X = pTheta^2 + pR^2 Delta - Theta - P Ph + partR (zero on the null
// cone; kept so the flow is the reference's off the cone as well).
// Compiled with -ffp-contract=fast.
#include <cstring>

#include "kerr_dev.h"

namespace bt {
namespace {

using namespace kerr;

// Cash-Karp tableau (Metric.h:41-49), rounded to float.
constexpr float fb21 = (float)b21, fb31 = (float)b31, fb32 = (float)b32, fb41 = (float)b41,
                fb42 = (float)b42, fb43 = (float)b43, fb51 = (float)b51, fb52 = (float)b52,
                fb53 = (float)b53, fb54 = (float)b54, fb61 = (float)b61, fb62 = (float)b62,
                fb63 = (float)b63, fb64 = (float)b64, fb65 = (float)b65;
constexpr float fc1 = (float)c1, fc3 = (float)c3, fc4 = (float)c4, fc6 = (float)c6;
constexpr float fdc1 = (float)dc1, fdc3 = (float)dc3, fdc4 = (float)dc4, fdc5 = (float)dc5,
                fdc6 = (float)dc6;
constexpr double PIO2 = PI / 2;
constexpr double TWO_OVER_PI = 2 / PI;

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// x^p for x > 0 through the hardware log2/exp2 (the step-size control only).
__device__ __forceinline__ float powf_pos(float x, float p) {
    return __builtin_amdgcn_exp2f(p * __builtin_amdgcn_logf(x));
}

// sin and cos of theta (FP64) with FP32 polynomials: theta = n*pi/2 + v,
// |v| <= pi/4, reduced in FP64 so that sin stays relatively accurate near
// both poles.  Taylor series to v^9 / v^10: truncation < 2e-9 on |v| <= pi/4.
__device__ __forceinline__ void sincos_theta(double th, float& s, float& c) {
    double n = rint(th * TWO_OVER_PI);
    float v = (float)fma(-n, PIO2, th);
    float v2 = v * v;
    float sv = v * (1.f + v2 * (-1.f / 6 + v2 * (1.f / 120 + v2 * (-1.f / 5040 + v2 * (1.f / 362880)))));
    float cv = 1.f + v2 * (-0.5f + v2 * (1.f / 24 + v2 * (-1.f / 720 + v2 * (1.f / 40320 + v2 * (-1.f / 3628800)))));
    int qd = (int)n & 3;
    float s0 = (qd & 1) ? cv : sv;
    float c0 = (qd & 1) ? sv : cv;
    s = (qd & 2) ? -s0 : s0;
    c = ((qd + 1) & 2) ? -c0 : c0;
}

// Per-ray constants.
struct RayC {
    float a, asq, b, bsq, q, partR;
    double e;      // 1 - a^2
    double Poff;   // a^2 - a b
};

// Right-hand side at the stage point (r, theta) in FP64 and (pR, pTheta) in FP32.
// NOISE: xn[i] receives the magnitude of the terms summed into o[i], so that
// the FP32 rounding error of o[i] is of order u * xn[i] (u = 2^-24).  The
// null-constraint term X is a difference of terms of order q, r^2 and
// b^2 / sin^2 that cancel to (near) zero, so it dominates xn[3] and xn[4].
template <bool NOISE = false>
__device__ __forceinline__ void derivs_m(const RayC& C, double r_d, double th_d, float pr, float pt,
                                         float* o, float* xn = nullptr) {
    double rm1_d = r_d - 1.0;
    double D_d = fma(rm1_d, rm1_d, -C.e);
    double P_d = fma(r_d, r_d, C.Poff);
    float r = (float)r_d, rm1 = (float)rm1_d, D = (float)D_d, P = (float)P_d;
    float s, c;
    sincos_theta(th_d, s, c);
    float c2 = c * c;
    float rosq = r * r + C.asq * c2;
    float iR = rcp(rosq), iD = rcp(D), is = rcp(s);
    float is2 = is * is;
    float Ph = P * iD;
    float Theta = C.q - c2 * (C.bsq * is2 - C.asq);
    float prsq = pr * pr;
    float X = pt * pt - Theta + C.partR + prsq * D - P * Ph;
    float iR2X = iR * iR * X;
    o[0] = D * iR * pr;
    o[1] = iR * pt;
    o[2] = iR * (C.a * Ph - C.a + C.b * is2);
    o[3] = iR * (2.f * r * Ph - rm1 * (Ph * Ph + prsq)) + r * iR2X;
    o[4] = iR * c * (C.bsq * is2 * is - C.asq * s) - C.asq * c * s * iR2X;
    if (NOISE) {
        const float Xa = pt * pt + fabsf(Theta) + fabsf(C.partR) + prsq * fabsf(D) + fabsf(P * Ph) +
                         C.bsq * is2 * c2;
        const float iR2Xa = iR * iR * Xa;
        xn[0] = fabsf(o[0]);
        xn[1] = fabsf(o[1]);
        xn[2] = iR * (fabsf(C.a * Ph) + fabsf(C.a) + fabsf(C.b * is2));
        xn[3] = iR * (fabsf(2.f * r * Ph) + fabsf(rm1) * (Ph * Ph + prsq)) + fabsf(r) * iR2Xa;
        xn[4] = iR * fabsf(c) * (C.bsq * is2 * fabsf(is) + C.asq * fabsf(s)) + C.asq * fabsf(c * s) * iR2Xa;
    }
}

// Rounding-noise accounting of the FP32 step control.  The FP32 error
// estimate err_i = h * sum_j dc_j k_j carries an absolute rounding error of
// about u |h| sum_j |dc_j| xn_i, so errmax carries
//   nu = u |h| S max_i(xn_i / scal_i) / eps,   S = sum_j |dc_j| = 0.1155,
// (xn from the first stage stands for all six).  The FP64 path decides with
// an errmax free of that noise; the two step sequences stay together only
// while the FP32 decisions agree with the FP64 ones:
//  * an attempt whose errmax lies within margin * (nu + 5 errmax rho) of
//    the accept threshold 1 might be decided the other way in FP64; rho is
//    the relative deviation the noise has left in h (an accepted step's
//    new h moves by 0.2 nu / errmax, a rejected one's by 0.25 of it and of
//    the old deviation; errmax scales as h^5);
//  * the noise summed over a ray's attempts bounds how far its FP32 state
//    and step sizes may have wandered from the FP64 ones;
//  * the rounding the accepted theta and phi increments carry, u |h|
//    (xn_theta + xn_phi) per step, summed over the ray (its drift), bounds
//    the FP32 part of the final angles' error: a ray whose decisions all
//    agree with FP64 can still end an arcsecond off when its path winds
//    far in phi (measured: near-equatorial rays with 2.5e-6 rad of drift
//    ended 5e-6 rad off, rays below 1e-6 well under an arcsecond).
// Any of these hands the ray to the exact FP64 path from its initial
// state.
// Besides nu, errmax carries the plain FP32 rounding of its own evaluation
// (the scale |y| + |h k1| from the rounded state, its reciprocal, the
// products: kErrRound relative), and a new h that of its FP32 storage and
// of the hardware exp2/log2 power (kHRound relative); both enter the margin
// and rho, so that a decision within FP32's own rounding of the threshold
// is not taken in FP32 either.
constexpr float kU32 = 5.9604645e-8f;   // 2^-24
constexpr float kSumAbsDc = 0.11554158f;
constexpr float kErrRound = 4.f * kU32;
constexpr float kHRound = 3.f * kU32;

// Rays confined to the equatorial plane (Carter constant q = 0 up to
// rounding): theta sits at a double root of Theta, a turning point it never
// leaves in FP64, while FP32 rounding of the theta terms kicks it off the
// plane.  They are few and short, so they restart in the exact FP64 path.
constexpr double kPlanarQ = 1e-9;
// Rays with b near 0 (|b| < prm.polar_b) pass within about |b| / sqrt(q)
// of a pole, and the b / sin^2 and b^2 cos / sin^3 terms make the error
// estimate of a step whose stage points straddle the pole hypersensitive to
// the step size: with |b| of 7e-6 to 7e-5, an FP32 step size 8e-5 off the
// FP64 one (within the rho bound) moved errmax by 4% instead of the 0.04%
// that errmax ~ h^5 allows, the next step sizes parted, and the rays ended
// 2-145 arcseconds off (measured attempt by attempt).  They are few (|b|
// below 1e-3 is a thin band of the sky), so they restart in the exact FP64
// path as well.

// One Cash-Karp attempt with the reference's step control (Metric.h:253-297)
// in FP32; an accepted step adds the FP32 increment to the FP64 state.
// Returns false, leaving y, z and h untouched, when the accept/reject
// decision is within the noise margin of the threshold; nu receives the
// noise bound of errmax, rho is updated as described above, and an accepted
// step adds its angle increments' rounding to drift.
__device__ __forceinline__ bool attempt_m(const RayC& C, double* y, double& z, float& h, float inv_eps,
                                          float margin, float& rho, float& nu, float& drift) {
    float k1[5], k2[5], k3[5], k4[5], k5[5], k6[5], xn[5];
    const float pr0 = (float)y[3], pt0 = (float)y[4];
    const float hh = h;
    derivs_m<true>(C, y[0], y[1], pr0, pt0, k1, xn);
    float d[5];
#define STAGE_IN(expr)                                                                      \
    _Pragma("unroll") for (int i = 0; i < 5; i++) d[i] = hh * (expr);
#define STAGE_EVAL(out) derivs_m(C, y[0] + (double)d[0], y[1] + (double)d[1], pr0 + d[3], pt0 + d[4], out)
    STAGE_IN(fb21 * k1[i]);
    STAGE_EVAL(k2);
    STAGE_IN(fb31 * k1[i] + fb32 * k2[i]);
    STAGE_EVAL(k3);
    STAGE_IN(fb41 * k1[i] + fb42 * k2[i] + fb43 * k3[i]);
    STAGE_EVAL(k4);
    STAGE_IN(fb51 * k1[i] + fb52 * k2[i] + fb53 * k3[i] + fb54 * k4[i]);
    STAGE_EVAL(k5);
    STAGE_IN(fb61 * k1[i] + fb62 * k2[i] + fb63 * k3[i] + fb64 * k4[i] + fb65 * k5[i]);
    STAGE_EVAL(k6);
#undef STAGE_IN
#undef STAGE_EVAL
    float errmax = 0.f, xs = 0.f;
#pragma unroll
    for (int i = 0; i < 5; i++) {
        d[i] = hh * (fc1 * k1[i] + fc3 * k3[i] + fc4 * k4[i] + fc6 * k6[i]);
        float err = hh * (fdc1 * k1[i] + fdc3 * k3[i] + fdc4 * k4[i] + fdc5 * k5[i] + fdc6 * k6[i]);
        float iscal = rcp(fabsf((float)y[i]) + fabsf(k1[i] * hh) + (float)TINY);
        errmax = fmaxf(errmax, fabsf(err) * iscal);
        xs = fmaxf(xs, xn[i] * iscal);
    }
    errmax *= inv_eps;
    nu = kU32 * kSumAbsDc * fabsf(hh) * xs * inv_eps;
    if (fabsf(errmax - 1.0f) < margin * (nu + errmax * (5.f * rho + kErrRound))) return false;
    const bool accept = errmax <= 1.0f;
    const float pw = powf_pos(errmax, accept ? (float)PGROW : (float)PSHRNK);   // one call for both branches
    if (accept) {
        drift += kU32 * fabsf(hh) * (xn[1] + xn[2]);
        z += (double)hh;
#pragma unroll
        for (int i = 0; i < 5; i++) y[i] += (double)d[i];
        if (errmax > (float)ERRCON) {
            h = (float)SAFETY * hh * pw;
            rho = 0.2f * (nu * rcp(errmax) + kErrRound) + kHRound;
        } else {
            h = (float)ADAPTIVE * hh;
            rho += kU32;
        }
    } else {
        h = fminf((float)SAFETY * hh * pw, 0.1f * hh);
        rho = 0.25f * (rho + nu * rcp(errmax) + kErrRound) + kHRound;
    }
    return true;
}

// Continues the rays queued in Q (the warm-up queue, walked a second time
// with its own head) from their warm-up state in cont, skipping rays that
// ended in the warm-up; rays whose step collapses, that near the attempt
// cap, that lie in the equatorial plane, have b near 0 or whose FP32 step control may
// part from the FP64 one go to Q64 (the exact FP64 path from the start).
template <class Src>
__global__ __launch_bounds__(256, 5) void rk_mixed(Src S, ParamsK prm, QueueK Q, unsigned* head,
                                                QueueK Q64, ContK cont, StatsK* st, int restart_at,
                                                unsigned chunk) {
    const unsigned lane = lane_id();
    const unsigned qlen = Q.counters[0];
    unsigned res_next = 0, res_end = 0;   // this wave's reserved queue positions
    const float inv_eps = (float)(1.0 / prm.eps);
    const float hmin = (float)prm.hmin_fp64;
    const float margin = prm.decision_margin;
    const float budget = prm.noise_budget > 0 ? prm.noise_budget : __builtin_inff();
    const float dbudget = prm.drift_budget > 0 ? prm.drift_budget : __builtin_inff();

    int64_t slot = -1;
    int sg = 0;
    double r_cap = 0;
    bool exhausted = false;
    double y[5];
    double z = 0;
    float h = 0;
    RayC C;
    if (!Src::kMulti) {
        const PoseK& P = S.pose(0);
        C.a = (float)P.a;
        C.asq = (float)P.asq;
        C.e = 1.0 - P.asq;
        r_cap = P.r_cap;
    }
    int nstp = 0;
    bool planar = false;
    float rho = 0.f, nsum = 0.f, drift = 0.f;
    int cur_prio = 0;
    unsigned long long attempts = 0, restarts = 0;

    while (true) {
        bool need = (slot < 0) && !exhausted;
        unsigned long long m = __ballot(need);
        if (m) {
            unsigned my = claim_positions(m, lane, head, chunk, res_next, res_end);
            if (need) {
                if (my < qlen) {
                    slot = Q.slots[my];
                    nstp = cont.nstp[slot];
                }
                if (my >= qlen) {
                    exhausted = true;
                } else if (nstp < 0) {
                    slot = -1;   // ended in the warm-up; claim again next round
                } else {
                    if (Src::kMulti) {
                        sg = S.seg(slot);
                        const PoseK& P = S.pose(sg);
                        C.a = (float)P.a;
                        C.asq = (float)P.asq;
                        C.e = 1.0 - P.asq;
                        r_cap = P.r_cap;
                    }
                    const double pa = Src::kMulti ? S.pose(sg).a : S.pose(0).a;
                    const double pasq = Src::kMulti ? S.pose(sg).asq : S.pose(0).asq;
#pragma unroll
                    for (int i = 0; i < 5; i++) y[i] = cont.y[i * cont.stride + slot];
                    double b = Q.b[slot], q = Q.q[slot];
                    planar = !(fabs(q) > kPlanarQ) || fabs(b) < prm.polar_b;
                    C.b = (float)b;
                    C.bsq = (float)(b * b);
                    C.q = (float)q;
                    C.partR = (float)(q + (pa - b) * (pa - b));
                    C.Poff = pasq - pa * b;
                    z = cont.y[5 * cont.stride + slot];
                    h = (float)cont.h[slot];
                    rho = kU32;   // h rounded to FP32
                    nsum = 0.f;
                    drift = 0.f;
                }
            }
        }
        // lanes that drew a ray which ended in the warm-up hold no slot but
        // are not exhausted: they claim again next round
        if (__ballot(slot >= 0 || !exhausted) == 0) break;
        update_wave_prio(0, slot >= 0, nstp, cur_prio);
        bool restart = false, flagged = false, done = false;
        if (slot >= 0 && planar) {
            restart = done = true;
        } else if (slot >= 0) {
            float nu;
            const bool decided = attempt_m(C, y, z, h, inv_eps, margin, rho, nu, drift);
            attempts++;
            nsum += nu;
            flagged = !decided || !(nsum <= budget) || !(drift <= dbudget);
            const int step = nstp;
            nstp++;
            bool esc = !flagged && (z <= prm.z_end || (prm.r_escape > 0 && y[0] > prm.r_escape));
            bool cap = !flagged && !esc && r_cap > 0 && y[0] < r_cap;
            restart = !flagged && !esc && !cap && (nstp >= restart_at || !(fabsf(h) >= hmin));
            done = esc || cap || restart || flagged;
            if (esc || cap) {
                double th = -1, ph = -1;
                if (esc) {
                    th = y[1];
                    ph = y[2];
                    wrap_to_pi(th, ph);
                }
                ray_store(S.rio(sg), S.local(sg, slot), th, ph, step);
            }
        }
        queue_push(Q64, restart || flagged, (uint32_t)slot, lane);
        if (restart || flagged) restarts++;
        if (done) slot = -1;
    }
    const unsigned long long cnt[2] = {attempts, restarts};
    unsigned long long* const dst[2] = {&st->attempts, &st->fallback};
    block_add_stats<2>(cnt, dst);
}

}  // namespace

// Below this tolerance FP32 cannot carry the step control: the rounding of
// the FP32 stage sums (unit roundoff 6e-8, relative to |y| + |h y'|) is no
// longer small against eps, so the error estimate turns into noise, steps
// collapse and the FP32 right-hand side's rounding accumulates into the
// angles.  The mixed mode then traces every ray in FP64.
constexpr double kMixedEpsMin = 1e-6;

namespace {
// phase 0: snapshot the integrated-ray count; phase 1: add the rays
// integrated since to the fallback count.
__global__ void stats_mark(StatsK* s, unsigned long long* snap, int phase) {
    if (phase == 0)
        snap[0] = s->integrated;
    else
        s->fallback += s->integrated - snap[0];
}
}  // namespace

template <class Src>
int trace_mixed_src(Context* c, const Src& S, const ParamsK& prm, int64_t n, StatsK* d_stats) {
    if (n <= 0) return BT_OK;
    if (n > 0x7fffffffLL) return fail(BT_ERR_INVALID, "trace_mixed: too many rays in one batch");
    if (prm.eps < kMixedEpsMin) {
        // every ray in FP64; the rays it integrates are counted as fallback rays
        unsigned long long* snap = nullptr;
        int rc;
        if (d_stats) {
            void* p;
            if ((rc = c->buffer("stats_snap", sizeof(unsigned long long), &p))) return rc;
            snap = (unsigned long long*)p;
            stats_mark<<<1, 1, 0, c->stream>>>(d_stats, snap, 0);
            BT_HIP(hipGetLastError());
        }
        if ((rc = trace_f64_src(c, S, prm, n, d_stats))) return rc;
        if (d_stats) {
            stats_mark<<<1, 1, 0, c->stream>>>(d_stats, snap, 1);
            BT_HIP(hipGetLastError());
        }
        return BT_OK;
    }
    // Qw: warm-up (FP64) then FP32 queue; Qc: near-critical rays, the exact
    // FP64 path from the start, traced on the auxiliary stream while the
    // main stream runs the warm-up and the FP32 bulk; Qr: rays the FP32
    // kernel hands back (equatorial, step collapse, attempt cap, noise
    // accounting), the exact FP64 path from the start on the main stream.
    QueueK Qw, Qc, Qr;
    int rc;
    if ((rc = queue_alloc(c, n, &Qw))) return rc;
    void* p;
    Qc = Qr = Qw;
    if ((rc = c->buffer("q_slots_crit", sizeof(uint32_t) * n, &p))) return rc;
    Qc.slots = (uint32_t*)p;
    Qc.counters = Qw.counters + 2 * kCtrLine;   // its own line: length, head
    if ((rc = c->buffer("q_slots_restart", sizeof(uint32_t) * n, &p))) return rc;
    Qr.slots = (uint32_t*)p;
    Qr.counters = Qw.counters + 3 * kCtrLine;
    unsigned* head32 = Qw.counters + kCtrLine;   // the bulk's head on Qw
    ContK cont;
    cont.stride = n;
    if ((rc = c->buffer("cont_y", sizeof(double) * 6 * n, &p))) return rc;
    cont.y = (double*)p;
    if ((rc = c->buffer("cont_h", sizeof(double) * n, &p))) return rc;
    cont.h = (double*)p;
    if ((rc = c->buffer("cont_nstp", sizeof(int32_t) * n, &p))) return rc;
    cont.nstp = (int32_t*)p;

    // rays that near the attempt cap restart in exact FP64, so that a capped
    // ray ends exactly as in the FP64 path
    const int margin = prm.max_steps / 16 > 16 ? prm.max_steps / 16 : 16;
    const int restart_at = prm.max_steps - margin > 1 ? prm.max_steps - margin : 1;
    if ((rc = trace_setup(c, S, prm, n, Qw, &Qc, d_stats))) return rc;
    // fork: critical rays on the auxiliary stream, at most one workgroup per CU
    BT_HIP(hipEventRecord(c->ev_fork, c->stream));
    BT_HIP(hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    const int64_t crit_lanes = (int64_t)c->num_cus * c->tune.block;
    // the critical rays' waves issue ahead of the bulk's (base priority 1)
    rc = trace_f64_queue(c, S, prm, n < crit_lanes ? n : crit_lanes, Qc, d_stats, "rk_f64_critical",
                         nullptr, c->aux, 1);
    // join (also on error, so that nothing is left running on the aux stream)
    auto join = [&]() -> hipError_t {
        const hipError_t e = hipEventRecord(c->ev_join, c->aux);
        const hipError_t w = hipStreamWaitEvent(c->stream, c->ev_join, 0);
        return e != hipSuccess ? e : w;
    };
    if (rc) {
        (void)join();
        return rc;
    }
    if ((rc = trace_f64_queue(c, S, prm, n, Qw, d_stats, "rk_f64_warmup", &cont))) {
        (void)join();
        return rc;
    }
    {
        KTimer t(c, "rk_mixed");
        Launch L;
        if ((rc = persistent_launch(c, rk_mixed<Src>, n, &L))) {
            (void)join();
            return rc;
        }
        rk_mixed<Src><<<L.blocks, L.threads, 0, c->stream>>>(S, prm, Qw, head32, Qr, cont, d_stats, restart_at,
                                                             c->tune.chunk);
        if (hipGetLastError() != hipSuccess) {
            (void)join();
            return fail(BT_ERR_HIP, "rk_mixed launch failed");
        }
    }
    rc = trace_f64_queue(c, S, prm, n, Qr, d_stats, "rk_f64_fallback", nullptr);
    const hipError_t je = join();
    if (rc) return rc;
    BT_HIP(je);
    return BT_OK;
}

template int trace_mixed_src<MultiSrc>(Context*, const MultiSrc&, const ParamsK&, int64_t, StatsK*);

int trace_mixed(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
                StatsK* d_stats) {
    SingleSrc S{pose, io};
    return trace_mixed_src(c, S, prm, n, d_stats);
}

int trace_batch(Context* c, const BatchHost& b, const ParamsK& prm, StatsK* d_stats) {
    const int nseg = (int)b.poses.size();
    if (nseg == 0) return BT_OK;
    if ((int)b.ios.size() != nseg || (int)b.counts.size() != nseg)
        return fail(BT_ERR_INVALID, "trace_batch: inconsistent segment lists");
    std::vector<int64_t> start(nseg + 1, 0);
    for (int s = 0; s < nseg; s++) start[s + 1] = start[s] + b.counts[s];
    const int64_t n = start[nseg];
    if (n == 0) return BT_OK;
    void* p;
    int rc;
    MultiSrc S;
    S.nseg = nseg;
    // the segment table, through pinned staging: one copy, no synchronisation
    const size_t o_ios = (sizeof(PoseK) * nseg + 255) & ~(size_t)255;
    const size_t o_start = o_ios + ((sizeof(RayIO) * nseg + 255) & ~(size_t)255);
    const size_t tbytes = o_start + sizeof(int64_t) * (nseg + 1);
    void* hs;
    if ((rc = c->stage(tbytes, &hs))) return rc;
    std::memcpy(hs, b.poses.data(), sizeof(PoseK) * nseg);
    std::memcpy((char*)hs + o_ios, b.ios.data(), sizeof(RayIO) * nseg);
    std::memcpy((char*)hs + o_start, start.data(), sizeof(int64_t) * (nseg + 1));
    if ((rc = c->buffer("batch_table", tbytes, &p))) return rc;
    S.poses = (const PoseK*)p;
    S.ios = (const RayIO*)((char*)p + o_ios);
    S.start = (const int64_t*)((char*)p + o_start);
    BT_HIP(hipMemcpyAsync(p, hs, tbytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = c->stage_done())) return rc;
    return prm.precision == BT_PREC_MIXED ? trace_mixed_src(c, S, prm, n, d_stats)
                                          : trace_f64_src(c, S, prm, n, d_stats);
}

}  // namespace bt
