// kerr_f64.hip -- the plain FP64 ray path: the drop-in for
// Grid::integration_wrapper (Grid.h:514-556) and Metric.h's integrator.
//
// Two kernels per batch:
//   ic_capture_f64  one lane per ray: initial conditions + a-priori capture
//                   test; captured rays are written out, escaping rays are
//                   appended (staged per wave in LDS) to a work queue with
//                   their constants of motion.
//   rk_f64          persistent wavefronts; one lane traces one ray with
//                   (b, q) and the 5-vector state in registers.  Lanes whose
//                   ray finished are refilled from the queue after a wave
//                   ballot, so a wave never idles on a long ray's neighbours.
// Compiled with -ffp-contract=off: results equal the reference's host code
// up to the last-ulp differences of the device sin/cos/pow.
#include "kerr_dev.h"

namespace bt {
namespace {

using namespace kerr;

// rk_f64 modes: the exact attempt run to the end (the plain FP64 path, and
// every ray the mixed path sends to FP64); attempt_w through the step-size
// growth only (the mixed path's warm-up).
constexpr int kExact = 0, kWarm = 1;

__global__ void zero_counters(unsigned* c, int n) {
    int i = threadIdx.x;
    if (i < n) c[i] = 0;
}

// Wave-staged append: the wave's lanes with `take` write ray k into the
// wave's LDS staging list (nbuf entries so far, wave-uniform); the list goes
// to the queue with one fetch-and-add once it may not hold another round
// (or, with `flush`, whatever it holds).  Every lane of the wave calls it.
// One atomic per CAP - 63 or more rays and no block barrier: a shared
// queue head takes one fetch-and-add every few nanoseconds at best, and a
// block-wide append would hold every wave of the block at each round.
template <unsigned CAP>
__device__ __forceinline__ void wave_append(const QueueK& Qq, uint32_t* buf, unsigned& nbuf, bool take,
                                            uint32_t k, unsigned lane, bool flush) {
    const unsigned long long m = __ballot(take);
    if (take) buf[nbuf + (unsigned)__popcll(m & ((1ull << lane) - 1))] = k;
    nbuf += (unsigned)__popcll(m);
    if (nbuf > CAP - 64 || (flush && nbuf)) {
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(&Qq.counters[0], nbuf);
        base = __shfl(base, 0);
        for (unsigned i = lane; i < nbuf; i += 64) Qq.slots[base + i] = buf[i];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        nbuf = 0;
    }
}

// Initial conditions and a-priori capture test, grid-stride over the rays
// (counters stay in registers until the end).  SPLIT = false: every
// escaping ray goes to Q.  SPLIT = true (mixed precision): rays whose
// (b, q) lie within crit_eps of the critical curve (Euclidean distance in
// the (b, q) plane) go to Q64 and are traced in FP64 throughout; the rest
// go to Q.  Queue entries and per-ray arrays use the global ray index;
// each wave stages its queue entries in LDS (wave_append), so the order of
// the entries is not the order of the rays.
constexpr unsigned kStageQ = 1024, kStageC = 256;

template <bool SPLIT, class Src>
__global__ __launch_bounds__(256) void ic_capture(Src S, ParamsK prm, int64_t n, QueueK Q,
                                                  QueueK Q64, StatsK* st) {
    __shared__ uint32_t s_stage_q[4][kStageQ];
    __shared__ uint32_t s_stage_c[SPLIT ? 4 : 1][SPLIT ? kStageC : 1];
    const unsigned lane = lane_id();
    const unsigned wave = (threadIdx.x >> 6) & 3;
    unsigned nq = 0, nc = 0;
    unsigned long long c_rays = 0, c_int = 0, c_nw = 0, c_gd = 0, c_cr = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k0 = (int64_t)blockIdx.x * blockDim.x; k0 < n; k0 += stride) {
        const int64_t k = k0 + threadIdx.x;
        unsigned newton = 0, golden = 0;
        bool live = k < n;
        bool escapes = false, crit = false;
        double pR = 0, pT = 0, b = 0, q = 0;
        if (live) {
            const int sg = S.seg(k);
            const PoseK& P = S.pose(sg);
            const RayIO& io = S.rio(sg);
            const int64_t kl = S.local(sg, k);
            double th, ph;
            ray_direction(io, kl, th, ph);
            initial_conditions(P, th, ph, pR, pT, b, q);
            double r0 = 2.0;
            escapes = prm.apriori ? check_celest(P, pR, b, q, prm.newton_cap, newton, golden,
                                                 SPLIT ? &r0 : nullptr)
                                  : true;
            if (!escapes) ray_store(io, kl, -1, -1, 0);
            if (SPLIT && escapes) {
                if (!prm.apriori && !P.a_zero) check_bq(P, b, q, prm.newton_cap, newton, &r0);
                crit = !far_from_critical(P, b, q, r0, prm.crit_eps) &&
                       critical_distance(P, b, q, r0) < prm.crit_eps;
            }
        }
        if (escapes) {
            Q.pr[k] = pR;
            Q.pt[k] = pT;
            Q.b[k] = b;
            Q.q[k] = q;
        }
        const bool last = k0 + stride >= n;
        wave_append<kStageQ>(Q, s_stage_q[wave], nq, escapes && !crit, (uint32_t)k, lane, last);
        if (SPLIT) wave_append<kStageC>(Q64, s_stage_c[wave], nc, crit, (uint32_t)k, lane, last);
        c_rays += live;
        c_int += escapes;
        c_nw += newton;
        c_gd += golden;
        c_cr += crit;
    }
    const unsigned long long cnt[5] = {c_rays, c_int, c_nw, c_gd, c_cr};
    unsigned long long* const dst[5] = {&st->rays, &st->integrated, &st->newton, &st->golden, &st->fallback};
    block_add_stats<5>(cnt, dst);
}

// WARM = false: every ray runs to its end (the plain FP64 path).  WARM =
// true (mixed precision): a ray runs in FP64 (attempt_w) through the initial
// step-size growth, where the FP64 error estimate is far below ERRCON and an
// FP32 estimate would only see rounding noise; after its first attempt with
// errmax > ERRCON its state goes to cont.  cont.nstp[ray] = -1 marks a ray
// that ended here.
// The exact attempt fits 2 waves per SIMD (about 250 VGPRs).  The warm-up
// attempt of one pose needs about 170 and is held to 168, which fits 3 with
// no spills; the multi-pose form (its pose in registers as well) would
// spill there and keeps 2.
template <int MODE, class Src>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == kWarm && !Src::kMulti ? 3 : 2))) void rk_f64(
    Src S, ParamsK prm, QueueK Q, StatsK* st, unsigned chunk, ContK cont, int prio) {
    constexpr bool WARM = MODE == kWarm;
    const unsigned lane = lane_id();
    int cur_prio = 0;
    update_wave_prio(prio, false, 0, cur_prio);
    const unsigned qlen = Q.counters[0];
    unsigned* head = &Q.counters[1];
    unsigned res_next = 0, res_end = 0;   // this wave's reserved queue positions

    int64_t slot = -1;
    int sg = 0;                 // segment (pose) of the lane's ray
    LanePose lp;
    if (!Src::kMulti) lp.load(S.pose(0));
    bool exhausted = false;
    double var[5];
    double z = 0, h = 0, b = 0, q = 0;
    int nstp = 0;
    unsigned long long attempts = 0;

    while (true) {
        bool need = (slot < 0) && !exhausted;
        unsigned long long m = __ballot(need);
        if (m) {
            unsigned my = claim_positions(m, lane, head, chunk, res_next, res_end);
            if (need) {
                if (my < qlen) {
                    slot = Q.slots[my];
                    if (Src::kMulti) {
                        sg = S.seg(slot);
                        lp.load(S.pose(sg));
                    }
                    b = Q.b[slot];
                    q = Q.q[slot];
                    // the exact mode reads the camera position from the pose
                    // rather than holding it (registers)
                    const PoseK& P0 = S.pose(sg);
                    var[0] = WARM ? lp.r : P0.r;
                    var[1] = WARM ? lp.theta : P0.theta;
                    var[2] = WARM ? lp.phi : P0.phi;
                    var[3] = Q.pr[slot];
                    var[4] = Q.pt[slot];
                    z = 0.0;
                    h = prm.h0;
                    nstp = 0;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(slot >= 0) == 0) break;
        update_wave_prio(prio, slot >= 0, nstp, cur_prio);
        if (slot >= 0) {
            double em;
            if (WARM) {
                RayW C;
                C.set(lp.a, lp.asq, b, q);
                em = attempt_w(C, var, z, h, prm.eps);
            } else {
                em = attempt(lp.a, lp.asq, var, z, h, prm.eps, b, q);
            }
            attempts++;
            const int step = nstp;
            nstp++;
            bool esc = z <= prm.z_end || (prm.r_escape > 0 && var[0] > prm.r_escape);
            bool cap = !esc && lp.r_cap > 0 && var[0] < lp.r_cap;
            bool out = !esc && !cap && nstp >= prm.max_steps;
            if (esc || cap || out) {
                double th, ph;
                if (cap) {
                    th = -1;
                    ph = -1;
                } else {
                    // out of attempts: varStart was never overwritten (Metric.h:298-302)
                    th = esc ? var[1] : (WARM ? lp.theta : S.pose(sg).theta);
                    ph = esc ? var[2] : (WARM ? lp.phi : S.pose(sg).phi);
                    wrap_to_pi(th, ph);
                }
                ray_store(S.rio(sg), S.local(sg, slot), th, ph, step);
                if (WARM) cont.nstp[slot] = -1;
                slot = -1;
            } else if (WARM && em > ERRCON) {
#pragma unroll
                for (int i = 0; i < 5; i++) cont.y[i * cont.stride + slot] = var[i];
                cont.y[5 * cont.stride + slot] = z;
                cont.h[slot] = h;
                cont.nstp[slot] = nstp;
                slot = -1;
            }
        }
    }
    const unsigned long long cnt[2] = {attempts, attempts};
    unsigned long long* const dst[2] = {&st->attempts, &st->attempts_fp64};
    block_add_stats<2>(cnt, dst);
}

// Operand pair i of the shared-division self-check: signs, exponents
// (2^-200 .. 2^200) and significands from a counter hash, with one class in
// four special: denominators or numerators whose significand is all ones
// but for a few low bits, zero numerators of either sign, and denominators
// that are powers of two.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double operand(unsigned long long h, unsigned long long g) {
    const unsigned long long sign = h >> 63;
    const unsigned long long expo = 1023 - 200 + (g % 401);
    return __longlong_as_double((long long)((sign << 63) | (expo << 52) | (h & 0x000fffffffffffffull)));
}

__global__ void selfcheck_div(int64_t n, unsigned long long seed, unsigned long long* bad) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool miss = false;
    if (i < n) {
        const unsigned long long h1 = mix64(seed + 4 * (unsigned long long)i), h2 = mix64(h1 + 1),
                                 h3 = mix64(h1 + 2), h4 = mix64(h1 + 3);
        double x = operand(h1, h3), d = operand(h2, h4);
        const unsigned cls = (unsigned)(h4 >> 60);
        const unsigned long long ones = 0x000fffffffffffffull - (h3 & 0xff);
        if (cls == 12) d = __longlong_as_double((__double_as_longlong(d) & ~0x000fffffffffffffll) | (long long)ones);
        if (cls == 13) x = __longlong_as_double((__double_as_longlong(x) & ~0x000fffffffffffffll) | (long long)ones);
        if (cls == 14) x = (h3 & 1) ? -0.0 : 0.0;
        if (cls == 15) d = __longlong_as_double(__double_as_longlong(d) & ~0x000fffffffffffffll);
        const double ref = x / d;
        const SharedDiv by(d);
        miss = __double_as_longlong(by.any(x)) != __double_as_longlong(ref);
        const SharedDiv byp(fabs(d));
        miss = miss || __double_as_longlong(byp.pos(x)) != __double_as_longlong(x / fabs(d));
    }
    const unsigned long long m = __ballot(miss);
    if (lane_id() == 0 && m) atomicAdd(bad, (unsigned long long)__popcll(m));
}

}  // namespace

int zero_words(unsigned* p, int n, hipStream_t st) {
    if (n <= 0) return BT_OK;
    if (n > 1024) return fail(BT_ERR_INVALID, "zero_words: at most 1024 words");
    zero_counters<<<1, n <= 64 ? 64 : (n + 63) / 64 * 64, 0, st>>>(p, n);
    BT_HIP(hipGetLastError());
    return BT_OK;
}

int queue_alloc(Context* c, int64_t n, QueueK* Q) {
    void* p;
    int rc;
    if ((rc = c->buffer("q_pr", sizeof(double) * n, &p))) return rc;
    Q->pr = (double*)p;
    if ((rc = c->buffer("q_pt", sizeof(double) * n, &p))) return rc;
    Q->pt = (double*)p;
    if ((rc = c->buffer("q_b", sizeof(double) * n, &p))) return rc;
    Q->b = (double*)p;
    if ((rc = c->buffer("q_q", sizeof(double) * n, &p))) return rc;
    Q->q = (double*)p;
    if ((rc = c->buffer("q_slots", sizeof(uint32_t) * n, &p))) return rc;
    Q->slots = (uint32_t*)p;
    if ((rc = c->buffer("q_ctr", sizeof(unsigned) * kQueueCtrWords, &p))) return rc;
    Q->counters = (unsigned*)p;
    return BT_OK;
}

template <class Src>
int trace_setup(Context* c, const Src& S, const ParamsK& prm, int64_t n, const QueueK& Q,
                const QueueK* Q64, StatsK* d_stats) {
    if (int rc = zero_words(Q.counters, kQueueCtrWords, c->stream)) return rc;   // the q_ctr block
    KTimer t(c, Q64 ? "ic_split" : "ic_capture_f64");
    // one resident generation of blocks, grid-stride: each block and wave
    // then makes its few queue and statistics atomics once
    int occ = 0;
    const hipError_t oe = Q64 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, ic_capture<true, Src>, 256, 0)
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, ic_capture<false, Src>, 256, 0);
    if (oe != hipSuccess || occ < 1) occ = 1;
    int64_t blocks = (n + 255) / 256;
    if (blocks > (int64_t)c->num_cus * occ) blocks = (int64_t)c->num_cus * occ;
    if (Q64)
        ic_capture<true, Src><<<(unsigned)blocks, 256, 0, c->stream>>>(S, prm, n, Q, *Q64, d_stats);
    else
        ic_capture<false, Src><<<(unsigned)blocks, 256, 0, c->stream>>>(S, prm, n, Q, Q, d_stats);
    BT_HIP(hipGetLastError());
    return BT_OK;
}

template <class Src>
int trace_f64_queue(Context* c, const Src& S, const ParamsK& prm, int64_t n_max, const QueueK& Q,
                    StatsK* d_stats, const char* timer, const ContK* cont, hipStream_t stream, int prio) {
    if (!stream) stream = c->stream;
    KTimer t(c, timer, stream);
    Launch L;
    if (cont) {
        if (int rc = persistent_launch(c, rk_f64<kWarm, Src>, n_max, &L)) return rc;
        rk_f64<kWarm, Src><<<L.blocks, L.threads, 0, stream>>>(S, prm, Q, d_stats, c->tune.chunk, *cont, prio);
    } else {
        if (int rc = persistent_launch(c, rk_f64<kExact, Src>, n_max, &L)) return rc;
        rk_f64<kExact, Src><<<L.blocks, L.threads, 0, stream>>>(S, prm, Q, d_stats, c->tune.chunk, ContK{},
                                                                prio);
    }
    BT_HIP(hipGetLastError());
    return BT_OK;
}

template int trace_setup<SingleSrc>(Context*, const SingleSrc&, const ParamsK&, int64_t, const QueueK&,
                                    const QueueK*, StatsK*);
template int trace_setup<MultiSrc>(Context*, const MultiSrc&, const ParamsK&, int64_t, const QueueK&,
                                   const QueueK*, StatsK*);
template int trace_f64_queue<SingleSrc>(Context*, const SingleSrc&, const ParamsK&, int64_t, const QueueK&,
                                        StatsK*, const char*, const ContK*, hipStream_t, int);
template int trace_f64_queue<MultiSrc>(Context*, const MultiSrc&, const ParamsK&, int64_t, const QueueK&,
                                       StatsK*, const char*, const ContK*, hipStream_t, int);

template <class Src>
int trace_f64_src(Context* c, const Src& S, const ParamsK& prm, int64_t n, StatsK* d_stats) {
    if (n <= 0) return BT_OK;
    if (n > 0x7fffffffLL) return fail(BT_ERR_INVALID, "trace_f64: too many rays in one batch");
    QueueK Q;
    int rc;
    if ((rc = queue_alloc(c, n, &Q))) return rc;
    if ((rc = trace_setup(c, S, prm, n, Q, nullptr, d_stats))) return rc;
    return trace_f64_queue(c, S, prm, n, Q, d_stats, "rk_f64", nullptr);
}

template int trace_f64_src<MultiSrc>(Context*, const MultiSrc&, const ParamsK&, int64_t, StatsK*);

int trace_f64(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
              StatsK* d_stats) {
    SingleSrc S{pose, io};
    return trace_f64_src(c, S, prm, n, d_stats);
}

}  // namespace bt

extern "C" int bt_selfcheck_division(bt_context* h, int64_t n, uint64_t seed, int64_t* mismatches) {
    using namespace bt;
    if (!h || !mismatches || n < 0 || n > (int64_t(1) << 30))
        return fail(BT_ERR_INVALID, "bt_selfcheck_division: null argument or n outside [0, 2^30]");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    Context& c = h->impl;
    BT_HIP(hipSetDevice(c.device));
    void* p;
    if (int rc = c.buffer("selfcheck_div", sizeof(unsigned long long), &p)) return rc;
    unsigned long long* bad = (unsigned long long*)p;
    if (int rc = zero_words((unsigned*)bad, 2, c.stream)) return rc;
    if (n > 0) {
        selfcheck_div<<<(unsigned)((n + 255) / 256), 256, 0, c.stream>>>(n, seed, bad);
        BT_HIP(hipGetLastError());
    }
    unsigned long long out = 0;
    BT_HIP(hipMemcpyAsync(&out, bad, sizeof(out), hipMemcpyDeviceToHost, c.stream));
    BT_HIP(hipStreamSynchronize(c.stream));
    *mismatches = (int64_t)out;
    return BT_OK;
}
