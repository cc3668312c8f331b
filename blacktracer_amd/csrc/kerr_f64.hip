// kerr_f64.hip -- the plain FP64 ray path: the drop-in for
// Grid::integration_wrapper (Grid.h:514-556) and Metric.h's integrator.
//
// Two kernels per batch:
//   ic_capture_f64  one lane per ray: initial conditions + a-priori capture
//                   test; captured rays are written out, escaping rays are
//                   appended (wave-aggregated) to a work queue with their
//                   constants of motion.
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

struct QueueK {
    double* pr;
    double* pt;
    double* b;
    double* q;
    uint32_t* slots;    // queue of ray indices
    unsigned* counters; // [0] = queue length, [1] = head
};

__global__ void zero_counters(unsigned* c, int n) {
    int i = threadIdx.x;
    if (i < n) c[i] = 0;
}

__global__ __launch_bounds__(256) void ic_capture_f64(PoseK P, ParamsK prm, int64_t n, RayIO io,
                                                      QueueK Q, StatsK* st) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned lane = lane_id();
    unsigned newton = 0, golden = 0;
    bool live = k < n;
    bool escapes = false;
    double pR = 0, pT = 0, b = 0, q = 0;
    if (live) {
        double th, ph;
        ray_direction(io, k, th, ph);
        initial_conditions(P, th, ph, pR, pT, b, q);
        escapes = prm.apriori ? check_celest(P, pR, b, q, prm.newton_cap, newton, golden) : true;
        if (!escapes) ray_store(io, k, -1, -1, 0);
    }
    unsigned long long m = __ballot(escapes);
    unsigned base = 0;
    if (m) {
        int leader = __ffsll((long long)m) - 1;
        if (lane == (unsigned)leader) base = atomicAdd(&Q.counters[0], (unsigned)__popcll(m));
        base = __shfl(base, leader);
    }
    if (escapes) {
        unsigned pos = base + (unsigned)__popcll(m & ((1ull << lane) - 1));
        Q.slots[pos] = (uint32_t)k;
        Q.pr[k] = pR;
        Q.pt[k] = pT;
        Q.b[k] = b;
        Q.q[k] = q;
    }
    unsigned long long s_rays = wave_sum(live ? 1ull : 0ull);
    unsigned long long s_int = wave_sum(escapes ? 1ull : 0ull);
    unsigned long long s_nw = wave_sum((unsigned long long)newton);
    unsigned long long s_gd = wave_sum((unsigned long long)golden);
    if (lane == 0 && s_rays) {
        atomicAdd(&st->rays, s_rays);
        atomicAdd(&st->integrated, s_int);
        atomicAdd(&st->newton, s_nw);
        atomicAdd(&st->golden, s_gd);
    }
}

__global__ __launch_bounds__(256) void rk_f64(PoseK P, ParamsK prm, RayIO io, QueueK Q,
                                              StatsK* st) {
    const unsigned lane = lane_id();
    const unsigned long long lt_mask = (1ull << lane) - 1;
    const unsigned qlen = Q.counters[0];
    unsigned* head = &Q.counters[1];
    const double a = P.a, asq = P.asq;

    int64_t slot = -1;
    bool exhausted = false;
    double var[5];
    double z = 0, h = 0, b = 0, q = 0;
    int nstp = 0;
    unsigned long long attempts = 0;

    while (true) {
        bool need = (slot < 0) && !exhausted;
        unsigned long long m = __ballot(need);
        if (m) {
            int leader = __ffsll((long long)m) - 1;
            unsigned base = 0;
            if (lane == (unsigned)leader) base = atomicAdd(head, (unsigned)__popcll(m));
            base = __shfl(base, leader);
            if (need) {
                unsigned my = base + (unsigned)__popcll(m & lt_mask);
                if (my < qlen) {
                    slot = Q.slots[my];
                    var[0] = P.r;
                    var[1] = P.theta;
                    var[2] = P.phi;
                    var[3] = Q.pr[slot];
                    var[4] = Q.pt[slot];
                    b = Q.b[slot];
                    q = Q.q[slot];
                    z = 0.0;
                    h = prm.h0;
                    nstp = 0;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(slot >= 0) == 0) break;
        if (slot >= 0) {
            attempt(a, asq, var, z, h, prm.eps, b, q);
            attempts++;
            const int step = nstp;
            nstp++;
            bool esc = z <= prm.z_end || (prm.r_escape > 0 && var[0] > prm.r_escape);
            bool cap = !esc && P.r_cap > 0 && var[0] < P.r_cap;
            bool out = !esc && !cap && nstp >= prm.max_steps;
            if (esc || cap || out) {
                double th, ph;
                if (cap) {
                    th = -1;
                    ph = -1;
                } else {
                    // out of attempts: varStart was never overwritten (Metric.h:298-302)
                    th = esc ? var[1] : P.theta;
                    ph = esc ? var[2] : P.phi;
                    wrap_to_pi(th, ph);
                }
                ray_store(io, slot, th, ph, step);
                slot = -1;
            }
        }
    }
    unsigned long long s = wave_sum(attempts);
    if (lane == 0 && s) {
        atomicAdd(&st->attempts, s);
        atomicAdd(&st->attempts_fp64, s);
    }
}

}  // namespace

int trace_f64(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
              StatsK* d_stats) {
    if (n <= 0) return BT_OK;
    if (n > 0x7fffffffLL) return fail(BT_ERR_INVALID, "trace_f64: too many rays in one batch");
    QueueK Q;
    void* p;
    int rc;
    if ((rc = c->buffer("f64_pr", sizeof(double) * n, &p))) return rc;
    Q.pr = (double*)p;
    if ((rc = c->buffer("f64_pt", sizeof(double) * n, &p))) return rc;
    Q.pt = (double*)p;
    if ((rc = c->buffer("f64_b", sizeof(double) * n, &p))) return rc;
    Q.b = (double*)p;
    if ((rc = c->buffer("f64_q", sizeof(double) * n, &p))) return rc;
    Q.q = (double*)p;
    if ((rc = c->buffer("f64_slots", sizeof(uint32_t) * n, &p))) return rc;
    Q.slots = (uint32_t*)p;
    if ((rc = c->buffer("f64_ctr", 64, &p))) return rc;
    Q.counters = (unsigned*)p;

    zero_counters<<<1, 64, 0, c->stream>>>(Q.counters, 2);
    {
        KTimer t(c, "ic_capture_f64");
        int64_t blocks = (n + 255) / 256;
        ic_capture_f64<<<(unsigned)blocks, 256, 0, c->stream>>>(pose, prm, n, io, Q, d_stats);
    }
    BT_HIP(hipGetLastError());
    {
        KTimer t(c, "rk_f64");
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, rk_f64, 256, 0) != hipSuccess ||
            occ < 1)
            occ = 2;
        int64_t waves_needed = (n + 63) / 64;
        int64_t blocks = (int64_t)c->num_cus * occ;   // exactly one resident wave set
        int64_t need_blocks = (waves_needed + 3) / 4;
        if (need_blocks < blocks) blocks = need_blocks;
        rk_f64<<<(unsigned)blocks, 256, 0, c->stream>>>(pose, prm, io, Q, d_stats);
    }
    BT_HIP(hipGetLastError());
    return BT_OK;
}

}  // namespace bt
