// grid_flow.hip -- the batched adaptive build on the FP64 path as one
// persistent dataflow kernel: adaptiveBlockIntegration (Grid.h:465-505) of
// every grid of a batch, with no host synchronisation and no launch per
// level.
//
// Work comes in two kinds:
//   rays   a global queue of (grid, key) entries; a lane takes one, derives
//          its initial conditions (Grid.h:514-556) and runs the exact FP64
//          Cash-Karp attempts (Metric.h:253-297) to its end, writing the
//          result straight into its grid;
//   tasks  a global queue of wave-sized jobs: a seed job computes the
//          initial conditions and a-priori capture test (Metric.h:110-174)
//          of up to 64 start rays (poles and start lattice, Grid.h:295-338)
//          and queues the escaping ones; a refine job runs refineCheck
//          (Grid.h:410-445) on up to 64 blocks of one level, claims the new
//          points with an atomicOr on the presence bitmask (fillVector's
//          dedupe, Grid.h:453-459), runs the capture test on them, queues
//          the escaping ones and appends the children to the next level's
//          block list.
// Each grid counts the rays and jobs of its open phase (its start rays, or
// one level's refinement and the rays it creates).  The lane or wave that
// completes the phase's last item closes it: it fills the pole rows
// (Grid.h:303-315) after the start rays, or swaps the block lists, and
// queues the next level's refine jobs, or marks the final level's blocks
// (Grid.h:497-505) and retires the grid.  Grids advance independently, so
// one grid's slowest rays overlap the other grids' work.  The refinement
// set, and so the grid, does not depend on the order in which items run:
// the result equals the level-by-level build's bit for bit.
//
// Every wave exits once all grids have retired.  A wave that finds no work
// sleeps; one idle for longer than kIdleLimit reports an error and every
// wave exits (a safeguard: a correct build never waits that long).
// Compiled with -ffp-contract=off.
#include <cstring>
#include <vector>

#include "grid_dev.h"
#include "kerr_dev.h"

namespace bt {

namespace {

using namespace kerr;

struct FlowGrid {
    GridK G;
    PoseK P;
    uint32_t* keys;
    uint32_t* list[2];
    unsigned nkeys, nb0, npole;
    int gap0;
};

// Mutable per-grid state, one 64-byte line per grid.
struct alignas(64) FlowState {
    unsigned pending;   // outstanding rays and jobs of the open phase
    unsigned nnext;     // blocks appended to the next level's list
    unsigned rays;      // rays traced (captured rays included)
    int level;          // level of the open refine phase; -1 while the start rays run
    unsigned done;
    unsigned pad[11];
};

// Queue and build control, one 128-byte line per group of fields with its
// own traffic: thousands of waves poll the job line and the done line, and
// a shared line would serialise every one of those reads behind the others
// at one L2 channel.
struct FlowCtl {
    unsigned task_head, task_tail;   // jobs: tickets taken by fetch-and-add, positions written by closers
    unsigned pad0[30];
    unsigned long long ray_ht;       // ray queue: head (low word, tickets taken) and tail (high word, reserved)
    unsigned pad1[30];
    unsigned grids_done;
    unsigned error;                  // 1 ray queue full, 2 task queue full, 3 idle limit
    unsigned pad3[30];
};

struct FlowArgs {
    const FlowGrid* grids;
    FlowState* st;
    FlowCtl* ctl;
    unsigned long long* rayq;
    unsigned long long* taskq;
    StatsK* stats;
    ParamsK prm;
    double thresh;
    unsigned rcap, tcap, epoch;
    int ngrids;
};

// Queue entries carry the build's epoch in their top 16 bits, so entries
// left by an earlier build are never taken for this one's.
// ray:  epoch(16) | grid(8) | unused(8) | key(32)
// task: epoch(16) | refine(1) | grid(7) | level(5) | count - 1 (6) | first(29)
__device__ __forceinline__ unsigned long long ray_entry(unsigned epoch, unsigned g, uint32_t key) {
    return ((unsigned long long)epoch << 48) | ((unsigned long long)g << 40) | key;
}
__host__ __device__ __forceinline__ unsigned long long task_entry(unsigned epoch, bool refine, unsigned g,
                                                                  unsigned level, unsigned count,
                                                                  unsigned first) {
    return ((unsigned long long)epoch << 48) | ((unsigned long long)(refine ? 1 : 0) << 47) |
           ((unsigned long long)g << 40) | ((unsigned long long)level << 35) |
           ((unsigned long long)(count - 1) << 29) | first;
}
constexpr unsigned kFirstMax = 1u << 29;
constexpr unsigned kNoTicket = 0xffffffffu;
constexpr unsigned long long kIdleLimit = 1000000000ull;   // wall-clock ticks (100 MHz): 10 s

__device__ __forceinline__ unsigned long long load_u64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned load_u32(const unsigned* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_u32(unsigned* p, unsigned v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void store_f64(double* p, double v) {
    __hip_atomic_store((unsigned long long*)p, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_i32(int32_t* p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int32_t load_i32(const int32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Ordering of the build's cross-wave traffic.  Every value one wave hands
// to another (grid results, pole-row copies, block lists, queue entries,
// counters) is written and read with agent-scope atomics, which the
// hardware keeps coherent across the XCDs; so a wave publishes by waiting
// for its stores to complete before the counter update that releases them,
// and a taker needs only to keep its later loads after that update: both
// are workgroup-scope fences (waits), with no L2 write-back or
// invalidation, which an agent-scope fence would cost on every completion.
__device__ __forceinline__ void publish() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup"); }
__device__ __forceinline__ void take() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup"); }

__device__ __forceinline__ void key_direction(const GridK& G, uint32_t key, double& th, double& ph) {
    const uint32_t i = key / (uint32_t)G.M, j = key % (uint32_t)G.M;
    th = (double)i / (G.N - 1) * PI / (2 - G.eq);
    ph = (double)j / G.M * PI2;
}

__device__ __forceinline__ void store_result(const GridK& G, uint32_t key, double th, double ph, int step) {
    store_f64(&G.cel[2 * (size_t)key], th);
    store_f64(&G.cel[2 * (size_t)key + 1], ph);
    store_i32(&G.steps[key], step);
}

// raytrace()'s pole rows (Grid.h:303-315) with coherent accesses (see
// pole_row_point).
__device__ __forceinline__ void pole_row_point_coh(const GridK& G, int gap, uint32_t* blocks, uint32_t t) {
    const uint32_t j = t * (uint32_t)gap;
    const uint32_t M = (uint32_t)G.M;
    store_f64(&G.cel[2 * j], cel_load<true>(&G.cel[0]));
    store_f64(&G.cel[2 * j + 1], cel_load<true>(&G.cel[1]));
    store_i32(&G.steps[j], load_i32(&G.steps[0]));
    set_present(G, j);
    store_u32(&blocks[t], j);
    if (G.eq) {
        const uint32_t s = (uint32_t)(G.N - 1) * M;
        store_f64(&G.cel[2 * (s + j)], cel_load<true>(&G.cel[2 * s]));
        store_f64(&G.cel[2 * (s + j) + 1], cel_load<true>(&G.cel[2 * s + 1]));
        store_i32(&G.steps[s + j], load_i32(&G.steps[0]));   // the north pole's count (Grid.h:312)
        set_present(G, s + j);
    }
}

// Per-wave statistics, kept in registers (per lane) until the wave exits.
struct WaveStats {
    unsigned rays = 0, integrated = 0, attempts = 0, newton = 0, golden = 0;
};

// Appends the escaping rays (lanes with `take`) of grid g to the ray queue,
// counting them into the grid's open phase first.  Every lane calls it.
__device__ __forceinline__ void push_rays(const FlowArgs& A, unsigned g, bool take, uint32_t key, unsigned lane) {
    const unsigned long long m = __ballot(take);
    if (!m) return;
    const unsigned cnt = (unsigned)__popcll(m);
    unsigned base = 0;
    if (lane == 0) {
        atomicAdd(&A.st[g].pending, cnt);
        base = (unsigned)(atomicAdd(&A.ctl->ray_ht, (unsigned long long)cnt << 32) >> 32);
        if (base + cnt > A.rcap || base + cnt < base) atomicOr(&A.ctl->error, 1u);
    }
    base = __shfl(base, 0);
    if (take && base + cnt <= A.rcap) {
        const unsigned pos = base + (unsigned)__popcll(m & ((1ull << lane) - 1));
        __hip_atomic_store(&A.rayq[pos], ray_entry(A.epoch, g, key), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Initial conditions and the a-priori capture test of the lanes' points of
// grid g (ic_capture's work for one wave): captured rays are written out,
// escaping ones queued.  Every lane calls it.
__device__ __forceinline__ void seed_points(const FlowArgs& A, unsigned g, bool live, uint32_t key,
                                            unsigned lane, WaveStats& ws) {
    const FlowGrid& F = A.grids[g];
    bool escapes = false;
    if (live) {
        double th, ph, pR, pT, b, q;
        key_direction(F.G, key, th, ph);
        initial_conditions(F.P, th, ph, pR, pT, b, q);
        unsigned newton = 0, golden = 0;
        escapes = A.prm.apriori ? check_celest(F.P, pR, b, q, A.prm.newton_cap, newton, golden) : true;
        if (!escapes) store_result(F.G, key, -1, -1, 0);
        ws.newton += newton;
        ws.golden += golden;
        ws.rays++;
        ws.integrated += escapes;
    }
    const unsigned long long ml = __ballot(live);
    if (lane == 0 && ml) atomicAdd(&A.st[g].rays, (unsigned)__popcll(ml));
    push_rays(A, g, escapes, key, lane);
}

// Closes grid g's open phase (its last item has completed).  Every lane of
// the wave calls it with the same g.
__device__ void close_grid(const FlowArgs& A, unsigned g, unsigned lane) {
    take();   // after the phase's last completion
    const FlowGrid& F = A.grids[g];
    FlowState& S = A.st[g];
    int level = (int)load_u32((const unsigned*)&S.level);
    unsigned nb;
    if (level < 0) {
        // the start rays are done: raytrace()'s pole rows
        for (unsigned t = lane; t < F.npole; t += 64) pole_row_point_coh(F.G, F.gap0, F.list[0], t);
        level = F.G.start_level;
        nb = F.nb0;
    } else {
        level++;
        nb = load_u32(&S.nnext);
    }
    const uint32_t* list = F.list[(level - F.G.start_level) & 1];
    if (lane == 0) {
        store_u32(&S.nnext, 0u);
        store_u32((unsigned*)&S.level, (unsigned)level);
    }
    if (level >= F.G.max_level || nb == 0) {
        // adaptiveBlockIntegration's exits (Grid.h:497-505)
        if (nb > 0)
            for (unsigned t = lane; t < nb; t += 64) F.G.level[load_u32(&list[t])] = (int8_t)level;
        publish();
        if (lane == 0) {
            store_u32(&S.done, 1u);
            atomicAdd(&A.ctl->grids_done, 1u);
        }
        return;
    }
    const unsigned nchunks = (nb + 63) / 64;
    if (lane == 0) store_u32(&S.pending, nchunks);
    publish();
    unsigned base = 0;
    if (lane == 0) {
        base = atomicAdd(&A.ctl->task_tail, nchunks);
        if (base + nchunks > A.tcap || base + nchunks < base) atomicOr(&A.ctl->error, 2u);
    }
    base = __shfl(base, 0);
    if (base + nchunks <= A.tcap)
        for (unsigned t = lane; t < nchunks; t += 64) {
            const unsigned first = t * 64;
            const unsigned cnt = nb - first < 64 ? nb - first : 64;
            __hip_atomic_store(&A.taskq[base + t], task_entry(A.epoch, true, g, (unsigned)level, cnt, first),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
}

// Completion of one item of grid g's open phase by this wave (lane 0
// decrements); closes the phase when it was the last.  Every lane calls it.
__device__ __forceinline__ void complete_item(const FlowArgs& A, unsigned g, unsigned lane) {
    publish();
    unsigned old = 0;
    if (lane == 0) old = atomicSub(&A.st[g].pending, 1u);
    old = __shfl(old, 0);
    if (old == 1) close_grid(A, g, lane);
}

// One task (wave-cooperative).
__device__ void run_task(const FlowArgs& A, unsigned long long e, unsigned lane, uint32_t* lds_pts,
                         WaveStats& ws) {
    const bool refine = (e >> 47) & 1;
    const unsigned g = (unsigned)(e >> 40) & 0x7f;
    const int level = (int)((e >> 35) & 0x1f);
    const unsigned cnt = (unsigned)((e >> 29) & 0x3f) + 1;
    const unsigned first = (unsigned)(e & (kFirstMax - 1));
    const FlowGrid& F = A.grids[g];
    const bool live = lane < cnt;
    if (!refine) {
        seed_points(A, g, live, live ? F.keys[first + lane] : 0u, lane, ws);
    } else {
        const uint32_t* list = F.list[(level - F.G.start_level) & 1];
        uint32_t* next = F.list[(level - F.G.start_level + 1) & 1];
        BlockRefine R;
        bool ref = false;
        unsigned newmask = 0;
        if (live) {
            ref = refine_test<true>(F.G, load_u32(&list[first + lane]), level, A.thresh, R);
            if (ref) {
#pragma unroll
                for (int m = 0; m < 5; m++) {
                    const uint32_t bit = 1u << (R.pts[m] & 31);
                    if (!(atomicOr(&F.G.present[R.pts[m] >> 5], bit) & bit)) newmask |= 1u << m;
                }
            }
        }
        // the refined blocks' children, one reservation for the wave
        const unsigned long long rm = __ballot(ref);
        unsigned nbase = 0;
        if (lane == 0 && rm) nbase = atomicAdd(&A.st[g].nnext, 4u * (unsigned)__popcll(rm));
        nbase = __shfl(nbase, 0) + 4u * (unsigned)__popcll(rm & ((1ull << lane) - 1));
        if (ref) {
#pragma unroll
            for (int m = 0; m < 4; m++) store_u32(&next[nbase + m], R.kids[m]);
        }
        // the new points, compacted through LDS, then seeded 64 at a time
        const unsigned nnew = (unsigned)__popc(newmask);
        unsigned inc = nnew;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned o = __shfl_up(inc, off);
            if ((int)lane >= off) inc += o;
        }
        const unsigned tot = __shfl(inc, 63);
        unsigned w = inc - nnew;
#pragma unroll
        for (int m = 0; m < 5; m++)
            if ((newmask >> m) & 1u) lds_pts[w++] = R.pts[m];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (unsigned r0 = 0; r0 < tot; r0 += 64) {
            const bool lv = r0 + lane < tot;
            seed_points(A, g, lv, lv ? lds_pts[r0 + lane] : 0u, lane, ws);
        }
        __builtin_amdgcn_wave_barrier();
    }
    complete_item(A, g, lane);
}

// Two waves per SIMD: the exact attempt's registers plus the queue and
// task state come to more than 256, which the bound spills (120 bytes per
// lane) rather than drop to one wave.
__global__ __launch_bounds__(256, 2) void grid_flow(FlowArgs A) {
    __shared__ uint32_t s_pts[4][320];
    const unsigned lane = lane_id();
    uint32_t* lds_pts = s_pts[(threadIdx.x >> 6) & 3];
    const ParamsK& prm = A.prm;
    WaveStats ws;
    int cur_prio = 0;

    // the first wave lifts the start token of every grid
    if (blockIdx.x == 0 && threadIdx.x < 64)
        for (int g = 0; g < A.ngrids; g++) complete_item(A, (unsigned)g, lane);

    // this lane's ray
    int rg = -1;                 // grid of the lane's ray, -1 = none
    uint32_t rkey = 0;
    double a = 0, asq = 0, r_cap = 0;   // the ray's pose (the rest is read where needed)
    double var[5];
    double z = 0, h = 0, b = 0, q = 0;
    int nstp = 0;
    // the wave's tickets (wave-uniform): a task position, and a run of ray
    // positions [ray_next, ray_end) not yet handed to lanes
    unsigned task_t = kNoTicket;
    unsigned ray_next = 0, ray_end = 0;
    unsigned tail_seen = 0;              // the ray queue's tail as last read
    unsigned long long idle_since = 0;
    unsigned iter = 0, nap = 1;          // iterations; idle back-off

    while (true) {
        iter++;
        const bool busy = __ballot(rg >= 0) != 0;   // some lane runs a ray
        if ((iter & 31) == 0 && load_u32(&A.ctl->error)) break;
        // New tickets are taken by a wave that runs no ray, and by a busy one
        // every 4th round: tickets are handed out by fetch-and-add, which
        // never fails, and a waiting wave reads only its own task slot and,
        // for rays, the queue's tail.
        const bool look = !busy || (iter & 3) == 0;
        bool worked = false;
        // 1. the task ticket: taken when the wave holds none, its slot read
        // every round (a refine task is on its grid's critical path)
        if (look && task_t == kNoTicket) {
            unsigned t = 0;
            if (lane == 0) t = atomicAdd(&A.ctl->task_head, 1u);
            task_t = __shfl(t, 0);
        }
        if (task_t < A.tcap) {
            unsigned long long e = 0;
            if (lane == 0) e = load_u64(&A.taskq[task_t]);
            e = __shfl(e, 0);
            if ((unsigned)(e >> 48) == A.epoch) {
                task_t = kNoTicket;
                take();
                run_task(A, e, lane, lds_pts, ws);
                worked = true;
            }
        }
        // 2. rays: a wave with no open run takes as many consecutive
        // positions as it has free lanes (a busy wave once 16 are free); its
        // free lanes take the run's positions as the producers' tail passes
        // them, so rays arrive packed into waves
        {
            const unsigned long long m = __ballot(rg < 0);
            const unsigned nfree = (unsigned)__popcll(m);
            if (look && ray_next == ray_end && (nfree >= 16 || (!busy && nfree))) {
                unsigned base = 0;
                if (lane == 0) base = (unsigned)atomicAdd(&A.ctl->ray_ht, (unsigned long long)nfree);
                base = __shfl(base, 0);
                ray_next = base;
                ray_end = base + nfree;
            }
            if (ray_next < ray_end && ray_next >= tail_seen && (look || (iter & 1) == 0)) {
                unsigned tl = 0;
                if (lane == 0) tl = (unsigned)(load_u64(&A.ctl->ray_ht) >> 32);
                tail_seen = min(__shfl(tl, 0), A.rcap);
            }
            const unsigned avail = ray_next < tail_seen ? min(ray_end, tail_seen) - ray_next : 0;
            const unsigned n = min(avail, nfree);
            if (n) {
                const unsigned rank = (unsigned)__popcll(m & ((1ull << lane) - 1));
                if (rg < 0 && rank < n) {
                    // the producer reserved the position before writing it
                    const unsigned pos = ray_next + rank;
                    unsigned long long e = load_u64(&A.rayq[pos]);
                    const unsigned long long t0 = wall_clock64();
                    while ((unsigned)(e >> 48) != A.epoch) {
                        if (wall_clock64() - t0 > kIdleLimit) break;
                        __builtin_amdgcn_s_sleep(1);
                        e = load_u64(&A.rayq[pos]);
                    }
                    if ((unsigned)(e >> 48) != A.epoch) {
                        atomicOr(&A.ctl->error, 3u);
                    } else {
                        rg = (int)((e >> 40) & 0xff);
                        rkey = (uint32_t)e;
                        const FlowGrid& F = A.grids[rg];
                        a = F.P.a;
                        asq = F.P.asq;
                        r_cap = F.P.r_cap;
                        double th, ph;
                        key_direction(F.G, rkey, th, ph);
                        initial_conditions(F.P, th, ph, var[3], var[4], b, q);
                        var[0] = F.P.r;
                        var[1] = F.P.theta;
                        var[2] = F.P.phi;
                        z = 0.0;
                        h = prm.h0;
                        nstp = 0;
                    }
                }
                ray_next += n;
                worked = true;
            }
        }
        const bool any = __ballot(rg >= 0) != 0;
        if (!any) {
            if (worked) {
                idle_since = 0;
                nap = 1;
                continue;
            }
            // every 4th idle round: both words of the done line in one read
            if ((iter & 3) == 0) {
                const unsigned long long de = load_u64((const unsigned long long*)&A.ctl->grids_done);
                if ((unsigned)(de >> 32) || (unsigned)de >= (unsigned)A.ngrids) break;
            }
            const unsigned long long now = wall_clock64();
            if (!idle_since) idle_since = now;
            if (now - idle_since > kIdleLimit) {
                if (lane == 0) atomicOr(&A.ctl->error, 3u);
                break;
            }
            // back off: 1, 2, 4, 8, 16 sleeps of 32 units of 64 cycles
            // (about 14 us at most at 2.4 GHz)
            for (unsigned k = 0; k < nap; k++) __builtin_amdgcn_s_sleep(32);
            if (nap < 16) nap *= 2;
            continue;
        }
        idle_since = 0;
        nap = 1;
        // 3. one attempt of every lane's ray (rk_f64's exact mode)
        update_wave_prio(0, rg >= 0, nstp, cur_prio);
        bool finished = false;
        if (rg >= 0) {
            attempt(a, asq, var, z, h, prm.eps, b, q);
            ws.attempts++;
            const int step = nstp;
            nstp++;
            const bool esc = z <= prm.z_end || (prm.r_escape > 0 && var[0] > prm.r_escape);
            const bool cap = !esc && r_cap > 0 && var[0] < r_cap;
            const bool out = !esc && !cap && nstp >= prm.max_steps;
            if (esc || cap || out) {
                double th, ph;
                if (cap) {
                    th = -1;
                    ph = -1;
                } else {
                    // out of attempts: varStart was never overwritten (Metric.h:298-302)
                    th = esc ? var[1] : A.grids[rg].P.theta;
                    ph = esc ? var[2] : A.grids[rg].P.phi;
                    wrap_to_pi(th, ph);
                }
                store_result(A.grids[rg].G, rkey, th, ph, step);
                finished = true;
            }
        }
        // 4. finished rays leave their grid's open phase; a lane that ends
        // a phase has the wave close it
        bool closer = false;
        int cg = rg;
        if (__ballot(finished)) publish();
        if (finished) {
            closer = atomicSub(&A.st[rg].pending, 1u) == 1u;
            rg = -1;
        }
        unsigned long long cm = __ballot(closer);
        while (cm) {
            const int leader = __ffsll((long long)cm) - 1;
            const unsigned g = (unsigned)__shfl(cg, leader);
            close_grid(A, g, lane);
            cm &= cm - 1;
        }
    }
    const unsigned long long cnt[6] = {ws.attempts, ws.attempts, ws.rays, ws.integrated, ws.newton, ws.golden};
    unsigned long long* const dst[6] = {&A.stats->attempts, &A.stats->attempts_fp64, &A.stats->rays,
                                        &A.stats->integrated, &A.stats->newton, &A.stats->golden};
    block_add_stats<6>(cnt, dst);
}

__global__ void zero_u64(unsigned long long* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

}  // namespace

int grid_flow_run(Context* c, const std::vector<FlowGridHost>& gh, const ParamsK& prm, double thresh,
                  StatsK* d_stats, std::vector<int64_t>* rays) {
    const int G = (int)gh.size();
    if (G == 0) return BT_OK;
    if (G > 128) return fail(BT_ERR_INVALID, "grid flow: at most 128 grids");
    // queue capacities: every grid point is traced at most once; a level's
    // refine jobs number at most ceil(blocks / 64), and the blocks of all
    // levels at most 4/3 of the grid's points
    uint64_t rcap = 0, tcap = 0;
    std::vector<FlowGrid> fg(G);
    std::vector<unsigned long long> seeds;
    std::vector<unsigned> nseed(G);
    for (int g = 0; g < G; g++) {
        const FlowGridHost& h = gh[g];
        const uint64_t nm = (uint64_t)h.G.N * (uint64_t)h.G.M;
        if (nm >= kFirstMax) return fail(BT_ERR_INVALID, "grid flow: grid too large");
        if (h.G.max_level - h.G.start_level > 31) return fail(BT_ERR_INVALID, "grid flow: too many levels");
        rcap += nm;
        tcap += nm / 48 + 64 + (h.nkeys + 63) / 64;
        FlowGrid& f = fg[g];
        f.G = h.G;
        f.P = h.P;
        f.keys = h.keys;
        f.list[0] = h.list0;
        f.list[1] = h.list1;
        f.nkeys = h.nkeys;
        f.nb0 = h.nb0;
        f.npole = h.npole;
        f.gap0 = h.gap0;
        nseed[g] = (h.nkeys + 63) / 64;
    }
    if (rcap >= (1ull << 32) || tcap >= (1ull << 32))
        return fail(BT_ERR_INVALID, "grid flow: batch too large for one build");
    void* p;
    int rc;
    if ((rc = c->buffer("flow_rayq", sizeof(unsigned long long) * rcap, &p))) return rc;
    unsigned long long* rayq = (unsigned long long*)p;
    if ((rc = c->buffer("flow_taskq", sizeof(unsigned long long) * tcap, &p))) return rc;
    unsigned long long* taskq = (unsigned long long*)p;
    // a queue buffer the context just (re)allocated holds arbitrary bytes,
    // and epochs repeat after 65535 builds: clear the queues then
    c->flow_epoch = (c->flow_epoch + 1) & 0xffffu;
    const bool fresh = rayq != c->flow_rq_seen || taskq != c->flow_tq_seen || rcap > c->flow_rcap ||
                       tcap > c->flow_tcap || c->flow_epoch == 0;
    if (c->flow_epoch == 0) c->flow_epoch = 1;
    if (fresh) {
        zero_u64<<<1024, 256, 0, c->stream>>>(rayq, (size_t)rcap);
        zero_u64<<<256, 256, 0, c->stream>>>(taskq, (size_t)tcap);
        BT_HIP(hipGetLastError());
        c->flow_rq_seen = rayq;
        c->flow_tq_seen = taskq;
        c->flow_rcap = rcap;
        c->flow_tcap = tcap;
    }
    const unsigned epoch = c->flow_epoch;
    for (int g = 0; g < G; g++)
        for (unsigned s = 0; s < nseed[g]; s++) {
            const unsigned first = s * 64;
            const unsigned cnt = gh[g].nkeys - first < 64 ? gh[g].nkeys - first : 64;
            seeds.push_back(task_entry(epoch, false, (unsigned)g, 0, cnt, first));
        }
    if (seeds.size() > tcap) return fail(BT_ERR_INVALID, "grid flow: task queue too small");
    std::vector<FlowState> fs(G);
    for (int g = 0; g < G; g++) {
        std::memset(&fs[g], 0, sizeof(FlowState));
        fs[g].pending = nseed[g] + 1;   // + the start token, lifted by the kernel
        fs[g].level = -1;
    }
    FlowCtl ctl;
    std::memset(&ctl, 0, sizeof(ctl));
    ctl.task_tail = (unsigned)seeds.size();
    // tables through pinned staging, one copy each, no synchronisation
    const size_t o_st = (sizeof(FlowGrid) * G + 255) & ~(size_t)255;
    const size_t o_ctl = o_st + ((sizeof(FlowState) * G + 255) & ~(size_t)255);
    const size_t tbytes = o_ctl + sizeof(FlowCtl);
    if ((rc = c->buffer("flow_tables", tbytes, &p))) return rc;
    char* d_tab = (char*)p;
    void* hs;
    const size_t sbytes = sizeof(unsigned long long) * seeds.size();
    if ((rc = c->stage(tbytes + sbytes, &hs))) return rc;
    std::memcpy(hs, fg.data(), sizeof(FlowGrid) * G);
    std::memcpy((char*)hs + o_st, fs.data(), sizeof(FlowState) * G);
    std::memcpy((char*)hs + o_ctl, &ctl, sizeof(ctl));
    std::memcpy((char*)hs + tbytes, seeds.data(), sbytes);
    BT_HIP(hipMemcpyAsync(d_tab, hs, tbytes, hipMemcpyHostToDevice, c->stream));
    BT_HIP(hipMemcpyAsync(taskq, (char*)hs + tbytes, sbytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = c->stage_done())) return rc;

    FlowArgs A;
    A.grids = (const FlowGrid*)d_tab;
    A.st = (FlowState*)(d_tab + o_st);
    A.ctl = (FlowCtl*)(d_tab + o_ctl);
    A.rayq = rayq;
    A.taskq = taskq;
    A.stats = d_stats;
    A.prm = prm;
    A.thresh = thresh;
    A.rcap = (unsigned)rcap;
    A.tcap = (unsigned)tcap;
    A.epoch = epoch;
    A.ngrids = G;
    {
        KTimer t(c, "grid_flow");
        Launch L;
        if ((rc = persistent_launch(c, grid_flow, (int64_t)1 << 40, &L))) return rc;
        grid_flow<<<L.blocks, L.threads, 0, c->stream>>>(A);
        BT_HIP(hipGetLastError());
    }
    // the grids' ray counts and the outcome
    std::vector<FlowState> out(G);
    FlowCtl octl;
    BT_HIP(hipMemcpyAsync(out.data(), A.st, sizeof(FlowState) * G, hipMemcpyDeviceToHost, c->stream));
    BT_HIP(hipMemcpyAsync(&octl, A.ctl, sizeof(octl), hipMemcpyDeviceToHost, c->stream));
    BT_HIP(hipStreamSynchronize(c->stream));
    if (octl.error)
        return fail(BT_ERR_HIP, octl.error == 3 ? "grid flow: a wave found no work for too long"
                                                : "grid flow: work queue overflow");
    if ((int)octl.grids_done != G) return fail(BT_ERR_HIP, "grid flow: not every grid completed");
    if (rays) {
        rays->resize(G);
        for (int g = 0; g < G; g++) (*rays)[g] = out[g].rays;
    }
    return BT_OK;
}

}  // namespace bt
