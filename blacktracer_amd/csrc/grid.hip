// grid.hip -- device-resident adaptive grid build: the drop-in for
// Grid::Grid(maxLevel, startLevel, angle, cam, bh) (Grid.h:605-624).
//
// The grid lives in HBM as dense row-major arrays (cell (i, j) at i*M + j):
//   cel     double2  celestial (theta', phi'); (-2,-2) = never traced
//   steps   int32    step index of the last Cash-Karp attempt
//   level   int8     level of the block whose key is (i, j), -1 = none
//   present bitmask  the reference's CamToCel key set
// Per refinement level (Grid.h:465-505) one kernel evaluates refineCheck for
// every block of the level, claims the five new points of a refined block
// with an atomicOr on the bitmask (the dedupe of fillVector, Grid.h:453-459)
// and appends the four children; the ray kernels then trace the claimed
// points straight into the grid.  The refinement set does not depend on
// the order in which blocks are visited, so it equals the reference's.
//
// fixTvertices (Grid.h:103-177) is order dependent (the reference iterates
// an unordered_map).  The pinned order is level ascending, then block key
// ascending.  Within one level, fixes on different horizontal lines touch
// disjoint points, and so do fixes on different vertical lines except that
// a meridian c and its antipode c + M/2 meet through the pole branches of
// checkAdjacentBlock (Grid.h:146-153).  The north pole branch also ties
// row 0 to the meridians: a top edge on row 0 reads the antipodal
// meridian's points in rows 1..gap-1, which that meridian's first segment
// rewrites.  So per level one thread first walks the row-0 blocks in key
// order (left meridian segment, top edge, right meridian segment, as each
// block does), then one thread walks each other horizontal line and one
// each meridian pair (from row gap down), in key order -- which reproduces
// the sequential pinned order exactly.
//
// Compiled with -ffp-contract=off (refinement compares and Hermite
// midpoints round as in the reference).
#include <cmath>
#include <cstring>
#include <vector>

#include "grid_dev.h"

namespace bt {
namespace {

constexpr double PI = 3.141592653589793238;
constexpr double PI2 = 3.141592653589793238 * 2.;

__global__ void init_grid(GridK G, int64_t nm, int64_t nwords) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = t; i < nm; i += stride) {
        G.cel[2 * i] = -2.0;
        G.cel[2 * i + 1] = -2.0;
        G.steps[i] = 0;
        G.level[i] = -1;
    }
    for (int64_t i = t; i < nwords; i += stride) G.present[i] = 0u;
}

__global__ void zero_u32(unsigned* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

__global__ void zero_span(unsigned* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Batched builds keep each grid's refinement counters in a 128-byte line of
// their own: the counters take an atomic per new point, and the grids of a
// batch refine in one launch.
constexpr int kCtrStride = 32;

__global__ void pole_keys(uint32_t* keys, uint32_t south) {
    if (threadIdx.x == 0) {
        keys[0] = 0u;
        keys[1] = south;
    }
}

__global__ void pole_rows(GridK G, int gap, uint32_t* blocks) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint32_t)G.M / (uint32_t)gap) return;
    pole_row_point(G, gap, blocks, t);
}

// integrateFirst(gap), Grid.h:325-338, lattice point t.  The lattice index
// grows with t, so the lanes of a wave whose points share a presence word
// are contiguous: a segmented OR over the wave gathers each run's bits into
// its first lane, which sets them with one atomic (a gap-1 lattice would
// otherwise put 32 atomics on every word).  Every lane of the wave calls
// this (the OR is a wave collective); lanes with live false touch no memory.
__device__ __forceinline__ void lattice_point(const GridK& G, int gap, uint32_t nrows, uint32_t ncols,
                                              uint32_t* trace, uint32_t* blocks, int mark_level, uint32_t t,
                                              bool live) {
    uint32_t word = 0xFFFFFFFFu, bits = 0u;
    if (live) {
        uint32_t i = (uint32_t)gap * (t / ncols + 1);
        uint32_t j = (uint32_t)gap * (t % ncols);
        uint32_t idx = i * (uint32_t)G.M + j;
        trace[t] = idx;
        word = idx >> 5;
        bits = 1u << (idx & 31);
        if (i != (uint32_t)(G.N - 1)) {
            if (mark_level >= 0)
                G.level[idx] = (int8_t)mark_level;
            else
                blocks[t] = idx;   // the excluded row is the last one, so t stays dense
        }
    }
    const int lane = (int)(threadIdx.x & 63u);
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t ow = __shfl_down(word, off, 64);
        uint32_t ob = __shfl_down(bits, off, 64);
        if (lane + off < 64 && ow == word) bits |= ob;
    }
    uint32_t pw = __shfl_up(word, 1, 64);
    if (live && (lane == 0 || pw != word)) atomicOr(&G.present[word], bits);
}

__global__ void start_lattice(GridK G, int gap, uint32_t nrows, uint32_t ncols, uint32_t* trace,
                              uint32_t* blocks, int mark_level) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    lattice_point(G, gap, nrows, ncols, trace, blocks, mark_level, t, t < nrows * ncols);
}

// The set-up of every grid of a batch (init_grid + pole_keys, start_lattice,
// pole_rows), one launch per step for all grids.  Job g owns the items
// [first_g, first_{g+1}) of each launch; its lattice items start on a
// multiple of 64, so that no wavefront spans two grids.
struct SetupJob {
    GridK G;
    uint32_t* keys;     // [north | south | lattice] ray keys
    uint32_t* blocks;   // [pole-row blocks | lattice blocks]
    int64_t init0, lat0, pole0;
    int64_t nm, nwords;
    int gap, mark_level;
    uint32_t nrows, ncols, npole;
};

template <int STEP>
__device__ __forceinline__ const SetupJob& setup_job(const SetupJob* jobs, int njobs, int64_t t) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int64_t f = STEP == 0 ? jobs[mid].init0 : STEP == 1 ? jobs[mid].lat0 : jobs[mid].pole0;
        if (f <= t) lo = mid;
        else hi = mid - 1;
    }
    return jobs[lo];
}

// A job's init items start on a multiple of kInitSpan, so a workgroup of
// 256 threads covers kInitSpan items of one grid: the job is found once per
// workgroup (a uniform search) and each thread fills kInitSpan / 256 points.
constexpr int kInitSpan = 4096;

__global__ __launch_bounds__(256) void init_grid_multi(const SetupJob* jobs, int njobs) {
    const int64_t t0 = (int64_t)blockIdx.x * kInitSpan;
    const SetupJob& J = setup_job<0>(jobs, njobs, t0);
    const GridK& G = J.G;
    const int64_t base = t0 - J.init0;
    for (int k = 0; k < kInitSpan / 256; k++) {
        const int64_t i = base + k * 256 + threadIdx.x;
        if (i >= J.nm) break;
        reinterpret_cast<double2*>(G.cel)[i] = make_double2(-2.0, -2.0);   // cel is 256-byte aligned
        G.steps[i] = 0;
        G.level[i] = -1;
        if (i < J.nwords) G.present[i] = 0u;
        if (i == 0) {
            J.keys[0] = 0u;
            J.keys[1] = (uint32_t)(G.N - 1) * (uint32_t)G.M;
        }
    }
}

// items is a multiple of 64: a wavefront lies wholly inside one job's range
// or wholly past the end.
__global__ void start_lattice_multi(const SetupJob* jobs, int njobs, int64_t items) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = t < items;
    const SetupJob& J = setup_job<1>(jobs, njobs, in ? t : items - 1);
    const int64_t lt = t - J.lat0;
    const bool live = in && lt < (int64_t)J.nrows * J.ncols;
    lattice_point(J.G, J.gap, J.nrows, J.ncols, J.keys + 1 + J.G.eq, J.blocks + J.npole, J.mark_level,
                  live ? (uint32_t)lt : 0u, live);
}

__global__ void pole_rows_multi(const SetupJob* jobs, int njobs, int64_t items) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= items) return;
    const SetupJob& J = setup_job<2>(jobs, njobs, t);
    const int64_t lt = t - J.pole0;
    if (lt >= (int64_t)J.npole) return;
    pole_row_point(J.G, J.gap, J.blocks, (uint32_t)lt);
}

__device__ __forceinline__ void claim(const GridK& G, uint32_t i, uint32_t j, uint32_t* trace,
                                      unsigned* ntrace) {
    uint32_t idx = i * (uint32_t)G.M + j;
    uint32_t bit = 1u << (idx & 31);
    uint32_t old = atomicOr(&G.present[idx >> 5], bit);
    if (!(old & bit)) trace[atomicAdd(ntrace, 1u)] = idx;
}

// adaptiveBlockIntegration's loop body (Grid.h:476-497) with refineCheck
// (Grid.h:410-445) for one block of one level.
__device__ __forceinline__ void refine_block(const GridK& G, uint32_t key, int level, double thresh,
                                             uint32_t* next_blocks, uint32_t* trace, unsigned* counters);

__device__ __forceinline__ void refine_block_wave(const GridK& G, bool live, uint32_t key, int level,
                                                  double thresh, uint32_t* next_blocks, uint32_t* trace,
                                                  unsigned* counters, bool uniform);

__global__ void refine_level(GridK G, const uint32_t* blocks, unsigned nb, int level,
                             double thresh, uint32_t* next_blocks, uint32_t* trace,
                             unsigned* counters) {
    unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = q < nb;
    refine_block_wave(G, live, live ? blocks[q] : 0u, level, thresh, next_blocks, trace, counters, true);
}

// One level of every grid of a batch in one launch: job g covers threads
// [item0, item0 + nb) of the grid.
struct RefineJob {
    GridK G;
    const uint32_t* blocks;
    uint32_t* next;
    uint32_t* keys;
    unsigned* ctr;
    unsigned nb, item0;
    int level;
};

__global__ __launch_bounds__(256) void refine_level_multi(const RefineJob* jobs, int njobs, unsigned items,
                                                          double thresh) {
    const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned tc = t < items ? t : items - 1;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].item0 <= tc) lo = mid;
        else hi = mid - 1;
    }
    const RefineJob J = jobs[lo];   // by value: the job's pointers stay in registers
    const unsigned q = tc - J.item0;
    const bool live = t < items && q < J.nb;
    // a wave inside one grid's job shares that grid's counters
    const bool uniform = __ballot(lo != __shfl(lo, 0)) == 0;
    refine_block_wave(J.G, live, live ? J.blocks[q] : 0u, J.level, thresh, J.next, J.keys, J.ctr, uniform);
}

__device__ __forceinline__ void refine_block(const GridK& G, uint32_t key, int level, double thresh,
                                             uint32_t* next_blocks, uint32_t* trace, unsigned* counters) {
    BlockRefine R;
    if (!refine_test(G, key, level, thresh, R)) return;
    const uint32_t M = (uint32_t)G.M;
#pragma unroll
    for (int m = 0; m < 5; m++) claim(G, R.pts[m] / M, R.pts[m] % M, trace, &counters[0]);
    unsigned base = atomicAdd(&counters[1], 4u);
#pragma unroll
    for (int m = 0; m < 4; m++) next_blocks[base + m] = R.kids[m];
}

// refine_block for a whole wave (every lane calls it; `live` lanes hold a
// block).  With `uniform` (all lanes share the counters) the new points and
// the next level's blocks are appended with one atomic per wave each;
// otherwise lane by lane.  Which slot a point gets does not matter: the
// next level traces every point and refines every block.
__device__ __forceinline__ void refine_block_wave(const GridK& G, bool live, uint32_t key, int level,
                                                  double thresh, uint32_t* next_blocks, uint32_t* trace,
                                                  unsigned* counters, bool uniform) {
    if (!uniform) {
        if (live) refine_block(G, key, level, thresh, next_blocks, trace, counters);
        return;
    }
    const unsigned lane = __lane_id();
    BlockRefine R;
    bool refine = false;
    unsigned newmask = 0;
    if (live) {
        refine = refine_test(G, key, level, thresh, R);
        if (refine) {
#pragma unroll
            for (int m = 0; m < 5; m++) {
                const uint32_t bit = 1u << (R.pts[m] & 31);
                if (!(atomicOr(&G.present[R.pts[m] >> 5], bit) & bit)) newmask |= 1u << m;
            }
        }
    }
    // new points: one reservation for the wave
    const unsigned nnew = (unsigned)__popc(newmask);
    unsigned inc = nnew;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        unsigned o = __shfl_up(inc, off);
        if ((int)lane >= off) inc += o;
    }
    const unsigned tot = __shfl(inc, 63);
    unsigned base = 0;
    if (lane == 0 && tot) base = atomicAdd(&counters[0], tot);
    base = __shfl(base, 0) + inc - nnew;
#pragma unroll
    for (int m = 0; m < 5; m++)
        if ((newmask >> m) & 1u) trace[base++] = R.pts[m];
    // the refined blocks' four children
    const unsigned long long rm = __ballot(refine);
    unsigned nbase = 0;
    if (lane == 0 && rm) nbase = atomicAdd(&counters[1], 4u * (unsigned)__popcll(rm));
    nbase = __shfl(nbase, 0) + 4u * (unsigned)__popcll(rm & ((1ull << lane) - 1));
    if (refine) {
#pragma unroll
        for (int m = 0; m < 4; m++) next_blocks[nbase + m] = R.kids[m];
    }
}

// Grid.h:503-504
__global__ void final_level(GridK G, const uint32_t* blocks, unsigned nb, int level) {
    unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nb) G.level[blocks[q]] = (int8_t)level;
}

// ---------------------------------------------------------------- T-vertices

// Grid.h:183-204 with tension = bias = 0, on one component.
__device__ __forceinline__ double hermite1(double v, double x0, double x1, double x2, double x3) {
    double const v2 = v * v;
    double const v3 = v * v2;
    double const aa = (1.0 + 0.0) * (1.0 - 0.0) / 2.0;
    double const bb = (1.0 - 0.0) * (1.0 - 0.0) / 2.0;
    double const u0 = 2.0 * v3 - 3.0 * v2 + 1.0;
    double const u1 = v3 - 2.0 * v2 + v;
    double const u2 = v3 - v2;
    double const u3 = -2.0 * v3 + 3.0 * v2;
    double m0 = (x1 - x0) * aa + (x2 - x1) * bb;
    double m1 = (x2 - x1) * aa + (x3 - x2) * bb;
    return x1 * u0 + m0 * u1 + m1 * u2 + x2 * u3;
}

// Metric.h:333-364 with factor 5 on n phi values.
__device__ __forceinline__ void fix_2pi(double* p, int n) {
    const double hi = PI2 * (1. - 1. / (double)5.f);
    const double lo = PI2 * (1. / (double)5.f);
    bool cross = false;
    for (int q = 0; q < n; q++)
        if (p[q] > hi) cross = true;
    if (!cross) return;
    for (int q = 0; q < n; q++)
        if (p[q] < lo) p[q] += PI2;
}

struct Edge {
    uint32_t ii, jj, ii2, jj2;
    int level, gap;
};

// checkAdjacentBlock (Grid.h:132-177): the reference's pre-order recursion
// run from an explicit stack.  GROUP issues all loads of a node together
// (the line walkers, whose nodes are mostly present); without it the node's
// presence is tested first (the pole walker, which meets many absent nodes).
template <bool GROUP>
__device__ void check_adjacent(const GridK& G, Edge root, int udlr) {
    const uint32_t M = (uint32_t)G.M, N = (uint32_t)G.N;
    double* cel = G.cel;
    Edge st[40];
    int sp = 0;
    st[sp++] = root;
    while (sp > 0) {
        Edge f = st[--sp];
        uint32_t i = f.ii + udlr * f.gap / 2;
        uint32_t j = f.jj + (1 - udlr) * f.gap / 2;
        uint32_t jprev = (f.jj - (1 - udlr) * f.gap + M) % M;
        uint32_t jnext = (f.jj + (1 - udlr) * 2 * f.gap) % M;
        uint32_t iprev = f.ii - udlr * f.gap;
        uint32_t inext = f.ii + 2 * udlr * f.gap;
        bool half = false;
        if (f.ii == 0) {
            jprev = (f.jj + M / 2) % M;
            iprev = (uint32_t)f.gap;
        } else if (inext > N - 1) {
            inext = f.ii;
            if (G.eq)
                jnext = (f.jj + M / 2) % M;
            else
                half = true;
        }
        if (!GROUP && !present_at(G, i, j)) continue;
        // every load of the node at once (out-of-grid neighbours read the
        // edge's own first point instead and are masked out), so that a
        // walker pays one memory round trip per node
        const uint32_t i1 = f.ii * M + f.jj, i2 = f.ii2 * M + f.jj2;
        const bool in_m = i < N && j < M, in_p = iprev < N && jprev < M, in_n = inext < N && jnext < M;
        const uint32_t mid = in_m ? i * M + j : i1;
        const uint32_t ip = in_p ? iprev * M + jprev : i1;
        const uint32_t in = in_n ? inext * M + jnext : i1;
        const uint32_t wm = G.present[mid >> 5], wp = G.present[ip >> 5], wn = G.present[in >> 5];
        const double t0 = cel[2 * ip], p0 = cel[2 * ip + 1];
        const double t1 = cel[2 * i1], p1 = cel[2 * i1 + 1];
        const double t2 = cel[2 * i2], p2 = cel[2 * i2 + 1];
        const double t3 = cel[2 * in], p3 = cel[2 * in + 1];
        // an empty statement that uses every loaded value keeps the compiler
        // from sinking the loads below the presence test
        if (GROUP)
            __asm__ volatile("" : : "v"(wm), "v"(wp), "v"(wn), "v"(t0), "v"(p0), "v"(t1), "v"(p1), "v"(t2),
                             "v"(p2), "v"(t3), "v"(p3));
        if (!(in_m && ((wm >> (mid & 31)) & 1u))) continue;
        const bool pp = in_p && ((wp >> (ip & 31)) & 1u), pn = in_n && ((wn >> (in & 31)) & 1u);
        if (pp && pn && (t0 != -1 || p0 != -1) && (t3 != -1 || p3 != -1)) {
            double ph[4] = {p0, p1, p2, p3};
            fix_2pi(ph, 4);
            cel[2 * mid] = hermite1(0.5, t0, t1, t2, half ? PI - t3 : t3);
            cel[2 * mid + 1] = hermite1(0.5, ph[0], ph[1], ph[2], ph[3]);
        } else {
            double ph[2] = {p1, p2};
            fix_2pi(ph, 2);
            cel[2 * mid] = 1. / 2. * (t2 + t1);
            cel[2 * mid + 1] = 1. / 2. * (ph[1] + ph[0]);
        }
        if (f.level + 1 == G.max_level) continue;
        if (sp + 2 > 40) continue;   // unreachable: depth <= max_level
        st[sp++] = Edge{i, j, f.ii2, f.jj2, f.level + 1, f.gap / 2};
        st[sp++] = Edge{f.ii, f.jj, i, j, f.level + 1, f.gap / 2};
    }
}

__device__ __forceinline__ bool block_fixable(const GridK& G, uint32_t i, uint32_t j, int L) {
    uint32_t idx = i * (uint32_t)G.M + j;
    return G.level[idx] == L && !(G.cel[2 * idx + 1] < 0);
}

// fixTvertices (Grid.h:103-119) for every block of level L, in key order
// along each line (see the header comment).  One wavefront per line: the
// lanes test 64 candidate blocks at a time (block of this level with a
// valid corner, and the edge midpoint present -- nothing a fix on this
// level changes), then lane 0 runs checkAdjacentBlock for the hits in
// ascending key order (ballot bit order), which keeps the sequential
// order of the fixes.
// The row-0 blocks of level L in key order, with the three edges each
// repairs that touch the north pole rows (its bottom edge is on row gap and
// is left to that row's walker).  One wavefront.
__device__ __forceinline__ void fix_pole(const GridK& G, int L) {
    const unsigned lane = threadIdx.x & 63;
    const uint32_t M = (uint32_t)G.M;
    const int gap = 1 << (G.max_level - L);
    const uint32_t ug = (uint32_t)gap;
    if (ug > (uint32_t)G.N - 1) return;
    for (uint32_t j0 = 0; j0 < M; j0 += 64 * ug) {
        const uint32_t j = j0 + lane * ug;
        unsigned long long m = __ballot(j < M && block_fixable(G, 0, j, L));
        while (m) {
            const int src = __ffsll((long long)m) - 1;
            m &= m - 1;
            const uint32_t jj = j0 + (uint32_t)src * ug, l = (jj + ug) % M;
            if (lane == 0) {
                check_adjacent<false>(G, Edge{0, jj, ug, jj, L, gap}, 1);
                check_adjacent<false>(G, Edge{0, jj, 0, l, L, gap}, 0);
                check_adjacent<false>(G, Edge{0, l, ug, l, L, gap}, 1);
            }
        }
    }
}

__device__ __forceinline__ void fix_line(const GridK& G, int L, unsigned nrows, unsigned npairs, unsigned t) {
    const unsigned lane = threadIdx.x & 63;
    const uint32_t M = (uint32_t)G.M, N = (uint32_t)G.N;
    const int gap = 1 << (G.max_level - L);
    const uint32_t ug = (uint32_t)gap, half = ug / 2;
    if (t == 0) return;   // row 0: fix_pole
    if (t < nrows) {
        const uint32_t y = t * ug;
        for (int pass = 0; pass < 2; pass++) {
            // pass 0: bottom edges (k_j, k_l) of blocks keyed on row y - gap;
            // pass 1: top edges (ij, i_l) of blocks keyed on row y
            if (pass == 0 && y < ug) continue;
            if (pass == 1 && y + ug > N - 1) continue;
            const uint32_t row = pass == 0 ? y - ug : y;
            for (uint32_t j0 = 0; j0 < M; j0 += 64 * ug) {
                const uint32_t j = j0 + lane * ug;
                bool hit = j < M && block_fixable(G, row, j, L) && present_at(G, y, j + half);
                unsigned long long m = __ballot(hit);
                while (m) {
                    const int src = __ffsll((long long)m) - 1;
                    m &= m - 1;
                    const uint32_t jj = j0 + (uint32_t)src * ug;
                    if (lane == 0) check_adjacent<true>(G, Edge{y, jj, y, (jj + ug) % M, L, gap}, 0);
                }
            }
        }
    } else if (t - nrows < npairs) {
        const uint32_t c = (t - nrows) * ug;
        const uint32_t c2 = c + M / 2;
        uint32_t cand[4] = {c, (c - ug + M) % M, c2, (c2 - ug + M) % M};
        // ascending j (4-element sort network)
#define BT_SWAP(x, y)                   \
    if (cand[x] > cand[y]) {            \
        uint32_t tmp_ = cand[x];        \
        cand[x] = cand[y];              \
        cand[y] = tmp_;                 \
    }
        BT_SWAP(0, 1) BT_SWAP(2, 3) BT_SWAP(0, 2) BT_SWAP(1, 3) BT_SWAP(1, 2)
#undef BT_SWAP
        // lane = 4 * (row offset) + candidate: bit order = rows ascending,
        // then candidates ascending, the sequential loop's order
        const uint32_t s = lane & 3;
        const uint32_t js = cand[s];
        const bool left = (js == c || js == c2);
        const uint32_t col = left ? js : (js + ug) % M;   // the edge's column
        for (uint32_t i0 = 0; i0 + ug <= N - 1; i0 += 16 * ug) {
            const uint32_t i = i0 + (lane >> 2) * ug;
            // the segment from row 0 is fix_pole's
            bool hit = i > 0 && i + ug <= N - 1 && block_fixable(G, i, js, L) && present_at(G, i + half, col);
            unsigned long long m = __ballot(hit);
            while (m) {
                const int src = __ffsll((long long)m) - 1;
                m &= m - 1;
                const uint32_t ii = i0 + (uint32_t)(src >> 2) * ug;
                const uint32_t cc = __shfl(col, src);
                if (lane == 0) check_adjacent<true>(G, Edge{ii, cc, ii + ug, cc, L, gap}, 1);
            }
        }
    }
}

__global__ __launch_bounds__(256) void fix_level(GridK G, int L, unsigned nrows, unsigned npairs) {
    fix_line(G, L, nrows, npairs, (blockIdx.x * blockDim.x + threadIdx.x) >> 6);
}

__global__ __launch_bounds__(64) void fix_level_pole(GridK G, int L) { fix_pole(G, L); }

// The same for several grids in one launch: grid g owns line items
// [item0[g], item0[g+1]) (rows first, then meridian pairs, as above).
struct FixJob {
    GridK G;
    unsigned nrows, npairs, item0;
    int L;
};

// fix_pole for job blockIdx.x (one wavefront each)
__global__ __launch_bounds__(64) void fix_pole_multi(const FixJob* jobs) {
    const FixJob& J = jobs[blockIdx.x];
    fix_pole(J.G, J.L);
}

__global__ __launch_bounds__(256) void fix_level_multi(const FixJob* jobs, int njobs) {
    const unsigned t = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].item0 <= t) lo = mid;
        else hi = mid - 1;
    }
    const FixJob& J = jobs[lo];
    if (t < J.item0) return;
    fix_line(J.G, J.L, J.nrows, J.npairs, t - J.item0);
}

unsigned blocks_for(int64_t n, int tpb) { return (unsigned)((n + tpb - 1) / tpb); }

}  // namespace

int grid_fix_tvertices(Context* c, const GridK& G) {
    for (int L = G.start_level; L < G.max_level; L++) {
        int gap = 1 << (G.max_level - L);
        unsigned nrows = (unsigned)((G.N - 1) / gap + 1);
        unsigned npairs = (unsigned)((G.M / 2) / gap);
        unsigned n = nrows + npairs;
        KTimer t(c, "fix_tvertices");
        fix_level_pole<<<1, 64, 0, c->stream>>>(G, L);
        fix_level<<<blocks_for((int64_t)n * 64, 256), 256, 0, c->stream>>>(G, L, nrows, npairs);
        BT_HIP(hipGetLastError());
    }
    return BT_OK;
}

// fixTvertices of several grids, one launch per level for all of them.
int grid_fix_tvertices_multi(Context* c, const std::vector<bt_grid*>& gs) {
    int lmin = 99, lmax = 0;
    for (bt_grid* g : gs) {
        lmin = std::min(lmin, g->k.start_level);
        lmax = std::max(lmax, g->k.max_level);
    }
    // every level's jobs, uploaded in one copy before the first launch
    std::vector<FixJob> all;
    std::vector<std::pair<size_t, unsigned>> lv;   // (first job, items) per level
    std::vector<FixJob> jobs;
    for (int L = lmin; L < lmax; L++) {
        jobs.clear();
        unsigned items = 0;
        for (bt_grid* g : gs) {
            const GridK& G = g->k;
            if (L < G.start_level || L >= G.max_level) continue;
            int gap = 1 << (G.max_level - L);
            FixJob j;
            j.G = G;
            j.L = L;
            j.nrows = (unsigned)((G.N - 1) / gap + 1);
            j.npairs = (unsigned)((G.M / 2) / gap);
            j.item0 = items;
            items += j.nrows + j.npairs;
            jobs.push_back(j);
        }
        if (jobs.empty()) continue;
        lv.push_back({all.size(), items});
        all.insert(all.end(), jobs.begin(), jobs.end());
    }
    if (all.empty()) return BT_OK;
    void* p;
    if (int rc = c->buffer("fix_jobs", sizeof(FixJob) * all.size(), &p)) return rc;
    FixJob* d_all = (FixJob*)p;
    BT_HIP(hipMemcpyAsync(d_all, all.data(), sizeof(FixJob) * all.size(), hipMemcpyHostToDevice, c->stream));
    for (size_t v = 0; v < lv.size(); v++) {
        const size_t first = lv[v].first, njobs = (v + 1 < lv.size() ? lv[v + 1].first : all.size()) - first;
        KTimer t(c, "fix_tvertices");
        fix_pole_multi<<<(unsigned)njobs, 64, 0, c->stream>>>(d_all + first);
        fix_level_multi<<<blocks_for((int64_t)lv[v].second * 64, 256), 256, 0, c->stream>>>(d_all + first,
                                                                                           (int)njobs);
        BT_HIP(hipGetLastError());
    }
    // the pageable job table is read by the copy: synchronise before it goes
    BT_HIP(hipStreamSynchronize(c->stream));
    return BT_OK;
}

int grid_build(Context* c, bt_grid* g, const bt_grid_params* gp) {
    const int N = g->N, M = g->M, eq = g->eq;
    const int max_level = g->max_level, start_level = g->start_level;
    const int64_t NM = (int64_t)N * M;
    GridK& G = g->k;
    int rc;
    void* p;
    const double thresh = gp ? gp->refine_thresh : 0.015;
    const int fix = gp ? gp->fix_tvertices : 1;

    if ((rc = c->buffer("grid_blocks_a", sizeof(uint32_t) * (NM + 64), &p))) return rc;
    uint32_t* blocks = (uint32_t*)p;
    if ((rc = c->buffer("grid_blocks_b", sizeof(uint32_t) * (NM + 64), &p))) return rc;
    uint32_t* next = (uint32_t*)p;
    if ((rc = c->buffer("grid_trace", sizeof(uint32_t) * (NM + 64), &p))) return rc;
    uint32_t* trace_keys = (uint32_t*)p;
    if ((rc = c->buffer("grid_ctr", 64, &p))) return rc;
    unsigned* ctr = (unsigned*)p;
    if ((rc = c->buffer("grid_stats", sizeof(StatsK), &p))) return rc;
    StatsK* d_stats = (StatsK*)p;
    if (int zr = zero_words((unsigned*)d_stats, (int)(sizeof(StatsK) / 4), c->stream)) return zr;

    const int64_t nwords = (NM + 31) / 32;
    init_grid<<<(unsigned)std::min<int64_t>(blocks_for(NM, 256), 4096), 256, 0, c->stream>>>(G, NM, nwords);
    BT_HIP(hipGetLastError());

    RayIO io;
    std::memset(&io, 0, sizeof(io));
    io.N = N;
    io.M = M;
    io.eq = eq;
    io.out_th = G.cel;
    io.out_ph = G.cel + 1;
    io.ostride = 2;
    io.out_step = G.steps;
    g->rays = 0;

    // raytrace(): the pole rays (Grid.h:295-301) and integrateFirst(gap)
    // (Grid.h:325-338) are independent, so they are traced in one batch:
    // keys [poles | start lattice]; the pole rows are filled afterwards.
    pole_keys<<<1, 64, 0, c->stream>>>(trace_keys, (uint32_t)(N - 1) * (uint32_t)M);
    BT_HIP(hipGetLastError());
    io.keys = trace_keys;
    const int gap0 = 1 << (max_level - start_level);
    const unsigned npole = (unsigned)(M / gap0);
    unsigned nrows = 0;
    for (int i = gap0; i < N - eq; i += gap0) nrows++;
    const unsigned ncols = (unsigned)(M / gap0);
    const int64_t nlat = (int64_t)nrows * ncols;
    const bool uniform = (max_level == start_level);
    start_lattice<<<blocks_for(nlat, 256), 256, 0, c->stream>>>(
        G, gap0, nrows, ncols, trace_keys + 1 + eq, blocks + npole, uniform ? start_level : -1);
    BT_HIP(hipGetLastError());
    if ((rc = trace(c, g->pose, g->prm, 1 + eq + nlat, io, d_stats))) return rc;
    g->rays += 1 + eq + nlat;
    pole_rows<<<blocks_for(npole, 256), 256, 0, c->stream>>>(G, gap0, blocks);
    BT_HIP(hipGetLastError());
    unsigned nb = npole;
    if (!uniform) {
        unsigned lat_blocks = nrows * ncols;
        if (!eq) lat_blocks -= ncols;   // row N-1 is traced but is no block
        nb += lat_blocks;
    }

    // adaptiveBlockIntegration(startLevel)
    int level = start_level;
    bool early = false;
    while (level < max_level) {
        if (nb == 0) {
            early = true;
            break;
        }
        zero_u32<<<1, 64, 0, c->stream>>>(ctr, 2);
        {
            KTimer t(c, "refine_level");
            refine_level<<<blocks_for(nb, 256), 256, 0, c->stream>>>(G, blocks, nb, level, thresh,
                                                                      next, trace_keys, ctr);
        }
        BT_HIP(hipGetLastError());
        unsigned h_ctr[2];
        BT_HIP(hipMemcpyAsync(h_ctr, ctr, sizeof(h_ctr), hipMemcpyDeviceToHost, c->stream));
        BT_HIP(hipStreamSynchronize(c->stream));
        if ((rc = trace(c, g->pose, g->prm, h_ctr[0], io, d_stats))) return rc;
        g->rays += h_ctr[0];
        level++;
        std::swap(blocks, next);
        nb = h_ctr[1];
    }
    if (!early && nb > 0) {
        final_level<<<blocks_for(nb, 256), 256, 0, c->stream>>>(G, blocks, nb, level);
        BT_HIP(hipGetLastError());
    }
    if (fix) {
        if ((rc = grid_fix_tvertices(c, G))) return rc;
    }
    StatsK s;
    BT_HIP(hipMemcpyAsync(&s, d_stats, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    BT_HIP(hipStreamSynchronize(c->stream));
    stats_to_abi(s, &g->stats);
    return BT_OK;
}

// The same build for several grids at once.  The grids advance level by
// level in lockstep; at each step the rays every grid needs (poles, start
// lattice, then each refinement level) go into one multi-pose batch, traced
// by one set of persistent launches with a shared work queue.  Refinement,
// final levels and T-vertex repair stay per grid (they are cheap and
// per-grid ordered).  All grids share the integration parameters of gs[0].
int grid_build_batch(Context* c, const std::vector<bt_grid*>& gs, const bt_grid_params* gp,
                     StatsK* h_stats) {
    const int G = (int)gs.size();
    if (G == 0) return BT_OK;
    if (G > 128) return fail(BT_ERR_INVALID, "grid batch: at most 128 grids");
    const double thresh = gp ? gp->refine_thresh : 0.015;
    const int fix = gp ? gp->fix_tvertices : 1;
    struct Per {
        uint32_t* blocks;
        uint32_t* next;
        uint32_t* keys;
        unsigned nb;
        int level;
        bool active;
        RayIO io;
    };
    std::vector<Per> per(G);
    int rc;
    void* p;
    if ((rc = c->buffer("gbatch_ctr", sizeof(unsigned) * kCtrStride * 128, &p))) return rc;
    unsigned* ctr = (unsigned*)p;
    if ((rc = c->buffer("gbatch_stats", sizeof(StatsK), &p))) return rc;
    StatsK* d_stats = (StatsK*)p;
    if (int zr = zero_words((unsigned*)d_stats, (int)(sizeof(StatsK) / 4), c->stream)) return zr;
    const ParamsK& prm = gs[0]->prm;

    // raytrace() and integrateFirst(gap) of every grid (Grid.h:295-301,
    // 325-338): per grid the keys [poles | start lattice]; the set-up of all
    // grids goes through one job table, one launch per step.
    std::vector<SetupJob> sj(G);
    std::vector<int64_t> counts(G);
    int64_t n_init = 0, n_lat = 0, n_pole = 0;
    for (int k = 0; k < G; k++) {
        bt_grid* g = gs[k];
        GridK& Gk = g->k;
        const int64_t NM = (int64_t)g->N * g->M;
        std::string sfx = "_" + std::to_string(k);
        if ((rc = c->buffer("gbatch_a" + sfx, sizeof(uint32_t) * (NM + 64), &p))) return rc;
        per[k].blocks = (uint32_t*)p;
        if ((rc = c->buffer("gbatch_b" + sfx, sizeof(uint32_t) * (NM + 64), &p))) return rc;
        per[k].next = (uint32_t*)p;
        if ((rc = c->buffer("gbatch_t" + sfx, sizeof(uint32_t) * (NM + 64), &p))) return rc;
        per[k].keys = (uint32_t*)p;
        const int N = g->N, M = g->M, eq = g->eq;
        const int gap0 = 1 << (g->max_level - g->start_level);
        const unsigned npole = (unsigned)(M / gap0);
        unsigned nrows = 0;
        for (int i = gap0; i < N - eq; i += gap0) nrows++;
        const unsigned ncols = (unsigned)(M / gap0);
        const int64_t nlat = (int64_t)nrows * ncols;
        const bool uniform = (g->max_level == g->start_level);
        SetupJob& J = sj[k];
        J.G = Gk;
        J.keys = per[k].keys;
        J.blocks = per[k].blocks;
        J.nm = NM;
        J.nwords = (NM + 31) / 32;
        J.gap = gap0;
        J.mark_level = uniform ? g->start_level : -1;
        J.nrows = nrows;
        J.ncols = ncols;
        J.npole = npole;
        J.init0 = n_init;
        J.lat0 = n_lat;
        J.pole0 = n_pole;
        n_init += (NM + kInitSpan - 1) / kInitSpan * kInitSpan;
        n_lat += (nlat + 63) / 64 * 64;
        n_pole += npole;
        counts[k] = 1 + eq + nlat;
        per[k].nb = npole;
        if (!uniform) {
            unsigned lat_blocks = nrows * ncols;
            if (!eq) lat_blocks -= ncols;   // row N-1 is traced but is no block
            per[k].nb += lat_blocks;
        }
        per[k].level = g->start_level;
        per[k].active = true;
        RayIO& io = per[k].io;
        std::memset(&io, 0, sizeof(io));
        io.N = g->N;
        io.M = g->M;
        io.eq = g->eq;
        io.out_th = Gk.cel;
        io.out_ph = Gk.cel + 1;
        io.ostride = 2;
        io.out_step = Gk.steps;
        io.keys = per[k].keys;
        g->rays = 0;
    }
    // the job table, through pinned staging: one copy, no synchronisation
    const size_t sj_bytes = sizeof(SetupJob) * G;
    void* hs;
    if ((rc = c->stage(sj_bytes, &hs))) return rc;
    std::memcpy(hs, sj.data(), sj_bytes);
    if ((rc = c->buffer("gbatch_setup", sj_bytes, &p))) return rc;
    const SetupJob* d_sj = (const SetupJob*)p;
    BT_HIP(hipMemcpyAsync(p, hs, sj_bytes, hipMemcpyHostToDevice, c->stream));
    if ((rc = c->stage_done())) return rc;
    init_grid_multi<<<(unsigned)(n_init / kInitSpan), 256, 0, c->stream>>>(d_sj, G);
    BT_HIP(hipGetLastError());
    start_lattice_multi<<<blocks_for(n_lat, 256), 256, 0, c->stream>>>(d_sj, G, n_lat);
    BT_HIP(hipGetLastError());
    if (prm.precision == BT_PREC_FP64) {
        // the FP64 path: every level of every grid in one dataflow kernel
        std::vector<FlowGridHost> fh(G);
        for (int k = 0; k < G; k++) {
            FlowGridHost& f = fh[k];
            f.G = gs[k]->k;
            f.P = gs[k]->pose;
            f.keys = per[k].keys;
            f.list0 = per[k].blocks;
            f.list1 = per[k].next;
            f.nkeys = (unsigned)counts[k];
            f.nb0 = per[k].nb;
            f.npole = sj[k].npole;
            f.gap0 = sj[k].gap;
        }
        std::vector<int64_t> rays;
        if ((rc = grid_flow_run(c, fh, prm, thresh, d_stats, &rays))) return rc;
        for (int k = 0; k < G; k++) gs[k]->rays = rays[k];
        if (fix && (rc = grid_fix_tvertices_multi(c, gs))) return rc;
        StatsK s;
        BT_HIP(hipMemcpyAsync(&s, d_stats, sizeof(s), hipMemcpyDeviceToHost, c->stream));
        BT_HIP(hipStreamSynchronize(c->stream));
        if (h_stats) *h_stats = s;
        return BT_OK;
    }
    auto run_batch = [&](const std::vector<int64_t>& counts) -> int {
        BatchHost b;
        for (int k = 0; k < G; k++) {
            if (counts[k] <= 0) continue;
            b.poses.push_back(gs[k]->pose);
            b.ios.push_back(per[k].io);
            b.counts.push_back(counts[k]);
            gs[k]->rays += counts[k];
        }
        return trace_batch(c, b, prm, d_stats);
    };
    // the poles and start lattices of all grids in one batch; the pole rows
    // are filled once the pole rays are done
    if ((rc = run_batch(counts))) return rc;
    pole_rows_multi<<<blocks_for(n_pole, 256), 256, 0, c->stream>>>(d_sj, G, n_pole);
    BT_HIP(hipGetLastError());
    // adaptiveBlockIntegration(startLevel), all grids one level per round
    std::vector<unsigned> h_ctr((size_t)kCtrStride * G);
    std::vector<RefineJob> rjobs;
    if ((rc = c->buffer("gbatch_rjobs", sizeof(RefineJob) * 128, &p))) return rc;
    RefineJob* d_rjobs = (RefineJob*)p;
    while (true) {
        bool any = false;
        for (int k = 0; k < G; k++) {
            Per& q = per[k];
            if (!q.active) continue;
            if (q.level >= gs[k]->max_level || q.nb == 0) {
                // adaptiveBlockIntegration's exits (Grid.h:497-505)
                if (q.nb > 0 && q.level >= gs[k]->max_level) {
                    final_level<<<blocks_for(q.nb, 256), 256, 0, c->stream>>>(gs[k]->k, q.blocks, q.nb, q.level);
                    BT_HIP(hipGetLastError());
                }
                q.active = false;
                continue;
            }
            any = true;
        }
        if (!any) break;
        zero_span<<<(kCtrStride * G + 255) / 256, 256, 0, c->stream>>>(ctr, kCtrStride * G);
        BT_HIP(hipGetLastError());
        rjobs.clear();
        unsigned items = 0;
        for (int k = 0; k < G; k++) {
            Per& q = per[k];
            if (!q.active) continue;
            RefineJob j;
            j.G = gs[k]->k;
            j.blocks = q.blocks;
            j.next = q.next;
            j.keys = q.keys;
            j.ctr = ctr + kCtrStride * k;
            j.nb = q.nb;
            j.item0 = items;
            j.level = q.level;
            items += q.nb;
            rjobs.push_back(j);
        }
        // the previous level's launch has finished (the stream was synchronised)
        BT_HIP(hipMemcpyAsync(d_rjobs, rjobs.data(), sizeof(RefineJob) * rjobs.size(), hipMemcpyHostToDevice,
                              c->stream));
        {
            KTimer t(c, "refine_level");
            refine_level_multi<<<blocks_for(items, 256), 256, 0, c->stream>>>(d_rjobs, (int)rjobs.size(), items,
                                                                              thresh);
            BT_HIP(hipGetLastError());
        }
        BT_HIP(hipMemcpyAsync(h_ctr.data(), ctr, sizeof(unsigned) * h_ctr.size(), hipMemcpyDeviceToHost, c->stream));
        BT_HIP(hipStreamSynchronize(c->stream));
        for (int k = 0; k < G; k++) counts[k] = per[k].active ? (int64_t)h_ctr[(size_t)kCtrStride * k] : 0;
        if ((rc = run_batch(counts))) return rc;
        for (int k = 0; k < G; k++) {
            Per& q = per[k];
            if (!q.active) continue;
            q.level++;
            std::swap(q.blocks, q.next);
            q.nb = h_ctr[(size_t)kCtrStride * k + 1];
        }
    }
    if (fix && (rc = grid_fix_tvertices_multi(c, gs))) return rc;
    StatsK s;
    BT_HIP(hipMemcpyAsync(&s, d_stats, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    BT_HIP(hipStreamSynchronize(c->stream));
    if (h_stats) *h_stats = s;
    return BT_OK;
}

size_t grid_device_bytes(int N, int M) {
    const int64_t NM = (int64_t)N * M;
    const int64_t nwords = (NM + 31) / 32;
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    // the grid's block (see grid_alloc) + its FP32 render grid
    const int sym = (float)M / (float)N > 3 ? 1 : 0;
    const size_t gn1 = sym ? 2 * (size_t)N - 1 : (size_t)N;
    return up(sizeof(double) * 2 * NM) + up(sizeof(int32_t) * NM) + up((size_t)NM) + up(sizeof(uint32_t) * nwords) +
           sizeof(float) * 2 * gn1 * M;
}

int grid_alloc(Context* c, bt_grid* g) {
    const int64_t NM = (int64_t)g->N * g->M;
    const int64_t nwords = (NM + 31) / 32;
    GridK& G = g->k;
    G.N = g->N;
    G.M = g->M;
    G.eq = g->eq;
    G.max_level = g->max_level;
    G.start_level = g->start_level;
    // one block: cel | steps | level | present, each 256-byte aligned
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t b_cel = up(sizeof(double) * 2 * NM), b_steps = up(sizeof(int32_t) * NM),
                 b_level = up((size_t)NM), b_present = up(sizeof(uint32_t) * nwords);
    void* p;
    size_t got;
    if (int rc = c->pool_alloc(b_cel + b_steps + b_level + b_present, &p, &got)) return rc;
    g->block = p;
    g->block_bytes = got;
    char* base = (char*)p;
    G.cel = (double*)base;
    G.steps = (int32_t*)(base + b_cel);
    G.level = (int8_t*)(base + b_cel + b_steps);
    G.present = (uint32_t*)(base + b_cel + b_steps + b_level);
    g->bytes = b_cel + b_steps + b_level + b_present;
    return BT_OK;
}

void grid_release(bt_grid* g) {
    if (g->gridf) (void)hipFree(g->gridf);
    g->gridf = nullptr;
    if (g->block) {
        if (g->ctx) g->ctx->pool_free(g->block, g->block_bytes);
        else (void)hipFree(g->block);
    }
    g->block = nullptr;
    g->block_bytes = 0;
    GridK& G = g->k;
    G.cel = nullptr;
    G.steps = nullptr;
    G.level = nullptr;
    G.present = nullptr;
}

}  // namespace bt
