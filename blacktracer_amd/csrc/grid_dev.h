// grid_dev.h -- device helpers of the adaptive grid shared by the level-by-
// level build (grid.hip) and the dataflow build (grid_flow.hip).
// Compiled with -ffp-contract=off in both (the refinement test compares
// squared differences as the reference does).
#pragma once

#include "bt_grid.h"

namespace bt {

__device__ __forceinline__ bool present_at(const GridK& G, uint32_t i, uint32_t j) {
    if (i >= (uint32_t)G.N || j >= (uint32_t)G.M) return false;
    uint32_t idx = i * (uint32_t)G.M + j;
    return (G.present[idx >> 5] >> (idx & 31)) & 1u;
}

__device__ __forceinline__ void set_present(const GridK& G, uint32_t idx) {
    atomicOr(&G.present[idx >> 5], 1u << (idx & 31));
}

// raytrace(), Grid.h:303-315: copy the pole values along the pole rows at
// the start-level spacing and queue the pole-row blocks (column t < M / gap).
__device__ __forceinline__ void pole_row_point(const GridK& G, int gap, uint32_t* blocks, uint32_t t) {
    uint32_t j = t * (uint32_t)gap;
    uint32_t M = (uint32_t)G.M;
    G.cel[2 * j] = G.cel[0];
    G.cel[2 * j + 1] = G.cel[1];
    G.steps[j] = G.steps[0];
    set_present(G, j);
    blocks[t] = j;
    if (G.eq) {
        uint32_t s = (uint32_t)(G.N - 1) * M;
        G.cel[2 * (s + j)] = G.cel[2 * s];
        G.cel[2 * (s + j) + 1] = G.cel[2 * s + 1];
        G.steps[s + j] = G.steps[0];   // the north pole's count (Grid.h:312)
        set_present(G, s + j);
    }
}

// One block of adaptiveBlockIntegration's loop (Grid.h:476-497): refineCheck
// (Grid.h:410-445) on the block whose top-left key is `key` at `level`.
// When it refines, pts holds the five points fillVector asks for (left
// edge midpoint, centre, top edge midpoint, bottom edge midpoint, right
// edge midpoint) and kids the four child blocks' keys; otherwise the block
// is marked with its level.
struct BlockRefine {
    uint32_t pts[5];
    uint32_t kids[4];
};

// Grid values as COHERENT reads them: agent-scope atomic loads, which see
// values other waves stored with agent-scope atomic stores (the dataflow
// build, where the XCDs' L2 caches are not coherent with each other);
// otherwise plain loads (the level-by-level build, where kernel boundaries
// order the accesses).
template <bool COHERENT>
__device__ __forceinline__ double cel_load(const double* p) {
    if (COHERENT)
        return __longlong_as_double((long long)__hip_atomic_load((const unsigned long long*)p, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT));
    return *p;
}

template <bool COHERENT = false>
__device__ __forceinline__ bool refine_test(const GridK& G, uint32_t key, int level, double thresh,
                                            BlockRefine& R) {
    const uint32_t M = (uint32_t)G.M;
    const uint32_t gap = 1u << (G.max_level - level);
    const uint32_t i = key / M;
    const uint32_t j = key % M;
    const uint32_t k = i + gap / 2;
    const uint32_t l = j + gap / 2;
    const uint32_t kk = i + gap, ll = (j + gap) % M;
    const double* c = G.cel;
    double th1 = cel_load<COHERENT>(&c[2 * (i * M + j)]), th2 = cel_load<COHERENT>(&c[2 * (kk * M + j)]);
    double th3 = cel_load<COHERENT>(&c[2 * (i * M + ll)]), th4 = cel_load<COHERENT>(&c[2 * (kk * M + ll)]);
    double ph1 = cel_load<COHERENT>(&c[2 * (i * M + j) + 1]), ph2 = cel_load<COHERENT>(&c[2 * (kk * M + j) + 1]);
    double ph3 = cel_load<COHERENT>(&c[2 * (i * M + ll) + 1]), ph4 = cel_load<COHERENT>(&c[2 * (kk * M + ll) + 1]);
    double diag = (th1 - th4) * (th1 - th4) + (ph1 - ph4) * (ph1 - ph4);
    double diag2 = (th2 - th3) * (th2 - th3) + (ph2 - ph3) * (ph2 - ph3);
    double mx = (diag < diag2) ? diag2 : diag;   // std::max
    const bool refine = (level < 6 && mx > 1E-10) || (mx > thresh);
    if (!refine) {
        G.level[i * M + j] = (int8_t)level;
        return false;
    }
    R.pts[0] = k * M + j;
    R.pts[1] = k * M + l;
    R.pts[2] = i * M + l;
    R.pts[3] = (i + gap) * M + l;
    R.pts[4] = k * M + (j + gap) % M;
    R.kids[0] = i * M + j;
    R.kids[1] = k * M + j;
    R.kids[2] = k * M + l;
    R.kids[3] = i * M + l;
    return true;
}

}  // namespace bt
