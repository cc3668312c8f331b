// bt_grid.h -- the device-resident adaptive grid (one camera pose).
#pragma once

#include <vector>

#include "bt_internal.h"

namespace bt {

// Dense row-major grid arrays in HBM (cell (i, j) at i*M + j).
struct GridK {
    double* cel;        // 2 per cell: theta', phi'
    int32_t* steps;
    int8_t* level;
    uint32_t* present;  // bit per cell
    int N, M, eq, max_level, start_level;
};

int grid_alloc(Context* c, bt_grid* g);
// Device bytes of an N x M grid with its FP32 render grid.
size_t grid_device_bytes(int N, int M);
int grid_build(Context* c, bt_grid* g, const bt_grid_params* gp);
int grid_fix_tvertices(Context* c, const GridK& G);
// Builds all grids (allocated, same integration parameters) together; the
// rays of each level of every grid share one multi-pose launch.  h_stats
// (nullable) receives the batch's counters.
int grid_build_batch(Context* c, const std::vector<bt_grid*>& gs, const bt_grid_params* gp,
                     StatsK* h_stats);
void grid_release(bt_grid* g);

// One grid of a dataflow build (grid_flow.hip), set up by the batch build:
// its start ray keys [north | south | lattice], the start level's block
// list [pole-row blocks | lattice blocks] and a second list of the same
// size.
struct FlowGridHost {
    GridK G;
    PoseK P;
    uint32_t* keys;
    uint32_t* list0;
    uint32_t* list1;
    unsigned nkeys, nb0, npole;
    int gap0;
};
// Builds the grids' adaptive levels (raytrace's pole rows included) in one
// persistent kernel on the FP64 path; rays (nullable) receives each grid's
// traced-ray count.  The grids' arrays and start lattices are set up.
int grid_flow_run(Context* c, const std::vector<FlowGridHost>& gh, const ParamsK& prm, double thresh,
                  StatsK* d_stats, std::vector<int64_t>* rays);

}  // namespace bt

struct bt_grid {
    bt::Context* ctx = nullptr;
    int N = 0, M = 0, eq = 0, max_level = 0, start_level = 0;
    int64_t rays = 0;
    size_t bytes = 0;
    void* block = nullptr;      // the device block holding the arrays (context pool)
    size_t block_bytes = 0;
    // FP32 dense grid the render pass reads (makeGrid's product), made once
    // on first use; guarded by gf_mu (frames on several streams may ask).
    mutable float* gridf = nullptr;
    mutable int gridf_rows = 0;
    mutable std::mutex gf_mu;
    bt::GridK k{};
    bt::PoseK pose{};
    bt::ParamsK prm{};
    bt_trace_stats stats{};
};
