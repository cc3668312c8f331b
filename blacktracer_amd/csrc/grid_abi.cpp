// grid_abi.cpp -- C ABI of the adaptive grid (Grid.h:605-654).
#include <cmath>
#include <cstring>
#include <vector>

#include "bt_grid.h"
#include "psh.h"

using namespace bt;

extern "C" {

void bt_grid_params_default(bt_grid_params* p) {
    if (!p) return;
    p->refine_thresh = 0.015;   // PRECCELEST, Grid.h:24
    p->fix_tvertices = 1;
}

}  // extern "C"

// A grid object with its pose, parameters and dense arrays allocated.
static int grid_create(Context& c, const bt_pose* pose, const bt_integ_params* prm, int max_level,
                       int start_level, int angle, bt_grid** out) {
    *out = nullptr;
    if (max_level < 1 || max_level > 12 || start_level < 1 || start_level > max_level)
        return fail(BT_ERR_INVALID, "grid build: need 1 <= start_level <= max_level <= 12");
    bt_grid* g = new bt_grid();
    g->ctx = &c;
    int rc;
    if ((rc = make_pose(pose, &g->pose)) || (rc = make_params(prm, &g->prm))) {
        delete g;
        return rc;
    }
    if (prm && prm->r_capture_factor > 0)
        g->pose.r_cap = prm->r_capture_factor * (1. + std::sqrt(1. - g->pose.asq));
    g->eq = angle ? 1 : 0;
    g->max_level = max_level;
    g->start_level = start_level;
    // Grid.h:612-615
    g->N = (int)(uint32_t)std::round(std::pow(2, max_level) / (2 - g->eq) + 1);
    g->M = (2 - g->eq) * 2 * (g->N - 1);
    if ((rc = grid_alloc(&c, g))) {
        grid_release(g);
        delete g;
        return rc;
    }
    *out = g;
    return BT_OK;
}

static void grid_destroy(bt_grid* g) {
    if (!g) return;
    grid_release(g);
    delete g;
}

extern "C" {

int bt_grid_build(bt_context* h, const bt_pose* pose, const bt_integ_params* prm, int max_level,
                  int start_level, int angle, const bt_grid_params* gp, bt_grid** out) {
    if (!h || !out) return fail(BT_ERR_INVALID, "bt_grid_build: null argument");
    *out = nullptr;
    Context& c = h->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    bt_grid* g;
    int rc;
    if ((rc = grid_create(c, pose, prm, max_level, start_level, angle, &g))) return rc;
    if ((rc = grid_build(&c, g, gp))) {
        (void)hipStreamSynchronize(c.stream);
        grid_destroy(g);
        return rc;
    }
    *out = g;
    return BT_OK;
}

int bt_grid_build_batch(bt_context* h, int n, const bt_pose* poses, const bt_integ_params* prm,
                        const int* max_levels, const int* start_levels, int angle,
                        const bt_grid_params* gp, bt_grid** out, bt_trace_stats* stats) {
    if (!h || !poses || !max_levels || !start_levels || !out)
        return fail(BT_ERR_INVALID, "bt_grid_build_batch: null argument");
    if (n < 1 || n > 128) return fail(BT_ERR_INVALID, "bt_grid_build_batch: need 1 <= n <= 128");
    for (int k = 0; k < n; k++) out[k] = nullptr;
    Context& c = h->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    std::vector<bt_grid*> gs(n, nullptr);
    int rc = BT_OK;
    for (int k = 0; k < n && !rc; k++)
        rc = grid_create(c, &poses[k], prm, max_levels[k], start_levels[k], angle, &gs[k]);
    StatsK s{};
    if (!rc) rc = grid_build_batch(&c, gs, gp, &s);
    if (rc) {
        (void)hipStreamSynchronize(c.stream);
        for (bt_grid* g : gs) grid_destroy(g);
        return rc;
    }
    if (stats) stats_to_abi(s, stats);
    for (int k = 0; k < n; k++) {
        gs[k]->stats.rays = gs[k]->rays;
        out[k] = gs[k];
    }
    return BT_OK;
}

int bt_grid_info(const bt_grid* g, int* N, int* M, int* max_level, int* start_level,
                 int* equafactor, int64_t* rays_traced) {
    if (!g) return fail(BT_ERR_INVALID, "grid is null");
    if (N) *N = g->N;
    if (M) *M = g->M;
    if (max_level) *max_level = g->max_level;
    if (start_level) *start_level = g->start_level;
    if (equafactor) *equafactor = g->eq;
    if (rays_traced) *rays_traced = g->rays;
    return BT_OK;
}

int bt_grid_stats(const bt_grid* g, bt_trace_stats* stats) {
    if (!g || !stats) return fail(BT_ERR_INVALID, "null argument");
    *stats = g->stats;
    return BT_OK;
}

int bt_grid_export(const bt_grid* g, double* cel, int32_t* steps, int8_t* levels) {
    if (!g) return fail(BT_ERR_INVALID, "grid is null");
    Context& c = *g->ctx;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    const int64_t NM = (int64_t)g->N * g->M;
    if (cel)
        BT_HIP(hipMemcpyAsync(cel, g->k.cel, sizeof(double) * 2 * NM, hipMemcpyDeviceToHost, c.stream));
    if (steps)
        BT_HIP(hipMemcpyAsync(steps, g->k.steps, sizeof(int32_t) * NM, hipMemcpyDeviceToHost, c.stream));
    if (levels)
        BT_HIP(hipMemcpyAsync(levels, g->k.level, NM, hipMemcpyDeviceToHost, c.stream));
    BT_HIP(hipStreamSynchronize(c.stream));
    return BT_OK;
}

int bt_grid_export_psh(const bt_grid* g, int* hash_width, int* offset_width, float* hash_table,
                       int32_t* offset_table, int32_t* hash_pos_tag) {
    if (!g || !hash_width || !offset_width) return fail(BT_ERR_INVALID, "null argument");
    Context& c = *g->ctx;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    const int64_t NM = (int64_t)g->N * g->M;
    const int64_t nwords = (NM + 31) / 32;
    std::vector<double> cel(2 * NM);
    std::vector<uint32_t> present(nwords);
    BT_HIP(hipMemcpyAsync(cel.data(), g->k.cel, sizeof(double) * 2 * NM, hipMemcpyDeviceToHost, c.stream));
    BT_HIP(hipMemcpyAsync(present.data(), g->k.present, sizeof(uint32_t) * nwords,
                          hipMemcpyDeviceToHost, c.stream));
    BT_HIP(hipStreamSynchronize(c.stream));
    // Grid.h:645-651: every CamToCel entry, as float2
    std::vector<int32_t> keys;
    std::vector<float> vals;
    for (int64_t idx = 0; idx < NM; idx++) {
        if (!((present[idx >> 5] >> (idx & 31)) & 1u)) continue;
        keys.push_back((int32_t)(idx / g->M));
        keys.push_back((int32_t)(idx % g->M));
        vals.push_back((float)cel[2 * idx]);
        vals.push_back((float)cel[2 * idx + 1]);
    }
    PshTable t;
    int rc = psh_build(keys, vals, &t);
    if (rc) return rc;
    *hash_width = t.hash_width;
    *offset_width = t.offset_width;
    if (hash_table) std::memcpy(hash_table, t.hash_table.data(), sizeof(float) * t.hash_table.size());
    if (offset_table)
        std::memcpy(offset_table, t.offset_table.data(), sizeof(int32_t) * t.offset_table.size());
    if (hash_pos_tag)
        std::memcpy(hash_pos_tag, t.hash_pos_tag.data(), sizeof(int32_t) * t.hash_pos_tag.size());
    return BT_OK;
}

void bt_grid_free(bt_grid* g) {
    if (!g) return;
    if (g->ctx) {
        std::lock_guard<std::recursive_mutex> lk(g->ctx->mu);
        (void)hipSetDevice(g->ctx->device);
        (void)hipStreamSynchronize(g->ctx->stream);
        grid_release(g);
    }
    delete g;
}

}  // extern "C"
