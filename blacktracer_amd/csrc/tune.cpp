// tune.cpp -- autotuner of the persistent integrators' launch shape.
//
// Three knobs (Context::tune): threads per workgroup, resident waves per
// SIMD (which, for a given batch, sets how many rays each lane traces in
// turn) and the queue chunk a wave reserves per atomic.  bt_autotune times
// a calibration batch of the caller's pose -- the directions of the
// camera's equal-angle sky (the lattice every grid build traces: theta =
// i pi / (N - 1), phi = j 2 pi / M, M = 2 (N - 1)) at about n_rays points,
// traced with the caller's integrator settings -- for candidate shapes by
// coordinate descent (chunk, then block, then waves), keeps the fastest in
// the context and reports it.  The best shape depends on the batch size
// (small batches are bound by their slowest rays, large ones by
// throughput), so callers calibrate at the size they trace.  Results do not
// depend on the launch shape: every ray's arithmetic is the same whichever
// lane traces it.
#include <cmath>
#include <cstdint>
#include <vector>

#include "bt_internal.h"

using namespace bt;

namespace {

int valid(const bt_tune* t) {
    if (t->block != 64 && t->block != 128 && t->block != 256)
        return fail(BT_ERR_INVALID, "tune: block must be 64, 128 or 256");
    if (t->chunk < 64 || t->chunk > (1 << 20))
        return fail(BT_ERR_INVALID, "tune: chunk must be in [64, 2^20]");
    if (t->waves < 0 || t->waves > 8) return fail(BT_ERR_INVALID, "tune: waves must be 0..8");
    return BT_OK;
}

}  // namespace

extern "C" {

int bt_tune_get(bt_context* h, bt_tune* out) {
    if (!h || !out) return fail(BT_ERR_INVALID, "bt_tune_get: null argument");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    out->chunk = (int)h->impl.tune.chunk;
    out->block = (int)h->impl.tune.block;
    out->waves = h->impl.tune.waves;
    out->rays_per_lane = 0;
    return BT_OK;
}

int bt_tune_set(bt_context* h, const bt_tune* t) {
    if (!h || !t) return fail(BT_ERR_INVALID, "bt_tune_set: null argument");
    int rc;
    if ((rc = valid(t))) return rc;
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    h->impl.tune.chunk = (unsigned)t->chunk;
    h->impl.tune.block = (unsigned)t->block;
    h->impl.tune.waves = t->waves;
    return BT_OK;
}

int bt_autotune(bt_context* h, const bt_pose* pose, const bt_integ_params* prm, int64_t n_rays,
                bt_tune* out, double* best_ms) {
    if (!h || !pose) return fail(BT_ERR_INVALID, "bt_autotune: null argument");
    if (n_rays <= 0) n_rays = 1 << 20;
    if (n_rays > (1 << 24)) return fail(BT_ERR_INVALID, "bt_autotune: at most 2^24 calibration rays");
    Context& c = h->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    // the equal-angle sky lattice with N rows and M = 2 (N - 1) columns
    // closest to n_rays points
    int64_t rows = 2;
    while ((rows + 1) * 2 * rows <= n_rays) rows++;
    const int64_t cols = 2 * (rows - 1);
    const size_t n = (size_t)(rows * cols);
    std::vector<double> th(n), ph(n);
    for (int64_t i = 0; i < rows; i++)
        for (int64_t j = 0; j < cols; j++) {
            th[(size_t)(i * cols + j)] = (double)i / (double)(rows - 1) * 3.141592653589793238;
            ph[(size_t)(i * cols + j)] = (double)j / (double)cols * 2.0 * 3.141592653589793238;
        }
    // pristine inputs (the trace works in place) and the working arrays
    void *d_in, *d_th, *d_ph, *d_st;
    int rc;
    if ((rc = c.buffer("tune_in", sizeof(double) * 2 * n, &d_in)) ||
        (rc = c.buffer("tune_th", sizeof(double) * n, &d_th)) ||
        (rc = c.buffer("tune_ph", sizeof(double) * n, &d_ph)) ||
        (rc = c.buffer("tune_st", sizeof(int32_t) * n, &d_st)))
        return rc;
    BT_HIP(hipMemcpyAsync(d_in, th.data(), sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    BT_HIP(hipMemcpyAsync((double*)d_in + n, ph.data(), sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    BT_HIP(hipStreamSynchronize(c.stream));
    hipEvent_t e0, e1;
    BT_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(BT_ERR_HIP, "bt_autotune: event creation failed");
    }
    const Tune saved = c.tune;
    // best of three runs (the first also warms caches and clocks)
    auto time_of = [&](const Tune& t, float* ms_out, int64_t* lanes) -> int {
        c.tune = t;
        c.launch_lanes = 0;
        float best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            BT_HIP(hipMemcpyAsync(d_th, d_in, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
            BT_HIP(hipMemcpyAsync(d_ph, (double*)d_in + n, sizeof(double) * n, hipMemcpyDeviceToDevice, c.stream));
            BT_HIP(hipEventRecord(e0, c.stream));
            int r = bt_trace_rays_device(h, pose, prm, (int64_t)n, (double*)d_th, (double*)d_ph, (int32_t*)d_st,
                                         nullptr);
            if (r) return r;
            BT_HIP(hipEventRecord(e1, c.stream));
            BT_HIP(hipEventSynchronize(e1));
            float ms = 0;
            BT_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0 && ms < best) best = ms;
        }
        *ms_out = best;
        *lanes = c.launch_lanes;
        return BT_OK;
    };
    Tune best = saved;
    best.chunk = 128;
    best.block = 256;
    best.waves = 0;
    float best_t = 0;
    int64_t best_lanes = 0;
    rc = time_of(best, &best_t, &best_lanes);
    const unsigned chunks[] = {64, 128, 256, 512, 1024};
    const unsigned blocks[] = {64, 128, 256};
    const int waves[] = {0, 1, 2, 3, 4};
    for (int pass = 0; pass < 3 && !rc; pass++) {
        const int count = pass == 0 ? 5 : pass == 1 ? 3 : 5;
        for (int k = 0; k < count && !rc; k++) {
            Tune t = best;
            if (pass == 0) t.chunk = chunks[k];
            if (pass == 1) t.block = blocks[k];
            if (pass == 2) t.waves = waves[k];
            if (t.chunk == best.chunk && t.block == best.block && t.waves == best.waves) continue;
            float ms = 0;
            int64_t lanes = 0;
            rc = time_of(t, &ms, &lanes);
            if (!rc && ms < best_t) {
                best = t;
                best_t = ms;
                best_lanes = lanes;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        c.tune = saved;
        return rc;
    }
    c.tune = best;
    if (out) {
        out->chunk = (int)best.chunk;
        out->block = (int)best.block;
        out->waves = best.waves;
        // rays per lane of the widest persistent launch at this shape
        out->rays_per_lane = best_lanes > 0 ? (double)n / (double)best_lanes : 0.0;
    }
    if (best_ms) *best_ms = best_t;
    return BT_OK;
}

}  // extern "C"
