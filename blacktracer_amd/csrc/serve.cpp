// serve.cpp -- continuous-batching render server over device-resident grids.
//
// No counterpart in the reference, whose frame loop (MakeImageKernel.cu:
// 975-1204) renders one frame at a time and re-uploads every input per
// makeImage call.  Here:
//   * a client thread submits render requests (pose, orbit offset,
//     resolution) and polls completions;
//   * a builder thread builds the grid of a pose not yet cached (the
//     adaptive Grid of that pose, kept in HBM with its FP32 render grid)
//     and the view matrix of a resolution not yet seen;
//   * a dispatcher thread admits pending requests in arrival order whenever
//     a slot is free and the request's frame working set fits the HBM
//     budget next to the cached grids (least recently used idle grids are
//     evicted to make room); each slot is a context of its own (stream and
//     work buffers), so admitted frames overlap on the device while others
//     are still running, and a request that arrives meanwhile is admitted as
//     soon as a slot and memory free up;
//   * completion is an event on the slot's stream after the frame's copy to
//     pinned staging memory; the dispatcher then copies it to the client's
//     buffer and records the latency (arrival -> frame in host memory).
// The HBM budget counts every device byte the server holds: cached grids
// (with their FP32 render grids), views, each slot's work buffers at their
// high-water mark (a slot context keeps its buffers between frames, so a
// slot is charged for the largest frame it rendered until they are
// trimmed), and a grid build's arrays and work buffers while it runs (the
// builder context gives its work buffers back after every build, and freed
// grids go back to the device, not to a pool).  Grids and slot buffers are
// freed outside the server lock, after their bytes have been chosen under
// it; their bytes stay counted until they are gone.
// Every frame is rendered by the same kernels on the same inputs whatever
// else is in flight, so a request's result is identical alone and in a mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "bt_grid.h"
#include "bt_render.h"

using namespace bt;
using Clock = std::chrono::steady_clock;

namespace {

struct GridKey {
    double a, r, theta;
    bool operator<(const GridKey& o) const {
        if (a != o.a) return a < o.a;
        if (r != o.r) return r < o.r;
        return theta < o.theta;
    }
};

struct CachedGrid {
    bt_grid* g = nullptr;
    size_t bytes = 0;
    int refs = 0;          // frames in flight reading it
    bool ready = false;
    bool queued = false;   // waiting for the builder
    int status = BT_OK;
    uint64_t last_use = 0;
};

struct CachedView {
    bt_view* v = nullptr;
    bool ready = false;
    bool queued = false;
    int status = BT_OK;
};

struct Pending {
    bt_request r;
    Clock::time_point arrive;
    size_t bytes;
    bool hit_checked = false;
    bool hit = false;
};

struct Slot {
    bt_context* ctx = nullptr;
    bool busy = false;
    Pending req;
    CachedGrid* grid = nullptr;
    hipEvent_t done = nullptr;
    uint8_t* staging = nullptr;   // pinned host memory
    size_t staging_bytes = 0;
    size_t charged = 0;           // bytes of the slot context's work buffers counted in `used`
    bool trimming = false;        // its buffers are being freed: not admitted to meanwhile
    Clock::time_point admit;
};

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct bt_server {
    bt_context* owner = nullptr;        // the caller's context (device, views)
    bt_context* builder = nullptr;      // grid builds run here
    const bt_scene* scene = nullptr;
    bt_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv;         // dispatcher and drain wait on it
    std::condition_variable cv_build;   // builder waits on it
    std::deque<Pending> pending;
    std::vector<Slot> slots;
    std::map<GridKey, CachedGrid> grids;
    std::map<std::pair<int, int>, CachedView> views;
    std::deque<GridKey> build_grids;
    std::deque<std::pair<int, int>> build_views;
    std::vector<bt_completion> done;
    std::vector<double> lat;
    size_t used = 0, peak = 0;
    int64_t submitted = 0, completed = 0, failed = 0, builds = 0, hits = 0, evictions = 0;
    int inflight = 0, peak_inflight = 0;
    size_t building = 0;                // bytes the builder has reserved for a grid in progress
    size_t releasing = 0;               // bytes chosen for freeing, still counted in `used`
    uint64_t tick = 0;
    bool stop = false;
    bool have_first = false;
    Clock::time_point first_arrive, last_done;
    std::thread dispatcher, builder_thread;
};

namespace {

void complete(bt_server* s, const Pending& p, int status, Clock::time_point admit, Clock::time_point now,
              bool hit) {
    bt_completion c;
    c.tag = p.r.tag;
    c.status = status;
    c.latency_ms = ms_between(p.arrive, now);
    c.queue_ms = ms_between(p.arrive, admit);
    c.frame_ms = ms_between(admit, now);
    c.grid_hit = hit ? 1 : 0;
    s->done.push_back(c);
    if (status == BT_OK) {
        s->completed++;
        s->lat.push_back(c.latency_ms);
    } else {
        s->failed++;
    }
    s->last_done = now;
}

// What make_room chose to free: idle grids (taken out of the cache) and
// idle slots whose work buffers go.  Their bytes stay in `used` (and in
// `releasing`) until release() has freed them.
struct Victims {
    std::vector<std::pair<bt_grid*, size_t>> grids;
    std::vector<Slot*> slots;
    size_t bytes = 0;
    bool empty() const { return grids.empty() && slots.empty(); }
};

// Device bytes a slot context holds after rendering a w x h frame: the
// frame's work buffers and the server's output buffer, each named buffer
// allocated with an eighth of slack and 256 bytes (Context::buffer).
size_t slot_bytes(int step, int w, int h) {
    const size_t f = frame_bytes(w, h, step) + (size_t)4 * w * h;
    return f + f / 8 + (size_t)64 * 256;
}

// A grid build's device bytes while it runs: the grid's arrays with its
// FP32 render grid, and the work buffers of the build (queues, continuation
// state, block lists: at most 160 B per grid point, allocated with an
// eighth of slack).
size_t build_bytes(int N, int M) {
    const size_t NM = (size_t)N * M;
    return grid_device_bytes(N, M) + 180 * NM + (size_t)64 * 256;
}

// Chooses idle cached grids (least recently used first), then the work
// buffers of idle slots other than `keep`, until `need` more bytes fit the
// budget once they are freed.  Caller holds s->mu; the caller frees the
// victims with release() before it allocates.
bool make_room(bt_server* s, size_t need, Victims& v, const Slot* keep = nullptr) {
    while (s->used - s->releasing + need > s->cfg.hbm_budget) {
        CachedGrid* victim = nullptr;
        GridKey vk{};
        for (auto& kv : s->grids) {
            CachedGrid& cg = kv.second;
            if (!cg.ready || cg.refs > 0 || !cg.g) continue;
            if (!victim || cg.last_use < victim->last_use) {
                victim = &cg;
                vk = kv.first;
            }
        }
        if (victim) {
            v.grids.emplace_back(victim->g, victim->bytes);
            v.bytes += victim->bytes;
            s->releasing += victim->bytes;
            s->grids.erase(vk);
            s->evictions++;
            continue;
        }
        Slot* idle = nullptr;
        for (Slot& sl : s->slots)
            if (&sl != keep && !sl.busy && !sl.trimming && sl.charged > 0) {
                idle = &sl;
                break;
            }
        if (!idle) return false;
        idle->trimming = true;
        v.slots.push_back(idle);
        v.bytes += idle->charged;
        s->releasing += idle->charged;
    }
    return true;
}

// Frees an idle cached grid: no frame holds it (refs == 0 since its last
// frame retired) and its build has completed, so no stream reads it.  Its
// blocks go back to the device, not to the builder context's pool, and the
// builder context's lock (held for the whole of a build) is not taken.
void free_idle_grid(bt_grid* g) {
    g->ctx = nullptr;
    grid_release(g);
    delete g;
}

// Frees the victims with s->mu released (lk is its lock), then takes their
// bytes out of `used`.  A slot being trimmed is not admitted to meanwhile.
void release(bt_server* s, Victims& v, std::unique_lock<std::mutex>& lk) {
    if (v.empty()) return;
    lk.unlock();
    for (auto& gb : v.grids) free_idle_grid(gb.first);
    for (Slot* sl : v.slots) {
        Context& c = sl->ctx->impl;
        std::lock_guard<std::recursive_mutex> cl(c.mu);
        c.release_buffers();
    }
    lk.lock();
    for (Slot* sl : v.slots) {
        sl->charged = 0;
        sl->trimming = false;
    }
    s->used -= v.bytes;
    s->releasing -= v.bytes;
    v = Victims();
    s->cv.notify_all();
}

void builder_main(bt_server* s) {
    (void)hipSetDevice(s->owner->impl.device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv_build.wait(lk, [&] { return s->stop || !s->build_grids.empty() || !s->build_views.empty(); });
        if (s->stop) return;
        if (!s->build_views.empty()) {
            auto wh = s->build_views.front();
            s->build_views.pop_front();
            lk.unlock();
            bt_view* v = nullptr;
            int rc = bt_view_create(s->owner, s->cfg.view_angle, s->cfg.off_up, s->cfg.off_right, wh.first, wh.second, 0,
                                    &v);
            lk.lock();
            CachedView& cv = s->views[wh];
            cv.queued = false;
            cv.status = rc;
            cv.v = v;
            cv.ready = rc == BT_OK;
            if (rc == BT_OK) s->used += v->bytes;
            s->cv.notify_all();
            continue;
        }
        GridKey key = s->build_grids.front();
        s->build_grids.pop_front();
        // reserve the grid's bytes first (Grid.h:612-615 sizes): wait while
        // frames in flight hold the memory, fail if nothing can free it
        const int N = (int)(uint32_t)std::round(std::pow(2, s->cfg.max_level) / (2 - 1) + 1);   // angle = 1
        const size_t est = build_bytes(N, 2 * (N - 1));
        bool fits;
        while (true) {
            Victims v;
            fits = make_room(s, est, v);
            const bool freed = !v.empty();
            release(s, v, lk);
            if (s->stop) return;
            if (s->used + est <= s->cfg.hbm_budget) {
                fits = true;
                break;
            }
            if (freed) continue;   // the victims are gone: look again at once
            // nothing in flight and nothing being freed: it can never fit
            if (!fits && s->inflight == 0 && s->releasing == 0) break;
            s->cv.wait(lk);        // frames retire, or another thread finishes freeing
        }
        if (!fits) {
            CachedGrid& cg = s->grids[key];
            cg.queued = false;
            cg.status = fail(BT_ERR_BUDGET, "server: a grid does not fit the HBM budget");
            s->cv.notify_all();
            continue;
        }
        s->used += est;
        s->building = est;
        s->peak = std::max(s->peak, s->used);
        lk.unlock();
        bt_pose pose{key.a, key.theta, 0.0, key.r, 0.0, 0.0, 1.0, 0.0, 0};
        bt_grid_params gp;
        bt_grid_params_default(&gp);
        bt_grid* g = nullptr;
        int rc = bt_grid_build(s->builder, &pose, &s->cfg.integ, s->cfg.max_level, s->cfg.start_level, 1, &gp, &g);
        GridF gf;
        {
            Context& bc = s->builder->impl;
            std::lock_guard<std::recursive_mutex> blk(bc.mu);
            if (rc == BT_OK) rc = grid_f32_of(bc, g, &gf);
            // the build's work buffers go back to the device (the budget
            // counted them only while the build ran)
            bc.release_buffers();
        }
        size_t bytes = 0;
        if (rc == BT_OK) bytes = g->bytes + sizeof(float) * 2 * (size_t)g->gridf_rows * g->M;
        lk.lock();
        s->used -= est;
        s->building = 0;
        CachedGrid& cg = s->grids[key];
        cg.queued = false;
        cg.status = rc;
        if (rc == BT_OK) {
            cg.g = g;
            cg.bytes = bytes;
            cg.ready = true;
            cg.last_use = ++s->tick;
            s->used += bytes;
            s->peak = std::max(s->peak, s->used);
            s->builds++;
        } else if (g) {
            lk.unlock();
            free_idle_grid(g);
            lk.lock();
        }
        s->cv.notify_all();
    }
}

// Launches request p on slot sl.  Caller holds s->mu.
int launch(bt_server* s, Slot& sl, CachedGrid& cg, const bt_view* v) {
    Context& c = sl.ctx->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    const bt_request& r = sl.req.r;
    const size_t nbytes = (size_t)4 * r.width * r.height;
    if (sl.staging_bytes < nbytes) {
        if (sl.staging) (void)hipHostFree(sl.staging);
        sl.staging = nullptr;
        sl.staging_bytes = 0;
        if (hipHostMalloc((void**)&sl.staging, nbytes, 0) != hipSuccess)
            return fail(BT_ERR_NOMEM, "server: pinned staging allocation failed");
        sl.staging_bytes = nbytes;
    }
    GridF gf;
    int rc;
    if ((rc = grid_f32_of(c, cg.g, &gf))) return rc;
    void* dout;
    if ((rc = c.buffer("srv_out", nbytes, &dout))) return rc;
    bt_render_flags f;
    bt_render_flags_default(&f);
    f.star = r.star;
    f.map = r.map;
    f.lr = r.lr;
    f.offset = r.offset;
    float cam[7];
    cam7_of(cg.g->pose, cam);
    if ((rc = render_frame(&c, gf, s->scene->k(), ViewK{v->view, v->W, v->H, v->pix_area}, cam, f, s->cfg.step, (uint8_t*)dout,
                           nullptr)))
        return rc;
    BT_HIP(hipMemcpyAsync(sl.staging, dout, nbytes, hipMemcpyDeviceToHost, c.stream));
    BT_HIP(hipEventRecord(sl.done, c.stream));
    return BT_OK;
}

void dispatcher_main(bt_server* s) {
    (void)hipSetDevice(s->owner->impl.device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool progress = false;
        // 1. retire finished frames
        for (Slot& sl : s->slots) {
            if (!sl.busy || hipEventQuery(sl.done) != hipSuccess) continue;
            // the client's buffer is written before the completion is posted
            std::memcpy(sl.req.r.out_bgra, sl.staging, (size_t)4 * sl.req.r.width * sl.req.r.height);
            Clock::time_point now = Clock::now();
            sl.grid->refs--;
            sl.grid->last_use = ++s->tick;
            s->inflight--;   // the slot keeps its buffers, and their charge
            sl.busy = false;
            complete(s, sl.req, BT_OK, sl.admit, now, sl.req.hit);
            progress = true;
        }
        // 2. admit pending requests, oldest first, while slots and memory allow
        bool rescan = false;
        for (auto it = s->pending.begin(); it != s->pending.end();) {
            // a free slot whose buffers already hold this frame (the smallest
            // such), else the free slot holding the most
            const size_t need_slot = slot_bytes(s->cfg.step, it->r.width, it->r.height);
            Slot* free_slot = nullptr;
            for (Slot& sl : s->slots) {
                if (sl.busy || sl.trimming) continue;
                if (!free_slot) {
                    free_slot = &sl;
                    continue;
                }
                const bool a_fits = sl.charged >= need_slot, b_fits = free_slot->charged >= need_slot;
                if (a_fits != b_fits ? a_fits : (a_fits ? sl.charged < free_slot->charged
                                                        : sl.charged > free_slot->charged))
                    free_slot = &sl;
            }
            if (!free_slot) break;
            GridKey key{it->r.a, it->r.r, it->r.theta};
            auto git = s->grids.find(key);
            if (git == s->grids.end()) {
                CachedGrid& cg = s->grids[key];
                cg.queued = true;
                s->build_grids.push_back(key);
                s->cv_build.notify_one();
                if (!it->hit_checked) it->hit_checked = true;
                ++it;
                continue;
            }
            CachedGrid& cg = git->second;
            if (!it->hit_checked) {
                it->hit_checked = true;
                it->hit = cg.ready;
                if (it->hit) s->hits++;
            }
            std::pair<int, int> wh{it->r.width, it->r.height};
            auto vit = s->views.find(wh);
            if (vit == s->views.end()) {
                CachedView& cv = s->views[wh];
                cv.queued = true;
                s->build_views.push_back(wh);
                s->cv_build.notify_one();
                ++it;
                continue;
            }
            if (cg.status != BT_OK || vit->second.status != BT_OK) {
                int st = cg.status != BT_OK ? cg.status : vit->second.status;
                complete(s, *it, st, Clock::now(), Clock::now(), false);
                if (cg.status != BT_OK) s->grids.erase(git);   // let a later request retry
                if (vit->second.status != BT_OK) s->views.erase(vit);
                it = s->pending.erase(it);
                progress = true;
                continue;
            }
            if (!cg.ready || !vit->second.ready) {
                ++it;
                continue;
            }
            // HBM admission: cached grids and views, every slot's buffers
            // and a build in progress within the budget
            const size_t extra = need_slot > free_slot->charged ? need_slot - free_slot->charged : 0;
            Victims vic;
            cg.refs++;   // pin it while choosing victims
            bool fits = make_room(s, extra, vic, free_slot);
            cg.refs--;
            if (!vic.empty()) {
                // free them with the lock released, then scan again
                release(s, vic, lk);
                rescan = true;
                progress = true;
                break;
            }
            if (fits && s->used + extra > s->cfg.hbm_budget) break;   // another thread is still freeing
            if (!fits) {
                if (s->inflight == 0 && s->building == 0 && s->releasing == 0) {
                    // nothing in flight, no grid being built and every idle
                    // grid gone: this request (frame + its grid + the views)
                    // can never fit
                    complete(s, *it, BT_ERR_BUDGET, Clock::now(), Clock::now(), it->hit);
                    it = s->pending.erase(it);
                    progress = true;
                    continue;
                }
                break;   // wait for frames to drain (keeps arrival order)
            }
            Slot& sl = *free_slot;
            sl.req = *it;
            sl.grid = &cg;
            sl.admit = Clock::now();
            cg.refs++;
            cg.last_use = ++s->tick;
            s->used += extra;
            sl.charged += extra;
            int rc = launch(s, sl, cg, vit->second.v);
            {
                // charge what the slot's buffers hold now (at most the estimate)
                const size_t held = sl.ctx->impl.buffer_bytes();
                s->used = s->used - sl.charged + std::max(sl.charged, held);
                sl.charged = std::max(sl.charged, held);
            }
            s->peak = std::max(s->peak, s->used);
            if (rc != BT_OK) {
                cg.refs--;
                complete(s, *it, rc, sl.admit, Clock::now(), it->hit);
            } else {
                sl.busy = true;
                s->inflight++;
                s->peak_inflight = std::max(s->peak_inflight, s->inflight);
            }
            it = s->pending.erase(it);
            progress = true;
        }
        if (progress) s->cv.notify_all();
        if (s->stop && s->inflight == 0) return;
        if (rescan) continue;
        if (!progress) s->cv.wait_for(lk, std::chrono::microseconds(s->inflight ? 50 : 2000));
    }
}

}  // namespace

extern "C" {

void bt_server_config_default(bt_server_config* c) {
    if (!c) return;
    c->hbm_budget = size_t(16) << 30;
    c->slots = 4;
    c->max_level = 10;
    c->start_level = 1;
    bt_integ_params_default(&c->integ);
    c->step = 1;
    c->view_angle = 3.141592653589793238 / 2;   // BlackTracer.cpp:119-120
    c->off_up = 0.0;
    c->off_right = 0.25 * 3.141592653589793238 / 4;
}

int bt_server_footprint(const bt_server_config* cfg, int width, int height, bt_server_bytes* out) {
    if (!cfg || !out || width <= 0 || height <= 0 || cfg->max_level < 1 || cfg->max_level > 12 || cfg->step < 1 ||
        cfg->step > 3)
        return fail(BT_ERR_INVALID, "bt_server_footprint: bad arguments");
    const int N = (int)(uint32_t)std::round(std::pow(2, cfg->max_level) + 1);   // full sky (angle = 1)
    const int M = 2 * (N - 1);
    out->slot = slot_bytes(cfg->step, width, height);
    out->view = sizeof(float) * (2 * (size_t)(width + 1) * (height + 1) + (size_t)width * height);
    out->grid = grid_device_bytes(N, M);
    out->build = build_bytes(N, M);
    return BT_OK;
}

int bt_server_create(bt_context* h, const bt_scene* scene, const bt_server_config* cfg, bt_server** out) {
    if (!h || !scene || !out) return fail(BT_ERR_INVALID, "bt_server_create: null argument");
    *out = nullptr;
    bt_server_config c;
    bt_server_config_default(&c);
    if (cfg) c = *cfg;
    if (c.slots < 1 || c.slots > 16) return fail(BT_ERR_INVALID, "bt_server_create: slots must be 1..16");
    if (c.step < 1 || c.step > 3) return fail(BT_ERR_INVALID, "bt_server_create: step must be 1..3");
    if (c.max_level < 1 || c.max_level > 12 || c.start_level < 1 || c.start_level > c.max_level)
        return fail(BT_ERR_INVALID, "bt_server_create: bad grid levels");
    ParamsK pk;
    int rc;
    if ((rc = make_params(&c.integ, &pk))) return rc;
    bt_server* s = new bt_server();
    s->owner = h;
    s->scene = scene;
    s->cfg = c;
    if ((rc = bt_context_create(h->impl.device, &s->builder))) {
        delete s;
        return rc;
    }
    s->builder->impl.tune = h->impl.tune;
    s->builder->impl.pool_cap = 0;   // freed grids go back to the device: the budget sees them go
    s->slots.resize(c.slots);
    for (Slot& sl : s->slots) {
        if ((rc = bt_context_create(h->impl.device, &sl.ctx)) ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            if (!rc) rc = fail(BT_ERR_HIP, "bt_server_create: event creation failed");
            for (Slot& t : s->slots) {
                if (t.done) (void)hipEventDestroy(t.done);
                if (t.ctx) bt_context_destroy(t.ctx);
            }
            bt_context_destroy(s->builder);
            delete s;
            return rc;
        }
    }
    s->builder_thread = std::thread(builder_main, s);
    s->dispatcher = std::thread(dispatcher_main, s);
    *out = s;
    return BT_OK;
}

int bt_server_submit(bt_server* s, const bt_request* r) {
    if (!s || !r || !r->out_bgra) return fail(BT_ERR_INVALID, "bt_server_submit: null argument");
    if (r->width <= 0 || r->height <= 0 || r->width > 16384 || r->height > 16384)
        return fail(BT_ERR_INVALID, "bt_server_submit: size out of range");
    if (!(std::fabs(r->a) < 1) || !(r->r > 0) || !(r->offset >= 0.f) || !std::isfinite(r->offset))
        return fail(BT_ERR_INVALID, "bt_server_submit: bad pose or offset");
    Pending p;
    p.r = *r;
    p.arrive = Clock::now();
    p.bytes = frame_bytes(r->width, r->height, s->cfg.step);
    if (p.bytes > s->cfg.hbm_budget)
        return fail(BT_ERR_BUDGET, "bt_server_submit: the frame alone exceeds the HBM budget");
    std::lock_guard<std::mutex> lk(s->mu);
    if (!s->have_first) {
        s->have_first = true;
        s->first_arrive = p.arrive;
    }
    s->pending.push_back(p);
    s->submitted++;
    s->cv.notify_all();
    return BT_OK;
}

int bt_server_poll(bt_server* s, bt_completion* out, int max, int* n) {
    if (!s || !n || (max > 0 && !out)) return fail(BT_ERR_INVALID, "bt_server_poll: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int k = std::min<int>(max, (int)s->done.size());
    for (int i = 0; i < k; i++) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return BT_OK;
}

int bt_server_drain(bt_server* s, double timeout_s) {
    if (!s) return fail(BT_ERR_INVALID, "bt_server_drain: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto pred = [&] { return s->completed + s->failed == s->submitted; };
    if (timeout_s <= 0) {
        s->cv.wait(lk, pred);
        return BT_OK;
    }
    if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
        return fail(BT_ERR_STATE, "bt_server_drain: timed out");
    return BT_OK;
}

int bt_server_stats_get(bt_server* s, bt_server_stats* st) {
    if (!s || !st) return fail(BT_ERR_INVALID, "bt_server_stats_get: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->grid_builds = s->builds;
    st->grid_hits = s->hits;
    st->evictions = s->evictions;
    st->peak_inflight = s->peak_inflight;
    st->peak_bytes = s->peak;
    if (!s->lat.empty()) {
        std::vector<double> l = s->lat;
        std::sort(l.begin(), l.end());
        auto q = [&](double p) {
            size_t i = (size_t)std::ceil(p * l.size());
            return l[std::min(l.size() - 1, i > 0 ? i - 1 : 0)];
        };
        st->p50_ms = q(0.5);
        st->p99_ms = q(0.99);
        double sum = 0;
        for (double x : l) sum += x;
        st->mean_ms = sum / l.size();
        st->wall_s = std::chrono::duration<double>(s->last_done - s->first_arrive).count();
        st->frames_per_s = st->wall_s > 0 ? s->completed / st->wall_s : 0;
    }
    return BT_OK;
}

void bt_server_destroy(bt_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->pending.clear();
        s->cv.notify_all();
        s->cv_build.notify_all();
    }
    if (s->dispatcher.joinable()) s->dispatcher.join();
    if (s->builder_thread.joinable()) s->builder_thread.join();
    for (Slot& sl : s->slots) {
        if (sl.ctx) (void)hipStreamSynchronize(sl.ctx->impl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.staging) (void)hipHostFree(sl.staging);
        if (sl.ctx) bt_context_destroy(sl.ctx);
    }
    for (auto& kv : s->grids)
        if (kv.second.g) free_idle_grid(kv.second.g);
    for (auto& kv : s->views)
        if (kv.second.v) bt_view_free(kv.second.v);
    if (s->builder) bt_context_destroy(s->builder);
    delete s;
}

}  // extern "C"
