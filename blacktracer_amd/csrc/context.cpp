// context.cpp -- bt_context (device, stream, work buffers, HIP-event
// timers), error reporting, pose set-up and the ray-batch entry points of
// the C ABI.
#include <cmath>
#include <cstring>
#include <string>

#include "bt_internal.h"

namespace bt {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

// ------------------------------------------------------------ Context
int Context::buffer(const std::string& name, size_t bytes, void** out) {
    Buffer& b = scratch[name];
    if (b.bytes < bytes) {
        if (b.ptr) {
            // the stream may still read the old buffer
            (void)hipStreamSynchronize(stream);
            (void)hipFree(b.ptr);
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes + bytes / 8 + 256;
        if (hipMalloc(&b.ptr, want) != hipSuccess) {
            b.ptr = nullptr;
            return fail(BT_ERR_NOMEM, "hipMalloc failed for work buffer " + name);
        }
        b.bytes = want;
    }
    *out = b.ptr;
    return BT_OK;
}

size_t Context::buffer_bytes() const {
    size_t t = 0;
    for (const auto& kv : scratch) t += kv.second.bytes;
    return t;
}

void Context::release_buffers() {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamSynchronize(aux);
    for (auto& kv : scratch)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    scratch.clear();
    flow_rq_seen = flow_tq_seen = nullptr;
}

int Context::stage(size_t bytes, void** out) {
    if (ev_stage && hipEventSynchronize(ev_stage) != hipSuccess)
        return fail(BT_ERR_HIP, "staging: waiting for the previous copies failed");
    if (host_stage_bytes < bytes) {
        if (host_stage) (void)hipHostFree(host_stage);
        host_stage = nullptr;
        host_stage_bytes = 0;
        const size_t want = bytes + bytes / 2 + 4096;
        if (hipHostMalloc(&host_stage, want, 0) != hipSuccess) {
            host_stage = nullptr;
            return fail(BT_ERR_NOMEM, "hipHostMalloc failed for the staging area");
        }
        host_stage_bytes = want;
    }
    *out = host_stage;
    return BT_OK;
}

int Context::stage_done() {
    if (hipEventRecord(ev_stage, stream) != hipSuccess) return fail(BT_ERR_HIP, "staging: event record failed");
    return BT_OK;
}

int Context::pool_alloc(size_t bytes, void** out, size_t* got) {
    auto it = pool.lower_bound(bytes);
    if (it != pool.end() && it->first <= 2 * bytes) {
        *out = it->second;
        *got = it->first;
        pool_bytes -= it->first;
        pool.erase(it);
        return BT_OK;
    }
    if (hipMalloc(out, bytes) != hipSuccess) {
        // give the pooled blocks back and try once more
        for (auto& kv : pool) (void)hipFree(kv.second);
        pool.clear();
        pool_bytes = 0;
        if (hipMalloc(out, bytes) != hipSuccess) {
            *out = nullptr;
            return fail(BT_ERR_NOMEM, "hipMalloc failed for a grid of " + std::to_string(bytes) + " bytes");
        }
    }
    *got = bytes;
    return BT_OK;
}

void Context::pool_free(void* p, size_t bytes) {
    if (!p) return;
    if (pool_bytes + bytes > pool_cap) {
        (void)hipFree(p);
        return;
    }
    pool.emplace(bytes, p);
    pool_bytes += bytes;
}

int Context::timer_begin(const char* name, int* slot, hipStream_t st) {
    PendingTimer t;
    t.name = name;
    t.stream = st ? st : stream;
    for (hipEvent_t* e : {&t.start, &t.stop}) {
        if (!free_events.empty()) {
            *e = free_events.back();
            free_events.pop_back();
        } else {
            BT_HIP(hipEventCreate(e));
        }
    }
    BT_HIP(hipEventRecord(t.start, t.stream));
    pending.push_back(t);
    *slot = (int)pending.size() - 1;
    return BT_OK;
}

int Context::timer_end(int slot) {
    BT_HIP(hipEventRecord(pending[slot].stop, pending[slot].stream));
    return BT_OK;
}

int Context::flush_timers() {
    if (pending.empty()) return BT_OK;
    BT_HIP(hipStreamSynchronize(stream));
    BT_HIP(hipStreamSynchronize(aux));
    for (auto& t : pending) {
        float ms = 0.f;
        BT_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        auto& acc = timers[t.name];
        acc.first += 1;
        acc.second += ms;
        free_events.push_back(t.start);
        free_events.push_back(t.stop);
    }
    pending.clear();
    return BT_OK;
}

// ------------------------------------------------------------ pose
// Metric.h:64-108, evaluated on the host exactly as the reference does.
static double sq(double x) { return x * x; }
static double h_delta(double a, double r) { return sq(r) - 2. * r + a * a; }
static double h_sigma(double a, double r, double th) {
    double asq = a * a;
    return std::sqrt(sq(sq(r) + asq) - asq * h_delta(a, r) * sq(std::sin(th)));
}
static double h_w(double a, double r, double th) { return 2. * a * r / sq(h_sigma(a, r, th)); }
static double h_ro(double a, double r, double th) {
    return std::sqrt(sq(r) + a * a * sq(std::cos(th)));
}
static double h_wbar(double a, double r, double th) {
    return h_sigma(a, r, th) * std::sin(th) / h_ro(a, r, th);
}
static double h_alpha(double a, double r, double th) {
    return h_ro(a, r, th) * std::sqrt(h_delta(a, r)) / h_sigma(a, r, th);
}
static double h_b0(double a, double r0) {
    double asq = a * a;
    return -(r0 * r0 * r0 - 3. * sq(r0) + asq * r0 + asq) / (a * (r0 - 1.));
}

// The steepest slope |dq0/db| of the critical curve {(b0(r), q0(r)) : r in
// [r1, r2]} (for a = 0 the curve q = 27 - b^2, slope 2|b| <= 2 sqrt(27)),
// sampled densely with a 5% margin; infinity where b0 is not monotonic
// (then the curve is no graph over b).  Used only to rule rays out of the
// near-critical set cheaply: a point at vertical distance D from a graph of
// slope at most S lies at least D / sqrt(1 + S^2) from it.
static double crit_curve_slope(double a, double r1, double r2) {
    if (a == 0) return 1.05 * 2. * std::sqrt(27.);
    const int n = 4096;
    double smax = 0, prev_db = 0;
    for (int k = 0; k <= n; k++) {
        const double r = r1 + (r2 - r1) * k / n;
        const double asq = a * a;
        const double db = (asq + a - 2. * r * (r * r - 3. * r + 3.)) / (a * (r - 1.) * (r - 1.));
        const double N = -r * r * r * (r * r * r - 6. * r * r + 9. * r - 4. * asq);
        const double Nd = -(6. * std::pow(r, 5) - 30. * std::pow(r, 4) + 36. * r * r * r - 12. * asq * r * r);
        const double dq = (Nd * (r - 1.) - 2. * N) / (asq * std::pow(r - 1., 3));
        if (!(std::fabs(db) > 1e-12) || (k > 0 && db * prev_db <= 0)) return HUGE_VAL;
        prev_db = db;
        smax = std::max(smax, std::fabs(dq / db));
    }
    return 1.05 * smax;
}

double host_calc_speed(double a, double r, double theta) {
    double omega = 1. / (a + std::pow(r, 1.5));
    return h_wbar(a, r, theta) / h_alpha(a, r, theta) * (omega - h_w(a, r, theta));
}

int make_pose(const bt_pose* in, PoseK* P) {
    if (!in) return fail(BT_ERR_INVALID, "pose is null");
    if (!(in->r > 0) || !std::isfinite(in->a) || !std::isfinite(in->theta) ||
        !std::isfinite(in->phi) || !std::isfinite(in->r))
        return fail(BT_ERR_INVALID, "pose: non-finite or non-positive radius/angles");
    if (!(std::fabs(in->a) < 1.0)) return fail(BT_ERR_INVALID, "pose: |a| must be < 1");
    std::memset(P, 0, sizeof(*P));
    const double a = in->a;
    P->a = a;
    P->asq = a * a;
    P->r = in->r;
    P->theta = in->theta;
    P->phi = in->phi;
    if (in->user_speed) {
        // Camera.h:27-38
        P->speed = in->speed;
        P->bphi = 1;
        P->btheta = 0;
        P->br = 0;
    } else {
        // Camera.h:40-52
        P->bphi = in->bphi;
        P->btheta = in->btheta;
        P->br = in->br;
        P->speed = host_calc_speed(a, in->r, in->theta);
    }
    // Camera.h:54-61
    P->alpha = h_alpha(a, P->r, P->theta);
    P->w = h_w(a, P->r, P->theta);
    P->wbar = h_wbar(a, P->r, P->theta);
    P->Delta = h_delta(a, P->r);
    P->ro = h_ro(a, P->r, P->theta);
    if (!(P->Delta > 0)) return fail(BT_ERR_INVALID, "pose: camera inside the horizon");
    // Grid.h:528-547 per-pose factors
    const double sp = P->speed;
    P->ww = P->w * P->wbar;
    P->neg_s = -(double)sqrtf(1 - sp * sp);
    P->sqrtD = (double)sqrtf(P->Delta);
    P->k = std::sqrt(1 - P->btheta * P->btheta);
    P->brbt_k = P->br * P->btheta / P->k;
    P->bpbt_k = P->bphi * P->btheta / P->k;
    P->cos2S = std::cos(P->theta) * std::cos(P->theta);
    P->sin2S = std::sin(P->theta) * std::sin(P->theta);
    // Metric.h:146-147,164-165
    double r1 = 2. * (1. + std::cos(2. * std::acos(-a) / 3.));
    double r2 = 2. * (1. + std::cos(2. * std::acos(a) / 3.));
    P->b1 = h_b0(a, r2);
    P->b2 = h_b0(a, r1);
    P->r_ph1 = r1;
    P->r_ph2 = r2;
    P->crit_slope = crit_curve_slope(a, r1, r2);
    P->gr = (std::sqrt(5.0) + 1.) / 2.;
    P->a_zero = (a == 0) ? 1 : 0;
    P->r_cap = 0;
    return BT_OK;
}

static int sgnf(float v) { return (0 < v) - (v < 0); }

int make_params(const bt_integ_params* in, ParamsK* out) {
    bt_integ_params d;
    if (!in) {
        bt_integ_params_default(&d);
        in = &d;
    }
    if (!(in->eps > 0) || !(in->h1 > 0) || in->max_steps <= 0 || in->newton_cap <= 0)
        return fail(BT_ERR_INVALID, "integ params: eps, h1, max_steps, newton_cap must be > 0");
    if (in->precision != BT_PREC_FP64 && in->precision != BT_PREC_MIXED)
        return fail(BT_ERR_INVALID, "integ params: unknown precision");
    out->eps = in->eps;
    out->h1 = in->h1;
    out->z_end = in->z_end;
    out->r_escape = in->r_escape;
    out->max_steps = in->max_steps;
    out->apriori = in->apriori_capture ? 1 : 0;
    out->newton_cap = in->newton_cap;
    out->precision = in->precision;
    out->crit_eps = in->crit_eps;
    out->hmin_fp64 = in->hmin_fp64;
    if (!(in->noise_budget >= 0) || !(in->decision_margin >= 0) || !(in->drift_budget >= 0) ||
        !(in->polar_b >= 0))
        return fail(BT_ERR_INVALID,
                    "integ params: noise_budget, decision_margin, drift_budget and polar_b must be >= 0");
    out->noise_budget = (float)in->noise_budget;
    out->decision_margin = (float)in->decision_margin;
    out->drift_budget = (float)in->drift_budget;
    out->polar_b = in->polar_b;
    out->h0 = in->h1 * sgnf((float)in->z_end);   // Metric.h:287
    return BT_OK;
}

void stats_to_abi(const StatsK& s, bt_trace_stats* o) {
    o->rays = (int64_t)s.rays;
    o->integrated = (int64_t)s.integrated;
    o->attempts = (int64_t)s.attempts;
    o->attempts_fp64 = (int64_t)s.attempts_fp64;
    o->newton = (int64_t)s.newton;
    o->golden = (int64_t)s.golden;
    o->fallback_rays = (int64_t)s.fallback;
}

}  // namespace bt

using namespace bt;

extern "C" {

const char* bt_last_error(void) { return g_last_error.c_str(); }

int bt_abi_version(void) { return BT_ABI_VERSION; }

int bt_context_create(int device, bt_context** out) {
    if (!out) return fail(BT_ERR_INVALID, "out is null");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(BT_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= count) return fail(BT_ERR_DEVICE, "device index out of range");
    BT_HIP(hipSetDevice(device));
    bt_context* h = new bt_context();
    Context& c = h->impl;
    c.device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c.num_cus = prop.multiProcessorCount;
    // the no-allocation entry point's scratch address on this device (once
    // per device; a failure is reported by that entry point)
    (void)noalloc_setup(device);
    // two streams at the default priority: the main one and an auxiliary one
    // for work that overlaps it (joined back before a call returns)
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c.aux, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c.ev_hand, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c.ev_back, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c.ev_stage, hipEventDisableTiming) != hipSuccess) {
        delete h;
        return fail(BT_ERR_HIP, "hipStreamCreate failed");
    }
    *out = h;
    return BT_OK;
}

void bt_context_destroy(bt_context* h) {
    if (!h) return;
    Context& c = h->impl;
    (void)hipSetDevice(c.device);
    (void)hipStreamSynchronize(c.stream);
    (void)hipStreamSynchronize(c.aux);
    for (auto& kv : c.scratch)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    for (auto& kv : c.pool) (void)hipFree(kv.second);
    for (auto& t : c.pending) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto e : c.free_events) (void)hipEventDestroy(e);
    (void)hipEventDestroy(c.ev_fork);
    (void)hipEventDestroy(c.ev_join);
    (void)hipEventDestroy(c.ev_hand);
    (void)hipEventDestroy(c.ev_back);
    (void)hipEventDestroy(c.ev_stage);
    if (c.host_stage) (void)hipHostFree(c.host_stage);
    (void)hipStreamDestroy(c.aux);
    (void)hipStreamDestroy(c.stream);
    delete h;
}

int bt_context_sync(bt_context* h) {
    if (!h) return fail(BT_ERR_INVALID, "context is null");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    BT_HIP(hipSetDevice(h->impl.device));
    BT_HIP(hipStreamSynchronize(h->impl.stream));
    return BT_OK;
}

int bt_timing_enable(bt_context* h, int enable) {
    if (!h) return fail(BT_ERR_INVALID, "context is null");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    h->impl.timing = enable != 0;
    return BT_OK;
}

int bt_timer_get(bt_context* h, const char* kernel, int64_t* launches, double* total_ms) {
    if (!h || !kernel) return fail(BT_ERR_INVALID, "null argument");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    BT_HIP(hipSetDevice(h->impl.device));
    int rc = h->impl.flush_timers();
    if (rc) return rc;
    auto it = h->impl.timers.find(kernel);
    if (launches) *launches = it == h->impl.timers.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == h->impl.timers.end() ? 0.0 : it->second.second;
    return BT_OK;
}

int bt_timer_reset(bt_context* h) {
    if (!h) return fail(BT_ERR_INVALID, "context is null");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    int rc = h->impl.flush_timers();
    h->impl.timers.clear();
    return rc;
}

int bt_camera_init(const bt_pose* pose, double* out6, float* cam7) {
    PoseK P;
    int rc = make_pose(pose, &P);
    if (rc) return rc;
    if (out6) {
        out6[0] = P.speed;
        out6[1] = P.alpha;
        out6[2] = P.w;
        out6[3] = P.wbar;
        out6[4] = P.Delta;
        out6[5] = P.ro;
    }
    if (cam7) {
        // Camera.h:67-78
        cam7[0] = (float)P.speed;
        cam7[1] = (float)P.alpha;
        cam7[2] = (float)P.w;
        cam7[3] = (float)P.wbar;
        cam7[4] = (float)P.br;
        cam7[5] = (float)P.btheta;
        cam7[6] = (float)P.bphi;
    }
    return BT_OK;
}

void bt_integ_params_default(bt_integ_params* p) {
    if (!p) return;
    p->eps = 1e-5;
    p->h1 = 0.01;
    p->z_end = -10000000;
    p->max_steps = 1000;
    p->apriori_capture = 1;
    p->r_escape = 0.0;
    p->r_capture_factor = 0.0;
    p->newton_cap = 100000;
    p->precision = BT_PREC_FP64;
    p->crit_eps = 0.03;
    p->hmin_fp64 = 1e-6;
    p->noise_budget = 0.13;
    p->decision_margin = 4.0;
    p->drift_budget = 1.2e-6;
    p->polar_b = 1e-3;
}

static int trace_common(bt_context* h, const bt_pose* pose, const bt_integ_params* prm,
                        int64_t n, double* d_th, double* d_ph, int32_t* d_step,
                        bt_trace_stats* stats) {
    Context& c = h->impl;
    PoseK P;
    ParamsK K;
    int rc;
    if ((rc = make_pose(pose, &P))) return rc;
    if ((rc = make_params(prm, &K))) return rc;
    if (prm && prm->r_capture_factor > 0)
        P.r_cap = prm->r_capture_factor * (1. + std::sqrt(1. - P.asq));
    void* p;
    if ((rc = c.buffer("trace_stats", sizeof(StatsK), &p))) return rc;
    StatsK* d_stats = (StatsK*)p;
    if (int zr = zero_words((unsigned*)d_stats, (int)(sizeof(StatsK) / 4), c.stream)) return zr;
    RayIO io;
    std::memset(&io, 0, sizeof(io));
    io.in_th = d_th;
    io.in_ph = d_ph;
    io.out_th = d_th;
    io.out_ph = d_ph;
    io.ostride = 1;
    io.out_step = d_step;
    if ((rc = trace(&c, P, K, n, io, d_stats))) return rc;
    if (stats) {
        StatsK s;
        BT_HIP(hipMemcpyAsync(&s, d_stats, sizeof(s), hipMemcpyDeviceToHost, c.stream));
        BT_HIP(hipStreamSynchronize(c.stream));
        stats_to_abi(s, stats);
    }
    return BT_OK;
}

int bt_trace_rays_device(bt_context* h, const bt_pose* pose, const bt_integ_params* prm,
                         int64_t n, double* d_theta, double* d_phi, int32_t* d_step,
                         bt_trace_stats* stats) {
    if (!h || n < 0 || (n > 0 && (!d_theta || !d_phi || !d_step)))
        return fail(BT_ERR_INVALID, "bt_trace_rays_device: null argument or n < 0");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    BT_HIP(hipSetDevice(h->impl.device));
    return trace_common(h, pose, prm, n, d_theta, d_phi, d_step, stats);
}

int bt_trace_rays(bt_context* h, const bt_pose* pose, const bt_integ_params* prm, int64_t n,
                  double* theta, double* phi, int32_t* step, bt_trace_stats* stats) {
    if (!h || n < 0 || (n > 0 && (!theta || !phi || !step)))
        return fail(BT_ERR_INVALID, "bt_trace_rays: null argument or n < 0");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    Context& c = h->impl;
    BT_HIP(hipSetDevice(c.device));
    if (n == 0) return BT_OK;
    void *dt, *dp, *ds;
    int rc;
    if ((rc = c.buffer("io_theta", sizeof(double) * n, &dt))) return rc;
    if ((rc = c.buffer("io_phi", sizeof(double) * n, &dp))) return rc;
    if ((rc = c.buffer("io_step", sizeof(int32_t) * n, &ds))) return rc;
    BT_HIP(hipMemcpyAsync(dt, theta, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    BT_HIP(hipMemcpyAsync(dp, phi, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    if ((rc = trace_common(h, pose, prm, n, (double*)dt, (double*)dp, (int32_t*)ds, stats)))
        return rc;
    BT_HIP(hipMemcpyAsync(theta, dt, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    BT_HIP(hipMemcpyAsync(phi, dp, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    BT_HIP(hipMemcpyAsync(step, ds, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c.stream));
    BT_HIP(hipStreamSynchronize(c.stream));
    return BT_OK;
}

}  // extern "C"

extern "C" {

int bt_device_alloc(bt_context* h, size_t bytes, void** out) {
    if (!h || !out || bytes == 0) return fail(BT_ERR_INVALID, "bt_device_alloc: bad argument");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    BT_HIP(hipSetDevice(h->impl.device));
    if (hipMalloc(out, bytes) != hipSuccess) {
        *out = nullptr;
        return fail(BT_ERR_NOMEM, "bt_device_alloc: hipMalloc failed");
    }
    return BT_OK;
}

void bt_device_free(bt_context* h, void* ptr) {
    if (!h || !ptr) return;
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    (void)hipSetDevice(h->impl.device);
    (void)hipStreamSynchronize(h->impl.stream);
    (void)hipFree(ptr);
}

int bt_memcpy(bt_context* h, void* dst, const void* src, size_t bytes, int to_device) {
    if (!h || (bytes && (!dst || !src))) return fail(BT_ERR_INVALID, "bt_memcpy: bad argument");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    BT_HIP(hipSetDevice(h->impl.device));
    BT_HIP(hipMemcpyAsync(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost,
                          h->impl.stream));
    BT_HIP(hipStreamSynchronize(h->impl.stream));
    return BT_OK;
}

}  // extern "C"
