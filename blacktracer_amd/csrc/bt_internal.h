// bt_internal.h -- shared host-side definitions of the blacktracer_amd
// library: context, error reporting, device-side POD argument blocks and
// the launchers each translation unit exports to the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bt_abi.h"

namespace bt {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define BT_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::bt::fail(BT_ERR_HIP, std::string(#call) + ": " +             \
                                              hipGetErrorString(e_));             \
    } while (0)

// ---------------------------------------------------------------- pose
// Everything a kernel needs about one pose, computed once on the host with
// the host libm so that per-pose values are bit-identical to the CPU path
// (Camera.h:40-61, Grid.h:524-547, Metric.h:145-151).
struct PoseK {
    double a, asq;
    double r, theta, phi;
    double speed, br, btheta, bphi;
    double alpha, w, wbar, ro, Delta;
    double ww;         // w * wbar
    double neg_s;      // -(double)sqrtf((float)(1 - speed*speed))   (Grid.h:529)
    double sqrtD;      // (double)sqrtf((float)Delta)                  (Grid.h:542)
    double k;          // sqrt(1 - btheta*btheta)                      (Grid.h:532)
    double brbt_k;     // br*btheta/k
    double bpbt_k;     // bphi*btheta/k
    double cos2S;      // cos(thetaS)*cos(thetaS)                      (Grid.h:547)
    double sin2S;      // sin(thetaS)*sin(thetaS)
    double b1, b2;     // _b0(_r2), _b0(_r1)                           (Metric.h:164-165)
    double r_cap;      // capture radius (0: off)
    double gr;         // golden ratio (sqrt(5)+1)/2                   (Metric.h:111)
    int    a_zero;     // checkRup returns false when a == 0           (Metric.h:128)
    int    pad;
};

struct ParamsK {
    double eps, h1, z_end, r_escape;
    int    max_steps, apriori, newton_cap, precision;
    double crit_eps, hmin_fp64;
    double h0;         // h1 * sgn((float)z_end)                      (Metric.h:287)
};

// Device counters, one set per launch.
struct StatsK {
    unsigned long long rays, integrated, attempts, attempts_fp64, newton, golden, fallback;
};

int make_pose(const bt_pose* in, PoseK* out);
int make_params(const bt_integ_params* in, ParamsK* out);
double host_calc_speed(double a, double r, double theta);

// Where a batch of rays comes from and where results go.
// Input: explicit (theta, phi) arrays, or grid keys (row-major cell index)
// mapped to theta = i/(N-1)*PI/(2-eq), phi = j/M*2PI (Grid.h:386-390).
// Output: theta'/phi' at out_th[o*ostride], out_ph[o*ostride] and
// step at out_step[o], where o = keys ? keys[k] : k.
struct RayIO {
    const double* in_th;
    const double* in_ph;
    const uint32_t* keys;
    int N, M, eq;
    double* out_th;
    double* out_ph;
    int ostride;
    int32_t* out_step;
};

// ---------------------------------------------------------------- context
struct PendingTimer {
    std::string name;
    hipEvent_t start, stop;
};

struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct Context {
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    bool timing = false;
    std::vector<PendingTimer> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, std::pair<int64_t, double>> timers;
    std::map<std::string, Buffer> scratch;   // grow-only named work buffers
    unsigned long long* d_counters = nullptr;  // small device scratch for counters

    // Returns a device buffer of at least `bytes` for `name`.
    int buffer(const std::string& name, size_t bytes, void** out);
    int timer_begin(const char* name, int* slot);
    int timer_end(int slot);
    int flush_timers();
};

// Scoped per-launch timer.
struct KTimer {
    Context* c;
    int slot = -1;
    KTimer(Context* ctx, const char* name) : c(ctx) {
        if (c->timing) c->timer_begin(name, &slot);
    }
    ~KTimer() {
        if (slot >= 0) c->timer_end(slot);
    }
};

// ---------------------------------------------------------------- launchers
// kerr_f64.hip: the plain FP64 path (Metric.h / Grid.h semantics).
// Traces n rays described by io; d_stats accumulates counters (device).
int trace_f64(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
              StatsK* d_stats);
// kerr_mixed.hip: FP32 bulk with FP64 set-up, fallback and accumulation.
int trace_mixed(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
                StatsK* d_stats);

inline int trace(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
                 StatsK* d_stats) {
    return prm.precision == BT_PREC_MIXED ? trace_mixed(c, pose, prm, n, io, d_stats)
                                          : trace_f64(c, pose, prm, n, io, d_stats);
}

void stats_to_abi(const StatsK& s, bt_trace_stats* out);

}  // namespace bt

// Opaque ABI handles.
struct bt_context {
    bt::Context impl;
};
