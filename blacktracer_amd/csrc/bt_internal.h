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
    double r_ph1, r_ph2;  // prograde / retrograde equatorial photon-orbit radii (Metric.h:146-147)
    double crit_slope;    // max |dq0/db| along the critical curve (inf: not a graph over b)
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
    float  noise_budget, decision_margin, drift_budget;
    double polar_b;
};

// Device counters, one set per launch.
struct StatsK {
    unsigned long long rays, integrated, attempts, attempts_fp64, newton, golden, fallback;
};

int make_pose(const bt_pose* in, PoseK* out);
// noalloc.hip: looks up the device's copy of the no-allocation entry
// point's scratch array (once per device, with that device current).
int noalloc_setup(int device);
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

// Where the rays of one launch come from.  SingleSrc: one pose, one RayIO
// (kernel arguments).  MultiSrc: a batch of segments, each with its own
// pose and RayIO; global ray k belongs to the segment s with
// start[s] <= k < start[s+1] and is ray k - start[s] of that segment.
struct SingleSrc {
    static constexpr bool kMulti = false;
    PoseK P;
    RayIO io;
    __device__ __forceinline__ int seg(int64_t) const { return 0; }
    __device__ __forceinline__ const PoseK& pose(int) const { return P; }
    __device__ __forceinline__ const RayIO& rio(int) const { return io; }
    __device__ __forceinline__ int64_t local(int, int64_t k) const { return k; }
};

struct MultiSrc {
    static constexpr bool kMulti = true;
    const PoseK* poses;
    const RayIO* ios;
    const int64_t* start;   // nseg + 1 offsets
    int nseg;
    __device__ __forceinline__ int seg(int64_t k) const {
        int lo = 0, hi = nseg - 1;   // largest s with start[s] <= k
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (start[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        return lo;
    }
    __device__ __forceinline__ const PoseK& pose(int s) const { return poses[s]; }
    __device__ __forceinline__ const RayIO& rio(int s) const { return ios[s]; }
    __device__ __forceinline__ int64_t local(int s, int64_t k) const { return k - start[s]; }
};

// Host description of a multi-pose batch (one segment per pose).
struct BatchHost {
    std::vector<PoseK> poses;
    std::vector<RayIO> ios;
    std::vector<int64_t> counts;
};

// ---------------------------------------------------------------- context
struct PendingTimer {
    std::string name;
    hipEvent_t start, stop;
    hipStream_t stream;
};

struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Launch-shape knobs of the persistent integrators.
struct Tune {
    unsigned chunk = 128;   // queue positions a wave reserves per atomic (>= 64)
    unsigned block = 256;   // threads per workgroup (64, 128 or 256)
    int waves = 0;          // resident waves per SIMD (0: as many as registers allow)
};

struct Context {
    int device = 0;
    Tune tune;
    int64_t launch_lanes = 0;   // widest persistent launch since last reset (lanes)
    int num_cus = 256;
    hipStream_t stream = nullptr;
    hipStream_t aux = nullptr;                 // overlapping work, joined before a call returns
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_hand = nullptr, ev_back = nullptr;   // mid-call hand-offs between the two streams
    std::recursive_mutex mu;
    bool timing = false;
    std::vector<PendingTimer> pending;
    std::vector<hipEvent_t> free_events;
    std::map<std::string, std::pair<int64_t, double>> timers;
    std::map<std::string, Buffer> scratch;   // grow-only named work buffers
    unsigned long long* d_counters = nullptr;  // small device scratch for counters
    // Device blocks of freed grids, kept for the next grid of a similar size
    // (a grid build would otherwise pay hipMalloc/hipFree of its arrays).
    std::multimap<size_t, void*> pool;
    size_t pool_bytes = 0;
    static constexpr size_t kPoolCap = size_t(32) << 30;
    size_t pool_cap = kPoolCap;   // pooled bytes kept at most (0: every freed block is freed)
    // Dataflow grid builds (grid_flow.hip): the epoch that tags their queue
    // entries, and the queue buffers as last cleared (a buffer that is new,
    // regrown or reallocated is cleared before its first use).
    unsigned flow_epoch = 0;
    void* flow_rq_seen = nullptr;
    void* flow_tq_seen = nullptr;
    uint64_t flow_rcap = 0, flow_tcap = 0;

    // A block of at least `bytes` (reused when one of up to 2x that size is
    // pooled); *got receives its actual size.
    int pool_alloc(size_t bytes, void** out, size_t* got);
    // Returns a block to the pool (or frees it beyond kPoolCap); the caller
    // has synchronised the stream.
    void pool_free(void* p, size_t bytes);

    // Returns a device buffer of at least `bytes` for `name`.
    int buffer(const std::string& name, size_t bytes, void** out);
    // Device bytes the named buffers hold now.
    size_t buffer_bytes() const;
    // Frees every named buffer (after the stream has drained); the next
    // call that needs one allocates it again.
    void release_buffers();

    // Pinned host staging for small tables copied to the device on the main
    // stream: stage() waits until the copies from the previous staging have
    // run (stage_done() marks them on the stream), so host-to-device copies
    // of these tables need no synchronisation of their own.
    void* host_stage = nullptr;
    size_t host_stage_bytes = 0;
    hipEvent_t ev_stage = nullptr;
    int stage(size_t bytes, void** out);
    int stage_done();
    int timer_begin(const char* name, int* slot, hipStream_t st = nullptr);
    int timer_end(int slot);
    int flush_timers();
};

// Scoped per-launch timer.
struct KTimer {
    Context* c;
    int slot = -1;
    KTimer(Context* ctx, const char* name, hipStream_t st = nullptr) : c(ctx) {
        if (c->timing) c->timer_begin(name, &slot, st);
    }
    ~KTimer() {
        if (slot >= 0) c->timer_end(slot);
    }
};

// Work queue of escaping rays: FP64 constants of motion indexed by ray,
// and a queue of ray indices with its length/head counters.
struct QueueK {
    double* pr;
    double* pt;
    double* b;
    double* q;
    uint32_t* slots;    // queue of ray indices
    unsigned* counters; // [0] = queue length, [1] = head
};

// The queue counters of one trace: kCtrLines lines of kCtrLine words (128
// bytes), one per queue or head that kernels running at the same time
// update with atomics (the mixed path: its warm-up queue, the bulk's head
// on it, the near-critical queue, the restart queue), so that their
// atomics do not queue behind each other at one cache line.
constexpr int kCtrLine = 32;
constexpr int kCtrLines = 4;
constexpr int kQueueCtrWords = kCtrLine * kCtrLines;

// Per-ray state handed from one integrator kernel to another (indexed by
// ray): y holds six arrays of `stride` doubles (r, theta, phi, pR, pTheta, z).
struct ContK {
    double* y;
    double* h;
    int32_t* nstp;
    int64_t stride;
};

// ---------------------------------------------------------------- launchers
// kerr_f64.hip: the plain FP64 path (Metric.h / Grid.h semantics).
// Traces n rays described by io; d_stats accumulates counters (device).
int trace_f64(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
              StatsK* d_stats);
// Building blocks shared by the FP64 and mixed paths, for one pose
// (SingleSrc) or a multi-pose batch (MultiSrc); ray indices are global.
// queue_alloc: per-ray arrays for n rays (context buffers).  trace_setup:
// zeroes the counters, computes initial conditions and the capture test;
// escaping rays go to Q, or, when Q64 is given, near-critical rays go to
// Q64 instead.  trace_f64_queue: traces every ray queued in Q in FP64 from
// its initial state (at most n_max rays); with cont given (warm-up mode) a
// ray leaves after its first attempt with errmax > ERRCON and its state
// goes to cont (cont.nstp = -1 for a ray that ended during the warm-up);
// prio is the waves' base issue priority (0..3).
// Zeroes n (<= 1024) 32-bit words with a kernel (no memset node, so the
// call can sit in a captured graph).
int zero_words(unsigned* p, int n, hipStream_t st);
int queue_alloc(Context* c, int64_t n, QueueK* Q);
template <class Src>
int trace_setup(Context* c, const Src& S, const ParamsK& prm, int64_t n, const QueueK& Q,
                const QueueK* Q64, StatsK* d_stats);
template <class Src>
int trace_f64_queue(Context* c, const Src& S, const ParamsK& prm, int64_t n_max, const QueueK& Q,
                    StatsK* d_stats, const char* timer, const ContK* cont,
                    hipStream_t stream = nullptr, int prio = 0);
template <class Src>
int trace_f64_src(Context* c, const Src& S, const ParamsK& prm, int64_t n, StatsK* d_stats);
// kerr_mixed.hip: FP32 bulk with FP64 set-up, warm-up, fallback and accumulation.
template <class Src>
int trace_mixed_src(Context* c, const Src& S, const ParamsK& prm, int64_t n, StatsK* d_stats);
int trace_mixed(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
                StatsK* d_stats);

inline int trace(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
                 StatsK* d_stats) {
    return prm.precision == BT_PREC_MIXED ? trace_mixed(c, pose, prm, n, io, d_stats)
                                          : trace_f64(c, pose, prm, n, io, d_stats);
}

// Traces every segment of a multi-pose batch in one set of launches
// (kerr_mixed.hip): the poses and RayIOs go to device memory, the rays of
// all segments share one work queue.
int trace_batch(Context* c, const BatchHost& b, const ParamsK& prm, StatsK* d_stats);

void stats_to_abi(const StatsK& s, bt_trace_stats* out);

}  // namespace bt

// Opaque ABI handles.
struct bt_context {
    bt::Context impl;
};
