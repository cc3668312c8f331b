// noalloc.hip -- bt_trace_rays_device_noalloc: the FP64 ray-batch entry
// point (bt_trace_rays_device, i.e. Grid::integration_wrapper on device
// arrays) for callers that cannot allocate device memory: code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc.
//
// Its kernels are the plain FP64 path's (ic_capture, rk_f64); their work
// queue and counters, which bt_trace_rays_device takes from the context's
// grow-only buffers, live here in one file-scope device array of fixed
// size instead.  The array exists once per GPU (each device holds its own
// copy until the process ends); its address on a device is looked up once,
// when the first context on that device is created, and only that address
// is handed to kernels.  Calls that overlap on one GPU share the array and
// race; the library's context mutex serialises calls through one context
// only.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "bt_internal.h"

namespace bt {
namespace {

// Layout of the array for BT_NOALLOC_MAX_RAYS rays: the queue's four FP64
// arrays (pR, pTheta, b, q), its slot list, its counter lines and the
// launch's statistics block, each at a 256-byte boundary.
constexpr int64_t kMaxRays = BT_NOALLOC_MAX_RAYS;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }
constexpr size_t kOffPr = 0;
constexpr size_t kOffPt = kOffPr + align_up(sizeof(double) * kMaxRays);
constexpr size_t kOffB = kOffPt + align_up(sizeof(double) * kMaxRays);
constexpr size_t kOffQ = kOffB + align_up(sizeof(double) * kMaxRays);
constexpr size_t kOffSlots = kOffQ + align_up(sizeof(double) * kMaxRays);
constexpr size_t kOffCounters = kOffSlots + align_up(sizeof(uint32_t) * kMaxRays);
constexpr size_t kOffStats = kOffCounters + align_up(sizeof(unsigned) * kQueueCtrWords);
constexpr size_t kScratchBytes = kOffStats + align_up(sizeof(StatsK));
static_assert(kScratchBytes < (size_t(40) << 20), "the no-allocation scratch stays near 36 MiB");

// Never used from host code as an address: host code uses the per-device
// pointer looked up below.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];          // device address of scratch on each device
int g_state[kMaxDevices];          // 0 not looked up, 1 ready, -1 lookup failed
hipError_t g_err[kMaxDevices];

}  // namespace

// Looks up this device's copy of the scratch array, once per device; called
// by bt_context_create with the device current.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return BT_OK;   // the entry point reports it
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_state[device] != 0) return BT_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {   // recorded; the entry point reports it
        g_state[device] = -1;
        g_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        return BT_ERR_HIP;
    }
    g_ptr[device] = p;
    g_state[device] = 1;
    return BT_OK;
}

}  // namespace bt

using namespace bt;

extern "C" int bt_trace_rays_device_noalloc(bt_context* h, const bt_pose* pose, const bt_integ_params* prm,
                                            int64_t n, double* d_theta, double* d_phi, int32_t* d_step) {
    if (!h || n < 0 || (n > 0 && (!d_theta || !d_phi || !d_step)))
        return fail(BT_ERR_INVALID, "bt_trace_rays_device_noalloc: null argument or n < 0");
    if (n > kMaxRays)
        return fail(BT_ERR_INVALID, "bt_trace_rays_device_noalloc: n = " + std::to_string(n) +
                                        " exceeds BT_NOALLOC_MAX_RAYS (" + std::to_string(kMaxRays) + ")");
    if (prm && prm->precision != BT_PREC_FP64)
        return fail(BT_ERR_INVALID, "bt_trace_rays_device_noalloc: only the FP64 path (BT_PREC_FP64)");
    std::lock_guard<std::recursive_mutex> lk(h->impl.mu);
    Context& c = h->impl;
    BT_HIP(hipSetDevice(c.device));
    if (c.device >= kMaxDevices) return fail(BT_ERR_DEVICE, "bt_trace_rays_device_noalloc: device index too large");
    char* base;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (g_state[c.device] != 1)
            return fail(BT_ERR_HIP, std::string("bt_trace_rays_device_noalloc: scratch address unavailable: ") +
                                        hipGetErrorString(g_err[c.device]));
        base = (char*)g_ptr[c.device];
    }
    PoseK P;
    ParamsK K;
    int rc;
    if ((rc = make_pose(pose, &P))) return rc;
    if ((rc = make_params(prm, &K))) return rc;
    if (prm && prm->r_capture_factor > 0)
        P.r_cap = prm->r_capture_factor * (1. + std::sqrt(1. - P.asq));
    if (n == 0) return BT_OK;

    QueueK Q;
    Q.pr = (double*)(base + kOffPr);
    Q.pt = (double*)(base + kOffPt);
    Q.b = (double*)(base + kOffB);
    Q.q = (double*)(base + kOffQ);
    Q.slots = (uint32_t*)(base + kOffSlots);
    Q.counters = (unsigned*)(base + kOffCounters);
    StatsK* d_stats = (StatsK*)(base + kOffStats);

    RayIO io;
    std::memset(&io, 0, sizeof(io));
    io.in_th = d_theta;
    io.in_ph = d_phi;
    io.out_th = d_theta;
    io.out_ph = d_phi;
    io.ostride = 1;
    io.out_step = d_step;
    SingleSrc S{P, io};

    // kernel timers would create events: off for this call
    struct TimingOff {
        Context& c;
        bool was;
        explicit TimingOff(Context& cc) : c(cc), was(cc.timing) { c.timing = false; }
        ~TimingOff() { c.timing = was; }
    } off(c);
    // the statistics block starts at zero (trace_setup clears the counters)
    if ((rc = zero_words((unsigned*)d_stats, (int)(sizeof(StatsK) / 4), c.stream))) return rc;
    if ((rc = trace_setup(&c, S, K, n, Q, nullptr, d_stats))) return rc;
    return trace_f64_queue(&c, S, K, n, Q, d_stats, "rk_f64", nullptr);
}
