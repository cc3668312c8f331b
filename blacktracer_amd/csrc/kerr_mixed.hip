// kerr_mixed.hip -- mixed-precision ray path (FP32 Cash-Karp bulk).
#include "bt_internal.h"

namespace bt {

int trace_mixed(Context* c, const PoseK& pose, const ParamsK& prm, int64_t n, const RayIO& io,
                StatsK* d_stats) {
    (void)c; (void)pose; (void)prm; (void)n; (void)io; (void)d_stats;
    return fail(BT_ERR_STATE, "mixed-precision path not built yet");
}

}  // namespace bt
