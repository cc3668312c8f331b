// render_abi.cpp -- C ABI of the render pass (makeImage).
#include "bt_internal.h"
using namespace bt;
extern "C" {
void bt_render_flags_default(bt_render_flags* f) {
    if (!f) return;
    f->star = 1; f->map = 1; f->lr = 1; f->hor = 0.f; f->ver = 0.f; f->offset = 0.f;
}
}

extern "C" int bt_make_image(bt_context* ctx, const float* stars, const int32_t* star_tree,
                             int star_size, const float* cam_param, const float* magnitude,
                             int tree_level, int M, int N, int step, const uint8_t* sky_bgr,
                             int sky_rows, int sky_cols, int G, float grid_start, float grid_step,
                             const float* hit, const float* viewthing, float view_angle, int GM,
                             int GN, int gridlvl, const int32_t* offset_tables,
                             const float* hash_tables, const int32_t* hash_pos_tag,
                             const int32_t* table_sizes, int otsize, int htsize,
                             const bt_render_flags* flags, uint8_t* out_bgra) {
    return fail(BT_ERR_STATE, "render pass not built yet");
}
