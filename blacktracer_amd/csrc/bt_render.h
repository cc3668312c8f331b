// bt_render.h -- device-resident inputs of the render pass and its launcher.
#pragma once

#include "bt_internal.h"

namespace bt {

// Dense FP32 celestial grid (GN rows x GM columns of float2).
struct GridF {
    const float* grid;
    int GN, GM, gridlvl;
};

// Stars (StarProcessor layout) and the summed-area sky table.
struct SceneK {
    const int32_t* tree;
    int tree_level;
    const float* star_pos;   // 2 per star: theta, phi
    const float* star_mag;   // 2 per star: magnitude, B-V
    int S;
    const float* table;      // rows * cols float4
    int rows, cols;
    // binned copy of the stars: kBinT x kBinP cells over (theta, phi), stars
    // sorted by cell (then by tree order); cell c holds [bin_start[c], bin_start[c+1])
    const int32_t* bin_start;
    const float* bin_pos;
    const float* bin_mag;
    // every star's magnitude lies in [-10, 60] (no NaN): with a pixel's
    // lensing in range, no tap of any star needs the reference's slow
    // expression
    int mag_tame;
    const int32_t* bin_idx;   // tree-order index of each binned star
};

constexpr int kBinT = 256;
constexpr int kBinP = 512;
constexpr float kBinDT = 3.141592653589793238f / kBinT;
constexpr float kBinDP = 6.283185307179586476f / kBinP;

// Host: the binned star arrays for SceneK (bin_start kBinT*kBinP+1 ints,
// pos/mag 2*S floats each, bin_idx S ints), from the tree-ordered arrays.
void star_bins(const float* pos, const float* mag, int S, int32_t* bin_start, float* bin_pos, float* bin_mag,
               int32_t* bin_idx);

// Pixel-corner camera-sky directions ((H+1) x (W+1) float2).
struct ViewK {
    const float* view;
    int W, H;
    const float* pix_area;   // per-pixel view solid angle, or null (made per frame)
};

// Optional device copies of intermediates (for parity tests).
struct FrameDebug {
    float* thphi;
    int* gap;
    uint8_t* bh;
    float* area;
    float* star;
};

int render_sum_table(Context* c, const uint8_t* d_sky, int rows, int cols, float* d_table);
int render_grid_from_psh(Context* c, const float* d_ht, const int32_t* d_ot, const int32_t* d_tag, int hw,
                         int ow, int GN, int GM, float* d_grid, int* GN1_out);
int render_grid_from_cel(Context* c, const double* d_cel, int GN, int GM, float* d_grid, int* GN1_out);
// Shadow-border directions of multi-grid blending (MakeImageKernel.cu:1257).
constexpr int kAngleNum = 1000;
// The second grid of a blended (G > 1) frame; render_frame fills border, a
// device array of 2 + 2 * kAngleNum float2, from the two grids.
struct Blend {
    GridF b;
    float alpha;
    float* border;
};

// Star trails of a frame sequence (MakeImageKernel.cu:98-174,618-629):
// per star, the pixels of its images in the last two frames.
constexpr int kTrailNum = 30;    // images kept per star and frame (MakeImageKernel.cu:1255)
constexpr int kTrailHits = 32;   // images recorded per star and frame
struct TrailK {
    int S;              // stars
    int cur;            // cache half this frame writes (the other is the previous frame's)
    int2* cache;        // [2][S][kTrailNum] (row, column)
    int* count;         // [2][S]
    unsigned* nhits;    // [S] images this frame
    uint32_t* hit_pix;  // [S][kTrailHits] pixel index row * W + column
    float* hit_data;    // [S][kTrailHits][3] part, frac, blackbody index
    float* grad;        // (H+1)(W+1) float2
    float* trail;       // W*H float3
    unsigned* stats;    // [1]: images beyond kTrailHits this frame (dropped)
};
// Per-frame trail and diffraction inputs of render_frame.
struct TrailFrame {
    TrailK k;
    bool first;              // frame 0: no previous images
    int W, H;                // the frame size the buffers were made for
    const uint8_t* sprite;   // diffraction sprite, fs x fs RGB8, or null
    int fs;
};

int render_frame(Context* c, const GridF& g, const SceneK& s, const ViewK& v, const float* cam7,
                 const bt_render_flags& f, int step, uint8_t* d_out, FrameDebug* dbg,
                 const Blend* blend = nullptr, const TrailFrame* tf = nullptr);

// Host-side inputs (Viewer.h, StarProcessor.h).
int viewer_matrix(double view_angle, double off_up, double off_right, int w, int h, int sphere, float* out);
int star_tree(const float* stars4, int S, int tree_level, int32_t* tree, float* pos, float* mag);
// Camera::getParamArray (Camera.h:67-78) of a pose.
void cam7_of(const PoseK& P, float* cam7);

}  // namespace bt

// Device-resident render inputs behind the bt_scene / bt_view handles.
struct bt_scene {
    bt::Context* ctx = nullptr;
    int32_t* tree = nullptr;
    float* pos = nullptr;
    float* mag = nullptr;
    float* table = nullptr;
    int32_t* bin_start = nullptr;
    float* bin_pos = nullptr;
    float* bin_mag = nullptr;
    int32_t* bin_idx = nullptr;
    int tree_level = 0, S = 0, rows = 0, cols = 0;
    int mag_tame = 1;
    size_t bytes = 0;
    bt::SceneK k() const {
        return bt::SceneK{tree, tree_level, pos, mag, S, table, rows, cols, bin_start, bin_pos, bin_mag, mag_tame,
                          bin_idx};
    }
};

struct bt_view {
    bt::Context* ctx = nullptr;
    float* view = nullptr;       // (H+1)(W+1) corners, then W*H pixel solid angles
    float* pix_area = nullptr;   // inside the view block
    int W = 0, H = 0;
    size_t bytes = 0;
};

namespace bt {
struct GridF;
// The FP32 render grid of a device grid (made once per grid, thread-safe).
int grid_f32_of(Context& c, const bt_grid* g, GridF* out);
// Device bytes a frame of w x h pixels uses at star-splat step `step`.
size_t frame_bytes(int w, int h, int step);
// Fills out (W*H floats) with each pixel's view solid angle, on c's stream.
int view_pixel_areas(Context* c, const float* view, int W, int H, float* out);
}  // namespace bt
