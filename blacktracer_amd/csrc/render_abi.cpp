// render_abi.cpp -- C ABI of the render pass: the makeImage drop-in
// (MakeImageKernel.cu:1206-1236) and the device-resident scene / view /
// frame objects used for serving.  Also the host-side inputs the reference
// computes on the CPU: Viewer::makeCameraView (Viewer.h:109-133) and
// StarProcessor::makeTree (StarProcessor.h:172-212).
#include <cmath>
#include <cstring>
#include <vector>

#include "bt_grid.h"
#include "bt_render.h"

namespace bt {

static const double kPI = 3.141592653589793238;   // Const.h:2-4
static const double kPI1_2 = kPI / 2.;
static const double kPI2 = kPI * 2.;

// Viewer.h:20-33 + 109-133 (perspective) and 82-88 (sphere)
int viewer_matrix(double viewangle, double offsetVer, double offsetHor, int pixw, int pixh, int sphere,
                  float* view) {
    if (pixw <= 0 || pixh <= 0 || !view) return fail(BT_ERR_INVALID, "viewer: bad size or null output");
    const size_t n = (size_t)(pixh + 1) * (pixw + 1);
    std::memset(view, 0, sizeof(float) * 2 * n);
    if (sphere) {
        for (int i = 0; i < pixh + 1; i++)
            for (int j = 0; j < pixw + 1; j++) {
                size_t o = (size_t)i * (pixw + 1) + j;
                view[2 * o] = (float)(1.f * i * kPI / pixh);
                view[2 * o + 1] = (float)(1.f * j * kPI2 / pixw);
            }
        return BT_OK;
    }
    double viewAngleWide = viewangle;
    double viewAngleUp = 1. * pixh * viewAngleWide / (1. * pixw);
    if (std::fabs(offsetVer) > (kPI - viewAngleUp) / 2.)
        return fail(BT_ERR_INVALID, "viewer: offsetUp too big");   // Viewer.h:110-113
    float yleft = (float)std::tan(viewAngleWide / 2.f);
    float step = (2.f * yleft) / pixw;
    float zup = yleft * pixh / pixw;
    for (int i = 0; i < pixh + 1; i++) {
        float zval = zup - i * step;
        float halfH = i * 1.f / (pixh / 2.f);
        float theta = halfH == 1.f ? (float)(kPI1_2 - offsetVer)
                                   : (float)(std::atan(1. / zval) + (halfH > 1.f) * kPI - offsetVer);
        for (int j = 0; j < pixw + 1; j++) {
            float yval = yleft - j * step;
            float halfV = j * 1.f / (pixw / 2.f);
            float phi = halfV == 1.f ? (float)(kPI - offsetHor)
                                     : (float)(std::atan(1. / yval) + (halfV > 1.f) * kPI + kPI1_2 - offsetHor);
            size_t o = (size_t)(pixw + 1) * i + pixw - j;
            view[2 * o] = theta;
            view[2 * o + 1] = phi;
        }
    }
    return BT_OK;
}

namespace {
struct TreeBuilder {
    const float* s4;
    int level_max;
    int32_t* tree;
    float* pos;
    float* mag;
    // StarProcessor.h:172-212, stars as an index list
    void build(std::vector<int>& idx, int level, float thphi[2], float size[2], int writePos) {
        const int n = (int)idx.size();
        int searchPos = 0;
        if (writePos != 0 && ((writePos + 1) & writePos) != 0) searchPos = tree[writePos - 1];
        tree[writePos] = searchPos + n;
        if (level == level_max) {
            for (int i = 0; i < n; i++) {
                const float* s = s4 + 4 * (size_t)idx[i];
                pos[(i + searchPos) * 2] = s[0];
                pos[(i + searchPos) * 2 + 1] = s[1];
                mag[(i + searchPos) * 2] = s[2];
                mag[(i + searchPos) * 2 + 1] = s[3];
            }
            return;
        }
        level++;
        int nn = level % 2;
        size[nn] = size[nn] * .5f;
        std::vector<int> lu, rd;
        lu.reserve(n);
        rd.reserve(n);
        float check = thphi[nn] + size[nn];
        for (int i : idx) (s4[4 * (size_t)i + nn] < check ? lu : rd).push_back(i);
        std::vector<int>().swap(idx);
        float sizeNew[2] = {size[0], size[1]};
        float thphiNew[2] = {thphi[0], thphi[1]};
        build(lu, level, thphiNew, sizeNew, writePos * 2 + 1);
        float thphiNew2[2] = {thphi[0], thphi[1]};
        float sizeNew2[2] = {size[0], size[1]};
        thphiNew2[nn] += size[nn];
        build(rd, level, thphiNew2, sizeNew2, writePos * 2 + 2);
    }
};
}  // namespace

int star_tree(const float* stars4, int S, int tree_level, int32_t* tree, float* pos, float* mag) {
    if (S < 0 || tree_level < 1 || tree_level > 20 || (S > 0 && !stars4) || !tree || (S > 0 && (!pos || !mag)))
        return fail(BT_ERR_INVALID, "star_tree: bad arguments");
    TreeBuilder b{stars4, tree_level, tree, pos, mag};
    std::vector<int> idx(S);
    for (int i = 0; i < S; i++) idx[i] = i;
    float thphi[2] = {0, 0};
    float size[2] = {(float)kPI, (float)kPI2};
    b.build(idx, 0, thphi, size, 0);
    return BT_OK;
}

static int bin_host(float v, float d, int n) {
    float x = v / d;
    int b = (x == x) ? (x >= 2147483648.0f ? 2147483647 : (x <= -2147483648.0f ? (-2147483647 - 1) : (int)x)) : 0;
    return b < 0 ? 0 : (b > n - 1 ? n - 1 : b);
}

// Counting sort of the stars into kBinT x kBinP (theta, phi) cells; stable,
// so stars keep their tree order inside a cell.  The cell of a value is
// computed exactly as the device computes the cell range of a bounding box.
void star_bins(const float* pos, const float* mag, int S, int32_t* bin_start, float* bin_pos, float* bin_mag,
               int32_t* bin_idx) {
    const int NB = kBinT * kBinP;
    std::vector<int32_t> cell(S > 0 ? S : 1);
    for (int c = 0; c <= NB; c++) bin_start[c] = 0;
    for (int q = 0; q < S; q++) {
        int c = bin_host(pos[2 * q], kBinDT, kBinT) * kBinP + bin_host(pos[2 * q + 1], kBinDP, kBinP);
        cell[q] = c;
        bin_start[c + 1]++;
    }
    for (int c = 0; c < NB; c++) bin_start[c + 1] += bin_start[c];
    std::vector<int32_t> fill(bin_start, bin_start + NB);
    for (int q = 0; q < S; q++) {
        int d = fill[cell[q]]++;
        bin_pos[2 * d] = pos[2 * q];
        bin_pos[2 * d + 1] = pos[2 * q + 1];
        bin_mag[2 * d] = mag[2 * q];
        bin_mag[2 * d + 1] = mag[2 * q + 1];
        bin_idx[d] = q;
    }
}

void cam7_of(const PoseK& P, float* cam7) {
    // Camera::getParamArray, Camera.h:67-78
    cam7[0] = (float)P.speed;
    cam7[1] = (float)P.alpha;
    cam7[2] = (float)P.w;
    cam7[3] = (float)P.wbar;
    cam7[4] = (float)P.br;
    cam7[5] = (float)P.btheta;
    cam7[6] = (float)P.bphi;
}

}  // namespace bt

using namespace bt;

static void scene_release(bt_scene* s) {
    if (s->tree) (void)hipFree(s->tree);
    if (s->pos) (void)hipFree(s->pos);
    if (s->mag) (void)hipFree(s->mag);
    if (s->table) (void)hipFree(s->table);
    if (s->bin_start) (void)hipFree(s->bin_start);
    if (s->bin_pos) (void)hipFree(s->bin_pos);
    if (s->bin_mag) (void)hipFree(s->bin_mag);
    if (s->bin_idx) (void)hipFree(s->bin_idx);
    s->bin_idx = nullptr;
    s->tree = nullptr;
    s->bin_start = nullptr;
    s->pos = s->mag = s->table = s->bin_pos = s->bin_mag = nullptr;
}

// Uploads an already built star tree and builds the summed sky table on the device.
static int scene_upload(Context& c, bt_scene* s, const int32_t* tree, int tree_level, const float* pos,
                        const float* mag, int S, const uint8_t* sky, int rows, int cols) {
    if (rows <= 0 || cols <= 0 || !sky) return fail(BT_ERR_INVALID, "scene: sky image required");
    const size_t tsz = (size_t)((1 << (tree_level + 1)) - 1);
    s->tree_level = tree_level;
    s->S = S;
    s->rows = rows;
    s->cols = cols;
    BT_HIP(hipMalloc((void**)&s->tree, sizeof(int32_t) * tsz));
    BT_HIP(hipMalloc((void**)&s->pos, sizeof(float) * 2 * (S > 0 ? S : 1)));
    BT_HIP(hipMalloc((void**)&s->mag, sizeof(float) * 2 * (S > 0 ? S : 1)));
    BT_HIP(hipMalloc((void**)&s->table, sizeof(float) * 4 * (size_t)rows * cols));
    const size_t nb = (size_t)kBinT * kBinP + 1;
    BT_HIP(hipMalloc((void**)&s->bin_start, sizeof(int32_t) * nb));
    BT_HIP(hipMalloc((void**)&s->bin_pos, sizeof(float) * 2 * (S > 0 ? S : 1)));
    BT_HIP(hipMalloc((void**)&s->bin_mag, sizeof(float) * 2 * (S > 0 ? S : 1)));
    BT_HIP(hipMalloc((void**)&s->bin_idx, sizeof(int32_t) * (S > 0 ? S : 1)));
    s->bytes = sizeof(int32_t) * (tsz + nb + (size_t)S) + sizeof(float) * 8 * (size_t)S +
               sizeof(float) * 4 * (size_t)rows * cols;
    s->mag_tame = 1;
    for (int q = 0; q < S; q++)
        if (!(mag[2 * q] >= -10.f && mag[2 * q] <= 60.f)) s->mag_tame = 0;
    std::vector<int32_t> bstart(nb);
    std::vector<float> bpos(2 * (size_t)(S > 0 ? S : 1)), bmag(2 * (size_t)(S > 0 ? S : 1));
    std::vector<int32_t> bidx((size_t)(S > 0 ? S : 1));
    star_bins(pos, mag, S, bstart.data(), bpos.data(), bmag.data(), bidx.data());
    BT_HIP(hipMemcpy(s->tree, tree, sizeof(int32_t) * tsz, hipMemcpyHostToDevice));
    BT_HIP(hipMemcpy(s->bin_start, bstart.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice));
    if (S > 0) {
        BT_HIP(hipMemcpy(s->pos, pos, sizeof(float) * 2 * S, hipMemcpyHostToDevice));
        BT_HIP(hipMemcpy(s->mag, mag, sizeof(float) * 2 * S, hipMemcpyHostToDevice));
        BT_HIP(hipMemcpy(s->bin_pos, bpos.data(), sizeof(float) * 2 * S, hipMemcpyHostToDevice));
        BT_HIP(hipMemcpy(s->bin_mag, bmag.data(), sizeof(float) * 2 * S, hipMemcpyHostToDevice));
        BT_HIP(hipMemcpy(s->bin_idx, bidx.data(), sizeof(int32_t) * S, hipMemcpyHostToDevice));
    }
    void* dsky;
    int rc;
    if ((rc = c.buffer("scene_sky_upload", (size_t)3 * rows * cols, &dsky))) return rc;
    BT_HIP(hipMemcpyAsync(dsky, sky, (size_t)3 * rows * cols, hipMemcpyHostToDevice, c.stream));
    if ((rc = render_sum_table(&c, (const uint8_t*)dsky, rows, cols, s->table))) return rc;
    BT_HIP(hipStreamSynchronize(c.stream));
    return BT_OK;
}

extern "C" {

void bt_render_flags_default(bt_render_flags* f) {
    if (!f) return;
    f->star = 1;
    f->map = 1;
    f->lr = 1;
    f->hor = 0.f;
    f->ver = 0.f;
    f->offset = 0.f;
}

int bt_viewer_matrix(double view_angle, double off_up, double off_right, int width, int height, int sphere,
                     float* out) {
    return viewer_matrix(view_angle, off_up, off_right, width, height, sphere, out);
}

int bt_star_tree(const float* stars4, int star_count, int tree_level, int32_t* tree, float* star_pos,
                 float* star_mag) {
    return star_tree(stars4, star_count, tree_level, tree, star_pos, star_mag);
}

// float -> int with truncation, saturating out of range, NaN -> 0.
static int cvt_i32_host(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return -2147483647 - 1;
    return (int)x;
}

// StarProcessor::addLowLightStarsToImage (StarProcessor.h:108-158), on the
// host as in the reference (a one-time edit of the sky image before the
// summed table is made).  The splat buffer starts at zero (the reference's
// cv::Mat is uninitialised) and the colour index stops at the table's last
// entry, 390 (the reference's clamp, 1172, reads past its table).
int bt_sky_add_low_stars(uint8_t* sky_bgr, int rows, int cols, const float* stars4, int star_count,
                         float magnitude_cut) {
    static const int kRGB[1173] = {
#include "blackbody_rgb.inc"
    };
    if (!sky_bgr || rows < 1 || cols < 1 || star_count < 0 || (star_count > 0 && !stars4))
        return fail(BT_ERR_INVALID, "bt_sky_add_low_stars: bad arguments");
    const size_t npix = (size_t)rows * cols;
    std::vector<int> light(3 * npix, 0);
    constexpr int kStep = 2;
    constexpr float kReach = (kStep + .5f) * (kStep + .5f);
    constexpr double kPi = 3.141592653589793238;   // Const.h:2
    for (int s = 0; s < star_count; s++) {
        const float* st = stars4 + 4 * (size_t)s;
        if (!(st[2] > magnitude_cut)) continue;   // bright stars go to the star tree (StarProcessor.h:39-46)
        const float pt = (float)(st[0] / kPi * 1.f * rows);          // double arithmetic, stored as float
        const float pp = (float)(st[1] / (kPi * 2.) * 1.f * cols);
        const int row0 = (int)pt, col0 = (int)pp;
        const float bv = st[3];
        const float tidx = 46.f * ((1.f / ((0.92f * bv) + 1.7f)) + (1.f / ((0.92f * bv) + 0.62f))) - 10.f;
        const int k = std::max(0, std::min((int)std::floor(tidx), 390));
        const float cb = (float)(kRGB[3 * k + 2] * kRGB[3 * k + 2]);
        const float cg = (float)(kRGB[3 * k + 1] * kRGB[3 * k + 1]);
        const float cr = (float)(kRGB[3 * k] * kRGB[3 * k]);
        const int u_lo = std::max(0, kStep - row0);
        const int u_hi = row0 + kStep >= rows ? 2 * kStep - (row0 + kStep - rows + 1) : 2 * kStep;
        const float fy = pt - row0, fx = pp - col0;
        for (int u = u_lo; u <= u_hi; u++)
            for (int v = 0; v <= 2 * kStep; v++) {
                const float dy = u - kStep + .5f - fy, dx = v - kStep + .5f - fx;
                const float d = dy * dy + dx * dx;
                if (d > kReach) continue;
                // 400 * 10^(-0.4 (m - 2.5 log10 g(d))), g(d) = exp(-2 d) / pi
                const float g = (float)std::exp((double)(-2.f * d)) * 0.318310f;   // StarProcessor.h:96-98
                const float e = -.4f * (st[2] - 2.5f * (float)std::log10((double)g));
                const float bright = 400.f * (float)std::pow(10.0, (double)e);
                int* px = &light[3 * ((size_t)(row0 - kStep + u) * cols + (size_t)((col0 - kStep + v + cols) % cols))];
                px[0] = cvt_i32_host(bright * cb);
                px[1] = cvt_i32_host(bright * cg);
                px[2] = cvt_i32_host(bright * cr);
            }
    }
    for (size_t q = 0; q < 3 * npix; q++) {
        const int o = sky_bgr[q];
        sky_bgr[q] = (uint8_t)std::min(255, (int)std::sqrt((double)(light[q] + o * o)));
    }
    return BT_OK;
}

int bt_scene_create(bt_context* h, const float* stars4, int star_count, int tree_level, const uint8_t* sky_bgr,
                    int sky_rows, int sky_cols, bt_scene** out) {
    if (!h || !out) return fail(BT_ERR_INVALID, "bt_scene_create: null argument");
    *out = nullptr;
    // the device star search lists at most 2^(L+1) - 1 < 2048 nodes
    if (tree_level < 1 || tree_level > 10 || star_count < 0)
        return fail(BT_ERR_INVALID, "bt_scene_create: tree_level in [1, 10], star_count >= 0");
    Context& c = h->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    std::vector<int32_t> tree((size_t)((1 << (tree_level + 1)) - 1));
    std::vector<float> pos(2 * (size_t)star_count), mag(2 * (size_t)star_count);
    int rc = star_tree(stars4, star_count, tree_level, tree.data(), pos.data(), mag.data());
    if (rc) return rc;
    bt_scene* s = new bt_scene();
    s->ctx = &c;
    rc = scene_upload(c, s, tree.data(), tree_level, pos.data(), mag.data(), star_count, sky_bgr, sky_rows, sky_cols);
    if (rc) {
        (void)hipStreamSynchronize(c.stream);
        scene_release(s);
        delete s;
        return rc;
    }
    *out = s;
    return BT_OK;
}

int bt_scene_create_tree(bt_context* h, const int32_t* tree, int tree_level, const float* pos, const float* mag,
                         int star_count, const uint8_t* sky_bgr, int sky_rows, int sky_cols, bt_scene** out) {
    if (!h || !out || !tree || (star_count > 0 && (!pos || !mag)))
        return fail(BT_ERR_INVALID, "bt_scene_create_tree: null argument");
    *out = nullptr;
    if (tree_level < 1 || tree_level > 10 || star_count < 0)
        return fail(BT_ERR_INVALID, "bt_scene_create_tree: tree_level in [1, 10], star_count >= 0");
    // the tree's last node holds the star count (cumulative end indices)
    if (tree[(1 << (tree_level + 1)) - 2] != star_count)
        return fail(BT_ERR_INVALID, "bt_scene_create_tree: tree does not index star_count stars");
    Context& c = h->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    bt_scene* s = new bt_scene();
    s->ctx = &c;
    int rc = scene_upload(c, s, tree, tree_level, pos, mag, star_count, sky_bgr, sky_rows, sky_cols);
    if (rc) {
        (void)hipStreamSynchronize(c.stream);
        scene_release(s);
        delete s;
        return rc;
    }
    *out = s;
    return BT_OK;
}

int bt_scene_export_table(const bt_scene* s, float* table) {
    if (!s || !table) return fail(BT_ERR_INVALID, "null argument");
    Context& c = *s->ctx;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    BT_HIP(hipMemcpyAsync(table, s->table, sizeof(float) * 4 * (size_t)s->rows * s->cols, hipMemcpyDeviceToHost,
                          c.stream));
    BT_HIP(hipStreamSynchronize(c.stream));
    return BT_OK;
}

void bt_scene_free(bt_scene* s) {
    if (!s) return;
    if (s->ctx) {
        std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
        (void)hipSetDevice(s->ctx->device);
        (void)hipStreamSynchronize(s->ctx->stream);
        scene_release(s);
    }
    delete s;
}

int bt_view_create(bt_context* h, double view_angle, double off_up, double off_right, int width, int height,
                   int sphere, bt_view** out) {
    if (!h || !out) return fail(BT_ERR_INVALID, "bt_view_create: null argument");
    *out = nullptr;
    if (width <= 0 || height <= 0 || width > 16384 || height > 16384)
        return fail(BT_ERR_INVALID, "bt_view_create: size out of range");
    Context& c = h->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    const size_t n = (size_t)(width + 1) * (height + 1);
    std::vector<float> host(2 * n);
    int rc = viewer_matrix(view_angle, off_up, off_right, width, height, sphere, host.data());
    if (rc) return rc;
    bt_view* v = new bt_view();
    v->ctx = &c;
    v->W = width;
    v->H = height;
    v->bytes = sizeof(float) * (2 * n + (size_t)width * height);
    if (hipMalloc((void**)&v->view, v->bytes) != hipSuccess) {
        delete v;
        return fail(BT_ERR_NOMEM, "bt_view_create: hipMalloc failed");
    }
    v->pix_area = v->view + 2 * n;
    hipError_t e = hipMemcpyAsync(v->view, host.data(), sizeof(float) * 2 * n, hipMemcpyHostToDevice, c.stream);
    if (e == hipSuccess) {
        rc = view_pixel_areas(&c, v->view, width, height, v->pix_area);
        e = rc ? hipErrorUnknown : hipStreamSynchronize(c.stream);
    }
    if (e != hipSuccess) {
        (void)hipFree(v->view);
        delete v;
        return rc ? rc : fail(BT_ERR_HIP, std::string("bt_view_create: ") + hipGetErrorString(e));
    }
    *out = v;
    return BT_OK;
}

void bt_view_free(bt_view* v) {
    if (!v) return;
    if (v->ctx) {
        std::lock_guard<std::recursive_mutex> lk(v->ctx->mu);
        (void)hipSetDevice(v->ctx->device);
        (void)hipStreamSynchronize(v->ctx->stream);
        if (v->view) (void)hipFree(v->view);
    }
    delete v;
}

}  // extern "C"

namespace bt {
// The FP32 dense grid the render kernels read, converted from the FP64
// device grid once per grid (on c's stream, which is synchronised before
// the grid is shared) and kept with it.
int grid_f32_of(Context& c, const bt_grid* g, GridF* out) {
    std::lock_guard<std::mutex> lk(g->gf_mu);
    if (!g->gridf) {
        int sym = (float)g->M / (float)g->N > 3 ? 1 : 0;
        int GN1 = sym ? 2 * g->N - 1 : g->N;
        float* p = nullptr;
        if (hipMalloc((void**)&p, sizeof(float) * 2 * (size_t)GN1 * g->M) != hipSuccess)
            return fail(BT_ERR_NOMEM, "hipMalloc failed for the FP32 render grid");
        int gn1 = 0;
        int rc = render_grid_from_cel(&c, g->k.cel, g->N, g->M, p, &gn1);
        if (rc == BT_OK && hipStreamSynchronize(c.stream) != hipSuccess) rc = fail(BT_ERR_HIP, "grid conversion failed");
        if (rc) {
            (void)hipFree(p);
            return rc;
        }
        g->gridf = p;
        g->gridf_rows = gn1;
    }
    *out = GridF{g->gridf, g->gridf_rows, g->M, g->max_level};
    return BT_OK;
}
}  // namespace bt

extern "C" {

}  // extern "C"

// Trail state of a frame sequence: per star, the pixels of its images in the
// last two frames, and the frame's trail, gradient and image buffers.
struct bt_trails {
    bt::Context* ctx = nullptr;
    int S = 0, W = 0, H = 0, fs = 0;
    int64_t frame = 0;
    void* block = nullptr;
    uint8_t* sprite = nullptr;
    bt::TrailK k{};
};

namespace {

// One frame from a device-resident grid, or (g1 non-null) blended between
// grids g and g1 at alpha with cam = 14 floats (the two poses' parameter
// arrays), as the G > 1 frame of the reference.
int frame_impl(bt_context* h, const bt_grid* g, const bt_grid* g1, float alpha, const bt_scene* s, const bt_view* v,
               const float* cam_in, const bt_render_flags* flags, int step, uint8_t* out_bgra, void* d_out_bgra,
               float* thphi, int32_t* gap, uint8_t* bh, float* area, float* star, float* border_out,
               bt_trails* tr = nullptr, float* trail_out = nullptr) {
    if (!h || !g || !s || !v) return fail(BT_ERR_INVALID, "bt_render_frame: null handle");
    if (tr && (tr->ctx != &h->impl || tr->W != v->W || tr->H != v->H || tr->S != s->S))
        return fail(BT_ERR_INVALID, "bt_render_frame_trails: the trail state was made for another context, size or star set");
    if (!out_bgra && !d_out_bgra) return fail(BT_ERR_INVALID, "bt_render_frame: no output buffer");
    Context& c = h->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    bt_render_flags f;
    bt_render_flags_default(&f);
    if (flags) f = *flags;
    float cam[7];
    if (!g1) {
        if (cam_in)
            std::memcpy(cam, cam_in, sizeof(cam));
        else
            cam7_of(g->pose, cam);
    } else {
        // camUpdate (MakeImageKernel.cu:386-389)
        float c0[7], c1[7];
        if (cam_in) {
            std::memcpy(c0, cam_in, sizeof(c0));
            std::memcpy(c1, cam_in + 7, sizeof(c1));
        } else {
            cam7_of(g->pose, c0);
            cam7_of(g1->pose, c1);
        }
        for (int q = 0; q < 7; q++) cam[q] = (1.f - alpha) * c0[q] + alpha * c1[q];
    }
    GridF gf;
    int rc;
    if ((rc = grid_f32_of(c, g, &gf))) return rc;
    Blend blend{};
    if (g1) {
        if (g1->N != g->N || g1->M != g->M || g1->max_level != g->max_level)
            return fail(BT_ERR_INVALID, "bt_render_frame_blend: the grids differ in size or level");
        if ((rc = grid_f32_of(c, g1, &blend.b))) return rc;
        void* p;
        if ((rc = c.buffer("r_blend_border", sizeof(float) * 2 * (2 + 2 * kAngleNum), &p))) return rc;
        blend.border = (float*)p;
        blend.alpha = alpha;
    }
    const size_t NP = (size_t)v->W * v->H, NC = (size_t)(v->W + 1) * (v->H + 1);
    void* dout = d_out_bgra;
    if (!dout && (rc = c.buffer("r_out", 4 * NP, &dout))) return rc;
    FrameDebug dbg{nullptr, nullptr, nullptr, nullptr, nullptr};
    bool want = thphi || gap || bh || area || star;
    if (want) {
        void* p;
        size_t off = 0;
        size_t tot = sizeof(float) * 2 * NC + sizeof(int) * NC + NP + sizeof(float) * NP + sizeof(float) * 3 * NP + 64;
        if ((rc = c.buffer("r_dbg", tot, &p))) return rc;
        char* base = (char*)p;
        dbg.thphi = (float*)(base + off);
        off += sizeof(float) * 2 * NC;
        dbg.gap = (int*)(base + off);
        off += sizeof(int) * NC;
        dbg.area = (float*)(base + off);
        off += sizeof(float) * NP;
        dbg.star = (float*)(base + off);
        off += sizeof(float) * 3 * NP;
        dbg.bh = (uint8_t*)(base + off);
    }
    TrailFrame tf{};
    if (tr) {
        tf.k = tr->k;
        tf.k.cur = (int)(tr->frame % 2);
        tf.first = tr->frame == 0;
        tf.W = tr->W;
        tf.H = tr->H;
        tf.sprite = tr->sprite;
        tf.fs = tr->fs;
    }
    if ((rc = render_frame(&c, gf, s->k(), ViewK{v->view, v->W, v->H, v->pix_area}, cam, f, step, (uint8_t*)dout,
                           want ? &dbg : nullptr, g1 ? &blend : nullptr, tr ? &tf : nullptr)))
        return rc;
    if (tr) {
        // the reference's frame counter advances with every displayed frame
        tr->frame++;
        if (trail_out && f.star)
            BT_HIP(hipMemcpyAsync(trail_out, tr->k.trail, sizeof(float) * 3 * NP, hipMemcpyDeviceToHost, c.stream));
        else if (trail_out)
            std::memset(trail_out, 0, sizeof(float) * 3 * NP);   // no star pass: no trails this frame
    }
    if (border_out && g1)
        BT_HIP(hipMemcpyAsync(border_out, blend.border, sizeof(float) * 2 * (2 + 2 * kAngleNum), hipMemcpyDeviceToHost,
                              c.stream));
    if (out_bgra) BT_HIP(hipMemcpyAsync(out_bgra, dout, 4 * NP, hipMemcpyDeviceToHost, c.stream));
    if (thphi) BT_HIP(hipMemcpyAsync(thphi, dbg.thphi, sizeof(float) * 2 * NC, hipMemcpyDeviceToHost, c.stream));
    if (gap) BT_HIP(hipMemcpyAsync(gap, dbg.gap, sizeof(int) * NC, hipMemcpyDeviceToHost, c.stream));
    if (bh) BT_HIP(hipMemcpyAsync(bh, dbg.bh, NP, hipMemcpyDeviceToHost, c.stream));
    if (area) BT_HIP(hipMemcpyAsync(area, dbg.area, sizeof(float) * NP, hipMemcpyDeviceToHost, c.stream));
    if (star) BT_HIP(hipMemcpyAsync(star, dbg.star, sizeof(float) * 3 * NP, hipMemcpyDeviceToHost, c.stream));
    if (out_bgra || want || border_out || trail_out) BT_HIP(hipStreamSynchronize(c.stream));
    return BT_OK;
}

}  // namespace

extern "C" {

int bt_render_frame_stages(bt_context* h, const bt_grid* g, const bt_scene* s, const bt_view* v, const float* cam7,
                           const bt_render_flags* flags, int step, uint8_t* out_bgra, void* d_out_bgra,
                           float* thphi, int32_t* gap, uint8_t* bh, float* area, float* star) {
    return frame_impl(h, g, nullptr, 0.f, s, v, cam7, flags, step, out_bgra, d_out_bgra, thphi, gap, bh, area, star,
                      nullptr);
}

int bt_render_frame_blend(bt_context* h, const bt_grid* g0, const bt_grid* g1, float alpha, const bt_scene* s,
                          const bt_view* v, const float* cam14, const bt_render_flags* flags, int step,
                          uint8_t* out_bgra, void* d_out_bgra, float* thphi, float* border) {
    if (!g1) return fail(BT_ERR_INVALID, "bt_render_frame_blend: second grid is null");
    if (!(alpha >= 0.f && alpha <= 1.f)) return fail(BT_ERR_INVALID, "bt_render_frame_blend: alpha must be in [0, 1]");
    return frame_impl(h, g0, g1, alpha, s, v, cam14, flags, step, out_bgra, d_out_bgra, thphi, nullptr, nullptr,
                      nullptr, nullptr, border);
}

size_t bt_frame_bytes(int width, int height, int step) {
    if (width <= 0 || height <= 0 || step < 1) return 0;
    return bt::frame_bytes(width, height, step);
}

int bt_render_frame(bt_context* h, const bt_grid* g, const bt_scene* s, const bt_view* v, const float* cam7,
                    const bt_render_flags* flags, int step, uint8_t* out_bgra, void* d_out_bgra) {
    return bt_render_frame_stages(h, g, s, v, cam7, flags, step, out_bgra, d_out_bgra, nullptr, nullptr, nullptr,
                                  nullptr, nullptr);
}

int bt_trails_create(bt_context* h, const bt_scene* s, int width, int height, const uint8_t* diffraction_rgb,
                     bt_trails** out) {
    if (!h || !s || !out || width <= 0 || height <= 0) return fail(BT_ERR_INVALID, "bt_trails_create: bad arguments");
    const int fs = width / 16;   // the reference's sprite size (MakeImageKernel.cu:1284)
    if (diffraction_rgb && fs < 1) return fail(BT_ERR_INVALID, "bt_trails_create: diffraction needs width >= 16");
    Context& c = h->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    const size_t S = (size_t)s->S, NP = (size_t)width * height, NC = (size_t)(width + 1) * (height + 1);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_cache = up(sizeof(int2) * 2 * S * kTrailNum), b_count = up(sizeof(int) * 2 * S),
                 b_nhits = up(sizeof(unsigned) * S), b_pix = up(sizeof(uint32_t) * S * kTrailHits),
                 b_data = up(sizeof(float) * 3 * S * kTrailHits), b_grad = up(sizeof(float) * 2 * NC),
                 b_trail = up(sizeof(float) * 3 * NP), b_stats = 256,
                 b_sprite = diffraction_rgb ? up((size_t)3 * fs * fs) : 0;
    const size_t total = b_cache + b_count + b_nhits + b_pix + b_data + b_grad + b_trail + b_stats + b_sprite;
    char* base = nullptr;
    if (hipMalloc((void**)&base, total) != hipSuccess) return fail(BT_ERR_NOMEM, "bt_trails_create: hipMalloc failed");
    bt_trails* t = new bt_trails;
    t->ctx = &c;
    t->S = (int)S;
    t->W = width;
    t->H = height;
    t->fs = fs;
    t->block = base;
    size_t off = 0;
    t->k.S = (int)S;
    t->k.cache = (int2*)(base + off);
    off += b_cache;
    t->k.count = (int*)(base + off);
    off += b_count;
    t->k.nhits = (unsigned*)(base + off);
    off += b_nhits;
    t->k.hit_pix = (uint32_t*)(base + off);
    off += b_pix;
    t->k.hit_data = (float*)(base + off);
    off += b_data;
    t->k.grad = (float*)(base + off);
    off += b_grad;
    t->k.trail = (float*)(base + off);
    off += b_trail;
    t->k.stats = (unsigned*)(base + off);
    off += b_stats;
    if (diffraction_rgb) {
        t->sprite = (uint8_t*)(base + off);
        hipError_t e = hipMemcpyAsync(t->sprite, diffraction_rgb, (size_t)3 * fs * fs, hipMemcpyHostToDevice, c.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        if (e != hipSuccess) {
            (void)hipFree(base);
            delete t;
            return fail(BT_ERR_HIP, std::string("bt_trails_create: ") + hipGetErrorString(e));
        }
    }
    *out = t;
    return BT_OK;
}

void bt_trails_reset(bt_trails* t) {
    if (t) t->frame = 0;
}

int bt_trails_stats(bt_trails* t, int64_t* frames, int64_t* dropped) {
    if (!t) return fail(BT_ERR_INVALID, "bt_trails_stats: null handle");
    std::lock_guard<std::recursive_mutex> lk(t->ctx->mu);
    BT_HIP(hipSetDevice(t->ctx->device));
    unsigned st[2] = {0, 0};
    BT_HIP(hipMemcpyAsync(st, t->k.stats, sizeof(st), hipMemcpyDeviceToHost, t->ctx->stream));
    BT_HIP(hipStreamSynchronize(t->ctx->stream));
    if (frames) *frames = t->frame;
    if (dropped) *dropped = t->frame > 0 ? st[1] : 0;
    return BT_OK;
}

void bt_trails_free(bt_trails* t) {
    if (!t) return;
    if (t->ctx) {
        std::lock_guard<std::recursive_mutex> lk(t->ctx->mu);
        (void)hipSetDevice(t->ctx->device);
        (void)hipStreamSynchronize(t->ctx->stream);
        if (t->block) (void)hipFree(t->block);
    }
    delete t;
}

int bt_render_frame_trails(bt_context* h, const bt_grid* g, const bt_scene* s, const bt_view* v, const float* cam7,
                           const bt_render_flags* flags, int step, bt_trails* trails, uint8_t* out_bgra,
                           void* d_out_bgra, float* star, float* trail) {
    if (!trails) return fail(BT_ERR_INVALID, "bt_render_frame_trails: null trail state");
    return frame_impl(h, g, nullptr, 0.f, s, v, cam7, flags, step, out_bgra, d_out_bgra, nullptr, nullptr, nullptr,
                      nullptr, star, nullptr, trails, trail);
}

int bt_make_image(bt_context* h, const float* stars, const int32_t* star_tree_in, int star_size,
                  const float* cam_param, const float* magnitude, int tree_level, int M, int N, int step,
                  const uint8_t* sky_bgr, int sky_rows, int sky_cols, int G, float grid_start, float grid_step,
                  const float* hit, const float* viewthing, float view_angle, int GM, int GN, int gridlvl,
                  const int32_t* offset_tables, const float* hash_tables, const int32_t* hash_pos_tag,
                  const int32_t* table_sizes, int otsize, int htsize, const bt_render_flags* flags,
                  uint8_t* out_bgra) {
    (void)grid_start;
    (void)grid_step;
    (void)hit;
    (void)view_angle;
    if (!h) return fail(BT_ERR_INVALID, "bt_make_image: context is null");
    if (G != 1) return fail(BT_ERR_INVALID, "bt_make_image: only G == 1 (one grid) is supported");
    if (M <= 0 || N <= 0 || GM <= 0 || GN <= 0 || tree_level < 1 || tree_level > 10 || star_size < 0)
        return fail(BT_ERR_INVALID, "bt_make_image: bad sizes");
    if (!cam_param || !viewthing || !offset_tables || !hash_tables || !hash_pos_tag || !table_sizes ||
        !out_bgra || !star_tree_in || (star_size > 0 && (!stars || !magnitude)))
        return fail(BT_ERR_INVALID, "bt_make_image: null input");
    const int hw = table_sizes[0], ow = table_sizes[1];
    if (hw <= 0 || ow <= 0 || (size_t)hw * hw > (size_t)htsize || (size_t)ow * ow > (size_t)otsize)
        return fail(BT_ERR_INVALID, "bt_make_image: table sizes inconsistent with htsize/otsize");
    Context& c = h->impl;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    BT_HIP(hipSetDevice(c.device));
    int rc;
    // scene: the caller's star tree + summed table of the sky
    bt_scene scene;
    scene.ctx = &c;
    if ((rc = scene_upload(c, &scene, star_tree_in, tree_level, stars, magnitude, star_size, sky_bgr, sky_rows,
                           sky_cols))) {
        (void)hipStreamSynchronize(c.stream);
        scene_release(&scene);
        return rc;
    }
    // PSH tables -> dense grid (makeGrid)
    void *dht, *dot, *dtag, *dgrid, *dview, *dout;
    int sym = (float)GM / (float)GN > 3 ? 1 : 0;
    int GN1 = sym ? 2 * GN - 1 : GN;
    const size_t NP = (size_t)M * N, NC = (size_t)(M + 1) * (N + 1);
    if ((rc = c.buffer("mi_ht", sizeof(float) * 2 * (size_t)htsize, &dht)) ||
        (rc = c.buffer("mi_ot", sizeof(int32_t) * 2 * (size_t)otsize, &dot)) ||
        (rc = c.buffer("mi_tag", sizeof(int32_t) * 2 * (size_t)htsize, &dtag)) ||
        (rc = c.buffer("mi_grid", sizeof(float) * 2 * (size_t)GN1 * GM, &dgrid)) ||
        (rc = c.buffer("mi_view", sizeof(float) * 2 * NC, &dview)) || (rc = c.buffer("mi_out", 4 * NP, &dout))) {
        (void)hipStreamSynchronize(c.stream);
        scene_release(&scene);
        return rc;
    }
    BT_HIP(hipMemcpyAsync(dht, hash_tables, sizeof(float) * 2 * (size_t)hw * hw, hipMemcpyHostToDevice, c.stream));
    BT_HIP(hipMemcpyAsync(dot, offset_tables, sizeof(int32_t) * 2 * (size_t)ow * ow, hipMemcpyHostToDevice, c.stream));
    BT_HIP(hipMemcpyAsync(dtag, hash_pos_tag, sizeof(int32_t) * 2 * (size_t)hw * hw, hipMemcpyHostToDevice, c.stream));
    BT_HIP(hipMemcpyAsync(dview, viewthing, sizeof(float) * 2 * NC, hipMemcpyHostToDevice, c.stream));
    int gn1 = 0;
    if ((rc = render_grid_from_psh(&c, (const float*)dht, (const int32_t*)dot, (const int32_t*)dtag, hw, ow, GN, GM,
                                   (float*)dgrid, &gn1)))
        goto done;
    {
        bt_render_flags f;
        bt_render_flags_default(&f);
        if (flags) f = *flags;
        rc = render_frame(&c, GridF{(const float*)dgrid, gn1, GM, gridlvl}, scene.k(), ViewK{(const float*)dview, M, N, nullptr},
                          cam_param, f, step, (uint8_t*)dout, nullptr);
        if (rc) goto done;
    }
    {
        hipError_t e = hipMemcpyAsync(out_bgra, dout, 4 * NP, hipMemcpyDeviceToHost, c.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        if (e != hipSuccess) rc = fail(BT_ERR_HIP, std::string("bt_make_image: ") + hipGetErrorString(e));
    }
done:
    (void)hipStreamSynchronize(c.stream);
    scene_release(&scene);
    return rc;
}

}  // extern "C"
