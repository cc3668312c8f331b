// render.hip -- the per-pixel render pass (Path B): the drop-in for the
// G == 1 frame of cudaPrep (MakeImageKernel.cu:1402-1529).
//
// Stages (one kernel each, all on the context stream):
//   pix_interp   per pixel corner: quadtree descent of the dense grid and
//                bicubic Hermite interpolation of the celestial angle
//                (pixInterpolation, interpolatePix, kernel.cuh:736-968)
//   bh_area      shadow flag and celestial solid angle per pixel
//                (findBlackPixels + findArea)
//   smooth_h/v   box-filtered solid angle (smoothAreaH/V)
//   star_map     per pixel: star-tree search, point-in-quad, sub-pixel
//                position, blackbody colour and the 3x3 Gaussian taps
//                (distortStarMap); taps are stored tap-major so the gather
//                of star_sum reads coalesced rows
//   star_sum     gather of the neighbours' taps + makePix
//   env_map      summed-area sky sampling, lensing and redshift
//                (distortEnvironmentMap)
//   composite    addStarsAndBackground
// Pixel (i, j) is processed by thread x = j, y = i so that a wavefront
// (16 x 4 pixels of a 16 x 16 block) reads consecutive row addresses.
// Compiled with -ffp-contract=off: results equal the CPU oracle bit for bit.
#include <cmath>

#include "bt_render.h"
#include "render_dev.h"

namespace bt {
namespace {

using namespace rdr;

__constant__ int kTempRGB[1173] = {
#include "blackbody_rgb.inc"
};

struct Cam7 {
    float v[7];
    const float* colours;    // star_colours table, 391 x 6 floats
    const float* pix_area;   // per pixel: camera-sky solid angle of its view quad
};

__device__ __forceinline__ F2 gload(const float* grid, int GN, int GM, int i, int j) {
    if (i < 0 || i >= GN || j < 0 || j >= GM) return F2{-2.f, -2.f};
    const float2 v = reinterpret_cast<const float2*>(grid)[(size_t)i * GM + j];
    return F2{v.x, v.y};
}

struct GridV {
    const float* g;
    int GN, GM;
    __device__ __forceinline__ F2 at(int i, int j) const { return gload(g, GN, GM, i, j); }
};

// kernel.cuh:774-788
__device__ F2 interpolate_linear(float percDown, float percRight, const F2* c) {
    float phi[4] = {c[0].y, c[1].y, c[2].y, c[3].y};
    float theta[4] = {c[0].x, c[1].x, c[2].x, c[3].x};
    pi_check(phi, 0.2f);
    float leftT = theta[0] + percDown * (theta[2] - theta[0]);
    float leftP = phi[0] + percDown * (phi[2] - phi[0]);
    float rightT = theta[1] + percDown * (theta[3] - theta[1]);
    float rightP = phi[1] + percDown * (phi[3] - phi[1]);
    float upT = theta[0] + percRight * (theta[1] - theta[0]);
    float upP = phi[0] + percRight * (phi[1] - phi[0]);
    float downT = theta[2] + percRight * (theta[3] - theta[2]);
    float downP = phi[2] + percRight * (phi[3] - phi[2]);
    return intersection(upT, upP, downT, downP, leftT, leftP, rightT, rightP);
}

// The reference's findPoint may recurse once into interpolateHermite
// (count 0 -> 1) and never deeper; the two levels are written out here.
__device__ F2 interpolate_hermite_l1(const GridV& g, int i, int j, int gap, float percDown,
                                     float percRight, F2* c);

// kernel.cuh:817-889 for recursion depth `count`.
template <int COUNT>
__device__ F2 find_point(const GridV& g, int i, int j, int offver, int offhor, int gap) {
    const int GM = g.GM, GN = g.GN;
    F2 gridpt = g.at(i, j);
    if (gridpt.x == -2 && gridpt.y == -2) {
        int j2 = (j + offhor * gap + GM) % GM;
        int i2 = i + offver * gap;
        F2 ij2 = g.at(i2, j2);
        if (ij2.x == -1 && ij2.y == -1) return F2{-1, -1};
        if (ij2.x != -2 && ij2.y != -2) {
            int j0 = (j - offhor * gap + GM) % GM;
            int i0 = (i - offver * gap);
            F2 ij0 = g.at(i0, j0);
            if (ij0.x < 0) return F2{-1, -1};
            int jprev = (j - 3 * offhor * gap + GM) % GM;
            int jnext = (j + 3 * offhor * gap + GM) % GM;
            int iprev = i - offver * 3 * gap;
            int inext = i + offver * 3 * gap;
            if (offver != 0) {
                if (i2 == 0) {
                    jnext = (j0 + GM / 2) % GM;
                    inext = i0;
                } else if (i0 == GN - 1) {
                    jprev = (j0 + GM / 2) % GM;
                    iprev = i2;
                } else if (i2 == GN - 1) {
                    inext = i0;
                    jnext = (j0 + GM / 2) % GM;
                }
            }
            F2 ijprev = g.at(iprev, jprev);
            F2 ijnext = g.at(inext, jnext);
            if (ijprev.x > -2 && ijnext.x > -2) {
                F2 pt[4] = {ijprev, ij0, ij2, ijnext};
                if (pt[0].x != -1 && pt[3].x != -1) {
                    pi_check_tot(pt, 0.2f, 4);
                    return hermite(0.5f, pt[0], pt[1], pt[2], pt[3]);
                }
            }
            F2 pt[2] = {ij2, ij0};
            pi_check_tot(pt, 0.2f, 2);
            return F2{.5f * (pt[0].x + pt[1].x), .5f * (pt[0].y + pt[1].y)};
        }
        if (COUNT > 0) return F2{-1.f, -1.f};
        int j0 = (j + gap) % GM;
        int j1 = (j - gap + GM) % GM;
        if (i - gap < 0) return F2{-1, -1};
        F2 cc[12];
        cc[0] = g.at(i + gap, j0);
        cc[1] = g.at(i - gap, j0);
        cc[2] = g.at(i - gap, j1);
        cc[3] = g.at(i + gap, j1);
        for (int q = 0; q < 4; q++)
            if (cc[q].x == -1 || cc[q].x == -2) return F2{-1, -1};
        return interpolate_hermite_l1(g, i - gap, j1, 2 * gap, .5f, .5f, cc);
    }
    return gridpt;
}

// kernel.cuh:891-938 for recursion depth COUNT.
template <int COUNT>
__device__ F2 interpolate_hermite(const GridV& g, int i, int j, int gap, float percDown,
                                  float percRight, F2* c) {
    const int GM = g.GM, GN = g.GN;
    int k = i + gap;
    int l = (j + gap) % GM;
    int imin1 = i - gap;
    int kplus1 = k + gap;
    int jmin1 = (j - gap + GM) % GM;
    int lplus1 = (l + gap) % GM;
    int jx = j, jy = j, lx = l, ly = l;
    if (i == 0) {
        jx = (j + GM / 2) % GM;
        lx = (jx + gap) % GM;
        imin1 = k;
    } else if (k == GN - 1) {
        jy = (j + GM / 2) % GM;
        ly = (jy + gap) % GM;
        kplus1 = i;
    }
    c[4] = find_point<COUNT>(g, i, jmin1, 0, -1, gap);
    c[5] = find_point<COUNT>(g, i, lplus1, 0, 1, gap);
    c[6] = find_point<COUNT>(g, k, jmin1, 0, -1, gap);
    c[7] = find_point<COUNT>(g, k, lplus1, 0, 1, gap);
    c[8] = find_point<COUNT>(g, imin1, jx, -1, 0, gap);
    c[9] = find_point<COUNT>(g, imin1, lx, -1, 0, gap);
    c[10] = find_point<COUNT>(g, kplus1, jy, 1, 0, gap);
    c[11] = find_point<COUNT>(g, kplus1, ly, 1, 0, gap);
    for (int q = 4; q < 12; q++)
        if (c[q].x == -1) return interpolate_linear(percDown, percRight, c);
    pi_check_tot(c, 0.2f, 12);
    F2 up = hermite(percRight, c[4], c[0], c[1], c[5]);
    F2 down = hermite(percRight, c[6], c[2], c[3], c[7]);
    F2 upup = F2{c[8].x + (c[9].x - c[8].x) * percRight, c[8].y + (c[9].y - c[8].y) * percRight};
    F2 downdown = F2{c[10].x + (c[11].x - c[10].x) * percRight, c[10].y + (c[11].y - c[10].y) * percRight};
    return hermite(percDown, upup, up, down, downdown);
}

__device__ F2 interpolate_hermite_l1(const GridV& g, int i, int j, int gap, float percDown,
                                     float percRight, F2* c) {
    return interpolate_hermite<1>(g, i, j, gap, percDown, percRight, c);
}

// MakeImageKernel.cu:328-351 with findBlock (kernel.cuh:736-752) and
// interpolateSpline (kernel.cuh:940-968).
__device__ F2 interpolate_pix(const GridV& g, float theta, float phi, int gridlvl, int& gapsave) {
    const int GM = g.GM;
    int half = (phi < BT_FPI) ? 0 : 1;
    int a = 0;
    int b = half * GM / 2;
    int gap = GM / 2;
    for (int s = 0; s < gridlvl + 1; s++) {
        int ngap = gap / 2;
        int k = a + ngap;
        int l = b + ngap;
        if (gap <= 1 || g.at(k, l).x == -2.f) break;
        float thHalf = BT_FPI2 * k / (1.f * GM);
        float phHalf = BT_FPI2 * l / (1.f * GM);
        if (thHalf <= theta) a = k;
        if (phHalf <= phi) b = l;
        gap = ngap;
    }
    gapsave = gap;
    int k = a + gap;
    int l = b + gap;
    float factor = BT_FPI2 / (1.f * GM);
    float thetaUp = factor * a, phiLeft = factor * b, thetaDown = factor * k, phiRight = factor * l;
    l = l % GM;
    F2 c[12];
    c[0] = g.at(a, b);
    c[1] = g.at(a, l);
    c[2] = g.at(k, b);
    c[3] = g.at(k, l);
    F2 r;
    bool done = false;
    if (thetaUp == theta) {
        if (phiLeft == phi) {
            r = c[0];
            done = true;
        } else if (phiRight == phi) {
            r = c[1];
            done = true;
        } else if (a == 0.f) {
            r = c[0];
            done = true;
        }
    } else if (thetaDown == theta) {
        if (phiLeft == phi) {
            r = c[2];
            done = true;
        } else if (phiRight == phi) {
            r = c[3];
            done = true;
        }
    }
    if (!done) {
        bool captured = false;
        for (int q = 0; q < 4; q++)
            if (c[q].x == -1 && c[q].y == -1) captured = true;
        if (captured) {
            r = F2{-1.f, -1.f};
        } else {
            float percDown = (theta - thetaUp) / (thetaDown - thetaUp);
            float percRight = (phi - phiLeft) / (phiRight - phiLeft);
            r = interpolate_hermite<0>(g, a, b, gap, percDown, percRight, c);
        }
    }
    if (!(r.x == -1 && r.y == -1)) wrap_to_pi(r.x, r.y);
    return r;
}

// ------------------------------------------------------------------ kernels

__global__ __launch_bounds__(256) void pix_interp(GridV g, int gridlvl, const float* view, int W, int H,
                                                  float hor, float ver, float* thphi, int* gap) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > H || j > W) return;
    size_t c = (size_t)i * (W + 1) + j;
    float theta = view[2 * c] + ver;
    float phi = fmodf(view[2 * c + 1] + hor + BT_FPI2, BT_FPI2);
    int gs = 0;
    F2 r = interpolate_pix(g, theta, phi, gridlvl, gs);
    thphi[2 * c] = r.x;
    thphi[2 * c + 1] = r.y;
    gap[c] = gs;
}

// ---- multi-grid blending (G > 1): a frame between the poses of two grids
// A and B at fraction alpha (MakeImageKernel.cu:386-456,466-510).
//
// border (float2): [0] = (min, max) of pi/GN * i over the points captured in
// either grid, [1] the same of pi/GN * j (findBhCenter), both starting at
// (100, 0) (MakeImageKernel.cu:1308-1310); then for direction s, [2+2s] the
// shadow border of A and [3+2s] that of B (findBhBorders), box-smoothed
// four times (smoothBorder).  GN is the rows of the whole-sky grid.

// min i, max i, min j, max j of the captured points (ints: pi/GN * i is
// non-decreasing in i, so its extremes are those of i)
__global__ void bh_extent_init(int* ext) {
    if (threadIdx.x == 0) {
        ext[0] = 0x7fffffff;
        ext[1] = -1;
        ext[2] = 0x7fffffff;
        ext[3] = -1;
    }
}

__global__ __launch_bounds__(256) void bh_extent(GridV a, GridV b, int* ext) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    const bool cap = i < a.GN && j < a.GM && (a.at(i, j).x == -1.f || b.at(i, j).x == -1.f);
    int v[4] = {cap ? i : 0x7fffffff, cap ? i : -1, cap ? j : 0x7fffffff, cap ? j : -1};
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v[0] = min(v[0], __shfl_xor(v[0], off));
        v[1] = max(v[1], __shfl_xor(v[1], off));
        v[2] = min(v[2], __shfl_xor(v[2], off));
        v[3] = max(v[3], __shfl_xor(v[3], off));
    }
    if ((threadIdx.y * blockDim.x + threadIdx.x) % 64 == 0 && v[1] >= 0) {
        atomicMin(&ext[0], v[0]);
        atomicMax(&ext[1], v[1]);
        atomicMin(&ext[2], v[2]);
        atomicMax(&ext[3], v[3]);
    }
}

// findBhCenter's result from the extents, then findBhBorders: from the
// centre, step one grid unit at a time along direction s until a captured
// point is followed by a visible one.  The reference walks without a bound
// (and reads outside its grid when the walk leaves it); here the walk stops
// at the grid's edge.
__global__ __launch_bounds__(256) void bh_borders(GridV a, GridV b, const int* ext, float2* border) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * kAngleNum) return;
    const float gridsize = BT_FPI / (1.f * a.GN);
    float2 b0 = make_float2(100.f, 0.f), b1 = make_float2(100.f, 0.f);
    if (ext[1] >= 0) {
        b0 = make_float2(fminf(100.f, gridsize * (float)ext[0]), fmaxf(0.f, gridsize * (float)ext[1]));
        b1 = make_float2(fminf(100.f, gridsize * (float)ext[2]), fmaxf(0.f, gridsize * (float)ext[3]));
    }
    if (i == 0) {
        border[0] = b0;
        border[1] = b1;
    }
    const int ii = i / 2;
    const float angle = BT_FPI2 / (1.f * kAngleNum) * 1.f * ii;
    const float thetaChange = -f_sin(angle);
    const float phiChange = f_cos(angle);
    float ptx = .5f * b0.x + .5f * b0.y, pty = .5f * b1.x + .5f * b1.y;
    ptx = ptx / BT_FPI2 * (float)a.GM;
    pty = pty / BT_FPI2 * (float)a.GM;
    const GridV& g = (i % 2) ? b : a;
    float gA = -2.f, gB = -2.f;
    while (!(gA > 0 && gB == -1.f)) {
        gB = gA;
        ptx += thetaChange;
        pty += phiChange;
        const int gx = cvt_i32(ptx), gy = cvt_i32(pty);
        if (gx < 0 || gx >= g.GN || gy < 0 || gy >= g.GM) break;
        gA = g.at(gx, gy).x;
    }
    border[2 + i] = make_float2((ptx - thetaChange) * BT_FPI2 / (1.f * a.GM), (pty - phiChange) * BT_FPI2 / (1.f * a.GM));
}

// smoothBorder (MakeImageKernel.cu:444-456): mean of each border point and
// its two neighbours of the same grid
__global__ __launch_bounds__(256) void smooth_border(const float2* src, float2* dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * kAngleNum) return;
    if (i == 0) {
        dst[0] = src[0];
        dst[1] = src[1];
    }
    const int prev = (i - 2 + 2 * kAngleNum) % (2 * kAngleNum);
    const int next = (i + 2) % (2 * kAngleNum);
    dst[i + 2] = make_float2(1.f / 3.f * (src[prev + 2].x + src[i + 2].x + src[next + 2].x),
                             1.f / 3.f * (src[prev + 2].y + src[i + 2].y + src[next + 2].y));
}

// pixInterpolation with Gr > 1 (MakeImageKernel.cu:466-510).  Inside the
// stretch circle around the shadow the two grids are sampled at positions
// warped so that their shadow borders meet the blended border.  The
// reference's direction slot can reach angleNum when the angle rounds to
// 2 pi (one entry past its border array); here it is clamped.
__global__ __launch_bounds__(256) void pix_interp_blend(GridV ga, GridV gb, int gridlvl, const float2* border,
                                                        float alpha, const float* view, int W, int H, float hor,
                                                        float ver, float* thphi, int* gap) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > H || j > W) return;
    size_t c = (size_t)i * (W + 1) + j;
    float theta = view[2 * c] + ver;
    float phi = fmodf(view[2 * c + 1] + hor + BT_FPI2, BT_FPI2);
    const float2 b0 = border[0], b1 = border[1];
    const float cx = .5f * b0.x + .5f * b0.y, cy = .5f * b1.x + .5f * b1.y;
    const float stretchRad = fmaxf(b0.y - b0.x, b1.x - b1.y) * 0.75f;
    const float centerdist = (theta - cx) * (theta - cx) + (phi - cy) * (phi - cy);
    int gs = 0;
    F2 A, B;
    if (centerdist < stretchRad * stretchRad) {
        float angle = f_atan2(cx - theta, phi - cy);
        angle = fmodf(angle + BT_FPI2, BT_FPI2);
        int slot = cvt_i32(angle / BT_FPI2 * (float)kAngleNum);
        slot = slot < 0 ? 0 : (slot >= kAngleNum ? kAngleNum - 1 : slot);
        const float2 bA = border[2 * slot + 2], bB = border[2 * slot + 3];
        const float nx = (1.f - alpha) * bA.x + alpha * bB.x;
        const float ny = (1.f - alpha) * bA.y + alpha * bB.y;
        if (centerdist <= (nx - cx) * (nx - cx) + (ny - cy) * (ny - cy)) {
            // inside the blended shadow; the reference leaves gap unwritten
            // here, and only shadow pixels touch this corner
            thphi[2 * c] = -1.f;
            thphi[2 * c + 1] = -1.f;
            gap[c] = 0;
            return;
        }
        const float tStoB = (cx - stretchRad * f_sin(angle) - nx);
        const float pStoB = (cy + stretchRad * f_cos(angle) - ny);
        // 1E-5 is a double in the reference: the comparison is in double
        const float thetaPerc = (double)fabsf(tStoB) < 1E-5 ? 0.f : 1.f - (theta - nx) / tStoB;
        const float phiPerc = (double)fabsf(pStoB) < 1E-5 ? 0.f : 1.f - (phi - ny) / pStoB;
        const float thetaA = theta - thetaPerc * (nx - bA.x);
        const float phiA = phi - phiPerc * (ny - bA.y);
        const float thetaB = theta - thetaPerc * (nx - bB.x);
        const float phiB = phi - phiPerc * (ny - bB.y);
        A = interpolate_pix(ga, thetaA, phiA, gridlvl, gs);
        B = interpolate_pix(gb, thetaB, phiB, gridlvl, gs);
    } else {
        A = interpolate_pix(ga, theta, phi, gridlvl, gs);
        B = interpolate_pix(gb, theta, phi, gridlvl, gs);
    }
    F2 r;
    if (A.x == -1.f || B.x == -1.f) {
        r = F2{-1.f, -1.f};
    } else {
        if (A.y < .2f * BT_FPI2 && B.y > .8f * BT_FPI2) A.y += BT_FPI2;
        if (B.y < .2f * BT_FPI2 && A.y > .8f * BT_FPI2) B.y += BT_FPI2;
        r = F2{(1.f - alpha) * A.x + alpha * B.x, fmodf((1.f - alpha) * A.y + alpha * B.y, BT_FPI2)};
    }
    thphi[2 * c] = r.x;
    thphi[2 * c + 1] = r.y;
    gap[c] = gs;   // the second lookup's, as the reference's writes land in that order
}

// kernel.cuh:652-672; false when a corner is negative (shadow).
__device__ __forceinline__ bool corners(const float* thphi, int W, int ind, float* t, float* p,
                                        bool& picheck, float offset) {
    const int M1 = W + 1;
    const float2* tp = reinterpret_cast<const float2*>(thphi);
    float2 c0 = tp[ind + M1], c1 = tp[ind], c2 = tp[ind + 1], c3 = tp[ind + M1 + 1];
    t[0] = c0.x; t[1] = c1.x; t[2] = c2.x; t[3] = c3.x;
    p[0] = c0.y; p[1] = c1.y; p[2] = c2.y; p[3] = c3.y;
    for (int q = 0; q < 4; q++) {
        if (p[q] < 0 || t[q] < 0) return false;
        p[q] = fmodf(p[q] + offset, BT_FPI2);
    }
    picheck = pi_check(p, .2f);
    return true;
}

__global__ __launch_bounds__(256) void bh_area(const float* thphi, int W, int H, uint8_t* bh, float* area) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= H || j >= W) return;
    float t[4], p[4];
    bool pc = false;
    bool ok = corners(thphi, W, i * (W + 1) + j, t, p, pc, 0.0f);
    size_t ij = (size_t)i * W + j;
    bh[ij] = ok ? 0 : 1;
    area[ij] = ok ? calc_area3(t[0], t[1], t[2], p[0], p[1], p[2]) + calc_area3(t[0], t[2], t[3], p[0], p[2], p[3])
                  : 0.f;
}

__device__ __forceinline__ int filter_size(const int* gap, int W, int i, int j) {
    const int M1 = W + 1;
    int fs = gap[i * M1 + j];
    fs = max(fs, gap[i * M1 + M1 + j]);
    fs = max(fs, gap[i * M1 + j + 1]);
    fs = max(fs, gap[i * M1 + M1 + j + 1]);
    return fs / 2;
}

// MakeImageKernel.cu:570-591
__global__ __launch_bounds__(256) void smooth_h(const float* area, const uint8_t* bh, const int* gap, int W,
                                                int H, float* area_s) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= H || j >= W) return;
    size_t ij = (size_t)i * W + j;
    if (bh[ij] == 1) {
        area_s[ij] = 0.f;
        return;
    }
    int fs = filter_size(gap, W, i, j);
    float sum = 0.f;
    int count = 0;
    for (int h = -fs; h <= fs; h++) {
        int jj = ((j + h) % W + W) % W;
        if (bh[(size_t)i * W + jj] == 0) {
            sum += area[(size_t)i * W + jj];
            count++;
        }
    }
    area_s[ij] = sum / (1.f * count);
}

// MakeImageKernel.cu:593-616
__global__ __launch_bounds__(256) void smooth_v(const float* area_s, const uint8_t* bh, const int* gap, int W,
                                                int H, float* area) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= H || j >= W) return;
    size_t ij = (size_t)i * W + j;
    if (bh[ij] == 1) return;
    int fs = filter_size(gap, W, i, j);
    int minv = max(0, i - fs);
    int maxv = min(H - 1, i + fs);
    float sum = 0.f;
    int count = 0;
    for (int h = minv; h <= maxv; h++)
        if (bh[(size_t)h * W + j] == 0) {
            sum += area_s[(size_t)h * W + j];
            count++;
        }
    area[ij] = sum / (1.f * count);
}

// kernel.cuh:674-692
__device__ void lensing_redshift(const float* t, const float* p, int W, int ind, const Cam7& cam,
                                 const float* view, float& frac, float& redshft, float solidAngle) {
    const int M1 = W + 1;
    if (solidAngle == 0.f)
        solidAngle = calc_area3(t[0], t[1], t[2], p[0], p[1], p[2]) + calc_area3(t[0], t[2], t[3], p[0], p[2], p[3]);
    const float2* v = reinterpret_cast<const float2*>(view);
    float2 v0 = v[ind], v1 = v[ind + 1], v2 = v[ind + M1 + 1], v3 = v[ind + M1];
    float ver4[4] = {v0.x, v1.x, v2.x, v3.x};
    float hor4[4] = {v0.y, v1.y, v2.y, v3.y};
    // calc_area4 of these corners, made once per view (view_pixel_areas)
    const int pi = ind / M1, pj = ind - pi * M1;
    const float pixArea = cam.pix_area[(size_t)pi * W + pj];
    frac = pixArea / solidAngle;
    float thetaCam = (ver4[0] + ver4[1] + ver4[2] + ver4[3]) * .25f;
    float phiCam = (hor4[0] + hor4[1] + hor4[2] + hor4[3]) * .25f;
    redshft = redshift(thetaCam, phiCam, cam.v);
}

// The camera-sky solid angle of every pixel's view quad (the pixArea of
// kernel.cuh:679-680): it depends on the view alone, so a view makes it once.
__global__ __launch_bounds__(256) void view_pixel_areas(const float* view, int W, int H, float* out) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= H || j >= W) return;
    const int M1 = W + 1, ind = i * M1 + j;
    const float2* v = reinterpret_cast<const float2*>(view);
    float2 v0 = v[ind], v1 = v[ind + 1], v2 = v[ind + M1 + 1], v3 = v[ind + M1];
    float ver4[4] = {v0.x, v1.x, v2.x, v3.x};
    float hor4[4] = {v0.y, v1.y, v2.y, v3.y};
    out[(size_t)i * W + j] = calc_area4(ver4, hor4);
}

// ---- star pass ---------------------------------------------------------
// distortStarMap (MakeImageKernel.cu:206-326).
//
// Candidates.  The reference tests every star of the star-tree nodes that
// overlap a pixel's quad bounding box (searchTree, MakeImageKernel.cu:27-96):
// ~390 stars per level-8 leaf for a quad of ~1e-3 rad.  Any superset of the
// stars inside the quad yields the same inside-star set, so pixels read
// candidates from a 256 x 512 (theta, phi) binning of the stars instead:
// one contiguous range per bin row, ~1 star per bin.  The exception is a
// quad that crosses phi = 2 pi (picheck): there the reference retests
// candidates with phi < 0.4 pi shifted by 2 pi, so which stars are
// candidates matters, and those pixels walk the reference's tree.
//
// Work split.  Near the shadow a pixel can project onto a large part of the
// sky with tens of thousands of candidates; one lane would serialise them and
// stall its wavefront.  star_map_light takes pixels with at most kLightStars
// candidates (about 92% of the pixels of a lensed frame, 6% of the
// candidates), one lane each, and queues the others for star_map_wave, where
// one wavefront splits a pixel's candidates and reduces the taps in a fixed
// order.
//
// Per tap the reference evaluates 100 * 10^(-0.4 (m - 2.5 log10 x)) with
// x = frac * gaussian(d) and converts HSP -> RGB at brightness P * that.
// For finite x > 0 this equals 100 * x * 10^(-0.4 m), and HSP -> RGB is
// linear in P, so the colour is converted once per star and each tap is a
// few multiplies; non-finite or non-positive x (which yield 0 or NaN in the
// reference) take the reference's own expression.  The star light agrees
// with the reference arithmetic to a few ulps up to summation order; frames
// to within 1/255.
constexpr int kLightStars = 16;

// Appends `value` to list for every lane with `take` set: one atomic per
// wavefront.  Every active lane of the wave must call it.
__device__ __forceinline__ void wave_append(uint32_t* list, unsigned* count, bool take, uint32_t value) {
    const unsigned lane = __lane_id();
    unsigned long long m = __ballot(take);
    if (!m) return;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == (unsigned)leader) base = atomicAdd(count, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (take) list[base + (unsigned)__popcll(m & ((1ull << lane) - 1))] = value;
}

struct StarPix {
    const float* colours;   // star_colours table (6 floats per blackbody index)
    float t[4], p[4];
    float mn[2], mx[2];
    int sgn;
    bool picheck;
    bool odd;   // lensing out of the range where every tap takes the fast form
    float frac, redshft, red;
};

// Per blackbody index: rgb_to_hsp of the table colour (H, S, P) and
// hsp_to_rgb of that at the same P (R, G, B) -- what every star of that
// colour would compute (kernel.cuh:451-468), once.
__global__ void star_colours(float* out) {
    int index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index > 390) return;
    float hh, ss, pp, r1, g1, b1;
    rgb_to_hsp((float)kTempRGB[3 * index] / 255.f, (float)kTempRGB[3 * index + 1] / 255.f,
               (float)kTempRGB[3 * index + 2] / 255.f, hh, ss, pp);
    hsp_to_rgb(hh, ss, pp, r1, g1, b1);
    float* o = out + 6 * index;
    o[0] = hh;
    o[1] = ss;
    o[2] = pp;
    o[3] = r1;
    o[4] = g1;
    o[5] = b1;
}

// Corners, bounding box and orientation of pixel (i, j) (no lensing).
__device__ void star_pixel_geom(const float* thphi, int W, int i, int j, float offset, const Cam7& cam,
                                StarPix& P) {
    P.colours = cam.colours;
    const int ind = i * (W + 1) + j;
    P.picheck = false;
    corners(thphi, W, ind, P.t, P.p, P.picheck, offset);
    const float* t = P.t;
    const float* p = P.p;
    P.mx[0] = fmaxf(fmaxf(t[0], t[1]), fmaxf(t[2], t[3]));
    P.mx[1] = fmaxf(fmaxf(p[0], p[1]), fmaxf(p[2], p[3]));
    P.mn[0] = fminf(fminf(t[0], t[1]), fminf(t[2], t[3]));
    P.mn[1] = fminf(fminf(p[0], p[1]), fminf(p[2], p[3]));
    float orient = (t[1] - t[0]) * (p[1] + p[0]) + (t[2] - t[1]) * (p[2] + p[1]) + (t[3] - t[2]) * (p[3] + p[2]) +
                   (t[0] - t[3]) * (p[0] + p[3]);
    P.sgn = orient < 0 ? -1 : 1;
}

__device__ bool star_pixel(const float* thphi, const uint8_t* bh, int W, int i, int j, float offset, int lr,
                           const float* area, const float* view, const Cam7& cam, StarPix& P) {
    const size_t ij = (size_t)i * W + j;
    if (bh[ij] != 0) return false;
    P.colours = cam.colours;
    const int ind = i * (W + 1) + j;
    P.picheck = false;
    corners(thphi, W, ind, P.t, P.p, P.picheck, offset);
    const float* t = P.t;
    const float* p = P.p;
    P.mx[0] = fmaxf(fmaxf(t[0], t[1]), fmaxf(t[2], t[3]));
    P.mx[1] = fmaxf(fmaxf(p[0], p[1]), fmaxf(p[2], p[3]));
    P.mn[0] = fminf(fminf(t[0], t[1]), fminf(t[2], t[3]));
    P.mn[1] = fminf(fminf(p[0], p[1]), fminf(p[2], p[3]));
    float orient = (t[1] - t[0]) * (p[1] + p[0]) + (t[2] - t[1]) * (p[2] + p[1]) + (t[3] - t[2]) * (p[3] + p[2]) +
                   (t[0] - t[3]) * (p[0] + p[3]);
    P.sgn = orient < 0 ? -1 : 1;
    P.redshft = 1.f;
    P.frac = 1.f;
    if (lr) lensing_redshift(t, p, W, ind, cam, view, P.frac, P.redshft, area[ij]);
    P.red = 4.f * log10f(P.redshft);   // device library FP32 (a few ulps)
    return true;
}

__device__ __forceinline__ void node_range(const int32_t* tree, int node, int& b, int& e) {
    b = (node != 0 && ((node + 1) & node) != 0) ? tree[node - 1] : 0;
    e = tree[node];
}

// searchTree (MakeImageKernel.cu:27-96) with picheck = 0: visit(node) is
// called for every listed node in the reference's list order.
template <class F>
__device__ void search_tree(const int32_t* tree, int treeLevel, const float* mn, const float* mx, F&& visit) {
    float nodeStart[2] = {0.f, 0.f};
    float nodeSize[2] = {BT_FPI, BT_FPI2};
    int node = 0;
    unsigned bitMask = 1u << treeLevel;
    int level = 0;
    int lvl = 0;
    while (bitMask != 0) {
        bitMask &= ~(1u << (treeLevel - level));
        for (lvl = level + 1; lvl <= treeLevel; lvl++) {
            int star_n = tree[node];
            if (node != 0 && ((node + 1) & node) != 0) star_n -= tree[node - 1];
            int tp = lvl & 1;
            float x_overlap = fmaxf(0.f, fminf(mx[0], nodeStart[0] + nodeSize[0]) - fmaxf(mn[0], nodeStart[0]));
            float y_overlap = fmaxf(0.f, fminf(mx[1], nodeStart[1] + nodeSize[1]) - fmaxf(mn[1], nodeStart[1]));
            float overlapArea = x_overlap * y_overlap;
            bool size = overlapArea / (nodeSize[0] * nodeSize[1]) > 0.8f;
            nodeSize[tp] = nodeSize[tp] * .5f;
            if (star_n == 0) {
                node = node * 2 + 1;
                break;
            }
            float check = nodeStart[tp] + nodeSize[tp];
            bool lu = mn[tp] < check;
            bool rd = mx[tp] >= check;
            if (lvl == treeLevel || (rd && lu && size)) {
                if (rd) visit(node * 2 + 2);
                if (lu) visit(node * 2 + 1);
                node = node * 2 + 1;
                break;
            } else {
                node = node * 2 + 1;
                if (rd) bitMask |= 1u << (treeLevel - lvl);
                if (!lu) break;
            }
        }
        level = treeLevel - (__ffs((int)bitMask)) + 1;
        if (level >= 0) {
            int diff = lvl - level;
            for (int q = 0; q < diff; q++) {
                int tp = (lvl - q) & 1;
                if (!(node & 1)) nodeStart[tp] -= nodeSize[tp];
                nodeSize[tp] = nodeSize[tp] * 2.f;
                node = (node - 1) / 2;
            }
            node++;
            int tp = level & 1;
            nodeStart[tp] += nodeSize[tp];
        }
    }
}

__device__ __forceinline__ int bin_of(float v, float d, int n) { return min(max(cvt_i32(v / d), 0), n - 1); }

// The phi bins [q0, q1] of bin row r (theta in [r, r + 1] * kBinDT) that can
// hold a star passing star_in_polygon, within the pixel's bounding-box bins
// [p0, p1]; q0 > q1 when none can.  Each edge test of star_in_polygon,
// A (sp - p_b) - B (st - t_b) > 0 with A = sgn (t_a - t_b) and
// B = sgn (p_a - p_b), bounds phi by a line in theta whose extremes over the
// row are at its two ends; the four bounds are intersected.  Conservative:
// the row is widened by 1e-3 of a bin in theta, the span by one bin on each
// side (the rounding of the FP32 edge tests is far below a bin), and an edge
// that is steep in this frame (|B| > 1000 |A|, where that rounding divided by
// |A| could reach a bin) bounds nothing.  On a strongly lensed pixel (a thin
// arc of the Einstein ring) this tests a small fraction of its bounding
// box's stars.
__device__ __forceinline__ void row_span(const StarPix& P, int r, int p0, int p1, int& q0, int& q1) {
    const float ta = ((float)r - 1e-3f) * kBinDT, tb = ((float)r + 1.001f) * kBinDT;
    float lo = P.mn[1], hi = P.mx[1];
    const float sg = (float)P.sgn;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int kb = (k + 1) & 3;
        const float A = sg * (P.t[k] - P.t[kb]);
        const float B = sg * (P.p[k] - P.p[kb]);
        if (!(fabsf(B) <= 1000.f * fabsf(A))) continue;   // also A == 0, NaN
        const float g = B / A;
        const float la = P.p[kb] + g * (ta - P.t[kb]);
        const float lb = P.p[kb] + g * (tb - P.t[kb]);
        if (A > 0.f)
            lo = fmaxf(lo, fminf(la, lb));
        else
            hi = fminf(hi, fmaxf(la, lb));
    }
    if (!(lo <= hi)) {
        q0 = 1;
        q1 = 0;
        return;
    }
    q0 = max(p0, bin_of(lo, kBinDP, kBinP) - 1);
    q1 = min(p1, bin_of(hi, kBinDP, kBinP) + 1);
}

// f(begin, end) for every candidate range of a pixel; ranges index
// ranges_pos/ranges_mag (tree order for picheck pixels, bin order otherwise).
template <class F>
__device__ void for_ranges(const StarPix& P, const SceneK& s, F&& f) {
    if (P.picheck) {
        search_tree(s.tree, s.tree_level, P.mn, P.mx, [&](int node) {
            int b, e;
            node_range(s.tree, node, b, e);
            f(b, e);
        });
    } else {
        const int r0 = bin_of(P.mn[0], kBinDT, kBinT), r1 = bin_of(P.mx[0], kBinDT, kBinT);
        const int p0 = bin_of(P.mn[1], kBinDP, kBinP), p1 = bin_of(P.mx[1], kBinDP, kBinP);
        for (int r = r0; r <= r1; r++) {
            int q0, q1;
            row_span(P, r, p0, p1, q0, q1);
            if (q0 <= q1) f(s.bin_start[r * kBinP + q0], s.bin_start[r * kBinP + q1 + 1]);
        }
    }
}

// The inside test of one candidate star (MakeImageKernel.cu:285-292): st,
// sp receive its position (phi shifted by 2 pi for the retest of a quad
// across phi = 2 pi).
__device__ __forceinline__ bool star_test(const StarPix& P, const float* pos, int q, float& st, float& sp) {
    const float2 sp2 = reinterpret_cast<const float2*>(pos)[q];
    st = sp2.x;
    sp = sp2.y;
    bool in = star_in_polygon(P.t, P.p, st, sp, P.sgn);
    if (P.picheck && !in && sp < BT_FPI2 * .2f) {
        sp += BT_FPI2;
        in = star_in_polygon(P.t, P.p, st, sp, P.sgn);
    }
    return in;
}

// 10^x in FP32 to a few ulps: 2^(x log2 10), the product carried in two
// parts (log2 10 = L1 + L2, the rounding of x L1 recovered by an FMA), so
// that the exponent's rounding (large for large |x|) does not reach the
// result.
__device__ __forceinline__ float exp10_f32(float x) {
    constexpr float L1 = 3.32192802429199219f;     // log2(10) rounded to float
    constexpr float L2 = 7.05797954e-08f;          // log2(10) - L1
    const float hi = x * L1;
    const float lo = fmaf(x, L1, -hi) + x * L2;
    return exp2f(hi) * (1.f + lo * 0.693147180559945309f);
}

// Whether every tap of every star takes the fast form in this pixel: the
// tap weight x = frac * gaussian(d) lies in (0, 3.4e38] for frac in
// [1e-30, 1e30] (gaussian(d) in [9e-4, 1] over the footprint), and
// E = 10^(-0.4 (m + red)) stays finite for m >= -10, |red| <= 80.
__device__ __forceinline__ bool star_pixel_odd(const StarPix& P, const SceneK& s) {
    return !(s.mag_tame && P.frac >= 1e-30f && P.frac <= 1e30f && P.red >= -80.f && P.red <= 80.f);
}

// The light of one star inside the pixel (MakeImageKernel.cu:293-322).
// Per tap the reference evaluates 100 * 10^(-0.4 (part - 2.5 log10 x)) with
// x = frac * gaussian(d); for finite x > 0 and finite E = 10^(-0.4 part)
// that is 100 x E.  ODD = false (pixels where star_pixel_odd is false):
// every tap takes that fast form.  ODD = true: each tap checks, and takes
// the reference's own expression (kernel.cuh:390, MakeImageKernel.cu:310-318)
// where the fast form does not apply.
template <int STEP, bool ODD>
__device__ __forceinline__ void star_splat(const StarPix& P, const float* mag, int q, float st, float sp,
                                           float (&acc)[(2 * STEP + 1) * (2 * STEP + 1)][3]) {
    constexpr int FW = 2 * STEP + 1;
    interpolate_star(P.t[0], P.t[1], P.t[2], P.t[3], P.p[0], P.p[1], P.p[2], P.p[3], st, sp, P.sgn);
    const float2 mg = reinterpret_cast<const float2*>(mag)[q];
    const float part = mg.x + P.red;
    float temp = 46.f / P.redshft * ((1.f / ((0.92f * mg.y) + 1.7f)) + (1.f / ((0.92f * mg.y) + 0.62f))) - 10.f;
    // clamp to the table's last entry (the reference's 1170 reads past it)
    int index = max(0, min(cvt_i32(floorf(temp)), 390));
    // the blackbody colour's HSP and its RGB at unit brightness, per index
    const float* col = P.colours + 6 * index;
    const float r1 = col[3], g1 = col[4], b1 = col[5];
    const float E = ODD ? f_exp10(-.4f * part) : exp10_f32(-.4f * part);
    const float maxDistSq = (STEP + .5f) * (STEP + .5f);
#pragma unroll
    for (int u = 0; u < FW; u++)
#pragma unroll
        for (int v = 0; v < FW; v++) {
            float dt = -STEP + u + .5f - st;
            float dp = -STEP + v + .5f - sp;
            float dist = dt * dt + dp * dp;
            if (dist > maxDistSq) continue;
            float x = P.frac * gaussian(dist);
            float rr, gg, bb;
            if (!ODD || (x > 0.f && x <= 3.4e38f && fabsf(E) <= 3.4e38f)) {
                float bright = 100.f * x * E;
                rr = r1 * bright;
                gg = g1 * bright;
                bb = b1 * bright;
            } else {
                float appMag = part - 2.5f * f_log10(x);
                float brightness = 100 * f_exp10(-.4f * appMag);
                hsp_to_rgb(col[0], col[1], col[2] * brightness, rr, gg, bb);
            }
            rr *= 255.f;
            gg *= 255.f;
            bb *= 255.f;
            acc[FW * u + v][0] += rr * rr;
            acc[FW * u + v][1] += gg * gg;
            acc[FW * u + v][2] += bb * bb;
        }
}

// One candidate star of one pixel (the body of MakeImageKernel.cu:285-322).
template <int STEP, bool ODD>
__device__ __forceinline__ void star_accumulate(const StarPix& P, const float* pos, const float* mag, int q,
                                                float (&acc)[(2 * STEP + 1) * (2 * STEP + 1)][3]) {
    float st, sp;
    if (star_test(P, pos, q, st, sp)) star_splat<STEP, ODD>(P, mag, q, st, sp, acc);
}

// Pixels with more than kLightStars candidates are queued with their
// candidate count and cut into chunks of kChunk candidates, one wavefront
// per chunk; their chunk indices are taken by an atomic when the pixel is
// queued (contiguous per pixel).  A pixel's chunk partials are summed in
// chunk order, so the result does not depend on which waves ran which
// chunks.  Every queued star is then
// one of a few candidates of one lane, whatever the pixel it lands in: the
// long per-star chain (bisection, taps) of a pixel with tens of thousands of
// candidates is spread over many waves instead of serialised on a few lanes.
constexpr int kChunk = 256;

// A queued count with this bit set is a pixel whose candidates come from
// the star tree (a quad across phi = 2 pi).  The tree is walked once for
// such a pixel (star_tree_ranges) and its listed node ranges are stored, so
// that every chunk of the pixel walks the stored ranges as a binned pixel
// walks its bin rows.
constexpr uint32_t kTreePixel = 0x80000000u;
// A queued count with this bit set is an odd pixel (star_pixel_odd): its
// chunks are traced by the ODD instance of star_map_chunk.
constexpr uint32_t kOddPixel = 0x40000000u;

__device__ __forceinline__ uint32_t queued_count(uint32_t v) { return v & ~(kTreePixel | kOddPixel); }
__device__ __forceinline__ uint32_t queued_chunks(uint32_t v) { return (queued_count(v) + kChunk - 1) / kChunk; }

// owner of a chunk index that holds no work (taken by a pixel whose chunks
// did not all fit `cap`, or whose tree ranges did not fit `rcap`; that
// pixel is traced by its own lane instead)
constexpr uint32_t kNoOwner = 0xffffffffu;

struct StarQueue {
    uint32_t* pix;         // queued pixel indices
    uint32_t* cnt;         // their candidate counts (| kTreePixel)
    uint32_t* nnode;       // tree pixels: listed node count
    uint32_t* rbase;       // tree pixels: first stored range
    int2* ranges;          // stored node ranges (begin, end) of tree pixels
    unsigned rcap;         // capacity of `ranges`
    uint32_t* cstart;      // first chunk of each queued pixel
    uint32_t* owner;       // queued pixel of each chunk (or kNoOwner)
    float* lens;           // per queued pixel: frac, redshift, 4 log10(redshift)
    uint32_t* odd;         // pixels for star_map_odd
    unsigned* ctr;         // [0] queued pixels, [1] chunk indices taken, [2] odd pixels, [3] stored ranges
    float* partial;        // per chunk: FW*FW*3 partial taps
    unsigned cap;          // chunk capacity of `partial`
    uint8_t* lit;          // per pixel: 1 when its taps are written this frame (0: all its taps are zero)
};

// All candidates of a pixel in order, into acc (the reference's loop),
// kLightStars at a time: the window's candidate indices are collected into
// this thread's LDS slots `cand`, then traced, so that the kernel holds one
// copy of the splat whichever of the tree or the bins lists the candidates.
template <int STEP, bool ODD>
__device__ __forceinline__ void star_seq(const StarPix& P, const SceneK& s, int count, int* cand,
                                         float (&acc)[(2 * STEP + 1) * (2 * STEP + 1)][3]) {
    const float* pos = P.picheck ? s.star_pos : s.bin_pos;
    const float* mag = P.picheck ? s.star_mag : s.bin_mag;
    for (int w = 0; w < count; w += kLightStars) {
        int nq = 0, g = 0;
        for_ranges(P, s, [&](int b, int e) {
            const int len = e - b;
            if (g + len > w && g < w + kLightStars) {
                const int q0 = b + max(0, w - g), q1 = b + min(len, w + kLightStars - g);
                for (int q = q0; q < q1; q++) cand[nq++] = q;
            }
            g += len;
        });
        for (int k = 0; k < nq; k++) star_accumulate<STEP, ODD>(P, pos, mag, cand[k], acc);
    }
}

template <int STEP>
__device__ __forceinline__ void store_taps(float* taps, size_t NP, size_t ij,
                                           const float (&acc)[(2 * STEP + 1) * (2 * STEP + 1)][3]) {
    // tap-major layout: taps[(tap * 3 + c) * NP + ij]
#pragma unroll
    for (int u = 0; u < (2 * STEP + 1) * (2 * STEP + 1); u++) {
        taps[(size_t)(u * 3 + 0) * NP + ij] = acc[u][0];
        taps[(size_t)(u * 3 + 1) * NP + ij] = acc[u][1];
        taps[(size_t)(u * 3 + 2) * NP + ij] = acc[u][2];
    }
}

template <int STEP>
__global__ __launch_bounds__(256) void star_map_light(const float* thphi, const uint8_t* bh, SceneK s, Cam7 cam,
                                                      int W, int H, float offset, int lr, const float* area,
                                                      const float* view, float* taps, StarQueue Q) {
    constexpr int FW = 2 * STEP + 1;
    __shared__ int s_cand[256][kLightStars];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    // every lane stays to the end: the queue append below scans over the wave
    const bool valid = i < H && j < W;
    const size_t NP = (size_t)W * H;
    const size_t ij = valid ? (size_t)i * W + j : 0;
    float acc[FW * FW][3];
#pragma unroll
    for (int u = 0; u < FW * FW; u++) acc[u][0] = acc[u][1] = acc[u][2] = 0.f;
    StarPix P;
    int count = 0, nodes = 0;
    bool odd = false;
    if (valid && star_pixel(thphi, bh, W, i, j, offset, lr, area, view, cam, P)) {
        for_ranges(P, s, [&](int b, int e) {
            count += e - b;
            nodes++;
        });
        odd = count > 0 && star_pixel_odd(P, s);
    }
    // pixels of many candidates go to the chunks; odd pixels of few to
    // star_map_odd
    const bool queue = count > kLightStars;
    bool to_odd = odd && !queue;
    bool self = valid && !queue && !odd;   // this lane writes the pixel's taps
    bool seq = self && count > 0;          // ... after tracing its candidates
    // wave-aggregated append of (pixel, count) and of the pixel's chunks
    const unsigned lane = __lane_id();
    const unsigned long long m = __ballot(queue);
    if (m) {
        const int leader = __ffsll((long long)m) - 1;
        const unsigned nchunk = queue ? ((unsigned)count + kChunk - 1) / kChunk : 0u;
        const unsigned nrng = queue && P.picheck ? (unsigned)nodes : 0u;
        unsigned inc = nchunk, rinc = nrng;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            unsigned o = __shfl_up(inc, off), ro = __shfl_up(rinc, off);
            if ((int)lane >= off) {
                inc += o;
                rinc += ro;
            }
        }
        const unsigned tot = __shfl(inc, 63), rtot = __shfl(rinc, 63);
        unsigned base = 0, cbase = 0, rbase = 0;
        if (lane == (unsigned)leader) {
            base = atomicAdd(&Q.ctr[0], (unsigned)__popcll(m));
            cbase = atomicAdd(&Q.ctr[1], tot);
            if (rtot) rbase = atomicAdd(&Q.ctr[3], rtot);
        }
        base = __shfl(base, leader);
        cbase = __shfl(cbase, leader);
        rbase = __shfl(rbase, leader);
        if (queue) {
            const unsigned k = base + (unsigned)__popcll(m & ((1ull << lane) - 1));
            const unsigned long long c0 = (unsigned long long)cbase + inc - nchunk;
            const unsigned long long r0 = (unsigned long long)rbase + rinc - nrng;
            Q.pix[k] = (uint32_t)ij;
            Q.nnode[k] = (uint32_t)nodes;
            Q.lens[3 * k] = P.frac;
            Q.lens[3 * k + 1] = P.redshft;
            Q.lens[3 * k + 2] = P.red;
            Q.cstart[k] = (uint32_t)c0;
            Q.rbase[k] = (uint32_t)r0;
            if (c0 + nchunk <= Q.cap && r0 + nrng <= Q.rcap) {
                Q.cnt[k] = (uint32_t)count | (P.picheck ? kTreePixel : 0u) | (odd ? kOddPixel : 0u);
                for (unsigned c = 0; c < nchunk; c++) Q.owner[c0 + c] = k;
            } else {
                // its chunks or tree ranges do not fit: no chunk work, this
                // lane (or star_map_odd) traces it
                Q.cnt[k] = 0;
                for (unsigned long long c = c0; c < Q.cap && c < c0 + nchunk; c++) Q.owner[c] = kNoOwner;
                if (odd)
                    to_odd = true;
                else
                    seq = self = true;
            }
        }
    }
    wave_append(Q.odd, &Q.ctr[2], to_odd, (uint32_t)ij);
    if (seq) star_seq<STEP, false>(P, s, count, s_cand[threadIdx.y * blockDim.x + threadIdx.x], acc);
    // a pixel without candidates has all-zero taps: they are not stored, and
    // star_sum skips them (the reference zeroes every tap of every pixel)
    if (valid) Q.lit[ij] = (queue || to_odd || seq) ? 1 : 0;
    if (seq) store_taps<STEP>(taps, NP, ij, acc);
}

// One lane per odd pixel (star_pixel_odd) of few candidates: all its
// candidates, each tap checked for the reference's slow expression.
template <int STEP>
__global__ __launch_bounds__(256) void star_map_odd(const float* thphi, const uint8_t* bh, SceneK s, Cam7 cam,
                                                    int W, int H, float offset, int lr, const float* area,
                                                    const float* view, float* taps, StarQueue Q) {
    constexpr int FW = 2 * STEP + 1;
    __shared__ int s_cand[256][kLightStars];
    const unsigned n = Q.ctr[2];
    const size_t NP = (size_t)W * H;
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t ij = Q.odd[k];
        float acc[FW * FW][3];
#pragma unroll
        for (int u = 0; u < FW * FW; u++) acc[u][0] = acc[u][1] = acc[u][2] = 0.f;
        StarPix P;
        if (star_pixel(thphi, bh, W, (int)(ij / (uint32_t)W), (int)(ij % (uint32_t)W), offset, lr, area, view,
                       cam, P)) {
            int count = 0;
            for_ranges(P, s, [&](int b, int e) { count += e - b; });
            star_seq<STEP, true>(P, s, count, s_cand[threadIdx.x], acc);
        }
        store_taps<STEP>(taps, NP, ij, acc);
    }
}

// One lane per queued tree pixel: walks the star tree once more and stores
// the listed node ranges, in list order, at the place star_map_light took
// for them.
__global__ __launch_bounds__(256) void star_tree_ranges(const float* thphi, int W, float offset, SceneK s,
                                                        Cam7 cam, StarQueue Q) {
    const unsigned n = Q.ctr[0];
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (!(Q.cnt[k] & kTreePixel)) continue;
        const uint32_t ij = Q.pix[k];
        StarPix P;
        star_pixel_geom(thphi, W, (int)(ij / (uint32_t)W), (int)(ij % (uint32_t)W), offset, cam, P);
        unsigned r = Q.rbase[k];
        for_ranges(P, s, [&](int b, int e) { Q.ranges[r++] = make_int2(b, e); });
    }
}

// One wavefront per chunk (grid-stride): lane l takes the chunk's
// candidates l, l + 64, ...; the taps are reduced over the wave by a fixed
// xor butterfly into the chunk's partial, summed per pixel by
// star_chunk_reduce.  For a binned pixel
// the wave walks the footprint's bin rows 64 at a time: each lane loads one
// row's range, a wave scan gives the rows' first candidate numbers (kept in
// LDS), and each lane finds the row of each of its candidates.
template <int STEP, bool ODD>
__global__ __launch_bounds__(256) void star_map_chunk(const float* thphi, const uint8_t* bh, SceneK s, Cam7 cam,
                                                      int W, int H, float offset, int lr, const float* area,
                                                      const float* view, float* taps, StarQueue Q) {
    constexpr int FW = 2 * STEP + 1;
    constexpr int NT = FW * FW * 3;
    constexpr int PER = (NT + 63) / 64;   // tap values per lane
    const unsigned lane = threadIdx.x & 63;
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned nwaves = (gridDim.x * blockDim.x) >> 6;
    const unsigned nch = min(Q.ctr[1], Q.cap);
    const unsigned CH = kChunk;
    __shared__ int s_beg[4][64];
    __shared__ unsigned s_pre[4][65];
    // stars found inside, staged until 64 can be splatted by a full wave
    __shared__ int s_iq[4][128];
    __shared__ float s_ist[4][128], s_isp[4][128];
    const unsigned wslot = threadIdx.x >> 6;
    for (unsigned c = wave; c < nch; c += nwaves) {
        const unsigned lo = Q.owner[c];
        if (lo == kNoOwner) continue;
        const uint32_t qc = Q.cnt[lo];
        if (((qc & kOddPixel) != 0) != ODD) continue;   // the other instance's chunk
        const uint32_t ij = Q.pix[lo];
        const unsigned g0 = (c - Q.cstart[lo]) * CH;
        const unsigned g1 = min(g0 + CH, queued_count(qc));
        const int i = (int)(ij / (uint32_t)W), j = (int)(ij % (uint32_t)W);
        StarPix P;
        star_pixel_geom(thphi, W, i, j, offset, cam, P);
        P.frac = Q.lens[3 * lo];
        P.redshft = Q.lens[3 * lo + 1];
        P.red = Q.lens[3 * lo + 2];
        const float* pos = P.picheck ? s.star_pos : s.bin_pos;
        const float* mag = P.picheck ? s.star_mag : s.bin_mag;
        float acc[FW * FW][3];
#pragma unroll
        for (int u = 0; u < FW * FW; u++) acc[u][0] = acc[u][1] = acc[u][2] = 0.f;
        const bool tree = (qc & kTreePixel) != 0;
        const uint32_t rbase = tree ? Q.rbase[lo] : 0u;
        {
            // The pixel's candidate ranges, 64 at a time: the stored tree
            // node ranges, or the bin rows r0..r1 cut to their phi spans.
            // Candidates are tested one per lane; the stars inside are
            // staged in LDS and splatted 64 at a time, so that the long
            // splat runs on full waves however few candidates pass.  One
            // loop alternates filling (up to 64 staged) and splatting, so
            // that the splat has one call site.
            const int r0 = bin_of(P.mn[0], kBinDT, kBinT), r1 = bin_of(P.mx[0], kBinDT, kBinT);
            const int p0 = bin_of(P.mn[1], kBinDP, kBinP), p1 = bin_of(P.mx[1], kBinDP, kBinP);
            const int nrange = tree ? (int)Q.nnode[lo] : r1 - r0 + 1;
            // wave-uniform state: current block, its candidate window
            // [gb, a1), the candidates before it, staged stars
            int kb = 0;
            unsigned base = 0, tot = 0, gb = 0, a1 = 0, nin = 0;
            bool loaded = false, done = false;
            int row = 0;   // per lane: range of the block holding its next candidate
            while (true) {
                while (!done && nin < 64) {
                    if (!loaded) {
                        if (kb >= nrange || base >= g1) {
                            done = true;
                            break;
                        }
                        const int k = kb + (int)lane;
                        int b = 0, e = 0;
                        if (k < nrange) {
                            if (tree) {
                                const int2 R = Q.ranges[rbase + (unsigned)k];
                                b = R.x;
                                e = R.y;
                            } else {
                                const int r = r0 + k;
                                int q0, q1;
                                row_span(P, r, p0, p1, q0, q1);
                                if (q0 <= q1) {
                                    b = s.bin_start[r * kBinP + q0];
                                    e = s.bin_start[r * kBinP + q1 + 1];
                                }
                            }
                        }
                        unsigned len = (unsigned)(e - b), inc = len;
#pragma unroll
                        for (int off = 1; off < 64; off <<= 1) {
                            unsigned o = __shfl_up(inc, off);
                            if ((int)lane >= off) inc += o;
                        }
                        __builtin_amdgcn_wave_barrier();
                        s_beg[wslot][lane] = b;
                        s_pre[wslot][lane + 1] = inc;
                        if (lane == 0) s_pre[wslot][0] = 0;
                        tot = __shfl(inc, 63);
                        __builtin_amdgcn_wave_barrier();
                        gb = max(base, g0);
                        a1 = min(base + tot, g1);
                        // range k with s_pre[k] <= x < s_pre[k+1] for the
                        // lane's first candidate, by bisection; later ones
                        // walk on from it (x grows)
                        row = 0;
                        if (gb + lane < a1) {
                            const unsigned x = gb + lane - base;
                            int hi2 = 63;
                            while (row < hi2) {
                                int mid = (row + hi2 + 1) >> 1;
                                if (s_pre[wslot][mid] <= x) row = mid;
                                else hi2 = mid - 1;
                            }
                        }
                        loaded = true;
                    }
                    if (gb >= a1) {
                        base += tot;
                        kb += 64;
                        loaded = false;
                        continue;
                    }
                    const unsigned g = gb + lane;
                    bool in = false;
                    int q = 0;
                    float st = 0.f, sp = 0.f;
                    if (g < a1) {
                        const unsigned x = g - base;   // candidate number within this block
                        while (s_pre[wslot][row + 1] <= x) row++;
                        q = s_beg[wslot][row] + (int)(x - s_pre[wslot][row]);
                        in = star_test(P, pos, q, st, sp);
                    }
                    const unsigned long long m = __ballot(in);
                    if (in) {
                        const unsigned k = nin + (unsigned)__popcll(m & ((1ull << lane) - 1));
                        s_iq[wslot][k] = q;
                        s_ist[wslot][k] = st;
                        s_isp[wslot][k] = sp;
                    }
                    nin += (unsigned)__popcll(m);
                    gb += 64;
                    __builtin_amdgcn_wave_barrier();
                }
                if (nin == 0) break;
                const unsigned take = nin < 64 ? nin : 64u;
                if (lane < take)
                    star_splat<STEP, ODD>(P, mag, s_iq[wslot][lane], s_ist[wslot][lane], s_isp[wslot][lane],
                                          acc);
                const unsigned rest = nin - take;
                int q2 = 0;
                float st2 = 0.f, sp2 = 0.f;
                if (lane < rest) {
                    q2 = s_iq[wslot][64 + lane];
                    st2 = s_ist[wslot][64 + lane];
                    sp2 = s_isp[wslot][64 + lane];
                }
                __builtin_amdgcn_wave_barrier();
                if (lane < rest) {
                    s_iq[wslot][lane] = q2;
                    s_ist[wslot][lane] = st2;
                    s_isp[wslot][lane] = sp2;
                }
                nin = rest;
                __builtin_amdgcn_wave_barrier();
                if (done && nin == 0) break;
            }
        }
        // tap value t (tap t / 3, colour t % 3) goes to lane t % 64, slot t / 64
        float mine[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) mine[k] = 0.f;
#pragma unroll
        for (int u = 0; u < FW * FW; u++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                float v = acc[u][cc];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
                if (lane == (unsigned)((3 * u + cc) & 63)) mine[(3 * u + cc) >> 6] = v;
            }
        // contiguous per chunk (one coalesced store); the per-tap planes
        // are written by star_chunk_reduce
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const unsigned t = (unsigned)k * 64 + lane;
            if (t < (unsigned)NT) Q.partial[(size_t)c * NT + t] = mine[k];
        }
    }
}

// One thread per (queued pixel, tap value): the pixel's chunk partials
// summed in chunk order.
template <int STEP>
__global__ __launch_bounds__(256) void star_chunk_reduce(int W, int H, float* taps, StarQueue Q) {
    constexpr int NT = (2 * STEP + 1) * (2 * STEP + 1) * 3;
    const unsigned n = Q.ctr[0];
    const size_t NP = (size_t)W * H;
    const size_t total = (size_t)n * NT;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += (size_t)gridDim.x * blockDim.x) {
        const unsigned k = (unsigned)(x / NT), t = (unsigned)(x % NT);
        const unsigned nchunk = queued_chunks(Q.cnt[k]);
        if (nchunk == 0) continue;   // traced by its own lane
        const unsigned c0 = Q.cstart[k];
        float v = 0.f;
        for (unsigned ch = 0; ch < nchunk; ch++) v += Q.partial[(size_t)(c0 + ch) * NT + t];
        taps[(size_t)t * NP + Q.pix[k]] = v;
    }
}

// sumStarLight (MakeImageKernel.cu:753-773, bounds-checked) + makePix (:802-835)
// makePix (MakeImageKernel.cu:802-835) of one pixel's summed star light.
__device__ __forceinline__ void make_pix(float br, float bg, float bb, uint8_t* img, size_t ij) {
    float r = sqrtf(br), g = sqrtf(bg), b = sqrtf(bb);
    float mxc = fmaxf(fmaxf(r, g), b);
    if (mxc > 255.f) {
        float hh, ss, pp;
        rgb_to_hsp(r / 255.f, g / 255.f, b / 255.f, hh, ss, pp);
        pp = 1.0f;
        hsp_to_rgb(hh, ss, pp, r, g, b);
        r *= 255;
        g *= 255;
        b *= 255;
    }
    uchar4 o;
    o.x = (uint8_t)min(255, cvt_i32(b + 0.f));
    o.y = (uint8_t)min(255, cvt_i32(g + 0.f));
    o.z = (uint8_t)min(255, cvt_i32(r + 0.f));
    o.w = 255;
    reinterpret_cast<uchar4*>(img)[ij] = o;
}

// sumStarLight (MakeImageKernel.cu:753-773) + makePix; trail (or null): the
// frame's star trails, added after the taps (factor 1).
template <int STEP>
__global__ __launch_bounds__(256) void star_sum(const float* taps, const uint8_t* lit, int W, int H, uint8_t* img,
                                                float* o_star, const float* trail) {
    constexpr int FW = 2 * STEP + 1;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= H || j >= W) return;
    const size_t NP = (size_t)W * H;
    float br = 0.f, bg = 0.f, bb = 0.f;
    int start = max(0, STEP - i);
    int stop = min(2 * STEP, STEP + H - i - 1);
    const float factor = 20.f;
    for (int u = start; u <= stop; u++)
        for (int v = 0; v <= 2 * STEP; v++) {
            size_t src = (size_t)(i + u - STEP) * W + ((j + v - STEP + W) % W);
            if (!lit[src]) continue;   // its taps are zero (and not stored)
            int tap = FW * FW - (FW * u + v + 1);
            br += factor * taps[(size_t)(tap * 3 + 0) * NP + src];
            bg += factor * taps[(size_t)(tap * 3 + 1) * NP + src];
            bb += factor * taps[(size_t)(tap * 3 + 2) * NP + src];
        }
    const size_t ij = (size_t)i * W + j;
    if (trail) {
        br += trail[3 * ij];
        bg += trail[3 * ij + 1];
        bb += trail[3 * ij + 2];
    }
    if (o_star) {
        o_star[3 * ij] = br;
        o_star[3 * ij + 1] = bg;
        o_star[3 * ij + 2] = bb;
    }
    make_pix(br, bg, bb, img, ij);
}

// ---- star trails and diffraction (MakeImageKernel.cu:98-174,517-537,
// 618-629,775-800): effects of a sequence of frames.
//
// Each frame records, per star, the pixels its images land in (pixels with
// fewer than S/100 candidate stars in their tree leaves), keeping the first
// kTrailNum; the next frame draws, for each image, a fading line from the
// previous frame's image of the same star that lies closest to the
// direction of celestial motion (the gradient field rotated by 90 degrees),
// within 45 degrees and M/25 pixels, unless the star moved by at most one
// pixel.  The reference appends images in atomic order and its lines
// overwrite each other in whatever order threads run; here a star's images
// are kept in pixel order and overlapping trails keep the per-channel
// maximum, so a frame does not depend on scheduling.

__global__ __launch_bounds__(256) void trail_clear(TrailK T, int frame0, size_t n3) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = t; q < (size_t)T.S; q += stride) {
        T.nhits[q] = 0;
        T.count[(size_t)T.cur * T.S + q] = 0;
        if (frame0) T.count[(size_t)(1 - T.cur) * T.S + q] = 0;
    }
    for (size_t p = t; p < n3; p += stride) T.trail[p] = 0.f;
    if (t == 0) T.stats[1] = 0;
}

// makeGradField (MakeImageKernel.cu:517-537) per pixel corner.  Corners in
// the shadow get (0, 0) (the reference leaves them unwritten), and the
// right-hand neighbour of the last corner of the last row, which the
// reference reads one element past its array, counts as invalid.
__global__ __launch_bounds__(256) void trail_grad(const float* thphi, int W, int H, float* grad) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > H || j > W) return;
    const int M1 = W + 1;
    const size_t NC = (size_t)M1 * (H + 1);
    auto th = [&](size_t c) { return c < NC ? thphi[2 * c] : -1.f; };
    const float up = th((size_t)min(H, i + 1) * M1 + j);
    const float down = th((size_t)max(0, i - 1) * M1 + j);
    const float left = th((size_t)i * M1 + max(0, j - 1));
    const float right = th((size_t)i * M1 + min(M1, j + 1));
    const float mid = th((size_t)i * M1 + j);
    float gx = 0.f, gy = 0.f;
    if (mid > 0) {
        float xdir = 0;
        if (up > 0 && down > 0) xdir = .5f * (up - mid) - .5f * (down - mid);
        float ydir = 0;
        if (left > 0 && right > 0) ydir = .5f * (right - mid) - .5f * (left - mid);
        gx = -ydir;
        gy = xdir;
    }
    grad[2 * ((size_t)i * M1 + j)] = gx;
    grad[2 * ((size_t)i * M1 + j) + 1] = gy;
}

// The images of this frame: per pixel with starsToCheck < S/100
// (MakeImageKernel.cu:273-282,301), every star inside its quad, with the
// quantities addTrails uses (part, frac, blackbody index).
__global__ __launch_bounds__(256) void trail_hits(const float* thphi, const uint8_t* bh, SceneK s, Cam7 cam,
                                                  int W, int H, float offset, int lr, const float* area,
                                                  const float* view, TrailK T) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= H || j >= W) return;
    StarPix P;
    if (!star_pixel(thphi, bh, W, i, j, offset, lr, area, view, cam, P)) return;
    int starsToCheck = 0;
    search_tree(s.tree, s.tree_level, P.mn, P.mx, [&](int node) {
        int b, e;
        node_range(s.tree, node, b, e);
        starsToCheck += e - b;
    });
    if (!(starsToCheck < s.S / 100)) return;
    const uint32_t pix = (uint32_t)i * (uint32_t)W + (uint32_t)j;
    // candidates as in the star pass: the binned stars of the quad's rows
    // (tree order for quads across phi = 2 pi); the same stars test inside
    const float* pos = P.picheck ? s.star_pos : s.bin_pos;
    const float* magv = P.picheck ? s.star_mag : s.bin_mag;
    for_ranges(P, s, [&](int b, int e) {
        for (int k = b; k < e; k++) {
            float st, sp;
            if (!star_test(P, pos, k, st, sp)) continue;
            const int q = P.picheck ? k : s.bin_idx[k];
            const float2 mg = reinterpret_cast<const float2*>(magv)[k];
            const float part = mg.x + P.red;
            float temp = 46.f / P.redshft * ((1.f / ((0.92f * mg.y) + 1.7f)) + (1.f / ((0.92f * mg.y) + 0.62f))) - 10.f;
            const int index = max(0, min(cvt_i32(floorf(temp)), 390));
            const unsigned slot = atomicAdd(&T.nhits[q], 1u);
            if (slot < (unsigned)kTrailHits) {
                const size_t h = (size_t)q * kTrailHits + slot;
                T.hit_pix[h] = pix;
                T.hit_data[3 * h] = part;
                T.hit_data[3 * h + 1] = P.frac;
                T.hit_data[3 * h + 2] = (float)index;
            } else {
                atomicAdd(&T.stats[1], 1u);
            }
        }
    });
}

// Per star: its images in pixel order; the first kTrailNum become this
// frame's cache (MakeImageKernel.cu:102-104).
__global__ __launch_bounds__(256) void trail_cache(TrailK T, int W) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= T.S) return;
    const int n = (int)min(T.nhits[q], (unsigned)kTrailHits);
    if (n == 0) return;
    uint32_t* hp = T.hit_pix + (size_t)q * kTrailHits;
    float* hd = T.hit_data + (size_t)q * kTrailHits * 3;
    for (int a = 1; a < n; a++) {   // insertion sort by pixel (a star has a few images)
        const uint32_t kp = hp[a];
        const float d0 = hd[3 * a], d1 = hd[3 * a + 1], d2 = hd[3 * a + 2];
        int b = a - 1;
        while (b >= 0 && hp[b] > kp) {
            hp[b + 1] = hp[b];
            hd[3 * (b + 1)] = hd[3 * b];
            hd[3 * (b + 1) + 1] = hd[3 * b + 1];
            hd[3 * (b + 1) + 2] = hd[3 * b + 2];
            b--;
        }
        hp[b + 1] = kp;
        hd[3 * (b + 1)] = d0;
        hd[3 * (b + 1) + 1] = d1;
        hd[3 * (b + 1) + 2] = d2;
    }
    const int keep = min(n, kTrailNum);
    int2* c = T.cache + ((size_t)T.cur * T.S + q) * kTrailNum;
    for (int k = 0; k < keep; k++) c[k] = make_int2((int)(hp[k] / (uint32_t)W), (int)(hp[k] % (uint32_t)W));
    T.count[(size_t)T.cur * T.S + q] = keep;
}

__device__ __forceinline__ void trail_put(float* trail, size_t ij, float v, int ch) {
    if (v >= 0.f) atomicMax(reinterpret_cast<int*>(trail) + 3 * ij + ch, __float_as_int(v));
}

// addTrails (MakeImageKernel.cu:98-174) for one image of one star.
__global__ __launch_bounds__(256) void trail_lines(TrailK T, Cam7 cam, const float* grad, int W, int H) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)T.S * kTrailHits) return;
    const int q = (int)(t / kTrailHits), slot = (int)(t % kTrailHits);
    if (slot >= (int)min(T.nhits[q], (unsigned)kTrailHits)) return;
    const size_t h = (size_t)q * kTrailHits + slot;
    const int M = W, M1 = W + 1;
    const int i = (int)(T.hit_pix[h] / (uint32_t)W), j = (int)(T.hit_pix[h] % (uint32_t)W);
    const float part = T.hit_data[3 * h], frac = T.hit_data[3 * h + 1];
    const float* col = cam.colours + 6 * (int)T.hit_data[3 * h + 2];
    const int prev_half = 1 - T.cur;
    const int pc = T.count[(size_t)prev_half * T.S + q];
    const int2* pcache = T.cache + ((size_t)prev_half * T.S + q) * kTrailNum;
    float traildist = (float)M;
    float angle = BT_FPI2;
    int2 prev = make_int2(0, 0);
    bool line = false;
    for (int w = 0; w < pc; w++) {
        const int2 pr = pcache[w];
        const int dx = pr.x - i, dy = pr.y - j;
        const int dxx = dx * dx, dyy = dy * dy;
        if (dxx <= 1 && dyy <= 1) {
            line = false;
            break;
        }
        const float grx = grad[2 * ((size_t)pr.x * M1 + pr.y)], gry = grad[2 * ((size_t)pr.x * M1 + pr.y) + 1];
        const float dist = sqrtf((float)(dxx + dyy));
        if (dist > M / 25.f) continue;
        const float div = (dist * sqrtf(grx * grx + gry * gry));
        const float a1 = f_acos((1.f * dx * grx + 1.f * dy * gry) / div);
        const float a2 = f_acos((1.f * dx * -grx + 1.f * dy * -gry) / div);
        const float a = fminf(a1, a2);
        if (a > angle || a > BT_FPI * .25f) continue;
        if (a < angle) {
            angle = a;
            traildist = dist;
            prev = pr;
            line = true;
        }
    }
    if (!line) return;
    const int deltax = i - prev.x, deltay = j - prev.y;
    const int sgnDeltaX = deltax < 0 ? -1 : 1, sgnDeltaY = deltay < 0 ? -1 : 1;
    const float deltaerr = deltay == 0 ? fabsf((float)deltax) : fabsf(deltax / (1.f * deltay));
    float error = 0.f;
    int y = prev.y, x = prev.x;
    const float appMag = part - 2.5f * f_log10(frac);
    const float bright0 = 100 * f_exp10(-.4f * appMag);
    // the reference's walk can miss (i, j) (a purely vertical move steps y
    // off the line) and then never ends; here it stops after as many steps
    // as the line is long or when it leaves the frame
    const int maxit = 2 * (abs(deltax) + abs(deltay)) + 2;
    for (int it = 0; it < maxit && (y != j || x != i); it++) {
        if (error < 1.f) {
            y += sgnDeltaY;
            error += deltaerr;
        }
        if (error >= 0.5f) {
            x += sgnDeltaX;
            error -= 1.f;
        }
        if (x < 0 || x >= H || y < 0 || y >= W) break;
        const float dist = (float)((x - i) * (x - i) + (y - j) * (y - j));
        const float fade = (traildist - sqrtf(dist)) * (traildist - sqrtf(dist)) / (traildist * traildist);
        const float brightness = bright0 * fade;
        float r, g, b;
        hsp_to_rgb(col[0], col[1], col[2] * brightness, r, g, b);
        r *= 255.f;
        g *= 255.f;
        b *= 255.f;
        const size_t ij = (size_t)x * M + y;
        trail_put(T.trail, ij, r * r, 0);
        trail_put(T.trail, ij, g * g, 1);
        trail_put(T.trail, ij, b * b, 2);
        if (dist <= 1.f) break;
    }
}

// addDiffraction (MakeImageKernel.cu:775-800): after sumStarLight the
// diffraction sprite is scattered around each pixel brighter than 255^2,
// then makePix runs.  Sources and thresholds are the light before
// diffraction (the reference reads them while other threads add to them).
// Here each output pixel gathers the contributions of the bright pixels
// whose sprite window covers it, in ascending pixel order, so the sums do
// not depend on scheduling (the reference adds with float atomics).

// Per frame row, the columns of its bright pixels in ascending order and
// their light over the reference's divisor 10000 * max^0.6 (one wavefront
// per row).
__global__ __launch_bounds__(256) void diff_rows(const float* light, int W, int H, uint32_t* cols, float* scaled,
                                                 int* nrow) {
    const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (row >= H) return;
    const unsigned lane = threadIdx.x & 63;
    int n = 0;
    for (int j0 = 0; j0 < W; j0 += 64) {
        const int j = j0 + (int)lane;
        bool hot = false;
        float l0 = 0.f, l1 = 0.f, l2 = 0.f, mx = 0.f;
        if (j < W) {
            const float* l = light + 3 * ((size_t)row * W + j);
            l0 = l[0];
            l1 = l[1];
            l2 = l[2];
            mx = fmaxf(fmaxf(l0, l1), l2);
            hot = mx > 65025.f;
        }
        const unsigned long long m = __ballot(hot);
        if (hot) {
            const size_t o = (size_t)row * W + n + __popcll(m & ((1ull << lane) - 1));
            const float div = 10000 * f_pow(mx, 0.6f);
            cols[o] = (uint32_t)j;
            scaled[3 * o] = l0 / div;
            scaled[3 * o + 1] = l1 / div;
            scaled[3 * o + 2] = l2 / div;
        }
        n += __popcll(m);
    }
    if (lane == 0) nrow[row] = n;
}

constexpr int kDiffCand = 512;   // bright pixels staged in LDS per round

// One 16 x 16 tile of output pixels per workgroup: lane 0 lists, in pixel
// order, the bright pixels whose window can reach the tile; every thread
// then adds the ones whose window covers its pixel.
__global__ __launch_bounds__(256) void diff_gather(const float* light, float* acc, const uint32_t* cols,
                                                   const float* scaled, const int* nrow, const uint8_t* sprite, int fs,
                                                   int W, int H) {
    __shared__ int c_r[kDiffCand], c_c[kDiffCand], c_x0[kDiffCand], c_x1[kDiffCand];
    __shared__ float c_l[kDiffCand][3];
    __shared__ int s_n, s_row, s_k, s_done;
    const int oj0 = blockIdx.x * 16, oi0 = blockIdx.y * 16;
    const int oj = oj0 + (int)threadIdx.x, oi = oi0 + (int)threadIdx.y;
    const int tid = threadIdx.y * 16 + threadIdx.x;
    const int fh = fs / 2;
    const bool mine = oi < H && oj < W;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    if (mine) {
        const float* l = light + 3 * ((size_t)oi * W + oj);
        a0 = l[0];
        a1 = l[1];
        a2 = l[2];
    }
    // rows whose windows can reach the tile
    const int r_lo = max(0, oi0 - (fs - 1 - fh)), r_hi = min(H - 1, oi0 + 15 + fh);
    if (tid == 0) {
        s_row = r_lo;
        s_k = 0;
        s_done = 0;
    }
    __syncthreads();
    while (true) {
        if (tid < 64) {
            // wavefront 0 lists the next candidates in pixel order, 64 row
            // entries at a time (a ballot keeps their order)
            int n = 0, r = s_row, k = s_k;
            const unsigned lane = (unsigned)tid;
            while (r <= r_hi) {
                const int nr = nrow[r];
                if (k >= nr) {
                    r++;
                    k = 0;
                    continue;
                }
                const int kk = k + (int)lane;
                bool take = false;
                int c = 0;
                if (kk < nr) {
                    c = (int)cols[(size_t)r * W + kk];
                    // the window's columns c - fh .. c - fh + fs - 1 (mod W) against the tile's
                    const int e = ((c - fh - oj0) % W + W) % W;
                    take = e <= 15 || e >= W - fs + 1;
                }
                const unsigned long long m = __ballot(take);
                if (n + __popcll(m) > kDiffCand) break;   // the rest in the next round
                if (take) {
                    const int d = n + __popcll(m & ((1ull << lane) - 1));
                    int startx = 0, endx = fs;
                    if (r < fh) startx = fh - r;
                    if (r > (H - fh)) endx = H - r + fh;
                    const float* l = scaled + 3 * ((size_t)r * W + kk);
                    c_r[d] = r;
                    c_c[d] = c;
                    c_x0[d] = startx;
                    c_x1[d] = endx;
                    c_l[d][0] = l[0];
                    c_l[d][1] = l[1];
                    c_l[d][2] = l[2];
                }
                n += __popcll(m);
                k += 64;
            }
            if (lane == 0) {
                s_n = n;
                s_row = r;
                s_k = k;
                s_done = r > r_hi;
            }
        }
        __syncthreads();
        const int n = s_n;
        if (mine) {
            for (int k = 0; k < n; k++) {
                const int qq = oi - (c_r[k] - fh);
                if (qq < c_x0[k] || qq >= c_x1[k]) continue;
                const int pp = ((oj - c_c[k] + fh) % W + W) % W;
                if (pp >= fs) continue;
                const uint8_t* sp = sprite + 3 * ((size_t)qq * fs + pp);
                const float dx = c_l[k][0] * (float)sp[0], dy = c_l[k][1] * (float)sp[1], dz = c_l[k][2] * (float)sp[2];
                a0 += fminf(65025.f, dx * dx);
                a1 += fminf(65025.f, dy * dy);
                a2 += fminf(65025.f, dz * dz);
            }
        }
        const int done = s_done;
        __syncthreads();
        if (done) break;
    }
    if (mine) {
        float* o = acc + 3 * ((size_t)oi * W + oj);
        o[0] = a0;
        o[1] = a1;
        o[2] = a2;
    }
}

__global__ __launch_bounds__(256) void copy_f32(const float* src, float* dst, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

__global__ __launch_bounds__(256) void make_pix_all(const float* light, int64_t NP, uint8_t* img) {
    const int64_t ij = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ij < NP) make_pix(light[3 * ij], light[3 * ij + 1], light[3 * ij + 2], img, (size_t)ij);
}

// distortEnvironmentMap (MakeImageKernel.cu:631-751)
// ---- distortEnvironmentMap (MakeImageKernel.cu:631-751) ----
// Pixels whose celestial footprint spans at most kHeavyCols sky columns
// (about 90% of them) are done one per lane by env_map; those with up to
// kLaneCols queue for env_map_lane (one per lane, compacted, so that a wave's
// lanes have similar work) and the rest for env_map_heavy, one wavefront
// per pixel, so that a pixel on the Einstein ring with thousands of columns
// does not stall its neighbours' wave.

constexpr int kHeavyCols = 32;

struct EnvPix {
    float t[4], p[4];
    int ind, pixMin, pixNum;
    float pixSize, thMax, thMin;
};

// Corners and column range of pixel (i, j); false for pixel index 0 (the
// reference's `ind > 0`, MakeImageKernel.cu:647), which stays black.
__device__ __forceinline__ bool env_setup(const float* thphi, int W, int i, int j, float offset, int rows,
                                          EnvPix& E) {
    bool picheck = false;
    E.ind = i * (W + 1) + j;
    corners(thphi, W, E.ind, E.t, E.p, picheck, offset);
    if (E.ind <= 0) return false;
    E.pixSize = BT_FPI / (float)rows;
    float phMax = fmaxf(fmaxf(E.p[0], E.p[1]), fmaxf(E.p[2], E.p[3]));
    float phMin = fminf(fminf(E.p[0], E.p[1]), fminf(E.p[2], E.p[3]));
    int pixMax = cvt_i32(phMax / E.pixSize);
    E.pixMin = cvt_i32(phMin / E.pixSize);
    E.pixNum = pixMax - E.pixMin + 1;
    if (E.pixNum > 1 && pixMax * E.pixSize > phMax) E.pixNum -= 1;
    E.thMax = fmaxf(fmaxf(E.t[0], E.t[1]), fmaxf(E.t[2], E.t[3]));
    E.thMin = fminf(fminf(E.t[0], E.t[1]), fminf(E.t[2], E.t[3]));
    return true;
}

// The quad's theta span on the column boundary at phi = cp, folded into the
// running (mn, mx) (the reference never resets it between boundaries).
__device__ __forceinline__ void env_boundary(const EnvPix& E, float cp, float& mn, float& mx) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float ap = E.p[q], bp = E.p[(q + 1) % 4];
        float at = E.t[q], bt = E.t[(q + 1) % 4];
        if ((cp - ap) * (cp - bp) <= 0) {
            float ct = at + (bt - at) / (bp - ap) * (cp - ap);
            mn = ct < mn ? ct : mn;
            mx = ct > mx ? ct : mx;
        }
    }
}

// Summed-table difference of column `col` over the running span.
__device__ __forceinline__ float4 env_column(const float4* table, int rows, int cols, float pixSize, int col,
                                             float mn, float mx) {
    int rmax = min(max(cvt_i32(mx / pixSize), 0), rows - 1);
    float4 mxc = table[(size_t)rmax * cols + col];
    int index2 = (cvt_i32(mn / pixSize) - 1) * cols + col;
    float4 mnc = make_float4(0.f, 0.f, 0.f, 0.f);
    if (index2 > 0) {
        if (index2 >= rows * cols) index2 = (rows - 1) * cols + col;
        mnc = table[index2];
    }
    return make_float4(mxc.x - mnc.x, mxc.y - mnc.y, mxc.z - mnc.z, mxc.w - mnc.w);
}

// Mean colour -> gamma -> HSP -> lensing/redshift -> BGRA (MakeImageKernel.cu:733-750).
__device__ __forceinline__ uchar4 env_finish(const EnvPix& E, float cx, float cy, float cz, float cw, int W,
                                             int lr, const Cam7& cam, const float* view, float area_ij) {
    // FP32 pow of the device library (a few ulps, as the reference's own
    // powf): these values only reach the 8-bit output
    cx = fminf(255.f, powf(cx / cw, 1.f / 2.2f));
    cy = fminf(255.f, powf(cy / cw, 1.f / 2.2f));
    cz = fminf(255.f, powf(cz / cw, 1.f / 2.2f));
    float Hh, Ss, Pp;
    rgb_to_hsp(cz / 255.f, cy / 255.f, cx / 255.f, Hh, Ss, Pp);
    if (lr) {
        float redshft, frac;
        lensing_redshift(E.t, E.p, W, E.ind, cam, view, frac, redshft, area_ij);
        Pp *= frac;
        Pp = powf(Pp, redshft);
    }
    hsp_to_rgb(Hh, Ss, fminf(1.f, Pp), cz, cy, cx);
    uchar4 o;
    o.x = (uint8_t)min(255, cvt_i32(cx * 255));
    o.y = (uint8_t)min(255, cvt_i32(cy * 255));
    o.z = (uint8_t)min(255, cvt_i32(cz * 255));
    o.w = 255;
    return o;
}

// Column sums of a pixel in the reference's order (one lane).
__device__ __forceinline__ float4 env_seq(const EnvPix& E, const float4* table, int rows, int cols) {
    float cx = 0.f, cy = 0.f, cz = 0.f, cw = 0.f;
    if (E.pixNum == 1) {
        int rmax = min(max(cvt_i32(E.thMax / E.pixSize), 0), rows - 1);
        int rmin = cvt_i32(E.thMin / E.pixSize) - 1;
        int col = min(max(E.pixMin, 0), cols - 1);
        float4 mxc = table[(size_t)rmax * cols + col];
        float4 mnc = make_float4(0.f, 0.f, 0.f, 0.f);
        if (rmin >= 0) mnc = table[(size_t)min(rmin, rows - 1) * cols + col];
        cx = mxc.x - mnc.x;
        cy = mxc.y - mnc.y;
        cz = mxc.z - mnc.z;
        cw = mxc.w - mnc.w;
    } else if (E.pixNum < cols) {
        float mx = -100, mn = cols + 1;
        for (int sidx = 0; sidx <= E.pixNum; sidx++) {
            env_boundary(E, (E.pixMin + sidx) * E.pixSize, mn, mx);
            if (sidx > 0) {
                float4 d = env_column(table, rows, cols, E.pixSize, (E.pixMin + sidx) % cols, mn, mx);
                cx += d.x;
                cy += d.y;
                cz += d.z;
                cw += d.w;
            }
        }
    } else {
        int thMaxPix = min(max(cvt_i32(E.thMax / E.pixSize), 0), rows - 1) * cols;
        int thMinPix = min(max(cvt_i32(E.thMin / E.pixSize), 0), rows - 1) * cols;
        for (int q = 0; q < E.pixNum; q++) {
            float4 mxc = table[thMaxPix + (E.pixMin + q) % cols];
            float4 mnc = table[thMinPix + (E.pixMin + q) % cols];
            cx += mxc.x - mnc.x;
            cy += mxc.y - mnc.y;
            cz += mxc.z - mnc.z;
            cw += mxc.w - mnc.w;
        }
    }
    return make_float4(cx, cy, cz, cw);
}

constexpr int kLaneCols = 256;

__global__ __launch_bounds__(256) void env_map(const float* thphi, const uint8_t* bh, SceneK s, Cam7 cam, int W,
                                               int H, float offset, int lr, const float* area,
                                               const float* view, uint8_t* img, uint32_t* qlane, uint32_t* qwave,
                                               unsigned* nq) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= H || j >= W) return;
    const size_t ij = (size_t)i * W + j;
    const float4* table = reinterpret_cast<const float4*>(s.table);
    uchar4 o = make_uchar4(0, 0, 0, 255);
    EnvPix E;
    const bool live = bh[ij] == 0 && env_setup(thphi, W, i, j, offset, s.rows, E);
    const int tier = !live || E.pixNum <= kHeavyCols ? 0 : E.pixNum <= kLaneCols ? 1 : 2;
    wave_append(qlane, &nq[0], tier == 1, (uint32_t)ij);
    wave_append(qwave, &nq[1], tier == 2, (uint32_t)ij);
    if (tier != 0) return;   // env_map_lane / env_map_heavy write this pixel
    if (live) {
        float4 c4 = env_seq(E, table, s.rows, s.cols);
        o = env_finish(E, c4.x, c4.y, c4.z, c4.w, W, lr, cam, view, area[ij]);
    }
    reinterpret_cast<uchar4*>(img)[ij] = o;
}

// One queued pixel per lane, in the reference's column order.
__global__ __launch_bounds__(256) void env_map_lane(const float* thphi, SceneK s, Cam7 cam, int W, float offset,
                                                    int lr, const float* area, const float* view, uint8_t* img,
                                                    const uint32_t* q, const unsigned* nq) {
    const unsigned n = *nq;
    const float4* table = reinterpret_cast<const float4*>(s.table);
    for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const uint32_t ij = q[e];
        EnvPix E;
        env_setup(thphi, W, (int)(ij / (uint32_t)W), (int)(ij % (uint32_t)W), offset, s.rows, E);
        float4 c4 = env_seq(E, table, s.rows, s.cols);
        reinterpret_cast<uchar4*>(img)[ij] = env_finish(E, c4.x, c4.y, c4.z, c4.w, W, lr, cam, view, area[ij]);
    }
}

__device__ __forceinline__ float wave_xor_sum(float v) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) v += __shfl_xor(v, off);
    return v;
}

// One wavefront per queued pixel (grid-stride over the queue).  Lanes take
// contiguous chunks of the column boundaries; the running theta span is an
// exact min/max, so a lane starts its chunk from the min/max of all earlier
// boundaries (a wave scan) and sees the spans the sequential loop sees.
// Column terms are summed per lane and reduced in a fixed order.
__global__ __launch_bounds__(256) void env_map_heavy(const float* thphi, SceneK s, Cam7 cam, int W,
                                                     float offset, int lr, const float* area, const float* view,
                                                     uint8_t* img, const uint32_t* heavy, const unsigned* nheavy) {
    const unsigned lane = threadIdx.x & 63;
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned nwaves = (gridDim.x * blockDim.x) >> 6;
    const unsigned n = *nheavy;
    const int rows = s.rows, cols = s.cols;
    const float4* table = reinterpret_cast<const float4*>(s.table);
    for (unsigned e = wave; e < n; e += nwaves) {
        const uint32_t ij = heavy[e];
        const int i = (int)(ij / (uint32_t)W), j = (int)(ij % (uint32_t)W);
        EnvPix E;
        env_setup(thphi, W, i, j, offset, rows, E);
        float cx = 0.f, cy = 0.f, cz = 0.f, cw = 0.f;
        if (E.pixNum < cols) {
            const int nb = E.pixNum + 1;   // boundaries 0..pixNum
            const int per = (nb + 63) / 64;
            const int lo = min((int)lane * per, nb), hi = min(lo + per, nb);
            float mx = -100, mn = cols + 1;
            for (int b = lo; b < hi; b++) env_boundary(E, (E.pixMin + b) * E.pixSize, mn, mx);
            // exclusive scan of (mn, mx) over lanes
            float pmn = mn, pmx = mx;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                float omn = __shfl_up(pmn, off), omx = __shfl_up(pmx, off);
                if ((int)lane >= off) {
                    pmn = fminf(pmn, omn);
                    pmx = fmaxf(pmx, omx);
                }
            }
            float emn = __shfl_up(pmn, 1), emx = __shfl_up(pmx, 1);
            if (lane == 0) {
                emn = cols + 1;
                emx = -100;
            }
            mn = emn;
            mx = emx;
            for (int b = lo; b < hi; b++) {
                env_boundary(E, (E.pixMin + b) * E.pixSize, mn, mx);
                if (b > 0) {
                    float4 d = env_column(table, rows, cols, E.pixSize, (E.pixMin + b) % cols, mn, mx);
                    cx += d.x;
                    cy += d.y;
                    cz += d.z;
                    cw += d.w;
                }
            }
        } else {
            int thMaxPix = min(max(cvt_i32(E.thMax / E.pixSize), 0), rows - 1) * cols;
            int thMinPix = min(max(cvt_i32(E.thMin / E.pixSize), 0), rows - 1) * cols;
            for (int q = lane; q < E.pixNum; q += 64) {
                float4 mxc = table[thMaxPix + (E.pixMin + q) % cols];
                float4 mnc = table[thMinPix + (E.pixMin + q) % cols];
                cx += mxc.x - mnc.x;
                cy += mxc.y - mnc.y;
                cz += mxc.z - mnc.z;
                cw += mxc.w - mnc.w;
            }
        }
        cx = wave_xor_sum(cx);
        cy = wave_xor_sum(cy);
        cz = wave_xor_sum(cz);
        cw = wave_xor_sum(cw);
        if (lane == 0)
            reinterpret_cast<uchar4*>(img)[ij] = env_finish(E, cx, cy, cz, cw, W, lr, cam, view, area[ij]);
    }
}

// addStarsAndBackground (MakeImageKernel.cu:837-853)
__global__ __launch_bounds__(256) void composite(const uint8_t* img_star, const uint8_t* img_map, int64_t n,
                                                 int star, int map, uint8_t* out) {
    int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    uchar4 o;
    if (star && map) {
        uchar4 s = reinterpret_cast<const uchar4*>(img_star)[q];
        uchar4 b = reinterpret_cast<const uchar4*>(img_map)[q];
        float p = 1.f;
        float sx = (float)s.x * s.x, sy = (float)s.y * s.y, sz = (float)s.z * s.z;
        float bx = (float)b.x * b.x, by = (float)b.y * b.y, bz = (float)b.z * b.z;
        o.x = (uint8_t)min(cvt_i32(sqrtf(p * sx + (2.f - p) * bx)), 255);
        o.y = (uint8_t)min(cvt_i32(sqrtf(p * sy + (2.f - p) * by)), 255);
        o.z = (uint8_t)min(cvt_i32(sqrtf(p * sz + (2.f - p) * bz)), 255);
        o.w = 255;
    } else if (map) {
        o = reinterpret_cast<const uchar4*>(img_map)[q];
    } else if (star) {
        o = reinterpret_cast<const uchar4*>(img_star)[q];
    } else {
        o = make_uchar4(0, 0, 0, 255);
    }
    reinterpret_cast<uchar4*>(out)[q] = o;
}

// Summed-area sky table (MakeImageKernel.cu:1312-1331): one thread per
// column, prefix sums down the rows in the reference's order.
__global__ __launch_bounds__(64) void sum_table(const uint8_t* sky, int rows, int cols, float* table) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= cols) return;
    float pxsz = BT_FPI2 / (1.f * cols);
    float prev0 = 0.f, prev1 = 0.f, prev2 = 0.f, prev3 = 0.f;
    float phi[4] = {pxsz * q, pxsz * q, pxsz * (q + 1), pxsz * (q + 1)};
    for (int p = 0; p < rows; p++) {
        float theta[4] = {pxsz * (p + 1), pxsz * p, pxsz * p, pxsz * (p + 1)};
        float area = calc_area4(theta, phi);
        const uint8_t* px = sky + 3 * ((size_t)p * cols + q);
        prev0 += f_pow((float)px[0], 2.2f) * area;
        prev1 += f_pow((float)px[1], 2.2f) * area;
        prev2 += f_pow((float)px[2], 2.2f) * area;
        prev3 += area;
        reinterpret_cast<float4*>(table)[(size_t)p * cols + q] = make_float4(prev0, prev1, prev2, prev3);
    }
}

// makeGrid (MakeImageKernel.cu:369-384) with hashLookup (kernel.cuh:705-733), G == 1.
__global__ void make_grid(const float* ht, const int32_t* ot, const int32_t* tag, int hw, int ow, int GN, int GM,
                          int sym, int GN1, float* grid) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= GN || j >= GM) return;
    int ix = (i + ow) % ow, iy = (j + ow) % ow;
    int ax = (i + hw) % hw + ot[2 * (ix * ow + iy)];
    int ay = (j + hw) % hw + ot[2 * (ix * ow + iy) + 1];
    int hx = (ax + hw) % hw, hy = (ay + hw) % hw;
    int sidx = hx * hw + hy;
    float vx = -2.f, vy = -2.f;
    if (tag[2 * sidx] == i && tag[2 * sidx + 1] == j) {
        vx = ht[2 * sidx];
        vy = ht[2 * sidx + 1];
    }
    float2* g2 = reinterpret_cast<float2*>(grid);
    g2[(size_t)i * GM + j] = make_float2(vx, vy);
    if (sym == 1) {
        if (vx != -1 && vx != -2) vx = BT_FPI - vx;
        int x = GN1 - i - 1;
        g2[(size_t)x * GM + j] = make_float2(vx, vy);
    }
}

// Dense FP32 grid from a device-built FP64 grid: the values saveAsGpuHash
// would store (Grid.h:645-650), without the hash round trip.
__global__ void grid_to_f32(const double* cel, int GN, int GM, int sym, int GN1, float* grid) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= GN || j >= GM) return;
    size_t o = (size_t)i * GM + j;
    float vx = (float)cel[2 * o], vy = (float)cel[2 * o + 1];
    float2* g2 = reinterpret_cast<float2*>(grid);
    g2[o] = make_float2(vx, vy);
    if (sym == 1) {
        if (vx != -1 && vx != -2) vx = BT_FPI - vx;
        int x = GN1 - i - 1;
        g2[(size_t)x * GM + j] = make_float2(vx, vy);
    }
}

__global__ void zero_u32_r(unsigned* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

dim3 grid2d(int W, int H) { return dim3((unsigned)((W + 15) / 16), (unsigned)((H + 15) / 16)); }

}  // namespace

int render_sum_table(Context* c, const uint8_t* d_sky, int rows, int cols, float* d_table) {
    KTimer t(c, "sum_table");
    sum_table<<<(unsigned)((cols + 63) / 64), 64, 0, c->stream>>>(d_sky, rows, cols, d_table);
    BT_HIP(hipGetLastError());
    return BT_OK;
}

int render_grid_from_psh(Context* c, const float* d_ht, const int32_t* d_ot, const int32_t* d_tag, int hw,
                         int ow, int GN, int GM, float* d_grid, int* GN1_out) {
    int sym = (float)GM / (float)GN > 3 ? 1 : 0;
    int GN1 = sym ? 2 * GN - 1 : GN;
    KTimer t(c, "make_grid");
    make_grid<<<grid2d(GM, GN), dim3(16, 16), 0, c->stream>>>(d_ht, d_ot, d_tag, hw, ow, GN, GM, sym, GN1, d_grid);
    BT_HIP(hipGetLastError());
    *GN1_out = GN1;
    return BT_OK;
}

int render_grid_from_cel(Context* c, const double* d_cel, int GN, int GM, float* d_grid, int* GN1_out) {
    int sym = (float)GM / (float)GN > 3 ? 1 : 0;
    int GN1 = sym ? 2 * GN - 1 : GN;
    KTimer t(c, "grid_to_f32");
    grid_to_f32<<<grid2d(GM, GN), dim3(16, 16), 0, c->stream>>>(d_cel, GN, GM, sym, GN1, d_grid);
    BT_HIP(hipGetLastError());
    *GN1_out = GN1;
    return BT_OK;
}

int view_pixel_areas(Context* c, const float* view, int W, int H, float* out) {
    view_pixel_areas<<<grid2d(W, H), dim3(16, 16), 0, c->stream>>>(view, W, H, out);
    BT_HIP(hipGetLastError());
    return BT_OK;
}

size_t frame_bytes(int w, int h, int step) {
    const size_t NP = (size_t)w * h, NC = (size_t)(w + 1) * (h + 1);
    const size_t fw = 2 * (size_t)step + 1;
    // thphi, gap | bh, area x2, star and map images, per-tap star light,
    // star chunk partials, star queues (6), odd pixels and lensing (3), env queues (2),
    // lit-pixel flags, tree ranges (NP / 4 + 65536 of 8 B), BGRA output
    return NC * (8 + 4) + NP * (1 + 4 + 4 + 4 + 4 + 12 * fw * fw + 12 * fw * fw + 28 + 12 + 8 + 2 + 4 + 1) + 4096 +
           65536 * 8;
}

int render_frame(Context* c, const GridF& g, const SceneK& s, const ViewK& v, const float* cam7,
                 const bt_render_flags& f, int step, uint8_t* d_out, FrameDebug* dbg, const Blend* blend,
                 const TrailFrame* tf) {
    const int W = v.W, H = v.H;
    const size_t NP = (size_t)W * H, NC = (size_t)(W + 1) * (H + 1);
    if (step != 1 && step != 2 && step != 3)
        return fail(BT_ERR_INVALID, "render: star splat step must be 1, 2 or 3");
    // the reference's frame offset is 2*pi*q/(...) >= 0 (MakeImageKernel.cu:1413); the sky
    // lookups assume phi + offset >= 0
    if (!(f.offset >= 0.f) || !std::isfinite(f.offset) || !std::isfinite(f.hor) || !std::isfinite(f.ver))
        return fail(BT_ERR_INVALID, "render: offset must be finite and >= 0; hor/ver finite");
    if (W <= 0 || H <= 0 || g.GN < 2 || g.GM < 2 || s.rows <= 0 || s.cols <= 0)
        return fail(BT_ERR_INVALID, "render: empty frame, grid or sky");
    if (tf && (tf->W != W || tf->H != H))
        return fail(BT_ERR_INVALID, "render: the trail state was made for another frame size");
    if (tf && tf->k.S != s.S) return fail(BT_ERR_INVALID, "render: the trail state was made for another star set");
    if (blend && (blend->b.GN != g.GN || blend->b.GM != g.GM || blend->b.gridlvl != g.gridlvl || !blend->border ||
                  !(blend->alpha >= 0.f && blend->alpha <= 1.f)))
        return fail(BT_ERR_INVALID, "render: blended grids must have equal sizes and levels; alpha in [0, 1]");
    void* p;
    int rc;
    if ((rc = c->buffer("r_thphi", sizeof(float) * 2 * NC, &p))) return rc;
    float* thphi = (float*)p;
    if ((rc = c->buffer("r_gap", sizeof(int) * NC, &p))) return rc;
    int* gap = (int*)p;
    if ((rc = c->buffer("r_bh", NP, &p))) return rc;
    uint8_t* bh = (uint8_t*)p;
    if ((rc = c->buffer("r_area", sizeof(float) * NP, &p))) return rc;
    float* area = (float*)p;
    if ((rc = c->buffer("r_area_s", sizeof(float) * NP, &p))) return rc;
    float* area_s = (float*)p;
    if ((rc = c->buffer("r_img_star", 4 * NP, &p))) return rc;
    uint8_t* img_star = (uint8_t*)p;
    if ((rc = c->buffer("r_img_map", 4 * NP, &p))) return rc;
    uint8_t* img_map = (uint8_t*)p;

    Cam7 cam;
    for (int q = 0; q < 7; q++) cam.v[q] = cam7[q];
    {
        // blackbody colours per index, made once per context
        bool fresh = c->scratch.find("r_star_colours") == c->scratch.end();
        if ((rc = c->buffer("r_star_colours", sizeof(float) * 6 * 391, &p))) return rc;
        cam.colours = (const float*)p;
        if (fresh) star_colours<<<2, 256, 0, c->stream>>>((float*)p);
    }
    GridV gv{g.grid, g.GN, g.GM};
    const dim3 blk(16, 16);
    if (v.pix_area) {
        cam.pix_area = v.pix_area;
    } else {
        if ((rc = c->buffer("r_pix_area", sizeof(float) * NP, &p))) return rc;
        cam.pix_area = (const float*)p;
        view_pixel_areas<<<grid2d(W, H), blk, 0, c->stream>>>(v.view, W, H, (float*)p);
    }
    if (!blend) {
        KTimer t(c, "pix_interp");
        pix_interp<<<grid2d(W + 1, H + 1), blk, 0, c->stream>>>(gv, g.gridlvl, v.view, W, H, f.hor, f.ver, thphi, gap);
    } else {
        GridV gb{blend->b.grid, blend->b.GN, blend->b.GM};
        float2* border = reinterpret_cast<float2*>(blend->border);
        {
            // findBhCenter, findBhBorders, smoothBorder x 4 (MakeImageKernel.cu:1429-1442)
            KTimer t(c, "blend_border");
            if ((rc = c->buffer("r_blend_tmp", sizeof(float2) * (2 + 2 * kAngleNum) + 64, &p))) return rc;
            float2* tmp = (float2*)p;
            int* ext = reinterpret_cast<int*>(tmp + 2 + 2 * kAngleNum);
            bh_extent_init<<<1, 64, 0, c->stream>>>(ext);
            bh_extent<<<grid2d(g.GM, g.GN), blk, 0, c->stream>>>(gv, gb, ext);
            const unsigned nb = (2 * kAngleNum + 255) / 256;
            bh_borders<<<nb, 256, 0, c->stream>>>(gv, gb, ext, border);
            smooth_border<<<nb, 256, 0, c->stream>>>(border, tmp);
            smooth_border<<<nb, 256, 0, c->stream>>>(tmp, border);
            smooth_border<<<nb, 256, 0, c->stream>>>(border, tmp);
            smooth_border<<<nb, 256, 0, c->stream>>>(tmp, border);
        }
        KTimer t(c, "pix_interp");
        pix_interp_blend<<<grid2d(W + 1, H + 1), blk, 0, c->stream>>>(gv, gb, g.gridlvl, border, blend->alpha, v.view,
                                                                      W, H, f.hor, f.ver, thphi, gap);
    }
    {
        KTimer t(c, "bh_area");
        bh_area<<<grid2d(W, H), blk, 0, c->stream>>>(thphi, W, H, bh, area);
    }
    {
        KTimer t(c, "smooth");
        smooth_h<<<grid2d(W, H), blk, 0, c->stream>>>(area, bh, gap, W, H, area_s);
        smooth_v<<<grid2d(W, H), blk, 0, c->stream>>>(area_s, bh, gap, W, H, area);
    }
    BT_HIP(hipGetLastError());
    // The sky pass runs on the auxiliary stream beside the star pass: both
    // read the per-pixel stages made above and write their own image, and
    // each has latency-bound kernels (the lensed pixels' queues) that leave
    // the other room.  Joined before the composite.
    const bool overlap = f.star && f.map;
    uint32_t *qlane = nullptr, *qwave = nullptr;
    unsigned* nq = nullptr;
    if (f.map) {
        if ((rc = c->buffer("r_env_q", sizeof(uint32_t) * 2 * (NP + 64), &p))) return rc;
        qlane = (uint32_t*)p;
        qwave = (uint32_t*)p + NP + 64;
        if ((rc = c->buffer("r_env_qn", 64, &p))) return rc;
        nq = (unsigned*)p;
    }
    // distortEnvironmentMap: pixels with few sky columns inline and one lane
    // each; then the wide ones, one wave each
    auto sky_light = [&](hipStream_t es) {
        zero_u32_r<<<1, 64, 0, es>>>(nq, 2);
        KTimer t(c, "env_map", es);
        env_map<<<grid2d(W, H), blk, 0, es>>>(thphi, bh, s, cam, W, H, f.offset, f.lr, area, v.view, img_map, qlane,
                                              qwave, nq);
        env_map_lane<<<(unsigned)c->num_cus * 8, 256, 0, es>>>(thphi, s, cam, W, f.offset, f.lr, area, v.view,
                                                              img_map, qlane, nq);
    };
    auto sky_heavy = [&](hipStream_t es) {
        KTimer t(c, "env_heavy", es);
        env_map_heavy<<<(unsigned)c->num_cus * 8, 256, 0, es>>>(thphi, s, cam, W, f.offset, f.lr, area, v.view,
                                                               img_map, qwave, nq + 1);
    };
    auto join = [&]() -> hipError_t {
        const hipError_t e = hipEventRecord(c->ev_join, c->aux);
        const hipError_t w = hipStreamWaitEvent(c->stream, c->ev_join, 0);
        return e != hipSuccess ? e : w;
    };
    if (overlap) {
        BT_HIP(hipEventRecord(c->ev_fork, c->stream));
        BT_HIP(hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    }
    if (f.star) {
      // the star pass; on an error the streams are joined before returning
      const int src = [&]() -> int {
        const int fw = 2 * step + 1;
        if ((rc = c->buffer("r_taps", sizeof(float) * 3 * fw * fw * NP, &p))) return rc;
        float* taps = (float*)p;
        StarQueue Q;
        if ((rc = c->buffer("r_star_lens", sizeof(float) * 3 * (NP + 64), &p))) return rc;
        Q.lens = (float*)p;
        if ((rc = c->buffer("r_star_odd", sizeof(uint32_t) * (NP + 64), &p))) return rc;
        Q.odd = (uint32_t*)p;
        if ((rc = c->buffer("r_star_q", sizeof(uint32_t) * 6 * (NP + 64), &p))) return rc;
        Q.pix = (uint32_t*)p;
        Q.cnt = Q.pix + NP + 64;
        Q.cstart = Q.cnt + NP + 64;
        Q.owner = Q.cstart + NP + 64;
        Q.nnode = Q.owner + NP + 64;
        Q.rbase = Q.nnode + NP + 64;
        // tree pixels are the few whose quad crosses phi = 2 pi
        Q.rcap = (unsigned)(NP / 4 + 65536);
        if ((rc = c->buffer("r_star_ranges", sizeof(int2) * (size_t)Q.rcap, &p))) return rc;
        Q.ranges = (int2*)p;
        if ((rc = c->buffer("r_star_qn", 64, &p))) return rc;
        Q.ctr = (unsigned*)p;
        if ((rc = c->buffer("r_star_lit", NP, &p))) return rc;
        Q.lit = (uint8_t*)p;
        const int NT = fw * fw * 3;
        // chunk indices beyond cap are never used: a pixel whose chunks do
        // not fit is traced by its own lane in star_map_light
        Q.cap = (unsigned)(NP + 64);
        if ((rc = c->buffer("r_star_part", sizeof(float) * NT * (size_t)Q.cap, &p))) return rc;
        Q.partial = (float*)p;
        zero_u32_r<<<1, 64, 0, c->stream>>>(Q.ctr, 4);
        const unsigned nwave = (unsigned)c->num_cus * 16, nred = (unsigned)c->num_cus * 4;
        if (tf) {
            // with both passes on, the trails go first on the auxiliary stream:
            // the main stream waits there (after the odd chunks) before the
            // chunk reduction, and so before star_sum reads the trails
            const hipStream_t ts = overlap ? c->aux : c->stream;
            KTimer t(c, "trails", ts);
            const TrailK& T = tf->k;
            trail_clear<<<(unsigned)c->num_cus * 4, 256, 0, ts>>>(T, tf->first ? 1 : 0, 3 * NP);
            trail_grad<<<grid2d(W + 1, H + 1), blk, 0, ts>>>(thphi, W, H, T.grad);
            trail_hits<<<grid2d(W, H), blk, 0, ts>>>(thphi, bh, s, cam, W, H, f.offset, f.lr, area, v.view, T);
            trail_cache<<<(unsigned)((T.S + 255) / 256), 256, 0, ts>>>(T, W);
            trail_lines<<<(unsigned)(((size_t)T.S * kTrailHits + 255) / 256), 256, 0, ts>>>(T, cam, T.grad, W, H);
            BT_HIP(hipGetLastError());
        }
#define BT_STAR_LAUNCH(S)                                                                                    \
    {                                                                                                       \
        {                                                                                                   \
            KTimer t(c, "star_light");                                                                      \
            star_map_light<S><<<grid2d(W, H), blk, 0, c->stream>>>(thphi, bh, s, cam, W, H, f.offset, f.lr,  \
                                                                  area, v.view, taps, Q);                   \
        }                                                                                                   \
        {                                                                                                   \
            KTimer t0(c, "star_odd");                                                                       \
            star_map_odd<S><<<nred, 256, 0, c->stream>>>(thphi, bh, s, cam, W, H, f.offset, f.lr, area,     \
                                                         v.view, taps, Q);                                  \
        }                                                                                                   \
        KTimer t(c, "star_heavy");                                                                          \
        {                                                                                                   \
            KTimer t1(c, "star_tree");                                                                      \
            star_tree_ranges<<<nred, 256, 0, c->stream>>>(thphi, W, f.offset, s, cam, Q);                   \
        }                                                                                                   \
        if (overlap) {                                                                                      \
            /* the odd pixels' chunks (few, long, latency-bound) on the auxiliary stream, between */       \
            /* the sky pass's light and heavy kernels, beside the other chunks */                           \
            BT_HIP(hipEventRecord(c->ev_hand, c->stream));                                                  \
            sky_light(c->aux);                                                                              \
            BT_HIP(hipStreamWaitEvent(c->aux, c->ev_hand, 0));                                              \
            star_map_chunk<S, true><<<nred, 256, 0, c->aux>>>(thphi, bh, s, cam, W, H, f.offset, f.lr,      \
                                                              area, v.view, taps, Q);                       \
            BT_HIP(hipEventRecord(c->ev_back, c->aux));                                                     \
            sky_heavy(c->aux);                                                                              \
        }                                                                                                   \
        {                                                                                                   \
            KTimer t2(c, "star_chunk");                                                                     \
            star_map_chunk<S, false><<<nwave, 256, 0, c->stream>>>(thphi, bh, s, cam, W, H, f.offset, f.lr,  \
                                                                   area, v.view, taps, Q);                  \
            if (!overlap)                                                                                   \
                star_map_chunk<S, true><<<nred, 256, 0, c->stream>>>(thphi, bh, s, cam, W, H, f.offset,     \
                                                                     f.lr, area, v.view, taps, Q);          \
        }                                                                                                   \
        if (overlap) BT_HIP(hipStreamWaitEvent(c->stream, c->ev_back, 0));                                  \
        KTimer t3(c, "star_reduce");                                                                        \
        star_chunk_reduce<S><<<nwave, 256, 0, c->stream>>>(W, H, taps, Q);                                  \
    }
        if (step == 1)
            BT_STAR_LAUNCH(1)
        else if (step == 2)
            BT_STAR_LAUNCH(2)
        else
            BT_STAR_LAUNCH(3)
#undef BT_STAR_LAUNCH
        BT_HIP(hipGetLastError());
        const bool diffract = tf && tf->sprite;
        float* light = nullptr;
        if (diffract) {
            if ((rc = c->buffer("r_diff_light", sizeof(float) * 3 * NP * 2, &p))) return rc;
            light = (float*)p;
        }
        {
            KTimer t(c, "star_sum");
            float* ostar = diffract ? light : (dbg ? dbg->star : nullptr);
            const float* trail = tf ? tf->k.trail : nullptr;
            if (step == 1)
                star_sum<1><<<grid2d(W, H), blk, 0, c->stream>>>(taps, Q.lit, W, H, img_star, ostar, trail);
            else if (step == 2)
                star_sum<2><<<grid2d(W, H), blk, 0, c->stream>>>(taps, Q.lit, W, H, img_star, ostar, trail);
            else
                star_sum<3><<<grid2d(W, H), blk, 0, c->stream>>>(taps, Q.lit, W, H, img_star, ostar, trail);
        }
        BT_HIP(hipGetLastError());
        if (diffract) {
            KTimer t(c, "diffraction");
            float* acc = light + 3 * NP;
            if ((rc = c->buffer("r_diff_rows", sizeof(uint32_t) * 4 * NP + sizeof(int) * (size_t)H, &p))) return rc;
            uint32_t* cols = (uint32_t*)p;
            float* scaled = reinterpret_cast<float*>(cols + NP);
            int* nrow = reinterpret_cast<int*>(cols + 4 * NP);
            const unsigned b3 = (unsigned)((3 * NP + 255) / 256), b1 = (unsigned)((NP + 255) / 256);
            diff_rows<<<(unsigned)((H + 3) / 4), 256, 0, c->stream>>>(light, W, H, cols, scaled, nrow);
            diff_gather<<<grid2d(W, H), blk, 0, c->stream>>>(light, acc, cols, scaled, nrow, tf->sprite, tf->fs, W,
                                                             H);
            make_pix_all<<<b1, 256, 0, c->stream>>>(acc, (int64_t)NP, img_star);
            if (dbg && dbg->star) copy_f32<<<b3, 256, 0, c->stream>>>(acc, dbg->star, (int64_t)(3 * NP));
            BT_HIP(hipGetLastError());
        }
        return BT_OK;
      }();
      if (src) {
          if (overlap) (void)join();
          return src;
      }
    }
    if (overlap) {
        BT_HIP(join());
    } else if (f.map) {
        sky_light(c->stream);
        sky_heavy(c->stream);
    }
    BT_HIP(hipGetLastError());
    {
        KTimer t(c, "composite");
        composite<<<(unsigned)((NP + 255) / 256), 256, 0, c->stream>>>(img_star, img_map, (int64_t)NP, f.star, f.map,
                                                                     d_out);
    }
    BT_HIP(hipGetLastError());
    if (dbg) {
        if (dbg->thphi) BT_HIP(hipMemcpyAsync(dbg->thphi, thphi, sizeof(float) * 2 * NC, hipMemcpyDeviceToDevice, c->stream));
        if (dbg->gap) BT_HIP(hipMemcpyAsync(dbg->gap, gap, sizeof(int) * NC, hipMemcpyDeviceToDevice, c->stream));
        if (dbg->bh) BT_HIP(hipMemcpyAsync(dbg->bh, bh, NP, hipMemcpyDeviceToDevice, c->stream));
        if (dbg->area) BT_HIP(hipMemcpyAsync(dbg->area, area, sizeof(float) * NP, hipMemcpyDeviceToDevice, c->stream));
    }
    return BT_OK;
}

}  // namespace bt
