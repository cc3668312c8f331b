// render_dev.h -- device helpers of the render pass, FP32 as in the
// reference (kernel.cuh), compiled with -ffp-contract=off.
//
// Transcendentals are evaluated in FP64 and rounded to FP32, so they are
// correctly rounded (up to ~2^-28 double-rounding cases) and independent of
// any one float libm; the reference used CUDA 10's sinf/cosf/atanf/powf/
// expf/log10f/exp10f, whose exact last bits are not reproducible here.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bt {
namespace rdr {

#define BT_FPI2 6.283185307179586476f   // kernel.cuh:24
#define BT_FPI 3.141592653589793238f    // kernel.cuh:25
#define BT_SQRT2PI 2.506628274631f      // kernel.cuh:23

struct F2 {
    float x, y;
};

// float -> int as CUDA's cvt.rzi.s32.f32: truncate, saturate, NaN -> 0
// (LLVM's fptosi is poison for NaN and out-of-range inputs).
__device__ __forceinline__ int cvt_i32(float x) {
    if (!(x == x)) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return -2147483647 - 1;
    return (int)x;
}

__device__ __forceinline__ float f_sin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float f_cos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float f_atan(float x) { return (float)atan((double)x); }
__device__ __forceinline__ float f_atan2(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ __forceinline__ float f_acos(float x) { return (float)acos((double)x); }
__device__ __forceinline__ float f_pow(float x, float y) { return (float)pow((double)x, (double)y); }
__device__ __forceinline__ float f_exp(float x) { return (float)exp((double)x); }
__device__ __forceinline__ float f_log10(float x) { return (float)log10((double)x); }
__device__ __forceinline__ float f_exp10(float x) { return (float)pow(10.0, (double)x); }

// kernel.cuh:586-610
__device__ inline float calc_area4(const float* t, const float* p) {
    float x[4], y[4], z[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        float sint = f_sin(t[q]);
        x[q] = sint * f_cos(p[q]);
        y[q] = sint * f_sin(p[q]);
        z[q] = f_cos(t[q]);
    }
    float dotpr1 = 1.f;
    dotpr1 += x[0] * x[2] + y[0] * y[2] + z[0] * z[2];
    float dotpr2 = dotpr1;
    dotpr1 += x[2] * x[1] + y[2] * y[1] + z[2] * z[1];
    dotpr1 += x[0] * x[1] + y[0] * y[1] + z[0] * z[1];
    dotpr2 += x[0] * x[3] + y[0] * y[3] + z[0] * z[3];
    dotpr2 += x[2] * x[3] + y[2] * y[3] + z[2] * z[3];
    float triprod1 = fabsf(x[0] * (y[1] * z[2] - y[2] * z[1]) - y[0] * (x[1] * z[2] - x[2] * z[1]) +
                           z[0] * (x[1] * y[2] - x[2] * y[1]));
    float triprod2 = fabsf(x[0] * (y[2] * z[3] - y[3] * z[2]) - y[0] * (x[2] * z[3] - x[3] * z[2]) +
                           z[0] * (x[2] * y[3] - x[3] * y[2]));
    return 2.f * (f_atan(triprod1 / dotpr1) + f_atan(triprod2 / dotpr2));
}

// kernel.cuh:612-649
__device__ inline float calc_area3(float t0, float t1, float t2, float p0, float p1, float p2) {
    const float t[3] = {t0, t1, t2}, p[3] = {p0, p1, p2};
    float xi[3], yi[3], zi[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        float sint = f_sin(t[q]);
        xi[q] = sint * f_cos(p[q]);
        yi[q] = sint * f_sin(p[q]);
        zi[q] = f_cos(t[q]);
    }
    float dot01 = xi[0] * xi[1] + yi[0] * yi[1] + zi[0] * zi[1];
    float dot02 = xi[0] * xi[2] + yi[0] * yi[2] + zi[0] * zi[2];
    float dot12 = xi[2] * xi[1] + yi[2] * yi[1] + zi[2] * zi[1];
    float x[3] = {xi[0], xi[1], xi[2]};
    float y[3] = {yi[0], yi[1], yi[2]};
    float z[3] = {zi[0], zi[1], zi[2]};
    if (dot01 < dot02 && dot01 < dot12) {
        x[0] = xi[2]; x[1] = xi[0]; x[2] = xi[1];
        y[0] = yi[2]; y[1] = yi[0]; y[2] = yi[1];
        z[0] = zi[2]; z[1] = zi[0]; z[2] = zi[1];
    } else if (dot02 < dot12 && dot02 < dot01) {
        x[0] = xi[1]; x[1] = xi[2]; x[2] = xi[0];
        y[0] = yi[1]; y[1] = yi[2]; y[2] = yi[0];
        z[0] = zi[1]; z[1] = zi[2]; z[2] = zi[0];
    }
    float dotpr1 = 1.f;
    dotpr1 += x[0] * x[2] + y[0] * y[2] + z[0] * z[2];
    dotpr1 += x[2] * x[1] + y[2] * y[1] + z[2] * z[1];
    dotpr1 += x[0] * x[1] + y[0] * y[1] + z[0] * z[1];
    float triprod1 = fabsf(x[0] * (y[1] * z[2] - y[2] * z[1]) - y[0] * (x[1] * z[2] - x[2] * z[1]) +
                           z[0] * (x[1] * y[2] - x[2] * y[1]));
    return 2.f * (f_atan(triprod1 / dotpr1));
}

// kernel.cuh:274-295
__device__ __forceinline__ bool pi_check(float* p, float factor) {
    float factor1 = BT_FPI2 * (1.f - factor);
    bool check = false;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (p[q] > factor1) check = true;
    if (!check) return false;
    check = false;
    float factor2 = BT_FPI2 * factor;
#pragma unroll
    for (int q = 0; q < 4; q++)
        if (p[q] < factor2) {
            p[q] += BT_FPI2;
            check = true;
        }
    return check;
}

// kernel.cuh:303-322
__device__ __forceinline__ void pi_check_tot(F2* tp, float factor, int size) {
    float factor1 = BT_FPI2 * (1.f - factor);
    bool check = false;
    for (int q = 0; q < size; q++)
        if (tp[q].y > factor1) check = true;
    if (!check) return;
    float factor2 = BT_FPI2 * factor;
    for (int q = 0; q < size; q++)
        if (tp[q].y < factor2) tp[q].y += BT_FPI2;
}

// kernel.cuh:694-703
__device__ __forceinline__ void wrap_to_pi(float& thetaW, float& phiW) {
    thetaW = fmodf(thetaW, BT_FPI2);
    while (thetaW < 0.f) thetaW += BT_FPI2;
    if (thetaW > BT_FPI) {
        thetaW -= 2.f * (thetaW - BT_FPI);
        phiW += BT_FPI;
    }
    while (phiW < 0.f) phiW += BT_FPI2;
    phiW = fmodf(phiW, BT_FPI2);
}

// kernel.cuh:790-815 (tension = bias = 0)
__device__ __forceinline__ F2 hermite(float v, F2 x0, F2 x1, F2 x2, F2 x3) {
    float const v2 = v * v;
    float const v3 = v * v2;
    float const aa = (1.f + 0.f) * (1.f - 0.f) / 2.f;
    float const bb = (1.f - 0.f) * (1.f - 0.f) / 2.f;
    float const m0T = aa * (x1.x - x0.x) + bb * (x2.x - x1.x);
    float const m0P = aa * (x1.y - x0.y) + bb * (x2.y - x1.y);
    float const m1T = aa * (x2.x - x1.x) + bb * (x3.x - x2.x);
    float const m1P = aa * (x2.y - x1.y) + bb * (x3.y - x2.y);
    float const u0 = 2.f * v3 - 3.f * v2 + 1.f;
    float const u1 = v3 - 2.f * v2 + v;
    float const u2 = v3 - v2;
    float const u3 = -2.f * v3 + 3.f * v2;
    return F2{u0 * x1.x + u1 * m0T + u2 * m1T + u3 * x2.x, u0 * x1.y + u1 * m0P + u2 * m1P + u3 * x2.y};
}

// kernel.cuh:754-772
__device__ __forceinline__ F2 intersection(float ax, float ay, float bx, float by, float cx, float cy,
                                           float dx, float dy) {
    double a1 = by - ay;
    double b1 = ax - bx;
    double c1 = a1 * (ax) + b1 * (ay);
    double a2 = dy - cy;
    double b2 = cx - dx;
    double c2 = a2 * (cx) + b2 * (cy);
    double determinant = a1 * b2 - a2 * b1;
    if (determinant == 0) return F2{-1, -1};
    double x = (b2 * c1 - b1 * c2) / determinant;
    double y = (a1 * c2 - a2 * c1) / determinant;
    return F2{(float)x, (float)y};
}

#define BT_Pr .299f
#define BT_Pg .587f
#define BT_Pb .114f

// kernel.cuh:388-427
__device__ inline void rgb_to_hsp(float R, float G, float B, float& H, float& S, float& P) {
    P = sqrtf(R * R * BT_Pr + G * G * BT_Pg + B * B * BT_Pb);
    if (R == G && R == B) {
        H = 0.f;
        S = 0.f;
        return;
    }
    if (R >= G && R >= B) {
        if (B >= G) {
            H = 6.f / 6.f - 1.f / 6.f * (B - G) / (R - G);
            S = 1.f - G / R;
        } else {
            H = 0.f / 6.f + 1.f / 6.f * (G - B) / (R - B);
            S = 1.f - B / R;
        }
    } else if (G >= R && G >= B) {
        if (R >= B) {
            H = 2.f / 6.f - 1.f / 6.f * (R - B) / (G - B);
            S = 1.f - B / G;
        } else {
            H = 2.f / 6.f + 1.f / 6.f * (B - R) / (G - R);
            S = 1.f - R / G;
        }
    } else {
        if (G >= R) {
            H = 4.f / 6.f - 1.f / 6.f * (G - R) / (B - R);
            S = 1.f - R / B;
        } else {
            H = 4.f / 6.f + 1.f / 6.f * (R - G) / (B - G);
            S = 1.f - G / B;
        }
    }
}

// kernel.cuh:440-524
__device__ inline void hsp_to_rgb(float H, float S, float P, float& R, float& G, float& B) {
    float part, minOverMax = 1.f - S;
    if (minOverMax > 0.f) {
        if (H < 1.f / 6.f) {
            H = 6.f * H;
            part = 1.f + H * (1.f / minOverMax - 1.f);
            B = P / sqrtf(BT_Pr / minOverMax / minOverMax + BT_Pg * part * part + BT_Pb);
            R = B / minOverMax;
            G = B + H * (R - B);
        } else if (H < 2.f / 6.f) {
            H = 6.f * (-H + 2.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            B = P / sqrtf(BT_Pg / minOverMax / minOverMax + BT_Pr * part * part + BT_Pb);
            G = B / minOverMax;
            R = B + H * (G - B);
        } else if (H < 3.f / 6.f) {
            H = 6.f * (H - 2.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            R = P / sqrtf(BT_Pg / minOverMax / minOverMax + BT_Pb * part * part + BT_Pr);
            G = R / minOverMax;
            B = R + H * (G - R);
        } else if (H < 4.f / 6.f) {
            H = 6.f * (-H + 4.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            R = P / sqrtf(BT_Pb / minOverMax / minOverMax + BT_Pg * part * part + BT_Pr);
            B = R / minOverMax;
            G = R + H * (B - R);
        } else if (H < 5.f / 6.f) {
            H = 6.f * (H - 4.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            G = P / sqrtf(BT_Pb / minOverMax / minOverMax + BT_Pr * part * part + BT_Pg);
            B = G / minOverMax;
            R = G + H * (B - G);
        } else {
            H = 6.f * (-H + 6.f / 6.f);
            part = 1.f + H * (1.f / minOverMax - 1.f);
            G = P / sqrtf(BT_Pr / minOverMax / minOverMax + BT_Pb * part * part + BT_Pg);
            R = G / minOverMax;
            B = G + H * (R - G);
        }
    } else {
        if (H < 1.f / 6.f) {
            H = 6.f * (H);
            R = sqrtf(P * P / (BT_Pr + BT_Pg * H * H));
            G = R * H;
            B = 0.f;
        } else if (H < 2.f / 6.f) {
            H = 6.f * (-H + 2.f / 6.f);
            G = sqrtf(P * P / (BT_Pg + BT_Pr * H * H));
            R = G * H;
            B = 0.f;
        } else if (H < .5f) {
            H = 6.f * (H - 2.f / 6.f);
            G = sqrtf(P * P / (BT_Pg + BT_Pb * H * H));
            B = G * H;
            R = 0.f;
        } else if (H < 4.f / 6.f) {
            H = 6.f * (-H + 4.f / 6.f);
            B = sqrtf(P * P / (BT_Pb + BT_Pg * H * H));
            G = B * H;
            R = 0.f;
        } else if (H < 5.f / 6.f) {
            H = 6.f * (H - 4.f / 6.f);
            B = sqrtf(P * P / (BT_Pb + BT_Pr * H * H));
            R = B * H;
            G = 0.f;
        } else {
            H = 6.f * (-H + 1.f);
            R = sqrtf(P * P / (BT_Pr + BT_Pb * H * H));
            B = R * H;
            G = 0.f;
        }
    }
}

// kernel.cuh:348-365; cam = {speed, alpha, w, wbar, br, btheta, bphi}
__device__ inline float redshift(float theta, float phi, const float* cam) {
    float xCam = f_sin(theta) * f_cos(phi);
    float zCam = f_cos(theta);
    float yCam = f_sin(theta) * f_sin(phi);
    float part = (1.f - cam[0] * yCam);
    float betaPart = sqrtf(1.f - cam[0] * cam[0]) / part;
    float xFido = -sqrtf(1.f - cam[0] * cam[0]) * xCam / part;
    float zFido = -sqrtf(1.f - cam[0] * cam[0]) * zCam / part;
    float yFido = (-yCam + cam[0]) / part;
    float k = sqrtf(1.f - cam[5] * cam[5]);
    float phiFido = -xFido * cam[4] / k + cam[6] * yFido + cam[6] * cam[5] / k * zFido;
    float eF = 1.f / (cam[1] + cam[2] * cam[3] * phiFido);
    float b = eF * cam[3] * phiFido;
    (void)zFido;
    return 1.f / (betaPart * (1.f - b * cam[2]) / cam[1]);
}

// kernel.cuh:245-266
__device__ __forceinline__ bool cross_ok(float t_a, float t_b, float p_a, float p_b, float st, float sp,
                                         int sgn) {
    float c1t = (float)sgn * (t_a - t_b);
    float c1p = (float)sgn * (p_a - p_b);
    float c2t = sgn ? st - t_b : st - t_a;
    float c2p = sgn ? sp - p_b : sp - p_a;
    return (c1t * c2p - c2t * c1p) > 0;
}
__device__ __forceinline__ bool star_in_polygon(const float* t, const float* p, float st, float sp,
                                                int sgn) {
    return cross_ok(t[0], t[1], p[0], p[1], st, sp, sgn) && cross_ok(t[1], t[2], p[1], p[2], st, sp, sgn) &&
           cross_ok(t[2], t[3], p[2], p[3], st, sp, sgn) && cross_ok(t[3], t[0], p[3], p[0], st, sp, sgn);
}

// kernel.cuh:112-200
__device__ inline void interpolate_star(float t0, float t1, float t2, float t3, float p0, float p1,
                                        float p2, float p3, float& start, float& starp, int sgn) {
    float error = 0.00001f;
    float midT = (t0 + t1 + t2 + t3) * .25f;
    float midP = (p0 + p1 + p2 + p3) * .25f;
    float starInPixY = 0.5f;
    float starInPixX = 0.5f;
    float perc = 0.5f;
#pragma unroll 1
    for (int q = 0; q < 10; q++) {
        if ((fabsf(start - midT) < error) && (fabsf(starp - midP) < error)) break;
        float half01T = (t0 + t1) * .5f;
        float half23T = (t2 + t3) * .5f;
        float half12T = (t2 + t1) * .5f;
        float half03T = (t0 + t3) * .5f;
        float half01P = (p0 + p1) * .5f;
        float half23P = (p2 + p3) * .5f;
        float half12P = (p2 + p1) * .5f;
        float half03P = (p0 + p3) * .5f;
        float line01to23T = half23T - half01T;
        float line03to12T = half12T - half03T;
        float line01to23P = half23P - half01P;
        float line03to12P = half12P - half03P;
        float line01toStarT = start - half01T;
        float line03toStarT = start - half03T;
        float line01toStarP = starp - half01P;
        float line03toStarP = starp - half03P;
        int a = (((line03to12T * line03toStarP) - (line03toStarT * line03to12P)) > 0.f) ? 1 : -1;
        int b = (((line01to23T * line01toStarP) - (line01toStarT * line01to23P)) > 0.f) ? 1 : -1;
        perc *= 0.5f;
        if (sgn * a > 0) {
            if (sgn * b > 0) {
                t2 = half12T; t0 = half01T; t3 = midT;
                p2 = half12P; p0 = half01P; p3 = midP;
                starInPixX -= perc; starInPixY -= perc;
            } else {
                t2 = midT; t1 = half01T; t3 = half03T;
                p2 = midP; p1 = half01P; p3 = half03P;
                starInPixX -= perc; starInPixY += perc;
            }
        } else {
            if (sgn * b > 0) {
                t1 = half12T; t3 = half23T; t0 = midT;
                p1 = half12P; p3 = half23P; p0 = midP;
                starInPixX += perc; starInPixY -= perc;
            } else {
                t0 = half03T; t1 = midT; t2 = half23T;
                p0 = half03P; p1 = midP; p2 = half23P;
                starInPixX += perc; starInPixY += perc;
            }
        }
        midT = (t0 + t1 + t2 + t3) * .25f;
        midP = (p0 + p1 + p2 + p3) * .25f;
    }
    start = starInPixY;
    starp = starInPixX;
}

// kernel.cuh:335-338.  FP32 expf (within 1 ulp): the star taps' sums are
// compared with a relative tolerance, and this is evaluated 9-49 times per
// star.
__device__ __forceinline__ float gaussian(float dist) {
    float sigma = 1.f / 2.5f;
    return expf(-.5f * dist / (sigma * sigma)) * 1.f / (sigma * BT_SQRT2PI);
}

}  // namespace rdr
}  // namespace bt
