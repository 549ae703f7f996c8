// Boys function F_0..F_L(T) on the device, FP32 and FP64.
//
// Method (both precisions): a Taylor table around equally spaced nodes T_k,
//   F_m(T_k + h) = sum_j F_{m+j}(T_k) (-h)^j / j!      (dF_m/dT = -F_{m+1}),
// for T below a switch point, the asymptotic form above it:
//   F_0 = sqrt(pi/T)/2,  F_{n+1} = ((2n+1) F_n - e^{-T}) / (2T).
// For L >= 2 only F_L comes from the table; the lower orders follow by the
// stable downward recursion F_{n-1} = (2T F_n + e^{-T}) / (2n-1), with
// e^{-T} = e^{-T_k} e^{-h} (table value times a Taylor polynomial), so no
// transcendental is evaluated below the switch.  The reference uses the same
// node spacing and switch for FP32 (gpu/boys.py:14-47, T_k = k/4, 36).
#pragma once
#include "wb_common.h"

namespace wb {

// FP32: nodes k/4, k = 0..144; row = F_0..F_12, e^{-T_k} (hi, 14 values),
// then their FP32 rounding remainders (lo, 14 values), padded to 36 floats:
// the table sits in LDS and lanes of a wave read neighbouring nodes; a
// 36-float stride starts consecutive rows 4 banks apart, so 8 consecutive
// nodes are read without bank conflicts (a 32-float stride would put every
// row on the same banks)
constexpr int B32_NODES = 145;
constexpr int B32_ROW = 36;
constexpr int B32_LO = 14;
constexpr float B32_SWITCH = 36.0f;
// FP64: nodes k/8, k = 0..320; row = F_0..F_14, e^{-T_k}
constexpr int B64_NODES = 321;
constexpr int B64_ROW = 16;
constexpr double B64_SWITCH = 40.0;

// host: fill the tables (FP64 values; caller rounds for FP32)
void boys_tables(std::vector<float>& t32, std::vector<double>& t64);

// Exact-ish x * (1/d) for a small odd integer d: one FMA correction makes the
// quotient the correctly rounded one in almost all cases, so the recursion
// carries no systematic bias from the rounded reciprocal.
__device__ __forceinline__ float div_small(float x, float d, float inv) {
    float y = x * inv;
    float r = fmaf(-y, d, x);
    return fmaf(r, inv, y);
}

// The leading table values and sqrt(pi)/2 carry their FP32 rounding
// remainders: a table value serves every (point, pair) near its node, so
// its rounding would bias all of them the same way.
template <int L>
__device__ __forceinline__ void boys_f32(float T, const float* __restrict__ tab, float* F) {
    if (T < B32_SWITCH) {
        const int k = (int)(T * 4.0f + 0.5f);
        const float x = 0.25f * (float)k - T;          // -h, exact
        const float* row = tab + k * B32_ROW;
        float p[7];
        p[0] = 1.0f;
        p[1] = x;
#pragma unroll
        for (int j = 2; j < 7; ++j) p[j] = p[j - 1] * x * (1.0f / (float)j);
        if (L <= 1) {
#pragma unroll
            for (int n = 0; n <= L; ++n) {
                float s = row[n + 6] * p[6];
#pragma unroll
                for (int j = 5; j >= 1; --j) s = fmaf(row[n + j], p[j], s);
                F[n] = row[n] + (row[B32_LO + n] + s);
            }
        } else {
            float s = row[L + 6] * p[6];
#pragma unroll
            for (int j = 5; j >= 1; --j) s = fmaf(row[L + j], p[j], s);
            F[L] = row[L] + (row[B32_LO + L] + s);
            float es = p[6];
#pragma unroll
            for (int j = 5; j >= 1; --j) es += p[j];
            // e^{-T} = e^{-T_k} e^{x} = (e_hi + e_lo)(1 + es)
            const float e = row[13] + fmaf(row[13], es, row[B32_LO + 13]);
            const float t2 = T + T;
#pragma unroll
            for (int n = L; n > 0; --n)
                F[n - 1] = div_small(fmaf(t2, F[n], e), (float)(2 * n - 1),
                                     1.0f / (float)(2 * n - 1));
        }
    } else {
        // 1/sqrt(T) to ~0.5 ulp: hardware estimate, then one Newton step on
        // the exact residual 1 - T s^2 (T s split into hi + lo by an FMA),
        // so the correction is not rounded against 1.0
        float s = __builtin_amdgcn_rsqf(T);
        const float th = T * s, tl = fmaf(T, s, -th);
        const float h = fmaf(-th, s, 1.0f) - tl * s;
        s = fmaf(0.5f * s, h, s);
        const float u = s * s;
        F[0] = fmaf(0.886226952f, s, -2.66231996e-08f * s);      // sqrt(pi)/2 hi + lo
#pragma unroll
        for (int n = 1; n <= L; ++n) F[n] = F[n - 1] * ((float)n - 0.5f) * u;
    }
}

template <int L>
__device__ __forceinline__ void boys_f64(double T, const double* __restrict__ tab, double* F) {
    if (T < B64_SWITCH) {
        const int k = (int)(T * 8.0 + 0.5);
        const double x = 0.125 * (double)k - T;
        const double* row = tab + k * B64_ROW;
        double p[9];
        p[0] = 1.0;
        p[1] = x;
#pragma unroll
        for (int j = 2; j < 9; ++j) p[j] = p[j - 1] * x * (1.0 / (double)j);
        double s = row[L + 8] * p[8];
#pragma unroll
        for (int j = 7; j >= 1; --j) s = fma(row[L + j], p[j], s);
        F[L] = row[L] + s;
        if (L > 0) {
            double es = p[8];
#pragma unroll
            for (int j = 7; j >= 1; --j) es += p[j];
            const double e = fma(row[15], es, row[15]);
            const double t2 = T + T;
#pragma unroll
            for (int n = L; n > 0; --n) F[n - 1] = fma(t2, F[n], e) * (1.0 / (double)(2 * n - 1));
        }
    } else {
        const double rs = 1.0 / sqrt(T);
        F[0] = 0.886226925452758014 * rs;
        if (L > 0) {
            const double e = exp(-T);
            const double h = 0.5 * rs * rs;
#pragma unroll
            for (int n = 0; n < L; ++n) F[n + 1] = ((double)(2 * n + 1) * F[n] - e) * h;
        }
    }
}

template <typename real> struct Boys;
template <> struct Boys<float> {
    template <int L> __device__ static void eval(float T, const float* t, float* F) { boys_f32<L>(T, t, F); }
};
template <> struct Boys<double> {
    template <int L> __device__ static void eval(double T, const double* t, double* F) { boys_f64<L>(T, t, F); }
};

}  // namespace wb
