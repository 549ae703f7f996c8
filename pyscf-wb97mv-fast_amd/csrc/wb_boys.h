// Boys function F_0..F_L(T) on the device, FP32 and FP64.
//
// Method (both precisions): a Taylor table around equally spaced nodes T_k,
//   F_m(T_k + h) = sum_j F_{m+j}(T_k) (-h)^j / j!      (dF_m/dT = -F_{m+1}),
// for T below a switch point, the asymptotic form above it:
//   F_0 = sqrt(pi/T)/2,  F_{n+1} = ((2n+1) F_n - e^{-T}) / (2T).
// For L >= 2 only F_L comes from the table; the lower orders follow by the
// stable downward recursion F_{n-1} = (2T F_n + e^{-T}) / (2n-1), with
// e^{-T} = e^{-T_k} e^{-h} (table value times a Taylor polynomial), so no
// transcendental is evaluated below the switch.  In the FP32 routine the
// recursion (and the asymptotic product for L >= 1) runs in FP64 from FP64
// leading values: each step is a convex combination of the errors of F_L
// and e^{-T}, so no order accumulates roundings, and every F_n is rounded
// to FP32 once (<= 0.5 ulp plus ~3e-8), at about the instruction count of
// an FP32 recursion with its divisions (FP64 FMA issues at the FP32 rate
// on CDNA4).  The reference uses the same node spacing and switch for FP32
// (gpu/boys.py:14-47, T_k = k/4, 36) and evaluates every order from a
// series of its own.
#pragma once
#include "wb_common.h"

namespace wb {

// FP32: nodes k/4, k = 0..144, one table per highest order L = 0..6 (a
// kernel class needs only its own).  Taylor coefficients are stored divided
// by j!, c_{n,j} = F_{n+j}(T_k) / j!, so the series is a plain Horner chain
// in x = T_k - T.  Row of table L:
//   L = 0:  c_{0,0..6}, lo_0                         (8 floats)
//   L = 1:  c_{0,0..6}, c_{1,0..6}, lo_0, lo_1       (16 floats)
//   L >= 2: F_L(T_k) and e^{-T_k} as two doubles, then c_{L,1..6}
//           (10 floats' worth; the doubles 8-byte aligned)
// lo: FP32 rounding remainder of the leading value.  Rows are padded to a
// multiple of 4 floats (16-byte loads) plus 4, so that rows of
// neighbouring nodes, which the lanes of a wave read together, start on
// different LDS banks.
constexpr int B32_NODES = 145;
constexpr float B32_SWITCH = 36.0f;
__host__ __device__ constexpr int b32_vals(int L) { return L == 0 ? 8 : (L == 1 ? 16 : 10); }
__host__ __device__ constexpr int b32_row(int L) { return (b32_vals(L) + 3) / 4 * 4 + 4; }
__host__ __device__ constexpr int b32_off(int L) {
    return L == 0 ? 0 : b32_off(L - 1) + B32_NODES * b32_row(L - 1);
}
constexpr int B32_LMAX = 6;
constexpr int B32_TOTAL = b32_off(B32_LMAX + 1);
// FP64: nodes k/8, k = 0..320; row = F_0..F_14, e^{-T_k}
constexpr int B64_NODES = 321;
constexpr int B64_ROW = 16;
constexpr double B64_SWITCH = 40.0;

// host: fill the tables (FP32: all orders at their offsets b32_off(L))
void boys_tables(std::vector<float>& t32, std::vector<double>& t64);

// Exact-ish x * (1/d) for a small odd integer d: one FMA correction makes the
// quotient the correctly rounded one in almost all cases, so the recursion
// carries no systematic bias from the rounded reciprocal.
__device__ __forceinline__ float div_small(float x, float d, float inv) {
    float y = x * inv;
    float r = fmaf(-y, d, x);
    return fmaf(r, inv, y);
}

// sum_{j=1..6} c[j] x^j (Horner)
__device__ __forceinline__ float taylor6(const float* c, float x) {
    float s = c[6];
#pragma unroll
    for (int j = 5; j >= 1; --j) s = fmaf(s, x, c[j]);
    return s * x;
}

// F_0..F_L(T); tab is the order-L table.  The leading table values carry
// their FP32 rounding remainders (L <= 1) or are FP64 (L >= 2), and so does
// sqrt(pi)/2: a table value serves every (point, pair) near its node, so its
// rounding would bias all of them the same way.
template <int L>
__device__ __forceinline__ void boys_f32(float T, const float* __restrict__ tab, float* F) {
    if (T < B32_SWITCH) {
        const int k = (int)(T * 4.0f + 0.5f);
        const float x = 0.25f * (float)k - T;          // -h, exact
        const float* row = tab + k * b32_row(L);
        if constexpr (L == 0) {
            F[0] = row[0] + (row[7] + taylor6(row, x));
        } else if constexpr (L == 1) {
            F[0] = row[0] + (row[14] + taylor6(row, x));
            F[1] = row[7] + (row[15] + taylor6(row + 7, x));
        } else {
            const double* rd = (const double*)row;        // F_L(T_k), e^{-T_k}
            const double fl = rd[0] + (double)taylor6(row + 3, x);
            // e^{x} - 1 to O(x^7)
            float es = 1.0f / 720.0f;
            es = fmaf(es, x, 1.0f / 120.0f);
            es = fmaf(es, x, 1.0f / 24.0f);
            es = fmaf(es, x, 1.0f / 6.0f);
            es = fmaf(es, x, 0.5f);
            es = fmaf(es, x, 1.0f);
            es *= x;
            // e^{-T} = e^{-T_k} e^{x}
            const double e = fma(rd[1], (double)es, rd[1]);
            const double t2 = 2.0 * (double)T;
            double f = fl;
            F[L] = (float)f;
#pragma unroll
            for (int n = L; n > 0; --n) {
                f = fma(t2, f, e) * (1.0 / (double)(2 * n - 1));
                F[n - 1] = (float)f;
            }
        }
    } else if constexpr (L == 0) {
        // 1/sqrt(T) to ~0.5 ulp: hardware estimate, then one Newton step on
        // the exact residual 1 - T s^2 (T s split into hi + lo by an FMA),
        // so the correction is not rounded against 1.0
        float s = __builtin_amdgcn_rsqf(T);
        const float th = T * s, tl = fmaf(T, s, -th);
        const float h = fmaf(-th, s, 1.0f) - tl * s;
        s = fmaf(0.5f * s, h, s);
        F[0] = fmaf(0.886226952f, s, -2.66231996e-08f * s);      // sqrt(pi)/2 hi + lo
    } else {
        // FP64 from the hardware estimate: one Newton step (error ~1.5 r^2,
        // r ~ 1e-7) gives 1/sqrt(T) and 1/T far below FP32 rounding; the
        // upward product F_n = F_{n-1} (n - 1/2) / T then rounds once per order
        const double sd = (double)__builtin_amdgcn_rsqf(T);
        const double r = fma(-(double)T, sd * sd, 1.0);
        const double s1 = fma(0.5 * sd, r, sd);
        const double u = s1 * s1;
        double f = 0.886226925452758014 * s1;
        F[0] = (float)f;
#pragma unroll
        for (int n = 1; n <= L; ++n) {
            f *= ((double)n - 0.5) * u;
            F[n] = (float)f;
        }
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
