// AO evaluation, FP64 MFMA GEMMs, GGA density and the VV10 double sum.
#include "wb_internal.h"
#include "md_gen.h"

namespace wb {

// ------------------------------------------------------------ AO values ---
// X[c][mu][g] (c = 0 values, 1..3 gradients), grid-major so that the 64
// lanes of a wave (64 consecutive points) store 512 contiguous bytes.
// Follows the evaluation of the reference's AoEvaluator (gpu/ao_eval.py:
// 255-329): radial sum in FP64 with exp in FP64, Cartesian powers, then
// cart -> sph (libcint order).
// One shell at one point: spherical values sv and (DERIV) gradients gs.
template <int l, int DERIV>
__device__ __forceinline__ void ao_shell_eval(double dx, double dy, double dz, double rad,
                                              double drad, double (&sv)[2 * l + 1],
                                              double (&gs)[3][2 * l + 1])
{
    constexpr int NC = (l + 1) * (l + 2) / 2;
    double px[l + 2], py[l + 2], pz[l + 2];
    px[0] = py[0] = pz[0] = 1.0;
#pragma unroll
    for (int k = 1; k < l + 2; ++k) {
        px[k] = px[k - 1] * dx;
        py[k] = py[k - 1] * dy;
        pz[k] = pz[k - 1] * dz;
    }
    double cv[NC], gc[3][NC];
    int c = 0;
#pragma unroll
    for (int lx = l; lx >= 0; --lx) {
#pragma unroll
        for (int ly = l - lx; ly >= 0; --ly) {
            const int lz = l - lx - ly;
            const double poly = px[lx] * py[ly] * pz[lz];
            cv[c] = poly * rad;
            if (DERIV) {
                gc[0][c] = py[ly] * pz[lz] * ((lx ? lx * px[lx - 1] : 0.0) * rad + px[lx + 1] * drad);
                gc[1][c] = px[lx] * pz[lz] * ((ly ? ly * py[ly - 1] : 0.0) * rad + py[ly + 1] * drad);
                gc[2][c] = px[lx] * py[ly] * ((lz ? lz * pz[lz - 1] : 0.0) * rad + pz[lz + 1] * drad);
            }
            ++c;
        }
    }
    if constexpr (l == 0) md_gsph_0<double>(cv, sv);
    else if constexpr (l == 1) md_gsph_1<double>(cv, sv);
    else if constexpr (l == 2) md_gsph_2<double>(cv, sv);
    else md_gsph_3<double>(cv, sv);
    if (DERIV) {
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            if constexpr (l == 0) md_gsph_0<double>(gc[ax], gs[ax]);
            else if constexpr (l == 1) md_gsph_1<double>(gc[ax], gs[ax]);
            else if constexpr (l == 2) md_gsph_2<double>(gc[ax], gs[ax]);
            else md_gsph_3<double>(gc[ax], gs[ax]);
        }
    }
}

// What a pass does with each shell's values at its point g:
//   AO_STORE0 / AO_STORE1: store values (and the three gradient planes);
//   AO_RHO: add sum_m T[mu][g] (value, gradient) to the density partials;
//   AO_PROJ: store B[mu][g] = v0 X + v1 dX/dx + v2 dX/dy + v3 dX/dz (v: the
//   GGA potential factors at g), the operand of the potential matrix.
enum AoMode { AO_STORE0 = 0, AO_STORE1, AO_RHO, AO_PROJ };

// returns, for AO_STORE0 / AO_STORE1, whether any value (or, with the
// gradients, any gradient component) of the shell at g exceeds AO_SIG_THR in
// magnitude
template <int l, int MODE>
__device__ __forceinline__ bool ao_shell(double dx, double dy, double dz, double rad,
                                         double drad, double* __restrict__ X, size_t plane,
                                         int ao0, int Ppad, int g, const double* __restrict__ T,
                                         const double (&v)[4], double (&acc)[4])
{
    constexpr int NS = 2 * l + 1;
    constexpr int DERIV = MODE != AO_STORE0;
    double sv[NS], gs[3][NS];
    ao_shell_eval<l, DERIV>(dx, dy, dz, rad, drad, sv, gs);
    bool big = false;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
        const size_t o = (size_t)(ao0 + m) * Ppad + g;
        if constexpr (MODE == AO_STORE0 || MODE == AO_STORE1) {
            X[o] = sv[m];
            big = big || fabs(sv[m]) > AO_SIG_THR;
            if constexpr (MODE == AO_STORE1) {
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    X[(ax + 1) * plane + o] = gs[ax][m];
                    big = big || fabs(gs[ax][m]) > AO_SIG_THR;
                }
            }
        } else if constexpr (MODE == AO_RHO) {
            const double t = T[o];
            acc[0] = fma(t, sv[m], acc[0]);
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) acc[1 + ax] = fma(t, gs[ax][m], acc[1 + ax]);
        } else {
            X[o] = v[0] * sv[m] + v[1] * gs[0][m] + v[2] * gs[1][m] + v[3] * gs[2][m];
        }
    }
    return big;
}

// grid: (points / 256, shell groups of spb); AO_RHO writes the group's
// partial sums part[group][4][Ppad].  sig (AO_STORE0 / AO_STORE1, may be
// null): sig[sub-block][shell] = 1 if a value (or stored gradient) of the
// shell exceeds AO_SIG_THR at a point of the 64-point sub-block (one wave =
// one sub-block), else 0.
template <int MODE>
__global__ __launch_bounds__(256)
void ao_kernel(const ShellInfo* __restrict__ shells, int nshell, int spb,
               const double* __restrict__ exps, const double* __restrict__ coefs,
               const double* __restrict__ xyz, int Ppad, double* __restrict__ X, size_t plane,
               const double* __restrict__ T, const double* __restrict__ vg,
               double* __restrict__ part, unsigned char* __restrict__ sig)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= Ppad) return;
    const double x = xyz[g], y = xyz[Ppad + g], z = xyz[2 * Ppad + g];
    double v[4] = {0, 0, 0, 0}, acc[4] = {0, 0, 0, 0};
    if constexpr (MODE == AO_PROJ) {
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = vg[(size_t)k * Ppad + g];
    }
    const int s0 = blockIdx.y * spb, s1 = min(nshell, s0 + spb);
    for (int is = s0; is < s1; ++is) {
        const ShellInfo sh = shells[is];
        const double dx = x - sh.cx, dy = y - sh.cy, dz = z - sh.cz;
        const double r2 = dx * dx + dy * dy + dz * dz;
        double rad = 0.0, drad = 0.0;
        for (int i = 0; i < sh.nprim; ++i) {
            const double a = exps[sh.prim0 + i];
            const double ar2 = a * r2;
            if (ar2 < 740.0) {
                const double ce = coefs[sh.prim0 + i] * exp(-ar2);
                rad += ce;
                drad -= 2.0 * a * ce;
            }
        }
        bool big;
        switch (sh.l) {
        case 0: big = ao_shell<0, MODE>(dx, dy, dz, rad, drad, X, plane, sh.ao0, Ppad, g, T, v, acc); break;
        case 1: big = ao_shell<1, MODE>(dx, dy, dz, rad, drad, X, plane, sh.ao0, Ppad, g, T, v, acc); break;
        case 2: big = ao_shell<2, MODE>(dx, dy, dz, rad, drad, X, plane, sh.ao0, Ppad, g, T, v, acc); break;
        default: big = ao_shell<3, MODE>(dx, dy, dz, rad, drad, X, plane, sh.ao0, Ppad, g, T, v, acc); break;
        }
        if constexpr (MODE == AO_STORE0 || MODE == AO_STORE1) {
            if (sig) {                                // uniform
                const bool any = __ballot(big) != 0ull;
                if ((threadIdx.x & 63) == 0) sig[(size_t)(g >> 6) * nshell + is] = any ? 1 : 0;
            }
        }
        (void)big;
    }
    if constexpr (MODE == AO_RHO) {
#pragma unroll
        for (int k = 0; k < 4; ++k) part[((size_t)blockIdx.y * 4 + k) * Ppad + g] = acc[k];
    }
}

template <int MODE>
static void launch_ao_pass(hipStream_t s, const Basis& b, const double* xyz, int Ppad, int spb,
                           double* X, const double* T, const double* vg, double* part,
                           unsigned char* sig = nullptr)
{
    dim3 grid((Ppad + 255) / 256, (b.nshell + spb - 1) / spb);
    hipLaunchKernelGGL(ao_kernel<MODE>, grid, dim3(256), 0, s, b.d_shells.as<ShellInfo>(),
                       b.nshell, spb, b.d_exps.as<double>(), b.d_coefs.as<double>(), xyz, Ppad, X,
                       (size_t)b.nao * Ppad, T, vg, part, sig);
    WB_HIP(hipGetLastError());
}

void launch_ao_eval_pts(hipStream_t s, const Basis& b, const double* xyz, int Ppad, int deriv,
                        double* X)
{
    if (deriv) launch_ao_pass<AO_STORE1>(s, b, xyz, Ppad, 16, X, nullptr, nullptr, nullptr);
    else launch_ao_pass<AO_STORE0>(s, b, xyz, Ppad, 16, X, nullptr, nullptr, nullptr);
}

// rho[0][g] = sum_mu T[mu][g] X[mu][g], rho[1+d][g] = 2 sum_mu T[mu][g] dX_d[mu][g]
// from the partials of the shell groups, summed in group order
__global__ __launch_bounds__(256)
void rho_reduce_kernel(int Ppad, int ngrp, const double* __restrict__ part,
                       double* __restrict__ rho)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= Ppad) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        double r = 0.0;
        for (int q = 0; q < ngrp; ++q) r += part[((size_t)q * 4 + k) * Ppad + g];
        rho[(size_t)k * Ppad + g] = k ? 2.0 * r : r;
    }
}

// out[g] (+)= scale * sum_mu A[mu][g] B[mu][g] (the kinetic energy density
// from T_d = D dX_d and dX_d, one axis per call)
__global__ __launch_bounds__(256)
void dot_planes_kernel(int nao, int Ppad, const double* __restrict__ A,
                       const double* __restrict__ B, double scale, int accumulate,
                       double* __restrict__ out)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= Ppad) return;
    double r = 0.0;
    for (int m = 0; m < nao; ++m) {
        const size_t o = (size_t)m * Ppad + g;
        r = fma(A[o], B[o], r);
    }
    out[g] = accumulate ? out[g] + scale * r : scale * r;
}

void launch_dot_planes(hipStream_t s, int nao, int Ppad, const double* A, const double* B,
                       double scale, bool accumulate, double* out)
{
    hipLaunchKernelGGL(dot_planes_kernel, dim3((Ppad + 255) / 256), dim3(256), 0, s, nao, Ppad,
                       A, B, scale, accumulate ? 1 : 0, out);
    WB_HIP(hipGetLastError());
}

// shell groups of the fused passes: about 8 per point block, for enough waves
static int ao_group(int nshell) { return std::max(1, (nshell + 7) / 8); }

size_t ao_rho_bytes(const Basis& b, int Ppad)
{
    const int spb = ao_group(b.nshell);
    return sizeof(double) * 4 * (size_t)Ppad * ((b.nshell + spb - 1) / spb);
}

void launch_ao_rho(hipStream_t s, const Basis& b, const Grid& g, const double* T, double* part,
                   double* rho)
{
    const int spb = ao_group(b.nshell), ngrp = (b.nshell + spb - 1) / spb;
    launch_ao_pass<AO_RHO>(s, b, g.d_xyz.as<double>(), g.Ppad, spb, nullptr, T, nullptr, part);
    hipLaunchKernelGGL(rho_reduce_kernel, dim3((g.Ppad + 255) / 256), dim3(256), 0, s, g.Ppad,
                       ngrp, part, rho);
    WB_HIP(hipGetLastError());
}

void launch_ao_proj(hipStream_t s, const Basis& b, const Grid& g, const double* vg, double* Bout)
{
    launch_ao_pass<AO_PROJ>(s, b, g.d_xyz.as<double>(), g.Ppad, ao_group(b.nshell), Bout,
                            nullptr, vg, nullptr);
}

void launch_ao_eval(hipStream_t s, const Basis& b, const Grid& g, int deriv, double* X)
{
    launch_ao_eval_pts(s, b, g.d_xyz.as<double>(), g.Ppad, deriv, X);
}

// ------------------------------------------------ AO significance lists ---
// For tile t of spt sub-blocks (t spt .. min((t + 1) spt, nsub) - 1): the
// AOs of the shells flagged significant in any of its sub-blocks, ascending,
// in lists[t][0 .. counts[t]); with inv, inv[t][mu] = the position of AO mu
// in that list, or -1.  One wave per tile; a wave-wide prefix sum of the
// shells' AO counts places each flagged shell's AOs.
__global__ __launch_bounds__(256)
void ao_list_kernel(int nsub, int spt, int nshell, const unsigned char* __restrict__ sig,
                    const int* __restrict__ sh_ao0, const int* __restrict__ sh_nao, int cap,
                    int* __restrict__ counts, int* __restrict__ lists, int* __restrict__ inv)
{
    const int t = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ntile = (nsub + spt - 1) / spt;
    if (t >= ntile) return;
    const int lane = threadIdx.x & 63;
    const int sb0 = t * spt, sb1 = min(nsub, sb0 + spt);
    int* out = lists + (size_t)t * cap;
    int* iv = inv ? inv + (size_t)t * cap : nullptr;
    int n = 0;
    for (int b0 = 0; b0 < nshell; b0 += 64) {
        const int sh = b0 + lane;
        int na = 0, a0 = 0, nall = 0;
        if (sh < nshell) {
            a0 = sh_ao0[sh];
            nall = sh_nao[sh];
            unsigned char f = 0;
            for (int sb = sb0; sb < sb1; ++sb) f |= sig[(size_t)sb * nshell + sh];
            na = f ? nall : 0;
        }
        int incl = na;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        const int off = n + incl - na;
        for (int m = 0; m < na; ++m) out[off + m] = a0 + m;
        if (iv)
            for (int m = 0; m < nall; ++m) iv[a0 + m] = na ? off + m : -1;
        n += __shfl(incl, 63, 64);
    }
    if (lane == 0) counts[t] = n;
}

AoLists ao_lists_layout(const Basis& b, int nsub, int kchunk)
{
    AoLists L;
    const int ntile = (nsub + 1) / 2;
    L.cap = b.nao;
    L.spt = std::max(1, kchunk / SUB);
    L.nsplit = (nsub + L.spt - 1) / L.spt;
    size_t o = 0;
    auto take = [&](size_t bytes) { const size_t r = o; o += (bytes + 255) / 256 * 256; return r; };
    L.sig = take((size_t)nsub * b.nshell);
    L.counts = take(sizeof(int) * (size_t)ntile);
    L.lists = take(sizeof(int) * (size_t)ntile * L.cap);
    L.scounts = take(sizeof(int) * (size_t)L.nsplit);
    L.slists = take(sizeof(int) * (size_t)L.nsplit * L.cap);
    L.sinv = take(sizeof(int) * (size_t)L.nsplit * L.cap);
    L.total = o;
    return L;
}

void launch_ao_eval_lists(hipStream_t s, const Basis& b, const Grid& g, int deriv, double* X,
                          void* ws, const AoLists& L)
{
    char* base = (char*)ws;
    unsigned char* sig = (unsigned char*)(base + L.sig);
    if (deriv)
        launch_ao_pass<AO_STORE1>(s, b, g.d_xyz.as<double>(), g.Ppad, 16, X, nullptr, nullptr,
                                  nullptr, sig);
    else
        launch_ao_pass<AO_STORE0>(s, b, g.d_xyz.as<double>(), g.Ppad, 16, X, nullptr, nullptr,
                                  nullptr, sig);
    const int ntile = (g.nsub + 1) / 2;
    hipLaunchKernelGGL(ao_list_kernel, dim3((ntile + 3) / 4), dim3(256), 0, s, g.nsub, 2,
                       b.nshell, sig, b.d_shao.as<int>(), b.d_shao.as<int>() + b.nshell, L.cap,
                       (int*)(base + L.counts), (int*)(base + L.lists), (int*)nullptr);
    hipLaunchKernelGGL(ao_list_kernel, dim3((L.nsplit + 3) / 4), dim3(256), 0, s, g.nsub, L.spt,
                       b.nshell, sig, b.d_shao.as<int>(), b.d_shao.as<int>() + b.nshell, L.cap,
                       (int*)(base + L.scounts), (int*)(base + L.slists), (int*)(base + L.sinv));
    WB_HIP(hipGetLastError());
}

// ---------------------------------------------------------- FP64 GEMMs ----
// v_mfma_f64_16x16x4_f64: lane l holds A[l%16][l/16], B[l/16][l%16] and
// C[4*r + l/16][l%16] in accumulator element r = 0..3.  A 256-thread block computes a 64x64
// tile, each wave a 32x32 quarter (2x2 MFMA tiles), BK = 16 per LDS stage.
typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int GB_M = 64, GB_N = 64, GB_K = 16;

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Loads a 64 (rows) x 16 (k) tile of a row-major matrix with k contiguous
// into S[k][row]; out-of-range entries are zero.
__device__ __forceinline__ void load_rows_k(const double* __restrict__ A, int lda, int rows,
                                            int kdim, int r0, int k0,
                                            double (*S)[GB_M + 1])
{
    const int t = threadIdx.x;
    const int r = t >> 2, kq = (t & 3) * 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int rr = r0 + r, kk = k0 + kq + i;
        S[kq + i][r] = (rr < rows && kk < kdim) ? A[(size_t)rr * lda + kk] : 0.0;
    }
}

template <typename OutT>
__global__ __launch_bounds__(256)
void gemm_nn_kernel(int M, int N, int K, const double* __restrict__ A,
                    const double* __restrict__ B, OutT* __restrict__ C)
{
    __shared__ double As[GB_K][GB_M + 1];
    __shared__ double Bs[GB_K][GB_N + 1];
    const int m0 = blockIdx.y * GB_M, n0 = blockIdx.x * GB_N;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    d4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (d4){0, 0, 0, 0};
    for (int k0 = 0; k0 < K; k0 += GB_K) {
        load_rows_k(A, K, M, K, m0, k0, As);
        {   // B[k][n], n contiguous
            const int kr = t >> 4, nq = (t & 15) * 4;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int kk = k0 + kr, nn = n0 + nq + i;
                Bs[kr][nq + i] = (kk < K && nn < N) ? B[(size_t)kk * N + nn] : 0.0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 4) {
            const int kr = kk + (lane >> 4), c = lane & 15;
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kr][wm + i * 16 + c];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[kr][wn + j * 16 + c];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = mfma64(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int mm = m0 + wm + i * 16 + 4 * r + (lane >> 4);
                const int nn = n0 + wn + j * 16 + (lane & 15);
                if (mm < M && nn < N) C[(size_t)mm * N + nn] = (OutT)acc[i][j][r];
            }
}

void gemm_nn(hipStream_t s, int M, int N, int K, const double* A, const double* B, void* C,
             bool out_f32)
{
    dim3 grid((N + GB_N - 1) / GB_N, (M + GB_M - 1) / GB_M);
    if (out_f32)
        hipLaunchKernelGGL(gemm_nn_kernel<float>, grid, dim3(256), 0, s, M, N, K, A, B,
                           (float*)C);
    else
        hipLaunchKernelGGL(gemm_nn_kernel<double>, grid, dim3(256), 0, s, M, N, K, A, B,
                           (double*)C);
    WB_HIP(hipGetLastError());
}

// Partial C tiles of one K split: Cp[z][M][N] = sum_{k in split z} A[m][k] B'[n][k],
// B'[n][k] = w[k] * sum_c Bc[n][k].
template <typename BT>
__global__ __launch_bounds__(256)
void gemm_nt_kernel(int M, int N, int K, int kchunk, const double* __restrict__ A,
                    const BT* __restrict__ B, int nb, size_t bstride,
                    const double* __restrict__ w, double* __restrict__ Cp)
{
    __shared__ double As[GB_K][GB_M + 1];
    __shared__ double Bs[GB_K][GB_N + 1];
    const int m0 = blockIdx.y * GB_M, n0 = blockIdx.x * GB_N, z = blockIdx.z;
    const int kb = z * kchunk, ke = min(K, kb + kchunk);
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    d4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = (d4){0, 0, 0, 0};
    for (int k0 = kb; k0 < ke; k0 += GB_K) {
        load_rows_k(A, K, M, ke, m0, k0, As);
        {
            const int r = t >> 2, kq = (t & 3) * 4;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int nn = n0 + r, kk = k0 + kq + i;
                double v = 0.0;
                if (nn < N && kk < ke) {
                    const BT* p = B + (size_t)nn * K + kk;
                    for (int c = 0; c < nb; ++c) v += (double)p[c * bstride];
                    if (w) v *= w[kk];
                }
                Bs[kq + i][r] = v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 4) {
            const int kr = kk + (lane >> 4), c = lane & 15;
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[kr][wm + i * 16 + c];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[kr][wn + j * 16 + c];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = mfma64(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    double* out = Cp + (size_t)z * M * N;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int mm = m0 + wm + i * 16 + 4 * r + (lane >> 4);
                const int nn = n0 + wn + j * 16 + (lane & 15);
                if (mm < M && nn < N) out[(size_t)mm * N + nn] = acc[i][j][r];
            }
}

// C = (accumulate ? C : 0) + Cp[0] + Cp[1] + ... in split order: a build
// cut into chunks of whole splits adds the same partials in the same order
__global__ void split_reduce_kernel(int MN, int nsplit, const double* __restrict__ Cp,
                                    double* __restrict__ C, int accumulate)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= MN) return;
    double v = accumulate ? C[i] : 0.0;
    for (int z = 0; z < nsplit; ++z) v += Cp[(size_t)z * MN + i];
    C[i] = v;
}

void gemm_nt_reduce(hipStream_t s, int M, int N, int K, const double* A, const void* B,
                    bool b_f32, int nb, size_t bstride, const double* w, double* C, DevBuf& ws)
{
    const int tiles = ((M + GB_M - 1) / GB_M) * ((N + GB_N - 1) / GB_N);
    int nsplit = std::max(1, std::min(4096 / std::max(tiles, 1), K / (GB_K * 32)));
    int kchunk = (K + nsplit - 1) / nsplit;
    kchunk = (kchunk + GB_K - 1) / GB_K * GB_K;
    nsplit = (K + kchunk - 1) / kchunk;
    double* Cp = (double*)ws.ensure(sizeof(double) * (size_t)nsplit * M * N);
    dim3 grid((N + GB_N - 1) / GB_N, (M + GB_M - 1) / GB_M, nsplit);
    if (b_f32)
        hipLaunchKernelGGL(gemm_nt_kernel<float>, grid, dim3(256), 0, s, M, N, K, kchunk, A,
                           (const float*)B, nb, bstride, w, Cp);
    else
        hipLaunchKernelGGL(gemm_nt_kernel<double>, grid, dim3(256), 0, s, M, N, K, kchunk, A,
                           (const double*)B, nb, bstride, w, Cp);
    WB_HIP(hipGetLastError());
    hipLaunchKernelGGL(split_reduce_kernel, dim3((M * N + 255) / 256), dim3(256), 0, s, M * N,
                       nsplit, Cp, C, 0);
    WB_HIP(hipGetLastError());
}

// ------------------------------------------------ tiled FP64 GEMMs -------
// Block of 4 waves in a 2 x 2 layout; a wave computes WM x WN MFMA tiles of
// 16 x 16, so the block tile is (32 WM) x (32 WN) (96 or 128 per side,
// chosen per call to minimize padding, e.g. nao = 648 -> 7 x 96 rather than
// 6 x 128).  BK = 16; the next K slice is fetched into registers while the
// current one is consumed from LDS (two LDS buffers, one barrier per slice).
// Global loads are 16 bytes per lane, 8 (k-contiguous operand) or BN/2
// (n-contiguous operand) consecutive lanes per row.  A k-contiguous slice
// sits in LDS as rows of eight 16-byte chunks, chunk ch of row r at position
// ch ^ ((r >> 1) & 7): the 16 rows an MFMA operand read touches then cover
// all 64 banks.  The k order inside a slice is permuted, identically for
// both operands: lane (c, kq) of MFMA step s takes k = 4 kq + s, i.e. two
// 128-bit LDS reads per row and slice.
// Block ids are remapped so that consecutive work items run on the same XCD
// (blocks are dealt round-robin over the 8 XCDs): the tiles of one K split
// then share its operand slices in one L2.
// (Water27, 648 x 218624 x 648: K GEMM 4.08 -> 3.24 ms, F GEMM 3.59 -> 3.11
// ms against the previous 8-byte-load, unswizzled version.)
constexpr int T_K = 16;
constexpr int NXCD = 8;
typedef double dv2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    const int per = (nblocks + NXCD - 1) / NXCD;
    const int xcd = bid % NXCD, local = bid / NXCD;
    // XCDs with an index below nblocks % NXCD get one block more
    const int rem = nblocks % NXCD;
    const int base = (rem == 0 || xcd < rem) ? xcd * per : rem * per + (xcd - rem) * (per - 1);
    return base + local;
}

// LDS position of 16-byte chunk ch (0..7) of row r of a k-contiguous slice
__device__ __forceinline__ int kslice_pos(int r, int ch) { return r * 8 + (ch ^ ((r >> 1) & 7)); }

// k-contiguous [rows][lda] operand, slice rows [r0, r0 + BR) x k [k0, k0 + 16):
// thread t holds chunks id = i 256 + t (row id / 8, chunk id % 8).  Rows past
// the end are clamped to the last row (they feed only output rows or columns
// that are never stored); chunks at or past kend read as zero (kend even).
template <int BR>
__device__ __forceinline__ void fetch_kslice(const double* __restrict__ A, int lda, int rows,
                                             int kend, int r0, int k0, dv2 (&v)[BR / 32])
{
#pragma unroll
    for (int i = 0; i < BR / 32; ++i) {
        const int id = i * 256 + threadIdx.x;
        const int r = min(r0 + (id >> 3), rows - 1), k = k0 + 2 * (id & 7);
        const dv2 t = *(const dv2*)(A + (size_t)r * lda + min(k, kend - 2));
        v[i] = k < kend ? t : (dv2){0.0, 0.0};
    }
}
template <int BR>
__device__ __forceinline__ void store_kslice(dv2* __restrict__ S, const dv2 (&v)[BR / 32])
{
#pragma unroll
    for (int i = 0; i < BR / 32; ++i) {
        const int id = i * 256 + threadIdx.x;
        S[kslice_pos(id >> 3, id & 7)] = v[i];
    }
}
// the two chunks (k = 4 kq .. 4 kq + 3) of row r that lane (c, kq) feeds to MFMA
__device__ __forceinline__ void read_kslice(const dv2* __restrict__ S, int r, int kq, dv2 (&x)[2])
{
    x[0] = S[kslice_pos(r, 2 * kq)];
    x[1] = S[kslice_pos(r, 2 * kq + 1)];
}

// n-contiguous [K][ldb] operand, slice k [k0, k0 + 16) x n [n0, n0 + BN) into
// LDS rows of BN + 4 doubles (a pitch of 8 mod 16 dwords puts rows 4 apart in
// opposite bank halves).  k rows past K are clamped to the last row (the
// k-contiguous operand is zero there); n past N is clamped likewise (columns
// never stored).
template <int BN>
__device__ __forceinline__ void fetch_nslice(const double* __restrict__ B, int ldb, int K, int N,
                                             int k0, int n0, dv2 (&v)[BN / 32])
{
    constexpr int CPR = BN / 2;
#pragma unroll
    for (int i = 0; i < BN / 32; ++i) {
        const int id = i * 256 + threadIdx.x;
        const int k = min(k0 + id / CPR, K - 1), n = min(n0 + 2 * (id % CPR), N - 2);
        v[i] = *(const dv2*)(B + (size_t)k * ldb + n);
    }
}
template <int BN>
__device__ __forceinline__ void store_nslice(double* __restrict__ S, const dv2 (&v)[BN / 32])
{
    constexpr int CPR = BN / 2, PITCH = BN + 4;
#pragma unroll
    for (int i = 0; i < BN / 32; ++i) {
        const int id = i * 256 + threadIdx.x;
        *(dv2*)(S + (id / CPR) * PITCH + 2 * (id % CPR)) = v[i];
    }
}

template <int WM, int WN, typename OutT>
__device__ __forceinline__ void store_tile(OutT* __restrict__ C, int ldc, int M, int N, int m0,
                                           int n0, int wm, int wn, const d4 (&acc)[WM][WN])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int mm = m0 + wm + i * 16 + 4 * r + (lane >> 4);
                const int nn = n0 + wn + j * 16 + (lane & 15);
                if (mm < M && nn < N) C[(size_t)mm * ldc + nn] = (OutT)acc[i][j][r];
            }
}

// C[M][N] = A[M][K] B[K][N] (A k-contiguous, B n-contiguous; K, N even);
// m tiles fastest (they share the B slice)
template <int WM, int WN, typename OutT>
__global__ __launch_bounds__(256, 2)
void gemm_nn_t_kernel(int M, int N, int K, const double* __restrict__ A,
                      const double* __restrict__ B, OutT* __restrict__ C, int mt)
{
    constexpr int BM = 32 * WM, BN = 32 * WN, PITCH = BN + 4;
    __shared__ dv2 As[2][BM * 8];
    __shared__ double Bs[2][T_K * PITCH];
    const int w = xcd_remap(blockIdx.x, gridDim.x);
    const int m0 = (w % mt) * BM, n0 = (w / mt) * BN;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave >> 1) * 16 * WM, wn = (wave & 1) * 16 * WN;
    const int c = lane & 15, kq = lane >> 4;
    d4 acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = (d4){0, 0, 0, 0};
    dv2 va[BM / 32], vb[BN / 32];
    fetch_kslice<BM>(A, K, M, K, m0, 0, va);
    fetch_nslice<BN>(B, N, K, N, 0, n0, vb);
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += T_K) {
        store_kslice<BM>(As[buf], va);
        store_nslice<BN>(Bs[buf], vb);
        __syncthreads();
        if (k0 + T_K < K) {
            fetch_kslice<BM>(A, K, M, K, m0, k0 + T_K, va);
            fetch_nslice<BN>(B, N, K, N, k0 + T_K, n0, vb);
        }
        dv2 a[WM][2];
#pragma unroll
        for (int i = 0; i < WM; ++i) read_kslice(As[buf], wm + 16 * i + c, kq, a[i]);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            double b[WN];
#pragma unroll
            for (int j = 0; j < WN; ++j) b[j] = Bs[buf][(4 * kq + s) * PITCH + wn + 16 * j + c];
#pragma unroll
            for (int i = 0; i < WM; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j) acc[i][j] = mfma64(a[i][s >> 1][s & 1], b[j], acc[i][j]);
        }
        buf ^= 1;
    }
    store_tile<WM, WN>(C, N, M, N, m0, n0, wm, wn, acc);
}

// Row-gathered NN GEMM over the AO lists of the 128-point n tiles:
//   C[m][n] = sum_{j < counts[t]} A[k_j][m] B[k_j][n],  k_j = lists[t][j],
// t = n / 128: A (rows: AOs, e.g. QD or D) and B (rows: AOs, the AO values
// X) are both read a whole row piece at a time, so the gathered loads stay
// 16-byte, 8 lanes per 128-byte line; rows past the list end are zero, and
// a last tile of 64 points reads clamped columns it never stores.  The
// sum runs over the AOs that are significant somewhere in the tile only
// (the others' values are below AO_SIG_THR): about a quarter of them on
// the benchmark molecules.  lda and ldb even, A and B 16-byte aligned.
template <int BR>
__device__ __forceinline__ void fetch_rows_gather(const double* __restrict__ A, int lda, int cols,
                                                  const int* __restrict__ lst, int cnt, int k0,
                                                  int c0, dv2 (&v)[BR / 32])
{
    constexpr int CPR = BR / 2;
#pragma unroll
    for (int i = 0; i < BR / 32; ++i) {
        const int id = i * 256 + threadIdx.x;
        const int k = k0 + id / CPR, c = min(c0 + 2 * (id % CPR), cols - 2);
        dv2 t = (dv2){0.0, 0.0};
        if (k < cnt) t = *(const dv2*)(A + (size_t)lst[k] * lda + c);
        v[i] = t;
    }
}

template <int WM, typename OutT>
__global__ __launch_bounds__(256, 2)
void gemm_nn_gather_kernel(int M, int N, const double* __restrict__ A, int lda,
                           const double* __restrict__ B, int ldb, OutT* __restrict__ C, int mt,
                           const int* __restrict__ counts, const int* __restrict__ lists,
                           int cap)
{
    constexpr int WN = 4;
    constexpr int BM = 32 * WM, BN = 32 * WN, PA = BM + 4, PB = BN + 4;
    __shared__ double As[2][T_K * PA];
    __shared__ double Bs[2][T_K * PB];
    const int w = xcd_remap(blockIdx.x, gridDim.x);
    const int nt = w / mt;
    const int m0 = (w % mt) * BM, n0 = nt * BN;
    const int cnt = counts[nt];
    const int* __restrict__ lst = lists + (size_t)nt * cap;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave >> 1) * 16 * WM, wn = (wave & 1) * 16 * WN;
    const int c = lane & 15, kq = lane >> 4;
    d4 acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = (d4){0, 0, 0, 0};
    dv2 va[BM / 32], vb[BN / 32];
    fetch_rows_gather<BM>(A, lda, M, lst, cnt, 0, m0, va);
    fetch_rows_gather<BN>(B, ldb, N, lst, cnt, 0, n0, vb);
    int buf = 0;
    for (int k0 = 0; k0 < cnt; k0 += T_K) {
#pragma unroll
        for (int i = 0; i < BM / 32; ++i) {
            const int id = i * 256 + threadIdx.x;
            *(dv2*)(&As[buf][(id / (BM / 2)) * PA + 2 * (id % (BM / 2))]) = va[i];
        }
#pragma unroll
        for (int i = 0; i < BN / 32; ++i) {
            const int id = i * 256 + threadIdx.x;
            *(dv2*)(&Bs[buf][(id / (BN / 2)) * PB + 2 * (id % (BN / 2))]) = vb[i];
        }
        __syncthreads();
        if (k0 + T_K < cnt) {
            fetch_rows_gather<BM>(A, lda, M, lst, cnt, k0 + T_K, m0, va);
            fetch_rows_gather<BN>(B, ldb, N, lst, cnt, k0 + T_K, n0, vb);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            double a[WM], b[WN];
#pragma unroll
            for (int i = 0; i < WM; ++i) a[i] = As[buf][(4 * kq + s) * PA + wm + 16 * i + c];
#pragma unroll
            for (int j = 0; j < WN; ++j) b[j] = Bs[buf][(4 * kq + s) * PB + wn + 16 * j + c];
#pragma unroll
            for (int i = 0; i < WM; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j) acc[i][j] = mfma64(a[i], b[j], acc[i][j]);
        }
        buf ^= 1;
    }
    store_tile<WM, WN>(C, N, M, N, m0, n0, wm, wn, acc);
}

// Cp[z][M][N] = sum_{k in split z} A[m][k] B[n][k] (both k-contiguous, K
// even); tiles fastest within a split (they share the A and B slices).
// GATHER: the rows of A are those of split z's AO list (glist[z][0 ..
// gcnt[z]), significant AOs only), and Cp[z] holds them compacted: row p of
// Cp[z] is AO glist[z][p]; m tiles past the count have nothing to do.
template <int BR>
__device__ __forceinline__ void fetch_kslice_rows(const double* __restrict__ A, int lda,
                                                  const int* __restrict__ rows, int nrows,
                                                  int kend, int r0, int k0, dv2 (&v)[BR / 32])
{
#pragma unroll
    for (int i = 0; i < BR / 32; ++i) {
        const int id = i * 256 + threadIdx.x;
        const int r = rows[min(r0 + (id >> 3), nrows - 1)], k = k0 + 2 * (id & 7);
        const dv2 t = *(const dv2*)(A + (size_t)r * lda + min(k, kend - 2));
        v[i] = k < kend ? t : (dv2){0.0, 0.0};
    }
}

template <int WM, int WN, bool GATHER>
__global__ __launch_bounds__(256, 2)
void gemm_nt_t_kernel(int M, int N, int K, int kchunk, int mt, int ntile,
                      const double* __restrict__ A, const double* __restrict__ B,
                      double* __restrict__ Cp, const int* __restrict__ gcnt,
                      const int* __restrict__ glist, int gcap)
{
    constexpr int BM = 32 * WM, BN = 32 * WN;
    __shared__ dv2 As[2][BM * 8];
    __shared__ dv2 Bs[2][BN * 8];
    const int w = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = w % ntile, z = w / ntile;
    const int m0 = (tile % mt) * BM, n0 = (tile / mt) * BN;
    const int kb = z * kchunk, ke = min(K, kb + kchunk);
    int Mz = M;
    const int* __restrict__ lz = nullptr;
    if constexpr (GATHER) {
        Mz = gcnt[z];
        if (m0 >= Mz) return;                     // uniform per block
        lz = glist + (size_t)z * gcap;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wm = (wave >> 1) * 16 * WM, wn = (wave & 1) * 16 * WN;
    const int c = lane & 15, kq = lane >> 4;
    d4 acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = (d4){0, 0, 0, 0};
    dv2 va[BM / 32], vb[BN / 32];
    auto fetch_a = [&](int k0) {
        if constexpr (GATHER) fetch_kslice_rows<BM>(A, K, lz, Mz, ke, m0, k0, va);
        else fetch_kslice<BM>(A, K, M, ke, m0, k0, va);
    };
    fetch_a(kb);
    fetch_kslice<BN>(B, K, N, ke, n0, kb, vb);
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += T_K) {
        store_kslice<BM>(As[buf], va);
        store_kslice<BN>(Bs[buf], vb);
        __syncthreads();
        if (k0 + T_K < ke) {
            fetch_a(k0 + T_K);
            fetch_kslice<BN>(B, K, N, ke, n0, k0 + T_K, vb);
        }
        dv2 a[WM][2], b[WN][2];
#pragma unroll
        for (int i = 0; i < WM; ++i) read_kslice(As[buf], wm + 16 * i + c, kq, a[i]);
#pragma unroll
        for (int j = 0; j < WN; ++j) read_kslice(Bs[buf], wn + 16 * j + c, kq, b[j]);
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < WM; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j)
                    acc[i][j] = mfma64(a[i][s >> 1][s & 1], b[j][s >> 1][s & 1], acc[i][j]);
        buf ^= 1;
    }
    store_tile<WM, WN>(Cp + (size_t)z * M * N, N, Mz, N, m0, n0, wm, wn, acc);
}

// C = (accumulate ? C : 0) + the compacted partials of the gathered split
// GEMM, in split order: row m of C takes row ginv[z][m] of Cp[z] where that
// is >= 0 (the AO was in split z's list), nothing from the others
__global__ void split_reduce_gather_kernel(int M, int N, int nsplit, const double* __restrict__ Cp,
                                           double* __restrict__ C, int accumulate,
                                           const int* __restrict__ ginv, int gcap)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= M * N) return;
    const int m = i / N, n = i - m * N;
    double v = accumulate ? C[i] : 0.0;
    for (int z = 0; z < nsplit; ++z) {
        const int p = ginv[(size_t)z * gcap + m];
        if (p >= 0) v += Cp[(size_t)z * M * N + (size_t)p * N + n];
    }
    C[i] = v;
}

// out[i] = w[i % K] * sum_c src[c][i]  (the G copies and the grid weights
// folded into one FP64 operand), i over nrow x K.  VEC: four consecutive
// elements per thread with 16-byte (FP32) / 32-byte (FP64) loads (K and
// the copy stride multiples of 4, buffers 16-byte aligned).
template <typename BT, bool VEC>
__global__ __launch_bounds__(256)
void fold_copies_kernel(size_t n, int K, const BT* __restrict__ src, int nb, size_t stride,
                        const double* __restrict__ w, double* __restrict__ out)
{
    if constexpr (VEC) {
        const size_t i = 4 * ((size_t)blockIdx.x * 256 + threadIdx.x);
        if (i >= n) return;
        double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
        for (int c = 0; c < nb; ++c) {
            const BT* p = src + c * stride + i;
            if constexpr (sizeof(BT) == 4) {
                const float4 q = *(const float4*)p;
                v0 += (double)q.x; v1 += (double)q.y; v2 += (double)q.z; v3 += (double)q.w;
            } else {
                const double2 q0 = *(const double2*)p, q1 = *(const double2*)(p + 2);
                v0 += q0.x; v1 += q0.y; v2 += q1.x; v3 += q1.y;
            }
        }
        if (w) {
            const size_t k = i % K;
            const double2 w0 = *(const double2*)(w + k), w1 = *(const double2*)(w + k + 2);
            v0 *= w0.x; v1 *= w0.y; v2 *= w1.x; v3 *= w1.y;
        }
        *(double2*)(out + i) = make_double2(v0, v1);
        *(double2*)(out + i + 2) = make_double2(v2, v3);
    } else {
        const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
        if (i >= n) return;
        double v = 0.0;
        for (int c = 0; c < nb; ++c) v += (double)src[c * stride + i];
        if (w) v *= w[i % K];
        out[i] = v;
    }
}

// The same fold over the exchange kernels' G copies, where a copy's rows
// in a sub-block region hold data only once written (masks, MdMasks
// layout; row0[r] = first row of row r's shell): never-written rows are
// skipped, not read.  Four consecutive points per thread (one sub-block).
template <typename BT>
__global__ __launch_bounds__(256)
void fold_masked_kernel(size_t n, int K, const BT* __restrict__ src, int nb, size_t stride,
                        const double* __restrict__ w, double* __restrict__ out,
                        const unsigned* __restrict__ masks, const int* __restrict__ row0,
                        int nsub)
{
    const size_t i = 4 * ((size_t)blockIdx.x * 256 + threadIdx.x);
    if (i >= n) return;
    const int r = (int)(i / K), p = (int)(i - (size_t)r * K);
    const int f = row0[r], sb = p / SUB;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    for (int c = 0; c < nb; ++c) {
        const unsigned word = masks[((size_t)c * nsub + sb) * MD_MASK_WORDS + (f >> 5)];
        if (!(word >> (f & 31) & 1u)) continue;
        const BT* q = src + c * stride + i;
        if constexpr (sizeof(BT) == 4) {
            const float4 x = *(const float4*)q;
            v0 += (double)x.x; v1 += (double)x.y; v2 += (double)x.z; v3 += (double)x.w;
        } else {
            const double2 x0 = *(const double2*)q, x1 = *(const double2*)(q + 2);
            v0 += x0.x; v1 += x0.y; v2 += x1.x; v3 += x1.y;
        }
    }
    const double2 w0 = *(const double2*)(w + p), w1 = *(const double2*)(w + p + 2);
    *(double2*)(out + i) = make_double2(v0 * w0.x, v1 * w0.y);
    *(double2*)(out + i + 2) = make_double2(v2 * w1.x, v3 * w1.y);
}

// 3 (96) or 4 (128) MFMA tiles per wave side: the one with less padding,
// unless the calling thread set a choice (gemm_tiles_scope, the autotuner's)
static thread_local int g_tiles = 0;
void set_gemm_tiles(int t) { g_tiles = (t == 3 || t == 4) ? t : 0; }
int get_gemm_tiles() { return g_tiles; }
static int wave_tiles(int D)
{
    if (g_tiles) return g_tiles;
    const int p96 = (D + 95) / 96 * 96, p128 = (D + 127) / 128 * 128;
    return p96 < p128 ? 3 : 4;
}

static bool aligned16(const void* p) { return ((uintptr_t)p % 16) == 0; }

void gemm_nn_fast(hipStream_t s, int M, int N, int K, const double* A, const double* B, void* C,
                  bool out_f32)
{
    if (K % 2 || N % 2 || !aligned16(A) || !aligned16(B)) {   // 16-byte loads impossible
        gemm_nn(s, M, N, K, A, B, C, out_f32);
        return;
    }
    const int wm = wave_tiles(M);
    const int mt = (M + 32 * wm - 1) / (32 * wm), nt = (N + 127) / 128;
    dim3 grid(mt * nt);
#define WB_NN(WMv, T) hipLaunchKernelGGL((gemm_nn_t_kernel<WMv, 4, T>), grid, dim3(256), 0, s, \
                                          M, N, K, A, B, (T*)C, mt)
    if (wm == 3) { if (out_f32) WB_NN(3, float); else WB_NN(3, double); }
    else { if (out_f32) WB_NN(4, float); else WB_NN(4, double); }
#undef WB_NN
    WB_HIP(hipGetLastError());
}

bool gemm_nn_gather_ok(int M, int N, const double* A, const double* B)
{
    return M % 2 == 0 && M >= 2 && N % 64 == 0 && N >= 64 && aligned16(A) && aligned16(B);
}

void gemm_nn_gather(hipStream_t s, int M, int N, const double* A, const double* B, void* C,
                    bool out_f32, const void* ws, const AoLists& L)
{
    if (!gemm_nn_gather_ok(M, N, A, B))
        throw Error(WB_INVALID, "gathered GEMM: M even, N a multiple of 64, aligned operands");
    const char* base = (const char*)ws;
    const int* counts = (const int*)(base + L.counts);
    const int* lists = (const int*)(base + L.lists);
    const int wm = wave_tiles(M);
    const int mt = (M + 32 * wm - 1) / (32 * wm), nt = (N + 127) / 128;
    dim3 grid(mt * nt);
#define WB_NG(WMv, T) hipLaunchKernelGGL((gemm_nn_gather_kernel<WMv, T>), grid, dim3(256), 0, s, \
                                          M, N, A, M, B, N, (T*)C, mt, counts, lists, L.cap)
    if (wm == 3) { if (out_f32) WB_NG(3, float); else WB_NG(3, double); }
    else { if (out_f32) WB_NG(4, float); else WB_NG(4, double); }
#undef WB_NG
    WB_HIP(hipGetLastError());
}

// split of the K GEMM over the point axis: <= 96 chunks of a multiple of
// 64 points (whole sub-blocks, so that a K build cut into chunks of whole
// splits sums the same partials)
int nt_split(int K, int& kchunk)
{
    int nsplit = std::max(1, std::min(96, K / (T_K * 64)));
    kchunk = (K + nsplit - 1) / nsplit;
    kchunk = (kchunk + SUB - 1) / SUB * SUB;
    return (K + kchunk - 1) / kchunk;
}

size_t gemm_nt_ws_bytes(int M, int N, int K, bool fold)
{
    int kchunk = 0;
    const int nsplit = nt_split(K, kchunk);
    return sizeof(double) * (size_t)nsplit * M * N + (fold ? sizeof(double) * (size_t)N * K : 0);
}

void gemm_nt_fast(hipStream_t s, int M, int N, int K, const double* A, const void* B, bool b_f32,
                  int nb, size_t bstride, const double* w, double* C, DevBuf& ws,
                  const unsigned* masks, const int* row0, int nsub, const NtGather* gather)
{
    if (masks && (K != nsub * SUB || !w || bstride % 4 || !aligned16(B) || !aligned16(w)))
        throw Error(WB_INVALID, "masked fold: B must be sub-block copies of 16-byte rows");
    if (K % 2 || !aligned16(A)) {                 // 16-byte loads impossible
        if (masks) throw Error(WB_INVALID, "masked fold: unaligned operand");
        gemm_nt_reduce(s, M, N, K, A, B, b_f32, nb, bstride, w, C, ws);
        return;
    }
    int kchunk = 0;
    const int nsplit = nt_split(K, kchunk);
    const bool fold = gemm_nt_folds(b_f32, nb, w, B, masks);
    const size_t part_bytes = sizeof(double) * (size_t)nsplit * M * N;
    char* base = (char*)ws.ensure(part_bytes + (fold ? sizeof(double) * (size_t)N * K : 0));
    gemm_nt_core(s, M, N, K, kchunk, A, B, b_f32, nb, bstride, w, C, false, (double*)base,
                 fold ? (double*)(base + part_bytes) : nullptr, masks, row0, nsub, gather);
}

NtGather nt_gather(const AoLists& L, const void* ws)
{
    const char* base = (const char*)ws;
    NtGather g;
    g.counts = (const int*)(base + L.scounts);
    g.lists = (const int*)(base + L.slists);
    g.inv = (const int*)(base + L.sinv);
    g.cap = L.cap;
    g.nsplit = L.nsplit;
    return g;
}

bool gemm_nt_folds(bool b_f32, int nb, const double* w, const void* B, const unsigned* masks)
{
    return b_f32 || nb != 1 || w != nullptr || !aligned16(B) || masks;
}

void gemm_nt_core(hipStream_t s, int M, int N, int K, int kchunk, const double* A, const void* B,
                  bool b_f32, int nb, size_t bstride, const double* w, double* C, bool accumulate,
                  double* Cp, double* Bf, const unsigned* masks, const int* row0, int nsub,
                  const NtGather* gather)
{
    if (K % 2 || !aligned16(A) || kchunk % T_K)
        throw Error(WB_INVALID, "split GEMM: K even, A 16-byte aligned, splits of whole slices");
    if (masks && (K != nsub * SUB || !w || bstride % 4 || !aligned16(B) || !aligned16(w)))
        throw Error(WB_INVALID, "masked fold: B must be sub-block copies of 16-byte rows");
    // fold the copies / weights into one FP64 operand when needed
    const double* Bd = (const double*)B;
    const size_t nbk = (size_t)N * K;
    const int wm = wave_tiles(M), wn = wave_tiles(N);
    const int mt = (M + 32 * wm - 1) / (32 * wm), ntile = mt * ((N + 32 * wn - 1) / (32 * wn));
    const int nsplit = (K + kchunk - 1) / kchunk;
    if (gemm_nt_folds(b_f32, nb, w, B, masks)) {
        if (!Bf) throw Error(WB_INVALID, "split GEMM: folded operand buffer missing");
        if (masks) {
            const unsigned blocks = (unsigned)((nbk / 4 + 255) / 256);
            if (b_f32)
                hipLaunchKernelGGL(fold_masked_kernel<float>, dim3(blocks), dim3(256), 0, s, nbk, K,
                                   (const float*)B, nb, bstride, w, Bf, masks, row0, nsub);
            else
                hipLaunchKernelGGL(fold_masked_kernel<double>, dim3(blocks), dim3(256), 0, s, nbk, K,
                                   (const double*)B, nb, bstride, w, Bf, masks, row0, nsub);
            WB_HIP(hipGetLastError());
        } else {
            const bool vec = K % 4 == 0 && bstride % 4 == 0 && aligned16(B);
            const unsigned blocks = (unsigned)((vec ? (nbk / 4 + 255) / 256 : (nbk + 255) / 256));
#define WB_FOLD(T, V) hipLaunchKernelGGL((fold_copies_kernel<T, V>), dim3(blocks), dim3(256), 0, s, \
                                         nbk, K, (const T*)B, nb, bstride, w, Bf)
            if (b_f32) { if (vec) WB_FOLD(float, true); else WB_FOLD(float, false); }
            else { if (vec) WB_FOLD(double, true); else WB_FOLD(double, false); }
#undef WB_FOLD
            WB_HIP(hipGetLastError());
        }
        Bd = Bf;
    }
    dim3 grid(ntile * nsplit);
    const bool g = gather != nullptr;
    if (g && (gather->nsplit != nsplit || gather->cap < M))
        throw Error(WB_INVALID, "split GEMM: AO lists for another split or row count");
    const int* gc = g ? gather->counts : nullptr;
    const int* gl = g ? gather->lists : nullptr;
    const int gcap = g ? gather->cap : 0;
#define WB_NT(a, b)                                                                              \
    if (g) hipLaunchKernelGGL((gemm_nt_t_kernel<a, b, true>), grid, dim3(256), 0, s, M, N, K,    \
                              kchunk, mt, ntile, A, Bd, Cp, gc, gl, gcap);                       \
    else hipLaunchKernelGGL((gemm_nt_t_kernel<a, b, false>), grid, dim3(256), 0, s, M, N, K,     \
                            kchunk, mt, ntile, A, Bd, Cp, gc, gl, gcap);
    if (wm == 3 && wn == 3) { WB_NT(3, 3) }
    else if (wm == 3) { WB_NT(3, 4) }
    else if (wn == 3) { WB_NT(4, 3) }
    else { WB_NT(4, 4) }
#undef WB_NT
    WB_HIP(hipGetLastError());
    if (g)
        hipLaunchKernelGGL(split_reduce_gather_kernel, dim3((M * N + 255) / 256), dim3(256), 0, s,
                           M, N, nsplit, Cp, C, accumulate ? 1 : 0, gather->inv, gcap);
    else
        hipLaunchKernelGGL(split_reduce_kernel, dim3((M * N + 255) / 256), dim3(256), 0, s, M * N,
                           nsplit, Cp, C, accumulate ? 1 : 0);
    WB_HIP(hipGetLastError());
}

// --------------------------------------------------------- VV10 sum -------
// Tile pruning (tol > 0): a wave skips a 64-point tile when, from the tile
// statistics and its own points' (wave reductions),
//   g >= Rmin^2 W0min_I + kmin_I, g' >= Rmin^2 W0min_J + kmin_J,
//   F <= qsum / (g g' (g+g')), U <= 2F/g, W <= U Rmax^2
// bounds every contribution in both directions below tol (the reference's
// host-side tile-pair bound, gpu/vv10.py:21-31, 147-164).
template <typename real> struct VP4 { real x, y, z, w; };   // coords, W0
template <typename real> struct VP2 { real k, q; };         // kappa, q = rho w

// ------------------------------------------------ symmetric VV10 sum -----
// Each unordered pair of points is evaluated once and feeds both sides:
//   D = 1 / (g g' (g+g')),  T_i = q_j D,  T_j = q_i D,
//   U_i += T_i (1/g + 1/(g+g')),  U_j += T_j (1/g' + 1/(g+g')),  W += U R^2
// with 1/g = g' (g+g') D, 1/g' = g (g+g') D, 1/(g+g') = g g' D.
//
// Points are grouped by 512 (one block: 4 waves x 2 points per lane); tiles
// are 64 points, 8 per group.  Group b pairs with the tiles of groups
// b+1 .. b+h (mod B), h = (B-1)/2, and for even B group b < B/2 also with
// group b + B/2: every pair of distinct groups is covered exactly once.  A
// group's own tiles are summed one-sided (every ordered pair, i = j
// included).  nsplit blocks share one group's partner range.
//
// Within a tile the wave runs 64 steps; at step k lane l pairs its points
// with tile point (l + k) mod 64, read from LDS.  The tile point's
// accumulators travel with it: after every step they rotate one lane down
// (DPP wave rotate), so lane l holds the accumulators of the point it
// pairs with next.  Every 16 steps both sides' FP32 sums are moved into FP64
// (i side: registers; tile side: LDS, added up at the end of the tile over
// the 4 waves in a fixed order) and written as one record per (group,
// partner tile).  A final pass adds, per point, its i-side parts and the
// records of the groups that visited its tile, in a fixed order: no
// atomics, bit-reproducible.
constexpr int VS_IG = 512;           // points per group (block)
constexpr int VS_GT = VS_IG / 64;    // tiles per group
constexpr int VS_SUB = 16;           // steps per FP32 partial sum

__host__ __device__ inline int vs_groups(int n) { return (n + VS_IG - 1) / VS_IG; }
__host__ __device__ inline int vs_hmax(int B) { return B / 2; }
__host__ __device__ inline int vs_npartner(int B, int b) {
    return (B - 1) / 2 + ((B % 2 == 0 && b < B / 2) ? 1 : 0);
}

// lane l receives the value of lane (l + 1) mod 64 (DPP wave_rol:1)
__device__ __forceinline__ float vs_rot(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xF, 0xF, false));
}
__device__ __forceinline__ double vs_rot(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x134, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x134, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <typename real>
__device__ __forceinline__ real vs_rcp(real x);
// FP32: the hardware reciprocal (<= 1 ulp) without refinement: a Newton
// step costs 2 of the pair's ~30 instructions (VV10 sum 17.9 -> 16.9 ms on
// water27) for a mixed-form energy change of 1e-9 Ha (dE_nl 3.9e-10 ->
// 1.3e-9 against the FP64 form)
template <> __device__ __forceinline__ float vs_rcp<float>(float x) {
    return __builtin_amdgcn_rcpf(x);
}
template <> __device__ __forceinline__ double vs_rcp<double>(double x) { return 1.0 / x; }

// one pair (i = a, j = p): i-side sums (f, u, w) and, when SYM, j-side
// contributions added to (cf, cu, cw)
template <typename real, bool SYM>
__device__ __forceinline__ void vs_pair(const VP4<real>& a, const VP2<real>& ka,
                                        const VP4<real>& p, const VP2<real>& kp,
                                        real& f, real& u, real& w,
                                        real& cf, real& cu, real& cw)
{
    const real dx = a.x - p.x, dy = a.y - p.y, dz = a.z - p.z;
    const real R2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const real g = fma(R2, a.w, ka.k);
    const real gp = fma(R2, p.w, kp.k);
    const real gt = g + gp;
    const real gg = g * gp;
    const real D = vs_rcp<real>(gg * gt);
    const real h = gt * D;                 // 1 / (g g')
    const real it = gg * D;                // 1 / (g + g')
    const real Ti = kp.q * D;
    const real Ui = Ti * fma(gp, h, it);   // Ti (1/g + 1/(g+g'))
    f += Ti; u += Ui; w = fma(Ui, R2, w);
    if constexpr (SYM) {
        const real Tj = ka.q * D;
        const real Uj = Tj * fma(g, h, it);
        cf += Tj; cu += Uj; cw = fma(Uj, R2, cw);
    }
}

// stat (tol > 0, may be null): [0] += (wave, tile) visits tested for
// pruning, [1] += those skipped; one add per wave at the end
template <typename real>
__global__ __launch_bounds__(256)
void vv10_sym_kernel(const int* __restrict__ nptr, const VP4<real>* __restrict__ pt,
                     const VP2<real>* __restrict__ pk, const float4* __restrict__ ts, float tol,
                     int nsplit, size_t pcap, double* __restrict__ ipart, real* __restrict__ jrec,
                     unsigned long long* __restrict__ stat)
{
    // the tile twice over, so that point (l + k) mod 64 is entry l + k
    __shared__ VP4<real> s_pt[128];
    __shared__ VP2<real> s_pk[128];
    __shared__ real s_j[4][3][4][64];        // [wave][quantity][flush][slot]
    const int n = *nptr;
    const int B = vs_groups(n);
    const int ntiles = (n + 63) / 64;
    const int hmax = vs_hmax(B);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned n_test = 0, n_skip = 0;             // wave-uniform pruning counts
    // work unit vb = (group b, part): one per block, or several per block
    // when the grid is smaller than B x nsplit (in order, stride gridDim.x)
    for (int vb = blockIdx.x; vb < B * nsplit; vb += gridDim.x) {
        const int b = vb / nsplit, part = vb % nsplit;
        const int i0 = b * VS_IG + wave * 128 + lane, i1 = i0 + 64;
        const bool v0 = i0 < n, v1 = i1 < n;
        // padding points carry q = 0 (and kappa = 1): they add nothing to anyone
        auto ld4 = [&](int i, bool v) { VP4<real> r{0, 0, 0, 0}; if (v) r = pt[i]; return r; };
        auto ld2 = [&](int i, bool v) { VP2<real> r{1, 0}; if (v) r = pk[i]; return r; };
        const VP4<real> a = ld4(i0, v0), c = ld4(i1, v1);
        const VP2<real> ka = ld2(i0, v0), kc = ld2(i1, v1);
        // wave statistics of its points (pruning)
        float cx = 0, cy = 0, cz = 0, rI = 0, wminI = 0, kminI = 0, qI = 0;
        if (tol > 0.f) {
            const float cnt = (float)__popcll(__ballot(v0)) + (float)__popcll(__ballot(v1));
            float sx = (v0 ? (float)a.x : 0.f) + (v1 ? (float)c.x : 0.f);
            float sy = (v0 ? (float)a.y : 0.f) + (v1 ? (float)c.y : 0.f);
            float sz = (v0 ? (float)a.z : 0.f) + (v1 ? (float)c.z : 0.f);
            float sq = (float)ka.q + (float)kc.q;
            for (int o = 32; o > 0; o >>= 1) {
                sx += __shfl_xor(sx, o, 64);
                sy += __shfl_xor(sy, o, 64);
                sz += __shfl_xor(sz, o, 64);
                sq += __shfl_xor(sq, o, 64);
            }
            const float cn = fmaxf(cnt, 1.f);
            cx = sx / cn; cy = sy / cn; cz = sz / cn; qI = sq * 1.0001f;
            auto dist = [&](const VP4<real>& p) {
                const float ex = (float)p.x - cx, ey = (float)p.y - cy, ez = (float)p.z - cz;
                return sqrtf(ex * ex + ey * ey + ez * ez);
            };
            float r = fmaxf(v0 ? dist(a) : 0.f, v1 ? dist(c) : 0.f);
            float wm = fminf(v0 ? (float)a.w : 3e38f, v1 ? (float)c.w : 3e38f);
            float km = fminf(v0 ? (float)ka.k : 3e38f, v1 ? (float)kc.k : 3e38f);
            for (int o = 32; o > 0; o >>= 1) {
                r = fmaxf(r, __shfl_xor(r, o, 64));
                wm = fminf(wm, __shfl_xor(wm, o, 64));
                km = fminf(km, __shfl_xor(km, o, 64));
            }
            rI = r * 1.0001f + 1e-5f; wminI = wm * 0.9999f; kminI = km * 0.9999f;
        }
        double F0 = 0, U0 = 0, W0 = 0, F1 = 0, U1 = 0, W1 = 0;

        // the block's work list: own tiles (part 0, one-sided), then its share
        // of the partner tiles
        const int np = vs_npartner(B, b);
        const int tot = np * VS_GT;
        const int lo = (int)((long long)tot * part / nsplit);
        const int hi = (int)((long long)tot * (part + 1) / nsplit);
        const int nown = part == 0 ? VS_GT : 0;
        for (int it = -nown; it < hi - lo; ++it) {
            const bool own = it < 0;
            int t, d = 0, u = 0;
            if (own) {
                t = b * VS_GT + (it + VS_GT);
            } else {
                const int x = lo + it;
                d = x / VS_GT + 1;
                u = x % VS_GT;
                t = ((b + d) % B) * VS_GT + u;
            }
            if (t >= ntiles) continue;           // uniform per block
            __syncthreads();
            if (threadIdx.x < 128) {
                const int j = t * 64 + (threadIdx.x & 63);
                s_pt[threadIdx.x] = ld4(j, j < n);
                s_pk[threadIdx.x] = ld2(j, j < n);
            }
            __syncthreads();
            bool skip = false;
            if (tol > 0.f) {
                const float4 tc = ts[2 * t], st = ts[2 * t + 1];   // c.xyz r | wmin kmin qsum
                const float dd = sqrtf((tc.x - cx) * (tc.x - cx) + (tc.y - cy) * (tc.y - cy) +
                                       (tc.z - cz) * (tc.z - cz));
                const float rmin = fmaxf(dd - rI - tc.w, 0.f);
                const float rmax = dd + rI + tc.w;
                const float gI = rmin * rmin * wminI + kminI;
                const float gJ = rmin * rmin * st.x + st.y;
                const float den = gI * gJ * (gI + gJ);
                const float bI = st.z / den, bJ = qI / den;
                const float bound = fmaxf(fmaxf(bI, 2.f * bI / gI * fmaxf(1.f, rmax * rmax)),
                                          fmaxf(bJ, 2.f * bJ / gJ * fmaxf(1.f, rmax * rmax)));
                skip = bound < tol;              // uniform per wave
                ++n_test;
                n_skip += skip ? 1u : 0u;
            }
            if (own) {
                if (!skip) {
#pragma unroll 1
                    for (int m = 0; m < 64 / VS_SUB; ++m) {
                        real f0 = 0, u0 = 0, w0 = 0, f1 = 0, u1 = 0, w1 = 0, z = 0;
                        const VP4<real>* tp = s_pt + ((lane + m * VS_SUB) & 63);
                        const VP2<real>* tk = s_pk + ((lane + m * VS_SUB) & 63);
#pragma unroll
                        for (int kk = 0; kk < VS_SUB; ++kk) {
                            const VP4<real> p = tp[kk];
                            const VP2<real> kp = tk[kk];
                            vs_pair<real, false>(a, ka, p, kp, f0, u0, w0, z, z, z);
                            vs_pair<real, false>(c, kc, p, kp, f1, u1, w1, z, z, z);
                        }
                        F0 += (double)f0; U0 += (double)u0; W0 += (double)w0;
                        F1 += (double)f1; U1 += (double)u1; W1 += (double)w1;
                    }
                }
                continue;
            }
            real (*sj)[4][64] = s_j[wave];
            if (skip) {
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    sj[0][m][lane] = 0; sj[1][m][lane] = 0; sj[2][m][lane] = 0;
                }
            } else {
#pragma unroll 1
                for (int m = 0; m < 64 / VS_SUB; ++m) {
                    real f0 = 0, u0 = 0, w0 = 0, f1 = 0, u1 = 0, w1 = 0;
                    real cf = 0, cu = 0, cw = 0;     // accumulators of point (lane + k) mod 64
                    const VP4<real>* tp = s_pt + ((lane + m * VS_SUB) & 63);
                    const VP2<real>* tk = s_pk + ((lane + m * VS_SUB) & 63);
#pragma unroll
                    for (int kk = 0; kk < VS_SUB; ++kk) {
                        const VP4<real> p = tp[kk];
                        const VP2<real> kp = tk[kk];
                        vs_pair<real, true>(a, ka, p, kp, f0, u0, w0, cf, cu, cw);
                        vs_pair<real, true>(c, kc, p, kp, f1, u1, w1, cf, cu, cw);
                        cf = vs_rot(cf); cu = vs_rot(cu); cw = vs_rot(cw);
                    }
                    // lane l now holds the sums of point (l + 16 (m + 1)) mod 64
                    const int slot = (lane + (m + 1) * VS_SUB) & 63;
                    sj[0][m][slot] = cf; sj[1][m][slot] = cu; sj[2][m][slot] = cw;
                    F0 += (double)f0; U0 += (double)u0; W0 += (double)w0;
                    F1 += (double)f1; U1 += (double)u1; W1 += (double)w1;
                }
            }
            __syncthreads();
            if (threadIdx.x < 192) {
                const int q = threadIdx.x >> 6, slot = threadIdx.x & 63;
                double acc = 0;
#pragma unroll
                for (int wv = 0; wv < 4; ++wv)
#pragma unroll
                    for (int m = 0; m < 4; ++m) acc += (double)s_j[wv][q][m][slot];
                real* rec = jrec + (((size_t)b * hmax + (d - 1)) * VS_GT + u) * 192;
                rec[q * 64 + slot] = (real)acc;
            }
        }
        double* ip = ipart + (size_t)part * 3 * pcap;
        if (v0) { ip[i0] = F0; ip[pcap + i0] = U0; ip[2 * pcap + i0] = W0; }
        if (v1) { ip[i1] = F1; ip[pcap + i1] = U1; ip[2 * pcap + i1] = W1; }
    }
    if (stat && tol > 0.f && lane == 0) {
        atomicAdd(stat, (unsigned long long)n_test);
        atomicAdd(stat + 1, (unsigned long long)n_skip);
    }
}

// per point: i-side parts in part order, then the records of the groups
// that visited its tile, by increasing distance d
template <typename real>
__global__ __launch_bounds__(256)
void vv10_sym_reduce(const int* __restrict__ nptr, int nsplit, size_t pcap,
                     const double* __restrict__ ipart, const real* __restrict__ jrec,
                     double* __restrict__ Fo, double* __restrict__ Uo, double* __restrict__ Wo)
{
    const int n = *nptr;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int B = vs_groups(n), hmax = vs_hmax(B);
    const int t = p >> 6, slot = p & 63, g = t / VS_GT, u = t % VS_GT;
    double F = 0, U = 0, W = 0;
    for (int s = 0; s < nsplit; ++s) {
        const double* ip = ipart + (size_t)s * 3 * pcap;
        F += ip[p]; U += ip[pcap + p]; W += ip[2 * pcap + p];
    }
    for (int d = 1; d <= hmax; ++d) {
        const int b = (g - d + B) % B;
        if (d > vs_npartner(B, b)) continue;
        const real* rec = jrec + (((size_t)b * hmax + (d - 1)) * VS_GT + u) * 192;
        F += (double)rec[slot]; U += (double)rec[64 + slot]; W += (double)rec[128 + slot];
    }
    Fo[p] = -1.5 * F; Uo[p] = U; Wo[p] = W;
}

// workspace bytes of vv10_sym for up to ncap points
size_t vv10_sym_bytes(int ncap, bool f64, int* nsplit_out)
{
    const int Bm = vs_groups(ncap);
    const int nsplit = std::max(1, std::min(16, (2048 + Bm - 1) / std::max(Bm, 1)));
    if (nsplit_out) *nsplit_out = nsplit;
    const size_t rs = f64 ? 8 : 4;
    const size_t pcap = (size_t)Bm * VS_IG;
    const size_t nrec = (size_t)Bm * std::max(vs_hmax(Bm), 1) * VS_GT;
    return (sizeof(double) * 3 * pcap * nsplit + 255) / 256 * 256 + rs * 192 * nrec;
}

// symmetric sum for n = *nptr <= ncap compacted points; ws holds
// vv10_sym_bytes(ncap) bytes
void launch_vv10_sym(hipStream_t s, const int* nptr, int ncap, const void* pts, const void* pk,
                     const float4* tiles, float tol, bool f64, void* ws, double* F, double* U,
                     double* W, unsigned long long* stat, int cap16)
{
    if (stat) zero_bytes(s, stat, 2 * sizeof(unsigned long long));
    int nsplit = 1;
    vv10_sym_bytes(ncap, f64, &nsplit);
    const int Bm = vs_groups(ncap);
    const size_t pcap = (size_t)Bm * VS_IG;
    double* ipart = (double*)ws;
    void* jrec = (char*)ws + (sizeof(double) * 3 * pcap * nsplit + 255) / 256 * 256;
    int blocks = Bm * nsplit;
    if (cap16 > 0) {
        int dev = 0, cus = 0;
        WB_HIP(hipGetDevice(&dev));
        WB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        blocks = std::min(blocks, std::max(1, cus * cap16 / 16));
    }
    const dim3 grid(blocks), red((ncap + 255) / 256);
    if (f64) {
        hipLaunchKernelGGL(vv10_sym_kernel<double>, grid, dim3(256), 0, s, nptr,
                           (const VP4<double>*)pts, (const VP2<double>*)pk, tiles, tol, nsplit,
                           pcap, ipart, (double*)jrec, stat);
        hipLaunchKernelGGL(vv10_sym_reduce<double>, red, dim3(256), 0, s, nptr, nsplit, pcap,
                           ipart, (const double*)jrec, F, U, W);
    } else {
        hipLaunchKernelGGL(vv10_sym_kernel<float>, grid, dim3(256), 0, s, nptr,
                           (const VP4<float>*)pts, (const VP2<float>*)pk, tiles, tol, nsplit,
                           pcap, ipart, (float*)jrec, stat);
        hipLaunchKernelGGL(vv10_sym_reduce<float>, red, dim3(256), 0, s, nptr, nsplit, pcap,
                           ipart, (const float*)jrec, F, U, W);
    }
    WB_HIP(hipGetLastError());
}

// ------------------------------------------------ VV10 flow helpers -----
// Per-point quantities of numint._vv10nlc (gpu/vv10.py:243-265), FP64,
// for points with rho >= thresh and w > 0; the kept points are compacted in
// grid order (one wave per 64-point sub-block; offsets from a scan).
__global__ __launch_bounds__(256)
void vv_count_kernel(int nsub, int Ppad, const double* __restrict__ rho,
                     const double* __restrict__ w, double thresh, int* __restrict__ cnt)
{
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= nsub) return;
    const int g = s * SUB + (threadIdx.x & 63);
    const bool keep = w[g] > 0.0 && rho[g] >= thresh;
    const unsigned long long b = __ballot(keep);
    if ((threadIdx.x & 63) == 0) cnt[s] = __popcll(b);
}

// exclusive scan of cnt[0..n) into off[0..n], off[n] = total, by one wave:
// a block of 1024 threads waits for a CU with 16 free wave slots, which a
// concurrent exchange launch on the other stream keeps from it for
// milliseconds (measured: 3.2 ms of the VV10 flow); one wave takes the first
// free slot
__global__ __launch_bounds__(64)
void scan_kernel(int n, const int* __restrict__ cnt, int* __restrict__ off)
{
    const int t = threadIdx.x;
    const int per = (n + 63) / 64;
    const int b = min(n, t * per), e = min(n, b + per);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += cnt[i];
    int inc = sum;                                   // inclusive wave scan
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (t >= o) inc += v;
    }
    int run = inc - sum;
    for (int i = b; i < e; ++i) { off[i] = run; run += cnt[i]; }
    if (t == 63) off[n] = inc;
}

template <typename real>
__global__ __launch_bounds__(256)
void vv_point_kernel(int nsub, int Ppad, const double* __restrict__ xyz,
                     const double* __restrict__ rho, const double* __restrict__ w,
                     double thresh, double b, double C, double ox, double oy, double oz,
                     const int* __restrict__ off, VP4<real>* __restrict__ pt,
                     VP2<real>* __restrict__ pk, int* __restrict__ idx,
                     double* __restrict__ aux /* [4][Ppad]: dW0dR, dW0dG, dKdR, - */)
{
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= nsub) return;
    const int lane = threadIdx.x & 63;
    const int g = s * SUB + lane;
    const double R = rho[g];
    const bool keep = w[g] > 0.0 && R >= thresh;
    const unsigned long long bal = __ballot(keep);
    if (!keep) return;
    const int pos = off[s] + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
    const double gx = rho[Ppad + g], gy = rho[2 * Ppad + g], gz = rho[3 * Ppad + g];
    const double G = gx * gx + gy * gy + gz * gz;
    const double pi43 = 4.0 * M_PI / 3.0;
    const double kvv = b * 1.5 * M_PI * pow(9.0 * M_PI, -1.0 / 6.0);
    const double t = G / (R * R);
    const double w0tmp = C * t * t;
    const double W0 = sqrt(w0tmp + pi43 * R);
    const double kap = kvv * pow(R, 1.0 / 6.0);
    pt[pos] = VP4<real>{(real)(xyz[g] - ox), (real)(xyz[Ppad + g] - oy),
                        (real)(xyz[2 * Ppad + g] - oz), (real)W0};
    pk[pos] = VP2<real>{(real)kap, (real)(R * w[g])};
    idx[pos] = g;
    aux[pos] = (0.5 * pi43 * R - 2.0 * w0tmp) / W0;
    aux[Ppad + pos] = w0tmp * R / (G * W0);
    aux[2 * Ppad + pos] = kap / 6.0;
}

// per 64-point tile of the compacted set: centre, radius, min W0, min kappa, sum q
template <typename real>
__global__ __launch_bounds__(256)
void vv_tile_kernel(const int* __restrict__ nptr, const VP4<real>* __restrict__ pt,
                    const VP2<real>* __restrict__ pk, int maxtiles, float4* __restrict__ ts)
{
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int n = *nptr;
    const int ntiles = (n + 63) / 64;
    if (t >= ntiles || t >= maxtiles) return;
    const int lane = threadIdx.x & 63;
    const int i = t * 64 + lane;
    const bool v = i < n;
    const VP4<real> p = v ? pt[i] : pt[t * 64];
    const VP2<real> k = v ? pk[i] : pk[t * 64];
    const float cnt = (float)__popcll(__ballot(v));
    float sx = v ? (float)p.x : 0.f, sy = v ? (float)p.y : 0.f, sz = v ? (float)p.z : 0.f;
    double qs = v ? (double)k.q : 0.0;
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
        sz += __shfl_xor(sz, o, 64);
        qs += __shfl_xor(qs, o, 64);
    }
    const float cx = sx / cnt, cy = sy / cnt, cz = sz / cnt;
    const float ex = (float)p.x - cx, ey = (float)p.y - cy, ez = (float)p.z - cz;
    float r = sqrtf(ex * ex + ey * ey + ez * ez);
    float wm = (float)p.w, km = (float)k.k;
    for (int o = 32; o > 0; o >>= 1) {
        r = fmaxf(r, __shfl_xor(r, o, 64));
        wm = fminf(wm, __shfl_xor(wm, o, 64));
        km = fminf(km, __shfl_xor(km, o, 64));
    }
    if (lane == 0) {
        ts[2 * t] = make_float4(cx, cy, cz, r * 1.0001f + 1e-5f);
        ts[2 * t + 1] = make_float4(wm * 0.9999f, km * 0.9999f, (float)(qs * 1.0001), 0.f);
    }
}

// exc/v0/v1 on the kept points -> weighted GGA potential wv[4][Ppad] and the
// energy density e[g] = rho w exc (zero elsewhere)
__global__ __launch_bounds__(256)
void vv_pot_kernel(const int* __restrict__ nptr, const int* __restrict__ idx, int Ppad,
                   const double* __restrict__ F, const double* __restrict__ U,
                   const double* __restrict__ W, const double* __restrict__ aux,
                   const double* __restrict__ rho, const double* __restrict__ w, double beta,
                   double* __restrict__ wv, double* __restrict__ e)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= *nptr) return;
    const int g = idx[i];
    const double exc = beta + 0.5 * F[i];
    const double v0 = beta + F[i] + 1.5 * (U[i] * aux[2 * Ppad + i] + W[i] * aux[i]);
    const double v1 = 1.5 * W[i] * aux[Ppad + i];
    const double wg = w[g];
    wv[g] = 0.5 * v0 * wg;
    wv[Ppad + g] = 2.0 * v1 * rho[Ppad + g] * wg;
    wv[2 * Ppad + g] = 2.0 * v1 * rho[2 * Ppad + g] * wg;
    wv[3 * Ppad + g] = 2.0 * v1 * rho[3 * Ppad + g] * wg;
    e[g] = rho[g] * wg * exc;
}

// B[nu][g] = sum_c wv[c][g] X[c][nu][g]
// out[0] = sum rho w, out[1] = sum e  (one block, fixed order)
__global__ __launch_bounds__(1024)
void vv_sums_kernel(int Ppad, const double* __restrict__ rho, const double* __restrict__ w,
                    const double* __restrict__ e, double* __restrict__ out)
{
    __shared__ double s0[1024], s1[1024];
    const int t = threadIdx.x;
    double a = 0, c = 0;
    for (int g = t; g < Ppad; g += 1024) {
        a += rho[g] * w[g];
        c += e[g];
    }
    s0[t] = a;
    s1[t] = c;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (t < o) { s0[t] += s0[t + o]; s1[t] += s1[t + o]; }
        __syncthreads();
    }
    if (t == 0) { out[0] = s0[0]; out[1] = s1[0]; }
}

__global__ void fill_kernel(size_t n, double* __restrict__ p, double v)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill32_kernel(size_t n, unsigned* __restrict__ p)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0u;
}
// n16 16-byte words, four per thread (grid-stride over 1024-word chunks)
__global__ __launch_bounds__(256)
void fill128_kernel(size_t n16, uint4* __restrict__ p)
{
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    for (size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 1024) {
        p[i] = z;
        if (i + 256 < n16) p[i + 256] = z;
        if (i + 512 < n16) p[i + 512] = z;
        if (i + 768 < n16) p[i + 768] = z;
    }
}
__global__ void transpose_kernel(int n, const double* __restrict__ a, double* __restrict__ b)
{
    const int i = blockIdx.y, j = blockIdx.x * 256 + threadIdx.x;
    if (j < n) b[(size_t)j * n + i] = a[(size_t)i * n + j];
}
// C = A + A^T (n x n)
__global__ void symadd_kernel(int n, const double* __restrict__ a, double* __restrict__ c)
{
    const int i = blockIdx.y, j = blockIdx.x * 256 + threadIdx.x;
    if (j < n) c[(size_t)i * n + j] = a[(size_t)i * n + j] + a[(size_t)j * n + i];
}

// a fill kernel rather than hipMemsetAsync, so that callers may capture it
// in a graph
void zero_bytes(hipStream_t s, void* p, size_t bytes)
{
    if (bytes == 0) return;
    if (((uintptr_t)p % 16) == 0 && bytes % 16 == 0) {
        const size_t n16 = bytes / 16;
        const size_t blocks = std::min<size_t>((n16 + 1023) / 1024, 65536);
        hipLaunchKernelGGL(fill128_kernel, dim3((unsigned)blocks), dim3(256), 0, s, n16, (uint4*)p);
    } else {
        const size_t n = bytes / 4;
        hipLaunchKernelGGL(fill32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n,
                           (unsigned*)p);
    }
    WB_HIP(hipGetLastError());
}
void transpose(hipStream_t s, int n, const double* a, double* b)
{
    hipLaunchKernelGGL(transpose_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, a, b);
    WB_HIP(hipGetLastError());
}
void symadd(hipStream_t s, int n, const double* a, double* c)
{
    hipLaunchKernelGGL(symadd_kernel, dim3((n + 255) / 256, n), dim3(256), 0, s, n, a, c);
    WB_HIP(hipGetLastError());
}

// y += a x (n doubles)
__global__ void axpy_kernel(size_t n, double a, const double* __restrict__ x, double* __restrict__ y)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = fma(a, x[i], y[i]);
}
void launch_axpy(hipStream_t s, size_t n, double a, const double* x, double* y)
{
    hipLaunchKernelGGL(axpy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, a, x, y);
    WB_HIP(hipGetLastError());
}

// Stratmann-Scuseria-Frisch partition (Chem. Phys. Lett. 257, 213 (1996)):
// for the point r owned by atom o, w = P_o / sum_a P_a with the cell function
// P_a = prod_{b != a} s(mu_ab), mu_ab = (|r - A| - |r - B|) / |A - B|,
// s = (1 - g(mu / a)) / 2, g(x) = (35x - 35x^3 + 21x^5 - 5x^7) / 16 on |x| < 1
// and sign(x) outside (a = 0.64).  Only atoms with |r - A| - |r - N| <
// a |A - N| (N the nearest atom) can have P_a > 0; the others are skipped.
// One thread per point; the atom index is uniform across the wave, so the
// atom coordinates and the distance tables come through scalar loads.
constexpr double SSF_A = 0.64;
__global__ __launch_bounds__(256) void ssf_partition_kernel(
    int npts, const double* __restrict__ xyz, const int* __restrict__ owner, int natm,
    const double* __restrict__ atoms, const double* __restrict__ rab,
    const double* __restrict__ inv_rab, double* __restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= npts) return;
    const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    const int own = owner[i];
    auto dist = [&](int b) {
        const double dx = x - atoms[3 * b], dy = y - atoms[3 * b + 1], dz = z - atoms[3 * b + 2];
        return sqrt(dx * dx + dy * dy + dz * dz);
    };
    double dn = dist(0);
    int near = 0;
    for (int b = 1; b < natm; ++b) {
        const double d = dist(b);
        if (d < dn) { dn = d; near = b; }
    }
    double tot = 0.0, pown = 0.0;
    for (int a = 0; a < natm; ++a) {
        const double da = dist(a);
        if (!(da - dn < SSF_A * rab[(size_t)near * natm + a] + 1e-12)) continue;
        double P = 1.0;
        const double* ir = inv_rab + (size_t)a * natm;
        for (int b = 0; b < natm; ++b) {
            if (b == a) continue;
            const double mu = (da - dist(b)) * ir[b];
            if (mu <= -SSF_A) continue;                  // s = 1
            if (mu >= SSF_A) { P = 0.0; break; }         // s = 0
            const double nu = mu / SSF_A, n2 = nu * nu;
            const double g = nu * (35.0 + n2 * (-35.0 + n2 * (21.0 - 5.0 * n2))) / 16.0;
            P *= 0.5 * (1.0 - g);
        }
        tot += P;
        if (a == own) pown = P;
    }
    out[i] = pown / tot;
}
void launch_ssf_partition(hipStream_t s, int npts, const double* xyz, const int* owner, int natm,
                          const double* atoms, const double* rab, const double* inv_rab,
                          double* out)
{
    if (npts > 0)
        hipLaunchKernelGGL(ssf_partition_kernel, dim3((npts + 255) / 256), dim3(256), 0, s, npts,
                           xyz, owner, natm, atoms, rab, inv_rab, out);
    WB_HIP(hipGetLastError());
}

// Atom-centred points on the device: raw point i lies on radial shell sh
// (shell_first[sh] <= i < shell_first[sh + 1]) of atom shell_atom[sh], at
// atom + r u_k with u_k, wa_k the k-th point of the shell's angular rule;
// its weight before the cell partition is wr wa_k.  Products and sums are
// rounded one by one (no contraction into fused multiply-adds), as a host
// build computes them.
__global__ __launch_bounds__(256) void grid_points_kernel(
    long long nraw, int nshell, const long long* __restrict__ shell_first,
    const int* __restrict__ shell_atom, const double* __restrict__ shell_r,
    const double* __restrict__ shell_w, const int* __restrict__ shell_rule,
    const int* __restrict__ rule_off, const double* __restrict__ rule_xyzw,
    const double* __restrict__ atoms, double* __restrict__ xyz, int* __restrict__ owner,
    double* __restrict__ wpre)
{
#pragma clang fp contract(off)
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= nraw) return;
    int lo = 0, hi = nshell - 1;                      // last shell with first <= i
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (shell_first[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const int a = shell_atom[lo];
    const double r = shell_r[lo];
    const double* u = rule_xyzw + 4 * ((size_t)rule_off[shell_rule[lo]] + (size_t)(i - shell_first[lo]));
    xyz[3 * i] = atoms[3 * a] + r * u[0];
    xyz[3 * i + 1] = atoms[3 * a + 1] + r * u[1];
    xyz[3 * i + 2] = atoms[3 * a + 2] + r * u[2];
    owner[i] = a;
    wpre[i] = shell_w[lo] * u[3];
}

// Order-preserving compaction of the points whose weight wpre x part is
// above w_cut, in three passes: kept points per block of 256, an exclusive
// scan of the block counts (one block), the scatter.
constexpr int GB_BLOCK = 256;
__global__ __launch_bounds__(GB_BLOCK) void grid_keep_count_kernel(
    long long n, const double* __restrict__ wpre, const double* __restrict__ part, double w_cut,
    int* __restrict__ bcount)
{
    __shared__ int s_cnt[GB_BLOCK / 64];
    const long long i = (long long)blockIdx.x * GB_BLOCK + threadIdx.x;
    const bool keep = i < n && __dmul_rn(wpre[i], part[i]) > w_cut;
    const unsigned long long m = __ballot(keep);
    if ((threadIdx.x & 63) == 0) s_cnt[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int k = 0; k < GB_BLOCK / 64; ++k) t += s_cnt[k];
        bcount[blockIdx.x] = t;
    }
}

// exclusive scan of nb block counts in place, the total at bcount[nb]
__global__ __launch_bounds__(1024) void grid_keep_scan_kernel(int nb, int* __restrict__ bcount)
{
    __shared__ int s_part[1024];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = t * per, b1 = min(nb, b0 + per);
    int sum = 0;
    for (int b = b0; b < b1; ++b) sum += bcount[b];
    s_part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {        // inclusive scan of the thread sums
        const int v = t >= off ? s_part[t - off] : 0;
        __syncthreads();
        s_part[t] += v;
        __syncthreads();
    }
    int run = t > 0 ? s_part[t - 1] : 0;
    for (int b = b0; b < b1; ++b) {
        const int c = bcount[b];
        bcount[b] = run;
        run += c;
    }
    if (t == 1023) bcount[nb] = s_part[1023];
}

__global__ __launch_bounds__(GB_BLOCK) void grid_keep_scatter_kernel(
    long long n, const double* __restrict__ xyz, const double* __restrict__ wpre,
    const double* __restrict__ part, double w_cut, const int* __restrict__ boff,
    double* __restrict__ xyz_out, double* __restrict__ w_out)
{
    __shared__ int s_cnt[GB_BLOCK / 64];
    const long long i = (long long)blockIdx.x * GB_BLOCK + threadIdx.x;
    const double w = i < n ? __dmul_rn(wpre[i], part[i]) : 0.0;
    const bool keep = i < n && w > w_cut;
    const unsigned long long m = __ballot(keep);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) s_cnt[wv] = __popcll(m);
    __syncthreads();
    if (!keep) return;
    int pos = boff[blockIdx.x];
    for (int k = 0; k < wv; ++k) pos += s_cnt[k];
    pos += __popcll(m & ((1ull << lane) - 1ull));
    xyz_out[3 * (size_t)pos] = xyz[3 * i];
    xyz_out[3 * (size_t)pos + 1] = xyz[3 * i + 1];
    xyz_out[3 * (size_t)pos + 2] = xyz[3 * i + 2];
    w_out[pos] = w;
}

void launch_grid_points(hipStream_t s, long long nraw, int nshell, const long long* shell_first,
                        const int* shell_atom, const double* shell_r, const double* shell_w,
                        const int* shell_rule, const int* rule_off, const double* rule_xyzw,
                        const double* atoms, double* xyz, int* owner, double* wpre)
{
    if (nraw > 0)
        hipLaunchKernelGGL(grid_points_kernel, dim3((unsigned)((nraw + 255) / 256)), dim3(256), 0,
                           s, nraw, nshell, shell_first, shell_atom, shell_r, shell_w, shell_rule,
                           rule_off, rule_xyzw, atoms, xyz, owner, wpre);
    WB_HIP(hipGetLastError());
}

int grid_keep_blocks(long long n) { return (int)((n + GB_BLOCK - 1) / GB_BLOCK); }

void launch_grid_keep(hipStream_t s, long long n, const double* xyz, const double* wpre,
                      const double* part, double w_cut, int* bcount /* blocks + 1 */,
                      double* xyz_out, double* w_out)
{
    const int nb = grid_keep_blocks(n);
    if (nb == 0) {
        zero_bytes(s, bcount, sizeof(int));
        return;
    }
    hipLaunchKernelGGL(grid_keep_count_kernel, dim3(nb), dim3(GB_BLOCK), 0, s, n, wpre, part,
                       w_cut, bcount);
    hipLaunchKernelGGL(grid_keep_scan_kernel, dim3(1), dim3(1024), 0, s, nb, bcount);
    hipLaunchKernelGGL(grid_keep_scatter_kernel, dim3(nb), dim3(GB_BLOCK), 0, s, n, xyz, wpre,
                       part, w_cut, bcount, xyz_out, w_out);
    WB_HIP(hipGetLastError());
}

// wv = (w v0 / 2, w v1, w v2, w v3, w v4 / 2) from the chain-ruled potential
// v [5][Ppad]: the operand factors of the semilocal potential matrix
__global__ void xc_weights_kernel(int Ppad, const double* __restrict__ w,
                                  const double* __restrict__ v, double* __restrict__ wv)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= Ppad) return;
    const double wg = w[g];
#pragma unroll
    for (int k = 0; k < 5; ++k)
        wv[(size_t)k * Ppad + g] = (k == 0 || k == 4 ? 0.5 : 1.0) * wg * v[(size_t)k * Ppad + g];
}
void launch_xc_weights(hipStream_t s, int Ppad, const double* w, const double* v, double* wv)
{
    hipLaunchKernelGGL(xc_weights_kernel, dim3((Ppad + 255) / 256), dim3(256), 0, s, Ppad, w, v, wv);
    WB_HIP(hipGetLastError());
}

// byte offsets of the VV10 flow's buffers in its workspace
VvLayout vv_layout(int nsub, int Ppad, bool f64)
{
    VvLayout L;
    const size_t rs = f64 ? 8 : 4;
    size_t o = 0;
    auto take = [&](size_t bytes) { size_t r = o; o += (bytes + 255) / 256 * 256; return r; };
    L.cnt = take(sizeof(int) * (nsub + 1));
    L.off = take(sizeof(int) * (nsub + 1));
    L.pt = take(4 * rs * Ppad);
    L.pk = take(2 * rs * Ppad);
    L.idx = take(sizeof(int) * Ppad);
    L.aux = take(sizeof(double) * 3 * Ppad);
    L.ts = take(sizeof(float4) * 2 * ((Ppad + 63) / 64));
    L.F = take(sizeof(double) * Ppad);
    L.U = take(sizeof(double) * Ppad);
    L.W = take(sizeof(double) * Ppad);
    L.sym = take(vv10_sym_bytes(Ppad, f64, nullptr));
    L.stat = take(2 * sizeof(unsigned long long));
    L.total = o;
    return L;
}

void vv10_flow(hipStream_t s, int nsub, int Ppad, const double* xyz, const double* w,
               const double* origin, const double* rho, double b, double C, double tol, bool f64,
               DevBuf& ws, double* wv, double* e, double* sums, Ctx* prof, int cap16)
{
    // workspace layout
    const VvLayout L = vv_layout(nsub, Ppad, f64);
    const size_t o_cnt = L.cnt, o_off = L.off, o_pt = L.pt, o_pk = L.pk, o_idx = L.idx,
                 o_aux = L.aux, o_ts = L.ts, o_F = L.F, o_U = L.U, o_W = L.W, o_sym = L.sym;
    const size_t ntile_max = (Ppad + 63) / 64;
    char* base = (char*)ws.ensure(L.total);
    int* cnt = (int*)(base + o_cnt);
    int* off = (int*)(base + o_off);
    int* idx = (int*)(base + o_idx);
    double* aux = (double*)(base + o_aux);
    float4* ts = (float4*)(base + o_ts);
    double *F = (double*)(base + o_F), *U = (double*)(base + o_U), *W = (double*)(base + o_W);
    const double thresh = 1e-8;
    const int wb = (nsub + 3) / 4;
    hipLaunchKernelGGL(vv_count_kernel, dim3(wb), dim3(256), 0, s, nsub, Ppad, rho, w, thresh, cnt);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(64), 0, s, nsub, cnt, off);
    const int* nptr = off + nsub;
    const double pi = M_PI;
    const double beta = pow(3.0 / (b * b), 0.75) / 32.0;
    if (f64) {
        auto* pt = (VP4<double>*)(base + o_pt);
        auto* pk = (VP2<double>*)(base + o_pk);
        hipLaunchKernelGGL(vv_point_kernel<double>, dim3(wb), dim3(256), 0, s, nsub, Ppad, xyz,
                           rho, w, thresh, b, C, origin[0], origin[1], origin[2], off, pt, pk,
                           idx, aux);
        hipLaunchKernelGGL(vv_tile_kernel<double>, dim3((ntile_max + 3) / 4), dim3(256), 0, s,
                           nptr, pt, pk, (int)ntile_max, ts);
        launch_vv10_sym(s, nptr, Ppad, pt, pk, ts, (float)tol, true, base + o_sym, F, U, W,
                        (unsigned long long*)(base + L.stat), cap16);
    } else {
        auto* pt = (VP4<float>*)(base + o_pt);
        auto* pk = (VP2<float>*)(base + o_pk);
        hipLaunchKernelGGL(vv_point_kernel<float>, dim3(wb), dim3(256), 0, s, nsub, Ppad, xyz,
                           rho, w, thresh, b, C, origin[0], origin[1], origin[2], off, pt, pk,
                           idx, aux);
        hipLaunchKernelGGL(vv_tile_kernel<float>, dim3((ntile_max + 3) / 4), dim3(256), 0, s,
                           nptr, pt, pk, (int)ntile_max, ts);
        launch_vv10_sym(s, nptr, Ppad, pt, pk, ts, (float)tol, false, base + o_sym, F, U, W,
                        (unsigned long long*)(base + L.stat), cap16);
    }
    (void)pi;
    if (prof) prof->mark(PH_VV10);
    zero_bytes(s, wv, sizeof(double) * 4 * Ppad);
    zero_bytes(s, e, sizeof(double) * Ppad);
    hipLaunchKernelGGL(vv_pot_kernel, dim3((Ppad + 255) / 256), dim3(256), 0, s, nptr, idx, Ppad,
                       F, U, W, aux, rho, w, beta, wv, e);
    hipLaunchKernelGGL(vv_sums_kernel, dim3(1), dim3(1024), 0, s, Ppad, rho, w, e, sums);
    WB_HIP(hipGetLastError());
}

}  // namespace wb
