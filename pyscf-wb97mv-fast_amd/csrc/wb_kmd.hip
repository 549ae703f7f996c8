// Fused McMurchie-Davidson COSX exchange kernels and their screening.
//
// For one shell-pair class (la >= lb) and one operator (1/r or erf(w r)/r),
// a wave owns one sub-block of 64 grid points (one point per lane) and one
// chunk of that sub-block's screened item list.  Per task (shell pair A, B)
// it forms the Cartesian block
//     V_ab(g) = sum_prims sum_tuv Ex Ey Ez R_tuv(p, P - g)
// in registers (never stored), and contracts it with F (F = X Q D):
//     G[g, B] += A(g)^T F[g, A],   G[g, A] += A(g) F[g, B]  (A != B)
// An item is one shell of one side with the kept pairs of one block of
// shells of the other side (a bit mask).  The blocked side is the one with
// the larger l (B when la = lb); items are ordered by block, so the block's
// F values and G accumulators stay in registers along a run of items and
// its G rows are written once per run, the item shell's once per item.  Each chunk writes its own copy of G (no atomics,
// fixed order, bit-reproducible); the copies are summed inside the K GEMM.
// The reference does the same contraction with one warp per 32-point
// sub-block and one task per shell pair (gpu/sgx_k.py:335-400).
#include "wb_boys.h"
#include "wb_internal.h"
#include <type_traits>
#include "md_gen.h"

namespace wb {

template <int L> struct MdDispatch;

// Read-only data that every lane of a wave reads at the same address (task
// lists, pair groups, block offsets, primitive-pair records) is read through
// the constant address space: the compiler then loads it with scalar loads
// into SGPRs whatever it can prove about aliasing.  Without this, such loads
// become vector loads whenever the pointers do not arrive as no-alias kernel
// arguments (the many-molecule queue kernel reads them from a job table): the
// values then occupy VGPRs, 30-90 more per class, and the kernels run at a
// fraction of their occupancy.
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ cptr<T> as_cptr(const T* p)
{
    return (cptr<T>)p;
}

// Classes whose Cartesian block is formed in two parts of the A components
// (without z, with z), one pass of the primitive loop each (md_nparts,
// generated): the large blocks otherwise leave one or two waves per SIMD.
// A components of part P (all of them when the class is not split)
template <int LA, int LB, int P> struct MdPartA {
    static constexpr __host__ __device__ int n() {
        if constexpr (md_nparts(LA, LB) == 1) return (LA + 1) * (LA + 2) / 2;
        else return MdPart<LA, LB, P>::n();
    }
    static constexpr __host__ __device__ int at(int i) {
        if constexpr (md_nparts(LA, LB) == 1) return i;
        else return MdPart<LA, LB, P>::at(i);
    }
};

// ONEC: the one-centre variant (A, B and P coincide); PART: the part of a
// split class (V holds that part's rows in MdPartA order)
template <int LA, int LB, bool ONEC, int PART, typename real, typename vacc>
__device__ __forceinline__ void md_prim(real dx, real dy, real dz, const real* R0,
                                        cptr<real> E, vacc* V)
{
#define WB_MD_PRIM(a, b)                                                              \
    if constexpr (LA == a && LB == b) {                                               \
        if constexpr (ONEC) md_prim_##a##b##_oc<real, vacc>(dx, dy, dz, R0, E, V);    \
        else md_prim_##a##b##_tc<real, vacc>(dx, dy, dz, R0, E, V);                   \
    }
#define WB_MD_PART(a, b, p)                                                           \
    if constexpr (LA == a && LB == b && PART == p) {                                  \
        if constexpr (ONEC) md_prim_##a##b##_oc_p##p<real, vacc>(dx, dy, dz, R0, E, V); \
        else md_prim_##a##b##_tc_p##p<real, vacc>(dx, dy, dz, R0, E, V);              \
    }
    if constexpr (md_nparts(LA, LB) == 1) {
        WB_MD_PRIM(0, 0) WB_MD_PRIM(1, 0) WB_MD_PRIM(1, 1) WB_MD_PRIM(2, 0) WB_MD_PRIM(2, 1)
        WB_MD_PRIM(3, 0) WB_MD_PRIM(3, 1) WB_MD_PRIM(3, 3)
    } else {
        WB_MD_PART(2, 2, 0) WB_MD_PART(2, 2, 1) WB_MD_PART(3, 2, 0) WB_MD_PART(3, 2, 1)
    }
#undef WB_MD_PART
#undef WB_MD_PRIM
}

// the Cartesian blocks of one shell pair at this lane's grid point, summed
// over the pair's primitive pairs, for NOP operators at once (they share
// P - g, the E coefficients and every load; Vs holds NOP blocks)
template <int LA, int LB, bool ONEC, int NOP, int PART, typename real, typename vacc>
__device__ __forceinline__ void md_block(int pp0, int npp, double px, double py, double pz,
                                         cptr<char> recs,
                                         const real* __restrict__ tab, vacc* Vs)
{
    constexpr int L = LA + LB;
    constexpr bool F32 = sizeof(real) == 4;
    constexpr int CRL = md_crl(LA, LB, F32);
    constexpr int RB = md_rec_bytes(LA, LB, F32, NOP);
    constexpr int NV = MdPartA<LA, LB, PART>::n() * ((LB + 1) * (LB + 2) / 2);
    // unrolled by two: the second primitive pair's scalar loads issue with
    // the first's (measured: exchange -1.2 %)
#pragma unroll 2
    for (int m = pp0; m < pp0 + npp; ++m) {
        const cptr<char> rec = recs + (size_t)m * RB;
        const cptr<double> P = (cptr<double>)rec;
        const cptr<real> C = (cptr<real>)(rec + 24);
        const real dx = (real)(P[0] - px);
        const real dy = (real)(P[1] - py);
        const real dz = (real)(P[2] - pz);
        const real r2 = fma(dx, dx, fma(dy, dy, dz * dz));
#pragma unroll
        for (int o = 0; o < NOP; ++o) {
            const cptr<real> Co = C + o * CRL;
            real Fb[L + 1], R0[L + 1];
            if constexpr (F32) {
                Boys<real>::template eval<L>(fmaf(Co[0], r2, Co[1] * r2), tab, Fb);
#pragma unroll
                for (int n = 0; n <= L; ++n) R0[n] = fmaf(Co[2 + 2 * n], Fb[n], Co[3 + 2 * n] * Fb[n]);
            } else {
                Boys<real>::template eval<L>(Co[0] * r2, tab, Fb);
#pragma unroll
                for (int n = 0; n <= L; ++n) R0[n] = Co[1 + n] * Fb[n];
            }
            md_prim<LA, LB, ONEC, PART, real, vacc>(dx, dy, dz, R0, C + NOP * CRL, Vs + o * NV);
        }
    }
}

template <int l, typename real>
__device__ __forceinline__ void fcart(const real* Fs, real* Fc) {
    if constexpr (l == 0) md_fcart_0<real>(Fs, Fc);
    else if constexpr (l == 1) md_fcart_1<real>(Fs, Fc);
    else if constexpr (l == 2) md_fcart_2<real>(Fs, Fc);
    else md_fcart_3<real>(Fs, Fc);
}
template <int l, typename real>
__device__ __forceinline__ void gsph(const real* Gc, real* Gs) {
    if constexpr (l == 0) md_gsph_0<real>(Gc, Gs);
    else if constexpr (l == 1) md_gsph_1<real>(Gc, Gs);
    else if constexpr (l == 2) md_gsph_2<real>(Gc, Gs);
    else md_gsph_3<real>(Gc, Gs);
}

constexpr int MD_WAVES = 4;

// The Boys table of order L staged in LDS by the block (FP32 only)
template <int L, typename real>
__device__ __forceinline__ const real* md_stage_boys(const real* __restrict__ boys_g, real* s_tab)
{
    if constexpr (sizeof(real) == 4) {
        constexpr int TAB = B32_NODES * b32_row(L);
        const real* src = boys_g + b32_off(L);
        for (int i = threadIdx.x; i < TAB; i += blockDim.x) s_tab[i] = src[i];
        __syncthreads();
        return s_tab;
    }
    return boys_g;
}

// One wave's work in one class launch: sub-block s (one point per lane) and
// chunk j of s's task list, of one molecule (its F, G copies, task lists and
// records).  real: integral arithmetic; gacc: G accumulation (registers and
// memory); NOP operators at once (G of operator o at G + o opstride).  The
// single-molecule kernel and the many-molecule queue kernel run exactly this.
// The pointers are __restrict__ parameters (the qualifier on local variables
// is ignored).
template <int LA, int LB, int NOP, typename real, typename gacc>
__device__ __forceinline__ void md_wave(
    const double* __restrict__ xyz, const real* __restrict__ F, gacc* __restrict__ G,
    const size_t gstride, const size_t opstride, const cptr<int> counts,
    const cptr<int> ids, const cptr<unsigned> skips, const int cap,
    const cptr<PairGroup> groups, const cptr<int> bbao,
    const cptr<char> recs, const int Ppad, const int nsub, const int nchunk,
    const int zero_rows, unsigned* __restrict__ masks, const size_t mopstride, const int zrow,
    const int first, const real* __restrict__ tab, const int s, const int j)
{
    constexpr int L = LA + LB;
    constexpr int NCA = (LA + 1) * (LA + 2) / 2, NCB = (LB + 1) * (LB + 2) / 2;
    constexpr int NSA = 2 * LA + 1, NSB = 2 * LB + 1;
    constexpr int BB = md_bb(LA, LB);
    constexpr bool F32 = sizeof(real) == 4;
    // the primitive sum of the Cartesian block is FP64 when it fits
    // (<= 36 entries); the larger f-shell blocks belong to single-primitive
    // shells in the def2 sets and stay in the arithmetic type
    using vacc = typename std::conditional<(NCA * NCB <= 36), double, real>::type;
    (void)NSA; (void)NSB; (void)L; (void)F32;
    const int lane = threadIdx.x & 63;
    gacc* const Gc = G + (size_t)j * gstride;         // operator o: Gc + o opstride
    const int g = s * SUB + lane;
    // This wave alone owns rows x its 64 points of copy j in every class
    // launch.  With masks, G is never cleared: a bitmask of the shells this
    // region has received (lane w holds rows 32w .. 32w + 31, a shell marked
    // at its first row; carried between the class launches in masks) sends
    // the first read-modify-write of a shell to the zero row zrow instead.
    // Without masks, the first class launched clears the region.
    const bool tracked = masks != nullptr;
    unsigned* const mw = tracked ? masks + ((size_t)j * nsub + s) * 64 : nullptr;
    unsigned trow = 0;
    if (tracked) {
        if (first) {
#pragma unroll
            for (int o = 0; o < NOP; ++o) Gc[o * opstride + (size_t)zrow * Ppad + g] = (gacc)0;
        } else {
            trow = mw[lane];
        }
    } else if (zero_rows > 0) {
#pragma unroll
        for (int o = 0; o < NOP; ++o)
            for (int r = 0; r < zero_rows; ++r) Gc[o * opstride + (size_t)r * Ppad + g] = (gacc)0;
    }
    auto store_mask = [&]() {
#pragma unroll
        for (int o = 0; o < NOP; ++o) mw[o * mopstride + lane] = trow;
    };
    // r: a shell's first row (wave-uniform); the row a read of its row r + m
    // takes: itself once written, else the zero row
    auto src_row = [&](int r, int m) -> int {
        if (!tracked) return r + m;
        const unsigned wv = (unsigned)__builtin_amdgcn_readlane((int)trow, r >> 5);
        return (wv >> (r & 31) & 1u) ? r + m : zrow;
    };
    auto mark = [&](int r) {
        if (tracked) trow |= (lane == (r >> 5)) ? (1u << (r & 31)) : 0u;
    };
    const int cnt = counts[s];
    const int t0 = (int)(((long long)cnt * j) / nchunk);
    const int t1 = (int)(((long long)cnt * (j + 1)) / nchunk);
    if (t0 >= t1) {
        if (tracked && first) store_mask();
        return;
    }
    const cptr<int> tl = ids + (size_t)s * cap;
    const cptr<unsigned> sl = skips + (size_t)s * cap;
    const double px = xyz[g], py = xyz[Ppad + g], pz = xyz[2 * Ppad + g];
    // wait for the point's coordinates here, once: otherwise the first use
    // inside the item loop waits on the vector-memory counter in program
    // order, i.e. also on the next item's prefetched F and G rows
    __builtin_amdgcn_s_waitcnt(0);

    // the blocked side K (the shells of one block: F in Cartesian form and
    // G accumulators in registers) and the item side I (one shell per item)
    constexpr bool BA = md_block_a(LA, LB);
    constexpr int LK = BA ? LA : LB, LI = BA ? LB : LA;
    constexpr int NCK = (LK + 1) * (LK + 2) / 2, NCI = (LI + 1) * (LI + 2) / 2;
    constexpr int NSK = 2 * LK + 1, NSI = 2 * LI + 1;
    constexpr int NV = NCA * NCB;
    int curB = -1;
    real FkC[BB][NCK];
    gacc accK[NOP][BB][NCK];
    // G[K] += the block's accumulators; every position is loaded and stored,
    // untouched ones unchanged (absent positions use the spare rows):
    // branch-free, so the loads issue as one batch and no store waits for
    // more than its own row
    auto flush_k = [&]() {
        int bo[BB];
#pragma unroll
        for (int q = 0; q < BB; ++q) bo[q] = bbao[curB * BB + q];
        gacc old[NOP][BB][NSK];
#pragma unroll
        for (int o = 0; o < NOP; ++o)
#pragma unroll
            for (int q = 0; q < BB; ++q)
#pragma unroll
                for (int m = 0; m < NSK; ++m)
                    old[o][q][m] = Gc[o * opstride + (size_t)src_row(bo[q], m) * Ppad + g];
#pragma unroll
        for (int o = 0; o < NOP; ++o)
#pragma unroll
            for (int q = 0; q < BB; ++q) {
                gacc gs[NSK];
                gsph<LK, gacc>(accK[o][q], gs);
#pragma unroll
                for (int m = 0; m < NSK; ++m)
                    Gc[o * opstride + (size_t)(bo[q] + m) * Ppad + g] = old[o][q][m] + gs[m];
            }
#pragma unroll
        for (int q = 0; q < BB; ++q) mark(bo[q]);
    };
    // software pipeline: while an item computes, the next item's F rows and
    // (inside one block run, where nothing else writes them before its own
    // update) its G rows are already on their way
    int item = tl[t0];
    real fi_n[NSI];
    gacc gi_n[NOP][NSI];
    bool gi_ok = false;
    {
        const int x0 = groups[item >> 8].x0;
#pragma unroll
        for (int m = 0; m < NSI; ++m) fi_n[m] = F[(size_t)(x0 + m) * Ppad + g];
    }
    for (int t = t0; t < t1; ++t) {
        const cptr<PairGroup> gr = groups + (item >> 8);
        const unsigned mask = (unsigned)item & 0xffu;
        const unsigned skw = sl[t];                          // trailing primitive pairs skipped
        const int bblk = gr->bblk;
        const int x0 = gr->x0;
        real fi[NSI];
        gacc gi[NOP][NSI];
        const bool have_gi = gi_ok;
#pragma unroll
        for (int m = 0; m < NSI; ++m) fi[m] = fi_n[m];
#pragma unroll
        for (int o = 0; o < NOP; ++o)
#pragma unroll
            for (int m = 0; m < NSI; ++m) gi[o][m] = gi_n[o][m];
        int item_n = 0;
        if (t + 1 < t1) {                                    // uniform
            item_n = tl[t + 1];
            const cptr<PairGroup> gn = groups + (item_n >> 8);
            const int x0n = gn->x0;
#pragma unroll
            for (int m = 0; m < NSI; ++m) fi_n[m] = F[(size_t)(x0n + m) * Ppad + g];
            // item shells within a run are distinct and a run change
            // flushes the block first: only a same-run prefetch is valid
            gi_ok = gn->bblk == bblk;
            // (the shell's written state cannot change before its own
            // update: nothing else in the run writes it)
            if (gi_ok) {
#pragma unroll
                for (int o = 0; o < NOP; ++o)
#pragma unroll
                    for (int m = 0; m < NSI; ++m)
                        gi_n[o][m] = Gc[o * opstride + (size_t)src_row(x0n, m) * Ppad + g];
            }
        }
        if (bblk != curB) {
            if (curB >= 0) flush_k();
            curB = bblk;
            // the block's F rows, all loads back to back (absent positions
            // read a spare row; their values are never used: no mask bit)
            int bo[BB];
#pragma unroll
            for (int q = 0; q < BB; ++q) bo[q] = bbao[curB * BB + q];
            real fk[BB][NSK];
#pragma unroll
            for (int q = 0; q < BB; ++q)
#pragma unroll
                for (int m = 0; m < NSK; ++m) fk[q][m] = F[(size_t)(bo[q] + m) * Ppad + g];
#pragma unroll
            for (int q = 0; q < BB; ++q) {
                fcart<LK, real>(fk[q], FkC[q]);
#pragma unroll
                for (int o = 0; o < NOP; ++o)
#pragma unroll
                    for (int c = 0; c < NCK; ++c) accK[o][q][c] = (gacc)0;
            }
        }
        real FiC[NCI];
        fcart<LI, real>(fi, FiC);
        gacc accI[NOP][NCI];
#pragma unroll
        for (int o = 0; o < NOP; ++o)
#pragma unroll
            for (int c = 0; c < NCI; ++c) accI[o][c] = (gacc)0;
        bool anyI = false;
        // the group's packed task words: no dependent pair-record load
        int pw[BB];
#pragma unroll
        for (int q = 0; q < BB; ++q) pw[q] = gr->pw[q];
#pragma unroll
        for (int q = 0; q < BB; ++q) {
            if (!(mask >> q & 1u)) continue;                 // uniform
            const int pp0 = (int)((unsigned)pw[q] >> 10);
            const int npp = (int)(((unsigned)pw[q] >> 2) & 255u) + 1 - (int)(skw >> (4 * q) & 15u);
            const bool onec = pw[q] & 2, diag = pw[q] & 1;
            const real* FaC = BA ? FkC[q] : FiC;
            const real* FbC = BA ? FiC : FkC[q];
            // the block (or each part of it): formed over the primitive pairs,
            // then G[B] += V^T F[A], G[A] += V F[B]; for a diagonal pair
            // (A = B, blocked side B) only the first
            auto part_block = [&](auto pc) {
                constexpr int P = decltype(pc)::value;
                using PA = MdPartA<LA, LB, P>;
                constexpr int NVP = PA::n() * NCB;
                vacc Vs[NOP * NVP];
#pragma unroll
                for (int i = 0; i < NOP * NVP; ++i) Vs[i] = (vacc)0;
                // one-centre pairs take the variant whose FP32 E products all
                // carry their rounding remainders
                if (onec)
                    md_block<LA, LB, true, NOP, P, real, vacc>(pp0, npp, px, py, pz, recs, tab, Vs);
                else
                    md_block<LA, LB, false, NOP, P, real, vacc>(pp0, npp, px, py, pz, recs, tab, Vs);
#pragma unroll
                for (int o = 0; o < NOP; ++o) {
                    real V[NVP];
#pragma unroll
                    for (int i = 0; i < NVP; ++i) V[i] = (real)Vs[o * NVP + i];
                    gacc* gA = BA ? accK[o][q] : accI[o];
                    gacc* gB = BA ? accI[o] : accK[o][q];
#pragma unroll
                    for (int bb = 0; bb < NCB; ++bb) {
                        real v = (real)0;
#pragma unroll
                        for (int i = 0; i < PA::n(); ++i) v = fma(V[i * NCB + bb], FaC[PA::at(i)], v);
                        gB[bb] += (gacc)v;
                    }
                    if (BA || !diag) {                       // uniform
#pragma unroll
                        for (int i = 0; i < PA::n(); ++i) {
                            real v = (real)0;
#pragma unroll
                            for (int bb = 0; bb < NCB; ++bb) v = fma(V[i * NCB + bb], FbC[bb], v);
                            gA[PA::at(i)] += (gacc)v;
                        }
                    }
                }
            };
            part_block(std::integral_constant<int, 0>{});
            if constexpr (md_nparts(LA, LB) > 1) part_block(std::integral_constant<int, 1>{});
            anyI = anyI || BA || !diag;
        }
        if (anyI) {
            if (!have_gi) {
#pragma unroll
                for (int o = 0; o < NOP; ++o)
#pragma unroll
                    for (int m = 0; m < NSI; ++m)
                        gi[o][m] = Gc[o * opstride + (size_t)src_row(x0, m) * Ppad + g];
            }
#pragma unroll
            for (int o = 0; o < NOP; ++o) {
                gacc gs[NSI];
                gsph<LI, gacc>(accI[o], gs);
#pragma unroll
                for (int m = 0; m < NSI; ++m)
                    Gc[o * opstride + (size_t)(x0 + m) * Ppad + g] = gi[o][m] + gs[m];
            }
            mark(x0);
        }
        item = item_n;
    }
    if (curB >= 0) flush_k();
    if (tracked) store_mask();
}

// worker w of a class launch (nsub sub-blocks, nchunk waves each) -> its
// sub-block s and chunk j: chunk-major (w = j nsub + rank), and within each
// chunk's pass the sub-blocks by rank in order (the launch's order: most kept
// primitive pairs first, so the tail of the launch holds short waves), else in
// grid order.  Measured on (H2O)27 def2-SVP, exchange both operators: grid
// order 19.70 ms, this order 18.61 ms; rank-major (a sub-block's chunks back
// to back) 20.72 ms in grid order; mapping consecutive blocks to one XCD
// 20.12 ms (it gathers the dense regions' waves on a few XCDs).
__device__ __forceinline__ void md_worker(long long w, int nsub, cptr<int> order, int& s, int& j)
{
    const int r = (int)(w % nsub);
    j = (int)(w / nsub);
    s = order != nullptr ? order[r] : r;
}

// md_wave on the molecule a describes (its pointers passed as the
// __restrict__ parameters)
template <int LA, int LB, int NOP, typename real, typename gacc>
__device__ __forceinline__ void md_wave_job(const MdJob& a, const real* __restrict__ tab, int s,
                                            int j)
{
    md_wave<LA, LB, NOP, real, gacc>(a.xyz, (const real*)a.F, (gacc*)a.G, a.gstride, a.opstride,
                                     as_cptr(a.counts), as_cptr(a.ids), as_cptr(a.skips), a.cap,
                                     as_cptr(a.groups), as_cptr(a.bbao), as_cptr(a.recs),
                                     a.Ppad, a.nsub, a.nchunk, a.zero_rows, a.masks, a.mopstride,
                                     a.zrow, a.first, tab, s, j);
}

// one class, one molecule: one wave per worker (md_worker).  The pointers
// are kernel arguments of their own (not fields of a by-value MdJob): with the
// job as one struct argument the same body took 30-90 more VGPRs per class
// (the (d,d) kernel one wave per SIMD instead of three) and the exchange
// phase twice the time.
template <int LA, int LB, int NOP, typename real, typename gacc>
__global__ __launch_bounds__(64 * MD_WAVES)
void md_k_kernel(const double* __restrict__ xyz, const real* __restrict__ F,
                 gacc* __restrict__ G, size_t gstride, size_t opstride,
                 const int* __restrict__ counts, const int* __restrict__ ids,
                 const unsigned* __restrict__ skips, int cap,
                 const PairGroup* __restrict__ groups, const int* __restrict__ bbao,
                 const char* __restrict__ recs,
                 const real* __restrict__ boys_g, int Ppad, int nsub, int nchunk,
                 int zero_rows, unsigned* __restrict__ masks, size_t mopstride, int zrow,
                 int first, const int* __restrict__ order)
{
    constexpr int TAB = sizeof(real) == 4 ? B32_NODES * b32_row(LA + LB) : 1;
    __shared__ __align__(16) real s_tab[TAB];
    const real* tab = md_stage_boys<LA + LB, real>(boys_g, s_tab);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int worker = blockIdx.x * MD_WAVES + wave;
    if (worker >= nsub * nchunk) return;
    int s, j;
    md_worker(worker, nsub, as_cptr(order), s, j);
    md_wave<LA, LB, NOP, real, gacc>(xyz, F, G, gstride, opstride, as_cptr(counts), as_cptr(ids),
                                     as_cptr(skips), cap, as_cptr(groups), as_cptr(bbao),
                                     as_cptr(recs), Ppad, nsub, nchunk, zero_rows, masks,
                                     mopstride, zrow, first, tab, s, j);
}

// one class, many molecules: a persistent grid whose waves take work items
// from one device counter until all are taken; item i belongs to the
// molecule m with first[m] <= i < first[m + 1] and is its wave
// (i - first[m]) of the single-molecule launch above, so every molecule's G
// is exactly the one its own launch writes.  Every wave leaves the loop once
// the counter has passed total.
template <int LA, int LB, int NOP, typename real, typename gacc>
__global__ __launch_bounds__(64 * MD_WAVES)
void md_queue_kernel(const MdJob* __restrict__ jobs, const long long* __restrict__ first,
                     int njob, long long total, unsigned long long* __restrict__ head,
                     const real* __restrict__ boys_g)
{
    constexpr int TAB = sizeof(real) == 4 ? B32_NODES * b32_row(LA + LB) : 1;
    __shared__ __align__(16) real s_tab[TAB];
    const real* tab = md_stage_boys<LA + LB, real>(boys_g, s_tab);
    const int lane = threadIdx.x & 63;
    for (;;) {
        unsigned long long it = 0;
        if (lane == 0) it = atomicAdd(head, 1ull);
        const unsigned lo32 = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)it);
        const unsigned hi32 = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(it >> 32));
        const long long item = (long long)(((unsigned long long)hi32 << 32) | lo32);
        if (item >= total) break;
        int lo = 0, hi = njob - 1;                    // last job with first[m] <= item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (first[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const MdJob a = jobs[lo];
        int s, j;
        md_worker(item - first[lo], a.nsub, as_cptr(a.order), s, j);
        md_wave_job<LA, LB, NOP, real, gacc>(a, tab, s, j);
    }
}

MdJob md_job(const ClassData& c, const Grid& g, const void* F, void* G, size_t gstride,
             size_t opstride, int nchunk, const int* counts, const int* ids,
             const unsigned* skips, const void* recs, int zero_rows, const MdMasks& mk)
{
    MdJob j{};
    j.xyz = g.d_xyz.as<double>();
    j.F = F;
    j.G = G;
    j.gstride = gstride;
    j.opstride = opstride;
    j.counts = counts;
    j.ids = ids;
    j.skips = skips;
    j.groups = c.d_groups.as<PairGroup>();
    j.bbao = c.d_bbao.as<int>();
    j.recs = (const char*)recs;
    j.masks = mk.m;
    j.mopstride = mk.opstride;
    j.cap = c.ngroups;
    j.Ppad = g.Ppad;
    j.nsub = g.nsub;
    j.nchunk = nchunk;
    j.zero_rows = zero_rows;
    j.zrow = mk.zrow;
    j.first = mk.first ? 1 : 0;
    return j;
}

// resident blocks of a queue kernel on the current device (all its CUs)
template <typename K>
static int queue_blocks(K kernel)
{
    int dev = 0, cus = 0, per = 0;
    WB_HIP(hipGetDevice(&dev));
    WB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    WB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 64 * MD_WAVES, 0));
    return std::max(1, cus) * std::max(1, per);
}

template <int LA, int LB, int NOP, typename real, typename gacc>
static void md_launch_t(hipStream_t s, const MdJob* j, const MdJob* jobs, const long long* first,
                        int njob, long long total, unsigned long long* head, const void* boys)
{
    if (j) {
        const long long workers = md_items(*j);
        const int blocks = (int)((workers + MD_WAVES - 1) / MD_WAVES);
        const MdJob& a = *j;
        hipLaunchKernelGGL((md_k_kernel<LA, LB, NOP, real, gacc>), dim3(blocks),
                           dim3(64 * MD_WAVES), 0, s, a.xyz, (const real*)a.F, (gacc*)a.G,
                           a.gstride, a.opstride, a.counts, a.ids, a.skips, a.cap, a.groups,
                           a.bbao, a.recs, (const real*)boys, a.Ppad, a.nsub, a.nchunk,
                           a.zero_rows, a.masks, a.mopstride, a.zrow, a.first, a.order);
    } else {
        auto kern = md_queue_kernel<LA, LB, NOP, real, gacc>;
        static const int resident = queue_blocks(kern);     // one device kind per process
        const long long need = (total + MD_WAVES - 1) / MD_WAVES;
        const int blocks = (int)std::min<long long>(need, resident);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(64 * MD_WAVES), 0, s, jobs, first, njob,
                           total, head, (const real*)boys);
    }
    WB_HIP(hipGetLastError());
}

// one class launch, single molecule (j) or queue (jobs ...)
static void md_dispatch(hipStream_t s, int la, int lb, bool f32, bool g64, int nop,
                        const MdJob* j, const MdJob* jobs, const long long* first, int njob,
                        long long total, unsigned long long* head, const void* boys)
{
    if (nop != 1 && !(nop == 2 && md_fuse(la, lb)))
        throw Error(WB_INVALID, "operator count not available for this shell-pair class");
#define WB_LAUNCH(XA, XB, NOPV)                                                                    \
    if (f32 && !g64)                                                                               \
        md_launch_t<XA, XB, NOPV, float, float>(s, j, jobs, first, njob, total, head, boys);       \
    else if (f32)                                                                                  \
        md_launch_t<XA, XB, NOPV, float, double>(s, j, jobs, first, njob, total, head, boys);      \
    else                                                                                           \
        md_launch_t<XA, XB, NOPV, double, double>(s, j, jobs, first, njob, total, head, boys);
#define WB_CASE(XA, XB)                                                                    \
    if (la == XA && lb == XB) {                                                            \
        if constexpr (md_fuse(XA, XB)) {                                                   \
            if (nop == 2) { WB_LAUNCH(XA, XB, 2) return; }                                 \
        }                                                                                  \
        WB_LAUNCH(XA, XB, 1)                                                               \
        return;                                                                            \
    }
    WB_CASE(0, 0) WB_CASE(1, 0) WB_CASE(1, 1) WB_CASE(2, 0) WB_CASE(2, 1)
    WB_CASE(2, 2) WB_CASE(3, 0) WB_CASE(3, 1) WB_CASE(3, 2) WB_CASE(3, 3)
#undef WB_CASE
#undef WB_LAUNCH
    throw Error(WB_UNSUPPORTED, "shell-pair class beyond (f, f)");
}

void launch_md(hipStream_t s, int la, int lb, bool f32, bool g64, int nop, const MdJob& j,
               const void* boys)
{
    if (md_items(j) == 0) return;
    md_dispatch(s, la, lb, f32, g64, nop, &j, nullptr, nullptr, 0, 0, nullptr, boys);
}

void launch_md_queue(hipStream_t s, int la, int lb, bool f32, bool g64, int nop,
                     const MdJob* jobs, const long long* first, int njob, long long total,
                     unsigned long long* head, const void* boys)
{
    if (njob == 0 || total == 0) return;
    md_dispatch(s, la, lb, f32, g64, nop, nullptr, jobs, first, njob, total, head, boys);
}

// F_0..F_L(T) of the device Boys routines for a list of T (exposed for
// tests: the same code the exchange kernels inline)
template <int L, typename real>
__global__ __launch_bounds__(256)
void boys_eval_kernel(int n, const double* __restrict__ T, const real* __restrict__ tab,
                      double* __restrict__ out)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    real F[L + 1];
    Boys<real>::template eval<L>((real)T[i], tab, F);
#pragma unroll
    for (int m = 0; m <= L; ++m) out[(size_t)m * n + i] = (double)F[m];
}

void launch_boys_eval(hipStream_t s, int L, int n, const double* T, bool f32,
                      const void* b32, const void* b64, double* out)
{
    if (n == 0) return;
    const int blocks = (n + 255) / 256;
#define WB_BCASE(XL)                                                                     \
    if (L == XL) {                                                                       \
        if (f32)                                                                         \
            hipLaunchKernelGGL((boys_eval_kernel<XL, float>), dim3(blocks), dim3(256), 0, s, n, T, \
                               (const float*)b32 + b32_off(XL), out);                    \
        else                                                                             \
            hipLaunchKernelGGL((boys_eval_kernel<XL, double>), dim3(blocks), dim3(256), 0, s, n, T, \
                               (const double*)b64, out);                                 \
        WB_HIP(hipGetLastError());                                                       \
        return;                                                                          \
    }
    WB_BCASE(0) WB_BCASE(1) WB_BCASE(2) WB_BCASE(3) WB_BCASE(4) WB_BCASE(5) WB_BCASE(6)
#undef WB_BCASE
    throw Error(WB_INVALID, "Boys order beyond 6");
}

// ------------------------------------------------------------ screening ----
// fmax[s][sh] = max_{g in s} w_g * sum_{mu in sh} |F[mu][g]|
// The shells of a sub-block are split over FMAX_SPLIT waves (one point per
// lane); a wave takes 8 shells per pass: all their rows' loads first (they
// are independent), then 8 interleaved wave reductions.  Measured on
// (H2O)27 def2-SVP (F 566 MB, FP32): 0.38 ms with one shell and one
// reduction at a time per sub-block, 0.32 ms this way; what remains is the
// read of F in 256-byte row pieces.
constexpr int FMAX_SPLIT = 4;
template <typename real>
__global__ __launch_bounds__(256)
void fmax_kernel(const real* __restrict__ F, const double* __restrict__ w, int Ppad,
                 int nsub, const int* __restrict__ sh_ao0, const int* __restrict__ sh_nao,
                 int nshell, float* __restrict__ fmax)
{
    const int wid = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wid >= nsub * FMAX_SPLIT) return;
    const int s = wid / FMAX_SPLIT, q = wid % FMAX_SPLIT;
    const int shb = (int)((long long)nshell * q / FMAX_SPLIT);
    const int she = (int)((long long)nshell * (q + 1) / FMAX_SPLIT);
    const int lane = threadIdx.x & 63;
    const int g = s * SUB + lane;
    const double wg = w[g];
    for (int sh0 = shb; sh0 < she; sh0 += 8) {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int sh = sh0 + k;
            double acc = 0.0;
            if (sh < she) {                              // uniform
                const int a0 = sh_ao0[sh], n = sh_nao[sh];
                for (int m = 0; m < n; ++m) acc += fabs((double)F[(size_t)(a0 + m) * Ppad + g]);
            }
            v[k] = (float)(acc * wg);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int k = 0; k < 8; ++k) v[k] = fmaxf(v[k], __shfl_xor(v[k], o, 64));
        if (lane < 8 && sh0 + lane < she) {
            float r = v[0];
#pragma unroll
            for (int k = 1; k < 8; ++k) r = lane == k ? v[k] : r;
            fmax[(size_t)s * nshell + sh0 + lane] = r * 1.0001f;
        }
    }
}

void launch_fmax(hipStream_t s, const void* F, bool f32, const double* w, int nao, int Ppad,
                 int nsub, const int* sh_ao0, const int* sh_nao, int nshell, float* fmax)
{
    const int blocks = (nsub * FMAX_SPLIT + 3) / 4;
    if (f32)
        hipLaunchKernelGGL(fmax_kernel<float>, dim3(blocks), dim3(256), 0, s,
                           (const float*)F, w, Ppad, nsub, sh_ao0, sh_nao, nshell, fmax);
    else
        hipLaunchKernelGGL(fmax_kernel<double>, dim3(blocks), dim3(256), 0, s,
                           (const double*)F, w, Ppad, nsub, sh_ao0, sh_nao, nshell, fmax);
    WB_HIP(hipGetLastError());
}

// One wave per (sub-block): the kept groups of one class, in group order,
// and per kept item the trailing primitive pairs each of its tasks skips
// (skips: 4 bits per block position).
// est = beta * F0bound(q d_min^2) * max(fmax[A], fmax[B]); F0(x) <= min(1,
// sqrt(pi/(4x))).  A task is kept when est >= tol (the reference's
// shell-pair screening, gpu/sgx_k.py:866-924).  The pair's primitive pairs
// are stored most important first; pbc[m] = (centre, c_m) with c_m >=
// beta_m sqrt(pi) / (2 sqrt(q_m)), so primitive pair m contributes at most
// c_m fmax / d_m at a distance d_m > 0 from the sub-block: a kept task skips
// the longest tail whose bounds sum to at most tol / 2 (never its first
// primitive pair, at most 15).  tol <= 0 keeps every pair and primitive.
// work[0] += kept tasks, work[1] += their evaluated primitive pairs (integer
// atomics: the totals do not depend on the order)
__global__ __launch_bounds__(256)
void task_kernel(TaskLaunch tl, const float4* __restrict__ sub, int nsub,
                 const float* __restrict__ fmax, int nshell, float tol)
{
    // every class with pairs in one launch (blockIdx.y: the class's entry),
    // so the light classes' waves hide the heavy classes' gather latency
    const TaskClass& tc = tl.c[blockIdx.y];
    const PairScreen* __restrict__ scr = tc.scr;
    const PairGroup* __restrict__ groups = tc.groups;
    const int ngroups = tc.ngroups, bb = tc.bb;
    const float4* __restrict__ pbc = tc.pbc;
    int* __restrict__ counts = tc.counts;
    int* __restrict__ ids = tc.ids;
    unsigned* __restrict__ skips = tc.skips;
    unsigned long long* __restrict__ work = tc.work;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int s = blockIdx.x * 4 + wave;
    if (s >= nsub) return;
    const int lane = threadIdx.x & 63;
    const float4 sb = sub[s];
    const float* fm = fmax + (size_t)s * nshell;
    int* out = ids + (size_t)s * ngroups;
    unsigned* sout = skips + (size_t)s * ngroups;
    int n = 0;
    unsigned long long ntask = 0, npp = 0;
    for (int k0 = 0; k0 < ngroups; k0 += 64) {
        const int k = k0 + lane;
        unsigned mask = 0, skw = 0;
        if (k < ngroups) {
            const PairGroup gr = groups[k];
#pragma unroll
            for (int q = 0; q < MD_BBMAX; ++q) {
                const int p = gr.pidx[q];
                if (q >= bb || p < 0) continue;
                const PairScreen ps = scr[p];
                const int pp0 = ps.pp >> 9, pnpp = ps.pp & 511;
                bool keep = true;
                int skip = 0;
                if (tol > 0.f) {
                    const float fmx = fmaxf(fm[ps.sh & 0xffff], fm[ps.sh >> 16]);
                    const float dx = ps.cx - sb.x, dy = ps.cy - sb.y, dz = ps.cz - sb.z;
                    const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
                    float d = dist - sb.w - ps.delta;
                    d -= 1e-5f * (dist + sb.w + ps.delta) + 1e-6f;   // FP32 safety
                    float f0 = 1.0f;
                    if (d > 0.f) f0 = fminf(1.0f, 0.8862269254527580f / (sqrtf(ps.q) * d));
                    keep = ps.beta * f0 * fmx * 1.001f >= tol;
                    if (keep) {
                        float acc = 0.f;
                        for (int m = pp0 + pnpp - 1; m > pp0 && skip < 15; --m) {
                            const float4 P = pbc[m];
                            const float ex = P.x - sb.x, ey = P.y - sb.y, ez = P.z - sb.z;
                            const float r = sqrtf(ex * ex + ey * ey + ez * ez);
                            float dm = r - sb.w;
                            dm -= 1e-5f * (r + sb.w) + 1e-6f;        // FP32 safety
                            if (dm <= 0.f) break;
                            acc += P.w * fmx / dm;
                            if (acc * 1.001f > 0.5f * tol) break;
                            ++skip;
                        }
                    }
                }
                if (keep) {
                    mask |= 1u << q;
                    skw |= (unsigned)skip << (4 * q);
                    npp += (unsigned long long)(pnpp - skip);
                }
            }
        }
        ntask += (unsigned long long)__popc(mask);
        const bool keep = mask != 0;
        const unsigned long long bal = __ballot(keep);
        const int pos = n + __builtin_amdgcn_mbcnt_hi((unsigned)(bal >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0));
        if (keep) {
            out[pos] = (k << 8) | (int)mask;
            sout[pos] = skw;
        }
        n += __popcll(bal);
    }
    for (int o = 32; o > 0; o >>= 1) {
        npp += __shfl_xor(npp, o, 64);
        ntask += __shfl_xor(ntask, o, 64);
    }
    if (lane == 0) {
        counts[s] = n;
        if (tc.wsub) tc.wsub[s] = (int)(npp < 0x7fffffffull ? npp : 0x7fffffffull);
        atomicAdd(work, ntask);
        atomicAdd(work + 1, npp);
    }
}

// one block per class of tl: a counting sort of the sub-blocks into
// ORD_BINS bins of descending weight (linear in wsub / max wsub); within a
// bin the order is that of the LDS atomics
constexpr int ORD_BINS = 256;
constexpr int ORD_THREADS = 256;      // a block the size of an exchange block
__global__ __launch_bounds__(ORD_THREADS)
void order_kernel(TaskLaunch tl, int nsub)
{
    const TaskClass& tc = tl.c[blockIdx.x];
    const int* __restrict__ w = tc.wsub;
    int* __restrict__ order = tc.order;
    __shared__ int s_max;
    __shared__ int s_pos[ORD_BINS];
    const int t = threadIdx.x;
    if (t == 0) s_max = 0;
    for (int b = t; b < ORD_BINS; b += ORD_THREADS) s_pos[b] = 0;
    __syncthreads();
    int m = 0;
    for (int i = t; i < nsub; i += ORD_THREADS) m = max(m, w[i]);
    atomicMax(&s_max, m);
    __syncthreads();
    const long long mx = s_max;
    auto bin = [&](int x) {
        return mx > 0 ? (ORD_BINS - 1) - (int)((long long)x * (ORD_BINS - 1) / mx) : 0;
    };
    for (int i = t; i < nsub; i += ORD_THREADS) atomicAdd(&s_pos[bin(w[i])], 1);
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int b = 0; b < ORD_BINS; ++b) {
            const int c = s_pos[b];
            s_pos[b] = acc;
            acc += c;
        }
    }
    __syncthreads();
    for (int i = t; i < nsub; i += ORD_THREADS) order[atomicAdd(&s_pos[bin(w[i])], 1)] = i;
}

void launch_order(hipStream_t s, const TaskLaunch& tl, int nsub)
{
    if (tl.n == 0 || nsub == 0) return;
    for (int k = 0; k < tl.n; ++k)
        if (!tl.c[k].wsub || !tl.c[k].order)
            throw Error(WB_INVALID, "launch_order: every class needs wsub and order");
    hipLaunchKernelGGL(order_kernel, dim3(tl.n), dim3(ORD_THREADS), 0, s, tl, nsub);
    WB_HIP(hipGetLastError());
}

void launch_tasks(hipStream_t s, const TaskLaunch& tl, const Grid& g, const float* fmax,
                  int nshell, float tol)
{
    if (tl.n == 0) return;
    const int blocks = (g.nsub + 3) / 4;
    hipLaunchKernelGGL(task_kernel, dim3(blocks, tl.n), dim3(256), 0, s, tl, g.d_sub.as<float4>(),
                       g.nsub, fmax, nshell, tol);
    WB_HIP(hipGetLastError());
}

}  // namespace wb
