// Host build of the generated McMurchie-Davidson device code (md_gen.h), so
// the generator's Hermite recursion, E contraction and cart<->sph transforms
// can be checked on the CPU against the oracle.  Test infrastructure only.
#include <cmath>
#include <cstring>
#define __device__
#define __host__
#define __forceinline__ inline
#define __restrict__
using std::fma;
#include "md_gen.h"

template <int l> static void gsph(const double* c, double* s) {
    if constexpr (l == 0) md_gsph_0<double>(c, s);
    else if constexpr (l == 1) md_gsph_1<double>(c, s);
    else if constexpr (l == 2) md_gsph_2<double>(c, s);
    else md_gsph_3<double>(c, s);
}
template <int l> static void fcart(const double* s, double* c) {
    if constexpr (l == 0) md_fcart_0<double>(s, c);
    else if constexpr (l == 1) md_fcart_1<double>(s, c);
    else if constexpr (l == 2) md_fcart_2<double>(s, c);
    else md_fcart_3<double>(s, c);
}

// onec: the one-centre variant (valid only when A = B)
template <int LA, int LB>
static void prim(bool onec, double dx, double dy, double dz, const double* R0, const double* E,
                 double* V) {
#define C(a, b)                                                                  \
    if constexpr (LA == a && LB == b) {                                          \
        if (onec) md_prim_##a##b##_oc<double, double>(dx, dy, dz, R0, E, V);     \
        else md_prim_##a##b##_tc<double, double>(dx, dy, dz, R0, E, V);          \
    }
    C(0,0) C(1,0) C(1,1) C(2,0) C(2,1) C(3,0) C(3,1) C(3,3)
#undef C
    // split classes: the two parts of the A components, scattered back
#define P(a, b)                                                                        \
    if constexpr (LA == a && LB == b) {                                                \
        constexpr int NCB = (b + 1) * (b + 2) / 2;                                     \
        double V0[MdPart<a, b, 0>::n() * NCB] = {0}, V1[MdPart<a, b, 1>::n() * NCB] = {0}; \
        if (onec) {                                                                    \
            md_prim_##a##b##_oc_p0<double, double>(dx, dy, dz, R0, E, V0);             \
            md_prim_##a##b##_oc_p1<double, double>(dx, dy, dz, R0, E, V1);             \
        } else {                                                                       \
            md_prim_##a##b##_tc_p0<double, double>(dx, dy, dz, R0, E, V0);             \
            md_prim_##a##b##_tc_p1<double, double>(dx, dy, dz, R0, E, V1);             \
        }                                                                              \
        for (int i = 0; i < MdPart<a, b, 0>::n(); ++i)                                 \
            for (int j = 0; j < NCB; ++j) V[MdPart<a, b, 0>::at(i) * NCB + j] += V0[i * NCB + j]; \
        for (int i = 0; i < MdPart<a, b, 1>::n(); ++i)                                 \
            for (int j = 0; j < NCB; ++j) V[MdPart<a, b, 1>::at(i) * NCB + j] += V1[i * NCB + j]; \
    }
    P(2,2) P(3,2)
#undef P
}

// One shell pair (A: la, B: lb, la >= lb), one point; inputs per primitive
// pair m: d = P - C (3), R0 (L+1, pref folded), E record.  Output: the
// spherical block A_sph (nsa x nsb) via sph->cart of unit vectors, i.e. the
// same transforms the kernel applies to F and G.
template <int LA, int LB>
static void block(int onec, int npp, const double* d, const double* R0, const double* E, double* out) {
    constexpr int NCA = (LA + 1) * (LA + 2) / 2, NCB = (LB + 1) * (LB + 2) / 2;
    constexpr int NSA = 2 * LA + 1, NSB = 2 * LB + 1, L = LA + LB;
    constexpr int ERL = 3 * (LA + 1) * (LB + 1) * (L + 1);
    double V[NCA * NCB] = {0};
    for (int m = 0; m < npp; ++m)
        prim<LA, LB>(onec != 0, d[3 * m], d[3 * m + 1], d[3 * m + 2], R0 + m * (L + 1), E + m * ERL, V);
    // A_sph[mu][nu] = e_nu^T (T_B V^T T_A^T) e_mu, built column by column as the
    // kernel does: FaC = fcart(e_mu), gbc = V^T FaC, gbs = gsph(gbc)
    for (int mu = 0; mu < NSA; ++mu) {
        double e[NSA] = {0}, fa[NCA], gb[NCB], gs[NSB];
        e[mu] = 1.0;
        fcart<LA>(e, fa);
        for (int b = 0; b < NCB; ++b) {
            double v = 0;
            for (int a = 0; a < NCA; ++a) v += V[a * NCB + b] * fa[a];
            gb[b] = v;
        }
        gsph<LB>(gb, gs);
        for (int nu = 0; nu < NSB; ++nu) out[mu * NSB + nu] = gs[nu];
    }
}

extern "C" int md_block(int la, int lb, int onec, int npp, const double* d, const double* R0,
                        const double* E, double* out) {
#define C(a, b) if (la == a && lb == b) { block<a, b>(onec, npp, d, R0, E, out); return 0; }
    C(0,0) C(1,0) C(1,1) C(2,0) C(2,1) C(2,2) C(3,0) C(3,1) C(3,2) C(3,3)
#undef C
    return 1;
}
