// C ABI of the wb97mv_amd library and its host-side orchestration.
//
// Host preprocessing (FP64, once per basis / grid) and the per-call device
// pipelines:
//   K build:  X = AO(grid) -> F = X (QD) -> screening -> per class MD kernel
//             (G copies) -> K = X^T (w o sum G)           (gpu/sgx_k.py:663-751)
//   NLC:      X, dX = AO(grid) -> T = X D -> rho, grad rho -> VV10 pointwise,
//             compaction, double sum -> potential -> vmat   (gpu/vv10.py:268-342)
#include <algorithm>
#include <tuple>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <numeric>
#include "wb_boys.h"
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include "wb_internal.h"
#include <memory>
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <cstdlib>
#include <exception>
#include <thread>

namespace wb {

static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }

// ---------------------------------------------------------------- Boys ----
static long double boys_ld(int m, long double T)
{
    // F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))
    if (T == 0.0L) return 1.0L / (2 * m + 1);
    long double term = 1.0L / (2 * m + 1), sum = term;
    for (int k = 1; k < 2000; ++k) {
        term *= 2.0L * T / (2 * m + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-21L) break;
    }
    return expl(-T) * sum;
}

void boys_tables(std::vector<float>& t32, std::vector<double>& t64)
{
    t32.assign(B32_TOTAL, 0.f);
    long double fact[7] = {1, 1, 2, 6, 24, 120, 720};
    for (int L = 0; L <= B32_LMAX; ++L) {
        float* tab = t32.data() + b32_off(L);
        const int rs = b32_row(L);
        for (int k = 0; k < B32_NODES; ++k) {
            const long double T = 0.25L * k;
            float* row = tab + k * rs;
            if (L <= 1) {
                // every order its own series; leading values with remainders
                for (int n = 0; n <= L; ++n) {
                    float* c = row + 7 * n;
                    for (int j = 0; j <= 6; ++j)
                        c[j] = (float)(boys_ld(n + j, T) / fact[j]);
                    row[(L == 0 ? 7 : 14 + n)] = (float)(boys_ld(n, T) - (long double)c[0]);
                }
            } else {
                // F_L(T_k) and e^{-T_k} as doubles, then c_{L,1..6}
                const double lead[2] = {(double)boys_ld(L, T), (double)expl(-T)};
                std::memcpy(row, lead, sizeof(lead));
                for (int j = 1; j <= 6; ++j) row[3 + j] = (float)(boys_ld(L + j, T) / fact[j]);
            }
        }
    }
    t64.assign(B64_NODES * B64_ROW, 0.0);
    for (int k = 0; k < B64_NODES; ++k) {
        const long double T = 0.125L * k;
        for (int m = 0; m <= 14; ++m) t64[k * B64_ROW + m] = (double)boys_ld(m, T);
        t64[k * B64_ROW + 15] = (double)expl(-T);
    }
}

// ------------------------------------------------------ basis / pairs -----
static int cart_idx_list(int l, int comps[][3])
{
    int c = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            comps[c][0] = lx; comps[c][1] = ly; comps[c][2] = l - lx - ly;
            ++c;
        }
    return c;
}

// E[i][j][t] for one axis (McMurchie-Davidson, gpu/shellpairs.py:76-113)
static void hermite_axis(int la, int lb, double p, double xpa, double xpb,
                         double E[4][4][7])
{
    std::memset(E, 0, sizeof(double) * 4 * 4 * 7);
    const double h = 0.5 / p;
    E[0][0][0] = 1.0;
    for (int i = 0; i <= la; ++i)
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            int ii = i, jj = j;
            double x;
            if (j == 0) { ii = i - 1; x = xpa; } else { jj = j - 1; x = xpb; }
            const int tp = ii + jj;
            for (int t = 0; t <= i + j; ++t) {
                double v = 0.0;
                if (t > 0) v += h * E[ii][jj][t - 1];
                if (t <= tp) v += x * E[ii][jj][t];
                if (t + 1 <= tp) v += (t + 1) * E[ii][jj][t + 1];
                E[i][j][t] = v;
            }
        }
}

// sup_{s >= 0} (s + a)^i (s + b)^j exp(-c s^2): log-concave, so the maximum
// is at the root of i/(s+a) + j/(s+b) - 2 c s (bisection)
static double poly_gauss_sup(int i, int j, double a, double b, double c)
{
    if (i == 0 && j == 0) return 1.0;
    auto dlog = [&](double s) { return i / (s + a + 1e-300) + j / (s + b + 1e-300) - 2 * c * s; };
    double lo = 0.0, hi = 1.0;
    while (dlog(hi) > 0) hi *= 2;
    if (dlog(0.0) <= 0) hi = 0.0;
    for (int it = 0; it < 200 && hi > 0; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (dlog(mid) > 0) lo = mid; else hi = mid;
    }
    const double s = hi;
    return std::pow(s + a, i) * std::pow(s + b, j) * std::exp(-c * s * s);
}

static double rowsum_c2s_max(int l)
{
    // max over spherical components m of sum_c |T[m][c]| of the real solid
    // harmonics: s 1/(2 sqrt(pi)), p sqrt(3/(4 pi)), d the z^2 row
    // 4 sqrt(5/(16 pi)), f the z^3 row 8 sqrt(7/(16 pi))
    switch (l) {
    case 0: return 0.5 / std::sqrt(M_PI);
    case 1: return std::sqrt(3.0 / (4 * M_PI));
    case 2: return 4.0 * std::sqrt(5.0 / (16 * M_PI));
    default: return 8.0 * std::sqrt(7.0 / (16 * M_PI));
    }
}

// a bound stored in FP32: never below the FP64 value, and never zero (a
// factor below FP32's normal range, K_AB ~ 1e-50, becomes FLT_MIN)
static float round_up_f32(double v)
{
    float f = (float)v;
    if ((double)f < v) f = std::nextafter(f, INFINITY);
    return std::max(f, FLT_MIN);
}

static void build_basis(Basis& B, int nshell, const int* l, const int* nprim,
                        const double* center, const double* exps, const double* coefs,
                        double pair_cut)
{
    B.nshell = nshell;
    B.pair_cut = pair_cut;
    B.shells.resize(nshell);
    int p0 = 0, ao = 0;
    for (int s = 0; s < nshell; ++s) {
        if (l[s] < 0 || l[s] > LMAX) throw Error(WB_UNSUPPORTED, "shell angular momentum > 3");
        if (nprim[s] < 1 || nprim[s] > 64) throw Error(WB_UNSUPPORTED, "bad primitive count");
        ShellInfo& si = B.shells[s];
        si.cx = center[3 * s]; si.cy = center[3 * s + 1]; si.cz = center[3 * s + 2];
        si.l = l[s]; si.nprim = nprim[s]; si.prim0 = p0; si.ao0 = ao;
        p0 += nprim[s];
        ao += 2 * l[s] + 1;
    }
    B.nao = ao;
    B.exps.assign(exps, exps + p0);
    B.coefs.assign(coefs, coefs + p0);

    for (int la = 0; la <= LMAX; ++la)
        for (int lb = 0; lb <= la; ++lb) {
            ClassData& C = B.cls[class_id(la, lb)];
            C.la = la; C.lb = lb;
            C.erl = 3 * (la + 1) * (lb + 1) * (la + lb + 1);
        }
    // pairs: A = the shell of larger l (larger index on ties), sorted by (A, B)
    std::vector<std::pair<int, int>> order;
    for (int i = 0; i < nshell; ++i)
        for (int j = 0; j <= i; ++j) {
            int a = i, b = j;
            if (l[j] > l[i]) std::swap(a, b);
            order.emplace_back(a, b);
        }
    std::sort(order.begin(), order.end());
    // Per pair (independent, on all host threads): the primitive-pair data
    // (most important first), the pair record and its screening record;
    // then, in pair order, the class arrays (pp0 = running primitive count)
    struct PairOut {
        bool keep = false;
        PairRec pr{};
        PairScreen ps{};
        std::vector<double> P, p, pref, E;
        std::vector<float4> pbc;
        std::vector<int> pij;
    };
    std::vector<PairOut> outs(order.size());
    auto make_pair = [&](size_t idx) {
        const auto [a, b] = order[idx];
        PairOut& o = outs[idx];
        const ShellInfo &A = B.shells[a], &Bs = B.shells[b];
        const int la = A.l, lb = Bs.l;
        int comps_a[10][3], comps_b[10][3];
        const double AB[3] = {A.cx - Bs.cx, A.cy - Bs.cy, A.cz - Bs.cz};
        const double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
        // shell-pair magnitude cutoff (gpu/shellpairs.py:290-294)
        double mag = 0.0;
        for (int i = 0; i < A.nprim; ++i)
            for (int j = 0; j < Bs.nprim; ++j) {
                const double ea = exps[A.prim0 + i], eb = exps[Bs.prim0 + j];
                const double v = std::fabs(coefs[A.prim0 + i] * coefs[Bs.prim0 + j]) *
                                 std::exp(-ea * eb / (ea + eb) * ab2);
                mag = std::max(mag, v);
            }
        if (mag <= pair_cut) return;
        o.keep = true;
        PairRec& pr = o.pr;
        pr.a0 = A.ao0; pr.b0 = Bs.ao0;
        pr.shA = a; pr.shB = b;
        pr.diag = (a == b);
        pr.onec = (ab2 == 0.0);
        cart_idx_list(la, comps_a);
        cart_idx_list(lb, comps_b);
        PairScreen& ps = o.ps;
        double pmin = 1e300, beta = 0.0;
        double Pd[3] = {0, 0, 0};
        std::vector<double> Ps;
        const double rs = rowsum_c2s_max(la) * rowsum_c2s_max(lb);
        // primitive pairs, most important first: (far-field bound, index)
        struct Prim { double key, P[3], p, pref, beta; int ia, ib; double E[3][4][4][7]; };
        std::vector<Prim> prims;
        prims.reserve((size_t)A.nprim * Bs.nprim);
        for (int i = 0; i < A.nprim; ++i)
            for (int j = 0; j < Bs.nprim; ++j) {
                Prim pm{};
                pm.ia = i;
                pm.ib = j;
                const double ea = exps[A.prim0 + i], eb = exps[Bs.prim0 + j];
                const double p = ea + eb, mu = ea * eb / p;
                const double K = std::exp(-mu * ab2);
                const double cc = coefs[A.prim0 + i] * coefs[Bs.prim0 + j];
                const double P[3] = {(ea * A.cx + eb * Bs.cx) / p, (ea * A.cy + eb * Bs.cy) / p,
                                     (ea * A.cz + eb * Bs.cz) / p};
                const double PA[3] = {P[0] - A.cx, P[1] - A.cy, P[2] - A.cz};
                const double PB[3] = {P[0] - Bs.cx, P[1] - Bs.cy, P[2] - Bs.cz};
                for (int d = 0; d < 3; ++d) pm.P[d] = P[d];
                pm.p = p;
                pm.pref = cc * K * 2.0 * M_PI / p;
                for (int d = 0; d < 3; ++d) hermite_axis(la, lb, p, PA[d], PB[d], pm.E[d]);
                // screening: |chi_a chi_b| <= |cc| K M exp(-0.75 p |r-P|^2),
                // M = max over Cartesian components of prod_axes
                //     sup_s (s+|PA|)^i (s+|PB|)^j exp(-0.25 p s^2)
                double M = 0.0;
                for (int u = 0; u < ncart(la); ++u)
                    for (int v = 0; v < ncart(lb); ++v) {
                        double m = 1.0;
                        for (int d = 0; d < 3; ++d)
                            m *= poly_gauss_sup(comps_a[u][d], comps_b[v][d], std::fabs(PA[d]),
                                                std::fabs(PB[d]), 0.25 * p);
                        M = std::max(M, m);
                    }
                pm.beta = std::fabs(cc) * K * M * 2.0 * M_PI / (0.75 * p) * rs;
                // far from the pair the bound falls as beta / sqrt(p) / distance
                pm.key = pm.beta / std::sqrt(p);
                prims.push_back(pm);
            }
        std::stable_sort(prims.begin(), prims.end(),
                         [](const Prim& x, const Prim& y) { return x.key > y.key; });
        const int L = la + lb;
        for (const Prim& pm : prims) {
            o.P.insert(o.P.end(), pm.P, pm.P + 3);
            o.p.push_back(pm.p);
            o.pref.push_back(pm.pref);
            for (int d = 0; d < 3; ++d)
                for (int ii = 0; ii <= la; ++ii)
                    for (int jj = 0; jj <= lb; ++jj)
                        for (int t = 0; t <= L; ++t) o.E.push_back(pm.E[d][ii][jj][t]);
            o.pij.push_back(pm.ia);
            o.pij.push_back(pm.ib);
            o.pbc.push_back(make_float4((float)pm.P[0], (float)pm.P[1], (float)pm.P[2],
                                        round_up_f32(0.8862269254527580 * pm.beta /
                                                     std::sqrt(0.75 * pm.p) * (1 + 1e-5))));
            beta += pm.beta;
            if (pm.p < pmin) { pmin = pm.p; Pd[0] = pm.P[0]; Pd[1] = pm.P[1]; Pd[2] = pm.P[2]; }
            Ps.insert(Ps.end(), pm.P, pm.P + 3);
        }
        double delta = 0.0;
        for (size_t m = 0; m < Ps.size() / 3; ++m) {
            const double dx = Ps[3 * m] - Pd[0], dy = Ps[3 * m + 1] - Pd[1], dz = Ps[3 * m + 2] - Pd[2];
            delta = std::max(delta, std::sqrt(dx * dx + dy * dy + dz * dz));
        }
        ps.cx = (float)Pd[0]; ps.cy = (float)Pd[1]; ps.cz = (float)Pd[2];
        ps.delta = (float)(delta * (1 + 1e-6) + 1e-5);
        ps.beta = round_up_f32(beta * (1 + 1e-5));
        ps.q = (float)(0.75 * pmin * (1 - 1e-6));
        pr.npp = (int)prims.size();
    };
    {
        // WB_HOST_THREADS caps the host threads (the result does not depend on it)
        size_t cap = std::max(1u, std::thread::hardware_concurrency());
        if (const char* e = std::getenv("WB_HOST_THREADS")) cap = std::max(1, std::atoi(e));
        const size_t nthr = std::min<size_t>({16, cap, order.size() / 256 + 1});
        std::atomic<size_t> next{0};
        std::vector<std::exception_ptr> errs(nthr);
        auto worker = [&](size_t t) {
            try {
                for (;;) {
                    const size_t i0 = next.fetch_add(64);
                    if (i0 >= order.size()) break;
                    const size_t i1 = std::min(order.size(), i0 + 64);
                    for (size_t i = i0; i < i1; ++i) make_pair(i);
                }
            } catch (...) {
                errs[t] = std::current_exception();
            }
        };
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nthr; ++t) pool.emplace_back(worker, t);
        worker(0);
        for (auto& th : pool) th.join();
        for (auto& e : errs) if (e) std::rethrow_exception(e);
    }
    long long total = 0;
    for (size_t idx = 0; idx < order.size(); ++idx) {
        PairOut& o = outs[idx];
        if (!o.keep) continue;
        ++total;
        const ShellInfo &A = B.shells[o.pr.shA], &Bs = B.shells[o.pr.shB];
        ClassData& C = B.cls[class_id(A.l, Bs.l)];
        PairRec pr = o.pr;
        pr.pp0 = C.nprims;
        // packed: pp = pp0 << 9 | npp (npp <= 256: up to 16 x 16 primitives,
        // the reference's envelope), pp0 < 2^22 per class
        if (pr.shA >= 32768 || pr.shB >= 32768 || pr.pp0 >= (1 << 22) || pr.npp > 256)
            throw Error(WB_UNSUPPORTED, "shell pair outside the packed screening records (more "
                                        "than 256 primitive pairs in one pair, or 2^22 in its "
                                        "class)");
        PairScreen ps = o.ps;
        ps.sh = pr.shA | pr.shB << 16;
        ps.pp = pr.pp0 << 9 | pr.npp;
        C.P.insert(C.P.end(), o.P.begin(), o.P.end());
        C.p.insert(C.p.end(), o.p.begin(), o.p.end());
        C.pref.insert(C.pref.end(), o.pref.begin(), o.pref.end());
        C.E.insert(C.E.end(), o.E.begin(), o.E.end());
        C.pbc.insert(C.pbc.end(), o.pbc.begin(), o.pbc.end());
        C.pij.insert(C.pij.end(), o.pij.begin(), o.pij.end());
        C.nprims += pr.npp;
        C.pairs.push_back(pr);
        C.scr.push_back(ps);
        C.npairs++;
        o = PairOut();                                   // release as we go
    }
    B.npairs_total = total;
    // shell blocks and pair groups per class
    for (int k = 0; k < NCLASS; ++k) {
        ClassData& C = B.cls[k];
        if (!C.npairs) continue;
        const int bb = md_bb(C.la, C.lb);
        const bool ba = md_block_a(C.la, C.lb);
        const int lk = ba ? C.la : C.lb;
        std::vector<int> rank(nshell, -1);
        int nb = 0;
        for (int i = 0; i < nshell; ++i)
            if (l[i] == lk) rank[i] = nb++;
        const int nblk = (nb + bb - 1) / bb;
        // absent positions of the last block point at the spare rows after
        // the nao AO rows of F and G (MD_SPARE of them; never read back)
        C.bbao.assign((size_t)nblk * bb, B.nao);
        for (int i = 0; i < nshell; ++i)
            if (rank[i] >= 0) C.bbao[rank[i]] = B.shells[i].ao0;
        auto kshell = [&](const PairRec& pr) { return ba ? pr.shA : pr.shB; };
        auto ishell = [&](const PairRec& pr) { return ba ? pr.shB : pr.shA; };
        std::vector<int> ord(C.npairs);
        for (int i = 0; i < C.npairs; ++i) ord[i] = i;
        auto key = [&](int i) {
            const PairRec& pr = C.pairs[i];
            return std::make_tuple(rank[kshell(pr)] / bb, ishell(pr), kshell(pr));
        };
        std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return key(x) < key(y); });
        std::vector<PairRec> pairs(C.npairs);
        std::vector<PairScreen> scr(C.npairs);
        C.groups.clear();
        for (int i = 0; i < C.npairs; ++i) {
            pairs[i] = C.pairs[ord[i]];
            scr[i] = C.scr[ord[i]];
            const int ks = kshell(pairs[i]), is = ishell(pairs[i]);
            const int blk = rank[ks] / bb, pos = rank[ks] % bb;
            if (C.groups.empty() || C.groups.back().bblk != blk || C.groups.back().shX != is) {
                PairGroup gr{};
                gr.shX = is;
                gr.x0 = B.shells[is].ao0;
                gr.bblk = blk;
                for (int j = 0; j < MD_BBMAX; ++j) gr.pidx[j] = gr.pw[j] = -1;
                C.groups.push_back(gr);
            }
            const PairRec& pr = pairs[i];
            if (pr.pp0 >= (1 << 22) || pr.npp < 1 || pr.npp > 256)
                throw Error(WB_UNSUPPORTED, "shell-pair class too large for the packed task words");
            C.groups.back().pidx[pos] = i;
            C.groups.back().pw[pos] = (int)((unsigned)pr.pp0 << 10 | (unsigned)(pr.npp - 1) << 2 |
                                            (pr.onec ? 2u : 0u) | (pr.diag ? 1u : 0u));
            C.groups.back().npair++;
        }
        C.pairs.swap(pairs);
        C.scr.swap(scr);
        C.ngroups = (int)C.groups.size();
        if (C.ngroups >= (1 << 23)) throw Error(WB_UNSUPPORTED, "too many shell-pair groups");
    }
}

// synchronous upload (setup paths; the host buffer may die on return)
static void upload(DevBuf& d, const void* h, size_t bytes, hipStream_t s)
{
    d.ensure(bytes);
    if (bytes) {
        WB_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, s));
        WB_HIP(hipStreamSynchronize(s));
    }
}

static void upload_basis(Basis& B, hipStream_t s)
{
    upload(B.d_shells, B.shells.data(), sizeof(ShellInfo) * B.nshell, s);
    upload(B.d_exps, B.exps.data(), sizeof(double) * B.exps.size(), s);
    upload(B.d_coefs, B.coefs.data(), sizeof(double) * B.coefs.size(), s);
    std::vector<int> shao(2 * B.nshell);
    for (int i = 0; i < B.nshell; ++i) {
        shao[i] = B.shells[i].ao0;
        shao[B.nshell + i] = 2 * B.shells[i].l + 1;
    }
    upload(B.d_shao, shao.data(), sizeof(int) * shao.size(), s);
    std::vector<int> row0(B.nao);
    for (int i = 0; i < B.nshell; ++i)
        for (int m = 0; m < 2 * B.shells[i].l + 1; ++m) row0[B.shells[i].ao0 + m] = B.shells[i].ao0;
    upload(B.d_row0, row0.data(), sizeof(int) * row0.size(), s);

    for (auto& C : B.cls) {
        if (!C.npairs) continue;
        upload(C.d_pairs, C.pairs.data(), sizeof(PairRec) * C.npairs, s);
        upload(C.d_scr, C.scr.data(), sizeof(PairScreen) * C.npairs, s);
        upload(C.d_groups, C.groups.data(), sizeof(PairGroup) * C.ngroups, s);
        upload(C.d_bbao, C.bbao.data(), sizeof(int) * C.bbao.size(), s);
        upload(C.d_pbc, C.pbc.data(), sizeof(float4) * C.pbc.size(), s);
    }
    WB_HIP(hipStreamSynchronize(s));
}

// Primitive-pair records of an operator set (md_rec_bytes layout).  Operator
// constants [tp, c_0..c_L]: theta = w^2/(w^2+p) (1 for 1/r), tp = theta p,
// c_n = pref sqrt(theta)(-2 theta p)^n, formed in FP64; FP32 records hold
// (hi, lo) pairs so the kernel forms each constant product with one
// rounding (fma(hi, x, lo x)), and the E coefficients followed by their
// FP32 rounding remainders.
static const void* op_consts(ClassData& C, const std::vector<double>& omegas, bool f32,
                             hipStream_t s)
{
    std::lock_guard<std::mutex> lock(C.opc_mu);
    auto it = C.opc.find(omegas);
    if (it == C.opc.end()) {
        const int L = C.la + C.lb, nc = L + 2, erl = C.erl, nop = (int)omegas.size();
        const size_t rb32 = md_rec_bytes(C.la, C.lb, true, nop);
        const size_t rb64 = md_rec_bytes(C.la, C.lb, false, nop);
        std::vector<char> r32((size_t)C.nprims * rb32, 0), r64((size_t)C.nprims * rb64, 0);
        std::vector<double> c(nc);
        for (int m = 0; m < C.nprims; ++m) {
            char* q32 = r32.data() + m * rb32;
            char* q64 = r64.data() + m * rb64;
            std::memcpy(q32, &C.P[3 * (size_t)m], 24);
            std::memcpy(q64, &C.P[3 * (size_t)m], 24);
            float* f = (float*)(q32 + 24);
            double* d = (double*)(q64 + 24);
            const double p = C.p[m];
            for (int o = 0; o < nop; ++o) {
                const double omega = omegas[o];
                const double th = omega > 0 ? omega * omega / (omega * omega + p) : 1.0;
                c[0] = th * p;
                double cn = C.pref[m] * std::sqrt(th);
                for (int n = 0; n <= L; ++n) {
                    c[1 + n] = cn;
                    cn *= -2.0 * th * p;
                }
                for (int i = 0; i < nc; ++i) {
                    const float hi = (float)c[i];
                    f[2 * (o * nc + i)] = hi;
                    f[2 * (o * nc + i) + 1] = (float)(c[i] - (double)hi);
                    d[o * nc + i] = c[i];
                }
            }
            float* e32 = f + 2 * nc * nop;
            double* e64 = d + nc * nop;
            for (int i = 0; i < erl; ++i) {
                const double v = C.E[(size_t)m * erl + i];
                const float hi = (float)v;
                e32[i] = hi;
                e32[erl + i] = (float)(v - (double)hi);
                e64[i] = v;
            }
        }
        auto b32 = std::make_shared<DevBuf>(), b64 = std::make_shared<DevBuf>();
        upload(*b32, r32.data(), r32.size(), s);
        upload(*b64, r64.data(), r64.size(), s);
        it = C.opc.emplace(omegas, std::make_pair(b32, b64)).first;
    }
    return f32 ? it->second.first->p : it->second.second->p;
}

// ---------------------------------------------------------------- grid ----
// k-d split into leaves of exactly 64 points (the last may be short): each
// split cuts the longest extent at a multiple of 64 points.
static void kd_order(std::vector<int>& idx, int b, int e, const double* xyz)
{
    const int n = e - b;
    if (n <= SUB) return;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = b; i < e; ++i)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], xyz[3 * idx[i] + d]);
            hi[d] = std::max(hi[d], xyz[3 * idx[i] + d]);
        }
    int ax = 0;
    for (int d = 1; d < 3; ++d) if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
    const int nblk = (n + SUB - 1) / SUB;
    const int mid = b + (nblk / 2) * SUB;
    std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e,
                     [&](int x, int y) { return xyz[3 * x + ax] < xyz[3 * y + ax]; });
    kd_order(idx, b, mid, xyz);
    kd_order(idx, mid, e, xyz);
}

static void build_grid(Grid& G, int npts, const double* xyz, const double* w, hipStream_t s)
{
    G.npts = npts;
    G.nsub = std::max(1, (npts + SUB - 1) / SUB);
    G.Ppad = G.nsub * SUB;
    G.perm.resize(npts);
    std::iota(G.perm.begin(), G.perm.end(), 0);
    kd_order(G.perm, 0, npts, xyz);
    std::vector<double> hx(3 * (size_t)G.Ppad), hw(G.Ppad, 0.0);
    std::vector<float4> sub(G.nsub);
    double o[3] = {0, 0, 0};
    for (int i = 0; i < npts; ++i)
        for (int d = 0; d < 3; ++d) o[d] += xyz[3 * i + d];
    for (int d = 0; d < 3; ++d) G.origin[d] = npts ? o[d] / npts : 0.0;
    for (int sb = 0; sb < G.nsub; ++sb) {
        const int b = sb * SUB, e = std::min(npts, b + SUB);
        double c[3] = {0, 0, 0};
        for (int i = b; i < e; ++i)
            for (int d = 0; d < 3; ++d) c[d] += xyz[3 * G.perm[i] + d];
        const int cnt = std::max(1, e - b);
        for (int d = 0; d < 3; ++d) c[d] /= cnt;
        double r = 0.0;
        for (int i = b; i < b + SUB; ++i) {
            const int src = (i < npts) ? G.perm[i] : (npts ? G.perm[b < npts ? b : 0] : -1);
            for (int d = 0; d < 3; ++d) hx[(size_t)d * G.Ppad + i] = src >= 0 ? xyz[3 * src + d] : 0.0;
            hw[i] = (i < npts) ? w[G.perm[i]] : 0.0;
            if (i < e) {
                const double dx = xyz[3 * src] - c[0], dy = xyz[3 * src + 1] - c[1],
                             dz = xyz[3 * src + 2] - c[2];
                r = std::max(r, std::sqrt(dx * dx + dy * dy + dz * dz));
            }
        }
        sub[sb] = make_float4((float)c[0], (float)c[1], (float)c[2],
                              (float)(r * (1 + 1e-6) + 1e-5));
    }
    G.w_sorted = hw;
    upload(G.d_xyz, hx.data(), sizeof(double) * hx.size(), s);
    upload(G.d_w, hw.data(), sizeof(double) * hw.size(), s);
    upload(G.d_sub, sub.data(), sizeof(float4) * sub.size(), s);
    WB_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------ profiling ---
void Ctx::start()
{
    if (!profile) return;
    if (!t0) WB_HIP(hipEventCreate(&t0));
    WB_HIP(hipEventRecord(t0, stream));
    marks.clear();
}

void Ctx::mark(int phase)
{
    if (!profile) return;
    hipEvent_t e;
    WB_HIP(hipEventCreate(&e));
    WB_HIP(hipEventRecord(e, stream));
    marks.emplace_back(phase, e);
}

// accumulate the recorded phases into phase_ms (synchronizes)
static void collect(Ctx& c)
{
    if (!c.profile || c.marks.empty()) return;
    WB_HIP(hipEventSynchronize(c.marks.back().second));
    hipEvent_t prev = c.t0;
    for (auto& [ph, e] : c.marks) {
        float ms = 0.f;
        WB_HIP(hipEventElapsedTime(&ms, prev, e));
        c.phase_ms[ph] += ms;
        if (prev != c.t0) (void)hipEventDestroy(prev);
        prev = e;
    }
    if (prev != c.t0) (void)hipEventDestroy(prev);
    c.marks.clear();
}

// ------------------------------------------------------------- K build ----
// The workspace of one K build, carved from one arena (offsets 256-byte
// aligned).  The grid is processed in chunks of csub sub-blocks, a whole
// number of the K GEMM's point splits (kchunk points each, fixed by the
// whole grid), so the split partials and the order in which they are added
// are those of the one-chunk build: a tiled build equals the untiled one bit
// for bit.  Everything else is per point or per sub-block.  csub = nsub: one
// chunk, the chunk is the grid itself (no copies of its coordinates).
struct KPlan {
    int kchunk = 0, csub = 0, nchunk = 0;
    size_t nitem_sub = 0;           // item-list entries per sub-block, all classes
    size_t qdt = 0, x = 0, f = 0, fmax = 0, counts = 0, wsub = 0, order = 0, ids = 0, skips = 0,
           g = 0, mask = 0, cp = 0, fold = 0, xyz = 0, w = 0, sub = 0, lst = 0, total = 0;
};

static KPlan k_plan(const Basis& B, const Grid& G, int nop, bool f32, bool g64, int ncopy,
                    int csub)
{
    KPlan P;
    nt_split(G.Ppad, P.kchunk);
    P.csub = std::max(1, std::min(csub, G.nsub));
    P.nchunk = (G.nsub + P.csub - 1) / P.csub;
    for (const auto& C : B.cls) P.nitem_sub += (size_t)C.ngroups;
    const size_t nao = B.nao, Pc = (size_t)P.csub * SUB, zrow = nao + MD_SPARE;
    const size_t nsplit = (Pc + P.kchunk - 1) / P.kchunk;
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    P.qdt = take(8 * nao * nao);
    P.x = take(8 * nao * Pc);
    P.f = take((f32 ? 4 : 8) * (nao + MD_SPARE) * Pc);
    P.fmax = take(4 * (size_t)P.csub * B.nshell);
    P.counts = take(4 * (size_t)NCLASS * P.csub);
    P.wsub = take(4 * (size_t)NCLASS * P.csub);
    P.order = take(4 * (size_t)NCLASS * P.csub);
    P.ids = take(4 * std::max<size_t>(P.nitem_sub * P.csub, 1));
    P.skips = take(4 * std::max<size_t>(P.nitem_sub * P.csub, 1));
    P.g = take((g64 ? 8 : 4) * (zrow + 1) * Pc * ncopy * nop);
    P.mask = take(zrow <= 32 * MD_MASK_WORDS ? 4 * (size_t)ncopy * P.csub * MD_MASK_WORDS * nop : 0);
    P.cp = take(8 * nsplit * nao * nao);
    P.fold = take(8 * nao * Pc);
    P.lst = take(ao_lists_layout(B, P.csub, P.kchunk).total);
    if (P.nchunk > 1) {
        P.xyz = take(8 * 3 * Pc);
        P.w = take(8 * Pc);
        P.sub = take(sizeof(float4) * P.csub);
    }
    P.total = off;
    return P;
}

// the plan with the most sub-blocks per chunk whose workspace fits budget
// bytes (0: the whole grid in one chunk)
static KPlan k_plan_budget(const Basis& B, const Grid& G, int nop, bool f32, bool g64, int ncopy,
                           size_t budget)
{
    KPlan one = k_plan(B, G, nop, f32, g64, ncopy, G.nsub);
    if (budget == 0 || one.total <= budget) return one;
    const int step = one.kchunk / SUB;                  // sub-blocks per split
    const int nsteps = (G.nsub + step - 1) / step;
    int lo = 1, hi = nsteps;                            // chunk = lo..hi whole splits
    KPlan best = k_plan(B, G, nop, f32, g64, ncopy, step);
    if (best.total > budget)
        throw Error(WB_OOM, "K build needs at least " + std::to_string(best.total) +
                                " bytes of workspace (budget " + std::to_string(budget) + ")");
    while (lo < hi) {                                   // total grows with the chunk
        const int mid = (lo + hi + 1) / 2;
        KPlan p = k_plan(B, G, nop, f32, g64, ncopy, mid * step);
        if (p.total <= budget) { lo = mid; best = p; }
        else hi = mid - 1;
    }
    return best;
}

// the records of an operator set if they are on the device (no allocation
// of any kind: the allocation-free K build uses this)
static const void* op_records_find(ClassData& C, const double* omegas, int nop, bool f32)
{
    std::lock_guard<std::mutex> lock(C.opc_mu);
    for (const auto& kv : C.opc)
        if ((int)kv.first.size() == nop && std::equal(omegas, omegas + nop, kv.first.begin()))
            return f32 ? kv.second.first->p : kv.second.second->p;
    throw Error(WB_INVALID, "operator records for these omegas are not on the device yet: run "
                            "wb_k_build or wb_k_build_dev once with them on this basis");
}

// primitive-pair records of an operator set: built (and cached on the
// basis) when allowed, else only looked up
static const void* op_records(ClassData& C, const std::vector<double>& omegas, bool f32,
                              hipStream_t s, bool may_build)
{
    return may_build ? op_consts(C, omegas, f32, s) : op_records_find(C, omegas.data(),
                                                                       (int)omegas.size(), f32);
}

// operators: 1/r (omega = 0) and erf(omega r)/r (omega > 0); the
// short-range erfc operator of PySCF's omega < 0 is outside this path
static void k_check_omegas(int nop, const double* omegas)
{
    for (int op = 0; op < nop; ++op)
        if (!(omegas[op] >= 0.0 && omegas[op] < 1e30))
            throw Error(WB_UNSUPPORTED, "omega must be 0 (1/r) or positive (erf); the "
                                        "short-range erfc operator (omega < 0) is not supported");
}

// primitive-pair records of every class: both operators in one record set
// where the class runs them fused, else one set per operator (building them
// may allocate and synchronize; noalloc: looked up only)
struct KRecs {
    const void* rec1[NCLASS][8] = {};
    const void* rec2[NCLASS] = {};
};
static void k_records(Ctx& c, Basis& B, int nop, const double* omegas, bool f32, bool noalloc,
                      KRecs& R)
{
    hipStream_t s = c.stream;
    for (int k = 0; k < NCLASS; ++k) {
        ClassData& C = B.cls[k];
        if (!C.npairs) continue;
        if (nop == 2 && md_fuse(C.la, C.lb) && (c.fuse_mask >> k & 1u))
            R.rec2[k] = noalloc ? op_records_find(C, omegas, 2, f32)
                                : op_records(C, {omegas[0], omegas[1]}, f32, s, true);
        else
            for (int op = 0; op < nop; ++op)
                R.rec1[k][op] = noalloc ? op_records_find(C, omegas + op, 1, f32)
                                        : op_records(C, {omegas[op]}, f32, s, true);
    }
}

// One chunk of one molecule's K build between its phases: its points (the
// grid itself, or views into the arena), AO values X, F, task lists, G
// copies and written-shell masks
struct KChunk {
    Grid Gc;
    const Grid* gv = nullptr;
    int ns = 0, Ppad = 0, zrow = 0;
    bool tracked = false, first = true;
    double* X = nullptr;
    void* F = nullptr;
    int* counts = nullptr;
    int* order = nullptr;             // [class][sub-block], heavy first (null: identity)
    int* ids = nullptr;
    unsigned* skips = nullptr;
    size_t cls_off[NCLASS] = {};
    char* Gb = nullptr;
    unsigned* masks = nullptr;
    size_t gstride = 0, opstride = 0, gsz = 0, mopstride = 0;
    // the AO lists of the chunk (gathered F and K GEMMs), when gathering
    bool gather = false;
    AoLists AL;
    char* lws = nullptr;
};

// chunk ci of the build planned by P on arena: its points, X, F = QD^T X,
// the screening bounds and the task lists (work: the per-class counters,
// 2 NCLASS x u64; QDT: QD transposed)
template <typename Mark>
static void k_chunk_setup(Ctx& c, Basis& B, const Grid& G, const KPlan& P, char* arena,
                          unsigned long long* work, const double* QD, const double* QDT, int ci,
                          bool f32, bool g64, double tol, KChunk& K, Mark mark)
{
    hipStream_t s = c.stream;
    const int nao = B.nao, ncopy = c.ncopy;
    const int s0 = ci * P.csub, ns = std::min(P.csub, G.nsub - s0);
    const int Ppad = ns * SUB;
    K.ns = ns;
    K.Ppad = Ppad;
    K.zrow = nao + MD_SPARE;
    K.tracked = K.zrow <= 32 * MD_MASK_WORDS;
    K.first = true;
    // the chunk's points as a grid of their own (views into the arena)
    K.gv = &G;
    if (P.nchunk > 1) {
        Grid& Gc = K.Gc;
        Gc.nsub = ns;
        Gc.Ppad = Ppad;
        Gc.npts = std::min(Ppad, G.npts - s0 * SUB);
        Gc.d_xyz.borrow(arena + P.xyz, 8 * 3 * (size_t)Ppad);
        Gc.d_w.borrow(arena + P.w, 8 * (size_t)Ppad);
        Gc.d_sub.borrow(arena + P.sub, sizeof(float4) * ns);
        for (int d = 0; d < 3; ++d)
            WB_HIP(hipMemcpyAsync(Gc.d_xyz.as<double>() + (size_t)d * Ppad,
                                  G.d_xyz.as<double>() + (size_t)d * G.Ppad + (size_t)s0 * SUB,
                                  8 * (size_t)Ppad, hipMemcpyDeviceToDevice, s));
        WB_HIP(hipMemcpyAsync(Gc.d_w.p, G.d_w.as<double>() + (size_t)s0 * SUB,
                              8 * (size_t)Ppad, hipMemcpyDeviceToDevice, s));
        WB_HIP(hipMemcpyAsync(Gc.d_sub.p, G.d_sub.as<float4>() + s0, sizeof(float4) * ns,
                              hipMemcpyDeviceToDevice, s));
        K.gv = &Gc;
    }
    K.X = (double*)(arena + P.x);
    // F and G carry MD_SPARE spare rows after the AO rows for the padding
    // positions of the exchange kernels' B blocks
    K.F = arena + P.f;
    // F = QD^T X over each 128-point tile's significant AOs only (the rows
    // of QD and X gathered by the tile's AO list); the dense GEMM when the
    // shapes do not allow it
    const bool gather = gemm_nn_gather_ok(nao, Ppad, QD, K.X);
    const AoLists AL = ao_lists_layout(B, ns, P.kchunk);
    K.gather = gather;
    K.AL = AL;
    K.lws = arena + P.lst;
    if (gather) launch_ao_eval_lists(s, B, *K.gv, 0, K.X, K.lws, AL);
    else launch_ao_eval(s, B, *K.gv, 0, K.X);
    mark(PH_AO);
    if (gather) gemm_nn_gather(s, nao, Ppad, QD, K.X, K.F, f32, K.lws, AL);
    else gemm_nn_fast(s, nao, Ppad, nao, QDT, K.X, K.F, f32);
    mark(PH_FGEMM);
    if (c.rec_gate && ci == 0) WB_HIP(hipEventRecord(c.ev_gate, s));
    // screening and task lists
    float* fmax = (float*)(arena + P.fmax);
    launch_fmax(s, K.F, f32, K.gv->d_w.as<double>(), nao, Ppad, ns, B.d_shao.as<int>(),
                B.d_shao.as<int>() + B.nshell, B.nshell, fmax);
    K.counts = (int*)(arena + P.counts);
    zero_bytes(s, K.counts, sizeof(int) * (size_t)NCLASS * ns);
    K.ids = (int*)(arena + P.ids);
    K.skips = (unsigned*)(arena + P.skips);
    // the exchange launches take their sub-blocks heaviest first
    K.order = (int*)(arena + P.order);
    size_t toff = 0;
    TaskLaunch tl;
    for (int k = 0; k < NCLASS; ++k) {
        K.cls_off[k] = toff;
        const ClassData& C = B.cls[k];
        if (C.npairs) {
            TaskClass& t = tl.c[tl.n++];
            t.scr = C.d_scr.as<PairScreen>();
            t.groups = C.d_groups.as<PairGroup>();
            t.pbc = C.d_pbc.as<float4>();
            t.counts = K.counts + (size_t)k * ns;
            t.ids = K.ids + toff;
            t.skips = K.skips + toff;
            t.work = work + 2 * k;
            t.wsub = (int*)(arena + P.wsub) + (size_t)k * ns;
            t.order = K.order ? K.order + (size_t)k * ns : nullptr;
            t.ngroups = C.ngroups;
            t.bb = md_bb(C.la, C.lb);
        }
        toff += (size_t)C.ngroups * ns;
    }
    launch_tasks(s, tl, *K.gv, fmax, B.nshell, (float)tol);
    if (K.order) launch_order(s, tl, ns);
    // rows: nao, MD_SPARE spare rows (padding positions of shell blocks),
    // one zero row (read for a shell's first update in a region)
    K.gstride = (size_t)(K.zrow + 1) * Ppad;
    K.opstride = K.gstride * ncopy;                   // elements between operators' G
    K.gsz = g64 ? 8 : 4;
    K.Gb = arena + P.g;
    // G is never cleared: written-shell masks per (copy, sub-block) region
    // send each shell's first update to the zero row (rows < 2048; larger
    // bases clear G in the first class launch instead)
    K.mopstride = (size_t)ncopy * ns * MD_MASK_WORDS;
    K.masks = K.tracked ? (unsigned*)(arena + P.mask) : nullptr;
    mark(PH_SCREEN);
}

// the exchange-kernel shares of class k for the chunk, in the order the
// single-molecule build launches them: one share for both operators (the
// return value: the operator count of each share, 2) or one per operator
// (1); none without pairs.  Marks the chunk's first class.
static int k_class_jobs(const Ctx& c, const Basis& B, KChunk& K, const KRecs& R, int k, int nop,
                        MdJob* out, int& njob)
{
    njob = 0;
    const ClassData& C = B.cls[k];
    if (!C.npairs) return 1;
    const int ns = K.ns, ncopy = c.ncopy;
    const int* cnt = K.counts + (size_t)k * ns;
    const int* ids_k = K.ids + K.cls_off[k];
    const unsigned* skips_k = K.skips + K.cls_off[k];
    const int* ord = K.order ? K.order + (size_t)k * ns : nullptr;
    const int zrows = K.first ? K.zrow : 0;
    // waves per sub-block of this class (the first class launched
    // initializes every copy's region, so it uses them all)
    const int nch = (K.first || c.cls_waves[k] <= 0) ? ncopy : std::min(c.cls_waves[k], ncopy);
    int nopv = 1;
    if (R.rec2[k]) {
        // both operators in one pass over the tasks
        MdMasks mk{K.masks, K.mopstride, K.zrow, K.first};
        out[njob++] = md_job(C, *K.gv, K.F, K.Gb, K.gstride, K.opstride, nch, cnt, ids_k, skips_k,
                             R.rec2[k], zrows, mk);
        out[njob - 1].order = ord;
        nopv = 2;
    } else {
        for (int op = 0; op < nop; ++op) {
            MdMasks mk{K.tracked ? K.masks + K.mopstride * op : nullptr, K.mopstride, K.zrow,
                       K.first};
            out[njob++] = md_job(C, *K.gv, K.F, K.Gb + K.gsz * K.opstride * op, K.gstride,
                                 K.opstride, nch, cnt, ids_k, skips_k, R.rec1[k][op], zrows, mk);
            out[njob - 1].order = ord;
        }
    }
    K.first = false;
    return nopv;
}

// after the class launches: G of a chunk without any shell pair cleared,
// then K (+)= X^T (w o G) per operator
template <typename Mark>
static void k_chunk_finish(Ctx& c, Basis& B, const KPlan& P, char* arena, const KChunk& K,
                           int nop, bool g64, double* d_K, bool accumulate, Mark mark)
{
    hipStream_t s = c.stream;
    const int nao = B.nao;
    const size_t nn = (size_t)nao * nao;
    if (K.first) {                                    // no shell pair at all
        if (K.tracked) zero_bytes(s, K.masks, sizeof(unsigned) * K.mopstride * nop);
        else zero_bytes(s, K.Gb, K.gsz * K.opstride * nop);
    }
    mark(PH_MD);
    // K = X^T (w o G) over each point split's significant AO rows of X
    const NtGather ng = K.gather ? nt_gather(K.AL, K.lws) : NtGather{};
    for (int op = 0; op < nop; ++op) {
        gemm_nt_core(s, nao, nao, K.Ppad, P.kchunk, K.X, K.Gb + K.gsz * K.opstride * op, !g64,
                     c.ncopy, K.gstride, K.gv->d_w.as<double>(), d_K + (size_t)op * nn,
                     accumulate, (double*)(arena + P.cp), (double*)(arena + P.fold),
                     K.tracked ? K.masks + K.mopstride * op : nullptr, B.d_row0.as<int>(), K.ns,
                     K.gather ? &ng : nullptr);
        mark(PH_KGEMM);
    }
}

// The K build on the arena (P.total bytes at arena; work: the per-class
// counters, 2 NCLASS x u64).  noalloc: no allocating call at all (records
// looked up, no profiling events).
static void k_build_arena(Ctx& c, Basis& B, Grid& G, const double* d_QD, int nop,
                          const double* omegas, double tol, int prec, double* d_K,
                          const KPlan& P, char* arena, unsigned long long* work, bool noalloc)
{
    k_check_omegas(nop, omegas);
    const bool f32 = (prec == WB_PREC_MIXED);
    const bool g64 = !f32 || c.g64;
    hipStream_t s = c.stream;
    const bool prof = c.profile && !noalloc;
    GemmTilesScope tiles(c.gemm_tiles);
    if (prof) c.start();
    auto mark = [&](int ph) { if (prof) c.mark(ph); };
    KRecs R;
    k_records(c, B, nop, omegas, f32, noalloc, R);     // first: may allocate and synchronize
    double* QDT = (double*)(arena + P.qdt);
    transpose(s, B.nao, d_QD, QDT);
    zero_bytes(s, work, sizeof(unsigned long long) * 2 * NCLASS);
    const void* boys = f32 ? c.d_b32.p : c.d_b64.p;
    size_t ntask = 0;
    for (auto& C : B.cls) ntask += (size_t)C.npairs * G.nsub;
    for (int ci = 0; ci < P.nchunk; ++ci) {
        KChunk K;
        k_chunk_setup(c, B, G, P, arena, work, d_QD, QDT, ci, f32, g64, tol, K, mark);
        for (int k = 0; k < NCLASS; ++k) {
            MdJob js[8];
            int nj = 0;
            const int nopv = k_class_jobs(c, B, K, R, k, nop, js, nj);
            for (int i = 0; i < nj; ++i)
                launch_md(s, B.cls[k].la, B.cls[k].lb, f32, g64, nopv, js[i], boys);
        }
        k_chunk_finish(c, B, P, arena, K, nop, g64, d_K, ci > 0, mark);
    }
    if (!noalloc) c.tasks_total = (long long)ntask;
    if (prof) collect(c);
}

// ------------------------------------------- allocation-free K build ----
// Device scratch of wb_k_build_noalloc, a fixed maximum of 256 MiB.  Every
// GPU the library runs on holds its own copy of this array until the
// process ends.  In host code the array's name is only a host-side handle:
// its device address, per device, comes from hipGetSymbolAddress with that
// device current (resolve_scratch, called once per device when a context is
// created there), and only that address reaches kernels and copies.  All
// calls on one GPU share its array: a call clears what must start at zero
// (its counters, task counts and written-shell masks) itself, and calls that
// overlap on one GPU race.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)256 << 20;
constexpr size_t NOALLOC_WORK_BYTES = 256;          // the per-class counters, first
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[MAX_DEVICES] = {};

static void resolve_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    WB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(WB_DEVICE_ERROR, "no device address for the scratch array");
    g_scratch_ptr[device] = (char*)p;
}

static char* scratch_of(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= MAX_DEVICES || !g_scratch_ptr[device])
        throw Error(WB_INVALID, "no scratch on this device: create the context with wb_ctx_create");
    return g_scratch_ptr[device];
}

static void k_build_noalloc(Ctx& c, Basis& B, Grid& G, const double* d_QD, int nop,
                            const double* omegas, double tol, int prec, double* d_K)
{
    const bool f32 = prec == WB_PREC_MIXED;
    const size_t cap = NOALLOC_SCRATCH_BYTES - NOALLOC_WORK_BYTES;
    KPlan P;
    try {
        P = k_plan_budget(B, G, nop, f32, !f32 || c.g64, c.ncopy, cap);
    } catch (const Error& e) {
        if (e.code != WB_OOM) throw;
        throw Error(WB_UNSUPPORTED, std::string("allocation-free K build: ") + e.what() +
                                        " exceeds the fixed 256 MiB scratch (fewer G copies, "
                                        "or wb_k_build_dev, which allocates)");
    }
    char* base = scratch_of(c.device);
    k_build_arena(c, B, G, d_QD, nop, omegas, tol, prec, d_K, P, base + NOALLOC_WORK_BYTES,
                  (unsigned long long*)base, true);
}

// the K build with the context's workspaces (the arena sized by the
// context's budget)
static void k_build_dev(Ctx& c, Basis& B, Grid& G, const double* d_QD, int nop,
                        const double* omegas, double tol, int prec, double* d_K)
{
    const bool f32 = prec == WB_PREC_MIXED;
    const KPlan P = k_plan_budget(B, G, nop, f32, !f32 || c.g64, c.ncopy, c.k_budget);
    char* arena = (char*)c.ws_karena.ensure(P.total);
    auto* work = (unsigned long long*)c.ws_work.ensure(sizeof(unsigned long long) * 2 * NCLASS);
    k_build_arena(c, B, G, d_QD, nop, omegas, tol, prec, d_K, P, arena, work, false);
}

// kept tasks (sub-block x shell pair) of the last screening, all classes
static long long count_kept(Ctx& c)
{
    unsigned long long h[2 * NCLASS] = {0};
    if (!c.ws_work.p) return 0;
    WB_HIP(hipMemcpyAsync(h, c.ws_work.p, sizeof(h), hipMemcpyDeviceToHost, c.stream));
    WB_HIP(hipStreamSynchronize(c.stream));
    long long s = 0;
    for (int k = 0; k < NCLASS; ++k) s += (long long)h[2 * k];
    return s;
}

// ------------------------------------------------------------------ NLC ---
static void nlc_dev(Ctx& c, Basis& B, Grid& G, const double* d_dm, double b, double Cc,
                    double tol, int prec, double* d_out2, double* d_vmat)
{
    const bool f64 = (prec == WB_PREC_FP64);
    const int nao = B.nao, Ppad = G.Ppad;
    hipStream_t s = c.stream;
    c.start();
    const size_t plane = (size_t)nao * Ppad;
    // AO values only: the density and the potential operand evaluate the
    // gradients on the fly instead of storing (and twice reading) 3 planes
    double* X = (double*)c.ws_Xv.ensure(sizeof(double) * plane);
    // T = D X (D symmetric) over each 128-point tile's significant AOs
    const bool gather = gemm_nn_gather_ok(nao, Ppad, d_dm, X);
    int kc = 0;
    nt_split(Ppad, kc);
    const AoLists AL = ao_lists_layout(B, G.nsub, kc);
    void* lws = gather ? c.ws_list.ensure(AL.total) : nullptr;
    if (gather) launch_ao_eval_lists(s, B, G, 0, X, lws, AL);
    else launch_ao_eval(s, B, G, 0, X);
    c.mark(PH_NLC_AO);
    double* T = (double*)c.ws_T.ensure(sizeof(double) * plane);
    if (gather) gemm_nn_gather(s, nao, Ppad, d_dm, X, T, false, lws, AL);
    else gemm_nn_fast(s, nao, Ppad, nao, d_dm, X, T, false);
    double* rho = (double*)c.ws_rho.ensure(sizeof(double) * 4 * Ppad);
    double* part = (double*)c.ws_red.ensure(ao_rho_bytes(B, Ppad));
    launch_ao_rho(s, B, G, T, part, rho);
    c.mark(PH_NLC_RHO);
    double* wv = (double*)c.ws_vv.ensure(sizeof(double) * 5 * Ppad);
    vv10_flow(s, G.nsub, Ppad, G.d_xyz.as<double>(), G.d_w.as<double>(), G.origin, rho, b, Cc,
              tol, f64, c.ws_B, wv, wv + 4 * Ppad, d_out2, c.profile ? &c : nullptr,
              c.vv_cap16);
    c.vv_stat = (const unsigned long long*)(c.ws_B.as<char>() + vv_layout(G.nsub, Ppad, f64).stat);
    double* Bm = T;    // T is dead after the density: reuse for B
    launch_ao_proj(s, B, G, wv, Bm);
    double* V = (double*)c.ws_K.ensure(sizeof(double) * nao * nao);
    // X B^T over each point split's significant AO rows of X
    const NtGather ng = gather ? nt_gather(AL, lws) : NtGather{};
    gemm_nt_fast(s, nao, nao, Ppad, X, Bm, false, 1, 0, nullptr, V, c.ws_Kp, nullptr, nullptr, 0,
                 gather ? &ng : nullptr);
    symadd(s, nao, V, d_vmat);
    c.mark(PH_NLC_FOCK);
    collect(c);
}

}  // namespace wb

using namespace wb;

#define WB_TRY(...)                                                      \
    try {                                                                \
        __VA_ARGS__;                                                     \
        return WB_OK;                                                    \
    } catch (const wb::Error& e) {                                       \
        set_last_error(e.what());                                        \
        return e.code;                                                   \
    } catch (const std::bad_alloc&) {                                    \
        set_last_error("host allocation failed");                        \
        return WB_OOM;                                                   \
    } catch (const std::exception& e) {                                  \
        set_last_error(e.what());                                        \
        return WB_DEVICE_ERROR;                                          \
    }

extern "C" {

const char* wb_last_error(void) { return g_last_error.c_str(); }

int wb_version(void) { return WB_ABI_VERSION; }

int wb_ctx_create(int device, void* stream, wb_ctx_t* out)
{
    WB_TRY({
        if (!out) throw Error(WB_INVALID, "null output");
        auto* c = new Ctx();
        c->device = device;
        WB_HIP(hipSetDevice(device));
        resolve_scratch(device);
        if (stream) {
            c->stream = (hipStream_t)stream;
        } else {
            WB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            c->own_stream = true;
        }
        std::vector<float> t32;
        std::vector<double> t64;
        boys_tables(t32, t64);
        upload(c->d_b32, t32.data(), sizeof(float) * t32.size(), c->stream);
        upload(c->d_b64, t64.data(), sizeof(double) * t64.size(), c->stream);
        WB_HIP(hipStreamSynchronize(c->stream));
        *out = c;
    })
}

namespace wb {
size_t Ctx::ws_bytes() const
{
    size_t n = 0;
    for (const DevBuf* d : bufs) n += d->bytes;
    for (const auto& l : lanes) n += l->ws_bytes();
    if (side) n += side->ws_bytes();
    return n;
}
void Ctx::trim()
{
    for (DevBuf* d : bufs) d->release();
    for (auto& l : lanes) l->trim();
    if (side) side->trim();
}
Ctx::~Ctx()
{
    lanes.clear();
    side.reset();
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_gate) (void)hipEventDestroy(ev_gate);
    if (t0) (void)hipEventDestroy(t0);
    if (own_stream && stream) {
        (void)hipStreamSynchronize(stream);
        (void)hipStreamDestroy(stream);
    }
}

// a context on c's device with its own stream and workspaces (batch lanes,
// the VV10 branch of a concurrent cycle)
static std::unique_ptr<Ctx> make_sub_ctx(const Ctx& c)
{
    auto l = std::make_unique<Ctx>();
    l->device = c.device;
    WB_HIP(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking));
    l->own_stream = true;
    std::vector<float> t32;
    std::vector<double> t64;
    boys_tables(t32, t64);
    upload(l->d_b32, t32.data(), sizeof(float) * t32.size(), l->stream);
    upload(l->d_b64, t64.data(), sizeof(double) * t64.size(), l->stream);
    return l;
}
}  // namespace wb

int wb_ctx_mem(wb_ctx_t ctx, long long* bytes)
{
    WB_TRY({
        if (!ctx || !bytes) throw Error(WB_INVALID, "null argument");
        *bytes = (long long)((Ctx*)ctx)->ws_bytes();
    })
}

int wb_ctx_trim(wb_ctx_t ctx)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c) throw Error(WB_INVALID, "null context");
        WB_HIP(hipSetDevice(c->device));
        WB_HIP(hipStreamSynchronize(c->stream));
        for (auto& l : c->lanes) WB_HIP(hipStreamSynchronize(l->stream));
        if (c->side) WB_HIP(hipStreamSynchronize(c->side->stream));
        c->trim();
    })
}

int wb_ctx_destroy(wb_ctx_t ctx)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c) return WB_OK;
        WB_HIP(hipSetDevice(c->device));
        WB_HIP(hipStreamSynchronize(c->stream));
        delete c;                       // the destructor waits for the sub-contexts' streams
    })
}

int wb_ctx_set_copies(wb_ctx_t ctx, int ncopy)
{
    WB_TRY({
        if (ncopy < 1 || ncopy > 64) throw Error(WB_INVALID, "ncopy out of range");
        ((Ctx*)ctx)->ncopy = ncopy;
    })
}

int wb_ctx_get_class_waves(wb_ctx_t ctx, int* waves10)
{
    WB_TRY({
        if (!ctx || !waves10) throw Error(WB_INVALID, "null argument");
        std::copy(((Ctx*)ctx)->cls_waves, ((Ctx*)ctx)->cls_waves + NCLASS, waves10);
    })
}

int wb_ctx_set_class_waves(wb_ctx_t ctx, const int* waves10)
{
    WB_TRY({
        if (!ctx) throw Error(WB_INVALID, "null context");
        int w[NCLASS] = {0};
        if (waves10)
            for (int k = 0; k < NCLASS; ++k) {
                if (waves10[k] < 0 || waves10[k] > 64) throw Error(WB_INVALID, "waves out of range");
                w[k] = waves10[k];
            }
        std::copy(w, w + NCLASS, ((Ctx*)ctx)->cls_waves);
    })
}

int wb_ctx_set_k_budget(wb_ctx_t ctx, long long bytes)
{
    WB_TRY({
        if (!ctx || bytes < 0) throw Error(WB_INVALID, "null context or negative budget");
        ((Ctx*)ctx)->k_budget = (size_t)bytes;
    })
}

int wb_k_plan(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, int nop, int precision,
              long long* out3)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c || !basis || !grid || !out3) throw Error(WB_INVALID, "null argument");
        if (nop < 1 || nop > 8) throw Error(WB_INVALID, "nop out of range");
        const bool f32 = precision == WB_PREC_MIXED;
        const KPlan P = k_plan_budget(*(Basis*)basis, *(Grid*)grid, nop, f32, !f32 || c->g64,
                                      c->ncopy, c->k_budget);
        out3[0] = (long long)P.total;
        out3[1] = P.nchunk;
        out3[2] = P.csub;
    })
}

int wb_ctx_set_g64(wb_ctx_t ctx, int on)
{
    WB_TRY({ ((Ctx*)ctx)->g64 = on != 0; })
}

int wb_basis_create(wb_ctx_t ctx, int nshell, const int* l, const int* nprim,
                    const double* center, const double* exps, const double* coefs,
                    double pair_cut, wb_basis_t* out)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        WB_HIP(hipSetDevice(c->device));
        auto* b = new Basis();
        try {
            build_basis(*b, nshell, l, nprim, center, exps, coefs, pair_cut);
            upload_basis(*b, c->stream);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    })
}

int wb_basis_info(wb_basis_t basis, long long* out4)
{
    WB_TRY({
        auto* b = (Basis*)basis;
        long long np = 0;
        for (auto& C : b->cls) np += C.nprims;
        out4[0] = b->nao;
        out4[1] = b->nshell;
        out4[2] = b->npairs_total;
        out4[3] = np;
    })
}

int wb_basis_destroy(wb_basis_t basis)
{
    WB_TRY({ delete (Basis*)basis; })
}

int wb_basis_bytes(wb_basis_t basis, int nop, long long* out)
{
    WB_TRY({
        auto* b = (Basis*)basis;
        if (!b || !out) throw Error(WB_INVALID, "null argument");
        if (nop < 0 || nop > 8) throw Error(WB_INVALID, "nop out of range");
        size_t n = b->d_shells.bytes + b->d_exps.bytes + b->d_coefs.bytes + b->d_shao.bytes +
                   b->d_row0.bytes;
        for (auto& C : b->cls) {
            n += C.d_pairs.bytes + C.d_scr.bytes + C.d_groups.bytes + C.d_bbao.bytes +
                 C.d_pbc.bytes;
            std::lock_guard<std::mutex> lock(C.opc_mu);
            // records for nop operators cached: one set for all of them (a
            // fused class), or one set per operator
            int sets1 = 0;
            bool setn = false;
            for (const auto& kv : C.opc) {
                n += kv.second.first->bytes + kv.second.second->bytes;
                if (kv.first.size() == 1) ++sets1;
                if ((int)kv.first.size() == nop) setn = true;
            }
            const bool have = setn || sets1 >= nop;
            if (!have && nop > 0 && C.nprims > 0) {
                // what a K build of nop operators adds: one record set for
                // the operators together where the class runs them fused,
                // else one per operator, each in FP32 and FP64
                const size_t one = (size_t)md_rec_bytes(C.la, C.lb, true, 1) +
                                   md_rec_bytes(C.la, C.lb, false, 1);
                const size_t two = (size_t)md_rec_bytes(C.la, C.lb, true, 2) +
                                   md_rec_bytes(C.la, C.lb, false, 2);
                n += (size_t)C.nprims * ((nop == 2 && md_fuse(C.la, C.lb)) ? std::max(two, 2 * one)
                                                                          : nop * one);
            }
        }
        *out = (long long)n;
    })
}

int wb_grid_bytes(wb_grid_t grid, long long* out)
{
    WB_TRY({
        auto* g = (Grid*)grid;
        if (!g || !out) throw Error(WB_INVALID, "null argument");
        *out = (long long)(g->d_xyz.bytes + g->d_w.bytes + g->d_sub.bytes);
    })
}

int wb_grid_create(wb_ctx_t ctx, int npts, const double* coords, const double* weights,
                   wb_grid_t* out)
{
    WB_TRY({
        if (npts < 1) throw Error(WB_INVALID, "empty grid");
        auto* c = (Ctx*)ctx;
        WB_HIP(hipSetDevice(c->device));
        auto* g = new Grid();
        try {
            build_grid(*g, npts, coords, weights, c->stream);
        } catch (...) {
            delete g;
            throw;
        }
        *out = g;
    })
}

int wb_grid_info(wb_grid_t grid, long long* out3)
{
    WB_TRY({
        auto* g = (Grid*)grid;
        out3[0] = g->npts;
        out3[1] = g->Ppad;
        out3[2] = g->nsub;
    })
}

int wb_grid_destroy(wb_grid_t grid)
{
    WB_TRY({ delete (Grid*)grid; })
}

int wb_k_build_dev(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* d_qd,
                   int nop, const double* omegas, double tile_tol, int precision,
                   double* d_k)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (nop < 1 || nop > 8) throw Error(WB_INVALID, "nop out of range");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        WB_HIP(hipSetDevice(c->device));
        k_build_dev(*c, *(Basis*)basis, *(Grid*)grid, d_qd, nop, omegas, tile_tol, precision, d_k);
    })
}

int wb_k_build_noalloc(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* d_qd,
                       int nop, const double* omegas, double tile_tol, int precision,
                       double* d_k)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c || !basis || !grid || !d_qd || !omegas || !d_k) throw Error(WB_INVALID, "null argument");
        if (nop < 1 || nop > 8) throw Error(WB_INVALID, "nop out of range");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        WB_HIP(hipSetDevice(c->device));
        k_build_noalloc(*c, *(Basis*)basis, *(Grid*)grid, d_qd, nop, omegas, tile_tol, precision,
                        d_k);
    })
}

int wb_k_build(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* qd, int nop,
               const double* omegas, double tile_tol, int precision, double* k_out)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        auto* b = (Basis*)basis;
        if (nop < 1 || nop > 8) throw Error(WB_INVALID, "nop out of range");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        WB_HIP(hipSetDevice(c->device));
        const size_t nn = (size_t)b->nao * b->nao;
        double* d_in = (double*)c->ws_K.ensure(sizeof(double) * nn * (1 + nop));
        WB_HIP(hipMemcpyAsync(d_in, qd, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
        auto t0 = std::chrono::steady_clock::now();
        k_build_dev(*c, *b, *(Grid*)grid, d_in, nop, omegas, tile_tol, precision, d_in + nn);
        WB_HIP(hipMemcpyAsync(k_out, d_in + nn, sizeof(double) * nn * nop, hipMemcpyDeviceToHost,
                              c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
        c->ms_last = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    })
}

// Times the K build on the caller's molecule for each configuration of the
// context's run-time choices and keeps the fastest: the number of G copies
// (1, 2, 4, 8, 16, those whose G fits in a quarter of the free device
// memory), then, one class at a time, whether a class with a fused
// two-operator kernel runs it fused or as two passes.  Each configuration is
// timed as the minimum of `reps` builds (HIP events, after one untimed).
int wb_ctx_autotune(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* qd, int nop,
                    const double* omegas, double tile_tol, int precision, int reps,
                    double* report)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        auto* b = (Basis*)basis;
        auto* g = (Grid*)grid;
        if (!c || !b || !g || !qd || !omegas) throw Error(WB_INVALID, "null argument");
        if (nop < 1 || nop > 8) throw Error(WB_INVALID, "nop out of range");
        if (reps < 1 || reps > 100) throw Error(WB_INVALID, "reps must be 1..100");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        WB_HIP(hipSetDevice(c->device));
        const size_t nn = (size_t)b->nao * b->nao;
        hipEvent_t e0, e1;
        WB_HIP(hipEventCreate(&e0));
        WB_HIP(hipEventCreate(&e1));
        int trials = 0;
        bool waves_set = false;                // the per-class search below sets them
        int waves0[NCLASS];
        std::copy(c->cls_waves, c->cls_waves + NCLASS, waves0);
        auto time_cfg = [&](int ncopy, unsigned mask, int tiles) {
            c->ncopy = ncopy;
            c->fuse_mask = mask;
            c->gemm_tiles = tiles;
            if (!waves_set) std::fill(c->cls_waves, c->cls_waves + NCLASS, 0);
            double* d = (double*)c->ws_K.ensure(sizeof(double) * nn * (1 + nop));
            WB_HIP(hipMemcpyAsync(d, qd, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
            k_build_dev(*c, *b, *g, d, nop, omegas, tile_tol, precision, d + nn);
            float best = 1e30f;
            for (int r = 0; r < reps; ++r) {
                WB_HIP(hipEventRecord(e0, c->stream));
                k_build_dev(*c, *b, *g, d, nop, omegas, tile_tol, precision, d + nn);
                WB_HIP(hipEventRecord(e1, c->stream));
                WB_HIP(hipEventSynchronize(e1));
                float ms = 0.f;
                WB_HIP(hipEventElapsedTime(&ms, e0, e1));
                best = std::min(best, ms);
            }
            ++trials;
            return best;
        };
        const int ncopy0 = c->ncopy;
        const unsigned mask0 = c->fuse_mask;
        const int tiles0 = c->gemm_tiles;
        try {
            const float t_default = time_cfg(8, ~0u, 0);
            int best_n = 8, best_tiles = 0;
            unsigned best_mask = ~0u;
            float best_t = t_default;
            size_t free_b = 0, total_b = 0;
            WB_HIP(hipMemGetInfo(&free_b, &total_b));
            const bool g64 = precision == WB_PREC_FP64 || c->g64;
            for (int n : {1, 2, 4, 16}) {
                const size_t gbytes = (size_t)n * (b->nao + MD_SPARE + 1) * g->Ppad * nop *
                                      (g64 ? 8 : 4);
                if (gbytes > (free_b + c->ws_karena.bytes) / 4) continue;
                const float t = time_cfg(n, best_mask, 0);
                if (t < best_t) { best_t = t; best_n = n; }
            }
            if (nop == 2)
                for (int k = 0; k < NCLASS; ++k) {
                    const ClassData& C = b->cls[k];
                    if (!C.npairs || !md_fuse(C.la, C.lb)) continue;
                    const unsigned m = best_mask & ~(1u << k);
                    const float t = time_cfg(best_n, m, 0);
                    if (t < best_t) { best_t = t; best_mask = m; }
                }
            // GEMM wave tiles: both shapes besides the padding rule's choice
            for (int tl : {3, 4}) {
                const float t = time_cfg(best_n, best_mask, tl);
                if (t < best_t) { best_t = t; best_tiles = tl; }
            }
            // waves per sub-block, class by class (all but the first class
            // launched, which initializes every copy): halve while faster
            waves_set = true;
            std::fill(c->cls_waves, c->cls_waves + NCLASS, 0);
            bool first_cls = true;
            for (int k = 0; k < NCLASS; ++k) {
                if (!b->cls[k].npairs) continue;
                if (first_cls) { first_cls = false; continue; }
                for (int wv = best_n / 2; wv >= 1; wv /= 2) {
                    c->cls_waves[k] = wv;
                    const float t = time_cfg(best_n, best_mask, best_tiles);
                    if (t < best_t) best_t = t;
                    else { c->cls_waves[k] = wv * 2 >= best_n ? 0 : wv * 2; break; }
                }
            }
            c->ncopy = best_n;
            c->fuse_mask = best_mask;
            c->gemm_tiles = best_tiles;
            // the largest configuration tried sized G: let it regrow to the kept one
            WB_HIP(hipStreamSynchronize(c->stream));
            c->ws_karena.release();
            if (report) {
                report[0] = best_n;
                report[1] = (double)best_mask;
                report[2] = best_t;
                report[3] = t_default;
                report[4] = trials;
                report[5] = best_tiles;
            }
        } catch (...) {
            c->ncopy = ncopy0;
            c->fuse_mask = mask0;
            c->gemm_tiles = tiles0;
            std::copy(waves0, waves0 + NCLASS, c->cls_waves);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        WB_HIP(hipEventDestroy(e0));
        WB_HIP(hipEventDestroy(e1));
    })
}

int wb_ctx_get_tuning(wb_ctx_t ctx, int* ncopy, unsigned* fuse_mask, int* gemm_tiles)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c || !ncopy || !fuse_mask || !gemm_tiles) throw Error(WB_INVALID, "null argument");
        *ncopy = c->ncopy;
        *fuse_mask = c->fuse_mask;
        *gemm_tiles = c->gemm_tiles;
    })
}

int wb_ctx_set_tuning(wb_ctx_t ctx, int ncopy, unsigned fuse_mask, int gemm_tiles)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c) throw Error(WB_INVALID, "null context");
        if (ncopy < 1 || ncopy > 64) throw Error(WB_INVALID, "ncopy out of range");
        if (gemm_tiles != 0 && gemm_tiles != 3 && gemm_tiles != 4)
            throw Error(WB_INVALID, "gemm_tiles must be 0, 3 or 4");
        c->ncopy = ncopy;
        c->fuse_mask = fuse_mask;
        c->gemm_tiles = gemm_tiles;
    })
}

int wb_k_stats(wb_ctx_t ctx, wb_grid_t grid, long long* out2)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        out2[0] = c->tasks_total;
        out2[1] = count_kept(*c);
        (void)grid;
    })
}

int wb_nlc_vxc_dev(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* d_dm,
                   double b, double C, double tile_tol, int precision, double* d_out2,
                   double* d_vmat)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        WB_HIP(hipSetDevice(c->device));
        nlc_dev(*c, *(Basis*)basis, *(Grid*)grid, d_dm, b, C, tile_tol, precision, d_out2, d_vmat);
    })
}

// ------------------------------------------------------- semilocal XC ---
// The device side of nr_rks (reference gpu/xc.py:130-201): the densities on
// the grid and the potential matrix from the chain-ruled potential; the
// functional itself is evaluated by the caller (PySCF's libxc, as in the
// reference).  FP64 throughout.
// rho: [5][Ppad] (rho, d rho / dx, dy, dz, tau) in the grid's internal order
static void xc_rho_dev(Ctx& c, Basis& B, Grid& G, const double* d_dm, bool mgga, double* rho)
{
    const int nao = B.nao, Ppad = G.Ppad;
    hipStream_t s = c.stream;
    const size_t plane = (size_t)nao * Ppad;
    double* X = (double*)c.ws_Xv.ensure(sizeof(double) * plane * (mgga ? 4 : 1));
    // the GEMMs over each tile's significant AOs (values, and gradients when
    // stored), as in the VV10 flow
    const bool gather = gemm_nn_gather_ok(nao, Ppad, d_dm, X);
    int kc = 0;
    nt_split(Ppad, kc);
    const AoLists AL = ao_lists_layout(B, G.nsub, kc);
    void* lws = gather ? c.ws_list.ensure(AL.total) : nullptr;
    if (gather) launch_ao_eval_lists(s, B, G, mgga ? 1 : 0, X, lws, AL);
    else launch_ao_eval(s, B, G, mgga ? 1 : 0, X);
    auto nn = [&](const double* Xp, double* out) {             // out = D Xp (D symmetric)
        if (gather) gemm_nn_gather(s, nao, Ppad, d_dm, Xp, out, false, lws, AL);
        else gemm_nn_fast(s, nao, Ppad, nao, d_dm, Xp, out, false);
    };
    double* T = (double*)c.ws_T.ensure(sizeof(double) * plane);
    nn(X, T);
    double* part = (double*)c.ws_red.ensure(ao_rho_bytes(B, Ppad));
    launch_ao_rho(s, B, G, T, part, rho);
    if (mgga) {
        // tau = 1/2 sum_d sum_mu (D dX_d)_mu dX_d,mu
        for (int d = 0; d < 3; ++d) {
            const double* Xd = X + (size_t)(d + 1) * plane;
            nn(Xd, T);
            launch_dot_planes(s, nao, Ppad, T, Xd, 0.5, d > 0, rho + 4 * (size_t)Ppad);
        }
    } else {
        zero_bytes(s, rho + 4 * (size_t)Ppad, sizeof(double) * Ppad);
    }
}

// vmat = sum_g w [ v0 X X^T + v_d (dX_d X^T + X dX_d^T) + 1/2 v4 sum_d dX_d dX_d^T ],
// v: [5][Ppad] chain-ruled potential (internal order, padding 0)
static void xc_vmat_dev(Ctx& c, Basis& B, Grid& G, const double* v, bool mgga, double* d_vmat)
{
    const int nao = B.nao, Ppad = G.Ppad;
    hipStream_t s = c.stream;
    const size_t plane = (size_t)nao * Ppad;
    double* X = (double*)c.ws_Xv.ensure(sizeof(double) * plane * (mgga ? 4 : 1));
    // the GEMMs over each point split's significant AO rows (values, and
    // gradients when stored)
    const bool gather = gemm_nn_gather_ok(nao, Ppad, X, X);
    int kc = 0;
    nt_split(Ppad, kc);
    const AoLists AL = ao_lists_layout(B, G.nsub, kc);
    void* lws = gather ? c.ws_list.ensure(AL.total) : nullptr;
    if (gather) launch_ao_eval_lists(s, B, G, mgga ? 1 : 0, X, lws, AL);
    else launch_ao_eval(s, B, G, mgga ? 1 : 0, X);
    const NtGather ng = gather ? nt_gather(AL, lws) : NtGather{};
    const NtGather* gp = gather ? &ng : nullptr;
    // B = (w v0 / 2) X + w v_d dX_d (weights and the 1/2 of the symmetric sum)
    double* wv = (double*)c.ws_vv.ensure(sizeof(double) * 5 * Ppad);
    launch_xc_weights(s, Ppad, G.d_w.as<double>(), v, wv);
    double* Bm = (double*)c.ws_T.ensure(sizeof(double) * plane);
    launch_ao_proj(s, B, G, wv, Bm);
    double* V = (double*)c.ws_K.ensure(sizeof(double) * (size_t)nao * nao * (mgga ? 2 : 1));
    gemm_nt_fast(s, nao, nao, Ppad, X, Bm, false, 1, 0, nullptr, V, c.ws_Kp, nullptr, nullptr, 0,
                 gp);
    symadd(s, nao, V, d_vmat);                                  // vmat = V + V^T
    if (mgga) {
        // + sum_d dX_d diag(w v4 / 2) dX_d^T (added once: already symmetric)
        double* Vt = V + (size_t)nao * nao;
        for (int d = 0; d < 3; ++d) {
            const double* Xd = X + (size_t)(d + 1) * plane;
            gemm_nt_fast(s, nao, nao, Ppad, Xd, Xd, false, 1, 0, wv + 4 * (size_t)Ppad, Vt, c.ws_Kp,
                         nullptr, nullptr, 0, gp);
            launch_axpy(s, (size_t)nao * nao, 1.0, Vt, d_vmat);
        }
    }
}

// ---------------------------------------------------------- cycle batch ---
// Many molecules' cycles (K build and/or VV10) drained from one shared queue
// by `lanes` host threads, each driving a sub-context with its own stream and
// workspaces, so small molecules' launches overlap on the device and large
// ones keep it full.  Jobs are taken largest first.  Each job runs exactly
// the single-molecule code on a context of its own, so its outputs equal
// wb_k_build's / wb_nlc_vxc's for it alone, bit for bit.
static void cycle_dev(Ctx& c, Basis& B, Grid* kg, const double* d_qd, int nop,
                      const double* omegas, double k_tol, Grid* vg, const double* d_dm, double b,
                      double Cc, double vv_tol, int prec, double* d_k, double* d_out2,
                      double* d_vmat);

// one job on context w with host buffers: inputs up, the cycle (K build and
// VV10 concurrent when both are asked for), outputs down
static void run_cycle_job(Ctx& w, const wb_cycle_job& j, int nop, const double* omegas,
                          double k_tol, double b, double Cc, double vv_tol, int precision)
{
    Basis& B = *(Basis*)j.basis;
    const size_t nn = (size_t)B.nao * B.nao;
    double *dk = nullptr, *dv = nullptr;
    if (j.kgrid) {
        dk = (double*)w.ws_K.ensure(sizeof(double) * nn * (1 + nop));
        WB_HIP(hipMemcpyAsync(dk, j.qd, sizeof(double) * nn, hipMemcpyHostToDevice, w.stream));
    }
    if (j.nlcgrid) {
        dv = (double*)w.ws_io.ensure(sizeof(double) * (2 * nn + 2));
        WB_HIP(hipMemcpyAsync(dv, j.dm, sizeof(double) * nn, hipMemcpyHostToDevice, w.stream));
    }
    if (j.flags & WB_JOB_SEQUENTIAL) {
        if (j.kgrid) k_build_dev(w, B, *(Grid*)j.kgrid, dk, nop, omegas, k_tol, precision, dk + nn);
        if (j.nlcgrid)
            nlc_dev(w, B, *(Grid*)j.nlcgrid, dv, b, Cc, vv_tol, precision, dv + 2 * nn, dv + nn);
    } else {
        cycle_dev(w, B, (Grid*)j.kgrid, dk, nop, omegas, k_tol, (Grid*)j.nlcgrid, dv, b, Cc,
                  vv_tol, precision, dk ? dk + nn : nullptr, dv ? dv + 2 * nn : nullptr,
                  dv ? dv + nn : nullptr);
    }
    if (j.kgrid)
        WB_HIP(hipMemcpyAsync(j.k, dk + nn, sizeof(double) * nn * nop, hipMemcpyDeviceToHost,
                              w.stream));
    if (j.nlcgrid) {
        WB_HIP(hipMemcpyAsync(j.vmat, dv + nn, sizeof(double) * nn, hipMemcpyDeviceToHost,
                              w.stream));
        WB_HIP(hipMemcpyAsync(j.nlc, dv + 2 * nn, sizeof(double) * 2, hipMemcpyDeviceToHost,
                              w.stream));
    }
    WB_HIP(hipStreamSynchronize(w.stream));
}

// Bytes each workspace needs for one job (what k_build_dev, nlc_dev and
// run_cycle_job ensure(); a buffer both flows use on one context needs the
// larger size).  need: the job's context; side: its VV10 side context, which
// runs the VV10 flow when the job has both a K build and a VV10 term.
static void job_needs(const Ctx& w, const wb_cycle_job& j, int nop, int prec,
                      size_t need[WS_COUNT], size_t side[WS_COUNT])
{
    for (int b = 0; b < WS_COUNT; ++b) need[b] = side[b] = 0;
    const Basis& B = *(const Basis*)j.basis;
    const size_t nao = B.nao, nn = nao * nao;
    const bool f32 = prec == WB_PREC_MIXED;
    auto at_least = [](size_t* n, int b, size_t v) { n[b] = std::max(n[b], v); };
    if (j.kgrid) {
        const Grid& G = *(const Grid*)j.kgrid;
        at_least(need, WS_KARENA, k_plan_budget(B, G, nop, f32, !f32 || w.g64, w.ncopy,
                                                w.k_budget).total);
        at_least(need, WS_WORK, 8 * 2 * NCLASS);
        at_least(need, WS_K, 8 * nn * (1 + nop));
    }
    if (j.nlcgrid) {
        const Grid& G = *(const Grid*)j.nlcgrid;
        const size_t P = G.Ppad;
        // the VV10 flow runs on the side context unless the job is sequential
        // or has no K build
        size_t* v = (j.kgrid && !(j.flags & WB_JOB_SEQUENTIAL) && !w.profile) ? side : need;
        at_least(v, WS_XV, 8 * nao * P);
        {
            int kc = 0;
            nt_split((int)P, kc);
            at_least(v, WS_LIST, ao_lists_layout(B, G.nsub, kc).total);
        }
        at_least(v, WS_T, 8 * nao * P);
        at_least(v, WS_RHO, 8 * 4 * P);
        at_least(v, WS_RED, ao_rho_bytes(B, (int)P));
        at_least(v, WS_VV, 8 * 5 * P);
        at_least(v, WS_B, vv_layout(G.nsub, (int)P, !f32).total);
        at_least(v, WS_KP, gemm_nt_ws_bytes((int)nao, (int)nao, (int)P, false));
        at_least(v, WS_K, 8 * nn);
        at_least(need, WS_IO, 8 * (2 * nn + 2));
    }
}

// bytes the context (with its VV10 side context) holds after a job with
// these needs; batch lanes and sub-contexts of other kinds are not counted
static size_t held_after(const Ctx& w, const size_t need[WS_COUNT], const size_t side[WS_COUNT])
{
    size_t n = 0;
    for (int b = 0; b < WS_COUNT; ++b) {
        n += std::max(w.bufs[b]->bytes, need[b]);
        n += std::max(w.side ? w.side->bufs[b]->bytes : (size_t)0, side[b]);
    }
    return n;
}

static Ctx& batch_lane(Ctx& c, size_t i)
{
    while (c.lanes.size() <= i) c.lanes.push_back(make_sub_ctx(c));
    Ctx& l = *c.lanes[i];
    l.ncopy = c.ncopy;
    l.fuse_mask = c.fuse_mask;
    l.gemm_tiles = c.gemm_tiles;
    l.g64 = c.g64;
    l.k_budget = c.k_budget;
    std::copy(c.cls_waves, c.cls_waves + NCLASS, l.cls_waves);
    return l;
}

int wb_cycle(wb_ctx_t ctx, const wb_cycle_job* job, int nop, const double* omegas,
             double k_tol, double b, double C, double vv_tol, int precision)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c || !job || !job->basis) throw Error(WB_INVALID, "null argument");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        if (job->kgrid && (nop < 1 || nop > 8 || !omegas || !job->qd || !job->k))
            throw Error(WB_INVALID, "K build: nop out of range or null buffer");
        if (job->nlcgrid && (!job->dm || !job->nlc || !job->vmat))
            throw Error(WB_INVALID, "VV10: null buffer");
        WB_HIP(hipSetDevice(c->device));
        run_cycle_job(*c, *job, nop, omegas, k_tol, b, C, vv_tol, precision);
    })
}

int wb_job_bytes(wb_ctx_t ctx, const wb_cycle_job* job, int nop, int precision,
                 long long* after, long long* alone)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c || !job || !job->basis) throw Error(WB_INVALID, "null argument");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        if (job->kgrid && (nop < 1 || nop > 8)) throw Error(WB_INVALID, "nop out of range");
        size_t need[WS_COUNT], side[WS_COUNT];
        job_needs(*c, *job, nop, precision, need, side);
        size_t sum = 0;
        for (int k = 0; k < WS_COUNT; ++k) sum += need[k] + side[k];
        if (after) *after = (long long)held_after(*c, need, side);
        if (alone) *alone = (long long)sum;
    })
}

int wb_cycle_batch(wb_ctx_t ctx, int njob, const wb_cycle_job* jobs, int nop,
                   const double* omegas, double k_tol, double b, double C, double vv_tol,
                   int precision, int lanes, double* job_ms)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (njob < 0 || (njob > 0 && !jobs)) throw Error(WB_INVALID, "bad job list");
        if (lanes < 1 || lanes > 16) throw Error(WB_INVALID, "lanes must be 1..16");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        bool anyk = false;
        for (int i = 0; i < njob; ++i) {
            const wb_cycle_job& j = jobs[i];
            if (!j.basis) throw Error(WB_INVALID, "job " + std::to_string(i) + ": null basis");
            if (j.kgrid && (!j.qd || !j.k))
                throw Error(WB_INVALID, "job " + std::to_string(i) + ": K build needs qd and k");
            if (j.nlcgrid && (!j.dm || !j.nlc || !j.vmat))
                throw Error(WB_INVALID, "job " + std::to_string(i) + ": VV10 needs dm, nlc, vmat");
            anyk |= j.kgrid != nullptr;
        }
        if (anyk && (nop < 1 || nop > 8 || !omegas)) throw Error(WB_INVALID, "nop out of range");
        WB_HIP(hipSetDevice(c->device));
        WB_HIP(hipStreamSynchronize(c->stream));     // work queued on the caller's context first
        // largest first: nao^2 x (K points + VV10 points) as the size estimate
        std::vector<int> order(njob);
        std::vector<double> est(njob);
        for (int i = 0; i < njob; ++i) {
            const Basis& B = *(const Basis*)jobs[i].basis;
            const double pk = jobs[i].kgrid ? ((const Grid*)jobs[i].kgrid)->Ppad : 0;
            const double pv = jobs[i].nlcgrid ? ((const Grid*)jobs[i].nlcgrid)->Ppad : 0;
            est[i] = (double)B.nao * B.nao * (pk + pv);
            order[i] = i;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return est[x] > est[y]; });
        const size_t nl = (size_t)std::max(1, std::min(lanes, njob));
        std::vector<Ctx*> lc(nl);
        for (size_t i = 0; i < nl; ++i) lc[i] = &batch_lane(*c, i);
        std::atomic<int> next{0};
        std::atomic<bool> failed{false};
        std::vector<std::string> errs(nl);
        std::vector<int> errcode(nl, WB_OK);
        // Device memory: every lane's workspaces grow to the largest job it
        // has run.  A lane starts its next job only when the bytes it will
        // hold afterwards (exact, from the job's buffer needs) plus what the
        // other lanes hold or will hold fit the device memory that was free
        // at the start plus the lanes' own; an idle lane that holds memory
        // releases it when the job does not fit, and a job that does not fit
        // even alone runs with the device to itself.  A job that still runs
        // out of memory is retried once alone after every workspace of the
        // context and its lanes has been released.
        size_t free_b = 0, total_b = 0;
        WB_HIP(hipMemGetInfo(&free_b, &total_b));
        size_t lanes_b = 0;
        for (Ctx* l : lc) lanes_b += l->ws_bytes();
        const size_t reserve = std::min<size_t>(total_b / 32, (size_t)8 << 30);
        const size_t budget = free_b + lanes_b > reserve ? free_b + lanes_b - reserve : 0;
        std::mutex gate_mu;
        std::condition_variable gate_cv;
        std::vector<size_t> committed(nl);
        for (size_t t = 0; t < nl; ++t) committed[t] = lc[t]->ws_bytes();
        std::vector<char> exited(nl, 0);
        int running = 0;
        bool exclusive = false;
        auto others_of = [&](size_t li) {
            size_t o = 0;
            for (size_t t = 0; t < nl; ++t)
                if (t != li) o += committed[t];
            return o;
        };
        // wait for room; returns with the lane counted as running
        auto enter = [&](size_t li, Ctx& w, const size_t* need, const size_t* side) {
            std::unique_lock<std::mutex> lk(gate_mu);
            for (;;) {
                if (!exclusive) {
                    const size_t after = held_after(w, need, side), others = others_of(li);
                    if (others + after <= budget || (others == 0 && running == 0)) {
                        committed[li] = after;
                        ++running;
                        return;
                    }
                    // release memory held by idle lanes: this one, and lanes
                    // that have left the queue (their streams are synchronized)
                    bool freed = false;
                    if (w.ws_bytes() > 0) {
                        (void)hipStreamSynchronize(w.stream);
                        w.trim();
                        committed[li] = 0;
                        freed = true;
                    }
                    for (size_t t = 0; t < nl; ++t)
                        if (exited[t] && committed[t] > 0) {
                            lc[t]->trim();
                            committed[t] = 0;
                            freed = true;
                        }
                    if (freed) {
                        gate_cv.notify_all();
                        continue;
                    }
                }
                gate_cv.wait(lk);
            }
        };
        auto leave = [&](size_t li, Ctx& w) {
            std::lock_guard<std::mutex> lk(gate_mu);
            --running;
            committed[li] = w.ws_bytes();
            gate_cv.notify_all();
        };
        auto run_alone = [&](size_t li, Ctx& w, int i) {
            {
                std::unique_lock<std::mutex> lk(gate_mu);
                gate_cv.wait(lk, [&] { return !exclusive; });
                exclusive = true;
                gate_cv.wait(lk, [&] { return running == 0; });
                c->trim();                   // every lane is idle and synchronized
            }
            auto done = [&] {
                std::lock_guard<std::mutex> lk(gate_mu);
                exclusive = false;
                for (size_t t = 0; t < nl; ++t) committed[t] = lc[t]->ws_bytes();
                gate_cv.notify_all();
            };
            try {
                run_cycle_job(w, jobs[i], nop, omegas, k_tol, b, C, vv_tol, precision);
            } catch (...) {
                (void)hipStreamSynchronize(w.stream);
                done();
                throw;
            }
            done();
            (void)li;
        };
        // WB_BATCH_FAIL_ONCE=i: job i's first attempt reports out-of-memory
        // before it starts (exercises the retry path from the host)
        const char* fe = std::getenv("WB_BATCH_FAIL_ONCE");
        const int fail_once = fe ? std::atoi(fe) : -1;
        const auto t0 = std::chrono::steady_clock::now();
        auto worker = [&](size_t li) {
            Ctx& w = *lc[li];
            int cur = -1;
            try {
                WB_HIP(hipSetDevice(w.device));
                for (;;) {
                    const int k = next.fetch_add(1);
                    if (k >= njob || failed.load()) break;
                    const int i = cur = order[k];
                    size_t need[WS_COUNT], side[WS_COUNT];
                    job_needs(w, jobs[i], nop, precision, need, side);
                    enter(li, w, need, side);
                    bool oom = false;
                    try {
                        if (i == fail_once) throw Error(WB_OOM, "forced out-of-memory");
                        run_cycle_job(w, jobs[i], nop, omegas, k_tol, b, C, vv_tol, precision);
                    } catch (const Error& e) {
                        if (e.code != WB_OOM) {
                            (void)hipStreamSynchronize(w.stream);
                            leave(li, w);
                            throw;
                        }
                        oom = true;
                    } catch (...) {
                        (void)hipStreamSynchronize(w.stream);
                        leave(li, w);
                        throw;
                    }
                    if (oom) {
                        (void)hipGetLastError();             // clear the failed allocation
                        (void)hipStreamSynchronize(w.stream);
                        leave(li, w);
                        run_alone(li, w, i);
                    } else {
                        leave(li, w);
                    }
                    if (job_ms)
                        job_ms[i] = std::chrono::duration<double, std::milli>(
                                        std::chrono::steady_clock::now() - t0).count();
                }
            } catch (const Error& e) {
                errcode[li] = e.code;
                errs[li] = e.what();
            } catch (const std::bad_alloc&) {
                errcode[li] = WB_OOM;
                errs[li] = "host out of memory";
            } catch (const std::exception& e) {
                errcode[li] = WB_DEVICE_ERROR;
                errs[li] = e.what();
            }
            if (errcode[li] != WB_OK) {
                failed = true;
                errs[li] = "job " + std::to_string(cur) + ": " + errs[li];
            }
            std::lock_guard<std::mutex> lk(gate_mu);
            exited[li] = 1;
            gate_cv.notify_all();
        };
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nl; ++t) pool.emplace_back(worker, t);
        worker(0);
        for (auto& th : pool) th.join();
        for (size_t t = 0; t < nl; ++t)
            if (errcode[t] != WB_OK) throw Error(errcode[t], "cycle batch: " + errs[t]);
    })
}

// ------------------------------------------------------------ cycle ---
// One SCF cycle's device work: the K build on kg and the VV10 term on vg.
// With both (and phase timing off) the VV10 flow runs on the side context's
// stream, forked from and joined back to c's stream by events, so that its
// FP32 pair sum overlaps the K build's FP64 MFMA GEMMs and the K build's
// exchange kernels overlap the VV10 GEMMs.  The two flows share no buffer
// (each context has its own workspaces) and their arithmetic is unchanged:
// the outputs equal the sequential cycle's bit for bit.
static Ctx& side_ctx(Ctx& c)
{
    if (!c.side) c.side = make_sub_ctx(c);
    if (!c.ev_fork) WB_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    if (!c.ev_join) WB_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
    if (!c.ev_gate) WB_HIP(hipEventCreateWithFlags(&c.ev_gate, hipEventDisableTiming));
    c.side->ncopy = c.ncopy;
    c.side->fuse_mask = c.fuse_mask;
    c.side->g64 = c.g64;
    c.side->k_budget = c.k_budget;
    std::copy(c.cls_waves, c.cls_waves + NCLASS, c.side->cls_waves);
    return *c.side;
}

static void cycle_dev(Ctx& c, Basis& B, Grid* kg, const double* d_qd, int nop,
                      const double* omegas, double k_tol, Grid* vg, const double* d_dm, double b,
                      double Cc, double vv_tol, int prec, double* d_k, double* d_out2,
                      double* d_vmat)
{
    if (kg && vg && !c.profile) {
        Ctx& sd = side_ctx(c);
        // The VV10 branch starts at the cycle start in the FP64 form, and
        // after the K build's F GEMM in the mixed form (measured on (H2O)27
        // def2-SVP: two gathered FP64 GEMMs side by side take longer than
        // one after the other; mixed 45.6 -> 44.9 ms, FP64 96.1 -> 97.8 ms
        // if it waited there too).  The whole K build is enqueued before
        // the branch waits for its start events (the fork orders it after
        // everything before the cycle, also when no gate was recorded).
        // The VV10 pair sum runs on a capped grid beside the K build (4
        // blocks per CU in FP32, 3 per 4 CUs in FP64), so that exchange
        // blocks keep finding room: mixed 44.7 -> 44.3 ms, FP64 94.7 -> 87.0
        // ms (uncapped, the FP64 pair sum's long blocks held most CUs).
        struct Gate {
            Ctx& c;
            Ctx& sd;
            ~Gate() { c.rec_gate = false; sd.vv_cap16 = 0; }
        } gate{c, sd};
        const bool late = prec == WB_PREC_MIXED;
        sd.vv_cap16 = late ? 64 : 12;
        c.rec_gate = late;
        WB_HIP(hipEventRecord(c.ev_fork, c.stream));
        k_build_dev(c, B, *kg, d_qd, nop, omegas, k_tol, prec, d_k);
        c.rec_gate = false;
        WB_HIP(hipStreamWaitEvent(sd.stream, c.ev_fork, 0));
        if (late) WB_HIP(hipStreamWaitEvent(sd.stream, c.ev_gate, 0));
        nlc_dev(sd, B, *vg, d_dm, b, Cc, vv_tol, prec, d_out2, d_vmat);
        WB_HIP(hipEventRecord(c.ev_join, sd.stream));
        WB_HIP(hipStreamWaitEvent(c.stream, c.ev_join, 0));
        c.vv_stat = sd.vv_stat;                       // ordered after the join
        return;
    }
    if (kg) k_build_dev(c, B, *kg, d_qd, nop, omegas, k_tol, prec, d_k);
    if (vg) nlc_dev(c, B, *vg, d_dm, b, Cc, vv_tol, prec, d_out2, d_vmat);
}

int wb_cycle_dev(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t kgrid, const double* d_qd, int nop,
                 const double* omegas, double k_tol, wb_grid_t nlcgrid, const double* d_dm,
                 double b, double C, double vv_tol, int precision, double* d_k, double* d_out2,
                 double* d_vmat)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c || !basis) throw Error(WB_INVALID, "null argument");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        if (kgrid && (nop < 1 || nop > 8 || !omegas || !d_qd || !d_k))
            throw Error(WB_INVALID, "K build: nop out of range or null buffer");
        if (nlcgrid && (!d_dm || !d_out2 || !d_vmat)) throw Error(WB_INVALID, "VV10: null buffer");
        WB_HIP(hipSetDevice(c->device));
        cycle_dev(*c, *(Basis*)basis, (Grid*)kgrid, d_qd, nop, omegas, k_tol, (Grid*)nlcgrid,
                  d_dm, b, C, vv_tol, precision, d_k, d_out2, d_vmat);
    })
}

// ------------------------------------------------------- cycle queue ---
// K builds of many molecules with the exchange kernels of all of them in one
// launch per shell-pair class (launch_md_queue): every molecule's set-up
// (AO values, F, screening, task lists) runs on the context's stream into an
// arena of its own, then one queue launch per class drains the (molecule,
// sub-block, chunk) work items of all of them, then each molecule's K GEMMs
// run.  Each molecule's buffers and arithmetic are those of its own
// k_build_arena, so its K equals wb_k_build's bit for bit.
struct QueueJob {
    int idx = 0;                        // the caller's job index
    Basis* B = nullptr;
    const Grid* G = nullptr;
    KPlan P;
    size_t arena = 0, io = 0;           // offsets in the group's arena: the plan's, QD and K's
};

static size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// work buffer of a group of n jobs: per job the per-class counters, one
// queue counter per class, then the work-item descriptors (at most nop per
// job and class) and their prefix sums
static size_t queue_work_bytes(size_t n, int nop)
{
    return align256(8 * (2 * NCLASS * n + NCLASS)) + align256(sizeof(MdJob) * NCLASS * n * nop) +
           align256(8 * NCLASS * (n * nop + 1));
}

// one group: every job's one-chunk plan, QD and K at its offsets in the
// context's K arena
static void k_build_queue_group(Ctx& c, const std::vector<QueueJob>& g, const wb_cycle_job* jobs,
                                int nop, const double* omegas, double tol, int prec)
{
    const bool f32 = prec == WB_PREC_MIXED;
    const bool g64 = !f32 || c.g64;
    hipStream_t s = c.stream;
    const size_t n = g.size();
    GemmTilesScope tiles(c.gemm_tiles);
    std::vector<KRecs> R(n);
    for (size_t i = 0; i < n; ++i)           // first: building records may allocate and synchronize
        k_records(c, *g[i].B, nop, omegas, f32, false, R[i]);
    size_t arena_b = 0;
    for (const auto& q : g)
        arena_b = std::max({arena_b, q.arena + q.P.total,
                            q.io + 8 * (size_t)q.B->nao * q.B->nao * (1 + nop)});
    char* arena = (char*)c.ws_karena.ensure(arena_b);
    char* io = arena;
    char* wbuf = (char*)c.ws_work.ensure(queue_work_bytes(n, nop));
    auto* work = (unsigned long long*)wbuf;
    unsigned long long* heads = work + 2 * NCLASS * n;
    MdJob* d_desc = (MdJob*)(wbuf + align256(8 * (2 * NCLASS * n + NCLASS)));
    auto* d_first = (long long*)((char*)d_desc + align256(sizeof(MdJob) * NCLASS * n * nop));
    zero_bytes(s, work, 8 * (2 * NCLASS * n + NCLASS));
    // every molecule's QD up first (a copy from pageable memory can wait for
    // the stream), then every molecule's set-up
    for (size_t i = 0; i < n; ++i)
        WB_HIP(hipMemcpyAsync(io + g[i].io, jobs[g[i].idx].qd, 8 * (size_t)g[i].B->nao * g[i].B->nao,
                              hipMemcpyHostToDevice, s));
    std::vector<std::unique_ptr<KChunk>> K(n);
    for (size_t i = 0; i < n; ++i) {
        K[i].reset(new KChunk);
        const Basis& B = *g[i].B;
        double* qd = (double*)(io + g[i].io);
        char* a = arena + g[i].arena;
        double* QDT = (double*)(a + g[i].P.qdt);
        transpose(s, B.nao, qd, QDT);
        k_chunk_setup(c, *g[i].B, *g[i].G, g[i].P, a, work + 2 * NCLASS * i, qd, QDT, 0, f32,
                      g64, tol, *K[i], [](int) {});
    }
    // the class launches: every molecule's shares of class k in job order
    // (each molecule's in the order its own build launches them)
    std::vector<MdJob> desc;
    std::vector<long long> first;
    struct Launch { int la, lb, nopv, njob; size_t d0, f0; long long total; };
    std::vector<Launch> launches;
    for (int k = 0; k < NCLASS; ++k) {
        Launch L{0, 0, 0, 0, desc.size(), first.size(), 0};
        for (size_t i = 0; i < n; ++i) {
            MdJob js[8];
            int nj = 0;
            const int nopv = k_class_jobs(c, *g[i].B, *K[i], R[i], k, nop, js, nj);
            if (!nj) continue;
            if (L.njob && L.nopv != nopv)
                throw Error(WB_DEVICE_ERROR, "queue launch: operator counts differ within a class");
            L.la = g[i].B->cls[k].la;
            L.lb = g[i].B->cls[k].lb;
            L.nopv = nopv;
            for (int t = 0; t < nj; ++t) {
                first.push_back(L.total);
                desc.push_back(js[t]);
                L.total += md_items(js[t]);
                ++L.njob;
            }
        }
        if (!L.njob) {
            first.resize(L.f0);
            continue;
        }
        first.push_back(L.total);
        launches.push_back(L);
    }
    if (desc.size() > NCLASS * n * (size_t)nop || first.size() > NCLASS * (n * nop + 1))
        throw Error(WB_DEVICE_ERROR, "queue launch: more work-item descriptors than planned");
    // the host vectors live until the group's closing synchronization
    if (!desc.empty()) {
        WB_HIP(hipMemcpyAsync(d_desc, desc.data(), sizeof(MdJob) * desc.size(),
                              hipMemcpyHostToDevice, s));
        WB_HIP(hipMemcpyAsync(d_first, first.data(), sizeof(long long) * first.size(),
                              hipMemcpyHostToDevice, s));
    }
    const void* boys = f32 ? c.d_b32.p : c.d_b64.p;
    for (size_t l = 0; l < launches.size(); ++l) {
        const Launch& L = launches[l];
        launch_md_queue(s, L.la, L.lb, f32, g64, L.nopv, d_desc + L.d0, d_first + L.f0, L.njob,
                        L.total, heads + l, boys);
    }
    // every molecule's K GEMMs, then every K down
    for (size_t i = 0; i < n; ++i) {
        const size_t nn = (size_t)g[i].B->nao * g[i].B->nao;
        k_chunk_finish(c, *g[i].B, g[i].P, arena + g[i].arena, *K[i], nop, g64,
                       (double*)(io + g[i].io) + nn, false, [](int) {});
    }
    for (size_t i = 0; i < n; ++i) {
        const size_t nn = (size_t)g[i].B->nao * g[i].B->nao;
        WB_HIP(hipMemcpyAsync(jobs[g[i].idx].k, (double*)(io + g[i].io) + nn, 8 * nn * nop,
                              hipMemcpyDeviceToHost, s));
    }
    WB_HIP(hipStreamSynchronize(s));
}

// the arena block of one job in a queue group: its one-chunk plan, then QD
// and K
static size_t queue_job_bytes(const KPlan& P, const Basis& B, int nop)
{
    return align256(P.total) + align256(8 * (size_t)B.nao * B.nao * (1 + nop));
}

int wb_cycle_queue(wb_ctx_t ctx, int njob, const wb_cycle_job* jobs, int nop,
                   const double* omegas, double k_tol, double b, double C, double vv_tol,
                   int precision, double* job_ms)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c) throw Error(WB_INVALID, "null context");
        if (njob < 0 || (njob > 0 && !jobs)) throw Error(WB_INVALID, "bad job list");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        bool anyk = false, anyv = false;
        for (int i = 0; i < njob; ++i) {
            const wb_cycle_job& j = jobs[i];
            if (!j.basis) throw Error(WB_INVALID, "job " + std::to_string(i) + ": null basis");
            if (j.kgrid && (!j.qd || !j.k))
                throw Error(WB_INVALID, "job " + std::to_string(i) + ": K build needs qd and k");
            if (j.nlcgrid && (!j.dm || !j.nlc || !j.vmat))
                throw Error(WB_INVALID, "job " + std::to_string(i) + ": VV10 needs dm, nlc, vmat");
            anyk |= j.kgrid != nullptr;
            anyv |= j.nlcgrid != nullptr;
        }
        if (anyk && (nop < 1 || nop > 8 || !omegas)) throw Error(WB_INVALID, "nop out of range");
        if (anyk) k_check_omegas(nop, omegas);
        WB_HIP(hipSetDevice(c->device));
        WB_HIP(hipStreamSynchronize(c->stream));     // work queued on the caller's context first
        const bool f32 = precision == WB_PREC_MIXED;
        const bool g64 = !f32 || c->g64;
        // largest K build first (nao^2 x points)
        std::vector<int> order(njob);
        std::vector<double> est(njob, 0.0);
        for (int i = 0; i < njob; ++i) {
            order[i] = i;
            if (jobs[i].kgrid)
                est[i] = (double)((const Basis*)jobs[i].basis)->nao *
                         ((const Basis*)jobs[i].basis)->nao * ((const Grid*)jobs[i].kgrid)->Ppad;
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return est[x] > est[y]; });
        const auto t0 = std::chrono::steady_clock::now();
        auto now_ms = [&] {
            return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count();
        };
        std::vector<double> k_ms(njob, 0.0), v_ms(njob, 0.0);
        // the VV10 terms: the side context, on a host thread of their own, one
        // job after another, its workspaces sized for the largest first (so
        // they are not reallocated while the K groups run)
        Ctx* sd = anyv ? &side_ctx(*c) : nullptr;
        if (sd) {
            size_t vneed[WS_COUNT] = {0};
            for (int i = 0; i < njob; ++i) {
                if (!jobs[i].nlcgrid) continue;
                wb_cycle_job v = jobs[i];
                v.kgrid = nullptr;
                size_t need[WS_COUNT], side[WS_COUNT];
                job_needs(*sd, v, nop, precision, need, side);
                for (int w = 0; w < WS_COUNT; ++w) vneed[w] = std::max(vneed[w], need[w]);
            }
            for (int w = 0; w < WS_COUNT; ++w) sd->bufs[w]->ensure(vneed[w]);
        }
        // K groups within the device memory free now plus the K arena's own
        // (less a reserve) and within the context's K budget, or 64 GiB
        // without one; a job too large for a group runs its own K build,
        // tiled to the device's budget.  The plan comes first, so that the
        // arena is sized once for every group of the call (and for the next
        // call with the same jobs).
        size_t free_b = 0, total_b = 0;
        WB_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t reserve = std::min<size_t>(total_b / 32, (size_t)8 << 30);
        const size_t held = c->ws_karena.bytes + c->ws_work.bytes;
        const size_t dev_budget = free_b + held > reserve ? free_b + held - reserve : 0;
        const size_t budget = std::min(dev_budget, c->k_budget ? c->k_budget : (size_t)64 << 30);
        const size_t alone_budget = c->k_budget ? std::min(c->k_budget, dev_budget) : dev_budget;
        struct Step { std::vector<QueueJob> grp; int alone = -1; size_t bytes = 0; };
        std::vector<Step> plan;
        size_t arena_max = 0, nmax = 0;
        for (int i : order) {
            if (!jobs[i].kgrid) continue;
            QueueJob q;
            q.idx = i;
            q.B = (Basis*)jobs[i].basis;
            q.G = (const Grid*)jobs[i].kgrid;
            q.P = k_plan(*q.B, *q.G, nop, f32, g64, c->ncopy, q.G->nsub);
            const size_t jb = queue_job_bytes(q.P, *q.B, nop);
            if (jb + queue_work_bytes(1, nop) > budget) {
                Step st;
                st.alone = i;
                st.bytes = k_plan_budget(*q.B, *q.G, nop, f32, g64, c->ncopy, alone_budget).total;
                arena_max = std::max(arena_max, st.bytes);
                plan.push_back(std::move(st));
                continue;
            }
            if (plan.empty() || plan.back().alone >= 0 ||
                plan.back().bytes + jb + queue_work_bytes(plan.back().grp.size() + 1, nop) > budget)
                plan.emplace_back();
            Step& st = plan.back();
            q.arena = st.bytes;
            q.io = st.bytes + align256(q.P.total);
            st.bytes += jb;
            st.grp.push_back(q);
            arena_max = std::max(arena_max, st.bytes);
            nmax = std::max(nmax, st.grp.size());
        }
        c->ws_karena.ensure(arena_max);
        if (nmax) c->ws_work.ensure(queue_work_bytes(nmax, nop));
        std::string verr;
        int vcode = WB_OK;
        std::thread vt;
        if (sd) {
            vt = std::thread([&] {
                try {
                    WB_HIP(hipSetDevice(sd->device));
                    for (int i : order) {
                        if (!jobs[i].nlcgrid) continue;
                        wb_cycle_job v = jobs[i];
                        v.kgrid = nullptr;
                        run_cycle_job(*sd, v, nop, omegas, k_tol, b, C, vv_tol, precision);
                        v_ms[i] = now_ms();
                    }
                } catch (const Error& e) {
                    vcode = e.code;
                    verr = e.what();
                } catch (const std::exception& e) {
                    vcode = WB_DEVICE_ERROR;
                    verr = e.what();
                }
            });
        }
        try {
            for (const Step& st : plan) {
                if (st.alone < 0) {
                    k_build_queue_group(*c, st.grp, jobs, nop, omegas, k_tol, precision);
                    const double t = now_ms();
                    for (const auto& q : st.grp) k_ms[q.idx] = t;
                    continue;
                }
                wb_cycle_job kj = jobs[st.alone];
                kj.nlcgrid = nullptr;
                const size_t kb = c->k_budget;
                c->k_budget = alone_budget;
                try {
                    run_cycle_job(*c, kj, nop, omegas, k_tol, b, C, vv_tol, precision);
                } catch (...) {
                    c->k_budget = kb;
                    throw;
                }
                c->k_budget = kb;
                k_ms[st.alone] = now_ms();
            }
        } catch (...) {
            if (vt.joinable()) vt.join();
            throw;
        }
        if (vt.joinable()) vt.join();
        if (vcode != WB_OK) throw Error(vcode, "cycle queue, VV10: " + verr);
        if (job_ms)
            for (int i = 0; i < njob; ++i) job_ms[i] = std::max(k_ms[i], v_ms[i]);
    })
}

// ------------------------------------------------------- cycle graphs ---
// One SCF cycle's device work recorded once as a HIP graph.  The graph owns
// its workspaces (a private context on the owner's stream): an eager call on
// the owner context may regrow the owner's buffers without invalidating the
// recorded addresses.  Nothing in the recorded path allocates, copies from
// the host or synchronizes; buffers are cleared by fill kernels.
struct CycleGraph {
    Ctx w;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~CycleGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

int wb_cycle_graph_create(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t kgrid, const double* d_qd,
                          int nop, const double* omegas, double k_tol, wb_grid_t nlcgrid,
                          const double* d_dm, double b, double C, double vv_tol, int precision,
                          double* d_k, double* d_out2, double* d_vmat, wb_graph_t* out)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!out || !basis) throw Error(WB_INVALID, "null argument");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        if (kgrid && (nop < 1 || nop > 8 || !omegas || !d_qd || !d_k))
            throw Error(WB_INVALID, "K build: nop out of range or null buffer");
        if (nlcgrid && (!d_dm || !d_out2 || !d_vmat)) throw Error(WB_INVALID, "VV10: null buffer");
        if (!kgrid && !nlcgrid) throw Error(WB_INVALID, "nothing to record");
        WB_HIP(hipSetDevice(c->device));
        auto g = std::make_unique<CycleGraph>();
        Ctx& w = g->w;
        w.device = c->device;
        w.stream = c->stream;
        w.ncopy = c->ncopy;
        w.fuse_mask = c->fuse_mask;
        w.gemm_tiles = c->gemm_tiles;
        w.g64 = c->g64;
        w.k_budget = c->k_budget;
        std::copy(c->cls_waves, c->cls_waves + NCLASS, w.cls_waves);
        std::vector<float> t32;
        std::vector<double> t64;
        boys_tables(t32, t64);
        upload(w.d_b32, t32.data(), sizeof(float) * t32.size(), w.stream);
        upload(w.d_b64, t64.data(), sizeof(double) * t64.size(), w.stream);
        std::vector<double> om(omegas, omegas + (kgrid ? nop : 0));
        Basis& B = *(Basis*)basis;
        auto body = [&]() {
            cycle_dev(w, B, (Grid*)kgrid, d_qd, nop, om.data(), k_tol, (Grid*)nlcgrid, d_dm, b, C,
                      vv_tol, precision, d_k, d_out2, d_vmat);
        };
        body();                 // sizes every workspace and fills the constant caches
        WB_HIP(hipStreamSynchronize(w.stream));
        WB_HIP(hipStreamBeginCapture(w.stream, hipStreamCaptureModeThreadLocal));
        try {
            body();
        } catch (...) {
            hipGraph_t part = nullptr;
            (void)hipStreamEndCapture(w.stream, &part);
            if (part) (void)hipGraphDestroy(part);
            throw;
        }
        WB_HIP(hipStreamEndCapture(w.stream, &g->graph));
        WB_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        *out = (wb_graph_t)g.release();
    })
}

int wb_cycle_graph_launch(wb_graph_t graph)
{
    WB_TRY({
        auto* g = (CycleGraph*)graph;
        if (!g) throw Error(WB_INVALID, "null graph");
        WB_HIP(hipSetDevice(g->w.device));
        WB_HIP(hipGraphLaunch(g->exec, g->w.stream));
    })
}

int wb_cycle_graph_destroy(wb_graph_t graph)
{
    WB_TRY({
        auto* g = (CycleGraph*)graph;
        if (!g) return WB_OK;
        WB_HIP(hipSetDevice(g->w.device));
        WB_HIP(hipStreamSynchronize(g->w.stream));
        delete g;
    })
}

int wb_xc_rho(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* dm, int mgga,
              double* rho5)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        auto* bs = (Basis*)basis;
        auto* G = (Grid*)grid;
        if (!c || !bs || !G || !dm || !rho5) throw Error(WB_INVALID, "null argument");
        WB_HIP(hipSetDevice(c->device));
        const size_t nn = (size_t)bs->nao * bs->nao;
        double* din = (double*)c->ws_io.ensure(sizeof(double) * (nn + 5 * (size_t)G->Ppad));
        double* rho = din + nn;
        WB_HIP(hipMemcpyAsync(din, dm, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
        xc_rho_dev(*c, *bs, *G, din, mgga != 0, rho);
        std::vector<double> h(5 * (size_t)G->Ppad);
        WB_HIP(hipMemcpyAsync(h.data(), rho, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                              c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
        const int n = G->npts;
        for (int k = 0; k < 5; ++k)
            for (int j = 0; j < n; ++j) rho5[(size_t)k * n + G->perm[j]] = h[(size_t)k * G->Ppad + j];
    })
}

int wb_xc_vmat(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* v5, int mgga,
               double* vmat)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        auto* bs = (Basis*)basis;
        auto* G = (Grid*)grid;
        if (!c || !bs || !G || !v5 || !vmat) throw Error(WB_INVALID, "null argument");
        WB_HIP(hipSetDevice(c->device));
        const size_t nn = (size_t)bs->nao * bs->nao;
        const int n = G->npts;
        std::vector<double> h(5 * (size_t)G->Ppad, 0.0);
        for (int k = 0; k < 5; ++k)
            for (int j = 0; j < n; ++j) h[(size_t)k * G->Ppad + j] = v5[(size_t)k * n + G->perm[j]];
        double* din = (double*)c->ws_io.ensure(sizeof(double) * (nn + h.size()));
        double* dv = din + nn;
        WB_HIP(hipMemcpyAsync(dv, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                              c->stream));
        xc_vmat_dev(*c, *bs, *G, dv, mgga != 0, din);
        WB_HIP(hipMemcpyAsync(vmat, din, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
    })
}

int wb_nlc_vxc(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* dm, double b,
               double C, double tile_tol, int precision, double* out2, double* vmat)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        auto* bs = (Basis*)basis;
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "unknown precision");
        WB_HIP(hipSetDevice(c->device));
        const size_t nn = (size_t)bs->nao * bs->nao;
        double* din = (double*)c->ws_io.ensure(sizeof(double) * (2 * nn + 2));
        WB_HIP(hipMemcpyAsync(din, dm, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
        nlc_dev(*c, *bs, *(Grid*)grid, din, b, C, tile_tol, precision, din + 2 * nn, din + nn);
        WB_HIP(hipMemcpyAsync(vmat, din + nn, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipMemcpyAsync(out2, din + 2 * nn, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
    })
}

int wb_vv10_sum(wb_ctx_t ctx, int n, const double* coords, const double* q, const double* W0,
                const double* kappa, double tile_tol, int precision, double* F, double* U,
                double* W)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (n < 0) throw Error(WB_INVALID, "negative size");
        if (n == 0) return WB_OK;
        WB_HIP(hipSetDevice(c->device));
        const bool f64 = precision == WB_PREC_FP64;
        // points in their given order, tiles of 64 consecutive points
        double o[3] = {0, 0, 0};
        for (int i = 0; i < n; ++i) for (int d = 0; d < 3; ++d) o[d] += coords[3 * i + d];
        for (int d = 0; d < 3; ++d) o[d] /= n;
        const int ntiles = (n + 63) / 64;
        std::vector<float4> ts(2 * (size_t)ntiles);
        std::vector<double> pt(4 * (size_t)n), pk(2 * (size_t)n);
        for (int i = 0; i < n; ++i) {
            for (int d = 0; d < 3; ++d) pt[4 * i + d] = coords[3 * i + d] - o[d];
            pt[4 * i + 3] = W0[i];
            pk[2 * i] = kappa[i];
            pk[2 * i + 1] = q[i];
        }
        for (int t = 0; t < ntiles; ++t) {
            const int b = t * 64, e = std::min(n, b + 64);
            double c3[3] = {0, 0, 0}, r = 0, wm = 1e300, km = 1e300, qs = 0;
            for (int i = b; i < e; ++i) for (int d = 0; d < 3; ++d) c3[d] += pt[4 * i + d] / (e - b);
            for (int i = b; i < e; ++i) {
                double dd = 0;
                for (int d = 0; d < 3; ++d) dd += (pt[4 * i + d] - c3[d]) * (pt[4 * i + d] - c3[d]);
                r = std::max(r, std::sqrt(dd));
                wm = std::min(wm, W0[i]); km = std::min(km, kappa[i]); qs += q[i];
            }
            ts[2 * t] = make_float4((float)c3[0], (float)c3[1], (float)c3[2], (float)(r * 1.0001 + 1e-5));
            ts[2 * t + 1] = make_float4((float)(wm * 0.9999), (float)(km * 0.9999), (float)(qs * 1.0001), 0.f);
        }
        DevBuf dp, dk, dt, dout;
        if (f64) {
            upload(dp, pt.data(), sizeof(double) * pt.size(), c->stream);
            upload(dk, pk.data(), sizeof(double) * pk.size(), c->stream);
        } else {
            std::vector<float> p32(pt.begin(), pt.end()), k32(pk.begin(), pk.end());
            upload(dp, p32.data(), sizeof(float) * p32.size(), c->stream);
            upload(dk, k32.data(), sizeof(float) * k32.size(), c->stream);
            WB_HIP(hipStreamSynchronize(c->stream));
        }
        upload(dt, ts.data(), sizeof(float4) * ts.size(), c->stream);
        double* out = (double*)dout.ensure(sizeof(double) * 3 * (size_t)n);
        DevBuf dn, dws;
        upload(dn, &n, sizeof(int), c->stream);
        const size_t symb = vv10_sym_bytes(n, f64, nullptr);
        dws.ensure(symb + 2 * sizeof(unsigned long long));
        auto* stat = (unsigned long long*)(dws.as<char>() + symb);
        launch_vv10_sym(c->stream, dn.as<int>(), n, dp.p, dk.p, (const float4*)dt.p,
                        (float)tile_tol, f64, dws.p, out, out + n, out + 2 * (size_t)n, stat);
        WB_HIP(hipMemcpyAsync(c->vv_stat_host, stat, 2 * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, c->stream));
        c->vv_stat = nullptr;
        WB_HIP(hipMemcpyAsync(F, out, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipMemcpyAsync(U, out + n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipMemcpyAsync(W, out + 2 * (size_t)n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
    })
}

int wb_vv10_stats(wb_ctx_t ctx, long long* out2)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c || !out2) throw Error(WB_INVALID, "null argument");
        unsigned long long v[2] = {c->vv_stat_host[0], c->vv_stat_host[1]};
        if (c->vv_stat) {
            WB_HIP(hipSetDevice(c->device));
            WB_HIP(hipStreamSynchronize(c->stream));
            WB_HIP(hipMemcpy(v, c->vv_stat, sizeof(v), hipMemcpyDeviceToHost));
        }
        out2[0] = (long long)v[0];
        out2[1] = (long long)v[1];
    })
}

int wb_ao_eval(wb_ctx_t ctx, wb_basis_t basis, int npts, const double* coords, int deriv,
               double* out)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        auto* b = (Basis*)basis;
        if (npts < 1) throw Error(WB_INVALID, "no points");
        WB_HIP(hipSetDevice(c->device));
        const int Ppad = (npts + 63) / 64 * 64;
        std::vector<double> hx(3 * (size_t)Ppad, 0.0);
        for (int i = 0; i < Ppad; ++i)
            for (int d = 0; d < 3; ++d) hx[(size_t)d * Ppad + i] = coords[3 * (i < npts ? i : 0) + d];
        DevBuf dx, dX;
        upload(dx, hx.data(), sizeof(double) * hx.size(), c->stream);
        const int nc = deriv ? 4 : 1;
        const size_t plane = (size_t)b->nao * Ppad;
        double* X = (double*)dX.ensure(sizeof(double) * nc * plane);
        launch_ao_eval_pts(c->stream, *b, dx.as<double>(), Ppad, deriv, X);
        std::vector<double> h(nc * plane);
        WB_HIP(hipMemcpyAsync(h.data(), X, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
        for (int k = 0; k < nc; ++k)
            for (int m = 0; m < b->nao; ++m)
                for (int g = 0; g < npts; ++g)
                    out[((size_t)k * npts + g) * b->nao + m] = h[k * plane + (size_t)m * Ppad + g];
    })
}

// S_g = X^T diag(w) X on the context (nao x nao, device, into ws_K)
static double* grid_overlap_dev(Ctx& c, Basis& b, Grid& g)
{
    const int nao = b.nao, Ppad = g.Ppad;
    double* X = (double*)c.ws_X.ensure(sizeof(double) * nao * (size_t)Ppad);
    launch_ao_eval(c.stream, b, g, 0, X);
    // padding points carry weight 0
    double* S = (double*)c.ws_K.ensure(sizeof(double) * 3 * nao * (size_t)nao);
    gemm_nt_fast(c.stream, nao, nao, Ppad, X, X, false, 1, 0, g.d_w.as<double>(), S, c.ws_Kp);
    return S;
}

int wb_grid_overlap(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, double* out)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        auto* b = (Basis*)basis;
        auto* g = (Grid*)grid;
        if (!c || !b || !g || !out) throw Error(WB_INVALID, "null argument");
        WB_HIP(hipSetDevice(c->device));
        const int nao = b->nao;
        double* S = grid_overlap_dev(*c, *b, *g);
        WB_HIP(hipMemcpyAsync(out, S, sizeof(double) * nao * (size_t)nao, hipMemcpyDeviceToHost,
                              c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
    })
}

int wb_pjs_projector(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* S_an,
                     double* Q)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        auto* b = (Basis*)basis;
        auto* g = (Grid*)grid;
        if (!c || !b || !g || !S_an || !Q) throw Error(WB_INVALID, "null argument");
        WB_HIP(hipSetDevice(c->device));
        const int nao = b->nao;
        const size_t nn = (size_t)nao * nao;
        double* Sg = grid_overlap_dev(*c, *b, *g);        // ws_K: [Sg | rhs | out]
        double* rhs = Sg + nn;
        double* out = Sg + 2 * nn;
        WB_HIP(hipMemcpyAsync(rhs, S_an, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
        // LU with partial pivoting (PySCF's scipy.linalg.solve(S_g, S)); both
        // matrices symmetric, so row- and column-major storage agree on input;
        // the column-major solution X, read row-major, is X^T
        int* ipiv = (int*)c->ws_io.ensure(sizeof(int) * (nao + 1));
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success)
            throw Error(WB_DEVICE_ERROR, "rocblas_create_handle failed");
        struct Guard { rocblas_handle h; ~Guard() { rocblas_destroy_handle(h); } } guard{h};
        rocblas_set_stream(h, c->stream);
        if (rocsolver_dgetrf(h, nao, nao, Sg, nao, ipiv, ipiv + nao) != rocblas_status_success ||
            rocsolver_dgetrs(h, rocblas_operation_none, nao, nao, Sg, nao, ipiv, rhs, nao) !=
                rocblas_status_success)
            throw Error(WB_DEVICE_ERROR, "rocSOLVER LU solve failed");
        int info = 0;
        WB_HIP(hipMemcpyAsync(&info, ipiv + nao, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        transpose(c->stream, nao, rhs, out);
        WB_HIP(hipMemcpyAsync(Q, out, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
        if (info != 0)
            throw Error(WB_UNSUPPORTED, "grid overlap matrix is singular (LU pivot " +
                                            std::to_string(info) + "): the grid cannot fit the overlap");
    })
}

int wb_ctx_set_profile(wb_ctx_t ctx, int on)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        c->profile = on != 0;
        for (double& v : c->phase_ms) v = 0.0;
    })
}

int wb_profile_read(wb_ctx_t ctx, double* ms, int n, int reset)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        for (int i = 0; i < n && i < PH_COUNT; ++i) ms[i] = c->phase_ms[i];
        if (reset) for (double& v : c->phase_ms) v = 0.0;
    })
}

int wb_k_work(wb_ctx_t ctx, long long* out)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        unsigned long long h[2 * NCLASS] = {0};
        if (c->ws_work.p) {
            WB_HIP(hipMemcpyAsync(h, c->ws_work.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
            WB_HIP(hipStreamSynchronize(c->stream));
        }
        for (int i = 0; i < 2 * NCLASS; ++i) out[i] = (long long)h[i];
    })
}

int wb_test_gemm(wb_ctx_t ctx, int kind, int M, int N, int K, const double* A, const double* B,
                 const double* w, double* C)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        WB_HIP(hipSetDevice(c->device));
        DevBuf dA, dB, dw, dC, ws;
        const size_t nb = (size_t)K * N;
        upload(dA, A, sizeof(double) * (size_t)M * K, c->stream);
        upload(dB, B, sizeof(double) * nb, c->stream);
        if (w) upload(dw, w, sizeof(double) * K, c->stream);
        dC.ensure(sizeof(double) * (size_t)M * N);
        if (kind == 0)
            gemm_nn_fast(c->stream, M, N, K, dA.as<double>(), dB.as<double>(), dC.p, false);
        else if (kind == 1)
            gemm_nt_fast(c->stream, M, N, K, dA.as<double>(), dB.p, false, 1, 0,
                         w ? dw.as<double>() : nullptr, dC.as<double>(), ws);
        else if (kind == 2)
            gemm_nn(c->stream, M, N, K, dA.as<double>(), dB.as<double>(), dC.p, false);
        else
            gemm_nt_reduce(c->stream, M, N, K, dA.as<double>(), dB.p, false, 1, 0,
                           w ? dw.as<double>() : nullptr, dC.as<double>(), ws);
        WB_HIP(hipMemcpyAsync(C, dC.p, sizeof(double) * (size_t)M * N, hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
    })
}

int wb_grid_partition(wb_ctx_t ctx, int natm, const double* atoms, int npts,
                      const double* coords, const int* owner, double* out)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (natm < 1 || npts < 0 || !atoms || (npts > 0 && (!coords || !owner || !out)))
            throw Error(WB_INVALID, "bad grid partition arguments");
        for (int i = 0; i < npts; ++i)
            if (owner[i] < 0 || owner[i] >= natm)
                throw Error(WB_INVALID, "grid point owner outside 0..natm-1");
        // atom-pair distances and their inverses (0 on the diagonal)
        std::vector<double> rab((size_t)natm * natm), irab((size_t)natm * natm);
        for (int a = 0; a < natm; ++a)
            for (int b = 0; b < natm; ++b) {
                const double dx = atoms[3 * a] - atoms[3 * b], dy = atoms[3 * a + 1] - atoms[3 * b + 1],
                             dz = atoms[3 * a + 2] - atoms[3 * b + 2];
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (a != b && r == 0.0) throw Error(WB_INVALID, "two atoms at one position");
                rab[(size_t)a * natm + b] = r;
                irab[(size_t)a * natm + b] = a == b ? 0.0 : 1.0 / r;
            }
        WB_HIP(hipSetDevice(c->device));
        DevBuf dA, dR, dI, dX, dO, dW;
        upload(dA, atoms, sizeof(double) * 3 * natm, c->stream);
        upload(dR, rab.data(), sizeof(double) * rab.size(), c->stream);
        upload(dI, irab.data(), sizeof(double) * irab.size(), c->stream);
        upload(dX, coords, sizeof(double) * 3 * (size_t)npts, c->stream);
        upload(dO, owner, sizeof(int) * (size_t)npts, c->stream);
        dW.ensure(sizeof(double) * (size_t)std::max(npts, 1));
        launch_ssf_partition(c->stream, npts, dX.as<double>(), dO.as<int>(), natm,
                             dA.as<double>(), dR.as<double>(), dI.as<double>(), dW.as<double>());
        if (npts)
            WB_HIP(hipMemcpyAsync(out, dW.p, sizeof(double) * (size_t)npts, hipMemcpyDeviceToHost,
                                  c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
    })
}

int wb_grid_build(wb_ctx_t ctx, int natm, const double* atoms, int nshell, const int* shell_atom,
                  const double* shell_r, const double* shell_w, const int* shell_rule, int nrule,
                  const int* rule_npts, const double* rule_xyzw, double w_cut, long long capacity,
                  double* coords, double* weights, long long* npts)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (!c || natm < 1 || !atoms || nshell < 0 || nrule < 1 || !rule_npts || !rule_xyzw ||
            !npts || (nshell > 0 && (!shell_atom || !shell_r || !shell_w || !shell_rule)))
            throw Error(WB_INVALID, "bad grid build arguments");
        std::vector<int> rule_off(nrule + 1, 0);
        for (int k = 0; k < nrule; ++k) {
            if (rule_npts[k] < 1) throw Error(WB_INVALID, "empty angular rule");
            rule_off[k + 1] = rule_off[k] + rule_npts[k];
        }
        std::vector<long long> first(std::max(nshell, 1) + 1, 0);
        for (int sh = 0; sh < nshell; ++sh) {
            if (shell_atom[sh] < 0 || shell_atom[sh] >= natm)
                throw Error(WB_INVALID, "radial shell atom outside 0..natm-1");
            if (shell_rule[sh] < 0 || shell_rule[sh] >= nrule)
                throw Error(WB_INVALID, "radial shell rule outside 0..nrule-1");
            first[sh + 1] = first[sh] + rule_npts[shell_rule[sh]];
        }
        const long long nraw = nshell > 0 ? first[nshell] : 0;
        if (nraw > INT32_MAX) throw Error(WB_UNSUPPORTED, "more than 2^31 raw grid points");
        std::vector<double> rab((size_t)natm * natm), irab((size_t)natm * natm);
        for (int a = 0; a < natm; ++a)
            for (int b = 0; b < natm; ++b) {
                const double dx = atoms[3 * a] - atoms[3 * b], dy = atoms[3 * a + 1] - atoms[3 * b + 1],
                             dz = atoms[3 * a + 2] - atoms[3 * b + 2];
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (a != b && r == 0.0) throw Error(WB_INVALID, "two atoms at one position");
                rab[(size_t)a * natm + b] = r;
                irab[(size_t)a * natm + b] = a == b ? 0.0 : 1.0 / r;
            }
        WB_HIP(hipSetDevice(c->device));
        hipStream_t s = c->stream;
        DevBuf dA, dR, dI, dF, dSA, dSR, dSW, dSK, dRO, dRX, dX, dO, dWp, dP, dB, dXo, dWo;
        upload(dA, atoms, sizeof(double) * 3 * natm, s);
        upload(dR, rab.data(), sizeof(double) * rab.size(), s);
        upload(dI, irab.data(), sizeof(double) * irab.size(), s);
        upload(dRO, rule_off.data(), sizeof(int) * rule_off.size(), s);
        upload(dRX, rule_xyzw, sizeof(double) * 4 * (size_t)rule_off[nrule], s);
        const size_t n = (size_t)std::max<long long>(nraw, 1);
        if (nshell > 0) {
            upload(dF, first.data(), sizeof(long long) * (nshell + 1), s);
            upload(dSA, shell_atom, sizeof(int) * nshell, s);
            upload(dSR, shell_r, sizeof(double) * nshell, s);
            upload(dSW, shell_w, sizeof(double) * nshell, s);
            upload(dSK, shell_rule, sizeof(int) * nshell, s);
        }
        dX.ensure(sizeof(double) * 3 * n);
        dO.ensure(sizeof(int) * n);
        dWp.ensure(sizeof(double) * n);
        dP.ensure(sizeof(double) * n);
        dB.ensure(sizeof(int) * ((size_t)grid_keep_blocks(nraw) + 1));
        dXo.ensure(sizeof(double) * 3 * n);
        dWo.ensure(sizeof(double) * n);
        launch_grid_points(s, nraw, nshell, dF.as<long long>(), dSA.as<int>(), dSR.as<double>(),
                           dSW.as<double>(), dSK.as<int>(), dRO.as<int>(), dRX.as<double>(),
                           dA.as<double>(), dX.as<double>(), dO.as<int>(), dWp.as<double>());
        launch_ssf_partition(s, (int)nraw, dX.as<double>(), dO.as<int>(), natm, dA.as<double>(),
                             dR.as<double>(), dI.as<double>(), dP.as<double>());
        launch_grid_keep(s, nraw, dX.as<double>(), dWp.as<double>(), dP.as<double>(), w_cut,
                         dB.as<int>(), dXo.as<double>(), dWo.as<double>());
        int kept = 0;
        WB_HIP(hipMemcpyAsync(&kept, dB.as<int>() + grid_keep_blocks(nraw), sizeof(int),
                              hipMemcpyDeviceToHost, s));
        WB_HIP(hipStreamSynchronize(s));
        *npts = kept;
        if (kept > capacity)
            throw Error(WB_INVALID, "grid build: " + std::to_string(kept) +
                                        " points kept, more than the capacity " +
                                        std::to_string(capacity));
        if (kept > 0) {
            if (!coords || !weights) throw Error(WB_INVALID, "null output");
            WB_HIP(hipMemcpyAsync(coords, dXo.p, sizeof(double) * 3 * (size_t)kept,
                                  hipMemcpyDeviceToHost, s));
            WB_HIP(hipMemcpyAsync(weights, dWo.p, sizeof(double) * (size_t)kept,
                                  hipMemcpyDeviceToHost, s));
            WB_HIP(hipStreamSynchronize(s));
        }
    })
}

int wb_test_pair_bounds(int nshell, const int* l, const int* nprim, const double* center,
                        const double* exps, const double* coefs, double pair_cut,
                        long long* counts2, double* pairs, double* prims)
{
    WB_TRY({
        if (nshell < 1 || !l || !nprim || !center || !exps || !coefs || !counts2)
            throw Error(WB_INVALID, "null argument");
        auto B = std::make_unique<Basis>();
        build_basis(*B, nshell, l, nprim, center, exps, coefs, pair_cut);
        long long np = 0, nm = 0;
        for (const auto& C : B->cls) {
            np += C.npairs;
            nm += C.nprims;
        }
        counts2[0] = np;
        counts2[1] = nm;
        if (pairs && prims) {
            size_t ip = 0;
            long long mbase = 0;
            for (const auto& C : B->cls) {
                for (int i = 0; i < C.npairs; ++i) {
                    const PairRec& pr = C.pairs[i];
                    const PairScreen& ps = C.scr[i];
                    double* o = pairs + 10 * ip++;
                    o[0] = pr.shA; o[1] = pr.shB;
                    o[2] = ps.cx; o[3] = ps.cy; o[4] = ps.cz;
                    o[5] = ps.delta; o[6] = ps.beta; o[7] = ps.q;
                    o[8] = (double)(mbase + pr.pp0); o[9] = pr.npp;
                }
                for (int m = 0; m < C.nprims; ++m) {
                    const float4 q = C.pbc[m];
                    double* o = prims + 6 * (mbase + m);
                    o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w;
                    o[4] = C.pij[2 * (size_t)m]; o[5] = C.pij[2 * (size_t)m + 1];
                }
                mbase += C.nprims;
            }
        }
    })
}

int wb_boys_eval(wb_ctx_t ctx, int n, const double* T, int nmax, int precision, double* out)
{
    WB_TRY({
        auto* c = (Ctx*)ctx;
        if (nmax < 0 || nmax > B32_LMAX) throw Error(WB_INVALID, "nmax must be 0..6");
        if (n < 0 || (n > 0 && (!T || !out))) throw Error(WB_INVALID, "bad Boys arguments");
        if (precision != WB_PREC_MIXED && precision != WB_PREC_FP64)
            throw Error(WB_INVALID, "precision must be WB_PREC_MIXED or WB_PREC_FP64");
        WB_HIP(hipSetDevice(c->device));
        DevBuf dT, dF;
        upload(dT, T, sizeof(double) * n, c->stream);
        dF.ensure(sizeof(double) * (size_t)(nmax + 1) * n);
        launch_boys_eval(c->stream, nmax, n, dT.as<double>(), precision == WB_PREC_MIXED,
                         c->d_b32.p, c->d_b64.p, dF.as<double>());
        if (n)
            WB_HIP(hipMemcpyAsync(out, dF.p, sizeof(double) * (size_t)(nmax + 1) * n,
                                  hipMemcpyDeviceToHost, c->stream));
        WB_HIP(hipStreamSynchronize(c->stream));
    })
}

int wb_sync(wb_ctx_t ctx)
{
    WB_TRY({ WB_HIP(hipStreamSynchronize(((Ctx*)ctx)->stream)); })
}

float wb_last_ms(wb_ctx_t ctx) { return ((Ctx*)ctx)->ms_last; }

}  // extern "C"
