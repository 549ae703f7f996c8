/*
 * FP64 CPU oracle of the omega-B97M-V / COSX hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the tests, the smoke
 * check and the CPU-baseline leg of the benchmark link it; the product path
 * never does.
 *
 * It restates, in plain C (FP64, OpenMP over grid points), the arithmetic of
 * the reference's GPU path:
 *   - Boys F_n(T): series + downward recursion below T = 30, erf-based F_0 and
 *     upward recursion above (gpu/boys.py:14-47 states the function; the
 *     reference tabulates it from the same closed form, gpu/boys.py:66-99).
 *   - McMurchie-Davidson Hermite E coefficients (gpu/shellpairs.py:76-113),
 *     the Hermite Coulomb integrals R^n_tuv with the 1/r or erf(w r)/r kernel
 *     (theta = w^2/(w^2+p), R^n_000 = sqrt(theta)(-2 theta p)^n F_n(theta p
 *     |PC|^2); gpu/shellpairs.py:447-485 and gpu/sgx_k.py:783-811), and the
 *     3-centre ESP block A^g_{mu nu} = (mu | v(|r - g|) | nu)
 *     (gpu/shellpairs.py:499-542).
 *   - AO values X_{g mu} (and gradients) of spherical contracted GTOs
 *     (gpu/ao_eval.py:255-329).
 *   - The COSX exchange assembly K = X^T (w o G), G_g = A^g F_g, F = X (Q D)
 *     (gpu/sgx_k.py:663-751), without post-processing.
 *   - The VV10 double sum F/U/W (gpu/vv10.py:49-98; libdft VXC_vv10nlc),
 *     F scaled by -1.5 as returned there (gpu/vv10.py:239).
 *
 * Basis data arrive in a flat "shell table" (see wb_oracle.h) filled by the
 * Python side from the molecule's libcint-style arrays; cart->sph matrices
 * are passed in as data.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "wb_oracle.h"

#define LMAX 3
#define LTOT (2 * LMAX)
#define NC(l) (((l) + 1) * ((l) + 2) / 2)
#define NCMAX 10

/* ---------------------------------------------------------------- Boys -- */
void orc_boys(int nmax, double T, double *F)
{
    if (T < 1e-15) {
        for (int n = 0; n <= nmax; ++n) F[n] = 1.0 / (2 * n + 1) - T / (2 * n + 3);
        return;
    }
    const double e = exp(-T);
    if (T < 30.0) {
        /* F_m(T) = e^{-T} sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)) */
        double term = 1.0 / (2 * nmax + 1), sum = term;
        for (int k = 1; k < 400; ++k) {
            term *= 2.0 * T / (2 * nmax + 2 * k + 1);
            sum += term;
            if (term < sum * 1e-17) break;
        }
        F[nmax] = e * sum;
        for (int n = nmax; n > 0; --n) F[n - 1] = (2.0 * T * F[n] + e) / (2 * n - 1);
    } else {
        const double st = sqrt(T);
        F[0] = 0.5 * sqrt(M_PI) * erf(st) / st;
        for (int n = 0; n < nmax; ++n) F[n + 1] = ((2 * n + 1) * F[n] - e) / (2.0 * T);
    }
}

void orc_boys_batch(int nmax, int n, const double *T, double *F)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) orc_boys(nmax, T[i], F + (size_t)i * (nmax + 1));
}

/* ------------------------------------------------- Hermite E and R ------ */
/* E[i][j][t], i <= la, j <= lb, t <= i+j for one axis. */
static void hermite_e(int la, int lb, double p, double xpa, double xpb,
                      double E[LMAX + 1][LMAX + 1][LTOT + 1])
{
    memset(E, 0, sizeof(double) * (LMAX + 1) * (LMAX + 1) * (LTOT + 1));
    const double h = 0.5 / p;
    E[0][0][0] = 1.0;
    for (int i = 0; i <= la; ++i) {
        for (int j = 0; j <= lb; ++j) {
            if (i == 0 && j == 0) continue;
            double (*src)[LTOT + 1];
            double x;
            int ii = i, jj = j;
            if (j == 0) { ii = i - 1; x = xpa; }
            else { jj = j - 1; x = xpb; }
            src = &E[ii][jj];
            const int tmax = i + j, tprev = ii + jj;
            for (int t = 0; t <= tmax; ++t) {
                double v = 0.0;
                if (t > 0) v += h * (*src)[t - 1];
                if (t <= tprev) v += x * (*src)[t];
                if (t + 1 <= tprev) v += (t + 1) * (*src)[t + 1];
                E[i][j][t] = v;
            }
        }
    }
}

/* R[t][u][v] (n = 0) for t+u+v <= L; R000n = sqrt(th)(-2 th p)^n F_n. */
static void hermite_r(int L, const double *R000n, double x, double y, double z,
                      double R[LTOT + 1][LTOT + 1][LTOT + 1])
{
    /* work[n][t][u][v] built from level n+1 (Helgaker eq. 9.9.18-20) */
    static __thread double W[LTOT + 1][LTOT + 1][LTOT + 1][LTOT + 1];
    for (int n = L; n >= 0; --n) {
        const int m = L - n;   /* max t+u+v at this level */
        for (int t = 0; t <= m; ++t)
            for (int u = 0; u + t <= m; ++u)
                for (int v = 0; v + u + t <= m; ++v) {
                    double val;
                    if (t == 0 && u == 0 && v == 0) val = R000n[n];
                    else if (t > 0) {
                        val = x * W[n + 1][t - 1][u][v];
                        if (t > 1) val += (t - 1) * W[n + 1][t - 2][u][v];
                    } else if (u > 0) {
                        val = y * W[n + 1][t][u - 1][v];
                        if (u > 1) val += (u - 1) * W[n + 1][t][u - 2][v];
                    } else {
                        val = z * W[n + 1][t][u][v - 1];
                        if (v > 1) val += (v - 1) * W[n + 1][t][u][v - 2];
                    }
                    W[n][t][u][v] = val;
                }
    }
    for (int t = 0; t <= L; ++t)
        for (int u = 0; u + t <= L; ++u)
            for (int v = 0; v + u + t <= L; ++v) R[t][u][v] = W[0][t][u][v];
}

static void cart_comps(int l, int comps[][3])
{
    int c = 0;
    for (int lx = l; lx >= 0; --lx)
        for (int ly = l - lx; ly >= 0; --ly) {
            comps[c][0] = lx; comps[c][1] = ly; comps[c][2] = l - lx - ly;
            ++c;
        }
}

/* Contracted spherical ESP block of shells (sa, sb) at point C:
 * out[mu*nb + nu], mu in sa (2la+1), nu in sb (2lb+1). */
static void esp_block(const orc_basis *bs, int sa, int sb, const double *C,
                      double omega, double *out)
{
    const int la = bs->l[sa], lb = bs->l[sb];
    const int nca = NC(la), ncb = NC(lb), nsa = 2 * la + 1, nsb = 2 * lb + 1;
    const double *A = bs->center + 3 * sa, *B = bs->center + 3 * sb;
    const double ab2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1])
                     + (A[2] - B[2]) * (A[2] - B[2]);
    int ca[NCMAX][3], cb[NCMAX][3];
    cart_comps(la, ca);
    cart_comps(lb, cb);
    double V[NCMAX][NCMAX];
    memset(V, 0, sizeof(V));
    const int L = la + lb;
    double E[3][LMAX + 1][LMAX + 1][LTOT + 1];
    double R[LTOT + 1][LTOT + 1][LTOT + 1];
    double Fb[LTOT + 1], R0[LTOT + 1];
    for (int i = 0; i < bs->nprim[sa]; ++i) {
        const double a = bs->exps[bs->prim0[sa] + i], cai = bs->coefs[bs->prim0[sa] + i];
        for (int j = 0; j < bs->nprim[sb]; ++j) {
            const double b = bs->exps[bs->prim0[sb] + j], cbj = bs->coefs[bs->prim0[sb] + j];
            const double p = a + b, mu = a * b / p;
            const double Kab = exp(-mu * ab2);
            const double pref = cai * cbj * Kab * 2.0 * M_PI / p;
            double P[3];
            for (int d = 0; d < 3; ++d) P[d] = (a * A[d] + b * B[d]) / p;
            for (int d = 0; d < 3; ++d)
                hermite_e(la, lb, p, P[d] - A[d], P[d] - B[d], E[d]);
            double th = 1.0;
            if (omega > 0.0) th = omega * omega / (omega * omega + p);
            const double x = P[0] - C[0], y = P[1] - C[1], z = P[2] - C[2];
            const double T = th * p * (x * x + y * y + z * z);
            orc_boys(L, T, Fb);
            double q = sqrt(th);
            for (int n = 0; n <= L; ++n) { R0[n] = q * Fb[n]; q *= -2.0 * th * p; }
            hermite_r(L, R0, x, y, z, R);
            for (int u = 0; u < nca; ++u)
                for (int v = 0; v < ncb; ++v) {
                    const int ax = ca[u][0], ay = ca[u][1], az = ca[u][2];
                    const int bx = cb[v][0], by = cb[v][1], bz = cb[v][2];
                    double s = 0.0;
                    for (int t = 0; t <= ax + bx; ++t)
                        for (int uu = 0; uu <= ay + by; ++uu)
                            for (int vv = 0; vv <= az + bz; ++vv)
                                s += E[0][ax][bx][t] * E[1][ay][by][uu] * E[2][az][bz][vv]
                                   * R[t][uu][vv];
                    V[u][v] += pref * s;
                }
        }
    }
    const double *Ta = bs->c2s[la], *Tb = bs->c2s[lb];   /* (2l+1, ncart) */
    for (int m = 0; m < nsa; ++m)
        for (int n = 0; n < nsb; ++n) {
            double s = 0.0;
            for (int u = 0; u < nca; ++u) {
                const double tu = Ta[m * nca + u];
                if (tu == 0.0) continue;
                for (int v = 0; v < ncb; ++v) s += tu * Tb[n * ncb + v] * V[u][v];
            }
            out[m * nsb + n] = s;
        }
}

void orc_esp_dense(const orc_basis *bs, int npts, const double *coords,
                   double omega, double *out)
{
    const int nao = bs->nao, ns = bs->nshell;
#pragma omp parallel for schedule(dynamic)
    for (int g = 0; g < npts; ++g) {
        double blk[49];
        double *Ag = out + (size_t)g * nao * nao;
        for (int sa = 0; sa < ns; ++sa)
            for (int sb = 0; sb <= sa; ++sb) {
                esp_block(bs, sa, sb, coords + 3 * g, omega, blk);
                const int nsa = 2 * bs->l[sa] + 1, nsb = 2 * bs->l[sb] + 1;
                const int a0 = bs->ao0[sa], b0 = bs->ao0[sb];
                for (int m = 0; m < nsa; ++m)
                    for (int n = 0; n < nsb; ++n) {
                        Ag[(size_t)(a0 + m) * nao + b0 + n] = blk[m * nsb + n];
                        Ag[(size_t)(b0 + n) * nao + a0 + m] = blk[m * nsb + n];
                    }
            }
    }
}

/* ---------------------------------------------------------- AO values -- */
static void ao_point(const orc_basis *bs, const double *r, int deriv, double *X,
                     size_t stride)
{
    int comps[NCMAX][3];
    for (int s = 0; s < bs->nshell; ++s) {
        const int l = bs->l[s], nc = NC(l), nsph = 2 * l + 1;
        const double *A = bs->center + 3 * s;
        const double d[3] = {r[0] - A[0], r[1] - A[1], r[2] - A[2]};
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double rad = 0.0, drad = 0.0;   /* sum c e, sum -2 a c e */
        for (int i = 0; i < bs->nprim[s]; ++i) {
            const double a = bs->exps[bs->prim0[s] + i];
            const double ce = bs->coefs[bs->prim0[s] + i] * exp(-a * r2);
            rad += ce;
            drad += -2.0 * a * ce;
        }
        cart_comps(l, comps);
        double cv[NCMAX], cg[3][NCMAX];
        for (int c = 0; c < nc; ++c) {
            double pw[3][LMAX + 2];
            for (int ax = 0; ax < 3; ++ax) {
                pw[ax][0] = 1.0;
                for (int k = 1; k <= LMAX + 1; ++k) pw[ax][k] = pw[ax][k - 1] * d[ax];
            }
            const int *e = comps[c];
            const double poly = pw[0][e[0]] * pw[1][e[1]] * pw[2][e[2]];
            cv[c] = poly * rad;
            if (deriv) {
                for (int ax = 0; ax < 3; ++ax) {
                    const int b1 = (ax + 1) % 3, b2 = (ax + 2) % 3;
                    const double other = pw[b1][e[b1]] * pw[b2][e[b2]];
                    double dpoly = e[ax] > 0 ? e[ax] * pw[ax][e[ax] - 1] : 0.0;
                    cg[ax][c] = other * (dpoly * rad + pw[ax][e[ax] + 1] * drad);
                }
            }
        }
        const double *T = bs->c2s[l];
        for (int m = 0; m < nsph; ++m) {
            double v = 0.0, g[3] = {0, 0, 0};
            for (int c = 0; c < nc; ++c) {
                v += T[m * nc + c] * cv[c];
                if (deriv)
                    for (int ax = 0; ax < 3; ++ax) g[ax] += T[m * nc + c] * cg[ax][c];
            }
            X[bs->ao0[s] + m] = v;
            if (deriv)
                for (int ax = 0; ax < 3; ++ax) X[(ax + 1) * stride + bs->ao0[s] + m] = g[ax];
        }
    }
}

void orc_ao_eval(const orc_basis *bs, int npts, const double *coords, int deriv,
                 double *out)
{
    /* out: (1 + 3 deriv, npts, nao) */
    const size_t plane = (size_t)npts * bs->nao;
#pragma omp parallel for schedule(static)
    for (int g = 0; g < npts; ++g)
        ao_point(bs, coords + 3 * g, deriv, out + (size_t)g * bs->nao, plane);
}

/* ------------------------------------------------------- COSX K -------- */
static double block_bound(const orc_basis *bs, int sa, int sb)
{
    /* max over prim pairs of |c c| K_AB: the reference's pair cutoff measure
       (gpu/shellpairs.py:290-294) */
    const double *A = bs->center + 3 * sa, *B = bs->center + 3 * sb;
    const double ab2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1])
                     + (A[2] - B[2]) * (A[2] - B[2]);
    double m = 0.0;
    for (int i = 0; i < bs->nprim[sa]; ++i)
        for (int j = 0; j < bs->nprim[sb]; ++j) {
            const double a = bs->exps[bs->prim0[sa] + i], b = bs->exps[bs->prim0[sb] + j];
            const double v = fabs(bs->coefs[bs->prim0[sa] + i] * bs->coefs[bs->prim0[sb] + j])
                           * exp(-a * b / (a + b) * ab2);
            if (v > m) m = v;
        }
    return m;
}

void orc_k_build(const orc_basis *bs, int npts, const double *coords,
                 const double *weights, const double *QD, double omega,
                 double pair_cut, double *K)
{
    const int nao = bs->nao, ns = bs->nshell;
    /* shell pairs kept by the magnitude cutoff, sa >= sb */
    int *pa = malloc(sizeof(int) * ns * (ns + 1) / 2);
    int *pb = malloc(sizeof(int) * ns * (ns + 1) / 2);
    int np = 0;
    for (int sa = 0; sa < ns; ++sa)
        for (int sb = 0; sb <= sa; ++sb)
            if (block_bound(bs, sa, sb) > pair_cut) { pa[np] = sa; pb[np] = sb; ++np; }
    memset(K, 0, sizeof(double) * nao * nao);
#pragma omp parallel
    {
        double *Kp = calloc((size_t)nao * nao, sizeof(double));
        double *X = malloc(sizeof(double) * nao);
        double *F = malloc(sizeof(double) * nao);
        double *G = malloc(sizeof(double) * nao);
        double blk[49];
#pragma omp for schedule(dynamic, 4)
        for (int g = 0; g < npts; ++g) {
            const double *r = coords + 3 * g;
            ao_point(bs, r, 0, X, 0);
            for (int l = 0; l < nao; ++l) {
                double s = 0.0;
                for (int k = 0; k < nao; ++k) s += X[k] * QD[(size_t)k * nao + l];
                F[l] = s;
            }
            memset(G, 0, sizeof(double) * nao);
            for (int k = 0; k < np; ++k) {
                const int sa = pa[k], sb = pb[k];
                esp_block(bs, sa, sb, r, omega, blk);
                const int nsa = 2 * bs->l[sa] + 1, nsb = 2 * bs->l[sb] + 1;
                const int a0 = bs->ao0[sa], b0 = bs->ao0[sb];
                for (int m = 0; m < nsa; ++m)
                    for (int n = 0; n < nsb; ++n) {
                        const double v = blk[m * nsb + n];
                        G[b0 + n] += v * F[a0 + m];
                        if (sa != sb) G[a0 + m] += v * F[b0 + n];
                    }
            }
            const double w = weights[g];
            for (int m = 0; m < nao; ++m) {
                const double xw = X[m] * w;
                if (xw == 0.0) continue;
                double *row = Kp + (size_t)m * nao;
                for (int n = 0; n < nao; ++n) row[n] += xw * G[n];
            }
        }
#pragma omp critical
        for (size_t i = 0; i < (size_t)nao * nao; ++i) K[i] += Kp[i];
        free(Kp); free(X); free(F); free(G);
    }
    free(pa); free(pb);
}

/* Per-point screened K build, the CPU baseline's form of the (sub-block,
 * shell pair) screening of an SGX build: at point g the pair (A, B) is
 * skipped when w_g max(|F_A|, |F_B|) est(A, B, g) < tol.  est is an s-type
 * magnitude of the pair's ESP block from data kept per pair: c, the centre
 * of its most diffuse primitive pair, delta, the largest distance of a
 * primitive-pair centre from c, p_min, beta = sum_ij |c_i c_j|
 * exp(-mu_ij |AB|^2) 2 pi / p_ij, and the largest |P_ij A|, |P_ij B|:
 *   est = beta F_0(p_min max(0, |g c| - delta)^2)
 *         x (1 + |PA| + |g c| + delta)^la (1 + |PB| + |g c| + delta)^lb
 * (|erf(w r)/r| <= 1/r, so the 1/r form serves both operators; |F_A| the
 * largest |F| over shell A's AOs).  A cost model for timing, not a rigorous
 * bound; *kept counts the (point, pair) evaluations done. */
static double f0_bound(double T)
{
    return T < 1e-12 ? 1.0 : 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
}

typedef struct { double c[3], delta, pmin, beta, ra, rb; } pair_scr;

static void pair_screen(const orc_basis *bs, int sa, int sb, pair_scr *ps)
{
    const double *A = bs->center + 3 * sa, *B = bs->center + 3 * sb;
    const double ab2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1])
                     + (A[2] - B[2]) * (A[2] - B[2]);
    ps->pmin = 1e300;
    ps->beta = ps->delta = ps->ra = ps->rb = 0.0;
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < bs->nprim[sa]; ++i)
            for (int j = 0; j < bs->nprim[sb]; ++j) {
                const double a = bs->exps[bs->prim0[sa] + i], b = bs->exps[bs->prim0[sb] + j];
                const double p = a + b;
                double P[3], pa = 0, pb = 0, pc = 0;
                for (int d = 0; d < 3; ++d) {
                    P[d] = (a * A[d] + b * B[d]) / p;
                    pa += (P[d] - A[d]) * (P[d] - A[d]);
                    pb += (P[d] - B[d]) * (P[d] - B[d]);
                }
                if (pass == 0) {                       /* the most diffuse pair's centre */
                    if (p < ps->pmin) {
                        ps->pmin = p;
                        for (int d = 0; d < 3; ++d) ps->c[d] = P[d];
                    }
                    continue;
                }
                for (int d = 0; d < 3; ++d) pc += (P[d] - ps->c[d]) * (P[d] - ps->c[d]);
                if (sqrt(pc) > ps->delta) ps->delta = sqrt(pc);
                if (sqrt(pa) > ps->ra) ps->ra = sqrt(pa);
                if (sqrt(pb) > ps->rb) ps->rb = sqrt(pb);
                ps->beta += fabs(bs->coefs[bs->prim0[sa] + i] * bs->coefs[bs->prim0[sb] + j])
                          * exp(-a * b / p * ab2) * (2.0 * M_PI / p);
            }
}

static double pair_est(const pair_scr *ps, int la, int lb, const double *g)
{
    const double dx = g[0] - ps->c[0], dy = g[1] - ps->c[1], dz = g[2] - ps->c[2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const double e = d > ps->delta ? d - ps->delta : 0.0;
    double v = ps->beta * f0_bound(ps->pmin * e * e);
    const double fa = 1.0 + ps->ra + d + ps->delta, fb = 1.0 + ps->rb + d + ps->delta;
    for (int i = 0; i < la; ++i) v *= fa;
    for (int i = 0; i < lb; ++i) v *= fb;
    return v;
}

void orc_k_build_screened(const orc_basis *bs, int npts, const double *coords,
                          const double *weights, const double *QD, double omega,
                          double pair_cut, double tol, double *K, long long *kept)
{
    const int nao = bs->nao, ns = bs->nshell;
    int *pa = malloc(sizeof(int) * ns * (ns + 1) / 2);
    int *pb = malloc(sizeof(int) * ns * (ns + 1) / 2);
    int np = 0;
    for (int sa = 0; sa < ns; ++sa)
        for (int sb = 0; sb <= sa; ++sb)
            if (block_bound(bs, sa, sb) > pair_cut) { pa[np] = sa; pb[np] = sb; ++np; }
    pair_scr *scr = malloc(sizeof(pair_scr) * (np > 0 ? np : 1));
    for (int k = 0; k < np; ++k) pair_screen(bs, pa[k], pb[k], scr + k);
    memset(K, 0, sizeof(double) * nao * nao);
    long long nk = 0;
#pragma omp parallel reduction(+ : nk)
    {
        double *Kp = calloc((size_t)nao * nao, sizeof(double));
        double *X = malloc(sizeof(double) * nao);
        double *F = malloc(sizeof(double) * nao);
        double *G = malloc(sizeof(double) * nao);
        double *fs = malloc(sizeof(double) * ns);
        double blk[49];
#pragma omp for schedule(dynamic, 4)
        for (int g = 0; g < npts; ++g) {
            const double *r = coords + 3 * g;
            const double w = weights[g];
            ao_point(bs, r, 0, X, 0);
            for (int l = 0; l < nao; ++l) {
                double s = 0.0;
                for (int k = 0; k < nao; ++k) s += X[k] * QD[(size_t)k * nao + l];
                F[l] = s;
            }
            for (int sh = 0; sh < ns; ++sh) {
                double m = 0.0;
                for (int i = 0; i < 2 * bs->l[sh] + 1; ++i) {
                    const double v = fabs(F[bs->ao0[sh] + i]);
                    if (v > m) m = v;
                }
                fs[sh] = fabs(w) * m;
            }
            memset(G, 0, sizeof(double) * nao);
            for (int k = 0; k < np; ++k) {
                const int sa = pa[k], sb = pb[k];
                const double fm = fs[sa] > fs[sb] ? fs[sa] : fs[sb];
                if (fm * pair_est(scr + k, bs->l[sa], bs->l[sb], r) < tol) continue;
                ++nk;
                esp_block(bs, sa, sb, r, omega, blk);
                const int nsa = 2 * bs->l[sa] + 1, nsb = 2 * bs->l[sb] + 1;
                const int a0 = bs->ao0[sa], b0 = bs->ao0[sb];
                for (int m = 0; m < nsa; ++m)
                    for (int n = 0; n < nsb; ++n) {
                        const double v = blk[m * nsb + n];
                        G[b0 + n] += v * F[a0 + m];
                        if (sa != sb) G[a0 + m] += v * F[b0 + n];
                    }
            }
            for (int m = 0; m < nao; ++m) {
                const double xw = X[m] * w;
                if (xw == 0.0) continue;
                double *row = Kp + (size_t)m * nao;
                for (int n = 0; n < nao; ++n) row[n] += xw * G[n];
            }
        }
#pragma omp critical
        for (size_t i = 0; i < (size_t)nao * nao; ++i) K[i] += Kp[i];
        free(Kp); free(X); free(F); free(G); free(fs);
    }
    *kept = nk;
    free(pa); free(pb); free(scr);
}

/* ------------------------------------------------------------ VV10 ----- */
void orc_vv10_sum_outer(int n, const double *coords, const double *q, const double *W0,
                        const double *kappa, int nout, const int *outer, double *F,
                        double *U, double *W);
void orc_vv10_sum(int n, const double *coords, const double *q, const double *W0,
                  const double *kappa, double *F, double *U, double *W)
{
    orc_vv10_sum_outer(n, coords, q, W0, kappa, n, NULL, F, U, W);
}

/* F/U/W for the outer points listed in outer[0..nout) (all points when
 * outer == NULL), every point as inner point */
void orc_vv10_sum_outer(int n, const double *coords, const double *q, const double *W0,
                        const double *kappa, int nout, const int *outer, double *F,
                        double *U, double *W)
{
#pragma omp parallel for schedule(static)
    for (int o = 0; o < nout; ++o) {
        const int i = outer ? outer[o] : o;
        double f = 0.0, u = 0.0, w = 0.0;
        const double xi = coords[3 * i], yi = coords[3 * i + 1], zi = coords[3 * i + 2];
        for (int j = 0; j < n; ++j) {
            const double dx = xi - coords[3 * j], dy = yi - coords[3 * j + 1],
                         dz = zi - coords[3 * j + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double g = r2 * W0[i] + kappa[i];
            const double gp = r2 * W0[j] + kappa[j];
            const double gt = g + gp;
            const double T = q[j] / (g * gp * gt);
            const double Tu = T * (1.0 / g + 1.0 / gt);
            f += T;
            u += Tu;
            w += Tu * r2;
        }
        F[o] = -1.5 * f;
        U[o] = u;
        W[o] = w;
    }
}
