/* FP64 CPU oracle (test infrastructure only; see wb_oracle.c). */
#ifndef WB_ORACLE_H
#define WB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Segmented spherical basis, one contraction column per shell.
 * Coefficients carry the libcint normalization (radial function normalized,
 * int R^2 r^{2l+2} dr = 1); c2s[l] is the (2l+1, ncart) cart->sph matrix
 * (real harmonics normalized on the unit sphere, libcint cart order). */
typedef struct {
    int nshell, nao;
    const int *l, *nprim, *prim0, *ao0;   /* per shell */
    const double *center;                  /* (nshell, 3) bohr */
    const double *exps, *coefs;            /* flat primitives */
    const double *c2s[4];
} orc_basis;

void orc_boys(int nmax, double T, double *F);
void orc_boys_batch(int nmax, int n, const double *T, double *F);
/* (npts, nao, nao) dense A^g; omega = 0 -> 1/r, > 0 -> erf(omega r)/r */
void orc_esp_dense(const orc_basis *bs, int npts, const double *coords,
                   double omega, double *out);
/* (1 + 3*deriv, npts, nao) AO values (+ x, y, z gradients) */
void orc_ao_eval(const orc_basis *bs, int npts, const double *coords, int deriv,
                 double *out);
/* K_raw = X^T (w o (A F)), F = X QD; shell pairs with max |c c| K_AB <=
 * pair_cut are skipped (pass 0 for none).  No symmetrization. */
void orc_k_build(const orc_basis *bs, int npts, const double *coords,
                 const double *weights, const double *QD, double omega,
                 double pair_cut, double *K);
void orc_k_build_screened(const orc_basis *bs, int npts, const double *coords,
                          const double *weights, const double *QD, double omega,
                          double pair_cut, double tol, double *K, long long *kept);
/* VV10 kernel sums: F = -1.5 sum T, U, W as libdft returns them */
void orc_vv10_sum(int n, const double *coords, const double *q, const double *W0,
                  const double *kappa, double *F, double *U, double *W);

void orc_vv10_sum_outer(int n, const double *coords, const double *q, const double *W0,
                        const double *kappa, int nout, const int *outer, double *F,
                        double *U, double *W);

#ifdef __cplusplus
}
#endif
#endif
