/*
 * wb97mv_amd -- C ABI of the MI355X (gfx950) omega-B97M-V / COSX hot path.
 *
 * Plain pointers and sizes only.  Host pointers are caller-owned,
 * C-contiguous and not retained past the call; device state lives behind
 * the opaque handles.  Every function returns a status (WB_OK, ...); the
 * message of the last failure on the calling thread is wb_last_error().
 * Calls on different contexts may run concurrently from different host
 * threads; one context (one HIP stream) is not reentrant.
 *
 * Reference interfaces replaced (Cedrus810/pyscf-wb97mv-fast):
 *   COSX exchange K   pyscf.sgx.sgx_jk.get_k_only as hooked in
 *                     gpu/install.py:293-330 -> gpu/sgx_k.py:983-991
 *                     (GpuKBuilder.get_k, gpu/sgx_k.py:663-751)
 *   VV10              NumInt.nr_nlc_vxc as hooked in gpu/install.py:276-291
 *                     -> gpu/vv10.py:268-342 (gpu_nr_nlc_vxc)
 *   VV10 kernel sum   gpu/vv10.py:167-240 (Vv10Kernel.kernel_sum)
 *   AO values         gpu/ao_eval.py:186-222 (AoEvaluator.eval)
 *   shell pairs       gpu/shellpairs.py:184-437 (ShellPairs, built natively
 *                     inside wb_basis_create)
 */
#ifndef WB97MV_AMD_H
#define WB97MV_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* the library is built with hidden visibility: these are its only exports */
#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

#define WB_ABI_VERSION 1

/* status codes; the reference raises gpu.xc.Unsupported (gpu/xc.py:30-32)
 * and retries on OOM (gpu/install.py:99-109) */
#define WB_OK 0
#define WB_UNSUPPORTED 1
#define WB_OOM 2
#define WB_DEVICE_ERROR 3
#define WB_INVALID 4

/* precision of the integral / VV10 pair arithmetic; AO values, GEMMs,
 * weights and accumulations across tiles are FP64 in both */
#define WB_PREC_MIXED 0
#define WB_PREC_FP64 1

typedef void* wb_ctx_t;
typedef void* wb_basis_t;
typedef void* wb_grid_t;
typedef void* wb_graph_t;

const char* wb_last_error(void);
int wb_version(void);

/* Context on one device; stream = NULL creates a private non-blocking
 * stream (the reference's per-flow streams, gpu/install.py:118-135). */
int wb_ctx_create(int device, void* stream, wb_ctx_t* out);
int wb_ctx_destroy(wb_ctx_t ctx);
/* Device bytes held by the context's grow-only workspaces (its batch lanes
 * included), and their release (they regrow on the next call; the call waits
 * for the context's streams).  No counterpart in the reference (CuPy's pool);
 * used for admission by HBM footprint when serving many molecules. */
int wb_ctx_mem(wb_ctx_t ctx, long long* bytes);
int wb_ctx_trim(wb_ctx_t ctx);
/* number of G partial copies of the exchange kernel (default 8) */
int wb_ctx_set_copies(wb_ctx_t ctx, int ncopy);
/* Autotuner of the K build's run-time choices on the caller's molecule:
 * times (minimum of `reps` builds, HIP events) the G copy count over
 * {1, 2, 4, 8, 16} (those whose G fits a quarter of the free device memory),
 * then, class by class, fused vs two-pass exchange kernels for the classes
 * with a fused two-operator form, then the FP64 GEMMs' MFMA tiles per wave
 * side (3 or 4, against the padding rule's choice), then the waves per
 * sub-block of each class (wb_ctx_set_class_waves; halved while faster),
 * and keeps the fastest on the context.  report (optional, 6): best copies, fuse mask, best ms, ms
 * of the default (8 copies, all fused, tiles by padding), configurations
 * timed, GEMM tiles (0: by padding).  Results stay within the mixed form's
 * tolerance of the default's (the copies change the FP32 summation order).
 * No counterpart in the reference (fixed launch shapes).  get/set_tuning
 * read or restore a configuration (e.g. a cached one). */
int wb_ctx_autotune(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* qd, int nop,
                    const double* omegas, double tile_tol, int precision, int reps,
                    double* report);
int wb_ctx_get_tuning(wb_ctx_t ctx, int* ncopy, unsigned* fuse_mask, int* gemm_tiles);
int wb_ctx_set_tuning(wb_ctx_t ctx, int ncopy, unsigned fuse_mask, int gemm_tiles);
/* Waves per 64-point sub-block of each (la, lb) class's exchange kernel
 * (class c = la(la+1)/2 + lb; 10 values): a sub-block's task list is cut
 * into this many chunks, run by separate waves, each writing its own G copy
 * (0: the copy count, the default; at most the copy count; the first class
 * with pairs always uses every copy).  wb_ctx_autotune picks them, class by
 * class, after the copy count; NULL restores the default.  No counterpart in
 * the reference (one launch shape per class, gpu/sgx_k.py:764-773). */
int wb_ctx_get_class_waves(wb_ctx_t ctx, int* waves10);
int wb_ctx_set_class_waves(wb_ctx_t ctx, const int* waves10);
/* Workspace budget of the context's K builds in bytes (0, the default: no
 * limit).  A K build whose workspace (AO values, F, the G copies, task
 * lists, the K GEMM's partials) would exceed it runs over chunks of the
 * grid's 64-point sub-blocks, each a whole number of the K GEMM's point
 * splits, and accumulates K chunk by chunk in split order: the result is
 * bit for bit the one-chunk result.  A budget below one split's workspace
 * makes the K build fail with WB_OOM.  The reference tiles its grid by a
 * memory budget the same way (gpu/sgx_k.py:502-520, :558-570). */
int wb_ctx_set_k_budget(wb_ctx_t ctx, long long bytes);
/* The plan a K build on (basis, grid) would use with the context's copy
 * count and budget: out3 = {workspace bytes, chunks, sub-blocks per chunk}. */
int wb_k_plan(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, int nop, int precision,
              long long* out3);
/* mixed form: accumulate G in FP64 (1) or FP32 (0, default) */
int wb_ctx_set_g64(wb_ctx_t ctx, int on);
int wb_sync(wb_ctx_t ctx);
float wb_last_ms(wb_ctx_t ctx);
/* Phase timing with HIP events on the context's stream (on = 1); read
 * accumulated milliseconds of the phases {AO, F GEMM, screening, exchange
 * kernels, K GEMM, NLC AO, NLC density, VV10 pair sum (with pointwise and
 * compaction), NLC potential + Fock} into ms[0..n). */
int wb_ctx_set_profile(wb_ctx_t ctx, int on);
int wb_profile_read(wb_ctx_t ctx, double* ms, int n, int reset);
/* Work of the last K build per (la, lb) class c = la(la+1)/2 + lb:
 * out[2c] kept tasks (64-point sub-block x shell pair), out[2c+1] their
 * primitive pairs (x 64 points = primitive-pair evaluations). */
int wb_k_work(wb_ctx_t ctx, long long* out20);

/* Segmented spherical basis (a generally contracted shell is passed as one
 * shell per contraction column), l <= 3.  coefs carry the libcint
 * normalization (mol._libcint_ctr_coeff).  Shell pairs whose max
 * |c_a c_b| K_AB <= pair_cut are dropped (the reference uses 1e-14,
 * gpu/shellpairs.py:290-294).  Replaces ShellPack + ShellPairs
 * (gpu/ao_eval.py:57-131, gpu/shellpairs.py:187-437). */
int wb_basis_create(wb_ctx_t ctx, int nshell, const int* l, const int* nprim,
                    const double* center, const double* exps, const double* coefs,
                    double pair_cut, wb_basis_t* out);
/* out4 = {nao, nshell, kept shell pairs, primitive pairs} */
int wb_basis_info(wb_basis_t basis, long long* out4);
int wb_basis_destroy(wb_basis_t basis);
/* Device bytes the basis holds, plus the primitive-pair records a K build of
 * nop operators would still add to it (both precisions; 0 operators: what
 * it holds now).  No counterpart in the reference (its per-mol builder,
 * gpu/sgx_k.py:649-660, holds its arrays in CuPy's pool); for admission by
 * device footprint. */
int wb_basis_bytes(wb_basis_t basis, int nop, long long* out);

/* Integration grid (bohr), resident on the device; points are reordered
 * internally into spatially compact 64-point sub-blocks.  Replaces
 * _GridState (gpu/sgx_k.py:496-542). */
int wb_grid_create(wb_ctx_t ctx, int npts, const double* coords, const double* weights,
                   wb_grid_t* out);
/* out3 = {npts, padded points, sub-blocks} */
int wb_grid_info(wb_grid_t grid, long long* out3);
int wb_grid_destroy(wb_grid_t grid);
/* Device bytes the grid holds (coordinates, weights, sub-block spheres). */
int wb_grid_bytes(wb_grid_t grid, long long* out);

/* COSX exchange for nop operators (omega = 0: 1/r, > 0: erf(omega r)/r):
 * K_op = X^T (w o G_op), G_op[g] = A_op(g) F[g], F = X qd, with qd = Q dm
 * (nao x nao, row-major).  Returns the raw K (nop x nao x nao) before
 * get_k_only's Q^T and symmetrization (gpu/sgx_k.py:736-741).  Tasks
 * (64-point sub-block, shell pair) whose bound is below tile_tol are
 * skipped (0 keeps all; the reference default 1e-11). */
int wb_k_build(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* qd, int nop,
               const double* omegas, double tile_tol, int precision, double* k_out);
/* the same with device pointers (inputs resident in HBM) */
int wb_k_build_dev(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* d_qd,
                   int nop, const double* omegas, double tile_tol, int precision,
                   double* d_k);
/* wb_k_build_dev without any device allocation, for callers that cannot
 * allocate: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  No counterpart in the reference.  Same inputs,
 * outputs and result (bit for bit, with the context's copy count and G
 * precision); every buffer it needs comes from one fixed 256 MiB device
 * array per GPU, held from the library's first context on that GPU until
 * the process ends, and the grid runs over as many chunks as the array
 * needs (K accumulated in the one-chunk order).  A molecule whose smallest
 * chunk does not fit is WB_UNSUPPORTED, with the bytes it needs.  The
 * primitive-pair records of the operator set must already be on the device
 * (any earlier wb_k_build / wb_k_build_dev / wb_cycle* call with the same
 * omegas on this basis made them); otherwise WB_INVALID.  All calls on one
 * GPU share its array: calls that overlap on one GPU race.  wb_k_stats and
 * wb_k_work keep describing the last allocating K build. */
int wb_k_build_noalloc(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* d_qd,
                       int nop, const double* omegas, double tile_tol, int precision,
                       double* d_k);
/* out2 = {tasks considered, tasks kept} of the last K build on grid */
int wb_k_stats(wb_ctx_t ctx, wb_grid_t grid, long long* out2);

/* VV10 term of nr_nlc_vxc for one RKS density matrix on grid:
 * out2 = {nelec, excsum}, vmat (nao x nao).  (b, C) = nlc_coeff (6.0, 0.01
 * for wB97M-V).  tile_tol > 0 enables tile-pair pruning of the double sum. */
int wb_nlc_vxc(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* dm, double b,
               double C, double tile_tol, int precision, double* out2, double* vmat);
int wb_nlc_vxc_dev(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* d_dm,
                   double b, double C, double tile_tol, int precision, double* d_out2,
                   double* d_vmat);

/* Semilocal XC on grid (the device side of nr_rks, gpu/xc.py:130-201; the
 * functional is evaluated by the caller between the two calls).
 * wb_xc_rho: rho5 (5 x npts, the grid's point order) = rho, d rho/dx, /dy,
 * /dz and (mgga != 0) tau = 1/2 sum |grad phi|^2 for the density matrix dm
 * (nao x nao, symmetric); tau = 0 otherwise.
 * wb_xc_vmat: vmat (nao x nao) = sum_g w_g [ v0 phi phi^T + v_d (d_d phi phi^T
 * + phi d_d phi^T) + 1/2 v4 sum_d d_d phi d_d phi^T ] for the chain-ruled
 * potential v5 (5 x npts: v_rho, 2 v_sigma d rho / dx, /dy, /dz, v_tau; v4
 * is used only when mgga != 0).  FP64. */
int wb_xc_rho(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* dm, int mgga,
              double* rho5);
int wb_xc_vmat(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* v5, int mgga,
               double* vmat);

/* One SCF cycle's device work as a HIP graph (no counterpart in the
 * reference, which launches every kernel per call; for callers that replay
 * cycles on fixed buffers, such as a device-resident SCF loop, and skip the
 * per-kernel launch cost): the K build of wb_k_build_dev
 * (when kgrid is not null: d_qd -> d_k) followed by the VV10 term of
 * wb_nlc_vxc_dev (when nlcgrid is not null: d_dm -> d_out2, d_vmat), with
 * the context's copy count and G precision.  Records by running the cycle
 * once (outputs written), then capturing it on the context's stream; the
 * graph owns its workspaces.  The device buffers are fixed at recording:
 * write each cycle's qd / dm into the same buffers, then launch.  basis and
 * the grids must outlive the graph. */
int wb_cycle_graph_create(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t kgrid, const double* d_qd,
                          int nop, const double* omegas, double k_tol, wb_grid_t nlcgrid,
                          const double* d_dm, double b, double C, double vv_tol, int precision,
                          double* d_k, double* d_out2, double* d_vmat, wb_graph_t* out);
/* enqueue one replay on the context's stream (asynchronous: wb_sync waits) */
int wb_cycle_graph_launch(wb_graph_t graph);
int wb_cycle_graph_destroy(wb_graph_t graph);

/* One SCF cycle's device work, stream-ordered on the context's stream
 * (device pointers; wb_sync waits): the K build on kgrid (d_qd -> d_k, nop
 * operators) and/or the VV10 term on nlcgrid (d_dm -> d_out2 = (nelec, exc),
 * d_vmat).  With both, the VV10 flow runs concurrently on a second stream of
 * the context (its own workspaces), forked and joined with events (in the
 * mixed form from the end of the K build's F GEMM), its pair sum on a grid
 * narrow enough to leave the exchange kernels CU room; the outputs equal
 * those of wb_k_build_dev followed by wb_nlc_vxc_dev bit for bit.  Replaces the
 * get_k_only calls (gpu/sgx_k.py:983-991) and the nr_nlc_vxc call
 * (gpu/vv10.py:268-342) of one cycle for callers that issue them together
 * (device-resident SCF loops, serving). */
int wb_cycle_dev(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t kgrid, const double* d_qd, int nop,
                 const double* omegas, double k_tol, wb_grid_t nlcgrid, const double* d_dm,
                 double nlc_b, double nlc_C, double vv_tol, int precision, double* d_k,
                 double* d_out2, double* d_vmat);

/* One molecule's cycle in a batch: a K build (kgrid != NULL) and/or the VV10
 * term (nlcgrid != NULL).  Host buffers, C-contiguous, nao = the basis's. */
typedef struct wb_cycle_job {
    wb_basis_t basis;
    wb_grid_t kgrid;        /* K build on this grid, or NULL */
    const double* qd;       /* in: Q dm (nao, nao) */
    double* k;              /* out: (nop, nao, nao) raw K */
    wb_grid_t nlcgrid;      /* VV10 on this grid, or NULL */
    const double* dm;       /* in: dm (nao, nao) */
    double* nlc;            /* out: (nelec, exc) */
    double* vmat;           /* out: (nao, nao) */
    int flags;              /* WB_JOB_SEQUENTIAL: the two flows one after the other */
} wb_cycle_job;
/* wb_cycle_job.flags: run the K build and the VV10 flow of the job in
 * sequence on the job's context (default 0: concurrently, the VV10 flow on
 * the context's second stream).  Results are the same either way; with
 * other work sharing the device the sequence avoids extra contention. */
#define WB_JOB_SEQUENTIAL 1

/* One molecule's cycle (a wb_cycle_job) on the context with host buffers:
 * the inputs go up, the K build and the VV10 term run as in wb_cycle_dev
 * (concurrently when both are asked for), the outputs come down; the call
 * returns when they are in place.  Outputs equal wb_k_build's and
 * wb_nlc_vxc's bit for bit.  Replaces one cycle's get_k_only calls
 * (gpu/sgx_k.py:983-991) and nr_nlc_vxc call (gpu/vv10.py:268-342). */
int wb_cycle(wb_ctx_t ctx, const wb_cycle_job* job, int nop, const double* omegas,
             double k_tol, double nlc_b, double nlc_C, double vv_tol, int precision);

/* Device bytes of one job's workspaces: *after = what the context will
 * hold once the job has run on it (its grow-only workspaces as they are,
 * grown where the job needs more), *alone = what the job needs from empty
 * (after wb_ctx_trim).  Exact (the sizes the flows allocate); the basis and
 * grid data are the caller's and not included.  No counterpart in the
 * reference; used for admission by HBM footprint. */
int wb_job_bytes(wb_ctx_t ctx, const wb_cycle_job* job, int nop, int precision,
                 long long* after, long long* alone);

/* Cycles of many molecules of uneven size, drained largest first from one
 * shared queue by `lanes` (1..16) host threads, each with its own stream and
 * workspaces on the context's device (kept for the next call), so that the
 * jobs' work overlaps on the device.  Each job's outputs equal wb_k_build's
 * and wb_nlc_vxc's for it alone, bit for bit.  job_ms (optional, njob):
 * each job's completion time in ms from the start of the call.  The
 * reference has no batch path (one molecule per SCF; staging/schedule.py:
 * 99-452 drives it); a job is the work of one get_k_only call
 * (gpu/sgx_k.py:663-751) per operator and one nr_nlc_vxc call
 * (gpu/vv10.py:268-342).  A job that runs out of device memory is retried
 * once alone, after every workspace of the context and its lanes has been
 * released; any other job error stops the batch and names the job.  The
 * context must not be used from another thread during the call.  (Test
 * hook: the environment variable WB_BATCH_FAIL_ONCE=i makes job i's first
 * attempt report out-of-memory before it starts, exercising the retry.) */
int wb_cycle_batch(wb_ctx_t ctx, int njob, const wb_cycle_job* jobs, int nop,
                   const double* omegas, double k_tol, double nlc_b, double nlc_C,
                   double vv_tol, int precision, int lanes, double* job_ms);

/* Cycles of many molecules with the exchange kernels of all their K builds
 * in one launch per shell-pair class: each molecule's set-up (AO values, F,
 * screening, task lists) runs into a workspace block of its own, then one
 * persistent launch per (la, lb) class drains a device work queue of the
 * (molecule, sub-block, chunk) items of all of them, then each molecule's K
 * GEMMs run.  The VV10 terms run meanwhile on the context's second stream.
 * Molecules are taken largest first in groups whose workspaces fit the
 * context's K budget (wb_ctx_set_k_budget; 64 GiB when none is set) and
 * the device memory free at the call; one that fits no group runs its own
 * K build, tiled to those limits.  The K arena is sized once per call for
 * all its groups and kept (wb_ctx_trim releases it).  Each job's outputs equal
 * wb_k_build's and wb_nlc_vxc's for it alone, bit for bit.  job_ms
 * (optional, njob): each job's completion time in ms from the start of the
 * call.  The reference has no batch path (staging/schedule.py:99-452 runs
 * one molecule per SCF); a job is one get_k_only call per operator
 * (gpu/sgx_k.py:663-751) and one nr_nlc_vxc call (gpu/vv10.py:268-342). */
int wb_cycle_queue(wb_ctx_t ctx, int njob, const wb_cycle_job* jobs, int nop,
                   const double* omegas, double k_tol, double nlc_b, double nlc_C,
                   double vv_tol, int precision, double* job_ms);

/* VV10 kernel sums on given points (coords n x 3): F = -1.5 sum T, U, W as
 * libdft's VXC_vv10nlc returns them (gpu/vv10.py:167-240). */
int wb_vv10_sum(wb_ctx_t ctx, int n, const double* coords, const double* q, const double* W0,
                const double* kappa, double tile_tol, int precision, double* F, double* U,
                double* W);

/* Tile pruning of the last VV10 sum run through ctx (wb_vv10_sum, or the
 * VV10 flow of wb_nlc_vxc / wb_cycle): out2 = {(wave, tile) pairs tested
 * against the pruning bound, pairs skipped}; both 0 when the sum ran with
 * tile_tol = 0 (nothing is tested).  Waits for the context's stream.  The
 * reference prunes tile pairs on the host before its launch and reports
 * nothing (gpu/vv10.py:147-164, :179-221); this exposes what the device
 * skipped. */
int wb_vv10_stats(wb_ctx_t ctx, long long* out2);

/* AO values (deriv 0) or values + gradients (deriv 1), FP64, out
 * (1 + 3 deriv, npts, nao) as PySCF's eval_gto (gpu/ao_eval.py:186-222). */
int wb_ao_eval(wb_ctx_t ctx, wb_basis_t basis, int npts, const double* coords, int deriv,
               double* out);

/* Atomic partition weights of atom-centred grid points (host arrays): for
 * point i owned by atom owner[i], out[i] = P_owner / sum_a P_a with the
 * Stratmann-Scuseria-Frisch cells (a = 0.64), the scheme of the package's
 * own Becke grids.  Replaces the cell-partition step of the SGX / XC grid
 * build the reference runs on the host (grids.build(with_non0tab=True),
 * gpu/install.py:164; with_df.build(level=...), staging/schedule.py:391).
 * natm < 1, an owner outside 0..natm-1 or two atoms at one position is
 * WB_INVALID. */
int wb_grid_partition(wb_ctx_t ctx, int natm, const double* atoms, int npts,
                      const double* coords, const int* owner, double* out);

/* An atom-centred grid built on the device: the points of radial shell sh
 * (atom shell_atom[sh], radius shell_r[sh], radial weight shell_w[sh]) are
 * atom + r u_k for the points (u_k, wa_k) of angular rule shell_rule[sh]
 * (rule k: rule_npts[k] rows of x, y, z, w in rule_xyzw, rules in order),
 * weighted shell_w wa_k times the cell partition of wb_grid_partition;
 * points of weight <= w_cut are dropped.  Written in shell order, then rule
 * order, to coords (npts x 3) and weights (capacity points at most; *npts =
 * the count kept).  Products and sums are rounded one by one, so the grid
 * equals the host's construction of the same points bit for bit.  Replaces
 * the SGX / XC grid build the reference leaves to PySCF on the host
 * (with_df.build(level=...), staging/schedule.py:391; grids.build,
 * gpu/install.py:164). */
int wb_grid_build(wb_ctx_t ctx, int natm, const double* atoms, int nshell, const int* shell_atom,
                  const double* shell_r, const double* shell_w, const int* shell_rule, int nrule,
                  const int* rule_npts, const double* rule_xyzw, double w_cut, long long capacity,
                  double* coords, double* weights, long long* npts);

/* Overlap matrix on the grid, S_g = sum_g w_g chi_mu(r_g) chi_nu(r_g) (nao x
 * nao, host out), FP64 AO values and FP64 MFMA: the quadrature overlap that
 * COSX's overlap-fitted projector is built from (PySCF's fit_ovlp; the
 * reference reads the finished projector, sgx._pjs_data.
 * _overlap_correction_matrix, gpu/sgx_k.py:698-704). */
int wb_grid_overlap(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, double* out);

/* The overlap-fitted projector of COSX (PJS): Q = S_g^{-1} S_an (nao x nao,
 * host arrays), S_g the grid overlap (as wb_grid_overlap) and S_an the
 * analytic overlap (PySCF's int1e_ovlp), by LU with partial pivoting on the
 * device (rocSOLVER), as PySCF builds sgx._pjs_data._overlap_correction_matrix
 * with scipy.linalg.solve(S_g, S) on the host (read by the reference at
 * gpu/sgx_k.py:698-704).  A singular S_g (the grid cannot represent the
 * basis) is WB_UNSUPPORTED. */
int wb_pjs_projector(wb_ctx_t ctx, wb_basis_t basis, wb_grid_t grid, const double* S_an,
                     double* Q);

/* Boys function F_0..F_nmax (nmax <= 6) at n values T (host arrays), out
 * (nmax + 1, n), by the device routines the exchange kernels use: FP32
 * (WB_PREC_MIXED, T rounded to FP32) or FP64.  Replaces boys_eval(cp, T,
 * nmax) (gpu/boys.py:194-205); nmax outside 0..6 is WB_INVALID as the
 * reference's ValueError. */
int wb_boys_eval(wb_ctx_t ctx, int n, const double* T, int nmax, int precision, double* out);

/* Test hook (host only, no device work): the screening records of the
 * basis's shell pairs, built as wb_basis_create builds them, for checking
 * the bounds against exact integrals (the reference checks its host bound
 * the same way, tests/gpu/test_shellpairs.py:36-56).  counts2 = {pairs,
 * primitive pairs}; with pairs and prims non-NULL (sized from a first call
 * with NULLs) it also writes, per pair, {shell A, shell B, bound centre
 * x, y, z, spread delta, beta, q, first primitive pair, primitive pairs} --
 * |A_ab(g)| <= beta min(1, sqrt(pi) / (2 sqrt(q) d)), d = |g - centre| -
 * delta -- and per primitive pair {P x, y, z, c, primitive of A, primitive
 * of B} with |A^(ab)(g)| <= c / |g - P|, the primitive-tail bound.  FP32
 * values as the device holds them, widened to double. */
int wb_test_pair_bounds(int nshell, const int* l, const int* nprim, const double* center,
                        const double* exps, const double* coefs, double pair_cut,
                        long long* counts2, double* pairs, double* prims);

/* Test hook (no reference counterpart): the FP64 MFMA GEMMs on host data.
 * kind 0: C[M][N] = A[M][K] B[K][N]; kind 1: C[M][N] = sum_k A[M][k]
 * B[N][k] w[k] (w may be NULL); kinds 2 / 3: the same with the 64 x 64
 * tile kernels. */
int wb_test_gemm(wb_ctx_t ctx, int kind, int M, int N, int K, const double* A, const double* B,
                 const double* w, double* C);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif

#ifdef __cplusplus
}
#endif
#endif
