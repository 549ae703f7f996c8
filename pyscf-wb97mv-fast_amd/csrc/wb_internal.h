// Internal data structures and kernel launchers of the wb97mv_amd library.
#pragma once
#include <map>
#include <memory>
#include <mutex>
#include "wb_common.h"

namespace wb {

// ---------------------------------------------------------------- basis ----
struct ShellInfo {            // device copy, one per (segmented) shell
    double cx, cy, cz;
    int l, nprim, prim0, ao0;
};

struct PairRec {              // one shell pair of a class, la >= lb
    int a0, b0;               // AO offsets of shells A and B
    int pp0, npp;             // primitive pairs [pp0, pp0 + npp) in the class arrays
    int shA, shB;             // shell indices (screening)
    int diag, onec;           // diag: A == B; onec: A and B on one centre
};

// Shell blocks: the shells of one side of a class -- A, of angular
// momentum la, when la > lb, else B -- are cut, in shell order, into blocks
// of md_bb(la, lb); an exchange kernel keeps one block's F values and G
// accumulators in registers while it runs through the shells of the other
// side paired with that block (its items).
constexpr int MD_BBMAX = 8;
// spare rows after the nao AO rows of F and G: the padding positions of the
// last block of a class point at row nao and touch rows nao .. nao + 2l
constexpr int MD_SPARE = 7;
__host__ __device__ constexpr bool md_block_a(int la, int lb) { return la > lb; }
// (sizes measured on water27 def2-SVP: for (p, s) 4 beats 8 by 7 % of the
// exchange time -- 8 p shells in registers halve the occupancy -- and 3 and
// 2 by 1 % and 3 %)
__host__ __device__ constexpr int md_bb(int la, int lb) {
    return la == lb ? (la <= 1 ? 8 : 1)
                    : (la == 1 ? 4 : (la == 2 ? 2 : (lb == 0 ? 4 : (lb == 1 ? 2 : 1))));
}

// One record per primitive pair and operator set, for the exchange kernels:
//   P (3 doubles) | operator constants C per operator | Hermite E,
// padded to 16 bytes.  FP32: C = (hi, lo) pairs of [theta p, c_0..c_L], E =
// the erl leading values then their rounding remainders; FP64: plain.
__host__ __device__ constexpr int md_crl(int la, int lb, bool f32) { return (la + lb + 2) * (f32 ? 2 : 1); }
__host__ __device__ constexpr int md_erl(int la, int lb, bool f32) {
    return 3 * (la + 1) * (lb + 1) * (la + lb + 1) * (f32 ? 2 : 1);
}
// nop operators: P | C_0 .. C_{nop-1} | E
__host__ __device__ constexpr int md_rec_bytes(int la, int lb, bool f32, int nop) {
    return (24 + (nop * md_crl(la, lb, f32) + md_erl(la, lb, f32)) * (f32 ? 4 : 8) + 15) / 16 * 16;
}
// classes whose kernel evaluates both operators of a K build in one pass
// (the registers of two sets of blocks and accumulators fit: (s, s),
// (p, s), (d, s))
// (measured on water27: fusing (p, p) too -- 223 VGPRs, two waves per SIMD
// -- costs 0.9 ms of exchange time instead of saving it)
__host__ __device__ constexpr bool md_fuse(int la, int lb) { return lb == 0 && la <= 2; }

struct PairGroup {            // the pairs of one item shell and one block
    int shX, x0, bblk, npair; // item shell, its AO offset, block index, pairs
    int pidx[MD_BBMAX];       // pair index for the block position j, -1 if none
    // the same pairs packed for the exchange kernels (as unsigned bits):
    // pp0 << 10 | (npp - 1) << 2 | one-centre << 1 | diag, npp 1..256
    int pw[MD_BBMAX];
};

struct PairScreen {           // screening data per pair (pair order of its class)
    float cx, cy, cz;         // centre of the most diffuse primitive pair
    float delta;              // max distance of any primitive centre from (cx,cy,cz)
    float beta;               // sum of per-primitive bound prefactors (rounded up)
    float q;                  // 0.75 * min p (rounded down)
    int sh;                   // shA | shB << 16 (the pair's shells)
    int pp;                   // pp0 << 9 | npp (its primitive pairs, npp <= 256)
};

struct ClassData {
    int la = 0, lb = 0, npairs = 0, nprims = 0;
    std::vector<PairRec> pairs;
    std::vector<PairScreen> scr;
    std::vector<double> P;            // (nprims, 3) primitive-pair centres
    std::vector<double> p;            // (nprims,) exponents p
    std::vector<double> pref;         // (nprims,) c_a c_b K_AB 2 pi / p
    std::vector<double> E;            // (nprims, erl) Hermite E records
    int erl = 0;                      // E record length
    // per primitive pair (within a shell pair, most important first): centre
    // and far-field bound factor beta sqrt(pi) / (2 sqrt(0.75 p)) (rounded
    // up) of the per-task primitive screening
    std::vector<float4> pbc;
    std::vector<int> pij;             // (nprims, 2) primitive indices in shells A and B (host)
    // pairs are ordered by (B block, A, B); groups: runs of one (B block, A)
    std::vector<PairGroup> groups;
    std::vector<int> bbao;            // [nblk][md_bb]: AO offset of the shell at a block position, nao (spare rows) if none
    int ngroups = 0;
    DevBuf d_pairs, d_scr, d_groups, d_bbao, d_pbc;
    // operator set (omegas) -> primitive-pair records (FP32, FP64), built
    // on first use
    std::map<std::vector<double>, std::pair<std::shared_ptr<DevBuf>, std::shared_ptr<DevBuf>>> opc;
    std::mutex opc_mu;                 // batch lanes may share the basis
};

struct Basis {
    int nshell = 0, nao = 0;
    std::vector<ShellInfo> shells;
    std::vector<double> exps, coefs;
    DevBuf d_shells, d_exps, d_coefs, d_shao;   // d_shao: ao0[nshell], nao[nshell]
    DevBuf d_row0;                              // int [nao]: first AO of the AO's shell
    ClassData cls[NCLASS];
    double pair_cut = 1e-14;
    long long npairs_total = 0;
};

// ----------------------------------------------------------------- grid ----
struct Grid {
    int npts = 0, Ppad = 0, nsub = 0;
    std::vector<int> perm;             // sorted position -> caller's point index
    std::vector<double> w_sorted;      // host copy (weights in sorted order)
    double origin[3] = {0, 0, 0};      // centroid (VV10 FP32 frame)
    DevBuf d_xyz;                      // double [3][Ppad] SoA, padding = first point of block
    DevBuf d_w;                        // double [Ppad], padding 0
    DevBuf d_sub;                      // float4 [nsub] centre (relative to nothing: absolute) + radius
};

// --------------------------------------------------------------- context ---
// the context's grow-only workspaces, in a fixed order (Ctx::bufs)
enum WsBuf { WS_X, WS_XV, WS_KARENA, WS_KP, WS_K, WS_T, WS_RHO, WS_VV, WS_B, WS_RED, WS_IO,
             WS_WORK, WS_LIST, WS_COUNT };
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    DevBuf d_b32, d_b64;               // Boys tables
    // workspaces (grow-only); ws_karena: every buffer of a K build
    // (k_plan), in one allocation
    DevBuf ws_X, ws_Xv, ws_karena, ws_Kp, ws_K;
    DevBuf ws_T, ws_rho, ws_vv, ws_B, ws_red, ws_io;
    DevBuf ws_list;                    // AO lists of the VV10 flow's density GEMM
    // device bytes a K build may take for its workspace (0: no limit; the
    // grid is then processed in one chunk)
    size_t k_budget = 0;
    // statistics of the last K build
    long long tasks_total = 0, tasks_kept = 0;
    // VV10 pruning counters of the last VV10 sum run through this context
    // (device, in a workspace of the context that ran it; or host values)
    const unsigned long long* vv_stat = nullptr;
    unsigned long long vv_stat_host[2] = {0, 0};
    float ms_last = 0.f;
    int ncopy = 8;
    // classes whose two operators run in one pass (bit = class id; only where
    // md_fuse allows it); set by wb_ctx_autotune
    unsigned fuse_mask = ~0u;
    // MFMA tiles per wave side of the K build's GEMMs (0: by padding)
    int gemm_tiles = 0;
    // waves per sub-block of each (la, lb) class's exchange kernel: the
    // sub-block's task list is cut into this many chunks, each written to its
    // own G copy (0: ncopy, the most; at most ncopy; the first class launched
    // always uses ncopy, since it initializes every copy's region)
    int cls_waves[NCLASS] = {0};
    bool g64 = false;          // mixed form: G accumulated in FP64 (default FP32)
    // per-class work of the last K build: kept tasks, kept primitive pairs
    DevBuf ws_work;
    // phase timing (HIP events on this context's stream), off by default
    bool profile = false;
    std::vector<std::pair<int, hipEvent_t>> marks;     // (phase id, event after it)
    hipEvent_t t0 = nullptr;
    double phase_ms[16] = {0};
    void mark(int phase);
    void start();
    // sub-contexts of wb_cycle_batch (own streams and workspaces), made on first use
    std::vector<std::unique_ptr<Ctx>> lanes;
    // sub-context of the VV10 branch of a concurrent cycle, and the events
    // that fork it from and join it to this context's stream
    std::unique_ptr<Ctx> side;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // with rec_gate set, the K build records ev_gate on this stream after
    // its (first chunk's) F GEMM: where a concurrent cycle's VV10 branch
    // starts in the mixed form
    hipEvent_t ev_gate = nullptr;
    bool rec_gate = false;
    // on the VV10 side context of a concurrent cycle: the pair sum's grid,
    // in blocks per 16 CUs (0: one block per work unit)
    int vv_cap16 = 0;
    Ctx() = default;
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
    ~Ctx();            // waits for and destroys what the context owns
    // device bytes held by the workspaces (lanes included); release them all
    // (the caller makes sure no work is in flight)
    size_t ws_bytes() const;
    void trim();
    DevBuf* bufs[WS_COUNT] = {&ws_X, &ws_Xv, &ws_karena, &ws_Kp, &ws_K, &ws_T, &ws_rho,
                              &ws_vv, &ws_B, &ws_red, &ws_io, &ws_work, &ws_list};
};

enum Phase { PH_AO = 0, PH_FGEMM, PH_SCREEN, PH_MD, PH_KGEMM, PH_NLC_AO, PH_NLC_RHO,
             PH_VV10, PH_NLC_FOCK, PH_COUNT };

// --------------------------------------------------------- launchers --------
void launch_ao_eval(hipStream_t s, const Basis& b, const Grid& g, int deriv, double* X);
// AO values below this magnitude at every point of a 64-point sub-block
// leave the shell out of that sub-block's AO list (the reference prunes AO
// blocks with PySCF's non0tab the same way, gpu/ao_eval.py:134-153)
constexpr double AO_SIG_THR = 1e-16;
// AO lists of a grid in one workspace: sig [nsub][nshell] bytes; for the
// 128-point tiles (two sub-blocks each) of the gathered NN GEMM counts
// [ntile], lists [ntile][cap]; for the point splits of the split NT GEMM
// (kchunk points, spt sub-blocks each) scounts [nsplit], slists and sinv
// (the position of each AO in its split's list, or -1) [nsplit][cap]
struct AoLists {
    size_t sig = 0, counts = 0, lists = 0, scounts = 0, slists = 0, sinv = 0, total = 0;
    int cap = 0, spt = 0, nsplit = 0;
};
AoLists ao_lists_layout(const Basis& b, int nsub, int kchunk);
// X (values; with deriv, the gradient planes too) as launch_ao_eval, plus
// the lists of the AOs significant (value or stored gradient) in each tile
// and each point split, into ws (ao_lists_layout(...).total bytes)
void launch_ao_eval_lists(hipStream_t s, const Basis& b, const Grid& g, int deriv, double* X,
                          void* ws, const AoLists& L);
void launch_ao_eval_pts(hipStream_t s, const Basis& b, const double* xyz, int Ppad,
                        int deriv, double* X);

// MFMA tiles per wave side of the FP64 GEMMs launched from this host thread
// (3 or 4; 0: by padding); gemm_tiles_scope sets it for one flow
void set_gemm_tiles(int t);
int get_gemm_tiles();
struct GemmTilesScope {
    int prev;
    explicit GemmTilesScope(int t) : prev(get_gemm_tiles()) { set_gemm_tiles(t); }
    ~GemmTilesScope() { set_gemm_tiles(prev); }
};
// C[M][N] (+)= A[M][K] B[K][N], row-major; out float or double
void gemm_nn(hipStream_t s, int M, int N, int K, const double* A, const double* B,
             void* C, bool out_f32);
// C[M][N] = sum_k A[M][k] * Bsum[N][k] * (w ? w[k] : 1), Bsum = sum of nb copies
// (stride bstride) of type float (b_f32) or double; split over K, deterministic.
void gemm_nt_reduce(hipStream_t s, int M, int N, int K, const double* A,
                    const void* B, bool b_f32, int nb, size_t bstride,
                    const double* w, double* C, DevBuf& ws);

// C[M][N] = A^T B over the AO rows listed per 128-point n tile (A [K][M]
// and B [K][N] row-major by AO; launch_ao_eval_lists's ws and layout); M
// even, N a multiple of 64, A and B 16-byte aligned (gemm_nn_gather_ok)
bool gemm_nn_gather_ok(int M, int N, const double* A, const double* B);
void gemm_nn_gather(hipStream_t s, int M, int N, const double* A, const double* B, void* C,
                    bool out_f32, const void* ws, const AoLists& L);
// The split NT GEMM's rows of A (AOs) restricted per point split to the
// split's list of significant AOs (ao_lists_layout's slists / scounts /
// sinv): the split's partial holds those rows only, and the reduce adds each
// where its AO appears
struct NtGather {
    const int* counts = nullptr;
    const int* lists = nullptr;
    const int* inv = nullptr;
    int cap = 0, nsplit = 0;
};
NtGather nt_gather(const AoLists& L, const void* ws);

// MFMA tile versions (the production paths); gemm_nt_fast folds the copies
// and weights of B into one FP64 operand first when needed.  masks (with
// row0, the first row of each row's shell, and nsub): the exchange kernels'
// written-shell masks of the copies (MdMasks layout, one operator): rows of a
// copy's sub-block region that were never written are skipped (K = nsub * 64).
// gather: A's rows per split as NtGather lists (null: every row)
void gemm_nn_fast(hipStream_t s, int M, int N, int K, const double* A, const double* B,
                  void* C, bool out_f32);
void gemm_nt_fast(hipStream_t s, int M, int N, int K, const double* A, const void* B,
                  bool b_f32, int nb, size_t bstride, const double* w, double* C, DevBuf& ws,
                  const unsigned* masks = nullptr, const int* row0 = nullptr, int nsub = 0,
                  const NtGather* gather = nullptr);

// fmax[s][sh] = max over sub-block s of w_g sum_{mu in sh} |F[mu][g]|
void launch_fmax(hipStream_t s, const void* F, bool f32, const double* w, int nao,
                 int Ppad, int nsub, const int* sh_ao0, const int* sh_nao, int nshell,
                 float* fmax);
// one class's screening inputs and task-list outputs
struct TaskClass {
    const PairScreen* scr;
    const PairGroup* groups;
    const float4* pbc;
    int* counts;
    int* ids;
    unsigned* skips;
    unsigned long long* work;
    int* wsub;                 // kept primitive pairs per sub-block (its waves' work)
    int* order;                // sub-blocks by descending wsub (launch_order); may be null
    int ngroups, bb;
};
struct TaskLaunch {
    TaskClass c[NCLASS];
    int n = 0;                 // classes with pairs
};
// kept items (group << 8 | kept-position mask) per sub-block, and per item
// the trailing primitive pairs each kept task skips (4 bits per position),
// for every class of tl in one launch
void launch_tasks(hipStream_t s, const TaskLaunch& tl, const Grid& g, const float* fmax,
                  int nshell, float tol);
// per class of tl: order = the sub-blocks sorted by descending wsub (binned;
// ties in any order), so that an exchange launch dispatches its longest
// waves first and its tail holds short ones.  Results do not depend on it:
// every wave writes only its own (copy, sub-block) region.
void launch_order(hipStream_t s, const TaskLaunch& tl, int nsub);
// f32: FP32 integrals (else FP64); g64: accumulate G in FP64 (always with FP64 integrals)
// F_0..F_L at n values T (device arrays), out [L+1][n]
void launch_boys_eval(hipStream_t s, int L, int n, const double* T, bool f32,
                      const void* b32, const void* b64, double* out);
// recs: the class's primitive-pair records for nop operators (md_rec_bytes);
// nop = 2 only where md_fuse; G of operator o at G + o opstride (elements).
// Written-shell masks of the G regions (MD_MASK_WORDS words per copy and
// sub-block, [copy][sub-block][word], operator o at m + o opstride; a shell
// marked at its first row): a region's first read-modify-write of a shell
// reads the zero row zrow, so G is never cleared; first: the first class
// launched in a K build (starts from empty masks and writes the zero row).
// m = nullptr: no masks, and zero_rows > 0 makes each wave clear rows
// [0, zero_rows) of its region first.
constexpr int MD_MASK_WORDS = 64;                    // rows < 2048
struct MdMasks {
    unsigned* m = nullptr;
    size_t opstride = 0;
    int zrow = 0;
    bool first = false;
};
// One molecule's share of a class launch of the exchange kernel: its points,
// F, G copies (operator o at G + o opstride), the class's task lists of its
// sub-blocks (ids/skips: cap entries per sub-block), pair groups and
// primitive records.  nchunk (<= the number of G copies): waves per
// sub-block, chunk j of a sub-block's task list writing copy j.  masks null:
// the launch clears zero_rows rows of every region first.
struct MdJob {
    const double* xyz;
    const void* F;
    void* G;
    size_t gstride, opstride;
    const int* counts;
    const int* ids;
    const unsigned* skips;
    const PairGroup* groups;
    const int* bbao;
    const char* recs;
    unsigned* masks;
    const int* order;          // sub-block of each rank (null: identity order)
    size_t mopstride;
    int cap, Ppad, nsub, nchunk, zero_rows, zrow, first, pad_;
};
MdJob md_job(const ClassData& c, const Grid& g, const void* F, void* G, size_t gstride,
             size_t opstride, int nchunk, const int* counts, const int* ids,
             const unsigned* skips, const void* recs, int zero_rows, const MdMasks& mk);
// waves of one molecule's share (its work items in a queue launch)
inline long long md_items(const MdJob& j) { return (long long)j.nsub * j.nchunk; }
void launch_md(hipStream_t s, int la, int lb, bool f32, bool g64, int nop, const MdJob& j,
               const void* boys);
// One launch of class (la, lb) over njob molecules' shares (jobs, on the
// device): a grid sized to the device's resident waves takes work items from
// the counter head (zero at the launch); first[m] (njob + 1 entries, on the
// device): the first item of job m, first[njob] = total.  Each molecule's G
// equals the one its own launch_md writes, bit for bit.
void launch_md_queue(hipStream_t s, int la, int lb, bool f32, bool g64, int nop,
                     const MdJob* jobs, const long long* first, int njob, long long total,
                     unsigned long long* head, const void* boys);

// GGA density on the grid without stored AO gradients: rho[0][g] =
// sum_mu T[mu][g] X[mu][g], rho[1+d][g] = 2 sum_mu T[mu][g] dX_d[mu][g]
// (T = D X), the AO values and gradients evaluated on the fly; part:
// ao_rho_bytes scratch (partials per shell group, summed in fixed order)
size_t ao_rho_bytes(const Basis& b, int Ppad);
void launch_ao_rho(hipStream_t s, const Basis& b, const Grid& g, const double* T, double* part,
                   double* rho /* [4][Ppad] */);
// B[mu][g] = v0 X + v1 dX/dx + v2 dX/dy + v3 dX/dz, vg = [4][Ppad] (the GGA
// potential factors), AO values and gradients evaluated on the fly
void launch_ao_proj(hipStream_t s, const Basis& b, const Grid& g, const double* vg, double* Bout);

// out[g] (+)= scale * sum_mu A[mu][g] B[mu][g], A and B [nao][Ppad]
void launch_dot_planes(hipStream_t s, int nao, int Ppad, const double* A, const double* B,
                       double scale, bool accumulate, double* out);
// y += a x
void launch_axpy(hipStream_t s, size_t n, double a, const double* x, double* y);
// wv = (w v0 / 2, w v1, w v2, w v3, w v4 / 2), v and wv [5][Ppad]
void launch_xc_weights(hipStream_t s, int Ppad, const double* w, const double* v, double* wv);
// Stratmann-Scuseria-Frisch partition weight of each point's owner atom;
// xyz (npts, 3), atoms (natm, 3), rab / inv_rab (natm, natm) (0 on the diagonal)
void launch_ssf_partition(hipStream_t s, int npts, const double* xyz, const int* owner, int natm,
                          const double* atoms, const double* rab, const double* inv_rab,
                          double* out);
// atom-centred raw points (xyz [n][3], owner atom, radial x angular weight)
// from per-shell radial data and angular rules (x, y, z, w per point)
void launch_grid_points(hipStream_t s, long long nraw, int nshell, const long long* shell_first,
                        const int* shell_atom, const double* shell_r, const double* shell_w,
                        const int* shell_rule, const int* rule_off, const double* rule_xyzw,
                        const double* atoms, double* xyz, int* owner, double* wpre);
// the points whose weight wpre x part exceeds w_cut, in order: xyz_out,
// w_out; bcount: grid_keep_blocks(n) + 1 ints, the kept count at the end
int grid_keep_blocks(long long n);
void launch_grid_keep(hipStream_t s, long long n, const double* xyz, const double* wpre,
                      const double* part, double w_cut, int* bcount, double* xyz_out,
                      double* w_out);

// VV10
// VV10 double sum over n = *nptr <= ncap compacted points.  pts: (x, y, z,
// W0) and pk: (kappa, q) per point, float or double (f64); tiles: per
// 64-point tile two float4 (centre xyz, radius | W0min, kmin, qsum, 0); ws:
// vv10_sym_bytes(ncap) bytes.  Out: F (times -3/2), U, W per point.  stat
// (may be null; cleared here): with tol > 0, the (wave, tile) pairs tested
// for pruning and those skipped.
size_t vv10_sym_bytes(int ncap, bool f64, int* nsplit_out);
// cap16 > 0: a grid of cap16 blocks per 16 CUs, each block taking work units
// (group, part) in turn, so that a kernel running beside it keeps CU room;
// the same sums either way
void launch_vv10_sym(hipStream_t s, const int* nptr, int ncap, const void* pts, const void* pk,
                     const float4* tiles, float tol, bool f64, void* ws, double* F, double* U,
                     double* W, unsigned long long* stat, int cap16 = 0);

void zero_bytes(hipStream_t s, void* p, size_t bytes);
void transpose(hipStream_t s, int n, const double* a, double* b);
void symadd(hipStream_t s, int n, const double* a, double* c);
// The VV10 part of nr_nlc_vxc after the density: pointwise terms, compaction,
// double sum, potential factors wv = [4][Ppad] (the operand B = launch_ao_proj
// of them; vmat = X^T B + (X^T B)^T) and sums = (nelec, excsum).
struct VvLayout { size_t cnt, off, pt, pk, idx, aux, ts, F, U, W, sym, stat, total; };
VvLayout vv_layout(int nsub, int Ppad, bool f64);
// workspace bytes gemm_nt_fast takes (partials, plus the folded operand)
size_t gemm_nt_ws_bytes(int M, int N, int K, bool fold);
// the point-axis split gemm_nt_fast uses for K points: returns the number of
// splits, kchunk = points per split (a multiple of 64)
int nt_split(int K, int& kchunk);
// whether gemm_nt_core needs the folded operand buffer Bf (N x K doubles)
bool gemm_nt_folds(bool b_f32, int nb, const double* w, const void* B, const unsigned* masks);
// gemm_nt_fast on caller-provided buffers with a given split: C = (accumulate
// ? C : 0) + sum over splits z (in order) of A[:, split z] Bsum[:, split z]^T,
// Cp = ceil(K / kchunk) x M x N partials, Bf the folded operand (or null when
// gemm_nt_folds says none is needed).  A build over chunks of whole splits
// that accumulates them in order gives the one-call result bit for bit.
void gemm_nt_core(hipStream_t s, int M, int N, int K, int kchunk, const double* A, const void* B,
                  bool b_f32, int nb, size_t bstride, const double* w, double* C, bool accumulate,
                  double* Cp, double* Bf, const unsigned* masks, const int* row0, int nsub,
                  const NtGather* gather = nullptr);
struct Ctx;
void vv10_flow(hipStream_t s, int nsub, int Ppad, const double* xyz, const double* w,
               const double* origin, const double* rho, double b, double C, double tol, bool f64,
               DevBuf& ws, double* wv, double* e, double* sums, Ctx* prof, int cap16 = 0);

}  // namespace wb
