// Internal interface shared by the host files of the library (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "phf_common.h"
#include "../../include/phf_hip.h"

void lu_inverse_solve_batched(hipStream_t s, int n, int nso, const double* M, long mbs, const double* T, long sT,
                              double* X, long sX, double* det, int* info, int batch, const double* U = nullptr,
                              int npg = 1);
void lu_inverse_batched(hipStream_t s, int n, const double* M, long mbs, double* Minv, long ibs,
                        double* det, int* info, int batch);
int lu_inverse_max_n();
void eri_gemm_f64(hipStream_t s, const double* P, long ldp, const double* A, double* C, int Mp, int Nr, int K,
                  long ldc, int bm, const double* bias, const int* colmap);
int eri_gemm_f64_bm(int rows);
int mixed_segments(int prec);
int mixed_slices(int prec);
int mixed_elem_bytes(int prec);
void split_operand(hipStream_t s, int prec, const double* src, long lds, int rows, int kt, void* dst, double* inv);
void split_operand_map(hipStream_t s, int prec, const double* src, long lds, int rows, int kt, const int* cmap,
                       int kout, void* dst, double* inv, const double* amax = nullptr, int rows_data = -1);
void convert_f32(hipStream_t s, const double* src, float* dst, long n);
// Density pack + 16-bit/int8 split in one pass (prec 2, 3 or 4; NSO <= dens_split_max_nso()):
// slices of rows 2g, 2g+1 for g < npts from C0 and X_g = X + g sX; padding rows up to rows_total.
int dens_split_max_nso();
void dens_split(hipStream_t s, int prec, int nso, int nk, const double* C0, const double* X, long sX, int npts,
                const int* ocol, const int* padcols, int npad, int kout, long rows_total, void* dst, double* inv);
// The same pass writing the spin-free operand (n <= 64 labels): smap = tile
// row -> MO [128] then tile column -> MO [128]; four row sets of Mp rows
// (Mk = 3 Mp): P_aa + P_bb, P_aa - P_bb, P_x1, P_x2 (2 npts data rows each); kj columns.
void dens_split_rjk(hipStream_t s, int prec, int n, int nk, const double* C0, const double* X, long sX, int npts,
                    const int* smap, int kj, long Mp, long Mk, void* dst, double* inv);
// Staging maps of dens_split_rjk for the pairing mo (label s: alpha MO mo[s],
// beta MO mo[n + s]): smap[0..128) tile row -> MO, smap[128..256) tile column -> MO.
void dens_split_rjk_maps(const std::vector<int>& mo, int n, int* smap);
void eri_gemm_mixed(hipStream_t s, int prec, const void* P, long kp, const void* A, const double* invP,
                    const double* invA, double* C, int Mp, int Nr, int K, long ldc, int cfg, int splitk,
                    const double* bias, const int* colmap);
void eri_gemm_mixed_grouped(hipStream_t s, int prec, const H16Item* items, int nitems, int cfg, int grid,
                            bool atomic);
int h16_grouped_grid(int cfg, int ncu);
int h16_num_configs();
int h16_cfg_bm(int c);
int h16_cfg_bn(int c);
bool h16_cfg_fold64(int c);
// whether config c is the round-to-nearest 16-bit tile, and the largest NSO it
// serves (its accuracy limit; the contraction limit is in h16_cfg_serves)
bool h16_cfg_rn16(int c);
int h16_rn16_max_nso();
bool h16_cfg_serves(int c, int prec, int K);
int h16_slabs(int prec, int c, int K);
int h16_forced_config();
int h16_forced_splitk();
double h16_model_cost(int c, int sk, long Mp0, int Nr0, int ncu);

struct phf_ctx_s {
  int device = 0;
  int ncu = 256;
  hipStream_t stream = nullptr;
  bool owns_stream = true;    // false for the instances of a batch (they share the batch's stream)
  // a context with its own stream runs the three spin-block GEMMs of an
  // evaluation concurrently: two side streams, fork/join events
  hipStream_t side[2] = {};
  hipEvent_t fork_ev = nullptr, join_ev[2] = {};
  // Hc = C0^H h runs on side[0] during the pre-stages (hc_fork, hc_join)
  hipEvent_t hc_fork = nullptr, hc_join = nullptr;
  // spin-rotation grids: the T_g combine runs on side[1] beside the overlap
  // products (tq_fork, tq_join)
  hipEvent_t tq_fork = nullptr, tq_join = nullptr;
  // the determinant weights (wd_kernel) run on side[0] once the inverses are
  // done (wd_fork); hc_join then also covers them
  hipEvent_t wd_fork = nullptr;
  bool wd_forked = false;
  std::string err;
  // cached Hamiltonian
  int nso = 0;
  bool has_ham = false;
  bool cplx = false;
  int kpad = 0, nrpad = 0, kt = 0;
  double* dA = nullptr;     // [nrpad][kt] FP64
  double* dH1 = nullptr;    // hone, row-major complex128
  double* dHb = nullptr;    // hone as ERI-GEMM bias: [2][nrpad] (re plane, im plane), r = i*nso + k
  // reduced-precision copies of A, built on first use per mode:
  // [nrpad][nseg*kt] of FP32 / BF16 / FP16 and per-row inverse scales
  void* dAm[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double* dAinv[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int a_only = 0;            // > 0: only the reduced-precision copy of this mode is stored (dA released)
  // Spin-block structure (every MO of Ref purely alpha or beta): with the pairs
  // (i,k) and (j,l) ordered by spin class A is block diagonal, diag(D, X1, X2):
  // D rows/cols s_i = s_k, X1 rows (alpha, beta) x cols (beta, alpha), X2 the
  // reverse; all other entries of <ij||kl> vanish.  Blocks are stored in FP64
  // (dAb, block b at element blk_aoff[b], rows blk_rows[b] x blk_k[b]) and in
  // the reduced modes (dAbm/dAbinv, slices per block row).
  int ham_flags = 0;         // flags of the phf_set_hamiltonian* call
  bool blocked = false;      // blocks built for the current Hamiltonian
  bool use_blocks = true;    // evaluate through the blocks when built
  int blk_rows[3] = {0, 0, 0}, blk_k[3] = {0, 0, 0}, blk_kbeg[3] = {0, 0, 0}, blk_roff[3] = {0, 0, 0};
  long blk_aoff[3] = {0, 0, 0};
  int ktp = 0;               // permuted contraction length (sum of blk_k)
  int* d_rowmap = nullptr;   // [sum blk_rows] permuted A row -> natural i*nso + k, -1 padding
  int* d_colperm = nullptr;  // [ktp] permuted contraction column -> natural j*nso + l, -1 padding
  int* d_ocol = nullptr;     // [nso^2] natural j*nso + l -> permuted column (inverse of d_colperm)
  int* d_padcols = nullptr;  // permuted columns with d_colperm == -1
  int n_padcols = 0;
  double* dAb = nullptr;
  void* dAbm[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double* dAbinv[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<int> blk_spin; // MO spins (0 alpha, 1 beta) the blocks were built for
  // Spin-free form (spin-restricted Ref: every alpha MO has a beta partner
  // with the same spatial coefficients, label s = 0..n-1).  With the spatial
  // MO integrals J[(a,c),(b,d)] = (ac|bd) and K[(a,c),(b,d)] = (ad|bc) the
  // blocks are D = [[J-K, J], [J, J-K]], X1 = X2 = -K, so
  //   Gamma_aa + Gamma_bb = (2J - K)(P_aa + P_bb),  Gamma_aa - Gamma_bb = -K (P_aa - P_bb),
  //   Gamma_ab = -K P_x1,  Gamma_ba = -K P_x2
  // (P_aa(b,d) = P[(alpha b, alpha d)], P_bb likewise, P_x1(b,d) = P[(beta b, alpha d)],
  // P_x2(b,d) = P[(alpha b, beta d)]): four n^2 x n^2 products per density
  // instead of the blocks' six.  dAj holds [2J - K; K] in FP64 ([2][rj_kj][rj_kj]),
  // dAjm/dAjinv the reduced copies; d_rjmo the MOs of each label
  // ([0..n) alpha, [n..2n) beta); d_rjsmap the fused density kernel's staging
  // maps (tile row -> MO [128], tile column -> MO [128], -1 empty).
  bool rjk = false;          // built for the current Hamiltonian and Ref
  bool use_rjk = true;       // evaluate through it when built (and the blocks are on)
  int rj_n = 0, rj_kj = 0;
  double* dAj = nullptr;
  void* dAjm[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double* dAjinv[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int* d_rjmo = nullptr;
  int* d_rjsmap = nullptr;
  int* d_rjcol = nullptr;    // [4][rj_kj]: label pair a n + c -> natural column of (aa, bb, ab, ba) (-1 padding)
  std::vector<int> rj_mo;    // host copy of d_rjmo (the pairing it was built for)
  // workspace arena
  char* ws = nullptr;
  size_t ws_size = 0;
  // device-resident spin grid (phf_set_grid)
  bool has_grid = false;
  int g_nl = 0, g_ny = 0, g_npg = 0, g_nci = 0, g_ncisp = 0, g_ncipg = 0, g_qj = 0, g_sp = 0, g_nso = 0;
  long g_G = 0;
  double* dGr1p = nullptr;     // Ref^H R1[il]            [nl][nso x nso] column-major
  double* dGr2 = nullptr;      // R2[iy]                  [ny] row-major (as given)
  double* dGrpgref = nullptr;  // Rpg[ipg] Ref            [npg][nso x nso] column-major
  double* dGw = nullptr;       // weight x Wigner x weightpg  [G][nci][nci]
  // spin-rotation grids: every R1[il] and R2[iy] is u (x) I_n for a 2 x 2 u
  // (AO order [alpha | beta]); then R1 R2 = U_g (x) I_n and dGu holds U_g
  // ([G][2][2] complex), dGref is Ref (row-major as given)
  bool g_spinrot = false;
  double* dGu = nullptr;
  double* dGref = nullptr;
  double* dGP = nullptr;       // [s][t][ipg]: Ref_s^H (Rpg[ipg] Ref)_t  (NSO x NSO, column-major)
  // results of the last evaluation kept on the device (phf_eval_resident / phf_localg_vo)
  double* r_ysum = nullptr;   // [nci*nci] x (NSO x N) column-major, sum_g Wd Y_g
  double* r_xsum = nullptr;   // [nci*nci] x (NSO x N), sum_g Wd X_g
  int r_nocc = 0;
  bool r_valid = false;
  // small persistent device buffer for per-call arguments/results outside the arena
  char* small = nullptr;
  size_t small_size = 0;
  // pinned host staging of one evaluation's C0 upload and small results
  // (asynchronous copies; reused after each call's final synchronization)
  char* pin = nullptr;       // coherent, mapped host memory
  char* pin_dev = nullptr;   // its device address (kernels store the objective's results there)
  size_t pin_size = 0;
  bool pin_busy = false;   // copies from/to pin may be in flight (a call ended before its synchronization)
  // the two exchange blocks X1, X2 as one grouped launch (reduced forms):
  // timed once per (precision, rows) against two launches, 1 = grouped
  std::map<std::tuple<int, long, int, int>, int> xgroup;   // (precision, rows, block rows, block K)
  H16Item* d_xitems = nullptr;
  size_t xitems_cap = 0;
  std::vector<H16Item> h_xitems;   // the work list on the device, as uploaded
  // autotuned ERI GEMM tiles: (precision, rows, cols, K) -> (config, split-K)
  int autotune = 1;
  std::map<std::tuple<int, long, int, int>, std::pair<int, int>> tuned;
  std::string tune_log;
  // projected SD-space overlap grid (phf_sc_set_grid): per point g, column-major
  // R (N x N), Y = R C0 M^-1 (N x NOcc), B^T = R^H Y (N x NOcc), K = R^H Y R,
  // A = (I - Y) R; w_g and sum_g w_g; sc_ws is the workspace of phf_sc_apply*
  int sc_G = 0, sc_nso = 0, sc_nocc = 0;
  double *sc_R = nullptr, *sc_Y = nullptr, *sc_BT = nullptr, *sc_K = nullptr, *sc_A = nullptr;
  double *sc_w = nullptr, *sc_ovlp = nullptr;
  char* sc_ws = nullptr;
  size_t sc_ws_size = 0;
  // per-stage timing with HIP events on `stream` (phf_set_profiling)
  bool profiling = false;
  // graph replay of phf_eval_objective: one evaluation's launch sequence
  // (C0 upload from pinned memory through the copies back), captured after an
  // eager evaluation and replayed while og_key (shapes, buffers, generation,
  // the tile-pinning environment) is unchanged; gen is bumped by every setter
  // that changes what an evaluation launches
  hipGraph_t og_graph = nullptr;
  hipGraphExec_t og_exec = nullptr;
  std::string og_key, og_failed_key;
  long gen = 0;
  hipEvent_t ev[8] = {};
  double stage_ms[11] = {};
  int nstage = 0;
  // per spin block of the ERI GEMM (D, X1, X2; [q][0] before, [q][1] after
  // its launch on the stream it runs on) when profiling; nblk_timed blocks
  // were recorded by the last evaluation (X1 holds the grouped X1 + X2
  // launch when the exchange blocks run as one)
  hipEvent_t bev[3][2] = {};
  int nblk_timed = 0;
  bool rjk_grouped_last = false;   // the spin-free products ran as one grouped launch (bev[1] is empty)
  // workspace of phf_pcc_olap0 (the projected-CC grid consumer)
  char* pcc_ws = nullptr;
  size_t pcc_ws_size = 0;
};

inline long round_up(long x, long m) { return (x + m - 1) / m * m; }

// Events of the library are read only after a synchronization of their stream
// (timing) or waited on by another stream of the same device (fork/join): a
// device-scope release suffices.  The default system-scope fence writes back
// and invalidates the caches at every record, which showed as ~10 us gaps
// between kernels around each recorded event.
constexpr unsigned kTimingEventFlags = hipEventDisableSystemFence;
constexpr unsigned kSyncEventFlags = hipEventDisableTiming | hipEventDisableSystemFence;

// Bump allocator over a device block (base == nullptr measures).
struct Arena {
  char* base;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {   // nullptr for an empty buffer
    if (count == 0) return nullptr;
    off = (off + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(base ? base + off : nullptr);
    off += count * sizeof(T);
    return p;
  }
};

int phf_fail(phf_ctx* ctx, int code, const std::string& msg);
void phf_free_sc(phf_ctx* ctx);   // the projected SD-space overlap grid and workspace

#define HIPCK(expr)                                                                   \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return phf_fail(ctx, e_ == hipErrorOutOfMemory ? PHF_ERR_NOMEM : PHF_ERR_HIP,   \
                      std::string(#expr) + ": " + hipGetErrorString(e_));             \
  } while (0)

void fill_zero(hipStream_t s, double* p, long n);
// Wigner D^j_{nm}(alpha, beta, gamma) by the factorial sum (PHFTools.f90:350-376)
std::complex<double> phf_wignerfac(int j, int n, int m, double alpha, double beta, double gama);
int phf_ctx_init(phf_ctx* ctx, int device, hipStream_t shared_stream);
int phf_release_ws(phf_ctx* ctx);
int phf_keep_only_mixed(phf_ctx* ctx, int prec);
// Keep only the FP64 copy of the structure evaluations use (the spin-free form
// when built and on, else the spin blocks).
int phf_keep_used_structure(phf_ctx* ctx);
// Build the spin blocks of the cached real A for MO spins `spin` (size nso),
// after checking on the device that every entry outside them is exactly zero;
// returns PHF_OK without building when the check fails.  release_dense: free
// the dense A afterwards.  With ref (row-major complex, the Ref the
// Hamiltonian is expressed in) the spin-free form is built too when Ref is
// spin-restricted and A has that structure.
int phf_build_blocks(phf_ctx* ctx, const std::vector<int>& spin, bool release_dense, const double* ref = nullptr);
// MO spins of a real or complex Ref (nso x nso, row-major complex): 0 for a
// column with exactly zero beta rows, 1 for exactly zero alpha rows; empty if
// some column mixes spins.
std::vector<int> ref_spins(const double* ref, int nso);

// Device buffers of one evaluation (carved from the context's workspace).
struct EvalBufs {
  int N = 0, prec = 0, bm = 0, gchunk = 32, nchunk = 0;
  int V = 1;                  // density variants per grid point (4 for CmplxConj = 2)
  long Mp = 0, G = 0;
  double *C0 = nullptr, *Wpg = nullptr, *W2 = nullptr, *T = nullptr, *M = nullptr, *Minv = nullptr, *Det = nullptr;
  int* Info = nullptr;
  // FX holds (h + Gamma) X, then the adjoint Y = FX - X (Q - e I) in place
  double *X = nullptr, *Pt = nullptr, *Gam = nullptr, *FX = nullptr, *Q = nullptr, *Hc = nullptr, *E = nullptr;
  double *Wd = nullptr, *Part = nullptr, *Ysum = nullptr, *Xsum = nullptr, *HS = nullptr;
  double *Grad = nullptr, *Rdm = nullptr;
  void* Pm = nullptr;
  double* Pinv = nullptr;
  double* Amax = nullptr;     // per density row, its largest magnitude (bit patterns; reduced forms)
  bool blk = false;           // the GEMM runs over the spin blocks
  long kP = 0;                // row length (per slice) of the GEMM's density operand
  double* Pperm = nullptr;    // FP64 density rows in permuted column order (blk, FP64 form)
  // spin-free form: four operand row sets of Mp rows each (Mk = 3 Mp), set s
  // at rows s Mp: P_aa + P_bb, P_aa - P_bb, P_x1, P_x2 of every density (2 Gv
  // data rows, then zero padding), kP = rj_kj columns; Cjk the products in the
  // same row order (FP64), unpacked into Gam by rjk_unpack
  bool rjk = false;
  long Mk = 0;
  double* Cjk = nullptr;      // shares its memory with Pt (dead once the operand is packed)
  bool fused = false;         // the density pack writes the reduced operand directly (no Pt)
};

// Stages of one evaluation on ctx->stream.  eval_plan checks the cached state,
// sizes and carves the workspace (may allocate; not for graph capture).
// eval_pre runs K1/K2, the density pack and the reduced-precision split of
// the density rows once C0 = [I; Z] is in b.C0; eval_gemm the single ERI GEMM
// (autotuning a new 16-bit shape first, not for capture); eval_post K4/K5 and,
// unless resident, Gradmat/Rdmmat.  eval_pre/eval_post only launch kernels
// (eval_pre records marks[0..2] after the rotation chain, the X_g solve and the
// density pack when marks is given).
int eval_plan(phf_ctx* ctx, int nocc, int precision, int cmplxconj, EvalBufs& b);
void eval_pre(phf_ctx* ctx, const EvalBufs& b, hipEvent_t* marks = nullptr);
int eval_gemm(phf_ctx* ctx, const EvalBufs& b);
// Work list of a grouped 16-bit ERI GEMM over (context, buffers) units.
// blk_mask selects the spin blocks of blocked units (bit q: block q).
void build_h16_items(const std::vector<std::pair<const phf_ctx*, const EvalBufs*>>& units, int prec, int cfg,
                     int sk, int grid, std::vector<H16Item>& out, unsigned blk_mask = 7u);
// Order of a logical work list for a persistent launch of `grid` workgroups
// (each XCD takes a contiguous run of items per round).
void h16_xcd_order(const std::vector<H16Item>& logical, int grid, std::vector<H16Item>& out);
// True when build_h16_items makes the spin-free units' exchange sets write
// Gamma_ab, Gamma_ba + h directly (split-K 1 and at least one spin-free unit).
bool rjk_direct_items(const std::vector<std::pair<const phf_ctx*, const EvalBufs*>>& units, int sk);
void eval_post(phf_ctx* ctx, const EvalBufs& b, bool resident);
// Spin-free form: Gamma + h (b.Gam, natural column order) from the products in
// b.Cjk; eval_gemm runs it after its launches (the batch after its grouped one).
// all = false: Gamma_aa and Gamma_bb only (a grouped launch with split-K 1
// wrote Gamma_ab and Gamma_ba itself).
void rjk_unpack(phf_ctx* ctx, const EvalBufs& b, hipStream_t s, bool all);
// C0 = [I; Z] column-major complex (NSO x N) from Z ((NSO-N) x N, C order)
void build_c0(const double* z, int nso, int N, int cmplxconj, double* c0);
