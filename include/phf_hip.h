/*
 * phf_hip.h -- C ABI of the MI355X spin-projection grid kernel (libphf_hip.so).
 *
 * Drop-in replacement for the f2py module PHFTools_F built from
 * PHF_fort/PHFTools.f90 (PHF_fort/makefile:7) and called from PHF.py:70-76.
 * Every entry point returns PHF_OK (0) or one of the error codes below; it
 * never terminates the process (the reference STOPs, e.g. Wrappers.f90:311,
 * PHFTools.f90:199).  phf_last_error() gives the message of the last failure.
 *
 * Array conventions: plain host pointers; complex arrays are complex128
 * (interleaved re, im doubles); multi-index arrays are C (row-major) order,
 * i.e. exactly numpy's default layout of the arrays the reference passes.
 * Element (i,j,k,l) of an [a][b][c][d] array sits at ((i*b+j)*c+k)*d+l.
 * Index meanings are those of the Fortran dummies of the same name.
 */
#ifndef PHF_HIP_H
#define PHF_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phf_ctx_s phf_ctx;

enum {
  PHF_OK = 0,
  PHF_ERR_ARG = 1,            /* bad argument / shape (f2py raises ValueError) */
  PHF_ERR_HIP = 2,            /* HIP runtime failure */
  PHF_ERR_SINGULAR = 3,       /* exactly singular overlap at a grid point (Wrappers.f90:311) */
  PHF_ERR_NOT_HERMITIAN = 4,  /* Hmat/Smat fail the Hermiticity test (PHFTools.f90:195-200) */
  PHF_ERR_NO_HAMILTONIAN = 5, /* phf_buildhsg before phf_set_hamiltonian */
  PHF_ERR_UNSUPPORTED = 6,    /* option not implemented on this path */
  PHF_ERR_NOMEM = 7           /* device allocation failed */
};

/* Flags of phf_set_hamiltonian / phf_set_hamiltonian_factored.
 * PHF_HAM_COMPLEX      treat A as complex even if its imaginary part is zero
 *                      (dense entry point only).
 * PHF_HAM_DENSE        never use the spin-block structure: every evaluation
 *                      contracts the dense NSO^2 x NSO^2 A (MakeGamma as the
 *                      reference does it, PHFTools.f90:310-337).
 * PHF_HAM_BLOCKS_ONLY  when the structure is found, keep only the blocks and
 *                      release the dense A (factored entry point).
 * Without PHF_HAM_DENSE the library checks, once Ref is known (the factored
 * builder's Ref, or phf_set_grid's for the dense entry point), whether every
 * MO is purely alpha or beta and every entry of A outside the spin blocks is
 * exactly zero; then it stores A as three diagonal blocks (6 n^4 of the 16 n^4
 * entries for n alpha and n beta MOs) and evaluates through them.  The result
 * is the same contraction minus exact zeros. */
enum {
  PHF_HAM_COMPLEX = 1,
  PHF_HAM_DENSE = 2,
  PHF_HAM_BLOCKS_ONLY = 4
};

/* Precision of the ERI contraction (the overlap LU, determinant, densities,
 * the Fock contraction and the grid sum are always FP64).  The reduced forms
 * run on the FP32 / BF16 / FP16 matrix cores with FP32 accumulation folded
 * into FP64:
 *   PHF_PREC_FP32    one product of the FP32 roundings of A and rho
 *   PHF_PREC_BF16X3  A, rho split into three BF16 terms; six products
 *   PHF_PREC_F16X2   A, rho scaled by powers of two per row and split into
 *                    two FP16 terms; three products (hi*hi + hi*lo + lo*hi)
 *   PHF_PREC_I8X4    A, rho scaled by powers of two per row and split into
 *                    four int8 terms (28 bits); the ten products with
 *                    i + j <= 3 on the int8 matrix cores with exact int32
 *                    accumulation, folded into FP64 with weights 2^-7(i+j).
 *                    No rounding inside the matrix cores: the 16-bit cores
 *                    truncate their sums, which biases BF16X3/F16X2 at large
 *                    NSO. */
enum {
  PHF_PREC_FP64 = 0,
  PHF_PREC_FP32 = 1,
  PHF_PREC_BF16X3 = 2,
  PHF_PREC_F16X2 = 3,
  PHF_PREC_I8X4 = 4
};

/* Context: one GPU, one HIP stream, a cached device Hamiltonian and a
 * workspace arena.  No counterpart in the reference (its state lives in
 * Fortran locals). */
int phf_ctx_create(int device, phf_ctx** out);
int phf_ctx_destroy(phf_ctx* ctx);
/* Message of the last failure on ctx (or of the last failed phf_ctx_create
 * when ctx is NULL). */
const char* phf_last_error(const phf_ctx* ctx);

/* Upload the MO-basis Hamiltonian that BuildHSG receives as HOne / HTwo
 * (PHFTools.f90:17): hone complex [nso][nso], htwo complex [nso]^4 with
 * htwo[i][j][k][l] = <ij||kl>.  The device keeps A[(i,k),(j,l)] = V(i,j,k,l)
 * (real plane only when every imaginary part is exactly zero) until the next
 * call.  flags: PHF_HAM_* (bit 0 forces the complex form). */
int phf_set_hamiltonian(phf_ctx* ctx, const double* hone, const double* htwo, int nso, int flags);

/* Device-side construction of the same A from a factored spatial Hamiltonian,
 * for problems whose HTwo does not fit the host interface (complex128 HTwo
 * is 16 NSO^4 bytes: 69 GB at NSO = 256).  It reproduces the reference's
 * input pipeline (Main_PHF.py:22-33: spin-orbital blocks, Mulliken2Dirac
 * makeH.py:63-67, ao2mo ao2mo.py:4-10) for (pq|rs) = sum_L b[L][p][q] b[L][r][s]:
 *   hone = Ref^H (h (x) I2) Ref,  Bt_L = Ref^H (b_L (x) I2) Ref,
 *   <ij||kl> = sum_L Bt_L(i,k) Bt_L(j,l) - Bt_L(i,l) Bt_L(j,k).
 * h: double [n][n]; b: double [nl][n][n]; ref: complex [2n][2n] with zero
 * imaginary part (a complex Ref returns PHF_ERR_UNSUPPORTED).  flags: PHF_HAM_*. */
int phf_set_hamiltonian_factored(phf_ctx* ctx, const double* h, const double* b, int n, int nl,
                                 const double* ref, int flags);
/* Copy the device's MO-basis hone (complex [nso][nso]) to the host. */
int phf_get_hone(phf_ctx* ctx, double* hone);

/* phftools.buildhsg (PHFTools.f90:10-201; call site PHF.py:70-71).
 *   ref      complex [nso][nso]         Ref (MO coefficients, AO rows)
 *   z        complex [nso-nocc][nocc]   Thouless Z
 *   ngrid    int[2]                     (nl, ny)
 *   r1       complex [nr1][nso][nso], r2 complex [nr2][nso][nso],
 *   rpg      complex [npg][nso][nso]
 *   weightsp double [nl][ny]; weightpg double [npg][ncipg][ncipg]
 *   roota/rootb/rooty double[na]/[nb]/[ny_] (Wigner angles)
 *   qj, sp   J and spin-projection type (SP=2 swaps the angle indices)
 *   precision  PHF_PREC_*
 * Outputs (caller-allocated): hmat, smat complex [nci][nci];
 *   gradmat, rdmmat complex [nci][nci][nso][nso].
 * The MPI communicator argument of the reference (PHFTools.f90:41) has no
 * counterpart: one context is one GPU. */
int phf_buildhsg(phf_ctx* ctx, const double* ref, const double* z, const int* ngrid, int nci,
                 int cmplxconj, int ncisp, int nocc, const double* r1, int nr1, const double* r2,
                 int nr2, const double* rpg, int npg, const double* weightsp,
                 const double* weightpg, int ncipg, const double* roota, int na,
                 const double* rootb, int nb, const double* rooty, int ny_, int qj, int sp,
                 int precision, double* hmat, double* smat, double* gradmat, double* rdmmat);

/* The two halves of phf_buildhsg, for callers that evaluate many Z on one
 * grid (the optimiser loop of PHF.py:17-53 calls buildhsg with the same R1,
 * R2, Rpg, weights and Ref every time).  phf_set_grid uploads the grid and
 * Ref, forms Ref^H R1[il], Rpg[ipg] Ref and the weight x Wigner x weightpg
 * table on the device (arguments as phf_buildhsg); phf_buildhsg_z evaluates
 * one Z on the current grid.  No counterpart in the reference: its Fortran
 * recomputes everything per call. */
int phf_set_grid(phf_ctx* ctx, const double* ref, const int* ngrid, int nci, int ncisp, const double* r1,
                 int nr1, const double* r2, int nr2, const double* rpg, int npg, const double* weightsp,
                 const double* weightpg, int ncipg, const double* roota, int na, const double* rootb, int nb,
                 const double* rooty, int ny_, int qj, int sp);
int phf_buildhsg_z(phf_ctx* ctx, const double* z, int cmplxconj, int nocc, int precision, double* hmat,
                   double* smat, double* gradmat, double* rdmmat);

/* Device-resident objective (EandG, PHF.py:55-91, without moving Gradmat and
 * Rdmmat): phf_eval_resident evaluates Z like phf_buildhsg_z but returns only
 * hmat/smat and keeps the grid sums on the device; after the host solves the
 * nci x nci problem (EigenSolver, VecDecomp), phf_localg_vo returns the
 * (NSO-N) x N block G[N:, :N] of LocalG (gvo: complex [nso-nocc][nocc]),
 * which is the gradient dE/dZ*.  Valid until the next call that uses the
 * context.  phf_localg_vo does not take CmplxConj = 2 (use phf_buildhsg_z +
 * phf_localg there). */
int phf_eval_resident(phf_ctx* ctx, const double* z, int cmplxconj, int nocc, int precision, double* hmat,
                      double* smat);
/* One projected energy + gradient evaluation in one call when the grid has
 * nci = 1 (J = 0, one point-group irrep) and CmplxConj is 0 or 1: the 1 x 1
 * EigenSolver (PHFTools.py:8-18: E0 = Re H / |Re S|), VecDecomp and LocalG
 * run on the device after the grid sums, and only E0 and the gradient block
 * dE/dZ* = LocalG[N:, :N] (complex [nso-nocc][nocc]) cross PCIe.  The same
 * errors as phf_eval_resident, plus PHF_ERR_SINGULAR when |Re S| <= 1e-6 (the
 * EigenSolver would keep no state).  The sums stay resident as after
 * phf_eval_resident.  No counterpart in the reference (EandG, PHF.py:55-91,
 * does these steps on the host). */
int phf_eval_objective(phf_ctx* ctx, const double* z, int cmplxconj, int nocc, int precision, double* e0,
                       double* grad);
int phf_localg_vo(phf_ctx* ctx, const double* smat, const double* e0, const double* fsp, int ncisp,
                  const double* fpg, int ncipg, const double* fk, int nfk, int cmplxconj, double* gvo);

/* phftools.localg (PHFTools.f90:203-247; call site PHF.py:76).
 *   gradmat, rdmmat complex [nci][nci][nso][nso]; smat complex [nci][nci];
 *   e0 complex scalar (2 doubles); fsp complex[ncisp]; fpg complex[ncipg];
 *   fk complex[nfk] (nfk >= 2 when cmplxconj == 2).
 * Output g: complex [nso][nso]. */
int phf_localg(phf_ctx* ctx, const double* gradmat, const double* rdmmat, const double* smat,
               const double* e0, const double* fsp, int ncisp, const double* fpg, int ncipg,
               const double* fk, int nfk, int cmplxconj, int nso, int nci, double* g);

/* Stage timing with HIP events on the context's stream (no counterpart in the
 * reference, which only prints wall-clock times, Main_PHF.py:15,40,44).
 * After a profiled phf_buildhsg, phf_stage_times fills up to nmax values in
 * milliseconds: [0] all device work of the call, [1] rotation + overlap +
 * inverse + density pack + reduced-precision split, [2] ERI GEMM, [3]
 * epilogue + grid sums, [4] rotation chain, [5] overlap + LU inverse + X_g,
 * [6] density pack, [7] reduced-precision split (when the pack writes the
 * 16-bit/int8 slices itself, NSO <= 128, [6] covers both and [7] is ~0),
 * [8], [9], [10] the ERI GEMM launch of each spin block (D, X1, X2; [9]
 * holds X1 + X2 when the exchange blocks run as one grouped launch, and
 * [8] the whole GEMM for a dense A); in the spin-free form [8] the (2J - K)
 * launch or the grouped launch of both products, [9] the K launch (0 when
 * grouped), [10] the unpack into Gamma; it returns how many it wrote.
 * After a profiled phf_sc_apply / phf_sc_apply_packed(_dev): [0] all device
 * work, [1] stage-1 pair-transform GEMM, [2] stage-2 GEMM with the grid sum,
 * [3] P GEMM, [4] rank-term GEMM, [5] per-point t0/t1 kernel, [6] the rest
 * (factors, packing, outputs), [7] the number of GEMM launches.
 * After a profiled phf_pcc_olap0 / phf_pcc_olap: [0] all device work after
 * the uploads, [1] the zeroth-order per-point stages (rotations through
 * TransT2 and CCEnergy), [2] the CCSD residuals, [3] the particle-particle
 * ladder GEMM launches (1/2 <ab||ef> tau_efij), [4] the singles/doubles grid
 * sums, [5] the number of ladder launches (one per chunk of points). */
int phf_set_profiling(phf_ctx* ctx, int on);
int phf_stage_times(const phf_ctx* ctx, double* ms, int nmax);

/* Spin-block evaluation switch (no counterpart in the reference).
 * phf_set_blocked(ctx, 0) evaluates through the dense A although blocks exist
 * (fails if the dense A was released); phf_set_blocked(ctx, 1) returns to the
 * blocks.  phf_blocked reports 1 when evaluations run over the blocks. */
int phf_set_blocked(phf_ctx* ctx, int on);
int phf_blocked(const phf_ctx* ctx);

/* Spin-free evaluation switch (no counterpart in the reference).  For a
 * spin-restricted Ref (every alpha MO has a beta MO with the same spatial
 * coefficients) whose A has the matching structure (checked on the device to
 * 1e-12 of its largest entry), the blocks are D = [[J-K, J], [J, J-K]] and
 * X1 = X2 = -K over spatial MO pairs, and Gamma comes from (2J - K) applied to
 * rho_aa + rho_bb and K applied to rho_aa - rho_bb, rho_ab and rho_ba: four
 * n^2 x n^2 products per density instead of six.  Built with the blocks;
 * used by default (PHF_SPINFREE=0 at context creation, or
 * phf_set_spinfree(ctx, 0), keeps the blocks).  phf_spinfree reports 1 when
 * evaluations run through it. */
int phf_set_spinfree(phf_ctx* ctx, int on);
int phf_spinfree(const phf_ctx* ctx);

/* Tile autotuner of the reduced-precision ERI GEMM (no counterpart in the
 * reference).  on = 1 (default): the first call for a problem shape times
 * every tile/split-K candidate on the live operands and keeps the fastest;
 * on = 0: a cost model of MFMA rounds per CU decides.  Either call clears the
 * cache.  phf_tuning_report copies one JSON line per tuned shape (candidates,
 * times, choice) into buf and returns the length written. */
int phf_autotune(phf_ctx* ctx, int on);
int phf_tuning_report(const phf_ctx* ctx, char* buf, int len);

/* Projected SD-space overlap times vector for LinDep.
 *
 * phf_buildsc replaces phftools.buildsc (BuildSC, PHFTools.f90:464-537, with
 * BuildSC_Vec, BuildSC_dr.f90:8-1244; called at LinDep.py:73):
 *   t = sum_g w_g BuildSC_Vec(c0, R c1 R^H, ao2mo4(c2, R^H), Y_g) / sum_g w_g,
 *   R = R1[il] R2[iy] Rpg[ipg],  Y_g, s_g from RDM with the first nocc unit
 *   vectors as newMO,  w_g = weightsp(il,iy) s_g sum_{m,n,p,q} w,
 * where w is the reference's REAL weightpg(ipg,p,q) conj(fsp_m) fsp_n
 * conj(fpg_p) fpg_q (each complex product keeps its real part, :518-520;
 * the Wigner factor is one, :515).  Rk, fk, ncik, roota/b, rooty of the
 * reference signature are not used by it and not taken here.
 * Host arrays, numpy C order: c0 complex scalar; c1 complex[nso][nso];
 * c2 complex[nso]^4; r1 complex[nr1][nso][nso], r2 complex[nr2][nso][nso],
 * rpg complex[npg][nso][nso]; weightsp double[ngrid[0]][ngrid[1]]; weightpg
 * double[npg][ncipg][ncipg]; fsp complex[ncisp] (ncisp >= 2 qj + 1), fpg
 * complex[ncipg].  Outputs t0 (1), t1 (nso x nso), t2 (nso^4) complex,
 * caller-allocated.  1 <= nocc <= nso <= 64.  PHF_ERR_SINGULAR when
 * C0^H R C0 is exactly singular at a point (the reference STOPs in InvertC,
 * Wrappers.f90:311).
 *
 * phf_sc_set_grid + phf_sc_apply split the same call so the grid factors
 * stay on the device across the many products of a LinDep solve (no
 * counterpart in the reference, which rebuilds them per call): phf_sc_apply
 * takes nvec vectors at once (c0 complex[nvec], c1 complex[nvec][nso][nso],
 * c2 complex[nvec][nso]^4, outputs alike).
 * phf_sc_apply_packed takes nvec packed SD vectors (complex[nvec][dim],
 * dim = phf_sd_dim(nso, nocc) = 1 + O V + O(O-1)/2 V(V-1)/2) in the order of
 * Vec2Mat (LinDep.py:156-171) and returns Mat2Vec (LinDep.py:173-187) of the
 * product: LinDep.py:71-76's BuildSC for a block of vectors in one call.
 * phf_sc_apply_packed_dev is the same with device pointers (no copies). */
int phf_buildsc(phf_ctx* ctx, const double* c0, const double* c1, const double* c2, int nocc, int nso,
                const int* ngrid, int qj, const double* r1, int nr1, const double* r2, int nr2,
                const double* rpg, int npg, const double* weightsp, const double* weightpg, int ncipg,
                const double* fsp, int ncisp, const double* fpg, double* t0, double* t1, double* t2);
int phf_sc_set_grid(phf_ctx* ctx, int nso, int nocc, const int* ngrid, int qj, const double* r1, int nr1,
                    const double* r2, int nr2, const double* rpg, int npg, const double* weightsp,
                    const double* weightpg, int ncipg, const double* fsp, int ncisp, const double* fpg);
int phf_sc_apply(phf_ctx* ctx, int nvec, const double* c0, const double* c1, const double* c2, double* t0,
                 double* t1, double* t2);
int phf_sc_apply_packed(phf_ctx* ctx, int nvec, const double* vin, double* vout);
int phf_sc_apply_packed_dev(phf_ctx* ctx, int nvec, const double* vin, double* vout);
long phf_sd_dim(int nso, int nocc);

/* Batched InvertC + determinantC (Wrappers.f90:295-335, 195-231) on the
 * kernel that inverts the grid's overlaps M_g: m is complex[batch][n][n]
 * (element (b,i,j) = Matrix_b(i,j)), n <= 96.  Partial pivoting on the
 * largest |Re|+|Im| of the column, first index on ties (IZAMAX);
 * det[b] = product of the pivots with a sign flip per row interchange;
 * info[b] = the first exactly-zero pivot (1-based), 0 otherwise, in which
 * case minv[b] and det[b] are zero.  A singular matrix is not an error here
 * (the grid stages turn info != 0 into PHF_ERR_SINGULAR). */
int phf_invert_batched(phf_ctx* ctx, int n, int batch, const double* m, double* minv, double* det, int* info);
/* phf_invert_batched for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments and results
 * (bit for bit) as phf_invert_batched; it calls no hipMalloc/hipFree: the
 * input, its column-major copy, the inverse, det and info are staged in a
 * file-scope device array of 64 MiB per GPU (held until the process ends).
 * Limit: 2 * 16 * n*n*batch + 20 * batch bytes (each of the four regions
 * rounded up to 256) must fit 64 MiB, e.g. n = 32: 32,000 matrices, n = 96:
 * 3,600; larger calls return PHF_ERR_UNSUPPORTED before any device work.
 * Every call on one GPU shares that GPU's array: calls that overlap in time
 * on one GPU (two contexts, two threads) race and must be serialised by the
 * caller.  Each call synchronises the context's stream before it returns. */
int phf_invert_batched_noalloc(phf_ctx* ctx, int n, int batch, const double* m, double* minv, double* det,
                               int* info);
/* Projected-CC grid consumer: Olap0 and EOlap0 of PbarHbarOlap
 * (SUCCSD_Boson_cmplx/FPUCC.f90:14-272, argument meaning as there) with the
 * lambda0 / eps0 part of EvalKernel (FPUCC.f90:529-602, TrunODE = 1):
 *   hone complex[nso][nso], htwo complex[nso]^4 (<pq||rs>, MO basis);
 *   t1 complex[nso-nocc][nocc] (T1(NOcc+1+a, 1+i)); t2 complex[v][v][o][o];
 *   npoints[2] = NPoints; r1 complex[nr1][nso][nso], r2 complex[nr2][nso][nso],
 *   rpg complex[npg][nso][nso], rk complex[ncik][nso][nso] (read when ncik = 2),
 *   all in the MO basis (Main_PCC.py:37-42); roota[nra], rootb[nrb],
 *   rooty[nry]; weightsp[npoints0][npoints1]; weightpg[npg][ncipg][ncipg];
 *   fsp complex[ncisp] (ncisp = 2 qj + 1), fpg complex[ncipg], fk complex[ncik].
 * Outputs olap0, eolap0: one complex each.  nso <= 64, at most 32 occupied and
 * 32 virtual orbitals.  Only the zeroth-order outputs (phf_pcc_olap gives
 * every output).  A singular occupied block of R e^{T1} (InvertC's STOP)
 * returns PHF_ERR_SINGULAR. */
int phf_pcc_olap0(phf_ctx* ctx, const double* hone, const double* htwo, const double* t1, const double* t2,
                  int nocc, int nso, double enuc, int qj, int sp, const int* npoints, int npg, int ncisp,
                  int ncipg, int ncik, const double* r1, int nr1, const double* r2, int nr2, const double* rpg,
                  const double* rk, const double* roota, int nra, const double* rootb, int nrb,
                  const double* rooty, int nry, const double* weightsp, const double* weightpg,
                  const double* fsp, const double* fpg, const double* fk, double* olap0, double* eolap0);
/* Every output of PbarHbarOlap (SUCCSD_Boson_cmplx/FPUCC.f90:14-272; inputs as
 * phf_pcc_olap0), TrunODE = 1: Olap0, OlapEx1, OlapEx2, EOlap0, EOlapEx1,
 * EOlapEx2 in the Fortran argument order, with EvalKernel's singles and
 * doubles (FPUCC.f90:529-602): the CCSD residual CCRes12 of every point's
 * rotated amplitudes (CCRes.f90:164-536), BuildKernSD's fills
 * (FPUCC_Tools.f90:291-436) and, for sp = 2, ProjSz (FPUCC_Tools.f90:150-233).
 * olapex1 / eolapex1 complex[nso-nocc][nocc] (OlapEx1(NOcc+1+a, 1+i));
 * olapex2 / eolapex2 complex[v][v][o][o] (OlapEx2(NOcc+1+a, NOcc+1+b, 1+i, 1+j)). */
int phf_pcc_olap(phf_ctx* ctx, const double* hone, const double* htwo, const double* t1, const double* t2,
                 int nocc, int nso, double enuc, int qj, int sp, const int* npoints, int npg, int ncisp,
                 int ncipg, int ncik, const double* r1, int nr1, const double* r2, int nr2, const double* rpg,
                 const double* rk, const double* roota, int nra, const double* rootb, int nrb,
                 const double* rooty, int nry, const double* weightsp, const double* weightpg,
                 const double* fsp, const double* fpg, const double* fk, double* olap0, double* olapex1,
                 double* olapex2, double* eolap0, double* eolapex1, double* eolapex2);
/* buildsc_dr.buildsc_vec (BuildSC_Vec, BuildSC_dr.f90:8-1244) and any other
 * weighted sum over given points: replaces the grid of phf_sc_set_grid with
 * npts points of rotation r (complex[npts][nso][nso], NULL for identities),
 * transition density y (complex[npts][nso][nso], every column used) and
 * weight w (complex[npts], NULL for ones); phf_sc_apply then returns
 * sum_g w_g BuildSC_Vec(c0, R c1 R^H, ao2mo4(c2, R^H), Y_g) / sum_g w_g.
 * With npts = 1, r = w = NULL it is BuildSC_Vec(c0, c1, c2, Y) itself. */
int phf_sc_set_points(phf_ctx* ctx, int nso, int npts, const double* r, const double* y, const double* w);

/* Batched evaluation of many independent instances (no counterpart in the
 * reference, which evaluates one Hamiltonian per process; each instance is
 * one EandG of PHF.py:55-91 at J = 0).  Instances may differ in size, N and
 * grid.  phf_batch_set_instance builds instance i's A on the device from its
 * factors (as phf_set_hamiltonian_factored) and uploads its J = 0 spin grid
 * (phf_set_grid arguments with nci = ncisp = ncipg = 1; weightpg may be NULL
 * for ones); storage = PHF_PREC_FP64 keeps the FP64 A, a reduced-precision
 * mode keeps only that mode's copy of A (the instance then evaluates in that
 * mode only); flags = PHF_HAM_DENSE keeps the dense A (plain contraction),
 * otherwise a spin-block structured A is stored as its blocks only.  phf_batch_eval takes the Z of every instance concatenated in
 * instance order (complex C order (2n - nocc) x nocc each) and returns per
 * instance the projected energy E0 (energies[i]), optionally the 1 x 1 Hmat,
 * Smat (hs[4 i .. 4 i + 3] = Re H, Im H, Re S, Im S) and the gradient block
 * dE/dZ* = LocalG[N:, :N] (grads, laid out like z).  An instance with a
 * singular overlap, |Re S| <= 1e-6 or a failed Hermiticity test
 * (PHFTools.f90:195-200) fails: every instance's outputs are still written,
 * status[i] (optional) receives each instance's code and the call returns
 * the first failure's code.  One evaluation is one
 * launch sequence: every instance's rotation/overlap/density stages, one
 * persistent grouped ERI GEMM over all instances (16-bit modes; FP64 and
 * FP32 launch per instance), every instance's epilogue and grid sums.  The
 * first evaluation of a precision autotunes the grouped tiles and records
 * the sequence in a hipGraph that later evaluations replay
 * (phf_batch_set_graph(b, 0) turns that off).  phf_batch_stats: [0] device
 * time of the last evaluation in ms, [1] its ERI GEMM time (-1 for a graph
 * replay), [2] 1 if it was a graph replay, [3] the number of instances that
 * evaluate through the spin-free form (phf_spinfree). */
typedef struct phf_batch_s phf_batch;
int phf_batch_create(int device, int ninst, phf_batch** out);
int phf_batch_destroy(phf_batch* b);
const char* phf_batch_last_error(const phf_batch* b);
int phf_batch_set_instance(phf_batch* b, int i, const double* h, const double* bfac, int n, int nl,
                           const double* ref, int nocc, const int* ngrid, const double* r1, int nr1,
                           const double* r2, int nr2, const double* rpg, int npg, const double* weightsp,
                           const double* weightpg, int storage, int flags);
int phf_batch_eval(phf_batch* b, const double* z, int cmplxconj, int precision, double* energies,
                   double* grads, double* hs, int* status);
int phf_batch_set_graph(phf_batch* b, int on);
int phf_batch_autotune(phf_batch* b, int on);
int phf_batch_stats(const phf_batch* b, double* ms, int nmax);
int phf_batch_tuning_report(const phf_batch* b, char* buf, int len);

#ifdef __cplusplus
}
#endif
#endif /* PHF_HIP_H */
