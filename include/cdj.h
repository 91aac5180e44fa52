/* cdj.h - C ABI of the MI355X (gfx950) triangular-dislocation engine.
 *
 * Drop-in replacement for the compute path of CutAndDisplaceJulia: TD (src/TD.jl), CalculateInfluenceMatrices3D
 * (src/CalculateInfluenceMatrices3D.jl) and SlipCalculator3D (src/SlipCalculator3D.jl).  Host Julia binds these
 * symbols with ccall; the Python mirror in cutanddisplacejulia_amd/api.py binds them with ctypes.
 *
 * Conventions (those of the reference's Julia arrays):
 *   - all host pointers are caller-owned; the library never keeps them after a call returns;
 *   - P1, P2, P3, normals and midpoints are N x 3, column-major (coordinate k of triangle j at [k*N + j]);
 *   - influence matrices are nobs x ntri column-major (receiver i, source j at [i + j*nobs]);
 *   - the collocation matrix A is 3N x 3N column-major with row blocks (Tn, Tss, Tds) and column blocks
 *     (Dn, Dss, Dds), A = -[DnTn DssTn DdsTn; DnTss DssTss DdsTss; DnTds DssTds DdsTds];
 *   - every function returns a status code (CDJ_OK == 0); cdj_last_error gives the message, using the
 *     reference's own error text where the reference raises one;
 *   - calls on one context are synchronous and must not overlap.
 *   - `_dev` entry points take device pointers (HBM-resident data) and a hipStream_t passed as void*.
 */
#ifndef CDJ_H
#define CDJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CDJ_OK 0
#define CDJ_ERR_DIMS 1           /* "Element size inputs must be the same dimensions" (TD.jl:424) */
#define CDJ_ERR_HALFSPACE 2      /* "Half-space solution: Z coordinates must be negative!" (TD.jl:617) */
#define CDJ_ERR_DEGENERATE 3     /* zero-length triangle edge (CreateMidPoint!.jl:41-43) */
#define CDJ_ERR_DEVICE 4         /* HIP runtime error */
#define CDJ_ERR_NOMEM 5          /* device allocation failed */
#define CDJ_ERR_NOT_CONVERGED 6  /* iterative refinement did not reach the tolerance; fallback result given */
#define CDJ_ERR_ARG 7            /* invalid argument (null pointer, bad option) */
#define CDJ_ERR_SINGULAR 8       /* exactly singular matrix (zero pivot) */

/* precision of the traction-matrix assembly and of the factorisation */
#define CDJ_PREC_FP64 0       /* FP64 assembly, FP64 LU (the reference's arithmetic) */
#define CDJ_PREC_MIXED32 1    /* FP32 LU on MFMA as preconditioner, FP64 GMRES-IR against the FP64 A */
#define CDJ_PREC_LEARNED 2    /* learned SPAI preconditioner, FP64 FGMRES; FP32-LU path as fallback */
#define CDJ_PREC_MIXED16 3    /* as MIXED32, but the LU's trailing updates take FP16 operands (per-column power-of-two
                                 scaling of U) accumulated in FP32 on MFMA */

/* arithmetic of the assembly (cdj_opts.assembly) */
#define CDJ_ASM_FP64 0        /* every entry in FP64 (the reference's arithmetic) */
#define CDJ_ASM_MIXED 1       /* FP32 far field, FP64 near field: a (receiver, source) pair is recomputed in FP64
                                 when its collocation point lies within near_factor x the source's longest edge L
                                 of the source's centroid (which includes every self entry), or within 0.1 L of the
                                 source's plane and 2 near_factor L of its centroid */

/* preconditioners of cdj_precond_solve_dev */
#define CDJ_PC_NONE 0
#define CDJ_PC_JACOBI 1       /* exact inverse of each element's 3x3 self block */
#define CDJ_PC_LEARNED 2      /* block Jacobi + MLP-predicted neighbour blocks */
#define CDJ_PC_LU32 3         /* FP32 LU */
#define CDJ_PC_LEARNED_BF16 4 /* as CDJ_PC_LEARNED, the network evaluated with BF16 weights and activations on MFMA
                                 (FP32 accumulation, biases and tanh in FP32) */

typedef struct cdj_ctx cdj_ctx;
typedef struct cdj_mlp cdj_mlp;

typedef struct cdj_opts {
  int32_t precision;   /* CDJ_PREC_* */
  int32_t tile;        /* triangles per workgroup column in the TD kernels; 0 = library default */
  double tol;          /* target FP64 relative residual ||b - A x|| / ||b|| for refined solves; 0 = 1e-10 */
  int32_t max_iter;    /* iteration cap of the refinement; 0 = default */
  int32_t assembly;    /* CDJ_ASM_*: arithmetic of the traction-matrix assembly */
  double near_factor;  /* near-field radius in longest triangle edges about the centroid (FP64 recompute); 0 = 2.0 */
} cdj_opts;

typedef struct cdj_report {
  int32_t path;            /* precision path that produced x (CDJ_PREC_*) */
  int32_t iterations;      /* refinement iterations (0 for a direct FP64 solve) */
  double residual;         /* final FP64 relative residual ||b - A x|| / ||b|| (max over right-hand sides) */
  double assembly_ms;      /* device time of the assembly */
  double solve_ms;         /* device time of the factorisation and solve */
  int64_t near_entries;    /* entries recomputed in FP64 by the near-field pass */
} cdj_report;

/* Context: one per process and device. */
int cdj_create(int device, cdj_ctx** ctx);
int cdj_destroy(cdj_ctx* ctx);
const char* cdj_last_error(const cdj_ctx* ctx);

/* Device time (ms, HIP events on the launch stream) of the last assembly run on this context (its kernels, from the
 * first to the last); -1 if none.  Used by benchmarks to measure the kernels alone. */
double cdj_last_kernel_ms(cdj_ctx* ctx);

/* Entries of the last traction assembly on this context that the mixed mode (CDJ_ASM_MIXED) recomputed in FP64
 * (0 after an FP64 assembly).  No counterpart in the reference (its assembly is FP64 throughout). */
int64_t cdj_last_near_entries(const cdj_ctx* ctx);

/* Mesh edges of the last FP64 half-space traction assembly on this context, whose free-surface responses were
 * evaluated once per edge and shared by the edge's triangles (0 when every pair evaluated its own sides: full
 * space, a triangle lying in z = 0, or cdj_assemble_dev_noalloc), and the device time (ms, HIP events) of that
 * edge-response launch (-1 when it ran in several receiver chunks or untimed).  For benchmarks; no counterpart in
 * the reference. */
int cdj_last_edge_stats(cdj_ctx* ctx, int64_t* edges, double* ms);

/* TD with influence matrices: replaces TD(X,Y,Z,P1,P2,P3,Dss,Dds,Dn,nu,G,DispFlag,StrainFlag,HSflag[,StrainInf,
 * DispInf]) (TD.jl:77-133, 135-216).  strain: 18 nobs x ntri matrices in StrainInf field order (exxDn, eyyDn,
 * ezzDn, exyDn, exzDn, eyzDn, then Dss, then Dds), NULL when strain_flag != 1; disp: 9 matrices in DispInf order
 * (DnUx, DnUy, DnUz, DssUx, ..., DdsUz), NULL when disp_flag != 1. */
int cdj_td_influence(cdj_ctx* ctx, const double* X, const double* Y, const double* Z, int64_t nobs,
                     const double* P1, const double* P2, const double* P3, int64_t ntri, const double* Dss,
                     const double* Dds, const double* Dn, double nu, double G, int32_t disp_flag,
                     int32_t strain_flag, int32_t hs_flag, double* const* strain, double* const* disp);

/* TD summed field: replaces TD(..., ::Strains, ::Disps) (TD.jl:219-410).  strain6 = {exx, eyy, ezz, exy, exz,
 * eyz}, disp3 = {Ux, Uy, Uz}, each of length nobs.  The sum over triangles is in a fixed order, so results are
 * reproducible run to run (the reference's threaded accumulation at TD.jl:322-324, 375-380 is not). */
int cdj_td_field(cdj_ctx* ctx, const double* X, const double* Y, const double* Z, int64_t nobs, const double* P1,
                 const double* P2, const double* P3, int64_t ntri, const double* Dss, const double* Dds,
                 const double* Dn, double nu, double G, int32_t disp_flag, int32_t strain_flag, int32_t hs_flag,
                 double* const* strain6, double* const* disp3);

/* Traction influence matrices: replaces CalculateInfluenceMatrices3D(FaceNormalVector,MidPoint,P1,P2,P3,nu,G,
 * lambda,FixedEls,HSFlag,n) (CalculateInfluenceMatrices3D.jl:2-48).  T: 9 matrices in TractionInf field order
 * (DnTn, DnTss, DnTds, DssTn, DssTss, DssTds, DdsTn, DdsTss, DdsTds).
 * fixed: NULL or n flags (FixedEls; 1 = fixed).  Without fixed elements each T[k] is N x N.  With nfix fixed
 * elements each T[k] is N x nf (nf = N - nfix, leading dimension N), as the reference's FixedEls branch returns
 * (CalculateInfluenceMatrices3D.jl:10-33, 145-198): a fixed element's row holds the displacement influence at its
 * collocation point (DnTn <- DnUx, DnTss <- DnUy, DnTds <- DnUz, ...), a free element's row the traction, where
 * the i-th free row takes element i's plane as the reference's ConvertStrainInfMatsToTraction does when handed
 * the free rows only; the fixed sources' columns are dropped.  The fixed form is assembled in FP64 whatever
 * opts->assembly. */
int cdj_traction_influence(cdj_ctx* ctx, const double* normals, const double* midpoints, const double* P1,
                           const double* P2, const double* P3, int64_t n, double nu, double G, double lambda,
                           int32_t hs_flag, const int32_t* fixed, double* const* T, const cdj_opts* opts);

/* Slip solve: replaces the base method of SlipCalculator3D (SlipCalculator3D.jl:1-33).  b: 3N x nrhs
 * column-major right-hand sides [Tn; Tss; Tds] as built by SetupTractionVector (and KnockOutFixedRows);
 * x: 3N x nrhs solution [Dn; Dss; Dds]; A_out: 3N x 3N (the assembled A) or NULL.
 * fixed: NULL or N flags (FixedEls).  With nfix > 0 fixed elements, A is the 3N x 3nf matrix of the reference's
 * FixedEls branch (nf = N - nfix; see cdj_traction_influence), x (3nf x nrhs) its least-squares solution
 * [Dn; Dss; Dds] of the free elements (Julia's `\` on a non-square A), A_out 3N x 3nf; FP64 whatever
 * opts->precision, through the augmented system of cdj_lstsq (3N + 3nf <= 65535); report->residual is then the
 * least-squares residual ||b - A x|| / ||b||. */
int cdj_slip_calculator(cdj_ctx* ctx, const double* P1, const double* P2, const double* P3, int64_t n, double nu,
                        double G, double lambda, const double* midpoints, const double* normals, int32_t hs_flag,
                        const double* b, int64_t nrhs, const int32_t* fixed, double* x, double* A_out,
                        const cdj_opts* opts, cdj_report* report);

/* Traction influence rows at m given receivers: points (m x 3) with the planes of `normals` (m x 3), against n
 * source triangles; T: 9 m x n matrices in TractionInf order.  This is ConvertStrainInfMatsToTraction
 * (CalculateInfluenceMatrices3D.jl:73-142) applied to TD strains at arbitrary points - the free rows of the
 * fixed-element form (CalculateInfluenceMatrices3D.jl:10-28), where row i takes element i's plane. */
int cdj_traction_rows(cdj_ctx* ctx, const double* points, const double* normals, int64_t m, const double* P1,
                      const double* P2, const double* P3, int64_t n, double nu, double G, double lambda,
                      int32_t hs_flag, double* const* T, const cdj_opts* opts);

/* Least squares min ||A x - b|| for A m x k (k <= m, column-major), the `A\b` of SlipCalculator3D.jl:22 when
 * fixed elements make A non-square (Julia: pivoted QR).  Solved as [[I, A], [A^T, 0]] [r; x] = [b; 0] with the
 * FP64 LU; m + k <= 65535. */
int cdj_lstsq(cdj_ctx* ctx, const double* A, int64_t m, int64_t k, const double* b, int64_t nrhs, double* x,
              cdj_report* report);

/* Dense solve of A x = b (the `A\b` of SlipCalculator3D.jl:22): blocked LU with partial pivoting on the GPU. */
int cdj_solve(cdj_ctx* ctx, const double* A, const double* b, int64_t n, int64_t nrhs, double* x,
              const cdj_opts* opts, cdj_report* report);

/* ---- device-pointer entry points (inputs already in HBM) ----
 * `stream` NULL (0) selects the context's stream, a blocking stream: it is ordered with work on the legacy null
 * stream (the default stream of a framework that passes 0), so inputs written there are complete before the
 * entry point reads them and outputs are complete before later null-stream work reads them. */

/* Assemble A (3N x 3N, leading dimension lda) on `stream` from device arrays. */
int cdj_assemble_dev(cdj_ctx* ctx, const double* normals, const double* midpoints, const double* P1,
                     const double* P2, const double* P3, int64_t n, double nu, double G, double lambda,
                     int32_t hs_flag, double* A, int64_t lda, const cdj_opts* opts, void* stream);

/* cdj_assemble_dev without any device allocation, for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Same arguments, layout,
 * results as cdj_assemble_dev; its scratch (receiver and triangle records, and for CDJ_ASM_MIXED the
 * FP32 records, near-field table and counter) is a fixed array of CDJ_NOALLOC_SCRATCH_BYTES in device memory, one per
 * GPU, held until the process ends; in half space it evaluates every pair's three free-surface sides (the allocating
 * entry point shares them per mesh edge), equal to it within rounding.  Needs 256 + 8 (RECV + RS) n bytes, RECV = 12, RS = 32 (full space) or 160
 * (half space), plus 4 RS n (rounded up to 256) + 64 n for CDJ_ASM_MIXED: at most 95324 triangles in full space and
 * 24385 in half space (FP64 assembly); larger n is rejected with CDJ_ERR_ARG.  Every call on one GPU uses the same
 * array, whatever the context: calls that overlap on one GPU (other threads, other streams) race. */
#define CDJ_NOALLOC_SCRATCH_BYTES (32ll << 20)
int cdj_assemble_dev_noalloc(cdj_ctx* ctx, const double* normals, const double* midpoints, const double* P1,
                             const double* P2, const double* P3, int64_t n, double nu, double G, double lambda,
                             int32_t hs_flag, double* A, int64_t lda, const cdj_opts* opts, void* stream);

/* TD influence on device arrays; same layout as cdj_td_influence. */
int cdj_td_influence_dev(cdj_ctx* ctx, const double* X, const double* Y, const double* Z, int64_t nobs,
                         const double* P1, const double* P2, const double* P3, int64_t ntri, const double* Dss,
                         const double* Dds, const double* Dn, double nu, double G, int32_t disp_flag,
                         int32_t strain_flag, int32_t hs_flag, double* const* strain, double* const* disp,
                         void* stream);

/* LU factorisation in place (FP64, partial pivoting; ipiv 0-based, length n) and solve with the factors. */
int cdj_lu_factor_dev(cdj_ctx* ctx, double* A, int64_t n, int64_t lda, int32_t* ipiv, void* stream);
/* Many independent FP64 LUs at once (no counterpart in the reference, which factors one matrix per `A\b`; this is
 * the factorization behind cdj_batched_slip): matrix j is A[j] (device, n[j] x n[j], leading dimension lda[j]),
 * factored in place with pivots ipiv[j] (device) exactly as cdj_lu_factor_dev factors it alone; the matrices'
 * panels share cooperative launches.  info (host, count entries) receives 0 or 1 + the first zero pivot's index
 * for each matrix; a singular matrix is not an error here.  Synchronises the stream. */
int cdj_lu_factor_batched_dev(cdj_ctx* ctx, double* const* A, const int64_t* n, const int64_t* lda, int64_t count,
                              int32_t* const* ipiv, int32_t* info, void* stream);
int cdj_lu_solve_dev(cdj_ctx* ctx, const double* LU, int64_t n, int64_t lda, const int32_t* ipiv, double* B,
                     int64_t nrhs, int64_t ldb, void* stream);

/* ---- learned preconditioner (no counterpart in the reference: the build's learned solve path) ----
 * A per-triangle MLP over the k nearest triangles' geometry (relative centroids, normals, strike directions,
 * areas in the triangle's own frame) predicts the 3x3 blocks of a sparse approximate inverse row block;
 * M = block Jacobi + (h_i / G) MLP(features_i).  Trained on the GPU on (x, b = A x) pairs of assembled
 * matrices, loss mean ||M b - x||^2. */
int cdj_mlp_create(cdj_ctx* ctx, int32_t k, int32_t hidden, uint64_t seed, cdj_mlp** out);
int cdj_mlp_destroy(cdj_mlp* mlp);
int64_t cdj_mlp_nparams(const cdj_mlp* mlp);
int cdj_mlp_get_params(cdj_ctx* ctx, const cdj_mlp* mlp, float* params);
int cdj_mlp_set_params(cdj_ctx* ctx, cdj_mlp* mlp, const float* params);

/* `iters` Adam steps on one mesh (device geometry, device A of 3N x 3N, leading dimension lda) with
 * `nsamples` random slip vectors; *loss receives the last step's mean loss. */
int cdj_mlp_train_dev(cdj_ctx* ctx, cdj_mlp* mlp, const double* P1, const double* P2, const double* P3, int64_t n,
                      double G, const double* A, int64_t lda, int32_t nsamples, int32_t iters, float lr,
                      uint64_t seed, double* loss, void* stream);

/* Loss and parameter gradient for host inputs (gradient checks): geometry P1..P3 (n x 3), Jacobi blocks jac
 * (n x 9, row-major 3x3 per element), samples B, X (3n x S).  Also returns the features (n x k*10), neighbour
 * lists (n x k) and per-element geometry records (n x 16) the GPU used. */
int cdj_mlp_loss_grad(cdj_ctx* ctx, cdj_mlp* mlp, const double* P1, const double* P2, const double* P3, int64_t n,
                      double G, const double* jac, const double* B, const double* X, int32_t nsamples, double* loss,
                      float* grad, float* feat_out, int32_t* nbr_out, float* geo_out);

/* FP64 FGMRES on device A with the chosen preconditioner (CDJ_PC_*) to opts->tol; b, x: 3N vectors. */
int cdj_precond_solve_dev(cdj_ctx* ctx, const cdj_mlp* mlp, int32_t pc, const double* P1, const double* P2,
                          const double* P3, int64_t n, double G, const double* A, int64_t lda, const double* b,
                          double* x, const cdj_opts* opts, cdj_report* report, void* stream);

/* ---- batched slip solves (no single counterpart in the reference: one SlipCalculator3D base-method call,
 * SlipCalculator3D.jl:1-33, per mesh, with several right-hand sides each) ----
 * All meshes' collocation matrices are assembled in one persistent work-queue launch per pass (FP64, or the
 * FP32 far / FP64 near passes under opts->assembly = CDJ_ASM_MIXED), captured in a HIP graph together with the
 * per-triangle set-up; the source-chunk width and the number of persistent workgroups are chosen by timing the
 * candidates on the first batch of a shape (cached per half-space flag, mode and size).  A later call with the
 * same shape reuses the buffers and replays the graph.  Each mesh is then solved as cdj_slip_calculator would
 * (opts->precision).  Meshes must all be full space or all half space (hs_flag). */
typedef struct cdj_mesh {
  int64_t n;                   /* triangles */
  const double* P1;            /* n x 3, column-major */
  const double* P2;
  const double* P3;
  const double* normals;       /* n x 3 (CreateFaceNormalAndMidPoint) */
  const double* midpoints;     /* n x 3 */
  int64_t nrhs;                /* right-hand sides, >= 1 */
  const double* b;             /* 3n x nrhs, [Tn; Tss; Tds] per column */
  double* x;                   /* 3n x nrhs output, [Dn; Dss; Dds] */
} cdj_mesh;

int cdj_batched_slip(cdj_ctx* ctx, const cdj_mesh* meshes, int64_t nmesh, double nu, double G, double lambda,
                     int32_t hs_flag, const cdj_opts* opts, cdj_report* reports /* nmesh or NULL */,
                     cdj_report* total /* or NULL */);

/* Tile chosen by the autotuner for the current batch plan: sources per work item and persistent workgroups. */
int cdj_batch_tuning(const cdj_ctx* ctx, int32_t* tpb, int32_t* wgs);

/* The learned preconditioner's network on one mesh (host geometry, n x 3 column-major): Y receives the n x O
 * outputs (row-major, O = 144 for the 3 x 3 blocks of the 16 neighbours), evaluated in FP32 (bf16 == 0) or with BF16
 * weights and activations on MFMA (bf16 != 0).  No counterpart in the reference. */
int cdj_mlp_forward(cdj_ctx* ctx, const cdj_mlp* mlp, const double* P1, const double* P2, const double* P3, int64_t n,
                    int32_t bf16, float* Y);

/* Frictional slip: replaces SlipCalculator3D(P1,P2,P3,nu,G,lambda,MidPoint,FaceNormalVector,HSFlag,
 * MixedBoundaryConditionsFriction,FixedEls) (SlipCalculator3D.jl:36-233) for FixedEls all zero.  b: 3n tractions
 * [Tn; Tss; Tds] of the remote stresses plus the prescribed tractions; mu, Sf: n friction coefficients and
 * cohesions.  When no element can overcome its frictional strength the slip is zero (SlipCalculator3D.jl:52-59).
 * Otherwise the collocation matrix (scaled by mean half perimeter / G) is inverted on the GPU, the 5n x 5n
 * complementarity problem of SlipCalculator3D.jl:90-167 is built there and solved by a Fischer-Newton method
 * (the Newton systems factored by the GPU LU).  For elements with mu = 0 and Sf = 0 the four unknowns that the
 * problem fixes at zero or that only split the slip's sign are left out of the Newton solve (same solution; without
 * this the method stalls on those degenerate pairs).  report: iterations = Newton iterations, residual = final
 * 0.5 ||phi||^2 of the Fischer-Burmeister function, solve_ms = device time of the complementarity solve.
 * 5 n <= 65535. */
int cdj_slip_friction(cdj_ctx* ctx, const double* P1, const double* P2, const double* P3, int64_t n, double nu,
                      double G, double lambda, const double* midpoints, const double* normals, int32_t hs_flag,
                      const double* b, const double* mu, const double* Sf, double* Dn, double* Dss, double* Dds,
                      cdj_report* report);

/* Constant fluid volume: replaces SlipCalculator3D(P1,P2,P3,nu,G,lambda,MidPoint,FaceNormalVector,HSFlag,
 * MixedBoundaryConditionsFluidVolume,FractureFlag) (SlipCalculator3D.jl:242-475 with ComputePressurisedCrackDn.jl),
 * the single-fracture branch.  b: 3n tractions [Tn; Tss; Tds] of the remote stresses plus the prescribed
 * tractions; fracture_flag: n ints, 1 = crack element, 0 = free-surface element (more than one fracture, a flag
 * above 1, ends in the reference's error); volume: the crack's target volume V.  A pressure x Norm on the crack's
 * normal traction is searched for whose frictional slip (mu = 0.1, Sf = 0, as the reference fixes them) gives the
 * volume V: Norm from a uniform opening of V, rescaled until the volume at x = 10 lies in [V, 2.5 V], then Brent's
 * method on (V - volume(x))^2 over [0, 10] (abs_tol 0.001) and a golden-section search over [0, 50] when that
 * minimiser's volume is negative.  Each volume evaluation is one complementarity solve on the GPU with the
 * inverse and the 5n x 5n matrix built once.  The slip is NaN when its volume exceeds 2 |V| (as the reference).
 * pressure (may be NULL): the pressure x Norm found.  report: iterations = volume evaluations, near_entries =
 * Newton steps over all evaluations, residual = |V - volume| / |V|, solve_ms = device time of the whole call.
 * 5 n <= 65535. */
int cdj_slip_fluid_volume(cdj_ctx* ctx, const double* P1, const double* P2, const double* P3, int64_t n, double nu,
                          double G, double lambda, const double* midpoints, const double* normals, int32_t hs_flag,
                          const double* b, const int32_t* fracture_flag, double volume, double* Dn, double* Dss,
                          double* Dds, double* pressure, cdj_report* report);

/* ---- crack-tip edges and stress intensities (GetCrackTipElements3D.jl:1, StressIntensity3D.jl:1) ----
 * Edge blocks, in this order: P1P2, P1P3, P2P3 (the reference's FeP1P2S, FeP1P3S, FeP2P3S).
 * free_flags: n x 3 bytes (block-major: flags of block s at [s n, (s + 1) n)), 1 = free (crack-tip) edge.
 * geometry:   3 blocks of n x 13 doubles, column-major: FeLe, FeMd x y z, FeEv x y z, FeM2Ev x y z, FeM2ELe,
 *             IntAng (degrees), Area; NaN on rows whose edge is not free, except Area (every row).
 * K:          3 blocks of n x 4 doubles, column-major: K1, K2, K3, StrainEnergy (the value StressIntensity3D.jl
 *             stores under that name: sqrt(K1^2 + K2^2 + K3^2 / (1 - nu))); NaN off the free edges.
 * Free edges are edges of the mesh with no other edge of identical end points (EdgeConstraints.jl); the GPU finds
 * them with a hash table in O(n) instead of the reference's O(n^2) pairwise search.  n <= 7e8. */
int cdj_crack_tip_elements(cdj_ctx* ctx, const double* midpoints, const double* P1, const double* P2,
                           const double* P3, const double* normals, int64_t n, uint8_t* free_flags,
                           double* geometry);
int cdj_stress_intensity(cdj_ctx* ctx, const double* Dn, const double* Dss, const double* Dds, int64_t n, double G,
                         double nu, const double* normals, const uint8_t* free_flags, const double* geometry,
                         double* K);
/* Both steps on device arrays (HBM-resident inputs and outputs, same layouts), on `stream`; the device time of
 * the two steps is then available from cdj_last_kernel_ms. */
int cdj_stress_intensity_dev(cdj_ctx* ctx, const double* midpoints, const double* P1, const double* P2,
                             const double* P3, const double* normals, int64_t n, const double* Dn, const double* Dss,
                             const double* Dds, double G, double nu, uint8_t* free_flags, double* geometry, double* K,
                             void* stream);

/* ---- isotropic remeshing (host) ----
 * Replaces BuildCGAL.PolygonRemeshingCGAL (CGAL 4.13.1 isotropic_remeshing), which the reference's growth loops
 * call between OFFExport and OFFReader (test/test_PropagationTest.jl:151-561).  The published algorithm (Botsch &
 * Kobbelt 2004) with target edge length L: split edges longer than 4/3 L, collapse edges shorter than 4/5 L,
 * flip for valence 6 (4 on the boundary), tangential relaxation, projection onto the input surface; `iterations`
 * rounds.  The boundary keeps its shape (its vertices stay on the input's boundary polyline; corners turning by
 * more than 45 degrees stay put) and triangle orientation is kept.
 * points: npoints x 3, column-major (x, y, z); tris: ntris x 3, column-major, 0-based.  On success *out holds the
 * result, of *npoints_out points and *ntris_out triangles, until cdj_remesh_take copies it (same layouts; either
 * pointer may be NULL) and frees it.  No GPU is used.  CDJ_ERR_DEGENERATE for a triangle with a repeated corner. */
typedef struct cdj_remesh cdj_remesh;
int cdj_remesh_isotropic(const double* points, int64_t npoints, const int64_t* tris, int64_t ntris,
                         double target_edge_length, int32_t iterations, cdj_remesh** out, int64_t* npoints_out,
                         int64_t* ntris_out);
int cdj_remesh_take(cdj_remesh* r, double* points, int64_t* tris);

#ifdef __cplusplus
}
#endif

#endif /* CDJ_H */
