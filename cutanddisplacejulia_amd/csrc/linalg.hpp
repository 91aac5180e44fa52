// Dense linear algebra on gfx950: blocked LU with partial pivoting and triangular solves.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cdj {

static constexpr int LU_NB = 64;  // panel width

// Workspace sizes (bytes) for lu_factor.
size_t lu_workspace_bytes(int64_t n);

// In-place LU of the n x n column-major matrix A (leading dimension lda) with partial pivoting.
// ipiv (device, length n) receives 0-based pivot rows; info (device int) receives 0, 1 + the first zero
// pivot's index, or -1 if a panel's grid barrier wait overran (factors invalid).  ws must hold
// lu_workspace_bytes(n) bytes.  T = double or float.
// fp16_update (T = float only): the rank-256 trailing updates take FP16 operands (U scaled per column by a power
// of two) accumulated in FP32 on MFMA; panels, catch-up updates and triangular solves stay FP32.
// side: a second stream (not st) for the look-ahead, which runs most of each trailing update there beside the next
// block's panels on st; st waits for all of it before the call's last work.  NULL: no look-ahead.  Either way the
// factors are the same bit for bit.
template <class T>
hipError_t lu_factor(T* A, int64_t n, int64_t lda, int32_t* ipiv, int* info, void* ws, hipStream_t st,
                     bool fp16_update = false, hipStream_t side = nullptr);

// LU of an n x ncols matrix (ncols >= n): the extra columns are treated as right-hand sides, so on return they
// hold L^-1 P B.
template <class T>
hipError_t lu_factor_cols(T* A, int64_t n, int64_t ncols, int64_t lda, int32_t* ipiv, int* info, void* ws,
                          hipStream_t st, bool fp16_update = false, hipStream_t side = nullptr);

// One matrix of a batched factorization: as lu_factor_cols(A, n, ncols, lda, ipiv, info, ws, ...), with ws of
// lu_workspace_bytes(n) bytes (one per matrix).
template <class T>
struct LuJob {
  T* A;
  int64_t n, ncols, lda;
  int32_t* ipiv;
  int* info;
  void* ws;
};

// Factors every job's matrix as lu_factor_cols would, with the panels of all matrices at the same column sharing
// cooperative launches.  fp16_update applies to the jobs with ncols == n.  Only kernel launches are enqueued (no
// memsets, no host synchronisation), so the sequence can be captured in a graph once lu_batched_capacity<T>() has
// been called outside the capture.
template <class T>
hipError_t lu_factor_batched(const LuJob<T>* jobs, int njobs, hipStream_t st, bool fp16_update = false);

// Workgroups one batched panel launch may hold (0 on failure); one-time set-up of the panel kernel's LDS limit.
template <class T>
int lu_batched_capacity();

// Row permutation implied by the pivots (row i of P A is row perm[i] of A); perm_dev is a device array.
// Synchronises the stream.
hipError_t lu_permutation(const int32_t* ipiv, int64_t n, int32_t* perm_dev, hipStream_t st);

// Blocked triangular solves on the factors (B is n x nrhs, leading dimension ldb).
template <class T>
hipError_t tri_solve_lower_unit(const T* LU, int64_t n, int64_t lda, T* B, int64_t nrhs, int64_t ldb, hipStream_t st);
template <class T>
hipError_t tri_solve_upper(const T* LU, int64_t n, int64_t lda, T* B, int64_t nrhs, int64_t ldb, hipStream_t st);

// U X = B for many right-hand sides (B n x nrhs): blocked, the updates on MFMA.
template <class T>
hipError_t tri_solve_upper_many(const T* LU, int64_t n, int64_t lda, T* B, int64_t nrhs, int64_t ldb, hipStream_t st);

// Inverse of the n x n FP64 matrix in AI[:, 0:n] (an n x 2n buffer, leading dimension n): AI[:, n:2n] receives
// A^-1 (AI[:, 0:n] the LU factors).  ipiv: n ints; info as lu_factor; ws: lu_workspace_bytes(n).
hipError_t invert(double* AI, int64_t n, int32_t* ipiv, int* info, void* ws, hipStream_t st);

// Solve A X = B with the factors of lu_factor (B overwritten by X, n x nrhs, leading dimension ldb).
template <class T>
hipError_t lu_solve(const T* LU, int64_t n, int64_t lda, const int32_t* ipiv, T* B, int64_t nrhs, int64_t ldb,
                    hipStream_t st);

// y = alpha * A x + beta * y (A: m x n column-major, FP64 accumulate).  Used for residuals.
template <class TA>
void gemv(const TA* A, int64_t m, int64_t n, int64_t lda, const double* x, double* y, double alpha, double beta,
          hipStream_t st);

// Element-wise conversion between precisions.
void convert_d2f(const double* a, float* b, int64_t n, hipStream_t st);
void convert_f2d(const float* a, double* b, int64_t n, hipStream_t st);
void convert_matrix_d2f(const double* A, int64_t lda, float* B, int64_t ldb, int64_t m, int64_t n, hipStream_t st);

// Least-squares system matrix K = [[I, A], [A^T, 0]] ((m + k) x (m + k)) for min ||A x - b||: K [r; x] = [b; 0].
void build_kkt(const double* A, int64_t lda, int64_t m, int64_t k, double* K, int64_t ldk, hipStream_t st);

}  // namespace cdj
