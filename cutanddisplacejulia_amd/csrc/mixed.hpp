// Mixed-precision solve: FP64 residual GMRES-based iterative refinement preconditioned by a low-precision LU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

namespace cdj {

struct TrsvState {
  int* flags;             // device, one per 64-row block, zero-initialised once
  int* err;               // device error word (set on a timed-out wait)
  int epoch;              // increments per solve
  long long spin_limit;   // bound on every wait loop
};

struct GmresOpts {
  double tol = 1e-10;     // FP64 relative residual target
  int restart = 30;
  int max_iter = 200;
  int max_restarts = 20;
  long long spin_limit = 1LL << 20;  // about a second of polling: the chained blocks publish within microseconds
  bool zero_guess = false;  // start from x = 0 instead of x = M^-1 b
  // FP32 copy of A for the Arnoldi products (nullptr: FP64 A).  The residual at every restart stays FP64, so a
  // cycle on the FP32 operator is one step of iterative refinement; each cycle then stops once its estimated
  // residual has fallen by inner_rel (the FP32 operator cannot resolve more) or reached the target.
  const float* A32 = nullptr;
  int64_t lda32 = 0;
  double inner_rel = 1e-6;
};

// Preconditioner application z = M^-1 r on device vectors of length n (enqueued on the solver's stream).
using Precond = std::function<hipError_t(const double* r, double* z)>;

struct GmresResult {
  int iterations = 0;
  bool converged = false;
  bool trsv_timeout = false;
  double residual = 0.0;
  std::vector<double> history;
};

size_t gemv_part_bytes(int64_t m, int64_t n);

// y = alpha * A x + beta * yin (FP64 accumulate; A of precision TA); part holds gemv_part_bytes(m, n).
template <class TA>
void gemv_splitk(const TA* A, int64_t m, int64_t n, int64_t lda, const double* x, double* y, double alpha,
                 const double* yin, double beta, double* part, hipStream_t st);

template <class T>
hipError_t trsv_persistent(const T* LU, int64_t n, int64_t lda, T* x, bool lower, TrsvState& s, hipStream_t st);

// Largest n the persistent solve takes (all its 64-row blocks co-resident); 0 if the query failed.
template <class T>
int64_t trsv_persistent_max_n();

// z = U^-1 L^-1 P r with factors of precision T (work: n elements of T).
template <class T>
hipError_t apply_lu_precond(const T* LU, int64_t n, int64_t lda, const int32_t* perm, const double* r, double* z,
                            T* work, TrsvState& s, hipStream_t st);

// One system of a batched application of FP32 LU factors: X (+)= U^-1 L^-1 P R for nrhs <= APPLY_NR columns.
// R, X and W (FP32 work) are n x nrhs, column-major with leading dimension n.  A system with more right-hand sides
// is split into several jobs, which then run on different CUs.
constexpr int APPLY_NR = 2;
struct LuApplyJob {
  const float* LU;       // unit lower L and U in place, n x n, leading dimension lda
  const int32_t* perm;   // row i of P A is row perm[i] of A
  const double* R;
  double* X;
  float* W;
  int64_t n, lda;
  int32_t nrhs, accumulate;
};

// All jobs (device array) in one launch, one workgroup per job.
hipError_t lu_apply_batched(const LuApplyJob* jobs, int njobs, hipStream_t st);

// perm (device, n) from the 0-based pivot rows ipiv of an LU: row i of P A is row perm[i] of A.
struct PermJob {
  const int32_t* ipiv;
  int32_t* perm;
  int64_t n;
};
constexpr int64_t PERM_MAX = 16384;  // largest n (the pivots and the permutation are staged in LDS)
// One-time set-up of the kernel's LDS limit; returns the largest n supported (0 on failure).  Call it outside any
// graph capture.
int64_t pivots_to_perm_max();
// All jobs (device array) in one launch; nmax >= every job's n.  Kernel launches only (graph-capturable).
hipError_t pivots_to_perm(const PermJob* jobs, int njobs, int64_t nmax, hipStream_t st);

// R = B - A X (FP64, n x nrhs) and nrm2[q] = ||R[:, q]||^2, A read once per 8 columns; part holds
// residual_multi_bytes(n, nrhs).
size_t residual_multi_bytes(int64_t n, int64_t nrhs);
hipError_t residual_multi(const double* A, int64_t n, int64_t lda, const double* B, int64_t ldb, const double* X,
                          int64_t ldx, double* R, int64_t ldr, int64_t nrhs, double* part, double* nrm2,
                          hipStream_t st);

// nrm2[q] = ||R[:, q]||^2 in a fixed summation order.
void colnorm2(const double* R, int64_t ld, int64_t n, int64_t nrhs, double* nrm2, hipStream_t st);

// Right-preconditioned restarted FGMRES on A x = b (A FP64, n x n) with FP64 products and residuals, until
// ||b - A x|| / ||b|| <= o.tol or o.max_iter iterations.
hipError_t fgmres(const double* A, int64_t n, int64_t lda, const Precond& M, const double* b, double* x,
                  const GmresOpts& o, GmresResult& res, hipStream_t st);

// Solve A x = b (A FP64, n x n) by right-preconditioned restarted FGMRES with M = P^T L U in precision TF,
// until ||b - A x|| / ||b|| <= o.tol.  x receives the solution.
template <class TF>
hipError_t fgmres_ir(const double* A, int64_t n, int64_t lda, const TF* LU, int64_t ldlu, const int32_t* perm,
                     const double* b, double* x, const GmresOpts& o, GmresResult& res, hipStream_t st);

}  // namespace cdj
