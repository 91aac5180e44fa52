// Mixed-precision dense solve for gfx950: low-precision LU (FP32 MFMA) as a preconditioner of restarted
// FGMRES whose matrix-vector products, residuals and updates are FP64, iterated until the FP64 relative
// residual ||b - A x|| / ||b|| reaches the requested tolerance (GMRES-based iterative refinement).
//
//   gemv        : split-K FP64-accumulated matrix-vector product (the HBM-bound residual), deterministic
//                 two-pass reduction over column chunks.
//   trsv        : persistent triangular solves.  One workgroup per 64-row block; block b consumes the solved
//                 blocks c < b (c > b for the upper solve) as their ready flags appear, so the off-diagonal
//                 products overlap the sequential chain.  Flags carry a per-solve epoch, waits are bounded
//                 and a timed-out wait sets an error word and exits, so every wave always terminates.
//   precond     : z = U^-1 L^-1 P r in the factor precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "linalg.hpp"
#include "mixed.hpp"

namespace cdj {

namespace {

constexpr int TB = 64;  // rows per triangular block

unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// ---- split-K GEMV ---------------------------------------------------------------------------------------
template <class TA>
__global__ void __launch_bounds__(256) gemv_part_kernel(const TA* __restrict__ A, int64_t m, int64_t n, int64_t lda,
                                                        const double* __restrict__ x, int chunk,
                                                        double* __restrict__ part) {
  __shared__ double xs[1024];
  const int64_t c0 = (int64_t)blockIdx.y * chunk;
  const int64_t c1 = std::min<int64_t>(n, c0 + chunk);
  for (int64_t c = c0 + threadIdx.x; c < c1; c += 256) xs[c - c0] = x[c];
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  double s0 = 0.0, s1 = 0.0;
  int64_t c = c0;
  // unrolled so that many column loads are in flight per wave (the sums keep their order)
#pragma unroll 8
  for (; c + 1 < c1; c += 2) {
    s0 += (double)A[i + c * lda] * xs[c - c0];
    s1 += (double)A[i + (c + 1) * lda] * xs[c + 1 - c0];
  }
  if (c < c1) s0 += (double)A[i + c * lda] * xs[c - c0];
  part[(int64_t)blockIdx.y * m + i] = s0 + s1;
}

// The FP32 form for an even leading dimension: two rows per thread, read as one 8-byte pair per column (a wave's
// load then covers 512 contiguous bytes, as the FP64 form's does); each row's sum is formed exactly as above.
__global__ void __launch_bounds__(256) gemv_part2_f32_kernel(const float* __restrict__ A, int64_t m, int64_t n,
                                                             int64_t lda, const double* __restrict__ x, int chunk,
                                                             double* __restrict__ part) {
  __shared__ double xs[1024];
  const int64_t c0 = (int64_t)blockIdx.y * chunk;
  const int64_t c1 = std::min<int64_t>(n, c0 + chunk);
  for (int64_t c = c0 + threadIdx.x; c < c1; c += 256) xs[c - c0] = x[c];
  __syncthreads();
  const int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 2;
  if (i >= m) return;
  double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;  // row i: a0 + a1, row i + 1: b0 + b1
  int64_t c = c0;
  if (i + 1 < m) {
#pragma unroll 8
    for (; c + 1 < c1; c += 2) {
      const float2 p = *reinterpret_cast<const float2*>(A + i + c * lda);
      const float2 q = *reinterpret_cast<const float2*>(A + i + (c + 1) * lda);
      a0 += (double)p.x * xs[c - c0];
      b0 += (double)p.y * xs[c - c0];
      a1 += (double)q.x * xs[c + 1 - c0];
      b1 += (double)q.y * xs[c + 1 - c0];
    }
    if (c < c1) {
      const float2 p = *reinterpret_cast<const float2*>(A + i + c * lda);
      a0 += (double)p.x * xs[c - c0];
      b0 += (double)p.y * xs[c - c0];
    }
    part[(int64_t)blockIdx.y * m + i] = a0 + a1;
    part[(int64_t)blockIdx.y * m + i + 1] = b0 + b1;
  } else {
    for (; c + 1 < c1; c += 2) {
      a0 += (double)A[i + c * lda] * xs[c - c0];
      a1 += (double)A[i + (c + 1) * lda] * xs[c + 1 - c0];
    }
    if (c < c1) a0 += (double)A[i + c * lda] * xs[c - c0];
    part[(int64_t)blockIdx.y * m + i] = a0 + a1;
  }
}

__global__ void gemv_reduce_kernel(const double* __restrict__ part, int64_t m, int nchunk, double alpha,
                                   const double* __restrict__ yin, double beta, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int k = 0; k < nchunk; ++k) s += part[(int64_t)k * m + i];
  y[i] = alpha * s + (beta == 0.0 ? 0.0 : beta * yin[i]);
}

// ---- persistent triangular solves ----------------------------------------------------------------------
__device__ __forceinline__ bool wait_flag(const int* flag, int epoch, long long spin_limit, int* err) {
  long long it = 0;
  while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
    if (++it > spin_limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
      atomicOr(err, 1);
      return false;
    }
    __builtin_amdgcn_s_sleep(1);
  }
  return true;
}

// value of v in lane l (l wave-uniform), through scalar registers
__device__ __forceinline__ float lane_value(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned long long)b, l);
  const int hi = __builtin_amdgcn_readlane((int)((unsigned long long)b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// LOWER: unit lower (forward) solve; otherwise upper (backward) solve with the diagonal.
// x (length n, precision T) holds the right-hand side on entry and the solution on exit.
template <class T, bool LOWER>
__global__ void __launch_bounds__(256) trsv_persistent_kernel(const T* __restrict__ LU, int64_t n, int64_t lda,
                                                              T* x, int* flags, int epoch, long long spin_limit,
                                                              int* err) {
  __shared__ double accd[4][TB];
  __shared__ T xs[TB];
  __shared__ T rhs[TB];
  __shared__ T D[TB][TB + 1];  // this block's diagonal block of LU
  __shared__ int s_ok;
  const int nb = (int)((n + TB - 1) / TB);
  const int b = blockIdx.x;
  const int64_t r0 = (int64_t)b * TB;
  const int nr = (int)std::min<int64_t>(TB, n - r0);
  const int tid = threadIdx.x, row = tid % TB, grp = tid / TB;  // 4 groups of 64 threads
  // cooperative launch: acquire the previous kernels' writes (LU, x) before reading them
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  if (tid < nr) rhs[tid] = __hip_atomic_load(&x[r0 + tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int idx = tid; idx < nr * nr; idx += 256) {
    const int r = idx % nr, c = idx / nr;
    D[r][c] = LU[(r0 + r) + (r0 + c) * lda];
  }
  double acc = 0.0;  // per (row, grp) partial of sum_c L[r, c] x[c]
  // off-diagonal blocks in dependency order: c = 0..b-1 (lower), c = nb-1..b+1 (upper)
  const int ndep = LOWER ? b : nb - 1 - b;
  for (int k = 0; k < ndep; ++k) {
    const int c = LOWER ? k : nb - 1 - k;
    const int64_t q0 = (int64_t)c * TB;
    const int nq = (int)std::min<int64_t>(TB, n - q0);
    // this thread's 16 matrix entries of the dependency block do not depend on x: load them before the wait
    T a[TB / 4];
    const bool full = nq == TB && row < nr;
    if (full) {
#pragma unroll
      for (int t = 0; t < TB / 4; ++t) a[t] = LU[(r0 + row) + (q0 + grp + 4 * t) * lda];
    }
    if (tid == 0) s_ok = wait_flag(&flags[c], epoch, spin_limit, err) ? 1 : 0;
    __syncthreads();
    if (!s_ok) return;
    if (tid < nq) xs[tid] = __hip_atomic_load(&x[q0 + tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (full) {
#pragma unroll
      for (int t = 0; t < TB / 4; ++t) acc += (double)a[t] * (double)xs[grp + 4 * t];
    } else if (row < nr) {
      for (int q = grp; q < nq; q += 4) acc += (double)LU[(r0 + row) + (q0 + q) * lda] * (double)xs[q];
    }
    __syncthreads();
  }
  accd[grp][row] = acc;
  __syncthreads();
  if (tid < nr) rhs[tid] = (T)((double)rhs[tid] - (accd[0][tid] + accd[1][tid] + accd[2][tid] + accd[3][tid]));
  __syncthreads();
  // diagonal block: wave 0, lane i holds x[r0 + i]; step r broadcasts x[r0 + r] (the block is staged in LDS)
  if (tid < 64 && nr == TB) {
    // full block, steps unrolled: the broadcast is a register read of a fixed lane and the LDS reads of the
    // block do not wait for the chain; the division by the diagonal is a multiplication by its reciprocal
    // (formed before the chain) with one correction step
    T xi = rhs[tid];
    if (LOWER) {
#pragma unroll
      for (int r = 0; r < TB - 1; ++r) {
        const T xr = lane_value(xi, r);
        const T dr = D[tid][r];
        if (tid > r) xi -= dr * xr;
      }
    } else {
      const T d = D[tid][tid];
      const T rd = T(1) / d;
#pragma unroll
      for (int r = TB - 1; r >= 0; --r) {
        if (tid == r) {
          const T q = xi * rd;
          xi = fma(fma(-q, d, xi), rd, q);
        }
        const T xr = lane_value(xi, r);
        const T dr = D[tid][r];
        if (tid < r) xi -= dr * xr;
      }
    }
    __hip_atomic_store(&x[r0 + tid], xi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else if (tid < 64) {
    T xi = tid < nr ? rhs[tid] : T(0);
    if (LOWER) {
      for (int r = 0; r < nr - 1; ++r) {
        const T xr = __shfl(xi, r);
        if (tid > r && tid < nr) xi -= D[tid][r] * xr;
      }
    } else {
      for (int r = nr - 1; r >= 0; --r) {
        if (tid == r) xi = xi / D[r][r];
        const T xr = __shfl(xi, r);
        if (tid < r) xi -= D[tid][r] * xr;
      }
    }
    if (tid < nr) __hip_atomic_store(&x[r0 + tid], xi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __threadfence();
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&flags[b], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- vector kernels ------------------------------------------------------------------------------------
template <class TI, class TO>
__global__ void permute_cast_kernel(const int32_t* __restrict__ perm, const TI* __restrict__ in, TO* __restrict__ out,
                                    int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (TO)in[perm[i]];
}

template <class TI, class TO>
__global__ void cast_kernel(const TI* __restrict__ in, TO* __restrict__ out, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = (TO)in[i];
}

// partial dot products: out[block] = sum over the block's slice of a.b (one value per block)
__global__ void __launch_bounds__(256) dot_part_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                       int64_t n, double* __restrict__ out) {
  __shared__ double s[256];
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) v += a[i] * b[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) s[threadIdx.x] += s[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// multi-dot: out[j * nblk + block] = partial of V[:, j] . w for j < k
__global__ void __launch_bounds__(256) mdot_part_kernel(const double* __restrict__ V, int64_t ldv, int k,
                                                        const double* __restrict__ w, int64_t n,
                                                        double* __restrict__ out) {
  __shared__ double s[256];
  for (int j = 0; j < k; ++j) {
    double v = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
      v += V[i + j * ldv] * w[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[(int64_t)j * gridDim.x + blockIdx.x] = s[0];
    __syncthreads();
  }
}

// w -= V[:, 0:k] h   (h on device)
__global__ void mgs_update_kernel(const double* __restrict__ V, int64_t ldv, int k, const double* __restrict__ h,
                                  double* __restrict__ w, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = w[i];
  for (int j = 0; j < k; ++j) s -= V[i + j * ldv] * h[j];
  w[i] = s;
}

__global__ void axpby_kernel(double a, const double* __restrict__ x, double b, double* __restrict__ y, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  // b == 0 never reads y: y may be fresh, uninitialised memory (0 * NaN would poison the result)
  if (i < n) y[i] = b == 0.0 ? a * x[i] : a * x[i] + b * y[i];
}

// x += Z[:, 0:k] yk
__global__ void combine_kernel(const double* __restrict__ Z, int64_t ldz, int k, const double* __restrict__ yk,
                               double* __restrict__ x, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = x[i];
  for (int j = 0; j < k; ++j) s += Z[i + j * ldz] * yk[j];
  x[i] = s;
}

constexpr int DOT_BLOCKS = 240;

// out[j] = sum over q of part[j * nparts + q], in q order (the same order as a host summation)
__global__ void reduce_parts_kernel(const double* __restrict__ part, int nvec, int nparts, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nvec) return;
  double s = 0.0;
  // unrolled so that the loads are in flight together (the additions keep their order)
#pragma unroll 16
  for (int q = 0; q < nparts; ++q) s += part[(int64_t)j * nparts + q];
  out[j] = s;
}

// out[q] = sum_i R[i + q ld]^2, one workgroup per column, fixed summation order (bitwise reproducible)
__global__ void __launch_bounds__(256) colnorm2_kernel(const double* __restrict__ R, int64_t ld, int64_t n,
                                                       double* __restrict__ out) {
  __shared__ double s[256];
  const double* r = R + (int64_t)blockIdx.x * ld;
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) v += r[i] * r[i];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// ---- multi-column residual: R = B - A X for up to APPLY_NR columns, A read once ----------------------------
constexpr int RES_CHUNK = 256;  // columns of A per partial sum
constexpr int RES_NR = 8;       // right-hand sides per pass over A

__global__ void __launch_bounds__(256) gemv_nr_part_kernel(const double* __restrict__ A, int64_t m, int64_t n,
                                                           int64_t lda, const double* __restrict__ X, int64_t ldx,
                                                           int nr, double* __restrict__ part) {
  __shared__ double xs[RES_NR][RES_CHUNK];
  const int64_t c0 = (int64_t)blockIdx.y * RES_CHUNK;
  const int nc = (int)(n - c0 < RES_CHUNK ? n - c0 : RES_CHUNK);
  for (int t = threadIdx.x; t < nr * RES_CHUNK; t += 256) {
    const int q = t / RES_CHUNK, c = t % RES_CHUNK;
    xs[q][c] = c < nc ? X[c0 + c + q * ldx] : 0.0;
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  double acc[RES_NR];
#pragma unroll
  for (int q = 0; q < RES_NR; ++q) acc[q] = 0.0;
  const double* a = A + i + c0 * lda;
  int c = 0;
  // 16 independent column loads in flight per thread, then their products
  for (; c + 16 <= nc; c += 16) {
    double v[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) v[t] = a[(c + t) * lda];
#pragma unroll
    for (int t = 0; t < 16; ++t)
#pragma unroll
      for (int q = 0; q < RES_NR; ++q)
        if (q < nr) acc[q] += v[t] * xs[q][c + t];
  }
  for (; c < nc; ++c) {
    const double v = a[c * lda];
#pragma unroll
    for (int q = 0; q < RES_NR; ++q)
      if (q < nr) acc[q] += v * xs[q][c];
  }
  for (int q = 0; q < nr; ++q) part[((int64_t)blockIdx.y * nr + q) * m + i] = acc[q];
}

// R[i, q] = B[i, q] - sum over chunks of the partials, in chunk order
__global__ void __launch_bounds__(256) gemv_nr_reduce_kernel(const double* __restrict__ part, int64_t m, int nchunk,
                                                             int nr, const double* __restrict__ B, int64_t ldb,
                                                             double* __restrict__ R, int64_t ldr) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= m * nr) return;
  const int64_t q = t / m, i = t - q * m;
  double s = 0.0;
  for (int k = 0; k < nchunk; ++k) s += part[((int64_t)k * nr + q) * m + i];
  R[i + q * ldr] = B[i + q * ldb] - s;
}

// ---- row permutations from pivot sequences, one workgroup per system ---------------------------------------
// perm[i] = row of A that ends in row i of P A: the pivots' interchanges applied in order to the identity.  The
// sequence is inherently serial; it runs in LDS (the pivots staged first), one thread, all systems in parallel.
__global__ void __launch_bounds__(256) pivots_to_perm_kernel(const PermJob* __restrict__ jobs) {
  extern __shared__ int32_t psh[];
  const PermJob J = jobs[blockIdx.x];
  const int n = (int)J.n;
  int32_t* p = psh;
  int32_t* v = psh + n;
  for (int i = threadIdx.x; i < n; i += 256) {
    p[i] = i;
    v[i] = J.ipiv[i];
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 0; i < n; ++i) {
      const int j = v[i];
      const int32_t t = p[i];
      p[i] = p[j];
      p[j] = t;
    }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 256) J.perm[i] = p[i];
}

// ---- batched factor application: one workgroup per system -------------------------------------------------
// X (+)= U^-1 L^-1 P R for up to APPLY_NR columns of one system per workgroup.  The workgroup walks the
// 64-row blocks itself: the diagonal block is staged in LDS and solved by one wave per column, then all 1024
// threads update the remaining rows with the block's columns of the factor (coalesced column reads).  No other
// workgroup is involved, so there are no flags or waits, and many systems of uneven size share one launch.
constexpr int APPLY_NB = 64;
constexpr int APPLY_THREADS = 1024;

// W[i, q] -= sum_c F[i, k0 + c] xs[q][c] for one row i: a full 64-column block takes its factor entries in
// groups of 4 and the solved values as float4 LDS reads (one read serves four columns)
__device__ __forceinline__ void apply_row(const float* __restrict__ col, int64_t lda, int nb,
                                          const float (*xs)[APPLY_NB], float* W, int64_t i, int64_t n, int nr) {
  float acc[APPLY_NR];
#pragma unroll
  for (int q = 0; q < APPLY_NR; ++q) acc[q] = 0.f;
  if (nb == APPLY_NB) {
#pragma unroll 1
    for (int c0 = 0; c0 < APPLY_NB; c0 += 16) {
      float l[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) l[t] = col[(c0 + t) * lda];
#pragma unroll
      for (int t = 0; t < 16; t += 4)
#pragma unroll
        for (int q = 0; q < APPLY_NR; ++q) {
          const float4 xv = *reinterpret_cast<const float4*>(&xs[q][c0 + t]);
          acc[q] += l[t] * xv.x + l[t + 1] * xv.y + l[t + 2] * xv.z + l[t + 3] * xv.w;
        }
    }
  } else {
    for (int c = 0; c < nb; ++c) {
      const float l = col[c * lda];
#pragma unroll
      for (int q = 0; q < APPLY_NR; ++q) acc[q] += l * xs[q][c];
    }
  }
  for (int q = 0; q < nr; ++q) W[i + q * n] -= acc[q];
}

__global__ void __launch_bounds__(APPLY_THREADS) lu_apply_batched_kernel(const LuApplyJob* __restrict__ jobs) {
  __shared__ float D[APPLY_NB][APPLY_NB + 1];
  __shared__ __align__(16) float xs[APPLY_NR][APPLY_NB];
  const LuApplyJob J = jobs[blockIdx.x];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t n = J.n, lda = J.lda;
  const int nr = J.nrhs;
  const float* __restrict__ LU = J.LU;
  float* W = J.W;
  for (int64_t t = tid; t < n * nr; t += APPLY_THREADS) {
    const int64_t q = t / n, i = t - q * n;
    W[i + q * n] = (float)J.R[(int64_t)J.perm[i] + q * n];
  }
  __syncthreads();
  // forward: L unit lower
  for (int64_t k0 = 0; k0 < n; k0 += APPLY_NB) {
    const int nb = (int)(n - k0 < APPLY_NB ? n - k0 : APPLY_NB);
    for (int idx = tid; idx < nb * nb; idx += APPLY_THREADS) {
      const int r = idx % nb, c = idx / nb;
      D[r][c] = LU[(k0 + r) + (k0 + c) * lda];
    }
    __syncthreads();
    if (wave < nr) {
      float xi = lane < nb ? W[k0 + lane + wave * n] : 0.f;
      for (int r = 0; r < nb - 1; ++r) {
        const float xr = __shfl(xi, r);
        if (lane > r && lane < nb) xi -= D[lane][r] * xr;
      }
      if (lane < nb) {
        xs[wave][lane] = xi;
        W[k0 + lane + wave * n] = xi;
      }
    }
    __syncthreads();
    for (int64_t i = k0 + nb + tid; i < n; i += APPLY_THREADS) apply_row(LU + i + k0 * lda, lda, nb, xs, W, i, n, nr);
    __syncthreads();
  }
  // backward: U with its diagonal
  for (int64_t k0 = ((n - 1) / APPLY_NB) * APPLY_NB; k0 >= 0; k0 -= APPLY_NB) {
    const int nb = (int)(n - k0 < APPLY_NB ? n - k0 : APPLY_NB);
    for (int idx = tid; idx < nb * nb; idx += APPLY_THREADS) {
      const int r = idx % nb, c = idx / nb;
      D[r][c] = LU[(k0 + r) + (k0 + c) * lda];
    }
    __syncthreads();
    if (wave < nr) {
      float xi = lane < nb ? W[k0 + lane + wave * n] : 0.f;
      for (int r = nb - 1; r >= 0; --r) {
        if (lane == r) xi = xi / D[r][r];
        const float xr = __shfl(xi, r);
        if (lane < r) xi -= D[lane][r] * xr;
      }
      if (lane < nb) {
        xs[wave][lane] = xi;
        W[k0 + lane + wave * n] = xi;
      }
    }
    __syncthreads();
    for (int64_t i = tid; i < k0; i += APPLY_THREADS) apply_row(LU + i + k0 * lda, lda, nb, xs, W, i, n, nr);
    __syncthreads();
  }
  for (int64_t t = tid; t < n * nr; t += APPLY_THREADS) {
    const int64_t q = t / n, i = t - q * n;
    const double v = (double)W[i + q * n];
    J.X[i + q * n] = J.accumulate ? J.X[i + q * n] + v : v;
  }
}

// v = w / ||w|| with ||w||^2 read from device memory (zero vector when the norm is zero)
__global__ void scale_by_norm_kernel(const double* __restrict__ w, const double* __restrict__ nrm2,
                                     double* __restrict__ v, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hh = sqrt(*nrm2);
  v[i] = hh > 0.0 ? (1.0 / hh) * w[i] : 0.0;
}

}  // namespace

// ---- host-side building blocks -----------------------------------------------------------------------------
// Columns per partial sum: enough chunks that the launch has about 1024 workgroups (a small matrix otherwise leaves
// most CUs idle and the product latency-bound), between 64 and 1024 columns.
static int gemv_chunk(int64_t m, int64_t n) {
  const int64_t rb = (m + 255) / 256;
  const int64_t want = (1024 + rb - 1) / rb;
  int64_t chunk = (n + want - 1) / want;
  chunk = (chunk + 63) / 64 * 64;
  return (int)std::min<int64_t>(1024, std::max<int64_t>(64, chunk));
}

size_t gemv_part_bytes(int64_t m, int64_t n) {
  const int chunk = gemv_chunk(m, n);
  return sizeof(double) * (size_t)std::max<int64_t>(m, 1) * (size_t)std::max<int64_t>((n + chunk - 1) / chunk, 1);
}

template <class TA>
void gemv_splitk(const TA* A, int64_t m, int64_t n, int64_t lda, const double* x, double* y, double alpha,
                 const double* yin, double beta, double* part, hipStream_t st) {
  if (m <= 0) return;
  const int chunk = gemv_chunk(m, n);
  const int nchunk = (int)((n + chunk - 1) / chunk);
  if (nchunk > 0) {
    if (std::is_same<TA, float>::value && lda % 2 == 0 && reinterpret_cast<uintptr_t>(A) % 8 == 0)
      hipLaunchKernelGGL(gemv_part2_f32_kernel, dim3(nblk(m, 512), (unsigned)nchunk), dim3(256), 0, st,
                         reinterpret_cast<const float*>(A), m, n, lda, x, chunk, part);
    else
      hipLaunchKernelGGL(gemv_part_kernel<TA>, dim3(nblk(m, 256), (unsigned)nchunk), dim3(256), 0, st, A, m, n, lda,
                         x, chunk, part);
  }
  hipLaunchKernelGGL(gemv_reduce_kernel, dim3(nblk(m, 256)), dim3(256), 0, st, (const double*)part, m, nchunk, alpha,
                     yin, beta, y);
}

template <class T>
hipError_t trsv_persistent(const T* LU, int64_t n, int64_t lda, T* x, bool lower, TrsvState& s, hipStream_t st) {
  const int nb = (int)((n + TB - 1) / TB);
  s.epoch += 1;
  int epoch = s.epoch;
  long long lim = s.spin_limit;
  void* args[] = {(void*)&LU, &n, &lda, &x, &s.flags, &epoch, &lim, &s.err};
  if (lower)
    return hipLaunchCooperativeKernel((const void*)trsv_persistent_kernel<T, true>, dim3(nb), dim3(256), args, 0, st);
  return hipLaunchCooperativeKernel((const void*)trsv_persistent_kernel<T, false>, dim3(nb), dim3(256), args, 0, st);
}

template <class T>
int64_t trsv_persistent_max_n() {
  // one 64-row block per workgroup, all co-resident; thread-safe one-time query
  static const int64_t cap = [] {
    int dev = 0, cus = 0, lo = 0, up = 0;
    if (hipGetDevice(&dev) != hipSuccess) return (int64_t)0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return (int64_t)0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&lo, trsv_persistent_kernel<T, true>, 256, 0) != hipSuccess)
      return (int64_t)0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&up, trsv_persistent_kernel<T, false>, 256, 0) != hipSuccess)
      return (int64_t)0;
    return (int64_t)std::min(lo, up) * cus * TB;
  }();
  return cap;
}

template <class T>
hipError_t apply_lu_precond(const T* LU, int64_t n, int64_t lda, const int32_t* perm, const double* r, double* z,
                            T* work, TrsvState& s, hipStream_t st) {
  hipLaunchKernelGGL((permute_cast_kernel<double, T>), dim3(nblk(n, 256)), dim3(256), 0, st, perm, r, work, n);
  hipError_t e = trsv_persistent<T>(LU, n, lda, work, true, s, st);
  if (e != hipSuccess) return e;
  e = trsv_persistent<T>(LU, n, lda, work, false, s, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((cast_kernel<T, double>), dim3(nblk(n, 256)), dim3(256), 0, st, (const T*)work, z, n);
  return hipGetLastError();
}

// ---- many right-hand sides ----------------------------------------------------------------------------------
int64_t pivots_to_perm_max() {
  static const int64_t nmax = [] {
    const int64_t n = PERM_MAX;
    hipError_t r = hipFuncSetAttribute((const void*)pivots_to_perm_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(2 * sizeof(int32_t) * n));
    return r == hipSuccess ? n : (int64_t)0;
  }();
  return nmax;
}

hipError_t pivots_to_perm(const PermJob* jobs, int njobs, int64_t nmax, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  if (nmax > pivots_to_perm_max()) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pivots_to_perm_kernel, dim3((unsigned)njobs), dim3(256), 2 * sizeof(int32_t) * nmax, st, jobs);
  return hipGetLastError();
}

hipError_t lu_apply_batched(const LuApplyJob* jobs, int njobs, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(lu_apply_batched_kernel, dim3((unsigned)njobs), dim3(APPLY_THREADS), 0, st, jobs);
  return hipGetLastError();
}

void colnorm2(const double* R, int64_t ld, int64_t n, int64_t nrhs, double* out, hipStream_t st) {
  if (nrhs <= 0) return;
  hipLaunchKernelGGL(colnorm2_kernel, dim3((unsigned)nrhs), dim3(256), 0, st, R, ld, n, out);
}

size_t residual_multi_bytes(int64_t n, int64_t nrhs) {
  const int64_t nq = std::min<int64_t>(std::max<int64_t>(nrhs, 1), RES_NR);
  return sizeof(double) * (size_t)std::max<int64_t>(n, 1) * (size_t)nq *
         (size_t)std::max<int64_t>((n + RES_CHUNK - 1) / RES_CHUNK, 1);
}

hipError_t residual_multi(const double* A, int64_t n, int64_t lda, const double* B, int64_t ldb, const double* X,
                          int64_t ldx, double* R, int64_t ldr, int64_t nrhs, double* part, double* nrm2,
                          hipStream_t st) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  const int nchunk = (int)((n + RES_CHUNK - 1) / RES_CHUNK);
  for (int64_t q0 = 0; q0 < nrhs; q0 += RES_NR) {
    const int nr = (int)std::min<int64_t>(RES_NR, nrhs - q0);
    hipLaunchKernelGGL(gemv_nr_part_kernel, dim3(nblk(n, 256), (unsigned)nchunk), dim3(256), 0, st, A, n, n, lda,
                       X + q0 * ldx, ldx, nr, part);
    hipLaunchKernelGGL(gemv_nr_reduce_kernel, dim3(nblk(n * nr, 256)), dim3(256), 0, st, (const double*)part, n,
                       nchunk, nr, B + q0 * ldb, ldb, R + q0 * ldr, ldr);
  }
  colnorm2(R, ldr, n, nrhs, nrm2, st);
  return hipGetLastError();
}

// ---- FGMRES-IR driver ---------------------------------------------------------------------------------------
namespace {

struct Dev {
  std::vector<void*> bufs;
  ~Dev() {
    for (void* p : bufs) (void)hipFree(p);
  }
  template <class T> T* get(size_t count, hipError_t& e) {
    void* p = nullptr;
    e = hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1));
    if (e == hipSuccess) bufs.push_back(p);
    return static_cast<T*>(p);
  }
};

hipError_t host_dot(const double* a, const double* b, int64_t n, double* part, std::vector<double>& hp, double& out,
                    hipStream_t st) {
  hipLaunchKernelGGL(dot_part_kernel, dim3(DOT_BLOCKS), dim3(256), 0, st, a, b, n, part);
  hipError_t e = hipMemcpyAsync(hp.data(), part, sizeof(double) * DOT_BLOCKS, hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(st);
  double s = 0.0;
  for (int k = 0; k < DOT_BLOCKS; ++k) s += hp[k];
  out = s;
  return e;
}

}  // namespace

hipError_t fgmres(const double* A, int64_t n, int64_t lda, const Precond& M, const double* b, double* x,
                  const GmresOpts& o, GmresResult& res, hipStream_t st) {
  hipError_t e = hipSuccess;
  Dev d;
  const int m = o.restart;
  if (m < 1 || m > 1023) return hipErrorInvalidValue;  // one workgroup reduces the m + 1 Hessenberg entries
  double* V = d.get<double>((size_t)n * (m + 1), e);
  if (e) return e;
  double* Z = d.get<double>((size_t)n * m, e);
  if (e) return e;
  double* w = d.get<double>(n, e);
  if (e) return e;
  double* r = d.get<double>(n, e);
  if (e) return e;
  double* part = d.get<double>(std::max<size_t>(gemv_part_bytes(n, n) / sizeof(double), DOT_BLOCKS * (size_t)(m + 1)), e);
  if (e) return e;
  double* hdev = d.get<double>(2 * (m + 1) + 1, e);  // CGS2 coefficients (two passes) and the new norm^2
  if (e) return e;
  std::vector<double> hcol(2 * (m + 1) + 1);
  std::vector<double> hp(DOT_BLOCKS * (size_t)(m + 1));
  const unsigned g = nblk(n, 256);

  double bnorm2 = 0.0;
  e = host_dot(b, b, n, part, hp, bnorm2, st);
  if (e) return e;
  const double bnorm = std::sqrt(bnorm2) > 0 ? std::sqrt(bnorm2) : 1.0;
  res.iterations = 0;
  res.converged = false;
  res.history.clear();
  // initial guess: x = M^-1 b
  if (o.zero_guess) {
    e = hipMemsetAsync(x, 0, sizeof(double) * n, st);
  } else {
    e = M(b, x);
  }
  if (e) return e;
  for (int outer = 0; outer < o.max_restarts; ++outer) {
    // r = b - A x  (FP64)
    gemv_splitk<double>(A, n, n, lda, x, r, -1.0, b, 1.0, part, st);
    double rr = 0.0;
    e = host_dot(r, r, n, part, hp, rr, st);
    if (e) return e;
    double beta = std::sqrt(rr);
    res.residual = beta / bnorm;
    res.history.push_back(res.residual);
    if (res.residual <= o.tol) {
      res.converged = true;
      break;
    }
    if (res.iterations >= o.max_iter) break;
    // V[:,0] = r / beta
    hipLaunchKernelGGL(axpby_kernel, dim3(g), dim3(256), 0, st, 1.0 / beta, (const double*)r, 0.0, V, n);
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), gvec(m + 1, 0.0);
    gvec[0] = beta;
    int k = 0;
    for (; k < m && res.iterations < o.max_iter; ++k) {
      res.iterations++;
      // z_k = M^-1 v_k ; w = A z_k
      e = M(V + (int64_t)k * n, Z + (int64_t)k * n);
      if (e) return e;
      if (o.A32)
        gemv_splitk<float>(o.A32, n, n, o.lda32, Z + (int64_t)k * n, w, 1.0, nullptr, 0.0, part, st);
      else
        gemv_splitk<double>(A, n, n, lda, Z + (int64_t)k * n, w, 1.0, nullptr, 0.0, part, st);
      // classical Gram-Schmidt, twice (CGS2) for FP64 orthogonality; coefficients and the new norm stay on the
      // device, and the Hessenberg column comes back in one copy
      double* h1 = hdev;
      double* h2 = hdev + (m + 1);
      double* nrm2 = hdev + 2 * (m + 1);
      for (int pass = 0; pass < 2; ++pass) {
        double* hp_dev = pass == 0 ? h1 : h2;
        hipLaunchKernelGGL(mdot_part_kernel, dim3(DOT_BLOCKS), dim3(256), 0, st, (const double*)V, (int64_t)n, k + 1,
                           (const double*)w, n, part);
        hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(64 * ((k + 1 + 63) / 64)), 0, st, (const double*)part,
                           k + 1, DOT_BLOCKS, hp_dev);
        hipLaunchKernelGGL(mgs_update_kernel, dim3(g), dim3(256), 0, st, (const double*)V, (int64_t)n, k + 1,
                           (const double*)hp_dev, w, n);
      }
      hipLaunchKernelGGL(dot_part_kernel, dim3(DOT_BLOCKS), dim3(256), 0, st, (const double*)w, (const double*)w, n,
                         part);
      hipLaunchKernelGGL(reduce_parts_kernel, dim3(1), dim3(64), 0, st, (const double*)part, 1, DOT_BLOCKS, nrm2);
      hipLaunchKernelGGL(scale_by_norm_kernel, dim3(g), dim3(256), 0, st, (const double*)w, (const double*)nrm2,
                         V + (int64_t)(k + 1) * n, n);
      e = hipMemcpyAsync(hcol.data(), hdev, sizeof(double) * (2 * (m + 1) + 1), hipMemcpyDeviceToHost, st);
      if (e) return e;
      e = hipStreamSynchronize(st);
      if (e) return e;
      for (int j = 0; j <= k; ++j) H[(size_t)j * m + k] += hcol[j] + hcol[m + 1 + j];
      const double hh = std::sqrt(hcol[2 * (m + 1)]);
      H[(size_t)(k + 1) * m + k] = hh;
      // Givens rotations on column k
      for (int j = 0; j < k; ++j) {
        const double t = cs[j] * H[(size_t)j * m + k] + sn[j] * H[(size_t)(j + 1) * m + k];
        H[(size_t)(j + 1) * m + k] = -sn[j] * H[(size_t)j * m + k] + cs[j] * H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = t;
      }
      const double a = H[(size_t)k * m + k], bb = H[(size_t)(k + 1) * m + k];
      const double rho = std::hypot(a, bb);
      cs[k] = rho > 0 ? a / rho : 1.0;
      sn[k] = rho > 0 ? bb / rho : 0.0;
      H[(size_t)k * m + k] = rho;
      H[(size_t)(k + 1) * m + k] = 0.0;
      gvec[k + 1] = -sn[k] * gvec[k];
      gvec[k] = cs[k] * gvec[k];
      const double est = std::fabs(gvec[k + 1]) / bnorm;
      if (est <= o.tol * 0.5 || hh == 0.0 || (o.A32 && std::fabs(gvec[k + 1]) <= o.inner_rel * beta)) {
        ++k;
        break;
      }
    }
    // solve the k x k triangular system and update x += Z y
    std::vector<double> yk(k, 0.0);
    for (int i = k - 1; i >= 0; --i) {
      double s = gvec[i];
      for (int j = i + 1; j < k; ++j) s -= H[(size_t)i * m + j] * yk[j];
      yk[i] = s / H[(size_t)i * m + i];
    }
    if (k > 0) {
      e = hipMemcpyAsync(hdev, yk.data(), sizeof(double) * k, hipMemcpyHostToDevice, st);
      if (e) return e;
      hipLaunchKernelGGL(combine_kernel, dim3(g), dim3(256), 0, st, (const double*)Z, (int64_t)n, k,
                         (const double*)hdev, x, n);
    }
  }
  if (!res.converged) {
    gemv_splitk<double>(A, n, n, lda, x, r, -1.0, b, 1.0, part, st);
    double rr = 0.0;
    e = host_dot(r, r, n, part, hp, rr, st);
    if (e) return e;
    res.residual = std::sqrt(rr) / bnorm;
    res.history.push_back(res.residual);
    res.converged = res.residual <= o.tol;
  }
  e = hipStreamSynchronize(st);
  if (e) return e;
  return hipGetLastError();
}

template <class TF>
hipError_t fgmres_ir(const double* A, int64_t n, int64_t lda, const TF* LU, int64_t ldlu, const int32_t* perm,
                     const double* b, double* x, const GmresOpts& o, GmresResult& res, hipStream_t st) {
  hipError_t e = hipSuccess;
  Dev d;
  TF* work = d.get<TF>(n, e);
  if (e) return e;
  int* flags = d.get<int>((n + TB - 1) / TB + 1, e);
  if (e) return e;
  int* err = d.get<int>(1, e);
  if (e) return e;
  e = hipMemsetAsync(flags, 0, sizeof(int) * ((n + TB - 1) / TB + 1), st);
  if (e) return e;
  e = hipMemsetAsync(err, 0, sizeof(int), st);
  if (e) return e;
  TrsvState ts{flags, err, 0, o.spin_limit};
  Precond M = [&](const double* r, double* z) { return apply_lu_precond<TF>(LU, n, ldlu, perm, r, z, work, ts, st); };
  e = fgmres(A, n, lda, M, b, x, o, res, st);
  if (e) return e;
  int herr = 0;
  e = hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e) return e;
  e = hipStreamSynchronize(st);
  if (e) return e;
  res.trsv_timeout = herr != 0;
  if (res.trsv_timeout) res.converged = false;
  return hipSuccess;
}

template void gemv_splitk<double>(const double*, int64_t, int64_t, int64_t, const double*, double*, double,
                                  const double*, double, double*, hipStream_t);
template void gemv_splitk<float>(const float*, int64_t, int64_t, int64_t, const double*, double*, double,
                                 const double*, double, double*, hipStream_t);
template hipError_t trsv_persistent<float>(const float*, int64_t, int64_t, float*, bool, TrsvState&, hipStream_t);
template hipError_t trsv_persistent<double>(const double*, int64_t, int64_t, double*, bool, TrsvState&, hipStream_t);
template int64_t trsv_persistent_max_n<float>();
template int64_t trsv_persistent_max_n<double>();
template hipError_t apply_lu_precond<float>(const float*, int64_t, int64_t, const int32_t*, const double*, double*,
                                            float*, TrsvState&, hipStream_t);
template hipError_t apply_lu_precond<double>(const double*, int64_t, int64_t, const int32_t*, const double*, double*,
                                             double*, TrsvState&, hipStream_t);
template hipError_t fgmres_ir<float>(const double*, int64_t, int64_t, const float*, int64_t, const int32_t*,
                                     const double*, double*, const GmresOpts&, GmresResult&, hipStream_t);
template hipError_t fgmres_ir<double>(const double*, int64_t, int64_t, const double*, int64_t, const int32_t*,
                                      const double*, double*, const GmresOpts&, GmresResult&, hipStream_t);

}  // namespace cdj
