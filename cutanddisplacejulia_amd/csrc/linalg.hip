// Blocked right-looking LU with partial pivoting (the dgetrf behind `A\b`, SlipCalculator3D.jl:22) and
// blocked triangular solves, for gfx950.
//
//   panel   : one cooperative launch per 64-column panel.  The panel rows are split over the workgroups and held
//             in LDS; each column needs one grid barrier: every workgroup publishes its local pivot candidate
//             (value, row, row contents) and the owner of the diagonal row publishes that row, so after the
//             barrier all workgroups know the pivot, swap locally and apply the rank-1 update in LDS.
//   laswp   : the panel's row interchanges applied to the columns outside the panel.
//   trsm    : U12 = L11^-1 A12 with L11 staged in LDS.
//   update  : A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32), 64x64 tiles, K = 64
//             staged in LDS.  The product is formed as (U^T L^T)^T so each lane's accumulator holds a column
//             segment of A22, making the read-modify-write of A22 column-contiguous.
// Matrices may carry extra right-hand-side columns (ncols > n): they are swapped, solved and updated with the
// rest, so factoring [A | B] leaves L^-1 P B in place of B.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "linalg.hpp"

namespace cg = cooperative_groups;

namespace cdj {

namespace {

constexpr int PT = 256;        // panel threads
constexpr int MAX_P = 1024;    // max panel workgroups
constexpr int MAX_RPB = 128;   // max panel rows per workgroup (LDS: 128 x 64 x 8 = 64 KiB)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: A operand lane l = A[l % 16][l / 16], B operand lane l = B[l / 16][l % 16].  The accumulator
// layouts differ: f32 item v of lane l is D[4 (l / 16) + v][l % 16]; f64 item v of lane l is D[4 v + l / 16][l % 16].
// row(g, v) gives the D row of item v for lane group g = l / 16.
template <class T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4 acc;
  __device__ static acc step(double a, double b, acc c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int g, int v) { return 4 * v + g; }
};
template <> struct Mfma<float> {
  typedef f4 acc;
  __device__ static acc step(float a, float b, acc c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int g, int v) { return 4 * g + v; }
};

template <class T>
__global__ void __launch_bounds__(PT) lu_panel_kernel(T* A, int64_t lda, int64_t k, int64_t m, int nbw, int rpb,
                                                      int32_t* ipiv, int* info, double* cval, int* crow, T* crows,
                                                      T* rowj) {
  extern __shared__ unsigned char smem_raw[];
  T* sh = reinterpret_cast<T*>(smem_raw);  // rpb x LU_NB, row-major
  __shared__ T u[LU_NB];
  __shared__ double redv[PT];
  __shared__ int redr[PT];
  __shared__ int s_w, s_prow;
  cg::grid_group grid = cg::this_grid();
  const int P = gridDim.x, b = blockIdx.x, tid = threadIdx.x;
  const int64_t r0 = (int64_t)b * rpb;
  const int64_t r1 = r0 + rpb < m ? r0 + rpb : m;
  const int nr = r1 > r0 ? (int)(r1 - r0) : 0;
  for (int idx = tid; idx < nr * nbw; idx += PT) {
    const int r = idx % nr, c = idx / nr;
    sh[r * LU_NB + c] = A[(k + r0 + r) + (k + c) * lda];
  }
  __syncthreads();
  for (int j = 0; j < nbw; ++j) {
    const int par = j & 1;
    double bv = -1.0;
    int br = -1;
    for (int r = tid; r < nr; r += PT) {
      const int gr = (int)(r0 + r);
      if (gr < j) continue;
      const double v = fabs((double)sh[r * LU_NB + j]);
      if (v > bv) {
        bv = v;
        br = gr;
      }
    }
    redv[tid] = bv;
    redr[tid] = br;
    __syncthreads();
    for (int s = PT / 2; s > 0; s >>= 1) {
      if (tid < s) {
        const double v2 = redv[tid + s];
        const int r2 = redr[tid + s];
        if (r2 >= 0 && (redr[tid] < 0 || v2 > redv[tid] || (v2 == redv[tid] && r2 < redr[tid]))) {
          redv[tid] = v2;
          redr[tid] = r2;
        }
      }
      __syncthreads();
    }
    const int lbest = redr[0];
    if (tid == 0) {
      cval[par * P + b] = redv[0];
      crow[par * P + b] = lbest;
    }
    if (lbest >= 0)
      for (int c = tid; c < nbw; c += PT) crows[(int64_t)(par * P + b) * LU_NB + c] = sh[(lbest - r0) * LU_NB + c];
    if (j >= r0 && j < r1)
      for (int c = tid; c < nbw; c += PT) rowj[par * LU_NB + c] = sh[(j - r0) * LU_NB + c];
    __threadfence();
    grid.sync();
    if (tid == 0) {
      double best = -1.0;
      int w = -1, prow = -1;
      for (int q = 0; q < P; ++q) {
        const int rq = crow[par * P + q];
        if (rq < 0) continue;
        const double v = cval[par * P + q];
        if (w < 0 || v > best || (v == best && rq < prow)) {
          best = v;
          w = q;
          prow = rq;
        }
      }
      s_w = w;
      s_prow = prow;
      if (b == 0) {
        ipiv[k + j] = (int32_t)(k + prow);
        if (!(best > 0.0) && *info == 0) *info = (int)(k + j + 1);
      }
    }
    __syncthreads();
    const int w = s_w, prow = s_prow;
    for (int c = tid; c < nbw; c += PT) u[c] = crows[(int64_t)(par * P + w) * LU_NB + c];
    __syncthreads();
    if (prow != j) {
      if (prow >= r0 && prow < r1)
        for (int c = tid; c < nbw; c += PT) sh[(prow - r0) * LU_NB + c] = rowj[par * LU_NB + c];
      if (j >= r0 && j < r1)
        for (int c = tid; c < nbw; c += PT) sh[(j - r0) * LU_NB + c] = u[c];
    }
    __syncthreads();
    const T piv = u[j];
    if (piv != T(0))
      for (int r = tid; r < nr; r += PT)
        if (r0 + r > j) sh[r * LU_NB + j] = sh[r * LU_NB + j] / piv;
    __syncthreads();
    const int ncol = nbw - j - 1;
    if (ncol > 0)
      for (int idx = tid; idx < nr * ncol; idx += PT) {
        const int r = idx / ncol, c = j + 1 + idx % ncol;
        if (r0 + r > j) sh[r * LU_NB + c] -= sh[r * LU_NB + j] * u[c];
      }
    __syncthreads();
  }
  for (int idx = tid; idx < nr * nbw; idx += PT) {
    const int r = idx % nr, c = idx / nr;
    A[(k + r0 + r) + (k + c) * lda] = sh[r * LU_NB + c];
  }
}

// Row interchanges of one panel applied to columns [0, k) and [k + nbw, ncols).
template <class T>
__global__ void laswp_kernel(T* A, int64_t lda, int64_t ncols, int64_t k, int nbw, const int32_t* __restrict__ ipiv) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncols || (c >= k && c < k + nbw)) return;
  T* col = A + c * lda;
  for (int j = 0; j < nbw; ++j) {
    const int64_t p = ipiv[k + j];
    if (p != k + j) {
      const T t = col[k + j];
      col[k + j] = col[p];
      col[p] = t;
    }
  }
}

// A[k:k+nbw, c0:c0+ncols] = L11^-1 A[k:k+nbw, c0:...] with L11 unit lower (the panel's diagonal block).
template <class T>
__global__ void __launch_bounds__(256) trsm_kernel(T* A, int64_t lda, int64_t k, int nbw, int64_t c0, int64_t ncols) {
  __shared__ T L[LU_NB][LU_NB + 1];
  __shared__ T X[LU_NB][64 + 1];
  const int tid = threadIdx.x;
  const int64_t cb = c0 + (int64_t)blockIdx.x * 64;
  const int nc = (int)std::min<int64_t>(64, c0 + ncols - cb);
  for (int idx = tid; idx < nbw * nbw; idx += 256) {
    const int r = idx % nbw, c = idx / nbw;
    L[r][c] = A[(k + r) + (k + c) * lda];
  }
  for (int idx = tid; idx < nbw * nc; idx += 256) {
    const int r = idx % nbw, c = idx / nbw;
    X[r][c] = A[(k + r) + (cb + c) * lda];
  }
  __syncthreads();
  if (tid < nc) {
    for (int r = 1; r < nbw; ++r) {
      T s = X[r][tid];
      for (int q = 0; q < r; ++q) s -= L[r][q] * X[q][tid];
      X[r][tid] = s;
    }
  }
  __syncthreads();
  for (int idx = tid; idx < nbw * nc; idx += 256) {
    const int r = idx % nbw, c = idx / nbw;
    A[(k + r) + (cb + c) * lda] = X[r][c];
  }
}

// C[0:M, 0:N] -= L[0:M, 0:K] * U[0:K, 0:N], all column-major with leading dimension lda, K <= 64.
template <class T>
__global__ void __launch_bounds__(256) gemm_update_kernel(const T* __restrict__ L, const T* __restrict__ U,
                                                          T* __restrict__ C, int64_t lda, int64_t M, int64_t N, int K) {
  __shared__ T Ls[LU_NB][64 + 1];  // Ls[k][i] = L[i0 + i, k]
  __shared__ T Us[LU_NB][64 + 1];  // Us[k][j] = U[k, j0 + j]
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * 64, j0 = (int64_t)blockIdx.y * 64;
  for (int idx = tid; idx < 64 * LU_NB; idx += 256) {
    const int i = idx % 64, kk = idx / 64;
    Ls[kk][i] = (kk < K && i0 + i < M) ? L[(i0 + i) + (int64_t)kk * lda] : T(0);
  }
  for (int idx = tid; idx < 64 * LU_NB; idx += 256) {
    const int kk = idx % LU_NB, j = idx / LU_NB;
    Us[kk][j] = (kk < K && j0 + j < N) ? U[kk + (j0 + j) * lda] : T(0);
  }
  __syncthreads();
  const int wave = tid / 64, lane = tid % 64;
  const int wr = (wave % 2) * 32, wc = (wave / 2) * 32;
  const int li = lane % 16, lk = lane / 16;
  typename Mfma<T>::acc acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b2 = 0; b2 < 2; ++b2) acc[a][b2] = typename Mfma<T>::acc{0, 0, 0, 0};
  const int Kr = (K + 3) & ~3;
  for (int kk = 0; kk < Kr; kk += 4) {
    const T l0 = Ls[kk + lk][wr + li], l1 = Ls[kk + lk][wr + 16 + li];
    const T u0 = Us[kk + lk][wc + li], u1 = Us[kk + lk][wc + 16 + li];
    acc[0][0] = Mfma<T>::step(u0, l0, acc[0][0]);
    acc[0][1] = Mfma<T>::step(u0, l1, acc[0][1]);
    acc[1][0] = Mfma<T>::step(u1, l0, acc[1][0]);
    acc[1][1] = Mfma<T>::step(u1, l1, acc[1][1]);
  }
  // acc[tj][ti][v] = (L U)[row = wr + 16 ti + li, col = wc + 16 tj + Mfma::row(lk, v)]
#pragma unroll
  for (int tj = 0; tj < 2; ++tj)
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) {
      const int64_t row = i0 + wr + 16 * ti + li;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t col = j0 + wc + 16 * tj + Mfma<T>::row(lk, v);
        if (row < M && col < N) C[row + col * lda] -= acc[tj][ti][v];
      }
    }
}

// ---- triangular solves with nrhs <= 64 right-hand sides -------------------------------------------------
// Diagonal-block solve of rows [r0, r0+nb) for all right-hand sides: lower (unit) or upper.
template <class T, bool LOWER>
__global__ void __launch_bounds__(256) trsv_diag_kernel(const T* __restrict__ A, int64_t lda, int64_t r0, int nb,
                                                        T* B, int64_t ldb, int64_t nrhs) {
  __shared__ T D[LU_NB][LU_NB + 1];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int r = idx % nb, c = idx / nb;
    D[r][c] = A[(r0 + r) + (r0 + c) * lda];
  }
  __syncthreads();
  for (int64_t q = tid; q < nrhs; q += 256) {
    T* x = B + q * ldb + r0;
    if (LOWER) {
      for (int r = 1; r < nb; ++r) {
        T s = x[r];
        for (int c = 0; c < r; ++c) s -= D[r][c] * x[c];
        x[r] = s;
      }
    } else {
      for (int r = nb - 1; r >= 0; --r) {
        T s = x[r];
        for (int c = r + 1; c < nb; ++c) s -= D[r][c] * x[c];
        x[r] = s / D[r][r];
      }
    }
  }
}

// B[rows] -= A[rows, c0:c0+nb] * B[c0:c0+nb] for rows in [rb, re).
template <class T>
__global__ void __launch_bounds__(256) trsv_update_kernel(const T* __restrict__ A, int64_t lda, int64_t c0, int nb,
                                                          int64_t rb, int64_t re, T* B, int64_t ldb, int64_t nrhs) {
  const int64_t i = rb + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= re) return;
  for (int64_t q = 0; q < nrhs; ++q) {
    const T* x = B + q * ldb + c0;
    T s = T(0);
    for (int c = 0; c < nb; ++c) s += A[i + (c0 + c) * lda] * x[c];
    B[i + q * ldb] -= s;
  }
}

template <class T>
__global__ void permute_kernel(const int32_t* __restrict__ perm, const T* __restrict__ in, T* __restrict__ out, int64_t n,
                               int64_t nrhs, int64_t ld) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * nrhs) return;
  const int64_t q = t / n, i = t % n;
  out[i + q * ld] = in[perm[i] + q * ld];
}

template <class TA>
__global__ void __launch_bounds__(256) gemv_kernel(const TA* __restrict__ A, int64_t m, int64_t n, int64_t lda,
                                                   const double* __restrict__ x, double* __restrict__ y, double alpha,
                                                   double beta) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int64_t c = 0; c < n; ++c) s += (double)A[i + c * lda] * x[c];
  y[i] = alpha * s + (beta == 0.0 ? 0.0 : beta * y[i]);
}

__global__ void d2f_kernel(const double* a, float* b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void f2d_kernel(const float* a, double* b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
__global__ void d2f_matrix_kernel(const double* A, int64_t lda, float* B, int64_t ldb, int64_t m, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t c = blockIdx.y;
  if (i < m && c < n) B[i + c * ldb] = (float)A[i + c * lda];
}

unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

size_t lu_workspace_bytes(int64_t n) {
  (void)n;
  return sizeof(double) * (2 * MAX_P + 2 * (size_t)MAX_P * LU_NB + 2 * LU_NB) + sizeof(int) * 2 * MAX_P + 1024;
}

template <class T>
hipError_t lu_factor_cols(T* A, int64_t n, int64_t ncols, int64_t lda, int32_t* ipiv, int* info, void* ws,
                          hipStream_t st) {
  char* w = static_cast<char*>(ws);
  double* cval = reinterpret_cast<double*>(w);
  w += sizeof(double) * 2 * MAX_P;
  T* crows = reinterpret_cast<T*>(w);
  w += sizeof(double) * 2 * (size_t)MAX_P * LU_NB;
  T* rowj = reinterpret_cast<T*>(w);
  w += sizeof(double) * 2 * LU_NB;
  int* crow = reinterpret_cast<int*>(w);
  hipError_t e = hipMemsetAsync(info, 0, sizeof(int), st);
  if (e != hipSuccess) return e;
  for (int64_t k = 0; k < n; k += LU_NB) {
    const int nbw = (int)std::min<int64_t>(LU_NB, n - k);
    const int64_t m = n - k;
    int P = (int)std::min<int64_t>(256, (m + 63) / 64);
    int rpb = (int)((m + P - 1) / P);
    if (rpb > MAX_RPB) {
      P = (int)((m + MAX_RPB - 1) / MAX_RPB);
      rpb = (int)((m + P - 1) / P);
    }
    if (P > MAX_P) return hipErrorInvalidValue;
    size_t shm = sizeof(T) * (size_t)rpb * LU_NB;
    void* args[] = {&A, (void*)&lda, &k, (void*)&m, (void*)&nbw, &rpb, &ipiv, &info, &cval, &crow, &crows, &rowj};
    e = hipLaunchCooperativeKernel((const void*)lu_panel_kernel<T>, dim3(P), dim3(PT), args, shm, st);
    if (e != hipSuccess) return e;
    if (ncols > nbw) {
      hipLaunchKernelGGL(laswp_kernel<T>, dim3(nblk(ncols, 256)), dim3(256), 0, st, A, lda, ncols, k, nbw, ipiv);
    }
    const int64_t c0 = k + nbw;
    const int64_t nrest = ncols - c0;
    if (nrest > 0) {
      hipLaunchKernelGGL(trsm_kernel<T>, dim3(nblk(nrest, 64)), dim3(256), 0, st, A, lda, k, nbw, c0, nrest);
      const int64_t M = n - c0;
      if (M > 0)
        hipLaunchKernelGGL(gemm_update_kernel<T>, dim3(nblk(M, 64), nblk(nrest, 64)), dim3(256), 0, st,
                           (const T*)(A + c0 + k * lda), (const T*)(A + k + c0 * lda), A + c0 + c0 * lda, lda, M,
                           nrest, nbw);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <class T>
hipError_t lu_factor(T* A, int64_t n, int64_t lda, int32_t* ipiv, int* info, void* ws, hipStream_t st) {
  return lu_factor_cols<T>(A, n, n, lda, ipiv, info, ws, st);
}

// Permutation of the rows produced by the pivots: row i of P A is row perm[i] of A.
static std::vector<int32_t> perm_from_ipiv(const std::vector<int32_t>& ipiv) {
  const int64_t n = (int64_t)ipiv.size();
  std::vector<int32_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
  for (int64_t i = 0; i < n; ++i) std::swap(perm[i], perm[ipiv[i]]);
  return perm;
}

hipError_t lu_permutation(const int32_t* ipiv, int64_t n, int32_t* perm_dev, hipStream_t st) {
  std::vector<int32_t> h(n);
  hipError_t e = hipMemcpyAsync(h.data(), ipiv, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  std::vector<int32_t> p = perm_from_ipiv(h);
  e = hipMemcpyAsync(perm_dev, p.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

template <class T>
hipError_t tri_solve_lower_unit(const T* LU, int64_t n, int64_t lda, T* B, int64_t nrhs, int64_t ldb, hipStream_t st) {
  for (int64_t r0 = 0; r0 < n; r0 += LU_NB) {
    const int nb = (int)std::min<int64_t>(LU_NB, n - r0);
    hipLaunchKernelGGL((trsv_diag_kernel<T, true>), dim3(1), dim3(256), 0, st, LU, lda, r0, nb, B, ldb, nrhs);
    const int64_t rb = r0 + nb;
    if (rb < n)
      hipLaunchKernelGGL(trsv_update_kernel<T>, dim3(nblk(n - rb, 256)), dim3(256), 0, st, LU, lda, r0, nb, rb, n, B,
                         ldb, nrhs);
  }
  return hipGetLastError();
}

template <class T>
hipError_t tri_solve_upper(const T* LU, int64_t n, int64_t lda, T* B, int64_t nrhs, int64_t ldb, hipStream_t st) {
  const int64_t nblocks = (n + LU_NB - 1) / LU_NB;
  for (int64_t bi = nblocks - 1; bi >= 0; --bi) {
    const int64_t r0 = bi * LU_NB;
    const int nb = (int)std::min<int64_t>(LU_NB, n - r0);
    hipLaunchKernelGGL((trsv_diag_kernel<T, false>), dim3(1), dim3(256), 0, st, LU, lda, r0, nb, B, ldb, nrhs);
    if (r0 > 0)
      hipLaunchKernelGGL(trsv_update_kernel<T>, dim3(nblk(r0, 256)), dim3(256), 0, st, LU, lda, r0, nb, (int64_t)0, r0,
                         B, ldb, nrhs);
  }
  return hipGetLastError();
}

template <class T>
hipError_t lu_solve(const T* LU, int64_t n, int64_t lda, const int32_t* ipiv, T* B, int64_t nrhs, int64_t ldb,
                    hipStream_t st) {
  int32_t* perm = nullptr;
  T* tmp = nullptr;
  hipError_t e = hipMalloc(&perm, sizeof(int32_t) * std::max<int64_t>(n, 1));
  if (e != hipSuccess) return e;
  e = hipMalloc(&tmp, sizeof(T) * std::max<int64_t>(ldb * nrhs, 1));
  if (e == hipSuccess) e = lu_permutation(ipiv, n, perm, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(permute_kernel<T>, dim3(nblk(n * nrhs, 256)), dim3(256), 0, st, perm, (const T*)B, tmp, n, nrhs,
                       ldb);
    e = hipMemcpyAsync(B, tmp, sizeof(T) * ldb * nrhs, hipMemcpyDeviceToDevice, st);
  }
  if (e == hipSuccess) e = tri_solve_lower_unit<T>(LU, n, lda, B, nrhs, ldb, st);
  if (e == hipSuccess) e = tri_solve_upper<T>(LU, n, lda, B, nrhs, ldb, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(perm);
  if (tmp) (void)hipFree(tmp);
  return e;
}

template <class TA>
void gemv(const TA* A, int64_t m, int64_t n, int64_t lda, const double* x, double* y, double alpha, double beta,
          hipStream_t st) {
  if (m <= 0) return;
  hipLaunchKernelGGL(gemv_kernel<TA>, dim3(nblk(m, 256)), dim3(256), 0, st, A, m, n, lda, x, y, alpha, beta);
}

void convert_d2f(const double* a, float* b, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(d2f_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, a, b, n);
}
void convert_f2d(const float* a, double* b, int64_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(f2d_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, a, b, n);
}
void convert_matrix_d2f(const double* A, int64_t lda, float* B, int64_t ldb, int64_t m, int64_t n, hipStream_t st) {
  if (m > 0 && n > 0) hipLaunchKernelGGL(d2f_matrix_kernel, dim3(nblk(m, 256), (unsigned)n), dim3(256), 0, st, A, lda, B, ldb, m, n);
}

template hipError_t lu_factor<double>(double*, int64_t, int64_t, int32_t*, int*, void*, hipStream_t);
template hipError_t lu_factor<float>(float*, int64_t, int64_t, int32_t*, int*, void*, hipStream_t);
template hipError_t lu_factor_cols<double>(double*, int64_t, int64_t, int64_t, int32_t*, int*, void*, hipStream_t);
template hipError_t lu_factor_cols<float>(float*, int64_t, int64_t, int64_t, int32_t*, int*, void*, hipStream_t);
template hipError_t lu_solve<double>(const double*, int64_t, int64_t, const int32_t*, double*, int64_t, int64_t, hipStream_t);
template hipError_t lu_solve<float>(const float*, int64_t, int64_t, const int32_t*, float*, int64_t, int64_t, hipStream_t);
template hipError_t tri_solve_upper<double>(const double*, int64_t, int64_t, double*, int64_t, int64_t, hipStream_t);
template hipError_t tri_solve_upper<float>(const float*, int64_t, int64_t, float*, int64_t, int64_t, hipStream_t);
template hipError_t tri_solve_lower_unit<double>(const double*, int64_t, int64_t, double*, int64_t, int64_t, hipStream_t);
template hipError_t tri_solve_lower_unit<float>(const float*, int64_t, int64_t, float*, int64_t, int64_t, hipStream_t);
template void gemv<double>(const double*, int64_t, int64_t, int64_t, const double*, double*, double, double, hipStream_t);
template void gemv<float>(const float*, int64_t, int64_t, int64_t, const double*, double*, double, double, hipStream_t);

}  // namespace cdj
