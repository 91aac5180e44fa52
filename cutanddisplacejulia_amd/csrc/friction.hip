// Frictional slip on gfx950: the complementarity problem of SlipCalculator3D.jl's friction method and a
// Fischer-Newton solver whose Newton systems are factored by the GPU LU.
//
// The reference calls FischerNewton 0.1.4 (not vendored); this is the published Fischer-Newton method of Erleben
// (num4lcp, 2013): Fischer-Burmeister function phi(y, x) = sqrt(y^2 + x^2) - y - x with y = a x + b, the
// perturbation strategy for the generalized Jacobian J = diag(p) + diag(q) a, a sufficient-descent test with the
// gradient J'phi as fallback, and a projected Armijo line search x <- max(0, x + tau dx).  The Newton system is
// solved exactly (LU) instead of by restarted GMRES.  Scalars (merit, descent products) are reduced in one
// workgroup in a fixed order and read back once per trial step.
#include <hip/hip_runtime.h>

#include <cmath>

#include "friction.hpp"
#include "linalg.hpp"
#include "mixed.hpp"

namespace cdj {

namespace {

unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// one thread per entry of a
__global__ void __launch_bounds__(256) lcp_matrix_kernel(const double* __restrict__ C, int64_t n,
                                                         const double* __restrict__ mu, double* __restrict__ a) {
  const int64_t N5 = 5 * n, N3 = 3 * n;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= N5 * N5) return;
  const int64_t i = t % N5, j = t / N5;
  const int R = (int)(i / n), B = (int)(j / n);
  const int64_t ii = i % n, jj = j % n;
  double v = 0.0;
  if (R < 3) {
    const int64_t r = R * n + ii;
    if (B == 0)
      v = (C[r + jj * N3] - C[r + (n + jj) * N3] * mu[jj]) - C[r + (2 * n + jj) * N3] * mu[jj];
    else if (B == 1 || B == 2)
      v = C[r + (B * n + jj) * N3];
    else if (B == R + 2)  // the identity blocks of rows Dss (column block 3) and Dds (column block 4)
      v = ii == jj ? 1.0 : 0.0;
  } else {
    if (B == 0)
      v = ii == jj ? 2.0 * mu[jj] : 0.0;
    else if (B == R - 2)  // rows 3 and 4: -I in column blocks 1 and 2
      v = ii == jj ? -1.0 : 0.0;
  }
  a[i + j * N5] = v;
}

__global__ void __launch_bounds__(256) lcp_vector_kernel(const double* __restrict__ C, int64_t n,
                                                         const double* __restrict__ D, const double* __restrict__ Sf,
                                                         double* __restrict__ bb) {
  const int64_t N3 = 3 * n;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= 5 * n) return;
  if (i >= N3) {
    bb[i] = 2.0 * Sf[i % n];
    return;
  }
  double s1 = 0.0, s2 = 0.0;
  for (int64_t jj = 0; jj < n; ++jj) {
    s1 += C[i + (n + jj) * N3] * Sf[jj];
    s2 += C[i + (2 * n + jj) * N3] * Sf[jj];
  }
  bb[i] = (-s1 - s2) + D[i];
}

__device__ __forceinline__ double fb(double y, double x) { return sqrt(y * y + x * x) - y - x; }

__global__ void fb_kernel(const double* __restrict__ y, const double* __restrict__ x, int64_t N,
                          double* __restrict__ phi) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < N) phi[i] = fb(y[i], x[i]);
}

// perturbed generalized-Jacobian diagonals p, q and the Newton right-hand side -phi
__global__ void jac_diag_kernel(const double* __restrict__ y, const double* __restrict__ x,
                                const double* __restrict__ phi, int64_t N, double gamma, double* __restrict__ p,
                                double* __restrict__ q, double* __restrict__ rhs) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= N) return;
  double px = x[i];
  if (fabs(phi[i]) < gamma && fabs(px) < gamma) px = gamma * (px < 0 ? -1.0 : 1.0);
  const double nr = sqrt(y[i] * y[i] + px * px);
  p[i] = px / nr - 1.0;
  q[i] = y[i] / nr - 1.0;
  rhs[i] = -phi[i];
}

// J = diag(p) + diag(q) a
__global__ void __launch_bounds__(256) jac_kernel(const double* __restrict__ a, const double* __restrict__ p,
                                                  const double* __restrict__ q, int64_t N, double* __restrict__ J) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= N * N) return;
  const int64_t i = t % N, j = t / N;
  const double v = q[i] * a[t];
  J[t] = i == j ? p[i] + v : v;
}

// nabla = J' phi = p phi + a' (q phi); one wave per column of a
__global__ void __launch_bounds__(256) nabla_kernel(const double* __restrict__ a, const double* __restrict__ p,
                                                    const double* __restrict__ q, const double* __restrict__ phi,
                                                    int64_t N, double* __restrict__ nabla) {
  const int64_t j = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (j >= N) return;
  double s = 0.0;
  for (int64_t i = lane; i < N; i += 64) s += a[i + j * N] * (q[i] * phi[i]);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) nabla[j] = p[j] * phi[j] + s;
}

// out[0] = sum u v (mode 0) or max |u| (mode 1); one workgroup, fixed order
__global__ void __launch_bounds__(1024) reduce_kernel(const double* __restrict__ u, const double* __restrict__ v,
                                                      int64_t N, int mode, double* __restrict__ out) {
  __shared__ double part[1024];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int64_t i = t; i < N; i += 1024) s = mode == 0 ? s + u[i] * v[i] : fmax(s, fabs(u[i]));
  part[t] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) part[t] = mode == 0 ? part[t] + part[t + w] : fmax(part[t], part[t + w]);
    __syncthreads();
  }
  if (t == 0) *out = part[0];
}

__global__ void project_kernel(const double* __restrict__ x, const double* __restrict__ dx, double tau, int64_t N,
                               double* __restrict__ xk) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < N) xk[i] = fmax(0.0, x[i] + dx[i] * tau);
}

__global__ void neg_copy_kernel(const double* __restrict__ s, int64_t N, double* __restrict__ d) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < N) d[i] = -s[i];
}

__global__ void slip_kernel(const double* __restrict__ y, const double* __restrict__ x, int64_t n, double s,
                            double* __restrict__ D3) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  D3[i] = y[i] * s;
  D3[n + i] = (y[n + i] - x[3 * n + i]) * s;
  D3[2 * n + i] = (y[2 * n + i] - x[4 * n + i]) * s;
}

// y = a x + bb (split-K product: part holds gemv_part_bytes(N, N))
void residual(const double* a, const double* bb, const double* x, int64_t N, double* y, double* part,
              hipStream_t st) {
  gemv_splitk<double>(a, N, N, N, x, y, 1.0, bb, 1.0, part, st);
}

__global__ void scale_kernel(double* __restrict__ A, int64_t count, double s) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < count) A[i] *= s;
}

}  // namespace

void scale_matrix(double* A, int64_t count, double s, hipStream_t st) {
  hipLaunchKernelGGL(scale_kernel, dim3(nblk(count, 256)), dim3(256), 0, st, A, count, s);
}

void build_friction_lcp(const double* C, int64_t n, const double* D, const double* mu, const double* Sf, double* a,
                        double* bb, hipStream_t st) {
  const int64_t N5 = 5 * n;
  hipLaunchKernelGGL(lcp_matrix_kernel, dim3(nblk(N5 * N5, 256)), dim3(256), 0, st, C, n, mu, a);
  hipLaunchKernelGGL(lcp_vector_kernel, dim3(nblk(N5, 256)), dim3(256), 0, st, C, n, D, Sf, bb);
}

__global__ void add_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t N) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < N) y[i] += x[i];
}

void add_vector(double* y, const double* x, int64_t N, hipStream_t st) {
  hipLaunchKernelGGL(add_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, y, x, N);
}

std::vector<int32_t> lcp_keep(const double* mu, const double* Sf, int64_t n) {
  std::vector<int32_t> keep;
  keep.reserve(5 * n);
  for (int k = 0; k < 5; ++k)
    for (int64_t i = 0; i < n; ++i)
      if (k == 0 || mu[i] != 0.0 || Sf[i] != 0.0) keep.push_back((int32_t)(k * n + i));
  return keep;
}

__global__ void __launch_bounds__(256) lcp_gather_kernel(const double* __restrict__ a, const double* __restrict__ bb,
                                                         int64_t N, const int32_t* __restrict__ keep, int64_t Nr,
                                                         double* __restrict__ ar, double* __restrict__ br) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= Nr * Nr) return;
  const int64_t i = t % Nr, j = t / Nr;
  ar[t] = a[keep[i] + (int64_t)keep[j] * N];
  if (j == 0) br[i] = bb[keep[i]];
}

__global__ void lcp_scatter_kernel(const double* __restrict__ xr, const int32_t* __restrict__ keep, int64_t Nr,
                                   double* __restrict__ x) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < Nr) x[keep[k]] = xr[k];
}

__global__ void fill_zero_kernel(double* __restrict__ x, int64_t N) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < N) x[i] = 0.0;
}

void lcp_gather(const double* a, const double* bb, int64_t N, const int32_t* keep, int64_t Nr, double* ar,
                double* br, hipStream_t st) {
  hipLaunchKernelGGL(lcp_gather_kernel, dim3(nblk(Nr * Nr, 256)), dim3(256), 0, st, a, bb, N, keep, Nr, ar, br);
}

void lcp_scatter(const double* xr, const int32_t* keep, int64_t Nr, double* x, int64_t N, hipStream_t st) {
  hipLaunchKernelGGL(fill_zero_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, x, N);
  hipLaunchKernelGGL(lcp_scatter_kernel, dim3(nblk(Nr, 256)), dim3(256), 0, st, xr, keep, Nr, x);
}

void build_friction_lcp_vector(const double* C, int64_t n, const double* D, const double* Sf, double* bb,
                               hipStream_t st) {
  hipLaunchKernelGGL(lcp_vector_kernel, dim3(nblk(5 * n, 256)), dim3(256), 0, st, C, n, D, Sf, bb);
}

hipError_t fischer_newton(const double* a, const double* bb, int64_t N, double* x, const FnParams& P, FnResult& R,
                          double* J, int32_t* ipiv, int* info, void* ws, double* vec, double* red, double* part,
                          int* iwork, hipStream_t st) {
  // Newton systems: J's LU, then the two triangular solves as block-chained persistent kernels (one launch each)
  int32_t* perm = iwork;
  int* flags = iwork + N;
  int* terr = flags + (N + 63) / 64 + 1;
  hipError_t e0 = hipMemsetAsync(flags, 0, sizeof(int) * ((N + 63) / 64 + 2), st);
  if (e0 != hipSuccess) return e0;
  TrsvState ts{flags, terr, 0, 1ll << 22};
  double* y = vec;
  double* phi = vec + N;
  double* p = vec + 2 * N;
  double* q = vec + 3 * N;
  double* dx = vec + 4 * N;
  double* nabla = vec + 5 * N;
  double* xk = vec + 6 * N;
  double* work = vec + 7 * N;
  const unsigned g = nblk(N, 256);
  hipError_t e = hipMemsetAsync(x, 0, sizeof(double) * N, st);
  if (e != hipSuccess) return e;
  auto merit = [&](const double* xx, double& f) -> hipError_t {
    residual(a, bb, xx, N, y, part, st);
    hipLaunchKernelGGL(fb_kernel, dim3(g), dim3(256), 0, st, (const double*)y, xx, N, phi);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, st, (const double*)phi, (const double*)phi, N, 0, red);
    double h = 0.0;
    hipError_t ee = hipMemcpyAsync(&h, red, sizeof(double), hipMemcpyDeviceToHost, st);
    if (ee == hipSuccess) ee = hipStreamSynchronize(st);
    f = 0.5 * h;
    return ee;
  };
  R = FnResult{};
  double err = 1e20;
  int it = 0;
  while (it < P.max_iter) {
    if ((e = merit(x, err)) != hipSuccess) return e;
    if (err < P.tol_abs) break;
    hipLaunchKernelGGL(jac_diag_kernel, dim3(g), dim3(256), 0, st, (const double*)y, (const double*)x,
                       (const double*)phi, N, P.gamma, p, q, dx);
    hipLaunchKernelGGL(jac_kernel, dim3(nblk(N * N, 256)), dim3(256), 0, st, a, (const double*)p, (const double*)q, N,
                       J);
    hipLaunchKernelGGL(nabla_kernel, dim3(nblk(N, 4)), dim3(256), 0, st, a, (const double*)p, (const double*)q,
                       (const double*)phi, N, nabla);
    if ((e = lu_factor<double>(J, N, N, ipiv, info, ws, st)) != hipSuccess) return e;
    int h_info = 0;
    if ((e = hipMemcpyAsync(&h_info, info, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    if (h_info < 0) return hipErrorLaunchTimeOut;  // a panel's wait overran: factors invalid
    bool grad = h_info > 0;                          // singular Newton system
    if (!grad) {
      if ((e = lu_permutation(ipiv, N, perm, st)) != hipSuccess) return e;
      if ((e = apply_lu_precond<double>(J, N, N, perm, dx, dx, work, ts, st)) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, st, (const double*)dx, (const double*)dx, N, 1, red);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, st, (const double*)nabla, (const double*)nabla, N, 0,
                       red + 1);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, st, (const double*)nabla, (const double*)dx, N, 0,
                       red + 2);
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, st, (const double*)dx, (const double*)dx, N, 0, red + 3);
    double s[4];
    int h_terr = 0;
    if ((e = hipMemcpyAsync(s, red, sizeof(s), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(&h_terr, terr, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
    if (h_terr) return hipErrorLaunchTimeOut;  // a triangular solve's block wait overran: dx invalid
    const double dxmax = s[0], nn = s[1];
    double ndx = s[2];
    if (!grad && (dxmax < 2.220446049250313e-16 || std::sqrt(nn) < P.tol_abs)) break;
    if (grad || ndx > -P.rho * s[3]) {
      hipLaunchKernelGGL(neg_copy_kernel, dim3(g), dim3(256), 0, st, (const double*)nabla, N, dx);
      ndx = -nn;
      ++R.gradient_steps;
    }
    double tau = 1.0, fk = 0.0;
    const double f0 = err, gf = P.beta * ndx;
    for (;;) {
      hipLaunchKernelGGL(project_kernel, dim3(g), dim3(256), 0, st, (const double*)x, (const double*)dx, tau, N, xk);
      if ((e = merit(xk, fk)) != hipSuccess) return e;
      if (fk <= f0 + tau * gf || tau * tau < P.gamma) break;
      tau *= P.alpha;
    }
    if ((e = hipMemcpyAsync(x, xk, sizeof(double) * N, hipMemcpyDeviceToDevice, st)) != hipSuccess) return e;
    err = fk;
    ++it;
  }
  R.iterations = it;
  R.err = err;
  return hipGetLastError();
}

void friction_slip(const double* a, const double* bb, const double* x, int64_t n, double scale, double* y, double* D3,
                   double* part, hipStream_t st) {
  residual(a, bb, x, 5 * n, y, part, st);
  hipLaunchKernelGGL(slip_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, (const double*)y, x, n, scale, D3);
}

}  // namespace cdj
