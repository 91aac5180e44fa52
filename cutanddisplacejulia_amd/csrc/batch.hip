// Batched assembly: the collocation matrices of many meshes of uneven size in one persistent launch.
//
// Work items are (mesh, block of 256 receivers, chunk of `tpb` sources); their count per mesh is
// ceil(n / 256) * ceil(n / tpb) and item_start holds the prefix sums.  Each workgroup pulls items from one
// atomic counter until the list is exhausted, so a batch of 200- and 3000-triangle meshes keeps every CU busy
// without a launch per mesh.  Per-triangle set-up (receivers, TD records) works on the concatenated triangle
// arrays of the batch.  The mixed mode runs the FP32 pass over every item and then the FP64 near-field pass over
// the same item list.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "td_device.hpp"
#include "td_kernels.hpp"

namespace cdj {

namespace {

__device__ __forceinline__ int find_mesh(const int64_t* __restrict__ item_start, int nmesh, int64_t item) {
  int lo = 0, hi = nmesh - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (item_start[mid] <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Next work item for the whole workgroup (uniform); nitems or more ends the loop.
__device__ __forceinline__ int64_t next_item(unsigned long long* counter) {
  __shared__ unsigned long long s_item;
  if (threadIdx.x == 0) s_item = atomicAdd(counter, 1ull);
  __syncthreads();
  const int64_t it = (int64_t)s_item;
  __syncthreads();
  return it;
}

// A_m = -[T blocks]: entry (receiver il, source j) of block (mode m, traction q) at
// A_m[(q n + il) + (m n + j) lda], lda = 3 n (SlipCalculator3D.jl:13-16).
// FP32 instances are held to two waves per SIMD (the half-space one then spills 12 bytes per lane instead of
// running one wave per SIMD); FP64 keeps the register budget of one wave.
template <class T, class TO, bool HS>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 2 : 1) batch_far_kernel(const double* __restrict__ recv, const T* __restrict__ rec,
                                                        const BatchMeshDev* __restrict__ meshes,
                                                        const int64_t* __restrict__ item_start, int nmesh, int tpb,
                                                        T nu, T lam, T G, TO* __restrict__ A,
                                                        unsigned long long* counter) {
  constexpr int RS = HS ? REC_HS : REC_FS;
  const int64_t nitems = item_start[nmesh];
  for (;;) {
    const int64_t item = next_item(counter);
    if (item >= nitems) break;
    const int mi = find_mesh(item_start, nmesh, item);
    const BatchMeshDev M = meshes[mi];
    const int64_t q = item - item_start[mi];
    const int64_t nrb = (M.n + 255) / 256;
    const int64_t il = (q % nrb) * 256 + threadIdx.x;
    if (il >= M.n) continue;
    const int64_t j0 = (q / nrb) * tpb;
    const int64_t j1 = j0 + tpb < M.n ? j0 + tpb : M.n;
    const double* rv = recv + (M.tri0 + il) * RECV;
    const T Xi = (T)rv[0], Yi = (T)rv[1], Zi = (T)rv[2];
    const T n[3] = {(T)rv[3], (T)rv[4], (T)rv[5]};
    const T s[3] = {(T)rv[6], (T)rv[7], (T)rv[8]};
    const T d[3] = {(T)rv[9], (T)rv[10], (T)rv[11]};
    const int64_t lda = 3 * M.n;
    TO* Am = A + M.a_off;
    for (int64_t j = j0; j < j1; ++j) {
      T t[3][3];
      pair_traction<T, HS>(rec + (M.tri0 + j) * RS, Xi, Yi, Zi, nu, lam, G, n, s, d, t);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        TO* col = Am + (m * M.n + j) * lda + il;
        col[0] = (TO)(-t[m][0]);
        col[M.n] = (TO)(-t[m][1]);
        col[2 * M.n] = (TO)(-t[m][2]);
      }
    }
  }
}

// FP64 recompute of the near pairs (near_pair over the sources' near records) over the FP32 values; a wave with no
// near lane for a source skips it.
template <bool HS>
__global__ void __launch_bounds__(256) batch_near_kernel(const double* __restrict__ recv,
                                                         const double* __restrict__ rec,
                                                         const double* __restrict__ near,
                                                         const BatchMeshDev* __restrict__ meshes,
                                                         const int64_t* __restrict__ item_start, int nmesh, int tpb,
                                                         double nu, double lam, double G, double* __restrict__ A,
                                                         unsigned long long* counter,
                                                         unsigned long long* __restrict__ near_count) {
  constexpr int RS = HS ? REC_HS : REC_FS;
  const int64_t nitems = item_start[nmesh];
  unsigned long long cnt = 0;
  for (;;) {
    const int64_t item = next_item(counter);
    if (item >= nitems) break;
    const int mi = find_mesh(item_start, nmesh, item);
    const BatchMeshDev M = meshes[mi];
    const int64_t q = item - item_start[mi];
    const int64_t nrb = (M.n + 255) / 256;
    const int64_t il = (q % nrb) * 256 + threadIdx.x;
    const bool valid = il < M.n;
    const int64_t j0 = (q / nrb) * tpb;
    const int64_t j1 = j0 + tpb < M.n ? j0 + tpb : M.n;
    const double* rv = recv + (M.tri0 + (valid ? il : 0)) * RECV;
    const double Xi = rv[0], Yi = rv[1], Zi = rv[2];
    const double n[3] = {rv[3], rv[4], rv[5]};
    const double s[3] = {rv[6], rv[7], rv[8]};
    const double d[3] = {rv[9], rv[10], rv[11]};
    const int64_t lda = 3 * M.n;
    double* Am = A + M.a_off;
    for (int64_t j = j0; j < j1; ++j) {
      const bool nr = valid && near_pair(near + (M.tri0 + j) * NEAR_REC, Xi, Yi, Zi);
      if (!__any(nr)) continue;
      double t[3][3];
      pair_traction<double, HS>(rec + (M.tri0 + j) * RS, Xi, Yi, Zi, nu, lam, G, n, s, d, t);
      if (nr) {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          double* col = Am + (m * M.n + j) * lda + il;
          col[0] = -t[m][0];
          col[M.n] = -t[m][1];
          col[2 * M.n] = -t[m][2];
        }
        ++cnt;
      }
    }
  }
  if (cnt) atomicAdd(near_count, cnt * 9ull);
}

// Counters are cleared by a kernel (not a memset node) so the sequence can be captured in a graph and replayed.
__global__ void batch_clear_kernel(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

}  // namespace

void launch_batch_clear(unsigned long long* p, int n, hipStream_t st) {
  hipLaunchKernelGGL(batch_clear_kernel, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
}

void launch_batch_far(const double* recv, const void* rec, bool fp32, int hs, const BatchMeshDev* meshes,
                      const int64_t* item_start, int nmesh, int tpb, int wgs, double nu, double lam, double G,
                      double* A, unsigned long long* counter, hipStream_t st) {
  const dim3 grid(wgs), blk(256);
  if (fp32) {
    const float* r = static_cast<const float*>(rec);
    if (hs)
      hipLaunchKernelGGL((batch_far_kernel<float, double, true>), grid, blk, 0, st, recv, r, meshes, item_start,
                         nmesh, tpb, (float)nu, (float)lam, (float)G, A, counter);
    else
      hipLaunchKernelGGL((batch_far_kernel<float, double, false>), grid, blk, 0, st, recv, r, meshes, item_start,
                         nmesh, tpb, (float)nu, (float)lam, (float)G, A, counter);
  } else {
    const double* r = static_cast<const double*>(rec);
    if (hs)
      hipLaunchKernelGGL((batch_far_kernel<double, double, true>), grid, blk, 0, st, recv, r, meshes, item_start,
                         nmesh, tpb, nu, lam, G, A, counter);
    else
      hipLaunchKernelGGL((batch_far_kernel<double, double, false>), grid, blk, 0, st, recv, r, meshes, item_start,
                         nmesh, tpb, nu, lam, G, A, counter);
  }
}

void launch_batch_near(const double* recv, const double* rec, const double* near, int hs,
                       const BatchMeshDev* meshes, const int64_t* item_start, int nmesh, int tpb, int wgs, double nu,
                       double lam, double G, double* A, unsigned long long* counter, unsigned long long* near_count,
                       hipStream_t st) {
  const dim3 grid(wgs), blk(256);
  if (hs)
    hipLaunchKernelGGL(batch_near_kernel<true>, grid, blk, 0, st, recv, rec, near, meshes, item_start, nmesh, tpb, nu,
                       lam, G, A, counter, near_count);
  else
    hipLaunchKernelGGL(batch_near_kernel<false>, grid, blk, 0, st, recv, rec, near, meshes, item_start, nmesh, tpb,
                       nu, lam, G, A, counter, near_count);
}

}  // namespace cdj
