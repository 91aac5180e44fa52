// Batched assembly of many meshes' collocation matrices (persistent work-queue kernels).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cdj {

// One mesh of a batch: its first triangle in the concatenated triangle arrays, its triangle count and the
// offset (in doubles) of its matrix A (3n x 3n, leading dimension 3n) in the batch's matrix buffer.
struct BatchMeshDev {
  int64_t tri0, n, a_off;
};

void launch_batch_clear(unsigned long long* p, int n, hipStream_t st);

// Far pass over every work item: FP32 (rec = float records) or FP64 (rec = double records) arithmetic, FP64
// storage.  wgs persistent workgroups of 256 threads; *counter must be zero.
void launch_batch_far(const double* recv, const void* rec, bool fp32, int hs, const BatchMeshDev* meshes,
                      const int64_t* item_start, int nmesh, int tpb, int wgs, double nu, double lam, double G,
                      double* A, unsigned long long* counter, hipStream_t st);

// FP64 near-field pass (mixed mode); *near_count accumulates the entries replaced.
void launch_batch_near(const double* recv, const double* rec, const double* near, int hs,
                       const BatchMeshDev* meshes, const int64_t* item_start, int nmesh, int tpb, int wgs, double nu,
                       double lam, double G, double* A, unsigned long long* counter, unsigned long long* near_count,
                       hipStream_t st);

}  // namespace cdj
