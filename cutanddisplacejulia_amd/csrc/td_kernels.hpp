// Host-side launch interface of the TD kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cdj {

static constexpr int RECV = 12;  // receiver record: point, normal, strike dir, dip dir
static constexpr int REC_FS = 32;   // triangle record, full space (elements)
static constexpr int REC_HS = 160;  // triangle record, half space: real + image TD + 3 free-surface sides
static constexpr int REC_EDGE = 32;  // mesh-edge record of the half-space assembly (one free-surface side)
static constexpr int NEAR_REC = 8;   // near-field record of a source: centroid, (f L)^2, unit normal, L

// Output descriptor passed by value to the kernels.
//   influence: p[0..17] strain (StrainInf order), p[18..26] displacement (DispInf order); ld = nobs
//   traction:  p[3*m + q] block of mode m (Dn, Dss, Dds) x traction q (Tn, Tss, Tds), leading dimension ld,
//              every value multiplied by `sign`
template <class T> struct OutPtrs {
  T* p[27];
  int64_t ld;
  int strain_on;
  int disp_on;
  T sign;
};

void launch_setup(const double* P1, const double* P2, const double* P3, int64_t ntri, int hs, double* rec,
                  hipStream_t st);
void launch_rec_to_f32(const double* a, float* b, int64_t n, hipStream_t st);
void launch_receivers(const double* nrm, const double* mid, int64_t n, double* recv, hipStream_t st);
// receivers at given points (n x 3) with the planes of the given normals
void launch_receivers_at(const double* nrm, const double* pts, int64_t n, double* recv, hipStream_t st);
// *flag |= 1 when a point (Z[i * zstride], i < nobs) or a triangle corner has z > 0 (TD.jl:606-619)
void launch_check_hs(const double* Z, int64_t nobs, const double* P1, const double* P2, const double* P3,
                     int64_t ntri, int* flag, hipStream_t st, int64_t zstride = 1);
// p[0..n) = 0 (a kernel, not a memset node, so that the clear can sit in a captured graph)
void launch_zero_u32(unsigned* p, int64_t n, hipStream_t st);
template <class T>
void launch_influence(const T* X, const T* Y, const T* Z, int64_t nobs, const T* rec, int64_t ntri, int hs, int tpb,
                      const double* Dn, const double* Dss, const double* Dds, T nu, const OutPtrs<T>& out,
                      hipStream_t st);
template <class T>
void launch_field(const T* X, const T* Y, const T* Z, int64_t nobs, const T* rec, int64_t ntri, int hs, int tpb,
                  const double* Dn, const double* Dss, const double* Dds, T nu, int strain_on, int disp_on, T* part,
                  double* out, hipStream_t st);
template <class T>
void launch_traction(const double* recv, int64_t nrecv, const T* rec, int64_t ntri, int hs, int tpb, T nu, T lam, T G,
                     const OutPtrs<T>& out, hipStream_t st);
// FP32 arithmetic, FP64 storage (far field of the mixed assembly)
void launch_traction_f32_to_f64(const double* recv, int64_t nrecv, const float* rec, int64_t ntri, int hs, int tpb,
                                float nu, float lam, float G, const OutPtrs<double>& out, hipStream_t st);
// ---- half-space assembly with free-surface responses shared per mesh edge ----
// Device arrays of the edge topology (sides e = k ntri + j; see tdk_edge_insert / tdk_edge_number).
struct EdgeTopo {
  int* slots;      // edge_table_slots(ntri)
  int* own;        // edge_table_slots(ntri)
  int* slot_of;    // 3 ntri
  int* owner_idx;  // 3 ntri
  int* side_edge;  // 3 ntri: 2 * edge + (1 if the side runs against the edge's owner)
  int* edge_side;  // 3 ntri: owner side of each edge
  int* count;      // 2: number of edges, 1 if a triangle lies in z = 0
};
uint64_t edge_table_slots(int64_t ntri);
void launch_edge_topology(const double* P1, const double* P2, const double* P3, int64_t ntri, const EdgeTopo& e,
                          hipStream_t st);
// edge records (R_SIDE_SIZE doubles each, at most 3 ntri) from the triangle records `rec` (half space)
void launch_edge_setup(const double* P1, const double* P2, const double* P3, int64_t ntri, const double* rec,
                       const EdgeTopo& e, double* erec, hipStream_t st);
// Tb[(t * 9 + 3 kE + q) * ldT + i]: traction q on receiver i's plane of edge t's free-surface strain for the
// EFCS unit Burgers vector kE
void launch_edge_response(const double* recv, int64_t nrecv, const double* erec, int64_t nedge, double nu,
                          double lam, double G, double* Tb, int64_t ldT, hipStream_t st);
// half-space traction entries from the real and image TDs plus the edges' responses (no triangle in z = 0)
void launch_traction_combine(const double* recv, int64_t nrecv, const double* rec, int64_t ntri, int tpb, double nu,
                             double lam, double G, const int* side_edge, const double* Tb, int64_t ldT,
                             const OutPtrs<double>& out, hipStream_t st);

// Fixed-element rows: out block k (rows 0..n-1, columns 0..nsrc-1, times out.sign) from the compact traction
// rows Tf (9 blocks nf x nsrc) and displacement rows Ud (9 blocks nfix x nsrc); rowmap[r] >= 0 selects free row
// rowmap[r], otherwise fixed row -1 - rowmap[r].
void launch_fixed_scatter(const double* Tf, const double* Ud, const int32_t* rowmap, int64_t n, int64_t nf,
                          int64_t nfix, int64_t nsrc, const OutPtrs<double>& out, hipStream_t st);
// near[j * NEAR_REC] = {centroid, (factor * L)^2, unit normal, L}, L the longest edge (see near_pair)
void launch_near_info(const double* P1, const double* P2, const double* P3, int64_t ntri, double factor, double* near,
                      hipStream_t st);
// FP64 recompute of near pairs over the FP32 values; *count accumulates the number of entries replaced
void launch_nearfield(const double* recv, int64_t nrecv, const double* rec, const double* near, int64_t ntri, int hs,
                      int tpb, double nu, double lam, double G, const OutPtrs<double>& out, unsigned long long* count,
                      hipStream_t st);

}  // namespace cdj
