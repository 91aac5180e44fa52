// Crack-tip edges and stress intensities on gfx950.
//
//   edge_match : every edge of the mesh (3n: P1P2, P2P3, P3P1 of each triangle, EdgeConstraints.jl's types) is
//                keyed by its two end points in lexicographic order (CreateSortedEdgePoints.jl) and inserted into
//                an open-addressing hash table with one atomic compare-and-swap per probe.  An edge that finds an
//                equal key (exact IEEE equality of the six coordinates, MatchingRow.jl) marks itself and the
//                stored edge as shared; every other edge is a free (crack-tip) edge.  The reference compares all
//                pairs of triangles (EdgeConstraints.jl, O(n^2) with an n x n distance matrix); the table makes it
//                O(n).  Its midpoint-distance skip never drops a shared edge (the incentres of two triangles with
//                a common edge are closer than the sum of their half perimeters), so the flags are the same.
//   tip_geometry : one thread per triangle; the three edge blocks of GetCrackTipElements3D.jl's GetValues (edge
//                length and midpoint, mid-to-edge and edge directions with their orientation convention, interior
//                angle, area).
//   tip_k      : one thread per triangle; K1K2K3TriDislocation of StressIntensity3D.jl and the active formula of
//                StrainEnergyRelease.jl.
#include <hip/hip_runtime.h>

#include <cmath>

#include "crack_tip.hpp"

namespace cdj {

namespace {

constexpr double kPi = 3.14159265358979323846;

struct V3 {
  double x, y, z;
};
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// LinearAlgebra.normalize (normr.jl): multiplication by the reciprocal of the 2-norm
__device__ __forceinline__ V3 normalize(V3 a) {
  const double inv = 1.0 / sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  return {a.x * inv, a.y * inv, a.z * inv};
}
__device__ __forceinline__ V3 load3(const double* P, int64_t n, int64_t i) { return {P[i], P[n + i], P[2 * n + i]}; }

// Corner arrays of edge type k: 0 = (P1, P2), 1 = (P2, P3), 2 = (P3, P1)
__device__ __forceinline__ void edge_ends(const double* P1, const double* P2, const double* P3, int64_t n, int64_t e,
                                          V3& a, V3& b) {
  const int64_t t = e % n;
  const int k = (int)(e / n);
  const double* A = k == 0 ? P1 : (k == 1 ? P2 : P3);
  const double* B = k == 0 ? P2 : (k == 1 ? P3 : P1);
  a = load3(A, n, t);
  b = load3(B, n, t);
}

// The two end points in lexicographic (x, y, z) order (CreateSortedEdgePoints.jl), as six doubles
__device__ __forceinline__ void edge_key(const double* P1, const double* P2, const double* P3, int64_t n, int64_t e,
                                         double k[6]) {
  V3 a, b;
  edge_ends(P1, P2, P3, n, e, a, b);
  bool swap = false;
  if (a.x != b.x)
    swap = !(a.x < b.x);
  else if (a.y != b.y)
    swap = !(a.y < b.y);
  else if (a.z != b.z)
    swap = !(a.z < b.z);
  const V3 lo = swap ? b : a, hi = swap ? a : b;
  k[0] = lo.x;
  k[1] = lo.y;
  k[2] = lo.z;
  k[3] = hi.x;
  k[4] = hi.y;
  k[5] = hi.z;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) fill_int_kernel(int* p, int64_t n, int v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) edge_match_kernel(const double* __restrict__ P1, const double* __restrict__ P2,
                                                         const double* __restrict__ P3, int64_t n,
                                                         int* __restrict__ slots, uint64_t mask,
                                                         int* __restrict__ matched) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 3 * n) return;
  double k[6];
  edge_key(P1, P2, P3, n, e, k);
  uint64_t h = 0x243f6a8885a308d3ull;
  for (int c = 0; c < 6; ++c) {
    if (k[c] != k[c]) return;  // NaN never equals anything: the edge stays free
    h = mix64(h ^ (uint64_t)__double_as_longlong(k[c] + 0.0));  // +0.0 folds -0.0 onto 0.0 (they compare equal)
  }
  uint64_t p = h & mask;
  for (uint64_t probe = 0; probe <= mask; ++probe) {
    const int cur = atomicCAS(&slots[p], -1, (int)e);
    if (cur == -1) return;  // first edge with this key
    double q[6];
    edge_key(P1, P2, P3, n, cur, q);
    bool eq = true;
    for (int c = 0; c < 6; ++c) eq = eq && (q[c] == k[c]);
    if (eq) {
      matched[e] = 1;
      matched[cur] = 1;
      return;
    }
    p = (p + 1) & mask;
  }
}

// RotateObject3DAllignVectors(A, (0, 0, 1), v): the Rodrigues rotation taking A onto e_z (reflection through
// the origin when A == -e_z)
__device__ V3 rotate_to_ez(V3 A, V3 v) {
  const V3 B = {0.0, 0.0, 1.0};
  if (B.x == -A.x && B.y == -A.y && B.z == -A.z) return neg(v);
  V3 w = cross(A, B);
  const double wn = sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
  w = {w.x / wn, w.y / wn, w.z / wn};
  if (w.x != w.x || w.y != w.y || w.z != w.z) w = {0.0, 0.0, 0.0};
  const double W[3][3] = {{0.0, -w.z, w.y}, {w.z, 0.0, -w.x}, {-w.y, w.x, 0.0}};
  double W2[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) W2[r][c] = W[r][0] * W[0][c] + W[r][1] * W[1][c] + W[r][2] * W[2][c];
  const double ct = dot(A, B) / sqrt(dot(A, A)) / sqrt(dot(B, B));
  const double t = acos(ct);
  const double s = sin(t), omc = 1.0 - cos(t);
  double R[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) R[r][c] = (r == c ? 1.0 : 0.0) + W[r][c] * s + W2[r][c] * omc;
  return {R[0][0] * v.x + R[0][1] * v.y + R[0][2] * v.z, R[1][0] * v.x + R[1][1] * v.y + R[1][2] * v.z,
          R[2][0] * v.x + R[2][1] * v.y + R[2][2] * v.z};
}

// CalculateSSandDSDirs.jl for one normal: strike- and dip-slip direction cosines
__device__ void ss_ds_dirs(V3 nv, V3& ss, V3& ds) {
  const double inv = 1.0 / sqrt(nv.x * nv.x + nv.y * nv.y + 0.0 * 0.0);
  ss = {-(nv.y * inv), nv.x * inv, 0.0};
  const double dsz = cos(acos(nv.z) - (kPi / 2));
  double L = sin(acos(dsz));
  if (nv.z < 0) L = -L;
  const double a = atan2(nv.y, nv.x);
  ds = {-(cos(a) * L), -(sin(a) * L), dsz};
  const V3 vs = cross(V3{0.0, 0.0, 1.0}, nv);
  if (sqrt(dot(vs, vs)) == 0.0) {
    ds = {-1.0, 0.0, 0.0};
    ss = {0.0, nv.z < 1.0 ? -1.0 : 1.0, 0.0};
  }
}

__device__ __forceinline__ double heron(V3 a, V3 b, V3 c) {
  const V3 ab = sub(a, b), bc = sub(b, c), ca = sub(c, a);
  const double la = sqrt(ab.x * ab.x + ab.y * ab.y + ab.z * ab.z);
  const double lb = sqrt(bc.x * bc.x + bc.y * bc.y + bc.z * bc.z);
  const double lc = sqrt(ca.x * ca.x + ca.y * ca.y + ca.z * ca.z);
  const double s = (la + lb + lc) / 2.0;
  return sqrt(s * (s - la) * (s - lb) * (s - lc));
}

// Edge block s: (Pa, Pb, Pc) and the hash edge type of Pa-Pb
__device__ __forceinline__ void block_corners(int s, const double* P1, const double* P2, const double* P3,
                                              const double*& Pa, const double*& Pb, const double*& Pc, int& htype) {
  if (s == 0) {
    Pa = P1; Pb = P2; Pc = P3; htype = 0;
  } else if (s == 1) {
    Pa = P1; Pb = P3; Pc = P2; htype = 2;
  } else {
    Pa = P2; Pb = P3; Pc = P1; htype = 1;
  }
}

__global__ void __launch_bounds__(256) tip_geometry_kernel(const double* __restrict__ mid,
                                                           const double* __restrict__ P1,
                                                           const double* __restrict__ P2,
                                                           const double* __restrict__ P3,
                                                           const double* __restrict__ nrm, int64_t n,
                                                           const int* __restrict__ matched,
                                                           uint8_t* __restrict__ flags, double* __restrict__ geom) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double nan = __builtin_nan("");
  const V3 m = load3(mid, n, t), nv = load3(nrm, n, t);
  for (int s = 0; s < 3; ++s) {
    const double *Pa, *Pb, *Pc;
    int htype;
    block_corners(s, P1, P2, P3, Pa, Pb, Pc, htype);
    const V3 a = load3(Pa, n, t), b = load3(Pb, n, t), c = load3(Pc, n, t);
    const bool fr = matched[htype * n + t] == 0;
    double g[CT_GEOM];
    for (int q = 0; q < CT_GEOM; ++q) g[q] = nan;
    if (fr) {
      const V3 ab = sub(a, b);
      g[0] = sqrt(ab.x * ab.x + ab.y * ab.y + ab.z * ab.z);
      const V3 md = {(a.x + b.x) / 2, (a.y + b.y) / 2, (a.z + b.z) / 2};
      g[1] = md.x;
      g[2] = md.y;
      g[3] = md.z;
      const V3 dm = sub(md, m);
      g[10] = sqrt(dm.x * dm.x + dm.y * dm.y + dm.z * dm.z);
      const V3 m2e = normalize(dm);
      const V3 v = normalize(sub(a, c)), w = normalize(sub(b, c));
      g[11] = acos(dot(v, w)) * (180.0 / kPi);
      V3 ev = normalize(ab);
      // edge vector counter-clockwise from the mid-to-edge vector, looking along the normal
      const V3 m2r = rotate_to_ez(nv, m2e), evr = rotate_to_ez(nv, ev);
      if (m2r.x * evr.y + m2r.y * (-evr.x) < 0) ev = neg(ev);
      V3 m2 = cross(nv, ev);
      if (dot(m2e, m2) < 0) m2 = neg(m2);
      ev = cross(m2, nv);
      g[4] = ev.x;
      g[5] = ev.y;
      g[6] = ev.z;
      g[7] = m2.x;
      g[8] = m2.y;
      g[9] = m2.z;
    }
    g[12] = heron(a, b, c);
    double* G = geom + (int64_t)s * n * CT_GEOM;
    for (int q = 0; q < CT_GEOM; ++q) G[q * n + t] = g[q];
    flags[s * n + t] = fr ? 1 : 0;
  }
}

__global__ void __launch_bounds__(256) tip_k_kernel(const double* __restrict__ Dn, const double* __restrict__ Dss,
                                                    const double* __restrict__ Dds, int64_t n, double Gm, double nu,
                                                    const double* __restrict__ nrm, const uint8_t* __restrict__ flags,
                                                    const double* __restrict__ geom, double* __restrict__ K) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double nan = __builtin_nan("");
  const V3 nv = load3(nrm, n, t);
  for (int s = 0; s < 3; ++s) {
    double k1 = nan, k2 = nan, k3 = nan, se = nan;
    if (flags[s * n + t]) {
      const double* G = geom + (int64_t)s * n * CT_GEOM;
      const V3 ev = {G[4 * n + t], G[5 * n + t], G[6 * n + t]};
      const V3 m2 = {G[7 * n + t], G[8 * n + t], G[9 * n + t]};
      const double h = G[10 * n + t];
      V3 ss, ds;
      ss_ds_dirs(nv, ss, ds);
      const double dss = Dss[t], dds = Dds[t];
      const V3 dpl = {ss.x * dss + ds.x * dds, ss.y * dss + ds.y * dds, ss.z * dss + ds.z * dds};
      double sx = rotate_to_ez(nv, ss).x, mx = rotate_to_ez(nv, m2).x;
      if (fabs(mx) > 1) mx = 1.0;  // StressIntensity3D.jl's clamp (to +1 whatever the sign)
      if (fabs(sx) > 1) sx = 1.0;
      const double angle = fabs(acos(sx) - acos(mx));
      const double dir1 = (kPi / 2) + angle;
      const double cx = sin(dir1), cy = cos(dir1);
      double along = fabs(dds * cx + dss * cy);
      double mid2 = fabs(-dds * cy + dss * cx);
      if (dot(m2, dpl) <= 0) mid2 = -mid2;
      if (dot(ev, dpl) <= 0) along = -along;
      const double c = (Gm * sqrt(kPi)) / (2 * sqrt(h) * (1 - nu));
      k1 = (c * Dn[t]) / 1.834;
      k2 = (c * mid2) / 1.834;
      k3 = ((c * along) * (1 - nu)) / 1.834;
      se = sqrt(fabs(k1) * fabs(k1) + fabs(k2) * fabs(k2) + (1 / (1 - nu)) * (fabs(k3) * fabs(k3)));
    }
    double* Ks = K + (int64_t)s * n * CT_K;
    Ks[t] = k1;
    Ks[n + t] = k2;
    Ks[2 * n + t] = k3;
    Ks[3 * n + t] = se;
  }
}

unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

uint64_t crack_tip_slots(int64_t n) {
  uint64_t cap = 64;
  while (cap < 2 * 3 * (uint64_t)(n > 0 ? n : 1)) cap <<= 1;
  return cap;
}

hipError_t crack_tip_elements(const double* mid, const double* P1, const double* P2, const double* P3,
                              const double* nrm, int64_t n, int* slots, int* matched, uint8_t* flags, double* geom,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const uint64_t cap = crack_tip_slots(n);
  hipLaunchKernelGGL(fill_int_kernel, dim3(nblk((int64_t)cap, 256)), dim3(256), 0, st, slots, (int64_t)cap, -1);
  hipLaunchKernelGGL(fill_int_kernel, dim3(nblk(3 * n, 256)), dim3(256), 0, st, matched, 3 * n, 0);
  hipLaunchKernelGGL(edge_match_kernel, dim3(nblk(3 * n, 256)), dim3(256), 0, st, P1, P2, P3, n, slots, cap - 1,
                     matched);
  hipLaunchKernelGGL(tip_geometry_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, mid, P1, P2, P3, nrm, n,
                     (const int*)matched, flags, geom);
  return hipGetLastError();
}

hipError_t stress_intensity(const double* Dn, const double* Dss, const double* Dds, int64_t n, double G, double nu,
                            const double* nrm, const uint8_t* flags, const double* geom, double* K, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(tip_k_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, Dn, Dss, Dds, n, G, nu, nrm, flags, geom, K);
  return hipGetLastError();
}

}  // namespace cdj
