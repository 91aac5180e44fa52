// TD kernels for gfx950: triangle/receiver set-up, influence matrices, summed fields and the fused
// strain -> traction assembly of the collocation matrix.
#include <hip/hip_runtime.h>

#include "td_device.hpp"
#include "td_kernels.hpp"

namespace cdj {

// ---- per-triangle records (CalculateLocalTriCoords! TD.jl:712-732, GlobalToTDECoords TD.jl:870-908,
//      CalcTDVectsAndAngles TD.jl:962-1008, CalcSideVec TD.jl:935-960, AngSetup*FSC set-up) -------------
__device__ static void build_td(const double* P1, const double* P2, const double* P3, double* r) {
#pragma clang fp contract(off)
  double a[3], b[3], Vn[3], Vs[3], Vd[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = P2[k] - P1[k];
    b[k] = P3[k] - P1[k];
  }
  Vn[0] = a[1] * b[2] - a[2] * b[1];
  Vn[1] = a[2] * b[0] - a[0] * b[2];
  Vn[2] = a[0] * b[1] - a[1] * b[0];
  double nrm = sqrt(Vn[0] * Vn[0] + Vn[1] * Vn[1] + Vn[2] * Vn[2]);
  for (int k = 0; k < 3; ++k) Vn[k] = Vn[k] / nrm;
  // Vstrike = eZ x Vnorm, or eY * Vnorm_z for a horizontal triangle
  Vs[0] = 0.0 * Vn[2] - 1.0 * Vn[1];
  Vs[1] = 1.0 * Vn[0] - 0.0 * Vn[2];
  Vs[2] = 0.0 * Vn[1] - 0.0 * Vn[0];
  if (Vs[0] == 0.0 && Vs[1] == 0.0 && Vs[2] == 0.0) {
    Vs[0] = 0.0 * Vn[2];
    Vs[1] = 1.0 * Vn[2];
    Vs[2] = 0.0 * Vn[2];
  }
  nrm = sqrt(Vs[0] * Vs[0] + Vs[1] * Vs[1] + Vs[2] * Vs[2]);
  for (int k = 0; k < 3; ++k) Vs[k] = Vs[k] / nrm;
  Vd[0] = Vn[1] * Vs[2] - Vn[2] * Vs[1];
  Vd[1] = Vn[2] * Vs[0] - Vn[0] * Vs[2];
  Vd[2] = Vn[0] * Vs[1] - Vn[1] * Vs[0];
  double p1[3], p3[3];
  {
    const double d1[3] = {P1[0] - P2[0], P1[1] - P2[1], P1[2] - P2[2]};
    const double d3[3] = {P3[0] - P2[0], P3[1] - P2[1], P3[2] - P2[2]};
    p1[0] = (Vn[0] * d1[0]) + (Vn[1] * d1[1]) + (Vn[2] * d1[2]);
    p1[1] = (Vs[0] * d1[0]) + (Vs[1] * d1[1]) + (Vs[2] * d1[2]);
    p1[2] = (Vd[0] * d1[0]) + (Vd[1] * d1[1]) + (Vd[2] * d1[2]);
    p3[0] = (Vn[0] * d3[0]) + (Vn[1] * d3[1]) + (Vn[2] * d3[2]);
    p3[1] = (Vs[0] * d3[0]) + (Vs[1] * d3[1]) + (Vs[2] * d3[2]);
    p3[2] = (Vd[0] * d3[0]) + (Vd[1] * d3[1]) + (Vd[2] * d3[2]);
  }
  double e12[3], e13[3], e23[3];
  for (int k = 0; k < 3; ++k) {
    e12[k] = 0.0 - p1[k];
    e13[k] = p3[k] - p1[k];
    e23[k] = p3[k] - 0.0;
  }
  nrm = sqrt(e12[0] * e12[0] + e12[1] * e12[1] + e12[2] * e12[2]);
  for (int k = 0; k < 3; ++k) e12[k] = e12[k] / nrm;
  nrm = sqrt(e13[0] * e13[0] + e13[1] * e13[1] + e13[2] * e13[2]);
  for (int k = 0; k < 3; ++k) e13[k] = e13[k] / nrm;
  nrm = sqrt(e23[0] * e23[0] + e23[1] * e23[1] + e23[2] * e23[2]);
  for (int k = 0; k < 3; ++k) e23[k] = e23[k] / nrm;
  const double A = acos((e12[0] * e13[0]) + (e12[1] * e13[1]) + (e12[2] * e13[2]));
  const double B = acos((-e12[0] * e23[0]) + (-e12[1] * e23[1]) + (-e12[2] * e23[2]));
  const double C = acos((e23[0] * e13[0]) + (e23[1] * e13[1]) + (e23[2] * e13[2]));
  for (int k = 0; k < 3; ++k) {
    r[R_VN + k] = Vn[k];
    r[R_VS + k] = Vs[k];
    r[R_VD + k] = Vd[k];
    r[R_P2 + k] = P2[k];
  }
  r[R_P1Y] = p1[1];
  r[R_P1Z] = p1[2];
  r[R_P3Y] = p3[1];
  r[R_P3Z] = p3[2];
  r[R_E13Y] = e13[1];
  r[R_E13Z] = e13[2];
  r[R_E12Y] = e12[1];
  r[R_E12Z] = e12[2];
  r[R_E23Y] = e23[1];
  r[R_E23Z] = e23[2];
  r[R_CA] = cos(-kPi + A);
  r[R_SA] = sin(-kPi + A);
  r[R_CB] = cos(-kPi + B);
  r[R_SB] = sin(-kPi + B);
  r[R_CC] = cos(-kPi + C);
  r[R_SC] = sin(-kPi + C);
  r[R_FLAT] = (P1[2] == 0.0 && P2[2] == 0.0 && P3[2] == 0.0) ? 1.0 : 0.0;
  // trimodefinder denominator with p2 = (0, 0): (p2z - p3z)(p1y - p3y) + (p3y - p2y)(p1z - p3z)
  r[R_BASE] = ((0.0 - p3[2]) * (p1[1] - p3[1]) + (p3[1] - 0.0) * (p1[2] - p3[2]));
  r[30] = 0.0;
  r[31] = 0.0;
}

__device__ static double dcot(double x) { return 1.0 / tan(x); }

__device__ static void build_side(const double* PA, const double* PB, const double* rtd, double* s) {
#pragma clang fp contract(off)
  const double sv[3] = {PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2]};
  const double Sv = sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
  const double beta = acos((((-sv[0]) * 0.0) - sv[1] * 0.0 - sv[2] * 1.0) / Sv);
  for (int k = 0; k < 32; ++k) s[k] = 0.0;
  const bool skip = fabs(beta) < kEps || fabs(kPi - beta) < kEps || fabs(dcot(beta)) > 5e3;
  s[S_SKIP] = skip ? 1.0 : 0.0;
  if (skip) return;
  double p = sv[0], q = sv[1];
  const double n1 = sqrt(p * p + q * q + 0.0 * 0.0);
  p = p / n1;
  q = q / n1;
  for (int k = 0; k < 3; ++k) s[S_PA + k] = PA[k];
  s[S_EY1 + 0] = p;
  s[S_EY1 + 1] = q;
  // side vector in ADCS: rot3 with (ey1, ey2, ey3), ey2 = (q, -p, 0), ey3 = (-0, -0, -1)
  s[S_YB + 0] = (p * sv[0]) + (q * sv[1]) + (-0.0 * sv[2]);
  s[S_YB + 1] = (q * sv[0]) + ((-p) * sv[1]) + (-0.0 * sv[2]);
  s[S_YB + 2] = (0.0 * sv[0]) + (0.0 * sv[1]) + (-1.0 * sv[2]);
  s[S_AA] = -PA[2];
  s[S_AB] = -PB[2];
  // unit slip of mode m in EFCS is Vn / Vs / Vd; then EFCS -> ADCS
  const double* V[3] = {rtd + R_VN, rtd + R_VS, rtd + R_VD};
  for (int m = 0; m < 3; ++m) {
    const double e0 = V[m][0], e1 = V[m][1], e2 = V[m][2];
    s[S_B + 3 * m + 0] = (p * e0) + (q * e1) + (-0.0 * e2);
    s[S_B + 3 * m + 1] = (q * e0) + ((-p) * e1) + (-0.0 * e2);
    s[S_B + 3 * m + 2] = (0.0 * e0) + (0.0 * e1) + (-1.0 * e2);
  }
  s[S_COSB] = cos(beta);
  s[S_SINB] = sin(beta);
  s[S_COTB] = dcot(beta);
  s[S_COTB2] = dcot(beta / 2);
  const double ps = kPi - beta;
  s[S_COSPS] = cos(ps);
  s[S_SINPS] = sin(ps);
  s[S_COTPS] = dcot(ps);
  s[S_COTPS2] = dcot(ps / 2);
  const double pd = -kPi + beta;
  s[S_COSPD] = cos(pd);
  s[S_SINPD] = sin(pd);
  s[S_COTPD] = dcot(pd);
  s[S_COTPD2] = dcot(pd / 2);
}

__global__ void tdk_setup(const double* __restrict__ P1, const double* __restrict__ P2,
                          const double* __restrict__ P3, int64_t ntri, int hs, double* __restrict__ rec) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= ntri) return;
  const double a[3] = {P1[j], P1[ntri + j], P1[2 * ntri + j]};
  const double b[3] = {P2[j], P2[ntri + j], P2[2 * ntri + j]};
  const double c[3] = {P3[j], P3[ntri + j], P3[2 * ntri + j]};
  const int RS = hs ? REC_HS : REC_FS;
  double* r = rec + j * RS;
  build_td(a, b, c, r);
  if (hs) {
    const double ai[3] = {a[0], a[1], -a[2]};
    const double bi[3] = {b[0], b[1], -b[2]};
    const double ci[3] = {c[0], c[1], -c[2]};
    build_td(ai, bi, ci, r + OFF_IMG);
    build_side(a, b, r, r + OFF_SIDE);
    build_side(b, c, r, r + OFF_SIDE + R_SIDE_SIZE);
    build_side(c, a, r, r + OFF_SIDE + 2 * R_SIDE_SIZE);
  }
}

__global__ void tdk_rec_to_f32(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ---- receivers: collocation points and plane directions --------------------------------------------------
// SetupCollationPoints (CalculateInfluenceMatrices3D.jl:52-71) and CalculateSSandDSDirs.jl:37-98.
// recv[i*RECV + ...] = {x, y, z, nx, ny, nz, sx, sy, sz, dx, dy, dz}
// at != nullptr: the receiver points are given (n x 3) instead of derived from the midpoints.
__global__ void tdk_receivers(const double* __restrict__ nrm, const double* __restrict__ mid,
                              const double* __restrict__ at, int64_t n, double* __restrict__ recv) {
#pragma clang fp contract(off)
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double nx = nrm[i], ny = nrm[n + i], nz = nrm[2 * n + i];
  double* o = recv + i * RECV;
  if (at) {
    o[0] = at[i];
    o[1] = at[n + i];
    o[2] = at[2 * n + i];
  } else {
    o[0] = mid[i] - (nx * 1e-12);
    o[1] = mid[n + i] - (ny * 1e-12);
    o[2] = mid[2 * n + i] - (nz * 1e-12);
  }
  o[3] = nx;
  o[4] = ny;
  o[5] = nz;
  const double h = sqrt(nx * nx + ny * ny + 0.0 * 0.0);
  double sx = -(ny / h), sy = nx / h, sz = 0.0;
  const double DSCosAz = cos(acos(nz) - (kPi / 2));
  double L = sin(acos(DSCosAz));
  if (nz < 0) L = -L;
  const double a = atan2(ny, nx);
  double dx = -(cos(a) * L), dy = -(sin(a) * L), dz = DSCosAz;
  const double vs0 = 0.0 * nz - 1.0 * ny, vs1 = 1.0 * nx - 0.0 * nz, vs2 = 0.0 * ny - 0.0 * nx;
  if (vs0 == 0.0 && vs1 == 0.0 && vs2 == 0.0) {
    dx = -1.0;
    dy = 0.0;
    dz = 0.0;
    sx = 0.0;
    sy = nz < 1.0 ? -1.0 : 1.0;
    sz = 0.0;
  }
  o[6] = sx;
  o[7] = sy;
  o[8] = sz;
  o[9] = dx;
  o[10] = dy;
  o[11] = dz;
}

// ---- HS validity check (ViewInLoop TD.jl:606-619) --------------------------------------------------------
// Z[i * zstride] for i < nobs: the observation points' z coordinates (a plain array, or the receiver records)
__global__ void tdk_check_hs(const double* __restrict__ Z, int64_t zstride, int64_t nobs,
                             const double* __restrict__ P1, const double* __restrict__ P2,
                             const double* __restrict__ P3, int64_t ntri, int* __restrict__ flag) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < nobs) bad = Z[i * zstride] > 0;
  if (i < ntri) bad = bad || P1[2 * ntri + i] > 0 || P2[2 * ntri + i] > 0 || P3[2 * ntri + i] > 0;
  if (bad) atomicOr(flag, 1);
}

// ---- influence matrices ----------------------------------------------------------------------------------
// grid.x: 256-point tiles of observation points, grid.y: chunks of `tpb` triangles.  Each lane loops over the
// chunk; the triangle index is wave-uniform so its record comes through scalar loads.
template <class T, bool HS>
__global__ void __launch_bounds__(256) tdk_influence(const T* __restrict__ X, const T* __restrict__ Y,
                                                     const T* __restrict__ Z, int64_t nobs,
                                                     const T* __restrict__ rec, int64_t ntri, int tpb,
                                                     const double* __restrict__ Dn, const double* __restrict__ Dss,
                                                     const double* __restrict__ Dds, T nu, OutPtrs<T> out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nobs) return;
  const T Xi = X[i], Yi = Y[i], Zi = Z[i];
  const int64_t j0 = (int64_t)blockIdx.y * tpb;
  const int64_t j1 = j0 + tpb < ntri ? j0 + tpb : ntri;
  constexpr int RS = HS ? REC_HS : REC_FS;
  for (int64_t j = j0; j < j1; ++j) {
    const T* r = rec + j * RS;
    const T sl[3] = {(T)Dn[j], (T)Dss[j], (T)Dds[j]};
    const int64_t o = i + j * out.ld;
    if (out.strain_on) {
      T e[3][6];
      strain_pair<T, HS>(r, Xi, Yi, Zi, nu, e);
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int c = 0; c < 6; ++c) out.p[6 * m + c][o] = sl[m] == T(0) ? T(0) : sl[m] * e[m][c];
    }
    if (out.disp_on) {
      T u[3][3];
      disp_pair<T, HS>(r, Xi, Yi, Zi, nu, u);
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int c = 0; c < 3; ++c) out.p[18 + 3 * m + c][o] = sl[m] == T(0) ? T(0) : sl[m] * u[m][c];
    }
  }
}

// ---- summed field: per-chunk partial sums, then an in-order reduction ------------------------------------
// part[(chunk * ncomp + comp) * nobs + i]; comps: 6 strain (if on) then 3 displacement (if on).
template <class T, bool HS>
__global__ void __launch_bounds__(256) tdk_field_part(const T* __restrict__ X, const T* __restrict__ Y,
                                                      const T* __restrict__ Z, int64_t nobs,
                                                      const T* __restrict__ rec, int64_t ntri, int tpb,
                                                      const double* __restrict__ Dn, const double* __restrict__ Dss,
                                                      const double* __restrict__ Dds, T nu, int strain_on,
                                                      int disp_on, T* __restrict__ part) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nobs) return;
  const T Xi = X[i], Yi = Y[i], Zi = Z[i];
  const int64_t j0 = (int64_t)blockIdx.y * tpb;
  const int64_t j1 = j0 + tpb < ntri ? j0 + tpb : ntri;
  constexpr int RS = HS ? REC_HS : REC_FS;
  T se[6] = {0, 0, 0, 0, 0, 0}, su[3] = {0, 0, 0};
  for (int64_t j = j0; j < j1; ++j) {
    const T* r = rec + j * RS;
    const T sl[3] = {(T)Dn[j], (T)Dss[j], (T)Dds[j]};
    if (strain_on) {
      T e[3][6];
      strain_pair<T, HS>(r, Xi, Yi, Zi, nu, e);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const T a = sl[0] == T(0) ? T(0) : sl[0] * e[0][c];
        const T b = sl[1] == T(0) ? T(0) : sl[1] * e[1][c];
        const T d = sl[2] == T(0) ? T(0) : sl[2] * e[2][c];
        se[c] += (a + b) + d;
      }
    }
    if (disp_on) {
      T u[3][3];
      disp_pair<T, HS>(r, Xi, Yi, Zi, nu, u);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T a = sl[0] == T(0) ? T(0) : sl[0] * u[0][c];
        const T b = sl[1] == T(0) ? T(0) : sl[1] * u[1][c];
        const T d = sl[2] == T(0) ? T(0) : sl[2] * u[2][c];
        su[c] += (a + b) + d;
      }
    }
  }
  const int ncomp = (strain_on ? 6 : 0) + (disp_on ? 3 : 0);
  T* p = part + (int64_t)blockIdx.y * ncomp * nobs + i;
  int q = 0;
  if (strain_on)
    for (int c = 0; c < 6; ++c) p[(q++) * nobs] = se[c];
  if (disp_on)
    for (int c = 0; c < 3; ++c) p[(q++) * nobs] = su[c];
}

template <class T>
__global__ void tdk_field_reduce(const T* __restrict__ part, int64_t nobs, int ncomp, int nchunk,
                                 double* __restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nobs * ncomp) return;
  double s = 0.0;
  for (int k = 0; k < nchunk; ++k) s += (double)part[(int64_t)k * ncomp * nobs + t];
  out[t] = s;
}

// ---- fused strain -> stress -> traction assembly (ConvertStrainInfMatsToTraction,
//      CalculateInfluenceMatrices3D.jl:73-142) -------------------------------------------------------------
// out.p[3*m + q] is the block of mode m (Dn, Dss, Dds) and traction q (Tn, Tss, Tds); out.sign = -1 writes
// A = -[...] (SlipCalculator3D.jl:13-16).  Block pointers are already offset, leading dimension out.ld.
// FP32 instances are held to two waves per SIMD (the half-space one then spills 12 bytes per lane instead of
// running one wave per SIMD); FP64 keeps the register budget of one wave.
template <class T, class TO, bool HS>
__global__ void __launch_bounds__(256, sizeof(T) == 4 ? 2 : 1) tdk_traction(const double* __restrict__ recv, int64_t nrecv,
                                                    const T* __restrict__ rec, int64_t ntri, int tpb, T nu, T lam,
                                                    T G, OutPtrs<TO> out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nrecv) return;
  const double* rv = recv + i * RECV;
  const T Xi = (T)rv[0], Yi = (T)rv[1], Zi = (T)rv[2];
  const T n[3] = {(T)rv[3], (T)rv[4], (T)rv[5]};
  const T s[3] = {(T)rv[6], (T)rv[7], (T)rv[8]};
  const T d[3] = {(T)rv[9], (T)rv[10], (T)rv[11]};
  const int64_t j0 = (int64_t)blockIdx.y * tpb;
  const int64_t j1 = j0 + tpb < ntri ? j0 + tpb : ntri;
  constexpr int RS = HS ? REC_HS : REC_FS;
  const T sg = (T)out.sign;
  for (int64_t j = j0; j < j1; ++j) {
    const int64_t o = i + j * out.ld;
    T t[3][3];
    pair_traction<T, HS>(rec + j * RS, Xi, Yi, Zi, nu, lam, G, n, s, d, t);
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int q = 0; q < 3; ++q) out.p[3 * m + q][o] = (TO)(sg * t[m][q]);
  }
}

// Near-field pass of the mixed assembly: pairs that near_pair selects (near[j * NEAR_REC]: the source's centroid,
// near radius, plane and size) are recomputed in FP64 and overwrite the FP32 values.  A wave
// with no near lane for a source skips it, so the FP64 work is proportional to the near pairs' waves.
template <bool HS>
__global__ void __launch_bounds__(256) tdk_nearfield(const double* __restrict__ recv, int64_t nrecv,
                                                     const double* __restrict__ rec, const double* __restrict__ near,
                                                     int64_t ntri, int tpb, double nu, double lam, double G,
                                                     OutPtrs<double> out, unsigned long long* __restrict__ count) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool valid = i < nrecv;
  const double* rv = recv + (valid ? i : 0) * RECV;
  const double Xi = rv[0], Yi = rv[1], Zi = rv[2];
  const double n[3] = {rv[3], rv[4], rv[5]};
  const double s[3] = {rv[6], rv[7], rv[8]};
  const double d[3] = {rv[9], rv[10], rv[11]};
  const int64_t j0 = (int64_t)blockIdx.y * tpb;
  const int64_t j1 = j0 + tpb < ntri ? j0 + tpb : ntri;
  constexpr int RS = HS ? REC_HS : REC_FS;
  unsigned long long cnt = 0;
  for (int64_t j = j0; j < j1; ++j) {
    const bool nr = valid && near_pair(near + j * NEAR_REC, Xi, Yi, Zi);
    if (!__any(nr)) continue;
    // the same evaluation as the FP64 assembly (tdk_traction), so near entries equal its entries
    double t[3][3];
    pair_traction<double, HS>(rec + j * RS, Xi, Yi, Zi, nu, lam, G, n, s, d, t);
    if (nr) {
      const int64_t o = i + j * out.ld;
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int q = 0; q < 3; ++q) out.p[3 * m + q][o] = out.sign * t[m][q];
      ++cnt;
    }
  }
  if (cnt) atomicAdd(count, cnt * 9ull);
}

// near[j * NEAR_REC] = {centroid x, y, z, (near_factor * L)^2, unit normal x, y, z, L}, L the longest edge
__global__ void tdk_near_info(const double* __restrict__ P1, const double* __restrict__ P2,
                              const double* __restrict__ P3, int64_t ntri, double factor, double* __restrict__ near) {
  const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= ntri) return;
  double a[3], b[3], c[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = P1[k * ntri + j];
    b[k] = P2[k * ntri + j];
    c[k] = P3[k * ntri + j];
  }
  double l = 0.0;
  const double* pa[3] = {a, b, c};
  const double* pb[3] = {b, c, a};
  for (int e = 0; e < 3; ++e) {
    double s2 = 0.0;
    for (int k = 0; k < 3; ++k) s2 += (pb[e][k] - pa[e][k]) * (pb[e][k] - pa[e][k]);
    l = s2 > l ? s2 : l;
  }
  double* o = near + j * NEAR_REC;
  for (int k = 0; k < 3; ++k) o[k] = (a[k] + b[k] + c[k]) / 3.0;
  o[3] = factor * factor * l;
  const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double nn[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
  const double inv = 1.0 / sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
  for (int k = 0; k < 3; ++k) o[4 + k] = nn[k] * inv;
  o[7] = sqrt(l);
}

// ---- half-space assembly with the free-surface correction shared per mesh edge -------------------------
// A triangle's free-surface strain is the sum over its three sides of AngSetupStrainFSC (TD.jl:1839-1863), and a
// side traversed the other way round gives the same field with the opposite sign (the pair of angular
// dislocations swaps its ends and orientation).  Two triangles with a common edge therefore need one evaluation
// of that edge's pair, not two: the edge is evaluated once per receiver in its owner's direction, for the three
// unit Burgers vectors of EFCS, turned into tractions on the receiver plane, and each triangle adds +- the
// responses of its edges weighted by its slip directions.
//
// Side id e = k n + j: side k of triangle j with end points (P1, P2), (P2, P3), (P3, P1) for k = 0, 1, 2 (the
// order of build_side).  Sides with identical end points (exact IEEE equality, either direction) form one edge
// whose owner is the lowest side id, so the edge set and every value are the same run to run.
__device__ __forceinline__ void side_ends(const double* P1, const double* P2, const double* P3, int64_t n, int64_t e,
                                          double a[3], double b[3]) {
  const int64_t j = e % n;
  const int k = (int)(e / n);
  const double* A = k == 0 ? P1 : (k == 1 ? P2 : P3);
  const double* B = k == 0 ? P2 : (k == 1 ? P3 : P1);
  for (int c = 0; c < 3; ++c) {
    a[c] = A[c * n + j];
    b[c] = B[c * n + j];
  }
}

// end points in lexicographic order; returns false when a coordinate is NaN (such a side matches nothing)
__device__ __forceinline__ bool side_key(const double* P1, const double* P2, const double* P3, int64_t n, int64_t e,
                                         double k[6]) {
  double a[3], b[3];
  side_ends(P1, P2, P3, n, e, a, b);
  bool swap = false;
  if (a[0] != b[0])
    swap = !(a[0] < b[0]);
  else if (a[1] != b[1])
    swap = !(a[1] < b[1]);
  else
    swap = !(a[2] <= b[2]);
  for (int c = 0; c < 3; ++c) {
    k[c] = (swap ? b[c] : a[c]) + 0.0;  // +0.0 folds -0.0 onto 0.0 (they compare equal)
    k[3 + c] = (swap ? a[c] : b[c]) + 0.0;
  }
  bool ok = true;
  for (int c = 0; c < 6; ++c) ok = ok && k[c] == k[c];
  return ok;
}

__device__ __forceinline__ uint64_t key_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) tdk_fill_i32(int* __restrict__ p, int64_t n, int v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Open-addressing table keyed by the sorted end points: slots[p] holds the first side stored there, own[p] the
// lowest side id with that key; slot_of[e] = p (-1 for a side with a NaN coordinate).
__global__ void __launch_bounds__(256) tdk_edge_insert(const double* __restrict__ P1, const double* __restrict__ P2,
                                                       const double* __restrict__ P3, int64_t n,
                                                       int* __restrict__ slots, int* __restrict__ own, uint64_t mask,
                                                       int* __restrict__ slot_of) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= 3 * n) return;
  double k[6];
  if (!side_key(P1, P2, P3, n, e, k)) {
    slot_of[e] = -1;
    return;
  }
  uint64_t h = 0x243f6a8885a308d3ull;
  for (int c = 0; c < 6; ++c) h = key_mix(h ^ (uint64_t)__double_as_longlong(k[c]));
  uint64_t p = h & mask;
  for (uint64_t probe = 0; probe <= mask; ++probe) {
    const int cur = atomicCAS(&slots[p], -1, (int)e);
    bool eq = cur == -1;
    if (!eq) {
      double q[6];
      side_key(P1, P2, P3, n, cur, q);
      eq = true;
      for (int c = 0; c < 6; ++c) eq = eq && q[c] == k[c];
    }
    if (eq) {
      atomicMin(&own[p], (int)e);
      slot_of[e] = (int)p;
      return;
    }
    p = (p + 1) & mask;
  }
  slot_of[e] = -1;  // table full (cannot happen: it has at least twice as many slots as sides)
}

// One workgroup of 1024 threads numbers the edges in side-id order: owners get consecutive edge indices
// (edge_side[t] = owner side of edge t), then every side records its edge and direction:
// side_edge[e] = 2 t + (1 if side e runs against its owner).  count[0] = number of edges, count[1] = 1 if a
// triangle lies in z = 0.
__global__ void __launch_bounds__(1024) tdk_edge_number(const double* __restrict__ P1, const double* __restrict__ P2,
                                                        const double* __restrict__ P3, int64_t n,
                                                        const int* __restrict__ own, const int* __restrict__ slot_of,
                                                        int* __restrict__ owner_idx, int* __restrict__ side_edge,
                                                        int* __restrict__ edge_side, int* __restrict__ count) {
  __shared__ int wsum[16];
  __shared__ int base;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int64_t ns = 3 * n;
  if (tid == 0) {
    base = 0;
    count[1] = 0;
  }
  __syncthreads();
  for (int64_t c0 = 0; c0 < ns; c0 += 1024) {
    const int64_t e = c0 + tid;
    int owner = -1;
    if (e < ns) owner = slot_of[e] < 0 ? (int)e : own[slot_of[e]];
    const bool is_own = e < ns && owner == (int)e;
    const unsigned long long m = __ballot(is_own);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wid] = __popcll(m);
    __syncthreads();
    int off = base;
    for (int w = 0; w < wid; ++w) off += wsum[w];
    if (is_own) {
      owner_idx[e] = off + before;
      edge_side[off + before] = (int)e;
    }
    __syncthreads();
    if (tid == 0) {
      int tot = 0;
      for (int w = 0; w < 16; ++w) tot += wsum[w];
      base += tot;
    }
    __syncthreads();
  }
  // owners' indices are complete (the barrier above also orders the global writes within the workgroup)
  for (int64_t e = tid; e < ns; e += 1024) {
    const int o = slot_of[e] < 0 ? (int)e : own[slot_of[e]];
    double a[3], b[3], oa[3], ob[3];
    side_ends(P1, P2, P3, n, e, a, b);
    side_ends(P1, P2, P3, n, o, oa, ob);
    const bool rev = !(a[0] == oa[0] && a[1] == oa[1] && a[2] == oa[2]);
    side_edge[e] = 2 * owner_idx[o] + (rev ? 1 : 0);
  }
  // triangles lying in z = 0 (their image pass needs the fused kernel)
  for (int64_t j = tid; j < n; j += 1024)
    if (P1[2 * n + j] == 0.0 && P2[2 * n + j] == 0.0 && P3[2 * n + j] == 0.0) atomicOr(&count[1], 1);
  if (tid == 0) count[0] = base;
}

// Edge records (the side record of the owner side; its Burgers part is not used) for t < count.
__global__ void __launch_bounds__(128) tdk_edge_setup(const double* __restrict__ P1, const double* __restrict__ P2,
                                                      const double* __restrict__ P3, int64_t n,
                                                      const double* __restrict__ rec, const int* __restrict__ edge_side,
                                                      const int* __restrict__ count, double* __restrict__ erec) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= count[0]) return;
  const int64_t e = edge_side[t];
  double a[3], b[3];
  side_ends(P1, P2, P3, n, e, a, b);
  build_side(a, b, rec + (e % n) * REC_HS, erec + t * R_SIDE_SIZE);
}

// Edge responses: Tb[(t * 9 + 3 kE + q) * ldT + i] = traction q (Tn, Tss, Tds) on receiver i's plane of edge t's
// free-surface strain for the unit EFCS Burgers vector e_kE.
//
// A workgroup takes 64 receivers (one per lane) and blocks of kEdgeBlock edges.  The work of a block is split into
// (edge, ADCS Burgers component K) units: a unit evaluates three of the side's nine free-surface displacement
// outputs and their gradients (fsc_side_wk), turns the component's strain into its three tractions on the
// receiver's plane (edge_comp_traction) and leaves them in LDS.  After a barrier each (edge, kE) unit combines the
// three components' tractions into e_kE's (edge_unit_combine).  Units are dealt to the four waves in turn,
// so every wave evaluates each component equally often and the barrier waits on balanced waves.  One component per
// thread fits two waves per SIMD without accumulator-register copies; all nine outputs at once took ~470 registers
// per lane (one wave per SIMD, with 40 % of the issued instructions moving values to and from AGPRs).
constexpr int kEdgeBlock = 8;

__global__ void __launch_bounds__(256, 2) tdk_edge_response(const double* __restrict__ recv, int64_t nrecv,
                                                            const double* __restrict__ erec, int64_t nedge,
                                                            int64_t blk0, double nu, double lam, double G,
                                                            double* __restrict__ Tb, int64_t ldT) {
  __shared__ double Tl[kEdgeBlock][3][3][64];
  __shared__ double SGl[kEdgeBlock][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  const double* rv = recv + (i < nrecv ? i : nrecv - 1) * RECV;
  const double Xi = rv[0], Yi = rv[1], Zi = rv[2];
  const double k4 = (1.0 / (4.0 * kPi)) / (1.0 - nu);
  {
    const int64_t e0 = (blk0 + blockIdx.y) * kEdgeBlock;
    for (int u = wv; u < 3 * kEdgeBlock; u += 4) {
      const int eb = u / 3, K = u % 3;
      const int64_t t = e0 + eb;
      double t0 = 0.0, t1 = 0.0, t2 = 0.0;
      double sg = 1.0;
      if (t < nedge) {
        const double* s = erec + t * R_SIDE_SIZE;
        if (s[S_SKIP] == 0.0) {  // edge-uniform
          double w[6];
          if (K == 0)
            fsc_side_wk<0, double>(s, Xi, Yi, Zi, nu, w, sg);
          else if (K == 1)
            fsc_side_wk<1, double>(s, Xi, Yi, Zi, nu, w, sg);
          else
            fsc_side_wk<2, double>(s, Xi, Yi, Zi, nu, w, sg);
          // the receiver's frame is read here, not kept live through the evaluation above
          const double n[3] = {rv[3], rv[4], rv[5]}, sv[3] = {rv[6], rv[7], rv[8]}, d[3] = {rv[9], rv[10], rv[11]};
          edge_comp_traction(s, w, lam, G, n, sv, d, t0, t1, t2);
        }
      }
      Tl[eb][K][0][lane] = t0;
      Tl[eb][K][1][lane] = t1;
      Tl[eb][K][2][lane] = t2;
      if (K == 0) SGl[eb][lane] = sg;
    }
    __syncthreads();
    for (int u = wv; u < 3 * kEdgeBlock; u += 4) {
      const int eb = u / 3, kE = u % 3;
      const int64_t t = e0 + eb;
      if (t >= nedge) continue;
      const double* s = erec + t * R_SIDE_SIZE;
      double o0 = 0.0, o1 = 0.0, o2 = 0.0;
      if (s[S_SKIP] == 0.0)
        edge_unit_combine(s, kE, SGl[eb][lane], &Tl[eb][0][0][lane], &Tl[eb][1][0][lane], &Tl[eb][2][0][lane], 64, k4,
                          o0, o1, o2);
      if (i < nrecv) {
        double* T = Tb + (t * 9 + 3 * kE) * ldT + i;
        T[0] = o0;
        T[ldT] = o1;
        T[2 * ldT] = o2;
      }
    }
  }
}

// Traction entries of the half-space pairs: the real and image TDs evaluated here, the free-surface part as
// +- sum over the triangle's edges of sum_kE V_m[kE] Tb[edge][kE][q], V_m the unit slip vector of mode m
// (Vnorm, Vstrike, Vdip).  Not for triangles lying in z = 0.
__global__ void __launch_bounds__(256, 2) tdk_traction_combine(const double* __restrict__ recv, int64_t nrecv,
                                                            const double* __restrict__ rec, int64_t ntri, int tpb,
                                                            double nu, double lam, double G,
                                                            const int* __restrict__ side_edge,
                                                            const double* __restrict__ Tb, int64_t ldT,
                                                            OutPtrs<double> out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nrecv) return;
  const double* rv = recv + i * RECV;
  const double Xi = rv[0], Yi = rv[1], Zi = rv[2];
  const double n[3] = {rv[3], rv[4], rv[5]}, s[3] = {rv[6], rv[7], rv[8]}, d[3] = {rv[9], rv[10], rv[11]};
  const int64_t j0 = (int64_t)blockIdx.y * tpb;
  const int64_t j1 = j0 + tpb < ntri ? j0 + tpb : ntri;
  for (int64_t j = j0; j < j1; ++j) {
    const double* r = rec + j * REC_HS;
    double t[3][3];
    td_part_traction(r, Xi, Yi, Zi, nu, lam, G, n, s, d, t, true);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int se = side_edge[k * ntri + j];
      const double* T = Tb + (int64_t)(se >> 1) * 9 * ldT + i;
      const double sgn = (se & 1) ? -1.0 : 1.0;
      double v[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) v[q] = T[q * ldT];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const double* V = r + (m == 0 ? R_VN : (m == 1 ? R_VS : R_VD));
#pragma unroll
        for (int q = 0; q < 3; ++q) t[m][q] += sgn * (V[0] * v[q] + V[1] * v[3 + q] + V[2] * v[6 + q]);
      }
    }
    td_part_traction(r + OFF_IMG, Xi, Yi, Zi, nu, lam, G, n, s, d, t, false);
    const int64_t o = i + j * out.ld;
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int q = 0; q < 3; ++q) out.p[3 * m + q][o] = out.sign * t[m][q];
  }
}

// ---- fixed elements: traction rows at free receivers, displacement rows at fixed ones ------------------
// (PutDispInfsIntoInfMat / ConcatInfMats, CalculateInfluenceMatrices3D.jl:145-198).  Block k = 3 m + q of
// the output (mode m, traction q) takes rows of Tf block k at free receivers and rows of Ud block k (the
// DispInf field at the same position: DnTn <- DnUx, DnTss <- DnUy, ...) at fixed ones.  Only free sources
// were evaluated, so the columns are already those the reference keeps.  rowmap[r] = index among the free
// receivers, or -1 - index among the fixed ones.  Tf: 9 blocks nf x nsrc (ld nf); Ud: 9 blocks nfix x nsrc.
__global__ void __launch_bounds__(256) tdk_fixed_scatter(const double* __restrict__ Tf, const double* __restrict__ Ud,
                                                         const int32_t* __restrict__ rowmap, int64_t n, int64_t nf,
                                                         int64_t nfix, int64_t nsrc, OutPtrs<double> out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t map = rowmap[r];
  for (int64_t c = blockIdx.y; c < nsrc; c += gridDim.y) {
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double v = map >= 0 ? Tf[(int64_t)k * nf * nsrc + map + c * nf]
                                : Ud[(int64_t)k * nfix * nsrc + (-1 - map) + c * nfix];
      out.p[k][r + c * out.ld] = out.sign * v;
    }
  }
}

// ---- launch helpers --------------------------------------------------------------------------------------
uint64_t edge_table_slots(int64_t ntri) {
  uint64_t s = 64;
  while (s < (uint64_t)(6 * ntri)) s <<= 1;
  return s;
}

void launch_edge_topology(const double* P1, const double* P2, const double* P3, int64_t ntri, const EdgeTopo& e,
                          hipStream_t st) {
  if (ntri <= 0) return;
  const int64_t ns = 3 * ntri;
  const uint64_t S = edge_table_slots(ntri);
  hipLaunchKernelGGL(tdk_fill_i32, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, st, e.slots, (int64_t)S, -1);
  hipLaunchKernelGGL(tdk_fill_i32, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, st, e.own, (int64_t)S,
                     0x7fffffff);
  hipLaunchKernelGGL(tdk_edge_insert, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st, P1, P2, P3, ntri, e.slots,
                     e.own, S - 1, e.slot_of);
  hipLaunchKernelGGL(tdk_edge_number, dim3(1), dim3(1024), 0, st, P1, P2, P3, ntri, e.own, e.slot_of, e.owner_idx,
                     e.side_edge, e.edge_side, e.count);
}

void launch_edge_setup(const double* P1, const double* P2, const double* P3, int64_t ntri, const double* rec,
                       const EdgeTopo& e, double* erec, hipStream_t st) {
  if (ntri <= 0) return;
  hipLaunchKernelGGL(tdk_edge_setup, dim3((unsigned)((3 * ntri + 127) / 128)), dim3(128), 0, st, P1, P2, P3, ntri, rec,
                     e.edge_side, e.count, erec);
}

void launch_edge_response(const double* recv, int64_t nrecv, const double* erec, int64_t nedge, double nu,
                          double lam, double G, double* Tb, int64_t ldT, hipStream_t st) {
  if (nrecv <= 0 || nedge <= 0) return;
  const int64_t nblk = (nedge + kEdgeBlock - 1) / kEdgeBlock;
  for (int64_t b0 = 0; b0 < nblk; b0 += 65535) {
    const int64_t nb = nblk - b0 < 65535 ? nblk - b0 : 65535;
    hipLaunchKernelGGL(tdk_edge_response, dim3((unsigned)((nrecv + 63) / 64), (unsigned)nb), dim3(256), 0, st, recv,
                       nrecv, erec, nedge, b0, nu, lam, G, Tb, ldT);
  }
}

void launch_traction_combine(const double* recv, int64_t nrecv, const double* rec, int64_t ntri, int tpb, double nu,
                             double lam, double G, const int* side_edge, const double* Tb, int64_t ldT,
                             const OutPtrs<double>& out, hipStream_t st) {
  if (nrecv <= 0 || ntri <= 0) return;
  dim3 grid((unsigned)((nrecv + 255) / 256), (unsigned)((ntri + tpb - 1) / tpb));
  hipLaunchKernelGGL(tdk_traction_combine, grid, dim3(256), 0, st, recv, nrecv, rec, ntri, tpb, nu, lam, G, side_edge,
                     Tb, ldT, out);
}

void launch_fixed_scatter(const double* Tf, const double* Ud, const int32_t* rowmap, int64_t n, int64_t nf,
                          int64_t nfix, int64_t nsrc, const OutPtrs<double>& out, hipStream_t st) {
  if (n <= 0 || nsrc <= 0) return;
  const unsigned gy = (unsigned)(nsrc < 65535 ? nsrc : 65535);
  hipLaunchKernelGGL(tdk_fixed_scatter, dim3((unsigned)((n + 255) / 256), gy), dim3(256), 0, st, Tf, Ud, rowmap, n,
                     nf, nfix, nsrc, out);
}

void launch_setup(const double* P1, const double* P2, const double* P3, int64_t ntri, int hs, double* rec,
                  hipStream_t st) {
  if (ntri <= 0) return;
  hipLaunchKernelGGL(tdk_setup, dim3((unsigned)((ntri + 127) / 128)), dim3(128), 0, st, P1, P2, P3, ntri, hs, rec);
}

void launch_rec_to_f32(const double* a, float* b, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(tdk_rec_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, b, n);
}

void launch_receivers(const double* nrm, const double* mid, int64_t n, double* recv, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(tdk_receivers, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, nrm, mid,
                     (const double*)nullptr, n, recv);
}

void launch_receivers_at(const double* nrm, const double* pts, int64_t n, double* recv, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(tdk_receivers, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, st, nrm, (const double*)nullptr,
                     pts, n, recv);
}

void launch_check_hs(const double* Z, int64_t nobs, const double* P1, const double* P2, const double* P3,
                     int64_t ntri, int* flag, hipStream_t st, int64_t zstride) {
  const int64_t n = nobs > ntri ? nobs : ntri;
  if (n <= 0) return;
  hipLaunchKernelGGL(tdk_check_hs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Z, zstride, nobs, P1, P2, P3,
                     ntri, flag);
}

__global__ void tdk_zero_u32(unsigned* __restrict__ p, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

void launch_zero_u32(unsigned* p, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(tdk_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
}

template <class T>
void launch_influence(const T* X, const T* Y, const T* Z, int64_t nobs, const T* rec, int64_t ntri, int hs, int tpb,
                      const double* Dn, const double* Dss, const double* Dds, T nu, const OutPtrs<T>& out,
                      hipStream_t st) {
  if (nobs <= 0 || ntri <= 0) return;
  dim3 grid((unsigned)((nobs + 255) / 256), (unsigned)((ntri + tpb - 1) / tpb));
  if (hs)
    hipLaunchKernelGGL((tdk_influence<T, true>), grid, dim3(256), 0, st, X, Y, Z, nobs, rec, ntri, tpb, Dn, Dss, Dds,
                       nu, out);
  else
    hipLaunchKernelGGL((tdk_influence<T, false>), grid, dim3(256), 0, st, X, Y, Z, nobs, rec, ntri, tpb, Dn, Dss, Dds,
                       nu, out);
}

template <class T>
void launch_field(const T* X, const T* Y, const T* Z, int64_t nobs, const T* rec, int64_t ntri, int hs, int tpb,
                  const double* Dn, const double* Dss, const double* Dds, T nu, int strain_on, int disp_on, T* part,
                  double* out, hipStream_t st) {
  if (nobs <= 0) return;
  const int ncomp = (strain_on ? 6 : 0) + (disp_on ? 3 : 0);
  const int nchunk = ntri > 0 ? (int)((ntri + tpb - 1) / tpb) : 0;
  if (nchunk > 0) {
    dim3 grid((unsigned)((nobs + 255) / 256), (unsigned)nchunk);
    if (hs)
      hipLaunchKernelGGL((tdk_field_part<T, true>), grid, dim3(256), 0, st, X, Y, Z, nobs, rec, ntri, tpb, Dn, Dss,
                         Dds, nu, strain_on, disp_on, part);
    else
      hipLaunchKernelGGL((tdk_field_part<T, false>), grid, dim3(256), 0, st, X, Y, Z, nobs, rec, ntri, tpb, Dn, Dss,
                         Dds, nu, strain_on, disp_on, part);
  }
  const int64_t tot = nobs * ncomp;
  hipLaunchKernelGGL((tdk_field_reduce<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, part, nobs, ncomp,
                     nchunk, out);
}

template <class T>
void launch_traction(const double* recv, int64_t nrecv, const T* rec, int64_t ntri, int hs, int tpb, T nu, T lam, T G,
                     const OutPtrs<T>& out, hipStream_t st) {
  if (nrecv <= 0 || ntri <= 0) return;
  dim3 grid((unsigned)((nrecv + 255) / 256), (unsigned)((ntri + tpb - 1) / tpb));
  if (hs)
    hipLaunchKernelGGL((tdk_traction<T, T, true>), grid, dim3(256), 0, st, recv, nrecv, rec, ntri, tpb, nu, lam, G, out);
  else
    hipLaunchKernelGGL((tdk_traction<T, T, false>), grid, dim3(256), 0, st, recv, nrecv, rec, ntri, tpb, nu, lam, G, out);
}

void launch_traction_f32_to_f64(const double* recv, int64_t nrecv, const float* rec, int64_t ntri, int hs, int tpb,
                                float nu, float lam, float G, const OutPtrs<double>& out, hipStream_t st) {
  if (nrecv <= 0 || ntri <= 0) return;
  dim3 grid((unsigned)((nrecv + 255) / 256), (unsigned)((ntri + tpb - 1) / tpb));
  if (hs)
    hipLaunchKernelGGL((tdk_traction<float, double, true>), grid, dim3(256), 0, st, recv, nrecv, rec, ntri, tpb, nu,
                       lam, G, out);
  else
    hipLaunchKernelGGL((tdk_traction<float, double, false>), grid, dim3(256), 0, st, recv, nrecv, rec, ntri, tpb, nu,
                       lam, G, out);
}

void launch_near_info(const double* P1, const double* P2, const double* P3, int64_t ntri, double factor, double* near,
                      hipStream_t st) {
  if (ntri <= 0) return;
  hipLaunchKernelGGL(tdk_near_info, dim3((unsigned)((ntri + 127) / 128)), dim3(128), 0, st, P1, P2, P3, ntri, factor,
                     near);
}

void launch_nearfield(const double* recv, int64_t nrecv, const double* rec, const double* near, int64_t ntri, int hs,
                      int tpb, double nu, double lam, double G, const OutPtrs<double>& out, unsigned long long* count,
                      hipStream_t st) {
  if (nrecv <= 0 || ntri <= 0) return;
  dim3 grid((unsigned)((nrecv + 255) / 256), (unsigned)((ntri + tpb - 1) / tpb));
  if (hs)
    hipLaunchKernelGGL(tdk_nearfield<true>, grid, dim3(256), 0, st, recv, nrecv, rec, near, ntri, tpb, nu, lam, G, out,
                       count);
  else
    hipLaunchKernelGGL(tdk_nearfield<false>, grid, dim3(256), 0, st, recv, nrecv, rec, near, ntri, tpb, nu, lam, G,
                       out, count);
}

template void launch_influence<double>(const double*, const double*, const double*, int64_t, const double*, int64_t,
                                       int, int, const double*, const double*, const double*, double,
                                       const OutPtrs<double>&, hipStream_t);
template void launch_influence<float>(const float*, const float*, const float*, int64_t, const float*, int64_t, int,
                                      int, const double*, const double*, const double*, float, const OutPtrs<float>&,
                                      hipStream_t);
template void launch_field<double>(const double*, const double*, const double*, int64_t, const double*, int64_t, int,
                                   int, const double*, const double*, const double*, double, int, int, double*,
                                   double*, hipStream_t);
template void launch_traction<double>(const double*, int64_t, const double*, int64_t, int, int, double, double, double,
                                      const OutPtrs<double>&, hipStream_t);
template void launch_traction<float>(const double*, int64_t, const float*, int64_t, int, int, float, float, float,
                                     const OutPtrs<float>&, hipStream_t);

}  // namespace cdj
