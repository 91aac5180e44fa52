// Device-side triangular-dislocation (TD) math for gfx950.
//
// Layout and work split
//   * Every source triangle is reduced once, by tdk_setup, to a fixed-size record (FP64 or FP32) holding its
//     TDCS frame, vertex coordinates, angular-dislocation (AD) trig constants and, for the half-space, the
//     image triangle and the three free-surface-correction (FSC) side records.  A wavefront always works on
//     one triangle at a time, so the record address is wave-uniform and the record is read through the
//     scalar cache into SGPRs (a free broadcast to all 64 lanes).
//   * Each lane owns one observation point and accumulates the 3 slip modes x 6 strain components (or
//     3 x 3 displacements) of that pair in registers; the artefact-free configuration (TD.jl:1010-1079) and
//     the FSC configuration I/II (TD.jl:1943-1974) are per-lane *data* selects, not branches, so all lanes
//     run the same instruction stream.
//   * All influences are computed for unit slip and scaled by the caller's slip afterwards; a mode whose
//     slip is zero contributes exactly zero, as the reference's `if Dn != 0` guards do.
//   * FSC strains are obtained by forward-mode differentiation (3 directional derivatives carried with every
//     value) of the FSC displacement closed form: the strain closed form of the reference (AngDisStrainFSC,
//     TD.jl:2084-2472) is the symmetric gradient of its displacement closed form (AngDisDispFSC,
//     TD.jl:1412-1496) divided by 4*pi*(1-nu).
//
// Citations are to the reference's files; `TD.jl` is src/TD.jl.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "td_kernels.hpp"

namespace cdj {

static constexpr double kPi = 3.141592653589793;
static constexpr double kEps = 2.220446049250313e-16;

// ---- record layout (element offsets) --------------------------------------------------------------------
// TD part (one per real or image triangle), 32 elements
enum : int {
  R_VN = 0,     // normal (3)
  R_VS = 3,     // strike (3)
  R_VD = 6,     // dip (3)
  R_P2 = 9,     // EFCS P2, TDCS origin (3)
  R_P1Y = 12,   // TDCS p1 (y, z)
  R_P1Z = 13,
  R_P3Y = 14,   // TDCS p3 (y, z)
  R_P3Z = 15,
  R_E13Y = 16,  // unit side vectors in TDCS (y, z components only are used)
  R_E13Z = 17,
  R_E12Y = 18,
  R_E12Z = 19,
  R_E23Y = 20,
  R_E23Z = 21,
  R_CA = 22,    // cos/sin(-pi + A), (-pi + B), (-pi + C)
  R_SA = 23,
  R_CB = 24,
  R_SB = 25,
  R_CC = 26,
  R_SC = 27,
  R_FLAT = 28,  // 1 if all three vertices have z == 0
  R_BASE = 29,  // barycentric denominator of trimodefinder
  R_TD_SIZE = 32
};
// FSC side part, 32 elements
enum : int {
  S_SKIP = 0,   // 1 when the side is (near) vertical or |cot(beta)| > 5e3 (TD.jl:1270, 1878)
  S_PA = 1,     // EFCS start vertex (3)
  S_EY1 = 4,    // ADCS first axis (x, y); ey2 = (ey1y, -ey1x, 0), ey3 = (0, 0, -1)
  S_YB = 6,     // side vector in ADCS (3)
  S_AA = 9,     // -PA.z, -PB.z
  S_AB = 10,
  S_B = 11,     // unit-slip Burgers vectors in ADCS, b[mode][k] (9)
  S_COSB = 20,  // beta
  S_SINB = 21,
  S_COTB = 22,
  S_COTB2 = 23,
  S_COSPS = 24,  // pi - beta (strain configuration I)
  S_SINPS = 25,
  S_COTPS = 26,
  S_COTPS2 = 27,
  S_COSPD = 28,  // -pi + beta (displacement configuration I)
  S_SINPD = 29,
  S_COTPD = 30,
  S_COTPD2 = 31,
  R_SIDE_SIZE = 32
};
static_assert(REC_FS == R_TD_SIZE, "full-space record size");
static_assert(REC_HS == 2 * R_TD_SIZE + 3 * R_SIDE_SIZE, "half-space record size");
static_assert(REC_EDGE == R_SIDE_SIZE, "edge record size");
static constexpr int OFF_IMG = R_TD_SIZE;
static constexpr int OFF_SIDE = 2 * R_TD_SIZE;

// ---- forward-mode dual number: value + d/dy1, d/dy2, d/dy3 ---------------------------------------------
// DS<T, M> carries the derivative directions in the bit set M (bit k: d/dy(k+1)); the others are zero by
// construction and are neither read nor computed.  The free-surface code starts from the coordinates y1, y2, y3
// (one direction each) and its first products keep a few directions only: with full duals their zero parts cost
// real instructions, because IEEE arithmetic does not let the compiler fold x * 0 (x may be infinite or NaN) or
// x + 0 (x may be -0).  D3<T> is the dense case.
template <class T, int M> struct DS {
  T v, a, b, c;
};
template <class T> using D3 = DS<T, 7>;

template <int k, class T, int M> __device__ __forceinline__ T dcomp(const DS<T, M>& x) {
  return k == 0 ? x.a : (k == 1 ? x.b : x.c);
}
template <int k, int M1, int M2, class T> __device__ __forceinline__ T d_add(T x, T y) {
  constexpr bool X = (M1 >> k) & 1, Y = (M2 >> k) & 1;
  if constexpr (X && Y) return x + y;
  else if constexpr (X) return x;
  else if constexpr (Y) return y;
  else return T(0);  // never read
}
template <int k, int M1, int M2, class T> __device__ __forceinline__ T d_sub(T x, T y) {
  constexpr bool X = (M1 >> k) & 1, Y = (M2 >> k) & 1;
  if constexpr (X && Y) return x - y;
  else if constexpr (X) return x;
  else if constexpr (Y) return -y;
  else return T(0);
}
template <int k, int M1, int M2, class T> __device__ __forceinline__ T d_mul(T xv, T xd, T yv, T yd) {
  constexpr bool X = (M1 >> k) & 1, Y = (M2 >> k) & 1;
  if constexpr (X && Y) return xd * yv + xv * yd;
  else if constexpr (X) return xd * yv;
  else if constexpr (Y) return xv * yd;
  else return T(0);
}
template <class T, int M1, int M2> __device__ __forceinline__ DS<T, M1 | M2> operator+(DS<T, M1> x, DS<T, M2> y) {
  return {x.v + y.v, d_add<0, M1, M2>(x.a, y.a), d_add<1, M1, M2>(x.b, y.b), d_add<2, M1, M2>(x.c, y.c)};
}
template <class T, int M1, int M2> __device__ __forceinline__ DS<T, M1 | M2> operator-(DS<T, M1> x, DS<T, M2> y) {
  return {x.v - y.v, d_sub<0, M1, M2>(x.a, y.a), d_sub<1, M1, M2>(x.b, y.b), d_sub<2, M1, M2>(x.c, y.c)};
}
template <class T, int M1, int M2> __device__ __forceinline__ DS<T, M1 | M2> operator*(DS<T, M1> x, DS<T, M2> y) {
  return {x.v * y.v, d_mul<0, M1, M2>(x.v, x.a, y.v, y.a), d_mul<1, M1, M2>(x.v, x.b, y.v, y.b),
          d_mul<2, M1, M2>(x.v, x.c, y.v, y.c)};
}
// scaling a present direction; absent ones stay unread
template <int k, int M, class T> __device__ __forceinline__ T d_scale(T x, T s) {
  if constexpr ((M >> k) & 1) return x * s;
  else return T(0);
}
template <int k, int M, class T> __device__ __forceinline__ T d_neg(T x) {
  if constexpr ((M >> k) & 1) return -x;
  else return T(0);
}
template <class T, int M> __device__ __forceinline__ DS<T, M> operator-(DS<T, M> x) {
  return {-x.v, d_neg<0, M>(x.a), d_neg<1, M>(x.b), d_neg<2, M>(x.c)};
}
template <class T, int M> __device__ __forceinline__ DS<T, M> operator+(DS<T, M> x, T s) { return {x.v + s, x.a, x.b, x.c}; }
template <class T, int M> __device__ __forceinline__ DS<T, M> operator+(T s, DS<T, M> x) { return {x.v + s, x.a, x.b, x.c}; }
template <class T, int M> __device__ __forceinline__ DS<T, M> operator-(DS<T, M> x, T s) { return {x.v - s, x.a, x.b, x.c}; }
template <class T, int M> __device__ __forceinline__ DS<T, M> operator-(T s, DS<T, M> x) {
  return {s - x.v, d_neg<0, M>(x.a), d_neg<1, M>(x.b), d_neg<2, M>(x.c)};
}
template <class T, int M> __device__ __forceinline__ DS<T, M> operator*(DS<T, M> x, T s) {
  return {x.v * s, d_scale<0, M>(x.a, s), d_scale<1, M>(x.b, s), d_scale<2, M>(x.c, s)};
}
template <class T, int M> __device__ __forceinline__ DS<T, M> operator*(T s, DS<T, M> x) {
  return {x.v * s, d_scale<0, M>(x.a, s), d_scale<1, M>(x.b, s), d_scale<2, M>(x.c, s)};
}
// the dense form of a sparse dual (absent directions become explicit zeros)
template <class T, int M> __device__ __forceinline__ D3<T> dense(DS<T, M> x) {
  return {x.v, ((M >> 0) & 1) ? x.a : T(0), ((M >> 1) & 1) ? x.b : T(0), ((M >> 2) & 1) ? x.c : T(0)};
}
template <class T> __device__ __forceinline__ T dense(T x) { return x; }

// Reciprocals and square roots on the derivative path.  In FP64 an IEEE quotient is a sequence of about ten
// instructions (operand scaling, refinement, fix-up) and a square root about seventeen.  The forward-mode
// free-surface code takes them of finite, nonzero values in the normal range only (a zero or infinite operand is a
// point on a singular line of the closed form), so in FP64 it starts from the hardware's approximations and refines
// them: two Newton steps r <- r + r (1 - x r) for 1/x, and for sqrt(q) two steps of Goldschmidt's coupled iteration
// (g -> sqrt(q), h -> 1/(2 sqrt(q))) with a final Newton correction of g.  Both land within an ulp or two of the
// correctly rounded result.  Other scalar types divide.
template <class T> __device__ __forceinline__ T rcp_nr(T x) {
  if constexpr (std::is_same<T, double>::value) {
    T r = __builtin_amdgcn_rcp(x);
    T e = fma(-x, r, T(1));
    r = fma(r, e, r);
    e = fma(-x, r, T(1));
    return fma(r, e, r);
  } else {
    return T(1) / x;
  }
}
template <class T> __device__ __forceinline__ void sqrt_hrsqrt(T q, T& s, T& h) {
  if constexpr (std::is_same<T, double>::value) {
    const T y = __builtin_amdgcn_rsq(q);
    T g = q * y, hh = T(0.5) * y;
    T r = fma(-hh, g, T(0.5));
    g = fma(g, r, g);
    hh = fma(hh, r, hh);
    r = fma(-hh, g, T(0.5));
    g = fma(g, r, g);
    hh = fma(hh, r, hh);
    const T d = fma(-g, g, q);
    s = fma(d, hh, g);
    h = hh;
  } else {
    s = sqrt(q);
    h = T(0.5) / s;
  }
}
template <class T, int M> __device__ __forceinline__ DS<T, M> rcp(DS<T, M> x) {
  const T r = rcp_nr(x.v);
  const T m = -r * r;
  return {r, d_scale<0, M>(x.a, m), d_scale<1, M>(x.b, m), d_scale<2, M>(x.c, m)};
}
template <class T> __device__ __forceinline__ T rcp(T x) { return T(1) / x; }
template <class T, int M> __device__ __forceinline__ DS<T, M> dsqrt(DS<T, M> x) {
  T s, h;
  sqrt_hrsqrt(x.v, s, h);
  return {s, d_scale<0, M>(x.a, h), d_scale<1, M>(x.b, h), d_scale<2, M>(x.c, h)};
}
template <class T> __device__ __forceinline__ T dsqrt(T x) { return sqrt(x); }
// sqrt(q) and its reciprocal together (the reciprocal is 2 h, h from the same iteration)
template <class T, int M> __device__ __forceinline__ void dsqrt_rcp(DS<T, M> q, DS<T, M>& s, DS<T, M>& is) {
  T v, h;
  sqrt_hrsqrt(q.v, v, h);
  s = {v, d_scale<0, M>(q.a, h), d_scale<1, M>(q.b, h), d_scale<2, M>(q.c, h)};
  const T r = T(2) * h;
  const T m = -r * r;
  is = {r, d_scale<0, M>(s.a, m), d_scale<1, M>(s.b, m), d_scale<2, M>(s.c, m)};
}
template <class T> __device__ __forceinline__ void dsqrt_rcp(T q, T& s, T& is) {
  s = sqrt(q);
  is = T(1) / s;
}
template <class T, int M> __device__ __forceinline__ DS<T, M> dlog(DS<T, M> x) {
  const T r = rcp_nr(x.v);
  return {log(x.v), d_scale<0, M>(x.a, r), d_scale<1, M>(x.b, r), d_scale<2, M>(x.c, r)};
}
template <class T> __device__ __forceinline__ T dlog(T x) { return log(x); }
template <class T, int M> __device__ __forceinline__ DS<T, M> datan(DS<T, M> x) {
  const T r = rcp_nr(T(1) + x.v * x.v);
  return {atan(x.v), d_scale<0, M>(x.a, r), d_scale<1, M>(x.b, r), d_scale<2, M>(x.c, r)};
}
template <class T> __device__ __forceinline__ T datan(T x) { return atan(x); }
// 2*atan(-y2/D), the FSC Burgers' function (TD.jl:1430), with the derivative written as
// 2*(y2 dD - D dy2)/(D^2 + y2^2) so it stays finite where D crosses zero.
template <class T> __device__ __forceinline__ T burgers2(T y2, T D) { return T(2) * atan(-y2 / D); }
template <class T, int M1, int M2> __device__ __forceinline__ DS<T, M1 | M2> burgers2(DS<T, M1> y2, DS<T, M2> D) {
  const T f = T(2) * rcp_nr(D.v * D.v + y2.v * y2.v);
  // y2 dD - D dy2, per direction
  const DS<T, M1 | M2> g = DS<T, M2>{T(0), D.a, D.b, D.c} * y2.v - DS<T, M1>{T(0), y2.a, y2.b, y2.c} * D.v;
  return {T(2) * atan(-y2.v / D.v), d_scale<0, M1 | M2>(g.a, f), d_scale<1, M1 | M2>(g.b, f),
          d_scale<2, M1 | M2>(g.c, f)};
}

// The type of a value with the derivative directions M: itself for a plain scalar, DS<T, M> for a dual.
template <class S, int M> struct SparseOf {
  using type = S;
};
template <class T, int M0, int M> struct SparseOf<DS<T, M0>, M> {
  using type = DS<T, M>;
};
template <class S, int M> using Sp = typename SparseOf<S, M>::type;

// ---- small linear algebra -------------------------------------------------------------------------------
// T' = A T A^T with A given by its three columns c0, c1, c2 (TensorTransformation3D!.jl:35-53 with A in
// column-major order).  t = {xx, yy, zz, xy, xz, yz}.
template <class T>
__device__ __forceinline__ void tensor_rotate(T* t, const T* c0, const T* c1, const T* c2) {
  // A(i,a) = c_a[i]
  const T Txx = t[0], Tyy = t[1], Tzz = t[2], Txy = t[3], Txz = t[4], Tyz = t[5];
  T o[6];
  const int I[6] = {0, 1, 2, 0, 0, 1};
  const int J[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int i = I[q], j = J[q];
    const T a0 = c0[i], a1 = c1[i], a2 = c2[i];
    const T b0 = c0[j], b1 = c1[j], b2 = c2[j];
    o[q] = a0 * b0 * Txx + a1 * b1 * Tyy + a2 * b2 * Tzz + (a0 * b1 + a1 * b0) * Txy + (a0 * b2 + a2 * b0) * Txz +
           (a1 * b2 + a2 * b1) * Tyz;
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) t[q] = o[q];
}

// ADCS -> TDCS for TDSetupS: B = [1 0 0; 0 c s; 0 -s c] (TD.jl:1661-1662), T' = B T B^T.
template <class T> __device__ __forceinline__ void rotate_b(T* t, T c, T s) {
  const T Txx = t[0], Tyy = t[1], Tzz = t[2], Txy = t[3], Txz = t[4], Tyz = t[5];
  const T cc = c * c, ss = s * s, cs = c * s;
  t[0] = Txx;
  t[1] = cc * Tyy + T(2) * cs * Tyz + ss * Tzz;
  t[2] = ss * Tyy - T(2) * cs * Tyz + cc * Tzz;
  t[3] = c * Txy + s * Txz;
  t[4] = -s * Txy + c * Txz;
  t[5] = -cs * Tyy + (cc - ss) * Tyz + cs * Tzz;
}

// ---- angular dislocation, full space, strain: AngDisStrain TD.jl:1753-1828 ------------------------------
// o = {xxBx,xxBy,xxBz, yyBx,yyBy,yyBz, zzBx,zzBy,zzBz, xyBx,xyBy,xyBz, xzBx,xzBy,xzBz, yzBx,yzBy,yzBz,
//      rFi_rx, rFi_ry, rFi_rz}; divisions are replaced by products with 1/r, 1/(r-z), 1/W.
template <class T> __device__ __forceinline__ void ad_strain(T x, T y, T z, T cosA, T sinA, T nu, T* o) {
  const T E2 = T(2) * nu + T(1);
  const T eta = y * cosA - z * sinA;
  const T zeta = y * sinA + z * cosA;
  const T x2 = x * x, y2 = y * y, z2 = z * z;
  const T r2 = x2 + y2 + z2;
  // r, 1/r, 1/(r-z), 1/W through rcp_nr / sqrt_hrsqrt: where one of them is infinite the point lies on the
  // triangle's edge, whose pairs are NaN by the configuration rule (mode 0) whatever these return
  T r, hr;
  sqrt_hrsqrt(r2, r, hr);
  const T ir = T(2) * hr;
  const T ir2 = ir * ir;
  const T ir3 = ir2 * ir;
  const T rmz = r - z;
  const T irz = rcp_nr(rmz);
  const T W = zeta - r;
  const T iW = rcp_nr(W);
  const T iWr = iW * ir;         // 1/(W r)
  const T iW2 = iW * iW;         // 1/W^2
  const T iW2r = iW2 * ir;       // 1/(W^2 r)
  const T iW2r2 = iW2 * ir2;     // 1/(W^2 r^2)
  const T iWr3 = iW * ir3;       // 1/(W r^3)
  const T irz1 = ir * irz;       // 1/(r (r-z))
  const T irz2 = ir2 * irz * irz;  // 1/(r^2 (r-z)^2)
  const T irz3 = ir3 * irz;      // 1/(r^3 (r-z))
  const T C = (r * cosA - z) * iWr;
  const T S = (r * sinA - y) * iWr;
  const T i4pi = T(1.0 / (4.0 * kPi));
  // 1/(r - zeta) = -1/W
  o[18] = (-eta * ir * iW - y * irz1) * i4pi;
  o[19] = (x * irz1 + cosA * x * ir * iW) * i4pi;
  o[20] = (-sinA * x * ir * iW) * i4pi;
  const T Fx = x2 * iW2r2 - x2 * iWr3;
  o[0] = eta * iWr + eta * Fx + y * irz1 - x2 * y * irz2 - x2 * y * irz3;
  o[1] = (E2 * iWr + Fx) * cosA + E2 * irz1 - x2 * irz2 - x2 * irz3;
  o[2] = E2 * iWr + Fx;
  const T Gy = iWr + S * S - y2 * iWr3;
  o[3] = Gy * eta + E2 * y * irz1 - y2 * y * irz2 - y2 * y * irz3 - T(2) * nu * cosA * S;
  o[4] = irz1 - y2 * irz2 - y2 * irz3 + Gy * cosA;
  o[5] = Gy;
  const T Gz = iWr + C * C - z2 * iWr3;
  o[6] = eta * iWr + eta * C * C - eta * z2 * iWr3 + y * z * ir3 + T(2) * nu * sinA * C;
  o[7] = Gz * cosA + z * ir3;
  o[8] = Gz;
  o[9] = x * y2 * irz2 - nu * x * irz1 + x * y2 * irz3 - nu * x * cosA * iWr + eta * x * S * iWr + eta * x * y * iWr3;
  o[10] = x2 * y * irz2 - nu * y * irz1 + x2 * y * irz3 + nu * cosA * S + x2 * y * cosA * iWr3 + x2 * cosA * S * iWr;
  o[11] = nu * S + x2 * S * iWr + x2 * y * iWr3;
  o[12] = -x * y * ir3 + nu * x * sinA * iWr + eta * x * C * iWr + eta * x * z * iWr3;
  o[13] = -x2 * ir3 + nu * ir + nu * cosA * C + x2 * z * cosA * iWr3 + x2 * cosA * C * iWr;
  o[14] = nu * C + x2 * C * iWr + x2 * z * iWr3;
  const T yczs = y * cosA + z * sinA;
  o[15] = y2 * ir3 - nu * ir - nu * cosA * C + nu * sinA * S + eta * sinA * cosA * iW2 - eta * yczs * iW2r +
          eta * y * z * iW2r2 - eta * y * z * iWr3;
  o[16] = y * ir3 + sinA * cosA * cosA * iW2 - cosA * yczs * iW2r + y * z * cosA * iW2r2 - y * z * cosA * iWr3;
  o[17] = y * z * iWr3 - sinA * cosA * iW2 + yczs * iW2r - y * z * iW2r2;
}

// ---- angular dislocation, full space, displacement: AngDisDisp TD.jl:1177-1220 ---------------------------
// o = {ux,uy,uz, vx,vy,vz, wx,wy,wz}
template <class T> __device__ __forceinline__ void ad_disp(T x, T y, T z, T cosA, T sinA, T nu, T* o) {
  const T E2 = T(1) - T(2) * nu;
  const T eta = y * cosA - z * sinA;
  T zeta = y * sinA + z * cosA;
  const T r = sqrt(x * x + y * y + z * z);
  zeta = zeta > r ? r : zeta;  // TD.jl:1190-1195
  z = z > r ? r : z;
  const T rMzeta = r - zeta;
  const T rMz = r - z;
  const T ir = T(1) / r;
  const T iz = T(1) / rMzeta;
  const T iz2 = T(1) / rMz;
  const T lz = log(rMzeta);
  const T lz2 = log(rMz);
  o[0] = x * y * ir * iz2 - x * eta * ir * iz;
  o[3] = eta * sinA * iz - y * eta * ir * iz + y * y * ir * iz2 + E2 * (cosA * lz - lz2);
  o[6] = eta * cosA * iz - y * ir - eta * z * ir * iz - E2 * sinA * lz;
  o[1] = x * x * cosA * ir * iz - x * x * ir * iz2 - E2 * (cosA * lz - lz2);
  o[4] = y * cosA * ir * iz - sinA * cosA * iz - y * ir * iz2;
  o[7] = z * cosA * ir * iz - cosA * cosA * iz + ir;
  o[2] = E2 * lz - x * x * ir * iz;
  o[5] = sinA * iz - y * ir * iz;
  o[8] = cosA * iz - z * ir * iz;
}

// ---- angular dislocation, half-space harmonic part, displacement: AngDisDispFSC TD.jl:1412-1496 ---------
// S is T (displacement) or D3<T> (displacement plus its gradient).  Outputs o = 0..8 are
// {ux,uy,uz, vx,vy,vz, wx,wy,wz} (u,v,w = component; x,y,z suffix = Burgers component); the 1/(4 pi (1-nu))
// factor is applied by callers.  The quantities shared by all nine outputs are formed once (FscBase), each
// output separately (fsc_out), so a caller can evaluate the outputs one at a time and keep register use low.
template <class T, class S> struct FscBase {
  T cosB, sinB, cotB, icosB, nu, a, c1, c2, nu2, cot2;
  Sp<S, 1> y1;
  Sp<S, 2> y2;
  Sp<S, 4> y3b;
  S rb, irb, iW6, iW7, lW6, lW7, Fib;
};

// y1, y2, y3: the point in the angular dislocation's coordinates; as duals each carries its own direction only.
template <class T, class S>
__device__ __forceinline__ void fsc_base(Sp<S, 1> y1, Sp<S, 2> y2, Sp<S, 4> y3, T cosB, T sinB, T cotB, T cotB2,
                                         T icosB, T nu, T a, FscBase<T, S>& f) {
  f.cosB = cosB;
  f.sinB = sinB;
  f.cotB = cotB;
  f.icosB = icosB;
  f.nu = nu;
  f.a = a;
  f.c1 = T(1) - T(2) * nu;
  f.c2 = T(1) - nu;
  f.nu2 = T(2) * nu;
  f.cot2 = cotB * cotB;
  f.y1 = y1;
  f.y2 = y2;
  f.y3b = y3 + T(2) * a;
  const Sp<S, 5> z3b = f.y3b * cosB - y1 * sinB;
  dsqrt_rcp(y1 * y1 + y2 * y2 + f.y3b * f.y3b, f.rb, f.irb);
  const S W6 = f.rb + f.y3b;
  const S W7 = f.rb + z3b;
  f.iW6 = rcp(W6);
  f.iW7 = rcp(W7);
  f.lW6 = dlog(W6);
  f.lW7 = dlog(W7);
  f.Fib = burgers2(y2, y1 - W6 * cotB2);
}

// Products and sums of FscBase members that several outputs use (no division or transcendental among them).
template <class T, class S> struct FscDer {
  Sp<S, 5> z1b;
  Sp<S, 4> y3ma;
  Sp<S, 2> y22;
  Sp<S, 1> y11;
  S irb2, irb3, aR, rcy, rsy, cpa, q6, q7, p1;
};

template <class T, class S> __device__ __forceinline__ FscDer<T, S> fsc_der(const FscBase<T, S>& f) {
  FscDer<T, S> d;
  d.z1b = f.y1 * f.cosB + f.y3b * f.sinB;
  d.irb2 = f.irb * f.irb;
  d.irb3 = d.irb2 * f.irb;
  d.aR = f.a * f.irb;
  d.y3ma = f.y3b - f.a;
  d.rcy = f.rb * f.cosB + f.y3b;
  d.rsy = f.rb * f.sinB - f.y1;
  d.cpa = f.cosB + d.aR;
  d.y22 = f.y2 * f.y2;
  d.y11 = f.y1 * f.y1;
  d.q6 = d.y3ma * f.iW6;
  d.q7 = d.y3ma * f.iW7;
  d.p1 = T(1) + d.aR * f.icosB;
  return d;
}

template <class T, class S>
__device__ __forceinline__ S fsc_out_d(const FscBase<T, S>& f, const FscDer<T, S>& dq, int o) {
  const T cosB = f.cosB, sinB = f.sinB, cotB = f.cotB, icosB = f.icosB, nu = f.nu, a = f.a, c1 = f.c1, c2 = f.c2,
          nu2 = f.nu2, cot2 = f.cot2;
  const auto &y1 = f.y1;
  const auto &y2 = f.y2;
  const auto &y3b = f.y3b;
  const S &irb = f.irb, &iW6 = f.iW6, &iW7 = f.iW7, &lW6 = f.lW6, &lW7 = f.lW7, &Fib = f.Fib;
  const auto &z1b = dq.z1b;
  const auto &y3ma = dq.y3ma;
  const auto &y22 = dq.y22;
  const auto &y11 = dq.y11;
  const S &irb2 = dq.irb2, &irb3 = dq.irb3, &aR = dq.aR, &rcy = dq.rcy, &rsy = dq.rsy, &cpa = dq.cpa, &q6 = dq.q6,
          &q7 = dq.q7, &p1 = dq.p1;
  switch (o) {
    case 0:  // ux
      return dense(-T(2) * c2 * c1 * cot2 * Fib + c1 * y2 * iW6 * ((c1 - aR) * cotB - y1 * iW6 * (nu + aR)) +
             c1 * cosB * cotB * y2 * iW7 * cpa + a * cotB * y2 * y3ma * irb3 +
             y2 * y3ma * irb * iW6 * (-c1 * cotB + y1 * iW6 * (nu2 + aR) + a * y1 * irb2) +
             y2 * y3ma * irb * iW7 *
                 (cosB * iW7 * (rcy * (c1 * cosB - aR) * cotB + T(2) * c2 * cosB * rsy) - a * cosB * cotB * y3b * irb2));
    case 1:  // uy
      return dense(c1 * ((T(2) * c2 * cot2 + nu) * lW6 - (T(2) * c2 * cot2 + T(1)) * cosB * lW7) +
             c1 * iW6 * (-c1 * cotB * y1 + nu * y3b - a + a * cotB * y1 * irb + y11 * iW6 * (nu + aR)) -
             c1 * cotB * iW7 * (cosB * z1b - a * icosB * rsy * irb) - a * cotB * y1 * y3ma * irb3 +
             q6 * (nu2 + irb * (c1 * cotB * y1 + a) - y11 * irb * iW6 * (nu2 + aR) - a * y11 * irb3) +
             cotB * q7 * (-cosB * sinB + a * icosB * y1 * y3b * irb3 + rsy * irb * (T(2) * c2 * cosB - rcy * iW7 * p1)));
    case 2:  // uz
      return dense(c1 * (y2 * iW6 * (T(1) + aR) - cosB * y2 * iW7 * cpa) - y2 * y3ma * irb * (a * irb2 + iW6) +
             cosB * y2 * y3ma * irb * iW7 * (rcy * iW7 * cpa + a * y3b * irb2));
    case 3:  // vx
      return dense(c1 * ((T(2) * c2 * cot2 - nu) * lW6 - (T(2) * c2 * cot2 + c1) * cosB * lW7) -
             c1 * iW6 * (y1 * cotB * (c1 - aR) + nu * y3b - a + y22 * iW6 * (nu + aR)) -
             c1 * cotB * z1b * iW7 * cpa - a * cotB * y1 * y3ma * irb3 +
             q6 * (-nu2 + irb * (c1 * cotB * y1 - a) + y22 * irb * iW6 * (nu2 + aR) + a * y22 * irb3) +
             q7 * (cosB * cosB - irb * (c1 * cotB * z1b + a * cosB) + a * cotB * y3b * z1b * irb3 -
                   irb * iW7 * (cosB * cosB * y22 - a * cotB * z1b * irb * rcy)));
    case 4:  // vy
      return dense(T(2) * c2 * c1 * cot2 * Fib + c1 * y2 * iW6 * (-(c1 - aR) * cotB + y1 * iW6 * (nu + aR)) -
             c1 * cotB * y2 * iW7 * p1 - a * cotB * y2 * y3ma * irb3 +
             y2 * y3ma * irb * iW6 * (c1 * cotB - nu2 * y1 * iW6 - a * y1 * irb * (irb + iW6)) +
             cotB * y2 * y3ma * irb * iW7 * (-T(2) * c2 * cosB + rcy * iW7 * p1 + a * icosB * y3b * irb2));
    case 5:  // vz
      return dense(c1 * (-sinB * lW7 - y1 * iW6 * (T(1) + aR) + z1b * iW7 * cpa) + y1 * y3ma * irb * (a * irb2 + iW6) -
             q7 * (sinB * (cosB - aR) + z1b * irb * (T(1) + a * y3b * irb2) -
                   irb * iW7 * (cosB * sinB * y22 - a * z1b * irb * rcy)));
    case 6:  // wx
      return dense(T(2) * c2 * (c1 * cotB * Fib + y2 * iW6 * (nu2 + aR) - cosB * y2 * iW7 * cpa) +
             y2 * y3ma * irb * (nu2 * iW6 + a * irb2) + cosB * y2 * y3ma * irb * iW7 * (c1 - rcy * iW7 * cpa - a * y3b * irb2));
    case 7:  // wy
      return dense(-T(2) * c2 * c1 * cotB * (lW6 - cosB * lW7) - T(2) * c2 * y1 * iW6 * (nu2 + aR) +
             T(2) * c2 * z1b * iW7 * cpa + y3ma * irb * (c1 * cotB - nu2 * y1 * iW6 - a * y1 * irb2) -
             q7 * (cosB * sinB + cotB * rcy * irb * (T(2) * c2 * cosB - rcy * iW7) +
                   aR * (sinB - y3b * z1b * irb2 - z1b * rcy * irb * iW7)));
    default:  // wz
      return dense(T(2) * c2 * Fib + T(2) * c2 * sinB * y2 * iW7 * cpa +
             sinB * y2 * y3ma * irb * iW7 * (T(1) + rcy * iW7 * cpa + a * y3b * irb2));
  }
}

// One output with its derived quantities formed on the spot (callers that evaluate outputs one at a time keep
// the live set small this way).
template <class T, class S> __device__ __forceinline__ S fsc_out(const FscBase<T, S>& f, int o) {
  return fsc_out_d<T, S>(f, fsc_der<T, S>(f), o);
}

// All nine outputs (order {ux,uy,uz, vx,vy,vz, wx,wy,wz}).
template <class T, class S>
__device__ __forceinline__ void fsc_disp(S y1, S y2, S y3, T cosB, T sinB, T cotB, T cotB2, T icosB, T nu, T a, S* o) {
  FscBase<T, S> f;
  fsc_base<T, S>(y1, y2, y3, cosB, sinB, cotB, cotB2, icosB, nu, a, f);
  const int map[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
#pragma unroll
  for (int q = 0; q < 9; ++q) o[q] = fsc_out<T, S>(f, map[q]);
}

// ---- per-pair drivers -----------------------------------------------------------------------------------
// TDCS coordinates of an EFCS point (GlobalToTDECoords TD.jl:891) and trimodefinder (TD.jl:1010-1079).
// Contraction is disabled here so that the configuration decisions round exactly like the oracle's.
template <class T>
__device__ __forceinline__ void tdcs_point(const T* __restrict__ r, T X, T Y, T Z, T& x, T& y, T& z, int& mode) {
#pragma clang fp contract(off)
  const T dX = X - r[R_P2 + 0], dY = Y - r[R_P2 + 1], dZ = Z - r[R_P2 + 2];
  x = (r[R_VN + 0] * dX) + (r[R_VN + 1] * dY) + (r[R_VN + 2] * dZ);
  y = (r[R_VS + 0] * dX) + (r[R_VS + 1] * dY) + (r[R_VS + 2] * dZ);
  z = (r[R_VD + 0] * dX) + (r[R_VD + 1] * dY) + (r[R_VD + 2] * dZ);
  // trimodefinder(y, z, x, ...): its (x, y, z) are the TDCS (y, z, x); p2 = (0, 0)
  const T p1y = r[R_P1Y], p1z = r[R_P1Z], p3y = r[R_P3Y], p3z = r[R_P3Z];
  const T p2_2_m_p3_2 = T(0) - p3z;
  const T p3_1_m_p2_1 = p3y - T(0);
  const T p1_1_m_p3_1 = p1y - p3y;
  const T p3_2_m_p1_2 = p3z - p1z;
  const T Base = r[R_BASE];
  const T a = (p2_2_m_p3_2 * (y - p3y) + p3_1_m_p2_1 * (z - p3z)) / Base;
  const T b = (p3_2_m_p1_2 * (y - p3y) + p1_1_m_p3_1 * (z - p3z)) / Base;
  const T c = T(1) - a - b;
  const bool casen = (a <= T(0) && b > c && c > a) || (b <= T(0) && c > a && a > b) || (c <= T(0) && a > b && b > c);
  const bool casez = !casen && ((a == T(0) && b >= T(0) && c >= T(0)) || (a >= T(0) && b == T(0) && c >= T(0)) ||
                                (a >= T(0) && b >= T(0) && c == T(0)));
  mode = casen ? -1 : ((casez && x == T(0)) ? 0 : 1);
}

// One angular dislocation of TDSetupS (TD.jl:1623-1750) at unit slip for all three modes, added in TDCS.
template <class T>
__device__ __forceinline__ void ad_strain_add(T x, T y, T z, T cosA, T sinA, T ty, T tz, T CosT, T SinT, T nu,
                                              T acc[3][6]) {
  const T Y = CosT * (y - ty) - SinT * (z - tz);
  const T Z = SinT * (y - ty) + CosT * (z - tz);
  T e[21];
  ad_strain(x, Y, Z, cosA, sinA, nu, e);
  const T K = T(1.0 / (8.0 * kPi)) / (T(1) - nu);  // 1/(8 pi (1-nu))
  const T sK = sinA * K;                            // sinADE1
  const T rx = e[18], ry = e[19], rz = e[20];
  {
    T t[6] = {rx + K * e[0], K * e[3], K * e[6], T(0.5) * ry - K * e[9], T(0.5) * rz - K * e[12], K * e[15]};
    rotate_b(t, CosT, SinT);
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[0][c] += t[c];
  }
  // Dss: (by, bz) = (CosT, SinT); Dds: (by, bz) = (-SinT, CosT)   (RotateObject2D TD.jl:1637-1638)
#pragma unroll
  for (int m = 1; m < 3; ++m) {
    const T by = (m == 1) ? CosT : -SinT;
    const T bz = (m == 1) ? SinT : CosT;
    const T E4 = by * x * K;
    const T E5 = by * K;
    const T bzx = bz * x * sK;
    const T bzs = bz * sK;
    T t[6] = {-E4 * e[1] + bzx * e[2],
              by * ry - E4 * e[4] + bzx * e[5],
              -E4 * e[7] + bz * rz + bzx * e[8],
              T(0.5) * by * rx + E5 * e[10] - bzs * e[11],
              E5 * e[13] + T(0.5) * bz * rx - bzs * e[14],
              T(0.5) * by * rz - E4 * e[16] + T(0.5) * bz * ry - bzx * e[17]};
    rotate_b(t, CosT, SinT);
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[m][c] += t[c];
  }
}

// One angular dislocation of TDSetupD (TD.jl:1081-1173) at unit slip, added in TDCS.
template <class T>
__device__ __forceinline__ void ad_disp_add(T x, T y, T z, T cosA, T sinA, T ty, T tz, T CosT, T SinT, T nu,
                                            T acc[3][3]) {
  const T Y = CosT * (y - ty) - SinT * (z - tz);
  const T Z = SinT * (y - ty) + CosT * (z - tz);
  T u[9];
  ad_disp(x, Y, Z, cosA, sinA, nu, u);
  const T K = T(1.0 / (8.0 * kPi)) / (T(1) - nu);
  const T sK = sinA * K;
  {
    const T a = K * u[3], b = K * u[6];
    acc[0][0] += K * u[0];
    acc[0][1] += CosT * a + SinT * b;
    acc[0][2] += -SinT * a + CosT * b;
  }
#pragma unroll
  for (int m = 1; m < 3; ++m) {
    const T by = (m == 1) ? CosT : -SinT;
    const T bz = (m == 1) ? SinT : CosT;
    acc[m][0] += by * K * u[1] + bz * sK * u[2];
    const T a = by * x * K * u[4] + bz * x * sK * u[5];
    const T b = by * x * K * u[7] + bz * x * sK * u[8];
    acc[m][1] += CosT * a + SinT * b;
    acc[m][2] += -SinT * a + CosT * b;
  }
}

// Full-space TD strain of one triangle record at one point, TDCS accumulation (TDstrainFS TD.jl:1528-1595).
// Returns the configuration (casep 1, casen -1, casez 0).
template <class T>
__device__ __forceinline__ int td_strain_tdcs(const T* __restrict__ r, T X, T Y, T Z, T nu, T acc[3][6]) {
  T x, y, z;
  int mode;
  tdcs_point(r, X, Y, Z, x, y, z, mode);
  const T sg = mode < 0 ? T(-1) : T(1);  // side-vector signs of configuration I vs II
  // A at p1 with -e13 (I) / e13 (II); B at p2 with e12 / -e12; C at p3 with e23 / -e23
  ad_strain_add(x, y, z, r[R_CA], r[R_SA], r[R_P1Y], r[R_P1Z], -sg * r[R_E13Z], -sg * r[R_E13Y], nu, acc);
  ad_strain_add(x, y, z, r[R_CB], r[R_SB], T(0), T(0), sg * r[R_E12Z], sg * r[R_E12Y], nu, acc);
  ad_strain_add(x, y, z, r[R_CC], r[R_SC], r[R_P3Y], r[R_P3Z], sg * r[R_E23Z], sg * r[R_E23Y], nu, acc);
  return mode;
}

template <class T>
__device__ __forceinline__ int td_disp_tdcs(const T* __restrict__ r, T X, T Y, T Z, T nu, T acc[3][3], T& Fi) {
  T x, y, z;
  int mode;
  tdcs_point(r, X, Y, Z, x, y, z, mode);
  const T sg = mode < 0 ? T(-1) : T(1);
  ad_disp_add(x, y, z, r[R_CA], r[R_SA], r[R_P1Y], r[R_P1Z], -sg * r[R_E13Z], -sg * r[R_E13Y], nu, acc);
  ad_disp_add(x, y, z, r[R_CB], r[R_SB], T(0), T(0), sg * r[R_E12Z], sg * r[R_E12Y], nu, acc);
  ad_disp_add(x, y, z, r[R_CC], r[R_SC], r[R_P3Y], r[R_P3Z], sg * r[R_E23Z], sg * r[R_E23Y], nu, acc);
  // Burgers' function (TD.jl:820-835)
  const T a1 = -x, a2 = r[R_P1Y] - y, a3 = r[R_P1Z] - z;
  const T b2 = -y, b3 = -z;
  const T c2 = r[R_P3Y] - y, c3 = r[R_P3Z] - z;
  const T a12 = a1 * a1;
  const T na = sqrt(a12 + a2 * a2 + a3 * a3);
  const T nb = sqrt(a12 + b2 * b2 + b3 * b3);
  const T nc = sqrt(a12 + c2 * c2 + c3 * c3);
  const T one = a1 * (b2 * c3 - b3 * c2) - a2 * (a1 * c3 - b3 * a1) + a3 * (a1 * c2 - b2 * a1);
  const T ab = a1 * a1 + a2 * b2 + a3 * b3;
  const T ac = a1 * a1 + a2 * c2 + a3 * c3;
  const T bc = a1 * a1 + b2 * c2 + b3 * c3;
  const T two = na * nb * nc + ab * nc + ac * nb + bc * na;
  Fi = T(-2) * atan2(one, two) * T(1.0 / (4.0 * kPi));
  return mode;
}

// Near-field rule of the mixed assembly: a pair is recomputed in FP64 when the collocation point lies within
// f L of the source's centroid, or within 0.1 L of the source's plane and 2 f L of its centroid (L the source's
// longest edge, f the near factor).  In FP32 the angular-dislocation terms cancel most for points in or near the
// source's plane: there the error falls from 1e-4 of the column's largest entry at one edge length to 3.5e-6 at
// 2-4 and 8e-7 at 4-8 edge lengths, against at most 2.2e-6 off the plane at any distance.
__device__ __forceinline__ bool near_pair(const double* __restrict__ nj, double X, double Y, double Z) {
  const double dx = X - nj[0], dy = Y - nj[1], dz = Z - nj[2];
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double h = dx * nj[4] + dy * nj[5] + dz * nj[6];
  return r2 < nj[3] || (fabs(h) < 0.1 * nj[7] && r2 < 4.0 * nj[3]);
}

template <class T> __device__ __forceinline__ T qnan() { return __builtin_nan(""); }
template <> __device__ __forceinline__ float qnan<float>() { return __builtin_nanf(""); }

// Free-surface correction of one side at one point, strain (AngSetupStrainFSC TD.jl:1935-2075), before the
// Burgers vector enters, for ADCS Burgers component K only: w[c] = sum over the side's two angular dislocations
// (B minus A) of {d1u1, d2u2, d3u3, d2u1+d1u2, d3u1+d1u3, d3u2+d2u3} of the FSC displacement for component K (the
// strain-FSC closed form without its 1/(4 pi (1-nu))); sg = -1 in configuration I (beta x >= 0), whose v13, v23
// and first two Burgers components are negated (TD.jl:1951-1962), +1 in configuration II.  The side must not be
// a skipped one.  One component needs three of the nine displacement outputs and their gradients; the quantities
// they share (FscBase, FscDer) are formed per call.
template <int K, class T>
__device__ __forceinline__ void fsc_side_wk(const T* __restrict__ s, T X, T Y, T Z, T nu, T w[6], T& sg) {
  T x, y, z;
  {
#pragma clang fp contract(off)
    const T dX = X - s[S_PA + 0], dY = Y - s[S_PA + 1], dZ = Z - s[S_PA + 2];
    const T p = s[S_EY1 + 0], q = s[S_EY1 + 1];
    x = (p * dX) + (q * dY) + (T(-0.0) * dZ);
    y = (q * dX) + ((-p) * dY) + (T(-0.0) * dZ);
    z = (T(0) * dX) + (T(0) * dY) + (T(-1) * dZ);
  }
  const bool cfg1 = x >= T(0);  // beta * x >= 0 with 0 < beta < pi
  sg = cfg1 ? T(-1) : T(1);
  const T cB = cfg1 ? s[S_COSPS] : s[S_COSB];
  const T sB = cfg1 ? s[S_SINPS] : s[S_SINB];
  const T ctB = cfg1 ? s[S_COTPS] : s[S_COTB];
  const T ctB2 = cfg1 ? s[S_COTPS2] : s[S_COTB2];
  const T icB = T(1) / cB;
#pragma unroll
  for (int c = 0; c < 6; ++c) w[c] = T(0);
#pragma unroll 1
  for (int ad = 0; ad < 2; ++ad) {
    const T px = ad == 0 ? x : x - s[S_YB + 0];
    const T py = ad == 0 ? y : y - s[S_YB + 1];
    const T pz = ad == 0 ? z : z - s[S_YB + 2];
    const T aa = ad == 0 ? s[S_AA] : s[S_AB];
    const T sgn = ad == 0 ? T(-1) : T(1);
    FscBase<T, D3<T>> fb;
    fsc_base<T, D3<T>>(DS<T, 1>{sg * px, T(1), T(0), T(0)}, DS<T, 2>{sg * py, T(0), T(1), T(0)},
                       DS<T, 4>{pz, T(0), T(0), T(1)}, cB, sB, ctB, ctB2, icB, nu, aa, fb);
    const FscDer<T, D3<T>> dq = fsc_der<T, D3<T>>(fb);
    // output 3 c + K is component c of the displacement for Burgers component K; its gradient feeds v_cc
    // (diagonal) and the two shear sums it appears in
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const D3<T> u = fsc_out_d<T, D3<T>>(fb, dq, 3 * c + K);
      const T g0 = sgn * u.a, g1 = sgn * u.b, g2 = sgn * u.c;
      if (c == 0) {
        w[0] += g0;
        w[3] += g1;
        w[4] += sg * g2;
      } else if (c == 1) {
        w[3] += g0;
        w[1] += g1;
        w[5] += sg * g2;
      } else {
        w[4] += sg * g0;
        w[5] += sg * g1;
        w[2] += g2;
      }
    }
  }
}

// The same evaluation as fsc_side_wk for all three components at once, sharing FscBase; kept inline here, since
// sharing the loop changed the register allocation of the fused kernels for the worse (more spill).
template <class T>
__device__ __forceinline__ void fsc_strain_side(const T* __restrict__ s, T X, T Y, T Z, T nu, T eps[3][6]) {
  if (s[S_SKIP] != T(0)) return;  // triangle-uniform
  T x, y, z;
  {
#pragma clang fp contract(off)
    const T dX = X - s[S_PA + 0], dY = Y - s[S_PA + 1], dZ = Z - s[S_PA + 2];
    const T p = s[S_EY1 + 0], q = s[S_EY1 + 1];
    x = (p * dX) + (q * dY) + (T(-0.0) * dZ);
    y = (q * dX) + ((-p) * dY) + (T(-0.0) * dZ);
    z = (T(0) * dX) + (T(0) * dY) + (T(-1) * dZ);
  }
  const bool cfg1 = x >= T(0);  // beta * x >= 0 with 0 < beta < pi
  const T sg = cfg1 ? T(-1) : T(1);
  const T cB = cfg1 ? s[S_COSPS] : s[S_COSB];
  const T sB = cfg1 ? s[S_SINPS] : s[S_SINB];
  const T ctB = cfg1 ? s[S_COTPS] : s[S_COTB];
  const T ctB2 = cfg1 ? s[S_COTPS2] : s[S_COTB2];
  const T icB = T(1) / cB;
  const T k4 = T(1.0 / (4.0 * kPi)) / (T(1) - nu);
  // v = k4 * {d1u1, d2u2, d3u3, d2u1+d1u2, d3u1+d1u3, d3u2+d2u3} of the FSC displacement (the strain-FSC
  // closed form); w[k][c] = vB - vA per Burgers component k.  Configuration I negates v13, v23 and the first
  // two Burgers components (TD.jl:1951-1962).
  T w[3][6];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < 6; ++c) w[k][c] = T(0);
#pragma unroll 1
  for (int ad = 0; ad < 2; ++ad) {
    const T px = ad == 0 ? x : x - s[S_YB + 0];
    const T py = ad == 0 ? y : y - s[S_YB + 1];
    const T pz = ad == 0 ? z : z - s[S_YB + 2];
    const T aa = ad == 0 ? s[S_AA] : s[S_AB];
    const T sgn = ad == 0 ? T(-1) : T(1);
    FscBase<T, D3<T>> fb;
    fsc_base<T, D3<T>>(DS<T, 1>{sg * px, T(1), T(0), T(0)}, DS<T, 2>{sg * py, T(0), T(1), T(0)},
                       DS<T, 4>{pz, T(0), T(0), T(1)}, cB, sB, ctB, ctB2, icB, nu, aa, fb);
    // one output at a time: output q = 3 c + k is component c of the displacement for Burgers component k;
    // its gradient feeds v_cc (diagonal), and half of the two shear sums it appears in
#pragma unroll 1
    for (int q = 0; q < 9; ++q) {
      const D3<T> u = fsc_out<T, D3<T>>(fb, q);
      const int c = q / 3, k = q % 3;
      const T g0 = sgn * u.a, g1 = sgn * u.b, g2 = sgn * u.c;
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        if (kk != k) continue;
        if (c == 0) {
          w[kk][0] += g0;
          w[kk][3] += g1;
          w[kk][4] += sg * g2;
        } else if (c == 1) {
          w[kk][3] += g0;
          w[kk][1] += g1;
          w[kk][5] += sg * g2;
        } else {
          w[kk][4] += sg * g0;
          w[kk][5] += sg * g1;
          w[kk][2] += g2;
        }
      }
    }
  }
  const T c0[3] = {s[S_EY1 + 0], s[S_EY1 + 1], T(0)};
  const T c1[3] = {s[S_EY1 + 1], -s[S_EY1 + 0], T(0)};
  const T c2[3] = {T(-0.0), T(-0.0), T(-1)};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const T b1 = sg * s[S_B + 3 * m + 0] * k4, b2 = sg * s[S_B + 3 * m + 1] * k4, b3 = s[S_B + 3 * m + 2] * k4;
    T e[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const T h = c < 3 ? T(1) : T(0.5);
      e[c] = h * (b1 * w[0][c] + b2 * w[1][c] + b3 * w[2][c]);
    }
    tensor_rotate(e, c0, c1, c2);
#pragma unroll
    for (int c = 0; c < 6; ++c) eps[m][c] += e[c];
  }
}

// Free-surface correction of one side at one point, displacement, added in EFCS (AngSetupDispFSC
// TD.jl:1324-1405).
template <class T>
__device__ __forceinline__ void fsc_disp_side(const T* __restrict__ s, T X, T Y, T Z, T nu, T u[3][3]) {
  if (s[S_SKIP] != T(0)) return;
  T x, y, z;
  {
#pragma clang fp contract(off)
    const T dX = X - s[S_PA + 0], dY = Y - s[S_PA + 1], dZ = Z - s[S_PA + 2];
    const T p = s[S_EY1 + 0], q = s[S_EY1 + 1];
    x = (p * dX) + (q * dY) + (T(-0.0) * dZ);
    y = (q * dX) + ((-p) * dY) + (T(-0.0) * dZ);
    z = (T(0) * dX) + (T(0) * dY) + (T(-1) * dZ);
  }
  const bool cfg1 = x >= T(0);
  const T cB = cfg1 ? s[S_COSPD] : s[S_COSB];
  const T sB = cfg1 ? s[S_SINPD] : s[S_SINB];
  const T ctB = cfg1 ? s[S_COTPD] : s[S_COTB];
  const T ctB2 = cfg1 ? s[S_COTPD2] : s[S_COTB2];
  const T icB = T(1) / cB;
  const T k4 = T(1.0 / (4.0 * kPi)) / (T(1) - nu);
  T uA[9], uB[9];
  fsc_disp<T, T>(x, y, z, cB, sB, ctB, ctB2, icB, nu, s[S_AA], uA);
  fsc_disp<T, T>(x - s[S_YB + 0], y - s[S_YB + 1], z - s[S_YB + 2], cB, sB, ctB, ctB2, icB, nu, s[S_AB], uB);
  const T p = s[S_EY1 + 0], q = s[S_EY1 + 1];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const T b1 = s[S_B + 3 * m + 0] * k4, b2 = s[S_B + 3 * m + 1] * k4, b3 = s[S_B + 3 * m + 2] * k4;
    T w[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      w[c] = b1 * (uB[3 * c] - uA[3 * c]) + b2 * (uB[3 * c + 1] - uA[3 * c + 1]) + b3 * (uB[3 * c + 2] - uA[3 * c + 2]);
    // ADCS -> EFCS with [ey1 ey2 ey3] (symmetric)
    u[m][0] += p * w[0] + q * w[1] + T(-0.0) * w[2];
    u[m][1] += q * w[0] - p * w[1] + T(-0.0) * w[2];
    u[m][2] += T(0) * w[0] + T(0) * w[1] - w[2];
  }
}

// All of one triangle's strain influence at one EFCS point, unit slip, EFCS result
// (ComputeStrainInfluences TD.jl:670-708).
template <class T, bool HS>
__device__ __forceinline__ void strain_pair(const T* __restrict__ rec, T X, T Y, T Z, T nu, T eps[3][6]) {
  T acc[3][6];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[m][c] = T(0);
  const int mode = td_strain_tdcs(rec, X, Y, Z, nu, acc);
  const T bad = mode == 0 ? qnan<T>() : T(0);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[m][c] += bad;
    tensor_rotate(acc[m], rec + R_VN, rec + R_VS, rec + R_VD);
#pragma unroll
    for (int c = 0; c < 6; ++c) eps[m][c] = acc[m][c];
  }
  if constexpr (HS) {
#pragma unroll 1
    for (int k = 0; k < 3; ++k) fsc_strain_side(rec + OFF_SIDE + k * R_SIDE_SIZE, X, Y, Z, nu, eps);
    const T* ri = rec + OFF_IMG;
    const bool flat = ri[R_FLAT] != T(0);
    if (!flat) {
      // image contribution rotated from the image TDCS and added (TD.jl:1509-1516, 1597-1602)
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[m][c] = T(0);
      const int mi = td_strain_tdcs(ri, X, Y, Z, nu, acc);
      const T badi = mi == 0 ? qnan<T>() : T(0);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[m][c] += badi;
        tensor_rotate(acc[m], ri + R_VN, ri + R_VS, ri + R_VD);
#pragma unroll
        for (int c = 0; c < 6; ++c) eps[m][c] += acc[m][c];
      }
    } else {
      // Triangle lying in z = 0: the reference rotates the accumulated tensor into the image TDCS in place,
      // then its xz/yz sign flips rebind local copies (TD.jl:1518-1525, 1610-1617), so the caller keeps the
      // image-TDCS xz/yz and receives xx..xy of the rotated-back flipped copy.
      const T ex[3] = {ri[R_VN + 0], ri[R_VS + 0], ri[R_VD + 0]};
      const T ey[3] = {ri[R_VN + 1], ri[R_VS + 1], ri[R_VD + 1]};
      const T ez[3] = {ri[R_VN + 2], ri[R_VS + 2], ri[R_VD + 2]};
      T keep[3][2];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[m][c] = eps[m][c];
        tensor_rotate(acc[m], ex, ey, ez);
        keep[m][0] = acc[m][4];
        keep[m][1] = acc[m][5];
        acc[m][4] = -acc[m][4];
        acc[m][5] = -acc[m][5];
      }
      const int mi = td_strain_tdcs(ri, X, Y, Z, nu, acc);
      const T badi = mi == 0 ? qnan<T>() : T(0);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[m][c] += badi;
        tensor_rotate(acc[m], ri + R_VN, ri + R_VS, ri + R_VD);
        eps[m][0] = acc[m][0];
        eps[m][1] = acc[m][1];
        eps[m][2] = acc[m][2];
        eps[m][3] = acc[m][3];
        eps[m][4] = keep[m][0];
        eps[m][5] = keep[m][1];
      }
    }
  }
}

template <class T>
__device__ __forceinline__ void rot_to_efcs(const T* __restrict__ r, const T* v, T* o) {
  o[0] = r[R_VN + 0] * v[0] + r[R_VS + 0] * v[1] + r[R_VD + 0] * v[2];
  o[1] = r[R_VN + 1] * v[0] + r[R_VS + 1] * v[1] + r[R_VD + 1] * v[2];
  o[2] = r[R_VN + 2] * v[0] + r[R_VS + 2] * v[1] + r[R_VD + 2] * v[2];
}

// All of one triangle's displacement influence at one EFCS point (ComputeDispInfluences TD.jl:631-668).
template <class T, bool HS>
__device__ __forceinline__ void disp_pair(const T* __restrict__ rec, T X, T Y, T Z, T nu, T u[3][3]) {
  T acc[3][3];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[m][c] = T(0);
  T Fi;
  const int mode = td_disp_tdcs(rec, X, Y, Z, nu, acc, Fi);
  const T bad = mode == 0 ? qnan<T>() : T(0);
#pragma unroll
  for (int m = 0; m < 3; ++m) {
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[m][c] += bad;
    acc[m][m] += Fi;  // Dn -> x, Dss -> y, Dds -> z in TDCS
    rot_to_efcs(rec, acc[m], u[m]);
  }
  if constexpr (HS) {
#pragma unroll 1
    for (int k = 0; k < 3; ++k) fsc_disp_side(rec + OFF_SIDE + k * R_SIDE_SIZE, X, Y, Z, nu, u);
    const T* ri = rec + OFF_IMG;
    const bool flat = ri[R_FLAT] != T(0);
    if (!flat) {
#pragma unroll
      for (int m = 0; m < 3; ++m)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[m][c] = T(0);
      T Fii;
      const int mi = td_disp_tdcs(ri, X, Y, Z, nu, acc, Fii);
      const T badi = mi == 0 ? qnan<T>() : T(0);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[m][c] += badi;
        acc[m][m] += Fii;
        T o[3];
        rot_to_efcs(ri, acc[m], o);
        u[m][0] += o[0];
        u[m][1] += o[1];
        u[m][2] += o[2];
      }
    } else {
      // flat image triangle: the Uz sign flip rebinds a local copy (TD.jl:759-763, 854-864)
      T keep[3];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const T v0 = u[m][0], v1 = u[m][1], v2 = u[m][2];
        acc[m][0] = ri[R_VN + 0] * v0 + ri[R_VN + 1] * v1 + ri[R_VN + 2] * v2;
        acc[m][1] = ri[R_VS + 0] * v0 + ri[R_VS + 1] * v1 + ri[R_VS + 2] * v2;
        acc[m][2] = ri[R_VD + 0] * v0 + ri[R_VD + 1] * v1 + ri[R_VD + 2] * v2;
        keep[m] = acc[m][2];
        acc[m][2] = -acc[m][2];
      }
      T Fii;
      const int mi = td_disp_tdcs(ri, X, Y, Z, nu, acc, Fii);
      const T badi = mi == 0 ? qnan<T>() : T(0);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[m][c] += badi;
        acc[m][m] += Fii;
        T o[3];
        rot_to_efcs(ri, acc[m], o);
        u[m][0] = o[0];
        u[m][1] = o[1];
        u[m][2] = keep[m];
      }
    }
  }
}

// strain of one mode -> normal and shear tractions on the receiver plane (n, s, d): Hooke's law
// (HookesLaw3DStrain2Stress!.jl:1-66) then CalculateNormalTraction3D!.jl, TractionVectorCartesianComponents3D!.jl
// and CalculateTractionInChosenDirection3D!.jl as used by CalculateInfluenceMatrices3D.jl:92-110.
template <class T>
__device__ __forceinline__ void traction3(const T* e, T lam, T G, const T* n, const T* s, const T* d, T& Tn, T& Tss,
                                          T& Tds) {
  const T G2 = T(2) * G;
  const T cons = lam * (e[0] + e[1] + e[2]);
  const T Sxx = G2 * e[0] + cons, Syy = G2 * e[1] + cons, Szz = G2 * e[2] + cons;
  const T Sxy = G2 * e[3], Sxz = G2 * e[4], Syz = G2 * e[5];
  const T nx = n[0], ny = n[1], nz = n[2];
  Tn = Sxx * nx * nx + Syy * ny * ny + Szz * nz * nz + T(2) * Sxy * nx * ny + T(2) * Syz * ny * nz +
       T(2) * Sxz * nz * nx;
  const T tx = Sxx * nx + Sxy * ny + Sxz * nz;
  const T ty = Sxy * nx + Syy * ny + Syz * nz;
  const T tz = Sxz * nx + Syz * ny + Szz * nz;
  // ((I - n n^T) t) . dir
  const T px = tx * (T(1) - nx * nx) - ty * (nx * ny) - tz * (nz * nx);
  const T py = -tx * (nx * ny) + ty * (T(1) - ny * ny) - tz * (ny * nz);
  const T pz = -tx * (nz * nx) - ty * (ny * nz) + tz * (T(1) - nz * nz);
  Tss = px * s[0] + py * s[1] + pz * s[2];
  Tds = px * d[0] + py * d[1] + pz * d[2];
}

// Tractions (t0, t1, t2) = (normal, strike, dip) on the receiver plane of one side's free-surface strain for unit
// ADCS Burgers component K, before the factor 1/(4 pi (1-nu)) and the configuration's sign: w (fsc_side_wk) holds
// {d1u1, d2u2, d3u3, d2u1+d1u2, d3u1+d1u3, d3u2+d2u3} in the side's ADCS.  Traction is invariant under a rotation
// of both the strain and the plane, so the receiver's normal n and strike/dip directions sv, d are turned into the
// ADCS (columns ey1 = (p, q, 0), ey2 = (q, -p, 0), ey3 = (0, 0, -1) in EFCS) instead of the strain into EFCS
// (TensorTransformation3D!.jl:35-53 followed by the traction of CalculateInfluenceMatrices3D.jl:92-110).
template <class T>
__device__ __forceinline__ void edge_comp_traction(const T* __restrict__ s, const T w[6], T lam, T G, const T* n,
                                                   const T* sv, const T* d, T& t0, T& t1, T& t2) {
  const T p = s[S_EY1 + 0], q = s[S_EY1 + 1];
  const T na[3] = {p * n[0] + q * n[1], q * n[0] - p * n[1], -n[2]};
  const T sa[3] = {p * sv[0] + q * sv[1], q * sv[0] - p * sv[1], -sv[2]};
  const T da[3] = {p * d[0] + q * d[1], q * d[0] - p * d[1], -d[2]};
  const T e[6] = {w[0], w[1], w[2], T(0.5) * w[3], T(0.5) * w[4], T(0.5) * w[5]};
  traction3(e, lam, G, na, sa, da, t0, t1, t2);
}

// Tractions (o0, o1, o2) of one side's free-surface strain for the unit EFCS Burgers vector e_kE, from the component
// tractions tK[q * ts] (edge_comp_traction, K = 0, 1, 2) and the configuration sign sg; k4 = 1/(4 pi (1-nu)).
// e_kE in the side's ADCS is (ey1[kE], ey2[kE], ey3[kE]) = (p, q, 0), (q, -p, 0), (0, 0, -1) for kE = 0, 1, 2, and
// configuration I negates its first two components (TD.jl:1951-1962).
template <class T>
__device__ __forceinline__ void edge_unit_combine(const T* __restrict__ s, int kE, T sg, const T* t0, const T* t1,
                                                  const T* t2, int ts, T k4, T& o0, T& o1, T& o2) {
  const T p = s[S_EY1 + 0], q = s[S_EY1 + 1];
  const T g = sg * k4;
  if (kE == 2) {
    o0 = -k4 * t2[0];
    o1 = -k4 * t2[ts];
    o2 = -k4 * t2[2 * ts];
  } else {
    const T b1 = g * (kE == 0 ? p : q), b2 = g * (kE == 0 ? q : -p);
    o0 = b1 * t0[0] + b2 * t1[0];
    o1 = b1 * t0[ts] + b2 * t1[ts];
    o2 = b1 * t0[2 * ts] + b2 * t1[2 * ts];
  }
}

// One TD record's full-space strain (real, or image with the image record) at one point turned into the three
// tractions of each mode on the receiver plane and added to t[m][q] (first: stored instead).  Traction is linear in
// strain, so the parts of a half-space pair can be converted one by one (not for triangles lying in z = 0, whose
// image pass needs the accumulated strain; see strain_pair).
template <class T>
__device__ __forceinline__ void td_part_traction(const T* __restrict__ r, T X, T Y, T Z, T nu, T lam, T G, const T* n,
                                                 const T* s, const T* d, T t[3][3], bool first) {
  T acc[3][6];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[m][c] = T(0);
  const int mode = td_strain_tdcs(r, X, Y, Z, nu, acc);
  const T bad = mode == 0 ? qnan<T>() : T(0);
  // traction is invariant under a rotation of both the strain and the plane: the receiver's directions are turned
  // into the TDCS (components along Vnorm, Vstrike, Vdip) once, instead of each mode's strain into EFCS
  const T* Vn = r + R_VN;
  const T* Vs = r + R_VS;
  const T* Vd = r + R_VD;
  const T nt[3] = {Vn[0] * n[0] + Vn[1] * n[1] + Vn[2] * n[2], Vs[0] * n[0] + Vs[1] * n[1] + Vs[2] * n[2],
                   Vd[0] * n[0] + Vd[1] * n[1] + Vd[2] * n[2]};
  const T st[3] = {Vn[0] * s[0] + Vn[1] * s[1] + Vn[2] * s[2], Vs[0] * s[0] + Vs[1] * s[1] + Vs[2] * s[2],
                   Vd[0] * s[0] + Vd[1] * s[1] + Vd[2] * s[2]};
  const T dt[3] = {Vn[0] * d[0] + Vn[1] * d[1] + Vn[2] * d[2], Vs[0] * d[0] + Vs[1] * d[1] + Vs[2] * d[2],
                   Vd[0] * d[0] + Vd[1] * d[1] + Vd[2] * d[2]};
#pragma unroll
  for (int m = 0; m < 3; ++m) {
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[m][c] += bad;
    T a, b, c;
    traction3(acc[m], lam, G, nt, st, dt, a, b, c);
    t[m][0] = first ? a : t[m][0] + a;
    t[m][1] = first ? b : t[m][1] + b;
    t[m][2] = first ? c : t[m][2] + c;
  }
}

// The nine tractions t[m][q] (mode m, traction q) of one (receiver, source) pair on the receiver plane (n, s, d).
// Full space: from the TDCS strain, with the receiver's directions turned into the TDCS (td_part_traction).  Half
// space: from the EFCS strain of strain_pair, which adds the free-surface sides and the image (including the
// reference's flat-triangle branch) in EFCS.
template <class T, bool HS>
__device__ __forceinline__ void pair_traction(const T* __restrict__ rec, T X, T Y, T Z, T nu, T lam, T G, const T* n,
                                              const T* s, const T* d, T t[3][3]) {
  if constexpr (!HS) {
    td_part_traction(rec, X, Y, Z, nu, lam, G, n, s, d, t, true);
  } else {
    T e[3][6];
    strain_pair<T, HS>(rec, X, Y, Z, nu, e);
#pragma unroll
    for (int m = 0; m < 3; ++m) traction3(e[m], lam, G, n, s, d, t[m][0], t[m][1], t[m][2]);
  }
}

}  // namespace cdj
