// CPU FP64 restatement of CutAndDisplaceJulia's triangular-dislocation (TD) path.
//
// TEST INFRASTRUCTURE ONLY.  This is the oracle that the HIP product path is checked against; nothing in
// the product library links or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg load it.
//
// Every function follows the reference file:line it cites, in the same operation order, so that entry-level
// results agree with TD.jl to rounding.  The scalar type T is a template parameter so the same text can be
// instantiated with double (the oracle), with an operation-counting scalar (algorithmic flop counts) and
// with a forward-mode dual number (to pin the strain formulas against derivatives of the displacement
// formulas).
//
// The four angular-dislocation closed forms (AngDisDisp, AngDisStrain, AngDisDispFSC, AngDisStrainFSC) are
// the published expressions of Nikkhoo & Walter (2015, GJI 201:1119), which follow Comninou & Dundurs
// (1975); as formulas they have to be alike, and they are restated here expression by expression.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

namespace orc {

using std::acos;
using std::atan;
using std::atan2;
using std::cos;
using std::log;
using std::sin;
using std::sqrt;
using std::tan;

static constexpr double kPi = 3.141592653589793;  // Julia's `pi` rounded to Float64
static constexpr double kEps = 2.220446049250313e-16;  // Julia's eps()

template <class T> inline T pw2(T x) { return x * x; }
template <class T> inline T pw3(T x) { return x * x * x; }
template <class T> inline T pw5(T x) { return x * x * x * x * x; }
template <class T> inline T cot(T x) { return T(1.0) / tan(x); }

// RotateObject3DNewCoords!.jl:45-57 (scalar form): out_k = Ax1[k]*(X-Pa) + Ax2[k]*(Y-Pb) + Ax3[k]*(Z-Pc)
template <class T>
inline void rot3(T X, T Y, T Z, T Pa, T Pb, T Pc, const T* A1, const T* A2, const T* A3, T& ox, T& oy, T& oz) {
  const T x = (A1[0] * (X - Pa)) + (A2[0] * (Y - Pb)) + (A3[0] * (Z - Pc));
  const T y = (A1[1] * (X - Pa)) + (A2[1] * (Y - Pb)) + (A3[1] * (Z - Pc));
  const T z = (A1[2] * (X - Pa)) + (A2[2] * (Y - Pb)) + (A3[2] * (Z - Pc));
  ox = x;
  oy = y;
  oz = z;
}

// RotateObject2D!.jl:30-37 (scalar form) and RotateObject2D.jl:34-42 (same arithmetic once centred).
template <class T> inline void rot2(T X, T Y, T Pa, T Pb, T Ct, T St, T& ox, T& oy) {
  const T x = (Ct * (X - Pa)) + (-St * (Y - Pb));
  const T y = (St * (X - Pa)) + (Ct * (Y - Pb));
  ox = x;
  oy = y;
}

// TensorTransformation3D!.jl:35-53: T' = A T A^T, A given in Julia linear (column-major) order.
// t = {xx, yy, zz, xy, xz, yz}
template <class T> inline void tensor_tf(T* t, const T* A) {
  const T Txx = t[0], Tyy = t[1], Tzz = t[2], Txy = t[3], Txz = t[4], Tyz = t[5];
  const T txx = pw2(A[0]) * Txx + 2 * A[0] * A[3] * Txy + 2 * A[0] * A[6] * Txz + 2 * A[3] * A[6] * Tyz +
                pw2(A[3]) * Tyy + pw2(A[6]) * Tzz;
  const T tyy = pw2(A[1]) * Txx + 2 * A[1] * A[4] * Txy + 2 * A[1] * A[7] * Txz + 2 * A[4] * A[7] * Tyz +
                pw2(A[4]) * Tyy + pw2(A[7]) * Tzz;
  const T tzz = pw2(A[2]) * Txx + 2 * A[2] * A[5] * Txy + 2 * A[2] * A[8] * Txz + 2 * A[5] * A[8] * Tyz +
                pw2(A[5]) * Tyy + pw2(A[8]) * Tzz;
  const T txy = A[0] * A[1] * Txx + (A[0] * A[4] + A[1] * A[3]) * Txy + (A[0] * A[7] + A[1] * A[6]) * Txz +
                (A[3] * A[7] + A[4] * A[6]) * Tyz + A[3] * A[4] * Tyy + A[6] * A[7] * Tzz;
  const T txz = A[0] * A[2] * Txx + (A[0] * A[5] + A[2] * A[3]) * Txy + (A[0] * A[8] + A[2] * A[6]) * Txz +
                (A[3] * A[8] + A[5] * A[6]) * Tyz + A[3] * A[5] * Tyy + A[6] * A[8] * Tzz;
  const T tyz = A[1] * A[2] * Txx + (A[2] * A[4] + A[1] * A[5]) * Txy + (A[2] * A[7] + A[1] * A[8]) * Txz +
                (A[5] * A[7] + A[4] * A[8]) * Tyz + A[4] * A[5] * Tyy + A[7] * A[8] * Tzz;
  t[0] = txx;
  t[1] = tyy;
  t[2] = tzz;
  t[3] = txy;
  t[4] = txz;
  t[5] = tyz;
}

// cross!.jl:2-14
template <class T> inline void cross3(const T* a, const T* b, T* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// ---------------------------------------------------------------------------------------------------------
// Per-triangle frame, TDCS vertices, side vectors and angles.
// CalculateLocalTriCoords! (TD.jl:712-732), GlobalToTDECoords (TD.jl:870-908) for the vertices,
// CalcTDVectsAndAngles (TD.jl:962-1008).
template <class T> struct Tri {
  T P1[3], P2[3], P3[3];  // EFCS vertices
  T Vn[3], Vs[3], Vd[3];  // normal, strike, dip
  T p1[3], p2[3], p3[3];  // TDCS vertices (p2 = 0)
  T e12[3], e13[3], e23[3];
  T A, B, C;
  bool flat;  // all three vertices at z == 0 (TD.jl:759,1518)
};

template <class T> inline void tri_setup(const T* P1, const T* P2, const T* P3, Tri<T>& t) {
  for (int k = 0; k < 3; ++k) {
    t.P1[k] = P1[k];
    t.P2[k] = P2[k];
    t.P3[k] = P3[k];
  }
  // CalculateLocalTriCoords! TD.jl:719-730
  T a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = P2[k] - P1[k];
    b[k] = P3[k] - P1[k];
  }
  cross3(a, b, t.Vn);
  {
    const T nrm = sqrt(pw2(t.Vn[0]) + pw2(t.Vn[1]) + pw2(t.Vn[2]));
    for (int k = 0; k < 3; ++k) t.Vn[k] = t.Vn[k] / nrm;
  }
  const T eZ[3] = {T(0.0), T(0.0), T(1.0)};
  const T eY[3] = {T(0.0), T(1.0), T(0.0)};
  cross3(eZ, t.Vn, t.Vs);
  if (t.Vs[0] == 0.0 && t.Vs[1] == 0.0 && t.Vs[2] == 0.0) {
    for (int k = 0; k < 3; ++k) t.Vs[k] = eY[k] * t.Vn[2];
  }
  {
    const T nrm = sqrt(pw2(t.Vs[0]) + pw2(t.Vs[1]) + pw2(t.Vs[2]));
    for (int k = 0; k < 3; ++k) t.Vs[k] = t.Vs[k] / nrm;
  }
  cross3(t.Vn, t.Vs, t.Vd);
  // GlobalToTDECoords TD.jl:878-893: Vx=[Vn1,Vs1,Vd1], Vy=[Vn2,Vs2,Vd2], Vz=[Vn3,Vs3,Vd3]
  const T Vx[3] = {t.Vn[0], t.Vs[0], t.Vd[0]};
  const T Vy[3] = {t.Vn[1], t.Vs[1], t.Vd[1]};
  const T Vz[3] = {t.Vn[2], t.Vs[2], t.Vd[2]};
  rot3(P1[0], P1[1], P1[2], P2[0], P2[1], P2[2], Vx, Vy, Vz, t.p1[0], t.p1[1], t.p1[2]);
  t.p2[0] = t.p2[1] = t.p2[2] = T(0.0);
  rot3(P3[0], P3[1], P3[2], P2[0], P2[1], P2[2], Vx, Vy, Vz, t.p3[0], t.p3[1], t.p3[2]);
  // CalcTDVectsAndAngles TD.jl:972-1005
  for (int k = 0; k < 3; ++k) t.e12[k] = t.p2[k] - t.p1[k];
  {
    const T nrm = sqrt(pw2(t.e12[0]) + pw2(t.e12[1]) + pw2(t.e12[2]));
    for (int k = 0; k < 3; ++k) t.e12[k] = t.e12[k] / nrm;
  }
  for (int k = 0; k < 3; ++k) t.e13[k] = t.p3[k] - t.p1[k];
  {
    const T nrm = sqrt(pw2(t.e13[0]) + pw2(t.e13[1]) + pw2(t.e13[2]));
    for (int k = 0; k < 3; ++k) t.e13[k] = t.e13[k] / nrm;
  }
  for (int k = 0; k < 3; ++k) t.e23[k] = t.p3[k] - t.p2[k];
  {
    const T nrm = sqrt(pw2(t.e23[0]) + pw2(t.e23[1]) + pw2(t.e23[2]));
    for (int k = 0; k < 3; ++k) t.e23[k] = t.e23[k] / nrm;
  }
  t.A = acos((t.e12[0] * t.e13[0]) + (t.e12[1] * t.e13[1]) + (t.e12[2] * t.e13[2]));
  t.B = acos((-t.e12[0] * t.e23[0]) + (-t.e12[1] * t.e23[1]) + (-t.e12[2] * t.e23[2]));
  t.C = acos((t.e23[0] * t.e13[0]) + (t.e23[1] * t.e13[1]) + (t.e23[2] * t.e13[2]));
  t.flat = (P1[2] == 0.0 && P2[2] == 0.0 && P3[2] == 0.0);
}

// EFCS point -> TDCS (x = normal, y = strike, z = dip), origin at P2: GlobalToTDECoords TD.jl:891
template <class T> inline void to_tdcs(const Tri<T>& t, T X, T Y, T Z, T& x, T& y, T& z) {
  const T Vx[3] = {t.Vn[0], t.Vs[0], t.Vd[0]};
  const T Vy[3] = {t.Vn[1], t.Vs[1], t.Vd[1]};
  const T Vz[3] = {t.Vn[2], t.Vs[2], t.Vd[2]};
  rot3(X, Y, Z, t.P2[0], t.P2[1], t.P2[2], Vx, Vy, Vz, x, y, z);
}

// trimodefinder TD.jl:1010-1079, called as trimodefinder(y,z,x,p1[2:3],p2[2:3],p3[2:3]) (TD.jl:899).
// Returns +1 (casep, configuration I), -1 (casen, configuration II), 0 (casez: on a TD side, in its plane).
template <class T> inline int trimode(const Tri<T>& t, T x_t, T y_t, T z_t) {
  const T x = y_t, y = z_t, z = x_t;
  const T p1_1 = t.p1[1], p1_2 = t.p1[2], p2_1 = t.p2[1], p2_2 = t.p2[2], p3_1 = t.p3[1], p3_2 = t.p3[2];
  const T p2_2_m_p3_2 = p2_2 - p3_2;
  const T p3_1_m_p2_1 = p3_1 - p2_1;
  const T p1_1_m_p3_1 = p1_1 - p3_1;
  const T p1_2_m_p3_2 = p1_2 - p3_2;
  const T p3_2_m_p1_2 = p3_2 - p1_2;
  const T Base = (p2_2_m_p3_2 * p1_1_m_p3_1 + p3_1_m_p2_1 * p1_2_m_p3_2);
  const T a = (p2_2_m_p3_2 * (x - p3_1) + p3_1_m_p2_1 * (y - p3_2)) / Base;
  const T b = (p3_2_m_p1_2 * (x - p3_1) + p1_1_m_p3_1 * (y - p3_2)) / Base;
  const T c = T(1.0) - a - b;
  bool casen = false, casez = false;
  if (a <= 0 && b > c && c > a) casen = true;
  else if (b <= 0 && c > a && a > b) casen = true;
  else if (c <= 0 && a > b && b > c) casen = true;
  else if (a == 0 && b >= 0 && c >= 0) casez = true;
  else if (a >= 0 && b == 0 && c >= 0) casez = true;
  else if (a >= 0 && b >= 0 && c == 0) casez = true;
  if (casez && z != 0) casez = false;  // promoted to casep (TD.jl:1066-1069)
  if (casen) return -1;
  if (casez) return 0;
  return 1;
}

// ---------------------------------------------------------------------------------------------------------
// AngDisStrain TD.jl:1753-1828.  out = {xxBx,xxBy,xxBz, yyBx,yyBy,yyBz, zzBx,zzBy,zzBz, xyBx,xyBy,xyBz,
//                                        xzBx,xzBy,xzBz, yzBx,yzBy,yzBz, rFi_rx,rFi_ry,rFi_rz}
template <class T> inline void ang_dis_strain(T x, T y, T z, T cosA, T sinA, T nu, T E2, T* o) {
  const T eta = y * cosA - z * sinA;
  const T zeta = y * sinA + z * cosA;
  const T x2 = pw2(x), y2 = pw2(y), z2 = pw2(z);
  const T r2 = x2 + y2 + z2;
  const T r = sqrt(r2);
  const T r3 = r * r2;
  const T rmz = (r - z);
  const T rz = r * rmz;
  const T r2z2 = r2 * pw2(rmz);
  const T r3z = r3 * rmz;
  const T W = zeta - r;
  const T W2 = pw2(W);
  const T Wr = W * r;
  const T W2r = W2 * r;
  const T Wr3 = W * r3;
  const T W2r2 = W2 * r2;
  const T C = (r * cosA - z) / Wr;
  const T S = (r * sinA - y) / Wr;
  o[18] = (eta / r / (r - zeta) - y / r / rmz) / 4.0 / kPi;
  o[19] = (x / r / rmz - cosA * x / r / (r - zeta)) / 4.0 / kPi;
  o[20] = (sinA * x / r / (r - zeta)) / 4.0 / kPi;
  o[0] = (eta / Wr + eta * x2 / W2r2 - eta * x2 / Wr3 + y / rz - x2 * y / r2z2 - x2 * y / r3z);
  o[1] = ((E2 / Wr + x2 / W2r2 - x2 / Wr3) * cosA + E2 / rz - x2 / r2z2 - x2 / r3z);
  o[2] = (E2 / Wr + x2 / W2r2 - x2 / Wr3);
  o[3] = ((1.0 / Wr + pw2(S) - y2 / Wr3) * eta + E2 * y / rz - pw3(y) / r2z2 - pw3(y) / r3z - 2 * nu * cosA * S);
  o[4] = (1.0 / rz - y2 / r2z2 - y2 / r3z + (1.0 / Wr + pw2(S) - y2 / Wr3) * cosA);
  o[5] = (1.0 / Wr + pw2(S) - y2 / Wr3);
  o[6] = (eta / W / r + eta * pw2(C) - eta * z2 / Wr3 + y * z / r3 + 2.0 * nu * sinA * C);
  o[7] = ((1.0 / Wr + pw2(C) - z2 / Wr3) * cosA + z / r3);
  o[8] = (1.0 / Wr + pw2(C) - z2 / Wr3);
  o[9] = (x * y2 / r2z2 - nu * x / rz + x * y2 / r3z - nu * x * cosA / Wr + eta * x * S / Wr + eta * x * y / Wr3);
  o[10] = (x2 * y / r2z2 - nu * y / rz + x2 * y / r3z + nu * cosA * S + x2 * y * cosA / Wr3 + x2 * cosA * S / Wr);
  o[11] = (nu * S + x2 * S / Wr + x2 * y / Wr3);
  o[12] = (-x * y / r3 + nu * x * sinA / Wr + eta * x * C / Wr + eta * x * z / Wr3);
  o[13] = (-x2 / r3 + nu / r + nu * cosA * C + x2 * z * cosA / Wr3 + x2 * cosA * C / Wr);
  o[14] = (nu * C + x2 * C / Wr + x2 * z / Wr3);
  o[15] = (y2 / r3 - nu / r - nu * cosA * C + nu * sinA * S + eta * sinA * cosA / W2 -
           eta * (y * cosA + z * sinA) / W2r + eta * y * z / W2r2 - eta * y * z / Wr3);
  o[16] = (y / r3 + sinA * pw2(cosA) / W2 - cosA * (y * cosA + z * sinA) / W2r + y * z * cosA / W2r2 -
           y * z * cosA / Wr3);
  o[17] = (y * z / Wr3 - sinA * cosA / W2 + (y * cosA + z * sinA) / W2r - y * z / W2r2);
}

// AngDisDisp TD.jl:1177-1220.  out = {ux,uy,uz, vx,vy,vz, wx,wy,wz}
template <class T> inline void ang_dis_disp(T x, T y, T z, T cosA, T sinA, T E1, T E2, T cosA2, T sinADE1, T* o) {
  (void)E1;
  (void)cosA2;
  (void)sinADE1;
  const T eta = y * cosA - z * sinA;
  T zeta = y * sinA + z * cosA;
  const T r = sqrt(pw2(x) + pw2(y) + pw2(z));
  if (zeta > r) zeta = r;  // TD.jl:1190-1195
  if (z > r) z = r;
  const T rMzeta = r - zeta;
  const T rMz = r - z;
  o[0] = x * y / r / rMz - x * eta / r / rMzeta;
  o[3] = eta * sinA / rMzeta - y * eta / r / rMzeta + pw2(y) / r / rMz + E2 * (cosA * log(rMzeta) - log(rMz));
  o[6] = eta * cosA / rMzeta - y / r - eta * z / r / rMzeta - E2 * sinA * log(rMzeta);
  o[1] = pw2(x) * cosA / r / rMzeta - pw2(x) / r / rMz - E2 * (cosA * log(rMzeta) - log(rMz));
  o[4] = y * cosA / r / rMzeta - sinA * cosA / rMzeta - y / r / rMz;
  o[7] = z * cosA / r / rMzeta - pw2(cosA) / rMzeta + 1 / r;
  o[2] = E2 * log(rMzeta) - pw2(x) / r / rMzeta;
  o[5] = sinA / rMzeta - y / r / rMzeta;
  o[8] = cosA / rMzeta - z / r / rMzeta;
}

// AngDisDispFSC TD.jl:1412-1496.  out = {ux,uy,uz, vx,vy,vz, wx,wy,wz}; x/y/z suffix = Burgers component.
template <class T>
inline void ang_dis_disp_fsc(T y1, T y2, T y3, T cosB, T sinB, T cotB, T cotB2, T nu, T a, T* o) {
  const T y3b = y3+2.0*a;
  const T z1b = y1*cosB+y3b*sinB;
  const T z3b = -y1*sinB+y3b*cosB;
  const T r2b = pw2(y1)+pw2(y2)+pw2(y3b);
  const T rb = sqrt(r2b);
  const T c1 = (1-2*nu);
  const T c2 = (1-nu);
  const T Fib = 2*atan(-y2/(-(rb+y3b)*cotB2+y1));
  const T ux = (-2.0*c2*c1*Fib*pw2(cotB)+c1*y2/(rb+y3b)*((1.0-2.0*nu-a/rb)*cotB-y1/(rb+y3b)*(nu+a/rb))+c1*y2*cosB*cotB/(rb+z3b)*(cosB+a/rb)+a*y2*(y3b-a)*cotB/pw3(rb)+y2*(y3b-a)/(rb*(rb+y3b))*(-c1*cotB+y1/(rb+y3b)*(2.0*nu+a/rb)+a*y1/pw2(rb))+y2*(y3b-a)/(rb*(rb+z3b))*(cosB/(rb+z3b)*((rb*cosB+y3b)*(c1*cosB-a/rb)*cotB+2.0*c2*(rb*sinB-y1)*cosB)-a*y3b*cosB*cotB/pw2(rb)));
  const T vx = (c1*((2.0*c2*pw2(cotB)-nu)*log(rb+y3b)-(2.0*c2*pw2(cotB)+1.0-2.0*nu)*cosB*log(rb+z3b))-c1/(rb+y3b)*(y1*cotB*(1.0-2.0*nu-a/rb)+nu*y3b-a+pw2(y2)/(rb+y3b)*(nu+a/rb))-(1.0-2.0*nu)*z1b*cotB/(rb+z3b)*(cosB+a/rb)-a*y1*(y3b-a)*cotB/pw3(rb)+(y3b-a)/(rb+y3b)*(-2.0*nu+1.0/rb*(c1*y1*cotB-a)+pw2(y2)/(rb*(rb+y3b))*(2.0*nu+a/rb)+a*pw2(y2)/pw3(rb))+(y3b-a)/(rb+z3b)*(pw2(cosB)-1.0/rb*(c1*z1b*cotB+a*cosB)+a*y3b*z1b*cotB/pw3(rb)-1.0/(rb*(rb+z3b))*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*(rb*cosB+y3b))));
  const T wx = (2.0*c2*((c1*Fib*cotB)+(y2/(rb+y3b)*(2.0*nu+a/rb))-(y2*cosB/(rb+z3b)*(cosB+a/rb)))+y2*(y3b-a)/rb*(2.0*nu/(rb+y3b)+a/pw2(rb))+y2*(y3b-a)*cosB/(rb*(rb+z3b))*(1.0-2.0*nu-(rb*cosB+y3b)/(rb+z3b)*(cosB+a/rb)-a*y3b/pw2(rb)));
  const T uy = (c1*((2.0*c2*pw2(cotB)+nu)*log(rb+y3b)-(2.0*c2*pw2(cotB)+1.0)*cosB*log(rb+z3b))+c1/(rb+y3b)*(-c1*y1*cotB+nu*y3b-a+a*y1*cotB/rb+pw2(y1)/(rb+y3b)*(nu+a/rb))-(1.0-2.0*nu)*cotB/(rb+z3b)*(z1b*cosB-a*(rb*sinB-y1)/(rb*cosB))-a*y1*(y3b-a)*cotB/pw3(rb)+(y3b-a)/(rb+y3b)*(2.0*nu+1.0/rb*(c1*y1*cotB+a)-pw2(y1)/(rb*(rb+y3b))*(2.0*nu+a/rb)-a*pw2(y1)/pw3(rb))+(y3b-a)*cotB/(rb+z3b)*(-cosB*sinB+a*y1*y3b/(pw3(rb)*cosB)+(rb*sinB-y1)/rb*(2.0*c2*cosB-(rb*cosB+y3b)/(rb+z3b)*(1.0+a/(rb*cosB)))));
  const T vy = (2.0*c2*c1*Fib*pw2(cotB)+c1*y2/(rb+y3b)*(-(1.0-2.0*nu-a/rb)*cotB+y1/(rb+y3b)*(nu+a/rb))-c1*y2*cotB/(rb+z3b)*(1.0+a/(rb*cosB))-a*y2*(y3b-a)*cotB/pw3(rb)+y2*(y3b-a)/(rb*(rb+y3b))*(c1*cotB-2.0*nu*y1/(rb+y3b)-a*y1/rb*(1.0/rb+1.0/(rb+y3b)))+y2*(y3b-a)*cotB/(rb*(rb+z3b))*(-2.0*c2*cosB+(rb*cosB+y3b)/(rb+z3b)*(1.0+a/(rb*cosB))+a*y3b/(pw2(rb)*cosB)));
  const T wy = (-2*c2*c1*cotB*(log(rb+y3b)-cosB*log(rb+z3b))-2.0*c2*y1/(rb+y3b)*(2.0*nu+a/rb)+2.0*c2*z1b/(rb+z3b)*(cosB+a/rb)+(y3b-a)/rb*(c1*cotB-2.0*nu*y1/(rb+y3b)-a*y1/pw2(rb))-(y3b-a)/(rb+z3b)*(cosB*sinB+(rb*cosB+y3b)*cotB/rb*(2.0*c2*cosB-(rb*cosB+y3b)/(rb+z3b))+a/rb*(sinB-y3b*z1b/pw2(rb)-z1b*(rb*cosB+y3b)/(rb*(rb+z3b)))));
  const T uz = (c1*(y2/(rb+y3b)*(1.0+a/rb)-y2*cosB/(rb+z3b)*(cosB+a/rb))-y2*(y3b-a)/rb*(a/pw2(rb)+1.0/(rb+y3b))+y2*(y3b-a)*cosB/(rb*(rb+z3b))*((rb*cosB+y3b)/(rb+z3b)*(cosB+a/rb)+a*y3b/pw2(rb)));
  const T vz = (c1*(-sinB*log(rb+z3b)-y1/(rb+y3b)*(1.0+a/rb)+z1b/(rb+z3b)*(cosB+a/rb))+y1*(y3b-a)/rb*(a/pw2(rb)+1.0/(rb+y3b))-(y3b-a)/(rb+z3b)*(sinB*(cosB-a/rb)+z1b/rb*(1+a*y3b/pw2(rb))-1.0/(rb*(rb+z3b))*(pw2(y2)*cosB*sinB-a*z1b/rb*(rb*cosB+y3b))));
  const T wz = (2.0*c2*Fib+2.0*c2*(y2*sinB/(rb+z3b)*(cosB+a/rb))+y2*(y3b-a)*sinB/(rb*(rb+z3b))*(1.0+(rb*cosB+y3b)/(rb+z3b)*(cosB+a/rb)+a*y3b/pw2(rb)));
  o[0] = ux; o[1] = uy; o[2] = uz;
  o[3] = vx; o[4] = vy; o[5] = vz;
  o[6] = wx; o[7] = wy; o[8] = wz;
}

// AngDisStrainFSC TD.jl:2084-2472.  out = {v11x,v22x,v33x,v12x,v13x,v23x, v11y,...,v23y, v11z,...,v23z}
template <class T>
inline void ang_dis_strain_fsc(T y1, T y2, T y3, T cosB, T sinB, T cotB, T nu, T a, T* o) {
  const T y3b = y3+2.0*a;
  const T z1b = y1*cosB+y3b*sinB;
  const T z3b = -y1*sinB+y3b*cosB;
  const T rb2 = pw2(y1)+pw2(y2)+pw2(y3b);
  const T rb = sqrt(rb2);
  const T W1 = rb*cosB+y3b;
  const T W2 = cosB+a/rb;
  const T W3 = cosB+y3b/rb;
  const T W4 = nu+a/rb;
  const T W5 = 2.0*nu+a/rb;
  const T W6 = rb+y3b;
  const T W7 = rb+z3b;
  const T W8 = y3+a;
  const T W9 = 1.0+a/rb/cosB;
  const T N1 = 1.0-2.0*nu;
  const T rFib_ry2 = z1b/rb/(rb+z3b)-y1/rb/(rb+y3b);
  const T rFib_ry1 = y2/rb/(rb+y3b)-cosB*y2/rb/(rb+z3b);
  const T rFib_ry3 = -sinB*y2/rb/(rb+z3b);
  const T v11x = (1.0/4*((-2.0+2.0*nu)*N1*rFib_ry1*pw2(cotB)-N1*y2/pw2(W6)*((1.0-W5)*cotB-y1/W6*W4)/rb*y1+N1*y2/W6*(a/pw3(rb)*y1*cotB-1.0/W6*W4+pw2(y1)/pw2(W6)*W4/rb+pw2(y1)/W6*a/pw3(rb))-N1*y2*cosB*cotB/pw2(W7)*W2*(y1/rb-sinB)-N1*y2*cosB*cotB/W7*a/pw3(rb)*y1-3.0*a*y2*W8*cotB/pw5(rb)*y1-y2*W8/pw3(rb)/W6*(-N1*cotB+y1/W6*W5+a*y1/rb2)*y1-y2*W8/rb2/pw2(W6)*(-N1*cotB+y1/W6*W5+a*y1/rb2)*y1+y2*W8/rb/W6*(1.0/W6*W5-pw2(y1)/pw2(W6)*W5/rb-pw2(y1)/W6*a/pw3(rb)+a/rb2-2.0*a*pw2(y1)/pw2(rb2))-y2*W8/pw3(rb)/W7*(cosB/W7*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)-a*y3b*cosB*cotB/rb2)*y1-y2*W8/rb/pw2(W7)*(cosB/W7*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)-a*y3b*cosB*cotB/rb2)*(y1/rb-sinB)+y2*W8/rb/W7*(-cosB/pw2(W7)*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)*(y1/rb-sinB)+cosB/W7*(1.0/rb*cosB*y1*(N1*cosB-a/rb)*cotB+W1*a/pw3(rb)*y1*cotB+(2.0-2.0*nu)*(1.0/rb*sinB*y1-1.0)*cosB)+2.0*a*y3b*cosB*cotB/pw2(rb2)*y1))/kPi/(1.0-nu));
  const T v11y = (1.0/4*(N1*(((2.0-2.0*nu)*pw2(cotB)+nu)/rb*y1/W6-((2.0-2.0*nu)*pw2(cotB)+1.0)*cosB*(y1/rb-sinB)/W7)-N1/pw2(W6)*(-N1*y1*cotB+nu*y3b-a+a*y1*cotB/rb+pw2(y1)/W6*W4)/rb*y1+N1/W6*(-N1*cotB+a*cotB/rb-a*pw2(y1)*cotB/pw3(rb)+2.0*y1/W6*W4-pw3(y1)/pw2(W6)*W4/rb-pw3(y1)/W6*a/pw3(rb))+N1*cotB/pw2(W7)*(z1b*cosB-a*(rb*sinB-y1)/rb/cosB)*(y1/rb-sinB)-N1*cotB/W7*(pw2(cosB)-a*(1.0/rb*sinB*y1-1.0)/rb/cosB+a*(rb*sinB-y1)/pw3(rb)/cosB*y1)-a*W8*cotB/pw3(rb)+3.0*a*pw2(y1)*W8*cotB/pw5(rb)-W8/pw2(W6)*(2.0*nu+1.0/rb*(N1*y1*cotB+a)-pw2(y1)/rb/W6*W5-a*pw2(y1)/pw3(rb))/rb*y1+W8/W6*(-1.0/pw3(rb)*(N1*y1*cotB+a)*y1+1.0/rb*N1*cotB-2.0*y1/rb/W6*W5+pw3(y1)/pw3(rb)/W6*W5+pw3(y1)/rb2/pw2(W6)*W5+pw3(y1)/pw2(rb2)/W6*a-2.0*a/pw3(rb)*y1+3.0*a*pw3(y1)/pw5(rb))-W8*cotB/pw2(W7)*(-cosB*sinB+a*y1*y3b/pw3(rb)/cosB+(rb*sinB-y1)/rb*((2.0-2.0*nu)*cosB-W1/W7*W9))*(y1/rb-sinB)+W8*cotB/W7*(a*y3b/pw3(rb)/cosB-3.0*a*pw2(y1)*y3b/pw5(rb)/cosB+(1.0/rb*sinB*y1-1.0)/rb*((2.0-2.0*nu)*cosB-W1/W7*W9)-(rb*sinB-y1)/pw3(rb)*((2.0-2.0*nu)*cosB-W1/W7*W9)*y1+(rb*sinB-y1)/rb*(-1.0/rb*cosB*y1/W7*W9+W1/pw2(W7)*W9*(y1/rb-sinB)+W1/W7*a/pw3(rb)/cosB*y1)))/kPi/(1.0-nu));
  const T v11z = (1.0/4*(N1*(-y2/pw2(W6)*(1.0+a/rb)/rb*y1-y2/W6*a/pw3(rb)*y1+y2*cosB/pw2(W7)*W2*(y1/rb-sinB)+y2*cosB/W7*a/pw3(rb)*y1)+y2*W8/pw3(rb)*(a/rb2+1.0/W6)*y1-y2*W8/rb*(-2.0*a/pw2(rb2)*y1-1.0/pw2(W6)/rb*y1)-y2*W8*cosB/pw3(rb)/W7*(W1/W7*W2+a*y3b/rb2)*y1-y2*W8*cosB/rb/pw2(W7)*(W1/W7*W2+a*y3b/rb2)*(y1/rb-sinB)+y2*W8*cosB/rb/W7*(1.0/rb*cosB*y1/W7*W2-W1/pw2(W7)*W2*(y1/rb-sinB)-W1/W7*a/pw3(rb)*y1-2.0*a*y3b/pw2(rb2)*y1))/kPi/(1.0-nu));
  const T v22x = (1.0/4*(N1*(((2.0-2.0*nu)*pw2(cotB)-nu)/rb*y2/W6-((2.0-2.0*nu)*pw2(cotB)+1.0-2.0*nu)*cosB/rb*y2/W7)+N1/pw2(W6)*(y1*cotB*(1.0-W5)+nu*y3b-a+pw2(y2)/W6*W4)/rb*y2-N1/W6*(a*y1*cotB/pw3(rb)*y2+2.0*y2/W6*W4-pw3(y2)/pw2(W6)*W4/rb-pw3(y2)/W6*a/pw3(rb))+N1*z1b*cotB/pw2(W7)*W2/rb*y2+N1*z1b*cotB/W7*a/pw3(rb)*y2+3.0*a*y2*W8*cotB/pw5(rb)*y1-W8/pw2(W6)*(-2.0*nu+1.0/rb*(N1*y1*cotB-a)+pw2(y2)/rb/W6*W5+a*pw2(y2)/pw3(rb))/rb*y2+W8/W6*(-1.0/pw3(rb)*(N1*y1*cotB-a)*y2+2.0*y2/rb/W6*W5-pw3(y2)/pw3(rb)/W6*W5-pw3(y2)/rb2/pw2(W6)*W5-pw3(y2)/pw2(rb2)/W6*a+2.0*a/pw3(rb)*y2-3.0*a*pw3(y2)/pw5(rb))-W8/pw2(W7)*(pw2(cosB)-1.0/rb*(N1*z1b*cotB+a*cosB)+a*y3b*z1b*cotB/pw3(rb)-1.0/rb/W7*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1))/rb*y2+W8/W7*(1.0/pw3(rb)*(N1*z1b*cotB+a*cosB)*y2-3.0*a*y3b*z1b*cotB/pw5(rb)*y2+1.0/pw3(rb)/W7*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1)*y2+1.0/rb2/pw2(W7)*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1)*y2-1.0/rb/W7*(2.0*y2*pw2(cosB)+a*z1b*cotB/pw3(rb)*W1*y2-a*z1b*cotB/rb2*cosB*y2)))/kPi/(1.0-nu));
  const T v22y = (1.0/4*((2.0-2.0*nu)*N1*rFib_ry2*pw2(cotB)+N1/W6*((W5-1.0)*cotB+y1/W6*W4)-N1*pw2(y2)/pw2(W6)*((W5-1.0)*cotB+y1/W6*W4)/rb+N1*y2/W6*(-a/pw3(rb)*y2*cotB-y1/pw2(W6)*W4/rb*y2-y2/W6*a/pw3(rb)*y1)-N1*cotB/W7*W9+N1*pw2(y2)*cotB/pw2(W7)*W9/rb+N1*pw2(y2)*cotB/W7*a/pw3(rb)/cosB-a*W8*cotB/pw3(rb)+3.0*a*pw2(y2)*W8*cotB/pw5(rb)+W8/rb/W6*(N1*cotB-2.0*nu*y1/W6-a*y1/rb*(1.0/rb+1.0/W6))-pw2(y2)*W8/pw3(rb)/W6*(N1*cotB-2.0*nu*y1/W6-a*y1/rb*(1.0/rb+1.0/W6))-pw2(y2)*W8/rb2/pw2(W6)*(N1*cotB-2.0*nu*y1/W6-a*y1/rb*(1.0/rb+1.0/W6))+y2*W8/rb/W6*(2.0*nu*y1/pw2(W6)/rb*y2+a*y1/pw3(rb)*(1.0/rb+1.0/W6)*y2-a*y1/rb*(-1.0/pw3(rb)*y2-1.0/pw2(W6)/rb*y2))+W8*cotB/rb/W7*((-2.0+2.0*nu)*cosB+W1/W7*W9+a*y3b/rb2/cosB)-pw2(y2)*W8*cotB/pw3(rb)/W7*((-2.0+2.0*nu)*cosB+W1/W7*W9+a*y3b/rb2/cosB)-pw2(y2)*W8*cotB/rb2/pw2(W7)*((-2.0+2.0*nu)*cosB+W1/W7*W9+a*y3b/rb2/cosB)+y2*W8*cotB/rb/W7*(1.0/rb*cosB*y2/W7*W9-W1/pw2(W7)*W9/rb*y2-W1/W7*a/pw3(rb)/cosB*y2-2.0*a*y3b/pw2(rb2)/cosB*y2))/kPi/(1.0-nu));
  const T v22z = (1.0/4*(N1*(-sinB/rb*y2/W7+y2/pw2(W6)*(1.0+a/rb)/rb*y1+y2/W6*a/pw3(rb)*y1-z1b/pw2(W7)*W2/rb*y2-z1b/W7*a/pw3(rb)*y2)-y2*W8/pw3(rb)*(a/rb2+1.0/W6)*y1+y1*W8/rb*(-2.0*a/pw2(rb2)*y2-1.0/pw2(W6)/rb*y2)+W8/pw2(W7)*(sinB*(cosB-a/rb)+z1b/rb*(1.0+a*y3b/rb2)-1.0/rb/W7*(pw2(y2)*cosB*sinB-a*z1b/rb*W1))/rb*y2-W8/W7*(sinB*a/pw3(rb)*y2-z1b/pw3(rb)*(1.0+a*y3b/rb2)*y2-2.0*z1b/pw5(rb)*a*y3b*y2+1.0/pw3(rb)/W7*(pw2(y2)*cosB*sinB-a*z1b/rb*W1)*y2+1.0/rb2/pw2(W7)*(pw2(y2)*cosB*sinB-a*z1b/rb*W1)*y2-1.0/rb/W7*(2.0*y2*cosB*sinB+a*z1b/pw3(rb)*W1*y2-a*z1b/rb2*cosB*y2)))/kPi/(1.0-nu));
  const T v33x = (1.0/4*((2.0-2.0*nu)*(N1*rFib_ry3*cotB-y2/pw2(W6)*W5*(y3b/rb+1.0)-1.0/2.0*y2/W6*a/pw3(rb)*2.0*y3b+y2*cosB/pw2(W7)*W2*W3+1.0/2.0*y2*cosB/W7*a/pw3(rb)*2.0*y3b)+y2/rb*(2.0*nu/W6+a/rb2)-1.0/2.0*y2*W8/pw3(rb)*(2.0*nu/W6+a/rb2)*2.0*y3b+y2*W8/rb*(-2.0*nu/pw2(W6)*(y3b/rb+1.0)-a/pw2(rb2)*2.0*y3b)+y2*cosB/rb/W7*(1.0-2.0*nu-W1/W7*W2-a*y3b/rb2)-1.0/2.0*y2*W8*cosB/pw3(rb)/W7*(1.0-2.0*nu-W1/W7*W2-a*y3b/rb2)*2.0*y3b-y2*W8*cosB/rb/pw2(W7)*(1.0-2.0*nu-W1/W7*W2-a*y3b/rb2)*W3+y2*W8*cosB/rb/W7*(-(cosB*y3b/rb+1.0)/W7*W2+W1/pw2(W7)*W2*W3+1.0/2.0*W1/W7*a/pw3(rb)*2.0*y3b-a/rb2+a*y3b/pw2(rb2)*2.0*y3b))/kPi/(1.0-nu));
  const T v33y = (1.0/4*((-2.0+2.0*nu)*N1*cotB*((y3b/rb+1.0)/W6-cosB*W3/W7)+(2.0-2.0*nu)*y1/pw2(W6)*W5*(y3b/rb+1.0)+1.0/2.0*(2.0-2.0*nu)*y1/W6*a/pw3(rb)*2.0*y3b+(2.0-2.0*nu)*sinB/W7*W2-(2.0-2.0*nu)*z1b/pw2(W7)*W2*W3-1.0/2.0*(2.0-2.0*nu)*z1b/W7*a/pw3(rb)*2.0*y3b+1.0/rb*(N1*cotB-2.0*nu*y1/W6-a*y1/rb2)-1.0/2.0*W8/pw3(rb)*(N1*cotB-2.0*nu*y1/W6-a*y1/rb2)*2.0*y3b+W8/rb*(2.0*nu*y1/pw2(W6)*(y3b/rb+1.0)+a*y1/pw2(rb2)*2.0*y3b)-1.0/W7*(cosB*sinB+W1*cotB/rb*((2.0-2.0*nu)*cosB-W1/W7)+a/rb*(sinB-y3b*z1b/rb2-z1b*W1/rb/W7))+W8/pw2(W7)*(cosB*sinB+W1*cotB/rb*((2.0-2.0*nu)*cosB-W1/W7)+a/rb*(sinB-y3b*z1b/rb2-z1b*W1/rb/W7))*W3-W8/W7*((cosB*y3b/rb+1.0)*cotB/rb*((2.0-2.0*nu)*cosB-W1/W7)-1.0/2.0*W1*cotB/pw3(rb)*((2.0-2.0*nu)*cosB-W1/W7)*2.0*y3b+W1*cotB/rb*(-(cosB*y3b/rb+1.0)/W7+W1/pw2(W7)*W3)-1.0/2.0*a/pw3(rb)*(sinB-y3b*z1b/rb2-z1b*W1/rb/W7)*2.0*y3b+a/rb*(-z1b/rb2-y3b*sinB/rb2+y3b*z1b/pw2(rb2)*2.0*y3b-sinB*W1/rb/W7-z1b*(cosB*y3b/rb+1.0)/rb/W7+1.0/2.0*z1b*W1/pw3(rb)/W7*2.0*y3b+z1b*W1/rb/pw2(W7)*W3)))/kPi/(1.0-nu));
  const T v33z = (1.0/4*((2.0-2.0*nu)*rFib_ry3-(2.0-2.0*nu)*y2*sinB/pw2(W7)*W2*W3-1.0/2.0*(2.0-2.0*nu)*y2*sinB/W7*a/pw3(rb)*2.0*y3b+y2*sinB/rb/W7*(1.0+W1/W7*W2+a*y3b/rb2)-1.0/2.0*y2*W8*sinB/pw3(rb)/W7*(1.0+W1/W7*W2+a*y3b/rb2)*2.0*y3b-y2*W8*sinB/rb/pw2(W7)*(1.0+W1/W7*W2+a*y3b/rb2)*W3+y2*W8*sinB/rb/W7*((cosB*y3b/rb+1.0)/W7*W2-W1/pw2(W7)*W2*W3-1.0/2.0*W1/W7*a/pw3(rb)*2.0*y3b+a/rb2-a*y3b/pw2(rb2)*2.0*y3b))/kPi/(1.0-nu));
  const T v12x = (1.0/4*((-2.0+2.0*nu)*N1*rFib_ry2*pw2(cotB)+N1/W6*((1.0-W5)*cotB-y1/W6*W4)-N1*pw2(y2)/pw2(W6)*((1.0-W5)*cotB-y1/W6*W4)/rb+N1*y2/W6*(a/pw3(rb)*y2*cotB+y1/pw2(W6)*W4/rb*y2+y2/W6*a/pw3(rb)*y1)+N1*cosB*cotB/W7*W2-N1*pw2(y2)*cosB*cotB/pw2(W7)*W2/rb-N1*pw2(y2)*cosB*cotB/W7*a/pw3(rb)+a*W8*cotB/pw3(rb)-3.0*a*pw2(y2)*W8*cotB/pw5(rb)+W8/rb/W6*(-N1*cotB+y1/W6*W5+a*y1/rb2)-pw2(y2)*W8/pw3(rb)/W6*(-N1*cotB+y1/W6*W5+a*y1/rb2)-pw2(y2)*W8/rb2/pw2(W6)*(-N1*cotB+y1/W6*W5+a*y1/rb2)+y2*W8/rb/W6*(-y1/pw2(W6)*W5/rb*y2-y2/W6*a/pw3(rb)*y1-2.0*a*y1/pw2(rb2)*y2)+W8/rb/W7*(cosB/W7*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)-a*y3b*cosB*cotB/rb2)-pw2(y2)*W8/pw3(rb)/W7*(cosB/W7*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)-a*y3b*cosB*cotB/rb2)-pw2(y2)*W8/rb2/pw2(W7)*(cosB/W7*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)-a*y3b*cosB*cotB/rb2)+y2*W8/rb/W7*(-cosB/pw2(W7)*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)/rb*y2+cosB/W7*(1.0/rb*cosB*y2*(N1*cosB-a/rb)*cotB+W1*a/pw3(rb)*y2*cotB+(2.0-2.0*nu)/rb*sinB*y2*cosB)+2.0*a*y3b*cosB*cotB/pw2(rb2)*y2))/kPi/(1.0-nu))+(1.0/4*(N1*(((2.0-2.0*nu)*pw2(cotB)-nu)/rb*y1/W6-((2.0-2.0*nu)*pw2(cotB)+1.0-2.0*nu)*cosB*(y1/rb-sinB)/W7)+N1/pw2(W6)*(y1*cotB*(1.0-W5)+nu*y3b-a+pw2(y2)/W6*W4)/rb*y1-N1/W6*((1.0-W5)*cotB+a*pw2(y1)*cotB/pw3(rb)-pw2(y2)/pw2(W6)*W4/rb*y1-pw2(y2)/W6*a/pw3(rb)*y1)-N1*cosB*cotB/W7*W2+N1*z1b*cotB/pw2(W7)*W2*(y1/rb-sinB)+N1*z1b*cotB/W7*a/pw3(rb)*y1-a*W8*cotB/pw3(rb)+3.0*a*pw2(y1)*W8*cotB/pw5(rb)-W8/pw2(W6)*(-2.0*nu+1.0/rb*(N1*y1*cotB-a)+pw2(y2)/rb/W6*W5+a*pw2(y2)/pw3(rb))/rb*y1+W8/W6*(-1.0/pw3(rb)*(N1*y1*cotB-a)*y1+1.0/rb*N1*cotB-pw2(y2)/pw3(rb)/W6*W5*y1-pw2(y2)/rb2/pw2(W6)*W5*y1-pw2(y2)/pw2(rb2)/W6*a*y1-3.0*a*pw2(y2)/pw5(rb)*y1)-W8/pw2(W7)*(pw2(cosB)-1.0/rb*(N1*z1b*cotB+a*cosB)+a*y3b*z1b*cotB/pw3(rb)-1.0/rb/W7*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1))*(y1/rb-sinB)+W8/W7*(1.0/pw3(rb)*(N1*z1b*cotB+a*cosB)*y1-1.0/rb*N1*cosB*cotB+a*y3b*cosB*cotB/pw3(rb)-3.0*a*y3b*z1b*cotB/pw5(rb)*y1+1.0/pw3(rb)/W7*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1)*y1+1.0/rb/pw2(W7)*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1)*(y1/rb-sinB)-1.0/rb/W7*(-a*cosB*cotB/rb*W1+a*z1b*cotB/pw3(rb)*W1*y1-a*z1b*cotB/rb2*cosB*y1)))/kPi/(1.0-nu));
  const T v12y = (1.0/4*(N1*(((2.0-2.0*nu)*pw2(cotB)+nu)/rb*y2/W6-((2.0-2.0*nu)*pw2(cotB)+1.0)*cosB/rb*y2/W7)-N1/pw2(W6)*(-N1*y1*cotB+nu*y3b-a+a*y1*cotB/rb+pw2(y1)/W6*W4)/rb*y2+N1/W6*(-a*y1*cotB/pw3(rb)*y2-pw2(y1)/pw2(W6)*W4/rb*y2-pw2(y1)/W6*a/pw3(rb)*y2)+N1*cotB/pw2(W7)*(z1b*cosB-a*(rb*sinB-y1)/rb/cosB)/rb*y2-N1*cotB/W7*(-a/rb2*sinB*y2/cosB+a*(rb*sinB-y1)/pw3(rb)/cosB*y2)+3.0*a*y2*W8*cotB/pw5(rb)*y1-W8/pw2(W6)*(2.0*nu+1.0/rb*(N1*y1*cotB+a)-pw2(y1)/rb/W6*W5-a*pw2(y1)/pw3(rb))/rb*y2+W8/W6*(-1.0/pw3(rb)*(N1*y1*cotB+a)*y2+pw2(y1)/pw3(rb)/W6*W5*y2+pw2(y1)/rb2/pw2(W6)*W5*y2+pw2(y1)/pw2(rb2)/W6*a*y2+3.0*a*pw2(y1)/pw5(rb)*y2)-W8*cotB/pw2(W7)*(-cosB*sinB+a*y1*y3b/pw3(rb)/cosB+(rb*sinB-y1)/rb*((2.0-2.0*nu)*cosB-W1/W7*W9))/rb*y2+W8*cotB/W7*(-3.0*a*y1*y3b/pw5(rb)/cosB*y2+1.0/rb2*sinB*y2*((2.0-2.0*nu)*cosB-W1/W7*W9)-(rb*sinB-y1)/pw3(rb)*((2.0-2.0*nu)*cosB-W1/W7*W9)*y2+(rb*sinB-y1)/rb*(-1.0/rb*cosB*y2/W7*W9+W1/pw2(W7)*W9/rb*y2+W1/W7*a/pw3(rb)/cosB*y2)))/kPi/(1.0-nu))+(1.0/4*((2.0-2.0*nu)*N1*rFib_ry1*pw2(cotB)-N1*y2/pw2(W6)*((W5-1.0)*cotB+y1/W6*W4)/rb*y1+N1*y2/W6*(-a/pw3(rb)*y1*cotB+1.0/W6*W4-pw2(y1)/pw2(W6)*W4/rb-pw2(y1)/W6*a/pw3(rb))+N1*y2*cotB/pw2(W7)*W9*(y1/rb-sinB)+N1*y2*cotB/W7*a/pw3(rb)/cosB*y1+3.0*a*y2*W8*cotB/pw5(rb)*y1-y2*W8/pw3(rb)/W6*(N1*cotB-2.0*nu*y1/W6-a*y1/rb*(1.0/rb+1.0/W6))*y1-y2*W8/rb2/pw2(W6)*(N1*cotB-2.0*nu*y1/W6-a*y1/rb*(1.0/rb+1.0/W6))*y1+y2*W8/rb/W6*(-2.0*nu/W6+2.0*nu*pw2(y1)/pw2(W6)/rb-a/rb*(1.0/rb+1.0/W6)+a*pw2(y1)/pw3(rb)*(1.0/rb+1.0/W6)-a*y1/rb*(-1.0/pw3(rb)*y1-1.0/pw2(W6)/rb*y1))-y2*W8*cotB/pw3(rb)/W7*((-2.0+2.0*nu)*cosB+W1/W7*W9+a*y3b/rb2/cosB)*y1-y2*W8*cotB/rb/pw2(W7)*((-2.0+2.0*nu)*cosB+W1/W7*W9+a*y3b/rb2/cosB)*(y1/rb-sinB)+y2*W8*cotB/rb/W7*(1.0/rb*cosB*y1/W7*W9-W1/pw2(W7)*W9*(y1/rb-sinB)-W1/W7*a/pw3(rb)/cosB*y1-2.0*a*y3b/pw2(rb2)/cosB*y1))/kPi/(1.0-nu));
  const T v12z = (1.0/4*(N1*(1.0/W6*(1.0+a/rb)-pw2(y2)/pw2(W6)*(1.0+a/rb)/rb-pw2(y2)/W6*a/pw3(rb)-cosB/W7*W2+pw2(y2)*cosB/pw2(W7)*W2/rb+pw2(y2)*cosB/W7*a/pw3(rb))-W8/rb*(a/rb2+1.0/W6)+pw2(y2)*W8/pw3(rb)*(a/rb2+1.0/W6)-y2*W8/rb*(-2.0*a/pw2(rb2)*y2-1.0/pw2(W6)/rb*y2)+W8*cosB/rb/W7*(W1/W7*W2+a*y3b/rb2)-pw2(y2)*W8*cosB/pw3(rb)/W7*(W1/W7*W2+a*y3b/rb2)-pw2(y2)*W8*cosB/rb2/pw2(W7)*(W1/W7*W2+a*y3b/rb2)+y2*W8*cosB/rb/W7*(1.0/rb*cosB*y2/W7*W2-W1/pw2(W7)*W2/rb*y2-W1/W7*a/pw3(rb)*y2-2.0*a*y3b/pw2(rb2)*y2))/kPi/(1.0-nu))+(1.0/4*(N1*(-sinB*(y1/rb-sinB)/W7-1.0/W6*(1.0+a/rb)+pw2(y1)/pw2(W6)*(1.0+a/rb)/rb+pw2(y1)/W6*a/pw3(rb)+cosB/W7*W2-z1b/pw2(W7)*W2*(y1/rb-sinB)-z1b/W7*a/pw3(rb)*y1)+W8/rb*(a/rb2+1.0/W6)-pw2(y1)*W8/pw3(rb)*(a/rb2+1.0/W6)+y1*W8/rb*(-2.0*a/pw2(rb2)*y1-1.0/pw2(W6)/rb*y1)+W8/pw2(W7)*(sinB*(cosB-a/rb)+z1b/rb*(1.0+a*y3b/rb2)-1.0/rb/W7*(pw2(y2)*cosB*sinB-a*z1b/rb*W1))*(y1/rb-sinB)-W8/W7*(sinB*a/pw3(rb)*y1+cosB/rb*(1.0+a*y3b/rb2)-z1b/pw3(rb)*(1.0+a*y3b/rb2)*y1-2.0*z1b/pw5(rb)*a*y3b*y1+1.0/pw3(rb)/W7*(pw2(y2)*cosB*sinB-a*z1b/rb*W1)*y1+1.0/rb/pw2(W7)*(pw2(y2)*cosB*sinB-a*z1b/rb*W1)*(y1/rb-sinB)-1.0/rb/W7*(-a*cosB/rb*W1+a*z1b/pw3(rb)*W1*y1-a*z1b/rb2*cosB*y1)))/kPi/(1.0-nu));
  const T v13x = (1.0/4*((-2.0+2.0*nu)*N1*rFib_ry3*pw2(cotB)-N1*y2/pw2(W6)*((1.0-W5)*cotB-y1/W6*W4)*(y3b/rb+1.0)+N1*y2/W6*(1.0/2.0*a/pw3(rb)*2.0*y3b*cotB+y1/pw2(W6)*W4*(y3b/rb+1.0)+1.0/2.0*y1/W6*a/pw3(rb)*2.0*y3b)-N1*y2*cosB*cotB/pw2(W7)*W2*W3-1.0/2.0*N1*y2*cosB*cotB/W7*a/pw3(rb)*2.0*y3b+a/pw3(rb)*y2*cotB-3.0/2.0*a*y2*W8*cotB/pw5(rb)*2.0*y3b+y2/rb/W6*(-N1*cotB+y1/W6*W5+a*y1/rb2)-1.0/2.0*y2*W8/pw3(rb)/W6*(-N1*cotB+y1/W6*W5+a*y1/rb2)*2.0*y3b-y2*W8/rb/pw2(W6)*(-N1*cotB+y1/W6*W5+a*y1/rb2)*(y3b/rb+1.0)+y2*W8/rb/W6*(-y1/pw2(W6)*W5*(y3b/rb+1.0)-1.0/2.0*y1/W6*a/pw3(rb)*2.0*y3b-a*y1/pw2(rb2)*2.0*y3b)+y2/rb/W7*(cosB/W7*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)-a*y3b*cosB*cotB/rb2)-1.0/2.0*y2*W8/pw3(rb)/W7*(cosB/W7*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)-a*y3b*cosB*cotB/rb2)*2.0*y3b-y2*W8/rb/pw2(W7)*(cosB/W7*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)-a*y3b*cosB*cotB/rb2)*W3+y2*W8/rb/W7*(-cosB/pw2(W7)*(W1*(N1*cosB-a/rb)*cotB+(2.0-2.0*nu)*(rb*sinB-y1)*cosB)*W3+cosB/W7*((cosB*y3b/rb+1.0)*(N1*cosB-a/rb)*cotB+1.0/2.0*W1*a/pw3(rb)*2.0*y3b*cotB+1.0/2.0*(2.0-2.0*nu)/rb*sinB*2.0*y3b*cosB)-a*cosB*cotB/rb2+a*y3b*cosB*cotB/pw2(rb2)*2.0*y3b))/kPi/(1.0-nu))+(1.0/4*((2.0-2.0*nu)*(N1*rFib_ry1*cotB-y1/pw2(W6)*W5/rb*y2-y2/W6*a/pw3(rb)*y1+y2*cosB/pw2(W7)*W2*(y1/rb-sinB)+y2*cosB/W7*a/pw3(rb)*y1)-y2*W8/pw3(rb)*(2.0*nu/W6+a/rb2)*y1+y2*W8/rb*(-2.0*nu/pw2(W6)/rb*y1-2.0*a/pw2(rb2)*y1)-y2*W8*cosB/pw3(rb)/W7*(1.0-2.0*nu-W1/W7*W2-a*y3b/rb2)*y1-y2*W8*cosB/rb/pw2(W7)*(1.0-2.0*nu-W1/W7*W2-a*y3b/rb2)*(y1/rb-sinB)+y2*W8*cosB/rb/W7*(-1.0/rb*cosB*y1/W7*W2+W1/pw2(W7)*W2*(y1/rb-sinB)+W1/W7*a/pw3(rb)*y1+2.0*a*y3b/pw2(rb2)*y1))/kPi/(1.0-nu));
  const T v13y = (1.0/4*(N1*(((2.0-2.0*nu)*pw2(cotB)+nu)*(y3b/rb+1.0)/W6-((2.0-2.0*nu)*pw2(cotB)+1.0)*cosB*W3/W7)-N1/pw2(W6)*(-N1*y1*cotB+nu*y3b-a+a*y1*cotB/rb+pw2(y1)/W6*W4)*(y3b/rb+1.0)+N1/W6*(nu-1.0/2.0*a*y1*cotB/pw3(rb)*2.0*y3b-pw2(y1)/pw2(W6)*W4*(y3b/rb+1.0)-1.0/2.0*pw2(y1)/W6*a/pw3(rb)*2.0*y3b)+N1*cotB/pw2(W7)*(z1b*cosB-a*(rb*sinB-y1)/rb/cosB)*W3-N1*cotB/W7*(cosB*sinB-1.0/2.0*a/rb2*sinB*2.0*y3b/cosB+1.0/2.0*a*(rb*sinB-y1)/pw3(rb)/cosB*2.0*y3b)-a/pw3(rb)*y1*cotB+3.0/2.0*a*y1*W8*cotB/pw5(rb)*2.0*y3b+1.0/W6*(2.0*nu+1.0/rb*(N1*y1*cotB+a)-pw2(y1)/rb/W6*W5-a*pw2(y1)/pw3(rb))-W8/pw2(W6)*(2.0*nu+1.0/rb*(N1*y1*cotB+a)-pw2(y1)/rb/W6*W5-a*pw2(y1)/pw3(rb))*(y3b/rb+1.0)+W8/W6*(-1.0/2.0/pw3(rb)*(N1*y1*cotB+a)*2.0*y3b+1.0/2.0*pw2(y1)/pw3(rb)/W6*W5*2.0*y3b+pw2(y1)/rb/pw2(W6)*W5*(y3b/rb+1.0)+1.0/2.0*pw2(y1)/pw2(rb2)/W6*a*2.0*y3b+3.0/2.0*a*pw2(y1)/pw5(rb)*2.0*y3b)+cotB/W7*(-cosB*sinB+a*y1*y3b/pw3(rb)/cosB+(rb*sinB-y1)/rb*((2.0-2.0*nu)*cosB-W1/W7*W9))-W8*cotB/pw2(W7)*(-cosB*sinB+a*y1*y3b/pw3(rb)/cosB+(rb*sinB-y1)/rb*((2.0-2.0*nu)*cosB-W1/W7*W9))*W3+W8*cotB/W7*(a/pw3(rb)/cosB*y1-3.0/2.0*a*y1*y3b/pw5(rb)/cosB*2.0*y3b+1.0/2.0/rb2*sinB*2.0*y3b*((2.0-2.0*nu)*cosB-W1/W7*W9)-1.0/2.0*(rb*sinB-y1)/pw3(rb)*((2.0-2.0*nu)*cosB-W1/W7*W9)*2.0*y3b+(rb*sinB-y1)/rb*(-(cosB*y3b/rb+1.0)/W7*W9+W1/pw2(W7)*W9*W3+1.0/2.0*W1/W7*a/pw3(rb)/cosB*2.0*y3b)))/kPi/(1.0-nu))+(1.0/4*((-2.0+2.0*nu)*N1*cotB*(1.0/rb*y1/W6-cosB*(y1/rb-sinB)/W7)-(2.0-2.0*nu)/W6*W5+(2.0-2.0*nu)*pw2(y1)/pw2(W6)*W5/rb+(2.0-2.0*nu)*pw2(y1)/W6*a/pw3(rb)+(2.0-2.0*nu)*cosB/W7*W2-(2.0-2.0*nu)*z1b/pw2(W7)*W2*(y1/rb-sinB)-(2.0-2.0*nu)*z1b/W7*a/pw3(rb)*y1-W8/pw3(rb)*(N1*cotB-2.0*nu*y1/W6-a*y1/rb2)*y1+W8/rb*(-2.0*nu/W6+2.0*nu*pw2(y1)/pw2(W6)/rb-a/rb2+2.0*a*pw2(y1)/pw2(rb2))+W8/pw2(W7)*(cosB*sinB+W1*cotB/rb*((2.0-2.0*nu)*cosB-W1/W7)+a/rb*(sinB-y3b*z1b/rb2-z1b*W1/rb/W7))*(y1/rb-sinB)-W8/W7*(1.0/rb2*cosB*y1*cotB*((2.0-2.0*nu)*cosB-W1/W7)-W1*cotB/pw3(rb)*((2.0-2.0*nu)*cosB-W1/W7)*y1+W1*cotB/rb*(-1.0/rb*cosB*y1/W7+W1/pw2(W7)*(y1/rb-sinB))-a/pw3(rb)*(sinB-y3b*z1b/rb2-z1b*W1/rb/W7)*y1+a/rb*(-y3b*cosB/rb2+2.0*y3b*z1b/pw2(rb2)*y1-cosB*W1/rb/W7-z1b/rb2*cosB*y1/W7+z1b*W1/pw3(rb)/W7*y1+z1b*W1/rb/pw2(W7)*(y1/rb-sinB))))/kPi/(1.0-nu));
  const T v13z = (1.0/4*(N1*(-y2/pw2(W6)*(1.0+a/rb)*(y3b/rb+1.0)-1.0/2.0*y2/W6*a/pw3(rb)*2.0*y3b+y2*cosB/pw2(W7)*W2*W3+1.0/2.0*y2*cosB/W7*a/pw3(rb)*2.0*y3b)-y2/rb*(a/rb2+1.0/W6)+1.0/2.0*y2*W8/pw3(rb)*(a/rb2+1.0/W6)*2.0*y3b-y2*W8/rb*(-a/pw2(rb2)*2.0*y3b-1.0/pw2(W6)*(y3b/rb+1.0))+y2*cosB/rb/W7*(W1/W7*W2+a*y3b/rb2)-1.0/2.0*y2*W8*cosB/pw3(rb)/W7*(W1/W7*W2+a*y3b/rb2)*2.0*y3b-y2*W8*cosB/rb/pw2(W7)*(W1/W7*W2+a*y3b/rb2)*W3+y2*W8*cosB/rb/W7*((cosB*y3b/rb+1.0)/W7*W2-W1/pw2(W7)*W2*W3-1.0/2.0*W1/W7*a/pw3(rb)*2.0*y3b+a/rb2-a*y3b/pw2(rb2)*2.0*y3b))/kPi/(1.0-nu))+(1.0/4*((2.0-2.0*nu)*rFib_ry1-(2.0-2.0*nu)*y2*sinB/pw2(W7)*W2*(y1/rb-sinB)-(2.0-2.0*nu)*y2*sinB/W7*a/pw3(rb)*y1-y2*W8*sinB/pw3(rb)/W7*(1.0+W1/W7*W2+a*y3b/rb2)*y1-y2*W8*sinB/rb/pw2(W7)*(1.0+W1/W7*W2+a*y3b/rb2)*(y1/rb-sinB)+y2*W8*sinB/rb/W7*(1.0/rb*cosB*y1/W7*W2-W1/pw2(W7)*W2*(y1/rb-sinB)-W1/W7*a/pw3(rb)*y1-2.0*a*y3b/pw2(rb2)*y1))/kPi/(1.0-nu));
  const T v23x = (1.0/4*(N1*(((2.0-2.0*nu)*pw2(cotB)-nu)*(y3b/rb+1.0)/W6-((2.0-2.0*nu)*pw2(cotB)+1.0-2.0*nu)*cosB*W3/W7)+N1/pw2(W6)*(y1*cotB*(1.0-W5)+nu*y3b-a+pw2(y2)/W6*W4)*(y3b/rb+1.0)-N1/W6*(1.0/2.0*a*y1*cotB/pw3(rb)*2.0*y3b+nu-pw2(y2)/pw2(W6)*W4*(y3b/rb+1.0)-1.0/2.0*pw2(y2)/W6*a/pw3(rb)*2.0*y3b)-N1*sinB*cotB/W7*W2+N1*z1b*cotB/pw2(W7)*W2*W3+1.0/2.0*N1*z1b*cotB/W7*a/pw3(rb)*2.0*y3b-a/pw3(rb)*y1*cotB+3.0/2.0*a*y1*W8*cotB/pw5(rb)*2.0*y3b+1.0/W6*(-2.0*nu+1.0/rb*(N1*y1*cotB-a)+pw2(y2)/rb/W6*W5+a*pw2(y2)/pw3(rb))-W8/pw2(W6)*(-2.0*nu+1.0/rb*(N1*y1*cotB-a)+pw2(y2)/rb/W6*W5+a*pw2(y2)/pw3(rb))*(y3b/rb+1.0)+W8/W6*(-1.0/2.0/pw3(rb)*(N1*y1*cotB-a)*2.0*y3b-1.0/2.0*pw2(y2)/pw3(rb)/W6*W5*2.0*y3b-pw2(y2)/rb/pw2(W6)*W5*(y3b/rb+1.0)-1.0/2.0*pw2(y2)/pw2(rb2)/W6*a*2.0*y3b-3.0/2.0*a*pw2(y2)/pw5(rb)*2.0*y3b)+1.0/W7*(pw2(cosB)-1.0/rb*(N1*z1b*cotB+a*cosB)+a*y3b*z1b*cotB/pw3(rb)-1.0/rb/W7*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1))-W8/pw2(W7)*(pw2(cosB)-1.0/rb*(N1*z1b*cotB+a*cosB)+a*y3b*z1b*cotB/pw3(rb)-1.0/rb/W7*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1))*W3+W8/W7*(1.0/2.0/pw3(rb)*(N1*z1b*cotB+a*cosB)*2.0*y3b-1.0/rb*N1*sinB*cotB+a*z1b*cotB/pw3(rb)+a*y3b*sinB*cotB/pw3(rb)-3.0/2.0*a*y3b*z1b*cotB/pw5(rb)*2.0*y3b+1.0/2.0/pw3(rb)/W7*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1)*2.0*y3b+1.0/rb/pw2(W7)*(pw2(y2)*pw2(cosB)-a*z1b*cotB/rb*W1)*W3-1.0/rb/W7*(-a*sinB*cotB/rb*W1+1.0/2.0*a*z1b*cotB/pw3(rb)*W1*2.0*y3b-a*z1b*cotB/rb*(cosB*y3b/rb+1.0))))/kPi/(1.0-nu))+(1.0/4*((2.0-2.0*nu)*(N1*rFib_ry2*cotB+1.0/W6*W5-pw2(y2)/pw2(W6)*W5/rb-pw2(y2)/W6*a/pw3(rb)-cosB/W7*W2+pw2(y2)*cosB/pw2(W7)*W2/rb+pw2(y2)*cosB/W7*a/pw3(rb))+W8/rb*(2.0*nu/W6+a/rb2)-pw2(y2)*W8/pw3(rb)*(2.0*nu/W6+a/rb2)+y2*W8/rb*(-2.0*nu/pw2(W6)/rb*y2-2.0*a/pw2(rb2)*y2)+W8*cosB/rb/W7*(1.0-2.0*nu-W1/W7*W2-a*y3b/rb2)-pw2(y2)*W8*cosB/pw3(rb)/W7*(1.0-2.0*nu-W1/W7*W2-a*y3b/rb2)-pw2(y2)*W8*cosB/rb2/pw2(W7)*(1.0-2.0*nu-W1/W7*W2-a*y3b/rb2)+y2*W8*cosB/rb/W7*(-1.0/rb*cosB*y2/W7*W2+W1/pw2(W7)*W2/rb*y2+W1/W7*a/pw3(rb)*y2+2.0*a*y3b/pw2(rb2)*y2))/kPi/(1.0-nu));
  const T v23y = (1.0/4*((2.0-2.0*nu)*N1*rFib_ry3*pw2(cotB)-N1*y2/pw2(W6)*((W5-1.0)*cotB+y1/W6*W4)*(y3b/rb+1.0)+N1*y2/W6*(-1.0/2.0*a/pw3(rb)*2.0*y3b*cotB-y1/pw2(W6)*W4*(y3b/rb+1.0)-1.0/2.0*y1/W6*a/pw3(rb)*2.0*y3b)+N1*y2*cotB/pw2(W7)*W9*W3+1.0/2.0*N1*y2*cotB/W7*a/pw3(rb)/cosB*2.0*y3b-a/pw3(rb)*y2*cotB+3.0/2.0*a*y2*W8*cotB/pw5(rb)*2.0*y3b+y2/rb/W6*(N1*cotB-2.0*nu*y1/W6-a*y1/rb*(1.0/rb+1.0/W6))-1.0/2.0*y2*W8/pw3(rb)/W6*(N1*cotB-2.0*nu*y1/W6-a*y1/rb*(1.0/rb+1.0/W6))*2.0*y3b-y2*W8/rb/pw2(W6)*(N1*cotB-2.0*nu*y1/W6-a*y1/rb*(1.0/rb+1.0/W6))*(y3b/rb+1.0)+y2*W8/rb/W6*(2.0*nu*y1/pw2(W6)*(y3b/rb+1.0)+1.0/2.0*a*y1/pw3(rb)*(1.0/rb+1.0/W6)*2.0*y3b-a*y1/rb*(-1.0/2.0/pw3(rb)*2.0*y3b-1.0/pw2(W6)*(y3b/rb+1.0)))+y2*cotB/rb/W7*((-2.0+2.0*nu)*cosB+W1/W7*W9+a*y3b/rb2/cosB)-1.0/2.0*y2*W8*cotB/pw3(rb)/W7*((-2.0+2.0*nu)*cosB+W1/W7*W9+a*y3b/rb2/cosB)*2.0*y3b-y2*W8*cotB/rb/pw2(W7)*((-2.0+2.0*nu)*cosB+W1/W7*W9+a*y3b/rb2/cosB)*W3+y2*W8*cotB/rb/W7*((cosB*y3b/rb+1.0)/W7*W9-W1/pw2(W7)*W9*W3-1.0/2.0*W1/W7*a/pw3(rb)/cosB*2.0*y3b+a/rb2/cosB-a*y3b/pw2(rb2)/cosB*2.0*y3b))/kPi/(1.0-nu))+(1.0/4*((-2.0+2.0*nu)*N1*cotB*(1.0/rb*y2/W6-cosB/rb*y2/W7)+(2.0-2.0*nu)*y1/pw2(W6)*W5/rb*y2+(2.0-2.0*nu)*y1/W6*a/pw3(rb)*y2-(2.0-2.0*nu)*z1b/pw2(W7)*W2/rb*y2-(2.0-2.0*nu)*z1b/W7*a/pw3(rb)*y2-W8/pw3(rb)*(N1*cotB-2.0*nu*y1/W6-a*y1/rb2)*y2+W8/rb*(2.0*nu*y1/pw2(W6)/rb*y2+2.0*a*y1/pw2(rb2)*y2)+W8/pw2(W7)*(cosB*sinB+W1*cotB/rb*((2.0-2.0*nu)*cosB-W1/W7)+a/rb*(sinB-y3b*z1b/rb2-z1b*W1/rb/W7))/rb*y2-W8/W7*(1.0/rb2*cosB*y2*cotB*((2.0-2.0*nu)*cosB-W1/W7)-W1*cotB/pw3(rb)*((2.0-2.0*nu)*cosB-W1/W7)*y2+W1*cotB/rb*(-cosB/rb*y2/W7+W1/pw2(W7)/rb*y2)-a/pw3(rb)*(sinB-y3b*z1b/rb2-z1b*W1/rb/W7)*y2+a/rb*(2.0*y3b*z1b/pw2(rb2)*y2-z1b/rb2*cosB*y2/W7+z1b*W1/pw3(rb)/W7*y2+z1b*W1/rb2/pw2(W7)*y2)))/kPi/(1.0-nu));
  const T v23z = (1.0/4*(N1*(-sinB*W3/W7+y1/pw2(W6)*(1.0+a/rb)*(y3b/rb+1.0)+1.0/2.0*y1/W6*a/pw3(rb)*2.0*y3b+sinB/W7*W2-z1b/pw2(W7)*W2*W3-1.0/2.0*z1b/W7*a/pw3(rb)*2.0*y3b)+y1/rb*(a/rb2+1.0/W6)-1.0/2.0*y1*W8/pw3(rb)*(a/rb2+1.0/W6)*2.0*y3b+y1*W8/rb*(-a/pw2(rb2)*2.0*y3b-1.0/pw2(W6)*(y3b/rb+1.0))-1.0/W7*(sinB*(cosB-a/rb)+z1b/rb*(1.0+a*y3b/rb2)-1.0/rb/W7*(pw2(y2)*cosB*sinB-a*z1b/rb*W1))+W8/pw2(W7)*(sinB*(cosB-a/rb)+z1b/rb*(1.0+a*y3b/rb2)-1.0/rb/W7*(pw2(y2)*cosB*sinB-a*z1b/rb*W1))*W3-W8/W7*(1.0/2.0*sinB*a/pw3(rb)*2.0*y3b+sinB/rb*(1.0+a*y3b/rb2)-1.0/2.0*z1b/pw3(rb)*(1.0+a*y3b/rb2)*2.0*y3b+z1b/rb*(a/rb2-a*y3b/pw2(rb2)*2.0*y3b)+1.0/2.0/pw3(rb)/W7*(pw2(y2)*cosB*sinB-a*z1b/rb*W1)*2.0*y3b+1.0/rb/pw2(W7)*(pw2(y2)*cosB*sinB-a*z1b/rb*W1)*W3-1.0/rb/W7*(-a*sinB/rb*W1+1.0/2.0*a*z1b/pw3(rb)*W1*2.0*y3b-a*z1b/rb*(cosB*y3b/rb+1.0))))/kPi/(1.0-nu))+(1.0/4*((2.0-2.0*nu)*rFib_ry2+(2.0-2.0*nu)*sinB/W7*W2-(2.0-2.0*nu)*pw2(y2)*sinB/pw2(W7)*W2/rb-(2.0-2.0*nu)*pw2(y2)*sinB/W7*a/pw3(rb)+W8*sinB/rb/W7*(1.0+W1/W7*W2+a*y3b/rb2)-pw2(y2)*W8*sinB/pw3(rb)/W7*(1.0+W1/W7*W2+a*y3b/rb2)-pw2(y2)*W8*sinB/rb2/pw2(W7)*(1.0+W1/W7*W2+a*y3b/rb2)+y2*W8*sinB/rb/W7*(1.0/rb*cosB*y2/W7*W2-W1/pw2(W7)*W2/rb*y2-W1/W7*a/pw3(rb)*y2-2.0*a*y3b/pw2(rb2)*y2))/kPi/(1.0-nu));
  o[0] = v11x; o[1] = v22x; o[2] = v33x; o[3] = v12x; o[4] = v13x; o[5] = v23x;
  o[6] = v11y; o[7] = v22y; o[8] = v33y; o[9] = v12y; o[10] = v13y; o[11] = v23y;
  o[12] = v11z; o[13] = v22z; o[14] = v33z; o[15] = v12z; o[16] = v13z; o[17] = v23z;
}

// ---------------------------------------------------------------------------------------------------------
// TDSetupS TD.jl:1623-1750, for one point.  acc[m][c]: m = Dn,Dss,Dds; c = xx,yy,zz,xy,xz,yz (TDCS).
template <class T>
inline void td_setup_s(T x, T y, T z, T alpha, T Dn, T Dss, T Dds, T nu, const T* TriVertex, const T* SideVec,
                       T acc[3][6]) {
  const T CosT = SideVec[2];
  const T SinT = SideVec[1];
  T Dss1, Dds0, Dss0, Dds1;
  rot2(Dss, T(0.0), T(0.0), T(0.0), CosT, SinT, Dss1, Dds0);
  rot2(T(0.0), Dds, T(0.0), T(0.0), CosT, SinT, Dss0, Dds1);
  const T Ang = -kPi + alpha;
  const T cosA = cos(Ang);
  const T sinA = sin(Ang);
  const T E1 = (1.0 - nu);
  const T E2 = (2.0 * nu + 1.0);
  const T sinADE1 = sinA / 8 / kPi / (1 - nu);
  const T E3 = Dn / 8 / kPi / E1;
  const T E5Dss1 = Dss1 / 8 / kPi / E1;
  const T E5Dss0 = Dss0 / 8 / kPi / E1;
  const T Bm[9] = {T(1.0), T(0.0), T(0.0), T(0.0), CosT, -SinT, T(0.0), SinT, CosT};
  T Y, Z;
  rot2(y, z, TriVertex[1], TriVertex[2], CosT, SinT, Y, Z);
  T e[21];
  ang_dis_strain(x, Y, Z, cosA, sinA, nu, E2, e);
  const T exxDn = e[0], exxDss = e[1], exxDds = e[2];
  const T eyyDn = e[3], eyyDss = e[4], eyyDds = e[5];
  const T ezzDn = e[6], ezzDss = e[7], ezzDds = e[8];
  const T exyDn = e[9], exyDss = e[10], exyDds = e[11];
  const T exzDn = e[12], exzDss = e[13], exzDds = e[14];
  const T eyzDn = e[15], eyzDss = e[16], eyzDds = e[17];
  const T rFi_rx = e[18], rFi_ry = e[19], rFi_rz = e[20];
  if (Dn != 0) {
    T t[6] = {Dn * (rFi_rx) + E3 * exxDn, E3 * eyyDn, E3 * ezzDn, Dn * (rFi_ry) / 2 - E3 * exyDn,
              Dn * (rFi_rz) / 2 - E3 * exzDn, E3 * eyzDn};
    tensor_tf(t, Bm);
    for (int c = 0; c < 6; ++c) acc[0][c] = acc[0][c] + t[c];
  }
  if (Dss != 0) {
    const T E4Dss1 = Dss1 * x / 8 / kPi / E1;
    T t[6] = {(-E4Dss1 * exxDss) + (Dds0 * x * sinADE1 * exxDds),
              (Dss1 * (rFi_ry)-E4Dss1 * eyyDss) + (Dds0 * x * sinADE1 * eyyDds),
              (-E4Dss1 * ezzDss) + (Dds0 * (rFi_rz) + Dds0 * x * sinADE1 * ezzDds),
              (Dss1 * (rFi_rx) / 2 + E5Dss1 * exyDss) - (Dds0 * sinADE1 * exyDds),
              (E5Dss1 * exzDss) + (Dds0 * (rFi_rx) / 2 - Dds0 * sinADE1 * exzDds),
              (Dss1 * (rFi_rz) / 2 - E4Dss1 * eyzDss) + (Dds0 * (rFi_ry) / 2 - Dds0 * x * sinADE1 * eyzDds)};
    tensor_tf(t, Bm);
    for (int c = 0; c < 6; ++c) acc[1][c] = acc[1][c] + t[c];
  }
  if (Dds != 0) {
    const T E4Dss0 = Dss0 * x / 8 / kPi / E1;
    T t[6] = {(-E4Dss0 * exxDss) + (Dds1 * x * sinADE1 * exxDds),
              (Dss0 * (rFi_ry)-E4Dss0 * eyyDss) + (Dds1 * x * sinADE1 * eyyDds),
              (-E4Dss0 * ezzDss) + (Dds1 * (rFi_rz) + Dds1 * x * sinADE1 * ezzDds),
              (Dss0 * (rFi_rx) / 2 + E5Dss0 * exyDss) - (Dds1 * sinADE1 * exyDds),
              (E5Dss0 * exzDss) + (Dds1 * (rFi_rx) / 2 - Dds1 * sinADE1 * exzDds),
              (Dss0 * (rFi_rz) / 2 - E4Dss0 * eyzDss) + (Dds1 * (rFi_ry) / 2 - Dds1 * x * sinADE1 * eyzDds)};
    tensor_tf(t, Bm);
    for (int c = 0; c < 6; ++c) acc[2][c] = acc[2][c] + t[c];
  }
}

// TDSetupD TD.jl:1081-1173, for one point.  acc[m][c]: m = Dn,Dss,Dds; c = x,y,z (TDCS).
template <class T>
inline void td_setup_d(T x, T y, T z, T alpha, T Dn, T Dss, T Dds, T nu, const T* TriVertex, const T* SideVec,
                       T acc[3][3]) {
  const T CosT = SideVec[2];
  const T SinT = SideVec[1];
  T Dss1, Dds0, Dss0, Dds1;
  rot2(Dss, T(0.0), T(0.0), T(0.0), CosT, SinT, Dss1, Dds0);
  rot2(T(0.0), Dds, T(0.0), T(0.0), CosT, SinT, Dss0, Dds1);
  const T Ang = -kPi + alpha;
  const T cosA = cos(Ang);
  const T sinA = sin(Ang);
  const T E1 = (1.0 - nu);
  const T E2 = (1.0 - 2.0 * nu);
  const T cosA2 = pw2(cosA);
  const T sinADE1 = sinA / 8.0 / kPi / (1.0 - nu);
  const T Dn8p = Dn / 8.0 / kPi;
  T Y, Z;
  rot2(y, z, TriVertex[1], TriVertex[2], CosT, SinT, Y, Z);
  T u[9];
  ang_dis_disp(x, Y, Z, cosA, sinA, E1, E2, cosA2, sinADE1, u);
  const T ux = u[0], uy = u[1], uz = u[2], vx = u[3], vy = u[4], vz = u[5], wx = u[6], wy = u[7], wz = u[8];
  if (Dn != 0.0) {
    acc[0][0] = acc[0][0] + (Dn8p / E1 * ux);
    T a = Dn8p / E1 * vx, b = Dn8p / E1 * wx;
    rot2(a, b, T(0.0), T(0.0), CosT, -SinT, a, b);
    acc[0][1] = acc[0][1] + a;
    acc[0][2] = acc[0][2] + b;
  }
  if (Dss != 0.0) {
    acc[1][0] = acc[1][0] + (Dss1 / 8.0 / kPi / E1 * uy) + (Dds0 * sinADE1 * uz);
    T a = (Dss1 * x / 8.0 / kPi / E1 * vy) + (Dds0 * x * sinADE1 * vz);
    T b = (Dss1 * x / 8.0 / kPi / E1 * wy) + (Dds0 * x * sinADE1 * wz);
    rot2(a, b, T(0.0), T(0.0), CosT, -SinT, a, b);
    acc[1][1] = acc[1][1] + a;
    acc[1][2] = acc[1][2] + b;
  }
  if (Dds != 0.0) {
    acc[2][0] = acc[2][0] + (Dss0 / 8.0 / kPi / E1 * uy) + (Dds1 * sinADE1 * uz);
    T a = (Dss0 * x / 8.0 / kPi / E1 * vy) + (Dds1 * x * sinADE1 * vz);
    T b = (Dss0 * x / 8.0 / kPi / E1 * wy) + (Dds1 * x * sinADE1 * wz);
    rot2(a, b, T(0.0), T(0.0), CosT, -SinT, a, b);
    acc[2][1] = acc[2][1] + a;
    acc[2][2] = acc[2][2] + b;
  }
}

// Runs the three active angular dislocations of TDstrainFS / TDdispFS (TD.jl:1528-1569, 767-803) for one
// point: configuration I uses (A,p1,-e13),(B,p2,e12),(C,p3,e23); configuration II uses (A,p1,e13),
// (B,p2,-e12),(C,p3,-e23).
template <class T, class F> inline void three_ads(const Tri<T>& t, int mode, F&& f) {
  T ne13[3] = {-t.e13[0], -t.e13[1], -t.e13[2]};
  T ne12[3] = {-t.e12[0], -t.e12[1], -t.e12[2]};
  T ne23[3] = {-t.e23[0], -t.e23[1], -t.e23[2]};
  if (mode == 1) {
    f(t.A, t.p1, ne13);
    f(t.B, t.p2, t.e12);
    f(t.C, t.p3, t.e23);
  } else if (mode == -1) {
    f(t.A, t.p1, t.e13);
    f(t.B, t.p2, ne12);
    f(t.C, t.p3, ne23);
  }
}

// TDstrainFS TD.jl:1499-1620 for one EFCS point.  eps[m][c] holds the accumulated EFCS strain on entry
// (the caller's column values) and on exit.  For image == true it reproduces the reference's in-place
// rotation into the image TDCS and back (TD.jl:1509-1516, 1597-1602), including the flat-triangle case where
// the xz/yz sign flips rebind local copies (TD.jl:1518-1525, 1610-1617): the caller's xz/yz then keep the
// image-TDCS values of the accumulated tensor and its xx..xy receive the rotation of the sign-flipped copy.
template <class T>
inline void td_strain_fs(const Tri<T>& t, T X, T Y, T Z, T Dss, T Dds, T Dn, T nu, bool image, T eps[3][6]) {
  T loc[3][6];
  T keep_xz[3], keep_yz[3];
  const bool quirk = image && t.flat;
  if (image) {
    const T RotInv[9] = {t.Vn[0], t.Vs[0], t.Vd[0], t.Vn[1], t.Vs[1], t.Vd[1], t.Vn[2], t.Vs[2], t.Vd[2]};
    for (int m = 0; m < 3; ++m) {
      for (int c = 0; c < 6; ++c) loc[m][c] = eps[m][c];
      tensor_tf(loc[m], RotInv);
      if (quirk) {
        keep_xz[m] = loc[m][4];
        keep_yz[m] = loc[m][5];
        loc[m][4] = -loc[m][4];
        loc[m][5] = -loc[m][5];
      }
    }
  } else {
    for (int m = 0; m < 3; ++m)
      for (int c = 0; c < 6; ++c) loc[m][c] = eps[m][c];
  }
  T x, y, z;
  to_tdcs(t, X, Y, Z, x, y, z);
  const int mode = trimode(t, x, y, z);
  three_ads(t, mode, [&](T alpha, const T* tv, const T* sv) { td_setup_s(x, y, z, alpha, Dn, Dss, Dds, nu, tv, sv, loc); });
  if (mode == 0) {
    for (int m = 0; m < 3; ++m)
      for (int c = 0; c < 6; ++c) loc[m][c] = std::numeric_limits<double>::quiet_NaN();
  }
  const T RotMat[9] = {t.Vn[0], t.Vn[1], t.Vn[2], t.Vs[0], t.Vs[1], t.Vs[2], t.Vd[0], t.Vd[1], t.Vd[2]};
  for (int m = 0; m < 3; ++m) {
    tensor_tf(loc[m], RotMat);
    if (quirk) {
      for (int c = 0; c < 4; ++c) eps[m][c] = loc[m][c];
      eps[m][4] = keep_xz[m];
      eps[m][5] = keep_yz[m];
    } else {
      for (int c = 0; c < 6; ++c) eps[m][c] = loc[m][c];
    }
  }
}

// TDdispFS TD.jl:735-867 for one EFCS point; u[m][c] holds the caller's accumulated EFCS displacement.
// Flat image triangles: the Uz sign flip rebinds a local copy (TD.jl:759-763), so the caller's Uz keeps the
// image-TDCS z value of the accumulated vector while its Ux/Uy receive the rotation of the flipped copy.
template <class T>
inline void td_disp_fs(const Tri<T>& t, T X, T Y, T Z, T Dss, T Dds, T Dn, T nu, bool image, T u[3][3]) {
  T loc[3][3];
  T keep_z[3];
  const bool quirk = image && t.flat;
  if (image) {
    const T VxR[3] = {t.Vn[0], t.Vs[0], t.Vd[0]};
    const T VyR[3] = {t.Vn[1], t.Vs[1], t.Vd[1]};
    const T VzR[3] = {t.Vn[2], t.Vs[2], t.Vd[2]};
    for (int m = 0; m < 3; ++m) {
      rot3(u[m][0], u[m][1], u[m][2], T(0.0), T(0.0), T(0.0), VxR, VyR, VzR, loc[m][0], loc[m][1], loc[m][2]);
      if (quirk) {
        keep_z[m] = loc[m][2];
        loc[m][2] = -loc[m][2];
      }
    }
  } else {
    for (int m = 0; m < 3; ++m)
      for (int c = 0; c < 3; ++c) loc[m][c] = u[m][c];
  }
  T x, y, z;
  to_tdcs(t, X, Y, Z, x, y, z);
  const int mode = trimode(t, x, y, z);
  three_ads(t, mode, [&](T alpha, const T* tv, const T* sv) { td_setup_d(x, y, z, alpha, Dn, Dss, Dds, nu, tv, sv, loc); });
  if (mode == 0) {
    for (int m = 0; m < 3; ++m)
      for (int c = 0; c < 3; ++c) loc[m][c] = std::numeric_limits<double>::quiet_NaN();
  }
  // Burgers' function TD.jl:820-840
  const T a1 = -x, a2 = t.p1[1] - y, a3 = t.p1[2] - z;
  const T b1 = -x, b2 = -y, b3 = -z;
  const T c1 = -x, c2 = t.p3[1] - y, c3 = t.p3[2] - z;
  const T a12 = pw2(a1);
  const T na = sqrt((a12 + pw2(a2) + pw2(a3)));
  const T nb = sqrt((a12 + pw2(b2) + pw2(b3)));
  const T nc = sqrt((a12 + pw2(c2) + pw2(c3)));
  const T one = a1 * (b2 * c3 - b3 * c2) - a2 * (b1 * c3 - b3 * c1) + a3 * (b1 * c2 - b2 * c1);
  const T ab = a1 * b1 + a2 * b2 + a3 * b3;
  const T ac = a1 * c1 + a2 * c2 + a3 * c3;
  const T bc = b1 * c1 + b2 * c2 + b3 * c3;
  const T two = na * nb * nc + ab * nc + ac * nb + bc * na;
  const T Fi = -2.0 * atan2(one, two) / 4.0 / kPi;
  loc[0][0] = Dn * Fi + loc[0][0];
  loc[1][1] = Dss * Fi + loc[1][1];
  loc[2][2] = Dds * Fi + loc[2][2];
  for (int m = 0; m < 3; ++m) {
    T r0, r1, r2;
    rot3(loc[m][0], loc[m][1], loc[m][2], T(0.0), T(0.0), T(0.0), t.Vn, t.Vs, t.Vd, r0, r1, r2);
    if (quirk) {
      u[m][0] = r0;
      u[m][1] = r1;
      u[m][2] = keep_z[m];
    } else {
      u[m][0] = r0;
      u[m][1] = r1;
      u[m][2] = r2;
    }
  }
}

// Per-side constants of the free-surface correction: CalcSideVec (TD.jl:935-960), the skip test, the ADCS
// frame and the slip vectors in ADCS (AngSetupStrainFSC TD.jl:1876-1933, AngSetupDispFSC TD.jl:1267-1322).
template <class T> struct FscSide {
  bool skip;
  T PA[3], PB[3];
  T beta;
  T ey1[3], ey2[3], ey3[3];
  T y1B, y2B, y3B;
  T b[3][3];  // b[m] = ADCS slip vector of mode m (Dn, Dss, Dds)
};

template <class T>
inline void fsc_side_setup(const T* PA, const T* PB, const Tri<T>& t, T Dn, T Dss, T Dds, FscSide<T>& s) {
  T SideVec[3];
  for (int k = 0; k < 3; ++k) {
    SideVec[k] = PB[k] - PA[k];
    s.PA[k] = PA[k];
    s.PB[k] = PB[k];
  }
  const T eZ[3] = {T(0.0), T(0.0), T(1.0)};
  const T Sv = sqrt(pw2(SideVec[0]) + pw2(SideVec[1]) + pw2(SideVec[2]));
  T beta = -SideVec[0] * eZ[0] - SideVec[1] * eZ[1] - SideVec[2] * eZ[2];
  beta = acos(beta / Sv);
  s.beta = beta;
  using std::fabs;
  s.skip = (fabs(beta) < kEps || fabs(kPi - beta) < kEps || fabs(cot(beta)) > 5e3);
  if (s.skip) return;
  s.ey1[0] = SideVec[0];
  s.ey1[1] = SideVec[1];
  s.ey1[2] = T(0.0);
  const T nrmey1 = sqrt(pw2(s.ey1[0]) + pw2(s.ey1[1]) + pw2(s.ey1[2]));
  for (int k = 0; k < 3; ++k) {
    s.ey1[k] = s.ey1[k] / nrmey1;
    s.ey3[k] = -eZ[k];
  }
  cross3(s.ey3, s.ey1, s.ey2);
  rot3(SideVec[0], SideVec[1], SideVec[2], T(0.0), T(0.0), T(0.0), s.ey1, s.ey2, s.ey3, s.y1B, s.y2B, s.y3B);
  const T D[3][3] = {{Dn, T(0.0), T(0.0)}, {T(0.0), Dss, T(0.0)}, {T(0.0), T(0.0), Dds}};
  for (int m = 0; m < 3; ++m) {
    T e0, e1, e2;
    rot3(D[m][0], D[m][1], D[m][2], T(0.0), T(0.0), T(0.0), t.Vn, t.Vs, t.Vd, e0, e1, e2);
    rot3(e0, e1, e2, T(0.0), T(0.0), T(0.0), s.ey1, s.ey2, s.ey3, s.b[m][0], s.b[m][1], s.b[m][2]);
  }
}

// AngSetupStrainFSC TD.jl:1935-2075 for one EFCS point; adds into eps[m][c] (EFCS).
template <class T>
inline void strain_fsc_side(const FscSide<T>& s, T X, T Y, T Z, T Dn, T Dss, T Dds, T nu, T eps[3][6]) {
  if (s.skip) return;
  const T beta = s.beta;
  const T pib = kPi - beta;
  const T sinPib = sin(pib), cosPib = cos(pib), cotPib = cot(pib);
  const T sinB = sin(beta), cosB = cos(beta), cotB = cot(beta);
  T x, y, z;
  rot3(X, Y, Z, s.PA[0], s.PA[1], s.PA[2], s.ey1, s.ey2, s.ey3, x, y, z);
  const T Bx = beta * x;
  T vA[18], vB[18];
  T b[3][3];
  for (int m = 0; m < 3; ++m)
    for (int k = 0; k < 3; ++k) b[m][k] = s.b[m][k];
  if (Bx >= 0.0) {
    ang_dis_strain_fsc(-x, -y, z, cosPib, sinPib, cotPib, nu, -s.PA[2], vA);
    for (int k = 0; k < 3; ++k) {
      vA[6 * k + 4] = -vA[6 * k + 4];
      vA[6 * k + 5] = -vA[6 * k + 5];
    }
    for (int m = 0; m < 3; ++m) {
      b[m][0] = -b[m][0];
      b[m][1] = -b[m][1];
    }
    ang_dis_strain_fsc(-x + s.y1B, -y + s.y2B, z - s.y3B, cosPib, sinPib, cotPib, nu, -s.PB[2], vB);
    for (int k = 0; k < 3; ++k) {
      vB[6 * k + 4] = -vB[6 * k + 4];
      vB[6 * k + 5] = -vB[6 * k + 5];
    }
  } else {
    ang_dis_strain_fsc(x, y, z, cosB, sinB, cotB, nu, -s.PA[2], vA);
    ang_dis_strain_fsc(x - s.y1B, y - s.y2B, z - s.y3B, cosB, sinB, cotB, nu, -s.PB[2], vB);
  }
  const T AFlip[9] = {s.ey1[0], s.ey1[1], s.ey1[2], s.ey2[0], s.ey2[1], s.ey2[2], s.ey3[0], s.ey3[1], s.ey3[2]};
  const T Dm[3] = {Dn, Dss, Dds};
  for (int m = 0; m < 3; ++m) {
    if (Dm[m] == 0.0) continue;
    const T b1 = b[m][0], b2 = b[m][1], b3 = b[m][2];
    T e[6];
    for (int c = 0; c < 3; ++c) e[c] = -((b1 * vA[c]) + (b2 * vA[6 + c]) + (b3 * vA[12 + c]));
    for (int c = 3; c < 6; ++c) e[c] = -((b1 / 2.0 * vA[c]) + (b2 / 2.0 * vA[6 + c]) + (b3 / 2.0 * vA[12 + c]));
    for (int c = 0; c < 3; ++c) e[c] = e[c] + ((b1 * vB[c]) + (b2 * vB[6 + c]) + (b3 * vB[12 + c]));
    for (int c = 3; c < 6; ++c) e[c] = e[c] + ((b1 / 2.0 * vB[c]) + (b2 / 2.0 * vB[6 + c]) + (b3 / 2.0 * vB[12 + c]));
    tensor_tf(e, AFlip);
    for (int c = 0; c < 6; ++c) eps[m][c] = eps[m][c] + e[c];
  }
}

// AngSetupDispFSC TD.jl:1324-1405 for one EFCS point; adds into u[m][c] (EFCS).
template <class T>
inline void disp_fsc_side(const FscSide<T>& s, T X, T Y, T Z, T Dn, T Dss, T Dds, T nu, T u[3][3]) {
  if (s.skip) return;
  const T beta = s.beta;
  const T pib = -kPi + beta;
  const T sinPib = sin(pib), cosPib = cos(pib), cotPib = cot(pib), cotPib2 = cot(pib / 2);
  const T sinB = sin(beta), cosB = cos(beta), cotB = cot(beta), cotB2 = cot(beta / 2);
  T x, y, z;
  rot3(X, Y, Z, s.PA[0], s.PA[1], s.PA[2], s.ey1, s.ey2, s.ey3, x, y, z);
  const T Bx = beta * x;
  T uA[9], uB[9];
  if (Bx >= 0.0) {
    ang_dis_disp_fsc(x, y, z, cosPib, sinPib, cotPib, cotPib2, nu, -s.PA[2], uA);
    ang_dis_disp_fsc(x - s.y1B, y - s.y2B, z - s.y3B, cosPib, sinPib, cotPib, cotPib2, nu, -s.PB[2], uB);
  } else {
    ang_dis_disp_fsc(x, y, z, cosB, sinB, cotB, cotB2, nu, -s.PA[2], uA);
    ang_dis_disp_fsc(x - s.y1B, y - s.y2B, z - s.y3B, cosB, sinB, cotB, cotB2, nu, -s.PB[2], uB);
  }
  const T VxR[3] = {s.ey1[0], s.ey2[0], s.ey3[0]};
  const T VyR[3] = {s.ey1[1], s.ey2[1], s.ey3[1]};
  const T VzR[3] = {s.ey1[2], s.ey2[2], s.ey3[2]};
  const T Dm[3] = {Dn, Dss, Dds};
  for (int m = 0; m < 3; ++m) {
    if (Dm[m] == 0.0) continue;
    const T b1 = s.b[m][0], b2 = s.b[m][1], b3 = s.b[m][2];
    T w[3];
    for (int c = 0; c < 3; ++c)
      w[c] = -((b1 / 4.0 / kPi / (1 - nu) * uA[3 * c]) + (b2 / 4.0 / kPi / (1 - nu) * uA[3 * c + 1]) +
               (b3 / 4.0 / kPi / (1 - nu) * uA[3 * c + 2]));
    for (int c = 0; c < 3; ++c)
      w[c] = w[c] + ((b1 / 4.0 / kPi / (1 - nu) * uB[3 * c]) + (b2 / 4.0 / kPi / (1 - nu) * uB[3 * c + 1]) +
                     (b3 / 4.0 / kPi / (1 - nu) * uB[3 * c + 2]));
    T r0, r1, r2;
    rot3(w[0], w[1], w[2], T(0.0), T(0.0), T(0.0), VxR, VyR, VzR, r0, r1, r2);
    u[m][0] = u[m][0] + r0;
    u[m][1] = u[m][1] + r1;
    u[m][2] = u[m][2] + r2;
  }
}

// Everything one source triangle contributes at one observation point: ComputeStrainInfluences
// (TD.jl:670-708) and ComputeDispInfluences (TD.jl:631-668).
template <class T> struct TriFull {
  Tri<T> real, img;
  FscSide<T> side[3];
};

template <class T>
inline void tri_full_setup(const T* P1, const T* P2, const T* P3, T Dn, T Dss, T Dds, bool hs, TriFull<T>& f) {
  tri_setup(P1, P2, P3, f.real);
  if (hs) {
    // PrepForLoop TD.jl:428-435: image vertices have z -> -z
    const T P1i[3] = {P1[0], P1[1], -P1[2]};
    const T P2i[3] = {P2[0], P2[1], -P2[2]};
    const T P3i[3] = {P3[0], P3[1], -P3[2]};
    tri_setup(P1i, P2i, P3i, f.img);
    fsc_side_setup(P1, P2, f.real, Dn, Dss, Dds, f.side[0]);
    fsc_side_setup(P2, P3, f.real, Dn, Dss, Dds, f.side[1]);
    fsc_side_setup(P3, P1, f.real, Dn, Dss, Dds, f.side[2]);
  }
}

template <class T>
inline void strain_influence(const TriFull<T>& f, bool hs, T X, T Y, T Z, T Dss, T Dds, T Dn, T nu, T eps[3][6]) {
  for (int m = 0; m < 3; ++m)
    for (int c = 0; c < 6; ++c) eps[m][c] = T(0.0);
  td_strain_fs(f.real, X, Y, Z, Dss, Dds, Dn, nu, false, eps);
  if (hs) {
    for (int k = 0; k < 3; ++k) strain_fsc_side(f.side[k], X, Y, Z, Dn, Dss, Dds, nu, eps);
    td_strain_fs(f.img, X, Y, Z, Dss, Dds, Dn, nu, true, eps);
  }
}

template <class T>
inline void disp_influence(const TriFull<T>& f, bool hs, T X, T Y, T Z, T Dss, T Dds, T Dn, T nu, T u[3][3]) {
  for (int m = 0; m < 3; ++m)
    for (int c = 0; c < 3; ++c) u[m][c] = T(0.0);
  td_disp_fs(f.real, X, Y, Z, Dss, Dds, Dn, nu, false, u);
  if (hs) {
    for (int k = 0; k < 3; ++k) disp_fsc_side(f.side[k], X, Y, Z, Dn, Dss, Dds, nu, u);
    td_disp_fs(f.img, X, Y, Z, Dss, Dds, Dn, nu, true, u);
  }
}

}  // namespace orc
