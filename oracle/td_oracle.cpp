// C entry points of the CPU FP64 oracle (test infrastructure only; see td_oracle.hpp).
//
// Layouts follow the reference's Julia arrays: P1/P2/P3 are ntri x 3 column-major (P1[k*ntri + j] is
// coordinate k of triangle j); influence matrices are nobs x ntri column-major (entry (i, j) at i + j*nobs).
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "td_oracle.hpp"

using orc::TriFull;

namespace {

thread_local const char* g_err = "";

inline void load_tri(const double* P1, const double* P2, const double* P3, int64_t ntri, int64_t j, double* a,
                     double* b, double* c) {
  for (int k = 0; k < 3; ++k) {
    a[k] = P1[k * ntri + j];
    b[k] = P2[k * ntri + j];
    c[k] = P3[k * ntri + j];
  }
}

// Input validation of PrepForLoop (TD.jl:418-425) and ViewInLoop (TD.jl:606-619).
int check_inputs(const double* Z, int64_t nobs, const double* P1, const double* P2, const double* P3, int64_t ntri,
                 int hs) {
  if (nobs < 0 || ntri < 0) {
    g_err = "Element size inputs must be the same dimensions";
    return 1;
  }
  if (hs == 1 && ntri > 0) {
    bool above = false;
    for (int64_t j = 0; j < ntri; ++j)
      if (P1[2 * ntri + j] > 0 || P2[2 * ntri + j] > 0 || P3[2 * ntri + j] > 0) above = true;
    for (int64_t i = 0; i < nobs; ++i)
      if (Z[i] > 0) above = true;
    if (above) {
      g_err = "Half-space solution: Z coordinates must be negative!";
      return 2;
    }
  }
  return 0;
}

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err; }

// TD with preallocated influence matrices (TD.jl:135-216).  strain: 18 pointers in StrainInf field order
// (exxDn,eyyDn,ezzDn,exyDn,exzDn,eyzDn, ...Dss, ...Dds) or NULL; disp: 9 pointers in DispInf order
// (DnUx,DnUy,DnUz, DssUx,..., DdsUx,...) or NULL.  Parallel over source triangles like TD.jl:146.
int orc_td(const double* X, const double* Y, const double* Z, int64_t nobs, const double* P1, const double* P2,
           const double* P3, int64_t ntri, const double* Dss, const double* Dds, const double* Dn, double nu,
           double G, int disp_flag, int strain_flag, int hs_flag, double* const* strain, double* const* disp,
           int nthreads) {
  (void)G;
  int rc = check_inputs(Z, nobs, P1, P2, P3, ntri, hs_flag);
  if (rc) return rc;
  const bool hs = hs_flag == 1;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t j = 0; j < ntri; ++j) {
    double a[3], b[3], c[3];
    load_tri(P1, P2, P3, ntri, j, a, b, c);
    TriFull<double> f;
    orc::tri_full_setup(a, b, c, Dn[j], Dss[j], Dds[j], hs, f);
    for (int64_t i = 0; i < nobs; ++i) {
      const int64_t o = i + j * nobs;
      if (disp_flag == 1) {
        double u[3][3];
        orc::disp_influence(f, hs, X[i], Y[i], Z[i], Dss[j], Dds[j], Dn[j], nu, u);
        for (int m = 0; m < 3; ++m)
          for (int k = 0; k < 3; ++k) disp[3 * m + k][o] = u[m][k];
      }
      if (strain_flag == 1) {
        double e[3][6];
        orc::strain_influence(f, hs, X[i], Y[i], Z[i], Dss[j], Dds[j], Dn[j], nu, e);
        for (int m = 0; m < 3; ++m)
          for (int k = 0; k < 6; ++k) strain[6 * m + k][o] = e[m][k];
      }
    }
  }
  return 0;
}

// TD summed-field driver (TD.jl:219-410): strain6 = {exx,eyy,ezz,exy,exz,eyz}, disp3 = {Ux,Uy,Uz}.
// The reference accumulates per triangle as field .+= Dn .+ Dss .+ Dds (TD.jl:322-324, 375-380) with a race
// under several threads; here each observation point sums the triangles in index order, which is the
// single-thread order of the reference, and is deterministic.
int orc_td_field(const double* X, const double* Y, const double* Z, int64_t nobs, const double* P1,
                 const double* P2, const double* P3, int64_t ntri, const double* Dss, const double* Dds,
                 const double* Dn, double nu, double G, int disp_flag, int strain_flag, int hs_flag,
                 double* const* strain6, double* const* disp3, int nthreads) {
  (void)G;
  int rc = check_inputs(Z, nobs, P1, P2, P3, ntri, hs_flag);
  if (rc) return rc;
  const bool hs = hs_flag == 1;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
  std::vector<TriFull<double>> tris(ntri);
  for (int64_t j = 0; j < ntri; ++j) {
    double a[3], b[3], c[3];
    load_tri(P1, P2, P3, ntri, j, a, b, c);
    orc::tri_full_setup(a, b, c, Dn[j], Dss[j], Dds[j], hs, tris[j]);
  }
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
  for (int64_t i = 0; i < nobs; ++i) {
    double su[3] = {0, 0, 0}, se[6] = {0, 0, 0, 0, 0, 0};
    for (int64_t j = 0; j < ntri; ++j) {
      if (disp_flag == 1) {
        double u[3][3];
        orc::disp_influence(tris[j], hs, X[i], Y[i], Z[i], Dss[j], Dds[j], Dn[j], nu, u);
        for (int k = 0; k < 3; ++k) su[k] = su[k] + (u[0][k] + u[1][k] + u[2][k]);
      }
      if (strain_flag == 1) {
        double e[3][6];
        orc::strain_influence(tris[j], hs, X[i], Y[i], Z[i], Dss[j], Dds[j], Dn[j], nu, e);
        for (int k = 0; k < 6; ++k) se[k] = se[k] + (e[0][k] + e[1][k] + e[2][k]);
      }
    }
    if (disp_flag == 1)
      for (int k = 0; k < 3; ++k) disp3[k][i] = su[k];
    if (strain_flag == 1)
      for (int k = 0; k < 6; ++k) strain6[k][i] = se[k];
  }
  return 0;
}

// CalculateSSandDSDirs.jl:37-98.  normals, ss, ds: N x 3 column-major.
void orc_ss_ds_dirs(const double* nrm, int64_t n, double* ss, double* ds) {
  for (int64_t i = 0; i < n; ++i) {
    const double nx = nrm[i], ny = nrm[n + i], nz = nrm[2 * n + i];
    // normr of [nx ny 0]
    const double h = std::sqrt(nx * nx + ny * ny + 0.0 * 0.0);
    const double r0 = nx / h, r1 = ny / h;
    ss[i] = -r1;
    ss[n + i] = r0;
    ss[2 * n + i] = 0.0;
    const double DSCosAz = std::cos(std::acos(nz) - (orc::kPi / 2));
    double L = std::sin(std::acos(DSCosAz));
    if (nz < 0) L = -L;
    const double a = std::atan2(ny, nx);
    const double DSCosAx = std::cos(a) * L;
    const double DSCosAy = std::sin(a) * L;
    ds[i] = -DSCosAx;
    ds[n + i] = -DSCosAy;
    ds[2 * n + i] = DSCosAz;
    // flat-triangle convention (CalculateSSandDSDirs.jl:81-95): cross(eZ, n) == 0
    const double vs0 = 0.0 * nz - 1.0 * ny, vs1 = 1.0 * nx - 0.0 * nz, vs2 = 0.0 * ny - 0.0 * nx;
    if (vs0 == 0.0 && vs1 == 0.0 && vs2 == 0.0) {
      ds[i] = -1.0;
      ds[n + i] = 0.0;
      ds[2 * n + i] = 0.0;
      if (nz < 1.0) {
        ss[i] = 0.0;
        ss[n + i] = -1.0;
        ss[2 * n + i] = 0.0;
      } else {
        ss[i] = 0.0;
        ss[n + i] = 1.0;
        ss[2 * n + i] = 0.0;
      }
    }
  }
}

// Strain (6) -> traction (Tn, Tss, Tds) on a receiver plane: HookesLaw3DStrain2Stress!.jl:51-62,
// CalculateNormalTraction3D!.jl:37-56, TractionVectorCartesianComponents3D!.jl:43-50,
// CalculateTractionInChosenDirection3D!.jl:48-66.
}  // extern "C"

template <class T>
static inline void strain_to_traction(const T* e, T lam, T G, const T* nv, const T* s, const T* d, T* Tn, T* Tss,
                                      T* Tds) {
  const T G2 = 2.0 * G;
  const T cons = lam * (e[0] + e[1] + e[2]);
  const T Sxx = G2 * e[0] + cons, Syy = G2 * e[1] + cons, Szz = G2 * e[2] + cons;
  const T Sxy = G2 * e[3], Sxz = G2 * e[4], Syz = G2 * e[5];
  const T nx = nv[0], ny = nv[1], nz = nv[2];
  {
    const T CosAx2 = nx * nx, CosAy2 = ny * ny, CosAz2 = nz * nz;
    const T CosAxy = nx * ny, CosAyz = ny * nz, CosAzx = nz * nx;
    const T AA = Sxx * CosAx2, BB = Syy * CosAy2, CC = Szz * CosAz2;
    const T DD = 2 * Sxy * CosAxy, EE = 2 * Syz * CosAyz, FF = 2 * Sxz * CosAzx;
    *Tn = AA + BB + CC + DD + EE + FF;
  }
  const T Tx = Sxx * nx + Sxy * ny + Sxz * nz;
  const T Ty = Sxy * nx + Syy * ny + Syz * nz;
  const T Tz = Sxz * nx + Syz * ny + Szz * nz;
  const T CosAx2 = 1 - (nx * nx), CosAy2 = 1 - (ny * ny), CosAz2 = 1 - (nz * nz);
  const T CosAxy = nx * ny, CosAyz = ny * nz, CosAzx = nz * nx;
  const T* dirs[2] = {s, d};
  T* outs[2] = {Tss, Tds};
  for (int q = 0; q < 2; ++q) {
    const T* c = dirs[q];
    const T A = (Tx * CosAx2 - Ty * CosAxy - Tz * CosAzx) * c[0];
    const T B = (Tx * -CosAxy + Ty * CosAy2 - Tz * CosAyz) * c[1];
    const T C = (Tx * -CosAzx - Ty * CosAyz + Tz * CosAz2) * c[2];
    *outs[q] = A + B + C;
  }
}

extern "C" {

// CalculateInfluenceMatrices3D without fixed elements (CalculateInfluenceMatrices3D.jl:35-43):
// collocation points (SetupCollationPoints :52-71, unit slip) -> TD strain -> traction
// (ConvertStrainInfMatsToTraction :73-142).  T9: TractionInf field order
// (DnTn,DnTss,DnTds, DssTn,DssTss,DssTds, DdsTn,DdsTss,DdsTds), each N x N column-major.
int orc_traction_influence(const double* normals, const double* mid, const double* P1, const double* P2,
                           const double* P3, int64_t n, double nu, double G, double lam, int hs_flag,
                           double* const* T9, int nthreads) {
  std::vector<double> x(n), y(n), z(n), ones(n, 1.0);
  for (int64_t i = 0; i < n; ++i) {
    x[i] = mid[i] - (normals[i] * 1e-12);
    y[i] = mid[n + i] - (normals[n + i] * 1e-12);
    z[i] = mid[2 * n + i] - (normals[2 * n + i] * 1e-12);
  }
  int rc = check_inputs(z.data(), n, P1, P2, P3, n, hs_flag);
  if (rc) return rc;
  std::vector<double> ss(3 * n), ds(3 * n);
  orc_ss_ds_dirs(normals, n, ss.data(), ds.data());
  const bool hs = hs_flag == 1;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t j = 0; j < n; ++j) {
    double a[3], b[3], c[3];
    load_tri(P1, P2, P3, n, j, a, b, c);
    TriFull<double> f;
    orc::tri_full_setup(a, b, c, 1.0, 1.0, 1.0, hs, f);
    for (int64_t i = 0; i < n; ++i) {
      double e[3][6];
      orc::strain_influence(f, hs, x[i], y[i], z[i], 1.0, 1.0, 1.0, nu, e);
      const double nv[3] = {normals[i], normals[n + i], normals[2 * n + i]};
      const double sv[3] = {ss[i], ss[n + i], ss[2 * n + i]};
      const double dv[3] = {ds[i], ds[n + i], ds[2 * n + i]};
      const int64_t o = i + j * n;
      // mode order in e: Dn, Dss, Dds; T9 order: Dn*, Dss*, Dds* each (Tn, Tss, Tds)
      for (int m = 0; m < 3; ++m) strain_to_traction<double>(e[m], lam, G, nv, sv, dv, &T9[3 * m][o], &T9[3 * m + 1][o], &T9[3 * m + 2][o]);
    }
  }
  return 0;
}

// CalculateInfluenceMatrices3D with fixed elements (CalculateInfluenceMatrices3D.jl:10-33, 145-198).
//   * TD displacement at the fixed collocation points (:13-19) and TD strain at the free ones (:21-28), both for
//     every source with unit slip;
//   * the free strain rows go through ConvertStrainInfMatsToTraction with the full CosAx/CosAy/CosAz arrays
//     (:26), whose loops index the normals and the strike/dip cosines by row (CalculateNormalTraction3D!.jl:37,
//     CalculateTractionInChosenDirection3D!.jl:47): the t-th free row takes element t's plane;
//   * PutDispInfsIntoInfMat (:145-176) writes the fixed elements' rows from the DispInf field at the same
//     position (DnTn <- DnUx, DnTss <- DnUy, DnTds <- DnUz, ...) and the free rows from the traction, then
//     keeps the columns of the free sources.
// fixed: n flags (1 = fixed).  T9: TractionInf order, each n x nf column-major (nf = free count).
int orc_traction_influence_fixed(const double* normals, const double* mid, const double* P1, const double* P2,
                                 const double* P3, int64_t n, double nu, double G, double lam, int hs_flag,
                                 const int* fixed, double* const* T9, int nthreads) {
  std::vector<double> x(n), y(n), z(n);
  for (int64_t i = 0; i < n; ++i) {
    x[i] = mid[i] - (normals[i] * 1e-12);
    y[i] = mid[n + i] - (normals[n + i] * 1e-12);
    z[i] = mid[2 * n + i] - (normals[2 * n + i] * 1e-12);
  }
  int rc = check_inputs(z.data(), n, P1, P2, P3, n, hs_flag);
  if (rc) return rc;
  std::vector<double> ss(3 * n), ds(3 * n);
  orc_ss_ds_dirs(normals, n, ss.data(), ds.data());
  // row r -> free count t (traction with element t's plane) or fixed; source j -> output column
  std::vector<int64_t> rowfree(n, -1), col(n, -1);
  int64_t nf = 0;
  for (int64_t i = 0; i < n; ++i)
    if (fixed[i] != 1) {
      rowfree[i] = nf;
      col[i] = nf;
      ++nf;
    }
  const bool hs = hs_flag == 1;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t j = 0; j < n; ++j) {
    if (col[j] < 0) continue;
    double a[3], b[3], c[3];
    load_tri(P1, P2, P3, n, j, a, b, c);
    TriFull<double> f;
    orc::tri_full_setup(a, b, c, 1.0, 1.0, 1.0, hs, f);
    for (int64_t i = 0; i < n; ++i) {
      const int64_t o = i + col[j] * n;
      if (rowfree[i] < 0) {
        double u[3][3];
        orc::disp_influence(f, hs, x[i], y[i], z[i], 1.0, 1.0, 1.0, nu, u);
        for (int m = 0; m < 3; ++m)
          for (int q = 0; q < 3; ++q) T9[3 * m + q][o] = u[m][q];
        continue;
      }
      const int64_t t = rowfree[i];
      double e[3][6];
      orc::strain_influence(f, hs, x[i], y[i], z[i], 1.0, 1.0, 1.0, nu, e);
      const double nv[3] = {normals[t], normals[n + t], normals[2 * n + t]};
      const double sv[3] = {ss[t], ss[n + t], ss[2 * n + t]};
      const double dv[3] = {ds[t], ds[n + t], ds[2 * n + t]};
      for (int m = 0; m < 3; ++m)
        strain_to_traction<double>(e[m], lam, G, nv, sv, dv, &T9[3 * m][o], &T9[3 * m + 1][o], &T9[3 * m + 2][o]);
    }
  }
  return 0;
}

// CalculateNormalAndShearTractions3D.jl:37-58 for a per-element remote stress: returns Tn, Tds, Tss.
void orc_normal_shear_tractions(const double* sxx, const double* syy, const double* szz, const double* sxy,
                                const double* sxz, const double* syz, const double* normals, int64_t n, double* Tn,
                                double* Tds, double* Tss) {
  std::vector<double> ss(3 * n), ds(3 * n);
  orc_ss_ds_dirs(normals, n, ss.data(), ds.data());
  for (int64_t i = 0; i < n; ++i) {
    const double e[6] = {sxx[i], syy[i], szz[i], sxy[i], sxz[i], syz[i]};
    // stresses enter directly: use G = 1/2, lambda = 0 so Hooke's law is the identity
    const double nv[3] = {normals[i], normals[n + i], normals[2 * n + i]};
    const double sv[3] = {ss[i], ss[n + i], ss[2 * n + i]};
    const double dv[3] = {ds[i], ds[n + i], ds[2 * n + i]};
    strain_to_traction<double>(e, 0.0, 0.5, nv, sv, dv, &Tn[i], &Tss[i], &Tds[i]);
  }
}

// CreateFaceNormalAndMidPoint (P1,P2,P3 form): CreateTriangleNormal!.jl:14-30 and incentre midpoints
// CreateMidPoint!.jl:3-16.  Returns 3 if a triangle has a zero-length edge.
int orc_face_normal_midpoint(const double* P1, const double* P2, const double* P3, int64_t n, double* normals,
                             double* mid) {
  for (int64_t i = 0; i < n; ++i) {
    double Pa[3], Pb[3], Pc[3];
    load_tri(P1, P2, P3, n, i, Pa, Pb, Pc);
    const double U[3] = {Pb[0] - Pa[0], Pb[1] - Pa[1], Pb[2] - Pa[2]};
    const double V[3] = {Pc[0] - Pa[0], Pc[1] - Pa[1], Pc[2] - Pa[2]};
    const double Nx = (U[1] * V[2]) - (U[2] * V[1]);
    const double Ny = (U[2] * V[0]) - (U[0] * V[2]);
    const double Nz = (U[0] * V[1]) - (U[1] * V[0]);
    const double aMag = std::sqrt((Nx * Nx) + (Ny * Ny) + (Nz * Nz));
    normals[i] = Nx / aMag;
    normals[n + i] = Ny / aMag;
    normals[2 * n + i] = Nz / aMag;
    const double c = std::sqrt(orc::pw2(Pa[0] - Pb[0]) + orc::pw2(Pa[1] - Pb[1]) + orc::pw2(Pa[2] - Pb[2]));
    const double a = std::sqrt(orc::pw2(Pb[0] - Pc[0]) + orc::pw2(Pb[1] - Pc[1]) + orc::pw2(Pb[2] - Pc[2]));
    const double b = std::sqrt(orc::pw2(Pa[0] - Pc[0]) + orc::pw2(Pa[1] - Pc[1]) + orc::pw2(Pa[2] - Pc[2]));
    if (a == 0 || b == 0 || c == 0) {
      g_err = "Slither triangles or triangles with no area exist on your surface";
      return 3;
    }
    for (int k = 0; k < 3; ++k) mid[k * n + i] = ((Pa[k] * a) + (Pb[k] * b) + (Pc[k] * c)) / (a + b + c);
  }
  return 0;
}

// Single angular-dislocation closed forms, exposed so tests can pin them one by one.
void orc_ang_dis_strain_fsc(double y1, double y2, double y3, double beta, double nu, double a, double* out18) {
  orc::ang_dis_strain_fsc(y1, y2, y3, std::cos(beta), std::sin(beta), orc::cot(beta), nu, a, out18);
}
void orc_ang_dis_disp_fsc(double y1, double y2, double y3, double beta, double nu, double a, double* out9) {
  orc::ang_dis_disp_fsc(y1, y2, y3, std::cos(beta), std::sin(beta), orc::cot(beta), orc::cot(beta / 2), nu, a, out9);
}
void orc_ang_dis_strain(double x, double y, double z, double alpha, double nu, double* out21) {
  const double Ang = -orc::kPi + alpha;
  orc::ang_dis_strain(x, y, z, std::cos(Ang), std::sin(Ang), nu, 2.0 * nu + 1.0, out21);
}
void orc_ang_dis_disp(double x, double y, double z, double alpha, double nu, double* out9) {
  const double Ang = -orc::kPi + alpha;
  const double cosA = std::cos(Ang), sinA = std::sin(Ang);
  orc::ang_dis_disp(x, y, z, cosA, sinA, 1.0 - nu, 1.0 - 2.0 * nu, cosA * cosA, sinA / 8.0 / orc::kPi / (1.0 - nu), out9);
}

}  // extern "C"
