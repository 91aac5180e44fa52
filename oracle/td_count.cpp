// Algorithmic operation counts of the TD path (test/bench infrastructure only).
//
// The oracle's templates are instantiated with a scalar type that counts every arithmetic operation it
// performs: +, -, *, / each count 1 (no fused multiply-add exists in the oracle build), and sqrt, log,
// atan/atan2, acos, sin/cos/tan are counted separately.  Counting runs the real branch decisions
// (configuration I/II, skipped sides) on the real inputs, so per-pair averages follow the workload.
// Also: the traction assembly restricted to a block of receiver rows, used as the bounded CPU baseline.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace orc {
struct Cnt;
}
namespace orc {

struct Counts {
  double add = 0, mul = 0, div = 0, sqrt = 0, log = 0, atan = 0, trig = 0, cmp = 0;
};
static thread_local Counts* g_cnt = nullptr;

struct Cnt {
  double v;
  Cnt() : v(0) {}
  Cnt(double x) : v(x) {}
  explicit operator double() const { return v; }
};
static inline void add1() { if (g_cnt) g_cnt->add += 1; }
static inline void mul1() { if (g_cnt) g_cnt->mul += 1; }
static inline void div1() { if (g_cnt) g_cnt->div += 1; }
inline Cnt operator+(Cnt a, Cnt b) { add1(); return Cnt(a.v + b.v); }
inline Cnt operator-(Cnt a, Cnt b) { add1(); return Cnt(a.v - b.v); }
inline Cnt operator*(Cnt a, Cnt b) { mul1(); return Cnt(a.v * b.v); }
inline Cnt operator/(Cnt a, Cnt b) { div1(); return Cnt(a.v / b.v); }
inline Cnt operator-(Cnt a) { return Cnt(-a.v); }
inline Cnt& operator+=(Cnt& a, Cnt b) { a = a + b; return a; }
inline bool operator<(Cnt a, Cnt b) { return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { return a.v <= b.v; }
inline bool operator>=(Cnt a, Cnt b) { return a.v >= b.v; }
inline bool operator==(Cnt a, Cnt b) { return a.v == b.v; }
inline bool operator!=(Cnt a, Cnt b) { return a.v != b.v; }
inline Cnt sqrt(Cnt a) { if (g_cnt) g_cnt->sqrt += 1; return Cnt(std::sqrt(a.v)); }
inline Cnt log(Cnt a) { if (g_cnt) g_cnt->log += 1; return Cnt(std::log(a.v)); }
inline Cnt atan(Cnt a) { if (g_cnt) g_cnt->atan += 1; return Cnt(std::atan(a.v)); }
inline Cnt atan2(Cnt a, Cnt b) { if (g_cnt) g_cnt->atan += 1; return Cnt(std::atan2(a.v, b.v)); }
inline Cnt acos(Cnt a) { if (g_cnt) g_cnt->trig += 1; return Cnt(std::acos(a.v)); }
inline Cnt cos(Cnt a) { if (g_cnt) g_cnt->trig += 1; return Cnt(std::cos(a.v)); }
inline Cnt sin(Cnt a) { if (g_cnt) g_cnt->trig += 1; return Cnt(std::sin(a.v)); }
inline Cnt tan(Cnt a) { if (g_cnt) g_cnt->trig += 1; return Cnt(std::tan(a.v)); }
inline Cnt fabs(Cnt a) { return Cnt(std::fabs(a.v)); }

}  // namespace orc

namespace std {
template <> class numeric_limits<orc::Cnt> : public numeric_limits<double> {};
}  // namespace std

#include "td_oracle.hpp"

namespace {

template <class T>
inline void strain_to_traction_t(const T* e, T lam, T G, const T* nv, const T* s, const T* d, T* out3) {
  const T G2 = 2.0 * G;
  const T cons = lam * (e[0] + e[1] + e[2]);
  const T Sxx = G2 * e[0] + cons, Syy = G2 * e[1] + cons, Szz = G2 * e[2] + cons;
  const T Sxy = G2 * e[3], Sxz = G2 * e[4], Syz = G2 * e[5];
  const T nx = nv[0], ny = nv[1], nz = nv[2];
  out3[0] = Sxx * (nx * nx) + Syy * (ny * ny) + Szz * (nz * nz) + 2 * Sxy * (nx * ny) + 2 * Syz * (ny * nz) +
            2 * Sxz * (nz * nx);
  const T Tx = Sxx * nx + Sxy * ny + Sxz * nz;
  const T Ty = Sxy * nx + Syy * ny + Syz * nz;
  const T Tz = Sxz * nx + Syz * ny + Szz * nz;
  const T a2 = 1 - (nx * nx), b2 = 1 - (ny * ny), c2 = 1 - (nz * nz);
  const T xy = nx * ny, yz = ny * nz, zx = nz * nx;
  const T* dirs[2] = {s, d};
  for (int q = 0; q < 2; ++q) {
    const T* c = dirs[q];
    out3[1 + q] = (Tx * a2 - Ty * xy - Tz * zx) * c[0] + (Tx * -xy + Ty * b2 - Tz * yz) * c[1] +
                  (Tx * -zx - Ty * yz + Tz * c2) * c[2];
  }
}

}  // namespace

extern "C" void orc_ss_ds_dirs(const double* nrm, int64_t n, double* ss, double* ds);

extern "C" {

// Operation counts of the traction assembly (collocation point of receiver r, source triangle s) summed over
// npairs pairs.  out[8] = {add, mul, div, sqrt, log, atan, trig, pairs}.  Collocation points follow
// SetupCollationPoints (CalculateInfluenceMatrices3D.jl:52-71).
void orc_count_traction_flops(const double* normals, const double* mid, const double* P1, const double* P2,
                              const double* P3, int64_t n, double nu, double G, double lam, int hs_flag,
                              const int64_t* recv_idx, const int64_t* src_idx, int64_t npairs, int setup_too,
                              double* out) {
  using orc::Cnt;
  std::vector<double> ss(3 * n), ds(3 * n);
  orc_ss_ds_dirs(normals, n, ss.data(), ds.data());
  orc::Counts c;
  const bool hs = hs_flag == 1;
  for (int64_t p = 0; p < npairs; ++p) {
    const int64_t i = recv_idx[p], j = src_idx[p];
    Cnt a[3], b[3], cc[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = P1[k * n + j];
      b[k] = P2[k * n + j];
      cc[k] = P3[k * n + j];
    }
    orc::TriFull<Cnt> f;
    orc::g_cnt = setup_too ? &c : nullptr;
    orc::tri_full_setup(a, b, cc, Cnt(1.0), Cnt(1.0), Cnt(1.0), hs, f);
    orc::g_cnt = &c;
    const Cnt X = mid[i] - normals[i] * 1e-12, Y = mid[n + i] - normals[n + i] * 1e-12,
              Z = mid[2 * n + i] - normals[2 * n + i] * 1e-12;
    Cnt e[3][6];
    orc::strain_influence(f, hs, X, Y, Z, Cnt(1.0), Cnt(1.0), Cnt(1.0), Cnt(nu), e);
    const Cnt nv[3] = {normals[i], normals[n + i], normals[2 * n + i]};
    const Cnt sv[3] = {ss[i], ss[n + i], ss[2 * n + i]};
    const Cnt dv[3] = {ds[i], ds[n + i], ds[2 * n + i]};
    for (int m = 0; m < 3; ++m) {
      Cnt t[3];
      strain_to_traction_t<Cnt>(e[m], Cnt(lam), Cnt(G), nv, sv, dv, t);
    }
    orc::g_cnt = nullptr;
  }
  out[0] = c.add;
  out[1] = c.mul;
  out[2] = c.div;
  out[3] = c.sqrt;
  out[4] = c.log;
  out[5] = c.atan;
  out[6] = c.trig;
  out[7] = (double)npairs;
}

// Operation counts of the free-surface strain of single triangle sides (AngSetupStrainFSC TD.jl:1867-2081 for the
// three slip modes): side k of triangle j (side_idx = k n + j; sides P1P2, P2P3, P3P1) at the collocation point of
// receiver recv_idx, summed over npairs.  out[8] as orc_count_traction_flops, out[7] = npairs.
void orc_count_fsc_side_flops(const double* normals, const double* mid, const double* P1, const double* P2,
                              const double* P3, int64_t n, double nu, const int64_t* recv_idx, const int64_t* side_idx,
                              int64_t npairs, double* out) {
  using orc::Cnt;
  orc::Counts c;
  for (int64_t p = 0; p < npairs; ++p) {
    const int64_t i = recv_idx[p], j = side_idx[p] % n, k = side_idx[p] / n;
    Cnt a[3], b[3], cc[3];
    for (int q = 0; q < 3; ++q) {
      a[q] = P1[q * n + j];
      b[q] = P2[q * n + j];
      cc[q] = P3[q * n + j];
    }
    orc::TriFull<Cnt> f;
    orc::g_cnt = nullptr;
    orc::tri_full_setup(a, b, cc, Cnt(1.0), Cnt(1.0), Cnt(1.0), true, f);
    orc::g_cnt = &c;
    const Cnt X = mid[i] - normals[i] * 1e-12, Y = mid[n + i] - normals[n + i] * 1e-12,
              Z = mid[2 * n + i] - normals[2 * n + i] * 1e-12;
    Cnt e[3][6];
    for (int m = 0; m < 3; ++m)
      for (int q = 0; q < 6; ++q) e[m][q] = Cnt(0.0);
    orc::strain_fsc_side(f.side[k], X, Y, Z, Cnt(1.0), Cnt(1.0), Cnt(1.0), Cnt(nu), e);
    orc::g_cnt = nullptr;
  }
  out[0] = c.add;
  out[1] = c.mul;
  out[2] = c.div;
  out[3] = c.sqrt;
  out[4] = c.log;
  out[5] = c.atan;
  out[6] = c.trig;
  out[7] = (double)npairs;
}

// Traction assembly of receiver rows [r0, r1) against all n sources (the reference's CPU path for a block of
// rows; OpenMP over sources as TD.jl:146).  T9: 9 matrices (r1 - r0) x n, TractionInf order.
int orc_traction_influence_rows(const double* normals, const double* mid, const double* P1, const double* P2,
                                const double* P3, int64_t n, double nu, double G, double lam, int hs_flag, int64_t r0,
                                int64_t r1, double* const* T9, int nthreads) {
  std::vector<double> ss(3 * n), ds(3 * n);
  orc_ss_ds_dirs(normals, n, ss.data(), ds.data());
  const bool hs = hs_flag == 1;
  const int64_t nr = r1 - r0;
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t j = 0; j < n; ++j) {
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; ++k) {
      a[k] = P1[k * n + j];
      b[k] = P2[k * n + j];
      c[k] = P3[k * n + j];
    }
    orc::TriFull<double> f;
    orc::tri_full_setup(a, b, c, 1.0, 1.0, 1.0, hs, f);
    for (int64_t ii = 0; ii < nr; ++ii) {
      const int64_t i = r0 + ii;
      const double X = mid[i] - (normals[i] * 1e-12), Y = mid[n + i] - (normals[n + i] * 1e-12),
                   Z = mid[2 * n + i] - (normals[2 * n + i] * 1e-12);
      double e[3][6];
      orc::strain_influence(f, hs, X, Y, Z, 1.0, 1.0, 1.0, nu, e);
      const double nv[3] = {normals[i], normals[n + i], normals[2 * n + i]};
      const double sv[3] = {ss[i], ss[n + i], ss[2 * n + i]};
      const double dv[3] = {ds[i], ds[n + i], ds[2 * n + i]};
      const int64_t o = ii + j * nr;
      for (int m = 0; m < 3; ++m) {
        double t[3];
        strain_to_traction_t<double>(e[m], lam, G, nv, sv, dv, t);
        T9[3 * m][o] = t[0];
        T9[3 * m + 1][o] = t[1];
        T9[3 * m + 2][o] = t[2];
      }
    }
  }
  return 0;
}

}  // extern "C"
