// Univariate minimisation on an interval (host): the two methods SlipCalculator3D.jl's fluid-volume solve takes
// from Optim 1.11.0 (Brent() and GoldenSection(); SlipCalculator3D.jl:404, 431).  Optim is not vendored by the
// reference; these are the published methods: Brent (1973, "Algorithms for minimization without derivatives",
// ch. 5) parabolic interpolation with golden-section fallback, and golden-section search, both stopping when
// |x - midpoint| <= 2 tol - (upper - lower) / 2 with tol = rel_tol |x| + abs_tol.
#pragma once
#include <cmath>
#include <functional>

namespace cdj {

struct Minimum {
  double x = 0.0, f = 0.0;
  int calls = 0;
  bool ok = true;  // false when an evaluation reported failure (the search stops there)
};

constexpr double kGolden = 0.3819660112501051;  // (3 - sqrt(5)) / 2

// f returns false to abort the search (its value is then ignored)
inline Minimum brent_minimize(const std::function<bool(double, double&)>& f, double lo, double hi, double abs_tol,
                              double rel_tol = 1.4901161193847656e-08, int max_iter = 1000) {
  Minimum m;
  double x = lo + kGolden * (hi - lo), fx;
  if (!f(x, fx)) return Minimum{x, 0.0, 1, false};
  int calls = 1;
  double step = 0.0, old_step = 0.0, v = x, w = x, fv = fx, fw = fx;
  for (int it = 0; it < max_iter; ++it) {
    double p = 0.0, q = 0.0;
    const double tol = rel_tol * std::fabs(x) + abs_tol, mid = (hi + lo) / 2;
    if (std::fabs(x - mid) <= 2 * tol - (hi - lo) / 2) break;
    if (std::fabs(old_step) > tol) {  // parabola through (x, fx), (w, fw), (v, fv): its vertex is x + p / q
      const double r = (x - w) * (fx - fv);
      q = (x - v) * (fx - fw);
      p = (x - v) * q - (x - w) * r;
      q = 2 * (q - r);
      if (q > 0)
        p = -p;
      else
        q = -q;
    }
    if (std::fabs(p) < std::fabs(q * old_step / 2) && p > q * (lo - x) && p < q * (hi - x)) {
      old_step = step;
      step = p / q;
      if ((x + step - lo) < 2 * tol || (hi - (x + step)) < 2 * tol) step = x < mid ? tol : -tol;
    } else {
      old_step = x < mid ? hi - x : lo - x;
      step = kGolden * old_step;
    }
    const double u = std::fabs(step) >= tol ? x + step : x + (step > 0 ? tol : -tol);
    double fu;
    ++calls;
    if (!f(u, fu)) return Minimum{x, fx, calls, false};
    if (fu < fx) {
      if (u < x)
        hi = x;
      else
        lo = x;
      v = w;
      fv = fw;
      w = x;
      fw = fx;
      x = u;
      fx = fu;
    } else {
      if (u < x)
        lo = u;
      else
        hi = u;
      if (fu <= fw || w == x) {
        v = w;
        fv = fw;
        w = u;
        fw = fu;
      } else if (fu <= fv || v == x || v == w) {
        v = u;
        fv = fu;
      }
    }
  }
  m.x = x;
  m.f = fx;
  m.calls = calls;
  return m;
}

inline Minimum golden_section_minimize(const std::function<bool(double, double&)>& f, double lo, double hi,
                                       double abs_tol = 2.220446049250313e-16,
                                       double rel_tol = 1.4901161193847656e-08, int max_iter = 1000) {
  Minimum m;
  double x = lo + kGolden * (hi - lo), fx;
  if (!f(x, fx)) return Minimum{x, 0.0, 1, false};
  int calls = 1;
  for (int it = 0; it < max_iter; ++it) {
    const double tol = rel_tol * std::fabs(x) + abs_tol;
    if (std::fabs(x - (hi + lo) / 2) <= 2 * tol - (hi - lo) / 2) break;
    const bool right = hi - x > x - lo;
    const double u = right ? x + kGolden * (hi - x) : x - kGolden * (x - lo);
    double fu;
    ++calls;
    if (!f(u, fu)) return Minimum{x, fx, calls, false};
    if (fu < fx) {
      if (right)
        lo = x;
      else
        hi = x;
      x = u;
      fx = fu;
    } else if (right) {
      hi = u;
    } else {
      lo = u;
    }
  }
  m.x = x;
  m.f = fx;
  m.calls = calls;
  return m;
}

}  // namespace cdj
