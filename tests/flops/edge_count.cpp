// Operation counts of one half-space edge-response unit (one receiver, one mesh edge; the unit of the kernel
// tdk_edge_response) in the engine's forward-mode formulation.  The device header td_device.hpp is compiled for
// the host and its templates are instantiated with a scalar that counts +, -, *, / (one each) and, separately,
// sqrt, log and atan, leaving out what the compiler folds (see opc::C).  The free-surface code has no
// data-dependent branches (the configuration choice is a select), so the count of a unit that is not skipped is
// the same for every receiver and edge.
//
// Printed (JSON):
//   wk[K]    fsc_side_wk<K>: one ADCS Burgers component, the side's shared quantities formed in the call
//   traction edge_comp_traction for the three ADCS components and edge_unit_combine for the three EFCS unit
//            Burgers vectors: the unit's tractions from the components
//   shared   the part of fsc_side_wk that does not depend on K: the side frame and, per angular dislocation, the
//            point's ADCS offsets, FscBase and FscDer
//   single   all three components with the shared part formed once (the formulation's work for the unit, each
//            quantity evaluated once); wk[0] + wk[1] + wk[2] = single + 2 shared is checked by the caller
//   live_wk[K], live_single
//            the same passes counting only the operations whose results reach the pass's outputs (the strain
//            components w): the kernel needs the gradients of the free-surface displacement, not its values, so
//            the logarithms, arctangents and value chains that feed only displacement values are dead, and the
//            compiler removes them
//   live_traction  the traction part counted the same way (all of it is live)
// Test infrastructure (tests/test_flops.py); not part of the library.
#include <cmath>
#include <cstdio>
#include <vector>

namespace opc {
struct Counts {
  double add = 0, mul = 0, div = 0, sqrt = 0, log = 0, atan = 0;
};
static Counts* g = nullptr;
// Optional tape of the counted operations (kind, operand nodes; -1 for data and literals), for liveness.
struct Node {
  char kind;  // '+', '*', '/', 's' sqrt, 'l' log, 't' atan
  int a, b;
};
static std::vector<Node>* tape = nullptr;
// k marks a value known when the code is compiled (a literal, or arithmetic on literals only).  Operations whose
// result the compiler folds are not counted: arithmetic on two known values, a product with a known 0, a sum with
// a known 0, a product or quotient with a known +-1 (a sign change is a free source modifier).  The derivative
// seeds of the forward-mode values are such literals, so their first products fold.
struct C {
  double v;
  bool k;
  int n = -1;  // tape node that produced the value, -1 for data and literals
  C() : v(0), k(true) {}
  C(double x) : v(x), k(true) {}
};
inline C dat(double x) { C c(x); c.k = false; return c; }
inline C res(double x, bool known) { C c(x); c.k = known; return c; }
// an unknown result of a counted operation, recorded on the tape when one is active
inline C op(double x, char kind, const C& a, const C& b) {
  C c(x);
  c.k = false;
  if (tape) {
    tape->push_back({kind, a.n, b.n});
    c.n = (int)tape->size() - 1;
  }
  return c;
}
inline C neg(const C& a) { C c = res(-a.v, a.k); c.n = a.n; return c; }
inline bool zero(C a) { return a.k && a.v == 0.0; }
inline bool unit(C a) { return a.k && (a.v == 1.0 || a.v == -1.0); }
inline C operator-(C a) { return neg(a); }
inline C operator+(C a, C b) {
  if (a.k && b.k) return C(a.v + b.v);
  if (zero(a)) return b;
  if (zero(b)) return a;
  if (g) g->add++;
  return op(a.v + b.v, '+', a, b);
}
inline C operator-(C a, C b) { return a + (-b); }
inline C operator*(C a, C b) {
  if (a.k && b.k) return C(a.v * b.v);
  if (zero(a) || zero(b)) return C(0.0);
  if (unit(a)) return a.v > 0 ? b : -b;
  if (unit(b)) return b.v > 0 ? a : -a;
  if (g) g->mul++;
  return op(a.v * b.v, '*', a, b);
}
inline C operator/(C a, C b) {
  if (a.k && b.k) return C(a.v / b.v);
  if (zero(a)) return C(0.0);
  if (unit(b)) return b.v > 0 ? a : -a;
  if (g) g->div++;
  return op(a.v / b.v, '/', a, b);
}
inline C& operator+=(C& a, C b) { a = a + b; return a; }
inline bool operator>=(C a, C b) { return a.v >= b.v; }
inline bool operator<=(C a, C b) { return a.v <= b.v; }
inline bool operator<(C a, C b) { return a.v < b.v; }
inline bool operator>(C a, C b) { return a.v > b.v; }
inline bool operator==(C a, C b) { return a.v == b.v; }
inline bool operator!=(C a, C b) { return a.v != b.v; }
inline C sqrt(C a) { if (a.k) return C(std::sqrt(a.v)); if (g) g->sqrt++; return op(std::sqrt(a.v), 's', a, C()); }
inline C log(C a) { if (a.k) return C(std::log(a.v)); if (g) g->log++; return op(std::log(a.v), 'l', a, C()); }
inline C atan(C a) { if (a.k) return C(std::atan(a.v)); if (g) g->atan++; return op(std::atan(a.v), 't', a, C()); }
// operation counts of the tape nodes that the roots depend on
inline Counts live(const std::vector<Node>& t, const std::vector<int>& roots) {
  std::vector<char> mark(t.size(), 0);
  std::vector<int> stack;
  for (int r : roots)
    if (r >= 0) stack.push_back(r);
  while (!stack.empty()) {
    const int k = stack.back();
    stack.pop_back();
    if (mark[k]) continue;
    mark[k] = 1;
    if (t[k].a >= 0) stack.push_back(t[k].a);
    if (t[k].b >= 0) stack.push_back(t[k].b);
  }
  Counts c;
  for (size_t k = 0; k < t.size(); ++k) {
    if (!mark[k]) continue;
    switch (t[k].kind) {
      case '+': c.add++; break;
      case '*': c.mul++; break;
      case '/': c.div++; break;
      case 's': c.sqrt++; break;
      case 'l': c.log++; break;
      default: c.atan++; break;
    }
  }
  return c;
}
}  // namespace opc

#include "../../cutanddisplacejulia_amd/csrc/td_device.hpp"

using opc::C;
using namespace cdj;

// The frame and, per angular dislocation, offsets, FscBase and FscDer: the statements of fsc_side_wk before its
// output loop, in the same order.  With outputs, every component is added once (the single-pass formulation).
static void side_pass(const C* s, C X, C Y, C Z, C nu, bool outputs, C w[3][6]) {
  C x, y, z;
  {
    const C dX = X - s[S_PA + 0], dY = Y - s[S_PA + 1], dZ = Z - s[S_PA + 2];
    const C p = s[S_EY1 + 0], q = s[S_EY1 + 1];
    x = (p * dX) + (q * dY) + (C(-0.0) * dZ);
    y = (q * dX) + ((-p) * dY) + (C(-0.0) * dZ);
    z = (C(0) * dX) + (C(0) * dY) + (C(-1) * dZ);
  }
  const bool cfg1 = x >= C(0);
  const C sg = cfg1 ? C(-1) : C(1);
  const C cB = cfg1 ? s[S_COSPS] : s[S_COSB];
  const C sB = cfg1 ? s[S_SINPS] : s[S_SINB];
  const C ctB = cfg1 ? s[S_COTPS] : s[S_COTB];
  const C ctB2 = cfg1 ? s[S_COTPS2] : s[S_COTB2];
  const C icB = C(1) / cB;
  for (int ad = 0; ad < 2; ++ad) {
    const C px = ad == 0 ? x : x - s[S_YB + 0];
    const C py = ad == 0 ? y : y - s[S_YB + 1];
    const C pz = ad == 0 ? z : z - s[S_YB + 2];
    const C aa = ad == 0 ? s[S_AA] : s[S_AB];
    const C sgn = ad == 0 ? C(-1) : C(1);
    FscBase<C, D3<C>> fb;
    fsc_base<C, D3<C>>(DS<C, 1>{sg * px, C(1), C(0), C(0)}, DS<C, 2>{sg * py, C(0), C(1), C(0)},
                       DS<C, 4>{pz, C(0), C(0), C(1)}, cB, sB, ctB, ctB2, icB, nu, aa, fb);
    const FscDer<C, D3<C>> dq = fsc_der<C, D3<C>>(fb);
    if (!outputs) continue;
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) {
        const D3<C> u = fsc_out_d<C, D3<C>>(fb, dq, 3 * c + k);
        const C g0 = sgn * u.a, g1 = sgn * u.b, g2 = sgn * u.c;
        C* wk = w[k];
        if (c == 0) {
          wk[0] += g0;
          wk[3] += g1;
          wk[4] += sg * g2;
        } else if (c == 1) {
          wk[3] += g0;
          wk[1] += g1;
          wk[5] += sg * g2;
        } else {
          wk[4] += sg * g0;
          wk[5] += sg * g1;
          wk[2] += g2;
        }
      }
  }
}

static void print(const char* name, const opc::Counts& c, bool last = false) {
  std::printf("\"%s\": {\"add\": %.0f, \"mul\": %.0f, \"div\": %.0f, \"sqrt\": %.0f, \"log\": %.0f, \"atan\": %.0f}%s\n",
              name, c.add, c.mul, c.div, c.sqrt, c.log, c.atan, last ? "" : ",");
}

int main() {
  // a side record with plausible values (the count does not depend on them)
  C s[R_SIDE_SIZE];
  for (int k = 0; k < R_SIDE_SIZE; ++k) s[k] = opc::dat(0.3 + 0.01 * k);
  s[S_SKIP] = opc::dat(0.0);
  const C X = opc::dat(1.0), Y = opc::dat(2.0), Z = opc::dat(-3.0), nu = opc::dat(0.25);
  std::printf("{\n");
  for (int K = 0; K < 3; ++K) {
    opc::Counts c;
    opc::g = &c;
    C w[6], sg;
    if (K == 0) fsc_side_wk<0, C>(s, X, Y, Z, nu, w, sg);
    if (K == 1) fsc_side_wk<1, C>(s, X, Y, Z, nu, w, sg);
    if (K == 2) fsc_side_wk<2, C>(s, X, Y, Z, nu, w, sg);
    opc::g = nullptr;
    char name[8];
    std::snprintf(name, sizeof(name), "wk%d", K);
    print(name, c);
  }
  {
    C w0[6], w1[6], w2[6], n[3], sv[3], d[3];
    for (int c = 0; c < 6; ++c) w0[c] = w1[c] = w2[c] = opc::dat(0.1 * c + 0.05);
    for (int c = 0; c < 3; ++c) {
      n[c] = opc::dat(0.3 + 0.1 * c);
      sv[c] = opc::dat(0.5 - 0.1 * c);
      d[c] = opc::dat(0.2 + 0.2 * c);
    }
    opc::Counts c;
    opc::g = &c;
    C t[3][3], o0, o1, o2;
    const C* w[3] = {w0, w1, w2};
    for (int K = 0; K < 3; ++K)
      edge_comp_traction<C>(s, w[K], opc::dat(0.4), opc::dat(1.0), n, sv, d, t[K][0], t[K][1], t[K][2]);
    for (int kE = 0; kE < 3; ++kE)
      edge_unit_combine<C>(s, kE, opc::dat(-1.0), t[0], t[1], t[2], 1, opc::dat(0.1), o0, o1, o2);
    opc::g = nullptr;
    print("traction", c);
  }
  {
    opc::Counts c;
    opc::g = &c;
    C w[3][6];
    side_pass(s, X, Y, Z, nu, false, w);
    opc::g = nullptr;
    print("shared", c);
  }
  {
    opc::Counts c;
    C w[3][6];
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 6; ++q) w[k][q] = C(0);
    opc::g = &c;
    side_pass(s, X, Y, Z, nu, true, w);
    opc::g = nullptr;
    print("single", c);
  }
  // live counts
  for (int K = 0; K < 3; ++K) {
    std::vector<opc::Node> t;
    opc::tape = &t;
    C w[6], sg;
    if (K == 0) fsc_side_wk<0, C>(s, X, Y, Z, nu, w, sg);
    if (K == 1) fsc_side_wk<1, C>(s, X, Y, Z, nu, w, sg);
    if (K == 2) fsc_side_wk<2, C>(s, X, Y, Z, nu, w, sg);
    opc::tape = nullptr;
    std::vector<int> roots;
    for (int c = 0; c < 6; ++c) roots.push_back(w[c].n);
    roots.push_back(sg.n);
    char name[16];
    std::snprintf(name, sizeof(name), "live_wk%d", K);
    print(name, opc::live(t, roots));
  }
  {
    std::vector<opc::Node> t;
    C w[3][6];
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 6; ++q) w[k][q] = C(0);
    opc::tape = &t;
    side_pass(s, X, Y, Z, nu, true, w);
    opc::tape = nullptr;
    std::vector<int> roots;
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < 6; ++q) roots.push_back(w[k][q].n);
    print("live_single", opc::live(t, roots));
  }
  {
    std::vector<opc::Node> t;
    C w0[6], w1[6], w2[6], n[3], sv[3], d[3];
    for (int c = 0; c < 6; ++c) w0[c] = w1[c] = w2[c] = opc::dat(0.1 * c + 0.05);
    for (int c = 0; c < 3; ++c) {
      n[c] = opc::dat(0.3 + 0.1 * c);
      sv[c] = opc::dat(0.5 - 0.1 * c);
      d[c] = opc::dat(0.2 + 0.2 * c);
    }
    opc::tape = &t;
    C tt[3][3], o[3][3];
    const C* w[3] = {w0, w1, w2};
    for (int K = 0; K < 3; ++K)
      edge_comp_traction<C>(s, w[K], opc::dat(0.4), opc::dat(1.0), n, sv, d, tt[K][0], tt[K][1], tt[K][2]);
    for (int kE = 0; kE < 3; ++kE)
      edge_unit_combine<C>(s, kE, opc::dat(-1.0), tt[0], tt[1], tt[2], 1, opc::dat(0.1), o[kE][0], o[kE][1], o[kE][2]);
    opc::tape = nullptr;
    std::vector<int> roots;
    for (int kE = 0; kE < 3; ++kE)
      for (int q = 0; q < 3; ++q) roots.push_back(o[kE][q].n);
    print("live_traction", opc::live(t, roots), true);
  }
  std::printf("}\n");
  return 0;
}
