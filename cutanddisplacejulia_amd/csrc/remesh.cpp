// Isotropic remeshing of triangulated crack and fault surfaces, the step the reference's growth loops hand to CGAL
// (BuildCGAL.PolygonRemeshingCGAL, CGAL 4.13.1 Polygon_mesh_processing::isotropic_remeshing; called at
// test/test_PropagationTest.jl:151-561 between OFFExport and OFFReader).  CGAL is not part of the reference, so this
// is a restatement of the published algorithm it implements (Botsch & Kobbelt 2004, "A remeshing approach to
// multiresolution modeling"), host code:
//
//   repeat `iterations` times, with L the target edge length:
//     split every edge longer than 4/3 L at its midpoint (until none is),
//     collapse every edge shorter than 4/5 L whose collapse creates no edge longer than 4/3 L, keeps the mesh
//       manifold (link condition) and turns no triangle over,
//     flip interior edges where that brings the four vertices' valences closer to 6 (4 on the boundary),
//     move every interior vertex to the area-weighted centroid of its triangles, within its tangent plane,
//     project the interior vertices back onto the input surface (closest point, through a bounding-volume tree).
//
// The boundary (a crack's tip line) keeps its shape: a boundary edge is split at the point of the input's boundary
// polyline nearest its midpoint, collapsed only along itself onto one of its end points (the corner, if one is),
// never flipped, and boundary vertices are not relaxed, so every boundary vertex stays on the input's boundary
// polyline and corners (turning by more than 45 degrees) stay where they are.  Triangle orientation is preserved.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <unordered_map>
#include <vector>

namespace cdj {
namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm2(V3 a) { return dot(a, a); }

typedef std::array<int, 3> Tri;

inline uint64_t ekey(int a, int b) {
  const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
  return ((uint64_t)lo << 32) | hi;
}

// ---- closest point on the input surface ---------------------------------------------------------------------
// Closest point of triangle (a, b, c) to p, by the Voronoi regions of its vertices, edges and face (Ericson,
// Real-Time Collision Detection, 2005, section 5.1.5).
V3 closest_on_triangle(V3 p, V3 a, V3 b, V3 c) {
  const V3 ab = b - a, ac = c - a, ap = p - a;
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0 && d2 <= 0) return a;
  const V3 bp = p - b;
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0 && d4 <= d3) return b;
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) return a + (d1 / (d1 - d3)) * ab;
  const V3 cp = p - c;
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0 && d5 <= d6) return c;
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) return a + (d2 / (d2 - d6)) * ac;
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) return b + ((d4 - d3) / ((d4 - d3) + (d5 - d6))) * (c - b);
  const double den = 1.0 / (va + vb + vc);
  return a + (vb * den) * ab + (vc * den) * ac;
}

V3 closest_on_segment(V3 p, V3 a, V3 b) {
  const V3 ab = b - a;
  const double l2 = norm2(ab);
  if (!(l2 > 0)) return a;
  const double t = std::min(1.0, std::max(0.0, dot(p - a, ab) / l2));
  return a + t * ab;
}

// Bounding-volume tree over the input triangles, or over segments given as triangles (a, b, b) (median splits on
// the longest axis, leaves of up to 8).
class Bvh {
 public:
  Bvh(const std::vector<V3>& p, const std::vector<Tri>& t, bool segments = false)
      : p_(p), t_(t), seg_(segments) {
    idx_.resize(t.size());
    for (size_t i = 0; i < t.size(); ++i) idx_[i] = (int)i;
    cen_.resize(t.size());
    for (size_t i = 0; i < t.size(); ++i) cen_[i] = (1.0 / 3.0) * (p[t[i][0]] + p[t[i][1]] + p[t[i][2]]);
    if (!t.empty()) build(0, (int)t.size());
  }
  V3 closest(V3 q) const {
    double best = std::numeric_limits<double>::infinity();
    V3 bp = q;
    if (nodes_.empty()) return q;
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node& nd = nodes_[stack[--sp]];
      if (box_d2(nd, q) >= best) continue;
      if (nd.left < 0) {
        for (int i = nd.i0; i < nd.i1; ++i) {
          const Tri& f = t_[idx_[i]];
          const V3 c = seg_ ? closest_on_segment(q, p_[f[0]], p_[f[1]])
                            : closest_on_triangle(q, p_[f[0]], p_[f[1]], p_[f[2]]);
          const double d = norm2(c - q);
          if (d < best) {
            best = d;
            bp = c;
          }
        }
      } else if (sp + 2 <= 128) {
        const Node &l = nodes_[nd.left], &r = nodes_[nd.right];
        // nearer child last, so it is visited first
        if (box_d2(l, q) < box_d2(r, q)) {
          stack[sp++] = nd.right;
          stack[sp++] = nd.left;
        } else {
          stack[sp++] = nd.left;
          stack[sp++] = nd.right;
        }
      }
    }
    return bp;
  }

 private:
  struct Node {
    V3 lo, hi;
    int left, right, i0, i1;
  };
  static double box_d2(const Node& n, V3 q) {
    const double dx = std::max({n.lo.x - q.x, 0.0, q.x - n.hi.x});
    const double dy = std::max({n.lo.y - q.y, 0.0, q.y - n.hi.y});
    const double dz = std::max({n.lo.z - q.z, 0.0, q.z - n.hi.z});
    return dx * dx + dy * dy + dz * dz;
  }
  int build(int i0, int i1) {
    Node nd;
    const double inf = std::numeric_limits<double>::infinity();
    nd.lo = {inf, inf, inf};
    nd.hi = {-inf, -inf, -inf};
    V3 clo = nd.lo, chi = nd.hi;
    for (int i = i0; i < i1; ++i) {
      for (int k = 0; k < 3; ++k) {
        const V3 v = p_[t_[idx_[i]][k]];
        nd.lo = {std::min(nd.lo.x, v.x), std::min(nd.lo.y, v.y), std::min(nd.lo.z, v.z)};
        nd.hi = {std::max(nd.hi.x, v.x), std::max(nd.hi.y, v.y), std::max(nd.hi.z, v.z)};
      }
      const V3 c = cen_[idx_[i]];
      clo = {std::min(clo.x, c.x), std::min(clo.y, c.y), std::min(clo.z, c.z)};
      chi = {std::max(chi.x, c.x), std::max(chi.y, c.y), std::max(chi.z, c.z)};
    }
    nd.i0 = i0;
    nd.i1 = i1;
    nd.left = nd.right = -1;
    const int me = (int)nodes_.size();
    nodes_.push_back(nd);
    if (i1 - i0 > 8) {
      const V3 ext = chi - clo;
      const int ax = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
      const int mid = (i0 + i1) / 2;
      auto coord = [&](int i) { const V3 c = cen_[i]; return ax == 0 ? c.x : (ax == 1 ? c.y : c.z); };
      std::nth_element(idx_.begin() + i0, idx_.begin() + mid, idx_.begin() + i1,
                       [&](int a, int b) { return coord(a) < coord(b); });
      const int l = build(i0, mid);
      const int r = build(mid, i1);
      nodes_[me].left = l;
      nodes_[me].right = r;
    }
    return me;
  }
  const std::vector<V3>& p_;
  const std::vector<Tri>& t_;
  const bool seg_;
  std::vector<int> idx_;
  std::vector<V3> cen_;
  std::vector<Node> nodes_;
};

// ---- the working mesh and its connectivity --------------------------------------------------------------------
struct Edge {
  int a, b;       // end points
  int f0, f1;     // faces (f1 = -1 on the boundary); f0 holds the half-edge a -> b
  int k0, k1;     // local index of the half-edge in f0 / f1
  bool manifold;  // at most two faces
};

struct Mesh {
  std::vector<V3> p;
  std::vector<Tri> f;
  std::vector<char> fdead, vdead;
};

struct Topo {
  std::vector<Edge> e;
  std::unordered_map<uint64_t, int> emap;
  std::vector<int> voff, vfaces;  // incident faces per vertex (CSR)
  std::vector<int> nboff, nb;     // neighbours per vertex (CSR)
  std::vector<char> bnd, corner;  // boundary vertex; corner of the boundary (kept in place)

  int valence(int v) const { return nboff[v + 1] - nboff[v]; }
  bool adjacent(int a, int b) const { return emap.count(ekey(a, b)) != 0; }
};

void build_topo(const Mesh& m, Topo& t) {
  const int nv = (int)m.p.size(), nf = (int)m.f.size();
  t.e.clear();
  t.emap.clear();
  t.emap.reserve((size_t)nf * 2);
  std::vector<int> cnt(nv + 1, 0);
  for (int i = 0; i < nf; ++i) {
    if (m.fdead[i]) continue;
    for (int k = 0; k < 3; ++k) {
      const int a = m.f[i][k], b = m.f[i][(k + 1) % 3];
      ++cnt[a];
      const uint64_t key = ekey(a, b);
      auto it = t.emap.find(key);
      if (it == t.emap.end()) {
        t.emap.emplace(key, (int)t.e.size());
        t.e.push_back(Edge{a, b, i, -1, k, -1, true});
      } else {
        Edge& e = t.e[it->second];
        if (e.f1 < 0 && e.manifold) {
          e.f1 = i;
          e.k1 = k;
        } else {
          e.manifold = false;
        }
      }
    }
  }
  t.voff.assign(nv + 1, 0);
  for (int v = 0; v < nv; ++v) t.voff[v + 1] = t.voff[v] + cnt[v];
  t.vfaces.assign(t.voff[nv], 0);
  std::vector<int> fill(t.voff.begin(), t.voff.end() - 1);
  for (int i = 0; i < nf; ++i) {
    if (m.fdead[i]) continue;
    for (int k = 0; k < 3; ++k) t.vfaces[fill[m.f[i][k]]++] = i;
  }
  std::vector<int> ncnt(nv + 1, 0);
  for (const Edge& e : t.e) {
    ++ncnt[e.a];
    ++ncnt[e.b];
  }
  t.nboff.assign(nv + 1, 0);
  for (int v = 0; v < nv; ++v) t.nboff[v + 1] = t.nboff[v] + ncnt[v];
  t.nb.assign(t.nboff[nv], 0);
  std::vector<int> nfill(t.nboff.begin(), t.nboff.end() - 1);
  t.bnd.assign(nv, 0);
  t.corner.assign(nv, 0);
  std::vector<int> bcount(nv, 0);
  std::vector<std::array<int, 2>> bnbr(nv, {-1, -1});
  for (const Edge& e : t.e) {
    t.nb[nfill[e.a]++] = e.b;
    t.nb[nfill[e.b]++] = e.a;
    if (e.f1 < 0 || !e.manifold) {
      t.bnd[e.a] = t.bnd[e.b] = 1;
      if (bcount[e.a] < 2) bnbr[e.a][bcount[e.a]] = e.b;
      if (bcount[e.b] < 2) bnbr[e.b][bcount[e.b]] = e.a;
      ++bcount[e.a];
      ++bcount[e.b];
      if (!e.manifold) t.corner[e.a] = t.corner[e.b] = 1;
    }
  }
  // corners: boundary vertices that are not a plain pass-through of two boundary edges turning by <= 45 degrees
  const double c45 = std::cos(M_PI / 4);
  for (int v = 0; v < nv; ++v) {
    if (!t.bnd[v] || m.vdead[v]) continue;
    if (bcount[v] != 2) {
      t.corner[v] = 1;
      continue;
    }
    const V3 d0 = m.p[v] - m.p[bnbr[v][0]], d1 = m.p[bnbr[v][1]] - m.p[v];
    const double l = std::sqrt(norm2(d0) * norm2(d1));
    if (!(l > 0) || dot(d0, d1) < c45 * l) t.corner[v] = 1;
  }
}

V3 face_normal(const Mesh& m, const Tri& f) { return cross(m.p[f[1]] - m.p[f[0]], m.p[f[2]] - m.p[f[0]]); }

// ---- split ------------------------------------------------------------------------------------------------------
int split_pass(Mesh& m, double high, const Bvh& border) {
  Topo t;
  build_topo(m, t);
  const double h2 = high * high;
  std::vector<std::pair<double, int>> cand;
  for (int i = 0; i < (int)t.e.size(); ++i) {
    const Edge& e = t.e[i];
    if (!e.manifold) continue;
    const double l2 = norm2(m.p[e.a] - m.p[e.b]);
    if (l2 > h2) cand.push_back({l2, i});
  }
  std::sort(cand.begin(), cand.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
  std::vector<char> touched(m.f.size(), 0);
  int done = 0;
  for (const auto& c : cand) {
    const Edge& e = t.e[c.second];
    if (touched[e.f0] || (e.f1 >= 0 && touched[e.f1])) continue;
    const int mv = (int)m.p.size();
    // a boundary midpoint goes onto the input's boundary polyline (the edge may be a chord of it)
    const V3 mid = 0.5 * (m.p[e.a] + m.p[e.b]);
    m.p.push_back(e.f1 < 0 ? border.closest(mid) : mid);
    m.vdead.push_back(0);
    {
      const Tri f = m.f[e.f0];  // a -> b at k0
      const int a = f[e.k0], b = f[(e.k0 + 1) % 3], cc = f[(e.k0 + 2) % 3];
      m.f[e.f0] = Tri{a, mv, cc};
      m.f.push_back(Tri{mv, b, cc});
      m.fdead.push_back(0);
      touched[e.f0] = 1;
    }
    if (e.f1 >= 0) {
      const Tri f = m.f[e.f1];  // b -> a at k1
      const int b = f[e.k1], a = f[(e.k1 + 1) % 3], d = f[(e.k1 + 2) % 3];
      m.f[e.f1] = Tri{b, mv, d};
      m.f.push_back(Tri{mv, a, d});
      m.fdead.push_back(0);
      touched[e.f1] = 1;
    }
    touched.resize(m.f.size(), 1);
    ++done;
  }
  return done;
}

// ---- collapse ---------------------------------------------------------------------------------------------------
int collapse_pass(Mesh& m, double low, double high) {
  Topo t;
  build_topo(m, t);
  const double l2 = low * low, h2 = high * high;
  std::vector<std::pair<double, int>> cand;
  for (int i = 0; i < (int)t.e.size(); ++i) {
    const Edge& e = t.e[i];
    if (!e.manifold) continue;
    const double d2 = norm2(m.p[e.a] - m.p[e.b]);
    if (d2 < l2) cand.push_back({d2, i});
  }
  std::sort(cand.begin(), cand.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<char> touched(m.p.size(), 0);
  std::vector<int> common;
  int done = 0;
  for (const auto& c : cand) {
    const Edge& e = t.e[c.second];
    int a = e.a, b = e.b;
    if (touched[a] || touched[b] || m.vdead[a] || m.vdead[b]) continue;
    const bool bedge = e.f1 < 0;
    // where the merged vertex goes: onto a boundary (corner) vertex, along a boundary edge, or to the midpoint
    V3 P;
    if (t.bnd[a] && t.bnd[b]) {
      if (!bedge) continue;  // would pinch the surface across its interior
      if (t.corner[a] && t.corner[b]) continue;
      // onto an end point (the corner, if one is), so the boundary vertices stay on the input's polyline
      P = t.corner[b] ? m.p[b] : m.p[a];
    } else if (t.bnd[a]) {
      P = m.p[a];
    } else if (t.bnd[b]) {
      P = m.p[b];
    } else {
      P = 0.5 * (m.p[a] + m.p[b]);
    }
    // link condition: the common neighbours are exactly the edge's opposite vertices
    common.clear();
    for (int i = t.nboff[a]; i < t.nboff[a + 1]; ++i) {
      const int x = t.nb[i];
      for (int j = t.nboff[b]; j < t.nboff[b + 1]; ++j)
        if (t.nb[j] == x) common.push_back(x);
    }
    if ((int)common.size() != (bedge ? 1 : 2)) continue;
    // no vertex left with fewer than three neighbours (two on the boundary)
    if (t.valence(a) + t.valence(b) - (bedge ? 3 : 4) < (t.bnd[a] || t.bnd[b] ? 2 : 3)) continue;
    bool thin = false;
    for (int x : common) thin = thin || t.valence(x) <= (t.bnd[x] ? 2 : 3);
    if (thin) continue;
    // no edge longer than `high` and no triangle turned over or flattened
    bool ok = true;
    for (int s = 0; s < 2 && ok; ++s) {
      const int v = s == 0 ? a : b;
      for (int i = t.nboff[v]; i < t.nboff[v + 1] && ok; ++i) {
        const int x = t.nb[i];
        if (x != a && x != b && norm2(m.p[x] - P) > h2) ok = false;
      }
      for (int i = t.voff[v]; i < t.voff[v + 1] && ok; ++i) {
        const Tri& f = m.f[t.vfaces[i]];
        const bool both = (f[0] == a || f[1] == a || f[2] == a) && (f[0] == b || f[1] == b || f[2] == b);
        if (both) continue;
        const V3 n0 = face_normal(m, f);
        Tri g = f;
        for (int k = 0; k < 3; ++k)
          if (g[k] == v) g[k] = -1;
        V3 q[3];
        for (int k = 0; k < 3; ++k) q[k] = g[k] < 0 ? P : m.p[g[k]];
        const V3 n1 = cross(q[1] - q[0], q[2] - q[0]);
        if (dot(n0, n1) <= 0.1 * std::sqrt(norm2(n0) * norm2(n1)) || norm2(n1) <= 1e-6 * norm2(n0)) ok = false;
      }
    }
    if (!ok) continue;
    // keep the vertex at P when P is an end point (the boundary or corner one), else either
    if ((t.bnd[b] && !t.bnd[a]) || (t.bnd[a] && t.bnd[b] && t.corner[b])) std::swap(a, b);
    for (int i = t.nboff[a]; i < t.nboff[a + 1]; ++i) touched[t.nb[i]] = 1;
    for (int i = t.nboff[b]; i < t.nboff[b + 1]; ++i) touched[t.nb[i]] = 1;
    touched[a] = touched[b] = 1;
    m.p[a] = P;
    for (int i = t.voff[b]; i < t.voff[b + 1]; ++i) {
      const int fi = t.vfaces[i];
      Tri& f = m.f[fi];
      const bool hasa = f[0] == a || f[1] == a || f[2] == a;
      if (hasa) {
        m.fdead[fi] = 1;
      } else {
        for (int k = 0; k < 3; ++k)
          if (f[k] == b) f[k] = a;
      }
    }
    m.vdead[b] = 1;
    ++done;
  }
  return done;
}

// ---- valence-equalising flips -----------------------------------------------------------------------------------
int flip_pass(Mesh& m) {
  Topo t;
  build_topo(m, t);
  std::vector<int> val(m.p.size());
  for (size_t v = 0; v < m.p.size(); ++v) val[v] = t.valence((int)v);
  std::vector<char> touched(m.f.size(), 0);
  int done = 0;
  for (size_t i = 0; i < t.e.size(); ++i) {
    const Edge& e = t.e[i];
    if (e.f1 < 0 || !e.manifold || touched[e.f0] || touched[e.f1]) continue;
    const int a = e.a, b = e.b;
    if (m.f[e.f1][e.k1] != b) continue;  // neighbours wound inconsistently: leave the edge
    const int c = m.f[e.f0][(e.k0 + 2) % 3], d = m.f[e.f1][(e.k1 + 2) % 3];
    if (c == d || t.emap.count(ekey(c, d))) continue;
    auto tgt = [&](int v) { return t.bnd[v] ? 4 : 6; };
    const int before = std::abs(val[a] - tgt(a)) + std::abs(val[b] - tgt(b)) + std::abs(val[c] - tgt(c)) +
                       std::abs(val[d] - tgt(d));
    const int after = std::abs(val[a] - 1 - tgt(a)) + std::abs(val[b] - 1 - tgt(b)) +
                      std::abs(val[c] + 1 - tgt(c)) + std::abs(val[d] + 1 - tgt(d));
    if (after >= before) continue;
    if (val[a] <= 3 || val[b] <= 3) continue;
    const Tri g0{a, d, c}, g1{d, b, c};  // the quad a -> d -> b -> c split along c - d
    const V3 n = face_normal(m, m.f[e.f0]) + face_normal(m, m.f[e.f1]);
    const V3 n0 = face_normal(m, g0), n1 = face_normal(m, g1);
    const double nn = std::sqrt(norm2(n));
    if (dot(n, n0) <= 0.1 * nn * std::sqrt(norm2(n0)) || dot(n, n1) <= 0.1 * nn * std::sqrt(norm2(n1))) continue;
    m.f[e.f0] = g0;
    m.f[e.f1] = g1;
    touched[e.f0] = touched[e.f1] = 1;
    t.emap.emplace(ekey(c, d), -1);
    --val[a];
    --val[b];
    ++val[c];
    ++val[d];
    ++done;
  }
  return done;
}

// ---- tangential relaxation --------------------------------------------------------------------------------------
void relax(Mesh& m) {
  Topo t;
  build_topo(m, t);
  const int nv = (int)m.p.size();
  std::vector<V3> np = m.p;
  for (int v = 0; v < nv; ++v) {
    if (m.vdead[v] || t.bnd[v] || t.voff[v + 1] == t.voff[v]) continue;
    V3 q{0, 0, 0}, n{0, 0, 0};
    double w = 0;
    for (int i = t.voff[v]; i < t.voff[v + 1]; ++i) {
      const Tri& f = m.f[t.vfaces[i]];
      const V3 fn = face_normal(m, f);
      const double ar = std::sqrt(norm2(fn));
      q = q + (ar / 3.0) * (m.p[f[0]] + m.p[f[1]] + m.p[f[2]]);
      n = n + fn;
      w += ar;
    }
    if (!(w > 0) || !(norm2(n) > 0)) continue;
    q = (1.0 / w) * q;
    n = (1.0 / std::sqrt(norm2(n))) * n;
    np[v] = q + dot(n, m.p[v] - q) * n;
  }
  // undo the moves around any triangle that would turn over
  for (int pass = 0; pass < 3; ++pass) {
    bool bad = false;
    for (size_t i = 0; i < m.f.size(); ++i) {
      if (m.fdead[i]) continue;
      const Tri& f = m.f[i];
      const V3 n0 = face_normal(m, f);
      const V3 n1 = cross(np[f[1]] - np[f[0]], np[f[2]] - np[f[0]]);
      if (dot(n0, n1) <= 0.1 * std::sqrt(norm2(n0) * norm2(n1))) {
        for (int k = 0; k < 3; ++k) np[f[k]] = m.p[f[k]];
        bad = true;
      }
    }
    if (!bad) break;
  }
  m.p.swap(np);
}

void project(Mesh& m, const Bvh& bvh) {
  Topo t;
  build_topo(m, t);
  for (size_t v = 0; v < m.p.size(); ++v)
    if (!m.vdead[v] && !t.bnd[v]) m.p[v] = bvh.closest(m.p[v]);
}

struct RemeshResult {
  std::vector<double> points;   // np x 3, column-major
  std::vector<int64_t> tris;    // nt x 3, column-major, 0-based
};

}  // namespace
}  // namespace cdj

#include "../../include/cdj.h"

extern "C" {

struct cdj_remesh {
  cdj::RemeshResult r;
};

int cdj_remesh_isotropic(const double* points, int64_t npoints, const int64_t* tris, int64_t ntris,
                         double target_edge_length, int32_t iterations, cdj_remesh** out, int64_t* npoints_out,
                         int64_t* ntris_out) {
  using namespace cdj;
  if (!out || !npoints_out || !ntris_out) return CDJ_ERR_ARG;
  *out = nullptr;
  if (!points || !tris || npoints < 3 || ntris < 1 || npoints > (int64_t)1 << 30 || ntris > (int64_t)1 << 30)
    return CDJ_ERR_ARG;
  if (!(target_edge_length > 0) || !std::isfinite(target_edge_length) || iterations < 1) return CDJ_ERR_ARG;
  Mesh m;
  m.p.resize(npoints);
  for (int64_t i = 0; i < npoints; ++i) m.p[i] = V3{points[i], points[npoints + i], points[2 * npoints + i]};
  m.f.resize(ntris);
  for (int64_t i = 0; i < ntris; ++i)
    for (int k = 0; k < 3; ++k) {
      const int64_t v = tris[k * ntris + i];
      if (v < 0 || v >= npoints) return CDJ_ERR_ARG;
      m.f[i][k] = (int)v;
    }
  for (const Tri& f : m.f)
    if (f[0] == f[1] || f[1] == f[2] || f[0] == f[2]) return CDJ_ERR_DEGENERATE;
  m.fdead.assign(ntris, 0);
  m.vdead.assign(npoints, 0);
  const std::vector<V3> p0 = m.p;
  const std::vector<Tri> f0 = m.f;
  try {
    const Bvh bvh(p0, f0);
    std::vector<Tri> segs;  // the input's boundary edges, as degenerate triangles (a, b, b)
    {
      Mesh m0 = m;
      Topo t0;
      build_topo(m0, t0);
      for (const Edge& e : t0.e)
        if (e.f1 < 0 || !e.manifold) segs.push_back(Tri{e.a, e.b, e.b});
    }
    const Bvh border(p0, segs, true);
    const double L = target_edge_length, high = 4.0 / 3.0 * L, low = 4.0 / 5.0 * L;
    for (int it = 0; it < iterations; ++it) {
      for (int s = 0; s < 64 && split_pass(m, high, border) > 0; ++s) {
      }
      for (int s = 0; s < 64 && collapse_pass(m, low, high) > 0; ++s) {
      }
      for (int s = 0; s < 8 && flip_pass(m) > 0; ++s) {
      }
      relax(m);
      project(m, bvh);
    }
    // compact: live faces, referenced vertices
    std::vector<int> map(m.p.size(), -1);
    auto* h = new cdj_remesh;
    int64_t nv = 0, nf = 0;
    for (size_t i = 0; i < m.f.size(); ++i) {
      if (m.fdead[i]) continue;
      ++nf;
      for (int k = 0; k < 3; ++k)
        if (map[m.f[i][k]] < 0) map[m.f[i][k]] = (int)nv++;
    }
    h->r.points.assign(3 * nv, 0.0);
    for (size_t v = 0; v < m.p.size(); ++v)
      if (map[v] >= 0) {
        h->r.points[map[v]] = m.p[v].x;
        h->r.points[nv + map[v]] = m.p[v].y;
        h->r.points[2 * nv + map[v]] = m.p[v].z;
      }
    h->r.tris.assign(3 * nf, 0);
    int64_t j = 0;
    for (size_t i = 0; i < m.f.size(); ++i) {
      if (m.fdead[i]) continue;
      for (int k = 0; k < 3; ++k) h->r.tris[k * nf + j] = map[m.f[i][k]];
      ++j;
    }
    *out = h;
    *npoints_out = nv;
    *ntris_out = nf;
  } catch (const std::bad_alloc&) {
    return CDJ_ERR_NOMEM;
  }
  return CDJ_OK;
}

int cdj_remesh_take(cdj_remesh* r, double* points, int64_t* tris) {
  if (!r) return CDJ_ERR_ARG;
  if (points) std::memcpy(points, r->r.points.data(), sizeof(double) * r->r.points.size());
  if (tris) std::memcpy(tris, r->r.tris.data(), sizeof(int64_t) * r->r.tris.size());
  delete r;
  return CDJ_OK;
}

}  // extern "C"
