"""Pins of the CPU FP64 oracle against closed forms and invariants (no GPU).

* Okada (1985) surface solution, half space, displacement and strain, the reference's own test geometry
  (test/test_TDvsOkada.jl) - the reference asserts 1e-13; |.| is used here (the reference's signed max is
  one-sided).
* strain == sym(grad u) by central differences, full and half space.
* AngDisStrainFSC == sym(grad AngDisDispFSC) / (4 pi (1 - nu)) by central differences.
* Displacement jump across the TD interior equals the Burgers vector (full space).
* Far field of the TD decays like a point dislocation (strain ~ r^-3).
* Half space -> full space as depth -> infinity: the free-surface correction plus the image decay like D^-2
  (displacement) and D^-3 (strain).
* The half-space solution leaves the free surface traction-free (sigma_xz = sigma_yz = sigma_zz = 0 at z = 0).
* Face normals and incentre midpoints: the host's CreateFaceNormalAndMidPoint equals the oracle's, and both meet the
  geometry (unit normals orthogonal to the sides, the midpoint equidistant from the three sides).
* Rigid-motion covariance (full space): rotating and translating triangle and points rotates displacement and
  strain, for the same global Burgers vector.
* Far field = Volterra's integral of the Burgers vector against Kelvin's Green's function (Gauss-Legendre on the
  Duffy map), displacement directly and strain through its differences: an independent pin of the angular-dislocation
  sums that shares no formula with them.
* A square triangulated along either diagonal gives the same field under uniform slip (full and half space).
"""
import numpy as np
import pytest

from cutanddisplacejulia_amd import api
from oracle import analytic as A
from oracle import oracle as O


def test_okada_surface_halfspace():
    P1, P2, P3, md = A.okada_test_fault()
    xx = np.linspace(-10, 10, 50)
    X, Y = np.meshgrid(xx, xx)
    x, y = X.ravel(order="F"), Y.ravel(order="F")
    z = np.zeros_like(x)
    Dds, Dn, Dss = 2.0, 0.6, 0.5
    s, d = O.td_field(x, y, z, P1, P2, P3, [Dss] * 2, [Dds] * 2, [Dn] * 2, 0.25, 1.0, 1, 1, 1)
    tot = np.sqrt(Dds ** 2 + Dss ** 2)
    rake = 90 - np.degrees(np.arctan(Dss / Dds))
    ok = A.okada85_surface(x, y, md, 60, 45, 5, 5, rake, tot, Dn, 0.25)
    for mine, ref in zip(d + s, ok):
        assert np.max(np.abs(mine - ref)) < 1e-13


def test_okada_surface_refined_mesh():
    """The same rectangle split into 2 k^2 triangles still matches Okada (uniform slip)."""
    P1, P2, P3, md = A.okada_test_fault()
    # corners of the rectangle from the two reference triangles
    c = {tuple(np.round(p, 12)): p for p in np.vstack([P1, P2, P3])}
    pts = np.array(list(c.values()))
    top = pts[np.argsort(pts[:, 2])[-2:]]
    bot = pts[np.argsort(pts[:, 2])[:2]]
    # order so that top[0]-bot[0] is a dip edge
    if np.linalg.norm(top[0] - bot[0]) > np.linalg.norm(top[0] - bot[1]):
        bot = bot[::-1]
    k = 4
    grid = np.zeros((k + 1, k + 1, 3))
    for i in range(k + 1):
        for j in range(k + 1):
            a, b = i / k, j / k
            grid[i, j] = (1 - a) * ((1 - b) * top[0] + b * top[1]) + a * ((1 - b) * bot[0] + b * bot[1])
    Q1, Q2, Q3 = [], [], []
    for i in range(k):
        for j in range(k):
            p, q, r, s_ = grid[i, j], grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]
            Q1 += [p, p]
            Q2 += [q, r]
            Q3 += [r, s_]
    Q1, Q2, Q3 = map(np.array, (Q1, Q2, Q3))
    xx = np.linspace(-10, 10, 15)
    X, Y = np.meshgrid(xx, xx)
    x, y = X.ravel(), Y.ravel()
    z = np.zeros_like(x)
    n = Q1.shape[0]
    nv = np.cross(Q2 - Q1, Q3 - Q1)
    # uniform Dss along the strike; orientation of each triangle's frame must agree with the 2-triangle mesh
    nv0 = np.cross(P2[0] - P1[0], P3[0] - P1[0])
    sgn = np.sign(nv @ nv0)
    assert np.all(sgn > 0)
    s, d = O.td_field(x, y, z, Q1, Q2, Q3, [0.5] * n, [2.0] * n, [0.6] * n, 0.25, 1.0, 1, 1, 1)
    s0, d0 = O.td_field(x, y, z, P1, P2, P3, [0.5] * 2, [2.0] * 2, [0.6] * 2, 0.25, 1.0, 1, 1, 1)
    for a, b in zip(d + s, d0 + s0):
        assert np.max(np.abs(a - b)) < 1e-12 * max(1.0, np.max(np.abs(b)))


def test_okada_surface_benchmark_mesh():
    """The oracle on the half-space benchmark mesh itself (the Okada test rectangle with its top at depth 0.5, split
    into 5000 triangles) with uniform slip: at free-surface points its summed field equals Okada (1985) to
    1e-11 of each group's largest value (5000 summed contributions; the 2-triangle pin above holds 1e-13)."""
    from cutanddisplacejulia_amd import meshes
    c = meshes.config2(50)
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    rng = np.random.default_rng(0)
    X, Y, Z = rng.uniform(-10, 10, 8), rng.uniform(-10, 10, 8), np.zeros(8)
    Dds, Dn, Dss = 2.0, 0.6, 0.5
    so, do = O.td_field(X, Y, Z, P1, P2, P3, [Dss] * n, [Dds] * n, [Dn] * n, 0.25, 1.0, 1, 1, 1)
    tot = np.sqrt(Dds ** 2 + Dss ** 2)
    rake = 90 - np.degrees(np.arctan(Dss / Dds))
    ok = A.okada85_surface(X, Y, 0.5 + np.sin(np.deg2rad(45.0)) * 2.5, 60, 45, 5, 5, rake, tot, Dn, 0.25)
    mine = list(do) + list(so)
    for grp in (range(0, 3), range(3, 9)):
        scale = max(np.max(np.abs(ok[k])) for k in grp)
        for k in grp:
            assert np.max(np.abs(np.asarray(mine[k]) - ok[k])) <= 1e-11 * scale, k


@pytest.mark.parametrize("hs", [0, 1])
def test_strain_is_symmetric_gradient_of_displacement(hs):
    rng = np.random.default_rng(1)
    P1 = np.array([[-1.0, -0.5, -2.0]])
    P2 = np.array([[1.0, -0.3, -2.7]])
    P3 = np.array([[0.2, 1.1, -1.5]])
    pts = rng.uniform(-3, 3, (12, 3))
    pts[:, 2] = -np.abs(pts[:, 2]) - 0.1
    h = 1e-5
    one = [1.0]
    for p in pts:
        s, _ = O.td(p[0:1], p[1:2], p[2:3], P1, P2, P3, one, one, one, 0.25, 1.0, 0, 1, hs)
        g = np.zeros((3, 3, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            _, dp = O.td([p[0] + e[0]], [p[1] + e[1]], [p[2] + e[2]], P1, P2, P3, one, one, one, 0.25, 1.0, 1, 0, hs)
            _, dm = O.td([p[0] - e[0]], [p[1] - e[1]], [p[2] - e[2]], P1, P2, P3, one, one, one, 0.25, 1.0, 1, 0, hs)
            for m, md in enumerate(["Dn", "Dss", "Dds"]):
                for c, cc in enumerate("xyz"):
                    g[m, c, k] = (dp[md + "U" + cc][0, 0] - dm[md + "U" + cc][0, 0]) / (2 * h)
        for m, md in enumerate(["Dn", "Dss", "Dds"]):
            G = g[m]
            eps = 0.5 * (G + G.T)
            mine = np.array([s[k + md][0, 0] for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz")])
            fd = np.array([eps[0, 0], eps[1, 1], eps[2, 2], eps[0, 1], eps[0, 2], eps[1, 2]])
            assert np.max(np.abs(mine - fd)) <= 1e-7 * max(1e-3, np.max(np.abs(fd)))


def test_fsc_strain_is_gradient_of_fsc_displacement():
    rng = np.random.default_rng(3)
    nu = 0.25
    f = 1 / (4 * np.pi * (1 - nu))
    for _ in range(60):
        y = rng.uniform(-3, 3, 3)
        beta = rng.uniform(0.2, 2.9)
        a = rng.uniform(0.2, 3)
        v = O.ang_dis_strain_fsc(y[0], y[1], y[2], beta, nu, a)
        h = 1e-6
        g = np.zeros((3, 3, 3))
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            du = (O.ang_dis_disp_fsc(*(y + e), beta, nu, a) - O.ang_dis_disp_fsc(*(y - e), beta, nu, a)) / (2 * h)
            for c in range(3):
                for k in range(3):
                    g[k, c, d] = du[3 * c + k]
        for k in range(3):
            G = g[k] * f
            fd = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[0, 2] + G[2, 0], G[1, 2] + G[2, 1]])
            cf = v[6 * k:6 * k + 6]
            assert np.max(np.abs(cf - fd)) <= 1e-6 * max(1e-6, np.max(np.abs(cf)))


def test_displacement_jump_equals_burgers_vector():
    """Full space: u(just below) - u(just above) across the TD interior equals the slip vector."""
    P1 = np.array([[-1.0, -1.0, 0.0]])
    P2 = np.array([[1.0, -1.0, 0.0]])
    P3 = np.array([[0.0, 1.0, 0.0]])
    Dss, Dds, Dn = 0.3, -0.7, 0.5
    eps = 1e-9
    up = O.td([0.1], [-0.2], [eps], P1, P2, P3, [Dss], [Dds], [Dn], 0.25, 1.0, 1, 0, 0)[1]
    dn = O.td([0.1], [-0.2], [-eps], P1, P2, P3, [Dss], [Dds], [Dn], 0.25, 1.0, 1, 0, 0)[1]
    jump = np.array([sum(up[m + "U" + c][0, 0] - dn[m + "U" + c][0, 0] for m in ("Dn", "Dss", "Dds")) for c in "xyz"])
    # flat triangle with upward normal: Vnorm = +z, Vstrike = +y, Vdip = -x (TD.jl:719-730)
    expected = Dn * np.array([0, 0, 1.0]) + Dss * np.array([0, 1.0, 0]) + Dds * np.array([-1.0, 0, 0])
    assert np.allclose(np.abs(jump), np.abs(expected), atol=1e-6)
    assert np.allclose(np.abs(jump).sum(), np.abs(expected).sum(), atol=1e-6)


def test_far_field_decay():
    P1 = np.array([[-0.1, -0.1, -5.0]])
    P2 = np.array([[0.1, -0.1, -5.0]])
    P3 = np.array([[0.0, 0.1, -5.1]])
    d = np.array([1.0, 0.3, -0.4])
    d /= np.linalg.norm(d)
    vals = []
    for r in (20.0, 40.0, 80.0):
        p = np.array([0, 0, -5.0]) + r * d
        s, _ = O.td([p[0]], [p[1]], [p[2]], P1, P2, P3, [1.0], [0.0], [0.0], 0.25, 1.0, 0, 1, 0)
        vals.append(np.sqrt(sum(s[k][0, 0] ** 2 for k in s)))
    assert abs(np.log2(vals[0] / vals[1]) - 3) < 0.05
    assert abs(np.log2(vals[1] / vals[2]) - 3) < 0.03


def test_halfspace_tends_to_fullspace_with_depth():
    """A triangle and its observation points moved down by D: |u_HS - u_FS| / |u_FS| ~ D^-2 and the strain's
    ~ D^-3 (the image dislocation and the free-surface correction are 2D away)."""
    P = [np.array([[-0.6, -0.4, 0.1]]), np.array([[0.8, -0.2, 0.5]]), np.array([[0.0, 0.7, -0.4]])]
    pts = np.random.default_rng(7).uniform(-1.5, 1.5, (20, 3))
    rel = []
    for D in (40.0, 80.0, 160.0, 320.0):
        Q = [p + np.array([0.0, 0.0, -D]) for p in P]
        X, Y, Z = pts[:, 0], pts[:, 1], pts[:, 2] - D
        sh, dh = O.td(X, Y, Z, *Q, [1.0], [1.0], [1.0], 0.25, 1.0, 1, 1, 1)
        sf, df = O.td(X, Y, Z, *Q, [1.0], [1.0], [1.0], 0.25, 1.0, 1, 1, 0)
        du = max(np.max(np.abs(dh[k] - df[k])) for k in dh) / max(np.max(np.abs(df[k])) for k in df)
        de = max(np.max(np.abs(sh[k] - sf[k])) for k in sh) / max(np.max(np.abs(sf[k])) for k in sf)
        rel.append((du, de))
    assert rel[-1][0] < 1e-5 and rel[-1][1] < 1e-7
    for (u0, e0), (u1, e1) in zip(rel, rel[1:]):
        assert abs(np.log2(u0 / u1) - 2) < 0.08
        assert abs(np.log2(e0 / e1) - 3) < 0.12


def test_free_surface_is_traction_free():
    rng = np.random.default_rng(11)
    nu, G = 0.25, 1.0
    lam = 2 * G * nu / (1 - 2 * nu)
    for _ in range(6):
        c = np.concatenate([rng.uniform(-1, 1, 2), [-rng.uniform(0.1, 2)]])
        P = [c + rng.uniform(-0.6, 0.6, 3) for _ in range(3)]
        for p in P:
            p[2] = min(p[2], -0.02)
        xy = rng.uniform(-3, 3, (30, 2))
        s, _ = O.td(xy[:, 0], xy[:, 1], np.zeros(30), *(p[None, :] for p in P), [1.0], [1.0], [1.0], nu, G, 0, 1, 1)
        for m in ("Dn", "Dss", "Dds"):
            e = {k: s[k + m][:, 0] for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz")}
            tr = e["exx"] + e["eyy"] + e["ezz"]
            inplane = max(np.max(np.abs(v)) for v in (2 * G * e["exx"] + lam * tr, 2 * G * e["eyy"] + lam * tr,
                                                       2 * G * e["exy"]))
            surface = max(np.max(np.abs(v)) for v in (2 * G * e["ezz"] + lam * tr, 2 * G * e["exz"],
                                                       2 * G * e["eyz"]))
            assert surface <= 1e-12 * inplane


def test_face_normals_and_midpoints():
    rng = np.random.default_rng(13)
    P1, P2, P3 = (rng.uniform(-5, 5, (200, 3)) for _ in range(3))
    n_h, m_h = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    n_o, m_o = O.face_normal_midpoint(P1, P2, P3)
    assert np.max(np.abs(n_h - n_o)) <= 4 * np.finfo(float).eps
    assert np.max(np.abs(m_h - m_o)) <= 4 * np.finfo(float).eps * np.max(np.abs(m_o))
    assert np.allclose(np.linalg.norm(n_o, axis=1), 1.0, rtol=0, atol=1e-15)
    for A_, B_ in ((P1, P2), (P2, P3), (P3, P1)):
        side = B_ - A_
        assert np.max(np.abs(np.sum(n_o * side, axis=1)) / np.linalg.norm(side, axis=1)) <= 1e-14
    # right-handed: n along (P2 - P1) x (P3 - P1)
    assert np.all(np.sum(n_o * np.cross(P2 - P1, P3 - P1), axis=1) > 0)
    # incentre: in the triangle's plane, equidistant from the three sides, at the inradius 2 area / perimeter

    def dist(X, A_, B_):
        t = np.sum((X - A_) * (B_ - A_), axis=1) / np.sum((B_ - A_) ** 2, axis=1)
        return np.linalg.norm(X - A_ - t[:, None] * (B_ - A_), axis=1)
    area = 0.5 * np.linalg.norm(np.cross(P2 - P1, P3 - P1), axis=1)
    per = sum(np.linalg.norm(B_ - A_, axis=1) for A_, B_ in ((P1, P2), (P2, P3), (P3, P1)))
    r_in = 2 * area / per
    for A_, B_ in ((P1, P2), (P2, P3), (P3, P1)):
        assert np.max(np.abs(dist(m_o, A_, B_) - r_in) / r_in) <= 1e-9
    assert np.max(np.abs(np.sum(n_o * (m_o - P1), axis=1))) <= 1e-12


def test_halfspace_rejects_points_above_surface():
    with pytest.raises(O.OracleError, match="Z coordinates must be negative"):
        O.td([0.0], [0.0], [0.5], [[0, 0, -1.0]], [[1, 0, -1.0]], [[0, 1, -1.0]], [1.0], [1.0], [1.0], 0.25, 1.0, 0, 1, 1)


def test_size_mismatch_error():
    with pytest.raises(O.OracleError, match="same dimensions"):
        O.td([0.0], [0.0], [-0.5], [[0, 0, -1.0]], [[1, 0, -1.0]], [[0, 1, -1.0]], [1.0, 2.0], [1.0], [1.0], 0.25, 1.0,
             0, 1, 1)


def _tdcs_axes(P1, P2, P3):
    """Vnorm, Vstrike, Vdip of a triangle (the frame of TD.jl:712-732): the slip modes Dn, Dss, Dds act along them."""
    vn = np.cross(P2 - P1, P3 - P1)
    vn /= np.linalg.norm(vn)
    vs = np.cross([0.0, 0.0, 1.0], vn)
    if np.linalg.norm(vs) == 0.0:
        vs = np.array([0.0, vn[2], 0.0])
    vs /= np.linalg.norm(vs)
    return vn, vs, np.cross(vn, vs)


def _unit_modes(X, P1, P2, P3, hs=0):
    """Displacement (3 x 3: mode, component) and strain (3 x 6) of unit Dn, Dss, Dds at each point of X (m x 3)."""
    s, d = O.td(X[:, 0], X[:, 1], X[:, 2], P1[None], P2[None], P3[None], [1.0], [1.0], [1.0], 0.25, 1.0, 1, 1, hs)
    u = np.stack([np.stack([d[md + "U" + c][:, 0] for c in "xyz"], -1) for md in ("Dn", "Dss", "Dds")], 1)
    e = np.stack([np.stack([s[k + md][:, 0] for k in ("exx", "eyy", "ezz", "exy", "exz", "eyz")], -1)
                  for md in ("Dn", "Dss", "Dds")], 1)
    return u, e


def _sym(e6):
    exx, eyy, ezz, exy, exz, eyz = e6
    return np.array([[exx, exy, exz], [exy, eyy, eyz], [exz, eyz, ezz]])


def test_fullspace_rigid_motion_covariance():
    """Rotating and translating triangle and points rotates the fields: u'(R x + t) = R u(x) and
    e' = R e R^T for the same global Burgers vector (each mode's vector R b expressed in the moved triangle's own
    frame, whose strike and dip directions are not R times the old ones)."""
    rng = np.random.default_rng(21)
    P1, P2, P3 = np.array([-0.7, -0.5, 0.2]), np.array([0.9, -0.3, -0.1]), np.array([0.1, 0.8, 0.4])
    X = rng.uniform(-2, 2, (30, 3))
    u, e = _unit_modes(X, P1, P2, P3)
    axes = _tdcs_axes(P1, P2, P3)
    worst = 0.0
    for _ in range(4):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        a, b, c, d = q
        R = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                      [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                      [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])
        t = rng.uniform(-5, 5, 3)
        Q1, Q2, Q3 = (R @ P + t for P in (P1, P2, P3))
        u2, e2 = _unit_modes(X @ R.T + t, Q1, Q2, Q3)
        axes2 = _tdcs_axes(Q1, Q2, Q3)
        for m in range(3):
            w = np.array([np.dot(R @ axes[m], ax) for ax in axes2])  # R b_m in the moved frame
            um = np.einsum("k,pkc->pc", w, u2)
            em = np.einsum("k,pkc->pc", w, e2)
            for p in range(X.shape[0]):
                ru = R @ u[p, m]
                re = R @ _sym(e[p, m]) @ R.T
                worst = max(worst, np.max(np.abs(um[p] - ru)) / np.max(np.abs(u[:, m])),
                            np.max(np.abs(_sym(em[p]) - re)) / np.max(np.abs(e[:, m])))
    assert worst <= 1e-12, worst


def _volterra_disp(X, P1, P2, P3, du, nu, mu=1.0, ngl=16):
    """Volterra's formula u_k(x) = -int m_pq dG_kp/dx_q dS over the triangle, m = lam (du.n) I + mu (du n + n du),
    with Kelvin's full-space Green's function G_kp = [(3 - 4 nu) d_kp / r + r_k r_p / r^3] / (16 pi mu (1 - nu));
    Gauss-Legendre on the Duffy map of the triangle (ngl^2 points)."""
    lam = 2 * mu * nu / (1 - 2 * nu)
    n = np.cross(P2 - P1, P3 - P1)
    area2 = np.linalg.norm(n)
    n /= area2
    m = lam * np.dot(du, n) * np.eye(3) + mu * (np.outer(du, n) + np.outer(n, du))
    g, w = np.polynomial.legendre.leggauss(ngl)
    g, w = 0.5 * (g + 1), 0.5 * w
    C = 1.0 / (16 * np.pi * mu * (1 - nu))
    out = np.zeros((X.shape[0], 3))
    for s, ws in zip(g, w):
        for t, wt in zip(g, w):
            xi = P1 + s * (P2 - P1) + s * t * (P3 - P2)
            jac = area2 * s * ws * wt
            r = X - xi
            rr = np.linalg.norm(r, axis=1)
            # dG_kp/dx_q contracted with m_pq
            # d/dx_q (d_kp / r) = -d_kp r_q / r^3 and
            # d/dx_q (r_k r_p / r^3) = (d_kq r_p + d_pq r_k) / r^3 - 3 r_k r_p r_q / r^5, contracted with m_pq
            mr = r @ m.T
            rmr = np.sum(r * mr, axis=1)
            r3, r5 = rr[:, None] ** 3, rr[:, None] ** 5
            out -= C * jac * (-(3 - 4 * nu) * mr / r3 + (r @ m + np.trace(m) * r) / r3 - 3 * r * rmr[:, None] / r5)
    return out


def test_far_field_matches_volterra_quadrature():
    """Full space, points 3 to 12 side lengths away: each mode's displacement equals Volterra's integral of its own
    displacement jump (the unit axis, signed as measured across the triangle's interior) against Kelvin's Green's
    function, to 1e-10, and its strain the integral's symmetric gradient (fourth-order central differences)
    to 1e-8.  An independent pin of the full-space angular-dislocation sums (TD.jl:735-867, 1528-1750)."""
    nu = 0.25
    P1, P2, P3 = np.array([-0.7, -0.5, 0.2]), np.array([0.9, -0.3, -0.1]), np.array([0.1, 0.8, 0.4])
    axes = _tdcs_axes(P1, P2, P3)
    cen = (P1 + P2 + P3) / 3
    rng = np.random.default_rng(5)
    dirs = rng.normal(size=(12, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    X = cen + dirs * rng.uniform(5, 20, (12, 1))
    u, e = _unit_modes(X, P1, P2, P3)
    # the displacement jump of each unit mode across the interior
    h = 1e-9
    Xp = np.array([cen + h * axes[0], cen - h * axes[0]])
    uj, _ = _unit_modes(Xp, P1, P2, P3)
    for m in range(3):
        jump = uj[0, m] - uj[1, m]
        du = np.sign(np.dot(jump, axes[m])) * axes[m]
        assert np.max(np.abs(jump - du)) <= 1e-7
        uv = _volterra_disp(X, P1, P2, P3, du, nu)
        assert np.max(np.abs(uv - u[:, m])) <= 1e-10 * np.max(np.abs(u[:, m]))
        for p in range(X.shape[0]):
            d = 1e-3 * np.linalg.norm(X[p] - cen)
            grad = np.zeros((3, 3))
            for q in range(3):
                ev = np.zeros(3)
                ev[q] = d
                pts = np.array([X[p] + 2 * ev, X[p] + ev, X[p] - ev, X[p] - 2 * ev])
                v = _volterra_disp(pts, P1, P2, P3, du, nu)
                grad[:, q] = (-v[0] + 8 * v[1] - 8 * v[2] + v[3]) / (12 * d)
            es = 0.5 * (grad + grad.T)
            assert np.max(np.abs(es - _sym(e[p, m]))) <= 1e-8 * np.max(np.abs(e[:, m]))


@pytest.mark.parametrize("hs", [0, 1])
def test_two_triangulations_of_a_square_agree(hs):
    """A square split along either diagonal, uniform slip: the summed fields agree away from its plane."""
    c = np.array([0.3, -0.2, -2.0])
    a, b = np.array([1.0, 0.2, 0.3]), np.array([-0.1, 0.8, 0.4])
    b = b - np.dot(a, b) / np.dot(a, a) * a
    V = [c, c + a, c + a + b, c + b]
    T1 = [(V[0], V[1], V[2]), (V[0], V[2], V[3])]
    T2 = [(V[0], V[1], V[3]), (V[1], V[2], V[3])]
    rng = np.random.default_rng(9)
    X = rng.uniform(-2, 2, (40, 3)) + c
    X[:, 2] = -np.abs(X[:, 2] - c[2]) - 0.3 if hs else X[:, 2]
    nrm = np.cross(a, b)
    nrm /= np.linalg.norm(nrm)
    X = X[np.abs((X - c) @ nrm) > 0.05]
    res = []
    for tris in (T1, T2):
        P1, P2, P3 = (np.array([t[k] for t in tris]) for k in range(3))
        s, d = O.td_field(X[:, 0], X[:, 1], X[:, 2], P1, P2, P3, [0.4, 0.4], [-0.9, -0.9], [0.3, 0.3], 0.25, 1.0, 1,
                          1, hs)
        res.append((np.array(s), np.array(d)))
    for k in range(2):
        ref = res[0][k]
        assert np.max(np.abs(res[1][k] - ref)) <= 1e-12 * np.max(np.abs(ref))
