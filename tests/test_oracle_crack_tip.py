"""Pins of the crack-tip oracle (oracle/crack_tip.py) on the CPU.

* The reference's own case, test_TDvsStressIntensity.jl, on its own mesh (tests/golden/VertPenny-300-EqEdges.stl):
  a penny crack of radius 1 inclined 15 degrees to the loading plane, G = 5, nu = 0.25, remote sigma_zz = 1, slip
  from the oracle's collocation matrix (LAPACK); K_I..K_III on the free edges against Tada's closed form, with the
  reference's pass criterion (max |100 K / K_an| - 100 <= 11 over the finite values).
* Free-edge flags against an independent count of identical sorted end-point pairs (a dictionary over all edges).
* A single horizontal tip triangle worked by hand: with the mid-to-edge direction +x, K_II measures the slip
  component along +x and K_III the one along -y (the edge vector n x ... convention), and
  K_I = G sqrt(pi) / (2 sqrt(h) (1 - nu)) Dn / 1.834.
* The in-plane decomposition preserves the slip magnitude: (K_II^2 + (K_III / (1 - nu))^2) = c^2 (Dss^2 + Dds^2).
"""
import numpy as np

from oracle import analytic as AN
from oracle import crack_tip as CT
from oracle import oracle as O
from tests import stl

G_, NU_ = 5.0, 0.25


def _inclined_penny():
    P, T = stl.read_stl("VertPenny-300-EqEdges.stl")
    x = P[:, 1].copy()
    P[:, 1] = P[:, 3]
    P[:, 3] = x
    radius = P[:, 1].max()
    beta = 15.0
    bfv = 90.0 - beta
    c, s = np.cos(np.deg2rad(bfv)), np.sin(np.deg2rad(bfv))
    X, Z = P[:, 1].copy(), P[:, 3].copy()
    P[:, 1] = c * X - s * Z
    P[:, 3] = s * X + c * Z
    P1, P2, P3 = (P[T[:, k] - 1, 1:4].copy() for k in range(3))
    return P1, P2, P3, radius, beta, bfv


def _slip_oracle(P1, P2, P3, szz):
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    lam = 2 * G_ * NU_ / (1 - 2 * NU_)
    A = O.assemble_A(O.traction_influence(nrm, mid, P1, P2, P3, NU_, G_, lam, 0))
    n = P1.shape[0]
    Tn, Tds, Tss = O.normal_shear_tractions((0, 0, szz, 0, 0, 0), nrm)
    x = np.linalg.solve(A, np.concatenate([Tn, Tss, Tds]))
    return x[:n], x[n:2 * n], x[2 * n:], nrm, mid


def _collect(edges, key):
    return np.concatenate([e[key][e["FreeFlg"]] for e in edges])


def _pct(an, num):
    with np.errstate(divide="ignore", invalid="ignore"):
        p = (100.0 / an) * num
    p[np.abs(p) == np.inf] = np.nan
    return p


def test_oracle_inclined_penny_against_tada():
    P1, P2, P3, radius, beta, bfv = _inclined_penny()
    Dn, Dss, Dds, nrm, mid = _slip_oracle(P1, P2, P3, 1.0)
    edges = CT.stress_intensity(Dn, Dss, Dds, G_, NU_, nrm, CT.crack_tip_elements(mid, P1, P2, P3, nrm))
    md = np.vstack([e["FeMd"][e["FreeFlg"]] for e in edges])
    c, s = np.cos(np.deg2rad(-bfv)), np.sin(np.deg2rad(-bfv))
    xr = c * md[:, 0] - s * md[:, 2]
    theta = np.rad2deg(np.arctan2(md[:, 1], xr))
    K1an, K2an, K3an = AN.tada_inclined_penny_tension(1.0, beta, theta, 1, NU_)
    worst = []
    for an, key in ((K1an, "K1"), (K2an, "K2"), (K3an, "K3")):
        p = _pct(an, _collect(edges, key))
        worst.append(np.max(np.abs(p[~np.isnan(p)])) - 100)
    print("inclined penny: %d tip edges, max error K1 %.2f %%, K2 %.2f %%, K3 %.2f %%" % (md.shape[0], *worst))
    assert md.shape[0] > 30
    assert max(worst) <= 11


def test_free_edges_against_independent_count():
    from cutanddisplacejulia_amd import meshes
    P1, P2, P3 = meshes.penny(5, 1.0, jitter=0.02, seed=3)
    # open a hole: drop a few interior triangles so inner free edges exist too
    keep = np.ones(P1.shape[0], dtype=bool)
    keep[[10, 11, 40]] = False
    P1, P2, P3 = P1[keep], P2[keep], P3[keep]
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    f12, f23, f31 = CT.edge_constraints(P1, P2, P3, mid)
    count = {}
    keys = []
    for A, B in ((P1, P2), (P2, P3), (P3, P1)):
        ks = [tuple(sorted([tuple(a), tuple(b)])) for a, b in zip(A, B)]
        keys.append(ks)
        for k in ks:
            count[k] = count.get(k, 0) + 1
    for flags, ks in zip((f12, f23, f31), keys):
        assert np.array_equal(flags, np.array([count[k] == 1 for k in ks]))
    assert 0 < int(f12.sum() + f23.sum() + f31.sum()) < 3 * P1.shape[0]


def test_horizontal_tip_triangle_by_hand():
    # one triangle in z = 0 with its P1P2 edge at x = 1 (the other two edges are shared with a neighbour)
    P1 = np.array([[1.0, -0.5, 0.0], [1.0, -0.5, 0.0]])
    P2 = np.array([[1.0, 0.5, 0.0], [0.0, 0.0, 0.0]])
    P3 = np.array([[0.0, 0.0, 0.0], [0.5, -1.5, 0.0]])
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    assert np.allclose(nrm[0], [0, 0, -1]) or np.allclose(nrm[0], [0, 0, 1])
    if nrm[0, 2] < 0:  # orient both upwards
        P2, P3 = P3.copy(), P2.copy()
        nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    edges = CT.crack_tip_elements(mid, P1, P2, P3, nrm)
    e = {name: d for (name, _, _, _), d in zip(CT.EDGES, edges)}
    # the x = 1 edge is free and belongs to triangle 0
    free_edge = [name for name in e if e[name]["FreeFlg"][0] and np.allclose(e[name]["FeMd"][0], [1, 0, 0])]
    assert len(free_edge) == 1
    d = e[free_edge[0]]
    assert np.allclose(d["FeM2Ev"][0], [1, 0, 0], atol=1e-15)
    assert np.allclose(d["FeEv"][0], [0, -1, 0], atol=1e-15)
    h = d["FeM2ELe"][0]
    assert abs(h - (1.0 - mid[0, 0])) < 1e-15
    Dn, Dss, Dds = np.array([0.7, 0.0]), np.array([0.3, 0.0]), np.array([-0.2, 0.0])
    out = CT.stress_intensity(Dn, Dss, Dds, G_, NU_, nrm, edges)
    d = {name: o for (name, _, _, _), o in zip(CT.EDGES, out)}[free_edge[0]]
    c = (G_ * np.sqrt(np.pi)) / (2 * np.sqrt(h) * (1 - NU_))
    # flat upward normal: strike cosine (0, 1, 0), dip cosine (-1, 0, 0); slip in plane = (-Dds, Dss, 0)
    slip = np.array([-Dds[0], Dss[0], 0.0])
    assert abs(d["K1"][0] - c * Dn[0] / 1.834) <= 1e-14 * abs(d["K1"][0])
    assert abs(d["K2"][0] - c * slip @ np.array([1.0, 0, 0]) / 1.834) <= 1e-12
    assert abs(d["K3"][0] - c * slip @ np.array([0.0, -1.0, 0]) * (1 - NU_) / 1.834) <= 1e-12
    se = np.sqrt(d["K1"][0] ** 2 + d["K2"][0] ** 2 + d["K3"][0] ** 2 / (1 - NU_))
    assert abs(d["StrainEnergy"][0] - se) <= 1e-14 * se


def test_in_plane_decomposition_preserves_slip_magnitude():
    P1, P2, P3, *_ = _inclined_penny()
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    rng = np.random.default_rng(2)
    n = P1.shape[0]
    Dn, Dss, Dds = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    edges = CT.stress_intensity(Dn, Dss, Dds, G_, NU_, nrm, CT.crack_tip_elements(mid, P1, P2, P3, nrm))
    for e in edges:
        f = e["FreeFlg"]
        c = (G_ * np.sqrt(np.pi)) / (2 * np.sqrt(e["FeM2ELe"][f]) * (1 - NU_)) / 1.834
        lhs = e["K2"][f] ** 2 + (e["K3"][f] / (1 - NU_)) ** 2
        rhs = c ** 2 * (Dss[f] ** 2 + Dds[f] ** 2)
        assert np.allclose(lhs, rhs, rtol=1e-12, atol=0)
        assert np.all(np.isnan(e["K1"][~f]))
