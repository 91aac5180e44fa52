"""The oracle on the reference's own validation cases and meshes (tests/golden), with the reference's pass
criteria (CPU; no GPU): test_TDvsEshelbyPennyCrack.jl and test_TDvsMogi.jl.  The oracle's collocation matrix
is solved with numpy (LAPACK), the fixed-element system by least squares as Julia's `\\` does.
"""
import numpy as np

from oracle import analytic as AN
from oracle import oracle as O
from tests import gocad


def _pct(analytic, numeric):
    with np.errstate(divide="ignore", invalid="ignore"):
        p = (100.0 / np.asarray(analytic)) * np.asarray(numeric)
    p[np.isinf(p)] = np.nan
    return p


def _ok(p, lim):
    d = np.abs(p - 100.0)
    return not np.any(d[~np.isnan(d)] > lim)


def _elastic(G=1.0, nu=0.25):
    lam = 2 * G * nu / (1 - 2 * nu)
    return G, nu, lam


def test_oracle_eshelby_penny_crack():
    P, T = gocad.read_ts("CircleMesh_1a_500Faces.ts")
    P1, P2, P3 = gocad.p1p2p3(P, T)
    n = T.shape[0]
    G, nu, lam = _elastic()
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    A = O.assemble_A(O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, 0))
    z, o = np.zeros(n), np.ones(n)
    x_s = np.linalg.solve(A, np.concatenate([z, o, z]))   # b = [Tn; Tss; Tds], Tss = 1
    x_n = np.linalg.solve(A, np.concatenate([o, z, z]))   # Tn = 1
    Dss, Dds, Dn = x_s[n:2 * n], x_s[2 * n:], x_n[:n]
    rho = np.hypot(mid[:, 0], mid[:, 1])
    good = rho < 0.9
    us = AN.eshelby1957_penny_profile(G, nu, 0.0, 1.0, rho, 1.0)
    un = AN.eshelby1957_penny_profile(G, nu, 1.0, 0.0, rho, 1.0)
    assert _ok(_pct(us, np.sqrt(Dss ** 2 + Dds ** 2))[good], 11)
    assert _ok(_pct(un, Dn)[good], 11)


def test_oracle_mogi_with_fixed_elements():
    P, T = gocad.read_ts("SphereUniformDistributionON_46Faces.ts")
    PF, TF = gocad.read_ts("SphereFix4.ts")
    PF[:, 1:4] *= 3
    P, T, fixed = gocad.append(P, T, PF, TF)
    depth = 50.0
    P[:, 3] -= depth
    P1, P2, P3 = gocad.p1p2p3(P, T)
    n = T.shape[0]
    G, nu, lam = _elastic()
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    fx = fixed == 1
    free = ~fx
    nf = int(free.sum())
    pts = mid - nrm * 1e-12
    # free rows: traction at the free collocation points, row i with element i's plane (the reference's
    # ConvertStrainInfMatsToTraction on the free strain rows); fixed elements are appended, so these coincide
    assert np.all(free[:nf]) and np.all(fx[nf:])
    Tfull = O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, 1)
    ones = np.ones(n)
    _, do = O.td(pts[fx, 0], pts[fx, 1], pts[fx, 2], P1, P2, P3, ones, ones, ones, nu, G, 1, 0, 1)
    blocks = {}
    for k, name in enumerate(O.TRACTION_FIELDS):
        full = np.zeros((n, n))
        full[free] = Tfull[name][free]
        full[fx] = do[O.DISP_FIELDS[k]]
        blocks[name] = full[:, free]
    A = O.assemble_A(blocks)
    assert A.shape == (3 * n, 3 * nf)
    Tn = np.ones(n)
    Tn[fx] = 0.0
    b = np.concatenate([Tn, np.zeros(n), np.zeros(n)])
    x = np.linalg.lstsq(A, b, rcond=None)[0]
    Dn, Dss, Dds = x[:nf], x[nf:2 * nf], x[2 * nf:]
    assert Dn.max() - Dn.min() <= 0.05 / G
    X = np.arange(0.0, 101.0)
    so, dd = O.td_field(X, np.zeros_like(X), np.zeros_like(X), P1[free], P2[free], P3[free], Dss, Dds, Dn, nu, G,
                        1, 0, 1)
    ux, uy, uz = AN.mogi1958(depth, X, np.zeros_like(X), 1.0, 1.0, nu, G)
    for a, d in zip((ux, uy, uz), dd):
        assert _ok(_pct(a, d), 20)
