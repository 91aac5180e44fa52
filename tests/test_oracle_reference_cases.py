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


def _mogi_mesh():
    P, T = gocad.read_ts("SphereUniformDistributionON_46Faces.ts")
    PF, TF = gocad.read_ts("SphereFix4.ts")
    PF[:, 1:4] *= 3
    P, T, fixed = gocad.append(P, T, PF, TF)
    P[:, 3] -= 50.0
    P1, P2, P3 = gocad.p1p2p3(P, T)
    return P1, P2, P3, fixed


def _traction_numpy(strain, lam, G, nrm, ss, ds):
    """Hooke's law, then the traction on each row's plane (HookesLaw3DStrain2Stress!.jl:44-66,
    CalculateNormalTraction3D!.jl:37-59, TractionVectorCartesianComponents3D!.jl:43-53,
    CalculateTractionInChosenDirection3D!.jl:48-69), numpy over rows; strain is (rows, 6)."""
    e = strain
    tr = e[:, 0] + e[:, 1] + e[:, 2]
    S = np.zeros((e.shape[0], 3, 3))
    for a, b, c in ((0, 0, 0), (1, 1, 1), (2, 2, 2)):
        S[:, a, b] = 2 * G * e[:, c] + lam * tr
    for a, b, c in ((0, 1, 3), (0, 2, 4), (1, 2, 5)):
        S[:, a, b] = S[:, b, a] = 2 * G * e[:, c]
    t = np.einsum("rab,rb->ra", S, nrm)
    tn = np.einsum("ra,ra->r", t, nrm)
    shear = t - tn[:, None] * nrm
    return tn, np.einsum("ra,ra->r", shear, ss), np.einsum("ra,ra->r", shear, ds)


def test_oracle_fixed_system_matches_composition():
    """orc_traction_influence_fixed against the system composed from the oracle's plain traction matrix and TD
    displacements (the reference's own Mogi mesh, fixed elements appended, so every free row has its own plane)."""
    P1, P2, P3, fixed = _mogi_mesh()
    n = P1.shape[0]
    G, nu, lam = _elastic()
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    fx = fixed == 1
    free = ~fx
    pts = mid - nrm * 1e-12
    Tfull = O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, 1)
    ones = np.ones(n)
    _, do = O.td(pts[fx, 0], pts[fx, 1], pts[fx, 2], P1, P2, P3, ones, ones, ones, nu, G, 1, 0, 1)
    Tfix = O.traction_influence_fixed(nrm, mid, P1, P2, P3, nu, G, lam, 1, fixed)
    for k, name in enumerate(O.TRACTION_FIELDS):
        full = np.zeros((n, n))
        full[free] = Tfull[name][free]
        full[fx] = do[O.DISP_FIELDS[k]]
        assert np.array_equal(Tfix[name], full[:, free]), name


def test_oracle_fixed_rows_take_the_plane_of_their_free_count():
    """Interleaved fixed elements: the t-th free row is the traction of the strain at free element fr[t]'s
    collocation point on element t's plane (the reference passes the full normal arrays with the free strain rows
    only); checked against a numpy restatement of the strain -> traction conversion."""
    P1, P2, P3, _ = _mogi_mesh()          # a sphere and its inner sphere: the planes vary row to row
    n = P1.shape[0]
    fixed = np.zeros(n)
    fixed[::5] = 1
    fx = fixed == 1
    fr = np.flatnonzero(~fx)
    nf = fr.size
    G, nu, lam = _elastic()
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    pts = mid - nrm * 1e-12
    ones = np.ones(n)
    so, _ = O.td(pts[fr, 0], pts[fr, 1], pts[fr, 2], P1, P2, P3, ones, ones, ones, nu, G, 0, 1, 1)
    _, do = O.td(pts[fx, 0], pts[fx, 1], pts[fx, 2], P1, P2, P3, ones, ones, ones, nu, G, 1, 0, 1)
    ss, ds = O.ss_ds_dirs(nrm)
    Tfix = O.traction_influence_fixed(nrm, mid, P1, P2, P3, nu, G, lam, 1, fixed)
    for m, mode in enumerate(("Dn", "Dss", "Dds")):
        want = np.zeros((3, nf, nf))              # (traction, free row, column j = free source fr[j])
        for j in range(nf):
            e = np.stack([so[c + mode][:, fr[j]] for c in ("exx", "eyy", "ezz", "exy", "exz", "eyz")], axis=1)
            want[:, :, j] = _traction_numpy(e, lam, G, nrm[:nf], ss[:nf], ds[:nf])
        scale = np.max(np.abs(want))              # one scale per slip mode's three tractions
        for q in range(3):
            got = Tfix[O.TRACTION_FIELDS[3 * m + q]]
            assert np.max(np.abs(got[fr] - want[q])) <= 1e-12 * scale
            assert np.array_equal(got[fx], do[O.DISP_FIELDS[3 * m + q]][:, fr])
    # the quirk matters here: with each row's own plane the free rows differ
    own = O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, 1)
    assert np.max(np.abs(Tfix["DnTn"][fr] - own["DnTn"][np.ix_(fr, fr)])) > 1e-3
