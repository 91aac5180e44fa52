"""GPU parity of the TD kernels against the CPU FP64 oracle, through the C ABI (libcdjhip.so).

Tolerance (FP64): for every influence matrix, max|gpu - oracle| <= 1e-12 * max|oracle| over that matrix (a
normwise relative bound, since individual entries can be tiny after the cancellation between the three angular
dislocations), and NaN (on-edge, in-plane observation points, TD.jl:808-818, 1573-1594) exactly where the
oracle has NaN.
"""
import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshes
from oracle import analytic as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _cmp(g, o, tol=TOL):
    g = np.asarray(g)
    o = np.asarray(o)
    assert g.shape == o.shape
    assert np.array_equal(np.isnan(g), np.isnan(o))
    m = ~np.isnan(o)
    scale = np.max(np.abs(o[m])) if m.any() else 0.0
    err = np.max(np.abs(g[m] - o[m])) if m.any() else 0.0
    assert err <= tol * max(scale, 1e-300), (err, scale, err / max(scale, 1e-300))
    return err / max(scale, 1e-300)


def _cmp_group(gs, os_, tol=TOL):
    """Normwise over a group of fields (the six strain components, or the three displacement components): on the
    free surface the traction-free components exz, eyz vanish to rounding, so their own maximum is no scale."""
    scale = max(np.max(np.abs(np.asarray(o)[~np.isnan(o)]), initial=0.0) for o in os_)
    worst = 0.0
    for g, o in zip(gs, os_):
        g, o = np.asarray(g), np.asarray(o)
        assert g.shape == o.shape
        assert np.array_equal(np.isnan(g), np.isnan(o))
        m = ~np.isnan(o)
        err = np.max(np.abs(g[m] - o[m])) if m.any() else 0.0
        assert err <= tol * max(scale, 1e-300), (err, scale, err / max(scale, 1e-300))
        worst = max(worst, err / max(scale, 1e-300))
    return worst


def _small_fault(n=6):
    P1, P2, P3 = meshes.rect_fault(n, n, 60.0, 45.0, 5.0, 5.0, 0.5)
    return P1, P2, P3


def _obs(P1, P2, P3, k=12, z=0.0, seed=0):
    rng = np.random.default_rng(seed)
    X, Y, Z = meshes.grid_obs(k, 8.0, z)
    cen = (P1 + P2 + P3) / 3
    nrm = np.cross(P2 - P1, P3 - P1)
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    off = cen + 0.05 * nrm  # just off each triangle
    deep = rng.uniform(-6, 6, (40, 3))
    deep[:, 2] = -np.abs(deep[:, 2]) - 0.05
    X = np.concatenate([X, cen[:, 0], off[:, 0], deep[:, 0]])
    Y = np.concatenate([Y, cen[:, 1], off[:, 1], deep[:, 1]])
    Z = np.concatenate([Z, cen[:, 2], off[:, 2], deep[:, 2]])
    return X, Y, Z


@pytest.mark.parametrize("hs", [0, 1])
def test_td_influence_matches_oracle(hs):
    P1, P2, P3 = _small_fault(6)
    X, Y, Z = _obs(P1, P2, P3)
    n = P1.shape[0]
    rng = np.random.default_rng(5)
    Dss, Dds, Dn = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    Dn[3] = 0.0  # a zero-slip mode contributes exactly zero
    so, do = O.td(X, Y, Z, P1, P2, P3, Dss, Dds, Dn, 0.25, 1.0, 1, 1, hs)
    S, D = api.TD(X, Y, Z, P1, P2, P3, Dss, Dds, Dn, 0.25, 1.0, 1, 1, hs)
    worst = 0.0
    for k, name in zip(O.STRAIN_FIELDS, api.STRAIN_FIELDS):
        worst = max(worst, _cmp(getattr(S, name), so[k]))
    for k, name in zip(O.DISP_FIELDS, api.DISP_FIELDS):
        worst = max(worst, _cmp(getattr(D, name), do[k]))
    print("hs=%d worst normwise relative error %.3e" % (hs, worst))


def test_td_config1_fullspace():
    c = meshes.config1()
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    ones = np.ones(n)
    so, do = O.td(c["X"], c["Y"], c["Z"], P1, P2, P3, ones, ones, ones, 0.25, 1.0, 1, 1, 0)
    S, D = api.TD(c["X"], c["Y"], c["Z"], P1, P2, P3, ones, ones, ones, 0.25, 1.0, 1, 1, 0)
    for k, name in zip(O.STRAIN_FIELDS, api.STRAIN_FIELDS):
        _cmp(getattr(S, name), so[k])
    for k, name in zip(O.DISP_FIELDS, api.DISP_FIELDS):
        _cmp(getattr(D, name), do[k])


def test_td_field_okada():
    P1, P2, P3, md = A.okada_test_fault()
    xx = np.linspace(-10, 10, 50)
    X, Y = np.meshgrid(xx, xx)
    Z = np.zeros_like(X)
    Dds, Dn, Dss = 2.0, 0.6, 0.5
    S, D = api.TD(X, Y, Z, P1, P2, P3, [Dss] * 2, [Dds] * 2, [Dn] * 2, 0.25, 1.0, 1, 1, 1, api.Strains(), api.Disps())
    assert S.εxx.shape == X.shape
    tot = np.sqrt(Dds ** 2 + Dss ** 2)
    rake = 90 - np.degrees(np.arctan(Dss / Dds))
    ok = A.okada85_surface(X, Y, md, 60, 45, 5, 5, rake, tot, Dn, 0.25)
    mine = [D.Ux, D.Uy, D.Uz, S.εxx, S.εyy, S.εzz, S.εxy, S.εxz, S.εyz]
    for a, b in zip(mine, ok):
        assert np.max(np.abs(a - b)) < 1e-13
    so, do = O.td_field(X.ravel(order="F"), Y.ravel(order="F"), Z.ravel(order="F"), P1, P2, P3, [Dss] * 2,
                        [Dds] * 2, [Dn] * 2, 0.25, 1.0, 1, 1, 1)
    _cmp_group([a.ravel(order="F") for a in mine[:3]], do)
    _cmp_group([a.ravel(order="F") for a in mine[3:]], so)


def test_config2_full_mesh_matches_okada():
    """The half-space benchmark fault at full size (5000 triangles; the 100 x 100 free-surface grid over
    [-10, 10]^2): with uniform slip the summed TD field equals Okada's (1985) solution for the same rectangle, which
    pins the main, image and free-surface terms at the benchmark's scale.  The fault is the reference's Okada test
    rectangle (test_TDvsOkada.jl) with its top at depth 0.5."""
    c = meshes.config2(50)
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    assert n == 5000
    X = c["X"].reshape(100, 100, order="F")
    Y = c["Y"].reshape(100, 100, order="F")
    Z = c["Z"].reshape(100, 100, order="F")
    Dds, Dn, Dss = 2.0, 0.6, 0.5
    S, D = api.TD(X, Y, Z, P1, P2, P3, [Dss] * n, [Dds] * n, [Dn] * n, 0.25, 1.0, 1, 1, 1, api.Strains(),
                  api.Disps())
    tot = np.sqrt(Dds ** 2 + Dss ** 2)
    rake = 90 - np.degrees(np.arctan(Dss / Dds))
    centroid_depth = 0.5 + np.sin(np.deg2rad(45.0)) * 5.0 / 2
    ok = A.okada85_surface(X, Y, centroid_depth, 60, 45, 5, 5, rake, tot, Dn, 0.25)
    mine = [D.Ux, D.Uy, D.Uz, S.εxx, S.εyy, S.εzz, S.εxy, S.εxz, S.εyz]
    # displacements and strains each against their group's largest value (exz, eyz vanish on the free surface)
    for grp in (range(0, 3), range(3, 9)):
        scale = max(np.max(np.abs(ok[k])) for k in grp)
        for k in grp:
            assert np.max(np.abs(mine[k] - ok[k])) <= 1e-10 * scale, k


def test_td_field_deterministic():
    P1, P2, P3 = _small_fault(8)
    X, Y, Z = _obs(P1, P2, P3, 16)
    n = P1.shape[0]
    ones = np.ones(n)
    r1 = api.TD(X, Y, Z, P1, P2, P3, ones, ones, ones, 0.25, 1.0, 1, 1, 1, api.Strains(), api.Disps())
    r2 = api.TD(X, Y, Z, P1, P2, P3, ones, ones, ones, 0.25, 1.0, 1, 1, 1, api.Strains(), api.Disps())
    for a, b in zip([r1[0].εxx, r1[0].εyz, r1[1].Uz], [r2[0].εxx, r2[0].εyz, r2[1].Uz]):
        assert np.array_equal(a, b)
    so, do = O.td_field(X, Y, Z, P1, P2, P3, ones, ones, ones, 0.25, 1.0, 1, 1, 1)
    _cmp_group([r1[1].Ux, r1[1].Uy, r1[1].Uz], do)
    _cmp_group([r1[0].εxx, r1[0].εyy, r1[0].εzz, r1[0].εxy, r1[0].εxz, r1[0].εyz], so)


@pytest.mark.parametrize("hs", [0, 1])
def test_traction_influence_matches_oracle(hs):
    if hs:
        P1, P2, P3 = _small_fault(6)
    else:
        c = meshes.config1()
        P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    n = P1.shape[0]
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    To = O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, hs)
    Tg = api.CalculateInfluenceMatrices3D(nrm, mid, P1, P2, P3, nu, G, lam, np.zeros(n), hs, n)
    for name in api.TRACTION_FIELDS:
        _cmp(getattr(Tg, name), To[name])


def test_flat_halfspace_triangle_reproduces_reference_rebinding():
    """A half-space triangle lying in z = 0 follows the reference's local-rebinding behaviour (TD.jl:759-763,
    1518-1525): the oracle restates it, and the GPU must agree with the oracle."""
    P1 = np.array([[-1.0, -1.0, 0.0], [0.0, 0.0, -1.0]])
    P2 = np.array([[1.0, -1.0, 0.0], [1.0, 0.0, -1.5]])
    P3 = np.array([[0.0, 1.0, 0.0], [0.0, 1.0, -1.2]])
    X, Y, Z = meshes.grid_obs(9, 3.0, -0.3)
    ones = np.ones(2)
    so, do = O.td(X, Y, Z, P1, P2, P3, ones, ones, ones, 0.25, 1.0, 1, 1, 1)
    S, D = api.TD(X, Y, Z, P1, P2, P3, ones, ones, ones, 0.25, 1.0, 1, 1, 1)
    for k, name in zip(O.STRAIN_FIELDS, api.STRAIN_FIELDS):
        _cmp(getattr(S, name), so[k])
    for k, name in zip(O.DISP_FIELDS, api.DISP_FIELDS):
        _cmp(getattr(D, name), do[k])


def test_nan_on_triangle_edge():
    P1 = np.array([[-1.0, -1.0, -2.0]])
    P2 = np.array([[1.0, -1.0, -2.0]])
    P3 = np.array([[0.0, 1.0, -2.0]])
    X = np.array([0.0, 0.0, 0.3])
    Y = np.array([-1.0, -1.0, 0.2])
    Z = np.array([-2.0, -2.5, -1.0])  # first point on the P1-P2 edge, in the TD plane
    S, D = api.TD(X, Y, Z, P1, P2, P3, [1.0], [1.0], [1.0], 0.25, 1.0, 1, 1, 0)
    assert np.isnan(S.εxxDn[0, 0]) and np.isnan(D.DnUx[0, 0])
    assert np.isfinite(S.εxxDn[1:, 0]).all() and np.isfinite(D.DnUx[1:, 0]).all()


def test_errors_match_reference_messages():
    with pytest.raises(api.CdjError, match="Z coordinates must be negative"):
        api.TD([0.0], [0.0], [0.5], [0, 0, -1.0], [1, 0, -1.0], [0, 1, -1.0], [1.0], [1.0], [1.0], 0.25, 1.0, 0, 1, 1)
    with pytest.raises(api.CdjError, match="same dimensions"):
        api.TD([0.0], [0.0], [-0.5], [0, 0, -1.0], [1, 0, -1.0], [0, 1, -1.0], [1.0, 2.0], [1.0], [1.0], 0.25, 1.0,
               0, 1, 1)
