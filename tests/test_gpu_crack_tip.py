"""Crack-tip edges and stress intensities on the GPU (cdj_crack_tip_elements, cdj_stress_intensity and the fused
device entry point) against the oracle (oracle/crack_tip.py, pinned in tests/test_oracle_crack_tip.py).

Tolerances:
* free-edge flags: identical;
* edge geometry (lengths, midpoints, directions, interior angle, area) and K1: 1e-12 relative to the column's
  largest value;
* K2 and K3 split the in-plane slip by an angle taken as |acos(x1) - acos(x2)| of rotated unit-vector components
  (StressIntensity3D.jl).  Where such a component is within rounding of +-1 (tip edges parallel to the rotated
  x axis, common on structured meshes) acos amplifies a one-ulp difference to ~1e-8, in any implementation, so
  K2 and K3 are compared to 1e-7 of the block's largest value; their magnitude sqrt(K2^2 + (K3 / (1 - nu))^2),
  which does not depend on that angle, and StrainEnergy are compared to 1e-12.
"""
import ctypes

import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshes
from oracle import analytic as AN
from oracle import crack_tip as CT
from oracle import oracle as O
from tests import gocad, stl

pytestmark = pytest.mark.gpu

G_, NU_ = 5.0, 0.25
GEOM = ("FeLe", "FeMd", "FeEv", "FeM2Ev", "FeM2ELe", "IntAng")


def _rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(b)
    if not m.any():
        return 0.0
    scale = max(np.max(np.abs(b[m])), 1e-300)
    return float(np.max(np.abs(a[m] - b[m])) / scale)


def _compare(gpu, ref, nu=NU_):
    for g, r in zip(gpu, ref):
        assert np.array_equal(np.asarray(g.FreeFlg, dtype=bool), r["FreeFlg"])
        for k in GEOM:
            assert _rel(getattr(g, k), r[k]) <= 1e-12, k
        assert _rel(g.Area, r["Area"]) <= 1e-12
        if isinstance(g.K1, np.ndarray):
            assert _rel(g.K1, r["K1"]) <= 1e-12
            assert _rel(g.K2, r["K2"]) <= 1e-7 and _rel(g.K3, r["K3"]) <= 1e-7
            mg = np.sqrt(g.K2 ** 2 + (g.K3 / (1 - nu)) ** 2)
            mr = np.sqrt(r["K2"] ** 2 + (r["K3"] / (1 - nu)) ** 2)
            assert _rel(mg, mr) <= 1e-12
            assert _rel(g.StrainEnergy, r["StrainEnergy"]) <= 1e-12


def _inclined_penny():
    P, T = stl.read_stl("VertPenny-300-EqEdges.stl")
    x = P[:, 1].copy()
    P[:, 1] = P[:, 3]
    P[:, 3] = x
    bfv = 75.0
    c, s = np.cos(np.deg2rad(bfv)), np.sin(np.deg2rad(bfv))
    X, Z = P[:, 1].copy(), P[:, 3].copy()
    P[:, 1] = c * X - s * Z
    P[:, 3] = s * X + c * Z
    return tuple(P[T[:, k] - 1, 1:4].copy() for k in range(3))


def test_inclined_penny_matches_oracle_and_tada():
    """test_TDvsStressIntensity.jl end to end on the GPU: slip (SlipCalculator3D), tip edges, K vs Tada (11 %)."""
    P1, P2, P3 = _inclined_penny()
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G = api.ElasticConstantsCheck(G_, NU_)
    Dn, Dss, Dds, A, b = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 0,
                                              api.Stresses(np.zeros(n), np.zeros(n), np.ones(n), np.zeros(n),
                                                           np.zeros(n), np.zeros(n)), np.zeros(n))
    edges = api.StressIntensity3D(Dn, Dss, Dds, G, nu, nrm, *api.GetCrackTipElements3D(mid, P1, P2, P3, nrm))
    ref = CT.stress_intensity(Dn, Dss, Dds, G, nu, nrm, CT.crack_tip_elements(mid, P1, P2, P3, nrm))
    _compare(edges, ref)
    md = np.vstack([e.FeMd[e.FreeFlg] for e in edges])
    c, s = np.cos(np.deg2rad(-75.0)), np.sin(np.deg2rad(-75.0))
    theta = np.rad2deg(np.arctan2(md[:, 1], c * md[:, 0] - s * md[:, 2]))
    an = AN.tada_inclined_penny_tension(1.0, 15.0, theta, 1, nu)
    worst = []
    for a, key in zip(an, ("K1", "K2", "K3")):
        with np.errstate(divide="ignore", invalid="ignore"):
            p = (100.0 / a) * np.concatenate([getattr(e, key)[e.FreeFlg] for e in edges])
        p[np.abs(p) == np.inf] = np.nan
        worst.append(np.max(np.abs(p[~np.isnan(p)])) - 100)
    print("[gpu-test] inclined penny: %d tip edges, max error K1 %.2f %% K2 %.2f %% K3 %.2f %%" % (md.shape[0],
                                                                                                   *worst))
    assert max(worst) <= 11


def test_holed_penny_random_slip_matches_oracle():
    P1, P2, P3 = meshes.penny(8, 1.0, jitter=0.03, seed=11)
    keep = np.ones(P1.shape[0], dtype=bool)
    keep[[5, 6, 100, 101, 102, 250]] = False
    P1, P2, P3 = P1[keep], P2[keep], P3[keep]
    rot = meshes.random_rotation(np.random.default_rng(4))
    P1, P2, P3 = (P @ rot.T for P in (P1, P2, P3))
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    rng = np.random.default_rng(5)
    Dn, Dss, Dds = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    gpu = api.StressIntensity3D(Dn, Dss, Dds, G_, NU_, nrm, *api.GetCrackTipElements3D(mid, P1, P2, P3, nrm))
    ref = CT.stress_intensity(Dn, Dss, Dds, G_, NU_, nrm, CT.crack_tip_elements(mid, P1, P2, P3, nrm))
    _compare(gpu, ref)
    assert sum(int(e.FreeFlg.sum()) for e in gpu) > 6 * 8


def test_closed_surface_has_no_tip_edges():
    P, T = gocad.read_ts("SphereUniformDistributionON_46Faces.ts")
    P1, P2, P3 = gocad.p1p2p3(P, T)
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    edges = api.GetCrackTipElements3D(mid, P1, P2, P3, nrm)
    for e in edges:
        assert not e.FreeFlg.any()
        assert np.all(np.isnan(e.FeLe)) and np.all(np.isfinite(e.Area))


def test_duplicate_triangle_and_empty_mesh():
    P1 = np.array([[0.0, 0, 0], [0.0, 0, 0], [2.0, 0, 0]])
    P2 = np.array([[1.0, 0, 0], [1.0, 0, 0], [3.0, 0, 0]])
    P3 = np.array([[0.0, 1, 0], [0.0, 1, 0], [2.0, 1, 0]])
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    gpu = api.GetCrackTipElements3D(mid, P1, P2, P3, nrm)
    ref = CT.crack_tip_elements(mid, P1, P2, P3, nrm)
    _compare(gpu, ref)
    # the two coincident triangles share all three edges; the third triangle is isolated
    for e in gpu:
        assert list(e.FreeFlg) == [False, False, True]
    z = np.zeros((0, 3))
    assert all(e.FreeFlg.size == 0 for e in api.GetCrackTipElements3D(z, z, z, z, z))


def test_large_elliptical_crack_device_entry_point():
    """16224 triangles: the free edges are exactly the outer ring's edges (a dictionary count over all edges),
    and the fused device entry point reproduces the host entry points bitwise."""
    torch = pytest.importorskip("torch")
    from cutanddisplacejulia_amd import _lib, learned
    c = meshes.config5()
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    rng = np.random.default_rng(9)
    Dn, Dss, Dds = rng.normal(size=n), rng.normal(size=n), rng.normal(size=n)
    host = api.StressIntensity3D(Dn, Dss, Dds, G_, NU_, nrm, *api.GetCrackTipElements3D(mid, P1, P2, P3, nrm))
    count = {}
    keys = []
    for A, B in ((P1, P2), (P1, P3), (P2, P3)):
        ks = [tuple(sorted([tuple(a), tuple(b)])) for a, b in zip(A, B)]
        keys.append(ks)
        for k in ks:
            count[k] = count.get(k, 0) + 1
    for e, ks in zip(host, keys):
        assert np.array_equal(e.FreeFlg, np.array([count[k] == 1 for k in ks]))
    nfree = sum(int(e.FreeFlg.sum()) for e in host)
    print("[gpu-test] elliptical crack: %d tip edges of %d triangles" % (nfree, n))
    assert nfree == sum(1 for v in count.values() if v == 1)
    d = {k: learned.to_device_cols(v) for k, v in dict(P1=P1, P2=P2, P3=P3, N=nrm, M=mid).items()}
    dD = [torch.from_numpy(v).to("cuda:0") for v in (Dn, Dss, Dds)]
    flags = torch.empty(3 * n, dtype=torch.uint8, device="cuda:0")
    geom = torch.empty(3 * n * 13, dtype=torch.float64, device="cuda:0")
    K = torch.empty(3 * n * 4, dtype=torch.float64, device="cuda:0")
    L = _lib.load()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dp = [ctypes.c_void_p(t.data_ptr()) for t in (d["M"], d["P1"], d["P2"], d["P3"], d["N"])]
    _lib.check(L.cdj_stress_intensity_dev(_lib.ctx(), *dp, n, *[ctypes.c_void_p(t.data_ptr()) for t in dD], G_, NU_,
                                          ctypes.c_void_p(flags.data_ptr()), ctypes.c_void_p(geom.data_ptr()),
                                          ctypes.c_void_p(K.data_ptr()), sp))
    flags, geom, K = flags.cpu().numpy(), geom.cpu().numpy(), K.cpu().numpy()
    for s, e in enumerate(host):
        g = geom[s * n * 13:(s + 1) * n * 13].reshape(13, n)
        k = K[s * n * 4:(s + 1) * n * 4].reshape(4, n)
        assert np.array_equal(flags[s * n:(s + 1) * n].astype(bool), e.FreeFlg)
        for got, want in ((g[0], e.FeLe), (g[1:4].T, e.FeMd), (g[10], e.FeM2ELe), (g[12], e.Area), (k[0], e.K1),
                          (k[1], e.K2), (k[2], e.K3), (k[3], e.StrainEnergy)):
            assert np.array_equal(got, want, equal_nan=True)
