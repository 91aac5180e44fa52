"""Constant fluid volume on the GPU (cdj_slip_fluid_volume through api.SlipCalculator3D with
MixedBoundaryConditionsFluidVolume) against the oracle (oracle/fluid_volume.py, pinned in
tests/test_oracle_fluid_volume.py).

* A penny crack under fluid pressure alone and with a normal compression: the target volume is held (2e-3, the
  bound Brent's tolerance allows), and pressure and slip equal the oracle's (1e-8: both run the same minimiser on
  the same volumes, and each complementarity problem has one solution).
* Part of the penny flagged as free surface (flag 0): the reference's free-surface branch, against the oracle.
* The reference's own driver, test_DippingCrackConstantVolume.jl (its mesh, 75-degree dip, radius 1500 m at 2 km in
  a half-space, lithostatic stress and buoyancy): the compression keeps the crack shut over the first points
  Brent tries, so the reference's search stalls there and returns zero slip; GPU and oracle agree on that outcome.
* More than one fracture ends in the reference's error.
"""
import numpy as np
import pytest

from cutanddisplacejulia_amd import api
from oracle import fluid_volume as FV
from tests import stl
from tests.test_oracle_fluid_volume import G_, NU_, normal_stress, penny

pytestmark = pytest.mark.gpu


def _bc(stress, tractions, V):
    return api.MixedBoundaryConditionsFluidVolume(
        api.MixedBoundaryConditions(api.Stresses(*stress), api.Tractions(*tractions)), V)


def _compare(P1, P2, P3, nu, G, lam, mid, nrm, hs, stress, tractions, flag, V, tol=1e-8):
    Dn, Dss, Dds = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, hs, _bc(stress, tractions, V), flag)
    rep = dict(api.LAST_REPORT)
    on, oss, ods, info = FV.slip_fluid_volume(P1, P2, P3, nu, G, lam, mid, nrm, hs, stress, tractions, flag, V)
    print("[gpu-test] fluid volume: pressure %.9e (oracle %.9e), %d evaluations (oracle %d), %d Newton steps, "
          "volume misfit %.2e, %.1f ms" % (rep["pressure"], info["pressure"], rep["iterations"], info["evaluations"],
                                           rep["near_entries"], rep["residual"], rep["solve_ms"]))
    assert rep["pressure"] == pytest.approx(info["pressure"], rel=tol, abs=1e-300)
    ref = np.concatenate([on, oss, ods])
    got = np.concatenate([Dn, Dss, Dds])
    scale = max(np.max(np.abs(ref)), 1e-300)
    assert np.max(np.abs(got - ref)) <= tol * scale
    return Dn, Dss, Dds, rep, info


@pytest.mark.parametrize("s", [0.0, 0.004])
def test_penny_volume_matches_oracle(s):
    P1, P2, P3, nrm, mid = penny()
    n = P1.shape[0]
    lam = 2 * G_ * NU_ / (1 - 2 * NU_)
    V = np.pi * 0.01
    z = np.zeros(n)
    Dn, _, _, rep, info = _compare(P1, P2, P3, NU_, G_, lam, mid, nrm, 0, normal_stress(nrm, s), (z, z, z),
                                   np.ones(n), V)
    assert rep["residual"] < 2e-3 and abs(info["volume"] - V) / V < 2e-3
    assert np.all(Dn > 0)


def test_free_surface_flags_match_oracle():
    P1, P2, P3, nrm, mid = penny()
    n = P1.shape[0]
    lam = 2 * G_ * NU_ / (1 - 2 * NU_)
    flag = np.ones(n)
    flag[np.linalg.norm(mid, axis=1) > 0.8] = 0  # outer ring as free-surface elements
    z = np.zeros(n)
    _compare(P1, P2, P3, NU_, G_, lam, mid, nrm, 0, normal_stress(nrm, 0.0), (z, z, z), flag, np.pi * 0.005)


def test_reference_dipping_crack_case():
    P, T = stl.read_stl("VertPenny-300-EqEdges.stl")
    x = P[:, 1].copy()
    P[:, 1] = P[:, 3]
    P[:, 3] = x
    bfv = 90.0 - (-75.0)
    c, s = np.cos(np.deg2rad(bfv)), np.sin(np.deg2rad(bfv))
    Y, Z = P[:, 2].copy(), P[:, 3].copy()
    P[:, 2] = c * Y - s * Z
    P[:, 3] = s * Y + c * Z
    R = 1500.0
    P[:, 1:4] *= R
    P[:, 3] -= 2000.0
    P1, P2, P3 = (P[T[:, k] - 1, 1:4].copy() for k in range(3))
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    n = P1.shape[0]
    K, E, lam, nu, G = api.ElasticConstantsCheck(2.0e9, 0.25)
    g = 9.81
    W = (2900 * g) * mid[:, 2]
    z = np.zeros(n)
    stress = (W, W, W, z, z, z)
    tractions = (mid[:, 2] * (g * (2900 - 2600)), z, z)
    V = np.pi * R ** 2 * 2
    Dn, Dss, Dds, rep, info = _compare(P1, P2, P3, nu, G, lam, mid, nrm, 1, stress, tractions, np.ones(n), V)
    assert info["volume"] == 0.0 and not np.any(Dn)


def test_more_than_one_fracture_is_the_reference_error():
    P1, P2, P3, nrm, mid = penny()
    n = P1.shape[0]
    flag = np.ones(n)
    flag[: n // 2] = 2
    z = np.zeros(n)
    with pytest.raises(api.CdjError, match="Check scl here"):
        api.SlipCalculator3D(P1, P2, P3, NU_, G_, 0.4, mid, nrm, 0, _bc((z,) * 6, (z, z, z), 1.0), flag)
