"""Frictional slip on the GPU (cdj_slip_friction through api.SlipCalculator3D with MixedBoundaryConditionsFriction)
against the oracle (oracle/friction.py, pinned in tests/test_oracle_friction.py).

* The reference's own case, test_TDvsBurgmannLinearFrictionCrack.jl, with its pass criteria (36 % pointwise away from
  the tips, 1400 % summed) and the oracle's slip of the same problem (the complementarity problem has one
  solution: 1e-8 of the largest slip).
* Cohesion far above the driving shear: no slip, and the opening equals the normal-block solve (1e-9).
* No element able to overcome its frictional strength: zero slip without a solve (SlipCalculator3D.jl:52-59).
* A buried half-space fault with friction and cohesion, GPU vs oracle (1e-8).
"""
import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshes
from oracle import friction as FR
from oracle import oracle as O
from tests.test_oracle_friction import SCL, blade, burgmann_errors

pytestmark = pytest.mark.gpu


def _bc(stress, mu, Sf):
    n = stress[0].shape[0]
    z = np.zeros(n)
    return api.MixedBoundaryConditionsFriction(
        api.MixedBoundaryConditions(api.Stresses(*stress), api.Tractions(z, z, z)), api.FrictionParameters(mu, Sf))


def test_burgmann_friction_crack_matches_oracle_and_criteria():
    P1, P2, P3, nrm, mid, n, G, nu, lam, stress = blade()
    Sf = np.abs(mid[:, 0])
    DnF, DssF, DdsF = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 0, _bc(stress, 0.0, Sf), np.zeros(n))
    rep = dict(api.LAST_REPORT)
    Dn, Dss, Dds, _, _ = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 0, api.Stresses(*stress), np.zeros(n))
    worst, sums = burgmann_errors(mid, Dss, DssF, G, nu)
    print("[gpu-test] burgmann: %d Newton its, worst %.1f %% / %.1f %%, sums %.0f / %.0f, %.1f ms" % (
        rep["iterations"], *worst, *sums, rep["solve_ms"]))
    assert max(worst) <= 36 and max(sums) <= 1400
    z = np.zeros(n)
    on, oss, ods, info = FR.slip_friction(P1, P2, P3, nu, G, lam, mid, nrm, 0, stress, (z, z, z), 0.0, Sf)
    ref = np.concatenate([on, oss, ods])
    got = np.concatenate([DnF, DssF, DdsF])
    assert np.max(np.abs(got - ref)) <= 1e-8 * np.max(np.abs(ref))


def test_locked_shear_reduces_to_normal_block():
    P1, P2, P3, nrm, mid, n, G, nu, lam, _ = blade()
    z = np.zeros(n)
    stress = (z, np.full(n, 0.3 * SCL), z, np.full(n, SCL), z, z)
    Dn, Dss, Dds = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 0, _bc(stress, 0.0, 1e3 * SCL), z)
    A = O.assemble_A(O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, 0))
    Tn, Tds, Tss = O.normal_shear_tractions(stress, nrm)
    ref = np.linalg.solve(A[:n, :n], Tn)
    assert np.max(np.abs(Dn - ref)) <= 1e-9 * np.max(np.abs(ref))
    assert np.max(np.abs(Dss)) <= 1e-9 * np.max(np.abs(ref)) and np.max(np.abs(Dds)) <= 1e-9 * np.max(np.abs(ref))


def test_frictional_strength_not_overcome_gives_zero_slip():
    P1, P2, P3, nrm, mid, n, G, nu, lam, _ = blade()
    z = np.zeros(n)
    stress = (z, np.full(n, -10 * SCL), z, np.full(n, SCL), z, z)  # strong compression, mu = 0.6
    Dn, Dss, Dds = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 0, _bc(stress, 0.6, 0.0), z)
    assert not Dn.any() and not Dss.any() and not Dds.any()


def test_halfspace_fault_with_friction_matches_oracle():
    P1, P2, P3 = meshes.rect_fault(8, 6, strike=30.0, dip=60.0, length=2.0, width=1.5, top_depth=0.5)
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    rng = np.random.default_rng(2)
    z = np.zeros(n)
    stress = (np.full(n, -0.1), np.full(n, -0.05), np.full(n, -0.1), np.full(n, 0.3), z, np.full(n, 0.2))
    mu, Sf = 0.3 + 0.1 * rng.random(n), 0.02 * rng.random(n)
    Dn, Dss, Dds = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 1, _bc(stress, mu, Sf), z)
    on, oss, ods, info = FR.slip_friction(P1, P2, P3, nu, G, lam, mid, nrm, 1, stress, (z, z, z), mu, Sf)
    ref = np.concatenate([on, oss, ods])
    got = np.concatenate([Dn, Dss, Dds])
    print("[gpu-test] half-space friction: %d its (oracle %d), max slip %.3e" % (api.LAST_REPORT["iterations"],
                                                                                 info["iterations"], np.abs(ref).max()))
    assert np.abs(ref).max() > 0
    assert np.max(np.abs(got - ref)) <= 1e-8 * np.max(np.abs(ref))
