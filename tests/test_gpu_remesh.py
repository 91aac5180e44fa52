"""A growth-loop step end to end: the reference's penny mesh (tests/golden/CircleMesh_1a_500Faces.off, radius 1)
remeshed by cdj_remesh_isotropic (the CGAL step of test/test_PropagationTest.jl:151-561) to the edge length
GetDesiredEdgeLength gives for 1500 triangles, then SlipCalculator3D under remote tension sigma_zz = 1, the tip
edges and K_I (GetCrackTipElements3D, StressIntensity3D) on the GPU, against the penny closed form
K_I = 2 sigma sqrt(a / pi) (Tada, Paris & Irwin; Tada_StrIntInclinedPennyTension.jl at 90 degrees).

* the remeshed penny's K_I is within 5 % of the closed form on average over the tip edges (the input mesh's is
  20 % low), and its worst tip edge is better than the input mesh's.  The reference follows its remeshing with
  IsosceliseEdgeTrisNew (not restated here), which reshapes the tip triangles for the tip-element estimate of
  StressIntensity3D.jl; its 11 % per-edge criterion (test_TDvsStressIntensity.jl) is met on the equal-edged
  inclined penny in tests/test_gpu_crack_tip.py;
* a growth step: the input scaled to radius 1.3 and remeshed at the same target edge length has 1.5-1.9 times the
  triangles (the area grows 1.69 times), and its mean K_I is within 10 % of the closed form at a = 1.3, like the
  radius-1 crack's (the tip-element estimate's bias depends on the tip triangles' shape, which the isotropic
  remesher does not fix).
"""
import os

import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshio, remesh
from oracle import analytic as AN

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _k1(P, T):
    P1, P2, P3 = meshio.CreateP1P2P3(T, P)
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G = api.ElasticConstantsCheck(5.0, 0.25)
    z, o = np.zeros(n), np.ones(n)
    Dn, Dss, Dds, A, b = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 0, api.Stresses(z, z, o, z, z, z), z)
    edges = api.StressIntensity3D(Dn, Dss, Dds, G, nu, nrm, *api.GetCrackTipElements3D(mid, P1, P2, P3, nrm))
    return np.concatenate([e.K1[e.FreeFlg] for e in edges]), nu


def _penny():
    return meshio.OFFReader(os.path.join(GOLDEN, "CircleMesh_1a_500Faces.off"))


def test_remeshed_penny_meets_tada():
    P, T = _penny()
    P1, P2, P3 = meshio.CreateP1P2P3(T, P)
    L, _ = api.GetDesiredEdgeLength(P1, P2, P3, 1500)
    Q, U = remesh.PolygonRemeshing(P, T, L)
    k_in, nu = _k1(P, T)
    k_re, _ = _k1(Q, U)
    an = AN.tada_inclined_penny_tension(1.0, 90.0, 0.0, 1.0, nu)[0]
    e_in = np.max(np.abs(100 * k_in / an - 100))
    e_re = np.max(np.abs(100 * k_re / an - 100))
    print("[gpu-test] penny K_I vs Tada: input mesh %d triangles, %d tip edges, worst %.2f %%, mean %.2f %%; "
          "remeshed %d triangles, %d tip edges, worst %.2f %%, mean %.2f %%" % (
              T.shape[0], k_in.size, e_in, np.mean(100 * k_in / an - 100), U.shape[0], k_re.size, e_re,
              np.mean(100 * k_re / an - 100)))
    assert abs(np.mean(100 * k_re / an - 100)) <= 5
    assert e_re < e_in


def test_growth_step_remesh_and_solve():
    P, T = _penny()
    P1, P2, P3 = meshio.CreateP1P2P3(T, P)
    L, _ = api.GetDesiredEdgeLength(P1, P2, P3, 1500)
    Q, U = remesh.PolygonRemeshing(P, T, L)
    G = P.copy()
    G[:, 1:4] *= 1.3
    Qg, Ug = remesh.PolygonRemeshing(G, T, L)
    k0, nu = _k1(Q, U)
    k1, _ = _k1(Qg, Ug)
    e0 = np.mean(100 * k0 / AN.tada_inclined_penny_tension(1.0, 90.0, 0.0, 1.0, nu)[0] - 100)
    e1 = np.mean(100 * k1 / AN.tada_inclined_penny_tension(1.0, 90.0, 0.0, 1.3, nu)[0] - 100)
    print("[gpu-test] growth step: radius 1 -> 1.3, %d -> %d triangles, mean K_I error %.2f %% -> %.2f %%" % (
        U.shape[0], Ug.shape[0], e0, e1))
    assert 1.5 <= Ug.shape[0] / U.shape[0] <= 1.9
    assert abs(e0) <= 10 and abs(e1) <= 10
