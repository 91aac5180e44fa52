"""Batched slip solves: one persistent work-queue assembly over many uneven meshes, captured in a HIP graph.

Pinned by per-mesh equality with the unbatched path (cdj_slip_calculator through api.SlipCalculator3D-style
solves): FP64 assembly + FP64 LU agree to 1e-12 relative (same kernels' arithmetic, different launch shape), and
the mixed mode (FP32 far / FP64 near assembly, FP32 LU + FP64 GMRES-IR) agrees with the unbatched mixed path to
the refinement tolerance.  A second call with the same shape replays the captured graph.
"""
import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshes

pytestmark = pytest.mark.gpu

NU, G = 0.25, 1.0


def _unbatched(m, lam):
    P1, P2, P3 = m["P1"], m["P2"], m["P3"]
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    cols = []
    for sxx, syy, szz, sxy, sxz, syz in m["stresses"]:
        Dn, Dss, Dds, A, b = api.SlipCalculator3D(P1, P2, P3, NU, G, lam, mid, nrm, 0,
                                                  api.Stresses(sxx, syy, szz, sxy, sxz, syz), np.zeros(n))
        cols.append(np.concatenate([Dn, Dss, Dds]))
    return np.stack(cols, axis=1)


@pytest.fixture(scope="module")
def batch():
    return meshes.elliptical_cracks(count=6, n_min=100, n_max=500, rhs_max=3, seed=17)


@pytest.mark.parametrize("mode", ["fp64", "mixed", "mixed16"])
def test_batched_equals_unbatched(batch, mode):
    K, E, lam, nu, G_ = api.ElasticConstantsCheck(G, NU)
    api.set_precision(mode)
    try:
        xs, rep = api.batched_slip(batch, nu, G_, lam, 0)
        refs = [_unbatched(m, lam) for m in batch]
    finally:
        api.set_precision("fp64")
    tol = 1e-12 if mode == "fp64" else 1e-8
    for x, ref in zip(xs, refs):
        assert x.shape == ref.shape
        err = np.max(np.abs(x - ref)) / np.max(np.abs(ref))
        assert err <= tol, err
    print(mode, rep)
    assert rep["tile"]["sources_per_item"] > 0 and rep["tile"]["workgroups"] > 0
    if mode == "mixed":
        assert rep["near_entries"] > 0
    else:
        assert rep["near_entries"] == 0
    if mode != "fp64":
        # every mesh refined to the tolerance on the mixed path itself (no FP64 fallback)
        assert rep["residual"] <= 1e-10 and rep["iterations"] > 0
        assert rep["path"] == (3 if mode == "mixed16" else 1)


def test_batched_mixed_fallback_chain(batch):
    """A tolerance below FP64 rounding: the refinement rounds stall, FGMRES-IR stops at its iteration cap, and
    every mesh ends on the FP64 LU, whose solution equals the unbatched FP64 one."""
    K, E, lam, nu, G_ = api.ElasticConstantsCheck(G, NU)
    api.set_precision("mixed32")
    saved = dict(api._SETTINGS)
    api._SETTINGS.update(tol=1e-30, max_iter=3)
    try:
        xs, rep = api.batched_slip(batch, nu, G_, lam, 0)
    finally:
        api._SETTINGS.clear()
        api._SETTINGS.update(saved)
        api.set_precision("fp64")
    refs = [_unbatched(m, lam) for m in batch]
    for x, ref in zip(xs, refs):
        assert np.max(np.abs(x - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert rep["path"] == 0


def test_batched_graph_replay_with_new_geometry(batch):
    """Same shape, moved geometry: the captured graph is replayed on the new coordinates."""
    K, E, lam, nu, G_ = api.ElasticConstantsCheck(G, NU)
    x1, _ = api.batched_slip(batch, nu, G_, lam, 0)
    moved = []
    rng = np.random.default_rng(3)
    for m in batch:
        R = meshes.random_rotation(rng)
        moved.append(dict(m, P1=m["P1"] @ R.T, P2=m["P2"] @ R.T, P3=m["P3"] @ R.T))
    x2, _ = api.batched_slip(moved, nu, G_, lam, 0)
    refs = [_unbatched(m, lam) for m in moved]
    for a, ref in zip(x2, refs):
        assert np.max(np.abs(a - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert any(np.max(np.abs(a - b)) > 1e-6 for a, b in zip(x1, x2))
