"""Mixed-precision solve (FP32 LU on MFMA as preconditioner, FP64 FGMRES iterative refinement).

Bar: the FP64 relative residual ||b - A x|| / ||b|| against the FP64 A reaches 1e-10 (the requested tolerance),
and the forward error against the FP64 LU (LAPACK) solution is within cond(A) * 1e-10.
"""
import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def mixed():
    api.set_precision("mixed32")
    yield
    api.set_precision("fp64")
    api._SETTINGS["max_iter"] = 0


@pytest.mark.parametrize("n", [100, 777, 2000])
def test_mixed_dense_solve(mixed, n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + 2 * n ** 0.5 * np.eye(n)
    b = rng.normal(size=(n, 2))
    x = api.solve(A, b)
    rep = dict(api.LAST_REPORT)
    assert rep["path"] == 1, rep
    res = np.linalg.norm(b - A @ x, axis=0) / np.linalg.norm(b, axis=0)
    assert np.all(res <= 1e-10), res
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= np.linalg.cond(A) * 1e-10


def test_mixed_slip_calculator_config1(mixed):
    c = meshes.config1()
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    Dn, Dss, Dds, A, b = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 0,
                                              api.Stresses(0, 0, 1.0, 0, 0, 0), np.zeros(n))
    rep = dict(api.LAST_REPORT)
    assert rep["path"] == 1 and rep["residual"] <= 1e-10, rep
    x = np.concatenate([Dn, Dss, Dds])
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-10
    Ao = O.assemble_A(O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, 0))
    xo = np.linalg.solve(Ao, b)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8


def test_mixed_falls_back_to_fp64(mixed):
    """An ill-conditioned matrix with the refinement capped at one iteration cannot reach 1e-10: the FP64 LU
    produces the answer and the report says so."""
    n = 300
    rng = np.random.default_rng(7)
    U, _ = np.linalg.qr(rng.normal(size=(n, n)))
    Vt, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = U @ np.diag(np.logspace(0, -11, n)) @ Vt
    b = rng.normal(size=n)
    api._SETTINGS["max_iter"] = 1
    x = api.solve(A, b)
    rep = dict(api.LAST_REPORT)
    assert rep["path"] == 0, rep
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-4
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-3


@pytest.fixture
def mixed16():
    api.set_precision("mixed16")
    yield
    api.set_precision("fp64")
    api._SETTINGS["max_iter"] = 0


@pytest.mark.parametrize("n", [300, 1500])
def test_mixed16_dense_solve(mixed16, n):
    """LU with FP16-operand trailing updates (rank-256 updates from n > 256 on) as the refinement's
    preconditioner: the FP64 residual still reaches 1e-10, in more iterations than the FP32 LU needs."""
    rng = np.random.default_rng(n + 1)
    A = rng.normal(size=(n, n)) + 2 * n ** 0.5 * np.eye(n)
    A[:, ::7] *= 1e3  # columns of very different scale exercise the per-column scaling of U
    b = rng.normal(size=(n, 1))
    x = api.solve(A, b)
    rep = dict(api.LAST_REPORT)
    assert rep["path"] == 3, rep
    res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert res <= 1e-10, res
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= np.linalg.cond(A) * 1e-10


def test_mixed16_slip_calculator_halfspace(mixed16):
    P1, P2, P3 = meshes.rect_fault(14, 14, 60.0, 45.0, 5.0, 5.0, 0.5)
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    Dn, Dss, Dds, A, b = api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 1,
                                              api.Stresses(0, 0, 1.0, 0.3, 0, 0), np.zeros(n))
    rep = dict(api.LAST_REPORT)
    assert rep["path"] == 3 and rep["residual"] <= 1e-10, rep
    x = np.concatenate([Dn, Dss, Dds])
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-8


def test_mixed_solve_ignores_stale_device_memory(mixed):
    """Regression: the Krylov basis was started as y = a x + 0 y on freshly allocated memory, so NaNs left in
    reused device memory turned the second right-hand side's refinement into NaN (and the solve fell back to
    FP64).  Device memory is filled with NaN and released first; the mixed path must still converge."""
    import torch
    t = torch.full((1 << 28,), float("nan"), dtype=torch.float64, device="cuda:0")  # 2 GiB of NaN
    torch.cuda.synchronize()
    del t
    torch.cuda.empty_cache()
    n = 777
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + 2 * n ** 0.5 * np.eye(n)
    b = rng.normal(size=(n, 3))
    x = api.solve(A, b)
    rep = dict(api.LAST_REPORT)
    assert rep["path"] == 1, rep
    res = np.linalg.norm(b - A @ x, axis=0) / np.linalg.norm(b, axis=0)
    assert np.all(res <= 1e-10), res
