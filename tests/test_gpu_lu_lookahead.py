"""Look-ahead and the diagonal-pivot fast path in the blocked LU: the factors are the same bit for bit with and
without either.

With look-ahead the trailing update of an outer block is split between two streams (the next block's columns
first, the rest beside the next block's panels), and the panels' row interchanges in the other columns are
applied after the product; every entry still goes through the same operations. CDJ_LU_LOOKAHEAD=0 switches it off.

The fast path factors a panel without the grid-wide pivot exchange when partial pivoting would keep the diagonal
row at every column, and falls back to the exchange (after restoring the panel) when it would not.
CDJ_LU_FASTPANEL=0 switches it off. Both switches are read per call. Checked on matrices that never pivot (as the
collocation matrices), that pivot in some panels only, and that pivot everywhere, with pivot ties; on the FP64
factorization (device entry point), on the FP64 solve of [A | B] (extra right-hand-side columns ride along) and on
the FP32 / FP16-update factorizations behind the mixed solves, at sizes of several outer blocks. The FP64 factors
are also checked as P A = L U, the pivots against LAPACK's, and the solutions against LAPACK.
"""
import ctypes
import os

import numpy as np
import pytest

from cutanddisplacejulia_amd import api

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from cutanddisplacejulia_amd import _lib  # noqa: E402


def _matrix(n, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, n)) + 0.3 * n ** 0.5 * np.eye(n) * rng.choice([-1, 1], n)


def _with_env(fn, **flags):
    old = {k: os.environ.get(k) for k in flags}
    for k, v in flags.items():
        os.environ[k] = "1" if v else "0"
    try:
        return fn()
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _with_lookahead(flag, fn):
    return _with_env(fn, CDJ_LU_LOOKAHEAD=flag)


def _dominant(n, seed):
    """Diagonal entries larger than every column's other entries all along the elimination: never pivots."""
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, size=(n, n)) / n
    A[np.arange(n), np.arange(n)] = rng.uniform(1, 2, n) * rng.choice([-1, 1], n)
    return A


def _partly(n, seed):
    """Dominant, except a few columns where a row below holds the column's largest entry (and one exact tie)."""
    A = _dominant(n, seed)
    rng = np.random.default_rng(seed + 1)
    for j in rng.choice(n - 70, 6, replace=False):
        A[j + 1 + int(rng.integers(0, 69)), j] = 10.0
    # row 300 equal to row 250 up to column 250: at that column both hold the same value, an exact tie that the
    # diagonal row keeps (lowest position)
    A[300, :251] = A[250, :251]
    return A


def _factor_dev(A, **flags):
    n = A.shape[0]
    L = _lib.load()
    ctx = _lib.ctx()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def run():
        d = torch.from_numpy(np.asfortranarray(A).T.copy()).to("cuda:0")
        ipiv = torch.empty(n, dtype=torch.int32, device="cuda:0")
        _lib.check(L.cdj_lu_factor_dev(ctx, ctypes.c_void_p(d.data_ptr()), n, n, ctypes.c_void_p(ipiv.data_ptr()),
                                       sp))
        torch.cuda.synchronize()
        return d.cpu().numpy(), ipiv.cpu().numpy()

    return _with_env(run, **flags)


@pytest.mark.parametrize("kind,n", [("dominant", 1100), ("partly", 1100), ("random", 700), ("dominant", 2085)])
def test_fast_panel_factors_bitwise_equal_to_exchange(kind, n):
    print("[gpu-test] fast panel %s n=%d" % (kind, n), flush=True)
    A = {"dominant": _dominant, "partly": _partly, "random": lambda n, s: np.random.default_rng(s).normal(size=(n, n))}[
        kind](n, 3)
    f0, p0 = _factor_dev(A, CDJ_LU_FASTPANEL=False)
    f1, p1 = _factor_dev(A, CDJ_LU_FASTPANEL=True)
    assert np.array_equal(p0, p1)
    assert np.array_equal(f0, f1)
    import scipy.linalg as sl
    piv = sl.lu_factor(A)[1]
    if kind == "dominant":
        assert np.array_equal(p1, np.arange(n))
    if kind == "partly":
        assert 0 < np.sum(p1 != np.arange(n)) <= 12 and p1[250] == 250
    if kind != "random":  # well separated pivots: LAPACK's (a random matrix may hold near-ties that rounding decides)
        assert np.array_equal(p1, piv)


@pytest.mark.parametrize("n", [1000, 2085])
def test_fp64_factors_bitwise_equal_with_and_without_lookahead(n):
    print("[gpu-test] lookahead fp64 factors n=%d" % n, flush=True)
    A = _matrix(n, n)
    L = _lib.load()
    ctx = _lib.ctx()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def factor():
        # device copy in Julia (column-major) layout: the transpose's rows are A's columns
        d = torch.from_numpy(np.asfortranarray(A).T.copy()).to("cuda:0")
        ipiv = torch.empty(n, dtype=torch.int32, device="cuda:0")
        _lib.check(L.cdj_lu_factor_dev(ctx, ctypes.c_void_p(d.data_ptr()), n, n, ctypes.c_void_p(ipiv.data_ptr()),
                                       sp))
        torch.cuda.synchronize()
        return d.cpu().numpy(), ipiv.cpu().numpy()

    f0, p0 = _with_lookahead(False, factor)
    f1, p1 = _with_lookahead(True, factor)
    assert np.array_equal(p0, p1)
    assert np.array_equal(f0, f1)
    # and the factors are an LU of A: P A = L U
    LU = f1.T  # back to row-major A layout
    Lm = np.tril(LU, -1) + np.eye(n)
    Um = np.triu(LU)
    PA = A.copy()
    for i, p in enumerate(p1):
        PA[[i, p]] = PA[[p, i]]
    assert np.max(np.abs(Lm @ Um - PA)) <= 1e-12 * np.max(np.abs(A))


def test_fast_panel_zero_diagonal_falls_back():
    """A zero on the diagonal with nonzero entries below: partial pivoting swaps there, so the fast path must fall
    back in that panel; the factors and pivots are the exchange panel's."""
    n = 900
    A = _dominant(n, 21)
    A[100, 100] = 0.0
    print("[gpu-test] fast panel zero diagonal n=%d" % n, flush=True)
    f0, p0 = _factor_dev(A, CDJ_LU_FASTPANEL=False)
    f1, p1 = _factor_dev(A, CDJ_LU_FASTPANEL=True)
    assert p1[100] != 100
    assert np.array_equal(p0, p1) and np.array_equal(f0, f1)


def test_fast_panel_zero_column_reports_the_same_pivot():
    """An exactly zero column: no row can serve as pivot, partial pivoting keeps the diagonal and reports the first
    zero pivot; the fast path reports the same column."""
    n = 700
    A = _dominant(n, 22)
    A[:, 70] = 0.0
    print("[gpu-test] fast panel zero column n=%d" % n, flush=True)
    msgs = []
    for flag in (False, True):
        with pytest.raises(api.CdjError) as ei:
            _factor_dev(A, CDJ_LU_FASTPANEL=flag)
        msgs.append(str(ei.value))
    assert msgs[0] == msgs[1] and "zero pivot 71" in msgs[1], msgs


@pytest.mark.parametrize("mode", ["fp64", "mixed32", "mixed16"])
def test_fast_panel_solves_bitwise_equal_to_exchange(mode):
    n, nrhs = 1500, 2
    print("[gpu-test] fast panel solve %s n=%d" % (mode, n), flush=True)
    A = _partly(n, 11)
    b = np.random.default_rng(12).normal(size=(n, nrhs))
    api.set_precision(mode)
    try:
        x0 = _with_env(lambda: api.solve(A, b), CDJ_LU_FASTPANEL=False)
        x1 = _with_env(lambda: api.solve(A, b), CDJ_LU_FASTPANEL=True)
    finally:
        api.set_precision("fp64")
    assert np.array_equal(x0, x1)
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x1 - xr) / np.linalg.norm(xr) <= 1e2 * np.linalg.cond(A) * 2.2e-16


@pytest.mark.parametrize("mode", ["fp64", "mixed32", "mixed16"])
def test_solves_bitwise_equal_with_and_without_lookahead(mode):
    n, nrhs = 1800, 3
    print("[gpu-test] lookahead solve %s n=%d" % (mode, n), flush=True)
    A = _matrix(n, 7)
    b = np.random.default_rng(8).normal(size=(n, nrhs))
    api.set_precision(mode)
    try:
        x0 = _with_lookahead(False, lambda: api.solve(A, b))
        x1 = _with_lookahead(True, lambda: api.solve(A, b))
    finally:
        api.set_precision("fp64")
    assert np.array_equal(x0, x1)
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x1 - xr) / np.linalg.norm(xr) <= 1e2 * np.linalg.cond(A) * 2.2e-16
