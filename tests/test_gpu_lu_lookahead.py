"""Look-ahead in the blocked LU: the factors are the same bit for bit with and without it.

With look-ahead the trailing update of an outer block is split between two streams (the next block's columns
first, the rest beside the next block's panels), and the panels' row interchanges in the other columns are
applied after the product; every entry still goes through the same operations. CDJ_LU_LOOKAHEAD=0 switches it off
(read per call). Checked here on the FP64 factorization (device entry point), on the FP64 solve of [A | B] (extra
right-hand-side columns ride along in the split update) and on the FP32 / FP16-update factorizations behind the
mixed solves, at sizes of several outer blocks; the FP64 solution is also checked against LAPACK.
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


def _with_lookahead(flag, fn):
    old = os.environ.get("CDJ_LU_LOOKAHEAD")
    os.environ["CDJ_LU_LOOKAHEAD"] = "1" if flag else "0"
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["CDJ_LU_LOOKAHEAD"]
        else:
            os.environ["CDJ_LU_LOOKAHEAD"] = old


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
