"""cdj_assemble_dev_noalloc (the assembly with its scratch in a fixed device array) against cdj_assemble_dev (its
scratch from the context's allocator).  Full space in FP64 and the FP32-far / FP64-near mixed mode with its near-entry
counter run the same kernels, so the matrices must be identical bit for bit; in half space the allocating entry point
evaluates each mesh edge's free-surface part once and shares it, the no-allocation one evaluates every pair's three
sides, so the two agree to 1e-12 of each slip mode's largest entry.  Two calls in a row per case, the second with a
smaller mesh (the array keeps the first call's contents).  Device memory reported free by the
driver does not change across the no-allocation calls.  A mesh above the array's capacity is refused before any
device work."""
import ctypes

import numpy as np
import pytest

from cutanddisplacejulia_amd import _lib, api, meshes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from cutanddisplacejulia_amd import learned  # noqa: E402


def _mesh(rings, hs):
    P1, P2, P3 = meshes.penny(rings, 1.0, seed=11, normal=(0.3, 0.2, 1.0))
    if hs:
        P1, P2, P3 = (P + np.array([0.0, 0.0, -3.0]) for P in (P1, P2, P3))
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    return [learned.to_device_cols(v) for v in (nrm, mid, P1, P2, P3)], P1.shape[0]


def _assemble(fn, d, n, hs, mixed, A):
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    A.fill_(float("nan"))
    opts = _lib.Opts(tol=1e-10, assembly=_lib.CDJ_ASM_MIXED if mixed else _lib.CDJ_ASM_FP64, near_factor=2.0)
    rc = fn(_lib.ctx(), *[ctypes.c_void_p(t.data_ptr()) for t in d], n, nu, G, lam, int(hs),
            ctypes.c_void_p(A.data_ptr()), 3 * n, ctypes.byref(opts), None)
    _lib.check(rc)
    torch.cuda.synchronize()
    return A.cpu().numpy(), int(_lib.load().cdj_last_near_entries(_lib.ctx()))


@pytest.mark.parametrize("hs,mixed", [(0, False), (1, False), (0, True)])
def test_noalloc_matches_allocating_entry_point(hs, mixed):
    L = _lib.load()
    big, n_big = _mesh(18 if not hs else 12, hs)     # 1944 / 864 triangles
    small, n_small = _mesh(9 if not hs else 7, hs)   # 486 / 294 triangles
    print("[noalloc] hs %d mixed %d: n %d then %d" % (hs, mixed, n_big, n_small), flush=True)
    # outputs allocated, and code objects loaded by a first call, before the free memory is read
    outs = {n: torch.empty((3 * n, 3 * n), dtype=torch.float64, device="cuda:0") for n in (n_big, n_small)}
    _assemble(L.cdj_assemble_dev_noalloc, small, n_small, hs, mixed, outs[n_small])
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    got = [_assemble(L.cdj_assemble_dev_noalloc, d, n, hs, mixed, outs[n]) for d, n in ((big, n_big),
                                                                                      (small, n_small))]
    free1 = torch.cuda.mem_get_info()[0]
    ref = [_assemble(L.cdj_assemble_dev, d, n, hs, mixed, outs[n]) for d, n in ((big, n_big), (small, n_small))]
    for (a, na), (b, nb), n in zip(got, ref, (n_big, n_small)):
        assert not np.isnan(a).any()
        if hs:
            for m in range(3):  # column block of slip mode m (storage is A^T: rows of `a` are columns of A)
                blk_a, blk_b = a[m * n:(m + 1) * n], b[m * n:(m + 1) * n]
                assert np.max(np.abs(blk_a - blk_b)) <= 1e-12 * np.max(np.abs(blk_b))
        else:
            assert np.array_equal(a, b)
        assert na == nb and (na > 0) == mixed
    assert free1 == free0


def test_noalloc_refuses_a_mesh_above_its_capacity():
    L = _lib.load()
    n = 95325                                        # one triangle above the full-space FP64 capacity
    opts = _lib.Opts(tol=1e-10)
    rc = L.cdj_assemble_dev_noalloc(_lib.ctx(), None, None, None, None, None, n, 0.25, 1.0, 0.4, 0, None, 3 * n,
                                    ctypes.byref(opts), None)
    assert rc == _lib.CDJ_ERR_ARG
    assert "scratch" in _lib.last_error()
