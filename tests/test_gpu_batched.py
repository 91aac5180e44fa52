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


def _unbatched(m, lam, hs=0):
    P1, P2, P3 = m["P1"], m["P2"], m["P3"]
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    cols = []
    for sxx, syy, szz, sxy, sxz, syz in m["stresses"]:
        Dn, Dss, Dds, A, b = api.SlipCalculator3D(P1, P2, P3, NU, G, lam, mid, nrm, hs,
                                                  api.Stresses(sxx, syy, szz, sxy, sxz, syz), np.zeros(n))
        cols.append(np.concatenate([Dn, Dss, Dds]))
    return np.stack(cols, axis=1)


def _all_on_path(rep, path):
    """Every mesh's own report names `path` (the batch's path alone would hide a single fallback)."""
    assert [q["path"] for q in rep["meshes"]] == [path] * len(rep["meshes"])
    if path != 0:
        assert rep["fallbacks"] == 0 and all(q["residual"] <= 1e-10 for q in rep["meshes"])


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
        _all_on_path(rep, 3 if mode == "mixed16" else 1)


def test_batched_mixed_graph_replays(batch):
    """Mixed precision: the assembly and the solve's fixed part (FP32 LUs, permutations, first solve and one
    correction) are captured in graphs on the first call and replayed after.  Three calls with eager work between
    them (new geometry, then the first geometry again): each matches the unbatched mixed path, and the third
    repeats the first bit for bit."""
    K, E, lam, nu, G_ = api.ElasticConstantsCheck(G, NU)
    rng = np.random.default_rng(5)
    moved = []
    for m in batch:
        R = meshes.random_rotation(rng)
        moved.append(dict(m, P1=m["P1"] @ R.T, P2=m["P2"] @ R.T, P3=m["P3"] @ R.T))
    api.set_precision("mixed")
    try:
        x1, r1 = api.batched_slip(batch, nu, G_, lam, 0)
        x2, r2 = api.batched_slip(moved, nu, G_, lam, 0)
        x3, r3 = api.batched_slip(batch, nu, G_, lam, 0)
        ref_b = [_unbatched(m, lam) for m in batch]
        ref_m = [_unbatched(m, lam) for m in moved]
    finally:
        api.set_precision("fp64")
    for xs, refs in ((x1, ref_b), (x2, ref_m), (x3, ref_b)):
        for x, ref in zip(xs, refs):
            assert np.max(np.abs(x - ref)) <= 1e-8 * np.max(np.abs(ref))
    for a, b in zip(x1, x3):
        assert np.array_equal(a, b)
    for r in (r1, r2, r3):
        assert r["residual"] <= 1e-10 and r["path"] == 1
        _all_on_path(r, 1)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_batched_halfspace_many_loads(mode):
    """Half-space batch (cracks moved below the free surface) with one mesh carrying 11 remote loads: its factor
    applications split into several column groups and its residual into two passes over A."""
    K, E, lam, nu, G_ = api.ElasticConstantsCheck(G, NU)
    rng = np.random.default_rng(23)
    batch = meshes.elliptical_cracks(count=3, n_min=100, n_max=300, rhs_max=2, seed=29)
    down = np.array([0.0, 0.0, -16.0])
    batch = [dict(m, P1=m["P1"] + down, P2=m["P2"] + down, P3=m["P3"] + down) for m in batch]
    batch[1] = dict(batch[1], stresses=rng.normal(size=(11, 6)))
    assert all(np.max(m[k][:, 2]) < 0 for m in batch for k in ("P1", "P2", "P3"))
    api.set_precision(mode)
    try:
        xs, rep = api.batched_slip(batch, nu, G_, lam, 1)
        refs = [_unbatched(m, lam, hs=1) for m in batch]
    finally:
        api.set_precision("fp64")
    tol = 1e-12 if mode == "fp64" else 1e-8
    for x, ref in zip(xs, refs):
        assert x.shape == ref.shape
        assert np.max(np.abs(x - ref)) <= tol * np.max(np.abs(ref))
    if mode == "mixed":
        assert rep["residual"] <= 1e-10 and rep["path"] == 1
        _all_on_path(rep, 1)


def test_batched_mixed_large_mesh_host_permutations():
    """A mesh above 16384 unknowns (5766 triangles, 17298 unknowns) is too large for the device permutation kernel:
    the batch then runs its fixed part eagerly with the permutations from the host, and still matches the
    unbatched mixed solve."""
    K, E, lam, nu, G_ = api.ElasticConstantsCheck(G, NU)
    rng = np.random.default_rng(31)
    small = meshes.elliptical_cracks(count=1, n_min=100, n_max=150, rhs_max=2, seed=37)[0]
    P1, P2, P3 = meshes.penny(31, 1.0, seed=41, aspect=1.5)
    big = dict(P1=P1, P2=P2, P3=P3, stresses=rng.normal(size=(2, 6)))
    assert 3 * P1.shape[0] > 16384
    batch = [small, big]
    api.set_precision("mixed")
    try:
        xs, rep = api.batched_slip(batch, nu, G_, lam, 0)
        refs = [_unbatched(m, lam) for m in batch]
    finally:
        api.set_precision("fp64")
    for x, ref in zip(xs, refs):
        assert np.max(np.abs(x - ref)) <= 1e-8 * np.max(np.abs(ref))
    assert rep["residual"] <= 1e-10 and rep["path"] == 1
    _all_on_path(rep, 1)


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
    _all_on_path(rep, 0)
    assert rep["fallbacks"] == len(batch)


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


def _lu_single(L, ctx, sp, A, n):
    import ctypes
    import torch
    LU = A.clone()
    ipiv = torch.empty(max(n, 1), dtype=torch.int32, device="cuda:0")
    rc = L.cdj_lu_factor_dev(ctx, ctypes.c_void_p(LU.data_ptr()), n, n, ctypes.c_void_p(ipiv.data_ptr()), sp)
    return rc, LU, ipiv


def test_batched_lu_bitwise_equals_single():
    """cdj_lu_factor_batched_dev factors each matrix exactly as cdj_lu_factor_dev does alone (same operations in
    the same order), over sizes that need one and several panel workgroups, more matrices than one panel launch
    holds, integer-valued matrices with exact pivot ties across workgroups, diagonally dominant matrices (every panel
    taken by the diagonal-pivot fast path) and dominant ones with a forced row interchange (that panel falls back to
    the exchange), and a singular matrix."""
    import ctypes
    import torch
    from cutanddisplacejulia_amd import _lib
    L = _lib.load()
    ctx = _lib.ctx()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(11)
    sizes = [700, 1, 64, 65, 300, 257, 1100, 513, 0] + [int(s) for s in rng.integers(2, 200, 24)]
    mats = []
    for t, n in enumerate(sizes):
        if t % 3 == 1:  # small integers: many exactly equal candidates, including across workgroups
            a = rng.integers(-3, 4, size=(n, n)).astype(np.float64) + 7.0 * np.eye(n)
        elif t % 3 == 2:  # diagonally dominant: partial pivoting keeps the diagonal
            a = rng.uniform(-1, 1, size=(n, n)) / max(n, 1)
            a[np.arange(n), np.arange(n)] = rng.uniform(1, 2, n) * rng.choice([-1, 1], n)
            if t % 6 == 5 and n > 2:  # one column where a lower row wins
                j = int(rng.integers(0, n - 1))
                a[int(rng.integers(j + 1, n)), j] = 10.0
        else:
            a = rng.standard_normal((n, n))
        mats.append(torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).to("cuda:0"))
    sing = rng.standard_normal((40, 40))
    sing[:, 17] = 0.0  # a zero column stays zero through the elimination: an exact zero pivot at column 17
    sizes.append(40)
    mats.append(torch.from_numpy(np.asfortranarray(sing).ravel(order="F").copy()).to("cuda:0"))
    work = [m.clone() for m in mats]
    pivs = [torch.empty(max(n, 1), dtype=torch.int32, device="cuda:0") for n in sizes]
    cnt = len(sizes)
    Ap = (ctypes.c_void_p * cnt)(*[m.data_ptr() for m in work])
    Pp = (ctypes.c_void_p * cnt)(*[p.data_ptr() for p in pivs])
    nn = (ctypes.c_int64 * cnt)(*sizes)
    info = (ctypes.c_int32 * cnt)()
    _lib.check(L.cdj_lu_factor_batched_dev(ctx, Ap, nn, nn, cnt, Pp, info, sp))
    torch.cuda.synchronize()
    for t, n in enumerate(sizes):
        rc, LU, ipiv = _lu_single(L, ctx, sp, mats[t], n)
        torch.cuda.synchronize()
        if t == cnt - 1:
            assert rc == _lib.CDJ_ERR_SINGULAR and info[t] == 18
        else:
            assert rc == 0 and info[t] == 0, (t, n, rc, info[t])
        assert torch.equal(work[t], LU), (t, n)
        if n > 0:
            assert torch.equal(pivs[t][:n], ipiv[:n]), (t, n)
