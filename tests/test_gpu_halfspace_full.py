"""The benchmark's own matrix at its full size: the half-space collocation matrix of the 45-degree dipping fault
meshed into 5000 triangles (A 15000 x 15000 FP64), assembled on the device exactly as bench.py times it
(cdj_assemble_dev: each mesh edge's free-surface part evaluated once and shared by its two triangles), against the
oracle's rows (CalculateInfluenceMatrices3D.jl:2-48 restated, every triangle's three sides evaluated) on 256
receivers drawn at random.

Tolerance: max |gpu - oracle| <= 1e-12 x the largest oracle entry of the same slip mode (its three traction
blocks over the sampled rows).  One scale per mode, not per block: on a planar fault the blocks that couple opening
to shear and shear to normal traction (DnTss, DnTds, DssTn, DdsTn) are about 1e-6 of their mode's scale, and their
entries are free-surface terms of that size carrying the rounding noise (~1e-15) of the mode's O(10) entries, so a
per-block scale compares rounding noise.  The per-block figure and the entrywise relative error of entries above
1e-8 x the mode's maximum are printed.
"""
import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from cutanddisplacejulia_amd import _lib, learned  # noqa: E402
from oracle import oracle as O  # noqa: E402


@pytest.fixture(scope="module")
def fault():
    c = meshes.config2(50)
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    d = {k: learned.to_device_cols(v) for k, v in dict(P1=P1, P2=P2, P3=P3, N=nrm, M=mid).items()}
    A = torch.empty((3 * n, 3 * n), dtype=torch.float64, device="cuda:0")  # column-major: A[col, row]
    learned.assemble_dev(d["N"], d["M"], d["P1"], d["P2"], d["P3"], n, nu, G, lam, 1, A, 3 * n)
    torch.cuda.synchronize()
    return dict(P1=P1, P2=P2, P3=P3, n=n, nrm=nrm, mid=mid, nu=nu, G=G, lam=lam, A=A, edges=_edges())


def _edges():
    import ctypes
    e = ctypes.c_int64()
    _lib.check(_lib.load().cdj_last_edge_stats(_lib.ctx(), ctypes.byref(e), None))
    return e.value


def _rows(f, rows):
    """Rows `rows` (receiver indices) of the device A as the 9 TractionInf blocks (rows x n), A = -[...]."""
    n = f["n"]
    idx = torch.as_tensor(rows, device="cuda:0")
    out = {}
    for k, name in enumerate(O.TRACTION_FIELDS):
        m, q = divmod(k, 3)  # mode m = column block, traction q = row block
        blk = f["A"][m * n:(m + 1) * n, :]                      # columns of mode m (storage is A^T)
        out[name] = (-blk[:, q * n + idx]).t().contiguous().cpu().numpy()
    return out


def test_benchmark_matrix_matches_oracle_rows(fault):
    n = fault["n"]
    # the shared-edge path ran: 50 x 50 quads have 50 * 51 * 2 + 2500 = 7600 edges for 15000 triangle sides
    assert fault["edges"] == 7600
    rng = np.random.default_rng(20261204)
    starts = np.sort(rng.choice(np.arange(0, n - 32, 32), size=8, replace=False))
    rows = np.concatenate([np.arange(s, s + 32) for s in starts])
    got = _rows(fault, rows)
    worst, worst_block, dist = 0.0, 0.0, []
    for s in starts:
        ref = O.traction_influence_rows(fault["nrm"], fault["mid"], fault["P1"], fault["P2"], fault["P3"],
                                        fault["nu"], fault["G"], fault["lam"], 1, int(s), int(s) + 32)
        sel = slice(int(np.searchsorted(rows, s)), int(np.searchsorted(rows, s)) + 32)
        for m in range(3):
            names = O.TRACTION_FIELDS[3 * m:3 * m + 3]
            scale = max(np.max(np.abs(ref[k])) for k in names)
            for name in names:
                r, g = ref[name], got[name][sel]
                err = np.abs(g - r)
                worst = max(worst, float(np.max(err) / scale))
                worst_block = max(worst_block, float(np.max(err) / np.max(np.abs(r))))
                big = np.abs(r) > 1e-8 * scale
                dist.append(err[big] / np.abs(r[big]))
    rel = np.concatenate(dist)
    print("5000-triangle half-space A, %d sampled rows: max error %.2e of the slip mode's scale (%.2e of the block's "
          "own); entrywise relative error of entries > 1e-8 x mode max: median %.1e, 99%% %.1e, 99.99%% %.1e, "
          "max %.1e" % (rows.size, worst, worst_block, np.median(rel), np.quantile(rel, 0.99),
                        np.quantile(rel, 0.9999), rel.max()))
    assert worst <= 1e-12


def test_flat_free_surface_mesh_traction_matches_oracle():
    """Triangles lying exactly in z = 0 (the reference's flat-triangle image behaviour, TD.jl:1518-1525,
    1610-1617, whose image pass needs the accumulated strain, so every pair evaluates its own sides) and the same mesh
    moved below the surface (edges shared): the traction matrices match the oracle."""
    P1, P2, P3 = meshes.rect_fault(4, 4, 60.0, 0.0, 2.0, 2.0, 0.0)  # horizontal rectangle at z = 0
    for P in (P1, P2, P3):
        P[:, 2] = 0.0                                                  # exactly (cos 90 leaves ~1e-16)
    n = P1.shape[0]
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    for dz in (0.0, -0.7):
        Q1, Q2, Q3 = (P + np.array([0.0, 0.0, dz]) for P in (P1, P2, P3))
        assert np.all(Q1[:, 2] == dz)
        nrm, mid = api.CreateFaceNormalAndMidPoint(Q1, Q2, Q3)
        Ti = api.CalculateInfluenceMatrices3D(nrm, mid, Q1, Q2, Q3, nu, G, lam, np.zeros(n), 1, n)
        assert (_edges() > 0) == (dz != 0.0)
        To = O.traction_influence(nrm, mid, Q1, Q2, Q3, nu, G, lam, 1)
        for m in range(3):
            names = O.TRACTION_FIELDS[3 * m:3 * m + 3]
            scale = max(np.nanmax(np.abs(To[k])) for k in names)
            for k in names:
                g, r = getattr(Ti, k), To[k]
                assert np.array_equal(np.isnan(g), np.isnan(r))
                ok = ~np.isnan(r)
                assert np.max(np.abs(g[ok] - r[ok])) <= 1e-12 * scale, (dz, k)
