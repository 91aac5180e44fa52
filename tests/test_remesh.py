"""Isotropic remeshing for growth loops (cdj_remesh_isotropic, host code; no GPU) and the edge-length helpers.

The reference hands this step to CGAL 4.13.1 (BuildCGAL.PolygonRemeshingCGAL, test/test_PropagationTest.jl:151-561),
which is not part of it, so the remesher is pinned by what isotropic remeshing must deliver and by geometry:
* the reference's own penny mesh (tests/golden/CircleMesh_1a_500Faces.off) remeshed to the edge length that
  GetDesiredEdgeLength gives for 200 ... 6000 triangles: the triangle count lands within 10 % of the request,
  at least 90 % of the edges lie in [4/5 L, 4/3 L], the surface stays in its plane, every boundary vertex stays on
  the input's boundary polyline, the result is a manifold disc (V - E + F = 1) with the input's orientation, and
  the area can only shrink, by boundary collapses cutting chords of the convex boundary (within 2 %);
* a rectangular fault keeps its four corners, its plane and its area;
* a closed polyhedral sphere stays closed (every edge in two triangles, V - E + F = 2), outward, and every vertex
  lies on the input surface;
* AreaOfTriangle3D / GetDesiredEdgeLength (AreaOfTriangle3D.jl, GetDesiredEdgeLength.jl) on hand-worked triangles.
"""
import os
from collections import Counter

import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshes, meshio, remesh
from cutanddisplacejulia_amd._lib import CdjError

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _edges(T):
    return Counter(tuple(sorted(e)) for t in T for e in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])))


def _boundary_segments(P, T):
    """Boundary edges of the mesh (Points n x 4 / 1-based Triangles) as coordinate pairs."""
    ids = {int(i): k for k, i in enumerate(P[:, 0])}
    return [(P[ids[a], 1:4], P[ids[b], 1:4]) for (a, b), c in _edges(T).items() if c == 1]


def _dist_to_segments(x, segs):
    best = np.inf
    for a, b in segs:
        ab = b - a
        t = np.clip(np.dot(x - a, ab) / np.dot(ab, ab), 0.0, 1.0)
        best = min(best, float(np.linalg.norm(x - a - t * ab)))
    return best


def _area(P, T):
    P1, P2, P3 = meshio.CreateP1P2P3(T, P)
    a, _ = api.AreaOfTriangle3D(*P1.T, *P2.T, *P3.T)
    return a.sum(), P1, P2, P3


@pytest.fixture(scope="module")
def penny():
    return meshio.OFFReader(os.path.join(GOLDEN, "CircleMesh_1a_500Faces.off"))


@pytest.mark.parametrize("notris", [200, 500, 1500, 6000])
def test_penny_remesh_meets_target(penny, notris):
    P, T = penny
    a0, P1, P2, P3 = _area(P, T)
    L, _ = api.GetDesiredEdgeLength(P1, P2, P3, notris)
    Q, U = remesh.PolygonRemeshing(P, T, L)
    a1, Q1, Q2, Q3 = _area(Q, U)
    nt = U.shape[0]
    assert abs(nt - notris) <= 0.1 * notris
    E = np.concatenate([np.linalg.norm(Q1 - Q2, axis=1), np.linalg.norm(Q2 - Q3, axis=1),
                        np.linalg.norm(Q3 - Q1, axis=1)])
    assert np.mean((E >= 0.8 * L) & (E <= 4 / 3 * L)) >= 0.9
    assert np.max(np.abs(Q[:, 3])) == 0.0  # the input lies in z = 0
    # orientation: the input's normals all point the same way; so must the output's
    s0 = np.sign(np.cross(P2 - P1, P3 - P1)[:, 2])
    s1 = np.sign(np.cross(Q2 - Q1, Q3 - Q1)[:, 2])
    assert np.all(s0 == s0[0]) and np.all(s1 == s0[0])
    # a manifold disc: edges used once (boundary) or twice, V - E + F = 1, one boundary loop
    cnt = _edges(U)
    assert set(cnt.values()) <= {1, 2}
    assert Q.shape[0] - len(cnt) + nt == 1
    bnd = [e for e, c in cnt.items() if c == 1]
    deg = Counter(v for e in bnd for v in e)
    assert set(deg.values()) == {2}
    # boundary vertices on the input boundary polyline
    segs = _boundary_segments(P, T)
    ids = {int(i): k for k, i in enumerate(Q[:, 0])}
    bverts = {v for e in bnd for v in e}
    assert max(_dist_to_segments(Q[ids[v], 1:4], segs) for v in bverts) <= 1e-12
    # boundary vertices on the input's convex boundary: collapses along it can only cut chords
    assert a0 * 0.98 <= a1 <= a0 * (1 + 1e-12)


def test_rectangular_fault_keeps_corners_plane_and_area():
    c = meshes.config2(10)
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    uq, inv = np.unique(np.vstack([P1, P2, P3]), axis=0, return_inverse=True)
    inv = inv.ravel()
    T = np.stack([inv[:n], inv[n:2 * n], inv[2 * n:]], 1) + 1
    P = np.column_stack([np.arange(1, uq.shape[0] + 1, dtype=float), uq])
    a0, _, _, _ = _area(P, T)
    L, _ = api.GetDesiredEdgeLength(P1, P2, P3, 800)
    Q, U = remesh.PolygonRemeshing(P, T, L)
    a1, Q1, Q2, Q3 = _area(Q, U)
    assert abs(U.shape[0] - 800) <= 80
    assert abs(a1 - a0) <= 1e-12 * a0
    nrm = np.cross(P2[0] - P1[0], P3[0] - P1[0])
    nrm /= np.linalg.norm(nrm)
    assert np.max(np.abs((Q[:, 1:4] - P1[0]) @ nrm)) <= 1e-12
    for d in ([1, 1, 1.0], [-1, -1, -1.0], [1, -1, 0.0], [-1, 1, 0.0]):
        corner = uq[np.argmin(uq @ np.array(d))]
        assert np.any(np.all(np.abs(Q[:, 1:4] - corner) <= 1e-12, axis=1))


def test_closed_sphere_stays_closed_outward_and_on_the_surface():
    from tests import gocad
    P, T = gocad.read_ts("SphereUniformDistributionON_46Faces.ts")
    P1, P2, P3 = gocad.p1p2p3(P, T)
    c = np.vstack([P1, P2, P3]).mean(axis=0)
    L, _ = api.GetDesiredEdgeLength(P1, P2, P3, 600)
    uq, inv = np.unique(np.vstack([P1, P2, P3]), axis=0, return_inverse=True)
    inv = inv.ravel()
    n = P1.shape[0]
    Tin = np.stack([inv[:n], inv[n:2 * n], inv[2 * n:]], 1)
    Q, U = remesh.remesh_arrays(uq, Tin, L)
    cnt = _edges(U)
    assert set(cnt.values()) == {2}
    assert Q.shape[0] - len(cnt) + U.shape[0] == 2
    Q1, Q2, Q3 = Q[U[:, 0]], Q[U[:, 1]], Q[U[:, 2]]
    assert np.all(np.sum(np.cross(Q2 - Q1, Q3 - Q1) * ((Q1 + Q2 + Q3) / 3 - c), axis=1) > 0)
    # every vertex on some input triangle (in its plane, inside its barycentric range)
    A, B, C = P1, P2, P3
    nrm = np.cross(B - A, C - A)
    for x in Q:
        h = np.abs(np.sum((x - A) * nrm, axis=1)) / np.linalg.norm(nrm, axis=1)
        v0, v1, v2 = B - A, C - A, x - A
        d00, d01, d11 = np.sum(v0 * v0, 1), np.sum(v0 * v1, 1), np.sum(v1 * v1, 1)
        d20, d21 = np.sum(v2 * v0, 1), np.sum(v2 * v1, 1)
        den = d00 * d11 - d01 * d01
        v = (d11 * d20 - d01 * d21) / den
        w = (d00 * d21 - d01 * d20) / den
        inside = (v >= -1e-9) & (w >= -1e-9) & (v + w <= 1 + 1e-9)
        assert np.any(inside & (h <= 1e-12 * max(1.0, np.linalg.norm(x))))


def test_file_form_matches_array_form(penny, tmp_path):
    P, T = penny
    src = meshio.OFFExport(P, T, T.shape[0], P.shape[0], str(tmp_path / "BeforePolygonRemshing"))
    out = remesh.PolygonRemeshingCGAL(src, 0.1)
    Pf, Tf = meshio.OFFReader(out)
    Pa, Ta = remesh.PolygonRemeshing(*meshio.OFFReader(src), 0.1)
    assert np.array_equal(Tf, Ta)
    assert np.max(np.abs(Pf[:, 1:4] - Pa[:, 1:4])) <= 1e-12


def test_invalid_inputs_are_refused():
    pts = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0]])
    with pytest.raises(CdjError):
        remesh.remesh_arrays(pts, np.array([[0, 1, 1]]), 0.5)
    with pytest.raises(CdjError):
        remesh.remesh_arrays(pts, np.array([[0, 1, 3]]), 0.5)
    with pytest.raises(CdjError):
        remesh.remesh_arrays(pts, np.array([[0, 1, 2]]), 0.0)


def test_area_and_desired_edge_length():
    # the reference's documented example: the flat right triangle (0,0,0), (1,0,0), (0,1,0) has area 1/2
    a, hp = api.AreaOfTriangle3D(0, 0, 0, 1, 0, 0, 0, 1, 0)
    assert a[0] == pytest.approx(0.5, rel=1e-15) and hp[0] == pytest.approx((2 + np.sqrt(2)) / 2, rel=1e-15)
    # a mesh of equilateral triangles of side s asks for side s at its own triangle count
    s = 0.3
    P1 = np.array([[0.0, 0, 0], [s, 0, 0]])
    P2 = np.array([[s, 0, 0], [1.5 * s, s * np.sqrt(3) / 2, 0]])
    P3 = np.array([[s / 2, s * np.sqrt(3) / 2, 0], [s / 2, s * np.sqrt(3) / 2, 0]])
    L, Lmax = api.GetDesiredEdgeLength(P1, P2, P3, 2)
    assert L == pytest.approx(s, rel=1e-14) and Lmax == pytest.approx(s, rel=1e-14)
    assert api.GetDesiredEdgeLength(P1, P2, P3, 8)[0] == pytest.approx(s / 2, rel=1e-14)
    # collinear corners: Heron's product rounds to zero or below, never a negative area
    a, _ = api.AreaOfTriangle3D(0, 0, 0, 1, 0, 0, 2, 0, 0)
    assert a[0] == 0.0 or np.isnan(a[0])
