"""Mesh readers and writers (cutanddisplacejulia_amd.meshio) on the reference's own mesh files in tests/golden.

Pinned by: the test helpers' independent readers (tests/gocad.py, tests/stl.py) on the same files; the OFF file's
own header counts and its first coordinate line; write -> read round trips (13 decimals as written, identical
connectivity, CRLF layout of the reference's exporters); ATOM records and the multi-surface error.
"""
import os

import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshio
from tests import gocad, stl

G = gocad.GOLDEN


def test_gocad_reader_matches_helper_and_corners():
    for name in ("CircleMesh_1a_500Faces.ts", "SphereUniformDistributionON_46Faces.ts", "SphereFix4.ts"):
        P, T = meshio.GoCadAsciiReader(os.path.join(G, name))
        Pr, Tr = gocad.read_ts(name)
        assert np.array_equal(P, Pr) and np.array_equal(T, Tr)
        for a, b in zip(meshio.CreateP1P2P3(T, P), gocad.p1p2p3(Pr, Tr)):
            assert np.array_equal(a, b)
    P, T = meshio.GoCadAsciiReader(os.path.join(G, "CircleMesh_1a_500Faces.ts"))
    assert T.shape == (581, 3) and T.min() == 1 and T.max() == P.shape[0]


def test_stl_reader_matches_helper():
    P, T = meshio.STLReader(os.path.join(G, "VertPenny-300-EqEdges.stl"))
    Pr, Tr = stl.read_stl("VertPenny-300-EqEdges.stl")
    assert np.array_equal(P, Pr) and np.array_equal(T, Tr) and T.shape == (273, 3)


def test_off_reader_header_and_geometry():
    path = os.path.join(G, "CircleMesh_1a_500Faces.off")
    P, T = meshio.OFFReader(path)
    assert P.shape == (425, 4) and T.shape == (698, 3)
    assert np.array_equal(P[0, 1:], [0.6324519803032, 0.7557386118535, -0.0])
    assert T.min() == 1 and T.max() == 425
    P1, P2, P3 = meshio.CreateP1P2P3(T, P)
    area = 0.5 * np.linalg.norm(np.cross(P2 - P1, P3 - P1), axis=1)
    assert np.all(area > 0)
    # a disc of radius about 1 in z = 0 (polygonal rim: area within 5 % of pi)
    assert abs(area.sum() - np.pi) < 0.05 * np.pi and np.all(np.hypot(P[:, 1], P[:, 2]) <= 1.01)
    assert np.all(P[:, 3] == 0.0)


def test_off_and_stl_round_trips(tmp_path):
    P, T = meshio.GoCadAsciiReader(os.path.join(G, "CircleMesh_1a_500Faces.ts"))
    out = meshio.OFFExport(P, T, T.shape[0], P.shape[0], str(tmp_path / "disc"))
    with open(out, "rb") as f:
        head = f.read(40)
    assert head.startswith(b"OFF\r\n%d %d 0\r\n" % (P.shape[0], T.shape[0]))
    P2, T2 = meshio.OFFReader(out)
    assert np.array_equal(T2, T) and np.max(np.abs(P2 - P)) <= 5e-14
    P1, Q2, Q3 = meshio.CreateP1P2P3(T, P)
    nrm, _ = api.CreateFaceNormalAndMidPoint(P1, Q2, Q3)
    out = meshio.STLExport(P1, Q2, Q3, nrm, str(tmp_path / "disc"))
    S, TS = meshio.STLReader(out)
    R1, R2, R3 = meshio.CreateP1P2P3(TS, S)
    for a, b in ((R1, P1), (R2, Q2), (R3, Q3)):
        assert np.max(np.abs(a - b)) <= 5e-14


def test_gocad_atoms_and_multiple_surfaces(tmp_path):
    f = tmp_path / "atoms.ts"
    f.write_text("GOCAD TSurf 1\nTFACE\nVRTX 1 0 0 0\nVRTX 2 1 0 0\nVRTX 3 0 1 0\nATOM 4 2\nTRGL 1 2 3\n"
                 "TRGL 4 3 1\nEND\n")
    P, T = meshio.GoCadAsciiReader(str(f))
    assert P.shape == (4, 4) and np.array_equal(P[3], [4, 1, 0, 0])
    assert np.array_equal(T, [[1, 2, 3], [4, 3, 1]])
    g = tmp_path / "two.ts"
    g.write_text("GOCAD TSurf 1\nPROJECTION Unknown\nVRTX 1 0 0 0\n")
    with pytest.raises(meshio.MeshFormatError):
        meshio.GoCadAsciiReader(str(g))
