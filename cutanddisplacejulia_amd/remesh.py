"""Isotropic remeshing for growth loops (host), in place of BuildCGAL.PolygonRemeshingCGAL.

The reference's propagation loop (test/test_PropagationTest.jl:151-561) writes the mesh with OFFExport, remeshes the
file with CGAL 4.13.1's isotropic remeshing (BuildCGAL.PolygonRemeshingCGAL(OutputDirectory, target_edge_length),
not part of the reference) and reads the result back with OFFReader.  `PolygonRemeshingCGAL` keeps that file
interface; `PolygonRemeshing` works on the reference's Points / Triangles arrays.  Both run cdj_remesh_isotropic
(csrc/remesh.cpp): Botsch & Kobbelt's split / collapse / flip / relax / project rounds with the boundary kept in
shape.  No GPU is used.
"""
import ctypes
import os

import numpy as np

from . import _lib, meshio

ITERATIONS = 3  # rounds per call (the iteration count BuildCGAL passes to CGAL is not in the reference)


def remesh_arrays(points, tris, target_edge_length, iterations=ITERATIONS):
    """points: n x 3 coordinates; tris: m x 3 0-based corner indices.  Returns (points, tris) of the remeshed
    surface in the same layouts."""
    L = _lib.load()
    P = np.asfortranarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    T = np.asfortranarray(np.asarray(tris, dtype=np.int64).reshape(-1, 3))
    h = ctypes.c_void_p()
    npo, nto = ctypes.c_int64(), ctypes.c_int64()
    rc = L.cdj_remesh_isotropic(P.ctypes.data, P.shape[0], T.ctypes.data, T.shape[0], float(target_edge_length),
                                int(iterations), ctypes.byref(h), ctypes.byref(npo), ctypes.byref(nto))
    if rc:
        msg = {_lib.CDJ_ERR_DEGENERATE: "Slither triangles or triangles with no area exist on your surface"}.get(
            rc, "cdj_remesh_isotropic: invalid input (%d)" % rc)
        raise _lib.CdjError(rc, msg)
    Po = np.zeros((npo.value, 3), order="F")
    To = np.zeros((nto.value, 3), dtype=np.int64, order="F")
    _lib.check(L.cdj_remesh_take(h, Po.ctypes.data, To.ctypes.data))
    return Po, To


def PolygonRemeshing(Points, Triangles, target_edge_length, iterations=ITERATIONS):
    """The reference's arrays in and out: Points n x 4 (index, x, y, z), Triangles m x 3 with 1-based indices into
    Points[:, 0] (meshio conventions)."""
    Points = np.asarray(Points, dtype=np.float64)
    Triangles = np.asarray(Triangles)
    ids = Points[:, 0].astype(np.int64)
    pos = np.full(int(ids.max()) + 1, -1, dtype=np.int64)
    pos[ids] = np.arange(ids.size)
    tris = pos[Triangles.astype(np.int64)]
    if np.any(tris < 0):
        raise ValueError("Triangles refer to points that are not in Points")
    P, T = remesh_arrays(Points[:, 1:4], tris, target_edge_length, iterations)
    return np.column_stack([np.arange(1, P.shape[0] + 1, dtype=np.float64), P]), T + 1


def PolygonRemeshingCGAL(OutputDirectory, target_edge_length, iterations=ITERATIONS):
    """File form of BuildCGAL.PolygonRemeshingCGAL: remeshes the OFF file at OutputDirectory and returns the path
    of the remeshed OFF file, written beside it."""
    Points, Triangles = meshio.OFFReader(OutputDirectory)
    P, T = PolygonRemeshing(Points, Triangles, target_edge_length, iterations)
    base = os.path.splitext(os.path.abspath(OutputDirectory))[0] + "-Remeshed"
    return meshio.OFFExport(P, T, T.shape[0], P.shape[0], base)
