"""Mesh input and output with the reference's file conventions (host side, O(N)).

  GoCadAsciiReader   src/GoCadAsciiReader.jl   single-surface GoCad TSurf (.ts): VRTX / TRGL / ATOM records
  STLReader          src/STLReader.jl          ASCII STL: every `vertex` record is a point, facet k = points 3k+1..3k+3
  OFFReader          src/OFFReader.jl          OFF: counts on line 2, 0-based triangle indices
  OFFExport          src/OFFExport.jl          OFF with 13 decimals and CRLF line ends
  STLExport          src/STLExport.jl          ASCII STL with 13 decimals and CRLF line ends
  CreateP1P2P3       src/CreateP1P2P3.jl       triangle corner arrays from Points / Triangles

Points are n x 4 (index, x, y, z) and Triangles m x 3 with 1-based indices into Points[:, 0], as in the reference.
Numbers are parsed as the reference's readdlm does (correctly rounded decimal to binary64).
"""
import os

import numpy as np


class MeshFormatError(ValueError):
    pass


def _tokens(path):
    with open(path) as f:
        for line in f:
            t = line.split()
            if t:
                yield t


def GoCadAsciiReader(path):
    """Points (n x 4) and Triangles (m x 3, 1-based) of a single-surface GoCad ASCII file."""
    pts, tris, atoms = [], [], []
    for t in _tokens(path):
        if t[0] == "PROJECTION":
            raise MeshFormatError("More than one Gocad ascii surface in the file, this function can only deal with 1")
        if t[0] == "VRTX":
            pts.append([float(t[1]), float(t[2]), float(t[3]), float(t[4])])
        elif t[0] == "TRGL":
            tris.append([float(t[1]), float(t[2]), float(t[3])])
        elif t[0] == "ATOM":
            atoms.append([float(t[1]), float(t[2])])
    P = np.array(pts, dtype=np.float64).reshape(-1, 4)
    T = np.array(tris, dtype=np.float64).reshape(-1, 3)
    if atoms:
        # an ATOM record is a new vertex id placed on an existing vertex; rows are then sorted by id
        # (GoCadAsciiReader.jl: Atoms appended to Points, then sortslices)
        by_id = {}
        for row in P:
            by_id.setdefault(row[0], row[1:4])
        extra = np.array([[a[0], *by_id[a[1]]] for a in atoms])
        P = np.vstack([P, extra])
        P = P[np.lexsort(P.T[::-1])]
    # renumber: a row whose id differs from its position takes that position, and triangles follow
    # (GoCadAsciiReader.jl's loop, in its order)
    for i in range(P.shape[0]):
        if P[i, 0] != i + 1:
            T[T == P[i, 0]] = i + 1
            P[i, 0] = i + 1
    return P, T.astype(np.int64)


def STLReader(path):
    """Points (3m x 4) and Triangles (m x 3) of an ASCII STL file (no vertex merging)."""
    pts = [[float(t[1]), float(t[2]), float(t[3])] for t in _tokens(path) if t[0] == "vertex"]
    if len(pts) % 3:
        raise MeshFormatError("STL vertex count is not a multiple of 3")
    n = len(pts)
    P = np.column_stack([np.arange(1, n + 1, dtype=np.float64), np.array(pts, dtype=np.float64).reshape(-1, 3)])
    T = np.arange(1, n + 1, dtype=np.int64).reshape(-1, 3)
    return P, T


def OFFReader(path):
    """Points (n x 4) and Triangles (m x 3, 1-based) of an OFF file."""
    rows = list(_tokens(path))
    if not rows or rows[0][0] != "OFF":
        raise MeshFormatError("not an OFF file")
    n_points, n_tris = int(rows[1][0]), int(rows[1][1])
    P = np.array([[float(v) for v in r[:3]] for r in rows[2:2 + n_points]], dtype=np.float64).reshape(-1, 3)
    T = np.array([[int(v) for v in r[1:4]] for r in rows[2 + n_points:2 + n_points + n_tris]],
                 dtype=np.int64).reshape(-1, 3) + 1
    return np.column_stack([np.arange(1, n_points + 1, dtype=np.float64), P]), T


def OFFExport(Points, Triangles, n_tris, n_points, filename):
    """Write `filename`.off (OFFExport.jl layout) and return its absolute path."""
    path = os.path.abspath(filename + ".off")
    P = np.asarray(Points, dtype=np.float64)
    T = np.asarray(Triangles, dtype=np.int64)
    with open(path, "w", newline="") as f:
        f.write("OFF\r\n%d %d 0\r\n" % (n_points, n_tris))
        for i in range(n_points):
            f.write("%0.13f %0.13f %0.13f\r\n" % (P[i, 1], P[i, 2], P[i, 3]))
        for i in range(n_tris):
            f.write("3 %d %d %d\r\n" % (T[i, 0] - 1, T[i, 1] - 1, T[i, 2] - 1))
        f.write("\r\n")
    return path


def STLExport(P1, P2, P3, FaceNormalVector, filename):
    """Write `filename`.stl (STLExport.jl layout) and return its absolute path."""
    path = os.path.abspath(filename + ".stl")
    P1, P2, P3, N = (np.asarray(a, dtype=np.float64) for a in (P1, P2, P3, FaceNormalVector))
    with open(path, "w", newline="") as f:
        f.write("solid STL generated by Herr Davis\r\n")
        for i in range(P1.shape[0]):
            f.write("  facet normal %0.13f %0.13f %0.13f\r\n" % tuple(N[i]))
            f.write("    outer loop\r\n")
            for P in (P1, P2, P3):
                f.write("      vertex %0.13f %0.13f %0.13f\r\n" % tuple(P[i]))
            f.write("    endloop\r\n  endfacet\r\n")
        f.write("endsolid vcg\r\n")
    return path


def CreateP1P2P3(Triangles, Points):
    """Corner arrays (m x 3 each) of the triangles: the coordinates of the first Points row with that index (zeros
    when no row has it, as the reference leaves them)."""
    P = np.asarray(Points, dtype=np.float64)
    T = np.asarray(Triangles, dtype=np.int64)
    first = {}
    for j in range(P.shape[0]):
        first.setdefault(int(P[j, 0]), j)
    out = []
    for c in range(3):
        C = np.zeros((T.shape[0], 3))
        for i, v in enumerate(T[:, c]):
            j = first.get(int(v))
            if j is not None:
                C[i] = P[j, 1:4]
        out.append(C)
    return tuple(out)
