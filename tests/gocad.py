"""Minimal GoCad TSurf reader for the reference's test meshes in tests/golden (test infrastructure).

Restates GoCadAsciiReader.jl (VRTX / TRGL records; vertices renumbered 1..n in file order) and CreateP1P2P3.jl.
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_ts(name):
    """Points (n x 4: index, x, y, z) and Triangles (m x 3, 1-based) of a single-surface .ts file."""
    pts, tris = [], []
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] in ("VRTX", "PVRTX"):
                pts.append([float(t[1]), float(t[2]), float(t[3]), float(t[4])])
            elif t[0] == "TRGL":
                tris.append([int(t[1]), int(t[2]), int(t[3])])
    P = np.array(pts)
    T = np.array(tris)
    # renumber vertices to their row (GoCadAsciiReader.jl: Points[i,1] != i -> remap)
    remap = {int(idx): i + 1 for i, idx in enumerate(P[:, 0])}
    T = np.vectorize(lambda v: remap[v])(T)
    P[:, 0] = np.arange(1, len(P) + 1)
    return P, T


def p1p2p3(P, T):
    """Triangle corner arrays (CreateP1P2P3.jl)."""
    return P[T[:, 0] - 1, 1:4].copy(), P[T[:, 1] - 1, 1:4].copy(), P[T[:, 2] - 1, 1:4].copy()


def append(P1s, T1s, P2s, T2s):
    """DataAppender3D.jl: second mesh appended, its elements flagged 1."""
    P2s = P2s.copy()
    sz = len(P1s)
    P2s[:, 0] += sz
    return np.vstack([P1s, P2s]), np.vstack([T1s, T2s + sz]), np.concatenate([np.zeros(len(T1s)), np.ones(len(T2s))])
