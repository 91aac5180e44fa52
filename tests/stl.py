"""ASCII STL reader for the reference's test meshes in tests/golden (test infrastructure).

Restates STLReader.jl: every `vertex` record becomes a point in file order (no merging of shared vertices) and
facet k is the triangle (3k+1, 3k+2, 3k+3).
"""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_stl(name):
    """Points (n x 4: index, x, y, z) and Triangles (m x 3, 1-based)."""
    pts = []
    with open(os.path.join(GOLDEN, name)) as f:
        for line in f:
            t = line.split()
            if t and t[0] == "vertex":
                pts.append([float(t[1]), float(t[2]), float(t[3])])
    P = np.column_stack([np.arange(1, len(pts) + 1, dtype=np.float64), np.array(pts)])
    T = np.arange(1, len(pts) + 1).reshape(-1, 3)
    return P, T
