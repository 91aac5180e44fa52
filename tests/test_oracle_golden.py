"""The oracle's displacement and strain influences of single triangles against 60-digit values
(tests/golden/td_golden.json, written by tests/golden/make_td_golden.py: an mpmath restatement of the displacement
path, strains by central differences of it), at points 1e-2 ... 1e-6 side lengths from a side, off and in the
triangle's plane, near a vertex, 1e-12 ... 1e-4 off the plane at the centroid, far away, and on and just under
the free surface; full space, half space, a triangle 0.05 below the surface and one with a vertical side (whose
free-surface correction the reference skips, TD.jl:1265).

Tolerance, per point and slip mode, relative to the mode's largest component there: 256 eps kappa, with
kappa = max(1, (L/d)^2, (r/L)^3) (d: distance to the nearest side, r: to the centroid, L: the longest side).  Near
a side the angular-dislocation formulas (TD.jl:1177-1220, 1412-1496, 1753-1828, 2084-2471) cancel like
(r - zeta) ~ d^2 / r, so any double-precision evaluation of them carries ~eps (L/d)^2; far away the three
dislocations' fields cancel to the triangle's.  The observed error / (eps kappa) stays between 0.005 and 60 over
five decades of d, the signature of that cancellation rather than of a wrong term, which would give an error
independent of d.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

EPS = np.finfo(np.float64).eps
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "td_golden.json")


def _load():
    with open(GOLDEN) as f:
        g = json.load(f)
    return float(g["nu"]), g["cases"]


def _case_arrays(c):
    P = [np.array([[float(v) for v in c[k]]]) for k in ("P1", "P2", "P3")]
    X = np.array([float(v) for v in c["X"]])
    gu = np.array([[float(v) for v in r] for r in c["disp"]])
    ge = np.array([[float(v) for v in r] for r in c["strain"]])
    return P, X, gu, ge


def _errors(u, e, gu, ge):
    eu = max(np.max(np.abs(u[m] - gu[m])) / np.max(np.abs(gu[m])) for m in range(3))
    ee = max(np.max(np.abs(e[m] - ge[m])) / np.max(np.abs(ge[m])) for m in range(3))
    return eu, ee


def _oracle(P, X, nu, hs):
    s, d = O.td(X[0:1], X[1:2], X[2:3], *P, [1.0], [1.0], [1.0], nu, 1.0, 1, 1, hs)
    u = np.array([[d[k][0, 0] for k in O.DISP_FIELDS[3 * m:3 * m + 3]] for m in range(3)])
    e = np.array([[s[k][0, 0] for k in O.STRAIN_FIELDS[6 * m:6 * m + 6]] for m in range(3)])
    return u, e


def test_golden_file_covers_the_cases():
    nu, cases = _load()
    tris = {c["triangle"] for c in cases}
    assert tris == {"full_space", "half_space", "half_space_shallow", "half_space_vertical_side"}
    assert len(cases) >= 70
    assert max(c["kappa"] for c in cases) >= 1e12          # points 1e-6 side lengths from a side
    assert any(float(c["X"][2]) == 0.0 for c in cases if c["hs"])


def test_oracle_matches_60_digit_values():
    nu, cases = _load()
    worst = 0.0
    for c in cases:
        P, X, gu, ge = _case_arrays(c)
        u, e = _oracle(P, X, nu, c["hs"])
        eu, ee = _errors(u, e, gu, ge)
        tol = 256 * EPS * c["kappa"]
        worst = max(worst, eu / (EPS * c["kappa"]), ee / (EPS * c["kappa"]))
        assert eu <= tol and ee <= tol, (c["triangle"], c["label"], eu, ee, tol)
    print("worst error / (eps kappa) over %d points: %.1f" % (len(cases), worst))


def test_generic_points_agree_to_rounding():
    """Points at least 0.3 side lengths from every side and within 8 side lengths: <= 2e-13 of the mode's scale."""
    nu, cases = _load()
    n = 0
    for c in cases:
        if c["kappa"] > 16:
            continue
        P, X, gu, ge = _case_arrays(c)
        u, e = _oracle(P, X, nu, c["hs"])
        eu, ee = _errors(u, e, gu, ge)
        assert eu <= 2e-13 and ee <= 2e-13, (c["triangle"], c["label"], eu, ee)
        n += 1
    assert n >= 20


def test_golden_values_separate_full_and_half_space():
    """The full-space solution misses the half-space values by far more than the tolerance at every half-space
    point, so the comparison above pins the free-surface correction and the image triangle."""
    nu, cases = _load()
    for c in cases:
        if not c["hs"] or c["kappa"] > 1e4:
            continue
        P, X, gu, ge = _case_arrays(c)
        u, e = _oracle(P, X, nu, 0)
        eu, ee = _errors(u, e, gu, ge)
        assert max(eu, ee) > 1e4 * 256 * EPS * c["kappa"], (c["triangle"], c["label"], eu, ee)
