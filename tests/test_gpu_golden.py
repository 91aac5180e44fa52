"""The device TD kernels (displacement and strain influences, cdj_td through api.TD) against the 60-digit values
of tests/golden/td_golden.json: near sides (1e-2 ... 1e-6 side lengths), near a vertex, near the plane, far
away and on the free surface, full and half space.

Tolerance, per point and slip mode, relative to the mode's largest component: 1024 eps kappa, with
kappa = max(1, (L/d)^2, (r/L)^3) the conditioning of the angular-dislocation formulas at the point (d: distance to
the nearest side, r: to the centroid, L: the longest side).  The device evaluates the free-surface strain by
forward-mode differentiation and contracts products into FMAs, so its rounding differs from the oracle's; the
cancellation that sets the error is the same.
"""
import json
import os
from collections import defaultdict

import numpy as np
import pytest

from cutanddisplacejulia_amd import api

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "td_golden.json")


def test_device_td_matches_60_digit_values():
    with open(GOLDEN) as f:
        g = json.load(f)
    nu = float(g["nu"])
    groups = defaultdict(list)
    for c in g["cases"]:
        groups[c["triangle"]].append(c)
    worst, n = 0.0, 0
    for name, cases in groups.items():
        P = [np.array([[float(v) for v in cases[0][k]]]) for k in ("P1", "P2", "P3")]
        X = np.array([[float(v) for v in c["X"]] for c in cases])
        hs = cases[0]["hs"]
        S, D = api.TD(X[:, 0], X[:, 1], X[:, 2], *P, [1.0], [1.0], [1.0], nu, 1.0, 1, 1, hs)
        for i, c in enumerate(cases):
            gu = np.array([[float(v) for v in r] for r in c["disp"]])
            ge = np.array([[float(v) for v in r] for r in c["strain"]])
            u = np.array([[getattr(D, k)[i, 0] for k in api.DISP_FIELDS[3 * m:3 * m + 3]] for m in range(3)])
            e = np.array([[getattr(S, k)[i, 0] for k in api.STRAIN_FIELDS[6 * m:6 * m + 6]] for m in range(3)])
            eu = max(np.max(np.abs(u[m] - gu[m])) / np.max(np.abs(gu[m])) for m in range(3))
            ee = max(np.max(np.abs(e[m] - ge[m])) / np.max(np.abs(ge[m])) for m in range(3))
            worst = max(worst, eu / (EPS * c["kappa"]), ee / (EPS * c["kappa"]))
            tol = 1024 * EPS * c["kappa"]
            assert eu <= tol and ee <= tol, (name, c["label"], eu, ee, tol)
            n += 1
    print("device TD vs 60-digit values, %d points: worst error / (eps kappa) %.1f" % (n, worst))
