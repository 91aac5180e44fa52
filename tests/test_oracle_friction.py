"""Pins of the frictional-slip oracle (oracle/friction.py) on the CPU.

* The reference's own case, test_TDvsBurgmannLinearFrictionCrack.jl, on its own mesh
  (tests/golden/BladeFaultShort_500Tris_Vertical_NorthDip_0.5.ts, flattened and scaled as there): with zero
  friction coefficient and a cohesion rising linearly from the centre, the slip must be within 36 % of Burgmann et
  al. (1994) away from the tips and the summed error below 1400 %; the frictionless slip of the same crack within
  36 % of Pollard & Segall (1987).
* The LCP solution is complementary: x >= 0, y = a x + b >= 0 and x'y = 0 (to rounding).
* With cohesion far above the driving shear nothing slips and the opening solves the normal block of the matrix
  alone (numpy's solve of that block).
* Fischer-Newton on small random LCPs with a known solution (a = M M' + I, b chosen from a complementary pair).
"""
import numpy as np

from oracle import analytic as AN
from oracle import friction as FR
from oracle import oracle as O
from tests import gocad

SCL = 1e-4


def blade():
    P, T = gocad.read_ts("BladeFaultShort_500Tris_Vertical_NorthDip_0.5.ts")
    P[:, 2] = P[:, 2] * 0
    P[:, 1:4] = P[:, 1:4] * SCL
    P1, P2, P3 = gocad.p1p2p3(P, T)
    nrm, mid = O.face_normal_midpoint(P1, P2, P3)
    n = P1.shape[0]
    G, nu = 1.0, 0.25
    lam = 2 * G * nu / (1 - 2 * nu)
    stress = (np.zeros(n), np.full(n, -0.05 * SCL), np.zeros(n), np.full(n, SCL), np.zeros(n), np.zeros(n))
    return P1, P2, P3, nrm, mid, n, G, nu, lam, stress


def burgmann_errors(mid, Dss, DssF, G, nu):
    x = mid[:, 0]
    good = np.abs(mid[:, 2]) < SCL
    x, Dss, DssF = x[good], np.abs(Dss[good]), np.abs(DssF[good])
    an = AN.pollard_segall1987_slip(G, nu, 0.0, SCL, x, SCL)
    anf = AN.burgmann1994_linear_friction_slip(G, nu, SCL, x, SCL)
    with np.errstate(divide="ignore", invalid="ignore"):
        p, pf = (100.0 / an) * Dss, (100.0 / anf) * DssF
    p[np.abs(p) == np.inf] = np.nan
    pf[np.abs(pf) == np.inf] = np.nan
    sums = (np.nansum(p - 100), np.nansum(pf - 100))
    inner = np.abs(x) < 0.9 * SCL
    worst = (np.nanmax(np.abs(p[inner] - 100)), np.nanmax(np.abs(pf[inner] - 100)))
    return worst, sums


def test_oracle_burgmann_friction_crack():
    P1, P2, P3, nrm, mid, n, G, nu, lam, stress = blade()
    z = np.zeros(n)
    DnF, DssF, DdsF, info = FR.slip_friction(P1, P2, P3, nu, G, lam, mid, nrm, 0, stress, (z, z, z), 0.0,
                                             np.abs(mid[:, 0]))
    A = O.assemble_A(O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, 0))
    Tn, Tds, Tss = O.normal_shear_tractions(stress, nrm)
    x = np.linalg.solve(A, np.concatenate([Tn, Tss, Tds]))
    Dss = x[n:2 * n]
    worst, sums = burgmann_errors(mid, Dss, DssF, G, nu)
    print("burgmann: %d its, worst %.1f %% / %.1f %%, sums %.0f / %.0f" % (info["iterations"], *worst, *sums))
    assert max(worst) <= 36 and max(sums) <= 1400
    # complementarity of the LCP solution
    xx, yy = info["x"], info["y"]
    s = max(np.max(np.abs(xx)), np.max(np.abs(yy)))
    assert np.min(xx) >= 0 and np.min(yy) >= -1e-10 * s and np.max(np.abs(xx * yy)) <= 1e-10 * s * s


def test_locked_shear_reduces_to_the_normal_block():
    """Cohesion far above the driving shear: no element slips, and the opening under tension solves the normal
    block alone, A[Tn rows, Dn cols] Dn = Tn."""
    P1, P2, P3, nrm, mid, n, G, nu, lam, _ = blade()
    z = np.zeros(n)
    stress = (z, np.full(n, 0.3 * SCL), z, np.full(n, SCL), z, z)
    Dn, Dss, Dds, info = FR.slip_friction(P1, P2, P3, nu, G, lam, mid, nrm, 0, stress, (z, z, z), 0.0, 1e3 * SCL)
    A = O.assemble_A(O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, 0))
    Tn, Tds, Tss = O.normal_shear_tractions(stress, nrm)
    ref = np.linalg.solve(A[:n, :n], Tn)
    assert np.max(np.abs(Dn - ref)) <= 1e-9 * np.max(np.abs(ref))
    assert np.max(np.abs(Dss)) <= 1e-9 * np.max(np.abs(ref)) and np.max(np.abs(Dds)) <= 1e-9 * np.max(np.abs(ref))


def test_fischer_newton_known_solution():
    rng = np.random.default_rng(3)
    for N in (5, 40, 200):
        M = rng.normal(size=(N, N))
        A = M @ M.T / N + np.eye(N)
        xs = np.where(rng.random(N) < 0.5, rng.random(N) + 0.1, 0.0)
        ys = np.where(xs > 0, 0.0, rng.random(N) + 0.1)
        b = ys - A @ xs
        x, info = FR.fischer_newton(A, b)
        assert np.max(np.abs(x - xs)) <= 1e-10, (N, info)
