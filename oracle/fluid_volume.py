"""CPU restatement of the constant-fluid-volume slip solve (test infrastructure only: imported by tests/, never by
the product package).

Follows SlipCalculator3D.jl:242-475 (the MixedBoundaryConditionsFluidVolume method, single-fracture branch) and
ComputePressurisedCrackDn.jl:1-215: the base collocation matrix A and traction vector b, A scaled by
mean(half perimeter) / G and inverted, the frictional complementarity problem prepared once with the reference's
fixed parameters (mu = 0.1, Sf = 0), a normalising pressure Norm from the tractions of a uniform opening of the
target volume, Norm adjusted until the volume at x = 10 lies in [V, 2.5 V], then a Brent minimisation of
(V - volume(x))^2 over x in [0, 10] (abs_tol 0.001), a golden-section search over [0, 50] when that minimiser gives
a negative volume, and the slip at the minimiser (NaN when its volume exceeds 2 |V|).

volume(x): the fracture elements' normal traction gets + x Norm, D = A^-1 b, and when any opening is positive the
frictional problem with that D gives the slip (Dn, Dss, Dds); volume = sum(Dn * Area) over the fracture.

The reference minimises with Optim 1.11.0 (Brent(), GoldenSection(); not vendored).  Restated here are those two
published univariate methods: Brent's (1973, "Algorithms for minimization without derivatives", ch. 5) parabolic
interpolation with golden-section fallback, and the golden-section search, both with the stopping rule
|x - midpoint| <= 2 tol - (upper - lower) / 2, tol = rel_tol |x| + abs_tol, rel_tol = sqrt(eps).  The multi-fracture
branch of the reference ends in an error (SlipCalculator3D.jl:303); it is restated as that error.
"""
import numpy as np

from oracle import crack_tip as CT
from oracle import friction as F
from oracle import oracle as O

GOLDEN = 0.5 * (3.0 - np.sqrt(5.0))
REL_TOL = np.sqrt(np.finfo(float).eps)
MULTI_FRACTURE_ERROR = "Check scl here - what do we use for one crack below"
MU, SF = 0.1, 0.0  # SlipCalculator3D.jl:275-276
NORM_LOOP_CAP = 200


class VolumeError(RuntimeError):
    pass


def brent(f, lo, hi, rel_tol=REL_TOL, abs_tol=np.finfo(float).eps, max_iter=1000):
    """Minimiser of f on [lo, hi] by Brent's method; returns (x, f(x), evaluations)."""
    x = lo + GOLDEN * (hi - lo)
    fx = f(x)
    calls = 1
    step = old_step = 0.0
    v = w = x   # previous and second previous best points
    fv = fw = fx
    for _ in range(max_iter):
        p = q = 0.0
        tol = rel_tol * abs(x) + abs_tol
        mid = (hi + lo) / 2
        if abs(x - mid) <= 2 * tol - (hi - lo) / 2:
            break
        if abs(old_step) > tol:
            r = (x - w) * (fx - fv)
            q = (x - v) * (fx - fw)
            p = (x - v) * q - (x - w) * r
            q = 2 * (q - r)
            if q > 0:
                p = -p
            else:
                q = -q
        if abs(p) < abs(q * old_step / 2) and p > q * (lo - x) and p < q * (hi - x):
            old_step = step
            step = p / q
            if (x + step - lo) < 2 * tol or (hi - (x + step)) < 2 * tol:
                step = tol if x < mid else -tol
        else:
            old_step = (hi - x) if x < mid else (lo - x)
            step = GOLDEN * old_step
        u = x + step if abs(step) >= tol else x + (tol if step > 0 else -tol)
        fu = f(u)
        calls += 1
        if fu < fx:
            if u < x:
                hi = x
            else:
                lo = x
            v, fv = w, fw
            w, fw = x, fx
            x, fx = u, fu
        else:
            if u < x:
                lo = u
            else:
                hi = u
            if fu <= fw or w == x:
                v, fv = w, fw
                w, fw = u, fu
            elif fu <= fv or v == x or v == w:
                v, fv = u, fu
    return x, fx, calls


def golden_section(f, lo, hi, rel_tol=REL_TOL, abs_tol=np.finfo(float).eps, max_iter=1000):
    """Minimiser of f on [lo, hi] by golden-section search; returns (x, f(x), evaluations)."""
    x = lo + GOLDEN * (hi - lo)
    fx = f(x)
    calls = 1
    for _ in range(max_iter):
        tol = rel_tol * abs(x) + abs_tol
        if abs(x - (hi + lo) / 2) <= 2 * tol - (hi - lo) / 2:
            break
        if hi - x > x - lo:
            u = x + GOLDEN * (hi - x)
            fu = f(u)
            calls += 1
            if fu < fx:
                lo, x, fx = x, u, fu
            else:
                hi = u
        else:
            u = x - GOLDEN * (x - lo)
            fu = f(u)
            calls += 1
            if fu < fx:
                hi, x, fx = x, u, fu
            else:
                lo = u
    return x, fx, calls


class Problem:
    """The pieces SlipCalculator3D.jl:248-282 prepares once."""

    def __init__(self, P1, P2, P3, nu, G, lam, mid, nrm, hs_flag, stress6, tractions, flag, volume):
        n = P1.shape[0]
        self.n = n
        self.flag = np.asarray(flag).ravel()
        if np.any(self.flag > 1):
            raise VolumeError(MULTI_FRACTURE_ERROR)
        self.V = float(np.asarray(volume).ravel()[0])
        Tn, Tds, Tss = O.normal_shear_tractions(stress6, nrm)
        Tn, Tss, Tds = Tn + tractions[0], Tss + tractions[1], Tds + tractions[2]
        self.b = np.concatenate([Tn, Tss, Tds])
        A = O.assemble_A(O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, hs_flag))
        self.area, hp = CT.area_half_perimeter(P1, P2, P3)
        self.scl = np.mean(hp) / G
        self.A = A * self.scl
        self.C = np.linalg.inv(self.A)
        mu, Sf = np.full(n, MU), np.full(n, SF)
        self.a, self.bb0 = F.friction_lcp(self.C, np.zeros(3 * n), mu, Sf, n)
        self.frac = self.flag == 1
        self.free = self.flag == 0
        # single fracture: Norm from a uniform opening of the target volume (SlipCalculator3D.jl:349-356)
        h = (self.V / np.sum(self.area)) / self.scl
        self.norm = float(np.max((self.A @ np.concatenate([np.full(n, h), np.zeros(2 * n)]))[:n]))
        self.evaluations = 0

    def slip(self, x):
        """ComputePressurisedCrackDn(x::Tractions, ...): (Dn, Dss, Dds) at pressure x Norm."""
        n = self.n
        self.evaluations += 1
        B = self.b.copy()
        B[:n][self.frac] += x * self.norm
        D = self.C @ B
        Dn_ = D[:n]
        z = np.zeros(n)
        if self.free.any():
            if not np.any(Dn_[~self.free] > 0):
                return z, z.copy(), z.copy()
            bb = self.bb0.copy()
            arb = abs(np.min(bb[:n][self.free]) * 2)
            bb[:n][self.free] += arb
        else:
            if not np.any(Dn_ > 0):
                return z, z.copy(), z.copy()
            bb = self.bb0.copy()
            arb = 0.0
        bb[:3 * n] += D
        xs, _ = F.fischer_newton(self.a, bb)
        y = self.a @ xs + bb
        L = [slice(k * n, (k + 1) * n) for k in range(5)]
        Dn = y[L[0]] * self.scl
        Dss = (y[L[1]] - xs[L[3]]) * self.scl
        Dds = (y[L[2]] - xs[L[4]]) * self.scl
        if self.free.any():
            Dn = Dn - arb
        return Dn, Dss, Dds

    def volume(self, x):
        Dn, _, _ = self.slip(x)
        return float(np.sum(Dn[self.frac] * self.area[self.frac]))

    def misfit(self, x):
        return abs(self.V - self.volume(x)) ** 2


def slip_fluid_volume(P1, P2, P3, nu, G, lam, mid, nrm, hs_flag, stress6, tractions, flag, volume):
    """SlipCalculator3D(..., MixedBoundaryConditionsFluidVolume, FractureFlag) -> (Dn, Dss, Dds, info)."""
    pb = Problem(P1, P2, P3, nu, G, lam, mid, nrm, hs_flag, stress6, tractions, flag, volume)
    V = pb.V
    huge = pb.volume(10.0)
    it = 0
    while huge < V:
        pb.norm *= 1.05
        huge = pb.volume(10.0)
        it += 1
        if it > NORM_LOOP_CAP:
            raise VolumeError("the target volume is not reached at any pressure scale")
    reduce = 0.3
    it = 0
    while huge > V * 2.5:
        pb.norm *= (1 - reduce)
        huge = pb.volume(10.0)
        if huge < V:
            pb.norm /= (1 - reduce)
            reduce /= 2
            huge = pb.volume(10.0)
        it += 1
        if it > NORM_LOOP_CAP:
            raise VolumeError("the pressure scale does not settle")
    x, _, _ = brent(pb.misfit, 0.0, 10.0, abs_tol=0.001)
    method = "brent"
    vol_out = pb.volume(x)
    if (V - vol_out) > V:
        x, _, _ = golden_section(pb.misfit, 0.0, 50.0)
        method = "golden_section"
    Dn, Dss, Dds = pb.slip(x)
    vol = float(np.sum(Dn[pb.frac] * pb.area[pb.frac]))
    if abs(vol) > abs(V * 2):
        Dn, Dss, Dds = Dn * np.nan, Dss * np.nan, Dds * np.nan
    return Dn, Dss, Dds, {"x": x, "pressure": x * pb.norm, "norm": pb.norm, "volume": vol, "method": method,
                          "evaluations": pb.evaluations, "scale": pb.scl}
