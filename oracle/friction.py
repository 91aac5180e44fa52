"""CPU restatement of the frictional slip solve (test infrastructure only: imported by tests/, never by the
product package).

Follows SlipCalculator3D.jl:36-233 (the MixedBoundaryConditionsFriction method and its two helpers): remote stresses
plus tractions resolved on the elements, the early return when no element can overcome its frictional strength,
the base collocation matrix scaled by mean(half perimeter) / G, its inverse C, the 5n x 5n linear complementarity
problem of Kaven et al. (2012, eq. 28, as modified there) and the slip read back from its solution.

The reference solves the LCP with FischerNewton 0.1.4 (a package the reference does not vendor).  Restated here is
the published Fischer-Newton method of Erleben (num4lcp, "Numerical methods for linear complementarity problems in
physics-based animation", 2013): the Fischer-Burmeister function phi(y, x) = sqrt(y^2 + x^2) - y - x, the
perturbation strategy for its generalized Jacobian, a sufficient-descent check with the gradient as fallback (also
taken when the Newton system is exactly singular) and a projected Armijo backtracking line search.  The Newton system is solved exactly (LU) where num4lcp uses restarted
GMRES; the parameters are those of num4lcp's defaults except the iteration cap and tolerances named below.  The
LCP solution (x >= 0, y = a x + b >= 0, x'y = 0) is unique for these matrices, so any convergent method gives it.
"""
import numpy as np

from oracle import crack_tip as CT
from oracle import oracle as O

FN = dict(max_iter=100, tol_abs=1e-26, alpha=0.5, beta=1e-3, gamma=1e-28, rho=np.finfo(float).eps)


def fb_phi(y, x):
    return np.sqrt(y * y + x * x) - y - x


def fischer_newton(A, b, max_iter=FN["max_iter"], tol_abs=FN["tol_abs"], alpha=FN["alpha"], beta=FN["beta"],
                   gamma=FN["gamma"], rho=FN["rho"]):
    """x with x >= 0, a x + b >= 0 and x'(a x + b) = 0; returns (x, {iterations, err})."""
    N = b.shape[0]
    x = np.zeros(N)
    err = 1e20
    it = 0
    while it < max_iter:
        y = A @ x + b
        phi = fb_phi(y, x)
        err = 0.5 * (phi @ phi)
        if err < tol_abs:
            break
        S = (np.abs(phi) < gamma) & (np.abs(x) < gamma)
        px = x.copy()
        d = np.sign(x)
        d[d == 0] = 1.0
        px[S] = gamma * d[S]
        nr = np.sqrt(y * y + px * px)
        p = px / nr - 1.0
        q = y / nr - 1.0
        J = np.diag(p) + q[:, None] * A
        nabla = p * phi + A.T @ (q * phi)
        try:
            dx = np.linalg.solve(J, -phi)
            grad = False
        except np.linalg.LinAlgError:  # exactly singular Newton system: gradient direction
            dx = -nabla
            grad = True
        if not grad and (np.max(np.abs(dx)) < np.finfo(float).eps or np.linalg.norm(nabla) < tol_abs):
            break
        if grad or nabla @ dx > -rho * (dx @ dx):
            dx = -nabla
        tau = 1.0
        f0 = err
        gf = beta * (nabla @ dx)
        while True:
            xk = np.maximum(0.0, x + dx * tau)
            yk = A @ xk + b
            pk = fb_phi(yk, xk)
            fk = 0.5 * (pk @ pk)
            if fk <= f0 + tau * gf or tau * tau < gamma:
                break
            tau *= alpha
        x = xk
        it += 1
    return x, {"iterations": it, "err": err}


def friction_lcp(C, D, mu, Sf, n):
    """The LCP [y] = [a][x] + [b] of SlipCalculator3D.jl:90-167 (with the displacements D = C b added)."""
    L = [slice(k * n, (k + 1) * n) for k in range(5)]
    blk = lambda r, c: C[L[r], L[c]]
    dmu = np.diag(mu)
    ID, ZE = np.eye(n), np.zeros((n, n))
    a = np.block([
        [blk(0, 0) - blk(0, 1) @ dmu - blk(0, 2) @ dmu, blk(0, 1), blk(0, 2), ZE, ZE],
        [blk(1, 0) - blk(1, 1) @ dmu - blk(1, 2) @ dmu, blk(1, 1), blk(1, 2), ID, ZE],
        [blk(2, 0) - blk(2, 1) @ dmu - blk(2, 2) @ dmu, blk(2, 1), blk(2, 2), ZE, ID],
        [2 * dmu, -ID, ZE, ZE, ZE],
        [2 * dmu, ZE, -ID, ZE, ZE]])
    bb = np.concatenate([-blk(0, 1) @ Sf - blk(0, 2) @ Sf, -blk(1, 1) @ Sf - blk(1, 2) @ Sf,
                         -blk(2, 1) @ Sf - blk(2, 2) @ Sf, 2 * Sf, 2 * Sf])
    bb[:3 * n] += D
    return a, bb


def lcp_keep(mu, Sf, n):
    """The LCP unknowns left after removing, for every element with mu = 0 and Sf = 0, its two traction parts
    (x2, x3: rows 3 n + i and 4 n + i read y = -x, so both are 0) and its two negative slip parts (x4, x5: they enter
    only that element's own slip rows, so the slip's sign is carried by y2, y3).  The LCP on the rest has the same
    solution for every remaining unknown and no degenerate pair (x = y = 0), on which Fischer-Newton converges only
    linearly; with no such element nothing is removed."""
    mu, Sf = np.broadcast_to(mu, (n,)), np.broadcast_to(Sf, (n,))
    free = (mu == 0) & (Sf == 0)
    keep = np.ones(5 * n, dtype=bool)
    for k in range(1, 5):
        keep[k * n:(k + 1) * n][free] = False
    return keep


def slip_friction(P1, P2, P3, nu, G, lam, mid, nrm, hs_flag, stress6, tractions, mu, Sf):
    """SlipCalculator3D(..., MixedBoundaryConditionsFriction, FixedEls = 0) -> (Dn, Dss, Dds, info)."""
    n = P1.shape[0]
    Tn, Tds, Tss = O.normal_shear_tractions(stress6, nrm)
    Tn, Tss, Tds = Tn + tractions[0], Tss + tractions[1], Tds + tractions[2]
    mu = np.zeros(n) + mu
    Sf = np.zeros(n) + Sf
    fric = Tn * mu
    if np.all(-np.abs(Tss) > fric) and np.all(-np.abs(Tds) > fric):
        z = np.zeros(n)
        return z, z.copy(), z.copy(), {"locked": True}
    A = O.assemble_A(O.traction_influence(nrm, mid, P1, P2, P3, nu, G, lam, hs_flag))
    b = np.concatenate([Tn, Tss, Tds])
    _, hp = CT.area_half_perimeter(P1, P2, P3)
    scl = np.mean(hp) / G
    C = np.linalg.inv(A * scl)
    a, bb = friction_lcp(C, C @ b, mu, Sf, n)
    keep = lcp_keep(mu, Sf, n)
    x = np.zeros(5 * n)
    x[keep], info = fischer_newton(a[np.ix_(keep, keep)], bb[keep])
    y = a @ x + bb
    L = [slice(k * n, (k + 1) * n) for k in range(5)]
    Dn = y[L[0]] * scl
    Dss = (y[L[1]] - x[L[3]]) * scl
    Dds = (y[L[2]] - x[L[4]]) * scl
    info.update(locked=False, x=x, y=y, scale=scl)
    return Dn, Dss, Dds, info
