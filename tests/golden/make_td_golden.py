"""Writes tests/golden/td_golden.json: displacement and strain influences of single triangular dislocations, to 25
significant digits (the inputs are doubles, written exactly), at points near edges, vertices and the triangle's
plane, and elsewhere, in full space and half space.  Run: python tests/golden/make_td_golden.py

Displacement: an mpmath restatement at 60 significant digits of the reference's displacement path for one triangle
and unit slip of each mode (Dn, Dss, Dds):
  * full space, TDdispFS (TD.jl:735-867): local frame (CalculateLocalTriCoords! TD.jl:712-732), TDCS coordinates
    (GlobalToTDECoords TD.jl:870-908), configuration (trimodefinder TD.jl:1010-1079), three angular dislocations
    (TDSetupD TD.jl:1081-1173, AngDisDisp TD.jl:1177-1220) and the Burgers function (TD.jl:820-840);
  * half space: the real triangle, plus the free-surface correction of its three sides (AngSetupDispFSC
    TD.jl:1258-1409, AngDisDispFSC TD.jl:1412-1496, here in the factored form the GPU kernels use), plus the image
    triangle (vertices with z -> -z) in full space.
Strain: the symmetric gradient of that displacement by central differences with a step of 1e-20 at 60 digits
(truncation and rounding both near 1e-40), an independent route to the strain closed forms AngDisStrain
(TD.jl:1753-1828) and AngDisStrainFSC (TD.jl:2084-2471) that the oracle and the GPU evaluate.
Points are kept off the configuration boundaries and off the triangle's edges in its plane (NaN there).
"""
import json
import os

import mpmath as mp

mp.mp.dps = 60
PI = mp.pi
HERE = os.path.dirname(os.path.abspath(__file__))


def vec(*a):
    """A point whose coordinates are doubles (so that the oracle reads exactly the point evaluated here)."""
    return [mp.mpf(float(x)) for x in a]


def sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def unit(a):
    n = mp.sqrt(dot(a, a))
    return [a[0] / n, a[1] / n, a[2] / n]


def frame(P1, P2, P3):
    """Vnorm, Vstrike, Vdip (CalculateLocalTriCoords!), TDCS vertices with origin P2, side units and angles."""
    Vn = unit(cross(sub(P2, P1), sub(P3, P1)))
    Vs = cross([0, 0, 1], Vn)
    if all(v == 0 for v in Vs):
        Vs = [0, Vn[2], 0]
    Vs = unit(Vs)
    Vd = cross(Vn, Vs)
    tdcs = lambda P: [dot(Vn, sub(P, P2)), dot(Vs, sub(P, P2)), dot(Vd, sub(P, P2))]  # noqa: E731
    p1, p2, p3 = tdcs(P1), [mp.mpf(0)] * 3, tdcs(P3)
    e12, e13, e23 = unit(sub(p2, p1)), unit(sub(p3, p1)), unit(sub(p3, p2))
    A = mp.acos(dot(e12, e13))
    B = mp.acos(-dot(e12, e23))
    C = mp.acos(dot(e23, e13))
    return dict(P2=P2, Vn=Vn, Vs=Vs, Vd=Vd, p1=p1, p2=p2, p3=p3, e12=e12, e13=e13, e23=e23, A=A, B=B, C=C,
                tdcs=tdcs)


def trimode(t, x, y, z):
    """trimodefinder(y, z, x, p1[2:3], p2[2:3], p3[2:3]): +1 configuration I, -1 II, 0 on a side in the plane."""
    px, py = y, z
    p1, p2, p3 = t["p1"], t["p2"], t["p3"]
    base = (p2[2] - p3[2]) * (p1[1] - p3[1]) + (p3[1] - p2[1]) * (p1[2] - p3[2])
    a = ((p2[2] - p3[2]) * (px - p3[1]) + (p3[1] - p2[1]) * (py - p3[2])) / base
    b = ((p3[2] - p1[2]) * (px - p3[1]) + (p1[1] - p3[1]) * (py - p3[2])) / base
    c = 1 - a - b
    if (a <= 0 and b > c and c > a) or (b <= 0 and c > a and a > b) or (c <= 0 and a > b and b > c):
        return -1
    if ((a == 0 and b >= 0 and c >= 0) or (a >= 0 and b == 0 and c >= 0) or (a >= 0 and b >= 0 and c == 0)) \
            and x == 0:
        return 0
    return 1


def ang_dis_disp(x, y, z, cosA, sinA, E2):
    """AngDisDisp: {ux,uy,uz, vx,vy,vz, wx,wy,wz} without the Burgers function (x/y/z suffix: Burgers comp.)."""
    eta = y * cosA - z * sinA
    zeta = y * sinA + z * cosA
    r = mp.sqrt(x * x + y * y + z * z)
    zeta = min(zeta, r)
    z = min(z, r)
    rmzeta, rmz = r - zeta, r - z
    ux = x * y / r / rmz - x * eta / r / rmzeta
    vx = eta * sinA / rmzeta - y * eta / r / rmzeta + y * y / r / rmz + E2 * (cosA * mp.log(rmzeta) - mp.log(rmz))
    wx = eta * cosA / rmzeta - y / r - eta * z / r / rmzeta - E2 * sinA * mp.log(rmzeta)
    uy = x * x * cosA / r / rmzeta - x * x / r / rmz - E2 * (cosA * mp.log(rmzeta) - mp.log(rmz))
    vy = y * cosA / r / rmzeta - sinA * cosA / rmzeta - y / r / rmz
    wy = z * cosA / r / rmzeta - cosA * cosA / rmzeta + 1 / r
    uz = E2 * mp.log(rmzeta) - x * x / r / rmzeta
    vz = sinA / rmzeta - y / r / rmzeta
    wz = cosA / rmzeta - z / r / rmzeta
    return [ux, uy, uz, vx, vy, vz, wx, wy, wz]


def td_disp_fs(t, X, nu):
    """TDdispFS with unit slip of each mode: u[m] (EFCS) for m = Dn, Dss, Dds."""
    x, y, z = t["tdcs"](X)
    mode = trimode(t, x, y, z)
    if mode == 0:
        raise ValueError("point on a triangle side in its plane")
    s = 1 if mode == 1 else -1
    ads = [(t["A"], t["p1"], [-s * v for v in t["e13"]]), (t["B"], t["p2"], [s * v for v in t["e12"]]),
           (t["C"], t["p3"], [s * v for v in t["e23"]])]
    E1, E2 = 1 - nu, 1 - 2 * nu
    loc = [[mp.mpf(0)] * 3 for _ in range(3)]
    for alpha, tv, sv in ads:
        CosT, SinT = sv[2], sv[1]
        Ang = -PI + alpha
        cosA, sinA = mp.cos(Ang), mp.sin(Ang)
        sinADE1 = sinA / 8 / PI / E1
        Y = CosT * (y - tv[1]) - SinT * (z - tv[2])
        Z = SinT * (y - tv[1]) + CosT * (z - tv[2])
        ux, uy, uz, vx, vy, vz, wx, wy, wz = ang_dis_disp(x, Y, Z, cosA, sinA, E2)
        # unit slips in ADCS: Dss -> (CosT, SinT), Dds -> (-SinT, CosT) (RotateObject2D TD.jl:1099-1100)
        for m, (bss, bds, bn) in enumerate(((0, 0, 1), (CosT, SinT, 0), (-SinT, CosT, 0))):
            u0 = bn / 8 / PI / E1 * ux + bss / 8 / PI / E1 * uy + bds * sinADE1 * uz
            a = bn / 8 / PI / E1 * vx + bss * x / 8 / PI / E1 * vy + bds * x * sinADE1 * vz
            b = bn / 8 / PI / E1 * wx + bss * x / 8 / PI / E1 * wy + bds * x * sinADE1 * wz
            # back to TDCS: rotate by -T
            loc[m][0] += u0
            loc[m][1] += CosT * a + SinT * b
            loc[m][2] += -SinT * a + CosT * b
    # Burgers' function (solid angle) added to the slip's own component
    a_ = [-x, t["p1"][1] - y, t["p1"][2] - z]
    b_ = [-x, -y, -z]
    c_ = [-x, t["p3"][1] - y, t["p3"][2] - z]
    na, nb, nc = (mp.sqrt(dot(v, v)) for v in (a_, b_, c_))
    one = dot(a_, cross(b_, c_))
    two = na * nb * nc + dot(a_, b_) * nc + dot(a_, c_) * nb + dot(b_, c_) * na
    Fi = -2 * mp.atan2(one, two) / 4 / PI
    for m in range(3):
        loc[m][m] += Fi
    Vn, Vs, Vd = t["Vn"], t["Vs"], t["Vd"]
    return [[loc[m][0] * Vn[k] + loc[m][1] * Vs[k] + loc[m][2] * Vd[k] for k in range(3)] for m in range(3)]


def fsc_disp_ad(y1, y2, y3, cosB, sinB, cotB, cotB2, nu, a):
    """AngDisDispFSC: {ux,uy,uz, vx,vy,vz, wx,wy,wz} (x/y/z suffix: Burgers component), without 1/(4 pi (1-nu)),
    in the factored form of the GPU kernels (reciprocals of rb, rb + y3b, rb + z3b)."""
    icosB = 1 / cosB
    c1, c2, nu2, cot2 = 1 - 2 * nu, 1 - nu, 2 * nu, cotB * cotB
    y3b = y3 + 2 * a
    z3b = y3b * cosB - y1 * sinB
    rb = mp.sqrt(y1 * y1 + y2 * y2 + y3b * y3b)
    irb = 1 / rb
    W6, W7 = rb + y3b, rb + z3b
    iW6, iW7 = 1 / W6, 1 / W7
    lW6, lW7 = mp.log(W6), mp.log(W7)
    Fib = 2 * mp.atan(-y2 / (y1 - W6 * cotB2))
    z1b = y1 * cosB + y3b * sinB
    irb2 = irb * irb
    irb3 = irb2 * irb
    aR = a * irb
    y3ma = y3b - a
    rcy = rb * cosB + y3b
    rsy = rb * sinB - y1
    cpa = cosB + aR
    y22, y11 = y2 * y2, y1 * y1
    q6, q7 = y3ma * iW6, y3ma * iW7
    p1 = 1 + aR * icosB
    ux = (-2 * c2 * c1 * cot2 * Fib + c1 * y2 * iW6 * ((c1 - aR) * cotB - y1 * iW6 * (nu + aR)) +
          c1 * cosB * cotB * y2 * iW7 * cpa + a * cotB * y2 * y3ma * irb3 +
          y2 * y3ma * irb * iW6 * (-c1 * cotB + y1 * iW6 * (nu2 + aR) + a * y1 * irb2) +
          y2 * y3ma * irb * iW7 * (cosB * iW7 * (rcy * (c1 * cosB - aR) * cotB + 2 * c2 * cosB * rsy) -
                                   a * cosB * cotB * y3b * irb2))
    uy = (c1 * ((2 * c2 * cot2 + nu) * lW6 - (2 * c2 * cot2 + 1) * cosB * lW7) +
          c1 * iW6 * (-c1 * cotB * y1 + nu * y3b - a + a * cotB * y1 * irb + y11 * iW6 * (nu + aR)) -
          c1 * cotB * iW7 * (cosB * z1b - a * icosB * rsy * irb) - a * cotB * y1 * y3ma * irb3 +
          q6 * (nu2 + irb * (c1 * cotB * y1 + a) - y11 * irb * iW6 * (nu2 + aR) - a * y11 * irb3) +
          cotB * q7 * (-cosB * sinB + a * icosB * y1 * y3b * irb3 + rsy * irb * (2 * c2 * cosB - rcy * iW7 * p1)))
    uz = (c1 * (y2 * iW6 * (1 + aR) - cosB * y2 * iW7 * cpa) - y2 * y3ma * irb * (a * irb2 + iW6) +
          cosB * y2 * y3ma * irb * iW7 * (rcy * iW7 * cpa + a * y3b * irb2))
    vx = (c1 * ((2 * c2 * cot2 - nu) * lW6 - (2 * c2 * cot2 + c1) * cosB * lW7) -
          c1 * iW6 * (y1 * cotB * (c1 - aR) + nu * y3b - a + y22 * iW6 * (nu + aR)) -
          c1 * cotB * z1b * iW7 * cpa - a * cotB * y1 * y3ma * irb3 +
          q6 * (-nu2 + irb * (c1 * cotB * y1 - a) + y22 * irb * iW6 * (nu2 + aR) + a * y22 * irb3) +
          q7 * (cosB * cosB - irb * (c1 * cotB * z1b + a * cosB) + a * cotB * y3b * z1b * irb3 -
                irb * iW7 * (cosB * cosB * y22 - a * cotB * z1b * irb * rcy)))
    vy = (2 * c2 * c1 * cot2 * Fib + c1 * y2 * iW6 * (-(c1 - aR) * cotB + y1 * iW6 * (nu + aR)) -
          c1 * cotB * y2 * iW7 * p1 - a * cotB * y2 * y3ma * irb3 +
          y2 * y3ma * irb * iW6 * (c1 * cotB - nu2 * y1 * iW6 - a * y1 * irb * (irb + iW6)) +
          cotB * y2 * y3ma * irb * iW7 * (-2 * c2 * cosB + rcy * iW7 * p1 + a * icosB * y3b * irb2))
    vz = (c1 * (-sinB * lW7 - y1 * iW6 * (1 + aR) + z1b * iW7 * cpa) + y1 * y3ma * irb * (a * irb2 + iW6) -
          q7 * (sinB * (cosB - aR) + z1b * irb * (1 + a * y3b * irb2) -
                irb * iW7 * (cosB * sinB * y22 - a * z1b * irb * rcy)))
    wx = (2 * c2 * (c1 * cotB * Fib + y2 * iW6 * (nu2 + aR) - cosB * y2 * iW7 * cpa) +
          y2 * y3ma * irb * (nu2 * iW6 + a * irb2) +
          cosB * y2 * y3ma * irb * iW7 * (c1 - rcy * iW7 * cpa - a * y3b * irb2))
    wy = (-2 * c2 * c1 * cotB * (lW6 - cosB * lW7) - 2 * c2 * y1 * iW6 * (nu2 + aR) + 2 * c2 * z1b * iW7 * cpa +
          y3ma * irb * (c1 * cotB - nu2 * y1 * iW6 - a * y1 * irb2) -
          q7 * (cosB * sinB + cotB * rcy * irb * (2 * c2 * cosB - rcy * iW7) +
                aR * (sinB - y3b * z1b * irb2 - z1b * rcy * irb * iW7)))
    wz = 2 * c2 * Fib + 2 * c2 * sinB * y2 * iW7 * cpa + sinB * y2 * y3ma * irb * iW7 * (1 + rcy * iW7 * cpa +
                                                                                      a * y3b * irb2)
    return [ux, uy, uz, vx, vy, vz, wx, wy, wz]


def fsc_side_disp(PA, PB, t, X, nu):
    """AngSetupDispFSC of one side with unit slip of each mode: u[m] (EFCS), or None for a skipped side."""
    sv = sub(PB, PA)
    beta = mp.acos(-sv[2] / mp.sqrt(dot(sv, sv)))
    if abs(beta) < mp.mpf(2) ** -52 or abs(PI - beta) < mp.mpf(2) ** -52 or abs(1 / mp.tan(beta)) > 5e3:
        return None
    ey1 = unit([sv[0], sv[1], 0])
    ey3 = [0, 0, -1]
    ey2 = cross(ey3, ey1)
    E = [ey1, ey2, ey3]
    adcs = lambda v: [dot(E[0], v), dot(E[1], v), dot(E[2], v)]  # noqa: E731  (E is symmetric)
    x, y, z = adcs(sub(X, PA))
    yB = adcs(sv)
    B = beta if beta * x < 0 else -PI + beta
    cB, sB, ctB, ctB2 = mp.cos(B), mp.sin(B), 1 / mp.tan(B), 1 / mp.tan(B / 2)
    uA = fsc_disp_ad(x, y, z, cB, sB, ctB, ctB2, nu, -PA[2])
    uB = fsc_disp_ad(x - yB[0], y - yB[1], z - yB[2], cB, sB, ctB, ctB2, nu, -PB[2])
    out = []
    for V in (t["Vn"], t["Vs"], t["Vd"]):
        b = adcs(V)
        w = [sum(b[k] / (4 * PI * (1 - nu)) * (uB[3 * c + k] - uA[3 * c + k]) for k in range(3)) for c in range(3)]
        out.append(adcs(w))
    return out


def seg_dist(X, A, B):
    AB, AX = sub(B, A), sub(X, A)
    t = min(max(dot(AX, AB) / dot(AB, AB), 0), 1)
    D = [AX[k] - t * AB[k] for k in range(3)]
    return mp.sqrt(dot(D, D))


def conditioning(P, X):
    """kappa = max(1, (L/d)^2, (r/L)^3): d the distance to the nearest triangle side, r to the centroid, L the
    longest side.  Near a side the angular-dislocation terms cancel like (r - zeta) ~ d^2 / r; far away the three
    dislocations' fields cancel to the triangle's decaying one.  A double-precision evaluation of the formulas is
    accurate to a small multiple of eps x kappa."""
    L = max(mp.sqrt(dot(sub(P[i], P[j]), sub(P[i], P[j]))) for i, j in ((0, 1), (1, 2), (2, 0)))
    d = min(seg_dist(X, P[i], P[j]) for i, j in ((0, 1), (1, 2), (2, 0)))
    cen = [(P[0][k] + P[1][k] + P[2][k]) / 3 for k in range(3)]
    r = mp.sqrt(dot(sub(X, cen), sub(X, cen)))
    return max(mp.mpf(1), (L / d) ** 2, (r / L) ** 3)


def td_disp(P, X, nu, hs):
    t = frame(*P)
    u = td_disp_fs(t, X, nu)
    if hs:
        for PA, PB in ((P[0], P[1]), (P[1], P[2]), (P[2], P[0])):
            f = fsc_side_disp(PA, PB, t, X, nu)
            if f is not None:
                u = [[u[m][k] + f[m][k] for k in range(3)] for m in range(3)]
        img = frame(*[[Q[0], Q[1], -Q[2]] for Q in P])
        ui = td_disp_fs(img, X, nu)
        u = [[u[m][k] + ui[m][k] for k in range(3)] for m in range(3)]
    return u


def td_strain(P, X, nu, hs, h=mp.mpf("1e-20")):
    g = [[[mp.mpf(0)] * 3 for _ in range(3)] for _ in range(3)]   # g[m][i][j] = d u_i / d x_j
    for j in range(3):
        Xp = list(X)
        Xm = list(X)
        Xp[j] += h
        Xm[j] -= h
        up, um = td_disp(P, Xp, nu, hs), td_disp(P, Xm, nu, hs)
        for m in range(3):
            for i in range(3):
                g[m][i][j] = (up[m][i] - um[m][i]) / (2 * h)
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    return [[(g[m][i][j] + g[m][j][i]) / 2 for i, j in pairs] for m in range(3)]


def points(P, hs):
    """Labelled observation points around triangle P."""
    t = frame(*P)
    Vn = t["Vn"]
    P1, P2, P3 = P
    cen = [(P1[k] + P2[k] + P3[k]) / 3 for k in range(3)]
    mid12 = [(P1[k] + P2[k]) / 2 for k in range(3)]
    along = unit(sub(P2, P1))
    inward = unit(cross(Vn, along))                     # in the plane, perpendicular to side 1-2
    if dot(inward, sub(cen, mid12)) < 0:
        inward = [-v for v in inward]
    L = mp.sqrt(dot(sub(P2, P1), sub(P2, P1)))
    out = []

    def add(label, X):
        out.append((label, [mp.mpf(float(v)) for v in X]))     # evaluated exactly at the double the test reads

    for d in ("1e-2", "1e-3", "1e-4", "1e-5", "1e-6"):
        add("near edge, off plane %s L" % d, [mid12[k] + mp.mpf(d) * L * Vn[k] for k in range(3)])
    for d in ("1e-3", "1e-6"):
        add("near edge, in plane outside %s L" % d, [mid12[k] - mp.mpf(d) * L * inward[k] for k in range(3)])
        add("near edge, in plane inside %s L" % d,
            [mid12[k] + mp.mpf(d) * L * inward[k] + mp.mpf("1e-12") * Vn[k] for k in range(3)])
    for h in ("1e-12", "1e-8", "1e-4"):
        add("centroid, %s off plane" % h, [cen[k] + mp.mpf(h) * Vn[k] for k in range(3)])
    for s_ in ("0.5", "2", "6"):
        Xo = [mid12[k] - mp.mpf(s_) * L * inward[k] for k in range(3)]
        add("in plane outside at %s L" % s_, [Xo[k] + mp.mpf("1e-9") * Vn[k] for k in range(3)])
    add("near vertex 1e-4", [P2[k] + mp.mpf("1e-4") * L * (Vn[k] + inward[k]) for k in range(3)])
    add("generic 1", [cen[k] + mp.mpf("0.7") * L * (Vn[k] - along[k] / 2) for k in range(3)])
    add("generic 2", [cen[k] + mp.mpf("1.9") * L * (inward[k] + Vn[k] / 3) for k in range(3)])
    add("far 30 L", [cen[k] + mp.mpf(30) * L * unit([1, 2, -3])[k] for k in range(3)])
    if hs:
        add("just below the free surface", [cen[0] + mp.mpf("0.3"), cen[1] - mp.mpf("0.2"), mp.mpf("-1e-3")])
        add("on the free surface", [cen[0] - mp.mpf("0.4"), cen[1] + mp.mpf("0.5"), mp.mpf(0)])
    if hs:
        out = [(lab, X) for lab, X in out if X[2] <= 0]
    return out


def main():
    nu = mp.mpf("0.25")
    tris = {
        "full_space": ([vec("-0.7", "-0.5", "0.2"), vec("0.9", "-0.3", "-0.1"), vec("0.1", "0.8", "0.4")], 0),
        "half_space": ([vec("-0.6", "-0.4", "-1.2"), vec("0.8", "-0.2", "-0.7"), vec("0.0", "0.7", "-1.6")], 1),
        # a vertex 0.05 below the free surface: large free-surface corrections
        "half_space_shallow": ([vec("-0.5", "-0.3", "-0.05"), vec("0.7", "0.1", "-0.9"), vec("-0.2", "0.8", "-0.6")],
                               1),
        # side P1-P2 vertical: its free-surface correction is skipped (beta = 0, AngSetupDispFSC TD.jl:1265)
        "half_space_vertical_side": ([vec("0.1", "0.2", "-1.0"), vec("0.1", "0.2", "-2.0"), vec("1.0", "0.6", "-1.4")],
                                     1),
    }
    cases = []
    for name, (P, hs) in tris.items():
        for label, X in points(P, hs):
            u = td_disp(P, X, nu, hs)
            e = td_strain(P, X, nu, hs)
            cases.append({"triangle": name, "hs": hs, "label": label, "kappa": float(conditioning(P, X)),
                          "P1": [repr(float(v)) for v in P[0]], "P2": [repr(float(v)) for v in P[1]],
                          "P3": [repr(float(v)) for v in P[2]], "X": [repr(float(v)) for v in X],
                          "disp": [[mp.nstr(v, 25) for v in row] for row in u],
                          "strain": [[mp.nstr(v, 25) for v in row] for row in e]})
            print(name, label, flush=True)
    with open(os.path.join(HERE, "td_golden.json"), "w") as f:
        json.dump({"nu": "0.25", "digits": 60, "modes": ["Dn", "Dss", "Dds"],
                   "strain_components": ["xx", "yy", "zz", "xy", "xz", "yz"], "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
