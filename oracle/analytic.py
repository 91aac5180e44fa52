"""Closed-form elastic solutions used to pin the oracle (test infrastructure only).

okada85_surface: Okada (1985, BSSA 75:1135) surface displacements and horizontal strains of a finite
rectangular dislocation in a half-space, restated from the paper's equations (25)-(41) with Chinnery's
notation (24); output conventions follow Okada1985RectangularDislocation.jl:1-184 (the reference converts
Okada's tilts/strains into a TD-style strain tensor at z = 0: exx, eyy, ezz, exy, exz, eyz).

eshelby_penny_*: slip of a penny-shaped crack under uniform remote shear / normal traction (Eshelby 1957),
as checked by test/test_TDvsEshelbyPennyCrack.jl.
"""
import numpy as np

_EPS = np.finfo(float).eps


def _chinnery(f, x, p, L, W, q, dip, nu):
    return f(x, p, q, dip, nu) - f(x, p - W, q, dip, nu) - f(x - L, p, q, dip, nu) + f(x - L, p - W, q, dip, nu)


def _geom(xi, eta, q, dip):
    R = np.sqrt(xi ** 2 + eta ** 2 + q ** 2)
    yb = eta * np.cos(dip) + q * np.sin(dip)
    db = eta * np.sin(dip) - q * np.cos(dip)
    return R, yb, db


def _atan_term(xi, eta, q, R):
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.arctan(xi * eta / (q * R))
    return np.where(q != 0, t, 0.0)


# I-terms of the displacement field, eq. (28)-(29)
def _I5(xi, eta, q, dip, nu, R, db):
    X = np.sqrt(xi ** 2 + q ** 2)
    if np.cos(dip) > _EPS:
        with np.errstate(divide="ignore", invalid="ignore"):
            v = (1 - 2 * nu) * 2 / np.cos(dip) * np.arctan(
                (eta * (X + q * np.cos(dip)) + X * (R + X) * np.sin(dip)) / (xi * (R + X) * np.cos(dip)))
        return np.where(xi == 0, 0.0, v)
    return -(1 - 2 * nu) * xi * np.sin(dip) / (R + db)


def _I4(db, eta, q, dip, nu, R):
    if np.cos(dip) > _EPS:
        return (1 - 2 * nu) / np.cos(dip) * (np.log(R + db) - np.sin(dip) * np.log(R + eta))
    return -(1 - 2 * nu) * q / (R + db)


def _I3(eta, q, dip, nu, R):
    yb = eta * np.cos(dip) + q * np.sin(dip)
    db = eta * np.sin(dip) - q * np.cos(dip)
    if np.cos(dip) > _EPS:
        return (1 - 2 * nu) * (yb / (np.cos(dip) * (R + db)) - np.log(R + eta)) + \
            np.sin(dip) / np.cos(dip) * _I4(db, eta, q, dip, nu, R)
    return (1 - 2 * nu) / 2 * (eta / (R + db) + yb * q / (R + db) ** 2 - np.log(R + eta))


def _I2(eta, q, dip, nu, R):
    return (1 - 2 * nu) * (-np.log(R + eta)) - _I3(eta, q, dip, nu, R)


def _I1(xi, eta, q, dip, nu, R):
    db = eta * np.sin(dip) - q * np.cos(dip)
    if np.cos(dip) > _EPS:
        return (1 - 2 * nu) * (-xi / (np.cos(dip) * (R + db))) - np.sin(dip) / np.cos(dip) * _I5(xi, eta, q, dip, nu, R, db)
    return -(1 - 2 * nu) / 2 * xi * q / (R + db) ** 2


def _A(x, R):
    return (2 * R + x) / (R ** 3 * (R + x) ** 2)


# K-terms (tilts), eq. (40)-(41); J-terms (strains), eq. (34)-(35)
def _K3(xi, eta, q, dip, nu, R):
    _, yb, db = _geom(xi, eta, q, dip)
    if np.cos(dip) > _EPS:
        return (1 - 2 * nu) / np.cos(dip) * (q / (R * (R + eta)) - yb / (R * (R + db)))
    return (1 - 2 * nu) * np.sin(dip) / (R + db) * (xi ** 2 / (R * (R + db)) - 1)


def _K1(xi, eta, q, dip, nu, R):
    _, _, db = _geom(xi, eta, q, dip)
    if np.cos(dip) > _EPS:
        return (1 - 2 * nu) * xi / np.cos(dip) * (1 / (R * (R + db)) - np.sin(dip) / (R * (R + eta)))
    return (1 - 2 * nu) * xi * q / (R * (R + db) ** 2)


def _K2(xi, eta, q, dip, nu, R):
    return (1 - 2 * nu) * (-np.sin(dip) / R + q * np.cos(dip) / (R * (R + eta))) - _K3(xi, eta, q, dip, nu, R)


def _J1(xi, eta, q, dip, nu, R):
    _, _, db = _geom(xi, eta, q, dip)
    if np.cos(dip) > _EPS:
        return (1 - 2 * nu) / np.cos(dip) * (xi ** 2 / (R * (R + db) ** 2) - 1 / (R + db)) - \
            np.sin(dip) / np.cos(dip) * _K3(xi, eta, q, dip, nu, R)
    return (1 - 2 * nu) / 2 * q / (R + db) ** 2 * (2 * xi ** 2 / (R * (R + db)) - 1)


def _J2(xi, eta, q, dip, nu, R):
    _, yb, db = _geom(xi, eta, q, dip)
    if np.cos(dip) > _EPS:
        return (1 - 2 * nu) / np.cos(dip) * xi * yb / (R * (R + db) ** 2) - \
            np.sin(dip) / np.cos(dip) * _K1(xi, eta, q, dip, nu, R)
    return (1 - 2 * nu) / 2 * xi * np.sin(dip) / (R + db) ** 2 * (2 * q ** 2 / (R * (R + db)) - 1)


def _J3(xi, eta, q, dip, nu, R):
    return (1 - 2 * nu) * -xi / (R * (R + eta)) - _J2(xi, eta, q, dip, nu, R)


def _J4(xi, eta, q, dip, nu, R):
    return (1 - 2 * nu) * (-np.cos(dip) / R - q * np.sin(dip) / (R * (R + eta))) - _J1(xi, eta, q, dip, nu, R)


def _fn(body):
    def f(xi, eta, q, dip, nu):
        R, yb, db = _geom(xi, eta, q, dip)
        return body(xi, eta, q, dip, nu, R, yb, db)
    return f


s, c = np.sin, np.cos
# displacement, eq. (25)-(27)
ux_ss = _fn(lambda xi, eta, q, d, nu, R, yb, db: xi * q / (R * (R + eta)) + _I1(xi, eta, q, d, nu, R) * s(d) + _atan_term(xi, eta, q, R))
uy_ss = _fn(lambda xi, eta, q, d, nu, R, yb, db: yb * q / (R * (R + eta)) + q * c(d) / (R + eta) + _I2(eta, q, d, nu, R) * s(d))
uz_ss = _fn(lambda xi, eta, q, d, nu, R, yb, db: db * q / (R * (R + eta)) + q * s(d) / (R + eta) + _I4(db, eta, q, d, nu, R) * s(d))
ux_ds = _fn(lambda xi, eta, q, d, nu, R, yb, db: q / R - _I3(eta, q, d, nu, R) * s(d) * c(d))
uy_ds = _fn(lambda xi, eta, q, d, nu, R, yb, db: yb * q / (R * (R + xi)) - _I1(xi, eta, q, d, nu, R) * s(d) * c(d) + c(d) * _atan_term(xi, eta, q, R))
uz_ds = _fn(lambda xi, eta, q, d, nu, R, yb, db: db * q / (R * (R + xi)) - _I5(xi, eta, q, d, nu, R, db) * s(d) * c(d) + s(d) * _atan_term(xi, eta, q, R))
ux_tf = _fn(lambda xi, eta, q, d, nu, R, yb, db: q ** 2 / (R * (R + eta)) - _I3(eta, q, d, nu, R) * s(d) ** 2)
uy_tf = _fn(lambda xi, eta, q, d, nu, R, yb, db: -db * q / (R * (R + xi)) - s(d) * xi * q / (R * (R + eta)) - _I1(xi, eta, q, d, nu, R) * s(d) ** 2 + s(d) * _atan_term(xi, eta, q, R))
uz_tf = _fn(lambda xi, eta, q, d, nu, R, yb, db: yb * q / (R * (R + xi)) + c(d) * xi * q / (R * (R + eta)) - _I5(xi, eta, q, d, nu, R, db) * s(d) ** 2 - c(d) * _atan_term(xi, eta, q, R))
# tilt, eq. (37)-(39)
uzx_ss = _fn(lambda xi, eta, q, d, nu, R, yb, db: -xi * q ** 2 * _A(eta, R) * c(d) + (xi * q / R ** 3 - _K1(xi, eta, q, d, nu, R)) * s(d))
uzy_ss = _fn(lambda xi, eta, q, d, nu, R, yb, db: (db * q / R ** 3) * c(d) + (xi ** 2 * q * _A(eta, R) * c(d) - s(d) / R + yb * q / R ** 3 - _K2(xi, eta, q, d, nu, R)) * s(d))
uzx_ds = _fn(lambda xi, eta, q, d, nu, R, yb, db: db * q / R ** 3 + q * s(d) / (R * (R + eta)) + _K3(xi, eta, q, d, nu, R) * s(d) * c(d))
uzy_ds = _fn(lambda xi, eta, q, d, nu, R, yb, db: yb * db * q * _A(xi, R) - (2 * db / (R * (R + xi)) + xi * s(d) / (R * (R + eta))) * s(d) + _K1(xi, eta, q, d, nu, R) * s(d) * c(d))
uzx_tf = _fn(lambda xi, eta, q, d, nu, R, yb, db: q ** 2 / R ** 3 * s(d) - q ** 3 * _A(eta, R) * c(d) + _K3(xi, eta, q, d, nu, R) * s(d) ** 2)
uzy_tf = _fn(lambda xi, eta, q, d, nu, R, yb, db: (yb * s(d) + db * c(d)) * q ** 2 * _A(xi, R) + xi * q ** 2 * _A(eta, R) * s(d) * c(d) - (2 * q / (R * (R + xi)) - _K1(xi, eta, q, d, nu, R)) * s(d) ** 2)
# strain, eq. (31)-(33)
uxx_ss = _fn(lambda xi, eta, q, d, nu, R, yb, db: xi ** 2 * q * _A(eta, R) - _J1(xi, eta, q, d, nu, R) * s(d))
uxy_ss = _fn(lambda xi, eta, q, d, nu, R, yb, db: xi ** 3 * db / (R ** 3 * (eta ** 2 + q ** 2)) - (xi ** 3 * _A(eta, R) + _J2(xi, eta, q, d, nu, R)) * s(d))
uyx_ss = _fn(lambda xi, eta, q, d, nu, R, yb, db: xi * q / R ** 3 * c(d) + (xi * q ** 2 * _A(eta, R) - _J2(xi, eta, q, d, nu, R)) * s(d))
uyy_ss = _fn(lambda xi, eta, q, d, nu, R, yb, db: yb * q / R ** 3 * c(d) + (q ** 3 * _A(eta, R) * s(d) - 2 * q * s(d) / (R * (R + eta)) - (xi ** 2 + eta ** 2) / R ** 3 * c(d) - _J4(xi, eta, q, d, nu, R)) * s(d))
uxx_ds = _fn(lambda xi, eta, q, d, nu, R, yb, db: xi * q / R ** 3 + _J3(xi, eta, q, d, nu, R) * s(d) * c(d))
uxy_ds = _fn(lambda xi, eta, q, d, nu, R, yb, db: yb * q / R ** 3 - s(d) / R + _J1(xi, eta, q, d, nu, R) * s(d) * c(d))
uyx_ds = _fn(lambda xi, eta, q, d, nu, R, yb, db: yb * q / R ** 3 + q * c(d) / (R * (R + eta)) + _J1(xi, eta, q, d, nu, R) * s(d) * c(d))
uyy_ds = _fn(lambda xi, eta, q, d, nu, R, yb, db: yb ** 2 * q * _A(xi, R) - (2 * yb / (R * (R + xi)) + xi * c(d) / (R * (R + eta))) * s(d) + _J2(xi, eta, q, d, nu, R) * s(d) * c(d))
uxx_tf = _fn(lambda xi, eta, q, d, nu, R, yb, db: xi * q ** 2 * _A(eta, R) + _J3(xi, eta, q, d, nu, R) * s(d) ** 2)
uxy_tf = _fn(lambda xi, eta, q, d, nu, R, yb, db: -db * q / R ** 3 - xi ** 2 * q * _A(eta, R) * s(d) + _J1(xi, eta, q, d, nu, R) * s(d) ** 2)
uyx_tf = _fn(lambda xi, eta, q, d, nu, R, yb, db: q ** 2 / R ** 3 * c(d) + q ** 3 * _A(eta, R) * s(d) + _J1(xi, eta, q, d, nu, R) * s(d) ** 2)
uyy_tf = _fn(lambda xi, eta, q, d, nu, R, yb, db: (yb * c(d) - db * s(d)) * q ** 2 * _A(xi, R) - q * np.sin(2 * d) / (R * (R + xi)) - (xi * q ** 2 * _A(eta, R) - _J2(xi, eta, q, d, nu, R)) * s(d) ** 2)


def okada85_surface(E, N, depth, strike_deg, dip_deg, L, W, rake_deg, slip, U3, nu):
    """Surface (z = 0) field of a rectangular dislocation whose centroid is at (0, 0, -depth).

    Returns (ue, un, uz, exx, eyy, ezz, exy, exz, eyz) as numpy arrays shaped like E.
    """
    E = np.asarray(E, dtype=float)
    N = np.asarray(N, dtype=float)
    strike = np.deg2rad(strike_deg)
    dip = np.deg2rad(dip_deg)
    rake = np.deg2rad(rake_deg)
    U1 = np.cos(rake) * slip
    U2 = np.sin(rake) * slip
    d = depth + np.sin(dip) * W / 2
    ec = E + np.cos(strike) * np.cos(dip) * W / 2
    nc = N - np.sin(strike) * np.cos(dip) * W / 2
    x = np.cos(strike) * nc + np.sin(strike) * ec + L / 2
    y = np.sin(strike) * nc - np.cos(strike) * ec + np.cos(dip) * W
    p = y * np.cos(dip) + d * np.sin(dip)
    q = y * np.sin(dip) - d * np.cos(dip)

    def comb(fs, fd, ft):
        return (-U1 / (2 * np.pi) * _chinnery(fs, x, p, L, W, q, dip, nu)
                - U2 / (2 * np.pi) * _chinnery(fd, x, p, L, W, q, dip, nu)
                + U3 / (2 * np.pi) * _chinnery(ft, x, p, L, W, q, dip, nu))

    ux = comb(ux_ss, ux_ds, ux_tf)
    uy = comb(uy_ss, uy_ds, uy_tf)
    uz = comb(uz_ss, uz_ds, uz_tf)
    uxx = comb(uxx_ss, uxx_ds, uxx_tf)
    uxy = comb(uxy_ss, uxy_ds, uxy_tf)
    uyx = comb(uyx_ss, uyx_ds, uyx_tf)
    uyy = comb(uyy_ss, uyy_ds, uyy_tf)
    uzx = comb(uzx_ss, uzx_ds, uzx_tf)
    uzy = comb(uzy_ss, uzy_ds, uzy_tf)

    ss_, cs_ = np.sin(strike), np.cos(strike)
    ue = ss_ * ux - cs_ * uy
    un = cs_ * ux + ss_ * uy
    unn = cs_ ** 2 * uxx + np.sin(2 * strike) * (uxy + uyx) / 2 + ss_ ** 2 * uyy
    une = np.sin(2 * strike) * (uxx - uyy) / 2 + ss_ ** 2 * uyx - cs_ ** 2 * uxy
    uen = np.sin(2 * strike) * (uxx - uyy) / 2 - cs_ ** 2 * uyx + ss_ ** 2 * uxy
    uee = ss_ ** 2 * uxx - np.sin(2 * strike) * (uyx + uxy) / 2 + cs_ ** 2 * uyy
    uze = -ss_ * uzx + cs_ * uzy
    uzn = -cs_ * uzx - ss_ * uzy
    unz = -uzn
    uez = -uze
    uzz = -(uee + unn) * nu / (1 - nu)
    # conversion to the TD strain convention (Okada1985RectangularDislocation.jl:156-168)
    exy = -(uen + une) / 2
    eyz_ = -(uzn + unz) / 2
    exz_ = -(uze + uez) / 2
    return ue, un, uz, -uee, -unn, -uzz, exy, -exz_, -eyz_


def okada_test_fault(strike=60.0, dip=45.0, length=5.0, width=5.0, tip_depth=1.0):
    """The 2-triangle rectangle of test/test_TDvsOkada.jl:9-48 (vertices P1, P2, P3 as 2 x 3 arrays).

    Direction cosines follow CreateDirectionCosinesFromStrikeAndDip.jl (via CalculateSSandDSDirs and a
    rotation of -strike about z), and the flat rectangle is rotated with RotateObject3DNewCoords.
    """
    cd = np.cos(np.deg2rad(dip))
    # CalculateSSandDSDirs(cosd(dip), 0, cosd(dip)): strike = [-0, 1, 0], dip = [-DSCosAx, -DSCosAy, DSCosAz]
    nz = cd
    ds_cos_az = np.cos(np.arccos(nz) - np.pi / 2)
    Lh = np.sin(np.arccos(ds_cos_az))
    a = np.arctan2(0.0, cd)
    ssc = np.array([-0.0, 1.0, 0.0])
    dsc = np.array([-np.cos(a) * Lh, -np.sin(a) * Lh, ds_cos_az])
    ang = np.deg2rad(-strike)
    R = np.array([[np.cos(ang), -np.sin(ang), 0.0], [np.sin(ang), np.cos(ang), 0.0], [0.0, 0.0, 1.0]])
    ssc = R @ ssc
    dsc = R @ dsc
    fnv = np.cross(ssc, dsc)
    X = np.array([-1, 1, -1, -1, 1, 1], dtype=float) / 2 * width
    Y = np.array([1, -1, -1, -1, 1, 1], dtype=float) / 2 * length
    Z = np.zeros(6)
    x = dsc[0] * X + ssc[0] * Y + fnv[0] * Z
    y = dsc[1] * X + ssc[1] * Y + fnv[1] * Z
    z = dsc[2] * X + ssc[2] * Y + fnv[2] * Z
    z = z - z.max() - tip_depth
    pts = np.stack([x, y, z], axis=1)
    P1 = pts[[0, 1]]
    P2 = pts[[4, 3]]
    P3 = pts[[2, 5]]
    mid_depth = np.sin(np.deg2rad(dip)) * width / 2 + tip_depth
    return P1, P2, P3, mid_depth


def eshelby_penny_shear_slip(rho, a, tau, nu, mu):
    """Shear slip of a penny crack of radius a under remote shear traction tau (Eshelby 1957)."""
    return 4 * (1 - nu) * a * tau / (np.pi * mu * (2 - nu)) * np.sqrt(np.clip(1 - (rho / a) ** 2, 0, None)) * 2


def eshelby_penny_opening(rho, a, p, nu, mu):
    """Opening of a penny crack of radius a under internal pressure p (Sneddon / Eshelby)."""
    return 4 * (1 - nu) * p * a / (np.pi * mu) * np.sqrt(np.clip(1 - (rho / a) ** 2, 0, None))


def eshelby1957_penny_profile(G, nu, Tn, Ts, r, a):
    """Separation of the walls of a penny crack of radius a under uniform normal (Tn) and shear (Ts) traction:
    one-wall opening 2 a (1 - nu) Tn / (pi G) sqrt(1 - r^2/a^2) and shearing 4 (1 - nu) a Ts / (pi (2 - nu) G)
    sqrt(1 - r^2/a^2) (Segall 2010, eqs 4.73-4.74), combined as a vector magnitude and doubled
    (Eshelby1957_PennyCrackSlipProfile.jl)."""
    s = np.sqrt(np.clip(1.0 - (np.asarray(r) ** 2) / a ** 2, 0.0, None))
    ds = (4 * (1 - nu) * a * Ts) / (np.pi * (2 - nu) * G) * s
    dn = (2 * a * (1 - nu) * Tn) / (np.pi * G) * s
    return 2.0 * np.sqrt(ds ** 2 + dn ** 2)


def mogi1958(D, X, Y, radius, P, nu, G):
    """Surface displacement above a pressurised spherical cavity of the given radius at depth D (Mogi 1958):
    u = (radius^3 P (1 - nu) / G) (x, y, D) / R^3, R^2 = D^2 + x^2 + y^2 (Mogi1958_SphericalCavity.jl)."""
    X, Y = np.asarray(X, dtype=float), np.asarray(Y, dtype=float)
    R = np.sqrt(D ** 2 + X ** 2 + Y ** 2)
    c = radius ** 3 * P * (1 - nu) / G
    return X / R ** 3 * c, Y / R ** 3 * c, D / R ** 3 * c


def _quadrant(x):
    """Degrees reduced exactly to rem in [-45, 45] and the quadrant q in 0..3 with x = 90 q + rem (mod 360)."""
    r = np.fmod(np.asarray(x, dtype=np.float64), 360.0)
    q = np.round(r / 90.0)
    rem = r - 90.0 * q
    return np.deg2rad(rem), q.astype(np.int64) % 4


def sind(x):
    """Sine of an angle in degrees, exactly zero at multiples of 180 (as Julia's sind)."""
    rad, q = _quadrant(x)
    return np.choose(q, [np.sin(rad), np.cos(rad), -np.sin(rad), -np.cos(rad)])


def cosd(x):
    """Cosine of an angle in degrees, exactly zero at odd multiples of 90 (as Julia's cosd)."""
    rad, q = _quadrant(x)
    return np.choose(q, [np.cos(rad), -np.sin(rad), -np.cos(rad), np.sin(rad)])


def tada_inclined_penny_tension(szz, beta_deg, theta_deg, radius, nu):
    """Stress intensities around a penny-shaped crack of the given radius inclined by beta (degrees, from the
    loading plane) under remote tension szz (Tada, Paris & Irwin, The Stress Analysis of Cracks Handbook), as
    Tada_StrIntInclinedPennyTension.jl: K_I uniform, K_II ~ cos(theta), K_III ~ sin(theta).  Angles in degrees
    with exact zeros of sind/cosd, so K_II vanishes exactly at theta = +-90."""
    theta = np.asarray(theta_deg, dtype=np.float64)
    sb, cb = float(sind(beta_deg)), float(cosd(beta_deg))
    root = np.sqrt(np.pi * radius)
    K1 = np.ones_like(theta) * (2 / np.pi) * (szz * sb ** 2) * root
    K2 = (4 / (np.pi * (2 - nu))) * (szz * sb * cb) * cosd(theta) * root
    K3 = ((4 * (1 - nu)) / (np.pi * (2 - nu))) * (szz * sb * cb) * sind(theta) * root
    return K1, K2, K3


def pollard_segall1987_slip(G, nu, syy, sxy, x, a):
    """Slip of a 2-D crack of half length a under remote shear (and normal) stress (Pollard & Segall 1987), as
    PollardSegall1987_FractureSlipProfile.jl: 2 (1 - nu) / G sqrt(a^2 - x^2) (syy + sxy)."""
    x = np.asarray(x, dtype=np.float64)
    with np.errstate(invalid="ignore"):
        return (2 * (1 - nu) / G) * np.sqrt(a ** 2 - x ** 2) * (syy + sxy)


def burgmann1994_linear_friction_slip(G, nu, Ts, x, a):
    """Slip of a 2-D crack of half length a under remote shear Ts with a frictional strength rising linearly from
    zero at the centre to Ts at the tips (Burgmann et al. 1994), as Burgmann1994_FractureLinearFrictionSlipProfile.jl
    (x = 0 gives NaN, as there)."""
    x = np.abs(np.asarray(x, dtype=np.float64))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.sqrt(a ** 2 - x ** 2)
        ach = np.arccosh(a / x)
        return ((1 - nu) / G) * (2 * ((Ts * r) - ((Ts * (1 / np.pi)) * (r + (x ** 2 / a) * ach))))


def savage1984_gravity_valley(tectxx, rg, nu, a, b, u, v):
    """Stresses under a long symmetric valley carved into a half-plane loaded by gravity and a remote horizontal
    stress (Savage, Powers & Swolfs 1984, USGS OFR 84-827, the complex-potential solution of their RVT program), as
    Savage1984_GravityValleyStress.jl: the points w = u + i v (v < 0) of the mapping plane go to
    z = w + a b / (w - i a); returns the total stresses (Sxx, Syy, Sxy) and the mapped points (X, Y), tension
    positive, Y up.  As the reference, when any point lies within 1e-4 of w = -i a every point takes the Taylor
    expansion of the second potential about it."""
    u, v = np.asarray(u, dtype=np.float64), np.asarray(v, dtype=np.float64)
    I = 1j
    so = rg * b
    ai = I * a
    # potentials at w = -i a: gravity (eq. 4), tectonic (eq. 12) and the tectonic one's second derivative
    phia = -so * ((4 * a + b) * (1 - nu) + b) / (8 * (1 - nu) * (2 * a + b))
    phiat = -(tectxx * b) / (4 * (2 * a + b))
    d2phia = -(tectxx * b * (4 * a + b) * (b - 12 * a)) / (2 * a * (2 * a + b) * (4 * a + b) ** 3)
    w = u + I * v
    r = np.sqrt(u ** 2 + (v + a) ** 2)
    z = w + a * b / (w - ai)                                   # eq. 1
    wm, wp = w - ai, w + ai
    X, Y = z.real, z.imag
    dz = (wm ** 2 - a * b) / wm ** 2
    aw = -(I * (4 * a + b) / (8 * wm)) - a * b * (w - 3 * ai) / (8 * (1 - nu) * wm ** 3)     # eq. 3
    phi = -aw * so / dz + a * b * phia / (dz * wm ** 2)        # eq. 6
    awt = -(a * b * tectxx) / (2 * wm ** 2)                    # eq. 11
    phit = -awt / dz + a * b * phiat / (dz * wm ** 2)          # eq. 10
    sumt = 4 * phit.real + tectxx                              # eq. 8
    sumy = sumt + 4 * phi.real + rg * Y / (1 - nu)
    dlaw = (I * (4 * a + b) / (8 * wm ** 2) + 2 * a * b / (8 * (1 - nu) * wm ** 3)
            - 6 * I * a ** 2 * b / (8 * (1 - nu) * wm ** 4))
    dlphi = -so * dlaw / dz - (2 * a * b * (phi + phia)) / (dz * wm ** 3)
    d2phi = -(2 * phi / wm ** 2 + 4 * dlphi / wm + so * I * a ** 2 * b / (2 * (1 - nu) * wm ** 5)) / dz
    bw = -so * (I * (4 * a + b) / (8 * wm) + (1 - 2 * nu) * a * b * (w - 3 * ai) / (8 * (1 - nu) * wm ** 3))  # eq. 7
    psil = w * dlphi + bw + phi
    dlphit = -(a * b * tectxx * (4 * a + b) * wm) / (2 * (2 * a + b) * (wm ** 2 - a * b) ** 2)
    psilt = w * dlphit - awt + phit                            # part of eq. 9
    if np.any(r < 1.0e-4):
        psi2 = 0.5 * a * b * d2phi
        psi2t = 0.5 * a * b * d2phia
    else:
        psi2 = a * b * dlphi / wp - a * b * (phi - phia) / wp ** 2
        psi2t = a * b * dlphit / wp - a * b * (phit - phiat) / wp ** 2
    psi = -(psil + psi2) / dz
    psit = -(psilt + psi2t) / dz
    zbar = np.conj(z)
    strl = 2 * (zbar * dlphi / dz + psi)                       # eq. 5
    strlt = 2 * (zbar * dlphit / dz + psit)
    dif = strlt.real - tectxx + strl.real + rg * Y * (1 - 2 * nu) / (1 - nu)
    Sxy = (strl.imag + strlt.imag) / 2
    Sxx = (sumy - dif) / 2
    Syy = (sumy + dif) / 2
    return Sxx, Syy, Sxy, X, Y
