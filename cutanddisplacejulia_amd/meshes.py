"""Seeded synthetic meshes for the workloads (planar cracks and faults) in the reference's array layout.

Every generator returns P1, P2, P3 as N x 3 float64 arrays (one triangle per row).
"""
import numpy as np

BASE_SEED = 20261203


def _grid_tris(nu_, nv_, pts):
    """Split an (nu_+1) x (nv_+1) vertex grid (pts[i, j]) into 2 * nu_ * nv_ triangles along a fixed diagonal."""
    P1, P2, P3 = [], [], []
    for i in range(nu_):
        for j in range(nv_):
            a, b, c, d = pts[i, j], pts[i + 1, j], pts[i + 1, j + 1], pts[i, j + 1]
            P1 += [a, a]
            P2 += [b, c]
            P3 += [c, d]
    return np.array(P1), np.array(P2), np.array(P3)


def square_crack(n=8, half=1.0, jitter=0.05, seed=BASE_SEED + 1):
    """Planar square crack [-half, half]^2 at z = 0 split into n x n quads (2 n^2 triangles); interior vertices
    are moved in-plane by U(-jitter, jitter) per axis."""
    rng = np.random.default_rng(seed)
    u = np.linspace(-half, half, n + 1)
    pts = np.zeros((n + 1, n + 1, 3))
    for i in range(n + 1):
        for j in range(n + 1):
            x, y = u[i], u[j]
            if 0 < i < n and 0 < j < n:
                x += rng.uniform(-jitter, jitter)
                y += rng.uniform(-jitter, jitter)
            pts[i, j] = (x, y, 0.0)
    return _grid_tris(n, n, pts)


def strike_dip_cosines(strike_deg, dip_deg):
    """Strike and dip direction cosines of a plane (CreateDirectionCosinesFromStrikeAndDip.jl:1-13)."""
    cd = np.cos(np.deg2rad(dip_deg))
    ds_az = np.cos(np.arccos(cd) - np.pi / 2)
    L = np.sin(np.arccos(ds_az))
    a = np.arctan2(0.0, cd)
    ssc = np.array([-0.0, 1.0, 0.0])
    dsc = np.array([-np.cos(a) * L, -np.sin(a) * L, ds_az])
    ang = np.deg2rad(-strike_deg)
    R = np.array([[np.cos(ang), -np.sin(ang), 0.0], [np.sin(ang), np.cos(ang), 0.0], [0.0, 0.0, 1.0]])
    return R @ ssc, R @ dsc


def rect_fault(n_strike=50, n_dip=50, strike=60.0, dip=45.0, length=5.0, width=5.0, top_depth=0.5):
    """Planar rectangular fault (length along strike, width down dip) whose top edge is at z = -top_depth,
    meshed into n_strike x n_dip quads (2 n_strike n_dip triangles)."""
    ssc, dsc = strike_dip_cosines(strike, dip)
    s = np.linspace(-length / 2, length / 2, n_strike + 1)
    d = np.linspace(-width / 2, width / 2, n_dip + 1)
    pts = np.zeros((n_strike + 1, n_dip + 1, 3))
    for i in range(n_strike + 1):
        for j in range(n_dip + 1):
            pts[i, j] = s[i] * ssc + d[j] * dsc
    pts[..., 2] -= pts[..., 2].max() + top_depth
    return _grid_tris(n_strike, n_dip, pts)


def grid_obs(n=100, half=10.0, z=0.0):
    """n x n observation grid over [-half, half]^2 at height z, flattened column-major (meshgrid order)."""
    u = np.linspace(-half, half, n)
    X, Y = np.meshgrid(u, u)
    return X.ravel(order="F"), Y.ravel(order="F"), np.full(n * n, float(z))


def penny(n_rings=10, radius=1.0, seed=BASE_SEED + 4, jitter=0.0, center=(0.0, 0.0, 0.0), normal=(0, 0, 1),
          aspect=1.0):
    """Planar disc (or ellipse with semi-axes radius*aspect, radius) meshed with concentric rings: ring k has
    6 k vertices.  Returns P1, P2, P3 rotated so the disc normal is `normal` and centred at `center`."""
    rng = np.random.default_rng(seed)
    pts = [np.zeros(2)]
    rings = [[0]]
    for k in range(1, n_rings + 1):
        r = radius * k / n_rings
        m = 6 * k
        th = 2 * np.pi * np.arange(m) / m
        ring = []
        for t in th:
            p = np.array([r * np.cos(t), r * np.sin(t)])
            if k < n_rings and jitter > 0:
                p += rng.uniform(-jitter, jitter, 2) * radius / n_rings
            ring.append(len(pts))
            pts.append(p)
        rings.append(ring)
    pts = np.array(pts)
    tris = []
    for k in range(1, n_rings + 1):
        inner, outer = rings[k - 1], rings[k]
        mi, mo = len(inner), len(outer)
        # walk both rings by angle
        i = j = 0
        while i < mi or j < mo:
            ai = (i / mi) if mi > 1 else 1.0
            ao = j / mo
            if mi == 1:
                tris.append((inner[0], outer[j % mo], outer[(j + 1) % mo]))
                j += 1
                if j >= mo:
                    break
                continue
            if (ao + 1.0 / mo) <= (ai + 1.0 / mi) and j < mo:
                tris.append((inner[i % mi], outer[j % mo], outer[(j + 1) % mo]))
                j += 1
            else:
                tris.append((inner[i % mi], outer[j % mo], inner[(i + 1) % mi]))
                i += 1
    tris = np.array(tris)
    xy = pts.copy()
    xy[:, 0] *= aspect
    P = np.concatenate([xy, np.zeros((len(xy), 1))], axis=1)
    nrm = np.asarray(normal, dtype=float)
    nrm = nrm / np.linalg.norm(nrm)
    z = np.array([0.0, 0.0, 1.0])
    v = np.cross(z, nrm)
    s = np.linalg.norm(v)
    c = np.dot(z, nrm)
    if s > 1e-15:
        vx = np.array([[0, -v[2], v[1]], [v[2], 0, -v[0]], [-v[1], v[0], 0]])
        R = np.eye(3) + vx + vx @ vx * ((1 - c) / s ** 2)
    else:
        R = np.eye(3) if c > 0 else np.diag([1.0, -1.0, -1.0])
    P = P @ R.T + np.asarray(center, dtype=float)
    return P[tris[:, 0]], P[tris[:, 1]], P[tris[:, 2]]


def config1(seed=BASE_SEED + 1):
    """Full space, FP64: square crack z = 0 side 2, 8 x 8 quads -> 128 triangles, interior vertices jittered
    U(-0.05, 0.05); observation points = 128 centroids + 50 x 50 grid at z = 0.3 over [-2, 2]^2;
    mu = 1, nu = 0.25, remote szz = 1."""
    P1, P2, P3 = square_crack(8, 1.0, 0.05, seed)
    cen = (P1 + P2 + P3) / 3
    gx, gy, gz = grid_obs(50, 2.0, 0.3)
    X = np.concatenate([cen[:, 0], gx])
    Y = np.concatenate([cen[:, 1], gy])
    Z = np.concatenate([cen[:, 2], gz])
    return dict(P1=P1, P2=P2, P3=P3, X=X, Y=Y, Z=Z, nu=0.25, G=1.0, hs=0)


def config2(n=50):
    """Half space: 45-degree dipping rectangular fault (strike 60, 5 x 5), top at depth 0.5, n x n quads
    (5000 triangles for n = 50, A is 15000 x 15000); 100 x 100 free-surface grid over [-10, 10]^2; nu = 0.25."""
    P1, P2, P3 = rect_fault(n, n, 60.0, 45.0, 5.0, 5.0, 0.5)
    X, Y, Z = grid_obs(100, 10.0, 0.0)
    return dict(P1=P1, P2=P2, P3=P3, X=X, Y=Y, Z=Z, nu=0.25, G=1.0, hs=1)


def growth_sequence(steps=40, n_start=500, n_end=6400, radius=1.0, dip=45.0, jitter=0.03, seed=BASE_SEED + 4):
    """Penny crack of the given radius remeshed at `steps` growth steps whose triangle counts follow a geometric
    schedule from n_start to n_end (ring meshes: 6 r^2 triangles for r rings), interior vertices jittered by
    U(+-jitter) of the ring spacing, the disc dipping `dip` degrees (strike along y).  Each step also gets a
    remote-load gradient g ~ U(0, 0.5) for sigma_zz = 1 + g z.  Returns a list of dicts."""
    rng = np.random.default_rng(seed)
    out = []
    d = np.deg2rad(dip)
    normal = (np.sin(d), 0.0, np.cos(d))
    for s in range(steps):
        target = n_start * (n_end / n_start) ** (s / max(steps - 1, 1))
        rings = max(2, int(round(np.sqrt(target / 6.0))))
        P1, P2, P3 = penny(rings, radius, seed=seed + 1000 + s, jitter=jitter, center=(0.0, 0.0, -2.0),
                           normal=normal)
        out.append(dict(step=s + 1, P1=P1, P2=P2, P3=P3, g=float(rng.uniform(0.0, 0.5))))
    return out


def random_rotation(rng):
    """Uniformly random rotation matrix (QR of a Gaussian matrix with sign fix)."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def elliptical_cracks(count=64, n_min=200, n_max=3000, box=10.0, seed=BASE_SEED + 7, rhs_min=1, rhs_max=8):
    """Batch of planar elliptical cracks: semi-axes (aspect, 1) with aspect ~ U(1, 4), uniformly random orientation,
    centre ~ U(-box, box)^3, triangle count ~ U(n_min, n_max) (ring meshes: 6 r^2 triangles for r rings), and
    1-8 remote stress tensors per crack with N(0, 1) entries (symmetric).  Returns a list of dicts with P1, P2,
    P3 and `stresses` (k x 6: sxx, syy, szz, sxy, sxz, syz)."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(count):
        target = rng.uniform(n_min, n_max)
        rings = max(2, int(round(np.sqrt(target / 6.0))))
        aspect = rng.uniform(1.0, 4.0)
        R = random_rotation(rng)
        centre = rng.uniform(-box, box, 3)
        P1, P2, P3 = penny(rings, 1.0, seed=seed + 100 + c, aspect=aspect)
        P1, P2, P3 = (P @ R.T + centre for P in (P1, P2, P3))
        k = int(rng.integers(rhs_min, rhs_max + 1))
        out.append(dict(P1=P1, P2=P2, P3=P3, stresses=rng.normal(size=(k, 6)), aspect=aspect))
    return out


def config5(rings=52, a=2.0, b=1.0):
    """Large full-space case: planar elliptical crack with semi-axes a (x) and b (y) in z = 0, ring-meshed
    (6 rings^2 triangles: 16224 for 52 rings, A is 48672 x 48672 FP64), 200 x 200 observation grid at
    z = 0.5 over [-4, 4]^2; nu = 0.25, mu = 1, remote sigma_zz = 1."""
    P1, P2, P3 = penny(rings, b, aspect=a / b)
    X, Y, Z = grid_obs(200, 4.0, 0.5)
    return dict(P1=P1, P2=P2, P3=P3, X=X, Y=Y, Z=Z, nu=0.25, G=1.0, hs=0, a=a, b=b)
