"""CPU restatement of the crack-tip path: free (crack-tip) edges of a triangle mesh, their geometry, and the
stress intensities K_I, K_II, K_III estimated from the displacement discontinuity of the tip triangles.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's CPU-baseline leg, never by the product package.

Follows, in the same operation order:
* GetCrackTipElements3D.jl:1-221 (GetValues, the edge-vector convention fix and recomputation),
* EdgeConstraints.jl (free-edge flags, including the midpoint-distance skip), CreateSortedEdgePoints.jl,
  MatchingRow.jl, AreaOfTriangle3D.jl (Heron, with the half perimeter),
* StressIntensity3D.jl:1-244 (K1K2K3TriDislocation), RotateObject3DAllignVectors.jl,
  CalculateNormalShearTraction2D.jl, StrainEnergyRelease.jl (its active formula),
* normr.jl (LinearAlgebra.normalize: multiplication by the reciprocal of the 2-norm).
CalculateSSandDSDirs comes from the C oracle (oracle.ss_ds_dirs).

Per edge (P1P2, P1P3, P2P3 with the third corner P3, P2, P1) the result is a dict of the reference's
TriangleEdges fields; entries of non-free rows are NaN, FreeFlg is boolean, Area is filled for every row.
"""
import numpy as np

from oracle import oracle as O

EDGES = (("P1P2", 0, 1, 2), ("P1P3", 0, 2, 1), ("P2P3", 1, 2, 0))


def _normalize(v):
    """LinearAlgebra.normalize(v, 2) row by row: v * inv(norm(v))."""
    v = np.asarray(v, dtype=np.float64)
    nrm = np.sqrt(v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1] + v[..., 2] * v[..., 2])
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / nrm
    return v * inv[..., None]


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def rotate_align(A, B, v):
    """RotateObject3DAllignVectors(A, B, x, y, z, 0, 0, 0) for one vector v: the Rodrigues rotation taking A
    onto B (reflection through the origin when B == -A)."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    if np.all(B == -A):
        return -v
    w = _cross(A, B)
    with np.errstate(invalid="ignore", divide="ignore"):
        w = w / np.sqrt(np.sum(w * w))
    if np.any(np.isnan(w)):
        w = np.zeros(3)
    W = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    cos_t = np.dot(A, B) / np.sqrt(np.dot(A, A)) / np.sqrt(np.dot(B, B))
    with np.errstate(invalid="ignore"):
        t = np.arccos(cos_t)
    R = np.eye(3) + W * np.sin(t) + (W @ W) * (1.0 - np.cos(t))
    return R @ v


def sorted_edge_points(Pa, Pb):
    """CreateSortedEdgePoints.jl: the two end points of each edge in lexicographic (x, y, z) order -> n x 6."""
    out = np.empty((Pa.shape[0], 6))
    for i in range(Pa.shape[0]):
        a, b = Pa[i], Pb[i]
        swap = False
        for c in range(3):
            if a[c] != b[c]:
                swap = not (a[c] < b[c])
                break
        lo, hi = (b, a) if swap else (a, b)
        out[i, :3] = lo
        out[i, 3:] = hi
    return out


def area_half_perimeter(P1, P2, P3):
    """AreaOfTriangle3D.jl (Heron's formula) -> (Area, half perimeter)."""
    def dist(a, b):
        return np.sqrt((a[:, 0] - b[:, 0]) ** 2 + (a[:, 1] - b[:, 1]) ** 2 + (a[:, 2] - b[:, 2]) ** 2)
    a, b, c = dist(P1, P2), dist(P2, P3), dist(P3, P1)
    s = (a + b + c) / 2.0
    with np.errstate(invalid="ignore"):
        area = np.sqrt(s * (s - a) * (s - b) * (s - c))
    return area, s


def edge_constraints(P1, P2, P3, mid):
    """EdgeConstraints.jl: free-edge flags (P1P2, P2P3, P3P1).  An edge is not free when another edge of the
    mesh (of another triangle, or another side of the same one) has identical sorted end points; pairs whose
    midpoints are further apart than the sum of the two half perimeters are skipped."""
    n = P1.shape[0]
    S = [sorted_edge_points(P1, P2), sorted_edge_points(P2, P3), sorted_edge_points(P3, P1)]
    free = [np.ones(n, dtype=bool) for _ in range(3)]
    _, hp = area_half_perimeter(P1, P2, P3)
    d = mid[:, None, :] - mid[None, :, :]
    middist = np.sqrt(d[..., 0] ** 2 + d[..., 1] ** 2 + d[..., 2] ** 2)
    near = ~(middist > hp[:, None] + hp[None, :])

    def match(a, b):
        eq = np.ones((n, n), dtype=bool)
        for p in range(6):
            eq &= S[a][:, p][:, None] == S[b][:, p][None, :]
        return eq & near

    # different edge types, i == j included
    for a, b in ((0, 1), (0, 2), (1, 2)):
        m = match(a, b)
        free[a][m.any(axis=1)] = False
        free[b][m.any(axis=0)] = False
    # same edge type, i != j
    off = ~np.eye(n, dtype=bool)
    for a in range(3):
        m = match(a, a) & off
        hit = m.any(axis=1) | m.any(axis=0)
        free[a][hit] = False
    return free[0], free[1], free[2]


def _get_values(flg, Pa, Pb, Pc, mid, nrm):
    """GetValues of GetCrackTipElements3D.jl for one edge type."""
    n = Pa.shape[0]
    nan = np.full(n, np.nan)
    FeLe, FeM2ELe, IntAng = nan.copy(), nan.copy(), nan.copy()
    FeMd, FeEv, FeM2Ev = (np.full((n, 3), np.nan) for _ in range(3))
    I = flg
    FeLe[I] = np.sqrt((Pa[I, 0] - Pb[I, 0]) ** 2 + (Pa[I, 1] - Pb[I, 1]) ** 2 + (Pa[I, 2] - Pb[I, 2]) ** 2)
    FeMd[I] = (Pa[I] + Pb[I]) / 2
    FeM2ELe[I] = np.sqrt((FeMd[I, 0] - mid[I, 0]) ** 2 + (FeMd[I, 1] - mid[I, 1]) ** 2 +
                         (FeMd[I, 2] - mid[I, 2]) ** 2)
    FeM2Ev[I] = _normalize(FeMd[I] - mid[I])
    v = _normalize(Pa[I] - Pc[I])
    w = _normalize(Pb[I] - Pc[I])
    with np.errstate(invalid="ignore"):
        IntAng[I] = np.rad2deg(np.arccos(_dot(v, w)))
    FeEv[I] = _normalize(Pa[I] - Pb[I])
    idx = np.nonzero(I)[0]
    ez = np.array([0.0, 0.0, 1.0])
    # edge vector counter-clockwise from the mid-to-edge vector, looking along the normal
    for i in idx:
        m2 = rotate_align(nrm[i], ez, FeM2Ev[i])
        ev = rotate_align(nrm[i], ez, FeEv[i])
        if m2[0] * ev[1] + m2[1] * (-ev[0]) < 0:
            FeEv[i] = -FeEv[i]
    for i in idx:
        m2 = _cross(nrm[i], FeEv[i])
        if np.dot(FeM2Ev[i], m2) < 0:
            m2 = -m2
        FeM2Ev[i] = m2
    for i in idx:
        FeEv[i] = _cross(FeM2Ev[i], nrm[i])
    area, _ = area_half_perimeter(Pa, Pb, Pc)
    return dict(FeLe=FeLe, FeMd=FeMd, FeEv=FeEv, FeM2Ev=FeM2Ev, FreeFlg=flg.copy(), FeM2ELe=FeM2ELe,
                IntAng=IntAng, Area=area)


def crack_tip_elements(mid, P1, P2, P3, nrm):
    """GetCrackTipElements3D(MidPoint, P1, P2, P3, FaceNormalVector) -> (FeP1P2S, FeP1P3S, FeP2P3S)."""
    P1, P2, P3, mid, nrm = (np.asarray(a, dtype=np.float64) for a in (P1, P2, P3, mid, nrm))
    f12, f23, f31 = edge_constraints(P1, P2, P3, mid)
    P = (P1, P2, P3)
    flags = {"P1P2": f12, "P1P3": f31, "P2P3": f23}
    return tuple(_get_values(flags[name], P[a], P[b], P[c], mid, nrm) for name, a, b, c in EDGES)


def _k1k2k3(flg, G, nu, Dn, FeM2Ev, FeEv, FeM2ELe, nrm, Dss, Dds):
    """K1K2K3TriDislocation of StressIntensity3D.jl."""
    n = flg.shape[0]
    idx = np.nonzero(flg)[0]
    K1, K2, K3 = (np.full(n, np.nan) for _ in range(3))
    if idx.size == 0:
        return K1, K2, K3, np.full(n, np.nan)
    ss, ds = O.ss_ds_dirs(nrm[idx])
    Dpl = ss * Dss[idx][:, None] + ds * Dds[idx][:, None]
    ez = np.array([0.0, 0.0, 1.0])
    angle = np.zeros(idx.size)
    for q, i in enumerate(idx):
        dss = rotate_align(nrm[i], ez, ss[q])
        m2e = rotate_align(nrm[i], ez, FeM2Ev[i])
        mx, sx = m2e[0], dss[0]
        if abs(mx) > 1:
            mx = 1.0
        if abs(sx) > 1:
            sx = 1.0
        angle[q] = abs(np.arccos(sx) - np.arccos(mx))
    dir1 = np.pi / 2 + angle
    cx, cy = np.sin(dir1), np.cos(dir1)
    along = np.abs(Dds[idx] * cx + Dss[idx] * cy)
    mid2 = np.abs(-Dds[idx] * cy + Dss[idx] * cx)
    mid2 = np.where(_dot(FeM2Ev[idx], Dpl) <= 0, -mid2, mid2)
    along = np.where(_dot(FeEv[idx], Dpl) <= 0, -along, along)
    h = FeM2ELe[idx]
    c = (G * np.sqrt(np.pi)) / (2 * np.sqrt(h) * (1 - nu))
    K1[idx] = c * Dn[idx]
    K2[idx] = c * mid2
    K3[idx] = c * along * (1 - nu)
    K1, K2, K3 = K1 / 1.834, K2 / 1.834, K3 / 1.834
    se = np.sqrt(np.abs(K1) ** 2 + np.abs(K2) ** 2 + (1 / (1 - nu)) * np.abs(K3) ** 2)
    return K1, K2, K3, se


def stress_intensity(Dn, Dss, Dds, G, nu, nrm, edges):
    """StressIntensity3D(Dn, Dss, Dds, G, nu, FaceNormalVector, FeP1P2S, FeP1P3S, FeP2P3S): the same three
    edge dicts with K1, K2, K3 and StrainEnergy added (NaN off the free edges)."""
    Dn, Dss, Dds, nrm = (np.asarray(a, dtype=np.float64) for a in (Dn, Dss, Dds, nrm))
    out = []
    for e in edges:
        K1, K2, K3, se = _k1k2k3(e["FreeFlg"], G, nu, Dn, e["FeM2Ev"], e["FeEv"], e["FeM2ELe"], nrm, Dss, Dds)
        d = dict(e)
        d.update(K1=K1, K2=K2, K3=K3, StrainEnergy=se)
        out.append(d)
    return tuple(out)
