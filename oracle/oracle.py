"""ctypes binding of the CPU FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg, never
by the product package.  Array conventions are the reference's Julia ones: P1/P2/P3 and normals are N x 3 and
are passed column-major (Fortran order); influence matrices are nobs x N column-major.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.c_double
_i64 = ctypes.c_int64
_i = ctypes.c_int
_p = ctypes.c_void_p

STRAIN_FIELDS = ["exxDn", "eyyDn", "ezzDn", "exyDn", "exzDn", "eyzDn",
                 "exxDss", "eyyDss", "ezzDss", "exyDss", "exzDss", "eyzDss",
                 "exxDds", "eyyDds", "ezzDds", "exyDds", "exzDds", "eyzDds"]
DISP_FIELDS = ["DnUx", "DnUy", "DnUz", "DssUx", "DssUy", "DssUz", "DdsUx", "DdsUy", "DdsUz"]
TRACTION_FIELDS = ["DnTn", "DnTss", "DnTds", "DssTn", "DssTss", "DssTds", "DdsTn", "DdsTss", "DdsTds"]


class OracleError(RuntimeError):
    pass


def build():
    """Compile liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_td.restype = _i
        L.orc_td.argtypes = [_p, _p, _p, _i64, _p, _p, _p, _i64, _p, _p, _p, _d, _d, _i, _i, _i, _p, _p, _i]
        L.orc_td_field.restype = _i
        L.orc_td_field.argtypes = L.orc_td.argtypes
        L.orc_ss_ds_dirs.restype = None
        L.orc_ss_ds_dirs.argtypes = [_p, _i64, _p, _p]
        L.orc_traction_influence.restype = _i
        L.orc_traction_influence.argtypes = [_p, _p, _p, _p, _p, _i64, _d, _d, _d, _i, _p, _i]
        L.orc_normal_shear_tractions.restype = None
        L.orc_normal_shear_tractions.argtypes = [_p] * 7 + [_i64, _p, _p, _p]
        L.orc_face_normal_midpoint.restype = _i
        L.orc_face_normal_midpoint.argtypes = [_p, _p, _p, _i64, _p, _p]
        for name, n in (("orc_ang_dis_strain_fsc", 6), ("orc_ang_dis_disp_fsc", 6)):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [_d] * n + [_p]
        for name in ("orc_ang_dis_strain", "orc_ang_dis_disp"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [_d] * 5 + [_p]
        _LIB = L
    return _LIB


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _ptr(a):
    return a.ctypes.data_as(_p)


def _ptr_array(arrs):
    arr = (_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arr


def _tri_arrays(P1, P2, P3):
    P1, P2, P3 = (_f64(np.atleast_2d(P)) for P in (P1, P2, P3))
    return P1, P2, P3, P1.shape[0]


def td(X, Y, Z, P1, P2, P3, Dss, Dds, Dn, nu, G, disp_flag, strain_flag, hs_flag, nthreads=0):
    """Influence form of TD (TD.jl:77-216).  Returns (strain dict or None, disp dict or None)."""
    X, Y, Z = (_f64(np.ravel(v, order="F")) for v in (X, Y, Z))
    P1, P2, P3, n = _tri_arrays(P1, P2, P3)
    Dss, Dds, Dn = (_f64(np.ravel(v)) for v in (Dss, Dds, Dn))
    if not (len(Dss) == len(Dds) == len(Dn) == n and P2.shape[0] == n and P3.shape[0] == n):
        raise OracleError("Element size inputs must be the same dimensions")
    m = X.size
    strain = [np.zeros((m, n), order="F") for _ in range(18)] if strain_flag == 1 else None
    disp = [np.zeros((m, n), order="F") for _ in range(9)] if disp_flag == 1 else None
    rc = lib().orc_td(_ptr(X), _ptr(Y), _ptr(Z), m, _ptr(P1), _ptr(P2), _ptr(P3), n, _ptr(Dss), _ptr(Dds),
                      _ptr(Dn), float(nu), float(G), int(disp_flag), int(strain_flag), int(hs_flag),
                      _ptr_array(strain) if strain else None, _ptr_array(disp) if disp else None, int(nthreads))
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    s = dict(zip(STRAIN_FIELDS, strain)) if strain else None
    d = dict(zip(DISP_FIELDS, disp)) if disp else None
    return s, d


def td_field(X, Y, Z, P1, P2, P3, Dss, Dds, Dn, nu, G, disp_flag, strain_flag, hs_flag, nthreads=0):
    """Summed-field form of TD (TD.jl:219-410).  Returns (strain6 list or None, disp3 list or None)."""
    X, Y, Z = (_f64(np.ravel(v, order="F")) for v in (X, Y, Z))
    P1, P2, P3, n = _tri_arrays(P1, P2, P3)
    Dss, Dds, Dn = (_f64(np.ravel(v)) for v in (Dss, Dds, Dn))
    m = X.size
    strain = [np.zeros(m) for _ in range(6)] if strain_flag == 1 else None
    disp = [np.zeros(m) for _ in range(3)] if disp_flag == 1 else None
    rc = lib().orc_td_field(_ptr(X), _ptr(Y), _ptr(Z), m, _ptr(P1), _ptr(P2), _ptr(P3), n, _ptr(Dss), _ptr(Dds),
                            _ptr(Dn), float(nu), float(G), int(disp_flag), int(strain_flag), int(hs_flag),
                            _ptr_array(strain) if strain else None, _ptr_array(disp) if disp else None,
                            int(nthreads))
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return strain, disp


def face_normal_midpoint(P1, P2, P3):
    P1, P2, P3, n = _tri_arrays(P1, P2, P3)
    nrm = np.zeros((n, 3), order="F")
    mid = np.zeros((n, 3), order="F")
    rc = lib().orc_face_normal_midpoint(_ptr(P1), _ptr(P2), _ptr(P3), n, _ptr(nrm), _ptr(mid))
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return nrm, mid


def ss_ds_dirs(normals):
    normals = _f64(normals)
    n = normals.shape[0]
    ss = np.zeros((n, 3), order="F")
    ds = np.zeros((n, 3), order="F")
    lib().orc_ss_ds_dirs(_ptr(normals), n, _ptr(ss), _ptr(ds))
    return ss, ds


def traction_influence(normals, mid, P1, P2, P3, nu, G, lam, hs_flag, nthreads=0):
    """CalculateInfluenceMatrices3D without fixed elements -> dict of the 9 TractionInf matrices."""
    normals, mid = _f64(normals), _f64(mid)
    P1, P2, P3, n = _tri_arrays(P1, P2, P3)
    T = [np.zeros((n, n), order="F") for _ in range(9)]
    rc = lib().orc_traction_influence(_ptr(normals), _ptr(mid), _ptr(P1), _ptr(P2), _ptr(P3), n, float(nu),
                                      float(G), float(lam), int(hs_flag), _ptr_array(T), int(nthreads))
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return dict(zip(TRACTION_FIELDS, T))


def normal_shear_tractions(stress6, normals):
    """CalculateNormalAndShearTractions3D for per-element stresses -> (Tn, Tds, Tss)."""
    normals = _f64(normals)
    n = normals.shape[0]
    s = [_f64(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)).copy()) for v in stress6]
    Tn, Tds, Tss = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_normal_shear_tractions(*[_ptr(v) for v in s], _ptr(normals), n, _ptr(Tn), _ptr(Tds), _ptr(Tss))
    return Tn, Tds, Tss


def assemble_A(T):
    """SlipCalculator3D.jl:13-16: A = -[DnTn DssTn DdsTn; DnTss DssTss DdsTss; DnTds DssTds DdsTds]."""
    return -np.block([[T["DnTn"], T["DssTn"], T["DdsTn"]],
                      [T["DnTss"], T["DssTss"], T["DdsTss"]],
                      [T["DnTds"], T["DssTds"], T["DdsTds"]]])


def ang_dis_strain_fsc(y1, y2, y3, beta, nu, a):
    o = np.zeros(18)
    lib().orc_ang_dis_strain_fsc(y1, y2, y3, beta, nu, a, _ptr(o))
    return o


def ang_dis_disp_fsc(y1, y2, y3, beta, nu, a):
    o = np.zeros(9)
    lib().orc_ang_dis_disp_fsc(y1, y2, y3, beta, nu, a, _ptr(o))
    return o


def ang_dis_strain(x, y, z, alpha, nu):
    o = np.zeros(21)
    lib().orc_ang_dis_strain(x, y, z, alpha, nu, _ptr(o))
    return o


def ang_dis_disp(x, y, z, alpha, nu):
    o = np.zeros(9)
    lib().orc_ang_dis_disp(x, y, z, alpha, nu, _ptr(o))
    return o


def count_traction_flops(normals, mid, P1, P2, P3, nu, G, lam, hs_flag, recv_idx, src_idx, setup_too=False):
    """Operation counts of the traction assembly over the given (receiver, source) pairs.

    Returns dict(add, mul, div, sqrt, log, atan, trig, pairs); +, -, *, / count one each.
    """
    L = lib()
    f = L.orc_count_traction_flops
    f.restype = None
    f.argtypes = [_p, _p, _p, _p, _p, _i64, _d, _d, _d, _i, _p, _p, _i64, _i, _p]
    normals, mid = _f64(normals), _f64(mid)
    P1, P2, P3, n = _tri_arrays(P1, P2, P3)
    ri = np.ascontiguousarray(recv_idx, dtype=np.int64)
    si = np.ascontiguousarray(src_idx, dtype=np.int64)
    out = np.zeros(8)
    f(_ptr(normals), _ptr(mid), _ptr(P1), _ptr(P2), _ptr(P3), n, float(nu), float(G), float(lam), int(hs_flag),
      _ptr(ri), _ptr(si), ri.size, int(bool(setup_too)), _ptr(out))
    return dict(zip(["add", "mul", "div", "sqrt", "log", "atan", "trig", "pairs"], out.tolist()))


def count_fsc_side_flops(normals, mid, P1, P2, P3, nu, recv_idx, side_idx):
    """Operation counts of the free-surface strain of single triangle sides (all three slip modes) at collocation
    points: side_idx = k n + j is side k (P1P2, P2P3, P3P1) of triangle j.  Returns dict as count_traction_flops."""
    L = lib()
    f = L.orc_count_fsc_side_flops
    f.restype = None
    f.argtypes = [_p, _p, _p, _p, _p, _i64, _d, _p, _p, _i64, _p]
    normals, mid = _f64(normals), _f64(mid)
    P1, P2, P3, n = _tri_arrays(P1, P2, P3)
    ri = np.ascontiguousarray(recv_idx, dtype=np.int64)
    si = np.ascontiguousarray(side_idx, dtype=np.int64)
    out = np.zeros(8)
    f(_ptr(normals), _ptr(mid), _ptr(P1), _ptr(P2), _ptr(P3), n, float(nu), _ptr(ri), _ptr(si), ri.size, _ptr(out))
    return dict(zip(["add", "mul", "div", "sqrt", "log", "atan", "trig", "pairs"], out.tolist()))


def traction_influence_rows(normals, mid, P1, P2, P3, nu, G, lam, hs_flag, r0, r1, nthreads=0):
    """Traction influence of receiver rows [r0, r1) against all sources (bounded CPU-baseline sample)."""
    L = lib()
    f = L.orc_traction_influence_rows
    f.restype = _i
    f.argtypes = [_p, _p, _p, _p, _p, _i64, _d, _d, _d, _i, _i64, _i64, _p, _i]
    normals, mid = _f64(normals), _f64(mid)
    P1, P2, P3, n = _tri_arrays(P1, P2, P3)
    T = [np.zeros((r1 - r0, n), order="F") for _ in range(9)]
    rc = f(_ptr(normals), _ptr(mid), _ptr(P1), _ptr(P2), _ptr(P3), n, float(nu), float(G), float(lam), int(hs_flag),
           int(r0), int(r1), _ptr_array(T), int(nthreads))
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return dict(zip(TRACTION_FIELDS, T))


def traction_influence_fixed(normals, mid, P1, P2, P3, nu, G, lam, hs_flag, fixed, nthreads=0):
    """CalculateInfluenceMatrices3D with fixed elements -> dict of the 9 TractionInf matrices, each n x nf."""
    L = lib()
    f = L.orc_traction_influence_fixed
    f.restype = _i
    f.argtypes = [_p, _p, _p, _p, _p, _i64, _d, _d, _d, _i, _p, _p, _i]
    normals, mid = _f64(normals), _f64(mid)
    P1, P2, P3, n = _tri_arrays(P1, P2, P3)
    fx = np.ascontiguousarray((np.asarray(fixed).ravel() == 1).astype(np.int32))
    nf = n - int(fx.sum())
    T = [np.zeros((n, nf), order="F") for _ in range(9)]
    rc = f(_ptr(normals), _ptr(mid), _ptr(P1), _ptr(P2), _ptr(P3), n, float(nu), float(G), float(lam), int(hs_flag),
           _ptr(fx), _ptr_array(T), int(nthreads))
    if rc:
        raise OracleError(lib().orc_last_error().decode())
    return dict(zip(TRACTION_FIELDS, T))
