"""Python mirror of CutAndDisplaceJulia's TD / influence / slip-solve interface, running on the HIP engine.

Names, argument order and meaning follow the reference so that callers (and tests) read like the reference:
  TD                              src/TD.jl:77 (allocating), :135 (preallocated StrainInf/DispInf), :219 (summed)
  CalculateInfluenceMatrices3D    src/CalculateInfluenceMatrices3D.jl:2
  SlipCalculator3D                src/SlipCalculator3D.jl:1 (base method)
  SetupTractionVector             src/SetupTractionVector.jl:2-41
  CreateFaceNormalAndMidPoint     src/CreateFaceNormalAndMidPoint.jl (P1, P2, P3 form)
  ElasticConstantsCheck           src/ElasticConstantsCheck.jl:37-49 (G, nu form)
  GetCrackTipElements3D           src/GetCrackTipElements3D.jl:1
  StressIntensity3D               src/StressIntensity3D.jl:1
  TD_sum                          src/TD_sum.jl:1-41
  AreaOfTriangle3D                src/AreaOfTriangle3D.jl:52-131
  GetDesiredEdgeLength            src/GetDesiredEdgeLength.jl:1-12
Data types mirror src/CutAndDisplaceJulia.jl:46-91.  Arrays are numpy; N x 3 arrays hold one triangle per row
(P1[:, 0] are x coordinates), as in the reference.  Every TD / assembly / solve call runs in libcdjhip.so; the
O(N) host-side helpers here (normals, incentres, remote-stress tractions) prepare inputs as the reference's
callers do.
"""
import ctypes
from dataclasses import dataclass, field, fields
from typing import Any

import numpy as np

from . import _lib
from ._lib import CdjError, Opts, Report, check, ctx, f64, load, ptr, ptr_array

# ---- data types (CutAndDisplaceJulia.jl:46-91) -----------------------------------------------------------


@dataclass
class Stresses:
    σxx: Any = None
    σyy: Any = None
    σzz: Any = None
    σxy: Any = None
    σxz: Any = None
    σyz: Any = None


@dataclass
class Strains:
    εxx: Any = field(default_factory=list)
    εyy: Any = field(default_factory=list)
    εzz: Any = field(default_factory=list)
    εxy: Any = field(default_factory=list)
    εxz: Any = field(default_factory=list)
    εyz: Any = field(default_factory=list)


@dataclass
class Disps:
    Ux: Any = field(default_factory=list)
    Uy: Any = field(default_factory=list)
    Uz: Any = field(default_factory=list)


@dataclass
class Tractions:
    Tn: Any = None
    Tss: Any = None
    Tds: Any = None


@dataclass
class MixedBoundaryConditions:
    Stresses: Any = None
    Tractions: Any = None


@dataclass
class FrictionParameters:
    µ: Any = 0.0
    Sf: Any = 0.0


@dataclass
class MixedBoundaryConditionsFriction:
    MixedBoundaryConditions: Any = None
    FrictionParameters: Any = None


@dataclass
class MixedBoundaryConditionsFluidVolume:
    MixedBoundaryConditions: Any = None
    Volumes: Any = None


@dataclass
class TractionInf:
    DnTn: Any = None
    DnTss: Any = None
    DnTds: Any = None
    DssTn: Any = None
    DssTss: Any = None
    DssTds: Any = None
    DdsTn: Any = None
    DdsTss: Any = None
    DdsTds: Any = None


@dataclass
class StrainInf:
    εxxDn: Any = field(default_factory=list)
    εyyDn: Any = field(default_factory=list)
    εzzDn: Any = field(default_factory=list)
    εxyDn: Any = field(default_factory=list)
    εxzDn: Any = field(default_factory=list)
    εyzDn: Any = field(default_factory=list)
    εxxDss: Any = field(default_factory=list)
    εyyDss: Any = field(default_factory=list)
    εzzDss: Any = field(default_factory=list)
    εxyDss: Any = field(default_factory=list)
    εxzDss: Any = field(default_factory=list)
    εyzDss: Any = field(default_factory=list)
    εxxDds: Any = field(default_factory=list)
    εyyDds: Any = field(default_factory=list)
    εzzDds: Any = field(default_factory=list)
    εxyDds: Any = field(default_factory=list)
    εxzDds: Any = field(default_factory=list)
    εyzDds: Any = field(default_factory=list)


@dataclass
class DispInf:
    DnUx: Any = field(default_factory=list)
    DnUy: Any = field(default_factory=list)
    DnUz: Any = field(default_factory=list)
    DssUx: Any = field(default_factory=list)
    DssUy: Any = field(default_factory=list)
    DssUz: Any = field(default_factory=list)
    DdsUx: Any = field(default_factory=list)
    DdsUy: Any = field(default_factory=list)
    DdsUz: Any = field(default_factory=list)


STRAIN_FIELDS = [f.name for f in fields(StrainInf)]
DISP_FIELDS = [f.name for f in fields(DispInf)]
TRACTION_FIELDS = [f.name for f in fields(TractionInf)]

# Module-level solver settings (the reference's signatures have no room for them).
_SETTINGS = {"precision": _lib.CDJ_PREC_FP64, "assembly": _lib.CDJ_ASM_FP64, "tile": 0, "tol": 1e-10,
             "max_iter": 0, "near_factor": 2.0}
LAST_REPORT = {}

_MODES = {
    # the reference's arithmetic: FP64 assembly, FP64 LU
    "fp64": (_lib.CDJ_PREC_FP64, _lib.CDJ_ASM_FP64),
    # FP64 assembly; FP32 LU on MFMA as preconditioner of FP64 GMRES-IR against the FP64 matrix
    "mixed32": (_lib.CDJ_PREC_MIXED32, _lib.CDJ_ASM_FP64),
    # FP32 far field / FP64 near field assembly, then as "mixed32"; the residual is against that matrix
    "mixed": (_lib.CDJ_PREC_MIXED32, _lib.CDJ_ASM_MIXED),
    # FP64 assembly; LU with FP16-operand trailing updates as preconditioner of FP64 GMRES-IR
    "mixed16": (_lib.CDJ_PREC_MIXED16, _lib.CDJ_ASM_FP64),
}


def set_precision(mode, near_factor=None):
    """Select the arithmetic of CalculateInfluenceMatrices3D / SlipCalculator3D: 'fp64' (default), 'mixed32',
    'mixed16' or 'mixed' (see _MODES).  near_factor: FP64 radius in longest triangle edges about each source centroid."""
    _SETTINGS["precision"], _SETTINGS["assembly"] = _MODES[mode]
    if near_factor is not None:
        _SETTINGS["near_factor"] = float(near_factor)


def _opts():
    return Opts(precision=_SETTINGS["precision"], tile=_SETTINGS["tile"], tol=_SETTINGS["tol"],
                max_iter=_SETTINGS["max_iter"], assembly=_SETTINGS["assembly"], near_factor=_SETTINGS["near_factor"])


def _tri(P):
    P = np.asarray(P, dtype=np.float64)
    if P.ndim == 1:  # a single triangle given as a 3-vector (TD.jl:80)
        P = P.reshape(1, 3)
    return f64(P)


def _vec(v, n=None):
    v = np.asarray(v, dtype=np.float64).ravel(order="F")
    if n is not None and v.size == 1 and n != 1:
        v = np.full(n, float(v[0]))
    return f64(v)


# ---- TD ------------------------------------------------------------------------------------------------


def TD(X, Y, Z, P1List, P2List, P3List, Dss, Dds, Dn, ν, G, DispFlag, StrainFlag, HSflag,
       StrainInfMat=None, DispInfMat=None):
    """Triangular dislocation strains/displacements (Nikkhoo & Walter 2015) on the GPU.

    Forms (by the type of the last two arguments, as the reference's methods):
      * none                       -> (StrainInf, DispInf) influence matrices, len(X) x ntri each (TD.jl:77)
      * StrainInf, DispInf         -> fills the given matrices in place and returns them (TD.jl:135)
      * Strains, Disps             -> summed fields over all triangles, shaped like X (TD.jl:219)
    Raises CdjError with the reference's messages for mismatched sizes or points above a half-space.
    """
    shape = np.shape(X)
    x, y, z = _vec(X), _vec(Y), _vec(Z)
    P1, P2, P3 = _tri(P1List), _tri(P2List), _tri(P3List)
    n = P1.shape[0]
    dss, dds, dn = _vec(Dss), _vec(Dds), _vec(Dn)
    if not (P2.shape[0] == n and P3.shape[0] == n and dss.size == n and dds.size == n and dn.size == n
            and y.size == x.size and z.size == x.size):
        raise CdjError(_lib.CDJ_ERR_DIMS, "Element size inputs must be the same dimensions")
    m = x.size
    L = load()
    if isinstance(StrainInfMat, Strains) or isinstance(DispInfMat, Disps):
        s6 = [np.zeros(m) for _ in range(6)] if StrainFlag == 1 else None
        d3 = [np.zeros(m) for _ in range(3)] if DispFlag == 1 else None
        check(L.cdj_td_field(ctx(), ptr(x), ptr(y), ptr(z), m, ptr(P1), ptr(P2), ptr(P3), n, ptr(dss), ptr(dds),
                             ptr(dn), float(ν), float(G), int(DispFlag), int(StrainFlag), int(HSflag),
                             ptr_array(s6) if s6 else None, ptr_array(d3) if d3 else None))
        rs = (lambda a: a.reshape(shape, order="F")) if len(shape) > 1 and shape[1] != 1 else (lambda a: a)
        S = Strains(*[rs(a) for a in s6]) if s6 else Strains([], [], [], [], [], [])
        D = Disps(*[rs(a) for a in d3]) if d3 else Disps([], [], [])
        return S, D
    strain = None
    disp = None
    if StrainFlag == 1:
        if isinstance(StrainInfMat, StrainInf) and not isinstance(StrainInfMat.εxxDn, list):
            strain = [getattr(StrainInfMat, k) for k in STRAIN_FIELDS]
            for a in strain:
                if a.shape != (m, n) or not a.flags.f_contiguous or a.dtype != np.float64:
                    raise CdjError(_lib.CDJ_ERR_ARG, "StrainInf matrices must be Float64 %d x %d column-major" % (m, n))
        else:
            strain = [np.zeros((m, n), order="F") for _ in range(18)]
    if DispFlag == 1:
        if isinstance(DispInfMat, DispInf) and not isinstance(DispInfMat.DnUx, list):
            disp = [getattr(DispInfMat, k) for k in DISP_FIELDS]
            for a in disp:
                if a.shape != (m, n) or not a.flags.f_contiguous or a.dtype != np.float64:
                    raise CdjError(_lib.CDJ_ERR_ARG, "DispInf matrices must be Float64 %d x %d column-major" % (m, n))
        else:
            disp = [np.zeros((m, n), order="F") for _ in range(9)]
    check(L.cdj_td_influence(ctx(), ptr(x), ptr(y), ptr(z), m, ptr(P1), ptr(P2), ptr(P3), n, ptr(dss), ptr(dds),
                             ptr(dn), float(ν), float(G), int(DispFlag), int(StrainFlag), int(HSflag),
                             ptr_array(strain) if strain else None, ptr_array(disp) if disp else None))
    S = StrainInf(*strain) if strain else StrainInf()
    D = DispInf(*disp) if disp else DispInf()
    return S, D


def TD_sum(ExxDn, EyyDn, EzzDn, ExyDn, ExzDn, EyzDn,
           ExxDss, EyyDss, EzzDss, ExyDss, ExzDss, EyzDss,
           ExxDds, EyyDds, EzzDds, ExyDds, ExzDds, EyzDds,
           UxDn, UyDn, UzDn, UxDss, UyDss, UzDss, UxDds, UyDds, UzDds):
    """Sums the three slip modes' influence matrices over the elements (TD_sum.jl:1-41): each output is the
    row sum of Dn + Dss + Dds, an (m, 1) column; an empty strain (or displacement) set passes through unchanged, as
    the influence form returns for a flag of 0.  Returns (Exx, Eyy, Ezz, Exy, Exz, Eyz, Ux, Uy, Uz)."""
    def total(a, b, c):
        return np.sum(np.asarray(a) + np.asarray(b) + np.asarray(c), axis=1, keepdims=True)

    E = (ExxDn, EyyDn, EzzDn, ExyDn, ExzDn, EyzDn)
    if np.size(ExxDn) == 0:
        strains = E
    else:
        Es = (ExxDss, EyyDss, EzzDss, ExyDss, ExzDss, EyzDss)
        Ed = (ExxDds, EyyDds, EzzDds, ExyDds, ExzDds, EyzDds)
        strains = tuple(total(a, b, c) for a, b, c in zip(E, Es, Ed))
    if np.size(UxDn) == 0:
        disps = (UxDn, UyDn, UzDn)
    else:
        disps = tuple(total(a, b, c) for a, b, c in ((UxDn, UxDss, UxDds), (UyDn, UyDss, UyDds),
                                                       (UzDn, UzDss, UzDds)))
    return strains + disps


# ---- influence matrices and slip ------------------------------------------------------------------------


def CalculateInfluenceMatrices3D(FaceNormalVector, MidPoint, P1, P2, P3, ν, G, λ, FixedEls, HSFlag, n):
    """Traction influence matrices at the collocation points (CalculateInfluenceMatrices3D.jl:2-48).

    With fixed elements (any FixedEls == 1) the fixed receivers' rows hold displacement influences and the fixed
    sources' columns are dropped (CalculateInfluenceMatrices3D.jl:10-33, 145-198).
    """
    nrm, mid = f64(FaceNormalVector), f64(MidPoint)
    P1, P2, P3 = _tri(P1), _tri(P2), _tri(P3)
    fx = _fixed_flags(FixedEls, n)
    nf = n - (int(fx.sum()) if fx is not None else 0)
    T = [np.zeros((n, nf), order="F") for _ in range(9)]
    check(load().cdj_traction_influence(ctx(), ptr(nrm), ptr(mid), ptr(P1), ptr(P2), ptr(P3), int(n), float(ν),
                                        float(G), float(λ), int(HSFlag), ptr(fx), ptr_array(T),
                                        ctypes.byref(_opts())))
    return TractionInf(*T)


def _fixed_flags(FixedEls, n):
    """FixedEls as the int32 flags of the C ABI (1 = fixed), or None when there are none."""
    if FixedEls is None:
        return None
    fx = np.ascontiguousarray((np.asarray(FixedEls).ravel() == 1).astype(np.int32))
    if fx.shape[0] != n:
        raise CdjError(_lib.CDJ_ERR_DIMS, "Element size inputs must be the same dimensions")
    return fx if fx.any() else None


def SlipCalculator3D(P1, P2, P3, ν, G, λ, MidPoint, FaceNormalVector, HSFlag, BoundaryConditions, FixedEls):
    """Slip on the elements that satisfies the traction boundary conditions (SlipCalculator3D.jl:1-33).

    Returns (Dn, Dss, Dds, A, b) like the reference.  b may also be given directly as a 3N x nrhs array in place
    of BoundaryConditions, in which case Dn/Dss/Dds are N x nrhs.  With MixedBoundaryConditionsFriction the
    frictional method runs instead (SlipCalculator3D.jl:36-84) and, like the reference, returns (Dn, Dss, Dds);
    with MixedBoundaryConditionsFluidVolume the last argument is the FractureFlag vector and the constant-volume
    method runs (SlipCalculator3D.jl:242-475), also returning (Dn, Dss, Dds).
    """
    nrm = f64(FaceNormalVector)
    n = nrm.shape[0]
    if isinstance(BoundaryConditions, MixedBoundaryConditionsFriction):
        return _slip_friction(P1, P2, P3, ν, G, λ, MidPoint, nrm, HSFlag, BoundaryConditions, FixedEls)
    if isinstance(BoundaryConditions, MixedBoundaryConditionsFluidVolume):
        return _slip_fluid_volume(P1, P2, P3, ν, G, λ, MidPoint, nrm, HSFlag, BoundaryConditions, FixedEls)
    if isinstance(BoundaryConditions, np.ndarray):
        b = f64(BoundaryConditions.reshape(3 * n, -1, order="F"))
    else:
        Tn, Tds, Tss = SetupTractionVector(BoundaryConditions, nrm, n, G, λ)
        Tn, Tds, Tss = KnockOutFixedRows(FixedEls, Tn, Tds, Tss)
        b = f64(np.concatenate([Tn, Tss, Tds]).reshape(3 * n, 1))
    fx = _fixed_flags(FixedEls, n)
    nf = n - (int(fx.sum()) if fx is not None else 0)
    P1, P2, P3 = _tri(P1), _tri(P2), _tri(P3)
    mid = f64(MidPoint)
    A = np.zeros((3 * n, 3 * nf), order="F")
    D = np.zeros((3 * nf, b.shape[1]), order="F")
    rep = Report()
    check(load().cdj_slip_calculator(ctx(), ptr(P1), ptr(P2), ptr(P3), n, float(ν), float(G), float(λ), ptr(mid),
                                     ptr(nrm), int(HSFlag), ptr(b), b.shape[1], ptr(fx), ptr(D), ptr(A),
                                     ctypes.byref(_opts()), ctypes.byref(rep)))
    LAST_REPORT.clear()
    LAST_REPORT.update(rep.as_dict())
    squeeze = not isinstance(BoundaryConditions, np.ndarray)
    Dn, Dss, Dds = D[:nf], D[nf:2 * nf], D[2 * nf:3 * nf]
    if squeeze:
        Dn, Dss, Dds, b = Dn[:, 0], Dss[:, 0], Dds[:, 0], b[:, 0]
    return Dn, Dss, Dds, A, b


def _slip_friction(P1, P2, P3, ν, G, λ, MidPoint, nrm, HSFlag, BC, FixedEls):
    """SlipCalculator3D(..., ::MixedBoundaryConditionsFriction, FixedEls) (SlipCalculator3D.jl:36-84)."""
    n = nrm.shape[0]
    if FixedEls is not None and np.any(np.asarray(FixedEls).ravel() == 1):
        raise CdjError(_lib.CDJ_ERR_ARG, "fixed elements are not supported by the frictional solve")
    Tn, Tds, Tss = SetupTractionVector(BC.MixedBoundaryConditions, nrm, n, G, λ)
    b = f64(np.concatenate([Tn, Tss, Tds]))
    mu = _vec(BC.FrictionParameters.µ, n)
    Sf = _vec(BC.FrictionParameters.Sf, n)
    P1, P2, P3 = _tri(P1), _tri(P2), _tri(P3)
    Dn, Dss, Dds = np.zeros(n), np.zeros(n), np.zeros(n)
    rep = Report()
    check(load().cdj_slip_friction(ctx(), ptr(P1), ptr(P2), ptr(P3), n, float(ν), float(G), float(λ),
                                   ptr(f64(MidPoint)), ptr(nrm), int(HSFlag), ptr(b), ptr(mu), ptr(Sf), ptr(Dn),
                                   ptr(Dss), ptr(Dds), ctypes.byref(rep)))
    LAST_REPORT.clear()
    LAST_REPORT.update(rep.as_dict())
    return Dn, Dss, Dds


def _slip_fluid_volume(P1, P2, P3, ν, G, λ, MidPoint, nrm, HSFlag, BC, FractureFlag):
    """SlipCalculator3D(..., ::MixedBoundaryConditionsFluidVolume, FractureFlag) (SlipCalculator3D.jl:242-475).

    The pressure found is left in LAST_REPORT["pressure"]."""
    n = nrm.shape[0]
    Tn, Tds, Tss = SetupTractionVector(BC.MixedBoundaryConditions, nrm, n, G, λ)
    b = f64(np.concatenate([Tn, Tss, Tds]))
    flag = np.ascontiguousarray(np.asarray(FractureFlag).ravel(), dtype=np.int32)
    if flag.shape[0] != n:
        raise CdjError(_lib.CDJ_ERR_DIMS, "Element size inputs must be the same dimensions")
    V = float(np.asarray(BC.Volumes, dtype=np.float64).ravel()[0])
    P1, P2, P3 = _tri(P1), _tri(P2), _tri(P3)
    Dn, Dss, Dds = np.zeros(n), np.zeros(n), np.zeros(n)
    pressure = ctypes.c_double(0.0)
    rep = Report()
    check(load().cdj_slip_fluid_volume(ctx(), ptr(P1), ptr(P2), ptr(P3), n, float(ν), float(G), float(λ),
                                       ptr(f64(MidPoint)), ptr(nrm), int(HSFlag), ptr(b),
                                       flag.ctypes.data_as(ctypes.c_void_p), V, ptr(Dn), ptr(Dss), ptr(Dds),
                                       ctypes.byref(pressure), ctypes.byref(rep)))
    LAST_REPORT.clear()
    LAST_REPORT.update(rep.as_dict())
    LAST_REPORT["pressure"] = pressure.value
    return Dn, Dss, Dds


def solve(A, b):
    """x = A \\ b for square A (blocked LU with partial pivoting on the GPU)."""
    A = f64(A)
    n = A.shape[0]
    b2 = f64(np.asarray(b).reshape(n, -1, order="F"))
    x = np.zeros_like(b2, order="F")
    rep = Report()
    check(load().cdj_solve(ctx(), ptr(A), ptr(b2), n, b2.shape[1], ptr(x), ctypes.byref(_opts()), ctypes.byref(rep)))
    LAST_REPORT.clear()
    LAST_REPORT.update(rep.as_dict())
    if _SETTINGS["precision"] != _lib.CDJ_PREC_FP64 and rep.path == _lib.CDJ_PREC_FP64:
        LAST_REPORT["fallback"] = _lib.last_error()
    return x if np.ndim(b) > 1 else x[:, 0]


# ---- host-side input preparation (the reference's callers) ----------------------------------------------


def ElasticConstantsCheck(G, ν):
    """(K, E, λ, ν, G) from shear modulus and Poisson's ratio (ElasticConstantsCheck.jl:37-49)."""
    K = ((2 * G) * (1 + ν)) / (3 * (1 - (2 * ν)))
    E = (2 * G) * (1 + ν)
    λ = (2 * G * ν) / (1 - (2 * ν))
    return K, E, λ, ν, G


def CreateFaceNormalAndMidPoint(P1, P2, P3):
    """Unit normals (CreateTriangleNormal!.jl:14-30) and incentre midpoints (CreateMidPoint!.jl:3-16)."""
    P1, P2, P3 = (np.asarray(P, dtype=np.float64) for P in (P1, P2, P3))
    U = P2 - P1
    V = P3 - P1
    N = np.stack([U[:, 1] * V[:, 2] - U[:, 2] * V[:, 1], U[:, 2] * V[:, 0] - U[:, 0] * V[:, 2],
                  U[:, 0] * V[:, 1] - U[:, 1] * V[:, 0]], axis=1)
    N = N / np.sqrt(N[:, 0] * N[:, 0] + N[:, 1] * N[:, 1] + N[:, 2] * N[:, 2])[:, None]
    c = np.sqrt(((P1 - P2) ** 2).sum(1))
    a = np.sqrt(((P2 - P3) ** 2).sum(1))
    bb = np.sqrt(((P1 - P3) ** 2).sum(1))
    if np.any(a == 0) or np.any(bb == 0) or np.any(c == 0):
        raise CdjError(_lib.CDJ_ERR_DEGENERATE, "Slither triangles or triangles with no area exist on your surface")
    M = (P1 * a[:, None] + P2 * bb[:, None] + P3 * c[:, None]) / (a + bb + c)[:, None]
    return f64(N), f64(M)


def AreaOfTriangle3D(x1, y1, z1, x2, y2, z2, x3, y3, z3):
    """Triangle areas by Heron's formula and half perimeters (AreaOfTriangle3D.jl:52-131): side lengths P1-P2,
    P2-P3, P3-P1; a triangle whose Heron product rounds below zero gets NaN when its sides break the triangle
    inequality, else 0."""
    x1, y1, z1, x2, y2, z2, x3, y3, z3 = (np.atleast_1d(np.asarray(v, dtype=np.float64))
                                          for v in (x1, y1, z1, x2, y2, z2, x3, y3, z3))
    a = np.sqrt((x1 - x2) ** 2 + (y1 - y2) ** 2 + (z1 - z2) ** 2)
    b = np.sqrt((x2 - x3) ** 2 + (y2 - y3) ** 2 + (z2 - z3) ** 2)
    c = np.sqrt((x3 - x1) ** 2 + (y3 - y1) ** 2 + (z3 - z1) ** 2)
    hp = (a + b + c) / 2.0
    prod = hp * (hp - a) * (hp - b) * (hp - c)
    area = np.sqrt(np.maximum(prod, 0.0))
    bad = (prod < 0) & ((a + b < c) | (b + c < a) | (a + c < b))
    area[bad] = np.nan
    return area, hp


def GetDesiredEdgeLength(P1, P2, P3, NoTris):
    """Edge length of the equilateral triangles that tile the surface's area in NoTris triangles, and 2/3 of the
    largest half perimeter (GetDesiredEdgeLength.jl:1-12): the target passed to the remesher."""
    P1, P2, P3 = (np.atleast_2d(np.asarray(P, dtype=np.float64)) for P in (P1, P2, P3))
    area, hp = AreaOfTriangle3D(P1[:, 0], P1[:, 1], P1[:, 2], P2[:, 0], P2[:, 1], P2[:, 2], P3[:, 0], P3[:, 1],
                                P3[:, 2])
    desired = np.sum(area) / NoTris
    return float(np.sqrt(desired / (np.sqrt(3) / 4))), float(np.max(hp) * (2 / 3))


def CalculateSSandDSDirs(CosAx, CosAy, CosAz):
    """Strike-slip and dip-slip direction cosines of planes with normal (CosAx, CosAy, CosAz)
    (CalculateSSandDSDirs.jl:37-98)."""
    CosAx, CosAy, CosAz = (np.atleast_1d(np.asarray(v, dtype=np.float64)) for v in (CosAx, CosAy, CosAz))
    with np.errstate(invalid="ignore", divide="ignore"):
        h = np.sqrt(CosAx ** 2 + CosAy ** 2)
        ss = np.stack([-(CosAy / h), CosAx / h, np.zeros_like(CosAx)], axis=1)
    DSCosAz = np.cos(np.arccos(CosAz) - np.pi / 2)
    L = np.sin(np.arccos(DSCosAz))
    L = np.where(CosAz < 0, -L, L)
    a = np.arctan2(CosAy, CosAx)
    ds = np.stack([-(np.cos(a) * L), -(np.sin(a) * L), DSCosAz], axis=1)
    flat = (CosAx == 0) & (CosAy == 0)
    ds[flat] = [-1.0, 0.0, 0.0]
    ss[flat & (CosAz < 1.0)] = [0.0, -1.0, 0.0]
    ss[flat & ~(CosAz < 1.0)] = [0.0, 1.0, 0.0]
    return ss, ds


def CalculateNormalAndShearTractions3D(σxx, σyy, σzz, σxy, σxz, σyz, FaceNormalVector):
    """(Tn, Tds, Tss) on planes of the given normals (CalculateNormalAndShearTractions3D.jl:37-58)."""
    nrm = np.atleast_2d(np.asarray(FaceNormalVector, dtype=np.float64))
    nx, ny, nz = nrm[:, 0], nrm[:, 1], nrm[:, 2]
    ss, ds = CalculateSSandDSDirs(nx, ny, nz)
    Sxx, Syy, Szz, Sxy, Sxz, Syz = (np.broadcast_to(np.asarray(v, dtype=np.float64), nx.shape)
                                    for v in (σxx, σyy, σzz, σxy, σxz, σyz))
    Tn = Sxx * nx ** 2 + Syy * ny ** 2 + Szz * nz ** 2 + 2 * Sxy * nx * ny + 2 * Syz * ny * nz + 2 * Sxz * nz * nx
    Tx = Sxx * nx + Sxy * ny + Sxz * nz
    Ty = Sxy * nx + Syy * ny + Syz * nz
    Tz = Sxz * nx + Syz * ny + Szz * nz

    def along(d):
        A = (Tx * (1 - nx ** 2) - Ty * (nx * ny) - Tz * (nz * nx)) * d[:, 0]
        B = (Tx * -(nx * ny) + Ty * (1 - ny ** 2) - Tz * (ny * nz)) * d[:, 1]
        C = (Tx * -(nz * nx) - Ty * (ny * nz) + Tz * (1 - nz ** 2)) * d[:, 2]
        return A + B + C

    return Tn, along(ds), along(ss)


def SetupTractionVector(BC, FaceNormalVector, n, G, λ):
    """Boundary-condition tractions (Tn, Tds, Tss) on every element (SetupTractionVector.jl:2-41)."""
    if isinstance(BC, Stresses):
        return CalculateNormalAndShearTractions3D(BC.σxx, BC.σyy, BC.σzz, BC.σxy, BC.σxz, BC.σyz, FaceNormalVector)
    if isinstance(BC, Tractions):
        return (np.asarray(BC.Tn, dtype=np.float64) * np.ones(n), np.asarray(BC.Tds, dtype=np.float64) * np.ones(n),
                np.asarray(BC.Tss, dtype=np.float64) * np.ones(n))
    if isinstance(BC, MixedBoundaryConditions):
        sTn, sTds, sTss = SetupTractionVector(BC.Stresses, FaceNormalVector, n, G, λ)
        Tn, Tds, Tss = SetupTractionVector(BC.Tractions, FaceNormalVector, n, G, λ)
        return Tn + sTn, Tds + sTds, Tss + sTss
    raise TypeError("unsupported boundary condition type %r" % type(BC))


def KnockOutFixedRows(FixedEls, Tn, Tds, Tss):
    """Zero the tractions of fixed elements (KnockOutFixedRows.jl:1-17)."""
    Tn, Tds, Tss = (np.array(v, dtype=np.float64) for v in (Tn, Tds, Tss))
    if FixedEls is not None:
        f = np.asarray(FixedEls).ravel() == 1
        Tn[f] = 0.0
        Tds[f] = 0.0
        Tss[f] = 0.0
    return Tn, Tds, Tss


# ---- batched slip solves (many meshes, several remote loads each) -------------------------------------------


# ---- crack-tip edges and stress intensities (GetCrackTipElements3D.jl, StressIntensity3D.jl) --------------

_CT_GEOM = 13  # FeLe, FeMd xyz, FeEv xyz, FeM2Ev xyz, FeM2ELe, IntAng, Area
_CT_K = 4      # K1, K2, K3, StrainEnergy


@dataclass
class TriangleEdges:
    """TriangleEdges of CutAndDisplaceJulia.jl:57: one edge (P1P2, P1P3 or P2P3) of every triangle; NaN where the
    edge is not a free (crack-tip) edge, except Area."""
    FeLe: Any = None
    FeMd: Any = None
    FeEv: Any = None
    FeM2Ev: Any = None
    FreeFlg: Any = None
    FeM2ELe: Any = None
    IntAng: Any = None
    K1: Any = 0.0
    K2: Any = 0.0
    K3: Any = 0.0
    StrainEnergy: Any = 0.0
    Area: Any = None


def GetCrackTipElements3D(MidPoint, P1, P2, P3, FaceNormalVector):
    """Free edges of the mesh and their geometry -> (FeP1P2S, FeP1P3S, FeP2P3S) (GetCrackTipElements3D.jl:1)."""
    P1, P2, P3 = _tri(P1), _tri(P2), _tri(P3)
    mid, nrm = _tri(MidPoint), _tri(FaceNormalVector)
    n = P1.shape[0]
    if not (P2.shape[0] == P3.shape[0] == mid.shape[0] == nrm.shape[0] == n):
        raise CdjError(_lib.CDJ_ERR_DIMS, "Element size inputs must be the same dimensions")
    flags = np.zeros(3 * n, dtype=np.uint8)
    geom = np.zeros(3 * n * _CT_GEOM)
    check(load().cdj_crack_tip_elements(ctx(), ptr(mid), ptr(P1), ptr(P2), ptr(P3), ptr(nrm), n,
                                        flags.ctypes.data, ptr(geom)))
    out = []
    for s_ in range(3):
        g = geom[s_ * n * _CT_GEOM:(s_ + 1) * n * _CT_GEOM].reshape(_CT_GEOM, n).T
        out.append(TriangleEdges(FeLe=g[:, 0].copy(), FeMd=g[:, 1:4].copy(), FeEv=g[:, 4:7].copy(),
                                 FeM2Ev=g[:, 7:10].copy(), FreeFlg=flags[s_ * n:(s_ + 1) * n].astype(bool),
                                 FeM2ELe=g[:, 10].copy(), IntAng=g[:, 11].copy(), Area=g[:, 12].copy()))
    return tuple(out)


def StressIntensity3D(Dn, Dss, Dds, G, ν, FaceNormalVector, FeP1P2S, FeP1P3S, FeP2P3S):
    """K1, K2, K3 and StrainEnergy on the free edges, from the slip of the tip triangles; returns the three
    TriangleEdges with those fields filled (StressIntensity3D.jl:1)."""
    nrm = _tri(FaceNormalVector)
    n = nrm.shape[0]
    Dn, Dss, Dds = _vec(Dn, n), _vec(Dss, n), _vec(Dds, n)
    edges = (FeP1P2S, FeP1P3S, FeP2P3S)
    flags = np.concatenate([np.asarray(e.FreeFlg, dtype=np.uint8).ravel() for e in edges])
    geom = np.concatenate([np.column_stack([e.FeLe, e.FeMd, e.FeEv, e.FeM2Ev, e.FeM2ELe, e.IntAng, e.Area])
                           .ravel(order="F") for e in edges])
    if flags.size != 3 * n or geom.size != 3 * n * _CT_GEOM:
        raise CdjError(_lib.CDJ_ERR_DIMS, "Element size inputs must be the same dimensions")
    geom = f64(geom)
    K = np.zeros(3 * n * _CT_K)
    check(load().cdj_stress_intensity(ctx(), ptr(Dn), ptr(Dss), ptr(Dds), n, float(G), float(ν), ptr(nrm),
                                      flags.ctypes.data, ptr(geom), ptr(K)))
    out = []
    for s_, e in enumerate(edges):
        k = K[s_ * n * _CT_K:(s_ + 1) * n * _CT_K].reshape(_CT_K, n)
        out.append(TriangleEdges(FeLe=e.FeLe, FeMd=e.FeMd, FeEv=e.FeEv, FeM2Ev=e.FeM2Ev, FreeFlg=e.FreeFlg,
                                 FeM2ELe=e.FeM2ELe, IntAng=e.IntAng, K1=k[0].copy(), K2=k[1].copy(),
                                 K3=k[2].copy(), StrainEnergy=k[3].copy(), Area=e.Area))
    return tuple(out)


class _Mesh(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("P1", ctypes.c_void_p), ("P2", ctypes.c_void_p), ("P3", ctypes.c_void_p),
                ("normals", ctypes.c_void_p), ("midpoints", ctypes.c_void_p), ("nrhs", ctypes.c_int64),
                ("b", ctypes.c_void_p), ("x", ctypes.c_void_p)]


def batched_slip(meshes, nu, G, lam, hs_flag=0):
    """Slip of every mesh under each of its remote loads.

    meshes: list of dicts with P1, P2, P3 (n x 3) and `stresses` (k x 6 remote tensors sxx, syy, szz, sxy, sxz,
    syz), or `b` (3n x k right-hand sides [Tn; Tss; Tds]).  Returns a list of 3n x k arrays [Dn; Dss; Dds] and the
    total report (assembly time of the batched graph, summed solve time, near entries; "meshes": the per-mesh
    reports; "fallbacks": meshes of a mixed-precision batch that fell back to the FP64 LU).  The precision is the
    module setting (set_precision)."""
    L = load()
    keep = []
    arr = (_Mesh * len(meshes))()
    outs = []
    for i, m in enumerate(meshes):
        P1, P2, P3 = _tri(m["P1"]), _tri(m["P2"]), _tri(m["P3"])
        n = P1.shape[0]
        nrm, mid = CreateFaceNormalAndMidPoint(P1, P2, P3)
        if "b" in m:
            b = f64(np.asarray(m["b"]).reshape(3 * n, -1, order="F"))
        else:
            cols = []
            for sxx, syy, szz, sxy, sxz, syz in np.atleast_2d(m["stresses"]):
                Tn, Tds, Tss = SetupTractionVector(Stresses(sxx, syy, szz, sxy, sxz, syz), nrm, n, G, lam)
                cols.append(np.concatenate([Tn, Tss, Tds]))
            b = f64(np.stack(cols, axis=1))
        x = np.zeros_like(b, order="F")
        nrm, mid = f64(nrm), f64(mid)
        keep += [P1, P2, P3, nrm, mid, b, x]
        arr[i] = _Mesh(n, P1.ctypes.data, P2.ctypes.data, P3.ctypes.data, nrm.ctypes.data, mid.ctypes.data,
                       b.shape[1], b.ctypes.data, x.ctypes.data)
        outs.append(x)
    rep = Report()
    per = (Report * len(meshes))()
    check(L.cdj_batched_slip(ctx(), ctypes.byref(arr), len(meshes), float(nu), float(G), float(lam), int(hs_flag),
                             ctypes.byref(_opts()), per, ctypes.byref(rep)))
    tpb, wgs = ctypes.c_int32(), ctypes.c_int32()
    check(L.cdj_batch_tuning(ctx(), ctypes.byref(tpb), ctypes.byref(wgs)))
    r = rep.as_dict()
    # per-mesh path, iterations and residual: a mesh that fell back to the FP64 LU reports path FP64, residual -1
    r["meshes"] = [per[i].as_dict() for i in range(len(meshes))]
    r["fallbacks"] = sum(1 for q in r["meshes"] if q["path"] == _lib.CDJ_PREC_FP64) \
        if _SETTINGS["precision"] != _lib.CDJ_PREC_FP64 else 0
    r["tile"] = {"sources_per_item": tpb.value, "workgroups": wgs.value}
    LAST_REPORT.clear()
    LAST_REPORT.update(r)
    return outs, r
