"""Learned sparse-approximate-inverse preconditioner (GPU training and inference through the C ABI).

Device arrays are torch tensors (torch is used only for device memory and the stream); the network, its
training kernels and the solver are in libcdjhip.so.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import CdjError, Opts, Report, check, ctx, f64, load, ptr

PC = {"none": 0, "jacobi": 1, "learned": 2, "lu32": 3, "lu16": 3, "learned_bf16": 4}

_p = ctypes.c_void_p
_d = ctypes.c_double
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32


def _bind():
    L = load()
    if getattr(L, "_learned_bound", False):
        return L
    L.cdj_mlp_create.argtypes = [_p, _i32, _i32, ctypes.c_uint64, ctypes.POINTER(_p)]
    L.cdj_mlp_create.restype = ctypes.c_int
    L.cdj_mlp_destroy.argtypes = [_p]
    L.cdj_mlp_destroy.restype = ctypes.c_int
    L.cdj_mlp_nparams.argtypes = [_p]
    L.cdj_mlp_nparams.restype = ctypes.c_int64
    L.cdj_mlp_get_params.argtypes = [_p, _p, _p]
    L.cdj_mlp_get_params.restype = ctypes.c_int
    L.cdj_mlp_set_params.argtypes = [_p, _p, _p]
    L.cdj_mlp_set_params.restype = ctypes.c_int
    L.cdj_mlp_train_dev.argtypes = [_p, _p, _p, _p, _p, _i64, _d, _p, _i64, _i32, _i32, ctypes.c_float,
                                    ctypes.c_uint64, ctypes.POINTER(_d), _p]
    L.cdj_mlp_train_dev.restype = ctypes.c_int
    L.cdj_mlp_loss_grad.argtypes = [_p, _p, _p, _p, _p, _i64, _d, _p, _p, _p, _i32, ctypes.POINTER(_d), _p, _p, _p, _p]
    L.cdj_mlp_loss_grad.restype = ctypes.c_int
    L.cdj_precond_solve_dev.argtypes = [_p, _p, _i32, _p, _p, _p, _i64, _d, _p, _i64, _p, _p, _p, _p, _p]
    L.cdj_precond_solve_dev.restype = ctypes.c_int
    L._learned_bound = True
    return L


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def to_device_cols(a):
    """N x 3 (or any) numpy array -> column-major flattened FP64 torch tensor on cuda:0."""
    import torch
    return torch.from_numpy(np.asfortranarray(np.asarray(a, dtype=np.float64)).ravel(order="F").copy()).to("cuda:0")


class LearnedPreconditioner:
    """Per-triangle MLP (k neighbours, 2 hidden tanh layers) predicting SPAI blocks of A^-1."""

    def __init__(self, k=16, hidden=128, seed=1):
        L = _bind()
        h = _p()
        check(L.cdj_mlp_create(ctx(), int(k), int(hidden), int(seed), ctypes.byref(h)))
        self.h = h
        self.k = k
        self.hidden = hidden

    def __del__(self):
        try:
            if self.h:
                _bind().cdj_mlp_destroy(self.h)
        except Exception:
            pass

    @property
    def nparams(self):
        return int(_bind().cdj_mlp_nparams(self.h))

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        check(_bind().cdj_mlp_get_params(ctx(), self.h, p.ctypes.data_as(_p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        check(_bind().cdj_mlp_set_params(ctx(), self.h, p.ctypes.data_as(_p)))

    def train_dev(self, dP1, dP2, dP3, n, G, dA, lda, nsamples=16, iters=20, lr=1e-3, seed=0):
        """Adam steps on one mesh whose device geometry and assembled A (3n x 3n) are given; returns the loss."""
        loss = _d()
        check(_bind().cdj_mlp_train_dev(ctx(), self.h, _dptr(dP1), _dptr(dP2), _dptr(dP3), int(n), float(G),
                                        _dptr(dA), int(lda), int(nsamples), int(iters), float(lr), int(seed),
                                        ctypes.byref(loss), _stream()))
        return loss.value

    def forward(self, P1, P2, P3, bf16=False):
        """The network's n x O outputs for a mesh (host arrays), in FP32 or with BF16 weights and activations."""
        P1, P2, P3 = f64(P1), f64(P2), f64(P3)
        n = P1.shape[0]
        Y = np.zeros((n, 9 * self.k), dtype=np.float32)
        check(_bind().cdj_mlp_forward(ctx(), self.h, ptr(P1), ptr(P2), ptr(P3), n, int(bool(bf16)),
                                      Y.ctypes.data_as(_p)))
        return Y

    def loss_grad(self, P1, P2, P3, G, jac, B, X):
        """Loss, gradient and the GPU's features / neighbours / geometry records for host inputs."""
        P1, P2, P3 = f64(P1), f64(P2), f64(P3)
        n = P1.shape[0]
        jac = np.ascontiguousarray(jac, dtype=np.float64)
        B, X = f64(B), f64(X)
        S = B.shape[1]
        loss = _d()
        grad = np.zeros(self.nparams, dtype=np.float32)
        feat = np.zeros((n, 10 * self.k), dtype=np.float32)
        nbr = np.zeros((n, self.k), dtype=np.int32)
        geo = np.zeros((n, 16), dtype=np.float32)
        check(_bind().cdj_mlp_loss_grad(ctx(), self.h, ptr(P1), ptr(P2), ptr(P3), n, float(G), ptr(jac), ptr(B),
                                        ptr(X), S, ctypes.byref(loss), grad.ctypes.data_as(_p),
                                        feat.ctypes.data_as(_p), nbr.ctypes.data_as(_p), geo.ctypes.data_as(_p)))
        return loss.value, grad, feat, nbr, geo


def precond_solve_dev(dA, lda, db, dx, dP1, dP2, dP3, n, G, pc="learned", model=None, tol=1e-10, max_iter=0,
                      krylov32=False):
    """FGMRES on device A (3n x 3n) with the named preconditioner to an FP64 relative residual `tol`; returns the
    report dict.  krylov32: Arnoldi products on an FP32 copy of A, FP64 residual at every restart."""
    rep = Report()
    prec = _lib.CDJ_PREC_MIXED32 if krylov32 else _lib.CDJ_PREC_LEARNED
    if pc == "lu16":  # LU with FP16-operand trailing updates as the refinement's preconditioner
        prec = _lib.CDJ_PREC_MIXED16
    opts = Opts(precision=prec, tol=float(tol), max_iter=int(max_iter), near_factor=2.0)
    check(_bind().cdj_precond_solve_dev(ctx(), model.h if model is not None else None, PC[pc], _dptr(dP1),
                                        _dptr(dP2), _dptr(dP3), int(n), float(G), _dptr(dA), int(lda), _dptr(db),
                                        _dptr(dx), ctypes.byref(opts), ctypes.byref(rep), _stream()))
    return rep.as_dict()


def assemble_dev(dN, dM, dP1, dP2, dP3, n, nu, G, lam, hs, dA, lda, tile=0, mixed=False, near_factor=2.0):
    """Assemble A (3n x 3n, column-major, leading dimension lda) into the device tensor dA."""
    opts = Opts(tile=int(tile), tol=1e-10, assembly=_lib.CDJ_ASM_MIXED if mixed else _lib.CDJ_ASM_FP64,
                near_factor=float(near_factor))
    check(load().cdj_assemble_dev(ctx(), _dptr(dN), _dptr(dM), _dptr(dP1), _dptr(dP2), _dptr(dP3), int(n),
                                  float(nu), float(G), float(lam), int(hs), _dptr(dA), int(lda), ctypes.byref(opts),
                                  _stream()))
