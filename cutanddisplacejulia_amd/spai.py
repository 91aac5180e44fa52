"""Training and evaluation of the learned SPAI preconditioner on the penny-crack growth sequence.

The sequence: a radius-1 penny crack remeshed at 40 growth steps from 500 to 6400 triangles
(meshes.growth_sequence).  Each step's remote load is sigma_zz = 1 + g z, with g ~ U(0, 0.5).  The network is
trained on steps 1-30 with the project's own GPU kernels (cdj_mlp_train_dev).  The training pairs (x, b = A x)
come from the method's own assembly.  Steps 31-40 are held out for evaluation.

    python -m cutanddisplacejulia_amd.spai train [--out PATH] [--epochs E]
    python -m cutanddisplacejulia_amd.spai eval  [--params PATH] [--steps 31-40]

Device memory comes from torch tensors; every computation runs in libcdjhip.so.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import api, meshes
from . import learned

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PARAMS = os.path.join(HERE, "data", "spai_k16_h128.npy")
NU, G = 0.25, 1.0


def _torch():
    import torch
    return torch


def assemble(step, mixed=False, near_factor=2.0, hs=0):
    """Device geometry and the assembled 3N x 3N matrix (column-major, stored as torch [col, row]) of one step."""
    torch = _torch()
    P1, P2, P3 = step["P1"], step["P2"], step["P3"]
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G_ = api.ElasticConstantsCheck(G, NU)
    d = {k: learned.to_device_cols(v) for k, v in dict(P1=P1, P2=P2, P3=P3, N=nrm, M=mid).items()}
    A = torch.empty((3 * n, 3 * n), dtype=torch.float64, device="cuda:0")
    learned.assemble_dev(d["N"], d["M"], d["P1"], d["P2"], d["P3"], n, nu, G_, lam, hs, A, 3 * n, mixed=mixed,
                         near_factor=near_factor)
    return d, A, nrm, mid, lam


def rhs(step, nrm, mid, lam):
    """b = [Tn; Tss; Tds] for sigma_zz = 1 + g z at the collocation points (SetupTractionVector.jl:2-41)."""
    n = nrm.shape[0]
    szz = 1.0 + step["g"] * mid[:, 2]
    z = np.zeros(n)
    Tn, Tds, Tss = api.SetupTractionVector(api.Stresses(z, z, szz, z, z, z), nrm, n, G, lam)
    return np.concatenate([Tn, Tss, Tds])


def train(seq, epochs=8, k=16, hidden=128, seed=5, nsamples=16, iters=10, lr=1e-3, log=print):
    """Adam on steps `seq` (list of growth steps) with matrices assembled once and kept in HBM."""
    torch = _torch()
    model = learned.LearnedPreconditioner(k=k, hidden=hidden, seed=seed)
    cache = []
    for s in seq:
        d, A, _, _, _ = assemble(s)
        cache.append((s, d, A))
    torch.cuda.synchronize()
    losses = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        ep_loss = []
        for s, d, A in cache:
            n = s["P1"].shape[0]
            ep_loss.append(model.train_dev(d["P1"], d["P2"], d["P3"], n, G, A, 3 * n, nsamples=nsamples, iters=iters,
                                           lr=lr, seed=1000 * ep + s["step"]))
        losses.append(float(np.mean(ep_loss)))
        log("epoch %d mean loss %.4e (%.1f s)" % (ep, losses[-1], time.perf_counter() - t0))
    return model, losses


def _timed(fn):
    torch = _torch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, 1e3 * (time.perf_counter() - t0)


PCS = ("none", "jacobi", "learned", "learned_k32", "learned_bf16", "lu32", "lu16")


def evaluate_step(step, model, hs=0, pcs=PCS, tol=1e-10):
    """Solve one step with the plain FP64 path (assembly + LU) and the mixed / preconditioned paths.

    Returns a dict of wall times (ms, device work synchronised), iterations, FP64 residuals and forward errors
    against the FP64 LU solution.  "residual" is always ||b - A x|| / ||b|| with the FP64-assembled A; rows of
    the mixed assembly also give "residual_assembled", against the FP32-far / FP64-near matrix they solved."""
    import ctypes
    torch = _torch()
    from . import _lib
    L = _lib.load()
    ctx = _lib.ctx()
    n = step["P1"].shape[0]
    N3 = 3 * n
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {"step": step["step"], "n_tri": n, "N": N3}
    # plain FP64: assembly + LU on [A | b]
    (d, A, nrm, mid, lam), t_asm = _timed(lambda: assemble(step, hs=hs))
    b = torch.from_numpy(rhs(step, nrm, mid, lam)).to("cuda:0")
    A64 = A.clone()

    def lu():
        ipiv = torch.empty(N3, dtype=torch.int32, device="cuda:0")
        _lib.check(L.cdj_lu_factor_dev(ctx, ctypes.c_void_p(A.data_ptr()), N3, N3, ctypes.c_void_p(ipiv.data_ptr()),
                                       sp))
        x = b.clone()
        _lib.check(L.cdj_lu_solve_dev(ctx, ctypes.c_void_p(A.data_ptr()), N3, N3, ctypes.c_void_p(ipiv.data_ptr()),
                                      ctypes.c_void_p(x.data_ptr()), 1, N3, sp))
        return x

    x64, t_lu = _timed(lu)
    nb = torch.linalg.norm(b)
    res64 = float(torch.linalg.norm(b - torch.mv(A64.t(), x64)) / nb)
    out["fp64"] = {"assembly_ms": t_asm, "solve_ms": t_lu, "total_ms": t_asm + t_lu, "residual": res64}
    del A
    xref = x64
    Asolve = A64
    for mixed in (False, True):
        if mixed:
            (d, Asolve, nrm, mid, lam), t_asm = _timed(lambda: assemble(step, mixed=True, hs=hs))
            out["near_entries"] = int(L.cdj_last_near_entries(ctx))
            out["mixed_assembly_ms"] = t_asm
            out["mixed_matrix_diff"] = float(torch.max(torch.abs(Asolve - A64)) / torch.max(torch.abs(A64)))
        for name in pcs:
            pc, k32 = (name[:-4], True) if name.endswith("_k32") else (name, False)
            if pc.startswith("learned") and model is None:
                continue
            x = torch.zeros_like(b)
            rep, t = _timed(lambda: learned.precond_solve_dev(Asolve, N3, b, x, d["P1"], d["P2"], d["P3"], n, G,
                                                              pc=pc, model=model, tol=tol, krylov32=k32))
            res = float(torch.linalg.norm(b - torch.mv(A64.t(), x)) / nb)
            fwd = float(torch.linalg.norm(x - xref) / torch.linalg.norm(xref))
            key = ("mixed_" if mixed else "") + name
            out[key] = {"solve_ms": t, "iterations": rep["iterations"], "path": rep["path"], "residual": res,
                        "forward_err_vs_fp64": fwd,
                        "total_ms": t + (out["mixed_assembly_ms"] if mixed else out["fp64"]["assembly_ms"])}
            if mixed:
                out[key]["residual_assembled"] = float(torch.linalg.norm(b - torch.mv(Asolve.t(), x)) / nb)
    return out


def save_params(model, path):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, model.get_params())


def load_model(path=DEFAULT_PARAMS, k=16, hidden=128):
    p = np.load(path, allow_pickle=False)
    model = learned.LearnedPreconditioner(k=k, hidden=hidden, seed=0)
    model.set_params(p)
    return model


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--out", default=DEFAULT_PARAMS)
    ap.add_argument("--params", default=DEFAULT_PARAMS)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--first", type=int, default=31)
    ap.add_argument("--last", type=int, default=40)
    ap.add_argument("--hs", type=int, default=0)
    args = ap.parse_args(argv)
    seq = meshes.growth_sequence(40, 500, 6400)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if args.cmd == "train":
        model, losses = train(seq[:30], epochs=args.epochs, log=log)
        save_params(model, args.out)
        log(json.dumps({"saved": args.out, "losses": losses}))
    else:
        model = load_model(args.params) if os.path.exists(args.params) else None
        for s in seq[args.first - 1:args.last]:
            log(json.dumps(evaluate_step(s, model, hs=args.hs)))


if __name__ == "__main__":
    sys.exit(main())
