#!/usr/bin/env python
"""Benchmark of the triangular-dislocation hot path on one MI355X.

Workload (the half-space configuration the headline metric is quoted on): a 45-degree dipping rectangular fault
(strike 60, 5 x 5, top at depth 0.5) meshed into 50 x 50 quads = 5000 triangles; one step assembles the full
15000 x 15000 FP64 collocation matrix A (CalculateInfluenceMatrices3D + the block concatenation of
SlipCalculator3D) from HBM-resident geometry.  value = A entries per second over the timed steps.

Also reported: the FP64 LU solve of the same A (`A\\b`), the dominant kernel's roofline position (algorithmic
FP64 flops from the oracle's operation count per receiver/source pair) and the CPU baseline (the FP64 C++
restatement of TD.jl, OpenMP over source triangles like TD.jl:146) on a bounded sample of receiver rows.

python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-batched] [--no-large] [--friction] [--crack-tip]

The 64-crack batched assembly + solve and the 16224-triangle elliptical crack (A 48672^2 FP64: assembly, three
solvers to 1e-10, grid field) run by default (--no-batched, --no-large skip them); --friction and --crack-tip add
their cases; all are reported in the same JSON line.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TD influence entries/s; slip-solve time at 1e-10 resid, mixed/learned vs FP64"
# Algorithmic work of one (receiver, source) pair of the half-space traction assembly: +,-,*,/ of the FP64
# oracle instantiated with an operation-counting scalar, averaged over 4000 random pairs of this workload
# (tests/test_flops.py recomputes it).  Transcendentals (12 sqrt, 30 trig in set-up) are not included.  The kernel
# runs on the FP64 vector ALUs (the bound is VALU issue, 78.6 TFLOP/s on MI355X); "executed_frac" is the fraction
# of that peak in the FP64 operations the kernel actually issues (below).
FLOPS_PER_PAIR_HS = 43940.0
# Work of one (receiver, mesh edge) unit of the dominant kernel, tdk_edge_response: the free-surface strain of the
# edge's angular-dislocation pair for three unit Burgers vectors, turned into 9 tractions.  The reference evaluates
# it once per triangle side and pair; the engine once per mesh edge and receiver (7600 edges for the 15000 sides of
# this mesh), and each triangle then adds +- the responses of its edges (tdk_traction_combine).
#  * FLOPS_PER_EDGE_HS, the algorithmic work: +,-,*,/ of the engine's formulation (the forward-mode derivative of
#    the free-surface displacement, every shared quantity formed once), counted by compiling the device header for
#    the host with a counting scalar that leaves out what the compiler folds and what never reaches the unit's
#    outputs: only the gradients of the displacement are needed, so its values, logarithms and arctangents are
#    dead (tests/flops/edge_count.cpp; tests/test_flops.py recomputes it).  2 sqrt per unit are not included.
#  * KERNEL_FLOPS_PER_EDGE_HS: the same count for the kernel as written, which forms the shared quantities once
#    per Burgers component (three times) so that one component fits two waves per SIMD.
#  * REF_FLOPS_PER_EDGE_HS: the reference formulation of the same unit (the strain closed form AngDisStrainFSC for
#    one side and three slip modes), from the op-counting oracle over 6000 random (receiver, side) pairs of this
#    workload (tests/test_flops.py recomputes it).
FLOPS_PER_EDGE_HS = 4086.0
KERNEL_FLOPS_PER_EDGE_HS = 4536.0
REF_FLOPS_PER_EDGE_HS = 12428.0
# FP64 flops tdk_edge_response executes per (receiver, edge) on the same workload, from rocprofv3 --pmc
# SQ_INSTS_VALU_ADD_F64, _MUL_F64, _FMA_F64, _TRANS_F64 (x64 lanes, FMA twice): 1.740e11 per launch over
# 5000 x 7600 units.  It differs from KERNEL_FLOPS_PER_EDGE_HS by the Newton and Goldschmidt steps of the
# reciprocals and square roots (+), the compiler's common subexpressions within a unit (-) and the 56 idle
# lanes of the last receiver block (+).
EXECUTED_FLOPS_PER_EDGE_HS = 4579.7
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of tdk_edge_response on the 5000-triangle workload: rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE in separate runs, each scaled by what the counter reads for a 1 GiB device copy (fetch x2, write x1):
# 0.074 GB read + 2.975 GB written, against its algorithmic 72 B per unit (2.736 GB written); the excess is the
# write-back of 20 B per lane of register spill.
TRAFFIC_BYTES_EDGE_HS_5000 = 3.049e9


def _log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_baseline(P1, P2, P3, nrm, mid, nu, G, lam, rows, threads, reps=5):
    """The oracle (FP64 C++ port of TD.jl, OpenMP over source triangles) on `rows` receiver rows against all
    sources: median entries/s over `reps` timed runs after one warm-up, and the rows themselves (TractionInf
    blocks, rows x n) for the parity check of the GPU matrix."""
    from oracle import oracle as O
    n = P1.shape[0]
    O.traction_influence_rows(nrm, mid, P1, P2, P3, nu, G, lam, 1, 0, 8, nthreads=threads)  # warm-up
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        T = O.traction_influence_rows(nrm, mid, P1, P2, P3, nu, G, lam, 1, 0, rows, nthreads=threads)
        ts.append(time.perf_counter() - t0)
    dt = float(np.median(ts))
    return 9.0 * rows * n / dt, dt, float(np.sum(ts)), T


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def parity_rows(A_rows, T, n):
    """Largest |GPU - oracle| over the sampled rows relative to the largest oracle entry of the same slip mode (its
    three traction blocks; the blocks coupling opening to shear and shear to normal traction are ~1e-6 of that scale
    on a planar fault).  A_rows: the GPU matrix's rows of those receivers in its three row blocks, as
    {block q: (rows x 3n) array}."""
    names = ("DnTn", "DnTss", "DnTds", "DssTn", "DssTss", "DssTds", "DdsTn", "DdsTss", "DdsTds")
    worst = 0.0
    for m in range(3):
        scale = max(float(np.max(np.abs(T[names[3 * m + q]]))) for q in range(3))
        for q in range(3):
            g = -A_rows[q][:, m * n:(m + 1) * n]
            worst = max(worst, float(np.max(np.abs(g - T[names[3 * m + q]]))) / scale)
    return worst


def rocsolver_lu_ms(A_dev, reps=2):
    """torch.linalg.lu_factor (rocSOLVER getrf on ROCm) of the same FP64 matrix on the same GPU: best of `reps`."""
    import torch
    best = float("inf")
    torch.linalg.lu_factor(A_dev[:512, :512])
    for _ in range(reps):
        W = A_dev.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        torch.linalg.lu_factor(W)
        torch.cuda.synchronize()
        best = min(best, 1e3 * (time.perf_counter() - t0))
        del W
    return best


def cpu_lu_baseline(A_dev, b):
    """scipy (OpenBLAS) getrf + getrs of the same FP64 matrix on the host, all threads; median of 3 (ms)."""
    import scipy.linalg as sl
    A = A_dev.cpu().numpy().T.copy(order="F")  # device storage is column-major
    sl.lu_solve(sl.lu_factor(A[:512, :512]), b[:512])  # warm-up
    ts = []
    for _ in range(3):
        Aw = A.copy(order="F")
        t0 = time.perf_counter()
        lu = sl.lu_factor(Aw, overwrite_a=True, check_finite=False)
        sl.lu_solve(lu, b, check_finite=False)
        ts.append(time.perf_counter() - t0)
    return 1e3 * float(np.median(ts))


def slip_solve_comparison(step_no):
    """Slip solve of one penny-crack growth step (full space, remote sigma_zz = 1 + g z) to an FP64 relative
    residual of 1e-10: the plain FP64 path (assembly + LU) against learned-preconditioned FGMRES (trained network
    from cutanddisplacejulia_amd/data; trained here when absent), block Jacobi and the FP32 / FP16-update LU
    refinements, each on the FP64 matrix and on the FP32-far / FP64-near matrix."""
    from cutanddisplacejulia_amd import meshes, spai
    seq = meshes.growth_sequence(40, 500, 6400)
    if os.path.exists(spai.DEFAULT_PARAMS):
        model, trained = spai.load_model(), "loaded"
    else:
        model, _ = spai.train(seq[:30], epochs=2, log=_log)
        trained = "trained in bench (2 epochs)"
    step = seq[step_no - 1]
    spai.evaluate_step(seq[0], model)  # warm-up (allocations, code objects)
    r = spai.evaluate_step(step, model)
    _headline(r, "slip solve step %d (%d triangles)" % (step_no, r["n_tri"]))
    r["model"] = trained
    return r


def _headline(r, what):
    """The solve speed-ups of one workload, counted only over paths whose solution has an FP64 relative residual
    <= 1e-10 against the FP64-assembled matrix (rows of the FP32-far / FP64-near assembly solve a different matrix
    and report their FP64-matrix residual and forward error instead)."""
    plain = r["fp64"]["total_ms"]
    paths = {k: v for k, v in r.items() if isinstance(v, dict) and k != "fp64" and "residual" in v}
    ok = {k: v for k, v in paths.items() if v["residual"] <= 1e-10}
    best = min(ok, key=lambda k: ok[k]["total_ms"], default=None)
    learned_ok = [k for k in ok if k.startswith("learned")]
    ml = min(learned_ok, key=lambda k: ok[k]["total_ms"], default=None)
    r["fastest_1e10_path"] = best
    r["speedup_fastest_vs_fp64"] = plain / ok[best]["total_ms"] if best else None
    r["learned_best"] = ml
    r["speedup_learned_vs_fp64"] = plain / ok[ml]["total_ms"] if ml else None
    r["solves_per_s"] = {"fp64": 1e3 / plain, "learned": 1e3 / ok[ml]["total_ms"] if ml else None}
    mixed = [v for k, v in paths.items() if k.startswith("mixed_")]
    if mixed:
        r["mixed_assembly_fp64_residual"] = max(v["residual"] for v in mixed)
        r["mixed_assembly_forward_err"] = max(v["forward_err_vs_fp64"] for v in mixed)
    _log("%s: FP64 assembly+LU %.1f ms; fastest path at FP64 residual <= 1e-10: %s %.1f ms (%.1fx); learned %s "
         "%.1f ms (%.1fx); mixed-assembly rows: FP64-matrix residual %.1e, forward error %.1e" % (
             what, plain, best, ok[best]["total_ms"] if best else float("nan"), r["speedup_fastest_vs_fp64"] or 0,
             ml, ok[ml]["total_ms"] if ml else float("nan"), r["speedup_learned_vs_fp64"] or 0,
             r.get("mixed_assembly_fp64_residual", float("nan")), r.get("mixed_assembly_forward_err", float("nan"))))
    return r


def halfspace_solve_case(n_quads):
    """The half-space fault of the headline assembly (5000 triangles, A 15000^2) solved under remote sigma_zz = 1:
    FP64 assembly + FP64 LU against FP32-far / FP64-near assembly with the FP32 MFMA LU (and the FP16-update LU) as
    preconditioner of FP64 GMRES-IR, to an FP64 relative residual of 1e-10; also block Jacobi and the learned
    preconditioner (trained on full-space penny cracks), each on the FP64 matrix and on the mixed one."""
    from cutanddisplacejulia_amd import meshes, spai
    c = meshes.config2(n_quads)
    step = dict(P1=c["P1"], P2=c["P2"], P3=c["P3"], step=0, g=0.0)
    model = spai.load_model() if os.path.exists(spai.DEFAULT_PARAMS) else None
    spai.evaluate_step(dict(step, P1=c["P1"][:200], P2=c["P2"][:200], P3=c["P3"][:200]), model, hs=1,
                       pcs=("lu32",))  # warm-up
    r = spai.evaluate_step(step, model, hs=1, pcs=("jacobi", "learned_k32", "lu32", "lu16"))
    _headline(r, "half-space solve (%d triangles)" % r["n_tri"])
    return r


def batched_case():
    """64 planar elliptical cracks (N_i ~ U(200, 3000), 1-8 remote loads each, uniformly random orientation), all
    assembled by one persistent work-queue launch in a captured HIP graph with FP32 far / FP64 near entries, then
    solved together (the FP32 LUs of all meshes stepped in lockstep with shared panel, update and interchange
    launches; multi-right-hand-side FP64 refinement rounds to 1e-10).  The first call tunes the launch and captures
    the graph; the second replays it, and its figures are reported."""
    from cutanddisplacejulia_amd import api, meshes
    batch = meshes.elliptical_cracks()
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    ns = [m["P1"].shape[0] for m in batch]
    entries = float(sum(9 * n * n for n in ns))
    api.set_precision("mixed")
    try:
        t0 = time.perf_counter()
        api.batched_slip(batch, nu, G, lam, 0)
        first = time.perf_counter() - t0
        t0 = time.perf_counter()
        xs, rep = api.batched_slip(batch, nu, G, lam, 0)
        wall = time.perf_counter() - t0
    finally:
        api.set_precision("fp64")
    out = {"meshes": len(batch), "triangles": int(sum(ns)), "n_min": int(min(ns)), "n_max": int(max(ns)),
           "rhs": int(sum(np.atleast_2d(m["stresses"]).shape[0] for m in batch)), "entries": entries,
           "assembly_ms": rep["assembly_ms"], "assembly_entries_per_s": entries / (rep["assembly_ms"] * 1e-3),
           "near_entries": rep["near_entries"], "solve_ms": rep["solve_ms"], "residual": rep["residual"],
           "refinement_rounds": rep["iterations"],
           "solves_per_s": sum(np.atleast_2d(m["stresses"]).shape[0] for m in batch) / (rep["solve_ms"] * 1e-3),
           "wall_ms": 1e3 * wall, "first_call_ms": 1e3 * first, "tile": rep["tile"]}
    _log("batched (64 cracks, %d triangles, %.3g entries): assembly %.1f ms = %.3e entries/s, solves %.0f ms, "
         "worst residual %.1e, wall %.0f ms (first call %.0f ms)" % (out["triangles"], entries, rep["assembly_ms"],
                                                                    out["assembly_entries_per_s"], rep["solve_ms"],
                                                                    rep["residual"], 1e3 * wall, 1e3 * first))
    return out


def large_case():
    """Planar elliptical crack (a/b = 2) of 16224 triangles in full space, A 48672 x 48672 FP64 (19 GB): FP64
    assembly, then three solves to an FP64 relative residual of 1e-10 (plain FP64 LU; LU with FP16-operand
    trailing updates preconditioning FP64 GMRES-IR; learned SPAI FGMRES with FP32 Arnoldi products), then the
    solved crack's summed strain and displacement on a 200 x 200 grid (4e4 points)."""
    import torch
    from cutanddisplacejulia_amd import _lib, api, learned, meshes, spai
    c = meshes.config5()
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    N3 = 3 * n
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G = api.ElasticConstantsCheck(c["G"], c["nu"])
    d = {k: learned.to_device_cols(v) for k, v in dict(P1=P1, P2=P2, P3=P3, N=nrm, M=mid).items()}
    L = _lib.load()
    ctx = _lib.ctx()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    A = torch.empty((N3, N3), dtype=torch.float64, device="cuda:0")
    learned.assemble_dev(d["N"], d["M"], d["P1"], d["P2"], d["P3"], n, nu, G, lam, 0, A, N3)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    learned.assemble_dev(d["N"], d["M"], d["P1"], d["P2"], d["P3"], n, nu, G, lam, 0, A, N3)
    torch.cuda.synchronize()
    asm_ms = 1e3 * (time.perf_counter() - t0)
    Tn, Tds, Tss = api.SetupTractionVector(api.Stresses(0, 0, 1.0, 0, 0, 0), nrm, n, G, lam)
    b = torch.from_numpy(np.concatenate([Tn, Tss, Tds])).to("cuda:0")
    bn = float(torch.linalg.norm(b))

    def res(x):
        return float(torch.linalg.norm(b - torch.mv(A.t(), x)) / bn)

    out = {"triangles": n, "N": N3, "assembly_ms": asm_ms, "assembly_entries_per_s": float(N3) * N3 / (asm_ms * 1e-3)}
    _log("large: %d triangles, A %d^2 FP64 assembled in %.1f ms" % (n, N3, asm_ms))
    # plain FP64 LU + solve
    LU = A.clone()
    ipiv = torch.empty(N3, dtype=torch.int32, device="cuda:0")
    x64 = b.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(L.cdj_lu_factor_dev(ctx, ctypes.c_void_p(LU.data_ptr()), N3, N3, ctypes.c_void_p(ipiv.data_ptr()), sp))
    _lib.check(L.cdj_lu_solve_dev(ctx, ctypes.c_void_p(LU.data_ptr()), N3, N3, ctypes.c_void_p(ipiv.data_ptr()),
                                  ctypes.c_void_p(x64.data_ptr()), 1, N3, sp))
    torch.cuda.synchronize()
    t = 1e3 * (time.perf_counter() - t0)
    del LU
    out["fp64_lu"] = {"solve_ms": t, "tflops": (2.0 / 3.0) * N3 ** 3 / (t * 1e-3) / 1e12, "residual": res(x64)}
    _log("large: FP64 LU + solve %.0f ms (%.1f TFLOP/s), residual %.1e" % (t, out["fp64_lu"]["tflops"],
                                                                           out["fp64_lu"]["residual"]))
    model = spai.load_model()
    for name, kw in (("lu16_gmres_ir", dict(pc="lu16")), ("learned_fgmres_k32", dict(pc="learned", model=model,
                                                                                     krylov32=True))):
        x = torch.zeros_like(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = learned.precond_solve_dev(A, N3, b, x, d["P1"], d["P2"], d["P3"], n, G, **kw)
        torch.cuda.synchronize()
        t = 1e3 * (time.perf_counter() - t0)
        fe = float(torch.linalg.norm(x - x64) / torch.linalg.norm(x64))
        out[name] = {"solve_ms": t, "iterations": rep["iterations"], "residual": res(x), "forward_err_vs_fp64": fe}
        _log("large: %s %.0f ms, %d iterations, residual %.1e, vs FP64 %.1e" % (name, t, rep["iterations"],
                                                                               res(x), fe))
    # summed field of the solved slip on the grid (host arrays in and out)
    xh = x64.cpu().numpy()
    Dn, Dss, Dds = xh[:n], xh[n:2 * n], xh[2 * n:]
    api.TD(c["X"][:10], c["Y"][:10], c["Z"][:10], P1, P2, P3, Dss, Dds, Dn, nu, G, 1, 1, 0, api.Strains(),
           api.Disps())
    t0 = time.perf_counter()
    api.TD(c["X"], c["Y"], c["Z"], P1, P2, P3, Dss, Dds, Dn, nu, G, 1, 1, 0, api.Strains(), api.Disps())
    t = time.perf_counter() - t0
    pairs = float(c["X"].size) * n
    out["grid_field"] = {"points": int(c["X"].size), "pairs": pairs, "wall_ms": 1e3 * t, "pairs_per_s": pairs / t}
    _log("large: grid field %d points x %d triangles in %.1f ms (%.3e pairs/s, host arrays)" % (
        c["X"].size, n, 1e3 * t, pairs / t))
    del A
    torch.cuda.empty_cache()
    return out


def crack_tip_case():
    """Crack-tip edges and K_I..K_III (GetCrackTipElements3D + StressIntensity3D) from HBM-resident geometry and slip
    on penny meshes of 2400, 16224 and 998784 triangles (the device entry point, timed with HIP events on its
    stream), beside the oracle's restatement of the reference algorithm (pairwise edge matching) on the 2400-triangle
    mesh (host, one thread)."""
    import torch
    from cutanddisplacejulia_amd import _lib, api, learned, meshes
    from oracle import crack_tip as CT
    L = _lib.load()
    ctx = _lib.ctx()
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rng = np.random.default_rng(1)
    out = {}
    for rings in (20, 52, 408):
        P1, P2, P3 = meshes.penny(rings, 1.0)
        n = P1.shape[0]
        nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
        D = rng.normal(size=(3, n))
        d = [learned.to_device_cols(v) for v in (mid, P1, P2, P3, nrm)]
        dD = [torch.from_numpy(v.copy()).to("cuda:0") for v in D]
        flags = torch.empty(3 * n, dtype=torch.uint8, device="cuda:0")
        geom = torch.empty(3 * n * 13, dtype=torch.float64, device="cuda:0")
        K = torch.empty(3 * n * 4, dtype=torch.float64, device="cuda:0")
        args = ([ctypes.c_void_p(t.data_ptr()) for t in d] + [n] + [ctypes.c_void_p(t.data_ptr()) for t in dD] +
                [1.0, 0.25, ctypes.c_void_p(flags.data_ptr()), ctypes.c_void_p(geom.data_ptr()),
                 ctypes.c_void_p(K.data_ptr()), sp])
        ms = []
        for rep in range(4):
            _lib.check(L.cdj_stress_intensity_dev(ctx, *args))
            if rep:
                ms.append(L.cdj_last_kernel_ms(ctx))
        t = float(np.median(ms))
        nfree = int(flags.sum().item())
        out[str(n)] = {"triangles": n, "tip_edges": nfree, "device_ms": t, "triangles_per_s": n / (t * 1e-3)}
        _log("crack tip: %d triangles, %d tip edges, %.3f ms on the device (%.3e triangles/s)" % (n, nfree, t,
                                                                                               n / (t * 1e-3)))
        if rings == 20:
            t0 = time.perf_counter()
            e = CT.crack_tip_elements(mid, P1, P2, P3, nrm)
            CT.stress_intensity(D[0], D[1], D[2], 1.0, 0.25, nrm, e)
            tc = time.perf_counter() - t0
            out["cpu_oracle_2400"] = {"seconds": tc, "triangles_per_s": n / tc, "cores": 1,
                                      "kind": "port of the reference's pairwise algorithm (numpy)"}
            _log("crack tip: oracle (pairwise edge matching as the reference) %.3f s for %d triangles" % (tc, n))
    return out


def friction_case():
    """Frictional slip of the reference's Burgmann case (a 497-triangle vertical blade fault, mu = 0, cohesion
    rising linearly from the centre): the GPU path (inverse, 2485 x 2485 complementarity problem, Fischer-Newton
    with GPU-LU Newton steps) against the oracle's restatement of the same steps on the host (numpy/LAPACK, all
    cores), both from host arrays."""
    from cutanddisplacejulia_amd import api, meshio
    from oracle import friction as FR
    # the reference's Burgmann driver (test_TDvsBurgmann1994.jl): the blade mesh flattened to z = 0 and scaled
    scl = 1e-4
    P, T = meshio.GoCadAsciiReader(os.path.join(ROOT, "tests", "golden",
                                                "BladeFaultShort_500Tris_Vertical_NorthDip_0.5.ts"))
    P[:, 2] = P[:, 2] * 0
    P[:, 1:4] = P[:, 1:4] * scl
    P1, P2, P3 = meshio.CreateP1P2P3(T, P)
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    n = P1.shape[0]
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    stress = (np.zeros(n), np.full(n, -0.05 * scl), np.zeros(n), np.full(n, scl), np.zeros(n), np.zeros(n))
    z = np.zeros(n)
    Sf = np.abs(mid[:, 0])
    bc = api.MixedBoundaryConditionsFriction(api.MixedBoundaryConditions(api.Stresses(*stress), api.Tractions(z, z, z)),
                                             api.FrictionParameters(0.0, Sf))
    api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 0, bc, z)  # warm-up
    t0 = time.perf_counter()
    api.SlipCalculator3D(P1, P2, P3, nu, G, lam, mid, nrm, 0, bc, z)
    gpu = time.perf_counter() - t0
    rep = dict(api.LAST_REPORT)
    t0 = time.perf_counter()
    _, _, _, info = FR.slip_friction(P1, P2, P3, nu, G, lam, mid, nrm, 0, stress, (z, z, z), 0.0, Sf)
    cpu = time.perf_counter() - t0
    out = {"triangles": n, "lcp_size": 5 * n, "gpu_wall_ms": 1e3 * gpu, "gpu_lcp_ms": rep["solve_ms"],
           "newton_iterations": rep["iterations"], "cpu_oracle_s": cpu, "cpu_newton_iterations": info["iterations"]}
    _log("friction (Burgmann case, %d triangles, LCP %d): GPU %.0f ms wall (complementarity solve %.0f ms, %d Newton "
         "steps); oracle on the host %.2f s" % (n, 5 * n, 1e3 * gpu, rep["solve_ms"], rep["iterations"], cpu))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=50, help="quads per side of the fault mesh (50 -> 5000 triangles)")
    ap.add_argument("--cpu-rows", type=int, default=1000,
                    help="receiver rows per timed CPU-baseline run (5 runs, median); 1000 rows take ~1.1 s on 16 "
                         "cores")
    ap.add_argument("--no-cpu-lu", action="store_true", help="skip the host LU baseline")
    ap.add_argument("--no-rocsolver", action="store_true", help="skip the rocSOLVER getrf comparison")
    ap.add_argument("--solve-step", type=int, default=40,
                    help="penny-crack growth step of the slip-solve comparison (40: 6534 triangles)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the timed steps, copy this many MiB device-to-device (a kernel of known traffic, "
                         "to calibrate hardware byte counters against)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="skip the FP64 LU of the assembled matrix and the "
                    "slip-solve comparison")
    ap.add_argument("--no-compare", action="store_true", help="skip the slip-solve comparison only")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the 64-crack batched assembly + solve (graph replay), reported under 'batched'")
    ap.add_argument("--friction", action="store_true",
                    help="also time the frictional slip of the 497-triangle Burgmann case against the host oracle")
    ap.add_argument("--crack-tip", action="store_true",
                    help="also time the crack-tip edges + stress intensities on 2400 / 16224 / 998784-triangle meshes")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 16224-triangle elliptical crack (A 48672^2 FP64: assembly, three solvers, grid "
                         "field), reported under 'large'")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    from cutanddisplacejulia_amd import _lib, api, meshes

    c = meshes.config2(args.n)
    P1, P2, P3 = c["P1"], c["P2"], c["P3"]
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G = api.ElasticConstantsCheck(1.0, 0.25)
    N3 = 3 * n
    dev = torch.device("cuda:0")

    def dcol(a):  # N x 3 column-major on the device
        return torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).to(dev)

    dP1, dP2, dP3, dN, dM = (dcol(a) for a in (P1, P2, P3, nrm, mid))
    A = torch.empty((N3, N3), dtype=torch.float64, device=dev)  # column-major view: A[col, row] storage
    L = _lib.load()
    ctx = _lib.ctx()
    stream = torch.cuda.current_stream(dev)
    sp = ctypes.c_void_p(stream.cuda_stream)
    opts = _lib.Opts(tol=1e-10, near_factor=2.0)

    edges = ctypes.c_int64()
    edge_ms = ctypes.c_double()

    def step():
        rc = L.cdj_assemble_dev(ctx, ctypes.c_void_p(dN.data_ptr()), ctypes.c_void_p(dM.data_ptr()),
                                ctypes.c_void_p(dP1.data_ptr()), ctypes.c_void_p(dP2.data_ptr()),
                                ctypes.c_void_p(dP3.data_ptr()), n, nu, G, lam, 1, ctypes.c_void_p(A.data_ptr()), N3,
                                ctypes.byref(opts), sp)
        _lib.check(rc)
        _lib.check(L.cdj_last_edge_stats(ctx, ctypes.byref(edges), ctypes.byref(edge_ms)))
        return L.cdj_last_kernel_ms(ctx), edge_ms.value

    _log("workload: %d triangles, A %d x %d (half space)" % (n, N3, N3))
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    kms, ems = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        k, e = step()
        kms.append(k)
        ems.append(e)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    entries = float(N3) * N3
    value = entries * args.steps / elapsed
    kavg = float(np.mean(kms))       # all assembly kernels of a step
    eavg = float(np.mean(ems))       # tdk_edge_response alone
    pairs = float(n) * n
    units = float(edges.value) * n   # (receiver, edge) units of one edge-response launch
    achieved_tf = FLOPS_PER_EDGE_HS * units / (eavg * 1e-3) / 1e12
    reference_tf = REF_FLOPS_PER_EDGE_HS * units / (eavg * 1e-3) / 1e12
    formulation_tf = FLOPS_PER_PAIR_HS * pairs / (kavg * 1e-3) / 1e12
    _log("assembly: %.2f ms/step, kernels %.2f ms (edge responses %.2f ms over %d edges), %.3e entries/s; edge "
         "kernel %.1f TFLOP/s algorithmic" % (ms_per_step, kavg, eavg, edges.value, value, achieved_tf))

    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib * (1 << 20) // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        _log("copied %d MiB device-to-device (counter calibration)" % args.copy_mib)
        del src, dst

    lu15 = None
    if not args.no_solve:
        # FP64 LU of the assembled A with one right-hand side (remote sigma_zz = 1)
        Tn, Tds, Tss = api.SetupTractionVector(api.Stresses(0, 0, 1.0, 0, 0, 0), nrm, n, G, lam)
        b = np.concatenate([Tn, Tss, Tds])
        Ab = torch.empty((N3 + 1, N3), dtype=torch.float64, device=dev)
        Ab[:N3].copy_(A)
        Ab[N3].copy_(torch.from_numpy(b).to(dev))
        ipiv = torch.empty(N3, dtype=torch.int32, device=dev)
        A0 = A.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(L.cdj_lu_factor_dev(ctx, ctypes.c_void_p(Ab.data_ptr()), N3, N3, ctypes.c_void_p(ipiv.data_ptr()), sp))
        x = Ab[N3].clone()
        _lib.check(L.cdj_lu_solve_dev(ctx, ctypes.c_void_p(Ab.data_ptr()), N3, N3, ctypes.c_void_p(ipiv.data_ptr()),
                                      ctypes.c_void_p(x.data_ptr()), 1, N3, sp))
        torch.cuda.synchronize()
        t_lu = time.perf_counter() - t0
        # FP64 residual ||b - A x|| / ||b||  (A0 storage is column-major: A0[col, row])
        r = torch.from_numpy(b).to(dev) - torch.mv(A0.t(), x)
        res = float(torch.linalg.norm(r) / torch.linalg.norm(torch.from_numpy(b).to(dev)))
        lu_tf = (2.0 / 3.0) * N3 ** 3 / t_lu / 1e12
        _log("FP64 LU+solve: %.1f ms (%.1f TFLOP/s), residual %.2e" % (1e3 * t_lu, lu_tf, res))
        lu15 = {"fp64_lu_ms": 1e3 * t_lu, "fp64_lu_tflops": lu_tf, "fp64_residual": res, "n": N3}
        if args.no_rocsolver:
            lu15["rocsolver_getrf_ms"] = None
        else:
            try:
                rs = rocsolver_lu_ms(A0)
                lu15["rocsolver_getrf_ms"] = rs
                _log("rocSOLVER getrf (torch.linalg.lu_factor) of the same matrix: %.1f ms" % rs)
            except Exception as e:  # noqa: BLE001  (a library comparison only)
                lu15["rocsolver_getrf_ms"] = None
                _log("rocSOLVER comparison unavailable: %s" % e)
        if not args.no_cpu and not args.no_cpu_lu:
            cpu_lu = cpu_lu_baseline(A0, b)
            _log("CPU LU+solve (scipy/OpenBLAS, %s threads): %.0f ms" % (os.environ.get("OMP_NUM_THREADS", "all"),
                                                                         cpu_lu))
            lu15["cpu_lu_ms"] = cpu_lu
        del Ab, A0

    solve = None
    if not args.no_solve and not args.no_compare:
        solve = slip_solve_comparison(args.solve_step)

    hs_solve = None if (args.no_solve or args.no_compare) else halfspace_solve_case(args.n)
    batched = None if args.no_batched else batched_case()
    large = None if args.no_large else large_case()
    crack_tip = crack_tip_case() if args.crack_tip else None
    friction = friction_case() if args.friction else None

    cpu = None
    parity = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        rows = min(args.cpu_rows, n) if args.cpu_rows > 0 else n
        cv, cdt, ctot, T = cpu_baseline(P1, P2, P3, nrm, mid, nu, G, lam, rows, threads)
        _log("CPU baseline: %.3e entries/s (%d threads, %d rows, median %.2f s of 5)" % (cv, threads, rows, cdt))
        cpu = {"value": cv, "unit": "entries/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
               "sample": "%d receiver rows x %d sources of the same half-space assembly, median of 5 timed runs "
                         "(%.1f s of CPU work in total)" % (rows, n, ctot)}
        # the same rows of the GPU matrix (storage is A^T: A[col, row])
        A_rows = {q: A[:, q * n:q * n + rows].t().contiguous().cpu().numpy() for q in range(3)}
        parity = parity_rows(A_rows, T, n)
        _log("GPU matrix vs the oracle's rows 0..%d: max error %.2e of the slip mode's scale" % (rows - 1, parity))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "entries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "half-space collocation assembly: 45-degree dipping fault, %d triangles, "
                               "A %d x %d FP64" % (n, N3, N3), "n_tri": n, "hs": 1},
        "roofline": {"bound": "valu", "achieved": achieved_tf, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tf / PEAK_FP64_TFLOPS,
                     "traffic": TRAFFIC_BYTES_EDGE_HS_5000 if args.n == 50 else None,
                     "kernel": "tdk_edge_response", "kernel_ms": eavg, "flops_per_unit": FLOPS_PER_EDGE_HS,
                     "unit_of_work": "(receiver, mesh edge)", "units_per_launch": units,
                     "kernel_flops_per_unit": KERNEL_FLOPS_PER_EDGE_HS,
                     "executed_flops_per_unit": EXECUTED_FLOPS_PER_EDGE_HS if args.n == 50 else None,
                     "executed_frac": (EXECUTED_FLOPS_PER_EDGE_HS * units / (eavg * 1e-3) / 1e12 / PEAK_FP64_TFLOPS
                                       if args.n == 50 and EXECUTED_FLOPS_PER_EDGE_HS else None),
                     "reference_flops_per_unit": REF_FLOPS_PER_EDGE_HS,
                     "reference_formulation_tflops": reference_tf,
                     "assembly_kernels_ms": kavg,
                     "assembly_formulation_tflops": formulation_tf,
                     "note": "flops_per_unit counts the engine's forward-mode formulation with every shared quantity "
                             "formed once and only the operations that reach the unit's tractions; the kernel forms "
                             "the shared quantities once per Burgers component (kernel_flops_per_unit), and executes "
                             "executed_flops_per_unit (PMC, with the Newton steps of reciprocals and square roots).  "
                             "Under this FP64 load the chip runs at about 2.0 GHz (GRBM_GUI_ACTIVE per dispatch), "
                             "against the 2.4 GHz of the peak figure.  "
                             "reference_formulation_tflops and assembly_formulation_tflops "
                             "count the reference's closed forms (%g flops per edge unit, %g per pair) and exceed "
                             "the peak: the forward-mode strain is cheaper, and each mesh edge's free-surface part "
                             "is evaluated once, not once per triangle side" % (REF_FLOPS_PER_EDGE_HS,
                                                                                FLOPS_PER_PAIR_HS)},
        "cpu_baseline": cpu,
        "parity_vs_oracle_rows": parity,
        "lu_15000": lu15,
        "solve": solve,
        "halfspace_solve": hs_solve,
    }
    if batched is not None:
        out["batched"] = batched
    if large is not None:
        out["large"] = large
    if crack_tip is not None:
        out["crack_tip"] = crack_tip
    if friction is not None:
        out["friction"] = friction
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
