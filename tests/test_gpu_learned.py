"""Learned SPAI preconditioner on the GPU.

* the GPU features / neighbour lists equal a numpy computation from the same geometry;
* the GPU loss and gradient (FP32 kernels) equal the FP64 numpy restatement (tests/mlp_ref.py, itself pinned by
  finite differences) on the GPU's own features: loss to 1e-5 relative, gradient to 2e-3 normwise;
* training on a few growth steps lowers the loss, and on a held-out mesh the learned preconditioner reaches the
  1e-10 FP64 residual in fewer FGMRES iterations than block Jacobi, which needs fewer than no preconditioner.
"""
import numpy as np
import pytest

from cutanddisplacejulia_amd import api, meshes
from tests import mlp_ref

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
from cutanddisplacejulia_amd import learned  # noqa: E402


def _np_geometry(P1, P2, P3):
    u, v = P2 - P1, P3 - P1
    nn = np.cross(u, v)
    area = 0.5 * np.linalg.norm(nn, axis=1)
    nn /= np.linalg.norm(nn, axis=1)[:, None]
    s = np.stack([-nn[:, 1], nn[:, 0], np.zeros(len(nn))], axis=1)
    flat = np.linalg.norm(s, axis=1) == 0
    s[flat] = np.stack([np.zeros(flat.sum()), nn[flat, 2], np.zeros(flat.sum())], axis=1)
    s /= np.linalg.norm(s, axis=1)[:, None]
    d = np.cross(nn, s)
    h = (np.linalg.norm(u, axis=1) + np.linalg.norm(v, axis=1) + np.linalg.norm(P3 - P2, axis=1)) / 3
    c = (P1 + P2 + P3) / 3
    return c, nn, s, d, h, area


def test_features_match_numpy():
    P1, P2, P3 = meshes.penny(5, 1.0, jitter=0.03, normal=(0.3, 0.2, 0.9))
    n = P1.shape[0]
    k = 16
    m = learned.LearnedPreconditioner(k=k, hidden=16, seed=3)
    S = 2
    rng = np.random.default_rng(0)
    _, _, feat, nbr, geo = m.loss_grad(P1, P2, P3, 1.0, rng.normal(size=(n, 9)), rng.normal(size=(3 * n, S)),
                                       rng.normal(size=(3 * n, S)))
    c, nn, s, d, h, area = _np_geometry(P1, P2, P3)
    assert np.allclose(geo[:, 12], h, rtol=1e-6) and np.allclose(geo[:, 13], area, rtol=1e-5)
    for i in range(n):
        dist = np.sum((c - c[i]) ** 2, axis=1)
        dist[i] = -1
        order = np.argsort(dist, kind="stable")[:k]
        # the k-th distance may tie; compare as sets of distances
        assert nbr[i, 0] == i
        assert np.allclose(np.sort(dist[nbr[i]]), np.sort(dist[order]), rtol=1e-4, atol=1e-9)
        F = feat[i].reshape(k, 10)
        for q in range(k):
            j = nbr[i, q]
            ax = np.stack([nn[i], s[i], d[i]])
            assert np.allclose(F[q, 0:3], ax @ ((c[j] - c[i]) / h[i]), atol=2e-5)
            assert np.allclose(F[q, 3:6], ax @ nn[j], atol=2e-6)
            assert np.allclose(F[q, 6:9], ax @ s[j], atol=2e-6)
            assert np.isclose(F[q, 9], area[j] / h[i] ** 2, rtol=1e-5)


def test_loss_and_gradient_match_fp64_reference():
    P1, P2, P3 = meshes.penny(4, 1.0, jitter=0.03)
    n = P1.shape[0]
    k, H, S, G = 16, 32, 3, 1.7
    m = learned.LearnedPreconditioner(k=k, hidden=H, seed=11)
    params = m.get_params().astype(np.float64)
    # make the output layer non-trivial so every layer carries gradient
    rng = np.random.default_rng(1)
    params = params + rng.normal(scale=0.05, size=params.size)
    m.set_params(params.astype(np.float32))
    params = m.get_params().astype(np.float64)
    jac = rng.normal(size=(n, 9))
    B = rng.normal(size=(3 * n, S))
    X = rng.normal(size=(3 * n, S))
    loss, grad, feat, nbr, geo = m.loss_grad(P1, P2, P3, G, jac, B, X)
    hscale = geo[:, 12].astype(np.float64) / G
    Lr, gr = mlp_ref.loss_and_grad(params, feat.astype(np.float64), nbr, hscale, jac, B, X, k, H)
    assert abs(loss - Lr) <= 1e-5 * abs(Lr), (loss, Lr)
    assert np.linalg.norm(grad - gr) <= 2e-3 * np.linalg.norm(gr), np.linalg.norm(grad - gr) / np.linalg.norm(gr)


def _assembled(P1, P2, P3, nu=0.25, G=1.0):
    n = P1.shape[0]
    nrm, mid = api.CreateFaceNormalAndMidPoint(P1, P2, P3)
    K, E, lam, nu, G = api.ElasticConstantsCheck(G, nu)
    d = {k: learned.to_device_cols(v) for k, v in dict(P1=P1, P2=P2, P3=P3, N=nrm, M=mid).items()}
    A = torch.empty((3 * n, 3 * n), dtype=torch.float64, device="cuda:0")
    learned.assemble_dev(d["N"], d["M"], d["P1"], d["P2"], d["P3"], n, nu, G, lam, 0, A, 3 * n)
    return d, A, nrm, lam


def test_training_improves_gmres_on_held_out_mesh():
    seq = meshes.growth_sequence(steps=8, n_start=150, n_end=600)
    model = learned.LearnedPreconditioner(k=16, hidden=64, seed=5)
    losses = []
    for epoch in range(6):
        for s in seq[:6]:
            d, A, _, _ = _assembled(s["P1"], s["P2"], s["P3"])
            n = s["P1"].shape[0]
            losses.append(model.train_dev(d["P1"], d["P2"], d["P3"], n, 1.0, A, 3 * n, nsamples=16, iters=10,
                                          lr=2e-3, seed=100 * epoch + s["step"]))
    assert np.mean(losses[-6:]) < 0.8 * np.mean(losses[:6]), losses
    s = seq[-1]
    n = s["P1"].shape[0]
    d, A, nrm, lam = _assembled(s["P1"], s["P2"], s["P3"])
    Tn, Tds, Tss = api.SetupTractionVector(api.Stresses(0, 0, 1.0, 0, 0, 0), nrm, n, 1.0, lam)
    b = torch.from_numpy(np.concatenate([Tn, Tss, Tds])).to("cuda:0")
    its = {}
    for pc in ("none", "jacobi", "learned"):
        x = torch.zeros_like(b)
        rep = learned.precond_solve_dev(A, 3 * n, b, x, d["P1"], d["P2"], d["P3"], n, 1.0, pc=pc, model=model)
        assert rep["path"] == 2 and rep["residual"] <= 1e-10, (pc, rep)
        Ah = A.cpu().numpy().T  # storage is column-major
        res = np.linalg.norm(b.cpu().numpy() - Ah @ x.cpu().numpy()) / np.linalg.norm(b.cpu().numpy())
        assert res <= 1e-10, (pc, res)
        its[pc] = rep["iterations"]
    assert its["learned"] < its["jacobi"] <= its["none"], its


def test_bf16_network_matches_rounded_reference_and_preconditions():
    """The BF16 inference path (CDJ_PC_LEARNED_BF16): outputs equal the FP64 restatement with the same BF16
    roundings (tests/mlp_ref.forward_bf16) up to FP32 accumulation, stay within 5 % of the FP32 network, and the
    preconditioner reaches the 1e-10 FP64 residual in about as many iterations as the FP32 one."""
    from cutanddisplacejulia_amd import spai
    model = spai.load_model()
    seq = meshes.growth_sequence(40, 500, 6400)
    step = seq[33]
    P1, P2, P3 = step["P1"], step["P2"], step["P3"]
    n = P1.shape[0]
    y32 = model.forward(P1, P2, P3, bf16=False)
    y16 = model.forward(P1, P2, P3, bf16=True)
    rng = np.random.default_rng(0)
    _, _, feat, nbr, geo = model.loss_grad(P1, P2, P3, 1.0, rng.normal(size=(n, 9)), rng.normal(size=(3 * n, 1)),
                                           rng.normal(size=(3 * n, 1)))
    _, _, yref = mlp_ref.forward_bf16(model.get_params(), feat.astype(np.float64), model.k, model.hidden)
    scale = np.max(np.abs(yref))
    d = np.abs(y16 - yref)
    print("[gpu-test] bf16 network: max |gpu - ref| %.2e, mean %.2e, vs FP32 %.2e (scale %.2e)" % (
        d.max(), d.mean(), np.abs(y16 - y32).max(), scale))
    assert d.max() <= 5e-3 * scale and d.mean() <= 1e-4 * scale
    assert np.abs(y16 - y32).max() <= 5e-2 * scale
    d, A, nrm, mid, lam = spai.assemble(step)
    bb = torch.from_numpy(spai.rhs(step, nrm, mid, lam)).to("cuda:0")
    res = {}
    for pc in ("learned", "learned_bf16"):
        x = torch.zeros_like(bb)
        rep = learned.precond_solve_dev(A, 3 * n, bb, x, d["P1"], d["P2"], d["P3"], n, spai.G, pc=pc, model=model)
        r = float(torch.linalg.norm(bb - torch.mv(A.t(), x)) / torch.linalg.norm(bb))
        res[pc] = (rep["iterations"], r)
    print("[gpu-test] iterations: FP32 network %d, BF16 network %d" % (res["learned"][0], res["learned_bf16"][0]))
    assert res["learned_bf16"][1] <= 1e-10 and res["learned"][1] <= 1e-10
    assert res["learned_bf16"][0] <= res["learned"][0] + 5
