"""The FP64 numpy restatement of the learned preconditioner's loss gradient agrees with finite differences
(no GPU) - this pins the reference the GPU backward pass is checked against."""
import numpy as np

from tests import mlp_ref


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n, k, H, S = 12, 4, 6, 3
    F, O = 10 * k, 9 * k
    np_ = F * H + H + H * H + H + H * O + O
    params = rng.normal(scale=0.3, size=np_)
    feat = rng.normal(size=(n, F))
    nbr = np.stack([rng.permutation(n)[:k] for _ in range(n)])
    nbr[:, 0] = np.arange(n)
    hscale = rng.uniform(0.5, 1.5, n)
    jac = rng.normal(size=(n, 9))
    B = rng.normal(size=(3 * n, S))
    X = rng.normal(size=(3 * n, S))
    L, g = mlp_ref.loss_and_grad(params, feat, nbr, hscale, jac, B, X, k, H)
    assert abs(L - mlp_ref.loss_only(params, feat, nbr, hscale, jac, B, X, k, H)) < 1e-12 * abs(L)
    idx = rng.choice(np_, 40, replace=False)
    h = 1e-6
    for q in idx:
        e = np.zeros(np_)
        e[q] = h
        fd = (mlp_ref.loss_only(params + e, feat, nbr, hscale, jac, B, X, k, H)
              - mlp_ref.loss_only(params - e, feat, nbr, hscale, jac, B, X, k, H)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])), (q, fd, g[q])
