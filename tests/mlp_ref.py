"""FP64 numpy restatement of the learned-preconditioner network and its SPAI loss (test infrastructure).

Parameter layout (flat, FP32 on the GPU): W1 (F x H, row-major), b1 (H), W2 (H x H), b2 (H), W3 (H x O), b3 (O),
with F = 10 k, O = 9 k.  h1 = tanh(feat W1 + b1), h2 = tanh(h1 W2 + b2), Y = h2 W3 + b3.
Preconditioner row block of element i: z_i[a] = sum_b jac_i[a][b] r[b n + i]
                                             + (h_i / G) sum_q sum_b Y_i[q][a][b] r[b n + nbr_iq].
Loss over S samples (x_s, b_s = A x_s): L = sum_s sum_i |(M b_s)_i - x_s,i|^2 / (S n).
"""
import numpy as np


def unpack(params, k, H):
    F, O = 10 * k, 9 * k
    p = np.asarray(params, dtype=np.float64)
    o = 0
    W1 = p[o:o + F * H].reshape(F, H)
    o += F * H
    b1 = p[o:o + H]
    o += H
    W2 = p[o:o + H * H].reshape(H, H)
    o += H * H
    b2 = p[o:o + H]
    o += H
    W3 = p[o:o + H * O].reshape(H, O)
    o += H * O
    b3 = p[o:o + O]
    o += O
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def forward(params, feat, k, H):
    W1, b1, W2, b2, W3, b3 = unpack(params, k, H)
    h1 = np.tanh(feat @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    return h1, h2, h2 @ W3 + b3


def to_bf16(x):
    """Round to BF16 (round to nearest even through FP32), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(params, feat, k, H):
    """The network with each GEMM's operands (features or activations, and weights) rounded to BF16; products,
    sums, biases and tanh in FP64 (the GPU accumulates in FP32)."""
    W1, b1, W2, b2, W3, b3 = unpack(params, k, H)
    h1 = np.tanh(to_bf16(feat) @ to_bf16(W1) + b1)
    h2 = np.tanh(to_bf16(h1) @ to_bf16(W2) + b2)
    return h1, h2, to_bf16(h2) @ to_bf16(W3) + b3


def apply_M(Y, jac, nbr, hscale, r, n, k):
    """z = M r for one vector r (length 3n)."""
    R = r.reshape(3, n)
    J = jac.reshape(n, 3, 3)
    z = np.einsum("iab,bi->ai", J, R)
    Yb = Y.reshape(n, k, 3, 3)
    Rn = R[:, nbr]  # 3 x n x k
    z += hscale[None, :] * np.einsum("iqab,biq->ai", Yb, Rn)
    return z.reshape(-1)


def loss_and_grad(params, feat, nbr, hscale, jac, B, X, k, H):
    """Loss and gradient with respect to the flat parameter vector (FP64)."""
    n = feat.shape[0]
    S = B.shape[1]
    W1, b1, W2, b2, W3, b3 = unpack(params, k, H)
    h1, h2, Y = forward(params, feat, k, H)
    dY = np.zeros_like(Y)
    L = 0.0
    for s in range(S):
        z = apply_M(Y, jac, nbr, hscale, B[:, s], n, k)
        e = (z - X[:, s]).reshape(3, n)
        L += np.sum(e * e)
        Bn = B[:, s].reshape(3, n)[:, nbr]  # 3 x n x k  (component c, element i, neighbour q)
        dY += (2.0 * hscale[:, None, None, None] * np.einsum("ai,ciq->iqac", e, Bn)).reshape(n, 9 * k)
    w = 1.0 / (S * n)
    L *= w
    dY *= w
    gW3 = h2.T @ dY
    gb3 = dY.sum(0)
    dz2 = (dY @ W3.T) * (1 - h2 ** 2)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * (1 - h1 ** 2)
    gW1 = feat.T @ dz1
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return L, g


def loss_only(params, feat, nbr, hscale, jac, B, X, k, H):
    n = feat.shape[0]
    _, _, Y = forward(params, feat, k, H)
    L = 0.0
    for s in range(B.shape[1]):
        e = apply_M(Y, jac, nbr, hscale, B[:, s], n, k) - X[:, s]
        L += np.sum(e * e)
    return L / (B.shape[1] * n)
