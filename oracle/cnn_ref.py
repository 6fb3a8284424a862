"""FP64 numpy restatement of the orientation-proposal network (forward,
softmax cross-entropy, backward) with the parameter layout of bsx_cnn.

TEST INFRASTRUCTURE ONLY: the checker of the HIP training and inference
kernels (tests/); its backward pass is pinned by central finite differences.

Layout of the flat parameter vector: for layers l = 1..4 the conv weights
[c_l][c_{l-1}][3][3] then the biases [c_l]; then the FC weights
[nclasses][c_4 (S/8)^2] and biases [nclasses]. Conv: 3x3, zero padding 1,
strides 1, 2, 2, 2, ReLU after each conv.
"""
from __future__ import annotations

import numpy as np


def layout(S: int, ch, ncls: int):
    chans = [1] + list(ch)
    off = []
    o = 0
    for l in range(4):
        nw = chans[l + 1] * chans[l] * 9
        off.append((o, o + nw, chans[l + 1]))
        o += nw + chans[l + 1]
    feat = chans[4] * (S // 8) ** 2
    off.append((o, o + feat * ncls, ncls))
    o += feat * ncls + ncls
    return off, o, feat, chans


def _im2col(x, stride):
    B, C, H, W = x.shape
    Ho = (H + 2 - 3) // stride + 1
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    cols = np.empty((B, C, 3, 3, Ho, Ho))
    for ky in range(3):
        for kx in range(3):
            cols[:, :, ky, kx] = xp[:, :, ky:ky + stride * (Ho - 1) + 1:stride,
                                    kx:kx + stride * (Ho - 1) + 1:stride]
    return cols


def _col2im(dcols, shape, stride):
    B, C, H, W = shape
    Ho = dcols.shape[-1]
    dxp = np.zeros((B, C, H + 2, W + 2))
    for ky in range(3):
        for kx in range(3):
            dxp[:, :, ky:ky + stride * (Ho - 1) + 1:stride,
                kx:kx + stride * (Ho - 1) + 1:stride] += dcols[:, :, ky, kx]
    return dxp[:, :, 1:-1, 1:-1]


def forward(p: np.ndarray, x: np.ndarray, S: int, ch, ncls: int, round_act=None):
    """x: (B, S, S). Returns logits (B, ncls) and the cache for backward.
    round_act, if given, is applied to each conv layer's output (the storage
    rounding of an inference path)."""
    off, n, feat, chans = layout(S, ch, ncls)
    a = x.reshape(len(x), 1, S, S).astype(np.float64)
    cache = []
    for l in range(4):
        o, ob, co = off[l]
        W = p[o:ob].reshape(co, chans[l], 3, 3)
        b = p[ob:ob + co]
        stride = 1 if l == 0 else 2
        cols = _im2col(a, stride)
        z = np.einsum("oikl,bikluv->bouv", W, cols) + b[None, :, None, None]
        y = np.maximum(z, 0)
        if round_act is not None:
            y = round_act(y)
        cache.append((a, cols, y, stride))
        a = y
    o, ob, _ = off[4]
    Wf = p[o:ob].reshape(ncls, feat)
    bf = p[ob:ob + ncls]
    f = a.reshape(len(a), -1)
    logits = f @ Wf.T + bf
    return logits, (cache, f)


def backward(p: np.ndarray, cache, f, dl: np.ndarray, S: int, ch, ncls: int) -> np.ndarray:
    """Flat parameter gradient given dL/dlogits (B, ncls) and the forward cache."""
    off, n, feat, chans = layout(S, ch, ncls)
    g = np.zeros(n)
    o, ob, _ = off[4]
    Wf = p[o:ob].reshape(ncls, feat)
    g[o:ob] = (dl.T @ f).ravel()
    g[ob:ob + ncls] = dl.sum(axis=0)
    da = (dl @ Wf).reshape(cache[3][2].shape)
    for l in range(3, -1, -1):
        a_in, cols, y, stride = cache[l]
        o, ob, co = off[l]
        W = p[o:ob].reshape(co, chans[l], 3, 3)
        dz = da * (y > 0)
        g[o:ob] = np.einsum("bouv,bikluv->oikl", dz, cols).ravel()
        g[ob:ob + co] = dz.sum(axis=(0, 2, 3))
        if l > 0:
            dcols = np.einsum("oikl,bouv->bikluv", W, dz)
            da = _col2im(dcols, a_in.shape, stride)
    return g


def loss_and_grad(p: np.ndarray, x: np.ndarray, labels: np.ndarray, S: int, ch, ncls: int):
    """Mean cross-entropy, its gradient (flat, layout above) and per-sample losses."""
    logits, (cache, f) = forward(p, x, S, ch, ncls)
    B = len(x)
    m = logits.max(axis=1, keepdims=True)
    e = np.exp(logits - m)
    se = e.sum(axis=1, keepdims=True)
    losses = (np.log(se) + m)[:, 0] - logits[np.arange(B), labels]
    dl = e / se
    dl[np.arange(B), labels] -= 1
    dl /= B
    return losses.mean(), backward(p, cache, f, dl, S, ch, ncls), losses


def pair_loss(z: np.ndarray, B: int, tau: float):
    """Symmetric contrastive loss of 2B network outputs (particles 0..B-1,
    their references B..2B-1): unit rows u, cosines C = u_p u_r^T, per pair
    (row + column cross-entropy of C / tau at the matching entry) / 2.
    Returns (per-pair losses, dL/dz for L = their mean)."""
    nz = np.maximum(np.linalg.norm(z, axis=1, keepdims=True), 1e-12)
    u = z / nz
    up, ur = u[:B], u[B:]
    Cm = up @ ur.T / tau
    def softmax(a, axis):
        e = np.exp(a - a.max(axis=axis, keepdims=True))
        return e / e.sum(axis=axis, keepdims=True)
    P = softmax(Cm, 1)
    Q = softmax(Cm, 0)
    idx = np.arange(B)
    losses = 0.5 * (-np.log(P[idx, idx]) - np.log(Q[idx, idx]))
    dC = (P + Q - 2 * np.eye(B)) / (2 * B) / tau
    du = np.concatenate([dC @ ur, dC.T @ up])
    dz = (du - np.sum(du * u, axis=1, keepdims=True) * u) / nz
    return losses, dz


def pair_loss_and_grad(p: np.ndarray, x: np.ndarray, B: int, tau: float, S: int, ch, d: int):
    """Mean contrastive loss of the pairs (x: 2B inputs), its flat gradient and
    the per-pair losses."""
    z, (cache, f) = forward(p, x, S, ch, d)
    losses, dz = pair_loss(z, B, tau)
    return losses.mean(), backward(p, cache, f, dz, S, ch, d), losses
