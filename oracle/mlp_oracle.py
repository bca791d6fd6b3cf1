"""CPU oracle of the learned phase-init MLP (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker: the product path runs the HIP kernels of
nonenone_amd/csrc/mlp.hip and never calls anything here.

The network has no counterpart in the reference (its training data are the
reference algorithm's flash results, produced by this library's GPU flash).  The
oracle restates, in float64 numpy, the same network and loss as the HIP kernels:

    features  f = [(ln P - 15.5)/1.5, (T - 450)/150, z_0 .. z_{nc-1}]
    h1 = silu(W1 f + b1), h2 = silu(W2 h1 + b2), h3 = silu(W3 h2 + b3), o = W4 h3 + b4
    loss = mean_b [ -ln softmax(o[0:3])[label] + w * mask * mean_k (o[3+k] - t_k)^2 ]

Parameters are flat, per layer W (out x in, row-major) then b (out).  Pinned by
central finite differences of the loss (float64), the softmax cross-entropy
closed forms (per-state logit gradients sum to zero; loss ln 3 at equal logits),
and a planted mistake (a wrong SiLU derivative) that the finite-difference check
rejects.
"""
from __future__ import annotations

import numpy as np

HIDDEN = 64
CLASSES = 3


def layer_shapes(nc: int):
    nin, nout = nc + 2, CLASSES + nc
    return [(HIDDEN, nin), (HIDDEN, HIDDEN), (HIDDEN, HIDDEN), (nout, HIDDEN)]


def param_count(nc: int) -> int:
    return sum(o * i + o for o, i in layer_shapes(nc))


def unpack(params, nc: int):
    params = np.asarray(params, np.float64)
    out, off = [], 0
    for o, i in layer_shapes(nc):
        W = params[off:off + o * i].reshape(o, i)
        off += o * i
        b = params[off:off + o]
        off += o
        out.append((W, b))
    assert off == params.size
    return out


def features(P, T, z):
    P = np.maximum(np.asarray(P, np.float64), 1.0)
    return np.concatenate([((np.log(P) - 15.5) / 1.5)[:, None],
                           ((np.asarray(T, np.float64) - 450.0) / 150.0)[:, None],
                           np.asarray(z, np.float64)], axis=1)


def sigmoid(a):
    return 1.0 / (1.0 + np.exp(-a))


def silu(a):
    return a * sigmoid(a)


def silu_grad(a):
    s = sigmoid(a)
    return s * (1.0 + a * (1.0 - s))


def forward(params, X, nc: int):
    layers = unpack(params, nc)
    h = np.asarray(X, np.float64)
    pre, post = [], [h]
    for l, (W, b) in enumerate(layers):
        a = h @ W.T + b
        if l < 3:
            pre.append(a)
            h = silu(a)
            post.append(h)
        else:
            h = a
    return h, pre, post


def loss_grad(params, X, label, logk, mask, w: float, nc: int, dsilu=silu_grad):
    """Loss and its gradient w.r.t. the flat parameters (float64)."""
    n = X.shape[0]
    o, pre, post = forward(params, X, nc)
    logits = o[:, :CLASSES]
    m = logits.max(1, keepdims=True)
    e = np.exp(logits - m)
    se = e.sum(1, keepdims=True)
    label = np.asarray(label, np.int64)
    lse = (m + np.log(se))[:, 0]
    ce = lse - logits[np.arange(n), label]
    r = o[:, CLASSES:] - logk
    mk = np.asarray(mask, np.float64)
    sq = (r * r).mean(1) * mk
    loss = float((ce + w * sq).sum() / n)
    dO = np.zeros_like(o)
    p = e / se
    p[np.arange(n), label] -= 1.0
    dO[:, :CLASSES] = p / n
    dO[:, CLASSES:] = w * (2.0 / nc) * r * mk[:, None] / n
    layers = unpack(params, nc)
    grads = [None] * 4
    dA = dO
    for l in range(3, -1, -1):
        W, _ = layers[l]
        grads[l] = (dA.T @ post[l], dA.sum(0))
        if l > 0:
            dA = (dA @ W) * dsilu(pre[l - 1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def predict_bf16(params, X, nc: int):
    """BF16 inference semantics: weights, features and hidden activations rounded
    to BF16, products accumulated in float32."""
    layers = unpack(np.asarray(params, np.float32), nc)
    h = bf16(np.asarray(X, np.float32))
    for l, (W, b) in enumerate(layers):
        a = (h.astype(np.float32) @ bf16(W.astype(np.float32)).T + bf16(b.astype(np.float32)))
        a = a.astype(np.float32)
        h = bf16(silu(a.astype(np.float64)).astype(np.float32)) if l < 3 else a
    logits = h[:, :CLASSES].astype(np.float64)
    e = np.exp(logits - logits.max(1, keepdims=True))
    return e / e.sum(1, keepdims=True), h[:, CLASSES:]


def init_params(nc: int, seed: int = 0):
    """A random parameter vector (numpy PCG64; for tests that start from the oracle)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for o, i in layer_shapes(nc):
        lim = 1.0 / np.sqrt(i)
        parts.append(rng.uniform(-lim, lim, o * i + o))
    return np.concatenate(parts)
