"""CPU restatement of the refinement-prediction network (TEST INFRASTRUCTURE: used only by
tests/, ``__graft_entry__.smoke()`` and bench.py's CPU leg as the checker; never by the
product path).

The network has no counterpart in the reference: g1n0st/adaptive-gimp rebuilds its octree
from the current particles every substep (adaptive_gimp.py:323-337).  It is a
16 -> 64 -> 64 -> 4 perceptron with ReLU hidden layers and a softmax cross-entropy loss.
Parameters are one float32 vector, each matrix row-major (out, in):

    W1[64,16] b1[64] W2[64,64] b2[64] W3[4,64] b3[4]          (5508 values)

``forward`` / ``loss_grad`` compute in float64.  With ``bf16_rounding=True`` they round to bfloat16
(round to nearest even) exactly where the GPU kernels do:

* forward: the inputs, the three weight matrices and the two hidden activations (after the
  ReLU) are rounded before each product;
* backward: dz, dh2, dh1 are rounded where they enter a product; the bias gradients sum
  the unrounded values; rounding is treated as the identity when differentiating.

With ``bf16_rounding=False`` it is the exact network, which ``tests/test_learned_cpu.py`` pins by
central finite differences in float64.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 16, 64, 4
OFF_W1 = 0
OFF_B1 = OFF_W1 + NH * NIN
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
NPARAM = OFF_B3 + NOUT  # 5508


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(np.asarray(a, np.float32))
    u = a32.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(NH, NIN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(NH, NH),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(NOUT, NH), p[OFF_B3:NPARAM])


def init_params(seed=0):
    """He-uniform weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    for off, fan_in, n in ((OFF_W1, NIN, NH * NIN), (OFF_W2, NH, NH * NH), (OFF_W3, NH, NOUT * NH)):
        lim = np.sqrt(6.0 / fan_in)
        p[off:off + n] = rng.uniform(-lim, lim, n).astype(np.float32)
    return p


def _r(a, on):
    return bf16(a) if on else np.asarray(a, np.float64)


def forward(p, X, bf16_rounding=True):
    """Logits [rows, 4] and the intermediates needed by the backward pass."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    q = bf16_rounding
    if q:
        W1r, W2r, W3r = bf16(np.float32(W1)), bf16(np.float32(W2)), bf16(np.float32(W3))
    else:
        W1r, W2r, W3r = W1, W2, W3
    x = _r(np.asarray(X, np.float32), q) if q else np.asarray(X, np.float64)
    a1 = x @ W1r.T + b1
    h1 = _r(np.maximum(a1, 0.0), q)
    a2 = h1 @ W2r.T + b2
    h2 = _r(np.maximum(a2, 0.0), q)
    z = h2 @ W3r.T + b3
    return z, dict(x=x, a1=a1, h1=h1, a2=a2, h2=h2, W2=W2r, W3=W3r)


def loss_grad(p, X, y, class_weight=None, scale=None, bf16_rounding=True):
    """loss = scale * sum_r cw[y_r] * CE(z_r, y_r) and its gradient (float64, [5508]).
    ``scale`` defaults to 1 / sum_r cw[y_r] (the weighted mean)."""
    y = np.asarray(y, np.int64)
    cw = np.ones(NOUT) if class_weight is None else np.asarray(class_weight, np.float64)
    w = cw[y]
    if scale is None:
        scale = 1.0 / w.sum()
    q = bf16_rounding
    z, c = forward(p, X, q)
    zm = z - z.max(axis=1, keepdims=True)
    e = np.exp(zm)
    se = e.sum(axis=1, keepdims=True)
    sm = e / se
    loss = float(np.sum(scale * w * (np.log(se[:, 0]) - zm[np.arange(len(y)), y])))
    onehot = np.zeros_like(sm)
    onehot[np.arange(len(y)), y] = 1.0
    dz = (scale * w)[:, None] * (sm - onehot)
    dzr = _r(dz, q)
    g = np.zeros(NPARAM)
    g[OFF_B3:NPARAM] = dz.sum(axis=0)
    g[OFF_W3:OFF_B3] = (dzr.T @ c["h2"]).ravel()
    d2 = (dzr @ c["W3"]) * (c["a2"] > 0)
    d2r = _r(d2, q)
    g[OFF_B2:OFF_W3] = d2.sum(axis=0)
    g[OFF_W2:OFF_B2] = (d2r.T @ c["h1"]).ravel()
    d1 = (d2r @ c["W2"]) * (c["a1"] > 0)
    d1r = _r(d1, q)
    g[OFF_B1:OFF_W2] = d1.sum(axis=0)
    g[OFF_W1:OFF_B1] = (d1r.T @ c["x"]).ravel()
    return loss, g


def decide(z, tau, level, unit_level):
    """Predicted level per row from logits (the rule of the fused GPU predictor): with
    p = softmax(z) over classes (0: finest level, 1: one coarser, 2: two coarser, 3: none),
    the class is the first c whose cumulative probability p_0 + ... + p_c reaches tau
    (3 if none does); the predicted level is level-1-c, or -1 when that is coarser than
    the unit level or c == 3."""
    z = np.asarray(z, np.float64)
    e = np.exp(z - z.max(axis=1, keepdims=True))
    pr = e / e.sum(axis=1, keepdims=True)
    cum = np.cumsum(pr[:, :3], axis=1)
    c = np.full(len(z), 3)
    for k in (2, 1, 0):
        c = np.where(cum[:, k] >= tau, k, c)
    lev = level - 1 - c
    return np.where((c < 3) & (lev >= unit_level), lev, -1)
