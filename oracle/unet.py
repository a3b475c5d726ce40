"""CPU oracle of the learned preconditioner -- TEST INFRASTRUCTURE ONLY.

Plain numpy (float64) restatement of the 3D U-Net used as a flexible-CG
preconditioner, its backward pass and the Adam step.  No reference
counterpart exists for the network (the reference has no learned
preconditioner); Adam follows warp/_src/optim/adam.py:30-47.  The gradients
here are pinned by finite differences in tests/test_unet_oracle.py.

Layout: channels-last, a field on an n^3 grid is an array (n, n, n, C)
indexed [z, y, x, c] with x fastest (cell (x, y, z) -> x + n y + n^2 z).
Convolutions are 3x3x3, stride 1, zero padding; weights W[tap, ci, co] with
tap = (dz+1)*9 + (dy+1)*3 + (dx+1); y[v, co] = b[co] + sum_tap,ci
x[v + d(tap), ci] W[tap, ci, co].
"""

from __future__ import annotations

import numpy as np

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]

# (name, c_in, c_out, relu) in execution order; see forward() for the wiring.
LAYERS = [
    ("e0", 2, 16, True),
    ("e1", 16, 32, True),
    ("e2", 32, 64, True),
    ("e3", 64, 64, True),
    ("d2", 64, 32, True),
    ("d1", 32, 16, True),
    ("d0", 16, 1, False),
]


def _shift(x, d):
    """out[v] = x[v + d] with zeros outside (d = (dz, dy, dx))."""
    out = np.zeros_like(x)
    n = x.shape[:3]
    src = []
    dst = []
    for k in range(3):
        if d[k] >= 0:
            src.append(slice(d[k], n[k]))
            dst.append(slice(0, n[k] - d[k]))
        else:
            src.append(slice(0, n[k] + d[k]))
            dst.append(slice(-d[k], n[k]))
    out[tuple(dst)] = x[tuple(src)]
    return out


def conv3_fwd(x, W, b, relu):
    y = np.broadcast_to(b, x.shape[:3] + (W.shape[2],)).astype(np.float64).copy()
    for t, d in enumerate(TAPS):
        y += _shift(x, d) @ W[t]
    return np.maximum(y, 0.0) if relu else y


def conv3_bwd(x, W, y, dy, relu):
    """Gradients of a conv layer given its input x, output y and dL/dy."""
    if relu:
        dy = dy * (y > 0)
    db = dy.sum(axis=(0, 1, 2))
    dW = np.zeros_like(W)
    dx = np.zeros_like(x)
    for t, d in enumerate(TAPS):
        xs = _shift(x, d)
        dW[t] = xs.reshape(-1, x.shape[3]).T @ dy.reshape(-1, dy.shape[3])
        dx += _shift(dy @ W[t].T, tuple(-c for c in d))
    return dx, dW, db


def pool2(x):
    n = x.shape[0] // 2
    return x.reshape(n, 2, n, 2, n, 2, x.shape[3]).mean(axis=(1, 3, 5))


def pool2_bwd(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 0), 2, 1), 2, 2) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up2_bwd(dy):
    n = dy.shape[0] // 2
    return dy.reshape(n, 2, n, 2, n, 2, dy.shape[3]).sum(axis=(1, 3, 5))


def bf16(a):
    """Round to BF16 (via float32, round to nearest even); returns float64."""
    u = np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0):
    """He-normal weights, zero biases (float64)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co, relu in LAYERS:
        std = np.sqrt(2.0 / (27 * ci)) if relu else np.sqrt(1.0 / (27 * ci))
        p[name + ".W"] = rng.standard_normal((27, ci, co)) * std
        p[name + ".b"] = np.zeros(co)
    return p


def forward(p, inp):
    """inp: (n, n, n, 2) -> (out (n, n, n, 1), cache)."""
    c = {"inp": inp}
    c["a0"] = conv3_fwd(inp, p["e0.W"], p["e0.b"], True)
    c["p0"] = pool2(c["a0"])
    c["a1"] = conv3_fwd(c["p0"], p["e1.W"], p["e1.b"], True)
    c["p1"] = pool2(c["a1"])
    c["a2"] = conv3_fwd(c["p1"], p["e2.W"], p["e2.b"], True)
    c["p2"] = pool2(c["a2"])
    c["a3"] = conv3_fwd(c["p2"], p["e3.W"], p["e3.b"], True)
    c["u2"] = up2(c["a3"]) + c["a2"]
    c["b2"] = conv3_fwd(c["u2"], p["d2.W"], p["d2.b"], True)
    c["u1"] = up2(c["b2"]) + c["a1"]
    c["b1"] = conv3_fwd(c["u1"], p["d1.W"], p["d1.b"], True)
    c["u0"] = up2(c["b1"]) + c["a0"]
    c["out"] = conv3_fwd(c["u0"], p["d0.W"], p["d0.b"], False)
    return c["out"], c


def forward_bf16(p, inp):
    """The inference network as the BF16 kernels evaluate it: input, weights
    and every stored activation rounded to BF16 (pool of the stored values,
    upsample + skip rounded once), sums in float64, FP32-free final output."""
    W = {k: (bf16(v) if k.endswith(".W") else v) for k, v in p.items()}
    conv = lambda x, name, relu: conv3_fwd(x, W[name + ".W"], W[name + ".b"], relu)
    x = bf16(inp)
    a0 = bf16(conv(x, "e0", True))
    a1 = bf16(conv(bf16(pool2(a0)), "e1", True))
    a2 = bf16(conv(bf16(pool2(a1)), "e2", True))
    a3 = bf16(conv(bf16(pool2(a2)), "e3", True))
    b2 = bf16(conv(bf16(up2(a3) + a2), "d2", True))
    b1 = bf16(conv(bf16(up2(b2) + a1), "d1", True))
    return conv(bf16(up2(b1) + a0), "d0", False)


def backward(p, c, dout):
    """Gradients of all parameters given dL/dout."""
    g = {}
    du0, g["d0.W"], g["d0.b"] = conv3_bwd(c["u0"], p["d0.W"], c["out"], dout, False)
    da0 = du0.copy()
    du1, g["d1.W"], g["d1.b"] = conv3_bwd(c["u1"], p["d1.W"], c["b1"], up2_bwd(du0), True)
    da1 = du1.copy()
    du2, g["d2.W"], g["d2.b"] = conv3_bwd(c["u2"], p["d2.W"], c["b2"], up2_bwd(du1), True)
    da2 = du2.copy()
    dp2, g["e3.W"], g["e3.b"] = conv3_bwd(c["p2"], p["e3.W"], c["a3"], up2_bwd(du2), True)
    da2 += pool2_bwd(dp2)
    dp1, g["e2.W"], g["e2.b"] = conv3_bwd(c["p1"], p["e2.W"], c["a2"], da2, True)
    da1 += pool2_bwd(dp1)
    dp0, g["e1.W"], g["e1.b"] = conv3_bwd(c["p0"], p["e1.W"], c["a1"], da1, True)
    da0 += pool2_bwd(dp0)
    _, g["e0.W"], g["e0.b"] = conv3_bwd(c["inp"], p["e0.W"], c["a0"], da0, True)
    return g


def loss_mse(out, target):
    """L = 0.5 mean((out - target)^2); returns (L, dL/dout)."""
    d = out - target
    return 0.5 * float(np.mean(d * d)), d / d.size


def adam_step(param, grad, m, v, lr, beta1, beta2, t, eps):
    """One Adam update in place (warp/_src/optim/adam.py:30-47: bias correction
    with t + 1, eps added outside the square root)."""
    m[...] = beta1 * m + (1.0 - beta1) * grad
    v[...] = beta2 * v + (1.0 - beta2) * grad * grad
    mhat = m / (1.0 - beta1 ** (t + 1.0))
    vhat = v / (1.0 - beta2 ** (t + 1.0))
    param[...] = param - lr * mhat / (np.sqrt(vhat) + eps)


# ---------------------------------------------------------------------------
# Smoothing composite around the network and its one-step energy loss
# ---------------------------------------------------------------------------


def energy_loss(r, e, z, Az):
    """L = 1 - (r.z)^2 / ((r.e)(z.Az)) and dL/dz (A symmetric, A e = r)."""
    a, b, c = float(r @ z), float(z @ Az), float(r @ e)
    L = 1.0 - a * a / (c * b)
    g = (-2.0 * a / (c * b)) * r + (2.0 * a * a / (c * b * b)) * Az
    return L, g


def composite_forward(p, A, dinv, logk, r, nu, omega, gain, n):
    """z = nu Jacobi sweeps from 0; r1 = r - A z; z += gain s U(r1 / s, logk); nu sweeps."""
    z = np.zeros_like(r)
    for _ in range(nu):
        z = z + omega * dinv * (r - A @ z)
    r1 = r - A @ z
    s = float(np.sqrt(np.mean(r1 * r1)))
    inp = np.stack([r1 / s, logk], axis=-1).reshape(n, n, n, 2)
    out, cache = forward(p, inp)
    z = z + gain * s * out.reshape(-1)
    for _ in range(nu):
        z = z + omega * dinv * (r - A @ z)
    return z, cache, s


def composite_grads(p, A, dinv, logk, r, e, nu, omega, gain, n):
    """Loss and parameter gradients of the composite (adjoint of a sweep: g - A (omega D^-1 g))."""
    z, cache, s = composite_forward(p, A, dinv, logk, r, nu, omega, gain, n)
    L, g = energy_loss(r, e, z, A @ z)
    for _ in range(nu):
        g = g - A @ (omega * dinv * g)
    return L, backward(p, cache, (gain * s * g).reshape(n, n, n, 1))
