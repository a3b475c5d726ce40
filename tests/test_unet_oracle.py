"""Pin the numpy U-Net oracle (CPU only).

Forward convolution against torch.nn.functional.conv3d in float64 (an
independent implementation), gradients of the whole network against central
finite differences, pool / upsample against their adjoint identities, Adam
against its closed-form first step.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import unet


def test_conv_matches_torch_conv3d():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((6, 5, 7, 3))
    W = rng.standard_normal((27, 3, 4))
    b = rng.standard_normal(4)
    y = unet.conv3_fwd(x, W, b, relu=False)
    # torch: (N, C, D, H, W) with weight (Co, Ci, kz, ky, kx), cross-correlation
    w_t = torch.tensor(W.reshape(3, 3, 3, 3, 4)).permute(4, 3, 0, 1, 2)
    x_t = torch.tensor(x).permute(3, 0, 1, 2)[None]
    y_t = torch.nn.functional.conv3d(x_t, w_t, torch.tensor(b), padding=1)[0].permute(1, 2, 3, 0).numpy()
    np.testing.assert_allclose(y, y_t, rtol=1e-12, atol=1e-12)


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((8, 8, 8, 3))
    c = rng.standard_normal((4, 4, 4, 3))
    assert np.dot(unet.pool2(x).ravel(), c.ravel()) == pytest.approx(np.dot(x.ravel(), unet.pool2_bwd(c).ravel()))
    assert np.dot(unet.up2(c).ravel(), x.ravel()) == pytest.approx(np.dot(c.ravel(), unet.up2_bwd(x).ravel()))
    assert np.allclose(unet.pool2(unet.up2(c)), c)


def test_network_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    p = unet.init_params(3)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1  # non-zero biases exercise db
    inp = rng.standard_normal((8, 8, 8, 2))
    target = rng.standard_normal((8, 8, 8, 1))
    out, cache = unet.forward(p, inp)
    L, dout = unet.loss_mse(out, target)
    g = unet.backward(p, cache, dout)

    def loss_of(q):
        o, _ = unet.forward(q, inp)
        return unet.loss_mse(o, target)[0]

    h = 1e-6
    for name in p:
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(4, flat.size), replace=False):
            orig = flat[idx]
            flat[idx] = orig + h
            lp = loss_of(p)
            flat[idx] = orig - h
            lm = loss_of(p)
            flat[idx] = orig
            fd = (lp - lm) / (2 * h)
            an = g[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (name, idx, fd, an)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(4)
    param = rng.standard_normal(10)
    grad = rng.standard_normal(10)
    m = np.zeros(10)
    v = np.zeros(10)
    p0 = param.copy()
    # first step: bias correction makes mhat = g and vhat = g^2 for any beta
    unet.adam_step(param, grad, m, v, lr=0.01, beta1=0.9, beta2=0.999, t=0.0, eps=1e-8)
    np.testing.assert_allclose(param, p0 - 0.01 * grad / (np.abs(grad) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(m, 0.1 * grad)
    np.testing.assert_allclose(v, 0.001 * grad * grad)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-30, 30, 10000),
                        [1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, -(1.0 + 2.0**-8), 0.0, -0.0]])
    a32 = a.astype(np.float32)
    expect = torch.tensor(a32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(unet.bf16(a32), expect)


def _fv7_small(n, seed):
    """Small SPD 7-point operator (harmonic-mean faces, Dirichlet) as a dense matrix."""
    rng = np.random.default_rng(seed)
    k = np.exp(rng.standard_normal((n, n, n)))
    N = n**3
    A = np.zeros((N, N))
    idx = lambda x, y, z: x + n * (y + n * z)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                i = idx(x, y, z)
                for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
                    xx, yy, zz = x + d[0], y + d[1], z + d[2]
                    if 0 <= xx < n and 0 <= yy < n and 0 <= zz < n:
                        j = idx(xx, yy, zz)
                        t = 2 * k[z, y, x] * k[zz, yy, xx] / (k[z, y, x] + k[zz, yy, xx])
                        A[i, j] -= t
                        A[i, i] += t
                    else:
                        A[i, i] += 2 * k[z, y, x]
    return A, np.log(k).reshape(-1)


def test_energy_loss_gradient_and_bounds():
    rng = np.random.default_rng(8)
    A, _ = _fv7_small(4, 1)
    e = rng.standard_normal(A.shape[0])
    r = A @ e
    z = rng.standard_normal(A.shape[0])
    L, g = unet.energy_loss(r, e, z, A @ z)
    assert 0.0 <= L <= 1.0
    # the exact error is a perfect direction, at any scale and sign
    assert abs(unet.energy_loss(r, e, -3.0 * e, A @ (-3.0 * e))[0]) < 1e-12
    # L = 1 - ||e_min||_A^2 / ||e||_A^2 after the exact line search along z
    alpha = (r @ z) / (z @ A @ z)
    e1 = e - alpha * z
    assert L == pytest.approx((e1 @ A @ e1) / (e @ A @ e), rel=1e-10)
    h = 1e-6
    for i in rng.choice(z.size, 5, replace=False):
        zp, zm = z.copy(), z.copy()
        zp[i] += h
        zm[i] -= h
        fd = (unet.energy_loss(r, e, zp, A @ zp)[0] - unet.energy_loss(r, e, zm, A @ zm)[0]) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9)


@pytest.mark.parametrize("nu", [0, 1, 2])
def test_composite_gradients_match_finite_differences(nu):
    n = 8
    rng = np.random.default_rng(10 + nu)
    A, logk = _fv7_small(n, 3)
    dinv = 1.0 / np.diag(A)
    e = rng.standard_normal(n**3)
    r = A @ e
    p = unet.init_params(4)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    lk = (logk - logk.mean()) / logk.std()
    gain = 1.0 / np.mean(np.diag(A))
    L, g = unet.composite_grads(p, A, dinv, lk, r, e, nu, 2.0 / 3.0, gain, n)

    def loss_of(q):
        z, _, _ = unet.composite_forward(q, A, dinv, lk, r, nu, 2.0 / 3.0, gain, n)
        return unet.energy_loss(r, e, z, A @ z)[0]

    h = 1e-6
    for name in ("e0.W", "e3.W", "d0.W", "d0.b", "d1.b"):
        flat = p[name].reshape(-1)
        for idx in rng.choice(flat.size, size=min(3, flat.size), replace=False):
            orig = flat[idx]
            flat[idx] = orig + h
            lp = loss_of(p)
            flat[idx] = orig - h
            lm = loss_of(p)
            flat[idx] = orig
            fd = (lp - lm) / (2 * h)
            an = g[name].reshape(-1)[idx]
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(fd)) + 1e-8, (nu, name, idx, fd, an)
