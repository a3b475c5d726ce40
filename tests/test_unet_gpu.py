"""GPU U-Net kernels against the numpy oracle (float64): forward output,
every parameter gradient, the Adam step; and a short training run."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import unet as ounet
from warp_amd.optim import learned

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_backward_vs_oracle(device):
    n = 16
    rng = np.random.default_rng(7)
    p = ounet.init_params(5)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.1
    inp = rng.standard_normal((n, n, n, 2))
    target = rng.standard_normal((n, n, n, 1))
    out_ref, cache = ounet.forward(p, inp)
    L_ref, dout_ref = ounet.loss_mse(out_ref, target)
    g_ref = ounet.backward(p, cache, dout_ref)

    net = learned.UNet(n, device, params={k: v.astype(np.float32) for k, v in p.items()})
    inp_t = torch.tensor(inp.reshape(-1), dtype=torch.float32, device=device)
    tgt_t = torch.tensor(target.reshape(-1), dtype=torch.float32, device=device)
    out = net.forward(inp_t)
    loss, dout = net.loss(out, tgt_t)
    net.backward(dout)
    torch.cuda.synchronize()
    e_out = _rel(out.cpu().numpy(), out_ref.reshape(-1))
    print(f"[unet] forward rel. error {e_out:.2e}, loss {float(loss.item()):.6f} vs {L_ref:.6f}")
    assert e_out < 1e-5
    assert float(loss.item()) == pytest.approx(L_ref, rel=1e-5)
    for k in p:
        e = _rel(net.grads[k].cpu().numpy().reshape(-1), g_ref[k].reshape(-1))
        print(f"[unet] grad {k}: rel. error {e:.2e}")
        assert e < 1e-4, k

    # Adam step against the oracle recurrence
    pw = {k: v.astype(np.float64).copy() for k, v in p.items()}
    m = {k: np.zeros_like(v) for k, v in pw.items()}
    v = {k: np.zeros_like(v) for k, v in pw.items()}
    for k in pw:
        ounet.adam_step(pw[k], net.grads[k].cpu().numpy().astype(np.float64), m[k], v[k], 1e-3, 0.9, 0.999, 0.0, 1e-8)
    net.adam_step(lr=1e-3)
    for k in pw:
        np.testing.assert_allclose(net.params[k].cpu().numpy(), pw[k], rtol=1e-5, atol=1e-7)


def test_training_reduces_loss(device):
    n = 16
    rng = np.random.default_rng(3)
    net = learned.UNet(n, device, seed=1)
    inp = torch.tensor(rng.standard_normal(n**3 * 2), dtype=torch.float32, device=device)
    target = torch.tensor(rng.standard_normal(n**3), dtype=torch.float32, device=device)
    first = float(net.train_step(inp, target, lr=3e-3).item())
    for _ in range(60):
        last = float(net.train_step(inp, target, lr=3e-3).item())
    print(f"[unet] loss {first:.4f} -> {last:.4f}")
    assert last < 0.7 * first
