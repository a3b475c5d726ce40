"""Learned flexible CG: energy-loss kernel and trainer gradients against the
numpy oracle (FP64), stagnation fallback, and the gain of a trained network
over the same solver without it."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sps
import torch

from oracle import unet as ounet
from warp_amd import _native
from warp_amd.optim import learned

pytestmark = pytest.mark.gpu


def _csr(system):
    A = system.A
    N = A.nrow
    return sps.csr_matrix((A.values.double().cpu().numpy(), A.columns.cpu().numpy(), A.offsets.cpu().numpy()),
                          shape=(N, N))


def test_energy_loss_kernel_vs_oracle(device):
    rng = np.random.default_rng(0)
    n = 5000
    r, az, e, z = (rng.standard_normal(n) for _ in range(4))
    sc = np.array([r @ z, z @ az, r @ e])
    L_ref, g_ref = ounet.energy_loss(r, e, z, az)
    # keep every device tensor referenced until the kernel has run
    rt, azt, sct = (torch.tensor(a, dtype=torch.float64, device=device) for a in (r, az, sc))
    g = torch.empty(n, dtype=torch.float64, device=device)
    loss = torch.zeros(1, dtype=torch.float64, device=device)
    _native.call("warp_amd_energy_loss_grad", n, _native.ptr(rt), _native.ptr(azt), _native.ptr(sct),
                 _native.ptr(g), _native.ptr(loss), _native.stream_of(device))
    torch.cuda.synchronize()
    assert float(loss.item()) == pytest.approx(L_ref, rel=1e-13)
    np.testing.assert_allclose(g.cpu().numpy(), g_ref, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("nu,precision,tol", [(0, "fp32", 2e-3), (1, "fp32", 2e-3), (1, "bf16", 5e-2)])
def test_trainer_gradients_vs_oracle(device, nu, precision, tol):
    n = 32
    system = learned.FV7System(n, 1.5, 2.0, seed=11, device=device)
    A = _csr(system)
    rng = np.random.default_rng(3)
    e = rng.standard_normal(n**3)
    r = A @ e
    p = {k: v.astype(np.float64) for k, v in learned.init_params(2).items()}
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.05
    net = learned.UNet(n, device, params={k: v.astype(np.float32) for k, v in p.items()})
    trainer = learned.EnergyTrainer(net, smooth_steps=nu, precision=precision)
    loss = trainer.step(system, torch.tensor(r, device=device), torch.tensor(e, device=device), lr=0.0)
    torch.cuda.synchronize()
    dinv = system.dinv.cpu().numpy()
    lk = system.logk_std.double().cpu().numpy()
    # the oracle sees the FP32 parameters and FP32-rounded logk the kernels see
    p32 = {k: v.astype(np.float32).astype(np.float64) for k, v in p.items()}
    L_ref, g_ref = ounet.composite_grads(p32, A, dinv, lk, r, e, nu, trainer.omega, 1.0 / system.dmean, n)
    assert float(loss.item()) == pytest.approx(L_ref, rel=1e-4 if precision == "fp32" else 2e-2)
    for k in p:
        got = net.grads[k].double().cpu().numpy().reshape(g_ref[k].shape)
        rel = np.linalg.norm(got - g_ref[k]) / max(np.linalg.norm(g_ref[k]), 1e-30)
        print(f"[trainer {precision}] grad {k}: rel. error {rel:.2e}")
        assert rel < tol, (k, rel)


def test_zero_network_stagnates_then_falls_back(device):
    """A network with zero weights gives z = the smoothing composite only;
    with nu = 0 it gives z = 0: the stagnation detector hands over to the
    Jacobi-CG fallback, which still reaches the FP64-checked tolerance."""
    n = 32
    system = learned.FV7System(n, 1.5, 2.0, seed=21, device=device)
    p = {k: np.zeros_like(v) for k, v in learned.init_params(0).items()}
    net = learned.UNet(n, device, params=p)
    fcg = learned.LearnedFCG(system, net, smooth_steps=0)
    x = torch.zeros(n**3, dtype=torch.float64, device=device)
    out = fcg.solve(x, tol=1e-8, maxiter=200, stagnation_window=10)
    assert out["iterations"] <= 12
    assert out["fallback_iterations"] > 0
    assert out["true_relative_residual"] < 2e-8


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_training_lowers_the_energy_loss(device, precision):
    """A short run of the training pipeline (plain-CG pairs, then the FCG's own
    iterates) lowers the energy loss of held-out pairs, with the network's
    training arithmetic in FP32 kernels or on BF16 MFMA."""
    n = 32
    systems = [learned.FV7System(n, 1.5, 2.0, seed=100 + i, device=device) for i in range(4)]
    held = learned.FV7System(n, 1.5, 2.0, seed=999, device=device)
    pairs = learned.generate_energy_pairs(held, count=10)
    assert len(pairs) == 10
    net = learned.UNet(n, device, seed=0)
    trainer = learned.EnergyTrainer(net, smooth_steps=1)
    before = np.mean([trainer.loss(held, r, e) for r, e in pairs])
    result = learned.train_learned_preconditioner(net, systems, smooth_steps=1, steps=300, dagger_rounds=1,
                                                  dagger_steps=100, seed=0, precision=precision)
    assert result["pcg_pairs"] == 40 and result["fcg_pairs"] > 0
    after = np.mean([trainer.loss(held, r, e) for r, e in pairs])
    assert after < 0.8 * before, (before, after)


def test_shipped_network_gain_over_smoothing_alone(device):
    """The shipped 128^3 network against the same flexible CG with the network
    output removed (Jacobi smoothing only), on a held-out system: both reach an
    FP64-checked 1e-8, the network cuts the iteration count at least 4x."""
    import os

    ckpt = os.path.join(os.path.dirname(learned.__file__), "..", "data", "fv7_unet_128.npz")
    net, meta = learned.load_checkpoint(ckpt, 128, device)
    test = learned.FV7System(128, float(meta["sigma"]), float(meta["corr"]), seed=5003, device=device)
    x = torch.zeros(128**3, dtype=torch.float64, device=device)
    trained = learned.LearnedFCG(test, net, smooth_steps=int(meta["smooth_steps"])).solve(
        x, tol=1e-8, maxiter=2000, fallback=False, stagnation_window=1000)
    zero = learned.UNet(128, device, params={k: np.zeros_like(v) for k, v in learned.init_params(0).items()})
    x.zero_()
    smooth = learned.LearnedFCG(test, zero, smooth_steps=int(meta["smooth_steps"])).solve(
        x, tol=1e-8, maxiter=4000, fallback=False, stagnation_window=10000)
    assert trained["true_relative_residual"] <= 1e-8
    assert smooth["true_relative_residual"] <= 1e-8
    assert trained["iterations"] * 4 <= smooth["iterations"], (trained["iterations"], smooth["iterations"])


def test_fused_iteration_matches_unfused(device):
    """The fused flexible-CG iteration (3 fused SpMV kernels, FP32 sweep and
    preconditioned residual) follows the unfused one (FP64 diag_mv, residual,
    sweep, dots, direction, SpMV); the preconditioner differs only by FP32
    rounding, so the residual histories agree to ~1e-3."""
    n = 32
    system = learned.FV7System(n, 1.5, 2.0, seed=31, device=device)
    net = learned.UNet(n, device, seed=4)
    hist = {}
    for fused in (False, True):
        fcg = learned.LearnedFCG(system, net, smooth_steps=1, fused=fused)
        x = torch.zeros(n**3, dtype=torch.float64, device=device)
        out = fcg.solve(x, tol=1e-30, maxiter=8, use_graph=False, fallback=False, stagnation_window=1000)
        hist[fused] = np.array(out["history"])
    np.testing.assert_allclose(hist[True], hist[False], rtol=1e-3)
