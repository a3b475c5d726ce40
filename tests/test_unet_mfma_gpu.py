"""BF16 MFMA convolution against the numpy oracle with BF16-rounded operands.

Inputs and weights are rounded to BF16 in the oracle exactly as the kernels
round them, products are summed in float64 there and in FP32 on the GPU, so
FP32 outputs agree to the accumulation error (bounded by 2e-5 of the sum of
absolute terms) and BF16 outputs to one BF16 rounding step.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import unet as ounet
from warp_amd import _native

pytestmark = pytest.mark.gpu

PAIRS = [(2, 16), (16, 32), (32, 64), (64, 64), (64, 32), (32, 16), (16, 1), (16, 16), (1, 16), (16, 2)]
RELU, POOL, F32 = 1, 2, 4


def _bf16_tensor(a, device):
    t = torch.tensor(np.asarray(a, np.float32).reshape(-1), device=device)
    out = torch.empty(t.numel(), dtype=torch.int16, device=device)
    _native.call("warp_amd_f32_to_bf16", t.numel(), _native.ptr(t), _native.ptr(out), _native.stream_of(device))
    return out


def _to_f32(t, device):
    out = torch.empty(t.numel(), dtype=torch.float32, device=device)
    _native.call("warp_amd_bf16_to_f32", t.numel(), _native.ptr(t), _native.ptr(out), _native.stream_of(device))
    return out.cpu().numpy().astype(np.float64)


def _swizzle(W, ci, co, flip, device):
    Wt = torch.tensor(np.asarray(W, np.float32).reshape(-1), device=device)
    nbytes = _native.lib().warp_amd_conv3_mfma_weight_size(ci, co)
    wsw = torch.empty(nbytes, dtype=torch.uint8, device=device)
    _native.call("warp_amd_conv3_mfma_swizzle", ci, co, flip, _native.ptr(Wt), _native.ptr(wsw),
                 _native.stream_of(device))
    return wsw


def _conv(device, m, ci, co, x_bf, wsw, bias, flags, coarse_bf=None):
    s = _native.stream_of(device)
    b = torch.tensor(np.asarray(bias, np.float32), device=device)
    y = torch.full((m**3 * co,), -1, dtype=torch.int16, device=device)
    yf = torch.full((m**3 * co,), float("nan"), dtype=torch.float32, device=device)
    pooled = torch.full(((m // 2) ** 3 * co,), -1, dtype=torch.int16, device=device)
    _native.call("warp_amd_conv3_mfma", m, ci, co, 1 if coarse_bf is not None else 0, _native.ptr(x_bf),
                 _native.ptr(coarse_bf) if coarse_bf is not None else None, _native.ptr(wsw), _native.ptr(b), flags,
                 _native.ptr(y), _native.ptr(yf), _native.ptr(pooled), s)
    torch.cuda.synchronize()
    return y, yf, pooled


def _abs_bound(x, W, b):
    return ounet.conv3_fwd(np.abs(x), np.abs(W), np.abs(b), relu=False)


@pytest.mark.parametrize("ci,co", PAIRS)
@pytest.mark.parametrize("m", [8, 12])
def test_conv_fp32_output(device, ci, co, m):
    rng = np.random.default_rng(1000 * ci + co + m)
    x = ounet.bf16(rng.standard_normal((m, m, m, ci)))
    W = rng.standard_normal((27, ci, co)) / np.sqrt(27 * ci)
    b = rng.standard_normal(co) * 0.1
    Wb = ounet.bf16(W)
    ref = ounet.conv3_fwd(x, Wb, b, relu=False)
    wsw = _swizzle(W, ci, co, 0, device)
    _, yf, _ = _conv(device, m, ci, co, _bf16_tensor(x, device), wsw, b, F32)
    got = yf.cpu().numpy().reshape(ref.shape)
    bound = 2e-5 * _abs_bound(x, Wb, b) + 1e-30
    assert np.all(np.abs(got - ref) <= bound), float(np.max(np.abs(got - ref) / bound))


@pytest.mark.parametrize("ci,co", [p for p in PAIRS if p[1] % 16 == 0])
def test_conv_bf16_relu_pool(device, ci, co):
    m = 8
    rng = np.random.default_rng(7 * ci + co)
    x = ounet.bf16(rng.standard_normal((m, m, m, ci)))
    W = rng.standard_normal((27, ci, co)) / np.sqrt(27 * ci)
    b = rng.standard_normal(co) * 0.1
    Wb = ounet.bf16(W)
    ref = np.maximum(ounet.conv3_fwd(x, Wb, b, relu=False), 0.0)
    y, _, pooled = _conv(device, m, ci, co, _bf16_tensor(x, device), _swizzle(W, ci, co, 0, device), b, RELU | POOL)
    got = _to_f32(y, device).reshape(ref.shape)
    # one BF16 step (2^-8 relative) plus the FP32 accumulation error before rounding
    tol = 2.0**-8 * np.abs(ref) + 2e-5 * _abs_bound(x, Wb, b)
    assert np.all(np.abs(got - ref) <= tol)
    # the pool averages the stored BF16 values
    pref = ounet.pool2(got.reshape(ref.shape))
    pgot = _to_f32(pooled, device).reshape(pref.shape)
    assert np.all(np.abs(pgot - pref) <= 2.0**-8 * np.abs(pref) + 1e-30)


@pytest.mark.parametrize("ci,co", [(64, 32), (32, 16), (16, 1), (64, 64)])
def test_conv_upsample_add_input(device, ci, co):
    m = 8
    rng = np.random.default_rng(11 * ci + co)
    skip = ounet.bf16(rng.standard_normal((m, m, m, ci)))
    coarse = ounet.bf16(rng.standard_normal((m // 2, m // 2, m // 2, ci)))
    x = ounet.bf16(skip + ounet.up2(coarse))
    W = rng.standard_normal((27, ci, co)) / np.sqrt(27 * ci)
    b = rng.standard_normal(co) * 0.1
    Wb = ounet.bf16(W)
    ref = ounet.conv3_fwd(x, Wb, b, relu=False)
    _, yf, _ = _conv(device, m, ci, co, _bf16_tensor(skip, device), _swizzle(W, ci, co, 0, device), b, F32,
                     coarse_bf=_bf16_tensor(coarse, device))
    got = yf.cpu().numpy().reshape(ref.shape)
    assert np.all(np.abs(got - ref) <= 2e-5 * _abs_bound(x, Wb, b) + 1e-30)


@pytest.mark.parametrize("lin,lout", [(16, 32), (32, 64), (64, 64), (64, 32), (32, 16), (16, 1)])
def test_conv_backward_data_operator(device, lin, lout):
    """transpose_flip swizzle applied to dy equals the oracle's input gradient."""
    m = 8
    rng = np.random.default_rng(lin + 3 * lout)
    W = rng.standard_normal((27, lin, lout)) / np.sqrt(27 * lin)
    Wb = ounet.bf16(W)
    dy = ounet.bf16(rng.standard_normal((m, m, m, lout)))
    dx_ref, _, _ = ounet.conv3_bwd(np.zeros((m, m, m, lin)), Wb, None, dy, relu=False)
    wsw = _swizzle(W, lout, lin, 1, device)
    _, yf, _ = _conv(device, m, lout, lin, _bf16_tensor(dy, device), wsw, np.zeros(lin), F32)
    got = yf.cpu().numpy().reshape(dx_ref.shape)
    Wflip = np.transpose(Wb[::-1], (0, 2, 1))  # W'[t][co][ci] = W[26 - t][ci][co]
    bound = 2e-5 * ounet.conv3_fwd(np.abs(dy), np.abs(Wflip), np.zeros(lin), relu=False) + 1e-30
    assert np.all(np.abs(got - dx_ref) <= bound)


def test_unet_forward_bf16_vs_oracle(device):
    from warp_amd.optim import learned

    n = 32  # the coarsest of the 4 levels needs a side of at least 4
    rng = np.random.default_rng(21)
    p = ounet.init_params(3)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.05
    inp = rng.standard_normal((n, n, n, 2))
    ref = ounet.forward_bf16(p, inp)[..., 0]
    ref64, _ = ounet.forward(p, inp)
    net = learned.UNet(n, device, params={k: v.astype(np.float32) for k, v in p.items()})
    out = net.forward_bf16(_bf16_tensor(inp, device))
    got = out.cpu().numpy().astype(np.float64).reshape(ref.shape)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    # identical rounding points; only FP32-vs-FP64 accumulation can flip a BF16 rounding
    assert rel < 5e-3, rel
    # and the BF16 network stays close to the FP64 one
    rel64 = np.linalg.norm(got - ref64[..., 0]) / np.linalg.norm(ref64)
    assert rel64 < 5e-2, rel64


WGRAD_PAIRS = [(2, 16), (16, 32), (32, 64), (64, 64), (64, 32), (32, 16), (16, 1)]


def _wgrad(device, m, ci, co, x_bf, dy, coarse_bf=None):
    nbytes = _native.lib().warp_amd_conv3_mfma_wgrad_workspace_size(m, ci, co)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
    dW = torch.full((27 * ci * co,), float("nan"), dtype=torch.float32, device=device)
    db = torch.full((co,), float("nan"), dtype=torch.float32, device=device)
    dyt = torch.tensor(np.asarray(dy, np.float32).reshape(-1), device=device)
    _native.call("warp_amd_conv3_mfma_wgrad", m, ci, co, 1 if coarse_bf is not None else 0, _native.ptr(x_bf),
                 _native.ptr(coarse_bf) if coarse_bf is not None else None, _native.ptr(dyt), _native.ptr(dW),
                 _native.ptr(db), _native.ptr(ws), ws.numel(), _native.stream_of(device))
    torch.cuda.synchronize()
    return dW.cpu().numpy().reshape(27, ci, co).astype(np.float64), db.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("ci,co", WGRAD_PAIRS)
@pytest.mark.parametrize("m", [8, 12])
def test_wgrad_vs_oracle(device, ci, co, m):
    """Weight / bias gradients on MFMA against the FP64 oracle fed the same
    BF16-rounded input and output gradient; FP32 accumulation error bounded by
    2e-5 of the sum of absolute products."""
    rng = np.random.default_rng(500 * ci + co + m)
    x = ounet.bf16(rng.standard_normal((m, m, m, ci)))
    dy = rng.standard_normal((m, m, m, co))
    dyb = ounet.bf16(dy)
    W0 = np.zeros((27, ci, co))
    _, dW_ref, db_ref = ounet.conv3_bwd(x, W0, None, dyb, relu=False)
    _, dW_abs, db_abs = ounet.conv3_bwd(np.abs(x), W0, None, np.abs(dyb), relu=False)
    dW, db = _wgrad(device, m, ci, co, _bf16_tensor(x, device), dy)
    assert np.all(np.abs(dW - dW_ref) <= 2e-5 * dW_abs + 1e-30), float(np.max(np.abs(dW - dW_ref) / (dW_abs + 1e-30)))
    assert np.all(np.abs(db - db_ref) <= 2e-5 * db_abs + 1e-30)


@pytest.mark.parametrize("ci,co", [(64, 32), (32, 16), (16, 1)])
def test_wgrad_upsample_add_input(device, ci, co):
    m = 8
    rng = np.random.default_rng(41 * ci + co)
    skip = ounet.bf16(rng.standard_normal((m, m, m, ci)))
    coarse = ounet.bf16(rng.standard_normal((m // 2, m // 2, m // 2, ci)))
    x = ounet.bf16(skip + ounet.up2(coarse))
    dy = rng.standard_normal((m, m, m, co))
    dyb = ounet.bf16(dy)
    W0 = np.zeros((27, ci, co))
    _, dW_ref, db_ref = ounet.conv3_bwd(x, W0, None, dyb, relu=False)
    _, dW_abs, _ = ounet.conv3_bwd(np.abs(x), W0, None, np.abs(dyb), relu=False)
    dW, db = _wgrad(device, m, ci, co, _bf16_tensor(skip, device), dy, coarse_bf=_bf16_tensor(coarse, device))
    assert np.all(np.abs(dW - dW_ref) <= 2e-5 * dW_abs + 1e-30)
    np.testing.assert_allclose(db, db_ref, rtol=1e-5, atol=1e-6)


def test_wgrad_deterministic(device):
    """Partial sums added in a fixed order: two calls give the same bits."""
    m, ci, co = 16, 32, 64
    rng = np.random.default_rng(9)
    xb = _bf16_tensor(rng.standard_normal((m, m, m, ci)), device)
    dy = rng.standard_normal((m, m, m, co))
    a = _wgrad(device, m, ci, co, xb, dy)
    b = _wgrad(device, m, ci, co, xb, dy)
    assert a[0].tobytes() == b[0].tobytes() and a[1].tobytes() == b[1].tobytes()


def test_unet_backward_bf16_vs_oracle(device):
    """Every parameter gradient of the MFMA training backward against the FP64
    oracle's.  Each operator is pinned above with the oracle fed the same
    BF16 operands; through the whole chain the BF16 activations, ReLU masks
    and output gradients compound (on these random inputs 0.1-10 % per
    parameter, measured; 0.4-3 % on the trainer's own pairs, see
    test_trainer_gradients_vs_oracle): each within 15 %, the whole gradient's
    direction within 0.5 % (the FP32 backward stays within 1e-4)."""
    from warp_amd.optim import learned

    n = 32
    rng = np.random.default_rng(23)
    p = ounet.init_params(6)
    for k in p:
        if k.endswith(".b"):
            p[k] = rng.standard_normal(p[k].shape) * 0.05
    inp = rng.standard_normal((n, n, n, 2))
    dout = rng.standard_normal((n, n, n, 1))
    _, cache = ounet.forward(p, inp)
    g_ref = ounet.backward(p, cache, dout)
    net = learned.UNet(n, device, params={k: v.astype(np.float32) for k, v in p.items()})
    net.forward_bf16(_bf16_tensor(inp, device))
    net.backward_bf16(torch.tensor(dout.reshape(-1), dtype=torch.float32, device=device))
    torch.cuda.synchronize()
    rels, got_all, ref_all = {}, [], []
    for k in p:
        got = net.grads[k].double().cpu().numpy().reshape(g_ref[k].shape)
        rels[k] = np.linalg.norm(got - g_ref[k]) / max(np.linalg.norm(g_ref[k]), 1e-30)
        print(f"[unet bf16] grad {k}: rel. error {rels[k]:.2e}")
        got_all.append(got.reshape(-1))
        ref_all.append(g_ref[k].reshape(-1))
    g, r = np.concatenate(got_all), np.concatenate(ref_all)
    cos = float(g @ r / (np.linalg.norm(g) * np.linalg.norm(r)))
    print(f"[unet bf16] cosine of the whole gradient {cos:.5f}")
    assert all(v < 0.15 for v in rels.values()), rels
    assert cos > 0.995, cos


@pytest.mark.parametrize("nvox,offset", [(12288, 0), (1001, 0), (4096, 1)])
def test_pack_input_bf16_vs_torch_rounding(device, nvox, offset):
    """Network input (r * inv_scale, logk) as BF16 pairs: 16-byte path (aligned,
    nvox % 4 == 0) and the per-voxel path (odd count, misaligned views), both
    against torch's round-to-nearest-even BF16 conversion."""
    g = torch.Generator().manual_seed(nvox + offset)
    r_h = torch.randn(nvox + offset, generator=g, dtype=torch.float64) * 3.0
    k_h = torch.randn(nvox + offset, generator=g)
    r = r_h.to(device)[offset:]
    logk = k_h.to(device)[offset:]
    inv = torch.tensor([0.37], dtype=torch.float64, device=device)
    out = torch.empty(2 * nvox, dtype=torch.int16, device=device)
    _native.call("warp_amd_unet_pack_input_bf16", nvox, _native.ptr(r), _native.ptr(logk), _native.ptr(inv),
                 _native.ptr(out), _native.stream_of(device))
    torch.cuda.synchronize()
    want = torch.stack([(r_h[offset:] * 0.37).float().to(torch.bfloat16), k_h[offset:].to(torch.bfloat16)], 1)
    assert torch.equal(out.cpu(), want.view(torch.int16).reshape(-1))
