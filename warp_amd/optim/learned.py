"""Learned preconditioner: a 3D U-Net trained on the GPU, used inside flexible CG.

No reference counterpart (the reference offers Jacobi / identity
preconditioners only, warp/_src/optim/linear.py:236-280).  Components:

* :class:`UNet` -- 4-level 3D U-Net on an n^3 cell grid (channels 16 / 32 /
  64 / 64, 3x3x3 convolutions, ReLU, 2x average pooling, nearest upsampling
  with additive skips, linear 1-channel output).  FP32 masters with FP32
  forward / backward / Adam kernels for training; BF16 MFMA kernels for
  inference.
* :class:`FV7System` -- the operator family -div(k grad u) = f, 7-point
  finite volumes, k = exp(Gaussian random field), FP32 matrix.
* :class:`LearnedFCG` -- flexible CG (FP64 vectors) whose preconditioner is
  Jacobi smoothing around the network correction; graph-replayed, with a
  stagnation fallback and an FP64 true-residual check.
* :class:`EnergyTrainer`, :func:`train_learned_preconditioner` -- training
  on the one-step energy reduction 1 - (r.z)^2 / ((r.e)(z.Az)) of the
  preconditioned direction, first on (r, e) pairs from plain FP64 Jacobi-CG
  solves, then on pairs from the learned FCG's own iterates.

The network sees (r1 / rms(r1), standardised log k) and its output is scaled
back by rms(r1), so the preconditioner is scale-equivariant.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from .. import _native

# (name, c_in, c_out, relu, level) in execution order
LAYERS = [
    ("e0", 2, 16, True, 0),
    ("e1", 16, 32, True, 1),
    ("e2", 32, 64, True, 2),
    ("e3", 64, 64, True, 3),
    ("d2", 64, 32, True, 2),
    ("d1", 32, 16, True, 1),
    ("d0", 16, 1, False, 0),
]
LAYER_INDEX = {name: (ci, co, relu, lvl) for name, ci, co, relu, lvl in LAYERS}


def init_params(seed: int = 0) -> dict[str, np.ndarray]:
    """He-normal weights W[tap, ci, co] and zero biases (float32 numpy)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, ci, co, relu, _ in LAYERS:
        std = math.sqrt((2.0 if relu else 1.0) / (27 * ci))
        p[name + ".W"] = (rng.standard_normal((27, ci, co)) * std).astype(np.float32)
        p[name + ".b"] = np.zeros(co, np.float32)
    return p


class UNet:
    """U-Net on an n^3 grid (n divisible by 8), FP32 kernels."""

    def __init__(self, n: int, device, params: dict | None = None, seed: int = 0):
        if n % 8:
            raise ValueError("grid side must be divisible by 8")
        self.n = n
        self.dev = torch.device(device)
        src = params if params is not None else init_params(seed)
        self.params = {k: torch.as_tensor(np.asarray(v, np.float32), device=self.dev).contiguous() for k, v in src.items()}
        self.grads = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.t = 0
        self._buf: dict[str, torch.Tensor] = {}
        self._lib = _native.lib()
        self._wsw = None  # swizzled BF16 weights, rebuilt after every update

    # -- buffers -----------------------------------------------------------
    def side(self, level: int) -> int:
        return self.n >> level

    def buf(self, name: str, level: int, ch: int) -> torch.Tensor:
        count = self.side(level) ** 3 * ch
        t = self._buf.get(name)
        if t is None or t.numel() != count:
            t = torch.empty(count, dtype=torch.float32, device=self.dev)
            self._buf[name] = t
        return t

    def _s(self):
        return _native.stream_of(self.dev)

    # -- primitives ----------------------------------------------------------
    def _conv(self, name, x, y, accumulate=False):
        ci, co, relu, lvl = LAYER_INDEX[name]
        _native.call(
            "warp_amd_conv3_fwd_f32", self.side(lvl), ci, co, _native.ptr(x), _native.ptr(self.params[name + ".W"]),
            _native.ptr(self.params[name + ".b"]), _native.ptr(y), 1 if relu else 0, 1 if accumulate else 0, self._s(),
        )

    def _pool(self, lvl, ch, x, y):
        _native.call("warp_amd_pool2_f32", self.side(lvl), ch, _native.ptr(x), _native.ptr(y), self._s())

    def _up_add(self, lvl_coarse, ch, x, skip, y):
        _native.call("warp_amd_up2_add_f32", self.side(lvl_coarse), ch, _native.ptr(x), _native.ptr(skip),
                     _native.ptr(y), self._s())

    # -- forward / backward --------------------------------------------------
    def forward(self, inp: torch.Tensor) -> torch.Tensor:
        """inp: n^3 x 2 FP32 (flat); returns the n^3 output (flat view of an internal buffer)."""
        B = self.buf
        a0, p0 = B("a0", 0, 16), B("p0", 1, 16)
        a1, p1 = B("a1", 1, 32), B("p1", 2, 32)
        a2, p2 = B("a2", 2, 64), B("p2", 3, 64)
        a3 = B("a3", 3, 64)
        u2, b2 = B("u2", 2, 64), B("b2", 2, 32)
        u1, b1 = B("u1", 1, 32), B("b1", 1, 16)
        u0, out = B("u0", 0, 16), B("out", 0, 1)
        self._inp = inp
        self._conv("e0", inp, a0)
        self._pool(0, 16, a0, p0)
        self._conv("e1", p0, a1)
        self._pool(1, 32, a1, p1)
        self._conv("e2", p1, a2)
        self._pool(2, 64, a2, p2)
        self._conv("e3", p2, a3)
        self._up_add(3, 64, a3, a2, u2)
        self._conv("d2", u2, b2)
        self._up_add(2, 32, b2, a1, u1)
        self._conv("d1", u1, b1)
        self._up_add(1, 16, b1, a0, u0)
        self._conv("d0", u0, out)
        return out

    # -- BF16 MFMA inference ---------------------------------------------------
    def sync_bf16(self) -> None:
        """Refresh the swizzled BF16 weights from the FP32 masters (after training)."""
        s = self._s()
        self._wsw = {}
        for name, ci, co, _, _ in LAYERS:
            nbytes = self._lib.warp_amd_conv3_mfma_weight_size(ci, co)
            t = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
            _native.call("warp_amd_conv3_mfma_swizzle", ci, co, 0, _native.ptr(self.params[name + ".W"]),
                         _native.ptr(t), s)
            self._wsw[name] = t

    def _hbuf(self, name: str, level: int, ch: int) -> torch.Tensor:
        """BF16 activation buffer (stored as int16 bit patterns)."""
        count = self.side(level) ** 3 * ch
        key = "h:" + name
        t = self._buf.get(key)
        if t is None or t.numel() != count:
            t = torch.empty(count, dtype=torch.int16, device=self.dev)
            self._buf[key] = t
        return t

    def _mconv(self, name, x, y=None, pooled=None, coarse=None, yf=None):
        ci, co, relu, lvl = LAYER_INDEX[name]
        flags = (1 if relu else 0) | (2 if pooled is not None else 0) | (4 if yf is not None else 0)
        _native.call(
            "warp_amd_conv3_mfma", self.side(lvl), ci, co, 1 if coarse is not None else 0, _native.ptr(x),
            _native.ptr(coarse) if coarse is not None else None, _native.ptr(self._wsw[name]),
            _native.ptr(self.params[name + ".b"]), flags, _native.ptr(y) if y is not None else None,
            _native.ptr(yf) if yf is not None else None, _native.ptr(pooled) if pooled is not None else None,
            self._s(),
        )

    def forward_bf16(self, inp_bf16: torch.Tensor) -> torch.Tensor:
        """inp_bf16: BF16 [n^3][2] (int16 bit patterns); returns the FP32 n^3 output.

        Seven MFMA launches: pooling is fused into the encoder convolutions and
        the upsample + skip add into the decoder convolutions' operand loads.
        The grid side must be a multiple of 32 (coarsest level side >= 4)."""
        if self.n % 32:
            raise ValueError("the BF16 U-Net needs a grid side divisible by 32")
        if getattr(self, "_wsw", None) is None:
            self.sync_bf16()
        H = self._hbuf
        a0, p0 = H("a0", 0, 16), H("p0", 1, 16)
        a1, p1 = H("a1", 1, 32), H("p1", 2, 32)
        a2, p2 = H("a2", 2, 64), H("p2", 3, 64)
        a3 = H("a3", 3, 64)
        b2, b1 = H("b2", 2, 32), H("b1", 1, 16)
        out = self.buf("out16", 0, 1)
        self._inp_bf16 = inp_bf16
        self._mconv("e0", inp_bf16, y=a0, pooled=p0)
        self._mconv("e1", p0, y=a1, pooled=p1)
        self._mconv("e2", p1, y=a2, pooled=p2)
        self._mconv("e3", p2, y=a3)
        self._mconv("d2", a2, y=b2, coarse=a3)
        self._mconv("d1", a1, y=b1, coarse=b2)
        self._mconv("d0", a0, yf=out, coarse=b1)
        return out

    # -- BF16 MFMA training ------------------------------------------------------
    def sync_bf16_flipped(self) -> None:
        """Swizzled BF16 backward-data weights (each layer's flipped transpose)."""
        s = self._s()
        self._wsw_t = {}
        for name, ci, co, _, _ in LAYERS:
            if name == "e0":
                continue  # no gradient with respect to the network input
            nbytes = self._lib.warp_amd_conv3_mfma_weight_size(co, ci)
            t = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
            _native.call("warp_amd_conv3_mfma_swizzle", co, ci, 1, _native.ptr(self.params[name + ".W"]),
                         _native.ptr(t), s)
            self._wsw_t[name] = t

    def _wgrad_bf16(self, name, x, dy, coarse=None):
        """dW, db of a layer on MFMA from its BF16 input (x + upsample(coarse) for decoders) and FP32 dy."""
        ci, co, _, lvl = LAYER_INDEX[name]
        n = self.side(lvl)
        ws_bytes = self._lib.warp_amd_conv3_mfma_wgrad_workspace_size(n, ci, co)
        t = self._buf.get("wgrad_mfma_ws")
        if t is None or t.numel() < ws_bytes:
            t = torch.empty(ws_bytes, dtype=torch.uint8, device=self.dev)
            self._buf["wgrad_mfma_ws"] = t
        _native.call(
            "warp_amd_conv3_mfma_wgrad", n, ci, co, 1 if coarse is not None else 0, _native.ptr(x),
            _native.ptr(coarse) if coarse is not None else None, _native.ptr(dy), _native.ptr(self.grads[name + ".W"]),
            _native.ptr(self.grads[name + ".b"]), _native.ptr(t), t.numel(), self._s(),
        )

    def _bwd_data_bf16(self, name, dy, dx):
        """dx = the input gradient of a layer (FP32) from its FP32 output gradient dy, on MFMA."""
        ci, co, _, lvl = LAYER_INDEX[name]
        n = self.side(lvl)
        s = self._s()
        h = self._hbuf("gy_" + name, lvl, co)
        _native.call("warp_amd_f32_to_bf16", h.numel(), _native.ptr(dy), _native.ptr(h), s)
        _native.call("warp_amd_conv3_mfma", n, co, ci, 0, _native.ptr(h), None, _native.ptr(self._wsw_t[name]), None,
                     4, None, _native.ptr(dx), None, s)

    def _relu_bf16(self, y_bf16, dy, out):
        _native.call("warp_amd_relu_mask_bf16", dy.numel(), _native.ptr(y_bf16), _native.ptr(dy), _native.ptr(out),
                     self._s())
        return out

    def backward_bf16(self, dout: torch.Tensor) -> None:
        """Gradients of all parameters (into self.grads) for dL/dout after :meth:`forward_bf16`.

        Weight and bias gradients on MFMA from the BF16 layer inputs the
        forward formed (decoder inputs skip + upsample(coarse) re-formed the
        same way) and the output gradients rounded to BF16; input gradients
        by the MFMA convolution with flipped weights (FP32 out); ReLU masks
        from the stored BF16 activations; pooling / upsampling adjoints and
        the skip sums in FP32."""
        if getattr(self, "_wsw_t", None) is None:
            self.sync_bf16_flipped()
        H, B = self._hbuf, self.buf
        s = self._s()
        a0, p0 = H("a0", 0, 16), H("p0", 1, 16)
        a1, p1 = H("a1", 1, 32), H("p1", 2, 32)
        a2, p2 = H("a2", 2, 64), H("p2", 3, 64)
        a3 = H("a3", 3, 64)
        b2, b1 = H("b2", 2, 32), H("b1", 1, 16)
        # d0 (linear): input a0 + up(b1)
        self._wgrad_bf16("d0", a0, dout, coarse=b1)
        da0 = B("da0", 0, 16)
        self._bwd_data_bf16("d0", dout, da0)
        # d1: input a1 + up(b2), output b1
        g1 = B("g1", 1, 16)
        _native.call("warp_amd_up2_bwd_f32", self.side(1), 16, _native.ptr(da0), _native.ptr(g1), s)
        dy = self._relu_bf16(b1, g1, B("m1", 1, 16))
        self._wgrad_bf16("d1", a1, dy, coarse=b2)
        da1 = B("da1", 1, 32)
        self._bwd_data_bf16("d1", dy, da1)
        # d2: input a2 + up(a3), output b2
        g2 = B("g2", 2, 32)
        _native.call("warp_amd_up2_bwd_f32", self.side(2), 32, _native.ptr(da1), _native.ptr(g2), s)
        dy = self._relu_bf16(b2, g2, B("m2", 2, 32))
        self._wgrad_bf16("d2", a2, dy, coarse=a3)
        da2 = B("da2", 2, 64)
        self._bwd_data_bf16("d2", dy, da2)
        # e3: input p2, output a3
        g3 = B("g3", 3, 64)
        _native.call("warp_amd_up2_bwd_f32", self.side(3), 64, _native.ptr(da2), _native.ptr(g3), s)
        dy = self._relu_bf16(a3, g3, B("m3", 3, 64))
        self._wgrad_bf16("e3", p2, dy)
        dp2 = B("dp2", 3, 64)
        self._bwd_data_bf16("e3", dy, dp2)
        _native.call("warp_amd_pool2_bwd_f32", self.side(2), 64, _native.ptr(dp2), _native.ptr(da2), 1, s)
        # e2: input p1, output a2
        dy = self._relu_bf16(a2, da2, B("m2b", 2, 64))
        self._wgrad_bf16("e2", p1, dy)
        dp1 = B("dp1", 2, 32)
        self._bwd_data_bf16("e2", dy, dp1)
        _native.call("warp_amd_pool2_bwd_f32", self.side(1), 32, _native.ptr(dp1), _native.ptr(da1), 1, s)
        # e1: input p0, output a1
        dy = self._relu_bf16(a1, da1, B("m1b", 1, 32))
        self._wgrad_bf16("e1", p0, dy)
        dp0 = B("dp0", 1, 16)
        self._bwd_data_bf16("e1", dy, dp0)
        _native.call("warp_amd_pool2_bwd_f32", self.side(0), 16, _native.ptr(dp0), _native.ptr(da0), 1, s)
        # e0: input (BF16 network input), output a0
        dy = self._relu_bf16(a0, da0, B("m0", 0, 16))
        self._wgrad_bf16("e0", self._inp_bf16, dy)

    def _layer_grads(self, name, x, y, dy, masked):
        """Weight / bias gradients of one layer; returns dy masked by the ReLU."""
        ci, co, relu, lvl = LAYER_INDEX[name]
        n = self.side(lvl)
        if relu:
            _native.call("warp_amd_relu_mask_f32", n**3 * co, _native.ptr(y), _native.ptr(dy), _native.ptr(masked), self._s())
            dy = masked
        ws_bytes = self._lib.warp_amd_conv3_bwd_weight_workspace_size(n, ci, co)
        key = "wgrad_ws"
        t = self._buf.get(key)
        if t is None or t.numel() < ws_bytes:
            t = torch.empty(ws_bytes, dtype=torch.uint8, device=self.dev)
            self._buf[key] = t
        _native.call(
            "warp_amd_conv3_bwd_weight_f32", n, ci, co, _native.ptr(x), _native.ptr(dy),
            _native.ptr(self.grads[name + ".W"]), _native.ptr(self.grads[name + ".b"]), _native.ptr(t), t.numel(),
            self._s(),
        )
        return dy

    def _conv_bwd_data(self, name, dy, dx, accumulate=False):
        ci, co, _, lvl = LAYER_INDEX[name]
        _native.call(
            "warp_amd_conv3_bwd_data_f32", self.side(lvl), ci, co, _native.ptr(dy),
            _native.ptr(self.params[name + ".W"]), _native.ptr(dx), 1 if accumulate else 0, self._s(),
        )

    def backward(self, dout: torch.Tensor) -> None:
        """Gradients of all parameters (into self.grads) for dL/dout (after forward)."""
        B = self.buf
        s = self._s()
        # decoder, level 0: d0 (linear)
        self._layer_grads("d0", B("u0", 0, 16), None, dout, None)
        da0 = B("da0", 0, 16)
        self._conv_bwd_data("d0", dout, da0)
        # d1
        g1 = B("g1", 1, 16)
        _native.call("warp_amd_up2_bwd_f32", self.side(1), 16, _native.ptr(da0), _native.ptr(g1), s)
        m1 = B("m1", 1, 16)
        dy = self._layer_grads("d1", B("u1", 1, 32), B("b1", 1, 16), g1, m1)
        da1 = B("da1", 1, 32)
        self._conv_bwd_data("d1", dy, da1)
        # d2
        g2 = B("g2", 2, 32)
        _native.call("warp_amd_up2_bwd_f32", self.side(2), 32, _native.ptr(da1), _native.ptr(g2), s)
        m2 = B("m2", 2, 32)
        dy = self._layer_grads("d2", B("u2", 2, 64), B("b2", 2, 32), g2, m2)
        da2 = B("da2", 2, 64)
        self._conv_bwd_data("d2", dy, da2)
        # e3
        g3 = B("g3", 3, 64)
        _native.call("warp_amd_up2_bwd_f32", self.side(3), 64, _native.ptr(da2), _native.ptr(g3), s)
        m3 = B("m3", 3, 64)
        dy = self._layer_grads("e3", B("p2", 3, 64), B("a3", 3, 64), g3, m3)
        dp2 = B("dp2", 3, 64)
        self._conv_bwd_data("e3", dy, dp2)
        _native.call("warp_amd_pool2_bwd_f32", self.side(2), 64, _native.ptr(dp2), _native.ptr(da2), 1, s)
        # e2
        m2b = B("m2b", 2, 64)
        dy = self._layer_grads("e2", B("p1", 2, 32), B("a2", 2, 64), da2, m2b)
        dp1 = B("dp1", 2, 32)
        self._conv_bwd_data("e2", dy, dp1)
        _native.call("warp_amd_pool2_bwd_f32", self.side(1), 32, _native.ptr(dp1), _native.ptr(da1), 1, s)
        # e1
        m1b = B("m1b", 1, 32)
        dy = self._layer_grads("e1", B("p0", 1, 16), B("a1", 1, 32), da1, m1b)
        dp0 = B("dp0", 1, 16)
        self._conv_bwd_data("e1", dy, dp0)
        _native.call("warp_amd_pool2_bwd_f32", self.side(0), 16, _native.ptr(dp0), _native.ptr(da0), 1, s)
        # e0
        m0 = B("m0", 0, 16)
        self._layer_grads("e0", self._inp, B("a0", 0, 16), da0, m0)

    def loss(self, out: torch.Tensor, target: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        """Returns (device loss scalar, dL/dout) for L = 0.5 mean((out - target)^2)."""
        count = out.numel()
        dout = self.buf("dout", 0, 1)
        loss = self._buf.get("loss")
        if loss is None:
            loss = torch.zeros(1, dtype=torch.float64, device=self.dev)
            self._buf["loss"] = loss
        ws_bytes = self._lib.warp_amd_mse_workspace_size(count)
        ws = self._buf.get("mse_ws")
        if ws is None or ws.numel() < ws_bytes:
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self.dev)
            self._buf["mse_ws"] = ws
        _native.call("warp_amd_mse_loss_f32", count, _native.ptr(out), _native.ptr(target), _native.ptr(dout),
                     _native.ptr(loss), _native.ptr(ws), ws.numel(), self._s())
        return loss, dout

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        self._wsw = None
        self._wsw_t = None
        for k, p in self.params.items():
            _native.call("warp_amd_adam_step_f32", p.numel(), _native.ptr(self.grads[k]), _native.ptr(self.m[k]),
                         _native.ptr(self.v[k]), _native.ptr(p), lr, beta1, beta2, float(self.t), eps, self._s())
        self.t += 1

    def train_step(self, inp, target, lr=1e-3) -> torch.Tensor:
        out = self.forward(inp)
        loss, dout = self.loss(out, target)
        self.backward(dout)
        self.adam_step(lr)
        return loss

    # -- checkpoint ----------------------------------------------------------
    def state_dict(self) -> dict[str, np.ndarray]:
        d = {k: v.detach().cpu().numpy() for k, v in self.params.items()}
        d.update({"adam.m." + k: v.cpu().numpy() for k, v in self.m.items()})
        d.update({"adam.v." + k: v.cpu().numpy() for k, v in self.v.items()})
        d["adam.t"] = np.array([self.t])
        return d

    def save(self, path: str) -> None:
        np.savez(path, **self.state_dict())

    @classmethod
    def load(cls, path: str, n: int, device) -> "UNet":
        with np.load(path, allow_pickle=False) as f:
            params = {k: f[k] for k in f.files if k.split(".")[0] in LAYER_INDEX}
            net = cls(n, device, params=params)
            for k in net.params:
                if "adam.m." + k in f.files:
                    net.m[k].copy_(torch.from_numpy(f["adam.m." + k]))
                    net.v[k].copy_(torch.from_numpy(f["adam.v." + k]))
            if "adam.t" in f.files:
                net.t = int(f["adam.t"][0])
        return net


# ---------------------------------------------------------------------------
# Operator family, training pairs, training
# ---------------------------------------------------------------------------


class FV7System:
    """-div(k grad u) = f on an n^3 grid (7-point FV, k = exp(GRF)), FP32 matrix.

    ``b`` is FP64; ``logk_std`` is the standardised log-coefficient fed to
    the network.
    """

    def __init__(self, n: int, sigma: float, corr_len: float, seed: int, device):
        from .. import problems, sparse
        from .. import types as wt

        dev = torch.device(device)
        self.n = n
        off, col, val, b, k = problems.fv7_problem(n, sigma, corr_len, seed, seed + 1_000_003, device=dev)
        A = sparse.bsr_zeros(n**3, n**3, wt.float32, device=dev)
        A.offsets, A.columns, A.values, A.nnz = off, col, val, col.numel()
        self.A = A
        self.b = b.to(torch.float64)
        lk = torch.log(k).reshape(-1)
        self.logk_std = ((lk - lk.mean()) / lk.std().clamp_min(1e-12)).to(torch.float32).contiguous()
        self.device = dev
        from . import linear

        # FP64 inverse diagonal (Jacobi) and the mean diagonal, which sets the
        # scale of the network's correction
        self.dinv = linear.preconditioner(A, "diag")._prec_data.to(torch.float64).reshape(-1).contiguous()
        self.dmean = float(torch.mean(1.0 / self.dinv).item())

    def residual(self, x: torch.Tensor, r: torch.Tensor, rr: torch.Tensor, ws: torch.Tensor) -> None:
        A = self.A
        _native.call(
            "warp_amd_bsr_residual_f32_f64", A.nrow, 1, _native.ptr(A.offsets), _native.ptr(A.columns),
            _native.ptr(A.values), int(A.nnz), _native.ptr(x), _native.ptr(self.b), _native.ptr(r), None,
            _native.ptr(rr), _native.ptr(ws), ws.numel(), _native.stream_of(self.device),
        )


def generate_energy_pairs(system: FV7System, count: int = 10, tol: float = 1e-10, maxiter: int = 20000,
                          return_solution: bool = False):
    """(r, e) pairs with A e = r from one FP64 Jacobi-CG solve of ``system``.

    Snapshots x_k at k = 0 and at ``count - 1`` iterations spread
    geometrically over the solve; e_k = x* - x_k (x* the converged solution),
    r_k = b - A x_k, both divided by rms(r_k) (the energy loss is invariant
    to that common scale).  Returns a list of (r, e) FP32 n^3 tensors."""
    from .. import sparse
    from .. import types as wt
    from . import linear

    A32 = system.A
    dev = system.device
    N = A32.nrow
    A64 = sparse.bsr_zeros(N, N, wt.float64, device=dev)
    A64.offsets, A64.columns, A64.values, A64.nnz = A32.offsets, A32.columns, A32.values.to(torch.float64), A32.nnz
    M = linear.preconditioner(A64, "diag")
    x_star = torch.zeros(N, dtype=torch.float64, device=dev)
    it_total, _, _ = linear.cg(A64, system.b, x_star, tol=tol, atol=0.0, maxiter=maxiter, M=M, check_every=50)
    # count - 1 distinct iterations in [1, it_total), geometrically spaced (denser
    # where the geometric points collide at small k)
    want = sorted({max(1, int(round(it_total ** (j / (count - 1))))) for j in range(1, count)})
    want = [t for t in want if t < it_total]
    extra = 1
    while len(want) < min(count - 1, it_total - 1) and extra < it_total:
        if extra not in want:
            want.append(extra)
        extra += 1
    want = sorted(want)
    snaps = {0: torch.zeros(N, dtype=torch.float64, device=dev)}
    xw = torch.zeros(N, dtype=torch.float64, device=dev)

    def cb(it, err, atol):
        if it in want and it not in snaps:
            snaps[it] = xw.clone()

    if want:
        linear.cg(A64, system.b, xw, tol=tol, atol=0.0, maxiter=max(want) + 1, M=M, check_every=1, callback=cb)
    r = torch.empty(N, dtype=torch.float64, device=dev)
    rr = torch.zeros(1, dtype=torch.float64, device=dev)
    ws = torch.empty(_native.lib().warp_amd_bsr_residual_workspace_size(N), dtype=torch.uint8, device=dev)
    pairs = []
    for it in sorted(snaps):
        system.residual(snaps[it], r, rr, ws)
        rs = math.sqrt(float(rr.item()) / N)
        if rs <= 0.0:
            continue
        pairs.append(((r / rs).to(torch.float32), ((x_star - snaps[it]) / rs).to(torch.float32)))
    return (pairs, x_star) if return_solution else pairs


class EnergyTrainer:
    """Trains the U-Net inside the smoothing composite used by :class:`LearnedFCG`.

    Loss: the one-step energy reduction 1 - (r.z)^2 / ((r.e)(z.Az)) of the
    preconditioned direction z = M(r) -- the fraction of the A-norm error
    that an exact line search along z leaves.  Every step runs on library
    kernels: the composite forward (Jacobi sweeps, FP32 network), the loss
    gradient, its adjoint back through the post-smoothing sweeps
    (g <- g - omega A D^-1 g), the network backward and Adam.
    """

    def __init__(self, net: UNet, smooth_steps: int = 1, omega: float = 2.0 / 3.0, precision: str = "fp32"):
        if precision not in ("fp32", "bf16"):
            raise ValueError("precision must be 'fp32' or 'bf16'")
        self.net = net
        self.nu = int(smooth_steps)
        self.omega = float(omega)
        # "bf16": the network runs forward and backward on MFMA (BF16 operands,
        # FP32 accumulation, FP32 master weights and Adam); "fp32": FP32 kernels
        self.precision = precision
        dev = net.dev
        N = net.n**3
        self.N = N
        f64 = dict(dtype=torch.float64, device=dev)
        self.r = torch.zeros(N, **f64)
        self.e = torch.zeros(N, **f64)
        self.r1 = torch.zeros(N, **f64)
        self.z = torch.zeros(N, **f64)
        self.zt = torch.zeros(N, **f64)
        self.y = torch.zeros(N, **f64)
        self.g = torch.zeros(N, **f64)
        self.gt = torch.zeros(N, **f64)
        self.h = torch.zeros(N, **f64)
        self.sc = torch.zeros(16, **f64)  # 1 scale, 2 inv, 7 r1.r1, 10 r.z, 11 z.Az, 12 r.e, 13 loss, 14 g.g
        self.inp = torch.empty(N * 2, dtype=torch.int16 if precision == "bf16" else torch.float32, device=dev)
        self.dout = torch.empty(N, dtype=torch.float32, device=dev)
        lib = _native.lib()
        self.ws_mv = torch.empty(lib.warp_amd_bsr_residual_workspace_size(N), dtype=torch.uint8, device=dev)
        self.dot_bytes = lib.warp_amd_dot_workspace_size(N)
        self.ws_dot = torch.zeros(self.dot_bytes, dtype=torch.uint8, device=dev)

    def _sp(self, i):
        return self.sc.data_ptr() + 8 * i

    def _s(self):
        return _native.stream_of(self.net.dev)

    def _residual(self, system, b, x, out, slot):
        A = system.A
        _native.call("warp_amd_bsr_residual_f32_f64", A.nrow, 1, _native.ptr(A.offsets), _native.ptr(A.columns),
                     _native.ptr(A.values), int(A.nnz), _native.ptr(x), _native.ptr(b), _native.ptr(out), None,
                     self._sp(slot), _native.ptr(self.ws_mv), self.ws_mv.numel(), self._s())

    def _dot(self, a, b, slot):
        _native.call("warp_amd_dot_f64", self.N, _native.ptr(a), _native.ptr(b), self._sp(slot),
                     _native.ptr(self.ws_dot), self.dot_bytes, self._s())

    def forward(self, system: FV7System, r: torch.Tensor) -> torch.Tensor:
        """z = M(r) with the FP32 network (activations kept for backward); returns z."""
        s = self._s()
        A = system.A
        gain = 1.0 / system.dmean
        if self.nu == 0:
            self.z.zero_()
            self.r1.copy_(r)
            self._dot(r, r, 7)
        else:
            _native.call("warp_amd_diag_mv_f64", self.N, _native.ptr(system.dinv), _native.ptr(r), _native.ptr(self.z),
                         _native.ptr(self.z), self.omega, 0.0, s)
            for _ in range(self.nu - 1):
                self._sweep(system, r, self.z, self.zt)
                self.z, self.zt = self.zt, self.z
            self._residual(system, r, self.z, self.r1, 7)
        _native.call("warp_amd_rms_scales", self._sp(7), self.N, self._sp(1), self._sp(2), s)
        if self.precision == "bf16":
            _native.call("warp_amd_unet_pack_input_bf16", self.N, _native.ptr(self.r1), _native.ptr(system.logk_std),
                         self._sp(2), _native.ptr(self.inp), s)
            out = self.net.forward_bf16(self.inp)
        else:
            _native.call("warp_amd_unet_pack_input", self.N, _native.ptr(self.r1), _native.ptr(system.logk_std),
                         self._sp(2), _native.ptr(self.inp), s)
            out = self.net.forward(self.inp)
        _native.call("warp_amd_unet_unpack_add", self.N, _native.ptr(out), self._sp(1), gain, _native.ptr(self.z), s)
        for _ in range(self.nu):
            self._sweep(system, r, self.z, self.zt)
            self.z, self.zt = self.zt, self.z
        return self.z

    def _sweep(self, system, r, z_in, z_out):
        A = system.A
        _native.call("warp_amd_jacobi_sweep_f32_f64", A.nrow, _native.ptr(A.offsets), _native.ptr(A.columns),
                     _native.ptr(A.values), int(A.nnz), _native.ptr(z_in), _native.ptr(r), _native.ptr(system.dinv),
                     self.omega, _native.ptr(z_out), self._s())

    def step(self, system: FV7System, r32: torch.Tensor, e32: torch.Tensor, lr: float) -> torch.Tensor:
        """One Adam step on the pair (r, e) of ``system``; returns the loss (device FP64 view)."""
        loss = self.gradient(system, r32, e32)
        self.net.adam_step(lr)
        return loss

    def batch_step(self, batch: list, lr: float) -> torch.Tensor:
        """One Adam step on the mean gradient of [(system, r, e), ...]; returns the mean loss (device)."""
        if len(batch) == 1:
            return self.step(*batch[0], lr)
        s = self._s()
        if getattr(self, "_gacc", None) is None:
            self._gacc = {k: torch.zeros_like(v) for k, v in self.net.grads.items()}
        losses = []
        w = 1.0 / len(batch)
        for j, (system, r32, e32) in enumerate(batch):
            losses.append(self.gradient(system, r32, e32).clone())
            for k, g in self.net.grads.items():
                _native.call("warp_amd_axpby_f32", g.numel(), w, _native.ptr(g), 0.0 if j == 0 else 1.0,
                             _native.ptr(self._gacc[k]), s)
        for k, g in self.net.grads.items():
            g.copy_(self._gacc[k])
        self.net.adam_step(lr)
        return torch.cat(losses).mean().reshape(1)

    def gradient(self, system: FV7System, r32: torch.Tensor, e32: torch.Tensor) -> torch.Tensor:
        """Loss of the pair (device FP64 view) and its parameter gradients in net.grads."""
        s = self._s()
        A = system.A
        self.r.copy_(r32.reshape(-1))
        self.e.copy_(e32.reshape(-1))
        z = self.forward(system, self.r)
        _native.call("warp_amd_bsr_mv_f32_f64", A.nrow, 1, _native.ptr(A.offsets), _native.ptr(A.columns),
                     _native.ptr(A.values), int(A.nnz), _native.ptr(z), _native.ptr(self.y), self._sp(11),
                     _native.ptr(self.ws_mv), self.ws_mv.numel(), s)
        self._dot(self.r, z, 10)
        self._dot(self.r, self.e, 12)
        _native.call("warp_amd_energy_loss_grad", self.N, _native.ptr(self.r), _native.ptr(self.y), self._sp(10),
                     _native.ptr(self.g), self._sp(13), s)
        # adjoint of each post-smoothing sweep z <- z + omega D^-1 (r - A z): g <- g - A (omega D^-1 g)
        for _ in range(self.nu):
            _native.call("warp_amd_diag_mv_f64", self.N, _native.ptr(system.dinv), _native.ptr(self.g),
                         _native.ptr(self.h), _native.ptr(self.h), self.omega, 0.0, s)
            self._residual(system, self.g, self.h, self.gt, 14)
            self.g, self.gt = self.gt, self.g
        _native.call("warp_amd_scale_to_f32", self.N, _native.ptr(self.g), self._sp(1), 1.0 / system.dmean,
                     _native.ptr(self.dout), s)
        if self.precision == "bf16":
            self.net.backward_bf16(self.dout)
        else:
            self.net.backward(self.dout)
        return self.sc[13:14]

    def loss(self, system: FV7System, r32: torch.Tensor, e32: torch.Tensor) -> float:
        """Loss of one pair without a training step."""
        A = system.A
        self.r.copy_(r32.reshape(-1))
        self.e.copy_(e32.reshape(-1))
        z = self.forward(system, self.r)
        _native.call("warp_amd_bsr_mv_f32_f64", A.nrow, 1, _native.ptr(A.offsets), _native.ptr(A.columns),
                     _native.ptr(A.values), int(A.nnz), _native.ptr(z), _native.ptr(self.y), self._sp(11),
                     _native.ptr(self.ws_mv), self.ws_mv.numel(), self._s())
        self._dot(self.r, z, 10)
        self._dot(self.r, self.e, 12)
        _native.call("warp_amd_energy_loss_grad", self.N, _native.ptr(self.r), _native.ptr(self.y), self._sp(10),
                     _native.ptr(self.g), self._sp(13), self._s())
        return float(self.sc[13].item())

    def train(self, systems: list, pairs: list, steps: int, lr: float = 1e-3, lr_final: float = 1e-5, seed: int = 0,
              log=None, log_every: int = 250, batch: int = 1) -> list[float]:
        """Adam with cosine decay over ``steps`` steps, each on ``batch`` random pairs.

        pairs: list of (system index, r32, e32).  Returns the mean loss at
        each log point."""
        rng = np.random.default_rng(seed)
        hist = []
        acc = []
        for step in range(steps):
            cur = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * step / max(steps - 1, 1)))
            chosen = [pairs[int(i)] for i in rng.integers(len(pairs), size=batch)]
            acc.append(self.batch_step([(systems[si], r32, e32) for si, r32, e32 in chosen], cur))
            if (step + 1) % log_every == 0 or step == steps - 1:
                mean = float(torch.cat(acc).mean().item())
                acc = []
                hist.append(mean)
                if log:
                    log(f"[train] step {step + 1}/{steps}: mean loss {mean:.4f}")
        return hist


def train_learned_preconditioner(net: UNet, systems: list, smooth_steps: int = 1, omega: float = 2.0 / 3.0,
                                 pairs_per_solve: int = 10, steps: int = 2000, lr: float = 1e-3,
                                 dagger_rounds: int = 2, dagger_steps: int = 1000, dagger_lr: float = 5e-4,
                                 dagger_iters: int = 40, dagger_every: int = 8, seed: int = 0, log=None,
                                 batch: int = 1, precision: str = "fp32") -> dict:
    """Train ``net`` as the preconditioner of :class:`LearnedFCG` on ``systems``.

    1. (r, e) pairs from one FP64 Jacobi-CG solve per system
       (:func:`generate_energy_pairs`), ``steps`` Adam steps on the energy loss.
    2. ``dagger_rounds`` times: run the learned FCG on every training system,
       add (r_k, x* - x_k) of its own iterates (every ``dagger_every``-th of
       the first ``dagger_iters``) to the pairs, train ``dagger_steps`` more.
       Residuals met inside FCG differ from plain-CG residuals; without this
       step the trained network stalls FCG after a few iterations.
    Returns dict(pcg_pairs, fcg_pairs, loss_history, seconds)."""
    t0 = time.time()
    trainer = EnergyTrainer(net, smooth_steps, omega, precision=precision)
    pairs = []
    x_stars = []
    for si, sysm in enumerate(systems):
        ps, x_star = generate_energy_pairs(sysm, count=pairs_per_solve, return_solution=True)
        pairs += [(si, r, e) for r, e in ps]
        x_stars.append(x_star)
    n_pcg = len(pairs)
    if log:
        log(f"[train] {n_pcg} plain-CG pairs from {len(systems)} systems in {time.time() - t0:.1f}s")
    hist = trainer.train(systems, pairs, steps, lr=lr, seed=seed, log=log, batch=batch)
    n_fcg = 0
    for rd in range(dagger_rounds):
        new = []
        for si, sysm in enumerate(systems):
            fcg = LearnedFCG(sysm, net, smooth_steps=smooth_steps, omega=omega)
            new += [(si, r, e) for r, e in fcg.collect_pairs(x_stars[si], dagger_iters, dagger_every)]
        pairs += new
        n_fcg += len(new)
        if log:
            log(f"[train] round {rd}: +{len(new)} FCG pairs ({len(pairs)} in total)")
        hist += trainer.train(systems, pairs, dagger_steps, lr=dagger_lr, seed=seed + 1 + rd, log=log, batch=batch)
    return dict(pcg_pairs=n_pcg, fcg_pairs=n_fcg, loss_history=hist, seconds=time.time() - t0)


# ---------------------------------------------------------------------------
# Flexible CG with the learned preconditioner
# ---------------------------------------------------------------------------


class LearnedFCG:
    """Flexible CG (Notay 2000, truncation 1) in FP64 with a learned preconditioner.

    Preconditioner (all vectors FP64, matrix FP32):
      z  = nu damped Jacobi sweeps on A z = r from z = 0
      r1 = r - A z,  s = rms(r1)
      z += gain * s * U(r1 / s, log k)        (U: the BF16 MFMA U-Net; gain = 1 / mean diag)
      z  = nu more sweeps from that z
    With nu = 0 it is z = gain * s * U(r / s).  Every scalar stays on the
    device; two iterations (ping-pong buffers) are captured in one HIP graph
    and replayed, with a host check of ||r|| after each replay.  On
    stagnation the solve falls back to block-Jacobi CG with FP64 refinement
    from the current iterate.
    """

    def __init__(self, system: FV7System, net: UNet, smooth_steps: int = 1, omega: float = 2.0 / 3.0,
                 gain: float | None = None, precision: str = "bf16", fused: bool | None = None):
        if precision not in ("bf16", "fp32"):
            raise ValueError("precision must be 'bf16' or 'fp32'")
        self.sys = system
        self.net = net
        self.nu = int(smooth_steps)
        self.omega = float(omega)
        self.gain = 1.0 / system.dmean if gain is None else float(gain)
        self.precision = precision
        dev = system.device
        N = system.A.nrow
        self.N = N
        f64 = dict(dtype=torch.float64, device=dev)
        self.r = torch.zeros(N, **f64)
        self.r1 = torch.zeros(N, **f64)
        self.z = torch.zeros(N, **f64)
        self.zt = torch.zeros(N, **f64)
        self.p = [torch.zeros(N, **f64), torch.zeros(N, **f64)]
        self.q = [torch.zeros(N, **f64), torch.zeros(N, **f64)]
        # 0 rr, 1 scale, 2 inv, 3 zq, 4/5 pq[0/1], 6 pr, 7 r1.r1
        self.sc = torch.zeros(16, **f64)
        if precision == "bf16":
            self.inp = torch.empty(N * 2, dtype=torch.int16, device=dev)
        else:
            self.inp = torch.empty(N * 2, dtype=torch.float32, device=dev)
        lib = _native.lib()
        self.ws_mv = torch.empty(lib.warp_amd_bsr_residual_workspace_size(N), dtype=torch.uint8, device=dev)
        self.ws_up = torch.empty(lib.warp_amd_fcg_update_workspace_size(N), dtype=torch.uint8, device=dev)
        self.dot_bytes = lib.warp_amd_dot_workspace_size(N)
        self.ws_dot = torch.zeros(self.dot_bytes, dtype=torch.uint8, device=dev)
        self.dinv = system.dinv
        # one sweep each side, BF16 network: the fused kernels (16 launches per iteration)
        self.fused = (self.nu == 1 and precision == "bf16") if fused is None else bool(fused)
        if self.fused and (self.nu != 1 or precision != "bf16"):
            raise ValueError("the fused iteration needs smooth_steps=1 and precision='bf16'")
        self.ws_fused = torch.empty(lib.warp_amd_fcg_fused_workspace_size(N), dtype=torch.uint8, device=dev)
        self.w32 = torch.zeros(N, dtype=torch.float32, device=dev)  # omega D^-1 r (fused path)
        self.z32 = torch.zeros(N, dtype=torch.float32, device=dev)  # preconditioned residual (fused path)
        self._graph = None

    def _s(self):
        return _native.stream_of(self.sys.device)

    def _sp(self, i):
        return self.sc.data_ptr() + 8 * i

    def _dot(self, a, b, slot):
        _native.call("warp_amd_dot_f64", self.N, _native.ptr(a), _native.ptr(b), self._sp(slot), _native.ptr(self.ws_dot),
                     self.dot_bytes, self._s())

    def _sweep(self, z_in, z_out):
        A = self.sys.A
        _native.call("warp_amd_jacobi_sweep_f32_f64", A.nrow, _native.ptr(A.offsets), _native.ptr(A.columns),
                     _native.ptr(A.values), int(A.nnz), _native.ptr(z_in), _native.ptr(self.r), _native.ptr(self.dinv),
                     self.omega, _native.ptr(z_out), self._s())

    def _net(self, src, ss_slot) -> torch.Tensor:
        """Network output for the residual `src` whose squared norm is in slot ss_slot."""
        s = self._s()
        _native.call("warp_amd_rms_scales", self._sp(ss_slot), self.N, self._sp(1), self._sp(2), s)
        if self.precision == "bf16":
            _native.call("warp_amd_unet_pack_input_bf16", self.N, _native.ptr(src), _native.ptr(self.sys.logk_std),
                         self._sp(2), _native.ptr(self.inp), s)
            return self.net.forward_bf16(self.inp)
        _native.call("warp_amd_unet_pack_input", self.N, _native.ptr(src), _native.ptr(self.sys.logk_std),
                     self._sp(2), _native.ptr(self.inp), s)
        return self.net.forward(self.inp)

    def precondition(self) -> None:
        """z = M(r) (see the class docstring); uses self.r and slot 0 (r.r)."""
        s = self._s()
        A = self.sys.A
        if self.nu == 0:
            self.z.zero_()
            out = self._net(self.r, 0)
        else:
            # first sweep from zero: z = omega D^-1 r
            _native.call("warp_amd_diag_mv_f64", self.N, _native.ptr(self.dinv), _native.ptr(self.r),
                         _native.ptr(self.z), _native.ptr(self.z), self.omega, 0.0, s)
            for _ in range(self.nu - 1):
                self._sweep(self.z, self.zt)
                self.z, self.zt = self.zt, self.z
            _native.call("warp_amd_bsr_residual_f32_f64", A.nrow, 1, _native.ptr(A.offsets), _native.ptr(A.columns),
                         _native.ptr(A.values), int(A.nnz), _native.ptr(self.z), _native.ptr(self.r),
                         _native.ptr(self.r1), None, self._sp(7), _native.ptr(self.ws_mv), self.ws_mv.numel(), s)
            out = self._net(self.r1, 7)
        _native.call("warp_amd_unet_unpack_add", self.N, _native.ptr(out), self._sp(1), self.gain, _native.ptr(self.z), s)
        for _ in range(self.nu):
            self._sweep(self.z, self.zt)
            self.z, self.zt = self.zt, self.z

    def _iteration_fused(self, cur: int) -> None:
        prev = 1 - cur
        A = self.sys.A
        s = self._s()
        P = _native.ptr
        mat = (A.nrow, P(A.offsets), P(A.columns), P(A.values), int(A.nnz))
        ws = (P(self.ws_fused), self.ws_fused.numel(), s)
        _native.call("warp_amd_fcg_presmooth", *mat, P(self.r), P(self.w32), P(self.r1), self._sp(7), self._sp(1),
                     self._sp(2), *ws)
        _native.call("warp_amd_unet_pack_input_bf16", self.N, P(self.r1), P(self.sys.logk_std), self._sp(2),
                     P(self.inp), s)
        out = self.net.forward_bf16(self.inp)
        _native.call("warp_amd_fcg_postsmooth", *mat, P(self.r), P(self.dinv), self.omega, P(self.w32), P(out),
                     self._sp(1), self.gain, P(self.q[prev]), P(self.z32), self._sp(3), *ws)
        _native.call("warp_amd_fcg_direction_spmv", *mat, P(self.z32), P(self.p[prev]), self._sp(3),
                     self._sp(4 + prev), P(self.r), P(self.p[cur]), P(self.q[cur]), self._sp(4 + cur), self._sp(6), *ws)
        _native.call("warp_amd_fcg_update", self.N, self._sp(6), self._sp(4 + cur), P(self.p[cur]), P(self.q[cur]),
                     P(self.x), P(self.r), self._sp(0), P(self.dinv), self.omega, P(self.w32), P(self.ws_up),
                     self.ws_up.numel(), s)

    def iteration(self, cur: int) -> None:
        if self.fused:
            self._iteration_fused(cur)
            return
        prev = 1 - cur
        A = self.sys.A
        s = self._s()
        self.precondition()
        self._dot(self.z, self.q[prev], 3)
        _native.call("warp_amd_fcg_direction", self.N, _native.ptr(self.z), _native.ptr(self.p[prev]), self._sp(3),
                     self._sp(4 + prev), _native.ptr(self.p[cur]), s)
        _native.call("warp_amd_bsr_mv_f32_f64", A.nrow, 1, _native.ptr(A.offsets), _native.ptr(A.columns),
                     _native.ptr(A.values), int(A.nnz), _native.ptr(self.p[cur]), _native.ptr(self.q[cur]),
                     self._sp(4 + cur), _native.ptr(self.ws_mv), self.ws_mv.numel(), s)
        self._dot(self.p[cur], self.r, 6)
        _native.call("warp_amd_fcg_update", self.N, self._sp(6), self._sp(4 + cur), _native.ptr(self.p[cur]),
                     _native.ptr(self.q[cur]), _native.ptr(self.x), _native.ptr(self.r), self._sp(0), None, 0.0, None,
                     _native.ptr(self.ws_up), self.ws_up.numel(), s)

    def _start(self, x0: torch.Tensor | None = None) -> float:
        """Initial state (x = x0 or 0, r = b - A x, fresh directions); returns ||b||."""
        if getattr(self, "x", None) is None:
            self.x = torch.empty_like(self.r)  # a captured graph keeps this address
        if x0 is None:
            self.x.zero_()
        else:
            self.x.copy_(x0)
        self.sys.residual(self.x, self.r, self.sc[0:1], self.ws_mv)
        self._weight()
        for t in self.p + self.q:
            t.zero_()
        self.sc[4] = 1.0
        self.sc[5] = 1.0
        return float(torch.linalg.vector_norm(self.sys.b).item())

    def _weight(self) -> None:
        """w = omega D^-1 r for the fused path (kept current by its update kernel)."""
        if self.fused:
            _native.call("warp_amd_fcg_weight", self.N, _native.ptr(self.dinv), _native.ptr(self.r), self.omega,
                         _native.ptr(self.w32), self._s())

    def collect_pairs(self, x_star: torch.Tensor, iterations: int = 40, every: int = 8) -> list:
        """(r_k, x* - x_k) of this solver's own iterates from x = 0, every
        ``every``-th of the first ``iterations``, both divided by rms(r_k);
        FP32 n^3 tensors (eager, no graph)."""
        self._start()
        out = []
        N = self.N
        for k in range(1, iterations + 1):
            self.iteration((k - 1) % 2)
            if k % every == 0:
                rs = math.sqrt(max(float(self.sc[0].item()), 0.0) / N)
                if rs <= 0.0:
                    break
                out.append(((self.r / rs).to(torch.float32), ((x_star - self.x) / rs).to(torch.float32)))
        return out

    def solve(self, x: torch.Tensor, tol: float = 1e-8, maxiter: int = 500, use_graph: bool = True,
              stagnation_window: int = 20, stagnation_factor: float = 0.5, fallback: bool = True):
        """Solve A x = b from the initial guess x (FP64, updated in place).

        Returns dict(iterations, relative_residual (recursive), history,
        fallback_iterations, true_relative_residual (FP64 b - A x))."""
        bnorm = self._start(x)
        hist = [math.sqrt(float(self.sc[0].item())) / bnorm]
        it = 0
        if use_graph and self._graph is None:
            # warm up both halves eagerly (allocates every buffer), then capture
            saved = (self.x.clone(), self.r.clone(), self.sc.clone())
            self.iteration(0)
            self.iteration(1)
            torch.cuda.synchronize()
            self.x.copy_(saved[0]); self.r.copy_(saved[1]); self.sc.copy_(saved[2])
            for t in self.p + self.q:
                t.zero_()
            g = torch.cuda.CUDAGraph()
            cs = torch.cuda.Stream(self.sys.device)
            cs.wait_stream(torch.cuda.current_stream(self.sys.device))
            with torch.cuda.stream(cs):
                with torch.cuda.graph(g, stream=cs):
                    self.iteration(0)
                    self.iteration(1)
            torch.cuda.current_stream(self.sys.device).wait_stream(cs)
            self._graph = g
        stalled = False
        for _ in range(4):  # FCG, then residual replacement while the true residual is above tol
            while hist[-1] > tol and it < maxiter:
                if use_graph:
                    self._graph.replay()
                else:
                    self.iteration(0)
                    self.iteration(1)
                it += 2
                hist.append(math.sqrt(max(float(self.sc[0].item()), 0.0)) / bnorm)
                w = stagnation_window // 2
                if len(hist) > w and hist[-1] > stagnation_factor * hist[-1 - w]:
                    stalled = True
                    break
            if stalled or it >= maxiter:
                break
            # converged on the recursive residual: replace it by the FP64 true residual
            self.sys.residual(self.x, self.r, self.sc[0:1], self.ws_mv)
            self._weight()
            true_rel = math.sqrt(max(float(self.sc[0].item()), 0.0)) / bnorm
            if true_rel <= tol:
                break
            hist.append(true_rel)
            for t in self.p + self.q:
                t.zero_()
            self.sc[4] = 1.0
            self.sc[5] = 1.0
        x.copy_(self.x)
        fb_iters = 0
        if hist[-1] > tol and fallback:
            from . import refine

            res = refine.cg_refined(self.sys.A, self.sys.b, x, tol=tol, M="diag")
            fb_iters = res.inner_iterations
            hist.append(res.relative_residual)
        self.sys.residual(x, self.r1, self.sc[8:9], self.ws_mv)
        true_rel = math.sqrt(max(float(self.sc[8].item()), 0.0)) / bnorm
        return dict(iterations=it, relative_residual=hist[-1], history=hist, fallback_iterations=fb_iters,
                    true_relative_residual=true_rel)


# ---------------------------------------------------------------------------
# Checkpoints and command line
# ---------------------------------------------------------------------------


def save_checkpoint(path: str, net: UNet, meta: dict) -> None:
    """Weights, Adam state (m, v, t) and the training recipe in one .npz."""
    d = net.state_dict()
    for k, v in meta.items():
        d["meta." + k] = np.asarray(v)
    np.savez(path, **d)


def load_checkpoint(path: str, n: int, device) -> tuple[UNet, dict]:
    net = UNet.load(path, n, device)
    with np.load(path, allow_pickle=False) as f:
        meta = {k[5:]: f[k].item() if f[k].ndim == 0 else f[k] for k in f.files if k.startswith("meta.")}
    return net, meta


def evaluate(net: UNet, meta: dict, seeds, device, tol: float = 1e-8, log=None) -> list[dict]:
    """Learned FCG vs FP64 Jacobi-CG and FP32 Jacobi-CG + FP64 refinement on
    held-out systems; every solution checked by the FP64 true residual."""
    from .. import sparse
    from .. import types as wt
    from . import linear, refine

    n = net.n
    res = []
    for seed in seeds:
        sysm = FV7System(n, float(meta["sigma"]), float(meta["corr"]), int(seed), device)
        N = n**3
        A64 = sparse.bsr_zeros(N, N, wt.float64, device=device)
        A64.offsets, A64.columns, A64.values, A64.nnz = sysm.A.offsets, sysm.A.columns, sysm.A.values.double(), sysm.A.nnz
        M64 = linear.preconditioner(A64, "diag")
        x = torch.zeros(N, dtype=torch.float64, device=device)
        linear.cg(A64, sysm.b, x, tol=tol, atol=0.0, maxiter=10, M=M64)  # kernels warm
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it64, _, _ = linear.cg(A64, sysm.b, x, tol=tol, atol=0.0, maxiter=100000, M=M64)
        torch.cuda.synchronize()
        t64 = time.perf_counter() - t0
        r = torch.empty(N, dtype=torch.float64, device=device)
        rr = torch.zeros(1, dtype=torch.float64, device=device)
        ws = torch.empty(_native.lib().warp_amd_bsr_residual_workspace_size(N), dtype=torch.uint8, device=device)
        bn = float(torch.linalg.vector_norm(sysm.b).item())
        sysm.residual(x, r, rr, ws)
        true64 = math.sqrt(float(rr.item())) / bn
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        # inner 1e-4: two refinements reach 1e-8 (3e-5, the default tuned for 1e-10,
        # over-solves here: 69-87 ms against 66-80)
        ir = refine.cg_refined(sysm.A, sysm.b, x, tol=tol, inner_tol=1e-4, M="diag")
        torch.cuda.synchronize()
        tir = time.perf_counter() - t0
        sysm.residual(x, r, rr, ws)
        true_ir = math.sqrt(float(rr.item())) / bn
        fcg = LearnedFCG(sysm, net, smooth_steps=int(meta["smooth_steps"]), omega=float(meta["omega"]))
        # two iterations untimed (kernels warm), as the Jacobi-CG runs above get
        # theirs; the timed solve then holds everything a solve of this system
        # costs: the initial residual, every iteration and the FP64 check.  It
        # launches directly: replaying a captured graph measured no faster (16.7
        # vs 16.9 ms) and capturing costs 4-7 ms per system.
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fcg.solve(x, tol=tol, maxiter=2, use_graph=False, fallback=False)
        torch.cuda.synchronize()
        t_warm = time.perf_counter() - t0
        x.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fcg.solve(x, tol=tol, maxiter=2000, use_graph=False)
        torch.cuda.synchronize()
        tf = time.perf_counter() - t0
        rec = dict(seed=int(seed), jacobi64_it=it64, jacobi64_ms=t64 * 1e3, jacobi64_true=true64,
                   ir_inner_it=ir.inner_iterations, ir_ms=tir * 1e3, ir_true=true_ir,
                   fcg_it=out["iterations"], fcg_fallback_it=out["fallback_iterations"], fcg_ms=tf * 1e3,
                   fcg_warmup_ms=t_warm * 1e3, fcg_true=out["true_relative_residual"])
        res.append(rec)
        if log:
            log(f"[eval] seed {seed}: Jacobi-CG fp64 {it64} it {t64 * 1e3:.1f} ms ({true64:.1e}) | IR {ir.inner_iterations} it "
                f"{tir * 1e3:.1f} ms ({true_ir:.1e}) | FCG {out['iterations']}+{out['fallback_iterations']} it "
                f"{tf * 1e3:.1f} ms ({out['true_relative_residual']:.1e}) | speedup {min(t64, tir) / tf:.2f}x")
    return res


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(prog="python -m warp_amd.optim.learned")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="train the U-Net preconditioner on the GPU and write a checkpoint")
    t.add_argument("--n", type=int, default=128)
    t.add_argument("--sigma", type=float, default=1.5)
    t.add_argument("--corr", type=float, default=8.0)
    t.add_argument("--train-seeds", type=int, default=200)
    t.add_argument("--seed0", type=int, default=1000)
    t.add_argument("--pairs-per-solve", type=int, default=10)
    t.add_argument("--steps", type=int, default=4000)
    t.add_argument("--dagger-rounds", type=int, default=2)
    t.add_argument("--dagger-steps", type=int, default=2000)
    t.add_argument("--smooth-steps", type=int, default=1)
    t.add_argument("--init-seed", type=int, default=0)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--dagger-lr", type=float, default=5e-4)
    t.add_argument("--batch", type=int, default=1)
    t.add_argument("--dagger-iters", type=int, default=40)
    t.add_argument("--dagger-every", type=int, default=8)
    t.add_argument("--eval-seeds", type=int, default=5)
    t.add_argument("--eval-seed0", type=int, default=4000, help="validation seeds (held out from training and test)")
    t.add_argument("--precision", choices=("fp32", "bf16"), default="fp32",
                   help="network arithmetic in training: FP32 kernels or BF16 MFMA (FP32 accumulation)")
    t.add_argument("--out", required=True)
    e = sub.add_parser("evaluate", help="learned FCG vs Jacobi-CG on held-out seeds")
    e.add_argument("--ckpt", required=True)
    e.add_argument("--n", type=int, default=128)
    e.add_argument("--seeds", type=int, default=50)
    e.add_argument("--seed0", type=int, default=5000)
    args = ap.parse_args(argv)
    dev = torch.device("cuda", 0)

    def log(msg):
        print(msg, flush=True)

    if args.cmd == "train":
        t0 = time.time()
        systems = []
        for i in range(args.train_seeds):
            systems.append(FV7System(args.n, args.sigma, args.corr, args.seed0 + i, dev))
            if (i + 1) % 20 == 0:
                log(f"[train] {i + 1}/{args.train_seeds} systems ({time.time() - t0:.0f}s)")
        net = UNet(args.n, dev, seed=args.init_seed)
        info = train_learned_preconditioner(net, systems, smooth_steps=args.smooth_steps,
                                            pairs_per_solve=args.pairs_per_solve, steps=args.steps,
                                            dagger_rounds=args.dagger_rounds, dagger_steps=args.dagger_steps,
                                            seed=args.init_seed, log=log, lr=args.lr, dagger_lr=args.dagger_lr,
                                            batch=args.batch, dagger_iters=args.dagger_iters,
                                            dagger_every=args.dagger_every, precision=args.precision)
        meta = dict(n=args.n, sigma=args.sigma, corr=args.corr, train_seed0=args.seed0, train_seeds=args.train_seeds,
                    smooth_steps=args.smooth_steps, omega=2.0 / 3.0, steps=args.steps, dagger_rounds=args.dagger_rounds,
                    dagger_steps=args.dagger_steps, init_seed=args.init_seed, lr=args.lr, dagger_lr=args.dagger_lr,
                    batch=args.batch, dagger_iters=args.dagger_iters, dagger_every=args.dagger_every,
                    precision=args.precision, pcg_pairs=info["pcg_pairs"],
                    fcg_pairs=info["fcg_pairs"], train_seconds=info["seconds"])
        save_checkpoint(args.out, net, meta)
        log(f"[train] wrote {args.out} after {time.time() - t0:.0f}s")
        del systems
        if args.eval_seeds:
            evaluate(net, meta, range(args.eval_seed0, args.eval_seed0 + args.eval_seeds), dev, log=log)
        return 0
    net, meta = load_checkpoint(args.ckpt, args.n, dev)
    res = evaluate(net, meta, range(args.seed0, args.seed0 + args.seeds), dev, log=log)
    sp = sorted(min(r["jacobi64_ms"], r["ir_ms"]) / r["fcg_ms"] for r in res)
    log(f"[eval] median speedup {sp[len(sp) // 2]:.2f}x over {len(res)} held-out systems")
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(main())
