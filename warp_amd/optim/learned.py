"""Learned preconditioner: a 3D U-Net trained on the GPU, used inside flexible CG.

No reference counterpart (the reference offers Jacobi / identity
preconditioners only, warp/_src/optim/linear.py:236-280).  Components:

* :class:`UNet` -- 4-level 3D U-Net on an n^3 cell grid (channels 16 / 32 /
  64 / 64, 3x3x3 convolutions, ReLU, 2x average pooling, nearest upsampling
  with additive skips, linear 1-channel output).  Forward, backward and Adam
  run in the library's hand-written kernels; weights are FP32 masters.
* :func:`generate_pairs` -- training pairs (normalised residual, error) taken
  from FP64 CG solves of the operator family.
* :func:`train` -- Adam on the mean-squared error of the predicted error.

The network input is (r / rms(r), standardised log k) and its output is
scaled back by rms(r), so the preconditioner is scale-equivariant.
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
            params = {k: f[k] for k in f.files if not k.startswith("adam.")}
            net = cls(n, device, params=params)
            for k in net.params:
                if "adam.m." + k in f.files:
                    net.m[k].copy_(torch.from_numpy(f["adam.m." + k]))
                    net.v[k].copy_(torch.from_numpy(f["adam.v." + k]))
            if "adam.t" in f.files:
                net.t = int(f["adam.t"][0])
        return net
