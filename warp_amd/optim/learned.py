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

    def residual(self, x: torch.Tensor, r: torch.Tensor, rr: torch.Tensor, ws: torch.Tensor) -> None:
        A = self.A
        _native.call(
            "warp_amd_bsr_residual_f32_f64", A.nrow, 1, _native.ptr(A.offsets), _native.ptr(A.columns),
            _native.ptr(A.values), int(A.nnz), _native.ptr(x), _native.ptr(self.b), _native.ptr(r), None,
            _native.ptr(rr), _native.ptr(ws), ws.numel(), _native.stream_of(self.device),
        )


def generate_pairs(system: FV7System, count: int = 10, tol: float = 1e-10, maxiter: int = 20000):
    """(input, target) pairs from one FP64 Jacobi-CG solve of ``system``.

    Snapshots x_k are taken at ``count`` iterations spread geometrically over
    the solve; with x* the converged solution, e_k = x* - x_k and
    r_k = b - A x_k (FP64).  Input = (r_k / rms(r_k), standardised log k),
    target = e_k / rms(e_k).  Returns a list of (inp [n^3*2] f32, target [n^3] f32).
    """
    from .. import sparse
    from .. import types as wt
    from . import linear

    A32 = system.A
    dev = system.device
    N = A32.nrow
    A64 = sparse.bsr_zeros(N, N, wt.float64, device=dev)
    A64.offsets, A64.columns, A64.values, A64.nnz = A32.offsets, A32.columns, A32.values.to(torch.float64), A32.nnz
    M = linear.preconditioner(A64, "diag")
    x = torch.zeros(N, dtype=torch.float64, device=dev)
    # first pass: iteration count of the full solve
    it_total, _, _ = linear.cg(A64, system.b, x, tol=tol, atol=0.0, maxiter=maxiter, M=M, check_every=50)
    targets_it = sorted({max(1, int(round(it_total ** (j / count)))) for j in range(1, count + 1)})
    targets_it = [t for t in targets_it if t < it_total][:count]
    x_star = x.clone()
    snaps = {}

    def cb(it, err, atol):
        if it in targets_it and it not in snaps:
            snaps[it] = xw.clone()

    xw = torch.zeros(N, dtype=torch.float64, device=dev)
    linear.cg(A64, system.b, xw, tol=tol, atol=0.0, maxiter=max(targets_it) + 1 if targets_it else 1, M=M,
              check_every=1, callback=cb)
    r = torch.empty(N, dtype=torch.float64, device=dev)
    rr = torch.zeros(1, dtype=torch.float64, device=dev)
    ws = torch.empty(_native.lib().warp_amd_bsr_residual_workspace_size(N), dtype=torch.uint8, device=dev)
    pairs = []
    for it in sorted(snaps):
        xk = snaps[it]
        system.residual(xk, r, rr, ws)
        e = x_star - xk
        rs = torch.sqrt(rr / N).item()
        es = torch.sqrt(torch.mean(e * e)).item()
        if rs <= 0 or es <= 0:
            continue
        inp = torch.stack([(r / rs).to(torch.float32), system.logk_std], dim=1).reshape(-1).contiguous()
        pairs.append((inp, (e / es).to(torch.float32).contiguous()))
    return pairs


def train(net: UNet, pairs: list, epochs: int = 3, lr: float = 1e-3, lr_final: float = 1e-4, seed: int = 0,
          log=None) -> list[float]:
    """Adam on the per-pair MSE of the predicted (normalised) error; returns mean loss per epoch."""
    rng = np.random.default_rng(seed)
    history = []
    steps = epochs * len(pairs)
    step = 0
    for ep in range(epochs):
        order = rng.permutation(len(pairs))
        losses = []
        for i in order:
            # cosine decay from lr to lr_final
            cur = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * step / max(steps - 1, 1)))
            inp, tgt = pairs[i]
            losses.append(net.train_step(inp, tgt, lr=cur))
            step += 1
        mean = float(torch.stack(losses).mean().item())
        history.append(mean)
        if log:
            log(f"[train] epoch {ep}: mean loss {mean:.5f}")
    return history


# ---------------------------------------------------------------------------
# Flexible CG with the learned preconditioner
# ---------------------------------------------------------------------------


class LearnedFCG:
    """Flexible CG (Notay 2000, truncation 1) in FP64 with z = s U(r / s) (+ w D^-1 r).

    Every scalar stays on the device; two iterations (ping-pong buffers) are
    captured in one HIP graph and replayed, with a host check of ||r|| after
    each replay.  On stagnation the solve falls back to Jacobi CG with FP64
    refinement from the current iterate.
    """

    def __init__(self, system: FV7System, net: UNet, jacobi_weight: float = 0.0):
        self.sys = system
        self.net = net
        self.w = float(jacobi_weight)
        dev = system.device
        N = system.A.nrow
        self.N = N
        f64 = dict(dtype=torch.float64, device=dev)
        self.r = torch.zeros(N, **f64)
        self.z = torch.zeros(N, **f64)
        self.p = [torch.zeros(N, **f64), torch.zeros(N, **f64)]
        self.q = [torch.zeros(N, **f64), torch.zeros(N, **f64)]
        self.sc = torch.zeros(16, **f64)  # 0 rr, 1 scale, 2 inv, 3 zq, 4/5 pq[0/1], 6 pr
        self.inp = torch.empty(N * 2, dtype=torch.float32, device=dev)
        lib = _native.lib()
        self.ws_mv = torch.empty(lib.warp_amd_bsr_residual_workspace_size(N), dtype=torch.uint8, device=dev)
        self.ws_up = torch.empty(lib.warp_amd_fcg_update_workspace_size(N), dtype=torch.uint8, device=dev)
        self.dot_bytes = lib.warp_amd_dot_workspace_size(N)
        self.ws_dot = torch.zeros(self.dot_bytes, dtype=torch.uint8, device=dev)
        if self.w != 0.0:
            from . import linear

            self.dinv = linear.preconditioner(system.A, "diag")._prec_data.to(torch.float64).reshape(-1).contiguous()
        self._graph = None

    def _s(self):
        return _native.stream_of(self.sys.device)

    def _sp(self, i):
        return self.sc.data_ptr() + 8 * i

    def _dot(self, a, b, slot):
        _native.call("warp_amd_dot_f64", self.N, _native.ptr(a), _native.ptr(b), self._sp(slot), _native.ptr(self.ws_dot),
                     self.dot_bytes, self._s())

    def precondition(self) -> None:
        """z = s U(r / s) (+ w D^-1 r), s = rms(r) from the current rr."""
        s = self._s()
        _native.call("warp_amd_rms_scales", self._sp(0), self.N, self._sp(1), self._sp(2), s)
        _native.call("warp_amd_unet_pack_input", self.N, _native.ptr(self.r), _native.ptr(self.sys.logk_std),
                     self._sp(2), _native.ptr(self.inp), s)
        out = self.net.forward(self.inp)
        _native.call("warp_amd_unet_unpack_output", self.N, _native.ptr(out), self._sp(1), _native.ptr(self.z), s)
        if self.w != 0.0:
            _native.call("warp_amd_diag_mv_f64", self.N, _native.ptr(self.dinv), _native.ptr(self.r), _native.ptr(self.z),
                         _native.ptr(self.z), self.w, 1.0, s)

    def iteration(self, cur: int) -> None:
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
                     _native.ptr(self.q[cur]), _native.ptr(self.x), _native.ptr(self.r), self._sp(0),
                     _native.ptr(self.ws_up), self.ws_up.numel(), s)

    def solve(self, x: torch.Tensor, tol: float = 1e-8, maxiter: int = 500, use_graph: bool = True,
              stagnation_window: int = 20, stagnation_factor: float = 0.5, fallback: bool = True):
        """Solve A x = b from the initial guess x (FP64, updated in place).

        Returns dict(iterations, relative_residual (recursive), history,
        fallback_iterations)."""
        self.x = x
        self.sys.residual(x, self.r, self.sc[0:1], self.ws_mv)
        bnorm = float(torch.linalg.vector_norm(self.sys.b).item())
        for t in self.p + self.q:
            t.zero_()
        self.sc[4] = 1.0
        self.sc[5] = 1.0
        hist = [math.sqrt(float(self.sc[0].item())) / bnorm]
        it = 0
        if use_graph and self._graph is None:
            # warm up both halves eagerly (allocates every buffer), then capture
            saved = (x.clone(), self.r.clone(), self.sc.clone())
            self.iteration(0)
            self.iteration(1)
            torch.cuda.synchronize()
            x.copy_(saved[0]); self.r.copy_(saved[1]); self.sc.copy_(saved[2])
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
                break
        fb_iters = 0
        if hist[-1] > tol and fallback:
            from . import refine

            res = refine.cg_refined(self.sys.A, self.sys.b, x, tol=tol, M="diag")
            fb_iters = res.inner_iterations
            hist.append(res.relative_residual)
        return dict(iterations=it, relative_residual=hist[-1], history=hist, fallback_iterations=fb_iters)
