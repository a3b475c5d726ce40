"""Conjugate gradient on MI355X: the ``warp.optim.linear`` CG drop-in.

Mirrors ``LinearOperator``, ``aslinearoperator``, ``preconditioner``, ``CG`` and
``cg`` of the reference (warp/optim/linear.py:13-24 ->
warp/_src/optim/linear.py) with the same arguments, defaults, tolerance rules
and loop semantics.  Two execution routes, both on the gfx950 library:

* native: a compact BSR matrix with 1x1 or 3x3 blocks and no / Jacobi /
  block-Jacobi preconditioner is solved by the library's own host loop over
  three fused kernels per iteration (``warp_amd_cg_solve_*``);
* generic: any other ``LinearOperator`` (dense, diagonal, user ``matvec``)
  is driven from Python with the library's dot / update primitives.

Dot products and alpha / beta are FP64 in both routes; vectors stay in the
system scalar type.
"""

from __future__ import annotations

import ctypes
import math
import operator
from typing import Callable

import torch

from .. import _native, sparse
from .. import types as wt

__all__ = ["CG", "LinearOperator", "LinearSolverState", "aslinearoperator", "cg", "preconditioner"]

PREC_NONE = 0
PREC_DIAG = 1
PREC_BLOCK = 2


class LinearOperator:
    """Linear operator ``z = alpha * A x + beta * y`` (linear.py:33-131)."""

    def __init__(
        self,
        shape: tuple[int, int],
        dtype,
        device: torch.device,
        matvec: Callable,
        batch_offsets: torch.Tensor | None = None,
        max_batch_length: int | None = None,
    ):
        if max_batch_length is not None:
            if batch_offsets is None:
                raise ValueError("max_batch_length requires batch_offsets")
            if max_batch_length < 0:
                raise ValueError("max_batch_length must be non-negative")
        self._shape = shape
        self._dtype = dtype
        self._device = device
        self._matvec = matvec
        self._batch_offsets = batch_offsets
        self._max_batch_length = max_batch_length
        # native-route descriptors (set by aslinearoperator / preconditioner)
        self._bsr: sparse.BsrMatrix | None = None
        self._prec_kind: int | None = None
        self._prec_data: torch.Tensor | None = None

    @property
    def shape(self) -> tuple[int, int]:
        return self._shape

    @property
    def dtype(self):
        return self._dtype

    @property
    def device(self) -> torch.device:
        return self._device

    @property
    def matvec(self) -> Callable:
        return self._matvec

    @property
    def scalar_type(self):
        return wt.type_scalar_type(self._dtype)

    @property
    def batch_offsets(self):
        return self._batch_offsets

    @property
    def batch_count(self) -> int:
        return 1 if self._batch_offsets is None else self._batch_offsets.shape[0] - 1

    @property
    def max_batch_length(self):
        return self._max_batch_length


def _torch_scalar(t: torch.Tensor):
    return wt.type_to_scalar(t.dtype)


def _sfx(torch_dtype) -> str:
    if torch_dtype == torch.float32:
        return "f32"
    if torch_dtype == torch.float64:
        return "f64"
    raise TypeError(f"warp_amd linear solvers support float32 and float64, got {torch_dtype}")


def aslinearoperator(A, batch_offsets=None, max_batch_length=None) -> LinearOperator | None:
    """Wrap a BSR matrix, a dense 2-D tensor or a 1-D diagonal as a LinearOperator (linear.py:135-233)."""
    if A is None or isinstance(A, LinearOperator):
        return A

    if isinstance(A, (sparse.BsrMatrix, sparse._BsrScalingExpression)):
        mat, scale = sparse._extract_matrix_and_scale(A)

        def bsr_mv(x, y, z, alpha, beta):
            if z.data_ptr() != y.data_ptr() and beta != 0.0:
                z.copy_(y)
            sparse.bsr_mv(A, x, z, alpha, beta)

        op = LinearOperator(mat.shape, mat.dtype, mat.device, bsr_mv, batch_offsets, max_batch_length)
        if scale == 1.0:
            op._bsr = mat
        return op

    if isinstance(A, torch.Tensor):
        sfx = _sfx(A.dtype)
        if A.ndim == 2:

            def dense_mv(x, y, z, alpha, beta):
                _native.call(
                    f"warp_amd_dense_mv_{sfx}",
                    A.shape[0],
                    A.shape[1],
                    _native.ptr(A.contiguous()),
                    _native.ptr(x),
                    _native.ptr(y),
                    _native.ptr(z),
                    float(alpha),
                    float(beta),
                    _native.stream_of(A.device),
                )

            return LinearOperator(tuple(A.shape), _torch_scalar(A), A.device, dense_mv, batch_offsets, max_batch_length)
        if A.ndim == 1:
            op = LinearOperator((A.shape[0], A.shape[0]), _torch_scalar(A), A.device, None, batch_offsets, max_batch_length)
            op._matvec = _make_diag_mv(A, sfx)
            op._prec_kind = PREC_DIAG
            op._prec_data = A
            return op

    raise ValueError(f"Unable to create LinearOperator from {A}")


def _make_diag_mv(d: torch.Tensor, sfx: str):
    def diag_mv(x, y, z, alpha, beta):
        _native.call(
            f"warp_amd_diag_mv_{sfx}",
            d.numel(),
            _native.ptr(d),
            _native.ptr(x),
            _native.ptr(y),
            _native.ptr(z),
            float(alpha),
            float(beta),
            _native.stream_of(d.device),
        )

    return diag_mv


def _make_block_mv(inv_blocks: torch.Tensor, nrow: int, br: int, sfx: str):
    """z = alpha * blockdiag(inv_blocks) x + beta * y via the BSR product of a block-diagonal matrix."""
    offsets = torch.arange(nrow + 1, dtype=torch.int32, device=inv_blocks.device)
    columns = torch.arange(nrow, dtype=torch.int32, device=inv_blocks.device)

    def block_mv(x, y, z, alpha, beta):
        if z.data_ptr() != y.data_ptr() and beta != 0.0:
            z.copy_(y)
        _native.call(
            f"warp_amd_bsr_mv_{sfx}",
            nrow,
            nrow,
            br,
            br,
            _native.ptr(offsets),
            None,
            _native.ptr(columns),
            _native.ptr(inv_blocks),
            nrow,
            _native.ptr(x),
            _native.ptr(z),
            float(alpha),
            float(beta),
            _native.stream_of(inv_blocks.device),
        )

    return block_mv


def preconditioner(A, ptype: str = "diag") -> LinearOperator | None:
    """Build a preconditioner for ``A`` (linear.py:236-280).

    ``"diag"`` (Jacobi: inverse of the diagonal entries, 1 where zero),
    ``"diag_abs"`` (same with absolute values), ``"id"`` (None), and, as an
    extension for 3x3 (any square) block matrices, ``"block_diag"``: the
    inverse of every diagonal block.
    """
    if ptype == "id":
        return None
    if ptype not in ("diag", "diag_abs", "block_diag"):
        raise ValueError(f"Unsupported preconditioner type '{ptype}'")
    use_abs = 1 if ptype == "diag_abs" else 0

    if isinstance(A, sparse.BsrMatrix):
        br, bc = A.block_shape
        if br != bc:
            raise ValueError("Diagonal preconditioners need square blocks")
        sparse._require_gpu(A.device, "preconditioner")
        sfx = _sfx(A.scalar_type.torch_dtype)
        block = ptype == "block_diag" and br > 1
        out_shape = (A.nrow, br, br) if block else ((A.nrow,) if br == 1 else (A.nrow, br))
        inv = torch.empty(out_shape, dtype=A.scalar_type.torch_dtype, device=A.device)
        _native.call(
            f"warp_amd_bsr_jacobi_{sfx}",
            A.nrow,
            A.ncol,
            br,
            bc,
            _native.ptr(A.offsets),
            _native.ptr(A.row_counts),
            _native.ptr(A.columns),
            _native.ptr(A.values),
            use_abs,
            1 if block else 0,
            _native.ptr(inv),
            _native.stream_of(A.device),
        )
        n = A.nrow * br
        if block:
            op = LinearOperator((n, n), A.dtype, A.device, _make_block_mv(inv, A.nrow, br, sfx))
            op._prec_kind = PREC_BLOCK
        else:
            op = LinearOperator((n, n), A.dtype, A.device, _make_diag_mv(inv.view(-1), sfx))
            op._prec_kind = PREC_DIAG
        op._prec_data = inv
        return op

    if isinstance(A, torch.Tensor) and A.ndim == 2:
        if ptype == "block_diag":
            raise ValueError("block_diag preconditioner needs a BsrMatrix")
        sparse._require_gpu(A.device, "preconditioner")
        d = torch.diagonal(A).contiguous()
        den = d.abs() if use_abs else d
        inv = torch.where(d == 0, torch.ones_like(d), 1.0 / torch.where(d == 0, torch.ones_like(d), den))
        return aslinearoperator(inv)

    raise ValueError("Unsupported source matrix type for building diagonal preconditioner")


def _scalar_dof_count(x: torch.Tensor) -> int:
    return x.numel()


class LinearSolverState:
    """Pre-allocated solver state (linear.py:721-841)."""

    def __init__(
        self,
        A,
        b: torch.Tensor,
        x: torch.Tensor,
        tol: float | None = None,
        atol: float | None = None,
        maxiter: float | None = 0,
        M=None,
        callback: Callable | None = None,
        check_every: int = 10,
        use_cuda_graph: bool = True,
    ):
        self._A = aslinearoperator(A)
        self._M = aslinearoperator(M)
        self._b = b
        self._x = x
        self._tol = tol
        self._atol = atol
        self._callback = callback
        self._check_every = check_every
        self._use_cuda_graph = use_cuda_graph
        self._device = self._A.device
        self._scalar_type = self._A.scalar_type
        self._batch_count = self._A.batch_count
        if maxiter is None or maxiter == 0:
            maxiter = _scalar_dof_count(b) // self._batch_count
        self._maxiter = int(maxiter)
        self._allocate()

    def _allocate(self):
        raise NotImplementedError

    def _run(self, A, b, x, M):
        raise NotImplementedError

    def _check_compatible(self, A, b, x, M):
        if A is not self._A:
            if A.shape != self._A.shape:
                raise ValueError(f"Incompatible A.shape: expected {self._A.shape}, got {A.shape}")
            if A.dtype != self._A.dtype:
                raise ValueError(f"Incompatible A.dtype: expected {self._A.dtype}, got {A.dtype}")
            if A.device != self._A.device:
                raise ValueError(f"Incompatible A.device: expected {self._A.device}, got {A.device}")
        if b is not self._b:
            if b.shape != self._b.shape:
                raise ValueError(f"Incompatible b.shape: expected {tuple(self._b.shape)}, got {tuple(b.shape)}")
            if b.dtype != self._b.dtype:
                raise ValueError(f"Incompatible b.dtype: expected {self._b.dtype}, got {b.dtype}")
        if x is not self._x:
            if x.shape != self._x.shape:
                raise ValueError(f"Incompatible x.shape: expected {tuple(self._x.shape)}, got {tuple(x.shape)}")
            if x.dtype != self._x.dtype:
                raise ValueError(f"Incompatible x.dtype: expected {self._x.dtype}, got {x.dtype}")
        if M is not None and self._M is not None and M is not self._M:
            if M.shape != self._M.shape:
                raise ValueError(f"Incompatible M.shape: expected {self._M.shape}, got {M.shape}")

    def __call__(self, A=None, b=None, x=None, M=None):
        A_op = aslinearoperator(A) if A is not None else self._A
        M_op = aslinearoperator(M) if M is not None else self._M
        b_arr = b if b is not None else self._b
        x_arr = x if x is not None else self._x
        self._check_compatible(A_op, b_arr, x_arr, M_op)
        return self._run(A_op, b_arr, x_arr, M_op)


class CG(LinearSolverState):
    """Pre-allocated conjugate gradient solver (linear.py:847-1033).

    ``restart``: optional number of iterations between explicit residual
    recomputations ``r = b - A x``.
    """

    def __init__(
        self,
        A,
        b,
        x,
        tol=None,
        atol=None,
        maxiter=0,
        M=None,
        callback=None,
        check_every=10,
        use_cuda_graph=True,
        *,
        restart: int | None = None,
    ):
        if restart is not None:
            if isinstance(restart, bool):
                raise TypeError("restart must be an integer or None")
            try:
                restart = operator.index(restart)
            except TypeError as error:
                raise TypeError("restart must be an integer or None") from error
            if restart < 1:
                raise ValueError("restart must be positive")
        self._restart = restart
        super().__init__(A, b, x, tol, atol, maxiter, M, callback, check_every, use_cuda_graph)

    def _allocate(self):
        sparse._require_gpu(self._b.device, "cg")
        n = _scalar_dof_count(self._b)
        lib = _native.lib()
        self._ws_bytes = lib.warp_amd_cg_workspace_size(n, self._b.element_size())
        self._ws = torch.zeros(self._ws_bytes, dtype=torch.uint8, device=self._b.device)
        self._dot_ws = None
        self._batch_plan = None

    # -- native route ------------------------------------------------------
    def _native_ok(self, A: LinearOperator, b, x, M) -> bool:
        bsr = A._bsr
        if bsr is None or bsr.row_counts is not None or A.batch_offsets is not None:
            return False
        br, bc = bsr.block_shape
        if br != bc or br not in (1, 3):
            return False
        if M is not None and (M._prec_kind not in (PREC_DIAG, PREC_BLOCK) or M._prec_data is None):
            return False
        if M is not None and M._prec_data.numel() not in (bsr.nrow * br, bsr.nrow * br * br):
            return False
        return b.is_contiguous() and x.is_contiguous() and b.dtype == x.dtype == bsr.scalar_type.torch_dtype

    def _run_batched(self, A, b, x, M):
        return _batched_cg(self, A, b, x, M)

    def _run(self, A, b, x, M):
        if self._native_ok(A, b, x, M):
            return self._run_native(A, b, x, M)
        if A.batch_offsets is not None:
            return self._run_batched(A, b, x, M)
        return self._run_generic(A, b, x, M)

    def _run_native(self, A, b, x, M):
        bsr = A._bsr
        br = bsr.block_shape[0]
        sfx = _sfx(b.dtype)
        prec_kind = PREC_NONE if M is None else M._prec_kind
        prec_ptr = None if M is None else _native.ptr(M._prec_data)
        nan = float("nan")
        tol = nan if self._tol is None else float(self._tol)
        atol = nan if self._atol is None else float(self._atol)

        user_cb = self._callback
        cb = _native.CG_CALLBACK(0)
        if user_cb is not None and self._check_every > 0:

            def _trampoline(it, err, at, _user):
                user_cb(int(it), float(err), float(at))
                return 0

            cb = _native.CG_CALLBACK(_trampoline)

        iters = ctypes.c_int(0)
        rnorm = ctypes.c_double(0.0)
        atol_out = ctypes.c_double(0.0)
        want_host = self._check_every > 0
        _native.call(
            f"warp_amd_cg_solve_{sfx}",
            bsr.nrow,
            br,
            _native.ptr(bsr.offsets),
            _native.ptr(bsr.columns),
            _native.ptr(bsr.values),
            bsr._lane_nnz(wait=want_host),
            prec_kind,
            prec_ptr,
            _native.ptr(b),
            _native.ptr(x),
            tol,
            atol,
            self._maxiter,
            int(self._check_every),
            int(self._restart or 0),
            1 if self._use_cuda_graph else 0,
            cb,
            None,
            _native.ptr(self._ws),
            self._ws_bytes,
            ctypes.byref(iters) if want_host else None,
            ctypes.byref(rnorm) if want_host else None,
            ctypes.byref(atol_out) if want_host else None,
            _native.stream_of(b.device),
        )
        if want_host:
            return int(iters.value), float(rnorm.value), float(atol_out.value)
        return self._device_results()

    def _device_results(self):
        lib = _native.lib()
        o_it = lib.warp_amd_cg_state_offset(0)
        o_rr = lib.warp_amd_cg_state_offset(1)
        o_at = lib.warp_amd_cg_state_offset(2)
        it = self._ws[o_it : o_it + 4].view(torch.int32)
        rr = self._ws[o_rr : o_rr + 8].view(torch.float64)
        at = self._ws[o_at : o_at + 8].view(torch.float64)
        return it, rr, at

    # -- generic route -----------------------------------------------------
    def _run_generic(self, A, b, x, M):
        dev = b.device
        sfx = _sfx(b.dtype)
        n = _scalar_dof_count(b)
        stream = _native.stream_of(dev)
        lib = _native.lib()
        if self._dot_ws is None:
            self._dot_bytes = lib.warp_amd_dot_workspace_size(n)
            self._dot_ws = torch.zeros(self._dot_bytes, dtype=torch.uint8, device=dev)
            self._vecs = torch.zeros((4, n), dtype=b.dtype, device=dev)
            self._scal = torch.zeros(8, dtype=torch.float64, device=dev)
        r, z, p, Ap = self._vecs[0], self._vecs[1], self._vecs[2], self._vecs[3]
        bf = b.reshape(-1)
        xf = x.view(-1)
        sc = self._scal
        bb, atol_sq, rr, rz0, rz1, pAp = (sc[i : i + 1] for i in range(6))
        rz = [rz0, rz1]

        def dot(a, c, out):
            _native.call(
                f"warp_amd_dot_{sfx}", n, _native.ptr(a), _native.ptr(c), _native.ptr(out),
                _native.ptr(self._dot_ws), self._dot_bytes, stream,
            )

        nan = float("nan")
        dot(bf, bf, bb)
        _native.call(
            "warp_amd_cg_tolerance", _native.ptr(bb),
            nan if self._tol is None else float(self._tol),
            nan if self._atol is None else float(self._atol),
            1 if b.dtype == torch.float64 else 0, _native.ptr(atol_sq), stream,
        )

        cur = [0]

        def update_rr_rz(slot):
            if M is None:
                dot(r, r, rr)
                dot(r, r, rz[slot])
            else:
                M.matvec(r, z, z, 1.0, 0.0)
                dot(r, r, rr)
                dot(r, z, rz[slot])

        A.matvec(xf, bf, r, -1.0, 1.0)
        update_rr_rz(cur[0])
        p.copy_(r if M is None else z)

        def do_iteration(update_direction=True):
            A.matvec(p, Ap, Ap, 1.0, 0.0)
            dot(p, Ap, pAp)
            _native.call(
                f"warp_amd_cg_step_xr_{sfx}", n, _native.ptr(xf), _native.ptr(r), _native.ptr(p), _native.ptr(Ap),
                _native.ptr(rr), _native.ptr(atol_sq), _native.ptr(rz[cur[0]]), _native.ptr(pAp), stream,
            )
            if update_direction:
                nxt = 1 - cur[0]
                update_rr_rz(nxt)
                zz = r if M is None else z
                _native.call(
                    f"warp_amd_cg_step_p_{sfx}", n, _native.ptr(zz), _native.ptr(p), _native.ptr(rr),
                    _native.ptr(atol_sq), _native.ptr(rz[nxt]), _native.ptr(rz[cur[0]]), stream,
                )
                cur[0] = nxt

        restart = self._restart
        maxiter = self._maxiter
        check_every = self._check_every
        if restart is None:
            cycle_size = 1

            def do_cycle():
                do_iteration()
        else:
            cycle_size = max(1, min(restart, maxiter))
            if check_every > 0:
                check_every = max(check_every, cycle_size)

            def do_cycle():
                for _ in range(cycle_size - 1):
                    do_iteration()
                do_iteration(update_direction=False)
                A.matvec(xf, bf, r, -1.0, 1.0)
                update_rr_rz(cur[0])
                p.copy_(r if M is None else z)

        if check_every <= 0:
            # device-side state only: run every cycle (alpha/beta freeze at convergence)
            for _ in range(0, maxiter, cycle_size):
                do_cycle()
            iters = torch.full((1,), maxiter, dtype=torch.int32, device=dev)
            return iters, rr.clone(), atol_sq.clone()

        return _host_loop(do_cycle, cycle_size, rr, atol_sq, maxiter, self._callback, check_every)


def _batched_cg(state, A, b, x, M):
    """CG over the subproblems of ``A.batch_offsets`` (linear.py:893-1033 with
    batches): per-subproblem dots, tolerances, alpha / beta and freezing;
    inactive dofs are left unchanged."""
    dev = b.device
    sfx = _sfx(b.dtype)
    n = _scalar_dof_count(b)
    nent = b.shape[0] if b.dim() > 0 else 1
    dpe = n // max(nent, 1)
    stream = _native.stream_of(dev)
    lib = _native.lib()
    offs = A.batch_offsets
    if offs.dtype != torch.int32 or offs.device != dev or not offs.is_contiguous():
        raise TypeError("batch_offsets must be a contiguous int32 array on the solver's device")
    nb = offs.numel() - 1
    if nb < 1:
        raise ValueError("batch_offsets must hold at least two entries")
    plan = state._batch_plan
    if plan is None or plan[0] is not offs:
        chunk_start = torch.empty(nb + 1, dtype=torch.int32, device=dev)
        ws_bytes = lib.warp_amd_seg_dot_plan_workspace_size(nb)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        _native.call("warp_amd_seg_dot_plan", nb, _native.ptr(offs), _native.ptr(chunk_start), _native.ptr(ws),
                     ws_bytes, stream)
        max_chunks = lib.warp_amd_seg_dot_max_chunks(n, nb)
        partials = torch.empty(max(max_chunks, 1), dtype=torch.float64, device=dev)
        vecs = torch.zeros((4, n), dtype=b.dtype, device=dev)
        scal = torch.zeros((6, nb), dtype=torch.float64, device=dev)
        plan = state._batch_plan = (offs, chunk_start, max_chunks, partials, vecs, scal)
    _, chunk_start, max_chunks, partials, vecs, scal = plan
    r, z, p, Ap = vecs[0], vecs[1], vecs[2], vecs[3]
    bf = b.reshape(-1)
    xf = x.view(-1)
    bb, atol_sq, rr, rz0, rz1, pAp = (scal[i] for i in range(6))
    rz = [rz0, rz1]

    def dot(a, c, out):
        _native.call(f"warp_amd_seg_dot_{sfx}", n, nb, _native.ptr(offs), _native.ptr(chunk_start), max_chunks,
                     _native.ptr(a), _native.ptr(c), _native.ptr(partials), _native.ptr(out), stream)

    nan = float("nan")
    dot(bf, bf, bb)
    _native.call("warp_amd_cg_tolerance_batched", nb, _native.ptr(bb),
                 nan if state._tol is None else float(state._tol), nan if state._atol is None else float(state._atol),
                 1 if b.dtype == torch.float64 else 0, _native.ptr(atol_sq), stream)
    Ap.zero_()
    z.zero_()
    cur = [0]

    def update_rr_rz(slot):
        if M is None:
            dot(r, r, rr)
            dot(r, r, rz[slot])
        else:
            M.matvec(r, z, z, 1.0, 0.0)
            dot(r, r, rr)
            dot(r, z, rz[slot])

    A.matvec(xf, bf, r, -1.0, 1.0)
    update_rr_rz(cur[0])
    p.copy_(r if M is None else z)

    def do_iteration(update_direction=True):
        A.matvec(p, Ap, Ap, 1.0, 0.0)
        dot(p, Ap, pAp)
        _native.call(f"warp_amd_cg_step_xr_batched_{sfx}", nent, dpe, nb, _native.ptr(offs), _native.ptr(xf),
                     _native.ptr(r), _native.ptr(p), _native.ptr(Ap), _native.ptr(rr), _native.ptr(atol_sq),
                     _native.ptr(rz[cur[0]]), _native.ptr(pAp), stream)
        if update_direction:
            nxt = 1 - cur[0]
            update_rr_rz(nxt)
            zz = r if M is None else z
            _native.call(f"warp_amd_cg_step_p_batched_{sfx}", nent, dpe, nb, _native.ptr(offs), _native.ptr(zz),
                         _native.ptr(p), _native.ptr(rr), _native.ptr(atol_sq), _native.ptr(rz[nxt]),
                         _native.ptr(rz[cur[0]]), stream)
            cur[0] = nxt

    restart = state._restart
    maxiter = state._maxiter
    check_every = state._check_every
    if restart is None:
        cycle_size = 1

        def do_cycle():
            do_iteration()
    else:
        cycle_size = max(1, min(restart, maxiter))
        if check_every > 0:
            check_every = max(check_every, cycle_size)

        def do_cycle():
            for _ in range(cycle_size - 1):
                do_iteration()
            do_iteration(update_direction=False)
            A.matvec(xf, bf, r, -1.0, 1.0)
            update_rr_rz(cur[0])
            p.copy_(r if M is None else z)

    if check_every <= 0:
        for _ in range(0, maxiter, cycle_size):
            do_cycle()
        iters = torch.full((1,), maxiter, dtype=torch.int32, device=dev)
        return iters, rr.clone(), atol_sq.clone()
    return _host_loop(do_cycle, cycle_size, rr, atol_sq, maxiter, state._callback, check_every)


def _host_loop(do_cycle, cycle_size, rr, atol_sq, maxiter, callback, check_every):
    """Residual checks every ``check_every`` iterations (linear.py:2182-2243).
    ``rr`` / ``atol_sq`` hold one value per subproblem: the loop stops when
    every subproblem meets its own tolerance, and reports the largest residual
    and tolerance."""
    atol_sq_h = atol_sq.detach().to("cpu", torch.float64).reshape(-1)
    atol = math.sqrt(float(atol_sq_h.max()))
    check_every = max(check_every, cycle_size)
    cur_iter = 0

    def read():
        rr_h = rr.detach().to("cpu", torch.float64).reshape(-1)
        return float(rr_h.max()), bool((rr_h <= atol_sq_h).all())

    err_sq, done = read()
    if callback is not None:
        callback(cur_iter, math.sqrt(err_sq), atol)
    if done:
        return cur_iter, math.sqrt(err_sq), atol
    while True:
        do_cycle()
        cur_iter += cycle_size
        if cur_iter >= maxiter:
            break
        if (cur_iter % check_every) < cycle_size:
            err_sq, done = read()
            if done:
                break
            if callback is not None:
                callback(cur_iter, math.sqrt(err_sq), atol)
    err = math.sqrt(read()[0])
    if callback is not None:
        callback(cur_iter, err, atol)
    return cur_iter, err, atol


def cg(
    A,
    b: torch.Tensor,
    x: torch.Tensor,
    tol: float | None = None,
    atol: float | None = None,
    maxiter: float | None = 0,
    M=None,
    callback: Callable | None = None,
    check_every=10,
    use_cuda_graph=True,
    run: bool = True,
    *,
    restart: int | None = None,
):
    """Solve the SPD system ``A x = b`` with (preconditioned) CG (linear.py:1035-1121).

    Returns ``(iterations, residual_norm, absolute_tolerance)`` when
    ``check_every > 0``; device tensors ``(iterations, residual_norm_sq,
    absolute_tolerance_sq)`` when ``check_every == 0``; the :class:`CG`
    functor when ``run`` is False.  The stopping test is
    ``||r||^2 <= max(tol * ||b||, atol)^2`` on the recursive residual.
    """
    state = CG(
        A, b, x, tol=tol, atol=atol, maxiter=maxiter, M=M, callback=callback, check_every=check_every,
        use_cuda_graph=use_cuda_graph, restart=restart,
    )
    if run:
        return state()
    return state
