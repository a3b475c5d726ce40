"""Block-sparse row (BSR) matrices on MI355X: the ``warp.sparse`` drop-in.

Same names, argument order, defaults and validation messages as the
reference (warp/sparse.py:25-51 -> warp/_src/sparse.py); arrays are torch
tensors (see :mod:`warp_amd.types` for the layout).  All work runs in the
hand-written gfx950 library through ctypes; there is no CPU fallback.

Covered here: ``BsrMatrix`` (compact and padded topologies, the sticky
status of padded operations), ``bsr_zeros``, ``bsr_set_zero``,
``bsr_set_from_triplets`` (compact, masked, padded), ``bsr_from_triplets``,
``bsr_compress``, ``bsr_set_transpose`` / ``bsr_transposed``, ``bsr_assign`` /
``bsr_copy``, ``bsr_mv`` (plain and transposed) and ``bsr_get_diag``.
"""

from __future__ import annotations

import ctypes
import warnings

import torch

from . import _native
from . import types as wt
from .types import block_shape_of, type_is_matrix, type_scalar_type, type_size

__all__ = [
    "BSR_STATUS_ROW_CAPACITY_EXCEEDED",
    "BSR_STATUS_SUCCESS",
    "BsrMatrix",
    "bsr_assign",
    "bsr_axpy",
    "bsr_axpy_work_arrays",
    "bsr_block_index",
    "bsr_compress",
    "bsr_copy",
    "bsr_diag",
    "bsr_from_triplets",
    "bsr_get_diag",
    "bsr_identity",
    "bsr_matrix_t",
    "bsr_mm",
    "bsr_mm_work_arrays",
    "bsr_mv",
    "bsr_row_index",
    "bsr_scale",
    "bsr_set_diag",
    "bsr_set_from_triplets",
    "bsr_set_identity",
    "bsr_set_transpose",
    "bsr_set_zero",
    "bsr_transposed",
    "bsr_zeros",
]

BSR_STATUS_SUCCESS = 0
BSR_STATUS_ROW_CAPACITY_EXCEEDED = 1


def _bsr_status_message(status: int) -> str:
    """Text of a sparse status code (sparse.py:873-878)."""
    if status == BSR_STATUS_SUCCESS:
        return "success"
    if status == BSR_STATUS_ROW_CAPACITY_EXCEEDED:
        return "row capacity exceeded"
    return f"unknown status {status}"


def _warn_masked_arg_deprecated(func_name: str) -> None:
    warnings.warn(
        f"The `masked` argument to {func_name}() is deprecated; pass topology='masked' instead.",
        DeprecationWarning,
        stacklevel=3,
    )


# Scalar bit masks used to detect numerical-zero blocks (sparse.py:993-1006).
_zero_value_masks = {
    wt.float16: 0x7FFF,
    wt.float32: 0x7FFFFFFF,
    wt.float64: 0x7FFFFFFFFFFFFFFF,
}


def _require_gpu(device: torch.device, what: str) -> None:
    if device.type != "cuda":
        raise RuntimeError(f"{what}: warp_amd runs on ROCm GPU devices only, got device '{device}'")


def _to_device(device) -> torch.device:
    if device is None:
        return torch.device("cuda", torch.cuda.current_device()) if torch.cuda.is_available() else torch.device("cuda")
    dev = torch.device(device)
    if dev.type == "cuda" and dev.index is None and torch.cuda.is_available():
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


class BsrMatrix:
    """Compact BSR / CSR matrix (reference ``BsrMatrix``, warp/_src/sparse.py:103-463).

    Attributes:
        nrow, ncol: numbers of block rows / columns.
        nnz: host-side upper bound on stored blocks; exact after :meth:`nnz_sync`.
        offsets: int32 tensor of size ``nrow + 1``.
        row_counts: ``None`` for compact matrices.
        columns: int32 tensor of size ``>= nnz``.
        values: tensor of shape ``(>= nnz, br, bc)`` (or ``(>= nnz,)`` for CSR).
    """

    def __init__(self, block_type, device: torch.device):
        self._block_type = block_type
        self.nrow = 0
        self.ncol = 0
        self.nnz = 0
        self.offsets: torch.Tensor | None = None
        self.row_counts: torch.Tensor | None = None
        self.columns: torch.Tensor | None = None
        self.values: torch.Tensor | None = None
        self._device = device
        self._nnz_buf: torch.Tensor | None = None
        self._nnz_event = None

    # -- type information -------------------------------------------------
    @property
    def scalar_type(self) -> wt.ScalarType:
        return type_scalar_type(self._block_type)

    @property
    def block_shape(self) -> tuple[int, int]:
        return block_shape_of(self._block_type)

    @property
    def block_size(self) -> int:
        return type_size(self._block_type)

    @property
    def shape(self) -> tuple[int, int]:
        br, bc = self.block_shape
        return (self.nrow * br, self.ncol * bc)

    @property
    def dtype(self):
        return self._block_type

    @property
    def device(self) -> torch.device:
        return self.offsets.device

    @property
    def scalar_values(self) -> torch.Tensor:
        """``values`` viewed as a ``(capacity, br, bc)`` scalar tensor."""
        br, bc = self.block_shape
        return self.values.view(-1, br, bc)

    # -- nnz bookkeeping (sparse.py:210-340) -------------------------------
    def _setup_nnz_transfer(self):
        if self._nnz_buf is None and self.device.type == "cuda":
            self._nnz_buf = torch.zeros(1, dtype=torch.int32, pin_memory=True)
            self._nnz_event = torch.cuda.Event()
        return self._nnz_buf, self._nnz_event

    def _copy_nnz_async(self):
        buf, event = self._setup_nnz_transfer()
        if buf is not None:
            buf.copy_(self.offsets[self.nrow : self.nrow + 1], non_blocking=True)
            event.record(torch.cuda.current_stream(self.device))
        return buf, event

    def nnz_sync(self) -> int:
        """Wait for the pending ``offsets[nrow]`` transfer (or start one) and return the exact nnz."""
        if self._nnz_event is None or not getattr(self, "_nnz_pending", False):
            self._copy_nnz_async()
        if self._nnz_event is not None:
            self._nnz_event.synchronize()
            self.nnz = int(self._nnz_buf[0])
        else:
            self.nnz = int(self.offsets[self.nrow])
        self._nnz_pending = False
        return self.nnz

    # -- topology version: the cached mirror-symmetry verdict is valid for one
    # version; assigning offsets / columns / row_counts or any operation of
    # this module that rewrites them starts a new one
    def __setattr__(self, name, value):
        if name in ("offsets", "columns", "row_counts", "nrow", "ncol"):
            self.__dict__["_topology_version"] = self.__dict__.get("_topology_version", 0) + 1
        object.__setattr__(self, name, value)

    def _bump_topology(self) -> None:
        self.__dict__["_topology_version"] = self.__dict__.get("_topology_version", 0) + 1

    def _lane_nnz(self, wait: bool = False) -> int:
        """nnz used to choose the lanes per block row of the products: exact
        once the pending ``offsets[nrow]`` transfer has landed (a non-blocking
        event query), or after waiting for it when ``wait`` is set and no graph
        capture is running; otherwise the host-side upper bound.  After a
        compact assembly that bound is the triplet count, which would pick
        twice the lanes a 27-block row needs."""
        if getattr(self, "_nnz_pending", False) and self._nnz_event is not None:
            if self._nnz_event.query():
                self.nnz = int(self._nnz_buf[0])
                self._nnz_pending = False
            elif wait and not torch.cuda.is_current_stream_capturing():
                self.nnz_sync()
        return int(self.nnz)

    def notify_nnz_changed(self, nnz: int | None = None, nnz_capacity: int | None = None) -> None:
        """Notify that ``offsets`` changed outside this module (sparse.py:311-333)."""
        self._bump_topology()
        self._nnz_pending = False
        if nnz is None:
            if nnz_capacity is None:
                self.nnz_sync()
            else:
                self.nnz = nnz_capacity
        else:
            self.nnz = int(nnz)
        _bsr_ensure_fits(self, capacity=nnz_capacity)

    def uncompress_rows(self, out: torch.Tensor | None = None) -> torch.Tensor:
        """Row index of every stored block slot, ``-1`` outside the active blocks (sparse.py:164-193)."""
        if out is None:
            out = torch.empty(self.nnz, dtype=torch.int32, device=self.device)
        n = min(int(self.nnz), out.numel())
        if self.device.type != "cuda":
            # host-side bookkeeping of CPU-resident matrices (no device work involved)
            starts = self.offsets[: self.nrow].to(torch.int64)
            if self.row_counts is None:
                counts = self.offsets[1 : self.nrow + 1].to(torch.int64) - starts
            else:
                counts = torch.minimum(self.row_counts[: self.nrow].to(torch.int64),
                                       self.offsets[1 : self.nrow + 1].to(torch.int64) - starts)
            rows = torch.repeat_interleave(torch.arange(self.nrow, dtype=torch.int32), counts)
            first = torch.repeat_interleave(torch.cumsum(counts, 0) - counts, counts)
            pos = torch.repeat_interleave(starts, counts) + torch.arange(rows.numel()) - first
            out[:n].fill_(-1)
            keep = pos < n
            out[pos[keep]] = rows[keep]
            return out
        _native.call(
            "warp_amd_bsr_uncompress_rows", self.nrow, _native.ptr(self.offsets), _native.ptr(self.row_counts), n,
            _native.ptr(out), _native.stream_of(self.device),
        )
        return out

    # -- sticky status of padded operations (sparse.py:265-310) -----------
    def _ensure_status(self) -> torch.Tensor:
        status = getattr(self, "_status", None)
        if status is None or status.device != self.device:
            status = torch.zeros(1, dtype=torch.int32, device=self.device)
            self._status = status
        return status

    def _status_if_any(self) -> torch.Tensor | None:
        return getattr(self, "_status", None)

    def status_sync(self) -> int:
        """Status code of the padded operations since the last :meth:`clear_status` (synchronizes).

        The code is sticky: operations can raise it, successful ones do not clear it."""
        status = self._status_if_any()
        if status is None:
            return BSR_STATUS_SUCCESS
        if status.device.type == "cuda" and torch.cuda.is_current_stream_capturing():
            raise RuntimeError(
                "BsrMatrix.status_sync() cannot read sparse status during a live CUDA graph capture "
                "because it requires a device-to-host readback. Query the status outside the capture."
            )
        return int(status[0])

    def status_message(self) -> str:
        """Human-readable text of :meth:`status_sync`."""
        return _bsr_status_message(self.status_sync())

    def clear_status(self) -> None:
        """Reset the sparse status code to success."""
        status = self._status_if_any()
        if status is not None:
            status.zero_()

    # -- operators (sparse.py:411-462) -------------------------------------
    def transpose(self):
        """Transposed copy of this matrix (sparse.py:460-462)."""
        return bsr_transposed(self)

    def __add__(self, y):
        return bsr_axpy(y, bsr_copy(self))

    def __iadd__(self, y):
        return bsr_axpy(y, self)

    def __radd__(self, x):
        return bsr_axpy(x, bsr_copy(self))

    def __sub__(self, y):
        return bsr_axpy(y, bsr_copy(self), alpha=-1.0)

    def __rsub__(self, x):
        return bsr_axpy(x, bsr_copy(self), beta=-1.0)

    def __isub__(self, y):
        return bsr_axpy(y, self, alpha=-1.0)

    def __imul__(self, y):
        return bsr_scale(self, y)

    def __matmul__(self, x):
        if isinstance(x, torch.Tensor):
            return bsr_mv(self, x)
        return bsr_mm(self, x)

    def __rmatmul__(self, x):
        """x @ A = A^T x (reference BsrMatrix.__rmatmul__)."""
        if isinstance(x, torch.Tensor):
            return bsr_mv(self, x, transpose=True)
        return bsr_mm(x, self)

    def __imatmul__(self, y):
        return bsr_mm(self, y, self)

    def __mul__(self, s):
        return _BsrScalingExpression(self, s)

    def __rmul__(self, s):
        return _BsrScalingExpression(self, s)

    def __truediv__(self, s):
        return _BsrScalingExpression(self, 1.0 / s)

    def __neg__(self):
        return _BsrScalingExpression(self, -1.0)


class _BsrScalingExpression:
    """Lazy ``scale * A`` (sparse.py:1716-1790)."""

    def __init__(self, mat: BsrMatrix, scale: float):
        self.mat = mat
        self.scale = scale

    def __matmul__(self, x):
        if isinstance(x, torch.Tensor):
            return bsr_mv(self, x)
        return bsr_mm(self.mat, x, alpha=self.scale)

    def __rmatmul__(self, x):
        if isinstance(x, torch.Tensor):
            return bsr_mv(self, x, transpose=True)
        return bsr_mm(x, self.mat, alpha=self.scale)

    def __add__(self, y):
        return bsr_axpy(y, bsr_copy(self.mat), alpha=self.scale)

    def __radd__(self, x):
        return bsr_axpy(x, bsr_copy(self.mat), beta=self.scale)

    def __sub__(self, y):
        return bsr_axpy(y, bsr_copy(self.mat), alpha=-self.scale)

    def __rsub__(self, x):
        return bsr_axpy(x, bsr_copy(self.mat), beta=-self.scale)

    def __mul__(self, s):
        return _BsrScalingExpression(self.mat, self.scale * s)

    __rmul__ = __mul__

    def __neg__(self):
        return _BsrScalingExpression(self.mat, -self.scale)

    def __truediv__(self, s):
        return _BsrScalingExpression(self.mat, self.scale / s)

    def eval(self) -> BsrMatrix:
        """Materialize ``scale * A`` as a new matrix (sparse.py:1684-1685)."""
        return bsr_copy(self)

    def transpose(self):
        """Transposed copy, still scaled (sparse.py:1776-1778)."""
        return _BsrScalingExpression(self.mat.transpose(), self.scale)

    @property
    def nnz(self):
        return self.mat.nnz

    @property
    def offsets(self):
        return self.mat.offsets

    @property
    def row_counts(self):
        return self.mat.row_counts

    @property
    def columns(self):
        return self.mat.columns

    @property
    def scalar_type(self):
        return self.mat.scalar_type

    @property
    def block_shape(self):
        return self.mat.block_shape

    @property
    def block_size(self):
        return self.mat.block_size

    @property
    def nrow(self):
        return self.mat.nrow

    @property
    def ncol(self):
        return self.mat.ncol

    @property
    def shape(self):
        return self.mat.shape

    @property
    def dtype(self):
        return self.mat.dtype

    @property
    def device(self):
        return self.mat.device


def _extract_matrix_and_scale(A):
    if isinstance(A, BsrMatrix):
        return A, 1.0
    if isinstance(A, _BsrScalingExpression):
        return A.mat, A.scale
    raise ValueError("Argument cannot be interpreted as a BsrMatrix")


def _bsr_ensure_fits(bsr: BsrMatrix, nnz: int | None = None, capacity: int | None = None) -> None:
    """Grow columns / values to hold ``capacity`` blocks (sparse.py:749-766)."""
    if nnz is None:
        nnz = bsr.nnz
    else:
        bsr.nnz = int(nnz)
    if capacity is None:
        capacity = nnz
    elif capacity < nnz:
        raise ValueError(f"capacity must be at least nnz, got {capacity} and {nnz}")
    if bsr.columns.numel() < capacity:
        bsr.columns = torch.empty(capacity, dtype=torch.int32, device=bsr.columns.device)
    if bsr.values.shape[0] < capacity:
        bsr.values = torch.empty((capacity, *bsr.values.shape[1:]), dtype=bsr.values.dtype, device=bsr.values.device)


def _bsr_resize(bsr: BsrMatrix, rows_of_blocks=None, cols_of_blocks=None, preserve_offsets: bool = False) -> None:
    """New dimensions; with ``preserve_offsets`` retained rows keep their capacity and added rows get none (sparse.py:693-716)."""
    old_nrow = bsr.nrow
    old_offsets = bsr.offsets
    if rows_of_blocks is not None:
        bsr.nrow = int(rows_of_blocks)
    if cols_of_blocks is not None:
        bsr.ncol = int(cols_of_blocks)
    if bsr.offsets.numel() < bsr.nrow + 1:
        bsr.offsets = torch.empty(bsr.nrow + 1, dtype=torch.int32, device=old_offsets.device)
    if preserve_offsets and bsr.nrow > old_nrow:
        if bsr.offsets is not old_offsets:
            bsr.offsets[: old_nrow + 1].copy_(old_offsets[: old_nrow + 1])
        bsr.offsets[old_nrow + 1 : bsr.nrow + 1] = bsr.offsets[old_nrow]


def _bsr_ensure_independent_row_counts(bsr: BsrMatrix, preserve_row_counts: bool = True) -> None:
    """Give ``bsr`` its own ``row_counts`` of size ``nrow`` (sparse.py:724-742)."""
    old = bsr.row_counts
    if old is not None and old.numel() >= bsr.nrow:
        if not preserve_row_counts:
            old.zero_()
        return
    device = bsr.offsets.device
    if not preserve_row_counts or bsr.nrow == 0:
        bsr.row_counts = torch.zeros(bsr.nrow, dtype=torch.int32, device=device)
    else:
        bsr.row_counts = (bsr.offsets[1 : bsr.nrow + 1] - bsr.offsets[: bsr.nrow]).to(torch.int32)


def _check_row_capacity(bsr: BsrMatrix, topology: str, row_capacity, nnz_capacity):
    if nnz_capacity is not None:
        nnz_capacity = int(nnz_capacity)
        if nnz_capacity < 0:
            raise ValueError("nnz_capacity must be nonnegative")
    if row_capacity is not None:
        if topology != "padded":
            raise ValueError("row_capacity requires topology='padded'")
        if isinstance(row_capacity, torch.Tensor):
            if row_capacity.ndim != 1 or row_capacity.shape[0] != bsr.nrow:
                raise ValueError(f"row_capacity array must have shape ({bsr.nrow},), got {tuple(row_capacity.shape)}")
            if row_capacity.dtype != torch.int32:
                raise TypeError("row_capacity array must have dtype int32")
            if row_capacity.device != bsr.offsets.device:
                raise ValueError(
                    f"row_capacity array must reside on the destination matrix device, got {row_capacity.device} and {bsr.offsets.device}"
                )
        else:
            row_capacity = int(row_capacity)
            if row_capacity < 0:
                raise ValueError("row_capacity must be nonnegative")
    return row_capacity, nnz_capacity


def _bsr_set_zero_topology(bsr: BsrMatrix, topology: str, row_capacity, nnz_capacity, preserve_row_capacity=False):
    """Empty topology: compact, or padded with the given / kept row capacity (sparse.py:580-636)."""
    row_capacity, nnz_capacity = _check_row_capacity(bsr, topology, row_capacity, nnz_capacity)
    if topology == "padded":
        if row_capacity is not None:
            if isinstance(row_capacity, torch.Tensor):
                bsr.offsets[:1].zero_()
                torch.cumsum(row_capacity, 0, out=bsr.offsets[1 : bsr.nrow + 1])
                nnz = None
            else:
                bsr.offsets[: bsr.nrow + 1] = (
                    torch.arange(bsr.nrow + 1, dtype=torch.int32, device=bsr.offsets.device) * row_capacity
                )
                nnz = bsr.nrow * row_capacity
        elif preserve_row_capacity:
            nnz = bsr.nnz
        else:
            bsr.offsets.zero_()
            nnz = 0
        _bsr_ensure_independent_row_counts(bsr, preserve_row_counts=False)
        bsr._ensure_status()
    else:
        bsr.offsets.zero_()
        bsr.row_counts = None
        nnz = 0
    bsr.notify_nnz_changed(nnz=nnz, nnz_capacity=nnz_capacity)


def bsr_zeros(
    rows_of_blocks: int,
    cols_of_blocks: int,
    block_type,
    device=None,
    *,
    topology: str | None = None,
    row_capacity=None,
    nnz_capacity: int | None = None,
) -> BsrMatrix:
    """Construct an empty BSR or CSR matrix (sparse.py:639-690).

    ``topology="padded"`` (implied by ``row_capacity``) reserves
    ``row_capacity`` slots per row (an int, or an int32 tensor of per-row
    capacities) with ``row_counts`` zero; such matrices are filled in place
    (``bsr_set_from_triplets(..., topology="padded")``, ``bsr_compress``,
    ``bsr_set_transpose``).  ``nnz_capacity`` preallocates block storage."""
    if topology is None:
        topology = "padded" if row_capacity is not None else "compact"
    elif topology not in ("compact", "padded"):
        raise ValueError(f"Unsupported topology policy: {topology}")

    if not type_is_matrix(block_type):
        if type_is_vector_like(block_type):
            raise ValueError("BsrMatrix block type must be either warp matrix or scalar")
        block_type = wt.type_to_scalar(block_type)

    dev = _to_device(device)
    bsr = BsrMatrix(block_type, dev)
    bsr.nrow = int(rows_of_blocks)
    bsr.ncol = int(cols_of_blocks)
    bsr.nnz = 0
    torch_dtype = bsr.scalar_type.torch_dtype
    bsr.offsets = torch.zeros(bsr.nrow + 1, dtype=torch.int32, device=dev)
    bsr.columns = torch.empty(0, dtype=torch.int32, device=dev)
    if type_is_matrix(block_type):
        bsr.values = torch.empty((0, *block_type._shape_), dtype=torch_dtype, device=dev)
    else:
        bsr.values = torch.empty(0, dtype=torch_dtype, device=dev)
    _bsr_set_zero_topology(bsr, topology, row_capacity, nnz_capacity)
    if topology == "padded":
        # unused slots: column -1 and zero values
        bsr.columns.fill_(-1)
        bsr.values.zero_()
    return bsr


def type_is_vector_like(t) -> bool:
    return wt.type_is_vector(t)


def bsr_set_zero(
    bsr: BsrMatrix,
    rows_of_blocks: int | None = None,
    cols_of_blocks: int | None = None,
    *,
    topology: str = "compact",
    row_capacity=None,
    nnz_capacity: int | None = None,
):
    """Set a matrix to zero, possibly resizing it (sparse.py:769-810).

    ``"compact"`` drops all blocks; ``"padded"`` keeps (or resets, with
    ``row_capacity``) the row capacity and empties every row; ``"masked"``
    keeps the topology and zeroes the values."""
    bsr._bump_topology()
    if topology not in ("compact", "padded", "masked"):
        raise ValueError(f"Unsupported topology policy: {topology}")
    if row_capacity is not None and topology != "padded":
        raise ValueError("row_capacity requires topology='padded'")
    if topology == "masked":
        if rows_of_blocks is not None or cols_of_blocks is not None:
            raise ValueError("Cannot resize a matrix with topology='masked'")
        bsr.values.zero_()
        return
    preserve = topology == "padded" and row_capacity is None
    _bsr_resize(bsr, rows_of_blocks, cols_of_blocks, preserve_offsets=preserve)
    bsr._nnz_pending = False
    _bsr_set_zero_topology(bsr, topology, row_capacity, nnz_capacity, preserve_row_capacity=preserve)


def bsr_set_from_triplets(
    dest: BsrMatrix,
    rows: torch.Tensor,
    columns: torch.Tensor,
    values: torch.Tensor | None = None,
    count: torch.Tensor | None = None,
    prune_numerical_zeros: bool = True,
    masked: bool = False,
    topology: str | None = None,
):
    """Fill ``dest`` from COO triplets, summing duplicates (sparse.py:1232-1519).

    Triplets with out-of-range indices are dropped; with
    ``prune_numerical_zeros`` (default) blocks whose scalars are all +-0 are
    dropped too.  Duplicates are summed in ascending triplet order.

    ``topology="compact"`` (default) rebuilds the topology; ``dest.nnz`` is
    then the triplet count (an upper bound) until :meth:`BsrMatrix.nnz_sync`.
    ``"padded"`` writes each row into its existing capacity
    ``offsets[r]:offsets[r+1]`` and sets ``row_counts``; a row that does not
    fit is left empty and ``dest.status_sync()`` reports
    ``BSR_STATUS_ROW_CAPACITY_EXCEEDED``.  ``"masked"`` (or the deprecated
    ``masked=True``) keeps the topology of ``dest`` and drops triplets outside
    it; no zero block is pruned.
    """
    dest._bump_topology()
    if topology is None:
        topology = "masked" if masked else "compact"
    elif topology not in ("compact", "masked", "padded"):
        raise ValueError(f"Unsupported topology policy: {topology}")
    elif masked and topology != "masked":
        raise ValueError("Cannot pass masked=True with a non-masked topology policy")

    if rows.device != columns.device or rows.device != dest.device:
        raise ValueError(
            f"Rows and columns must reside on the destination matrix device, got {rows.device}, {columns.device} and {dest.device}"
        )
    if rows.shape[0] != columns.shape[0]:
        raise ValueError(
            f"Rows and columns arrays must have the same length, got {rows.shape[0]} and {columns.shape[0]}"
        )
    if rows.dtype != torch.int32 or columns.dtype != torch.int32:
        raise TypeError("Rows and columns arrays must be of type int32")
    if count is not None:
        if count.device != rows.device:
            raise ValueError(f"Count and rows must reside on the same device, got {count.device} and {rows.device}")
        if tuple(count.shape) != (1,):
            raise ValueError(f"Count array must be a single-element array, got {tuple(count.shape)}")
        if count.dtype != torch.int32:
            raise TypeError("Count array must be of type int32")
    if values is not None:
        if values.device != rows.device:
            raise ValueError(f"Values and rows must reside on the same device, got {values.device} and {rows.device}")
        if values.shape[0] != rows.shape[0]:
            raise ValueError(
                f"Values and rows arrays must have the same length, got {values.shape[0]} and {rows.shape[0]}"
            )
        if values.dtype != dest.scalar_type.torch_dtype:
            raise ValueError(
                f"Scalar type of values array ({values.dtype}) should correspond to that of matrix ({dest.scalar_type})"
            )
        if values.ndim == 1:
            if dest.block_size != 1:
                raise ValueError(
                    f"Values array type must correspond to that of the dest matrix, got {values.dtype} and {dest.dtype}"
                )
        elif values.ndim == 3:
            if tuple(values.shape[1:]) != dest.block_shape:
                raise ValueError(
                    f"Last two dimensions in values array ({tuple(values.shape[1:])}) should correspond to matrix block shape {dest.block_shape})"
                )
        else:
            raise ValueError(f"Number of dimension for values array should be 1 or 3, got {values.ndim}")
        if prune_numerical_zeros and not values.is_contiguous():
            raise ValueError("Values array should be contiguous for numerical zero pruning")

    if masked:
        _warn_masked_arg_deprecated("bsr_set_from_triplets")

    nnz = rows.shape[0]
    if nnz == 0:
        bsr_set_zero(dest, topology=topology)
        return

    _require_gpu(dest.device, "bsr_set_from_triplets")
    scalar_type = dest.scalar_type
    zero_mask = _zero_value_masks.get(scalar_type, 0) if prune_numerical_zeros and values is not None else 0
    if values is not None:
        values = values.contiguous()
    rows = rows.contiguous()
    columns = columns.contiguous()
    lib = _native.lib()

    if topology == "padded":
        # into the existing row capacity; overflowing rows raise the sticky status
        _bsr_ensure_independent_row_counts(dest, preserve_row_counts=False)
        status = dest._ensure_status()
        block_bytes = dest.block_size * scalar_type.size if values is not None else 0
        ws_bytes = lib.warp_amd_bsr_topology_workspace_size(dest.nrow, nnz, block_bytes)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dest.device)
        _native.call(
            "warp_amd_bsr_set_from_triplets_padded",
            dest.block_size,
            scalar_type.size if values is not None else 0,
            dest.nrow,
            dest.ncol,
            nnz,
            _native.ptr(count),
            _native.ptr(rows),
            _native.ptr(columns),
            _native.ptr(values),
            zero_mask,
            _native.ptr(dest.offsets),
            _native.ptr(dest.row_counts),
            _native.ptr(dest.columns),
            _native.ptr(dest.values) if values is not None else None,
            int(dest.values.shape[0]),
            _native.ptr(status),
            _native.ptr(ws),
            ws_bytes,
            _native.stream_of(dest.device),
        )
        return

    if topology == "masked":
        # existing topology kept; blocks outside it are dropped
        if values is None:
            dest.values.zero_()
            return
        block_bytes = dest.block_size * scalar_type.size
        ws_bytes = lib.warp_amd_bsr_topology_workspace_size(dest.nrow, nnz, block_bytes)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dest.device)
        _native.call(
            "warp_amd_bsr_set_from_triplets_masked",
            dest.block_size,
            scalar_type.size,
            dest.nrow,
            dest.ncol,
            nnz,
            _native.ptr(count),
            _native.ptr(rows),
            _native.ptr(columns),
            _native.ptr(values),
            _native.ptr(dest.offsets),
            _native.ptr(dest.row_counts),
            _native.ptr(dest.columns),
            _native.ptr(dest.values),
            int(dest.values.shape[0]),
            _native.ptr(ws),
            ws_bytes,
            _native.stream_of(dest.device),
        )
        return

    _bsr_ensure_fits(dest, nnz=nnz)
    ws_bytes = lib.warp_amd_bsr_from_triplets_workspace_size(dest.nrow, nnz)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dest.device)
    nnz_buf, nnz_event = dest._setup_nnz_transfer()
    stream = _native.stream_of(dest.device)
    _native.call(
        "warp_amd_bsr_from_triplets",
        dest.block_size,
        scalar_type.size if values is not None else 0,
        dest.nrow,
        dest.ncol,
        nnz,
        _native.ptr(count),
        _native.ptr(rows),
        _native.ptr(columns),
        _native.ptr(values),
        zero_mask,
        0,
        None,
        None,
        _native.ptr(dest.offsets),
        None,
        _native.ptr(dest.columns),
        _native.ptr(dest.values) if values is not None else None,
        nnz_buf.data_ptr(),
        _native.ptr(ws),
        ws_bytes,
        stream,
    )
    nnz_event.record(torch.cuda.current_stream(dest.device))
    dest._nnz_pending = True
    dest.row_counts = None


def bsr_from_triplets(
    rows_of_blocks: int,
    cols_of_blocks: int,
    rows: torch.Tensor,
    columns: torch.Tensor,
    values: torch.Tensor,
    prune_numerical_zeros: bool = True,
) -> BsrMatrix:
    """Construct a compact BSR matrix from COO triplets (sparse.py:1522-1568)."""
    if values.ndim == 3:
        block_type = wt.matrix(shape=tuple(values.shape[1:]), dtype=values.dtype)
    else:
        block_type = wt.type_to_scalar(values.dtype)
    A = bsr_zeros(rows_of_blocks, cols_of_blocks, block_type, device=values.device)
    bsr_set_from_triplets(A, rows, columns, values, prune_numerical_zeros=prune_numerical_zeros)
    return A


def bsr_compress(src: BsrMatrix, prune_numerical_zeros: bool = True, inplace: bool = False, topology: str | None = None):
    """Sort and coalesce the active blocks of each row of ``src``; returns ``src`` (sparse.py:1571-1672).

    Active blocks are ``offsets[r] : offsets[r] + row_counts[r]`` (or up to
    ``offsets[r+1]`` for compact matrices); negative columns are ignored and,
    with ``prune_numerical_zeros``, so are all-zero blocks.  Duplicates are
    summed left to right in storage order, as the reference's in-place row
    compression does.  ``topology="compact"`` (default) packs the rows and
    sets ``row_counts`` to None; ``"padded"`` keeps each row's capacity and
    updates ``row_counts``.  ``inplace`` is accepted for signature
    compatibility: both modes run the same native, non-differentiable path
    (as in the reference, callers should not rely on buffer identity).
    """
    if not isinstance(src, BsrMatrix):
        raise ValueError("bsr_compress() requires a concrete BsrMatrix")
    if topology is None:
        topology = "compact"
    elif topology not in ("compact", "padded"):
        raise ValueError(f"Unsupported topology policy: {topology}")
    _require_gpu(src.device, "bsr_compress")
    src._bump_topology()
    compact = topology == "compact"
    if not compact and src.row_counts is None:
        src.row_counts = (src.offsets[1 : src.nrow + 1] - src.offsets[: src.nrow]).to(torch.int32)
    if src.nrow == 0:
        if compact:
            src.row_counts = None
            src.nnz = 0
        return src
    scalar_type = src.scalar_type
    if scalar_type.torch_dtype not in (torch.float32, torch.float64):
        raise TypeError("bsr_compress supports float32 and float64 matrices")
    # storage positions that can hold active blocks
    T = min(int(src.offsets[src.nrow]), src.columns.numel(), src.values.shape[0])
    lib = _native.lib()
    block_bytes = src.block_size * scalar_type.size
    ws_bytes = lib.warp_amd_bsr_compress_workspace_size(src.nrow, T, block_bytes)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=src.device)
    zero_mask = _zero_value_masks.get(scalar_type, 0) if prune_numerical_zeros else 0
    nnz_buf, nnz_event = src._setup_nnz_transfer()
    _native.call(
        "warp_amd_bsr_compress",
        src.block_size,
        scalar_type.size,
        src.nrow,
        src.ncol,
        T,
        1 if prune_numerical_zeros else 0,
        zero_mask,
        1 if compact else 0,
        _native.ptr(src.offsets),
        _native.ptr(src.row_counts),
        _native.ptr(src.columns),
        _native.ptr(src.values),
        nnz_buf.data_ptr() if compact else None,
        _native.ptr(ws),
        ws_bytes,
        _native.stream_of(src.device),
    )
    if compact:
        nnz_event.record(torch.cuda.current_stream(src.device))
        src._nnz_pending = True
        src.row_counts = None
    return src


def _positions(A: BsrMatrix) -> int:
    """Storage positions that can hold active blocks of ``A``."""
    return max(0, min(int(A.nnz), A.columns.numel(), A.values.shape[0]))


def bsr_set_transpose(dest: BsrMatrix, src, masked: bool = False, topology: str | None = None):
    """Assign the transpose of ``src`` (a matrix or ``scale * matrix``) to ``dest`` (sparse.py:2636-2850).

    ``topology="compact"`` (default) rebuilds ``dest``: transposed rows list
    their blocks by ascending source row, nothing merged.  ``"masked"`` keeps
    the topology of ``dest`` (blocks missing from ``src`` become zero).
    ``"padded"`` writes each transposed row into the existing capacity of
    ``dest`` and raises ``BSR_STATUS_ROW_CAPACITY_EXCEEDED`` for rows that do
    not fit (left empty).
    """
    dest._bump_topology()
    src, src_scale = _extract_matrix_and_scale(src)
    if topology is None:
        topology = "masked" if masked else "compact"
    elif topology not in ("compact", "masked", "padded"):
        raise ValueError(f"Unsupported topology policy: {topology}")
    elif masked and topology != "masked":
        raise ValueError("Cannot pass masked=True with a non-masked topology policy")
    if masked:
        _warn_masked_arg_deprecated("bsr_set_transpose")
    if dest.values.device != src.values.device:
        raise ValueError(
            f"All arguments must reside on the same device, got {dest.values.device} and {src.values.device}"
        )
    if dest.scalar_type != src.scalar_type:
        raise ValueError(f"All arguments must have the same scalar type, got {dest.scalar_type} and {src.scalar_type}")
    transpose_block_shape = src.block_shape[::-1]
    if dest.block_shape != transpose_block_shape:
        raise ValueError(f"Destination block shape must be {transpose_block_shape}, got {dest.block_shape}")
    _require_gpu(dest.device, "bsr_set_transpose")
    if dest.scalar_type.torch_dtype not in (torch.float32, torch.float64):
        raise TypeError("bsr_set_transpose supports float32 and float64 matrices")
    lib = _native.lib()
    stream = _native.stream_of(dest.device)
    br, bc = src.block_shape
    ssize = src.scalar_type.size

    if topology in ("masked", "padded") and (dest.nrow != src.ncol or dest.ncol != src.nrow):
        raise ValueError(
            f"Destination matrix must have {src.ncol} rows and {src.nrow} columns, got {dest.nrow} and {dest.ncol}"
        )

    if topology == "masked":
        dest.values.zero_()
        _native.call(
            "warp_amd_bsr_transpose_masked_values", ssize, br, bc, dest.nrow, _native.ptr(dest.offsets),
            _native.ptr(dest.row_counts), _native.ptr(dest.columns), src.nrow, _native.ptr(src.offsets),
            _native.ptr(src.row_counts), _native.ptr(src.columns), _native.ptr(src.values), float(src_scale),
            _native.ptr(dest.values), stream,
        )
        return

    if topology == "padded":
        status = dest._ensure_status()
        T = _positions(src)
        if T == 0:
            bsr_set_zero(dest, topology="padded")
            return
        _bsr_ensure_independent_row_counts(dest)
        dest_slots = dest.columns.numel()
        block_map = torch.empty(max(dest_slots, T), dtype=torch.int32, device=dest.device)
        ws_bytes = lib.warp_amd_bsr_transpose_workspace_size(src.ncol, T)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dest.device)
        _native.call(
            "warp_amd_bsr_transpose", src.nrow, src.ncol, T, _native.ptr(src.offsets), _native.ptr(src.row_counts),
            _native.ptr(src.columns), _native.ptr(dest.offsets), _native.ptr(dest.row_counts),
            _native.ptr(dest.columns), _native.ptr(block_map), _native.ptr(status), _native.ptr(ws), ws_bytes, stream,
        )
        count = min(dest_slots, dest.values.shape[0])
    else:
        _bsr_resize(dest, rows_of_blocks=src.ncol, cols_of_blocks=src.nrow)
        T = _positions(src)
        if T == 0:
            bsr_set_zero(dest)
            return
        _bsr_ensure_fits(dest, nnz=T)
        block_map = torch.empty(T, dtype=torch.int32, device=dest.device)
        ws_bytes = lib.warp_amd_bsr_transpose_workspace_size(src.ncol, T)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dest.device)
        _native.call(
            "warp_amd_bsr_transpose", src.nrow, src.ncol, T, _native.ptr(src.offsets), _native.ptr(src.row_counts),
            _native.ptr(src.columns), _native.ptr(dest.offsets), None, _native.ptr(dest.columns),
            _native.ptr(block_map), None, _native.ptr(ws), ws_bytes, stream,
        )
        dest.row_counts = None
        dest._copy_nnz_async()
        dest._nnz_pending = True
        count = T
    _native.call(
        "warp_amd_bsr_transpose_values", ssize, br, bc, count, _native.ptr(dest.offsets[dest.nrow :]),
        _native.ptr(block_map), _native.ptr(src.values), float(src_scale), _native.ptr(dest.values), stream,
    )


def bsr_transposed(A) -> BsrMatrix:
    """New compact matrix holding the transpose of ``A`` (sparse.py:2853-2870)."""
    if A.block_shape == (1, 1):
        block_type = A.scalar_type
    else:
        block_type = wt.matrix(shape=A.block_shape[::-1], dtype=A.scalar_type)
    transposed = bsr_zeros(A.ncol, A.nrow, block_type, device=A.device)
    bsr_set_transpose(transposed, A)
    return transposed


def bsr_scale(x, alpha) -> BsrMatrix:
    """``x := alpha * x``; returns the matrix (sparse.py, ``bsr_scale``)."""
    x, scale = _extract_matrix_and_scale(x)
    alpha = float(alpha) * scale
    if alpha != 1.0 and x.nnz > 0:
        count = min(int(x.nnz), x.values.shape[0])
        if alpha == 0.0:
            x.values.zero_()
        else:
            _require_gpu(x.device, "bsr_scale")
            _native.call(
                "warp_amd_scale_values", x.scalar_type.size, count * x.block_size, _native.ptr(x.values), alpha,
                _native.stream_of(x.device),
            )
    return x


def _split_ratios(src_shape, dest_shape):
    """(src_sub, dest_sub) pieces per block along one axis, or None when neither divides the other."""
    if src_shape >= dest_shape:
        src_sub, dest_sub = src_shape // dest_shape, 1
    else:
        src_sub, dest_sub = 1, dest_shape // src_shape
    if src_sub * dest_shape != src_shape * dest_sub:
        return None
    return src_sub, dest_sub


def _reblock(src: BsrMatrix, dest_block_shape):
    rows = _split_ratios(src.block_shape[0], dest_block_shape[0])
    if rows is None:
        raise ValueError(
            f"Incompatible dest and src block shapes; block rows must evenly divide one another (Got {dest_block_shape[0]}, {src.block_shape[0]})"
        )
    cols = _split_ratios(src.block_shape[1], dest_block_shape[1])
    if cols is None:
        raise ValueError(
            f"Incompatible dest and src block shapes; block columns must evenly divide one another (Got {dest_block_shape[1]}, {src.block_shape[1]})"
        )
    (src_subrows, dest_subrows), (src_subcols, dest_subcols) = rows, cols
    dest_nrow = (src.nrow * src_subrows) // dest_subrows
    dest_ncol = (src.ncol * src_subcols) // dest_subcols
    if src.nrow * src_subrows != dest_nrow * dest_subrows or src.ncol * src_subcols != dest_ncol * dest_subcols:
        raise ValueError(
            f"The requested block shape {tuple(dest_block_shape)} does not evenly divide the source matrix of total size {src.shape}"
        )
    return src_subrows, src_subcols, dest_subrows, dest_subcols, dest_nrow, dest_ncol


def bsr_assign(dest: BsrMatrix, src, structure_only: bool = False, masked: bool = False, topology: str | None = None):
    """Copy ``src`` (a matrix or ``scale * matrix``) into ``dest`` (sparse.py:2258-2452).

    ``dest`` may use another scalar type or a block shape whose dimensions
    divide, or are divided by, those of ``src``.  ``"compact"`` (default)
    rebuilds the topology of ``dest``; ``"masked"`` keeps it (source blocks
    outside it are dropped, missing ones become zero); ``"padded"`` copies
    each source row (or, across block shapes, the blocks its source rows
    overlap) into the existing row capacity of ``dest``, raising
    ``BSR_STATUS_ROW_CAPACITY_EXCEEDED`` for rows that do not fit."""
    dest._bump_topology()
    src, src_scale = _extract_matrix_and_scale(src)
    if dest.values.device != src.values.device:
        raise ValueError("Source and destination matrices must reside on the same device")
    if topology is None:
        topology = "masked" if masked else "compact"
    elif topology not in ("compact", "masked", "padded"):
        raise ValueError(f"Unsupported topology policy: {topology}")
    elif masked and topology != "masked":
        raise ValueError("Cannot pass masked=True with a non-masked topology policy")
    if masked:
        _warn_masked_arg_deprecated("bsr_assign")
    src_subrows, src_subcols, dest_subrows, dest_subcols, dest_nrow, dest_ncol = _reblock(src, dest.block_shape)
    _require_gpu(dest.device, "bsr_assign")
    stream = _native.stream_of(dest.device)

    if topology == "padded":
        status = dest._ensure_status()
        if dest_nrow != dest.nrow or dest_ncol != dest.ncol:
            raise ValueError(
                f"Incompatible destination matrix size, expected ({dest_nrow}, {dest_ncol}), got ({dest.nrow}, {dest.ncol})"
            )
        if dest is src:
            if not structure_only and src_scale != 1.0:
                bsr_scale(dest, src_scale)
            return
        _bsr_ensure_independent_row_counts(dest)
        if dest.block_shape != src.block_shape:
            _assign_padded_reblock(dest, src, src_scale, structure_only, status, src_subrows, src_subcols,
                                   dest_subrows, dest_subcols)
            return
        with_values = not structure_only
        if with_values:
            for m in (src, dest):
                if m.scalar_type.torch_dtype not in (torch.float32, torch.float64):
                    raise TypeError("bsr_assign supports float32 and float64 values")
        _native.call(
            "warp_amd_bsr_assign_padded", src.scalar_type.size, dest.scalar_type.size, dest.nrow, src.block_size,
            _native.ptr(src.offsets), _native.ptr(src.row_counts), _native.ptr(src.columns),
            _native.ptr(src.values) if with_values else None, float(src_scale), _native.ptr(dest.offsets),
            _native.ptr(dest.row_counts), _native.ptr(dest.columns),
            _native.ptr(dest.values) if with_values else None, _native.ptr(status), stream,
        )
        return

    if topology == "masked":
        if dest_nrow != dest.nrow or dest_ncol != dest.ncol:
            raise ValueError(
                f"Incompatible destination matrix size, expected ({dest_nrow}, {dest_ncol}), got ({dest.nrow}, {dest.ncol})"
            )
    else:
        _bsr_resize(dest, rows_of_blocks=dest_nrow, cols_of_blocks=dest_ncol)

    if dest is src and topology != "masked":
        if not structure_only and src_scale != 1.0:
            bsr_scale(dest, src_scale)
        return

    T = _positions(src)
    same_shape = src_subrows == src_subcols == dest_subrows == dest_subcols == 1
    if topology != "masked" and same_shape and src.row_counts is None:
        # compact source, same block shape: the rebuilt topology is the
        # source's own (sorted unique rows), so it is copied instead
        _bsr_ensure_fits(dest, nnz=T)
        dest.offsets[: dest.nrow + 1].copy_(src.offsets[: src.nrow + 1])
        dest.columns[:T].copy_(src.columns[:T])
        dest.row_counts = None
        dest._nnz_pending = False
        if not structure_only:
            for m in (src, dest):
                if m.scalar_type.torch_dtype not in (torch.float32, torch.float64):
                    raise TypeError("bsr_assign supports float32 and float64 values")
            dest.values[T:].zero_()
            _native.call(
                "warp_amd_copy_scaled_values", src.scalar_type.size, dest.scalar_type.size, T * src.block_size,
                _native.ptr(src.values), float(src_scale), _native.ptr(dest.values), stream,
            )
        return
    if topology != "masked":
        n = T * src_subrows * src_subcols
        dest_rows = torch.empty(n, dtype=torch.int32, device=dest.device)
        dest_cols = torch.empty(n, dtype=torch.int32, device=dest.device)
        _native.call(
            "warp_amd_bsr_assign_list_blocks", T, src_subrows, src_subcols, dest_subrows, dest_subcols, src.nrow,
            _native.ptr(src.offsets), _native.ptr(src.row_counts), _native.ptr(src.columns), _native.ptr(dest_rows),
            _native.ptr(dest_cols), stream,
        )
        if n == 0:
            bsr_set_zero(dest)
        else:
            _bsr_ensure_fits(dest, nnz=n)
            lib = _native.lib()
            ws_bytes = lib.warp_amd_bsr_from_triplets_workspace_size(dest.nrow, n)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dest.device)
            nnz_buf, nnz_event = dest._setup_nnz_transfer()
            _native.call(
                "warp_amd_bsr_from_triplets", dest.block_size, 0, dest.nrow, dest.ncol, n, None,
                _native.ptr(dest_rows), _native.ptr(dest_cols), None, 0, 0, None, None, _native.ptr(dest.offsets),
                None, _native.ptr(dest.columns), None, nnz_buf.data_ptr(), _native.ptr(ws), ws_bytes, stream,
            )
            nnz_event.record(torch.cuda.current_stream(dest.device))
            dest._nnz_pending = True
            dest.row_counts = None
    if not structure_only:
        dest.values.zero_()
        _assign_values(dest, src, src_scale, T, src_subrows, src_subcols, dest_subrows, dest_subcols)


def _assign_padded_reblock(dest, src, src_scale, structure_only, status, src_subrows, src_subcols, dest_subrows,
                           dest_subcols):
    """Padded assignment across block shapes (sparse.py:2201-2255): the
    destination topology within its row capacity, then the sub-blocks copied
    into zeroed values."""
    _native.call(
        "warp_amd_bsr_assign_padded_reblock_topology", src_subrows, src_subcols, dest_subrows, dest_subcols, src.nrow,
        dest.nrow, dest.ncol, _native.ptr(src.offsets), _native.ptr(src.row_counts), _native.ptr(src.columns),
        _native.ptr(dest.offsets), _native.ptr(dest.row_counts), _native.ptr(dest.columns), _native.ptr(status),
        _native.stream_of(dest.device),
    )
    if not structure_only:
        dest.values.zero_()
        _assign_values(dest, src, src_scale, _positions(src), src_subrows, src_subcols, dest_subrows, dest_subcols)


def _assign_values(dest, src, src_scale, T, src_subrows, src_subcols, dest_subrows, dest_subcols):
    for m in (src, dest):
        if m.scalar_type.torch_dtype not in (torch.float32, torch.float64):
            raise TypeError("bsr_assign supports float32 and float64 values")
    sbr, sbc = src.block_shape
    dbr, dbc = dest.block_shape
    _native.call(
        "warp_amd_bsr_assign_copy_blocks", src.scalar_type.size, dest.scalar_type.size, T, src_subrows, src_subcols,
        dest_subrows, dest_subcols, src.nrow, _native.ptr(src.offsets), _native.ptr(src.row_counts),
        _native.ptr(src.columns), sbr, sbc, _native.ptr(src.values), float(src_scale), _native.ptr(dest.offsets),
        _native.ptr(dest.row_counts), _native.ptr(dest.columns), dbr, dbc, _native.ptr(dest.values),
        _native.stream_of(dest.device),
    )


def bsr_copy(A, scalar_type=None, block_shape=None, structure_only: bool = False, topology: str = "compact"):
    """Copy of ``A`` (a matrix or ``scale * matrix``), possibly with another scalar type or block shape (sparse.py:2455-2601).

    ``topology="padded"`` keeps the row capacity and ``row_counts`` of ``A``;
    with another block shape, each row gets the capacity of the source rows
    it covers (times the sub-columns per source block)."""
    src, src_scale = _extract_matrix_and_scale(A)
    if topology not in ("compact", "padded"):
        raise ValueError(f"Unsupported topology policy: {topology}")
    if scalar_type is None:
        scalar_type = src.scalar_type
    if block_shape is None:
        block_shape = src.block_shape
    block_shape = tuple(block_shape)
    block_type = scalar_type if block_shape == (1, 1) else wt.matrix(shape=block_shape, dtype=scalar_type)
    if topology == "padded" and block_shape != src.block_shape:
        # capacity of the covering source rows, then the padded reblocked assignment (sparse.py:2561-2588)
        src_subrows, src_subcols, dest_subrows, dest_subcols, nrow, ncol = _reblock(src, block_shape)
        _require_gpu(src.device, "bsr_copy")
        copy = bsr_zeros(nrow, ncol, block_type, device=src.device, topology="padded")
        status = copy._ensure_status()
        _bsr_ensure_fits(copy, nnz=int(src.nnz) * src_subrows * src_subcols)
        lib = _native.lib()
        ws_bytes = lib.warp_amd_bsr_reblock_capacity_workspace_size(nrow)
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=src.device)
        _native.call(
            "warp_amd_bsr_reblock_capacity", src_subrows, src_subcols, dest_subrows, dest_subcols, src.nrow, nrow,
            _native.ptr(src.offsets), _native.ptr(copy.offsets), _native.ptr(ws), ws_bytes,
            _native.stream_of(src.device),
        )
        copy.row_counts = torch.zeros(nrow, dtype=torch.int32, device=src.device)
        copy.columns.fill_(-1)
        _assign_padded_reblock(copy, src, src_scale, structure_only, status, src_subrows, src_subcols, dest_subrows,
                               dest_subcols)
        return copy
    if topology == "padded":
        copy = bsr_zeros(src.nrow, src.ncol, block_type, device=src.device, topology="padded")
        _bsr_ensure_fits(copy, nnz=src.nnz)
        copy.offsets[: src.nrow + 1].copy_(src.offsets[: src.nrow + 1])
        if src.row_counts is None:
            copy.row_counts = (src.offsets[1 : src.nrow + 1] - src.offsets[: src.nrow]).to(torch.int32)
        else:
            copy.row_counts = src.row_counts[: src.nrow].clone()
        n = min(int(src.nnz), src.columns.numel(), src.values.shape[0])
        if n > 0:
            copy.columns[:n].copy_(src.columns[:n])
            if not structure_only:
                copy.values[:n].copy_(src.values[:n])
                if src_scale != 1.0:
                    bsr_scale(copy, src_scale)
        return copy
    copy = bsr_zeros(src.nrow, src.ncol, block_type, device=src.device)
    bsr_assign(copy, A, structure_only=structure_only, topology="compact")
    return copy


class bsr_axpy_work_arrays:
    """Reusable temporaries of :func:`bsr_axpy` (sparse.py:3314-3337)."""

    def __init__(self):
        self._buffers = {}

    def _get(self, name, n, dtype, device):
        buf = self._buffers.get(name)
        if buf is None or buf.numel() < n or buf.dtype != dtype or buf.device != device:
            buf = torch.empty(max(n, 1), dtype=dtype, device=device)
            self._buffers[name] = buf
        return buf[:n]


class bsr_mm_work_arrays(bsr_axpy_work_arrays):
    """Reusable temporaries of :func:`bsr_mm` (sparse.py:3992-4064).  The product
    topology is recomputed on every call (``reuse_topology`` gives the same result)."""


def _check_same_block_type(x: BsrMatrix, y: BsrMatrix):
    if x.values.device != y.values.device:
        raise ValueError(f"All arguments must reside on the same device, got {x.values.device} and {y.values.device}")
    if x.scalar_type != y.scalar_type or x.block_shape != y.block_shape:
        raise ValueError(
            f"Matrices must have the same block type, got ({x.block_shape}, {x.scalar_type}) and ({y.block_shape}, {y.scalar_type})"
        )
    if x.nrow != y.nrow or x.ncol != y.ncol:
        raise ValueError(
            f"Matrices must have the same number of rows and columns, got ({x.nrow}, {x.ncol}) and ({y.nrow}, {y.ncol})"
        )


def _add_blocks(dest: BsrMatrix, rows, cols, values, n: int, scale: float):
    """dest[block (rows[i], cols[i])] += scale * values[i] for i < n."""
    _native.call(
        "warp_amd_bsr_add_blocks", dest.scalar_type.size, dest.block_size, n, _native.ptr(rows), _native.ptr(cols),
        _native.ptr(values), float(scale), _native.ptr(dest.offsets), _native.ptr(dest.row_counts),
        _native.ptr(dest.columns), _native.ptr(dest.values), _native.stream_of(dest.device),
    )


def _rescale_kept(dest: BsrMatrix, beta: float, n: int):
    """Values of the first ``n`` positions := 0 + beta * values, the bits a
    compact union rebuild gives the blocks it keeps."""
    _native.call("warp_amd_rescale_values", dest.scalar_type.size, n * dest.block_size, _native.ptr(dest.values),
                 float(beta), _native.stream_of(dest.device))


def _union_is_own_topology(dest: BsrMatrix, rows, cols, n: int) -> bool:
    """True when ``dest`` is compact and holds a block at every (rows[i],
    cols[i]), i < n (rows[i] < 0 ignored): the compact union topology is then
    ``dest``'s own and the rebuild can be skipped.  Reads one count back, so
    never under graph capture (then False)."""
    if dest.row_counts is not None or torch.cuda.is_current_stream_capturing():
        return False
    missing = torch.empty(1, dtype=torch.int32, device=dest.device)
    _native.call("warp_amd_bsr_count_missing_blocks", n, _native.ptr(rows), _native.ptr(cols),
                 _native.ptr(dest.offsets), None, _native.ptr(dest.columns), _native.ptr(missing),
                 _native.stream_of(dest.device))
    return int(missing.item()) == 0


def _topology_from_triplets(dest: BsrMatrix, rows, cols, n: int):
    """Compact topology of the (row, col) pairs into ``dest`` (offsets, columns, pending nnz)."""
    _bsr_ensure_fits(dest, nnz=n)
    lib = _native.lib()
    ws_bytes = lib.warp_amd_bsr_from_triplets_workspace_size(dest.nrow, n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dest.device)
    nnz_buf, nnz_event = dest._setup_nnz_transfer()
    _native.call(
        "warp_amd_bsr_from_triplets", dest.block_size, 0, dest.nrow, dest.ncol, n, None, _native.ptr(rows),
        _native.ptr(cols), None, 0, 0, None, None, _native.ptr(dest.offsets), None, _native.ptr(dest.columns), None,
        nnz_buf.data_ptr(), _native.ptr(ws), ws_bytes, _native.stream_of(dest.device),
    )
    nnz_event.record(torch.cuda.current_stream(dest.device))
    dest._nnz_pending = True
    dest.row_counts = None


def bsr_axpy(x, y: BsrMatrix | None = None, alpha=1.0, beta=1.0, masked: bool = False, work_arrays=None,
             topology: str | None = None) -> BsrMatrix:
    """``y := alpha * x + beta * y``; returns ``y`` (sparse.py:3340-3617).

    ``"compact"`` (default) gives ``y`` the union topology; ``"masked"`` keeps
    the topology of ``y`` (blocks of ``x`` outside it are dropped);
    ``"padded"`` merges each row into the existing row capacity of ``y`` and
    raises ``BSR_STATUS_ROW_CAPACITY_EXCEEDED`` (row left unchanged, emptied)
    where the union does not fit.  ``x`` and ``y`` may alias.  Blocks present
    in both get ``beta * y + alpha * x``."""
    x, x_scale = _extract_matrix_and_scale(x)
    alpha = float(alpha) * x_scale
    beta = float(beta)
    if topology is None:
        topology = "masked" if masked else "compact"
    elif topology not in ("compact", "masked", "padded"):
        raise ValueError(f"Unsupported topology policy: {topology}")
    elif masked and topology != "masked":
        raise ValueError("Cannot pass masked=True with a non-masked topology policy")
    if masked:
        _warn_masked_arg_deprecated("bsr_axpy")
    if y is None:
        if topology != "compact":
            raise ValueError("Left-hand-side 'y' matrix must be provided for this topology policy")
        y = bsr_zeros(x.nrow, x.ncol, block_type=x.dtype, device=x.values.device)
        beta = 0.0
    y._bump_topology()
    x_nnz, y_nnz = int(x.nnz), int(y.nnz)

    if topology == "padded":
        _check_same_block_type(x, y)
        status = y._ensure_status()
        if beta == 0.0:
            if x is y:
                return bsr_scale(y, alpha)
            _bsr_ensure_independent_row_counts(y)
            _require_gpu(y.device, "bsr_axpy")
            _native.call(
                "warp_amd_bsr_assign_padded", x.scalar_type.size, y.scalar_type.size, y.nrow, y.block_size,
                _native.ptr(x.offsets), _native.ptr(x.row_counts), _native.ptr(x.columns), _native.ptr(x.values),
                alpha, _native.ptr(y.offsets), _native.ptr(y.row_counts), _native.ptr(y.columns),
                _native.ptr(y.values), _native.ptr(status), _native.stream_of(y.device),
            )
            return y
        if alpha == 0.0 or x_nnz == 0:
            return bsr_scale(y, beta)
        if x is y:
            return bsr_scale(y, alpha + beta)
        _bsr_ensure_independent_row_counts(y)
        _require_gpu(y.device, "bsr_axpy")
        _native.call(
            "warp_amd_bsr_axpy_padded", y.scalar_type.size, y.block_size, y.nrow, alpha, beta,
            _native.ptr(x.offsets), _native.ptr(x.row_counts), _native.ptr(x.columns), _native.ptr(x.values),
            _native.ptr(y.offsets), _native.ptr(y.row_counts), _native.ptr(y.columns), _native.ptr(y.values),
            _native.ptr(status), _native.stream_of(y.device),
        )
        return y

    if beta == 0.0 or y_nnz == 0:
        bsr_assign(src=x, dest=y, topology=topology)
        return bsr_scale(y, alpha)
    if alpha == 0.0 or x_nnz == 0:
        return bsr_scale(y, beta)
    if x is y:
        return bsr_scale(y, alpha + beta)
    _check_same_block_type(x, y)
    _require_gpu(y.device, "bsr_axpy")
    if y.scalar_type.torch_dtype not in (torch.float32, torch.float64):
        raise TypeError("bsr_axpy supports float32 and float64 matrices")
    device = y.values.device
    xT, yT = _positions(x), _positions(y)

    if topology == "masked":
        bsr_scale(y, beta)
        rows = torch.empty(xT, dtype=torch.int32, device=device)
        x.uncompress_rows(out=rows)
        _add_blocks(y, rows, x.columns, x.values, xT, alpha)
        return y

    work = work_arrays if work_arrays is not None else bsr_axpy_work_arrays()
    n = xT + yT
    rows = work._get("rows", n, torch.int32, device)
    cols = work._get("cols", n, torch.int32, device)
    x.uncompress_rows(out=rows[yT:])
    cols[yT:].copy_(x.columns[:xT])
    if _union_is_own_topology(y, rows[yT:], cols[yT:], xT):
        # every block of x is already in y: same values as the rebuild below
        # (zero, + beta * old, + alpha * x), without sorting y's topology again
        _rescale_kept(y, beta, yT)
        _add_blocks(y, rows[yT:], cols[yT:], x.values, xT, alpha)
        return y
    y.uncompress_rows(out=rows[:yT])
    cols[:yT].copy_(y.columns[:yT])
    old_values = work._get("old_y_values", yT * y.block_size, y.values.dtype, device)
    old_values.copy_(y.values[:yT].reshape(-1))
    _topology_from_triplets(y, rows, cols, n)
    y.values.zero_()
    _add_blocks(y, rows, cols, old_values, yT, beta)
    _add_blocks(y, rows[yT:], cols[yT:], x.values, xT, alpha)
    return y


def bsr_mm(x, y, z: BsrMatrix | None = None, alpha=1.0, beta=0.0, masked: bool = False, work_arrays=None,
           reuse_topology: bool = False, tile_size: int = 0, max_new_nnz: int | None = None,
           topology: str | None = None) -> BsrMatrix:
    """``z := alpha * x @ y + beta * z``; returns ``z`` (sparse.py:4200-4640).

    The compact topology lists every product block, except that a row whose
    product count exceeds the column span it reaches takes that whole span
    (the reference's rule for deep products, so results match its topology).
    ``"masked"`` keeps the topology of ``z``; ``"padded"`` writes the exact
    product topology (plus ``z``'s blocks when ``beta != 0``) into the row
    capacity of ``z``, raising the row-capacity status where it does not fit.
    Each block is ``sum_k x[r, k] @ y[k, c]`` over the row's blocks of ``x`` in
    storage order.  ``tile_size > 0`` (a power of two <= 64) selects the tiled
    order: that many lanes per block each add the rank-one updates
    ``x[r, k][:, m] y[k, c][m, :]`` of every ``tile_size``-th (block, inner
    index) pair, combined by an xor butterfly; ``tile_size <= 0`` keeps the
    block-by-block order.
    ``max_new_nnz`` caps the product blocks listed (extra ones are dropped with
    a warning, as the reference does).  Operands may alias ``z``."""
    x, x_scale = _extract_matrix_and_scale(x)
    y, y_scale = _extract_matrix_and_scale(y)
    alpha = float(alpha) * x_scale * y_scale
    beta = float(beta)
    if topology is None:
        topology = "masked" if masked else "compact"
    elif topology not in ("compact", "masked", "padded"):
        raise ValueError(f"Unsupported topology policy: {topology}")
    elif masked and topology != "masked":
        raise ValueError("Cannot pass masked=True with a non-masked topology policy")
    if masked:
        _warn_masked_arg_deprecated("bsr_mm")
    if topology != "compact" and reuse_topology:
        raise ValueError("reuse_topology is only supported with topology='compact'")
    tile = int(tile_size) if tile_size > 0 else 0
    if tile and (tile & (tile - 1) or tile > 64):
        raise ValueError(f"tile_size must be a power of two <= 64, got {tile_size}")
    if z is None:
        if topology != "compact":
            raise ValueError("Left-hand-side 'z' matrix must be provided for this topology policy")
        zshape = (x.block_shape[0], y.block_shape[1])
        ztype = x.scalar_type if zshape == (1, 1) else wt.matrix(shape=zshape, dtype=x.scalar_type)
        z = bsr_zeros(x.nrow, y.ncol, block_type=ztype, device=x.values.device)
        beta = 0.0
    z._bump_topology()
    if x.values.device != y.values.device or x.values.device != z.values.device:
        raise ValueError(
            f"All arguments must reside on the same device, got {x.values.device}, {y.values.device} and {z.values.device}"
        )
    if x.scalar_type != y.scalar_type or x.scalar_type != z.scalar_type:
        raise ValueError(
            f"Matrices must have the same scalar type, got {x.scalar_type}, {y.scalar_type} and {z.scalar_type}"
        )
    if x.block_shape[0] != z.block_shape[0] or y.block_shape[1] != z.block_shape[1] or x.block_shape[1] != y.block_shape[0]:
        raise ValueError(
            f"Incompatible block sizes for matrix multiplication, got ({x.block_shape}, {y.block_shape}) and ({z.block_shape})"
        )
    if x.nrow != z.nrow or z.ncol != y.ncol or x.ncol != y.nrow:
        raise ValueError(
            f"Incompatible number of rows/columns for matrix multiplication, got ({x.nrow}, {x.ncol}) and ({y.nrow}, {y.ncol})"
        )
    if z.scalar_type.torch_dtype not in (torch.float32, torch.float64):
        raise TypeError("bsr_mm supports float32 and float64 matrices")
    _require_gpu(z.device, "bsr_mm")
    device = z.values.device
    stream = _native.stream_of(device)

    if alpha == 0.0 or x.nnz == 0 or y.nnz == 0:
        if topology == "padded":
            z._ensure_status()
        if topology == "masked" and beta == 0.0:
            z.values.zero_()
            return z
        return bsr_scale(z, beta)

    if topology == "masked":
        if z is x or z is y:
            raise ValueError("topology='masked' is not supported for aliased inputs")
        if beta == 0.0:
            z.values.zero_()
        _mm_values(x, y, z, alpha, _positions(z), beta if beta != 0.0 else 1.0, tile)
        return z

    # operands aliasing the output are read from copies
    if x is z:
        x = bsr_copy(x, topology="padded" if x.row_counts is not None else "compact")
    if y is z:
        y = bsr_copy(y, topology="padded" if y.row_counts is not None else "compact")

    if topology == "compact" and max_new_nnz is None:
        z.nnz_sync()
    work = work_arrays if work_arrays is not None else bsr_mm_work_arrays()
    padded = topology == "padded"
    keep_z = beta != 0.0
    zT = _positions(z) if keep_z else 0
    xT = _positions(x)
    lib = _native.lib()
    row_min = work._get("row_min", x.nrow + 1, torch.int32, device)
    mm_offsets = work._get("mm_offsets", xT + 1, torch.int32, device)
    ws_bytes = lib.warp_amd_bsr_mm_count_workspace_size(xT)
    ws = work._get("count_ws", ws_bytes, torch.uint8, device)
    total_buf = torch.zeros(1, dtype=torch.int32, pin_memory=True)
    _native.call(
        "warp_amd_bsr_mm_count", xT, x.nrow, y.nrow, y.ncol, 0 if padded else 1, _native.ptr(x.offsets),
        _native.ptr(x.row_counts), _native.ptr(x.columns), _native.ptr(y.offsets), _native.ptr(y.row_counts),
        _native.ptr(y.columns), _native.ptr(row_min), _native.ptr(mm_offsets),
        total_buf.data_ptr() if max_new_nnz is None else None, _native.ptr(ws), ws_bytes, stream,
    )
    if max_new_nnz is None:
        torch.cuda.current_stream(device).synchronize()
        n_new = int(total_buf[0])
        if n_new == 0:
            return bsr_scale(z, beta)
    else:
        n_new = int(max_new_nnz)
        if not torch.cuda.is_current_stream_capturing():
            total = int(mm_offsets[xT : xT + 1].cpu()[0])
            if total > n_new:
                warnings.warn(
                    f"Number of potential `bsr_mm` blocks ({total}) exceeded `max_new_nnz` ({n_new})",
                    RuntimeWarning, stacklevel=2,
                )
    n = zT + n_new
    rows = work._get("rows", n, torch.int32, device)
    cols = work._get("cols", n, torch.int32, device)
    _native.call(
        "warp_amd_bsr_mm_list", n_new, zT, xT, x.nrow, _native.ptr(x.offsets), _native.ptr(x.row_counts),
        _native.ptr(x.columns), _native.ptr(y.offsets), _native.ptr(y.columns), _native.ptr(row_min),
        _native.ptr(mm_offsets), _native.ptr(rows), _native.ptr(cols), stream,
    )
    if not padded and keep_z and max_new_nnz is None and _union_is_own_topology(z, rows[zT:], cols[zT:], n_new):
        # every product block is already in z: the compact rebuild would give
        # z's own topology back; same values (0 + beta * old, then the sums)
        _mm_values(x, y, z, alpha, zT, beta, tile)
        return z
    old_values = None
    if zT > 0:
        z.uncompress_rows(out=rows[:zT])
        cols[:zT].copy_(z.columns[:zT])
        old_values = work._get("old_z_values", zT * z.block_size, z.values.dtype, device)
        old_values.copy_(z.values[:zT].reshape(-1))
    if padded:
        status = z._ensure_status()
        _bsr_ensure_independent_row_counts(z)
        tws = lib.warp_amd_bsr_topology_workspace_size(z.nrow, n, 0)
        tw = work._get("topology_ws", tws, torch.uint8, device)
        _native.call(
            "warp_amd_bsr_set_from_triplets_padded", z.block_size, 0, z.nrow, z.ncol, n, None, _native.ptr(rows),
            _native.ptr(cols), None, 0, _native.ptr(z.offsets), _native.ptr(z.row_counts), _native.ptr(z.columns),
            None, 0, _native.ptr(status), _native.ptr(tw), tws, stream,
        )
        zpos = _positions(z)
    else:
        _topology_from_triplets(z, rows, cols, n)
        zpos = z.nnz_sync() if max_new_nnz is None else n
        _bsr_ensure_fits(z, nnz=zpos)
    z.values.zero_()
    if zT > 0:
        _add_blocks(z, rows, cols, old_values, zT, beta)
    _mm_values(x, y, z, alpha, zpos, 1.0, tile)
    return z


def _mm_values(x: BsrMatrix, y: BsrMatrix, z: BsrMatrix, alpha: float, T: int, beta: float = 1.0, tile: int = 0):
    """z = (0 + beta * z) + alpha * (x @ y) over z's first T positions (tile
    > 0: the tiled summation order with that many lanes per block)."""
    br, bk = x.block_shape
    bc = y.block_shape[1]
    head = (("warp_amd_bsr_mm_values_tiled", z.scalar_type.size, br, bk, bc, float(alpha), float(beta), T, tile)
            if tile else ("warp_amd_bsr_mm_values", z.scalar_type.size, br, bk, bc, float(alpha), float(beta), T))
    _native.call(
        *head, z.nrow,
        _native.ptr(z.offsets),
        _native.ptr(z.row_counts), _native.ptr(z.columns), _native.ptr(z.values), y.nrow, _native.ptr(x.offsets),
        _native.ptr(x.row_counts), _native.ptr(x.columns), _native.ptr(x.values), _native.ptr(y.offsets),
        _native.ptr(y.row_counts), _native.ptr(y.columns), _native.ptr(y.values), _native.stream_of(z.device),
    )


def bsr_set_diag(A: BsrMatrix, diag=None, rows_of_blocks: int | None = None, cols_of_blocks: int | None = None):
    """Make ``A`` block diagonal (sparse.py:2939-2995).  ``diag``: a tensor of
    blocks (one per diagonal block), one block value for all (a number or
    nested list, not a tensor), or None (values left as they are)."""
    A._bump_topology()
    if rows_of_blocks is None and cols_of_blocks is not None:
        rows_of_blocks = cols_of_blocks
    if cols_of_blocks is None and rows_of_blocks is not None:
        cols_of_blocks = rows_of_blocks
    if isinstance(diag, torch.Tensor) and rows_of_blocks is None:
        rows_of_blocks = cols_of_blocks = diag.shape[0]
    if rows_of_blocks is not None:
        _bsr_resize(A, rows_of_blocks, cols_of_blocks)
    nnz = min(A.nrow, A.ncol)
    _bsr_ensure_fits(A, nnz=nnz)
    _require_gpu(A.device, "bsr_set_diag")
    _native.call("warp_amd_bsr_diag_topology", A.nrow, nnz, _native.ptr(A.offsets), _native.ptr(A.columns),
                 _native.stream_of(A.device))
    A.row_counts = None
    A._nnz_pending = False
    A.nnz = nnz
    if isinstance(diag, torch.Tensor):
        A.values[:nnz].copy_(diag[:nnz].reshape(A.values[:nnz].shape))
    elif diag is not None:
        A.values[:nnz] = torch.as_tensor(diag, dtype=A.values.dtype).to(A.values.device)


def bsr_diag(diag=None, rows_of_blocks: int | None = None, cols_of_blocks: int | None = None, block_type=None,
             device=None) -> BsrMatrix:
    """New block-diagonal matrix (sparse.py:2998-3054)."""
    if rows_of_blocks is None and cols_of_blocks is not None:
        rows_of_blocks = cols_of_blocks
    if cols_of_blocks is None and rows_of_blocks is not None:
        cols_of_blocks = rows_of_blocks
    if isinstance(diag, torch.Tensor):
        if rows_of_blocks is None:
            rows_of_blocks = cols_of_blocks = diag.shape[0]
        device = diag.device
        if block_type is None:
            block_type = (wt.type_to_scalar(diag.dtype) if diag.ndim == 1
                          else wt.matrix(shape=tuple(diag.shape[1:]), dtype=wt.type_to_scalar(diag.dtype)))
    elif rows_of_blocks is None:
        raise ValueError(
            "rows_of_blocks and/or cols_of_blocks must be provided for constructing a diagonal matrix with uniform diagonal"
        )
    if block_type is None:
        if diag is None:
            raise ValueError("Either `diag` or `block_type` needs to be provided")
        d = torch.as_tensor(diag)
        if d.ndim == 1:
            raise ValueError(f"BsrMatrix block type must be either warp matrix or scalar; got a vector of length {d.shape[0]}")
        block_type = (wt.type_to_scalar(d.dtype) if d.ndim == 0
                      else wt.matrix(shape=tuple(d.shape), dtype=wt.type_to_scalar(d.dtype)))
    A = bsr_zeros(rows_of_blocks, cols_of_blocks, block_type=block_type, device=device)
    bsr_set_diag(A, diag)
    return A


def bsr_set_identity(A: BsrMatrix, rows_of_blocks: int | None = None) -> None:
    """Make ``A`` the identity (sparse.py:3057-3071)."""
    identity = 1.0 if A.block_shape == (1, 1) else torch.eye(A.block_shape[0]).tolist()
    bsr_set_diag(A, diag=identity, rows_of_blocks=rows_of_blocks, cols_of_blocks=rows_of_blocks)


def bsr_identity(rows_of_blocks: int, block_type, device=None) -> BsrMatrix:
    """New square identity matrix (sparse.py:3074-3093)."""
    A = bsr_zeros(rows_of_blocks, rows_of_blocks, block_type=block_type, device=device)
    bsr_set_identity(A)
    return A


def bsr_matrix_t(dtype):
    """Matrix type for blocks of ``dtype`` (sparse.py:493-520): a scalar or a
    matrix type; every block type shares the one :class:`BsrMatrix` class here."""
    if not type_is_matrix(dtype):
        try:
            wt.type_to_scalar(dtype)
        except TypeError:
            raise ValueError(f"BsrMatrix block type must be either warp matrix or scalar; got {dtype!r}") from None
    return BsrMatrix


def bsr_row_index(offsets: torch.Tensor, row_count: int, block_index, row_counts: torch.Tensor | None = None):
    """Row holding storage position ``block_index`` (a number or an integer
    tensor), -1 when the position holds no (active) block (sparse.py:1794-1828).
    Evaluated on tensors, not inside kernels."""
    idx = torch.as_tensor(block_index, device=offsets.device).to(torch.int64)
    off = offsets[: row_count + 1].to(torch.int64)
    rows = torch.searchsorted(off, idx, right=True) - 1
    rows = torch.where((idx >= 0) & (idx < off[row_count]), rows, torch.full_like(rows, -1))
    if row_counts is not None:
        safe = rows.clamp(min=0)
        beyond = idx >= off[safe] + row_counts.to(torch.int64)[safe]
        rows = torch.where((rows >= 0) & beyond, torch.full_like(rows, -1), rows)
    rows = rows.to(torch.int32)
    return int(rows) if rows.ndim == 0 else rows


def bsr_block_index(row, col, bsr_offsets: torch.Tensor, bsr_columns: torch.Tensor,
                    bsr_row_counts: torch.Tensor | None = None):
    """Storage position of block (``row``, ``col``), -1 when absent
    (sparse.py:1852-1900); rows must hold sorted columns.  ``row`` / ``col``
    are numbers or integer tensors of one shape.  Evaluated on tensors."""
    dev = bsr_offsets.device
    r = torch.as_tensor(row, device=dev).to(torch.int64)
    c = torch.as_tensor(col, device=dev).to(torch.int64)
    nrow = bsr_offsets.numel() - 1
    rs = r.clamp(0, max(nrow - 1, 0))
    beg = bsr_offsets.to(torch.int64)[rs]
    if bsr_row_counts is None:
        end = bsr_offsets.to(torch.int64)[rs + 1]
    else:
        end = beg + bsr_row_counts.to(torch.int64)[rs]
    # lower bound of c in columns[beg:end) by bisection, elementwise
    cols = bsr_columns.to(torch.int64)
    if cols.numel() == 0:
        out = torch.full_like(r, -1).to(torch.int32)
        return int(out) if out.ndim == 0 else out
    lo, hi = beg.clone(), end.clone()
    for _ in range(max(1, int(cols.numel()).bit_length())):
        active = lo < hi
        mid = torch.where(active, (lo + hi) // 2, lo)
        less = active & (cols[mid.clamp(max=cols.numel() - 1)] < c)
        lo = torch.where(less, mid + 1, lo)
        hi = torch.where(active & ~less, mid, hi)
    found = (r >= 0) & (r < nrow) & (lo < end) & (cols[lo.clamp(max=cols.numel() - 1)] == c)
    out = torch.where(found, lo, torch.full_like(lo, -1)).to(torch.int32)
    return int(out) if out.ndim == 0 else out


def _mirror_symmetric(A: BsrMatrix) -> bool:
    """True when ``A`` is compact and square with 1x1 or 3x3 blocks and every
    block (r, c) has a mirror (c, r): the transposed product can then gather
    without atomics.  The verdict is cached per topology version (see
    ``BsrMatrix.__setattr__``); computing it reads one count back, so under a
    graph capture only a cached verdict can say True."""
    br, bc = A.block_shape
    if A.row_counts is not None or A.nrow != A.ncol or br != bc or br not in (1, 3):
        return False
    key = (A.__dict__.get("_topology_version", 0), A.offsets.data_ptr(), A.columns.data_ptr())
    cached = A.__dict__.get("_mirror_cache")
    if cached is not None and cached[0] == key:
        return cached[1]
    if torch.cuda.is_current_stream_capturing():
        return False
    T = min(A.nnz_sync(), A.columns.numel())  # exact: stale tail slots hold no valid column
    verdict = False
    if T > 0:
        rows = A.uncompress_rows()
        missing = torch.empty(1, dtype=torch.int32, device=A.device)
        _native.call("warp_amd_bsr_count_missing_blocks", T, _native.ptr(A.columns), _native.ptr(rows),
                     _native.ptr(A.offsets), None, _native.ptr(A.columns), _native.ptr(missing),
                     _native.stream_of(A.device))
        verdict = int(missing.item()) == 0
    A.__dict__["_mirror_cache"] = (key, verdict)
    return verdict


def _vec_view(array: torch.Tensor, torch_dtype: torch.dtype, expected_scalar_count: int) -> torch.Tensor:
    """Flat scalar view of a vector array (sparse.py:4750-4789)."""
    if array.numel() != expected_scalar_count:
        raise ValueError(f"Invalid array scalar size, expected {expected_scalar_count}, got {array.numel()}")
    if array.dtype != torch_dtype:
        raise ValueError(f"Incompatible scalar types, expected {torch_dtype}, got {array.dtype}")
    if array.ndim > 2:
        raise ValueError(f"Incompatible array number of dimensions, expected 1 or 2, got {array.ndim}")
    if not array.is_contiguous():
        raise ValueError("Array must be contiguous")
    return array


class _MvPlan:
    """A matrix's constant arguments of bsr_mv, packed once (host struct of
    warp_amd_bsr_mv_packed), with what its x and y must match.  Valid for one
    topology version, one ``values`` tensor and one nnz.  Two copies of the
    struct differ in ``sweep`` only and are used in turn: consecutive products
    with one matrix traverse its rows in opposite directions (same bits), so
    each starts with the rows the previous one read last, still in the
    memory-side cache."""

    __slots__ = ("version", "values", "nnz", "dtype", "nx", "ny", "device", "dev_index", "args", "refs", "turn",
                 "fn", "stream")

    def __init__(self, A: "BsrMatrix", nnz_hint: int, torch_dtype: torch.dtype, nx: int, ny: int):
        br, bc = A.block_shape
        self.version = A.__dict__.get("_topology_version", 0)
        self.values = A.values
        self.nnz = A.nnz
        self.dtype = torch_dtype
        self.nx, self.ny = nx, ny
        self.device = A.values.device
        self.dev_index = A.values.get_device()
        a = _native.BsrMvArgs()
        a.nrow, a.ncol, a.block_rows, a.block_cols = A.nrow, A.ncol, br, bc
        a.offsets = A.offsets.data_ptr()
        a.row_counts = A.row_counts.data_ptr() if A.row_counts is not None else None
        a.columns = A.columns.data_ptr()
        a.values = A.values.data_ptr()
        a.nnz_hint = int(nnz_hint)
        a.lanes, a.rows_per_group = 0, 0
        a.scalar_size = 8 if torch_dtype == torch.float64 else 4
        a.sweep = 0
        b = _native.BsrMvArgs.from_buffer_copy(a)
        b.sweep = 1
        self.args = (a, b)
        self.refs = (ctypes.byref(a), ctypes.byref(b))
        self.turn = 0
        self.fn = _native.lib().warp_amd_bsr_mv_packed
        self.stream = _native.stream_getter(self.device)


def bsr_mv(
    A,
    x: torch.Tensor,
    y: torch.Tensor | None = None,
    alpha: float = 1.0,
    beta: float = 0.0,
    transpose: bool = False,
    work_buffer: torch.Tensor | None = None,
    tile_size: int = 0,
) -> torch.Tensor:
    """``y := alpha * A @ x + beta * y``; returns ``y`` (sparse.py:4792-4913).

    ``alpha == 0`` leaves ``x`` unread, ``beta == 0`` leaves ``y`` unread.
    ``x`` and ``y`` may alias (``work_buffer`` then holds the old ``y``).
    ``tile_size``: lanes per block row (a power of two <= 64) when positive;
    negative selects one lane per row (sums in storage order, like the
    reference row kernel); 0 chooses from the mean row length.
    """
    if type(A) is BsrMatrix and y is not None and not transpose and work_buffer is None and tile_size == 0:
        # repeated products with one matrix: its arguments are packed once
        # (_MvPlan), and x / y get the same checks in a cheaper form
        d = A.__dict__
        plan = d.get("_mv_plan")
        if (
            plan is not None
            and plan.version == d.get("_topology_version", 0)
            and plan.values is d.get("values")
            and plan.nnz == d.get("nnz")
            and not d.get("_nnz_pending", False)
            and x.dtype is plan.dtype
            and y.dtype is plan.dtype
            and x.numel() == plan.nx
            and y.numel() == plan.ny
            and x.ndim <= 2
            and y.ndim <= 2
            and x.is_contiguous()
            and y.is_contiguous()
            and x.get_device() == plan.dev_index
            and y.get_device() == plan.dev_index
        ):
            xp, yp = x.data_ptr(), y.data_ptr()
            if xp != yp:
                k = plan.turn
                plan.turn = k ^ 1
                st = plan.fn(plan.refs[k], xp, yp, float(alpha), float(beta), plan.stream())
                if st:
                    _native.check(st, "warp_amd_bsr_mv_packed")
                return y
    A, A_scale = _extract_matrix_and_scale(A)
    alpha = float(alpha) * A_scale

    # shapes of the operator applied: A, or A^T when transposed
    if transpose:
        bc, br = A.block_shape
        nrow, ncol = A.ncol, A.nrow
    else:
        br, bc = A.block_shape
        nrow, ncol = A.nrow, A.ncol
    torch_dtype = A.scalar_type.torch_dtype

    if y is None:
        shape = (nrow,) if br == 1 else (nrow, br)
        y = torch.empty(shape, dtype=torch_dtype, device=A.values.device)
        beta = 0.0
    beta = float(beta)

    if A.values.device != x.device or A.values.device != y.device:
        raise ValueError(
            f"A, x, and y must reside on the same device, got {A.values.device}, {x.device} and {y.device}"
        )

    if x.numel() > 0 and x.data_ptr() == y.data_ptr():
        if work_buffer is None:
            work_buffer = torch.empty_like(y)
        elif work_buffer.numel() < y.numel():
            raise ValueError(
                f"Work buffer size is insufficient, needs to be at least {y.numel()}, got {work_buffer.numel()}"
            )
        elif work_buffer.dtype != y.dtype:
            raise ValueError(f"Work buffer must have same data type as y, {y.dtype} vs {work_buffer.dtype}")
        work_buffer.view(-1)[: y.numel()].copy_(y.view(-1))
        x = work_buffer.view(-1)[: y.numel()]

    try:
        x_view = _vec_view(x, torch_dtype, ncol * bc)
    except ValueError as err:
        raise ValueError("Incompatible 'x' vector for bsr_mv") from err
    try:
        y_view = _vec_view(y, torch_dtype, nrow * br)
    except ValueError as err:
        raise ValueError("Incompatible 'y' vector for bsr_mv") from err

    _require_gpu(A.values.device, "bsr_mv")
    if torch_dtype not in (torch.float32, torch.float64):
        raise TypeError(f"bsr_mv supports float32 and float64 matrices, got {torch_dtype}")
    sfx = "f64" if torch_dtype == torch.float64 else "f32"
    nnz_hint = A._lane_nnz()
    if transpose:
        if _mirror_symmetric(A):
            # every block has its mirror: gather, deterministic
            _native.call(
                "warp_amd_bsr_mv_transpose_gather_" + sfx, A.nrow, A.block_shape[0], _native.ptr(A.offsets),
                _native.ptr(A.columns), _native.ptr(A.values), nnz_hint, _native.ptr(x_view), _native.ptr(y_view),
                alpha, beta, _native.stream_of(A.values.device),
            )
            return y
        # atomic scatter, as the reference: the last bits of y vary between runs
        _native.call(
            "warp_amd_bsr_mv_transpose_" + sfx, A.nrow, A.ncol, A.block_shape[0], A.block_shape[1],
            _native.ptr(A.offsets), _native.ptr(A.row_counts), _native.ptr(A.columns), _native.ptr(A.values),
            nnz_hint, _native.ptr(x_view), _native.ptr(y_view), alpha, beta, _native.stream_of(A.values.device),
        )
        return y
    if nnz_hint == 0 and nrow > 0:
        # no blocks (the matrix arrays may be empty): y = beta * y, y unread when beta is 0
        if beta == 0.0:
            y_view.zero_()
        elif beta != 1.0:
            _native.call("warp_amd_scale_values", y_view.element_size(), y_view.numel(), _native.ptr(y_view), beta,
                         _native.stream_of(y_view.device))
        return y
    fn = "warp_amd_bsr_mv_ex_" + sfx
    if tile_size == 0 and type(A) is BsrMatrix and not A.__dict__.get("_nnz_pending", False):
        plan = _MvPlan(A, nnz_hint, torch_dtype, x_view.numel(), y_view.numel())
        A.__dict__["_mv_plan"] = plan
        k = plan.turn
        plan.turn = k ^ 1
        _native.check(
            plan.fn(plan.refs[k], x_view.data_ptr(), y_view.data_ptr(), alpha, beta, _native.stream_of(plan.device)),
            "warp_amd_bsr_mv_packed",
        )
        return y
    # tile_size > 0: lanes per block row; < 0: one lane per row (row-sequential
    # sums, the reference row kernel's order); 0: chosen from the mean row length
    if tile_size > 0:
        if tile_size & (tile_size - 1) or tile_size > 64:
            raise ValueError(f"tile_size must be a power of two <= 64, got {tile_size}")
        lanes = int(tile_size)
    else:
        lanes = 1 if tile_size < 0 else 0
    _native.call(
        fn,
        nrow,
        ncol,
        br,
        bc,
        _native.ptr(A.offsets),
        _native.ptr(A.row_counts),
        _native.ptr(A.columns),
        _native.ptr(A.values),
        nnz_hint,
        _native.ptr(x_view),
        _native.ptr(y_view),
        alpha,
        beta,
        lanes,
        0,
        _native.stream_of(A.values.device),
    )
    return y


def bsr_get_diag(A, out: torch.Tensor | None = None) -> torch.Tensor:
    """Diagonal blocks of ``A`` (zero where absent) (sparse.py:2887-2924)."""
    A, scale = _extract_matrix_and_scale(A)
    dim = min(A.nrow, A.ncol)
    br, bc = A.block_shape
    shape = (dim,) if A.block_size == 1 else (dim, br, bc)
    torch_dtype = A.scalar_type.torch_dtype
    if out is None:
        out = torch.zeros(shape, dtype=torch_dtype, device=A.values.device)
    else:
        if out.dtype != torch_dtype or tuple(out.shape[1:]) != tuple(shape[1:]):
            raise ValueError(f"Output array must have type {A.dtype}, got {out.dtype}")
        if out.device != A.values.device:
            raise ValueError(f"Output array must reside on device {A.values.device}, got {out.device}")
        if out.shape[0] < dim:
            raise ValueError(f"Output array must be of length at least {dim}, got {out.shape[0]}")
    _require_gpu(A.values.device, "bsr_get_diag")
    fn = "warp_amd_bsr_get_diag_f64" if torch_dtype == torch.float64 else "warp_amd_bsr_get_diag_f32"
    _native.call(
        fn,
        A.nrow,
        A.ncol,
        br,
        bc,
        _native.ptr(A.offsets),
        _native.ptr(A.row_counts),
        _native.ptr(A.columns),
        _native.ptr(A.values),
        float(scale),
        _native.ptr(out),
        _native.stream_of(A.values.device),
    )
    return out
