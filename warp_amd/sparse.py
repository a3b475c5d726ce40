"""Block-sparse row (BSR) matrices on MI355X: the ``warp.sparse`` drop-in.

Same names, argument order, defaults and validation messages as the
reference (warp/sparse.py:25-51 -> warp/_src/sparse.py); arrays are torch
tensors (see :mod:`warp_amd.types` for the layout).  All work runs in the
hand-written gfx950 library through ctypes; there is no CPU fallback.

Covered here: ``BsrMatrix`` (compact topology), ``bsr_zeros``,
``bsr_set_zero``, ``bsr_set_from_triplets``, ``bsr_from_triplets``,
``bsr_mv`` and ``bsr_get_diag``.
"""

from __future__ import annotations

import torch

from . import _native
from . import types as wt
from .types import block_shape_of, type_is_matrix, type_scalar_type, type_size

BSR_STATUS_SUCCESS = 0
BSR_STATUS_ROW_CAPACITY_EXCEEDED = 1

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

    def notify_nnz_changed(self, nnz: int | None = None, nnz_capacity: int | None = None) -> None:
        """Notify that ``offsets`` changed outside this module (sparse.py:311-333)."""
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
        starts = self.offsets[: self.nrow].to(torch.int64)
        if self.row_counts is None:
            counts = self.offsets[1 : self.nrow + 1].to(torch.int64) - starts
        else:
            counts = self.row_counts[: self.nrow].to(torch.int64)
        rows = torch.repeat_interleave(torch.arange(self.nrow, device=self.device, dtype=torch.int32), counts)
        # position of each active block: its row start plus its rank within the row
        first = torch.repeat_interleave(torch.cumsum(counts, 0) - counts, counts)
        pos = torch.repeat_interleave(starts, counts) + torch.arange(rows.numel(), device=self.device) - first
        if out is None:
            out = torch.full((self.nnz,), -1, dtype=torch.int32, device=self.device)
        else:
            out[: self.nnz].fill_(-1)
        keep = pos < self.nnz
        out[pos[keep]] = rows[keep]
        return out

    # -- operators -------------------------------------------------------
    def __matmul__(self, x):
        if isinstance(x, torch.Tensor):
            return bsr_mv(self, x)
        return NotImplemented

    def __rmatmul__(self, x):
        """x @ A = A^T x (reference BsrMatrix.__rmatmul__)."""
        if isinstance(x, torch.Tensor):
            return bsr_mv(self, x, transpose=True)
        return NotImplemented

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
        return NotImplemented

    def __rmatmul__(self, x):
        if isinstance(x, torch.Tensor):
            return bsr_mv(self, x, transpose=True)
        return NotImplemented

    def __mul__(self, s):
        return _BsrScalingExpression(self.mat, self.scale * s)

    __rmul__ = __mul__

    def __neg__(self):
        return _BsrScalingExpression(self.mat, -self.scale)

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
    raise ValueError(f"Expected a BsrMatrix, got {type(A)}")


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
    and coalesced by :func:`bsr_compress`."""
    if topology is None:
        topology = "padded" if row_capacity is not None else "compact"
    elif topology not in ("compact", "padded"):
        raise ValueError(f"Unsupported topology policy: {topology}")
    if nnz_capacity is not None:
        nnz_capacity = int(nnz_capacity)
        if nnz_capacity < 0:
            raise ValueError("nnz_capacity must be nonnegative")
    if row_capacity is not None:
        if topology != "padded":
            raise ValueError("row_capacity requires topology='padded'")
        if isinstance(row_capacity, torch.Tensor):
            if row_capacity.ndim != 1 or row_capacity.shape[0] != int(rows_of_blocks):
                raise ValueError(
                    f"row_capacity array must have shape ({int(rows_of_blocks)},), got {tuple(row_capacity.shape)}"
                )
            if row_capacity.dtype != torch.int32:
                raise TypeError("row_capacity array must have dtype int32")
        else:
            row_capacity = int(row_capacity)
            if row_capacity < 0:
                raise ValueError("row_capacity must be nonnegative")

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
    if topology == "padded":
        # row r owns [offsets[r], offsets[r+1]), with row_counts[r] active blocks (none yet)
        if isinstance(row_capacity, torch.Tensor):
            bsr.offsets[1:] = torch.cumsum(row_capacity.to(dev), 0).to(torch.int32)
            nnz = int(bsr.offsets[bsr.nrow]) if nnz_capacity is None else nnz_capacity
        elif row_capacity is not None:
            bsr.offsets = torch.arange(bsr.nrow + 1, dtype=torch.int32, device=dev) * row_capacity
            nnz = bsr.nrow * row_capacity
        else:
            nnz = 0
        bsr.row_counts = torch.zeros(bsr.nrow, dtype=torch.int32, device=dev)
        capacity = max(nnz, nnz_capacity or 0)
        _bsr_ensure_fits(bsr, nnz=capacity)
        bsr.columns[:capacity].fill_(-1)
        bsr.values[:capacity].zero_()
        bsr.nnz = nnz
        return bsr
    _bsr_ensure_fits(bsr, nnz=int(nnz_capacity or 0))
    bsr.nnz = 0
    return bsr


def type_is_vector_like(t) -> bool:
    return wt.type_is_vector(t)


def bsr_set_zero(bsr: BsrMatrix, rows_of_blocks: int | None = None, cols_of_blocks: int | None = None, topology=None):
    """Remove all blocks (sparse.py:769-810, compact policy)."""
    if rows_of_blocks is not None:
        bsr.nrow = int(rows_of_blocks)
    if cols_of_blocks is not None:
        bsr.ncol = int(cols_of_blocks)
    if bsr.offsets.numel() < bsr.nrow + 1:
        bsr.offsets = torch.zeros(bsr.nrow + 1, dtype=torch.int32, device=bsr.offsets.device)
    else:
        bsr.offsets.zero_()
    bsr.row_counts = None
    bsr.nnz = 0
    bsr._nnz_pending = False


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
    dropped too.  Duplicates are summed in ascending triplet order.  After the
    call ``dest.nnz`` is the triplet count (an upper bound) until
    :meth:`BsrMatrix.nnz_sync`.
    """
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

    nnz = rows.shape[0]
    if nnz == 0:
        bsr_set_zero(dest)
        return

    if topology != "compact":
        raise NotImplementedError(f"'{topology}' topology policy is not provided by warp_amd")

    _require_gpu(dest.device, "bsr_set_from_triplets")
    _bsr_ensure_fits(dest, nnz=nnz)

    scalar_type = dest.scalar_type
    zero_mask = _zero_value_masks.get(scalar_type, 0) if prune_numerical_zeros and values is not None else 0
    if values is not None:
        values = values.contiguous()
    rows = rows.contiguous()
    columns = columns.contiguous()

    lib = _native.lib()
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
    if transpose:
        # atomic scatter, as the reference: the last bits of y vary between runs
        _native.call(
            "warp_amd_bsr_mv_transpose_" + sfx, A.nrow, A.ncol, A.block_shape[0], A.block_shape[1],
            _native.ptr(A.offsets), _native.ptr(A.row_counts), _native.ptr(A.columns), _native.ptr(A.values),
            int(A.nnz), _native.ptr(x_view), _native.ptr(y_view), alpha, beta, _native.stream_of(A.values.device),
        )
        return y
    fn = "warp_amd_bsr_mv_ex_" + sfx
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
        int(A.nnz),
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
