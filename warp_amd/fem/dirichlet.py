"""Dirichlet boundary conditions on an assembled BSR system (reference
warp/_src/fem/dirichlet.py; array_axpy from warp/_src/fem/linalg.py:425-461).

The projector ``P`` is block diagonal; after normalization each diagonal block
is an orthogonal projector onto the constrained directions of its node.  The
projected system is ``A' = (I - P) A (I - P) + P`` and
``b' = (I - P) (b - A P v) + P v``, built from the sparse kernels of
:mod:`warp_amd.sparse` (products, sums and assignments on the GPU).
"""

from __future__ import annotations

import torch

from .. import _native
from ..sparse import BsrMatrix, _require_gpu, bsr_assign, bsr_axpy, bsr_copy, bsr_mm, bsr_mv


def array_axpy(x: torch.Tensor, y: torch.Tensor, alpha: float = 1.0, beta: float = 1.0) -> None:
    """``y = alpha * x + beta * y`` (linalg.py:425-461)."""
    if x.shape != y.shape or x.device != y.device:
        raise ValueError("x and y arrays must have the same shape and device")
    if x.dtype != y.dtype or y.dtype not in (torch.float32, torch.float64):
        raise TypeError("array_axpy supports float32 and float64 arrays of one type")
    if not (x.is_contiguous() and y.is_contiguous()):
        raise ValueError("array_axpy needs contiguous arrays")
    _require_gpu(y.device, "array_axpy")
    _native.call(
        "warp_amd_axpby", y.element_size(), y.numel(), float(alpha), _native.ptr(x), float(beta), _native.ptr(y),
        _native.stream_of(y.device),
    )


def normalize_dirichlet_projector(projector_matrix: BsrMatrix, fixed_value: torch.Tensor | None = None) -> None:
    """Make every diagonal block of the projector an orthogonal projector (dirichlet.py:12-66).

    Eigenpairs with eigenvalue above ``1e-6 * trace`` are kept; ``fixed_value``
    (one value or vector per block row), when given, is rescaled the same way
    (``sum (v . b) / lambda v``)."""
    if projector_matrix.ncol != projector_matrix.nrow:
        raise ValueError("Projector must be a square block-diagonal matrix")
    br, bc = projector_matrix.block_shape
    if fixed_value is not None:
        if fixed_value.shape[0] != projector_matrix.nrow:
            raise ValueError("Fixed value array must be of length equal to the number of rows of blocks")
        if fixed_value.dtype != projector_matrix.values.dtype or not fixed_value.is_contiguous():
            raise ValueError("Fixed value array must be contiguous and of the projector's scalar type")
    _require_gpu(projector_matrix.device, "normalize_dirichlet_projector")
    _native.call(
        "warp_amd_normalize_dirichlet_projector", projector_matrix.scalar_type.size, br, projector_matrix.nrow,
        _native.ptr(projector_matrix.offsets), _native.ptr(projector_matrix.row_counts),
        _native.ptr(projector_matrix.columns), _native.ptr(projector_matrix.values), _native.ptr(fixed_value),
        _native.stream_of(projector_matrix.device),
    )


def project_system_rhs(system_matrix: BsrMatrix, system_rhs: torch.Tensor, projector_matrix: BsrMatrix,
                       fixed_value: torch.Tensor | None = None) -> None:
    """``rhs = (I - P) (rhs - A P v) + P v`` in place (dirichlet.py:69-90)."""
    rhs_tmp = system_rhs.clone()
    if fixed_value is None:
        system_rhs.zero_()
    else:
        bsr_mv(A=projector_matrix, x=fixed_value, y=system_rhs, alpha=1.0, beta=0.0)
    bsr_mv(A=system_matrix, x=system_rhs, y=rhs_tmp, alpha=-1.0, beta=1.0)
    # rhs_tmp = rhs - A P v, system_rhs = P v
    array_axpy(x=rhs_tmp, y=system_rhs, alpha=1.0, beta=1.0)
    bsr_mv(A=projector_matrix, x=rhs_tmp, y=system_rhs, alpha=-1.0, beta=1.0)


def project_system_matrix(system_matrix: BsrMatrix, projector_matrix: BsrMatrix) -> None:
    """``A = (I - P) A (I - P) + P`` in place (dirichlet.py:93-104)."""
    complement_system = bsr_copy(system_matrix)
    bsr_mm(x=projector_matrix, y=system_matrix, z=complement_system, alpha=-1.0, beta=1.0)
    bsr_assign(dest=system_matrix, src=complement_system)
    bsr_axpy(x=projector_matrix, y=system_matrix)
    bsr_mm(x=complement_system, y=projector_matrix, z=system_matrix, alpha=-1.0, beta=1.0)


def project_linear_system(system_matrix: BsrMatrix, system_rhs: torch.Tensor, projector_matrix: BsrMatrix,
                          fixed_value: torch.Tensor | None = None, normalize_projector: bool = True) -> None:
    """Project both sides of the system (dirichlet.py:107-124); with
    ``normalize_projector`` the projector (and fixed value) are normalized first."""
    if normalize_projector:
        normalize_dirichlet_projector(projector_matrix, fixed_value)
    project_system_rhs(system_matrix, system_rhs, projector_matrix, fixed_value)
    project_system_matrix(system_matrix, projector_matrix)
