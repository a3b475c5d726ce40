"""The part of ``warp.fem`` between assembly and solve: Dirichlet projection
of an assembled linear system (reference warp/fem/__init__.py:140-147 ->
warp/_src/fem/dirichlet.py)."""

from .dirichlet import (
    array_axpy,
    normalize_dirichlet_projector,
    project_linear_system,
    project_system_matrix,
    project_system_rhs,
)

__all__ = [
    "array_axpy",
    "normalize_dirichlet_projector",
    "project_linear_system",
    "project_system_matrix",
    "project_system_rhs",
]
