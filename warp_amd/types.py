"""Scalar, vector and matrix block types of the sparse API.

Mirrors the subset of ``warp.types`` that the sparse path names
(warp/_src/types.py: ``float32``/``float64``, ``vector``, ``matrix``): a BSR
matrix's ``block_type`` is a scalar type (CSR) or a matrix type with a
``_shape_``.  Arrays are torch tensors: an array of ``n`` blocks of shape
``(br, bc)`` is a tensor of shape ``(n, br, bc)``, an array of ``n`` vectors of
length ``k`` a tensor of shape ``(n, k)`` -- the same row-major memory layout
as the reference's ``wp.array`` of ``mat``/``vec`` (sparse.py:815-823).
"""

from __future__ import annotations

import functools

import torch


class ScalarType:
    def __init__(self, name: str, torch_dtype: torch.dtype, size: int):
        self.__name__ = name
        self.torch_dtype = torch_dtype
        self.size = size
        self._shape_ = None

    def __repr__(self) -> str:
        return self.__name__

    def __call__(self, v):
        return float(v) if self.torch_dtype.is_floating_point else int(v)


float16 = ScalarType("float16", torch.float16, 2)
float32 = ScalarType("float32", torch.float32, 4)
float64 = ScalarType("float64", torch.float64, 8)
int32 = ScalarType("int32", torch.int32, 4)

scalar_types = (float16, float32, float64, int32)

_TORCH_TO_SCALAR = {t.torch_dtype: t for t in scalar_types}


class _Composite:
    """Base of vector / matrix types."""

    _shape_: tuple[int, ...]
    _wp_scalar_type_: ScalarType
    _length_: int


@functools.cache
def matrix(shape: tuple[int, int], dtype=float32) -> type:
    """Matrix block type of the given shape and scalar type (warp.types.matrix)."""
    scalar = type_to_scalar(dtype)
    shape = (int(shape[0]), int(shape[1]))
    return type(
        f"mat{shape[0]}{shape[1]}{'d' if scalar is float64 else ''}",
        (_Composite,),
        {"_shape_": shape, "_wp_scalar_type_": scalar, "_length_": shape[0] * shape[1]},
    )


@functools.cache
def vector(length: int, dtype=float32) -> type:
    """Vector type of the given length and scalar type (warp.types.vector)."""
    scalar = type_to_scalar(dtype)
    return type(
        f"vec{length}{'d' if scalar is float64 else ''}",
        (_Composite,),
        {"_shape_": (int(length),), "_wp_scalar_type_": scalar, "_length_": int(length)},
    )


def type_to_scalar(dtype) -> ScalarType:
    """Map Python / torch / package scalar types to a ScalarType."""
    if isinstance(dtype, ScalarType):
        return dtype
    if dtype is float:
        return float32
    if dtype is int:
        return int32
    if isinstance(dtype, torch.dtype) and dtype in _TORCH_TO_SCALAR:
        return _TORCH_TO_SCALAR[dtype]
    raise TypeError(f"Unsupported scalar type {dtype!r}")


def type_is_matrix(t) -> bool:
    return isinstance(t, type) and issubclass(t, _Composite) and len(t._shape_) == 2


def type_is_vector(t) -> bool:
    return isinstance(t, type) and issubclass(t, _Composite) and len(t._shape_) == 1


def type_scalar_type(t) -> ScalarType:
    if isinstance(t, type) and issubclass(t, _Composite):
        return t._wp_scalar_type_
    return type_to_scalar(t)


def type_size(t) -> int:
    """Number of scalars in one element of type t."""
    if isinstance(t, type) and issubclass(t, _Composite):
        return t._length_
    return 1


def block_shape_of(t) -> tuple[int, int]:
    return t._shape_ if type_is_matrix(t) else (1, 1)


mat22 = matrix((2, 2), float32)
mat33 = matrix((3, 3), float32)
mat44 = matrix((4, 4), float32)
mat33d = matrix((3, 3), float64)
vec2 = vector(2, float32)
vec3 = vector(3, float32)
vec4 = vector(4, float32)
vec3d = vector(3, float64)
