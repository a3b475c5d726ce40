"""warp_amd: the sparse linear-algebra path of Warp, rebuilt for AMD MI355X (gfx950).

``warp_amd.sparse`` mirrors the public names of ``warp.sparse``: BSR assembly
from triplets (compact, padded and masked topologies), ``bsr_compress``,
transposes, ``bsr_assign`` / ``bsr_copy``, ``bsr_axpy``, ``bsr_mm``, ``bsr_mv``,
diagonal helpers and the row / block index lookups.  ``warp_amd.optim.linear``
mirrors the CG part of ``warp.optim.linear``; ``warp_amd.fem`` holds the
Dirichlet projection of ``warp.fem``.  Arrays are torch tensors on a ROCm
device; the compute runs in the hand-written HIP library ``_lib/libwarp_amd.so``.
"""

from . import types
from .types import float32, float64, int32, mat22, mat33, mat33d, mat44, matrix, vec2, vec3, vec3d, vec4, vector

__version__ = "0.1.0"


def library_path() -> str:
    from . import _native

    return _native.LIB_PATH


from . import sparse  # noqa: E402
from .optim import linear  # noqa: E402
