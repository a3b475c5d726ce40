"""warp_amd: the sparse linear-algebra path of Warp, rebuilt for AMD MI355X (gfx950).

``warp_amd.sparse`` mirrors ``warp.sparse`` (BSR assembly from triplets,
``bsr_mv``, ``bsr_get_diag``) and ``warp_amd.optim.linear`` mirrors the CG
part of ``warp.optim.linear``.  Arrays are torch tensors on a ROCm device;
the compute runs in the hand-written HIP library ``_lib/libwarp_amd.so``.
"""

from . import types
from .types import float32, float64, int32, mat33, mat33d, matrix, vec3, vec3d, vector

__version__ = "0.1.0"


def library_path() -> str:
    from . import _native

    return _native.LIB_PATH


from . import sparse  # noqa: E402
from .optim import linear  # noqa: E402
