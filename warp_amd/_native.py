"""ctypes binding of the gfx950 HIP library (``include/warp_amd.h``).

The library is loaded after ``torch`` so that both share the single HIP
runtime already mapped by torch.  There is no fallback: if the library is
missing, every GPU entry point raises.
"""

from __future__ import annotations

import ctypes
import functools
import os
import threading

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WARP_AMD_LIB", os.path.join(_HERE, "_lib", "libwarp_amd.so"))

_lib = None
_lock = threading.Lock()

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_size = ctypes.c_size_t
c_f32 = ctypes.c_float
c_f64 = ctypes.c_double
c_ptr = ctypes.c_void_p

CG_CALLBACK = ctypes.CFUNCTYPE(c_int, c_int, c_f64, c_f64, c_ptr)


class BsrMvArgs(ctypes.Structure):
    """warp_amd_bsr_mv_args: a matrix's constant arguments of bsr_mv, packed
    once so that each product passes six arguments (include/warp_amd.h)."""

    _fields_ = [
        ("nrow", c_int), ("ncol", c_int), ("block_rows", c_int), ("block_cols", c_int),
        ("offsets", c_ptr), ("row_counts", c_ptr), ("columns", c_ptr), ("values", c_ptr),
        ("nnz_hint", c_i64),
        ("lanes", c_int), ("rows_per_group", c_int), ("scalar_size", c_int), ("sweep", c_int),
    ]


# name -> (restype, argtypes)
_SIGNATURES = {
    "warp_amd_last_error": (ctypes.c_char_p, []),
    "warp_amd_clear_error": (None, []),
    "warp_amd_version": (c_int, []),
    "warp_amd_device_count": (c_int, [ctypes.POINTER(c_int)]),
    "warp_amd_malloc": (c_int, [ctypes.POINTER(c_ptr), c_size]),
    "warp_amd_free": (c_int, [c_ptr]),
    "warp_amd_malloc_host": (c_int, [ctypes.POINTER(c_ptr), c_size]),
    "warp_amd_free_host": (c_int, [c_ptr]),
    "warp_amd_memcpy": (c_int, [c_ptr, c_ptr, c_size, c_int, c_ptr]),
    "warp_amd_stream_synchronize": (c_int, [c_ptr]),
    "warp_amd_set_current_stream": (None, [c_ptr]),
    "warp_amd_get_current_stream": (c_ptr, []),
    "warp_amd_bsr_from_triplets_workspace_size": (c_size, [c_int, c_int]),
    "warp_amd_bsr_from_triplets": (
        c_int,
        [c_int, c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_u64, c_int]
        + [c_ptr] * 7
        + [c_ptr, c_size, c_ptr],
    ),
    "wp_bsr_matrix_from_triplets_device": (
        None,
        [c_int, c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_u64, ctypes.c_bool] + [c_ptr] * 7,
    ),
    "warp_amd_bsr_mv_lanes": (c_int, [c_int, c_int, c_int, c_i64]),
    "warp_amd_bsr_compress_workspace_size": (c_size, [c_int, c_int, c_int]),
    "warp_amd_bsr_compress": (
        c_int,
        [c_int, c_int, c_int, c_int, c_int, c_int, c_u64, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr],
    ),
    "warp_amd_bsr_topology_workspace_size": (c_size, [c_int, c_int, c_int]),
    "warp_amd_bsr_set_from_triplets_padded": (
        c_int,
        [c_int, c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_u64, c_ptr, c_ptr, c_ptr, c_ptr, c_i64]
        + [c_ptr, c_ptr, c_size, c_ptr],
    ),
    "warp_amd_bsr_set_from_triplets_masked": (
        c_int,
        [c_int, c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_i64]
        + [c_ptr, c_size, c_ptr],
    ),
    "warp_amd_bsr_transpose_workspace_size": (c_size, [c_int, c_int]),
    "warp_amd_bsr_transpose": (
        c_int, [c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr]
    ),
    "wp_bsr_transpose_device": (None, [c_int, c_int, c_int] + [c_ptr] * 8),
    "warp_amd_bsr_transpose_values": (c_int, [c_int, c_int, c_int, c_i64, c_ptr, c_ptr, c_ptr, c_f64, c_ptr, c_ptr]),
    "warp_amd_bsr_transpose_masked_values": (
        c_int, [c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_f64, c_ptr, c_ptr]
    ),
    "warp_amd_bsr_assign_list_blocks": (c_int, [c_int] * 6 + [c_ptr] * 5 + [c_ptr]),
    "warp_amd_bsr_assign_copy_blocks": (
        c_int, [c_int] * 8 + [c_ptr, c_ptr, c_ptr, c_int, c_int, c_ptr, c_f64, c_ptr, c_ptr, c_ptr, c_int, c_int, c_ptr, c_ptr]
    ),
    "warp_amd_bsr_assign_padded": (
        c_int, [c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_f64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]
    ),
    "warp_amd_bsr_assign_padded_reblock_topology": (c_int, [c_int] * 7 + [c_ptr] * 8),
    "warp_amd_bsr_reblock_capacity_workspace_size": (c_size, [c_int]),
    "warp_amd_bsr_reblock_capacity": (c_int, [c_int] * 6 + [c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
    "warp_amd_scale_values": (c_int, [c_int, c_i64, c_ptr, c_f64, c_ptr]),
    "warp_amd_bsr_uncompress_rows": (c_int, [c_int, c_ptr, c_ptr, c_int, c_ptr, c_ptr]),
    "warp_amd_rescale_values": (c_int, [c_int, c_i64, c_ptr, c_f64, c_ptr]),
    "warp_amd_copy_scaled_values": (c_int, [c_int, c_int, c_i64, c_ptr, c_f64, c_ptr, c_ptr]),
    "warp_amd_bsr_count_missing_blocks": (c_int, [c_i64] + [c_ptr] * 7),
    "warp_amd_bsr_add_blocks": (
        c_int, [c_int, c_int, c_i64, c_ptr, c_ptr, c_ptr, c_f64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]
    ),
    "warp_amd_bsr_axpy_padded": (
        c_int, [c_int, c_int, c_int, c_f64, c_f64] + [c_ptr] * 10
    ),
    "warp_amd_bsr_mm_count_workspace_size": (c_size, [c_int]),
    "warp_amd_bsr_mm_count": (c_int, [c_int] * 5 + [c_ptr] * 9 + [c_ptr, c_size, c_ptr]),
    "warp_amd_bsr_mm_list": (c_int, [c_int] * 4 + [c_ptr] * 9 + [c_ptr]),
    "warp_amd_bsr_mm_values_tiled": (
        c_int,
        [c_int, c_int, c_int, c_int, c_f64, c_f64, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_int]
        + [c_ptr] * 8 + [c_ptr],
    ),
    "warp_amd_bsr_mm_values": (
        c_int,
        [c_int, c_int, c_int, c_int, c_f64, c_f64, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_int]
        + [c_ptr] * 8 + [c_ptr],
    ),
    "warp_amd_bsr_diag_topology": (c_int, [c_int, c_int, c_ptr, c_ptr, c_ptr]),
    "warp_amd_axpby": (c_int, [c_int, c_i64, c_f64, c_ptr, c_f64, c_ptr, c_ptr]),
    "warp_amd_normalize_dirichlet_projector": (c_int, [c_int, c_int, c_int] + [c_ptr] * 6),
    "warp_amd_bsr_residual_workspace_size": (c_size, [c_int]),
    "warp_amd_bsr_residual_f32_f64": (c_int, [c_int, c_int, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
    "warp_amd_bsr_residual_f32_f64_noalloc": (c_int, [c_int, c_int, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "warp_amd_axpy_f32_to_f64": (c_int, [c_i64, c_f64, c_ptr, c_ptr, c_ptr]),
    "warp_amd_bsr_mv_f32_f64": (c_int, [c_int, c_int, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
    "warp_amd_rms_scales": (c_int, [c_ptr, c_i64, c_ptr, c_ptr, c_ptr]),
    "warp_amd_fcg_direction": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "warp_amd_fcg_update_workspace_size": (c_size, [c_i64]),
    "warp_amd_fcg_update": (
        c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_f64, c_ptr, c_ptr, c_size, c_ptr]
    ),
    "warp_amd_conv3_fwd_f32": (c_int, [c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_int, c_ptr]),
    "warp_amd_conv3_bwd_data_f32": (c_int, [c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_int, c_ptr]),
    "warp_amd_conv3_bwd_weight_workspace_size": (c_size, [c_int, c_int, c_int]),
    "warp_amd_conv3_bwd_weight_f32": (c_int, [c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
    "warp_amd_relu_mask_f32": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr]),
    "warp_amd_pool2_f32": (c_int, [c_int, c_int, c_ptr, c_ptr, c_ptr]),
    "warp_amd_pool2_bwd_f32": (c_int, [c_int, c_int, c_ptr, c_ptr, c_int, c_ptr]),
    "warp_amd_up2_add_f32": (c_int, [c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr]),
    "warp_amd_up2_bwd_f32": (c_int, [c_int, c_int, c_ptr, c_ptr, c_ptr]),
    "warp_amd_mse_workspace_size": (c_size, [c_i64]),
    "warp_amd_mse_loss_f32": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
    "warp_amd_adam_step_f32": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_f32, c_f32, c_f32, c_f32, c_f32, c_ptr]),
    "warp_amd_axpby_f32": (c_int, [c_i64, c_f32, c_ptr, c_f32, c_ptr, c_ptr]),
    "warp_amd_unet_pack_input": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "warp_amd_unet_unpack_output": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr]),
    "warp_amd_jacobi_sweep_f32_f64": (
        c_int, [c_int, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, c_ptr, c_f64, c_ptr, c_ptr]
    ),
    "warp_amd_unet_unpack_add": (c_int, [c_i64, c_ptr, c_ptr, c_f64, c_ptr, c_ptr]),
    "warp_amd_fcg_fused_workspace_size": (c_size, [c_int]),
    "warp_amd_fcg_weight": (c_int, [c_i64, c_ptr, c_ptr, c_f64, c_ptr, c_ptr]),
    "warp_amd_fcg_presmooth": (
        c_int, [c_int, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr]
    ),
    "warp_amd_fcg_postsmooth": (
        c_int,
        [c_int, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, c_f64, c_ptr, c_ptr, c_ptr, c_f64, c_ptr, c_ptr, c_ptr,
         c_ptr, c_size, c_ptr],
    ),
    "warp_amd_fcg_direction_spmv": (
        c_int,
        [c_int, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr,
         c_size, c_ptr],
    ),
    "warp_amd_energy_loss_grad": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "warp_amd_scale_to_f32": (c_int, [c_i64, c_ptr, c_ptr, c_f64, c_ptr, c_ptr]),
    "warp_amd_conv3_mfma_weight_size": (c_size, [c_int, c_int]),
    "warp_amd_conv3_mfma_swizzle": (c_int, [c_int, c_int, c_int, c_ptr, c_ptr, c_ptr]),
    "warp_amd_conv3_mfma": (
        c_int,
        [c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr],
    ),
    "warp_amd_conv3_mfma_wgrad_workspace_size": (c_size, [c_int, c_int, c_int]),
    "warp_amd_conv3_mfma_wgrad": (
        c_int,
        [c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr],
    ),
    "warp_amd_relu_mask_bf16": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr]),
    "warp_amd_f32_to_bf16": (c_int, [c_i64, c_ptr, c_ptr, c_ptr]),
    "warp_amd_bf16_to_f32": (c_int, [c_i64, c_ptr, c_ptr, c_ptr]),
    "warp_amd_unet_pack_input_bf16": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
    "warp_amd_cg_workspace_size": (c_size, [c_i64, c_int]),
    "warp_amd_cg_state_offset": (c_size, [c_int]),
    "warp_amd_dot_workspace_size": (c_size, [c_i64]),
    "warp_amd_cg_tolerance": (c_int, [c_ptr, c_f64, c_f64, c_int, c_ptr, c_ptr]),
    "warp_amd_seg_dot_max_chunks": (c_int, [c_i64, c_int]),
    "warp_amd_seg_dot_plan_workspace_size": (c_size, [c_int]),
    "warp_amd_seg_dot_plan": (c_int, [c_int, c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
    "warp_amd_cg_tolerance_batched": (c_int, [c_int, c_ptr, c_f64, c_f64, c_int, c_ptr, c_ptr]),
    "warp_amd_bsr_mv_packed": (c_int, [ctypes.POINTER(BsrMvArgs), c_ptr, c_ptr, c_f64, c_f64, c_ptr]),
}

for _sfx, _s in (("f32", c_f32), ("f64", c_f64)):
    _SIGNATURES.update(
        {
            f"warp_amd_bsr_mv_{_sfx}": (c_int, [c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, _s, _s, c_ptr]),
            f"warp_amd_bsr_mv_transpose_{_sfx}": (c_int, [c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, _s, _s, c_ptr]),
            f"warp_amd_bsr_mv_transpose_gather_{_sfx}": (c_int, [c_int, c_int, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, _s, _s, c_ptr]),
            f"warp_amd_bsr_mv_ex_{_sfx}": (c_int, [c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_i64, c_ptr, c_ptr, _s, _s, c_int, c_int, c_ptr]),
            f"warp_amd_bsr_get_diag_{_sfx}": (c_int, [c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, _s, c_ptr, c_ptr]),
            f"warp_amd_bsr_jacobi_{_sfx}": (c_int, [c_int, c_int, c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_int, c_int, c_ptr, c_ptr]),
            f"warp_amd_cg_solve_{_sfx}": (
                c_int,
                [c_int, c_int, c_ptr, c_ptr, c_ptr, c_i64, c_int, c_ptr, c_ptr, c_ptr, c_f64, c_f64, c_int, c_int, c_int,
                 c_int, CG_CALLBACK, c_ptr, c_ptr, c_size, ctypes.POINTER(c_int), ctypes.POINTER(c_f64), ctypes.POINTER(c_f64), c_ptr],
            ),
            f"warp_amd_dot_{_sfx}": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_size, c_ptr]),
            f"warp_amd_cg_step_xr_{_sfx}": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
            f"warp_amd_cg_step_p_{_sfx}": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
            f"warp_amd_seg_dot_{_sfx}": (c_int, [c_i64, c_int, c_ptr, c_ptr, c_int, c_ptr, c_ptr, c_ptr, c_ptr, c_ptr]),
            f"warp_amd_cg_step_xr_batched_{_sfx}": (c_int, [c_i64, c_int, c_int] + [c_ptr] * 10),
            f"warp_amd_cg_step_p_batched_{_sfx}": (c_int, [c_i64, c_int, c_int] + [c_ptr] * 8),
            f"warp_amd_diag_mv_{_sfx}": (c_int, [c_i64, c_ptr, c_ptr, c_ptr, c_ptr, _s, _s, c_ptr]),
            f"warp_amd_dense_mv_{_sfx}": (c_int, [c_int, c_int, c_ptr, c_ptr, c_ptr, c_ptr, _s, _s, c_ptr]),
        }
    )


class NativeError(RuntimeError):
    """Raised when a native entry point returns a non-zero status."""


def lib() -> ctypes.CDLL:
    """Load (once) and return the HIP library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NativeError(
                    f"warp_amd HIP library not found at {LIB_PATH}; build it with "
                    "`python -m warp_amd.csrc.build` (there is no CPU fallback)"
                )
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(handle, name, None)
                if fn is None:
                    continue
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    return sorted(_SIGNATURES)


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = lib().warp_amd_last_error()
        msg = msg.decode() if msg else "unknown error"
        raise NativeError(f"{what or 'warp_amd'} failed (status {status}): {msg}")


def call(name: str, *args):
    """Call an int-returning entry point and raise on failure."""
    check(getattr(lib(), name)(*args), name)


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_getter(device: torch.device):
    """A no-argument function returning the raw handle of ``device``'s current
    stream (what stream_of returns), for hot paths that query it per call."""
    index = device.index if device.index is not None else torch.cuda.current_device()
    if _raw_stream is not None:
        return functools.partial(_raw_stream, index)
    return lambda: torch.cuda.current_stream(index).cuda_stream


def stream_of(device: torch.device) -> int:
    """Raw handle of the current stream of ``device`` (no Stream object is built)."""
    if _raw_stream is not None:
        return _raw_stream(device.index if device.index is not None else torch.cuda.current_device())
    return torch.cuda.current_stream(device).cuda_stream
