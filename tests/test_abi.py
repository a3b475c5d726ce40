"""The C-ABI library loads and exports every symbol include/warp_amd.h
declares (CPU only: no compute call needs a GPU here)."""

from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "warp_amd.h")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M)
    return sorted({n for n in names if n.startswith(("warp_amd_", "wp_")) and not n.endswith("_callback")})


def test_header_parses():
    names = declared_functions()
    assert "warp_amd_bsr_mv_f32" in names
    assert "wp_bsr_matrix_from_triplets_device" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from warp_amd import _native

    lib = _native.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    # every declared function has a ctypes signature in the binding
    unbound = [n for n in declared_functions() if n not in _native._SIGNATURES]
    assert not unbound, f"no ctypes signature for: {unbound}"


def test_host_only_queries():
    from warp_amd import _native

    lib = _native.lib()
    assert lib.warp_amd_version() >= 100
    ws = lib.warp_amd_bsr_from_triplets_workspace_size(274625, 16777216)
    # keys + ping-pong + flags: at least 28 bytes per triplet
    assert ws >= 28 * 16777216
    assert lib.warp_amd_cg_workspace_size(1000, 4) >= 4 * 4000
    assert lib.warp_amd_dot_workspace_size(10) >= 8
    offs = [lib.warp_amd_cg_state_offset(i) for i in range(3)]
    assert len(set(offs)) == 3


def test_error_reporting_without_gpu():
    """Argument errors are reported before any device work."""
    from warp_amd import _native

    lib = _native.lib()
    st = lib.warp_amd_bsr_mv_f32(-1, 1, 1, 1, None, None, None, None, 0, None, None, 1.0, 0.0, None)
    assert st != 0
    assert b"invalid shape" in lib.warp_amd_last_error()
    lib.warp_amd_clear_error()
    assert lib.warp_amd_last_error() == b""


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import importlib

    from warp_amd import _native

    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(_native.NativeError, match="no CPU fallback"):
        _native.lib()


def test_cg_rejects_misaligned_workspace_without_gpu():
    """The CG solvers move workspace slices in 16-byte vectors: a workspace
    that is not 16-byte aligned is an argument error, raised before any
    device work (the pointers are never dereferenced)."""
    from warp_amd import _native

    lib = _native.lib()
    n = 64
    need = lib.warp_amd_cg_workspace_size(n, 4)
    fake = 1 << 20  # host-side check only: never dereferenced
    args = (n, 1, fake, fake, fake, 3 * n, 0, None, fake, fake, 1e-6, 0.0, 10, 1, 0, 0, _native.CG_CALLBACK(), None)
    st = lib.warp_amd_cg_solve_f32(*args, fake + 8, need, None, None, None, None)
    assert st == -2
    assert b"16-byte aligned" in lib.warp_amd_last_error()
    lib.warp_amd_clear_error()
