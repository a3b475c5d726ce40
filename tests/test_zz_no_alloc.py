"""The FP64 residual entry point that allocates nothing
(warp_amd_bsr_residual_f32_f64_noalloc: its partials live in the library's
device scratch) against the one that takes a caller workspace, through the
C ABI: r64, r32 and r.r bit for bit, on two calls in a row with the second
one smaller (the scratch keeps the first call's partials)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import warp_amd as wa
from warp_amd import _native, problems, sparse

pytestmark = pytest.mark.gpu


def _fv7(device, n, seed):
    off, col, val, b, _ = problems.fv7_problem(n, 2.3, n / 8.0, seed, seed + 1, device="cpu")
    A = sparse.bsr_zeros(n**3, n**3, wa.float32, device=device)
    A.offsets, A.columns, A.values, A.nnz = off.to(device), col.to(device), val.to(device), col.numel()
    return A, 1, b.to(torch.float64).to(device)


def _elasticity(device, n):
    rows, cols, vals, b, nn = problems.elasticity3d_q1(n, device="cpu")
    A = sparse.bsr_zeros(nn, nn, wa.mat33, device=device)
    sparse.bsr_set_from_triplets(A, rows.to(device), cols.to(device), vals.to(device), prune_numerical_zeros=False)
    A.nnz_sync()
    return A, 3, b.reshape(-1).to(torch.float64).to(device)


def _residual(A, br, x, b, device, with_workspace):
    n = b.numel()
    r64 = torch.full((n,), float("nan"), dtype=torch.float64, device=device)
    r32 = torch.full((n,), float("nan"), dtype=torch.float32, device=device)
    rr = torch.full((1,), float("nan"), dtype=torch.float64, device=device)
    args = [A.nrow, br, A.offsets.data_ptr(), A.columns.data_ptr(), A.values.data_ptr(), int(A.nnz), x.data_ptr(),
            b.data_ptr(), r64.data_ptr(), r32.data_ptr(), rr.data_ptr()]
    stream = _native.stream_of(device)
    if with_workspace:
        wsb = _native.lib().warp_amd_bsr_residual_workspace_size(A.nrow)
        ws = torch.empty(wsb, dtype=torch.uint8, device=device)
        _native.call("warp_amd_bsr_residual_f32_f64", *args, ws.data_ptr(), wsb, stream)
    else:
        _native.call("warp_amd_bsr_residual_f32_f64_noalloc", *args, stream)
    torch.cuda.synchronize()
    return r64.cpu().numpy(), r32.cpu().numpy(), rr.cpu().numpy()


@pytest.mark.parametrize("kind,sizes", [("fv7 1x1", (40, 24)), ("elasticity 3x3", (10, 6))])
def test_residual_noalloc_matches_workspace_entry(device, kind, sizes):
    for n in sizes:  # the second call is smaller than the first
        print(f"[no-alloc] {kind} n={n}", flush=True)
        A, br, b = _fv7(device, n, 3) if br_is_scalar(kind) else _elasticity(device, n)
        gen = torch.Generator(device=device).manual_seed(n)
        x = torch.randn(b.numel(), generator=gen, device=device, dtype=torch.float64)
        ref = _residual(A, br, x, b, device, True)
        got = _residual(A, br, x, b, device, False)
        for name, a, c in zip(("r64", "r32", "r.r"), got, ref):
            assert a.tobytes() == c.tobytes(), f"{kind} n={n}: {name} differs from the workspace entry point"
        assert np.isfinite(got[2][0]) and got[2][0] > 0.0


def br_is_scalar(kind: str) -> bool:
    return kind.endswith("1x1")


def test_residual_noalloc_rejects_sizes_beyond_the_scratch(device):
    """2,000,000 rows of 27 3x3 blocks need 125,000 partials (> 65,536): the
    call fails before any launch (the matrix arrays are never read)."""
    one_i = torch.zeros(1, dtype=torch.int32, device=device)
    one_f = torch.zeros(9, dtype=torch.float32, device=device)
    one_d = torch.zeros(3, dtype=torch.float64, device=device)
    nrow = 2_000_000
    with pytest.raises(_native.NativeError, match="scratch"):
        _native.call("warp_amd_bsr_residual_f32_f64_noalloc", nrow, 3, one_i.data_ptr(), one_i.data_ptr(),
                     one_f.data_ptr(), 27 * nrow, one_d.data_ptr(), one_d.data_ptr(), one_d.data_ptr(),
                     one_f.data_ptr(), one_d.data_ptr(), _native.stream_of(device))
    _native.lib().warp_amd_clear_error()
