"""FP64 iterative refinement around the FP32 CG (GPU).

The FP64 residual kernel is checked against the oracle's FP64 residual; the
refined solution is checked with the oracle's residual (independent of the
GPU kernels) and against scipy's FP64 direct solve, with the error stated.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl
import torch

import warp_amd as wa
from warp_amd import _native, problems, sparse
from warp_amd.optim import refine

pytestmark = pytest.mark.gpu


def _fv7(device, n, sigma, corr, seed):
    off, col, val, b, k = problems.fv7_problem(n, sigma, corr, seed, seed + 1, device="cpu")
    A = sparse.bsr_zeros(n**3, n**3, wa.float32, device=device)
    A.offsets, A.columns, A.values, A.nnz = off.to(device), col.to(device), val.to(device), col.numel()
    return A, off.numpy(), col.numpy(), val.numpy(), b.numpy().astype(np.float64)


def test_residual_kernel_vs_oracle(device, oracle_mod):
    A, off, col, val, b = _fv7(device, 12, 2.3, 3.0, 5)
    x = np.random.default_rng(0).standard_normal(b.size)
    ref = oracle_mod.residual_f32_f64(off, col, val, x, b)
    lib = _native.lib()
    n = b.size
    r64 = torch.empty(n, dtype=torch.float64, device=device)
    r32 = torch.empty(n, dtype=torch.float32, device=device)
    rr = torch.zeros(1, dtype=torch.float64, device=device)
    wsb = lib.warp_amd_bsr_residual_workspace_size(A.nrow)
    ws = torch.empty(wsb, dtype=torch.uint8, device=device)
    xt = torch.tensor(x, device=device)
    bt = torch.tensor(b, device=device)
    _native.call("warp_amd_bsr_residual_f32_f64", A.nrow, 1, A.offsets.data_ptr(), A.columns.data_ptr(),
                 A.values.data_ptr(), A.nnz, xt.data_ptr(), bt.data_ptr(), r64.data_ptr(), r32.data_ptr(),
                 rr.data_ptr(), ws.data_ptr(), wsb, torch.cuda.current_stream(device).cuda_stream)
    got = r64.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    np.testing.assert_array_equal(r32.cpu().numpy(), got.astype(np.float32))
    assert float(rr.item()) == pytest.approx(float(np.dot(ref, ref)), rel=1e-12)


@pytest.mark.parametrize("n,sigma,corr", [(16, 2.3, 4.0), (24, 1.5, 3.0)])
def test_refined_cg_reaches_fp64_tolerance(device, oracle_mod, n, sigma, corr):
    A, off, col, val, b = _fv7(device, n, sigma, corr, 11)
    x = torch.zeros(b.size, dtype=torch.float64, device=device)
    res = refine.cg_refined(A, torch.tensor(b, device=device), x, tol=1e-10, inner_tol=1e-4)
    xh = x.cpu().numpy()
    r = oracle_mod.residual_f32_f64(off, col, val, xh, b)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    A64 = sp.csr_matrix((val.astype(np.float64), col, off), shape=(b.size, b.size))
    xs = spl.spsolve(A64.tocsc(), b)
    err = np.linalg.norm(xh - xs) / np.linalg.norm(xs)
    print(f"[refine n={n} sigma={sigma}] outer {res.outer_iterations} inner {res.inner_iterations} "
          f"rel. residual (oracle) {rel:.2e}, rel. error vs FP64 direct {err:.2e}")
    assert res.relative_residual <= 1e-10
    assert rel <= 1.5e-10
    # FP64 accuracy well beyond what the FP32 solve alone reaches
    assert err < 1e-7


def test_refined_cg_elasticity(device, oracle_mod):
    rows, cols, vals, b, nn = problems.elasticity3d_q1(6, device="cpu")
    A = sparse.bsr_zeros(nn, nn, wa.mat33, device=device)
    sparse.bsr_set_from_triplets(A, rows.to(device), cols.to(device), vals.to(device), prune_numerical_zeros=False)
    A.nnz_sync()
    b64 = b.to(torch.float64).reshape(-1)
    x = torch.zeros_like(b64, device=device)
    res = refine.cg_refined(A, b64.to(device), x, tol=1e-10, M="block_diag")
    o = oracle_mod.bsr_from_triplets(nn, nn, rows.numpy(), cols.numpy(), vals.numpy(), prune=False)
    r = oracle_mod.residual_f32_f64(o["offsets"], o["columns"], o["values"], x.cpu().numpy(), b64.numpy())
    rel = np.linalg.norm(r) / np.linalg.norm(b64.numpy())
    print(f"[refine elasticity] outer {res.outer_iterations} inner {res.inner_iterations} rel. residual {rel:.2e}")
    assert rel <= 1.5e-10
