"""Dirichlet projection of an assembled system (reference warp/_src/fem/dirichlet.py)
against dense float64 restatements of its formulas:
A' = (I - P) A (I - P) + P and b' = (I - P)(b - A P v) + P v, with P's diagonal
blocks normalized to orthogonal projectors (eigenpairs above 1e-6 trace)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import warp_amd as wa
from warp_amd import fem, sparse

pytestmark = pytest.mark.gpu


def _dense(A):
    br, bc = A.block_shape
    off = A.offsets.cpu().numpy()
    rc = None if A.row_counts is None else A.row_counts.cpu().numpy()
    col = A.columns.cpu().numpy()
    val = A.values.cpu().numpy().astype(np.float64).reshape(-1, br, bc)
    D = np.zeros((A.nrow * br, A.ncol * bc))
    for r in range(A.nrow):
        end = off[r] + rc[r] if rc is not None else off[r + 1]
        for k in range(off[r], end):
            D[r * br:(r + 1) * br, col[k] * bc:(col[k] + 1) * bc] += val[k]
    return D


def _normalize_ref(P, v):
    lam, vec = np.linalg.eigh(0.5 * (P + P.T))
    eps = np.trace(P) * 1e-6
    Pn = np.zeros_like(P)
    vn = np.zeros_like(v)
    for k in range(P.shape[0]):
        if lam[k] > eps:
            e = vec[:, k]
            Pn += np.outer(e, e)
            vn += e @ v / lam[k] * e
    return Pn, vn


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_project_linear_system_vs_dense(device, dtype):
    rng = np.random.default_rng(3)
    n = 60
    rows = np.repeat(np.arange(n), 7)
    cols = np.clip(rows + rng.integers(-6, 7, rows.size), 0, n - 1)
    vals = rng.standard_normal((rows.size, 3, 3)).astype(dtype)
    A = sparse.bsr_zeros(n, n, wa.matrix((3, 3), wa.types.type_to_scalar(torch.from_numpy(vals[:1]).dtype)),
                         device=device)
    sparse.bsr_set_from_triplets(A, torch.tensor(rows, dtype=torch.int32, device=device),
                                 torch.tensor(cols.astype(np.int32), device=device), torch.from_numpy(vals).to(device))
    # projector: full clamp on some nodes, a scaled normal constraint on others, none elsewhere
    kind = rng.integers(0, 3, n)
    pblocks = np.zeros((n, 3, 3), dtype)
    for r in range(n):
        if kind[r] == 1:
            pblocks[r] = 2.0 * np.eye(3)
        elif kind[r] == 2:
            nv = rng.standard_normal(3)
            nv /= np.linalg.norm(nv)
            pblocks[r] = 3.0 * np.outer(nv, nv)
    keep = np.flatnonzero(kind > 0).astype(np.int32)
    P = sparse.bsr_zeros(n, n, A.dtype, device=device)
    sparse.bsr_set_from_triplets(P, torch.tensor(keep, device=device), torch.tensor(keep, device=device),
                                 torch.from_numpy(pblocks[keep]).to(device))
    v = rng.standard_normal((n, 3)).astype(dtype)
    b = rng.standard_normal((n, 3)).astype(dtype)

    Ad, Pd = _dense(A), _dense(P)
    Pn = np.zeros_like(Pd)
    vn = v.astype(np.float64).reshape(-1).copy()   # rows without a projector block keep their value
    for r in keep:
        s = slice(3 * r, 3 * r + 3)
        Pn[s, s], vn[s] = _normalize_ref(Pd[s, s], v[r].astype(np.float64))
    I = np.eye(3 * n)
    A_ref = (I - Pn) @ Ad @ (I - Pn) + Pn
    b_ref = (I - Pn) @ (b.reshape(-1) - Ad @ Pn @ vn) + Pn @ vn

    bt = torch.from_numpy(b).to(device)
    vt = torch.from_numpy(v).to(device)
    fem.project_linear_system(A, bt, P, vt)
    tol = dict(rtol=2e-4, atol=2e-4) if dtype == np.float32 else dict(rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(_dense(P), Pn, **tol)
    np.testing.assert_allclose(vt.cpu().numpy().reshape(-1), vn, **tol)
    np.testing.assert_allclose(_dense(A), A_ref, **tol)
    np.testing.assert_allclose(bt.cpu().numpy().reshape(-1), b_ref, **tol)


def test_projected_system_solves_with_cg(device):
    """After projection, CG solves the system and the constrained dofs hold the fixed values."""
    from warp_amd import problems
    from warp_amd.optim import linear

    r, c, v, b, nn = problems.elasticity3d_q1(6, device=device)
    A = sparse.bsr_zeros(nn, nn, wa.mat33, device=device)
    sparse.bsr_set_from_triplets(A, r, c, v, prune_numerical_zeros=False)
    p = 7
    nodes = torch.arange(nn, device=device)
    top = (nodes // (p * p)) == p - 1          # clamp the z = 1 face to a prescribed displacement
    idx = nodes[top].to(torch.int32)
    P = sparse.bsr_from_triplets(nn, nn, idx, idx, torch.eye(3, device=device).expand(idx.numel(), 3, 3).contiguous())
    fixed = torch.zeros(nn, 3, device=device)
    fixed[top, 2] = 0.01
    rhs = b.clone()
    fem.project_linear_system(A, rhs, P, fixed)
    x = torch.zeros_like(rhs)
    linear.cg(A, rhs, x, tol=1e-6, maxiter=2000)
    np.testing.assert_allclose(x[top].cpu().numpy(), fixed[top].cpu().numpy(), atol=1e-6)
    res = sparse.bsr_mv(A, x, alpha=-1.0)
    res += rhs
    assert float(res.norm() / rhs.norm()) < 1e-5
