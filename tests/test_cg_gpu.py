"""GPU conjugate gradient: reference solver scenarios and parity with the oracle.

Scenarios of the reference suite (test_linear_solvers.py:15-126, 129-253)
check err <= atol and the true residual (<= 2 atol in fp64, <= 32 atol in
fp32).  On FEM systems the native solver is compared with the oracle CG
that uses the same FP64-dot policy: same iteration count (checks happen at
the same multiples of check_every) and solutions within 1e-5 relative.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import warp_amd as wa
from warp_amd import problems, sparse
from warp_amd.optim import linear

pytestmark = pytest.mark.gpu


def _spd(n, seed, dtype, device):
    rng = np.random.default_rng(seed)
    C = rng.uniform(-100, 100, (n, n))
    f = rng.uniform(-100, 100, n)
    return torch.tensor(C @ C.T, dtype=dtype, device=device), torch.tensor(C @ f, dtype=dtype, device=device)


def _check_linear_solve(A, b, **kwargs):
    x = torch.zeros_like(b)
    niter, err, atol = linear.cg(A, b, x, use_cuda_graph=True, **kwargs)
    assert err <= atol
    # device-side convergence (check_every = 0) returns device tensors
    x.zero_()
    niter0, err0, atol0 = linear.cg(A, b, x, check_every=0, **kwargs)
    assert math.sqrt(float(err0.item())) <= math.sqrt(float(atol0.item()))
    # warm start
    niter_warm, err, atol = linear.cg(A, b, x, use_cuda_graph=False, **kwargs)
    assert err <= atol
    An = A.cpu().numpy().astype(np.float64) if isinstance(A, torch.Tensor) else None
    if An is not None:
        res = np.linalg.norm(An @ x.cpu().numpy().astype(np.float64) - b.cpu().numpy())
        assert res <= (2.0 if b.dtype == torch.float64 else 32.0) * atol


def test_cg_reference_dense_scenarios(device):
    """test_cg (test_linear_solvers.py:112-126), generic operator route."""
    A, b = _spd(64, 123, torch.float64, device)
    M = linear.preconditioner(A, "diag")
    _check_linear_solve(A, b, maxiter=1000)
    _check_linear_solve(A, b, M=M, maxiter=1000)
    A, b = _spd(16, 321, torch.float32, device)
    M = linear.preconditioner(A, "diag")
    _check_linear_solve(A, b, maxiter=1000)
    _check_linear_solve(A, b, M=M, maxiter=1000)
    rng = np.random.default_rng(321)
    A = torch.eye(5, dtype=torch.float32, device=device)
    b = torch.tensor(rng.uniform(-1, 1, 5), dtype=torch.float32, device=device)
    _check_linear_solve(A, b, maxiter=30)


def test_cg_restart_callbacks(device):
    """test_conjugate_solver_restart (test_linear_solvers.py:129-172) for cg."""
    rng = np.random.default_rng(2027)
    C = rng.standard_normal((12, 12))
    A_np = C.T @ C + 4.0 * np.eye(12)
    b_np = rng.standard_normal(12)
    A = torch.tensor(A_np, device=device)
    b = torch.tensor(b_np, device=device)
    M = linear.preconditioner(A, "diag")
    for prec in (None, M):
        its = []
        x = torch.zeros_like(b)
        niter, err, atol = linear.cg(A, b, x, tol=1e-10, maxiter=120, M=prec, callback=lambda i, r, t: its.append(i),
                                     check_every=1, restart=3)
        assert err <= atol and niter % 3 == 0
        assert its[0] == 0 and its[-1] == niter and all(i % 3 == 0 for i in its)
        assert np.linalg.norm(A_np @ x.cpu().numpy() - b_np) <= atol


def _restart_drift_system(batch_count):
    """_make_cg_restart_drift_system (test_linear_solvers.py:175-201): a
    130-dof mass + spring chain with every 17th dof fixed (identity row and
    column), stacked batch_count times block-diagonally; x0 = 0.1 N(0,1)
    (seed 7301) and b = mass x0, both zero on the fixed dofs."""
    dof_count = 130
    chain_count = dof_count - 1
    edge_count = chain_count - 1
    mass = 3.0 / chain_count
    w = 0.18**2 / (3.0 / edge_count)
    A1 = mass * np.eye(dof_count)
    e = np.arange(edge_count)
    np.add.at(A1, (e, e), w)
    np.add.at(A1, (e + 1, e + 1), w)
    A1[e, e + 1] -= w
    A1[e + 1, e] -= w
    fixed = np.arange(0, chain_count, 17)
    A1[fixed, :] = 0.0
    A1[:, fixed] = 0.0
    A1[fixed, fixed] = 1.0
    A = np.kron(np.eye(batch_count), A1).astype(np.float32)
    rng = np.random.default_rng(7301)
    x0 = (0.1 * rng.standard_normal((batch_count, dof_count))).astype(np.float32)
    b = (mass * x0).astype(np.float32)
    x0[:, fixed] = 0.0
    b[:, fixed] = 0.0
    return A, b.reshape(-1), x0.reshape(-1), dof_count


@pytest.mark.parametrize("batch_count", [1, 3])
def test_cg_restart_float32_drift(device, batch_count):
    """test_cg_restart_float32_drift (test_linear_solvers.py:204-253): FP32 CG
    restarted every 32 iterations, checked every 32, graph-replayed, to
    tol 1e-6; unpreconditioned and Jacobi.  Every subproblem's FP64 true
    residual must reach tol |b_i| (restarts recompute r = b - A x, so the
    recurrence's drift cannot stop the solve short)."""
    interval, tol = 32, 1.0e-6
    A_np, b_np, x0_np, ndof = _restart_drift_system(batch_count)
    A = torch.tensor(A_np, device=device)
    b = torch.tensor(b_np, device=device)
    M = linear.preconditioner(A, "diag")
    offs = None if batch_count == 1 else _batch_offsets([ndof] * batch_count, device)
    op = linear.aslinearoperator(A, batch_offsets=offs)
    for name, prec in (("none", None), ("Jacobi", M)):
        x = torch.tensor(x0_np, device=device)
        niter, err, atol = linear.cg(op, b, x, tol=tol, atol=0.0, maxiter=512, M=prec, check_every=interval,
                                     use_cuda_graph=True, restart=interval)
        err_t, atol_t = torch.as_tensor(err), torch.as_tensor(atol)
        assert bool((err_t <= atol_t).all()), (name, err, atol)
        assert niter % interval == 0 and niter <= 512, (name, niter)
        xn = x.cpu().numpy().astype(np.float64)
        for k in range(batch_count):
            sl = slice(k * ndof, (k + 1) * ndof)
            res = A_np[sl, sl].astype(np.float64) @ xn[sl] - b_np[sl].astype(np.float64)
            target = tol * np.linalg.norm(b_np[sl].astype(np.float64))
            assert np.linalg.norm(res) <= target, f"{name} batch {k}: {np.linalg.norm(res):.3e} > {target:.3e}"


def _lanes(A):
    from warp_amd import _native

    br, bc = A.block_shape
    return _native.lib().warp_amd_bsr_mv_lanes(A.nrow, br, bc, int(A.nnz))


def _bsr_from(device, nn, rows, cols, vals, prune=True):
    block = wa.mat33 if vals.ndim == 3 else wa.float32
    A = sparse.bsr_zeros(nn, nn, block, device=device)
    sparse.bsr_set_from_triplets(A, rows.to(device), cols.to(device), vals.to(device), prune_numerical_zeros=prune)
    A.nnz_sync()
    return A


@pytest.mark.parametrize("n", [32, 128])
def test_cg_poisson_vs_oracle(device, oracle_mod, n):
    """2D Q1 Poisson, Jacobi CG, tol 1e-6, maxiter 2000 (the CPU-runnable workload at n=128)."""
    rows, cols, vals, b, nd = problems.poisson2d_q1(n, 0)
    A = _bsr_from(device, nd, torch.from_numpy(rows), torch.from_numpy(cols), torch.from_numpy(vals))
    M = linear.preconditioner(A, "diag")
    bt = torch.from_numpy(b).to(device)
    x = torch.zeros_like(bt)
    its = []
    niter, err, atol = linear.cg(A, bt, x, tol=1e-6, maxiter=2000, M=M, callback=lambda i, r, t: its.append((i, r)))
    o = oracle_mod.bsr_from_triplets(nd, nd, rows, cols, vals)
    inv = oracle_mod.inverse_diag(oracle_mod.bsr_get_diag(o["offsets"], o["columns"], o["values"]))
    G = _lanes(A)
    # oracle in the kernels' SpMV order: same trajectory up to FP64 dot rounding
    xo, it_o, err_o, atol_o, hist = oracle_mod.cg(o["offsets"], o["columns"], o["values"], b, np.zeros(nd), tol=1e-6,
                                                 maxiter=2000, prec_kind=oracle_mod.PREC_DIAG, prec=inv, group=G)
    # oracle in the reference row order
    xr, it_r, err_r, _, _ = oracle_mod.cg(o["offsets"], o["columns"], o["values"], b, np.zeros(nd), tol=1e-6,
                                         maxiter=2000, prec_kind=oracle_mod.PREC_DIAG, prec=inv)
    print(f"[cg poisson n={n}] gpu iters {niter} err {err:.3e} | oracle iters {it_o} err {err_o:.3e} "
          f"| reference-order iters {it_r} | atol {atol:.3e}")
    assert err <= atol
    assert atol == pytest.approx(atol_o, rel=1e-12)
    assert niter == it_o
    xg = x.cpu().numpy()
    assert np.linalg.norm(xg - xo) <= 1e-5 * np.linalg.norm(xo)
    assert abs(niter - it_r) <= 10
    # FP64 true residual of the FP32 solution; error vs the FP64 direct solve matches the oracle's
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    A64 = sp.csr_matrix((o["values"].astype(np.float64), o["columns"], o["offsets"]), shape=(nd, nd))
    assert np.linalg.norm(b - A64 @ xg.astype(np.float64)) <= 32 * atol
    xs = spl.spsolve(A64.tocsc(), b.astype(np.float64))
    eg = np.linalg.norm(xg - xs) / np.linalg.norm(xs)
    er = np.linalg.norm(xr - xs) / np.linalg.norm(xs)
    print(f"[cg poisson n={n}] rel. error vs FP64 direct: gpu {eg:.3e}, reference-order oracle {er:.3e}")
    assert eg <= 1.05 * er + 1e-6
    assert its[0][0] == 0 and its[-1][0] == niter


@pytest.mark.parametrize("ptype", ["diag", "block_diag", "id"])
def test_cg_elasticity_vs_oracle(device, oracle_mod, ptype):
    n = 8
    rows, cols, vals, b, nn = problems.elasticity3d_q1(n, device="cpu")
    A = _bsr_from(device, nn, rows, cols, vals, prune=False)
    M = linear.preconditioner(A, ptype)
    bt = b.to(device)
    x = torch.zeros_like(bt)
    niter, err, atol = linear.cg(A, bt, x, tol=1e-6, atol=0.0, maxiter=5000, M=M)
    o = oracle_mod.bsr_from_triplets(nn, nn, rows.numpy(), cols.numpy(), vals.numpy(), prune=False)
    if ptype == "id":
        pk, pd = oracle_mod.PREC_NONE, None
    elif ptype == "diag":
        pk, pd = oracle_mod.PREC_DIAG, M._prec_data.cpu().numpy()
    else:
        pk, pd = oracle_mod.PREC_BLOCK, M._prec_data.cpu().numpy()
    xo, it_o, err_o, atol_o, _ = oracle_mod.cg(o["offsets"], o["columns"], o["values"], b.numpy(), np.zeros(3 * nn),
                                              tol=1e-6, atol=0.0, maxiter=5000, prec_kind=pk, prec=pd, group=_lanes(A))
    print(f"[cg elasticity {ptype}] gpu iters {niter} err {err:.3e} | oracle iters {it_o} err {err_o:.3e}")
    assert err <= atol
    assert niter == it_o
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    A64 = sp.bsr_matrix((o["values"].astype(np.float64), o["columns"], o["offsets"]), shape=(3 * nn, 3 * nn)).tocsr()
    xs = spl.spsolve(A64.tocsc(), b.numpy().astype(np.float64).reshape(-1))
    xg = x.cpu().numpy().reshape(-1).astype(np.float64)
    rel64 = np.linalg.norm(xg - xs) / np.linalg.norm(xs)
    print(f"[cg elasticity {ptype}] rel. error vs FP64 direct solve = {rel64:.3e}")
    assert np.linalg.norm(b.numpy().reshape(-1) - A64 @ xg) <= 32 * atol
    assert np.linalg.norm(xg - xo) <= 1e-5 * np.linalg.norm(xo)


def test_cg_native_check_every_zero_and_restart(device, oracle_mod):
    rows, cols, vals, b, nd = problems.poisson2d_q1(24, 0)
    A = _bsr_from(device, nd, torch.from_numpy(rows), torch.from_numpy(cols), torch.from_numpy(vals))
    M = linear.preconditioner(A, "diag")
    bt = torch.from_numpy(b).to(device)
    x = torch.zeros_like(bt)
    it, rr, at = linear.cg(A, bt, x, tol=1e-6, atol=0.0, maxiter=500, M=M, check_every=0)
    assert isinstance(it, torch.Tensor)
    assert float(rr.item()) <= float(at.item())
    x1 = torch.zeros_like(bt)
    n1, e1, a1 = linear.cg(A, bt, x1, tol=1e-6, atol=0.0, maxiter=500, M=M, check_every=1)
    # device-side convergence stops at the first converged iteration, like check_every=1
    assert int(it.item()) == n1
    np.testing.assert_array_equal(x.cpu().numpy(), x1.cpu().numpy())
    # restart: true residual at the cycle boundary
    x2 = torch.zeros_like(bt)
    n2, e2, a2 = linear.cg(A, bt, x2, tol=1e-6, atol=0.0, maxiter=512, M=M, check_every=32, restart=32)
    assert n2 % 32 == 0 and e2 <= a2
    o = oracle_mod.bsr_from_triplets(nd, nd, rows, cols, vals)
    import scipy.sparse as sp

    A64 = sp.csr_matrix((o["values"].astype(np.float64), o["columns"], o["offsets"]), shape=(nd, nd))
    assert np.linalg.norm(b - A64 @ x2.cpu().numpy()) <= 2 * a2


def test_cg_functor_reuse(device):
    rows, cols, vals, b, nd = problems.poisson2d_q1(16, 0)
    A = _bsr_from(device, nd, torch.from_numpy(rows), torch.from_numpy(cols), torch.from_numpy(vals))
    bt = torch.from_numpy(b).to(device)
    x = torch.zeros_like(bt)
    solver = linear.cg(A, bt, x, tol=1e-6, run=False)
    n1, e1, a1 = solver()
    x2 = torch.zeros_like(bt)
    n2, e2, a2 = solver(b=bt, x=x2)
    assert n1 == n2 and e1 == e2
    np.testing.assert_array_equal(x.cpu().numpy(), x2.cpu().numpy())
    with pytest.raises(ValueError, match="Incompatible b.shape"):
        solver(b=torch.zeros(nd + 1, device=device))


def test_cg_zero_rhs_and_identity(device):
    """b = 0 converges before the first iteration with x = 0; the identity
    converges in one iteration (textbook special cases)."""
    n = 50
    idx = torch.arange(n, dtype=torch.int32)
    A = _bsr_from(device, n, idx, idx, torch.ones(n))
    b = torch.zeros(n, device=device)
    x = torch.zeros(n, device=device)
    it, err, atol = linear.cg(A, b, x, tol=1e-6, maxiter=100)
    assert it == 0 and err == 0.0 and torch.count_nonzero(x) == 0
    b = torch.randn(n, device=device)
    it, err, atol = linear.cg(A, b, x, tol=1e-6, maxiter=100, check_every=1)
    assert it == 1 and err <= atol
    torch.testing.assert_close(x, b, rtol=1e-6, atol=0)


def _batch_offsets(sizes, device):
    return torch.tensor(np.concatenate([[0], np.cumsum(sizes)]), dtype=torch.int32, device=device)


def _spd_np(n, seed, dtype):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (q * rng.uniform(1.0, 10.0, n)) @ q.T
    return A.astype(dtype), rng.standard_normal(n).astype(dtype)


@pytest.mark.parametrize("sizes", [[20, 20, 20, 20], [8, 15, 10, 12]])
def test_batched_cg_dense(device, sizes):
    """test_batched_cg / test_batched_nonuniform (test_linear_solvers.py:451-489):
    block-diagonal SPD system, one subproblem per block; each block's residual
    within 32 tol |b_i| (FP32)."""
    rows = sum(sizes)
    A = np.zeros((rows, rows), np.float32)
    b = np.zeros(rows, np.float32)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for i, n in enumerate(sizes):
        A[off[i]:off[i + 1], off[i]:off[i + 1]], b[off[i]:off[i + 1]] = _spd_np(n, i, np.float32)
    op = linear.aslinearoperator(torch.tensor(A, device=device), batch_offsets=_batch_offsets(sizes, device),
                                 max_batch_length=max(sizes))
    assert op.batch_count == len(sizes) and op.max_batch_length == max(sizes)
    x = torch.zeros(rows, device=device)
    it, err, atol = linear.cg(op, torch.tensor(b, device=device), x, tol=1e-5, maxiter=1000)
    assert err <= atol
    xn = x.cpu().numpy().astype(np.float64)
    for i in range(len(sizes)):
        sl = slice(off[i], off[i + 1])
        res = np.linalg.norm(A[sl, sl].astype(np.float64) @ xn[sl] - b[sl])
        assert res <= 32 * 1e-5 * np.linalg.norm(b[sl].astype(np.float64)), i


def test_batched_cg_vector_entries_and_inactive_tail(device):
    """test_batched_vector_offsets / _inactive_tail / _vector_inactive_tail
    (test_linear_solvers.py:492-577): diagonal operators, entries of two scalars
    assigned to the subproblem of their first dof, inactive dofs untouched."""
    d = torch.tensor([2.0, 2.0, 5.0, 5.0], device=device)
    b = torch.tensor([[2.0, 4.0], [10.0, 15.0]], device=device)
    x = torch.zeros_like(b)
    op = linear.aslinearoperator(d, batch_offsets=_batch_offsets([2, 2], device), max_batch_length=2)
    linear.cg(op, b, x, tol=1e-7, maxiter=4, check_every=1, use_cuda_graph=False)
    np.testing.assert_allclose(x.cpu().numpy(), [[1.0, 2.0], [2.0, 3.0]], rtol=1e-5, atol=1e-5)
    # one subproblem over the first three dofs of five
    d = torch.tensor([2.0, 3.0, 4.0, 5.0, 6.0], dtype=torch.float64, device=device)
    b = torch.tensor([2.0, 6.0, 12.0, 100.0, 200.0], dtype=torch.float64, device=device)
    x = torch.tensor([0.0, 0.0, 0.0, 7.0, 8.0], dtype=torch.float64, device=device)
    op = linear.aslinearoperator(d, batch_offsets=_batch_offsets([3], device), max_batch_length=3)
    it, err, atol = linear.cg(op, b, x, tol=1e-8, maxiter=20, check_every=1)
    assert it > 1 and err <= atol
    np.testing.assert_allclose(x.cpu().numpy(), [1.0, 2.0, 3.0, 7.0, 8.0], rtol=1e-7, atol=1e-7)
    # two-scalar entries, the last entry inactive
    d = torch.tensor([2.0, 3.0, 4.0, 5.0, 6.0, 7.0], dtype=torch.float64, device=device)
    b = torch.tensor([[2.0, 6.0], [12.0, 20.0], [100.0, 200.0]], dtype=torch.float64, device=device)
    x = torch.tensor([[0.0, 0.0], [0.0, 0.0], [7.0, 8.0]], dtype=torch.float64, device=device)
    op = linear.aslinearoperator(d, batch_offsets=_batch_offsets([2, 2], device), max_batch_length=2)
    _, err, atol = linear.cg(op, b, x, tol=1e-8, maxiter=8, check_every=1, use_cuda_graph=False)
    assert err <= atol
    np.testing.assert_allclose(x.cpu().numpy(), [[1.0, 2.0], [3.0, 4.0], [7.0, 8.0]], rtol=1e-7, atol=1e-7)


def test_batched_cg_bsr_matches_single_solves(device):
    """Three 2D Poisson systems stacked block-diagonally in one BSR matrix,
    one subproblem each, Jacobi-preconditioned: every subproblem's solution
    equals its own single-system solve (the subproblems only differ in the FP64
    dot grouping) and the device-side results hold one value per subproblem."""
    mats = []
    for k, n in enumerate((12, 16, 20)):
        rows, cols, vals, b, nd = problems.poisson2d_q1(n, k)
        mats.append((rows, cols, vals, b, nd))
    off = np.concatenate([[0], np.cumsum([m[4] for m in mats])])
    R = np.concatenate([m[0] + off[i] * (m[0] >= 0) for i, m in enumerate(mats)]).astype(np.int32)
    C = np.concatenate([m[1] + off[i] * (m[1] >= 0) for i, m in enumerate(mats)]).astype(np.int32)
    R = np.where(np.concatenate([m[0] for m in mats]) < 0, -1, R)
    C = np.where(np.concatenate([m[1] for m in mats]) < 0, -1, C)
    V = np.concatenate([m[2] for m in mats])
    B = np.concatenate([m[3] for m in mats])
    ntot = int(off[-1])
    A = _bsr_from(device, ntot, torch.from_numpy(R), torch.from_numpy(C), torch.from_numpy(V))
    M = linear.preconditioner(A, "diag")
    op = linear.aslinearoperator(A, batch_offsets=torch.tensor(off, dtype=torch.int32, device=device))
    x = torch.zeros(ntot, device=device)
    it, err, atol = linear.cg(op, torch.from_numpy(B).to(device), x, tol=1e-6, maxiter=2000, M=M)
    assert err <= atol
    for i, (rows, cols, vals, b, nd) in enumerate(mats):
        Ai = _bsr_from(device, nd, torch.from_numpy(rows), torch.from_numpy(cols), torch.from_numpy(vals))
        xi = torch.zeros(nd, device=device)
        linear.cg(Ai, torch.from_numpy(b).to(device), xi, tol=1e-6, maxiter=2000, M=linear.preconditioner(Ai, "diag"))
        xs = x[off[i]:off[i + 1]].cpu().numpy()
        ref = xi.cpu().numpy()
        assert np.linalg.norm(xs - ref) <= 1e-4 * np.linalg.norm(ref), i
    x.zero_()
    its, rr, at = linear.cg(op, torch.from_numpy(B).to(device), x, tol=1e-6, maxiter=400, M=M, check_every=0)
    assert rr.shape == (3,) and at.shape == (3,) and bool((rr <= at).all())


@pytest.mark.parametrize("n,restart,check_every", [(64, None, 10), (64, 5, 10), (64, 7, 10), (64, None, 0),
                                                   (600, None, 10)])
def test_cg_graph_replay_matches_eager(device, n, restart, check_every):
    """use_cuda_graph replays the cycles between two checks as one HIP graph
    on a private stream (small systems); the iterates are the eager ones bit
    for bit, replay after replay: fused direction pass with device-side buffer
    parity at n = 64, restart cycles that divide check_every (graphed) or not
    (eager), device-only convergence (groups of 16 cycles), and a large system
    that always launches directly (n = 600)."""
    rows, cols, vals, b, nd = problems.poisson2d_q1(n, 3)
    A = _bsr_from(device, nd, torch.from_numpy(rows), torch.from_numpy(cols), torch.from_numpy(vals))
    M = linear.preconditioner(A, "diag")
    bt = torch.from_numpy(b).to(device)
    out = []
    for graph in (False, True, True):
        x = torch.zeros_like(bt)
        it, err, atol = linear.cg(A, bt, x, tol=1e-6, maxiter=3000, M=M, restart=restart, check_every=check_every,
                                  use_cuda_graph=graph)
        if check_every == 0:
            it, err = int(it.cpu()[0]), float(err.cpu()[0])
        out.append((it, err, x.cpu().numpy().tobytes()))
    assert out[0] == out[1] == out[2]
    assert out[0][0] > 10


def _seg_dot(device, a, b, offs):
    from warp_amd import _native

    lib = _native.lib()
    n, nb = a.numel(), offs.numel() - 1
    sfx = "f64" if a.dtype == torch.float64 else "f32"
    chunk = torch.empty(nb + 1, dtype=torch.int32, device=device)
    wsb = lib.warp_amd_seg_dot_plan_workspace_size(nb)
    ws = torch.empty(wsb, dtype=torch.uint8, device=device)
    s = torch.cuda.current_stream(device).cuda_stream
    _native.call("warp_amd_seg_dot_plan", nb, offs.data_ptr(), chunk.data_ptr(), ws.data_ptr(), wsb, s)
    mc = lib.warp_amd_seg_dot_max_chunks(n, nb)
    partials = torch.empty(max(mc, 1), dtype=torch.float64, device=device)
    out = torch.full((nb,), float("nan"), dtype=torch.float64, device=device)
    _native.call(f"warp_amd_seg_dot_{sfx}", n, nb, offs.data_ptr(), chunk.data_ptr(), mc, a.data_ptr(), b.data_ptr(),
                 partials.data_ptr(), out.data_ptr(), s)
    return out.cpu().numpy()


def test_segmented_dot_known_answers(device):
    """test_tiled_dot_large_single_batch / _batched_tree known answers
    (test_linear_solvers.py:605-680): inactive tail excluded, empty and odd-sized
    subproblems, and per-subproblem sums against FP64 sums of the same data."""
    length = 100_000
    active = length - 2
    a = torch.full((length,), 2.0, device=device)
    a[active:] = 100.0
    out = _seg_dot(device, a, a, _batch_offsets([active], device))
    assert out[0] == 4 * active
    sizes = [0, 1, 511, 512, 513, 65_537, 262_145]
    n = sum(sizes) + 17
    offs = _batch_offsets(sizes, device)
    two = torch.full((n,), 2.0, device=device)
    three = torch.full((n,), 3.0, device=device)
    np.testing.assert_array_equal(_seg_dot(device, two, three, offs), 6.0 * np.array(sizes))
    rng = np.random.default_rng(123)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    got = _seg_dot(device, torch.tensor(x, device=device), torch.tensor(y, device=device), offs)
    bo = np.concatenate([[0], np.cumsum(sizes)])
    ref = np.array([math.fsum(x[s:e] * y[s:e]) for s, e in zip(bo[:-1], bo[1:])])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)
