"""GPU parity at the benchmark workloads' own sizes, and on the assembly
routes that only very large row counts reach.

Every case runs the HIP path and the CPU oracle on the same seeded inputs:

* assembly: offsets, columns and summed values bit for bit (both sides sum
  duplicates sequentially in ascending triplet order); the summed values'
  error against an FP64 sum of the same triplets is printed and bounded by
  1e-6 relative to the sum of magnitudes;
* bsr_mv: bit for bit against the oracle's lane-group order, within 1e-5 of
  its row order (relative to (|A| |x|)_i), FP64 error printed;
* CG: the same iteration count as the oracle run in the kernels' SpMV order,
  FP64 true residual printed.

The routes of the assembly partition are picked from the row count: up to
2,097,152 rows the level-1 scatter is staged in LDS, above that it is not,
and above 33,554,432 rows the assembly buckets rows directly with global
atomics.  The fv7 face-loop and the sparse 40M-row cases reach the last two.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import warp_amd as wa
from warp_amd import _native, problems, sparse
from warp_amd.optim import linear

pytestmark = pytest.mark.gpu

STAGED_MAX_ROWS = 256 * 8192
TILED_MAX_ROWS = 4096 * 8192


def _say(msg: str) -> None:
    print(msg, flush=True)


def _fp64_sums(o, vals: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """FP64 sum and sum of magnitudes of each block's triplets, from the
    oracle's sorted triplet order (summed_offsets are inclusive run ends)."""
    idx = o["summed_indices"]
    ends = o["summed_offsets"].astype(np.int64)
    starts = np.concatenate([[0], ends[:-1]])
    v = vals.reshape(vals.shape[0], -1)[idx].astype(np.float64)
    s = np.add.reduceat(v, starts, axis=0) if idx.size else np.zeros((0, v.shape[1]))
    m = np.add.reduceat(np.abs(v), starts, axis=0) if idx.size else np.zeros((0, v.shape[1]))
    return s, m


def _check_assembly(A, o, vals: np.ndarray, label: str) -> None:
    nnz = A.nnz_sync()
    assert nnz == o["nnz"], (label, nnz, o["nnz"])
    assert np.array_equal(A.offsets.cpu().numpy(), o["offsets"]), f"{label}: offsets differ"
    assert np.array_equal(A.columns[:nnz].cpu().numpy(), o["columns"]), f"{label}: columns differ"
    got = A.values[:nnz].cpu().numpy()
    assert got.tobytes() == o["values"].tobytes(), f"{label}: summed values differ bitwise"
    s64, mag = _fp64_sums(o, vals)
    err = np.abs(got.reshape(nnz, -1).astype(np.float64) - s64) / np.maximum(mag, np.finfo(np.float64).tiny)
    worst = float(err.max()) if err.size else 0.0
    _say(f"[scale] {label}: {nnz} blocks bit-exact vs oracle; FP32 sums vs FP64: max {worst:.2e} of sum |v|")
    assert worst <= 1e-6, worst


def test_assembly_elasticity64_shuffled(device, oracle_mod):
    """The 16,777,216 shuffled element-loop triplets of the 64^3 hex mesh
    (3x3 fp32 values N(0,1), 7,189,057 distinct blocks)."""
    _say("[scale] assembly: 64^3 element loop, 16,777,216 shuffled 3x3 triplets")
    rows, cols, vals, nn = problems.elasticity_triplets_shuffled(64, device=device)
    A = sparse.bsr_zeros(nn, nn, wa.mat33, device=device)
    sparse.bsr_set_from_triplets(A, rows, cols, vals)
    r, c, v = rows.cpu().numpy(), cols.cpu().numpy(), vals.cpu().numpy()
    del rows, cols, vals
    o = oracle_mod.bsr_from_triplets(nn, nn, r, c, v)
    assert o["nnz"] == 193**3
    _check_assembly(A, o, v, "elasticity 64^3 shuffled")


def _fv7_face_triplets(n: int, seed: int):
    """Face loop of a 7-point finite-volume operator on an n^3 cell grid:
    each interior face (i, j) adds t to (i, i) and (j, j) and -t to (i, j)
    and (j, i), t ~ U(0.5, 1.5); the stream is shuffled."""
    rng = np.random.default_rng(seed)
    idx = np.arange(n**3, dtype=np.int64).reshape(n, n, n)  # [z, y, x]
    faces = [
        (idx[:, :, :-1].ravel(), idx[:, :, 1:].ravel()),
        (idx[:, :-1, :].ravel(), idx[:, 1:, :].ravel()),
        (idx[:-1, :, :].ravel(), idx[1:, :, :].ravel()),
    ]
    a = np.concatenate([f[0] for f in faces])
    b = np.concatenate([f[1] for f in faces])
    t = rng.uniform(0.5, 1.5, a.size).astype(np.float32)
    rows = np.concatenate([a, b, a, b]).astype(np.int32)
    cols = np.concatenate([a, b, b, a]).astype(np.int32)
    vals = np.concatenate([t, t, -t, -t])
    perm = rng.permutation(rows.size)
    return rows[perm], cols[perm], vals[perm]


def test_assembly_unstaged_route_fv7_face_loop(device, oracle_mod):
    """2,197,000 rows (> 2,097,152: more than 256 coarse buckets of 8192
    rows, so the level-1 scatter is not staged in LDS); 26,160,000 scalar
    triplets, every diagonal entry the sum of its cell's faces."""
    n = 130
    rows, cols, vals = _fv7_face_triplets(n, 17)
    N = n**3
    assert STAGED_MAX_ROWS < N <= TILED_MAX_ROWS
    _say(f"[scale] assembly: fv7 face loop {n}^3 ({N} rows), {rows.size} shuffled scalar triplets")
    A = sparse.bsr_zeros(N, N, wa.float32, device=device)
    sparse.bsr_set_from_triplets(A, torch.from_numpy(rows).to(device), torch.from_numpy(cols).to(device),
                                 torch.from_numpy(vals).to(device))
    o = oracle_mod.bsr_from_triplets(N, N, rows, cols, vals)
    assert o["nnz"] == N + 6 * n * n * (n - 1)
    _check_assembly(A, o, vals, f"fv7 face loop {n}^3")


@pytest.mark.parametrize("bshape", [(), (3, 3)])
def test_assembly_global_atomic_route(device, oracle_mod, bshape):
    """40,000,000 rows (> 33,554,432: buckets would exceed 8192 rows, so the
    rows are bucketed directly with global atomics) and 2,000,000 triplets
    over 500,000 keys (mean multiplicity 4), 3 % of them out of range."""
    nrow, ncol, T, U = 40_000_000, 1000, 2_000_000, 500_000
    assert nrow > TILED_MAX_ROWS
    _say(f"[scale] assembly: {nrow} rows, {T} triplets over {U} keys, block {bshape or '1x1'}")
    rng = np.random.default_rng(23 + len(bshape))
    key_r = rng.integers(0, nrow, U).astype(np.int32)
    key_c = rng.integers(0, ncol, U).astype(np.int32)
    which = rng.integers(0, U, T)
    rows, cols = key_r[which], key_c[which]
    bad = rng.random(T) < 0.03
    rows[bad & (rng.random(T) < 0.5)] = -1
    cols[bad & (rng.random(T) < 0.5)] = ncol + 5
    vals = rng.standard_normal((T, *bshape)).astype(np.float32)
    btype = wa.mat33 if bshape else wa.float32
    A = sparse.bsr_zeros(nrow, ncol, btype, device=device)
    sparse.bsr_set_from_triplets(A, torch.from_numpy(rows).to(device), torch.from_numpy(cols).to(device),
                                 torch.from_numpy(vals).to(device))
    o = oracle_mod.bsr_from_triplets(nrow, ncol, rows, cols, vals)
    _check_assembly(A, o, vals, f"{nrow}-row global-atomic route")


def _elasticity(n: int, device):
    rows, cols, vals, b, nn = problems.elasticity3d_q1(n, device=device)
    A = sparse.bsr_zeros(nn, nn, wa.mat33, device=device)
    sparse.bsr_set_from_triplets(A, rows, cols, vals, prune_numerical_zeros=False)
    nnz = A.nnz_sync()
    return A, b, nn, nnz


def _host_csr(A, nnz):
    return A.offsets.cpu().numpy(), A.columns[:nnz].cpu().numpy(), A.values[:nnz].cpu().numpy()


def test_bsr_mv_elasticity64_full(device, oracle_mod):
    """bsr_mv on the 64^3 elasticity matrix (274,625 block rows, 7,189,057
    3x3 blocks), without nnz_sync first (stale nnz) and after it."""
    import scipy.sparse as sp

    rows, cols, vals, _, nn = problems.elasticity3d_q1(64, device=device)
    A = sparse.bsr_zeros(nn, nn, wa.mat33, device=device)
    sparse.bsr_set_from_triplets(A, rows, cols, vals, prune_numerical_zeros=False)
    del rows, cols, vals
    gen = torch.Generator(device=device).manual_seed(29)
    x = torch.randn(nn * 3, generator=gen, device=device, dtype=torch.float32)
    assert A.nnz == 64**3 * 64 + 65**2  # host bound: the triplet count (element loop + clamped identities)
    torch.cuda.synchronize()
    # no nnz_sync: the pending transfer of offsets[nrow] has landed, and the
    # lane choice reads it without waiting
    y_stale = sparse.bsr_mv(A, x).reshape(-1).cpu().numpy()
    nnz = A.nnz_sync()
    assert nnz == 193**3
    y = sparse.bsr_mv(A, x).reshape(-1).cpu().numpy()
    off, col, val = _host_csr(A, nnz)
    xh = x.cpu().numpy()
    g = _native.lib().warp_amd_bsr_mv_lanes(nn, 3, 3, nnz)
    y_grp = oracle_mod.bsr_mv(off, col, val, xh, group=g)
    assert y.tobytes() == y_grp.tobytes(), "bsr_mv differs from the lane-group-order oracle"
    assert y_stale.tobytes() == y.tobytes(), "lanes chosen from the stale nnz changed the result"
    y_row = oracle_mod.bsr_mv(off, col, val, xh, group=0)
    Ab = sp.bsr_matrix((val.astype(np.float64), col, off), shape=(3 * nn, 3 * nn))
    Aabs = sp.bsr_matrix((np.abs(val).astype(np.float64), col, off), shape=(3 * nn, 3 * nn))
    scale = Aabs @ np.abs(xh.astype(np.float64))
    y64 = Ab @ xh.astype(np.float64)
    err_row = float(np.max(np.abs(y.astype(np.float64) - y_row) / scale))
    err64 = float(np.max(np.abs(y.astype(np.float64) - y64) / scale))
    _say(f"[scale] bsr_mv 64^3 elasticity: G={g}, bit-exact vs lane-group oracle; vs row order {err_row:.2e}, "
         f"vs FP64 {err64:.2e} (of (|A||x|)_i); relative 2-norm error vs FP64 "
         f"{np.linalg.norm(y - y64) / np.linalg.norm(y64):.2e}")
    assert err_row <= 1e-5 and err64 <= 1e-5


@pytest.mark.parametrize("ptype,prec_kind", [("diag", 1), ("block_diag", 2)])
def test_cg_elasticity32_vs_oracle(device, oracle_mod, ptype, prec_kind):
    """Jacobi and block-Jacobi CG to 1e-6 on the 32^3 elasticity system
    (107,811 dofs): the oracle, run in the kernels' SpMV order with FP64 dots,
    takes the same number of iterations; the FP64 true residual is printed."""
    A, b, nn, nnz = _elasticity(32, device)
    off, col, val = _host_csr(A, nnz)
    bt = b.reshape(-1).contiguous()
    M = linear.preconditioner(A, ptype)
    x = torch.zeros_like(bt)
    it, err, atol = linear.cg(A, bt, x, tol=1e-6, atol=0.0, maxiter=5000, M=M)
    g = _native.lib().warp_amd_bsr_mv_lanes(nn, 3, 3, nnz)
    prec = M._prec_data.cpu().numpy()
    xo, it_o, err_o, _, _ = oracle_mod.cg(off, col, val, bt.cpu().numpy(), np.zeros(3 * nn, np.float32), tol=1e-6,
                                          atol=0.0, maxiter=5000, prec_kind=prec_kind, prec=prec, group=g)
    xh = x.cpu().numpy()
    r = oracle_mod.residual_f32_f64(off, col, val, xh.astype(np.float64), bt.cpu().numpy().astype(np.float64))
    rel = float(np.linalg.norm(r) / np.linalg.norm(bt.cpu().numpy().astype(np.float64)))
    dx = float(np.linalg.norm(xh - xo) / np.linalg.norm(xo))
    _say(f"[scale] CG {ptype} 32^3 elasticity: {it} iterations (oracle {it_o}), |r| {err:.3e} <= atol {atol:.3e}; "
         f"FP64 true residual {rel:.2e}; solution vs oracle {dx:.2e}")
    assert it == it_o
    assert err <= atol
    assert dx <= 1e-5


def test_cg_wide_spmv_grid_vs_oracle(device, oracle_mod):
    """Jacobi-CG on the 208^3 7-point variable-coefficient operator (8,998,912
    rows: more SpMV tiles than the persistent grid takes, so each product runs
    one workgroup per tile and its partials are reduced in two fixed-order
    steps): 40 iterations match the oracle's 40 in the kernels' SpMV order to
    FP32 rounding of the reductions, and the recursive residual agrees."""
    n = 208
    N = n**3
    off, col, val, b, _ = problems.fv7_problem(n, 2.3, 16.0 * n / 256, 5, 6, device=device)
    A = sparse.bsr_zeros(N, N, wa.float32, device=device)
    A.offsets, A.columns, A.values, A.nnz = off, col, val, col.numel()
    nnz = int(col.numel())
    g = _native.lib().warp_amd_bsr_mv_lanes(N, 1, 1, nnz)
    assert N > 32768 * (256 // g) * 4, "the system must take the one-workgroup-per-tile SpMV"
    bt = b.to(torch.float32).contiguous()
    M = linear.preconditioner(A, "diag")
    x = torch.zeros_like(bt)
    iters = 40
    it, err, _ = linear.cg(A, bt, x, tol=1e-30, atol=0.0, maxiter=iters, M=M)
    offh, colh, valh = off.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()
    prec = M._prec_data.cpu().numpy()
    xo, it_o, err_o, _, _ = oracle_mod.cg(offh, colh, valh, bt.cpu().numpy(), np.zeros(N, np.float32), tol=1e-30,
                                          atol=0.0, maxiter=iters, prec_kind=oracle_mod.PREC_DIAG, prec=prec, group=g)
    xh = x.cpu().numpy()
    dx = float(np.linalg.norm(xh - xo) / np.linalg.norm(xo))
    _say(f"[scale] CG 208^3 7-point (wide SpMV grid, lanes {g}): {it} iterations (oracle {it_o}), |r| {err:.4e} "
         f"(oracle {err_o:.4e}); solution vs oracle {dx:.2e}")
    assert it == it_o == iters
    assert abs(err - err_o) <= 1e-3 * err_o
    assert dx <= 1e-4
