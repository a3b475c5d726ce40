"""Host-side logic of the drop-in API (CPU only): validation messages of the
reference, no CPU fallback, workload generators."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import warp_amd as wa
from warp_amd import problems, sparse
from warp_amd.optim import linear


def test_types():
    assert wa.mat33._shape_ == (3, 3)
    assert wa.types.type_scalar_type(wa.mat33d) is wa.float64
    assert wa.types.type_size(wa.matrix((3, 2), wa.float32)) == 6
    assert wa.types.type_to_scalar(float) is wa.float32


def test_triplets_validation_messages():
    A = sparse.bsr_zeros(4, 9, wa.matrix((3, 2), wa.float32), device="cpu")
    rows = torch.zeros(5, dtype=torch.float32)
    with pytest.raises(TypeError, match=r"Rows and columns arrays must be of type int32$"):
        sparse.bsr_set_from_triplets(A, rows, rows, torch.zeros(5, 3, 2))
    with pytest.raises(ValueError, match="same length"):
        sparse.bsr_set_from_triplets(A, torch.zeros(5, dtype=torch.int32), torch.zeros(4, dtype=torch.int32))
    with pytest.raises(ValueError, match="block shape"):
        sparse.bsr_set_from_triplets(
            A, torch.zeros(5, dtype=torch.int32), torch.zeros(5, dtype=torch.int32), torch.zeros(5, 3, 3)
        )
    # valid input on a CPU tensor: no CPU fallback
    with pytest.raises(RuntimeError, match="ROCm GPU devices only"):
        sparse.bsr_set_from_triplets(
            A, torch.zeros(5, dtype=torch.int32), torch.zeros(5, dtype=torch.int32), torch.ones(5, 3, 2)
        )
    # zero-length input resets the matrix without touching the device
    sparse.bsr_set_from_triplets(A, torch.zeros(0, dtype=torch.int32), torch.zeros(0, dtype=torch.int32))
    assert A.nnz == 0


def test_bsr_mv_validation_messages():
    A = sparse.bsr_zeros(2, 3, wa.mat33, device="cpu")
    x = torch.zeros(9)
    y = torch.zeros(6)
    A.ncol += 1
    with pytest.raises(ValueError, match="Incompatible 'x'"):
        sparse.bsr_mv(A, x, y)
    A.ncol -= 1
    A.nrow -= 1
    with pytest.raises(ValueError, match="Incompatible 'y'"):
        sparse.bsr_mv(A, x, y)
    A.nrow += 1
    with pytest.raises(RuntimeError, match="ROCm GPU devices only"):
        sparse.bsr_mv(A, x, y)


def test_preconditioner_and_solver_arguments():
    assert linear.preconditioner(None, "id") is None
    with pytest.raises(ValueError, match="Unsupported preconditioner type"):
        linear.preconditioner(torch.eye(3), "ilu")
    with pytest.raises(TypeError, match="restart must be an integer"):
        linear.cg(torch.eye(3), torch.zeros(3), torch.zeros(3), restart=True)
    with pytest.raises(ValueError, match="restart must be positive"):
        linear.cg(torch.eye(3), torch.zeros(3), torch.zeros(3), restart=0)
    with pytest.raises(ValueError, match="max_batch_length requires batch_offsets"):
        linear.LinearOperator((3, 3), wa.float32, torch.device("cpu"), None, max_batch_length=3)


def test_poisson2d_generator_counts():
    rows, cols, vals, b, n = problems.poisson2d_q1(128, 0)
    assert n == 127 * 127
    assert rows.shape[0] == 128 * 128 * 16
    ok = (rows >= 0) & (cols >= 0)
    keys = np.unique(rows[ok].astype(np.int64) * n + cols[ok])
    assert keys.size == (3 * 127 - 2) ** 2 == 143641


def test_hex_stiffness_properties():
    K = problems.hex8_elasticity_stiffness(1.0 / 64)
    assert np.allclose(K, K.T)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-10 * ev.max()) == 6  # 3 translations + 3 rotations
    assert ev.min() > -1e-12 * ev.max()
    # rigid motions are in the null space
    corners = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], dtype=float)
    for t in np.eye(3):
        assert np.abs(K @ np.tile(t, 8)).max() < 1e-10 * np.abs(K).max()
    rot = np.cross(np.array([0.0, 0.0, 1.0]), corners).reshape(-1)
    assert np.abs(K @ rot).max() < 1e-10 * np.abs(K).max()


def test_elasticity_generator_small():
    rows, cols, vals, b, nn = problems.elasticity3d_q1(4, device="cpu")
    assert nn == 125
    assert rows.shape[0] == 64 * 4**3 + 25
    keys = torch.unique(rows.to(torch.int64) * nn + cols.to(torch.int64))
    assert keys.numel() == (3 * 4 + 1) ** 3
    assert torch.all(b[torch.arange(0, nn, 5)] == 0)


def test_fv7_operator_properties():
    k = torch.exp(problems.gaussian_random_field(8, 1.5, 2.0, seed=1))
    assert abs(float(torch.log(k).std()) - 1.5) < 1e-3
    off, col, val = problems.fv7_poisson(k)
    n = 8
    assert off[-1].item() == 7 * n**3 - 6 * n**2
    import scipy.sparse as sp

    A = sp.csr_matrix((val.numpy().astype(np.float64), col.numpy(), off.numpy()), shape=(n**3, n**3))
    assert abs(A - A.T).max() < 1e-6 * abs(A).max()
    # strictly diagonally dominant on boundary rows, weakly inside: SPD
    assert np.linalg.eigvalsh(A.toarray()).min() > 0
    assert np.all(np.asarray(A.sum(axis=1)).ravel() >= -1e-5)


def test_grf_correlation_length():
    """The field's covariance is exp(-r^2 / (2 l^2)): the empirical
    correlation along each axis, averaged over four 128^3 fields with l = 8
    cells, is exp(-1/8) at lag l/2, exp(-1/2) at lag l and exp(-2) at lag 2l
    (a filter without the square root would give exp(-1/4) = 0.78 at lag l,
    a correlation length off by 2x exp(-1/32) or exp(-1/8) there)."""
    n, ell = 128, 8.0
    corr = {4: [], 8: [], 16: []}
    for seed in range(4):
        g = problems.gaussian_random_field(n, 1.0, ell, seed=100 + seed).to(torch.float64)
        assert abs(float(g.std()) - 1.0) < 1e-6 and abs(float(g.mean())) < 1e-6
        var = float((g * g).mean())
        for lag in corr:
            for axis in range(3):
                corr[lag].append(float((g * torch.roll(g, lag, dims=axis)).mean()) / var)
    for lag, vals in corr.items():
        want = math.exp(-0.5 * (lag / ell) ** 2)
        got = float(np.mean(vals))
        assert abs(got - want) < 0.04, (lag, got, want)


def test_learned_checkpoint_roundtrip(tmp_path):
    from warp_amd.optim import learned

    net = learned.UNet(32, "cpu", seed=3)
    net.m["e0.W"].fill_(0.25)
    net.t = 7
    meta = dict(n=32, sigma=1.5, corr=8.0, smooth_steps=1, omega=2.0 / 3.0)
    path = str(tmp_path / "ckpt.npz")
    learned.save_checkpoint(path, net, meta)
    back, meta2 = learned.load_checkpoint(path, 32, "cpu")
    assert set(back.params) == set(net.params)
    for k in net.params:
        assert torch.equal(back.params[k], net.params[k])
    assert torch.equal(back.m["e0.W"], net.m["e0.W"]) and back.t == 7
    assert meta2["smooth_steps"] == 1 and meta2["sigma"] == 1.5


def test_padded_bsr_zeros_and_uncompress_rows():
    from warp_amd import sparse

    A = sparse.bsr_zeros(3, 4, wa.float32, device="cpu", row_capacity=2)
    np.testing.assert_array_equal(A.offsets.numpy(), [0, 2, 4, 6])
    np.testing.assert_array_equal(A.row_counts.numpy(), [0, 0, 0])
    assert A.nnz == 6 and (A.columns.numpy() == -1).all()
    B = sparse.bsr_zeros(3, 4, wa.float32, device="cpu", row_capacity=torch.tensor([1, 0, 3], dtype=torch.int32))
    np.testing.assert_array_equal(B.offsets.numpy(), [0, 1, 1, 4])
    B.row_counts = torch.tensor([1, 0, 2], dtype=torch.int32)
    np.testing.assert_array_equal(B.uncompress_rows().numpy(), [0, 2, 2, -1])
    with pytest.raises(ValueError, match="row_capacity requires topology='padded'"):
        sparse.bsr_zeros(3, 4, wa.float32, device="cpu", topology="compact", row_capacity=2)
    with pytest.raises(ValueError, match="row_capacity must be nonnegative"):
        sparse.bsr_zeros(3, 4, wa.float32, device="cpu", row_capacity=-1)
    with pytest.raises(TypeError, match="row_capacity array must have dtype int32"):
        sparse.bsr_zeros(3, 4, wa.float32, device="cpu", row_capacity=torch.tensor([1, 2, 3]))
    with pytest.raises(ValueError, match="nnz_capacity must be nonnegative"):
        sparse.bsr_zeros(3, 4, wa.float32, device="cpu", nnz_capacity=-2)
    with pytest.raises(ValueError, match="Unsupported topology policy"):
        sparse.bsr_compress(A, topology="banded")
    with pytest.raises(ValueError, match="requires a concrete BsrMatrix"):
        sparse.bsr_compress(A * 2.0)
    with pytest.raises(RuntimeError, match="ROCm GPU devices only"):
        sparse.bsr_compress(A)


def test_row_and_block_index_helpers():
    # Known answers for sparse.py:1794-1900 on a 3-row matrix with an empty middle row.
    off = torch.tensor([0, 2, 2, 5], dtype=torch.int32)
    cols = torch.tensor([0, 3, 1, 2, 4], dtype=torch.int32)
    assert [sparse.bsr_row_index(off, 3, i) for i in range(-1, 7)] == [-1, 0, 0, 2, 2, 2, -1, -1]
    pairs = [(0, 0), (0, 3), (0, 1), (1, 0), (2, 1), (2, 2), (2, 4), (2, 3), (-1, 0)]
    assert [sparse.bsr_block_index(r, c, off, cols) for r, c in pairs] == [0, 1, -1, -1, 2, 3, 4, -1, -1]
    # Capacity-aware overloads ignore slack positions past row_counts.
    rc = torch.tensor([1, 0, 2], dtype=torch.int32)
    np.testing.assert_array_equal(sparse.bsr_row_index(off, 3, torch.arange(6), rc).numpy(), [0, -1, 2, 2, -1, -1])
    got = sparse.bsr_block_index(torch.tensor([0, 0, 2, 2]), torch.tensor([0, 3, 2, 4]), off, cols, rc)
    np.testing.assert_array_equal(got.numpy(), [0, -1, 3, -1])
    # Brute force on a random compact pattern: every (row, col) pair.
    g = torch.Generator().manual_seed(3)
    nrow, ncol = 17, 13
    mask = torch.rand(nrow, ncol, generator=g) < 0.3
    counts = mask.sum(1)
    off = torch.zeros(nrow + 1, dtype=torch.int32)
    off[1:] = torch.cumsum(counts, 0)
    cols = torch.nonzero(mask)[:, 1].to(torch.int32)
    rr, cc = torch.meshgrid(torch.arange(nrow), torch.arange(ncol), indexing="ij")
    got = sparse.bsr_block_index(rr.flatten(), cc.flatten(), off, cols)
    expect = torch.full((nrow * ncol,), -1, dtype=torch.int32)
    expect[mask.flatten()] = torch.arange(int(counts.sum()), dtype=torch.int32)
    np.testing.assert_array_equal(got.numpy(), expect.numpy())
    rows = sparse.bsr_row_index(off, nrow, torch.arange(int(counts.sum())))
    np.testing.assert_array_equal(rows.numpy(), torch.nonzero(mask)[:, 0].numpy())
    assert sparse.bsr_block_index(0, 0, torch.zeros(3, dtype=torch.int32), torch.zeros(0, dtype=torch.int32)) == -1


def test_public_names_match_reference_module():
    names = set(sparse.__all__)
    for name in names:
        assert hasattr(sparse, name), name
    assert sparse.bsr_matrix_t(wa.mat33) is sparse.BsrMatrix
    assert sparse.bsr_matrix_t(wa.float64) is sparse.BsrMatrix
    with pytest.raises(ValueError, match="matrix or scalar"):
        sparse.bsr_matrix_t(str)


def test_batched_host_loop_per_batch_tolerance():
    """test_batched_host_loop_per_batch_tolerance (test_linear_solvers.py:423-448):
    the loop stops when every subproblem meets its own tolerance."""
    r_norm_sq = torch.tensor([25.0, 4.0])
    atol_sq = torch.tensor([100.0, 1.0])
    cycles = []

    def do_cycle():
        cycles.append(1)
        r_norm_sq.copy_(torch.tensor([25.0, 0.25]))

    niter, err, atol = linear._host_loop(do_cycle, 1, r_norm_sq, atol_sq, 3, None, 1)
    assert niter == 1 and len(cycles) == 1 and err <= atol
