"""GPU parity of the sparse path against the CPU oracle (through the C ABI).

Structure (offsets, columns) and summed values are compared bit-exactly:
both sides sum duplicates sequentially in ascending triplet order.  bsr_mv
is compared to the oracle's row-order product within 1e-5 relative (FP32;
summation order differs) and to a dense FP64 product.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import warp_amd as wa
from warp_amd import problems, sparse

pytestmark = pytest.mark.gpu


def _t(a, device, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), device=device, dtype=dtype)


def _assemble(device, nrow, ncol, rows, cols, vals, prune=True, count=None):
    if vals is None:
        A = sparse.bsr_zeros(nrow, ncol, wa.float32, device=device)
    elif vals.ndim == 3:
        A = sparse.bsr_zeros(nrow, ncol, wa.matrix(vals.shape[1:], wa.types.type_to_scalar(_t(vals[:1], "cpu").dtype)), device=device)
    else:
        A = sparse.bsr_zeros(nrow, ncol, wa.types.type_to_scalar(_t(vals[:1], "cpu").dtype), device=device)
    cnt = None if count is None else torch.tensor([count], dtype=torch.int32, device=device)
    sparse.bsr_set_from_triplets(
        A, _t(rows, device, torch.int32), _t(cols, device, torch.int32), None if vals is None else _t(vals, device),
        count=cnt, prune_numerical_zeros=prune,
    )
    return A


def _check_against_oracle(oracle_mod, A, nrow, ncol, rows, cols, vals, prune=True, count=None):
    o = oracle_mod.bsr_from_triplets(nrow, ncol, rows, cols, vals, prune=prune, count=count)
    nnz = A.nnz_sync()
    assert nnz == o["nnz"]
    np.testing.assert_array_equal(A.offsets.cpu().numpy(), o["offsets"])
    np.testing.assert_array_equal(A.columns[:nnz].cpu().numpy(), o["columns"])
    if vals is not None:
        got = A.values[:nnz].cpu().numpy()
        assert got.tobytes() == o["values"].tobytes(), "summed values differ bitwise"


CASES = [
    # nrow, ncol, T, block shape, dtype, zero frac, out-of-range frac
    (8, 6, 100, (), np.float32, 0.0, 0.0),
    (4, 9, 50, (3, 2), np.float32, 0.1, 0.0),
    (300, 200, 20000, (3, 3), np.float32, 0.05, 0.1),
    (1000, 1000, 50000, (), np.float64, 0.02, 0.05),
    (50, 50, 3000, (3, 3), np.float64, 0.0, 0.0),
    (17, 19, 4096, (4, 4), np.float32, 0.1, 0.0),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_triplets_random(device, oracle_mod, case):
    nrow, ncol, T, bshape, dtype, zf, of = CASES[case]
    print(f"[triplets] case {case}: nrow={nrow} ncol={ncol} T={T} block={bshape}", flush=True)
    rng = np.random.default_rng(100 + case)
    rows = rng.integers(0, nrow, T).astype(np.int32)
    cols = rng.integers(0, ncol, T).astype(np.int32)
    oob = rng.random(T) < of
    rows[oob & (rng.random(T) < 0.5)] = -3
    cols[oob & (rng.random(T) < 0.5)] = ncol
    vals = rng.standard_normal((T, *bshape)).astype(dtype)
    vals[rng.random(T) < zf] = 0.0
    A = _assemble(device, nrow, ncol, rows, cols, vals)
    _check_against_oracle(oracle_mod, A, nrow, ncol, rows, cols, vals)
    # and without pruning
    A = _assemble(device, nrow, ncol, rows, cols, vals, prune=False)
    _check_against_oracle(oracle_mod, A, nrow, ncol, rows, cols, vals, prune=False)


def test_triplets_golden(device, oracle_mod):
    import os

    data = np.load(os.path.join(os.path.dirname(__file__), "golden", "triplets_ref.npz"), allow_pickle=False)
    names = sorted({k.split("/")[0] for k in data.files})
    for name in names:
        print(f"[triplets] golden {name}", flush=True)
        nrow, ncol, T, count = (int(v) for v in data[f"{name}/shape"])
        count = None if count < 0 else count
        rows, cols, vals = data[f"{name}/rows"], data[f"{name}/cols"], data[f"{name}/vals"]
        A = _assemble(device, nrow, ncol, rows, cols, vals, count=count)
        nnz = A.nnz_sync()
        np.testing.assert_array_equal(A.offsets.cpu().numpy(), data[f"{name}/offsets"], err_msg=name)
        np.testing.assert_array_equal(A.columns[:nnz].cpu().numpy(), data[f"{name}/columns"], err_msg=name)
        _check_against_oracle(oracle_mod, A, nrow, ncol, rows, cols, vals, count=count)


@pytest.mark.parametrize("per_row", [300, 5000, 20000])
def test_triplets_long_rows(device, oracle_mod, per_row):
    """Rows beyond one wave (256) and beyond one workgroup's LDS (4096)."""
    print(f"[triplets] long rows {per_row}", flush=True)
    rng = np.random.default_rng(per_row)
    nrow, ncol = 5, 100000
    rows = np.repeat(np.arange(nrow, dtype=np.int32), per_row)
    rows[::7] = 2  # one row much longer
    cols = rng.integers(0, ncol // 50 if per_row > 1000 else ncol, rows.size).astype(np.int32)
    perm = rng.permutation(rows.size)
    rows, cols = rows[perm], cols[perm]
    vals = rng.standard_normal(rows.size).astype(np.float32)
    A = _assemble(device, nrow, ncol, rows, cols, vals)
    _check_against_oracle(oracle_mod, A, nrow, ncol, rows, cols, vals)


ROW_CLASS_SIZES = [0, 1, 2, 63, 64, 65, 66, 100, 127, 128, 129, 200, 255, 256, 257, 300, 511, 512, 513, 600]


@pytest.mark.parametrize("bshape,dtype", [((), np.float32), ((3, 3), np.float32), ((2, 2), np.float64)])
def test_triplets_row_size_classes(device, oracle_mod, bshape, dtype):
    """Rows on both sides of each sort-path boundary (one key per lane: 64
    keys; two, four and eight keys per lane: 128, 256 and 512), heavy and no
    duplication, in
    shuffled order; then the transpose of the result, whose rows cross the
    same boundaries."""
    rng = np.random.default_rng(4242 + len(bshape))
    ncol = 4000
    rows, cols = [], []
    for r, n in enumerate(ROW_CLASS_SIZES * 2):
        rows.append(np.full(n, r, np.int32))
        # first copy of the list: about three triplets per unique column; second: all unique
        span = max(1, n // 3) if r < len(ROW_CLASS_SIZES) else ncol
        cols.append(rng.choice(span, size=n, replace=span < n).astype(np.int32))
    rows, cols = np.concatenate(rows), np.concatenate(cols)
    perm = rng.permutation(rows.size)
    rows, cols = rows[perm], cols[perm]
    nrow = 2 * len(ROW_CLASS_SIZES)
    vals = rng.standard_normal((rows.size, *bshape)).astype(dtype)
    print(f"[triplets] row classes block={bshape} dtype={np.dtype(dtype).name} T={rows.size}", flush=True)
    A = _assemble(device, nrow, ncol, rows, cols, vals)
    _check_against_oracle(oracle_mod, A, nrow, ncol, rows, cols, vals)
    A = _assemble(device, nrow, ncol, rows, cols, None)
    _check_against_oracle(oracle_mod, A, nrow, ncol, rows, cols, None)
    # columns of a transpose source holding up to ~300 entries each
    tr = np.repeat(np.arange(300, dtype=np.int32), 8)
    tc = rng.integers(0, 12, tr.size).astype(np.int32)
    tc[: ROW_CLASS_SIZES[-1]] = 11
    tv = rng.standard_normal((tr.size, *bshape)).astype(dtype)
    S = _assemble(device, 300, 12, tr, tc, tv)
    nnz = S.nnz_sync()
    off, scol = S.offsets.cpu().numpy(), S.columns[:nnz].cpu().numpy()
    dest = sparse.bsr_transposed(S)
    ref = oracle_mod.bsr_transpose(300, 12, off, scol)
    n = dest.nnz_sync()
    assert n == ref["columns"].size
    np.testing.assert_array_equal(dest.offsets.cpu().numpy(), ref["offsets"])
    np.testing.assert_array_equal(dest.columns.cpu().numpy()[:n], ref["columns"])
    sv = S.values[:nnz].cpu().numpy()
    want = sv[ref["block_map"]]
    if want.ndim == 3:
        want = want.transpose(0, 2, 1)
    np.testing.assert_array_equal(dest.values.cpu().numpy()[:n], want)


def test_triplets_known_answers(device):
    # test_bsr_stale_compact_tail (test_sparse.py:225-252)
    A = sparse.bsr_from_triplets(
        2, 2, _t([0, 0, 1, 1], device, torch.int32), _t([0, 0, 1, 1], device, torch.int32),
        _t(np.array([1.0, 2.0, 3.0, 4.0], np.float32), device),
    )
    assert A.nnz == 4
    np.testing.assert_array_equal(A.offsets.cpu().numpy(), [0, 1, 2])
    np.testing.assert_array_equal(A.uncompress_rows().cpu().numpy(), [0, 1, -1, -1])
    copied = sparse.bsr_copy(A, scalar_type=wa.float64)   # before the sync: stale upper bound
    assert A.nnz_sync() == 2
    assert A.columns.numel() == 4 and A.values.shape[0] == 4
    np.testing.assert_array_equal(A.values[:2].cpu().numpy(), [3.0, 7.0])
    assert copied.nnz_sync() == 2
    np.testing.assert_array_equal(copied.columns[:2].cpu().numpy(), [0, 1])
    np.testing.assert_array_equal(copied.values[:2].cpu().numpy(), np.array([3.0, 7.0]))
    assert copied.values.dtype == torch.float64
    # topology only (test_sparse.py:164-172)
    T = sparse.bsr_zeros(2, 3, wa.float32, device=device)
    sparse.bsr_set_from_triplets(T, _t([1, 0, 1], device, torch.int32), _t([2, 1, 2], device, torch.int32), values=None)
    assert T.nnz_sync() == 2
    assert T.row_counts is None
    np.testing.assert_array_equal(T.columns[:2].cpu().numpy(), [1, 2])
    # pruning (test_sparse.py:199-222)
    r, c = _t([1, 0, 2, 3], device, torch.int32), _t([0, 1, 2, 3], device, torch.int32)
    z = torch.zeros(4, device=device)
    assert sparse.bsr_from_triplets(12, 12, r, c, z, prune_numerical_zeros=False).nnz_sync() == 4
    assert sparse.bsr_from_triplets(12, 12, r, c, z, prune_numerical_zeros=True).nnz_sync() == 0
    neg = torch.full((4,), -0.0, device=device)
    assert sparse.bsr_from_triplets(12, 12, r, c, neg).nnz_sync() == 0


def test_triplets_count_and_reference_named_entry(device, oracle_mod):
    """`count` limits the triplets; the reference-named entry point builds the same topology."""
    import ctypes

    from warp_amd import _native

    rng = np.random.default_rng(77)
    nrow, ncol, T = 64, 64, 5000
    rows = rng.integers(0, nrow, T).astype(np.int32)
    cols = rng.integers(0, ncol, T).astype(np.int32)
    vals = rng.standard_normal((T, 3, 3)).astype(np.float32)
    A = _assemble(device, nrow, ncol, rows, cols, vals, count=1234)
    _check_against_oracle(oracle_mod, A, nrow, ncol, rows, cols, vals, count=1234)

    lib = _native.lib()
    d_rows, d_cols = _t(rows, device, torch.int32), _t(cols, device, torch.int32)
    off = torch.zeros(nrow + 1, dtype=torch.int32, device=device)
    colsout = torch.zeros(T, dtype=torch.int32, device=device)
    so = torch.zeros(T, dtype=torch.int32, device=device)
    si = torch.zeros(T, dtype=torch.int32, device=device)
    nnz_h = torch.zeros(1, dtype=torch.int32, pin_memory=True)
    lib.warp_amd_set_current_stream(torch.cuda.current_stream(device).cuda_stream)
    lib.warp_amd_clear_error()  # the entry returns void, like the reference: errors land here
    lib.wp_bsr_matrix_from_triplets_device(
        1, 0, nrow, ncol, T, None, d_rows.data_ptr(), d_cols.data_ptr(), None, 0, False,
        so.data_ptr(), si.data_ptr(), off.data_ptr(), None, colsout.data_ptr(), nnz_h.data_ptr(), None,
    )
    torch.cuda.synchronize()
    assert lib.warp_amd_last_error() in (b"", None)
    o = oracle_mod.bsr_from_triplets(nrow, ncol, rows, cols, None)
    assert int(nnz_h[0]) == o["nnz"]
    np.testing.assert_array_equal(off.cpu().numpy(), o["offsets"])
    np.testing.assert_array_equal(colsout[: o["nnz"]].cpu().numpy(), o["columns"])
    np.testing.assert_array_equal(so[: o["nnz"]].cpu().numpy(), o["summed_offsets"])
    np.testing.assert_array_equal(si[: int(o["summed_offsets"][-1])].cpu().numpy(), o["summed_indices"])


@pytest.mark.parametrize("bshape,dtype,prune", [((3, 3), np.float32, True), ((3, 3), np.float64, False),
                                                 ((), np.float32, True), ((4, 4), np.float32, True)])
def test_triplets_value_sums_with_summed_arrays(device, oracle_mod, bshape, dtype, prune):
    """Value sums through the C ABI with the caller's summed offsets / triplet
    indices asked for: columns, bitwise sums and both summed arrays equal the
    oracle's; a device count below T, out-of-range rows and zero blocks
    included."""
    from warp_amd import _native

    rng = np.random.default_rng(5150 + len(bshape) + (dtype == np.float64))
    nrow, ncol, T, count = 3000, 2500, 60000, 57001
    rows = rng.integers(0, nrow, T).astype(np.int32)
    cols = rng.integers(0, ncol, T).astype(np.int32)
    rows[rng.random(T) < 0.02] = nrow + 5  # out of range: dropped
    vals = rng.standard_normal((T, *bshape)).astype(dtype)
    vals[rng.random(T) < 0.05] = 0.0        # zero blocks: pruned when asked
    bsz = int(np.prod(bshape)) if bshape else 1
    sb = np.dtype(dtype).itemsize
    mask = {4: 0x7FFFFFFF, 8: 0x7FFFFFFFFFFFFFFF}[sb] if prune else 0
    o = oracle_mod.bsr_from_triplets(nrow, ncol, rows, cols, vals, prune=prune, count=count)
    lib = _native.lib()
    d_rows, d_cols, d_vals = _t(rows, device, torch.int32), _t(cols, device, torch.int32), _t(vals, device)
    d_count = torch.tensor([count], dtype=torch.int32, device=device)
    ws_bytes = lib.warp_amd_bsr_from_triplets_workspace_size(nrow, T)
    off = torch.zeros(nrow + 1, dtype=torch.int32, device=device)
    colsout = torch.full((T,), -1, dtype=torch.int32, device=device)
    so = torch.full((T,), -1, dtype=torch.int32, device=device)
    si = torch.full((T,), -1, dtype=torch.int32, device=device)
    out = torch.full((T, *bshape), float("nan"), dtype=d_vals.dtype, device=device)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=device)
    print(f"[triplets] C ABI block={bshape} dtype={np.dtype(dtype).name} prune={prune}", flush=True)
    _native.call(
        "warp_amd_bsr_from_triplets", bsz, sb, nrow, ncol, T, _native.ptr(d_count), _native.ptr(d_rows),
        _native.ptr(d_cols), _native.ptr(d_vals), mask, 0, _native.ptr(so), _native.ptr(si), _native.ptr(off),
        None, _native.ptr(colsout), _native.ptr(out), None, _native.ptr(ws), ws_bytes, _native.stream_of(device),
    )
    torch.cuda.synchronize()
    nnz = o["nnz"]
    valid = int(o["summed_offsets"][-1])
    np.testing.assert_array_equal(off.cpu().numpy(), o["offsets"])
    np.testing.assert_array_equal(colsout[:nnz].cpu().numpy(), o["columns"])
    np.testing.assert_array_equal(so[:nnz].cpu().numpy(), o["summed_offsets"])
    np.testing.assert_array_equal(si[:valid].cpu().numpy(), o["summed_indices"])
    assert out[:nnz].cpu().numpy().tobytes() == o["values"].tobytes(), "summed values differ bitwise"


def test_triplets_elasticity_pattern(device, oracle_mod):
    """Element-loop triplets (8x8 blocks per hex), shuffled, random values: the assembly workload at n=8."""
    rows, cols, vals, nn = problems.elasticity_triplets_shuffled(8, device="cpu")
    rows, cols, vals = rows.numpy(), cols.numpy(), vals.numpy()
    A = _assemble(device, nn, nn, rows, cols, vals)
    _check_against_oracle(oracle_mod, A, nn, nn, rows, cols, vals)
    assert A.nnz == (3 * 8 + 1) ** 3


def _lanes(A):
    from warp_amd import _native

    br, bc = A.block_shape
    return _native.lib().warp_amd_bsr_mv_lanes(A.nrow, br, bc, int(A.nnz))


def _dense_of(A):
    nnz = A.nnz_sync()
    br, bc = A.block_shape
    off = A.offsets.cpu().numpy()
    col = A.columns[:nnz].cpu().numpy()
    val = A.values[:nnz].cpu().numpy().astype(np.float64).reshape(nnz, br, bc)
    D = np.zeros((A.nrow * br, A.ncol * bc))
    for r in range(A.nrow):
        for k in range(off[r], off[r + 1]):
            D[r * br : (r + 1) * br, col[k] * bc : (col[k] + 1) * bc] += val[k]
    return D


@pytest.mark.parametrize(
    "bshape,dtype", [((1, 1), torch.float32), ((1, 3), torch.float32), ((3, 3), torch.float64), ((3, 3), torch.float32), ((3, 2), torch.float32)]
)
def test_bsr_mv_reference_scenario(device, bshape, dtype):
    """make_test_bsr_mv (test_sparse.py:830-903), including the transposed product y @ (alpha A)."""
    rng = np.random.default_rng(123)
    nrow, ncol, nnz = 2, 3, 6
    rows = _t(rng.integers(0, nrow, nnz).astype(np.int32), device)
    cols = _t(rng.integers(0, ncol, nnz).astype(np.int32), device)
    vals = _t(rng.random((nnz, *bshape)), device, dtype)
    block = wa.matrix(bshape, wa.types.type_to_scalar(dtype))
    A = sparse.bsr_zeros(nrow, ncol, block, device=device)
    sparse.bsr_set_from_triplets(A, rows, cols, vals)
    D = _dense_of(A)
    x = _t(rng.random(ncol * bshape[1]), device, dtype)
    y = _t(rng.random(nrow * bshape[0]), device, dtype)
    work = torch.empty_like(y)
    for alpha, beta in zip([-1.0, 0.0, 1.0], [2.0, -1.0, 0.0]):
        ref = alpha * D @ x.cpu().numpy().astype(np.float64) + beta * y.cpu().numpy().astype(np.float64)
        if beta == 0.0:
            y = A @ x
        else:
            sparse.bsr_mv(A, x, y, alpha, beta, work_buffer=work)
        np.testing.assert_allclose(y.cpu().numpy().reshape(-1), ref, atol=1e-4)
    # transposed product (test_sparse.py:878-881): x = y @ (A * alpha) = alpha A^T y
    alpha = -1.0
    ref = alpha * y.cpu().numpy().astype(np.float64).reshape(-1) @ D
    xt = y.reshape(-1) @ (A * alpha)
    np.testing.assert_allclose(xt.cpu().numpy().reshape(-1), ref, rtol=1e-4, atol=1e-4)
    # aliasing x == y on a square product A A^T
    D2 = D @ D.T
    r2, c2 = np.nonzero(np.ones((nrow, nrow)))
    blocks = np.stack([D2[i * bshape[0] : (i + 1) * bshape[0], j * bshape[0] : (j + 1) * bshape[0]] for i, j in zip(r2, c2)])
    AAt = sparse.bsr_zeros(nrow, nrow, wa.matrix((bshape[0], bshape[0]), wa.types.type_to_scalar(dtype)), device=device)
    sparse.bsr_set_from_triplets(AAt, _t(r2.astype(np.int32), device), _t(c2.astype(np.int32), device), _t(blocks, device, dtype))
    yy = y.clone().reshape(-1)
    ref = -1.0 * D2 @ yy.cpu().numpy().astype(np.float64) + 2.0 * yy.cpu().numpy().astype(np.float64)
    sparse.bsr_mv(AAt, yy, yy, -1.0, 2.0)
    np.testing.assert_allclose(yy.cpu().numpy(), ref, rtol=1e-4, atol=1e-4)
    # shape errors
    A.ncol += 1
    with pytest.raises(ValueError, match="Incompatible 'x'"):
        sparse.bsr_mv(A, x, y)
    A.ncol -= 1
    A.nrow -= 1
    with pytest.raises(ValueError, match="Incompatible 'y'"):
        sparse.bsr_mv(A, x, y)


@pytest.mark.parametrize("n,dtype", [(32, torch.float32), (128, torch.float32), (48, torch.float64)])
def test_bsr_mv_poisson_vs_oracle(device, oracle_mod, n, dtype):
    rows, cols, vals, b, nd = problems.poisson2d_q1(n, 0)
    vals = vals.astype(np.float64 if dtype == torch.float64 else np.float32)
    A = _assemble(device, nd, nd, rows, cols, vals)
    o = oracle_mod.bsr_from_triplets(nd, nd, rows, cols, vals)
    x = np.random.default_rng(1).standard_normal(nd).astype(vals.dtype)
    y0 = np.random.default_rng(2).standard_normal(nd).astype(vals.dtype)
    for alpha, beta in [(1.0, 0.0), (-1.0, 1.0), (0.5, -2.0), (0.0, 3.0)]:
        ref = oracle_mod.bsr_mv(o["offsets"], o["columns"], o["values"].reshape(-1, 1, 1), x, y0, alpha, beta)
        yg = _t(y0, device)
        sparse.bsr_mv(A, _t(x, device), yg, alpha, beta)
        got = yg.cpu().numpy()
        scale = np.abs(alpha) * oracle_mod.bsr_mv(o["offsets"], o["columns"], np.abs(o["values"]).reshape(-1, 1, 1), np.abs(x)) + np.abs(beta * y0)
        assert np.all(np.abs(got - ref) <= 1e-5 * scale + 1e-30)
        # bit-exact against the oracle run in the kernel's lane-group order
        G = _lanes(A)
        ref_g = oracle_mod.bsr_mv(o["offsets"], o["columns"], o["values"].reshape(-1, 1, 1), x, y0, alpha, beta, group=G)
        assert got.tobytes() == ref_g.tobytes(), f"G={G}"


@pytest.mark.parametrize("n", [6, 16])
def test_bsr_mv_elasticity_vs_oracle(device, oracle_mod, n):
    rows, cols, vals, b, nn = problems.elasticity3d_q1(n, device="cpu")
    rows, cols, vals = rows.numpy(), cols.numpy(), vals.numpy()
    A = _assemble(device, nn, nn, rows, cols, vals, prune=False)
    o = oracle_mod.bsr_from_triplets(nn, nn, rows, cols, vals, prune=False)
    x = np.random.default_rng(3).standard_normal(3 * nn).astype(np.float32)
    ref = oracle_mod.bsr_mv(o["offsets"], o["columns"], o["values"], x)
    scale = oracle_mod.bsr_mv(o["offsets"], o["columns"], np.abs(o["values"]), np.abs(x))
    got = sparse.bsr_mv(A, _t(x, device)).cpu().numpy().reshape(-1)
    assert np.all(np.abs(got - ref) <= 1e-5 * scale)
    ref_g = oracle_mod.bsr_mv(o["offsets"], o["columns"], o["values"], x, group=_lanes(A))
    assert got.tobytes() == ref_g.tobytes()
    # FP64 error of the FP32 product, stated
    ref64 = oracle_mod.bsr_mv(o["offsets"], o["columns"], o["values"].astype(np.float64), x.astype(np.float64))
    rel = np.linalg.norm(got - ref64) / np.linalg.norm(ref64)
    print(f"[bsr_mv 3x3 n={n}] rel. error vs FP64 = {rel:.3e}")
    assert rel < 1e-6


def test_bsr_get_diag_and_jacobi(device, oracle_mod):
    rng = np.random.default_rng(123)
    rows = np.array([0, 1, 2, 3, 2, 1], np.int32)
    cols = np.array([1, 1, 1, 3, 2, 2], np.int32)
    vals = rng.random((6, 3, 3)).astype(np.float32)
    A = _assemble(device, 4, 4, rows, cols, vals)
    d = sparse.bsr_get_diag(A).cpu().numpy()
    np.testing.assert_array_equal(d[0], np.zeros((3, 3)))
    np.testing.assert_array_equal(d[1], vals[1])
    np.testing.assert_array_equal(d[2], vals[4])
    np.testing.assert_array_equal(d[3], vals[3])
    sentinel = torch.full((4, 3, 3), 7.0, device=device)
    sparse.bsr_get_diag(A, out=sentinel)
    np.testing.assert_array_equal(sentinel.cpu().numpy(), d)
    d2 = sparse.bsr_get_diag(2.0 * A).cpu().numpy()
    np.testing.assert_allclose(d2, 2.0 * d)

    from warp_amd.optim import linear

    M = linear.preconditioner(A, "diag")
    expect = oracle_mod.inverse_diag(oracle_mod.bsr_get_diag(*(lambda o: (o["offsets"], o["columns"], o["values"]))(oracle_mod.bsr_from_triplets(4, 4, rows, cols, vals))))
    np.testing.assert_array_equal(M._prec_data.cpu().numpy().reshape(-1), expect)
    Mb = linear.preconditioner(A, "block_diag")
    inv = Mb._prec_data.cpu().numpy()
    for r in (1, 2, 3):
        np.testing.assert_allclose(inv[r] @ d[r].astype(np.float64), np.eye(3), atol=2e-4)
    # absent diagonal block: block Jacobi falls back to the point inverse (identity here)
    np.testing.assert_array_equal(inv[0], np.eye(3))


@pytest.mark.parametrize("bshape", [(1, 1), (3, 3)])
def test_bsr_mv_every_kernel_shape(device, oracle_mod, bshape):
    """Every lanes x rows-per-group kernel writes every output and matches the
    oracle's lane-order sum bit for bit (y starts as NaN, so a skipped store fails)."""
    from warp_amd import _native

    lib = _native.lib()
    rng = np.random.default_rng(31)
    nrow, ncol = 203, 190
    lengths = rng.integers(0, 40, nrow)
    lengths[::17] = 0
    rows = np.repeat(np.arange(nrow), lengths).astype(np.int32)
    cols = rng.integers(0, ncol, rows.size).astype(np.int32)
    br, bc = bshape
    vals = rng.standard_normal((rows.size, br, bc)).astype(np.float32)
    A = _assemble(device, nrow, ncol, rows, cols, vals if br > 1 else vals.reshape(-1))
    nnz = A.nnz_sync()
    o = oracle_mod.bsr_from_triplets(nrow, ncol, rows, cols, vals)
    x = rng.standard_normal(ncol * bc).astype(np.float32)
    xt = _t(x, device)
    rpgs = [1, 2, 4, 8] if br == 1 else [1, 2, 4]
    for lanes in [1, 2, 4, 8, 16, 32, 64]:
        ref = oracle_mod.bsr_mv(o["offsets"], o["columns"], o["values"], x, group=lanes)
        for rpg in rpgs:
            y = torch.full((nrow * br,), float("nan"), device=device)
            st = lib.warp_amd_bsr_mv_ex_f32(
                nrow, ncol, br, bc, A.offsets.data_ptr(), None, A.columns.data_ptr(), A.values.data_ptr(), nnz,
                xt.data_ptr(), y.data_ptr(), 1.0, 0.0, lanes, rpg, torch.cuda.current_stream(device).cuda_stream,
            )
            assert st == 0, lib.warp_amd_last_error()
            got = y.cpu().numpy()
            assert got.tobytes() == ref.tobytes(), f"lanes={lanes} rows_per_group={rpg}"


def test_bsr_mv_graph_capture(device):
    """bsr_mv captured in a HIP graph (as test_capturability, test_sparse.py:1005-1041):
    every replay, with eager work in between, equals the eager result."""
    rows, cols, vals, b, nn = problems.elasticity3d_q1(6, device="cpu")
    A = sparse.bsr_zeros(nn, nn, wa.mat33, device=device)
    sparse.bsr_set_from_triplets(A, rows.to(device), cols.to(device), vals.to(device), prune_numerical_zeros=False)
    A.nnz_sync()
    x = torch.randn(3 * nn, device=device)
    y = torch.zeros(3 * nn, device=device)
    eager = sparse.bsr_mv(A, x).reshape(-1).clone()
    s = torch.cuda.Stream(device)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            sparse.bsr_mv(A, x, y, alpha=1.0, beta=0.0)
    for k in range(4):
        y.fill_(float(k))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, eager), f"replay {k}"
        eager2 = sparse.bsr_mv(A, x).reshape(-1)
        assert torch.equal(eager2, eager)


@pytest.mark.parametrize("bshape,dtype", [((1, 1), torch.float32), ((3, 3), torch.float32), ((3, 3), torch.float64),
                                          ((2, 3), torch.float64)])
def test_bsr_mv_transpose_vs_dense(device, bshape, dtype):
    """alpha A^T x + beta y against the dense product; the atomic scatter
    (as in the reference) changes only the last bits between runs."""
    rng = np.random.default_rng(17)
    nrow, ncol, T = 300, 200, 4000
    rows = rng.integers(0, nrow, T).astype(np.int32)
    cols = rng.integers(0, ncol, T).astype(np.int32)
    vals = rng.standard_normal((T, *bshape))
    A = sparse.bsr_zeros(nrow, ncol, wa.matrix(bshape, wa.types.type_to_scalar(dtype)), device=device)
    sparse.bsr_set_from_triplets(A, _t(rows, device), _t(cols, device), _t(vals, device, dtype))
    D = _dense_of(A)
    x = rng.standard_normal(nrow * bshape[0])
    y0 = rng.standard_normal(ncol * bshape[1])
    tol = dict(rtol=1e-4, atol=1e-4) if dtype == torch.float32 else dict(rtol=1e-10, atol=1e-10)
    for alpha, beta in [(1.0, 0.0), (-0.5, 2.0), (0.0, 3.0), (2.0, 1.0)]:
        y = _t(y0, device, dtype)
        sparse.bsr_mv(A, _t(x, device, dtype), y, alpha, beta, transpose=True)
        ref = alpha * D.T @ x + beta * y0
        np.testing.assert_allclose(y.cpu().numpy().astype(np.float64).reshape(-1), ref, **tol)
    # y = None allocates the result with A's column blocks
    out = sparse.bsr_mv(A, _t(x, device, dtype), transpose=True)
    assert tuple(out.shape) == ((ncol,) if bshape[1] == 1 else (ncol, bshape[1]))
    np.testing.assert_allclose(out.cpu().numpy().astype(np.float64).reshape(-1), D.T @ x, **tol)


def _active_dense(A):
    """Dense matrix of the active blocks (row_counts honoured)."""
    br, bc = A.block_shape
    off = A.offsets.cpu().numpy()
    rc = None if A.row_counts is None else A.row_counts.cpu().numpy()
    col = A.columns.cpu().numpy()
    val = A.values.cpu().numpy().astype(np.float64).reshape(-1, br, bc)
    D = np.zeros((A.nrow * br, A.ncol * bc))
    cols = []
    for r in range(A.nrow):
        end = off[r] + rc[r] if rc is not None else off[r + 1]
        cols.append([int(c) for c in col[off[r]:end]])
        for k in range(off[r], end):
            D[r * br:(r + 1) * br, col[k] * bc:(col[k] + 1) * bc] += val[k]
    return D, cols


def test_bsr_compress_reference_scenarios(device):
    """test_sparse.py:343-379 (padded and compact in-place compression) and 1040-1075 (trailing capacity)."""
    cand = sparse.bsr_zeros(2, 3, wa.float32, device=device, row_capacity=_t(np.array([4, 3], np.int32), device))
    cand.row_counts = _t(np.array([4, 3], np.int32), device)
    cand.columns = _t(np.array([2, 0, 2, 1, 1, 2, 1], np.int32), device)
    cand.values = _t(np.array([5.0, 2.0, -1.0, 3.0, 4.0, 8.0, -4.0], np.float32), device)
    sparse.bsr_compress(cand, inplace=True, topology="padded")
    np.testing.assert_array_equal(cand.row_counts.cpu().numpy(), [3, 2])
    D, cols = _active_dense(cand)
    assert cols == [[0, 1, 2], [1, 2]]
    np.testing.assert_array_equal(D, [[2.0, 3.0, 4.0], [0.0, 0.0, 8.0]])

    long_cols = (np.arange(160, dtype=np.int32) * 37 + 11) % 97
    long_vals = (np.arange(160, dtype=np.float32) % 11.0) - 5.0
    expect = np.zeros(97, np.float32)
    for c, v in zip(long_cols, long_vals):
        expect[c] += v
    expect_cols = np.unique(long_cols[long_vals != 0.0]).astype(np.int32)
    long_row = sparse.bsr_zeros(2, 97, wa.float32, device=device,
                                row_capacity=_t(np.array([long_cols.size, 0], np.int32), device))
    long_row.row_counts = _t(np.array([long_cols.size, 0], np.int32), device)
    long_row.columns = _t(long_cols, device)
    long_row.values = _t(long_vals, device)
    sparse.bsr_compress(long_row, inplace=True, topology="compact")
    assert long_row.row_counts is None
    nnz = long_row.nnz_sync()
    assert nnz == expect_cols.size
    np.testing.assert_array_equal(long_row.offsets.cpu().numpy(), [0, nnz, nnz])
    np.testing.assert_array_equal(long_row.columns.cpu().numpy()[:nnz], expect_cols)
    np.testing.assert_allclose(long_row.values.cpu().numpy()[:nnz], expect[expect_cols])

    for inplace in (False, True):
        m = sparse.bsr_zeros(3, 3, wa.float32, device=device)
        m.nnz = 6
        m.offsets = _t(np.array([0, 1, 2, 3], np.int32), device)
        m.columns = _t(np.array([0, 1, 2, 0, 0, 0], np.int32), device)
        m.values = _t(np.array([3.0, 7.0, 11.0, 42.0, 42.0, 42.0], np.float32), device)
        sparse.bsr_compress(m, prune_numerical_zeros=False, inplace=inplace, topology="compact")
        assert m.row_counts is None and m.nnz_sync() == 3
        np.testing.assert_array_equal(m.offsets.cpu().numpy(), [0, 1, 2, 3])
        np.testing.assert_array_equal(m.columns.cpu().numpy()[:3], [0, 1, 2])
        np.testing.assert_allclose(m.values.cpu().numpy()[:3], [3.0, 7.0, 11.0])


def test_bsr_compress_prunes_and_keeps_capacity(device):
    """test_sparse.py:441-475 without the tape: a zero block is pruned, duplicates summed."""
    for topology in ("compact", "padded"):
        m = sparse.bsr_zeros(1, 3, wa.float32, device=device)
        m.nnz = 4
        m.offsets = _t(np.array([0, 4], np.int32), device)
        m.row_counts = _t(np.array([4], np.int32), device)
        m.columns = _t(np.array([1, 0, 1, 2], np.int32), device)
        m.values = _t(np.array([1.0, 2.0, 3.0, 0.0], np.float32), device)
        sparse.bsr_compress(m, topology=topology)
        D, cols = _active_dense(m)
        np.testing.assert_array_equal(D, [[2.0, 4.0, 0.0]])
        assert cols == [[0, 1]]
        if topology == "compact":
            assert m.row_counts is None and m.nnz_sync() == 2
            np.testing.assert_array_equal(m.offsets.cpu().numpy(), [0, 2])
        else:
            np.testing.assert_array_equal(m.offsets.cpu().numpy(), [0, 4])
            np.testing.assert_array_equal(m.row_counts.cpu().numpy(), [2])


@pytest.mark.parametrize("topology", ["compact", "padded"])
def test_bsr_compress_row_candidates_vs_oracle(device, oracle_mod, topology):
    """fem 'row_compress' construction: each row's candidate blocks written into a
    padded matrix in element order, then compressed.  Structure and values are
    bit-exact against the oracle assembly of the same (row, column, value)
    candidates in storage order (a stable sort and left-to-right sums)."""
    rng = np.random.default_rng(5)
    nrow, ncol = 500, 400
    cap = rng.integers(0, 60, nrow).astype(np.int32)
    counts = np.minimum(cap, rng.integers(0, 60, nrow)).astype(np.int32)
    A = sparse.bsr_zeros(nrow, ncol, wa.mat33, device=device, row_capacity=_t(cap, device))
    total = int(cap.sum())
    off = np.concatenate([[0], np.cumsum(cap)]).astype(np.int32)
    cols = np.full(total, -1, np.int32)
    vals = np.zeros((total, 3, 3), np.float32)
    rows_t, cols_t, vals_t = [], [], []
    for r in range(nrow):
        c = rng.integers(0, 25, counts[r]).astype(np.int32)   # many duplicates per row
        v = rng.standard_normal((counts[r], 3, 3)).astype(np.float32)
        v[rng.random(counts[r]) < 0.1] = 0.0                   # some zero blocks to prune
        cols[off[r]:off[r] + counts[r]] = c
        vals[off[r]:off[r] + counts[r]] = v
        rows_t.append(np.full(counts[r], r, np.int32)); cols_t.append(c); vals_t.append(v)
    A.row_counts = _t(counts, device)
    A.columns = _t(cols, device)
    A.values = _t(vals, device)
    sparse.bsr_compress(A, topology=topology)
    ref = oracle_mod.bsr_from_triplets(nrow, ncol, np.concatenate(rows_t), np.concatenate(cols_t),
                                       np.concatenate(vals_t))
    ref_cnt = np.diff(ref["offsets"])
    if topology == "compact":
        nnz = A.nnz_sync()
        np.testing.assert_array_equal(A.offsets.cpu().numpy(), ref["offsets"])
        np.testing.assert_array_equal(A.columns.cpu().numpy()[:nnz], ref["columns"])
        np.testing.assert_array_equal(A.values.cpu().numpy()[:nnz], ref["values"])
        assert (A.columns.cpu().numpy()[nnz:] == -1).all()
    else:
        np.testing.assert_array_equal(A.offsets.cpu().numpy(), off)
        np.testing.assert_array_equal(A.row_counts.cpu().numpy(), ref_cnt)
        gc = A.columns.cpu().numpy()
        gv = A.values.cpu().numpy()
        for r in range(nrow):
            k0, k1 = ref["offsets"][r], ref["offsets"][r + 1]
            np.testing.assert_array_equal(gc[off[r]:off[r] + ref_cnt[r]], ref["columns"][k0:k1])
            np.testing.assert_array_equal(gv[off[r]:off[r] + ref_cnt[r]], ref["values"][k0:k1])


# ---------------------------------------------------------------------------
# Padded / masked topologies, status codes, transposes, assignment
# ---------------------------------------------------------------------------

def _f(a, device):
    return _t(np.asarray(a, np.float32), device)


def _i(a, device):
    return _t(np.asarray(a, np.int32), device)


def _gapped_csr(device):
    """test_sparse.py:96-106: a 2x3 CSR with row capacity 3 filled by padded triplets."""
    A = sparse.bsr_zeros(2, 3, wa.float32, device=device, row_capacity=3)
    sparse.bsr_set_from_triplets(A, _i([0, 0, 1, 1], device), _i([0, 2, 1, 2], device), _f([1.0, 2.0, 3.0, 4.0], device),
                                 topology="padded")
    return A


def test_padded_gapped_layout_reference_scenario(device):
    """test_sparse.py:255-350: padded triplets, copies, assignment, transpose and
    their row-capacity status."""
    A = _gapped_csr(device)
    expected = np.array([[1.0, 0.0, 2.0], [0.0, 3.0, 4.0]])
    np.testing.assert_array_equal(A.offsets.cpu().numpy(), [0, 3, 6])
    np.testing.assert_array_equal(A.row_counts.cpu().numpy(), [2, 2])
    D, cols = _active_dense(A)
    assert cols == [[0, 2], [1, 2]]
    np.testing.assert_array_equal(D, expected)
    np.testing.assert_array_equal(A.uncompress_rows().cpu().numpy(), [0, 0, -1, 1, 1, -1])
    np.testing.assert_allclose(sparse.bsr_mv(A, _f([5.0, 7.0, 11.0], device)).cpu().numpy(), [27.0, 65.0])
    np.testing.assert_allclose(sparse.bsr_get_diag(A).cpu().numpy(), [1.0, 3.0])
    assert A.status_sync() == sparse.BSR_STATUS_SUCCESS and A.status_message() == "success"

    compact = sparse.bsr_copy(A)
    assert compact.row_counts is None
    np.testing.assert_array_equal(_active_dense(compact)[0], expected)
    padded_copy = sparse.bsr_copy(2.0 * A, topology="padded")
    np.testing.assert_array_equal(padded_copy.offsets.cpu().numpy(), [0, 3, 6])
    np.testing.assert_array_equal(padded_copy.row_counts.cpu().numpy(), [2, 2])
    np.testing.assert_array_equal(_active_dense(padded_copy)[0], 2.0 * expected)

    dest = sparse.bsr_zeros(2, 3, wa.float32, device=device, row_capacity=3)
    sparse.bsr_assign(dest, A, topology="padded")
    D, cols = _active_dense(dest)
    assert cols == [[0, 2], [1, 2]]
    np.testing.assert_array_equal(D, expected)
    small = sparse.bsr_zeros(2, 3, wa.float32, device=device, row_capacity=1)
    sparse.bsr_assign(small, A, topology="padded")
    assert small.status_sync() == sparse.BSR_STATUS_ROW_CAPACITY_EXCEEDED
    assert small.status_message() == "row capacity exceeded"
    small.clear_status()
    assert small.status_sync() == sparse.BSR_STATUS_SUCCESS

    tdest = sparse.bsr_zeros(3, 2, wa.float32, device=device, row_capacity=2)
    sparse.bsr_set_transpose(tdest, A, topology="padded")
    D, cols = _active_dense(tdest)
    assert cols == [[0], [1], [0, 1]]
    np.testing.assert_array_equal(D, expected.T)
    tsmall = sparse.bsr_zeros(3, 2, wa.float32, device=device, row_capacity=1)
    sparse.bsr_set_transpose(tsmall, A, topology="padded")
    assert tsmall.status_sync() == sparse.BSR_STATUS_ROW_CAPACITY_EXCEEDED
    np.testing.assert_array_equal(tsmall.row_counts.cpu().numpy(), [1, 1, 0])

    trip = sparse.bsr_zeros(2, 3, wa.float32, device=device, row_capacity=_i([2, 3], device))
    sparse.bsr_set_from_triplets(trip, _i([1, 0, 1, 0, 1], device), _i([2, 1, 2, 0, 1], device),
                                 _f([4.0, 1.0, 5.0, 2.0, 3.0], device), topology="padded")
    D, cols = _active_dense(trip)
    assert cols == [[0, 1], [1, 2]]
    np.testing.assert_array_equal(D, [[2.0, 1.0, 0.0], [0.0, 3.0, 9.0]])
    over = sparse.bsr_zeros(2, 3, wa.float32, device=device, row_capacity=1)
    sparse.bsr_set_from_triplets(over, _i([0, 0, 1], device), _i([0, 2, 1], device), _f([1.0, 2.0, 3.0], device),
                                 topology="padded")
    assert over.status_sync() == sparse.BSR_STATUS_ROW_CAPACITY_EXCEEDED
    np.testing.assert_array_equal(over.row_counts.cpu().numpy(), [0, 1])


def test_padded_alloc_reference_scenario(device):
    """test_sparse.py:1314-1390: per-row capacities, padded reset, nnz_capacity."""
    reserved = sparse.bsr_zeros(3, 4, wa.float32, device=device, row_capacity=_i([1, 0, 3], device))
    sparse.bsr_set_from_triplets(reserved, _i([0, 2, 2], device), _i([0, 1, 3], device), _f([1.0, 2.0, 3.0], device),
                                 topology="padded")
    assert reserved.nnz == 4
    np.testing.assert_array_equal(reserved.offsets.cpu().numpy(), [0, 1, 1, 4])
    np.testing.assert_array_equal(reserved.row_counts.cpu().numpy(), [1, 0, 2])
    np.testing.assert_array_equal(_active_dense(reserved)[0], [[1, 0, 0, 0], [0, 0, 0, 0], [0, 2, 0, 3]])
    sparse.bsr_set_zero(reserved, topology="padded")
    assert reserved.nnz == 4
    np.testing.assert_array_equal(reserved.offsets.cpu().numpy(), [0, 1, 1, 4])
    np.testing.assert_array_equal(reserved.row_counts.cpu().numpy(), [0, 0, 0])
    sparse.bsr_set_zero(reserved, topology="padded", row_capacity=_i([0, 2, 1], device))
    assert reserved.nnz_sync() == 3
    np.testing.assert_array_equal(reserved.offsets.cpu().numpy(), [0, 0, 2, 3])

    over = sparse.bsr_zeros(3, 4, wa.float32, device=device, row_capacity=_i([1, 0, 1], device), nnz_capacity=8)
    sparse.bsr_set_from_triplets(over, _i([0, 2], device), _i([1, 3], device), _f([2.0, 5.0], device),
                                 topology="padded")
    assert over.nnz == 8
    np.testing.assert_array_equal(over.offsets.cpu().numpy(), [0, 1, 1, 2])
    np.testing.assert_array_equal(over.row_counts.cpu().numpy(), [1, 0, 1])
    np.testing.assert_array_equal(_active_dense(over)[0], [[0, 2, 0, 0], [0, 0, 0, 0], [0, 0, 0, 5]])

    # test_sparse.py:165-180: topology only into a hand-made padded matrix
    t = sparse.bsr_zeros(2, 3, wa.float32, device=device)
    t.nnz = 5
    t.offsets = _i([0, 2, 5], device)
    t.row_counts = _i([0, 0], device)
    t.columns = _i([-9] * 5, device)
    t.values = torch.empty(5, dtype=torch.float32, device=device)
    sparse.bsr_set_from_triplets(t, _i([1, 0, 1], device), _i([2, 1, 2], device), values=None, topology="padded")
    np.testing.assert_array_equal(t.offsets.cpu().numpy(), [0, 2, 5])
    np.testing.assert_array_equal(t.row_counts.cpu().numpy(), [1, 1])
    assert _active_dense(t)[1] == [[1], [2]]


@pytest.mark.parametrize("bshape,dtype", [((1, 1), np.float32), ((3, 3), np.float64), ((2, 3), np.float32)])
def test_transpose_reference_scenario(device, bshape, dtype):
    """test_sparse.py:631-676: scaled transpose through the expression API, block-shape
    check, and the masked transpose into a smaller topology."""
    rng = np.random.default_rng(123)
    rows = np.array([0, 1, 2, 3, 2, 1], np.int32)
    cols = np.array([1, 4, 1, 3, 0, 2], np.int32)
    vals = rng.random(size=(6, *bshape)).astype(dtype)
    btype = wa.types.type_to_scalar(_t(vals[:1], "cpu").dtype)
    A = sparse.bsr_zeros(4, 5, btype if bshape == (1, 1) else wa.matrix(bshape, btype), device=device)
    v = vals.reshape(6) if bshape == (1, 1) else vals
    sparse.bsr_set_from_triplets(A, _i(rows, device), _i(cols, device), _t(v, device))
    ref = 2.0 * _active_dense(A)[0].T
    At = (2.0 * A).transpose().eval()
    np.testing.assert_allclose(_active_dense(At)[0], ref, rtol=1e-6)
    if bshape[0] != bshape[1]:
        with pytest.raises(ValueError, match="Destination block shape must be"):
            sparse.bsr_set_transpose(dest=A, src=A)
    sparse.bsr_set_from_triplets(A, _i(rows[:3], device), _i(cols[:3], device), _t(v[:3], device))
    nt = At.nnz_sync()
    At = sparse.bsr_from_triplets(At.nrow, At.ncol, At.uncompress_rows()[:3], At.columns[:3], At.values[:3])
    np.testing.assert_array_equal(At.uncompress_rows().cpu().numpy()[:3], [0, 1, 1])
    np.testing.assert_array_equal(At.columns.cpu().numpy()[:3], [2, 0, 2])
    assert nt == 6
    sparse.bsr_set_transpose(At, A, topology="masked")
    pruned = sparse.bsr_from_triplets(At.nrow, At.ncol, At.uncompress_rows(), At.columns[: At.nnz], At.values[: At.nnz])
    assert pruned.nnz_sync() == 2


@pytest.mark.parametrize("padded_src,padded_dest", [(False, False), (True, False), (False, True), (True, True)])
def test_transpose_vs_oracle(device, oracle_mod, padded_src, padded_dest):
    """Transposed topology bit-exact against the oracle (itself pinned on the
    reference's host transpose); values are exact copies times 2."""
    rng = np.random.default_rng(11 + 2 * padded_src + padded_dest)
    nrow, ncol = 700, 600
    counts = rng.integers(0, 40, nrow).astype(np.int32)
    cap = counts + (rng.integers(0, 4, nrow) if padded_src else 0)
    off = np.concatenate([[0], np.cumsum(cap)]).astype(np.int32)
    T = int(off[-1])
    cols = np.full(T, -1, np.int32)
    for r in range(nrow):
        cols[off[r]:off[r] + counts[r]] = np.sort(rng.choice(ncol, size=counts[r], replace=False))
    vals = rng.standard_normal((T, 3, 2)).astype(np.float32)
    A = sparse.bsr_zeros(nrow, ncol, wa.matrix((3, 2), wa.float32), device=device)
    A.nnz = T
    A.offsets, A.columns, A.values = _i(off, device), _i(cols, device), _t(vals, device)
    A.row_counts = _i(counts, device) if padded_src else None
    if padded_dest:
        tcap = rng.integers(0, 45, ncol).astype(np.int32)
        dest = sparse.bsr_zeros(ncol, nrow, wa.matrix((2, 3), wa.float32), device=device, row_capacity=_t(tcap, device))
        sparse.bsr_set_transpose(dest, 2.0 * A, topology="padded")
        toff = np.concatenate([[0], np.cumsum(tcap)]).astype(np.int32)
        ref = oracle_mod.bsr_transpose(nrow, ncol, off, cols, row_counts=counts if padded_src else None, capacity=toff)
        assert dest.status_sync() == ref["status"]
        np.testing.assert_array_equal(dest.row_counts.cpu().numpy(), ref["row_counts"])
        gc, gv = dest.columns.cpu().numpy(), dest.values.cpu().numpy()
        for r in range(ncol):
            n = ref["row_counts"][r]
            seg = slice(toff[r], toff[r] + n)
            np.testing.assert_array_equal(gc[seg], ref["columns"][seg])
            np.testing.assert_array_equal(gv[seg], 2.0 * vals[ref["block_map"][seg]].transpose(0, 2, 1))
    else:
        dest = sparse.bsr_transposed(2.0 * A)
        ref = oracle_mod.bsr_transpose(nrow, ncol, off, cols, row_counts=counts if padded_src else None)
        n = dest.nnz_sync()
        assert dest.row_counts is None and n == ref["columns"].size
        np.testing.assert_array_equal(dest.offsets.cpu().numpy(), ref["offsets"])
        np.testing.assert_array_equal(dest.columns.cpu().numpy()[:n], ref["columns"])
        np.testing.assert_array_equal(dest.values.cpu().numpy()[:n],
                                      2.0 * vals[ref["block_map"]].transpose(0, 2, 1))


def test_padded_triplets_vs_oracle(device, oracle_mod):
    """Padded assembly: each row is the oracle's compact row (bit-exact sums) when it
    fits its capacity, empty otherwise, with the sticky status raised."""
    rng = np.random.default_rng(21)
    nrow, ncol, T = 400, 300, 30000
    rows = rng.integers(0, nrow, T).astype(np.int32)
    cols = rng.integers(0, 60, T).astype(np.int32)
    vals = rng.standard_normal((T, 3, 3)).astype(np.float32)
    vals[rng.random(T) < 0.05] = 0.0
    ref = oracle_mod.bsr_from_triplets(nrow, ncol, rows, cols, vals)
    ref_cnt = np.diff(ref["offsets"])
    cap = (ref_cnt + rng.integers(-3, 5, nrow)).clip(0).astype(np.int32)
    A = sparse.bsr_zeros(nrow, ncol, wa.mat33, device=device, row_capacity=_t(cap, device))
    sparse.bsr_set_from_triplets(A, _t(rows, device), _t(cols, device), _t(vals, device), topology="padded")
    off = np.concatenate([[0], np.cumsum(cap)])
    fits = ref_cnt <= cap
    assert A.status_sync() == (0 if fits.all() else 1)
    np.testing.assert_array_equal(A.row_counts.cpu().numpy(), np.where(fits, ref_cnt, 0))
    gc, gv = A.columns.cpu().numpy(), A.values.cpu().numpy()
    for r in np.flatnonzero(fits):
        k0, k1 = ref["offsets"][r], ref["offsets"][r + 1]
        np.testing.assert_array_equal(gc[off[r]:off[r] + ref_cnt[r]], ref["columns"][k0:k1])
        assert gv[off[r]:off[r] + ref_cnt[r]].tobytes() == ref["values"][k0:k1].tobytes()
    over = np.flatnonzero(~fits & (cap > 0))
    assert not gv[off[over]].any()                      # overflowing rows keep cleared values


def test_masked_triplets_vs_oracle(device, oracle_mod):
    """Masked assembly: the topology is kept; each of its blocks holds the
    sequential sum of its triplets (bit-exact), zero when it has none."""
    rng = np.random.default_rng(22)
    nrow, ncol, T = 300, 200, 20000
    mask = oracle_mod.bsr_from_triplets(nrow, ncol, rng.integers(0, nrow, 3000).astype(np.int32),
                                        rng.integers(0, ncol, 3000).astype(np.int32))
    nm = mask["nnz"]
    A = sparse.bsr_zeros(nrow, ncol, wa.float64, device=device)
    A.nnz = nm
    A.offsets, A.columns = _t(mask["offsets"], device), _t(mask["columns"], device)
    A.values = torch.full((nm,), 7.0, dtype=torch.float64, device=device)
    rows = rng.integers(-2, nrow + 2, T).astype(np.int32)
    cols = rng.integers(0, ncol, T).astype(np.int32)
    vals = rng.standard_normal(T)
    count = T - 1000
    sparse.bsr_set_from_triplets(A, _t(rows, device), _t(cols, device), _t(vals, device),
                                 count=torch.tensor([count], dtype=torch.int32, device=device), topology="masked")
    # expected: triplets inside the mask, summed in ascending order (no pruning)
    sel = np.zeros(count, bool)
    mrow = np.repeat(np.arange(nrow), np.diff(mask["offsets"]))
    keys = set(zip(mrow.tolist(), mask["columns"].tolist()))
    for k in range(count):
        sel[k] = (int(rows[k]), int(cols[k])) in keys
    ref = oracle_mod.bsr_from_triplets(nrow, ncol, rows[:count][sel], cols[:count][sel], vals[:count][sel],
                                       prune=False)
    got = A.values.cpu().numpy()
    expect = np.zeros(nm)
    lookup = {(r, c): k for k, (r, c) in enumerate(zip(np.repeat(np.arange(nrow), np.diff(ref["offsets"])).tolist(),
                                                       ref["columns"].tolist()))}
    for k, key in enumerate(zip(mrow.tolist(), mask["columns"].tolist())):
        if key in lookup:
            expect[k] = ref["values"][lookup[key]]
    assert got.tobytes() == expect.tobytes()
    np.testing.assert_array_equal(A.offsets.cpu().numpy(), mask["offsets"])
    with pytest.warns(DeprecationWarning):
        sparse.bsr_set_from_triplets(A, _t(rows, device), _t(cols, device), _t(vals, device), masked=True)


@pytest.mark.parametrize("src_shape,dest_shape,dest_dtype", [
    ((3, 2), (1, 1), np.float64), ((3, 2), (3, 2), np.float32), ((1, 1), (3, 2), np.float32),
    ((3, 2), (6, 4), np.float64), ((2, 2), (1, 2), np.float32)])
def test_bsr_assign_reblock_vs_dense(device, src_shape, dest_shape, dest_dtype):
    """bsr_assign / bsr_copy with another block shape or scalar type reproduce the
    scaled dense matrix exactly (test_sparse.py:554-602 split/merge)."""
    rng = np.random.default_rng(31)
    nrow, ncol, T = 24, 36, 300
    rows = rng.integers(0, nrow, T).astype(np.int32)
    cols = rng.integers(0, ncol, T).astype(np.int32)
    vals = rng.standard_normal((T, *src_shape)).astype(np.float32)
    stype = wa.float32 if src_shape == (1, 1) else wa.matrix(src_shape, wa.float32)
    A = sparse.bsr_zeros(nrow, ncol, stype, device=device)
    sparse.bsr_set_from_triplets(A, _t(rows, device), _t(cols, device),
                                 _t(vals.reshape(T) if src_shape == (1, 1) else vals, device))
    D = _active_dense(A)[0]
    st = wa.float64 if dest_dtype == np.float64 else wa.float32
    B = sparse.bsr_copy(-0.5 * A, scalar_type=st, block_shape=dest_shape)
    assert B.block_shape == tuple(dest_shape)
    B.nnz_sync()
    Db, _ = _active_dense(B)
    np.testing.assert_array_equal(Db, -0.5 * D)
    M = sparse.bsr_zeros(B.nrow, B.ncol, B.dtype, device=device)
    sparse.bsr_assign(M, B, structure_only=True)
    assert M.nnz_sync() == B.nnz_sync()
    np.testing.assert_array_equal(M.columns.cpu().numpy()[: M.nnz], B.columns.cpu().numpy()[: B.nnz])
    # masked: keep only the diagonal blocks of the destination
    diag = sparse.bsr_from_triplets(B.nrow, B.ncol, _t(np.arange(min(B.nrow, B.ncol), dtype=np.int32), device),
                                    _t(np.arange(min(B.nrow, B.ncol), dtype=np.int32), device),
                                    torch.ones((min(B.nrow, B.ncol), *dest_shape) if dest_shape != (1, 1)
                                               else (min(B.nrow, B.ncol),), dtype=B.values.dtype, device=device),
                                    prune_numerical_zeros=False)
    sparse.bsr_assign(diag, A, topology="masked")
    Dd, _ = _active_dense(diag)
    br, bc = dest_shape
    keep = np.zeros_like(D, dtype=bool)
    for k in range(min(B.nrow, B.ncol)):
        keep[k * br:(k + 1) * br, k * bc:(k + 1) * bc] = True
    np.testing.assert_array_equal(Dd, np.where(keep, D, 0.0))
    # padded copy with the change of block shape: each row's capacity is that
    # of its covering source rows, its active columns those of the compact copy
    P = sparse.bsr_copy(-0.5 * A, scalar_type=st, block_shape=dest_shape, topology="padded")
    assert P.block_shape == tuple(dest_shape) and P.row_counts is not None
    assert P.status_sync() == sparse.BSR_STATUS_SUCCESS
    Dp, pcols = _active_dense(P)
    np.testing.assert_array_equal(Dp, -0.5 * D)
    _, bcols = _active_dense(B)
    assert pcols == bcols
    # capacity: source rows covering destination row r are those holding its
    # sub-rows; each source block spans src_subcols destination sub-columns
    src_subrows = max(1, src_shape[0] // br)
    dest_subrows = max(1, br // src_shape[0])
    src_subcols = max(1, src_shape[1] // bc)
    a_cap = np.diff(A.offsets.cpu().numpy())
    p_cap = np.diff(P.offsets.cpu().numpy())
    for r in range(P.nrow):
        first = r * dest_subrows // src_subrows
        last = min(A.nrow, (r * dest_subrows + dest_subrows + src_subrows - 1) // src_subrows)
        assert p_cap[r] == a_cap[first:last].sum() * src_subcols, r
    # padded assign across block shapes into capacities: the longest row one
    # block short (left empty, status raised), every other row one spare
    cap = np.array([len(c) for c in bcols], np.int32)
    small = int(np.argmax(cap))
    row_cap = cap + 1
    row_cap[small] = cap[small] - 1
    Q = sparse.bsr_zeros(B.nrow, B.ncol, B.dtype, device=device, row_capacity=_t(row_cap, device))
    sparse.bsr_assign(Q, -0.5 * A, topology="padded")
    assert Q.status_sync() == sparse.BSR_STATUS_ROW_CAPACITY_EXCEEDED
    Dq, qcols = _active_dense(Q)
    assert qcols[small] == [] and all(qcols[r] == bcols[r] for r in range(B.nrow) if r != small)
    rows_keep = np.ones(D.shape[0], bool)
    rows_keep[small * br:(small + 1) * br] = False
    np.testing.assert_array_equal(Dq[rows_keep], -0.5 * D[rows_keep])
    assert not Dq[~rows_keep].any()


def test_padded_reblock_copy_reference_scenario(device):
    """test_padded_bsr_capture_reblock_copy (test_sparse.py:1231-1274), eagerly:
    one 2x2 block split into 1x1 blocks keeps a padded topology."""
    src = sparse.bsr_zeros(1, 1, wa.mat22, device=device, row_capacity=1)
    sparse.bsr_set_from_triplets(src, _i([0], device), _i([0], device),
                                 _t(np.array([[[1.0, 2.0], [3.0, 4.0]]], np.float32), device), topology="padded")
    dest = sparse.bsr_copy(src, block_shape=(1, 1), topology="padded")
    np.testing.assert_array_equal(dest.offsets.cpu().numpy(), [0, 2, 4])
    np.testing.assert_array_equal(dest.row_counts.cpu().numpy(), [2, 2])
    np.testing.assert_array_equal(dest.columns.cpu().numpy()[:4], [0, 1, 0, 1])
    np.testing.assert_array_equal(dest.values.cpu().numpy()[:4], [1.0, 2.0, 3.0, 4.0])
    assert dest.status_sync() == sparse.BSR_STATUS_SUCCESS


# ---------------------------------------------------------------------------
# bsr_axpy / bsr_mm / block-diagonal matrices
# ---------------------------------------------------------------------------

def _from_triplets_np(device, nrow, ncol, rows, cols, vals):
    btype = wa.types.type_to_scalar(_t(vals[:1], "cpu").dtype)
    if vals.ndim == 3 and vals.shape[1:] != (1, 1):
        btype = wa.matrix(vals.shape[1:], btype)
    A = sparse.bsr_zeros(nrow, ncol, btype, device=device)
    v = vals.reshape(-1) if vals.ndim == 3 and vals.shape[1:] == (1, 1) else vals
    sparse.bsr_set_from_triplets(A, _i(rows, device), _i(cols, device), _t(v, device))
    return A


def _as_dict(A):
    n = A.nnz_sync()
    br, bc = A.block_shape
    return dict(nrow=A.nrow, ncol=A.ncol, offsets=A.offsets.cpu().numpy(), columns=A.columns[:n].cpu().numpy(),
                values=A.values[:n].cpu().numpy().reshape(n, br, bc))


SHAPES = [((1, 1), np.float32), ((3, 3), np.float64), ((1, 3), np.float32)]


@pytest.mark.parametrize("bshape,dtype", SHAPES)
def test_bsr_axpy_reference_scenario(device, bshape, dtype):
    """test_sparse.py:678-728: alpha/beta sweep, aliasing, masked, shape check."""
    rng = np.random.default_rng(123)
    nrow, ncol, nnz = 2, 3, 6
    mk = lambda: _from_triplets_np(device, nrow, ncol, rng.integers(0, nrow, nnz), rng.integers(0, ncol, nnz),  # noqa: E731
                                   rng.random((nnz, *bshape)).astype(dtype))
    x, y = mk(), mk()
    work = sparse.bsr_axpy_work_arrays()
    for alpha, beta in zip([-1.0, 0.0, 1.0], [2.0, -1.0, 0.0]):
        ref = alpha * _active_dense(x)[0] + beta * _active_dense(y)[0]
        sparse.bsr_axpy(x, y, alpha, beta, work_arrays=work)
        np.testing.assert_allclose(_active_dense(y)[0], ref, rtol=1e-5, atol=1e-6)
    ref = 3.0 * _active_dense(y)[0]
    y += y * 2.0
    np.testing.assert_allclose(_active_dense(y)[0], ref, rtol=1e-6)
    n = y.nnz_sync()
    y_mask = sparse.bsr_from_triplets(nrow, ncol, y.uncompress_rows()[:1], y.columns[:1], y.values[:1])
    sparse.bsr_axpy(y, y_mask, topology="masked")
    assert y_mask.nnz_sync() == 1
    np.testing.assert_allclose(y_mask.values[:1].cpu().numpy(), 2.0 * y.values[:1].cpu().numpy(), rtol=1e-6)
    assert n > 0
    y.ncol = y.ncol + 1
    with pytest.raises(ValueError, match="Matrices must have the same number of rows and columns"):
        sparse.bsr_axpy(x, y)


@pytest.mark.parametrize("bshape,dtype", SHAPES)
def test_bsr_mm_reference_scenario(device, bshape, dtype):
    """test_sparse.py:733-828: alpha/beta sweep, reused topology, masked product into a
    diagonal, operators, aliasing, shape checks; test_sparse.py:831-885 (A A^T)."""
    rng = np.random.default_rng(123)
    nnz = 6
    xs, ys = bshape, bshape[::-1]
    zs = (xs[0], ys[1])
    x = _from_triplets_np(device, 3, 2, rng.integers(0, 3, nnz), rng.integers(0, 2, nnz), rng.random((nnz, *xs)).astype(dtype))
    y = _from_triplets_np(device, 2, 3, rng.integers(0, 2, nnz), rng.integers(0, 3, nnz), rng.random((nnz, *ys)).astype(dtype))
    z = _from_triplets_np(device, 3, 3, rng.integers(0, 3, nnz), rng.integers(0, 3, nnz), rng.random((nnz, *zs)).astype(dtype))
    tol = dict(rtol=1e-5, atol=1e-5)
    work = sparse.bsr_mm_work_arrays()
    for alpha, beta in zip([-1.0, 0.0, 2.0], [2.0, -1.0, 0.0]):
        ref = alpha * (_active_dense(x)[0] @ _active_dense(y)[0]) + beta * _active_dense(z)[0]
        sparse.bsr_mm(x, y, z, alpha, beta, work_arrays=work)
        np.testing.assert_allclose(_active_dense(z)[0], ref, **tol)
    sparse.bsr_mm(x, y, z, alpha, beta, work_arrays=work, reuse_topology=True)
    np.testing.assert_allclose(_active_dense(z)[0], ref, **tol)

    zd = sparse.bsr_diag(rows_of_blocks=z.nrow, block_type=z.dtype, device=device)
    sparse.bsr_mm(x, y, zd, topology="masked")
    full = _active_dense(x)[0] @ _active_dense(y)[0]
    keep = np.zeros_like(full, bool)
    for k in range(3):
        keep[k * zs[0]:(k + 1) * zs[0], k * zs[1]:(k + 1) * zs[1]] = True
    np.testing.assert_allclose(_active_dense(zd)[0], np.where(keep, full, 0.0), **tol)
    prod = (alpha * x) @ y
    np.testing.assert_allclose(_active_dense(prod)[0], alpha * full, **tol)

    if xs[0] == xs[1]:
        a, b = -1.0, 2.0
        ref = a * (_active_dense(zd)[0] @ _active_dense(x)[0]) + b * _active_dense(x)[0]
        sparse.bsr_mm(zd, x, x, a, b)
        np.testing.assert_allclose(_active_dense(x)[0], ref, **tol)
        ref = a * (_active_dense(zd)[0] @ _active_dense(zd)[0]) + b * _active_dense(zd)[0]
        sparse.bsr_mm(zd, zd, zd, a, b)
        np.testing.assert_allclose(_active_dense(zd)[0], ref, **tol)
    else:
        with pytest.raises(ValueError, match="Incompatible block sizes"):
            sparse.bsr_mm(z, y)
    y.ncol = y.ncol * 2
    with pytest.raises(ValueError, match="Incompatible number of rows/columns"):
        sparse.bsr_mm(y, z)
    y.ncol = y.ncol // 2
    AAt = sparse.bsr_mm(y, sparse.bsr_transposed(y))
    Dy = _active_dense(y)[0]
    np.testing.assert_allclose(_active_dense(AAt)[0], Dy @ Dy.T, **tol)


def test_bsr_mm_deep_and_max_new_nnz(device):
    """test_sparse.py:907-1003: products of a wide dense block matrix with its
    transpose (deep rows), and the max_new_nnz cap."""
    rng = np.random.default_rng(123)
    nrow, ncol = 2, 200
    rows, cols = np.meshgrid(np.arange(nrow), np.arange(ncol))
    vals = rng.random((nrow * ncol, 3, 3)).astype(np.float32)
    A = _from_triplets_np(device, nrow, ncol, rows.ravel(), cols.ravel(), vals)
    At = sparse.bsr_transposed(A)
    D = _active_dense(A)[0]
    np.testing.assert_allclose(_active_dense(sparse.bsr_mm(A, At))[0], D @ D.T, rtol=1e-4)
    np.testing.assert_allclose(_active_dense(sparse.bsr_mm(At, A))[0], D.T @ D, rtol=1e-4, atol=1e-5)
    B = _from_triplets_np(device, 2, 2, [0, 0, 1, 1], [0, 1, 0, 1], np.array([1, 2, 3, 4], np.float32).reshape(4, 1, 1))
    C = sparse.bsr_zeros(2, 2, wa.float32, device=device)
    sparse.bsr_mm(B, B, C, max_new_nnz=4)
    assert C.nnz_sync() == 4
    sparse.bsr_set_zero(C)
    assert C.nnz_sync() == 0
    with pytest.warns(RuntimeWarning, match="exceeded"):
        sparse.bsr_mm(B, B, C, max_new_nnz=2)
    assert C.nnz_sync() == 2


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bsr_mm_and_axpy_vs_oracle(device, oracle_mod, dtype):
    """Compact product and sum bit-exact against the oracle (topology, including the
    deep-row span rule, and values in the stated summation order)."""
    rng = np.random.default_rng(41)

    def rand(nrow, ncol, nnz, bshape):
        return _from_triplets_np(device, nrow, ncol, rng.integers(0, nrow, nnz), rng.integers(0, ncol, nnz),
                                 rng.standard_normal((nnz, *bshape)).astype(dtype))
    x, y, z = rand(300, 200, 3000, (3, 3)), rand(200, 250, 2500, (3, 3)), rand(300, 250, 1500, (3, 3))
    ox, oy, oz = _as_dict(x), _as_dict(y), _as_dict(z)
    sparse.bsr_mm(x, y, z, alpha=-0.75, beta=1.5)
    ref = oracle_mod.bsr_mm(ox, oy, oz, alpha=-0.75, beta=1.5)
    got = _as_dict(z)
    np.testing.assert_array_equal(got["offsets"], ref["offsets"])
    np.testing.assert_array_equal(got["columns"], ref["columns"])
    assert got["values"].tobytes() == ref["values"].tobytes()
    # a gapped band matrix squared has deep rows (more products than the span of
    # columns reached): the span is listed whole, odd offsets as structural zeros
    n = 64
    r = np.repeat(np.arange(n), 4)
    c = np.clip(r + np.tile(np.arange(0, 8, 2), n), 0, n - 1)
    band = _from_triplets_np(device, n, n, r, c, rng.standard_normal((r.size, 1, 1)).astype(dtype))
    ob = _as_dict(band)
    sq = sparse.bsr_mm(band, band)
    ref = oracle_mod.bsr_mm(ob, ob)
    got = _as_dict(sq)
    assert (np.diff(got["offsets"]) > 4).any()      # spans wider than the products' columns
    np.testing.assert_array_equal(got["offsets"], ref["offsets"])
    np.testing.assert_array_equal(got["columns"], ref["columns"])
    assert got["values"].tobytes() == ref["values"].tobytes()
    w = rand(300, 250, 2000, (3, 3))
    ow, oz = _as_dict(w), _as_dict(z)
    sparse.bsr_axpy(w, z, alpha=0.5, beta=-2.0)
    ref = oracle_mod.bsr_axpy(ow, oz, alpha=0.5, beta=-2.0)
    got = _as_dict(z)
    np.testing.assert_array_equal(got["columns"], ref["columns"])
    assert got["values"].tobytes() == ref["values"].tobytes()


@pytest.mark.parametrize("shapes,dtype", [((1, 1, 1), np.float32), ((3, 3, 3), np.float32), ((3, 3, 3), np.float64),
                                          ((2, 4, 3), np.float32), ((4, 4, 4), np.float64)])
def test_bsr_mm_tiled_vs_oracle(device, oracle_mod, shapes, dtype):
    """bsr_mm(tile_size=G): values bit-exact against the oracle's tiled order
    (rank-one updates over (block, inner index) pairs dealt to G lanes, xor
    butterfly) for G = 1, 8, 64, the topology unchanged; tile sizes that are not
    powers of two <= 64 are refused."""
    br, bk, bc = shapes
    rng = np.random.default_rng(53 + br * bk)

    def rand(nrow, ncol, nnz, bshape):
        return _from_triplets_np(device, nrow, ncol, rng.integers(0, nrow, nnz), rng.integers(0, ncol, nnz),
                                 rng.standard_normal((nnz, *bshape)).astype(dtype))
    x, y, z0 = rand(120, 90, 1500, (br, bk)), rand(90, 110, 1200, (bk, bc)), rand(120, 110, 600, (br, bc))
    ox, oy, oz = _as_dict(x), _as_dict(y), _as_dict(z0)
    for tile in (1, 8, 64):
        z = sparse.bsr_copy(z0)
        sparse.bsr_mm(x, y, z, alpha=-0.75, beta=1.5, tile_size=tile)
        ref = oracle_mod.bsr_mm(ox, oy, oz, alpha=-0.75, beta=1.5, tile=tile)
        got = _as_dict(z)
        np.testing.assert_array_equal(got["offsets"], ref["offsets"])
        np.testing.assert_array_equal(got["columns"], ref["columns"])
        assert got["values"].tobytes() == ref["values"].tobytes(), tile
    with pytest.raises(ValueError, match="power of two"):
        sparse.bsr_mm(x, y, tile_size=48)


def test_bsr_mm_axpy_padded_and_masked_vs_oracle(device, oracle_mod):
    """Padded product / sum into row capacity (rows that fit equal the oracle's exact
    union, bit-exact; the others raise the status) and the masked product."""
    rng = np.random.default_rng(43)

    def rand(nrow, ncol, nnz):
        return _from_triplets_np(device, nrow, ncol, rng.integers(0, nrow, nnz), rng.integers(0, ncol, nnz),
                                 rng.standard_normal((nnz, 3, 3)).astype(np.float32))
    x, y, z0 = rand(200, 150, 1500), rand(150, 180, 1200), rand(200, 180, 600)
    ox, oy, oz = _as_dict(x), _as_dict(y), _as_dict(z0)
    ref = oracle_mod.bsr_mm(ox, oy, oz, alpha=2.0, beta=-1.0, deep_ranges=False)
    cnt = np.diff(ref["offsets"])
    cap = np.maximum(cnt + rng.integers(-2, 4, 200), np.diff(oz["offsets"])).astype(np.int32)
    z = sparse.bsr_zeros(200, 180, wa.mat33, device=device, row_capacity=_t(cap, device))
    sparse.bsr_assign(z, z0, topology="padded")     # the capacity holds z0's rows
    assert z.status_sync() == 0
    sparse.bsr_mm(x, y, z, alpha=2.0, beta=-1.0, topology="padded")
    fits = cnt <= cap
    assert z.status_sync() == (0 if fits.all() else 1)
    np.testing.assert_array_equal(z.row_counts.cpu().numpy(), np.where(fits, cnt, 0))
    off = np.concatenate([[0], np.cumsum(cap)])
    gc, gv = z.columns.cpu().numpy(), z.values.cpu().numpy()
    for r in np.flatnonzero(fits):
        k0, k1 = ref["offsets"][r], ref["offsets"][r + 1]
        np.testing.assert_array_equal(gc[off[r]:off[r] + cnt[r]], ref["columns"][k0:k1])
        assert gv[off[r]:off[r] + cnt[r]].tobytes() == ref["values"][k0:k1].tobytes()

    # padded axpy: y := 0.5 w + 2 y into y's capacity
    w = rand(200, 180, 800)
    ow = _as_dict(w)
    refa = oracle_mod.bsr_axpy(ow, oz, alpha=0.5, beta=2.0)
    cnta = np.diff(refa["offsets"])
    capa = (cnta + rng.integers(-1, 3, 200)).clip(np.diff(oz["offsets"])).astype(np.int32)
    ya = sparse.bsr_zeros(200, 180, wa.mat33, device=device, row_capacity=_t(capa, device))
    sparse.bsr_assign(ya, z0, topology="padded")
    sparse.bsr_axpy(w, ya, alpha=0.5, beta=2.0, topology="padded")
    fitsa = cnta <= capa
    assert ya.status_sync() == (0 if fitsa.all() else 1)
    offa = np.concatenate([[0], np.cumsum(capa)])
    gc, gv = ya.columns.cpu().numpy(), ya.values.cpu().numpy()
    for r in range(200):
        if fitsa[r]:
            k0, k1 = refa["offsets"][r], refa["offsets"][r + 1]
            np.testing.assert_array_equal(gc[offa[r]:offa[r] + cnta[r]], refa["columns"][k0:k1])
            assert gv[offa[r]:offa[r] + cnta[r]].tobytes() == refa["values"][k0:k1].tobytes()
    np.testing.assert_array_equal(ya.row_counts.cpu().numpy(), np.where(fitsa, cnta, 0))

    # masked product keeps z0's topology
    zm = sparse.bsr_copy(z0)
    sparse.bsr_mm(x, y, zm, alpha=1.5, beta=0.5, topology="masked")
    refm = oracle_mod.bsr_mm(ox, oy, oz, alpha=1.5, beta=0.5, z_topology=(oz["offsets"], oz["columns"]))
    got = _as_dict(zm)
    np.testing.assert_array_equal(got["columns"], oz["columns"])
    assert got["values"].tobytes() == refm["values"].tobytes()


def test_bsr_diag_and_identity(device):
    """bsr_diag / bsr_identity / bsr_set_diag (sparse.py:2939-3093)."""
    d = torch.arange(12, dtype=torch.float32, device=device).reshape(4, 3)
    blocks = torch.diag_embed(d)
    D = sparse.bsr_diag(blocks)
    assert D.nrow == D.ncol == 4 and D.nnz_sync() == 4
    np.testing.assert_array_equal(D.offsets.cpu().numpy(), [0, 1, 2, 3, 4])
    np.testing.assert_array_equal(D.columns[:4].cpu().numpy(), [0, 1, 2, 3])
    np.testing.assert_array_equal(_active_dense(D)[0], np.diag(np.arange(12.0)))
    I = sparse.bsr_identity(5, wa.mat33, device=device)
    np.testing.assert_array_equal(_active_dense(I)[0], np.eye(15))
    R = sparse.bsr_diag(2.0, rows_of_blocks=3, cols_of_blocks=5, device=device)
    np.testing.assert_array_equal(_active_dense(R)[0], 2.0 * np.eye(3, 5))


def test_bsr_get_set_diag_reference_scenario(device):
    """test_bsr_get_set_diag (test_sparse.py:486-551)."""
    rng = np.random.default_rng(123)
    nrow = ncol = 4
    rows = torch.tensor([0, 1, 2, 3, 2, 1], dtype=torch.int32, device=device)
    cols = torch.tensor([1, 1, 1, 3, 2, 2], dtype=torch.int32, device=device)
    vals_np = rng.random(size=(6, 3, 3))
    vals = torch.tensor(vals_np, dtype=torch.float32, device=device)
    A = sparse.bsr_zeros(nrow, ncol, wa.mat33, device=device)
    sparse.bsr_set_from_triplets(A, rows, cols, vals)
    diag = sparse.bsr_get_diag(A)
    d = diag.cpu().numpy()
    np.testing.assert_array_equal(d[0], np.zeros((3, 3)))
    for r, k in ((1, 1), (2, 4), (3, 3)):
        np.testing.assert_allclose(d[r], vals_np[k], atol=1e-5)
    sentinel = torch.full((nrow, 3, 3), 7.0, dtype=torch.float32, device=device)
    sparse.bsr_get_diag(A, out=sentinel)
    np.testing.assert_allclose(sentinel.cpu().numpy(), d, atol=1e-5)
    # set_diag / get_diag round trips
    D = sparse.bsr_diag(diag)
    sparse.bsr_get_diag(D, out=diag)
    np.testing.assert_array_equal(diag.cpu().numpy(), d)
    s_np = rng.random(size=nrow)
    Ds = sparse.bsr_diag(torch.tensor(s_np, dtype=torch.float32, device=device))
    np.testing.assert_allclose(sparse.bsr_get_diag(Ds).cpu().numpy(), s_np, atol=1e-6)
    np.testing.assert_allclose(sparse.bsr_get_diag(2.0 * Ds).cpu().numpy(), 2.0 * s_np, atol=1e-6)
    # uniform block diagonals
    with pytest.raises(ValueError, match="BsrMatrix block type must be either warp matrix or scalar"):
        sparse.bsr_diag(diag=vals_np[0, 0].tolist(), rows_of_blocks=nrow, cols_of_blocks=nrow + 1, device=device)
    U = sparse.bsr_diag(diag=vals_np[0].astype(np.float32).tolist(), rows_of_blocks=nrow, cols_of_blocks=nrow + 1,
                        device=device)
    assert U.values.shape[0] == nrow
    np.testing.assert_allclose(U.values.cpu().numpy(), np.broadcast_to(vals_np[0], (nrow, 3, 3)), atol=1e-6)
    Us = sparse.bsr_diag(diag=float(s_np[0]), rows_of_blocks=nrow, cols_of_blocks=nrow + 1, device=device)
    assert Us.values.shape[0] == nrow
    np.testing.assert_allclose(Us.values.cpu().numpy(), np.full(nrow, s_np[0]), atol=1e-6)
    I4 = sparse.bsr_identity(nrow, block_type=wa.matrix((4, 4), wa.float32), device=device)
    np.testing.assert_allclose(I4.values.cpu().numpy(), np.broadcast_to(np.eye(4), (nrow, 4, 4)), atol=1e-6)
    I1 = sparse.bsr_identity(nrow, block_type=wa.float64, device=device)
    np.testing.assert_array_equal(I1.values.cpu().numpy(), np.ones(nrow))


def test_bsr_split_merge_reference_scenario(device):
    """test_bsr_split_merge (test_sparse.py:554-600)."""
    rng = np.random.default_rng(123)
    nrow, ncol, n = 4, 8, 20
    rows = torch.tensor(rng.integers(0, high=nrow, size=n), dtype=torch.int32, device=device)
    cols = torch.tensor(rng.integers(0, high=ncol, size=n), dtype=torch.int32, device=device)
    vals = torch.tensor(rng.random(size=(n, 4, 2)), dtype=torch.float32, device=device)
    A = sparse.bsr_zeros(nrow, ncol, wa.matrix((4, 2), wa.float32), device=device)
    sparse.bsr_set_from_triplets(A, rows, cols, vals)
    ref = _dense_of(A)
    for shape, size in (((2, 2), 4), ((1, 1), 1), ((4, 4), 16), ((8, 8), 64)):
        B = sparse.bsr_copy(A, block_shape=shape)
        assert B.block_size == size
        np.testing.assert_allclose(_dense_of(B), ref, atol=1e-4)
    for bad in ((3, 3), (5, 5)):
        with pytest.raises(ValueError, match="Incompatible dest and src block shapes"):
            sparse.bsr_copy(A, block_shape=bad)
    with pytest.raises(
        ValueError,
        match=r"The requested block shape \(32, 32\) does not evenly divide the source matrix of total size \(16, 16\)",
    ):
        sparse.bsr_copy(A, block_shape=(32, 32))


def test_bsr_assign_masked_reference_scenario(device):
    """test_bsr_assign_masked (test_sparse.py:603-628): coarse diagonal by masked assign."""
    rng = np.random.default_rng(123)
    nrow, ncol, n = 16, 8, 20
    rows = torch.tensor(rng.integers(0, high=nrow, size=n), dtype=torch.int32, device=device)
    cols = torch.tensor(rng.integers(0, high=ncol, size=n), dtype=torch.int32, device=device)
    vals = torch.tensor(rng.random(size=(n, 1, 2)), dtype=torch.float32, device=device)
    A = sparse.bsr_from_triplets(nrow, ncol, rows, cols, vals)
    coarse = sparse.bsr_copy(A, block_shape=(4, 4))
    ref = _dense_of(sparse.bsr_diag(sparse.bsr_get_diag(coarse)))
    masked = sparse.bsr_diag(rows_of_blocks=4, block_type=coarse.dtype, device=device)
    sparse.bsr_assign(src=A, dest=masked, topology="masked")
    np.testing.assert_allclose(_dense_of(masked), ref, atol=1e-4)
    # the masked destination keeps its diagonal topology
    np.testing.assert_array_equal(masked.columns[:4].cpu().numpy(), [0, 1, 2, 3])


def test_bsr_copy_scale_reference_scenario(device):
    """TestSparse.test_bsr_copy_scale (test_sparse.py:1393-1415)."""
    nrow, bsize = 6, 2
    D = sparse.bsr_diag(diag=(np.eye(bsize) * 2.0).tolist(), rows_of_blocks=nrow, device=device)
    C = sparse.bsr_copy(D, scalar_type=wa.float64)
    assert C.scalar_type is wa.float64 and C.block_shape == (bsize, bsize)
    assert C.values.dtype == torch.float64
    sparse.bsr_scale(x=C, alpha=0.5)
    np.testing.assert_allclose(_dense_of(C), np.eye(nrow * bsize), atol=1e-4)
    sparse.bsr_scale(x=C, alpha=0.0)
    assert C.nrow == nrow and C.ncol == nrow and C.nnz == D.nnz
    pruned = sparse.bsr_from_triplets(C.nrow, C.ncol, C.uncompress_rows(), C.columns[: C.nnz], C.values[: C.nnz],
                                      prune_numerical_zeros=True)
    assert pruned.nnz_sync() == 0


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_bsr_mm_axpy_into_own_topology_vs_oracle(device, oracle_mod, dtype):
    """Compact product / sum whose blocks all exist in the destination already
    (the union topology is the destination's own): the rebuild is skipped and the
    result must still equal the oracle's compact one bit for bit -- topology,
    values, and kept zero blocks scaled by a negative beta (0 + beta * 0 is
    +0.0 in the rebuild, not -0.0)."""
    rng = np.random.default_rng(43)

    def rand(nrow, ncol, nnz, bshape):
        return _from_triplets_np(device, nrow, ncol, rng.integers(0, nrow, nnz), rng.integers(0, ncol, nnz),
                                 rng.standard_normal((nnz, *bshape)).astype(dtype))

    x, y = rand(120, 90, 600, (3, 3)), rand(90, 110, 500, (3, 3))
    z = sparse.bsr_mm(x, y)                        # z holds exactly the product topology
    nz = z.nnz_sync()
    z.values[:nz] = torch.as_tensor(rng.standard_normal((nz, 3, 3)).astype(dtype), device=device)
    ox, oy, oz = _as_dict(x), _as_dict(y), _as_dict(z)
    sparse.bsr_mm(x, y, z, alpha=-0.75, beta=1.0)
    ref = oracle_mod.bsr_mm(ox, oy, oz, alpha=-0.75, beta=1.0)
    got = _as_dict(z)
    np.testing.assert_array_equal(got["offsets"], ref["offsets"])
    np.testing.assert_array_equal(got["columns"], ref["columns"])
    assert got["values"].tobytes() == ref["values"].tobytes()
    # a diagonal projector times the matrix (the Dirichlet projection's products)
    d = rng.integers(0, 120, 30)
    P = _from_triplets_np(device, 120, 120, d, d, rng.standard_normal((d.size, 3, 3)).astype(dtype))
    A = rand(120, 120, 900, (3, 3))
    C = sparse.bsr_copy(A)
    oP, oA, oC = _as_dict(P), _as_dict(A), _as_dict(C)
    sparse.bsr_mm(P, A, C, alpha=-1.0, beta=1.0)
    ref = oracle_mod.bsr_mm(oP, oA, oC, alpha=-1.0, beta=1.0)
    got = _as_dict(C)
    np.testing.assert_array_equal(got["columns"], ref["columns"])
    assert got["values"].tobytes() == ref["values"].tobytes()
    oC = _as_dict(C)
    sparse.bsr_mm(C, P, A, alpha=-1.0, beta=1.0)   # P block-diagonal: C P lies inside A's topology
    ref = oracle_mod.bsr_mm(oC, oP, oA, alpha=-1.0, beta=1.0)
    got = _as_dict(A)
    np.testing.assert_array_equal(got["offsets"], ref["offsets"])
    np.testing.assert_array_equal(got["columns"], ref["columns"])
    assert got["values"].tobytes() == ref["values"].tobytes()
    # sum of a matrix into the union topology it already has
    w = rand(120, 110, 400, (3, 3))
    u = sparse.bsr_axpy(w, sparse.bsr_copy(z), alpha=1.0, beta=1.0)   # union of w and z
    nu = u.nnz_sync()
    u.values[:nu] = torch.as_tensor(rng.standard_normal((nu, 3, 3)).astype(dtype), device=device)
    u.values[:nu:7] = 0.0
    ow, ou = _as_dict(w), _as_dict(u)
    sparse.bsr_axpy(w, u, alpha=0.5, beta=-2.0)
    ref = oracle_mod.bsr_axpy(ow, ou, alpha=0.5, beta=-2.0)
    got = _as_dict(u)
    np.testing.assert_array_equal(got["offsets"], ref["offsets"])
    np.testing.assert_array_equal(got["columns"], ref["columns"])
    assert got["values"].tobytes() == ref["values"].tobytes()


def test_bsr_copy_same_shape_is_exact(device):
    """A compact same-shape copy keeps the source's topology and copies the
    values as (dest type)(scale * value), scale in the source type; a padded
    source is compacted."""
    rng = np.random.default_rng(47)
    rows, cols = rng.integers(0, 80, 700), rng.integers(0, 60, 700)
    A = _from_triplets_np(device, 80, 60, rows, cols, rng.standard_normal((700, 3, 3)).astype(np.float32))
    n = A.nnz_sync()
    A.values[3] = -0.0
    C = sparse.bsr_copy(A)
    assert C.nnz_sync() == n and C.row_counts is None
    np.testing.assert_array_equal(C.offsets.cpu().numpy(), A.offsets[:81].cpu().numpy())
    np.testing.assert_array_equal(C.columns[:n].cpu().numpy(), A.columns[:n].cpu().numpy())
    assert C.values[:n].cpu().numpy().tobytes() == A.values[:n].cpu().numpy().tobytes()
    D = sparse.bsr_copy(A * 2.5, scalar_type=wa.float64)
    expect = (np.float32(2.5) * A.values[:n].cpu().numpy()).astype(np.float64)
    assert D.values.dtype == torch.float64 and D.values[:n].cpu().numpy().tobytes() == expect.tobytes()
    # padded source: active blocks only, compacted
    P = sparse.bsr_copy(A, topology="padded")
    P.row_counts = torch.clamp(P.row_counts - 1, min=0)
    Q = sparse.bsr_copy(P)
    dq, _ = _active_dense(Q)
    dp, _ = _active_dense(P)
    np.testing.assert_array_equal(dq, dp)
    assert Q.row_counts is None and Q.nnz_sync() == int(P.row_counts.sum())


def test_degenerate_inputs(device, oracle_mod):
    """Empty and one-row cases: no triplets, a device count of zero, every
    triplet out of range, one row of many duplicates, products and sums with an
    empty operand, and bsr_mv of an empty matrix."""
    e32 = torch.zeros(0, dtype=torch.int32, device=device)
    A = sparse.bsr_from_triplets(7, 5, e32, e32, torch.zeros((0, 3, 3), device=device))
    assert A.nnz_sync() == 0
    np.testing.assert_array_equal(A.offsets.cpu().numpy(), np.zeros(8, np.int32))
    x = torch.randn(15, device=device)
    y = torch.full((21,), 3.0, device=device)
    sparse.bsr_mv(A, x, y, alpha=1.0, beta=0.5)
    np.testing.assert_array_equal(y.cpu().numpy(), np.full(21, 1.5, np.float32))
    # a device count of zero over non-empty arrays
    r = torch.tensor([0, 1, 2], dtype=torch.int32, device=device)
    B = sparse.bsr_zeros(3, 3, wa.float32, device=device)
    sparse.bsr_set_from_triplets(B, r, r, torch.ones(3, device=device), count=torch.zeros(1, dtype=torch.int32,
                                                                                          device=device))
    assert B.nnz_sync() == 0
    # every triplet outside the matrix
    bad = torch.tensor([-1, 3, 7, -5], dtype=torch.int32, device=device)
    C = sparse.bsr_from_triplets(3, 3, bad, torch.tensor([0, 0, 1, 1], dtype=torch.int32, device=device),
                                 torch.ones(4, device=device))
    assert C.nnz_sync() == 0
    # one row, 5000 duplicates of one block plus a second column: sums in triplet order
    rng = np.random.default_rng(53)
    rows = np.zeros(5000, np.int32)
    cols = (rng.random(5000) < 0.3).astype(np.int32)
    vals = rng.standard_normal((5000, 3, 3)).astype(np.float32)
    D = _from_triplets_np(device, 1, 2, rows, cols, vals)
    o = oracle_mod.bsr_from_triplets(1, 2, rows, cols, vals)
    assert D.nnz_sync() == 2
    np.testing.assert_array_equal(D.columns[:2].cpu().numpy(), o["columns"])
    assert D.values[:2].cpu().numpy().tobytes() == o["values"].tobytes()
    # products and sums with an empty operand
    Z = sparse.bsr_zeros(1, 2, wa.mat33, device=device)
    P = sparse.bsr_mm(D, sparse.bsr_zeros(2, 2, wa.mat33, device=device))
    assert P.nnz_sync() == 0
    S = sparse.bsr_axpy(Z, sparse.bsr_copy(D), alpha=2.0, beta=1.0)
    assert S.nnz_sync() == 2 and S.values[:2].cpu().numpy().tobytes() == o["values"].tobytes()
    T_ = sparse.bsr_transposed(A)
    assert (T_.nrow, T_.ncol) == (5, 7) and T_.nnz_sync() == 0
    # zero-row matrix
    E = sparse.bsr_zeros(0, 4, wa.float32, device=device)
    out = sparse.bsr_mv(E, torch.randn(4, device=device))
    assert out.numel() == 0


@pytest.mark.parametrize("bshape,dtype", [((1, 1), torch.float32), ((3, 3), torch.float32), ((3, 3), torch.float64)])
def test_bsr_mv_transpose_symmetric_pattern_is_deterministic(device, bshape, dtype):
    """A square matrix whose pattern is structurally symmetric (values are not)
    takes the gather form of A^T x: no atomics, the same bits on every run, and
    the bits of bsr_mv on the explicit transpose (same lanes, same order)."""
    rng = np.random.default_rng(19)
    n, T = 400, 3000
    r = rng.integers(0, n, T)
    c = rng.integers(0, n, T)
    rows = np.concatenate([r, c]).astype(np.int32)
    cols = np.concatenate([c, r]).astype(np.int32)
    vals = rng.standard_normal((2 * T, *bshape))
    A = sparse.bsr_zeros(n, n, wa.matrix(bshape, wa.types.type_to_scalar(dtype)) if bshape != (1, 1)
                         else wa.types.type_to_scalar(dtype), device=device)
    v = vals.reshape(-1) if bshape == (1, 1) else vals
    sparse.bsr_set_from_triplets(A, _t(rows, device), _t(cols, device), _t(v, device, dtype))
    A.nnz_sync()
    assert sparse._mirror_symmetric(A)
    AT = sparse.bsr_transposed(A)
    x = _t(rng.standard_normal(n * bshape[0]), device, dtype)
    y0 = rng.standard_normal(n * bshape[1])
    D = _dense_of(A)
    for alpha, beta in [(1.0, 0.0), (-0.5, 2.0), (2.0, 1.0)]:
        outs = []
        for _ in range(3):
            y = _t(y0, device, dtype)
            sparse.bsr_mv(A, x, y, alpha, beta, transpose=True)
            outs.append(y.cpu().numpy().tobytes())
        assert outs[0] == outs[1] == outs[2]
        yt = _t(y0, device, dtype)
        sparse.bsr_mv(AT, x, yt, alpha, beta)
        assert outs[0] == yt.cpu().numpy().tobytes()
        ref = alpha * D.T @ x.cpu().numpy().astype(np.float64) + beta * y0
        tol = dict(rtol=1e-4, atol=1e-4) if dtype == torch.float32 else dict(rtol=1e-10, atol=1e-10)
        np.testing.assert_allclose(np.frombuffer(outs[0], dtype=np.float32 if dtype == torch.float32
                                                 else np.float64).astype(np.float64), ref, **tol)
    # one block without its mirror: the atomic scatter (still correct to rounding)
    rows2 = np.append(rows, 0).astype(np.int32)
    cols2 = np.append(cols, n - 1).astype(np.int32)
    rows2[np.where((rows2 == n - 1) & (cols2 == 0))] = 1
    B = sparse.bsr_zeros(n, n, A.dtype, device=device)
    v2 = np.concatenate([v, v[:1]])
    sparse.bsr_set_from_triplets(B, _t(rows2, device), _t(cols2, device), _t(v2, device, dtype))
    DB = _dense_of(B)
    if not np.array_equal(DB != 0, (DB != 0).T):
        assert not sparse._mirror_symmetric(B)
