"""CPU oracle for the sparse-solver path -- TEST INFRASTRUCTURE ONLY.

Wraps ``warp_oracle.c`` (a plain-C restatement of the reference algorithms,
each function citing the reference lines it follows), holds a few numpy
restatements of index-only algorithms (the transposed topology), and, when it
has been built, loads ``_ref/libwarp_ref.so``: the reference's own CPU native
library compiled from its unmodified sources (``make ref``), used only to
check the restatements' topologies.

Only tests, the package smoke check and the bench's ``cpu_baseline`` leg use
this module, always as the checker / baseline and never as the product path.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libwarp_ref.so")

_lib = None
_ref = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


def build(ref: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_bsr_from_triplets.restype = _I
        L.orc_bsr_from_triplets.argtypes = [_I, _I, _I, _I, _I, _P, _P, _P, _P, ctypes.c_uint64, _P, _P, _P, _P, _P]
        for s, S in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            fn = getattr(L, f"orc_bsr_mv_{s}")
            fn.restype = None
            fn.argtypes = [_I, _I, _I, _P, _P, _P, _P, _P, _P, S, S]
            fn = getattr(L, f"orc_bsr_mv_grouped_{s}")
            fn.restype = None
            fn.argtypes = [_I, _I, _I, _P, _P, _P, _P, _P, _P, S, S, _I]
            fn = getattr(L, f"orc_bsr_get_diag_{s}")
            fn.restype = None
            fn.argtypes = [_I, _I, _I, _I, _P, _P, _P, _P, _P]
            fn = getattr(L, f"orc_inverse_diag_{s}")
            fn.restype = None
            fn.argtypes = [_I, _I, _I, _P, _I, _P]
        for s in ("f32", "f64", "f32d"):
            fn = getattr(L, f"orc_cg_{s}")
            fn.restype = None
            fn.argtypes = [_I, _I, _P, _P, _P, _P, _P, _D, _D, _I, _I, _P, _I, _I, _I, _P, _P, _I]
        L.orc_residual_f32_f64.restype = None
        L.orc_residual_f32_f64.argtypes = [_I, _I, _P, _P, _P, _P, _P, _P]
        _lib = L
    return _lib


def ref_lib() -> ctypes.CDLL | None:
    """The reference's CPU native library, or None when it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        L = ctypes.CDLL(REF_LIB)
        L.wp_bsr_matrix_from_triplets_host.restype = None
        L.wp_bsr_matrix_from_triplets_host.argtypes = [
            _I, _I, _I, _I, _I, _P, _P, _P, _P, ctypes.c_uint64, ctypes.c_bool, _P, _P, _P, _P, _P, _P, _P,
        ]
        L.wp_bsr_transpose_host.restype = None
        L.wp_bsr_transpose_host.argtypes = [_I, _I, _I] + [_P] * 8
        L.wp_bsr_compress_inplace_host.restype = None
        L.wp_bsr_compress_inplace_host.argtypes = [
            _I, _I, _I, _I, _I, ctypes.c_bool, ctypes.c_uint64, ctypes.c_bool, _P, _P, _P, _P, ctypes.c_bool, _P, _P,
        ]
        _ref = L
    return _ref


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


_ZERO_MASK = {4: 0x7FFFFFFF, 8: 0x7FFFFFFFFFFFFFFF, 2: 0x7FFF}


def bsr_from_triplets(nrow, ncol, rows, cols, values=None, prune=True, count=None):
    """Compact BSR from triplets (stable order; see warp_oracle.c).

    ``values``: (T,) or (T, br, bc) float32/float64, or None (topology only).
    Returns dict(offsets, columns, values, summed_offsets, summed_indices, nnz).
    """
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    T = rows.shape[0]
    if values is not None:
        values = np.ascontiguousarray(values)
        bshape = values.shape[1:] if values.ndim == 3 else ()
        block_size = int(np.prod(bshape)) if bshape else 1
        sb = values.dtype.itemsize
        mask = _ZERO_MASK[sb] if prune else 0
    else:
        bshape, block_size, sb, mask = (), 1, 0, 0
    so = np.zeros(max(T, 1), np.int32)
    si = np.zeros(max(T, 1), np.int32)
    off = np.zeros(nrow + 1, np.int32)
    colsout = np.zeros(max(T, 1), np.int32)
    out_vals = np.zeros((max(T, 1), *bshape), values.dtype) if values is not None else None
    cnt = None if count is None else np.array([count], np.int32)
    nnz = lib().orc_bsr_from_triplets(
        block_size, sb, nrow, ncol, T, _p(cnt), _p(rows), _p(cols), _p(values), mask, _p(so), _p(si), _p(off),
        _p(colsout), _p(out_vals),
    )
    valid = int(so[nnz - 1]) if nnz > 0 else 0
    return dict(
        offsets=off,
        columns=colsout[:nnz],
        values=None if out_vals is None else out_vals[:nnz],
        summed_offsets=so[:nnz],
        summed_indices=si[:valid],
        nnz=nnz,
    )


def ref_bsr_from_triplets(nrow, ncol, rows, cols, values=None, prune=True, count=None):
    """Topology from the reference's own host function (sparse.cpp:267-413)."""
    L = ref_lib()
    if L is None:
        return None
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    T = rows.shape[0]
    if values is not None:
        values = np.ascontiguousarray(values)
        block_size = int(np.prod(values.shape[1:])) if values.ndim == 3 else 1
        sb = values.dtype.itemsize
        mask = _ZERO_MASK[sb] if prune else 0
    else:
        block_size, sb, mask = 1, 0, 0
    so = np.zeros(max(T, 1), np.int32)
    si = np.zeros(max(T, 1), np.int32)
    off = np.zeros(nrow + 1, np.int32)
    colsout = np.zeros(max(T, 1), np.int32)
    nnz = np.zeros(1, np.int32)
    cnt = None if count is None else np.array([count], np.int32)
    L.wp_bsr_matrix_from_triplets_host(
        block_size, sb, nrow, ncol, T, _p(cnt), _p(rows), _p(cols), _p(values), mask, False, _p(so), _p(si), _p(off),
        None, _p(colsout), _p(nnz), None,
    )
    n = int(nnz[0])
    return dict(offsets=off, columns=colsout[:n], summed_offsets=so[:n], summed_indices=si, nnz=n)


def ref_bsr_compress_inplace(offsets, row_counts, columns, values, prune=True, make_compact=True):
    """The reference's own in-place row compression (wp_bsr_compress_inplace_host,
    sparse.cpp:521-607; rows: sparse_util.h:82-150) on copies of the arrays:
    per row, zero blocks pruned, a stable insertion sort by column, duplicates
    summed left to right.  Returns dict(offsets, row_counts, columns, values,
    nnz) or None when the reference library was not built."""
    L = ref_lib()
    if L is None:
        return None
    off = np.array(offsets, np.int32, copy=True)
    nrow = off.shape[0] - 1
    rc = None if row_counts is None else np.array(row_counts, np.int32, copy=True)
    cols = np.array(columns, np.int32, copy=True)
    vals = np.array(values, copy=True, order="C")
    block_size = int(np.prod(vals.shape[1:])) if vals.ndim > 1 else 1
    sb = vals.dtype.itemsize
    scalar_type = 1 if vals.dtype == np.float64 else 0
    nnz = np.zeros(1, np.int32)
    L.wp_bsr_compress_inplace_host(
        nrow, block_size, sb, scalar_type, int(cols.shape[0]), bool(prune), _ZERO_MASK[sb], bool(make_compact),
        _p(off), _p(rc), _p(cols), _p(vals), True, _p(nnz), None,
    )
    return dict(offsets=off, row_counts=rc, columns=cols, values=vals, nnz=int(nnz[0]))


def _active_positions(nrow, offsets, row_counts, T):
    """(positions, rows) of the active blocks below T; a row's active range
    never passes its capacity end (sparse.cpp:445-447)."""
    offsets = np.asarray(offsets, np.int64)
    T = int(min(T, offsets[nrow])) if nrow > 0 else 0
    if T <= 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    ends = offsets[1 : nrow + 1] if row_counts is None else np.minimum(
        offsets[:nrow] + np.asarray(row_counts, np.int64)[:nrow], offsets[1 : nrow + 1])
    p = np.arange(T, dtype=np.int64)
    r = np.searchsorted(offsets[: nrow + 1], p, side="right") - 1
    keep = p < ends[r]
    return p[keep], r[keep]


def bsr_transpose(nrow, ncol, offsets, columns, row_counts=None, T=None, capacity=None):
    """Transposed topology (reference wp_bsr_transpose_host, sparse.cpp:415-519;
    device twin sparse.cu:1400-1492).  The active blocks are ordered by
    (column, row) -- the host's comparator -- with ties (repeated blocks) by
    position, the order of the device's stable radix sort.

    capacity None: compact result dict(offsets, columns, block_map).
    capacity = transposed offsets read as row capacity: padded result
    dict(row_counts, columns, block_map, status) over the capacity, with
    columns -1 where unwritten and block_map -1 on unused slots; a row that
    does not fit is left empty and sets status 1.
    """
    columns = np.asarray(columns, np.int64)
    if T is None:
        T = int(np.asarray(offsets)[nrow])
    p, r = _active_positions(nrow, offsets, row_counts, T)
    c = columns[p]
    order = np.lexsort((p, r, c))
    sp, t_cols, t_rows = p[order], r[order], c[order]
    counts = np.bincount(t_rows, minlength=ncol).astype(np.int64)
    t_off = np.concatenate([[0], np.cumsum(counts)])
    if capacity is None:
        return dict(offsets=t_off.astype(np.int32), columns=t_cols.astype(np.int32), block_map=sp.astype(np.int32))
    capacity = np.asarray(capacity, np.int64)
    slots = int(capacity[ncol])
    out_cols = np.full(slots, -1, np.int32)
    out_map = np.full(slots, -1, np.int32)
    out_cnt = np.zeros(ncol, np.int32)
    status = 0
    for row in range(ncol):
        beg, cap = capacity[row], capacity[row + 1] - capacity[row]
        n = counts[row]
        if n > cap:
            status = 1
            continue
        out_cnt[row] = n
        out_cols[beg : beg + n] = t_cols[t_off[row] : t_off[row] + n]
        out_map[beg : beg + n] = sp[t_off[row] : t_off[row] + n]
    return dict(row_counts=out_cnt, columns=out_cols, block_map=out_map, status=status)


def ref_bsr_transpose(nrow, ncol, offsets, columns, row_counts=None, T=None, capacity=None):
    """The same, from the reference's own host function (sparse.cpp:415-519),
    or None when it was not built.  Same result layout as bsr_transpose."""
    L = ref_lib()
    if L is None:
        return None
    offsets = np.ascontiguousarray(offsets, np.int32)
    columns = np.ascontiguousarray(columns, np.int32)
    rc = None if row_counts is None else np.ascontiguousarray(row_counts, np.int32)
    if T is None:
        T = int(offsets[nrow])
    if capacity is None:
        t_off = np.zeros(ncol + 1, np.int32)
        t_cols = np.zeros(max(T, 1), np.int32)
        bmap = np.zeros(max(2 * T, 1), np.int32)
        L.wp_bsr_transpose_host(nrow, ncol, T, _p(offsets), _p(rc), _p(columns), _p(t_off), None, _p(t_cols),
                                _p(bmap), None)
        n = int(t_off[ncol])
        return dict(offsets=t_off, columns=t_cols[:n], block_map=bmap[:n])
    cap = np.ascontiguousarray(capacity, np.int32)
    slots = int(cap[ncol])
    t_cnt = np.zeros(ncol, np.int32)
    t_cols = np.full(max(slots, 1), -1, np.int32)
    bmap = np.zeros(max(slots, 2 * T, 1), np.int32)
    status = np.zeros(1, np.int32)
    L.wp_bsr_transpose_host(nrow, ncol, T, _p(offsets), _p(rc), _p(columns), _p(cap), _p(t_cnt), _p(t_cols),
                            _p(bmap), _p(status))
    return dict(row_counts=t_cnt, columns=t_cols[:slots], block_map=bmap[:slots], status=int(status[0]))


def _row_ids(offsets):
    offsets = np.asarray(offsets, np.int64)
    return np.repeat(np.arange(offsets.size - 1), np.diff(offsets))


def _find_blocks(offsets, columns, ncol, rows, cols):
    """Position of blocks (rows[i], cols[i]) in a compact row-sorted matrix, -1 if absent."""
    keys = _row_ids(offsets) * np.int64(ncol) + np.asarray(columns, np.int64)
    q = np.asarray(rows, np.int64) * np.int64(ncol) + np.asarray(cols, np.int64)
    pos = np.searchsorted(keys, q)
    ok = pos < keys.size
    ok[ok] = keys[pos[ok]] == q[ok]
    return np.where(ok, pos, -1)


def bsr_mm(x, y, z=None, alpha=1.0, beta=0.0, deep_ranges=True, z_topology=None, tile=0):
    """z := alpha x @ y + beta z on compact matrices given as dicts (nrow, ncol,
    offsets, columns, values[nnz, br, bc]) with sorted unique rows.

    Topology (reference bsr_mm, sparse.py:4200-4640 with bsr_mm_count_coeffs /
    _bsr_mm_list_coeffs, 3620-3735): every product block (row r of x, column c
    of the y rows it reaches), plus z's blocks when beta != 0; with deep_ranges
    a row whose product count exceeds the span of the columns it reaches lists
    that whole span instead.  z_topology = (offsets, columns) fixes the
    topology instead (masked product).

    Values (_bsr_mm_compute_values, 3764-3800, and the add of the old z): each
    block starts at 0 (+ beta z), then acc = sum over the row's x blocks in
    storage order of x_k @ y, each entry x[i,0] y[0,j] + x[i,1] y[1,j] + ...
    left to right, every product and sum rounded in the values' dtype; finally
    z += alpha acc.

    tile > 0 (the tiled-outer values kernel, sparse.py:3813-3909, with `tile`
    lanes per block): the row's (x block k, inner index m) pairs, m fastest,
    are dealt to lanes g = pair % tile; each lane adds its pairs' rank-one
    updates x_k[:, m] y[m, :] in order (product rounded, then added; pairs
    without a y block skipped), then the lanes combine by an xor butterfly
    (a[g] = a[g] + a[g ^ off], off = tile/2 .. 1) and lane 0 holds acc.  The
    reference's own lane sum (tile_sum) has no fixed order, so this is the
    GPU kernels' order, checked against dense products for the values."""
    dt = x["values"].dtype
    xo, xc, xv = np.asarray(x["offsets"], np.int64), np.asarray(x["columns"], np.int64), x["values"]
    yo, yc, yv = np.asarray(y["offsets"], np.int64), np.asarray(y["columns"], np.int64), y["values"]
    nrow, ncol = x["nrow"], y["ncol"]
    br, bk, bc = xv.shape[1], xv.shape[2], yv.shape[2]
    if z_topology is not None:
        out_off = np.asarray(z_topology[0], np.int64)
        out_cols = np.asarray(z_topology[1], np.int64)
    else:
        cand_rows, cand_cols = [], []
        if beta != 0.0 and z is not None:
            cand_rows.append(_row_ids(z["offsets"]))
            cand_cols.append(np.asarray(z["columns"], np.int64))
        for r in range(nrow):
            segs = [yc[yo[c]:yo[c + 1]] for c in xc[xo[r]:xo[r + 1]]]
            segs = [sg for sg in segs if sg.size]
            if not segs:
                continue
            count = sum(sg.size for sg in segs)
            lo, hi = min(sg[0] for sg in segs), max(sg[-1] for sg in segs)
            cols = np.arange(lo, hi + 1) if deep_ranges and count > hi - lo + 1 else np.concatenate(segs)
            cand_rows.append(np.full(cols.size, r))
            cand_cols.append(cols)
        rr = np.concatenate(cand_rows) if cand_rows else np.zeros(0, np.int64)
        cc = np.concatenate(cand_cols) if cand_cols else np.zeros(0, np.int64)
        keys = np.unique(rr * np.int64(ncol) + cc)
        out_cols = keys % ncol
        out_off = np.concatenate([[0], np.cumsum(np.bincount(keys // ncol, minlength=nrow))])
    out_rows = _row_ids(out_off)
    nb = out_cols.size
    vals = np.zeros((nb, br, bc), dt)
    if beta != 0.0 and z is not None:
        pos = _find_blocks(z["offsets"], z["columns"], ncol, out_rows, out_cols)
        has = pos >= 0
        vals[has] = vals[has] + np.asarray(beta, dt) * z["values"][pos[has]]
    acc = np.zeros((nb, br, bc), dt)
    xlen = np.diff(xo)
    if tile > 0 and nb:
        lanes = np.zeros((nb, tile, br, bc), dt)
        npairs = xlen[out_rows] * bk
        for c in range(int(npairs.max())):
            g, m = c % tile, c % bk
            live = np.flatnonzero(npairs > c)
            k = xo[out_rows[live]] + c // bk
            yb = _find_blocks(yo, yc, ncol, xc[k], out_cols[live])
            ok = yb >= 0
            live, k, yb = live[ok], k[ok], yb[ok]
            upd = xv[k][:, :, m][:, :, None] * yv[yb][:, m, :][:, None, :]
            lanes[live, g] = lanes[live, g] + upd
        off = tile // 2
        while off > 0:
            lanes = lanes + lanes[:, np.arange(tile) ^ off]
            off //= 2
        acc = lanes[:, 0]
    step = 0
    while nb and tile <= 0 and step < (xlen[out_rows].max() if nb else 0):
        live = np.flatnonzero(xlen[out_rows] > step)
        k = xo[out_rows[live]] + step
        yb = _find_blocks(yo, yc, ncol, xc[k], out_cols[live])
        ok = yb >= 0
        live, k, yb = live[ok], k[ok], yb[ok]
        a, b = xv[k], yv[yb]
        prod = np.empty((live.size, br, bc), dt)
        for i in range(br):
            for j in range(bc):
                t = a[:, i, 0] * b[:, 0, j]
                for m in range(1, bk):
                    t = t + a[:, i, m] * b[:, m, j]
                prod[:, i, j] = t
        acc[live] = acc[live] + prod
        step += 1
    vals = vals + np.asarray(alpha, dt) * acc
    return dict(offsets=out_off.astype(np.int32), columns=out_cols.astype(np.int32), values=vals)


def bsr_axpy(x, y, alpha=1.0, beta=1.0):
    """y := alpha x + beta y on compact matrices (dicts as for bsr_mm):
    union topology; blocks 0 + beta y, then + alpha x, each op rounded in the
    values' dtype (reference bsr_axpy, sparse.py:3480-3617, with
    _bsr_axpy_add_block)."""
    dt = y["values"].dtype
    nrow, ncol = y["nrow"], y["ncol"]
    ry, rx = _row_ids(y["offsets"]), _row_ids(x["offsets"])
    keys = np.unique(np.concatenate([ry * np.int64(ncol) + np.asarray(y["columns"], np.int64),
                                     rx * np.int64(ncol) + np.asarray(x["columns"], np.int64)]))
    cols = keys % ncol
    off = np.concatenate([[0], np.cumsum(np.bincount(keys // ncol, minlength=nrow))])
    rows = keys // ncol
    vals = np.zeros((keys.size, *y["values"].shape[1:]), dt)
    py = _find_blocks(y["offsets"], y["columns"], ncol, rows, cols)
    px = _find_blocks(x["offsets"], x["columns"], ncol, rows, cols)
    hy, hx = py >= 0, px >= 0
    vals[hy] = vals[hy] + np.asarray(beta, dt) * y["values"][py[hy]]
    vals[hx] = vals[hx] + np.asarray(alpha, dt) * x["values"][px[hx]]
    return dict(offsets=off.astype(np.int32), columns=cols.astype(np.int32), values=vals)


def bsr_mv(offsets, columns, values, x, y=None, alpha=1.0, beta=0.0, row_counts=None, group=0):
    """y = alpha A x + beta y.  group 0: reference row-kernel order
    (sparse.py:4627-4665); group G > 0: the GPU lane-group order."""
    values = np.ascontiguousarray(values)
    dt = values.dtype
    br, bc = (values.shape[1], values.shape[2]) if values.ndim == 3 else (1, 1)
    nrow = offsets.shape[0] - 1
    x = np.ascontiguousarray(x, dtype=dt).reshape(-1)
    yo = np.zeros(nrow * br, dt) if y is None else np.ascontiguousarray(y, dtype=dt).reshape(-1).copy()
    fn = lib().orc_bsr_mv_grouped_f64 if dt == np.float64 else lib().orc_bsr_mv_grouped_f32
    offsets = np.ascontiguousarray(offsets, np.int32)
    columns = np.ascontiguousarray(columns, np.int32)
    rc = None if row_counts is None else np.ascontiguousarray(row_counts, np.int32)
    fn(nrow, br, bc, _p(offsets), _p(rc), _p(columns), _p(values), _p(x), _p(yo), alpha, beta, int(group))
    return yo


def bsr_get_diag(offsets, columns, values, ncol=None, row_counts=None):
    values = np.ascontiguousarray(values)
    dt = values.dtype
    br, bc = (values.shape[1], values.shape[2]) if values.ndim == 3 else (1, 1)
    nrow = offsets.shape[0] - 1
    ncol = nrow if ncol is None else ncol
    dim = min(nrow, ncol)
    out = np.zeros((dim, br, bc), dt)
    fn = lib().orc_bsr_get_diag_f64 if dt == np.float64 else lib().orc_bsr_get_diag_f32
    offsets = np.ascontiguousarray(offsets, np.int32)
    columns = np.ascontiguousarray(columns, np.int32)
    rc = None if row_counts is None else np.ascontiguousarray(row_counts, np.int32)
    fn(nrow, ncol, br, bc, _p(offsets), _p(rc), _p(columns), _p(values), _p(out))
    return out if values.ndim == 3 else out.reshape(dim)


def inverse_diag(diag_blocks, use_abs=False):
    d = np.ascontiguousarray(diag_blocks)
    dt = d.dtype
    br, bc = (d.shape[1], d.shape[2]) if d.ndim == 3 else (1, 1)
    dim = d.shape[0]
    out = np.zeros(dim * br, dt)
    fn = lib().orc_inverse_diag_f64 if dt == np.float64 else lib().orc_inverse_diag_f32
    fn(dim, br, bc, _p(d), 1 if use_abs else 0, _p(out))
    return out


PREC_NONE, PREC_DIAG, PREC_BLOCK = 0, 1, 2


def cg(offsets, columns, values, b, x0, tol=None, atol=None, maxiter=0, prec_kind=PREC_NONE, prec=None,
       check_every=10, restart=0, dots="fp64", history_cap=4096, group=0):
    """Reference-semantics CG (see orc_cg in warp_oracle.c).

    dots: "fp64" (FP64 dots and alpha/beta, fp32 vectors) or "native" (all in
    the system scalar type, the reference's precision).  group: SpMV order
    (0 reference row order, G > 0 the GPU lane-group order).
    Returns (x, iterations, residual_norm, atol, history).
    """
    values = np.ascontiguousarray(values)
    dt = values.dtype
    br = values.shape[1] if values.ndim == 3 else 1
    nrow = offsets.shape[0] - 1
    b = np.ascontiguousarray(b, dtype=dt).reshape(-1)
    x = np.ascontiguousarray(x0, dtype=dt).reshape(-1).copy()
    if dt == np.float64:
        fn = lib().orc_cg_f64
    else:
        fn = lib().orc_cg_f32d if dots == "fp64" else lib().orc_cg_f32
    pr = None if prec is None else np.ascontiguousarray(prec, dtype=dt).reshape(-1)
    out = np.zeros(4)
    hist = np.zeros(history_cap)
    nan = float("nan")
    fn(
        nrow, br, _p(np.ascontiguousarray(offsets, np.int32)), _p(np.ascontiguousarray(columns, np.int32)),
        _p(values), _p(b), _p(x), nan if tol is None else float(tol), nan if atol is None else float(atol),
        int(maxiter), int(prec_kind), _p(pr), int(check_every), int(restart), int(group), _p(out), _p(hist),
        history_cap,
    )
    return x, int(out[0]), float(out[1]), float(out[2]), hist[: int(out[3])]


def residual_f32_f64(offsets, columns, values, x, b):
    values = np.ascontiguousarray(values, dtype=np.float32)
    br = values.shape[1] if values.ndim == 3 else 1
    nrow = offsets.shape[0] - 1
    r = np.zeros(nrow * br)
    lib().orc_residual_f32_f64(
        nrow, br, _p(np.ascontiguousarray(offsets, np.int32)), _p(np.ascontiguousarray(columns, np.int32)),
        _p(values), _p(np.ascontiguousarray(x, np.float64).reshape(-1)),
        _p(np.ascontiguousarray(b, np.float64).reshape(-1)), _p(r),
    )
    return r
