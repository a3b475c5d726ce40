"""Benchmark of the sparse-solver hot path on one MI355X.

Headline (`value`): bsr_mv HBM throughput on the 3D Q1 hex elasticity
matrix, 64^3 elements (65^3 nodes x 3 = 823,875 dofs, 3x3 fp32 BSR,
7,189,057 blocks).  One step = one y = A x over the whole matrix, inputs
resident in HBM: `steps` bsr_mv calls back to back, timed by HIP events in
20 groups of consecutive calls (consecutive calls traverse the rows in
opposite directions, as every product loop on one matrix does); value uses
the median over the groups of the time per call.
GB/s counts algorithmic bytes:
    blocks * (36 + 4) + 4 (nrow + 1) + 12 nrow (x) + 12 nrow (y).

Also reported (outside the timed steps): Jacobi-CG on the same system to
tol 1e-6, and triplet assembly of the shuffled 16.8M-triplet element-loop
stream.  cpu_baseline: the oracle's serial C bsr_mv (1 core) on the same
matrix, timed on the host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-extras]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30
COPY_REPS = 5
METRIC = "time to 1e-8 rel. residual: learned-FCG vs Jacobi-CG; bsr_mv HBM GB/s"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def spmv_bytes(nrow: int, nnzb: int, br: int, bc: int, scalar: int = 4) -> int:
    return nnzb * (br * bc * scalar + 4) + 4 * (nrow + 1) + scalar * bc * nrow + scalar * br * nrow


def load_traffic(name: str):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def time_cpu_spmv(offsets, columns, values, x, min_seconds: float = 10.0, max_reps: int = 2000):
    import numpy as np

    import oracle

    y = np.zeros(x.size, np.float32)
    fn = oracle.lib().orc_bsr_mv_f32
    p = oracle._p
    nrow = offsets.size - 1
    args = (nrow, 3, 3, p(offsets), None, p(columns), p(values), p(x), p(y), 1.0, 0.0)
    fn(*args)
    reps = 0
    t0 = time.perf_counter()
    while True:
        fn(*args)
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or reps >= max_reps:
            break
    return el / reps, reps


def cg_iter_bytes(spmv_b: int, ndof: int, scalar: int = 4) -> int:
    """Algorithmic bytes of one preconditioned CG iteration: the SpMV plus the
    fused minimum of 11 vector passes (read Ap, x, r, D, p; write x, r, p;
    re-read r, D, p for the direction)."""
    return spmv_b + 11 * scalar * ndof


def hbm_roofline(nbytes: float, seconds: float, what: str) -> dict:
    gbs = nbytes / seconds / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "alg_bytes": int(nbytes), "basis": what}


def time_cpu_cg(offsets, columns, values, b, prec_kind, prec, iters: int = 20):
    """The oracle's serial CG (1 core) for `iters` iterations (tolerance out of
    reach); returns (ms per iteration, iterations)."""
    import numpy as np

    import oracle

    t0 = time.perf_counter()
    _, it, _, _, _ = oracle.cg(offsets, columns, values, b, np.zeros(b.size, values.dtype), tol=1e-30, atol=1e-30,
                               maxiter=iters, prec_kind=prec_kind, prec=prec, check_every=iters)
    return (time.perf_counter() - t0) * 1e3 / max(it, 1), it


def extra_poisson2d(dev) -> dict:
    """2D Q1 Poisson, 128 x 128 cells (16,129 interior dofs): bsr_mv and
    Jacobi-CG to tol 1e-6 (maxiter 2000)."""
    import numpy as np
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    import torch

    import oracle
    from warp_amd import problems, sparse
    from warp_amd.optim import linear

    rp, cp, vp, bp, ndp = problems.poisson2d_q1(128, 0)
    Ap = sparse.bsr_from_triplets(ndp, ndp, torch.from_numpy(rp).to(dev), torch.from_numpy(cp).to(dev),
                                  torch.from_numpy(vp).to(dev))
    nnzp = Ap.nnz_sync()
    bpt = torch.from_numpy(bp).to(dev)
    xp = torch.zeros_like(bpt)
    yp = torch.empty_like(bpt)
    for _ in range(20):
        sparse.bsr_mv(Ap, xp, yp)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(500):
        sparse.bsr_mv(Ap, bpt, yp)
    e1.record()
    torch.cuda.synchronize()
    us_mv = e0.elapsed_time(e1) / 500 * 1e3
    # host cost of one call (Python + ctypes + launch), no GPU wait
    t0 = time.perf_counter()
    for _ in range(500):
        sparse.bsr_mv(Ap, bpt, yp)
    us_host = (time.perf_counter() - t0) / 500 * 1e6
    torch.cuda.synchronize()
    Mp = linear.preconditioner(Ap, "diag")
    linear.cg(Ap, bpt, xp, tol=1e-6, maxiter=20, M=Mp)
    ms = []
    for _ in range(5):
        xp.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        itp, errp, atolp = linear.cg(Ap, bpt, xp, tol=1e-6, maxiter=2000, M=Mp)
        torch.cuda.synchronize()
        ms.append((time.perf_counter() - t0) * 1e3)
    t_p = float(np.median(ms))
    off = Ap.offsets.cpu().numpy()
    col = Ap.columns[:nnzp].cpu().numpy()
    val = Ap.values[:nnzp].cpu().numpy()
    xh = xp.cpu().numpy().astype(np.float64)
    r = oracle.residual_f32_f64(off, col, val, xh, bp.astype(np.float64))
    true_rel = float(np.linalg.norm(r) / np.linalg.norm(bp))
    A64 = sp.csr_matrix((val.astype(np.float64), col, off), shape=(ndp, ndp))
    xs = spl.spsolve(A64.tocsc(), bp.astype(np.float64))
    err64 = float(np.linalg.norm(xh - xs) / np.linalg.norm(xs))
    # CPU baseline: the oracle's serial Jacobi-CG, whole solves
    inv = oracle.inverse_diag(oracle.bsr_get_diag(off, col, val))
    reps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < 3.0:
        _, it_c, _, _, _ = oracle.cg(off, col, val, bp, np.zeros(ndp, np.float32), tol=1e-6, maxiter=2000,
                                     prec_kind=oracle.PREC_DIAG, prec=inv)
        reps += 1
    cpu_ms = (time.perf_counter() - t0) * 1e3 / reps
    spmv_b = nnzp * 8 + 4 * (ndp + 1) + 8 * ndp
    iter_b = cg_iter_bytes(spmv_b, ndp)
    us_it = t_p * 1e3 / max(itp, 1)
    log(f"[bench] Poisson 2D Q1 (n={ndp}, nnz={nnzp}): bsr_mv {us_mv:.1f} us (host {us_host:.1f} us/call), "
        f"Jacobi-CG {itp} iterations {t_p:.2f} ms ({us_it:.1f} us/iteration), FP64 true residual {true_rel:.2e}, "
        f"vs FP64 direct solve {err64:.2e}; CPU 1 core {cpu_ms:.1f} ms ({it_c} iterations)")
    return {"workload": "2D Q1 Poisson 128x128 cells, 16,129 dofs, Jacobi-CG tol 1e-6 (atol = tol), maxiter 2000",
            "dofs": ndp, "nnz": nnzp, "bsr_mv_us": round(us_mv, 2), "bsr_mv_host_us_per_call": round(us_host, 2),
            "iterations": itp, "ms": round(t_p, 3), "us_per_iter": round(us_it, 2), "residual": errp, "atol": atolp,
            "fp64_true_rel_residual": true_rel, "fp64_rel_error_vs_direct": err64,
            "roofline": hbm_roofline(iter_b, us_it * 1e-6, "PCG iteration (1.3 MB: L2-resident, launch-bound)"),
            "cpu_baseline": {"value": round(cpu_ms, 3), "unit": "ms/solve", "cores": 1, "kind": "port",
                             "sample": f"{reps} whole serial Jacobi-CG solves of the same system ({it_c} iterations)"}}


def _assembly_fp64_error(o, vals) -> float:
    import numpy as np

    idx = o["summed_indices"]
    ends = o["summed_offsets"].astype(np.int64)
    starts = np.concatenate([[0], ends[:-1]])
    v = vals.reshape(vals.shape[0], -1)[idx].astype(np.float64)
    s64 = np.add.reduceat(v, starts, axis=0)
    mag = np.add.reduceat(np.abs(v), starts, axis=0)
    return float(np.max(np.abs(o["values"].reshape(len(ends), -1).astype(np.float64) - s64) / np.maximum(mag, 1e-300)))


def _time_assembly(A, r, c, v, reps: int = 9):
    """Median GPU time of bsr_set_from_triplets (HIP events around the call on
    its stream; the call returns before the device finishes) and median
    wall-clock time of the call plus the host read-back of the block count;
    returns (gpu seconds, wall seconds, blocks)."""
    import numpy as np
    import torch

    from warp_amd import sparse

    sparse.bsr_set_from_triplets(A, r, c, v)
    A.nnz_sync()
    gpu, wall = [], []
    stream = torch.cuda.current_stream(A.values.device)
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        sparse.bsr_set_from_triplets(A, r, c, v)
        e1.record(stream)
        u = A.nnz_sync()
        wall.append(time.perf_counter() - t0)
        torch.cuda.synchronize()
        gpu.append(e0.elapsed_time(e1) * 1e-3)
    return float(np.median(gpu)), float(np.median(wall)), u


def _cpu_assembly(nrow, r, c, v, with_ref: bool):
    """The oracle's serial stable assembly (topology + sums), and the
    reference's own host topology routine (sparse.cpp:267-413, compiled from
    its sources into oracle/_ref) when that library is present."""
    import oracle

    t0 = time.perf_counter()
    o = oracle.bsr_from_triplets(nrow, nrow, r, c, v)
    t_port = time.perf_counter() - t0
    t_ref = None
    if with_ref and oracle.ref_lib() is not None:
        t0 = time.perf_counter()
        oracle.ref_bsr_from_triplets(nrow, nrow, r, c, v)
        t_ref = time.perf_counter() - t0
    return o, t_port, t_ref


def extra_assembly(n: int, dev) -> dict:
    """bsr_set_from_triplets on the shuffled element-loop triplets of the n^3
    hex mesh (16,777,216 at n = 64)."""
    import torch

    import warp_amd as wa
    from warp_amd import problems, sparse

    r3, c3, v3, nn3 = problems.elasticity_triplets_shuffled(n, device=dev)
    A3 = sparse.bsr_zeros(nn3, nn3, wa.mat33, device=dev)
    t_a, t_wall, u3 = _time_assembly(A3, r3, c3, v3)
    T3 = r3.numel()
    asm_bytes = T3 * 44 + u3 * 40 + 4 * (nn3 + 1)
    rh, ch, vh = r3.cpu().numpy(), c3.cpu().numpy(), v3.cpu().numpy()
    got_off = A3.offsets.cpu().numpy()
    got_val = A3.values[:u3].cpu().numpy()
    del r3, c3, v3, A3
    torch.cuda.empty_cache()
    o, t_port, t_ref = _cpu_assembly(nn3, rh, ch, vh, with_ref=True)
    exact = bool((got_off == o["offsets"]).all() and got_val.tobytes() == o["values"].tobytes())
    err64 = _assembly_fp64_error(o, vh)
    log(f"[bench] assembly: {T3} triplets -> {u3} blocks in {t_a * 1e3:.3f} ms of GPU time ({t_wall * 1e3:.2f} ms "
        f"wall with the block-count read-back) ({T3 / t_a / 1e9:.2f} Gtriplets/s, "
        f"{asm_bytes / t_a / 1e9:.0f} GB/s); bit-exact vs oracle: {exact}; FP32 sums vs FP64 {err64:.2e}; "
        f"CPU 1 core {t_port * 1e3:.0f} ms (oracle), reference host topology "
        f"{'n/a' if t_ref is None else f'{t_ref * 1e3:.0f} ms'}")
    out = {"workload": f"bsr_set_from_triplets, shuffled element-loop triplets of the {n}^3 hex mesh, 3x3 fp32",
           "triplets": T3, "blocks": u3, "ms": round(t_a * 1e3, 3), "wall_ms_with_nnz_readback": round(t_wall * 1e3, 3),
           "timing": "ms: median HIP-event time of the call on its stream over 9 repetitions",
           "triplets_per_s": round(T3 / t_a),
           "alg_GBps": round(asm_bytes / t_a / 1e9, 1), "bit_exact_vs_oracle": exact,
           "fp64_max_err_over_sum_abs": err64,
           "roofline": hbm_roofline(asm_bytes, t_a, "T*(4+4+36) + U*(4+36) + 4*(nrow+1)"),
           "cpu_baseline": {"value": round(T3 / t_port), "unit": "triplets/s", "cores": 1, "kind": "port",
                            "sample": f"one serial stable-sort assembly with value sums of the same {T3} triplets "
                                      f"({t_port * 1e3:.0f} ms)"}}
    if t_ref is not None:
        out["cpu_reference_topology"] = {"value": round(T3 / t_ref), "unit": "triplets/s", "cores": 1,
                                         "kind": "reference",
                                         "sample": f"the reference's wp_bsr_matrix_from_triplets_host on the same "
                                                   f"triplets, topology only ({t_ref * 1e3:.0f} ms)"}
    return out


def extra_assembly_synthetic(n: int, dev) -> dict:
    """20,000,000 triplets over 2,500,000 random blocks of the same 27-point
    pattern (mean multiplicity 8)."""
    from warp_amd import problems, sparse
    import warp_amd as wa

    r6, c6, v6, nn6 = problems.synthetic_triplets(n, device=dev)
    A6 = sparse.bsr_zeros(nn6, nn6, wa.mat33, device=dev)
    t6, _, u6 = _time_assembly(A6, r6, c6, v6)
    T6 = r6.numel()
    b6 = T6 * 44 + u6 * 40 + 4 * (nn6 + 1)
    log(f"[bench] assembly (synthetic): {T6} triplets -> {u6} blocks in {t6 * 1e3:.2f} ms "
        f"({T6 / t6 / 1e9:.2f} Gtriplets/s)")
    return {"triplets": T6, "blocks": u6, "ms": round(t6 * 1e3, 3), "triplets_per_s": round(T6 / t6),
            "alg_GBps": round(b6 / t6 / 1e9, 1),
            "roofline": hbm_roofline(b6, t6, "T*(4+4+36) + U*(4+36) + 4*(nrow+1)")}


def extra_row_compress(n: int, dev) -> dict:
    """fem "row_compress" construction: every block row's candidate blocks
    written in element order into a padded matrix, then sorted and coalesced
    per row."""
    import numpy as np
    import torch

    import warp_amd as wa
    from warp_amd import problems, sparse

    re_, ce_, ve_, _, nn5 = problems.elasticity3d_q1(n, device=dev)
    order = torch.argsort(re_, stable=True)
    counts = torch.bincount(re_, minlength=nn5).to(torch.int32)
    cols_p = ce_[order].contiguous()
    vals_p = ve_[order].contiguous()
    del re_, ce_, ve_, order
    A5 = sparse.bsr_zeros(nn5, nn5, wa.mat33, device=dev, row_capacity=counts)
    ms_c = []
    for rep in range(6):
        A5.row_counts = counts.clone()
        A5.columns = cols_p.clone()
        A5.values = vals_p.clone()
        A5.offsets[1:] = torch.cumsum(counts, 0).to(torch.int32)
        A5.offsets[0] = 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sparse.bsr_compress(A5, prune_numerical_zeros=False)
        u5 = A5.nnz_sync()
        if rep:
            ms_c.append((time.perf_counter() - t0) * 1e3)
    t_c = float(np.median(ms_c))
    T5 = int(counts.sum())
    cb = T5 * 40 + u5 * 40 + 4 * (nn5 + 1)
    log(f"[bench] row-compress construction: {T5} candidates -> {u5} blocks in {t_c:.2f} ms "
        f"({T5 / t_c / 1e6:.2f} G candidates/s)")
    return {"candidates": T5, "blocks": u5, "ms": round(t_c, 3), "candidates_per_s": round(T5 / t_c * 1e3),
            "alg_GBps": round(cb / t_c / 1e6, 1),
            "roofline": hbm_roofline(cb, t_c * 1e-3, "T*(4+36) + U*(4+36) + 4*(nrow+1)")}


def extra_transpose_dirichlet(A, x, y, b, nn: int, nnzb: int, n: int, dev) -> dict:
    """Transpose of the elasticity matrix, the transposed product and its
    Dirichlet projection (the fem steps between assembly and solve): x = 0
    face, identity projector blocks, fixed value 0."""
    import numpy as np
    import torch

    from warp_amd import fem, sparse

    out = {}
    ms_t = []
    for rep in range(4):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        At = sparse.bsr_transposed(A)
        nt = At.nnz_sync()
        if rep:
            ms_t.append((time.perf_counter() - t0) * 1e3)
        del At
    t_t = float(np.median(ms_t))
    t_bytes = nnzb * (8 + 2 * 36) + 8 * (nn + 1)
    log(f"[bench] transpose: {nt} blocks in {t_t:.2f} ms ({t_bytes / t_t / 1e6:.0f} GB/s algorithmic)")
    out["transpose"] = {"blocks": nt, "ms": round(t_t, 3), "alg_GBps": round(t_bytes / t_t / 1e6, 1),
                        "roofline": hbm_roofline(t_bytes, t_t * 1e-3, "nnz*(4+4) + 8*(nrow+1) + 2 value passes")}
    # y = A^T x: the pattern is symmetric, so the deterministic gather; the
    # mirror check runs once per topology (first call), later calls only launch
    yT = torch.empty_like(y)
    sparse.bsr_mv(A, x, yT, transpose=True)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        sparse.bsr_mv(A, x, yT, transpose=True)
    e1.record()
    torch.cuda.synchronize()
    ms_mt = e0.elapsed_time(e1) / 20
    sym = sparse._mirror_symmetric(A)
    log(f"[bench] bsr_mv transpose=True: {ms_mt * 1e3:.1f} us per call "
        f"({'gather, deterministic' if sym else 'atomic scatter'})")
    out["transpose"]["bsr_mv_transpose_ms"] = round(ms_mt, 4)
    out["transpose"]["bsr_mv_transpose_path"] = "gather" if sym else "atomic"
    p1 = n + 1
    face = torch.nonzero(torch.arange(nn, device=dev) % p1 == 0).reshape(-1).to(torch.int32)
    eye = torch.eye(3, device=dev).expand(face.numel(), 3, 3).contiguous()
    ms_d = []
    for rep in range(3):
        Ad = sparse.bsr_copy(A)
        P = sparse.bsr_from_triplets(nn, nn, face, face, eye)
        bd = b.clone()
        fixed = torch.zeros_like(b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fem.project_linear_system(Ad, bd, P, fixed)
        nd = Ad.nnz_sync()
        if rep:
            ms_d.append((time.perf_counter() - t0) * 1e3)
        del Ad, P
    t_d = float(np.median(ms_d))
    log(f"[bench] Dirichlet projection (normalize + rhs + (I-P)A(I-P)+P): {nd} blocks in {t_d:.2f} ms")
    out["dirichlet_projection"] = {"face_nodes": int(face.numel()), "blocks": nd, "ms": round(t_d, 3)}
    return out


def extra_refinement(n4: int, dev) -> dict:
    """FP32 Jacobi-CG + FP64 iterative refinement to 1e-10 on the n4^3
    variable-coefficient 7-point system; the solution checked on the host by
    the oracle's FP64 residual."""
    import numpy as np
    import torch

    import oracle
    import warp_amd as wa
    from warp_amd import problems, sparse
    from warp_amd.optim import refine

    N4 = n4**3
    off4, col4, val4, b4, _ = problems.fv7_problem(n4, 2.3, 16.0 * n4 / 256, 5, 6, device=dev)
    A4 = sparse.bsr_zeros(N4, N4, wa.float32, device=dev)
    A4.offsets, A4.columns, A4.values, A4.nnz = off4, col4, val4, col4.numel()
    b64 = b4.to(torch.float64)
    x64 = torch.zeros(N4, dtype=torch.float64, device=dev)
    refine.cg_refined(A4, b64, x64, tol=1e-10, max_outer=1, inner_maxiter=10)  # kernels warm
    x64.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ir = refine.cg_refined(A4, b64, x64, tol=1e-10, M="diag")
    torch.cuda.synchronize()
    t_ir = time.perf_counter() - t0
    offh, colh, valh = off4.cpu().numpy(), col4.cpu().numpy(), val4.cpu().numpy()
    bh = b64.cpu().numpy()
    r_host = oracle.residual_f32_f64(offh, colh, valh, x64.cpu().numpy(), bh)
    true_rel = float(np.linalg.norm(r_host) / np.linalg.norm(bh))
    nnz4 = int(col4.numel())
    del A4, off4, col4, val4, b4, b64, x64, r_host
    torch.cuda.empty_cache()
    inv = oracle.inverse_diag(oracle.bsr_get_diag(offh, colh, valh))
    cpu_ms, cpu_it = time_cpu_cg(offh, colh, valh, bh.astype(np.float32), oracle.PREC_DIAG, inv, iters=10)
    ms_it = t_ir * 1e3 / max(ir.inner_iterations, 1)
    spmv_b = nnz4 * 8 + 4 * (N4 + 1) + 8 * N4
    iter_b = cg_iter_bytes(spmv_b, N4)
    log(f"[bench] refinement {n4}^3: {ir.outer_iterations} outer / {ir.inner_iterations} inner iterations, "
        f"{t_ir * 1e3:.1f} ms ({ms_it * 1e3:.0f} us per inner iteration, {iter_b / (ms_it * 1e-3) / 1e9:.0f} GB/s), "
        f"FP64 oracle residual {true_rel:.2e}; CPU 1 core {cpu_ms:.0f} ms per FP32 CG iteration")
    return {
        "workload": f"-div(k grad u)=f, 7-point FV {n4}^3, log-k GRF sigma 2.3, corr {16 * n4 // 256}, "
                    "FP32 Jacobi-CG inner (tol 3e-5) + FP64 refinement",
        "dofs": N4, "nnz": nnz4, "tol": 1e-10,
        "outer_iterations": ir.outer_iterations, "inner_iterations": ir.inner_iterations,
        "ms": round(t_ir * 1e3, 2), "ms_per_inner_iteration": round(ms_it, 4),
        "fp64_oracle_rel_residual": true_rel,
        "roofline": hbm_roofline(iter_b, ms_it * 1e-3, "inner FP32 PCG iteration (residual/refinement steps "
                                                        "counted in the time, not the bytes)"),
        "cpu_baseline": {"value": round(cpu_ms, 1), "unit": "ms/inner iteration", "cores": 1, "kind": "port",
                         "sample": f"{cpu_it} iterations of the oracle's serial FP32 Jacobi-CG on the same system"},
    }


# FLOPs of one application of the 4-level U-Net on an n^3 grid: 2*27*ci*co per
# voxel of each convolution's level (layers e0..d0 of warp_amd.optim.learned).
def unet_flops(n: int) -> int:
    from warp_amd.optim.learned import LAYERS

    return sum(2 * 27 * ci * co * (n >> lvl) ** 3 for _, ci, co, _, lvl in LAYERS)


def extra_learned(seeds: int, dev) -> dict:
    """Learned flexible CG vs Jacobi-CG on held-out 128^3 variable-coefficient
    systems, the BF16 MFMA U-Net's rate, and a 1-core CPU baseline of the
    plain method."""
    import numpy as np
    import torch

    import oracle

    ckpt = os.path.join(ROOT, "warp_amd", "data", "fv7_unet_128.npz")
    if seeds <= 0 or not os.path.exists(ckpt):
        return {"skipped": "no checkpoint" if seeds > 0 else "--fcg-seeds 0"}
    from warp_amd.optim import learned

    net, meta = learned.load_checkpoint(ckpt, 128, dev)
    recs = learned.evaluate(net, meta, range(5000, 5000 + seeds), dev, tol=1e-8, log=log)
    med = lambda key: float(np.median([r[key] for r in recs]))  # noqa: E731
    speed = [min(r["jacobi64_ms"], r["ir_ms"]) / r["fcg_ms"] for r in recs]
    # U-Net forward (7 BF16 MFMA launches) on its own, HIP events on its stream
    nvox = 128**3
    inp = torch.zeros(nvox * 2, dtype=torch.int16, device=dev)
    for _ in range(3):
        net.forward_bf16(inp)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        net.forward_bf16(inp)
    e1.record()
    torch.cuda.synchronize()
    unet_s = e0.elapsed_time(e1) / 20 * 1e-3
    # the same forward replayed from a HIP graph, as inside the captured FCG iterations
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream(dev)
    cap.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(cap):
        with torch.cuda.graph(graph, stream=cap):
            net.forward_bf16(inp)
    torch.cuda.current_stream(dev).wait_stream(cap)
    for _ in range(3):
        graph.replay()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(20):
        graph.replay()
    e1.record()
    torch.cuda.synchronize()
    unet_graph_s = e0.elapsed_time(e1) / 20 * 1e-3
    flops = unet_flops(128)
    tf = flops / unet_s / 1e12
    # CPU baseline: the oracle's serial FP64 Jacobi-CG on the first held-out system
    sysm = learned.FV7System(128, float(meta["sigma"]), float(meta["corr"]), 5000, dev)
    offh, colh = sysm.A.offsets.cpu().numpy(), sysm.A.columns.cpu().numpy()
    valh = sysm.A.values.cpu().numpy().astype(np.float64)
    inv = oracle.inverse_diag(oracle.bsr_get_diag(offh, colh, valh))
    cpu_ms, cpu_it = time_cpu_cg(offh, colh, valh, sysm.b.cpu().numpy(), oracle.PREC_DIAG, inv, iters=20)
    cpu_solve_ms = cpu_ms * med("jacobi64_it")
    N5 = nvox
    spmv64 = int(sysm.A.nnz) * 12 + 4 * (N5 + 1) + 16 * N5
    iter64 = cg_iter_bytes(spmv64, N5, 8)
    ms_it64 = med("jacobi64_ms") / max(med("jacobi64_it"), 1)
    log(f"[bench] U-Net forward 128^3 (BF16 MFMA, 7 launches): {unet_s * 1e6:.0f} us eager, "
        f"{unet_graph_s * 1e6:.0f} us replayed from a graph, {flops / 1e9:.1f} GFLOP, "
        f"{tf:.0f} TFLOP/s ({tf / 2500 * 100:.1f}% of 2.5 PFLOP/s); FP64 Jacobi-CG {ms_it64 * 1e3:.0f} us/iteration; "
        f"CPU 1 core {cpu_ms:.0f} ms/iteration")
    return {
        "workload": "-div(k grad u)=f, 7-point FV 128^3, log-k GRF sigma 1.5, corr 8, held-out seeds 5000+",
        "systems": len(recs),
        "tol": 1e-8,
        "median_speedup_vs_best_jacobi": round(float(np.median(speed)), 3),
        "min_speedup_vs_best_jacobi": round(float(np.min(speed)), 3),
        "median_fcg_ms": round(med("fcg_ms"), 2),
        "median_fcg_iterations": med("fcg_it"),
        "fallbacks": sum(1 for r in recs if r["fcg_fallback_it"] > 0),
        "median_jacobi64_ms": round(med("jacobi64_ms"), 2),
        "median_jacobi64_iterations": med("jacobi64_it"),
        "median_ir_ms": round(med("ir_ms"), 2),
        "max_true_rel_residual": max(max(r["fcg_true"], r["jacobi64_true"], r["ir_true"]) for r in recs),
        "timing": "per system: FCG and Jacobi-CG each run a few untimed iterations first (kernels warm); "
                  "the timed FCG solve includes its initial residual, every iteration and the FP64 check",
        "median_fcg_warmup_ms": round(med("fcg_warmup_ms"), 2),
        "checkpoint": os.path.relpath(ckpt, ROOT),
        "training": {"precision": str(meta.get("precision", "fp32")),
                     "seconds": round(float(meta.get("train_seconds", 0.0)), 1),
                     "pairs": {"plain_cg": int(meta.get("pcg_pairs", 0)), "fcg_iterates": int(meta.get("fcg_pairs", 0))},
                     "adam_steps": int(meta.get("steps", 0)) + int(meta.get("dagger_rounds", 0)) *
                     int(meta.get("dagger_steps", 0))},
        "unet": {"us_per_application": round(unet_s * 1e6, 1),
                 "us_per_application_graph_replay": round(unet_graph_s * 1e6, 1),
                 "gflop_per_application": round(flops / 1e9, 2),
                 "roofline": {"bound": "mfma", "achieved": round(tf, 1), "peak": 2500.0, "unit": "TFLOP/s",
                              "frac": round(tf / 2500.0, 4), "basis": "sum of 2*27*ci*co*voxels over the 7 layers"}},
        "jacobi64_roofline": hbm_roofline(iter64, ms_it64 * 1e-3, "FP64 PCG iteration, FP64 matrix values"),
        "cpu_baseline": {"value": round(cpu_solve_ms, 1), "unit": "ms/solve (plain Jacobi-CG to 1e-8)", "cores": 1,
                         "kind": "port",
                         "sample": f"{cpu_it} iterations of the oracle's serial FP64 Jacobi-CG on seed 5000 "
                                   f"({cpu_ms:.1f} ms each) x the median GPU iteration count"},
    }



def copy_calibration(dev) -> int:
    """A plain device-to-device copy of COPY_BYTES, repeated COPY_REPS times:
    the counters of its kernel give the scale of FETCH_SIZE / WRITE_SIZE."""
    import torch

    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(COPY_REPS):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / COPY_REPS
    print(json.dumps({"copy_bytes_read": COPY_BYTES, "copy_bytes_written": COPY_BYTES, "reps": COPY_REPS + 1,
                      "ms_per_copy": round(ms, 4), "GBps_read_plus_write": round(2 * COPY_BYTES / ms / 1e6, 1)}),
          flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=64, help="elements per side of the elasticity mesh")
    ap.add_argument("--no-extras", action="store_true", help="skip CG / assembly / CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--ir-n", type=int, default=256,
                    help="grid side of the FP64-refinement workload (variable-coefficient Poisson; 0: skip)")
    ap.add_argument("--fcg-seeds", type=int, default=10,
                    help="held-out 128^3 systems for the learned-FCG vs Jacobi-CG comparison (0: skip)")
    ap.add_argument("--eager", action="store_true",
                    help="time the bsr_mv calls with one event pair and stop there (for counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (calibrates FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (got --gpus {args.gpus}): the sparse path runs on a single MI355X",
            file=sys.stderr,
        )
        return 2

    import numpy as np
    import torch

    import warp_amd as wa
    from warp_amd import problems, sparse
    from warp_amd.optim import linear

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        return copy_calibration(dev)
    n = args.n
    log(f"[bench] generating 3D Q1 elasticity {n}^3 on {torch.cuda.get_device_name(dev)}")
    rows, cols, vals, b, nn = problems.elasticity3d_q1(n, device=dev)
    A = sparse.bsr_zeros(nn, nn, wa.mat33, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sparse.bsr_set_from_triplets(A, rows, cols, vals, prune_numerical_zeros=False)
    nnzb = A.nnz_sync()
    t_asm = time.perf_counter() - t0
    log(f"[bench] assembled {rows.numel()} triplets -> {nnzb} blocks in {t_asm * 1e3:.1f} ms")
    T_elem = rows.numel()
    del rows, cols, vals

    gen = torch.Generator(device=dev).manual_seed(11)
    x = torch.randn(nn * 3, generator=gen, device=dev, dtype=torch.float32)
    y = torch.empty(nn * 3, device=dev, dtype=torch.float32)

    # Timing: `steps` bsr_mv calls back to back on the current stream, HIP
    # events at the boundaries of 20 groups of consecutive calls; a short spin
    # kernel queued first keeps the host's enqueueing ahead of the GPU, so no
    # idle gap enters an interval.
    # Consecutive calls traverse the rows in opposite directions (the plan's
    # sweep bit; results bit-identical), as any loop of products with one
    # matrix does.  value = algorithmic bytes / median interval.
    for _ in range(max(args.warmup, 1)):
        sparse.bsr_mv(A, x, y)
    torch.cuda.synchronize()
    steps = max(args.steps, 1)
    if args.eager:
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            sparse.bsr_mv(A, x, y)
        e1.record()
        torch.cuda.synchronize()
        nbytes = spmv_bytes(nn, nnzb, 3, 3)
        ms_kernel = e0.elapsed_time(e1) / steps
        log(f"[bench] eager bsr_mv 3x3: {ms_kernel * 1e3:.1f} us/call ({nbytes / (ms_kernel * 1e-3) / 1e9:.0f} GB/s), "
            f"{steps} calls")
        return 0
    # the steps run in `blocks` groups of consecutive calls with an event at
    # every group boundary only (an event between two calls adds a gap of its
    # own); value uses the median over the groups of the time per call
    blocks = min(20, steps)
    bounds = [round(i * steps / blocks) for i in range(blocks + 1)]
    events = [torch.cuda.Event(enable_timing=True) for _ in range(blocks + 1)]
    stream = torch.cuda.current_stream(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    torch.cuda._sleep(2_000_000)
    events[0].record(stream)
    for grp in range(blocks):
        for _ in range(bounds[grp], bounds[grp + 1]):
            sparse.bsr_mv(A, x, y)
        events[grp + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_ms = [events[g].elapsed_time(events[g + 1]) / (bounds[g + 1] - bounds[g]) for g in range(blocks)]
    ms_kernel = float(np.median(per_ms))
    ms_mean = events[0].elapsed_time(events[blocks]) / steps
    ms_step = ms_kernel
    # every call's result equals a fresh call's, bit for bit
    y_ref = sparse.bsr_mv(A, x)
    torch.cuda.synchronize()
    if not torch.equal(y_ref.reshape(-1), y.reshape(-1)):
        raise RuntimeError("repeated bsr_mv differs from a fresh call")
    # the same product replayed from a HIP graph (one direction, captured once)
    cap_stream = torch.cuda.Stream(dev)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(cap_stream):
        with torch.cuda.graph(graph, stream=cap_stream):
            sparse.bsr_mv(A, x, y)
        for _ in range(5):
            graph.replay()
        ge = [torch.cuda.Event(enable_timing=True) for _ in range(41)]
        torch.cuda._sleep(2_000_000)
        ge[0].record(cap_stream)
        for i in range(40):
            graph.replay()
            ge[i + 1].record(cap_stream)
    torch.cuda.synchronize()
    ms_graph = float(np.median([ge[i].elapsed_time(ge[i + 1]) for i in range(40)]))
    if not torch.equal(y_ref.reshape(-1), y.reshape(-1)):
        raise RuntimeError("graph-replayed bsr_mv differs from the eager call")
    del graph
    nbytes = spmv_bytes(nn, nnzb, 3, 3)
    gbs = nbytes / (ms_kernel * 1e-3) / 1e9
    log(f"[bench] bsr_mv 3x3: median {ms_kernel * 1e3:.1f} us/call over {blocks} groups of {steps} calls "
        f"(mean {ms_mean * 1e3:.1f}, group min {min(per_ms) * 1e3:.1f}, max {max(per_ms) * 1e3:.1f}), {gbs:.0f} GB/s "
        f"({gbs / HBM_PEAK_GBS * 100:.1f}% of HBM peak); one-direction graph replay {ms_graph * 1e3:.1f} us")

    result = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "timed_region_wall_ms": round(wall * 1e3, 3),
        "kernel_us": {"median": round(ms_kernel * 1e3, 2), "mean": round(ms_mean * 1e3, 2),
                      "min": round(min(per_ms) * 1e3, 2), "max": round(max(per_ms) * 1e3, 2), "calls": steps,
                      "groups": blocks,
                      "graph_replay_one_direction_median": round(ms_graph * 1e3, 2)},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"bsr_mv y=Ax, 3D Q1 hex linear elasticity {n}^3 elements, 3x3 fp32 BSR",
            "block_rows": nn,
            "dofs": 3 * nn,
            "blocks": nnzb,
            "bytes_per_step": nbytes,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "traffic": load_traffic(f"bsr_mv_3x3_elasticity{n}"),
        },
        "cpu_baseline": None,
    }

    if not args.no_extras:
        import oracle

        # FP64 error of the headline product: against the FP64 CSR product on the host
        import scipy.sparse as sp

        offs = A.offsets.cpu().numpy()
        colsh = A.columns[:nnzb].cpu().numpy()
        valsh = np.ascontiguousarray(A.values[:nnzb].cpu().numpy())
        xh = x.cpu().numpy()
        yh = y.reshape(-1).cpu().numpy().astype(np.float64)
        Ab = sp.bsr_matrix((valsh.astype(np.float64), colsh, offs), shape=(3 * nn, 3 * nn))
        Aabs = sp.bsr_matrix((np.abs(valsh).astype(np.float64), colsh, offs), shape=(3 * nn, 3 * nn))
        y64 = Ab @ xh.astype(np.float64)
        scale = Aabs @ np.abs(xh.astype(np.float64))
        result["config"]["fp64_error"] = {
            "max_abs_over_AabsXabs": float(np.max(np.abs(yh - y64) / np.maximum(scale, 1e-300))),
            "rel_2norm": float(np.linalg.norm(yh - y64) / np.linalg.norm(y64)),
        }
        del Ab, Aabs, y64, scale

        # Jacobi and block-Jacobi CG on the same system, tol 1e-6 (reference default atol = tol)
        bt = b.reshape(-1).contiguous()
        b64h = bt.cpu().numpy().astype(np.float64)
        iter_bytes = cg_iter_bytes(nbytes, 3 * nn)
        for key, ptype, pkind in (("cg_jacobi", "diag", oracle.PREC_DIAG), ("cg_block_jacobi", "block_diag", oracle.PREC_BLOCK)):
            M = linear.preconditioner(A, ptype)
            xs = torch.zeros_like(bt)
            linear.cg(A, bt, xs, tol=1e-6, maxiter=20, M=M)  # warm-up of the solver kernels
            xs.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            it, err, atol = linear.cg(A, bt, xs, tol=1e-6, maxiter=20000, M=M)
            torch.cuda.synchronize()
            t_cg = time.perf_counter() - t0
            r = oracle.residual_f32_f64(offs, colsh, valsh, xs.cpu().numpy().astype(np.float64), b64h)
            true_rel = float(np.linalg.norm(r) / np.linalg.norm(b64h))
            cpu_ms, cpu_it = time_cpu_cg(offs, colsh, valsh, bt.cpu().numpy(), pkind, M._prec_data.cpu().numpy(),
                                         iters=20)
            ms_it = t_cg * 1e3 / max(it, 1)
            log(f"[bench] {ptype} CG: {it} iterations, {t_cg * 1e3:.1f} ms ({ms_it * 1e3:.1f} us/iteration, "
                f"{iter_bytes / (ms_it * 1e-3) / 1e9:.0f} GB/s), |r|={err:.3e} atol={atol:.3e}, FP64 true residual "
                f"{true_rel:.2e}; CPU 1 core {cpu_ms:.1f} ms/iteration")
            result[key] = {
                "iterations": it, "ms": round(t_cg * 1e3, 2), "ms_per_iter": round(ms_it, 4),
                "residual": err, "atol": atol, "fp64_true_rel_residual": true_rel,
                "roofline": hbm_roofline(iter_bytes, ms_it * 1e-3, "PCG iteration: bsr_mv bytes + 11 vector passes"),
                "cpu_baseline": {"value": round(cpu_ms, 3), "unit": "ms/iteration", "cores": 1, "kind": "port",
                                 "sample": f"{cpu_it} iterations of the oracle's serial CG on the same system"},
            }
            del M

        result["poisson2d_cg"] = extra_poisson2d(dev)
        result["assembly"] = extra_assembly(n, dev)
        result["assembly_synthetic"] = extra_assembly_synthetic(n, dev)
        result["row_compress"] = extra_row_compress(n, dev)
        result.update(extra_transpose_dirichlet(A, x, y, b, nn, nnzb, n, dev))
        if args.ir_n > 0:
            result["refinement"] = extra_refinement(args.ir_n, dev)
        result["learned_fcg"] = extra_learned(args.fcg_seeds, dev)

        # CPU baseline: the oracle's serial C bsr_mv on the same matrix
        sec, reps = time_cpu_spmv(offs, colsh, valsh, xh, min_seconds=args.cpu_seconds)
        cpu_gbs = nbytes / sec / 1e9
        log(f"[bench] CPU (1 core) bsr_mv: {sec * 1e3:.1f} ms/pass, {cpu_gbs:.2f} GB/s over {reps} passes")
        result["cpu_baseline"] = {
            "value": round(cpu_gbs, 3),
            "unit": "GB/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{reps} serial bsr_mv passes over the same {nnzb}-block matrix ({sec * 1e3:.1f} ms each)",
        }

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
