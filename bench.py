"""Benchmark of the sparse-solver hot path on one MI355X.

Headline (`value`): bsr_mv HBM throughput on the 3D Q1 hex elasticity
matrix, 64^3 elements (65^3 nodes x 3 = 823,875 dofs, 3x3 fp32 BSR,
7,189,057 blocks).  One step = one y = A x over the whole matrix, inputs
resident in HBM, timed as HIP-graph replays of 20 back-to-back calls (HIP
events on the replay stream).  GB/s counts algorithmic bytes:
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
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=64, help="elements per side of the elasticity mesh")
    ap.add_argument("--no-extras", action="store_true", help="skip CG / assembly / CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--ir-n", type=int, default=256,
                    help="grid side of the FP64-refinement workload (variable-coefficient Poisson; 0: skip)")
    ap.add_argument("--fcg-seeds", type=int, default=10,
                    help="held-out 128^3 systems for the learned-FCG vs Jacobi-CG comparison (0: skip)")
    ap.add_argument("--eager", action="store_true",
                    help="launch bsr_mv one call at a time instead of replaying a HIP graph (for counter runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (calibrates FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (got --gpus {args.gpus}); the sharded design is described in DESIGN.md",
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

    # Launch-overhead-free timing: PER_GRAPH back-to-back bsr_mv calls captured
    # in one HIP graph (as the reference's asv benchmark captures bsr_mv),
    # replayed steps / PER_GRAPH times between events on the replay stream.
    PER_GRAPH = 20
    for _ in range(max(args.warmup, 1)):
        sparse.bsr_mv(A, x, y)
    torch.cuda.synchronize()
    if args.eager:
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for _ in range(args.steps):
            sparse.bsr_mv(A, x, y)
        e1.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        nbytes = spmv_bytes(nn, nnzb, 3, 3)
        ms_kernel = e0.elapsed_time(e1) / args.steps
        log(f"[bench] eager bsr_mv 3x3: {ms_kernel * 1e3:.1f} us/call ({nbytes / (ms_kernel * 1e-3) / 1e9:.0f} GB/s), "
            f"{args.steps} calls")
        return 0
    cap_stream = torch.cuda.Stream(dev)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(cap_stream):
        with torch.cuda.graph(graph, stream=cap_stream):
            for _ in range(PER_GRAPH):
                sparse.bsr_mv(A, x, y)
    replays = max(1, args.steps // PER_GRAPH)
    steps = replays * PER_GRAPH
    for _ in range(2):
        graph.replay()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    with torch.cuda.stream(cap_stream):
        e0.record(cap_stream)
        for _ in range(replays):
            graph.replay()
        e1.record(cap_stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_kernel = e0.elapsed_time(e1) / steps
    ms_step = wall * 1e3 / steps
    # the graph's result equals an eager call
    y_ref = sparse.bsr_mv(A, x)
    torch.cuda.synchronize()
    if not torch.equal(y_ref.reshape(-1), y.reshape(-1)):
        raise RuntimeError("graph-replayed bsr_mv differs from the eager call")
    del graph
    nbytes = spmv_bytes(nn, nnzb, 3, 3)
    gbs = nbytes / (ms_kernel * 1e-3) / 1e9
    log(f"[bench] bsr_mv 3x3: {ms_kernel * 1e3:.1f} us/launch, {gbs:.0f} GB/s ({gbs / HBM_PEAK_GBS * 100:.1f}% of HBM peak)")

    result = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
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
        # Jacobi-CG on the same system
        M = linear.preconditioner(A, "diag")
        bt = b.reshape(-1).contiguous()
        xs = torch.zeros_like(bt)
        linear.cg(A, bt, xs, tol=1e-6, maxiter=20, M=M)  # warm-up of the solver kernels
        xs.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it, err, atol = linear.cg(A, bt, xs, tol=1e-6, maxiter=20000, M=M)
        torch.cuda.synchronize()
        t_cg = time.perf_counter() - t0
        log(f"[bench] Jacobi-CG: {it} iterations, {t_cg * 1e3:.1f} ms, |r|={err:.3e} atol={atol:.3e}")
        result["cg_jacobi"] = {"iterations": it, "ms": round(t_cg * 1e3, 2), "ms_per_iter": round(t_cg * 1e3 / max(it, 1), 4),
                               "residual": err, "atol": atol}
        Mb = linear.preconditioner(A, "block_diag")
        xs.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        itb, errb, atolb = linear.cg(A, bt, xs, tol=1e-6, maxiter=20000, M=Mb)
        torch.cuda.synchronize()
        t_cgb = time.perf_counter() - t0
        log(f"[bench] block-Jacobi-CG: {itb} iterations, {t_cgb * 1e3:.1f} ms")
        result["cg_block_jacobi"] = {"iterations": itb, "ms": round(t_cgb * 1e3, 2)}

        # 2D Q1 Poisson, 128 x 128 cells (16,129 interior dofs): Jacobi-CG, tol 1e-6, maxiter 2000
        rp, cp, vp, bp, ndp = problems.poisson2d_q1(128, 0)
        Ap = sparse.bsr_from_triplets(ndp, ndp, torch.from_numpy(rp).to(dev), torch.from_numpy(cp).to(dev),
                                      torch.from_numpy(vp).to(dev))
        nnzp = Ap.nnz_sync()
        bpt = torch.from_numpy(bp).to(dev)
        xp = torch.zeros_like(bpt)
        yp = torch.empty_like(bpt)
        for _ in range(20):
            sparse.bsr_mv(Ap, xp, yp)
        e0.record()
        for _ in range(500):
            sparse.bsr_mv(Ap, bpt, yp)
        e1.record()
        torch.cuda.synchronize()
        us_mv = e0.elapsed_time(e1) / 500 * 1e3
        Mp = linear.preconditioner(Ap, "diag")
        linear.cg(Ap, bpt, xp, tol=1e-6, maxiter=20, M=Mp)
        xp.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        itp, errp, atolp = linear.cg(Ap, bpt, xp, tol=1e-6, maxiter=2000, M=Mp)
        torch.cuda.synchronize()
        t_p = time.perf_counter() - t0
        log(f"[bench] Poisson 2D Q1 (n={ndp}, nnz={nnzp}): bsr_mv {us_mv:.1f} us, Jacobi-CG {itp} iterations "
            f"{t_p * 1e3:.1f} ms")
        result["poisson2d_cg"] = {"dofs": ndp, "nnz": nnzp, "bsr_mv_us": round(us_mv, 2), "iterations": itp,
                                  "ms": round(t_p * 1e3, 2), "us_per_iter": round(t_p * 1e6 / max(itp, 1), 2),
                                  "residual": errp, "atol": atolp}
        del Ap, Mp

        # assembly of the shuffled element-loop stream
        r3, c3, v3, nn3 = problems.elasticity_triplets_shuffled(n, device=dev)
        A3 = sparse.bsr_zeros(nn3, nn3, wa.mat33, device=dev)
        sparse.bsr_set_from_triplets(A3, r3, c3, v3)
        A3.nnz_sync()
        reps = 5
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            sparse.bsr_set_from_triplets(A3, r3, c3, v3)
            u3 = A3.nnz_sync()
        t_a = (time.perf_counter() - t0) / reps
        T3 = r3.numel()
        asm_bytes = T3 * 44 + u3 * 40 + 4 * (nn3 + 1)
        log(f"[bench] assembly: {T3} triplets -> {u3} blocks in {t_a * 1e3:.2f} ms ({T3 / t_a / 1e9:.2f} Gtriplets/s)")
        result["assembly"] = {"triplets": T3, "blocks": u3, "ms": round(t_a * 1e3, 3),
                              "triplets_per_s": round(T3 / t_a), "alg_GBps": round(asm_bytes / t_a / 1e9, 1)}
        del r3, c3, v3, A3

        # synthetic variant: 20,000,000 triplets over 2,500,000 random blocks of the same
        # 27-point pattern (mean multiplicity 8)
        r6, c6, v6, nn6 = problems.synthetic_triplets(n, device=dev)
        A6 = sparse.bsr_zeros(nn6, nn6, wa.mat33, device=dev)
        sparse.bsr_set_from_triplets(A6, r6, c6, v6)
        A6.nnz_sync()
        ms6 = []
        for _ in range(5):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sparse.bsr_set_from_triplets(A6, r6, c6, v6)
            u6 = A6.nnz_sync()
            ms6.append(time.perf_counter() - t0)
        t6 = float(np.median(ms6))
        T6 = r6.numel()
        b6 = T6 * 44 + u6 * 40 + 4 * (nn6 + 1)
        log(f"[bench] assembly (synthetic): {T6} triplets -> {u6} blocks in {t6 * 1e3:.2f} ms "
            f"({T6 / t6 / 1e9:.2f} Gtriplets/s)")
        result["assembly_synthetic"] = {"triplets": T6, "blocks": u6, "ms": round(t6 * 1e3, 3),
                                        "triplets_per_s": round(T6 / t6), "alg_GBps": round(b6 / t6 / 1e9, 1)}
        del r6, c6, v6, A6

        # fem "row_compress" construction: every block row's candidate blocks written
        # in element order into a padded matrix, then sorted and coalesced per row
        re_, ce_, ve_, _, nn5 = problems.elasticity3d_q1(n, device=dev)
        order = torch.argsort(re_, stable=True)
        counts = torch.bincount(re_, minlength=nn5).to(torch.int32)
        cols_p = ce_[order].contiguous()
        vals_p = ve_[order].contiguous()
        del re_, ce_, ve_, order
        A5 = sparse.bsr_zeros(nn5, nn5, wa.mat33, device=dev, row_capacity=counts)
        reps = 5
        ms_c = []
        for rep in range(reps + 1):
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
        log(f"[bench] row-compress construction: {T5} candidates -> {u5} blocks in {t_c:.2f} ms "
            f"({T5 / t_c / 1e6:.2f} G candidates/s)")
        result["row_compress"] = {"candidates": T5, "blocks": u5, "ms": round(t_c, 3),
                                  "candidates_per_s": round(T5 / t_c * 1e3),
                                  "alg_GBps": round((T5 * 40 + u5 * 40 + 4 * (nn5 + 1)) / t_c / 1e6, 1)}
        del A5, cols_p, vals_p, counts

        # transpose of the elasticity matrix and its Dirichlet projection (the fem steps
        # between assembly and solve): x = 0 face, identity projector blocks, fixed value 0
        from warp_amd import fem

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
        result["transpose"] = {"blocks": nt, "ms": round(t_t, 3), "alg_GBps": round(t_bytes / t_t / 1e6, 1)}
        # y = A^T x on the same matrix: its pattern is symmetric, so the deterministic
        # gather (after the mirror check, which reads one count back)
        yT = torch.empty_like(y)
        sparse.bsr_mv(A, x, yT, transpose=True)
        ms_mt = []
        for rep in range(6):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sparse.bsr_mv(A, x, yT, transpose=True)
            torch.cuda.synchronize()
            ms_mt.append((time.perf_counter() - t0) * 1e3)
        sym = sparse._mirror_symmetric(A)
        log(f"[bench] bsr_mv transpose=True: {float(np.median(ms_mt)):.3f} ms per call "
            f"({'gather, deterministic' if sym else 'atomic scatter'}, mirror check included)")
        result["transpose"]["bsr_mv_transpose_ms"] = round(float(np.median(ms_mt)), 3)
        result["transpose"]["bsr_mv_transpose_path"] = "gather" if sym else "atomic"
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
        result["dirichlet_projection"] = {"face_nodes": int(face.numel()), "blocks": nd, "ms": round(t_d, 3)}

        # FP32 Jacobi-CG + FP64 iterative refinement to 1e-10 on 256^3 variable-coefficient Poisson,
        # solution checked on the host by the oracle's FP64 residual
        if args.ir_n > 0:
            import oracle
            from warp_amd.optim import refine

            n4 = args.ir_n
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
            r_host = oracle.residual_f32_f64(off4.cpu().numpy(), col4.cpu().numpy(), val4.cpu().numpy(),
                                             x64.cpu().numpy(), b64.cpu().numpy())
            true_rel = float(np.linalg.norm(r_host) / np.linalg.norm(b64.cpu().numpy()))
            log(f"[bench] refinement {n4}^3: {ir.outer_iterations} outer / {ir.inner_iterations} inner iterations, "
                f"{t_ir * 1e3:.1f} ms, FP64 oracle residual {true_rel:.2e}")
            result["refinement"] = {
                "workload": f"-div(k grad u)=f, 7-point FV {n4}^3, log-k GRF sigma 2.3, corr {16 * n4 // 256}, "
                            "FP32 Jacobi-CG inner (tol 3e-5) + FP64 refinement",
                "dofs": N4, "nnz": int(col4.numel()), "tol": 1e-10,
                "outer_iterations": ir.outer_iterations, "inner_iterations": ir.inner_iterations,
                "ms": round(t_ir * 1e3, 2), "ms_per_inner_iteration": round(t_ir * 1e3 / max(ir.inner_iterations, 1), 4),
                "fp64_oracle_rel_residual": true_rel,
            }
            del A4, off4, col4, val4, b4, b64, x64, r_host

        # learned flexible CG vs Jacobi-CG on held-out 128^3 variable-coefficient Poisson systems
        ckpt = os.path.join(ROOT, "warp_amd", "data", "fv7_unet_128.npz")
        if args.fcg_seeds > 0 and os.path.exists(ckpt):
            from warp_amd.optim import learned

            net, meta = learned.load_checkpoint(ckpt, 128, dev)
            recs = learned.evaluate(net, meta, range(5000, 5000 + args.fcg_seeds), dev, tol=1e-8, log=log)
            med = lambda key: float(np.median([r[key] for r in recs]))
            speed = [min(r["jacobi64_ms"], r["ir_ms"]) / r["fcg_ms"] for r in recs]
            result["learned_fcg"] = {
                "workload": "-div(k grad u)=f, 7-point FV 128^3, log-k GRF sigma 1.5, corr 8, held-out seeds 5000+",
                "systems": len(recs),
                "tol": 1e-8,
                "median_speedup_vs_best_jacobi": round(float(np.median(speed)), 3),
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
            }
        else:
            result["learned_fcg"] = {"skipped": "no checkpoint" if args.fcg_seeds > 0 else "--fcg-seeds 0"}

        # CPU baseline: the oracle's serial C bsr_mv on the same matrix
        offs = A.offsets.cpu().numpy()
        colsh = A.columns[:nnzb].cpu().numpy()
        valsh = np.ascontiguousarray(A.values[:nnzb].cpu().numpy())
        xh = x.cpu().numpy()
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
