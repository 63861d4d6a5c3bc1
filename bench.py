"""Benchmark: contact-solves/s of the FP64 APGD Coulomb-friction solve on one MI355X.

Workload: 50,000 contacts between 20,000 rigid bodies (synthetic generator, seed 2),
Delassus operator W = H Minv H^T applied matrix-free (H^T, Minv and H on the
GPU; the residual uses the assembled-W row scaling max(1, tr(W_ii)/3)), Coulomb
law (De Saxce), cold start r0 = 0, solved to FB residual 1e-6. One step = one
complete solve with the problem and r resident in HBM.
value = n_contacts * steps / wall time.

Also reported: the roofline of the dominant kernel (the persistent solver
kernel) from its HIP-event time and its algorithmic bytes, the stand-alone
block-SpMV bandwidth, and the CPU oracle timed on this host's cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--tol 1e-6]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
COPY_BYTES = 1 << 30   # calibration copy: 1 GiB read + 1 GiB written, beyond the 256 MB MALL


def _load_package():
    if "sobogus_amd" in sys.modules:
        return sys.modules["sobogus_amd"]
    spec = importlib.util.spec_from_file_location("sobogus_amd", ROOT / "so-bogus_amd" / "__init__.py",
                                                  submodule_search_locations=[str(ROOT / "so-bogus_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["sobogus_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def wx_assembled_bytes(nnzb_lower: int, nc: int, s: int = 8) -> int:
    """Compulsory bytes of one symmetric W x (lower-triangle storage model)."""
    return nnzb_lower * (9 * s + 4) + 4 * (nc + 1) + 6 * nc * s


def wx_matfree_bytes(nc: int, nb: int, s: int = 8) -> int:
    """Compulsory bytes of one matrix-free H (Minv (H^T x)), 2 blocks per contact."""
    hx = 2 * nc * (18 * s + 4) + 4 * (nc + 1) + 6 * nb * s + 3 * nc * s
    minv = nb * 36 * s + 12 * nb * s
    return 2 * hx + minv


def measured_traffic(key: str, spmvs: int):
    """HBM traffic per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes.

    profiles/traffic.json holds, per kernel and workload, the calibrated bytes of
    one profiled launch and its operator-application count; the figure is scaled
    to this run's count. None when no pass was recorded for this configuration.
    """
    try:
        rec = json.loads((ROOT / "profiles" / "traffic.json").read_text())[key]
    except (OSError, KeyError, ValueError):
        return None, None
    return rec["traffic_bytes"] * spmvs / rec["spmvs"], rec["summary"]


def cpu_baseline(prob, tol: float, threads: int) -> dict:
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    O.set_threads(threads)
    op = O.OracleProblem(prob)
    t0 = time.perf_counter()
    res, r, info = op.pg_solve(np.zeros(3 * prob.nc), variant=O.PG_APGD, tol=tol, max_iters=100000, coulomb=True)
    dt = time.perf_counter() - t0
    return {"value": prob.nc / dt, "unit": "contact-solves/s", "cores": threads, "kind": "port",
            "sample": f"1 full APGD solve of the {prob.nc}-contact workload to FB {tol:g} (oracle, FP64, "
                      f"{info['iterations']} iterations, {dt:.2f} s, residual {res:.2e})"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--workload", default="contacts-50k")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--operator", choices=["matrix-free", "assembled"], default="matrix-free")
    ap.add_argument("--profile-run", action="store_true",
                    help="short run for rocprofv3: 1 warm-up + 2 solves + 2 calibration copies, no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is described in DESIGN.md, not implemented)", file=sys.stderr)
        return 2

    _load_package()
    from sobogus_amd import generator as G
    from sobogus_amd import runtime as R

    prob = G.workload(args.workload)
    ctx = R.Context(0)
    print(f"[bench] {ctx.arch} {ctx.n_cu} CUs; workload {args.workload}: {prob.nc} contacts, {prob.nb} bodies",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    dp = R.DualFrictionProblem(ctx, prob, assemble=True)
    nc, nb, hblocks, wblocks = dp.sizes()
    print(f"[bench] setup (W assembly + upload) {time.perf_counter() - t0:.2f} s, W lower blocks {wblocks}",
          file=sys.stderr, flush=True)
    op = R.OP_ASSEMBLED if args.operator == "assembled" else R.OP_MATRIX_FREE
    opts = R.pg_options(variant=R.PG_APGD, law=R.LAW_COULOMB, op=op, scaling=R.SCALING_DIAGONAL, tol=args.tol,
                        max_iters=100000)
    r = R.DeviceArray(ctx, 3 * nc)
    zero = np.zeros(3 * nc)

    def step():
        r.upload(zero)
        return dp.solve_pg(r, opts, on_device=True)

    if args.profile_run:
        for _ in range(3):
            res = step()
            print(f"[bench] profile solve: {res.iterations} iterations, {res.spmvs} W x, "
                  f"residual {res.residual:.3e}, {res.device_ms:.3f} ms", file=sys.stderr, flush=True)
        ctx.synchronize()
        ctx.bench_copy(COPY_BYTES, reps=2)
        del r, dp
        ctx.close()
        return 0
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    kernel_ms, results = [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        res = step()
        results.append(res)
        kernel_ms.append(res.device_ms)
    ctx.synchronize()
    wall = time.perf_counter() - t_start
    last = results[-1]
    # FB residual of the returned r with the assembled W and its row scaling
    final_res = dp.eval_with(r.download(), law=R.LAW_COULOMB, op=R.OP_ASSEMBLED)
    ms_per_step = 1000.0 * wall / args.steps
    value = nc * args.steps / wall
    print(f"[bench] {args.steps} solves: {ms_per_step:.2f} ms/solve, {last.iterations} iterations, "
          f"{last.spmvs} W x, residual {last.residual:.3e} (recomputed {final_res:.3e})", file=sys.stderr, flush=True)

    # roofline of the dominant kernel: the persistent solver (one launch per solve)
    k_ms = float(np.mean(kernel_ms))
    phases = ctx.last_phase_times()
    bytes_wx = wx_assembled_bytes(wblocks, nc)
    bytes_op = bytes_wx if args.operator == "assembled" else wx_matfree_bytes(nc, nb)
    algo_bytes = last.spmvs * bytes_op
    achieved = algo_bytes / (k_ms * 1e-3) / 1e9
    kname = f"pg_kernel<{'false' if args.operator == 'assembled' else 'true'}>"
    traffic, traffic_src = measured_traffic(f"{kname}|{args.workload}|{args.operator}|fp64", last.spmvs)
    # stand-alone block SpMV kernels (W x assembled, matrix-free H (Minv H^T x))
    ms_w = dp.bench_spmv(0, reps=50)
    ms_mf = dp.bench_spmv(3, reps=50)
    ms_copy = ctx.bench_copy(COPY_BYTES, reps=10)
    copy_gbs = 2 * COPY_BYTES / (ms_copy * 1e-3) / 1e9
    spmv = {
        "assembled_wx": {"ms": ms_w, "bytes": bytes_wx, "GBs": bytes_wx / (ms_w * 1e-3) / 1e9},
        "matrix_free_wx": {"ms": ms_mf, "bytes": wx_matfree_bytes(nc, nb),
                           "GBs": wx_matfree_bytes(nc, nb) / (ms_mf * 1e-3) / 1e9},
    }
    out = {
        "metric": "contact-solves/s to FB residual 1e-6, mixed vs FP64; block-SpMV GB/s vs peak",
        "value": value,
        "unit": "contact-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "n_contacts": nc, "n_bodies": nb, "solver": "APGD",
                   "law": "coulomb", "operator": args.operator, "tol": args.tol, "precision": "fp64"},
        "solve": {"iterations": last.iterations, "spmvs": last.spmvs, "ls_trials": last.ls_trials,
                  "nesterov_steps": last.nesterov_steps, "residual": last.residual,
                  "residual_recomputed": final_res, "kernel_ms": k_ms, "phase_us": phases},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": f"{kname} (persistent APGD)",
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "model": "operator applications x compulsory bytes of one W x"},
        "spmv": spmv,
        "stream_copy_GBs": copy_gbs,
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        print(f"[bench] CPU baseline on {threads} threads ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(prob, args.tol, threads)
    print(json.dumps(out), flush=True)
    # release device objects before interpreter shutdown
    del r, dp
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
