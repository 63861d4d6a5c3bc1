"""Benchmark: contact-solves/s to FB residual 1e-6 on one MI355X, mixed vs FP64.

Workload (default, "batch-512"): 512 independent friction problems from the
synthetic generator, n_c = floor(exp(U[ln 64, ln 20000])) contacts and
n_b = n_c / 2 bodies each (seeds 4000 + k; 1,839,956 contacts in all), Coulomb
law (De Saxce), matrix-free W = H Minv H^T, residual scaled by
max(1, tr(W_ii)/3), cold start r0 = 0, APGD to FB residual 1e-6 with at most 5000
iterations per problem. One step = the whole batch solved in one graph launch
(the solver kernel of every stage captured in a hipGraph), problem data and r
resident in HBM. The launch shape is the autotuner's pick on another batch of the
generator (seeds 5000 + k): nothing about the timed problems (their iteration
counts in particular) is known before the timed steps. value = sum of n_c over
problems whose FP64-evaluated residual is below 1e-6, times steps, over the wall
time of the timed steps.

Mixed precision: FP32 matrices, iterates, products and projection; FP64
residual, dot products and line-search decisions; then an FP64 stage that starts
from the FP32 result, verifies the tolerance and refines where it is not met. The
FP64 batch, the same batch tuned on itself (the autotuner then also records each
problem's iteration count), the plain FP64 path (512 single-problem solves one
after the other), the 50,000-contact single problem (FP64 and mixed; and to FB 1e-8
with APGD and the colored Gauss-Seidel solvers, assembled W and matrix-free,
against the oracle on the host), the derivative workload at 100,000 contacts, the
per-step set-up cost (host creation against the device value refresh), the batch
with r in host memory, and the stand-alone block SpMV bandwidth at 50k and at 1M
contacts (also with the caches flushed before every launch) are reported beside
it, with the CPU oracle timed on this host's cores on a sample of the batch.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--tol 1e-6]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
COPY_BYTES = 1 << 30   # calibration copy: 1 GiB read + 1 GiB written, beyond the 256 MB MALL
METRIC = "contact-solves/s to FB residual 1e-6, mixed vs FP64; block-SpMV GB/s vs peak"


def _load_package():
    if "sobogus_amd" in sys.modules:
        return sys.modules["sobogus_amd"]
    spec = importlib.util.spec_from_file_location("sobogus_amd", ROOT / "so-bogus_amd" / "__init__.py",
                                                  submodule_search_locations=[str(ROOT / "so-bogus_amd")])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["sobogus_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def wx_assembled_bytes(nnzb_lower: int, nc: int, s: int = 8) -> int:
    """Compulsory bytes of one symmetric W x (lower-triangle storage model)."""
    return nnzb_lower * (9 * s + 4) + 4 * (nc + 1) + 6 * nc * s


def wx_matfree_bytes(nc: int, nb: int, s: int = 8) -> int:
    """Compulsory bytes of one matrix-free H (Minv (H^T x)), 2 blocks per contact,
    int32 indices, s bytes per value (8: FP64, 4: FP32)."""
    hx = 2 * nc * (18 * s + 4) + 4 * (nc + 1) + 6 * nb * s + 3 * nc * s
    minv = nb * 36 * s + 12 * nb * s
    return 2 * hx + minv


def wx_factored_bytes(nc: int, nb: int, s: int = 8) -> int:
    """Compulsory bytes of one factored matrix-free product F (F^T x), F = H L with
    Minv = L L^T: the two tile passes of H x and H^T r, no Minv blocks."""
    hx = 2 * nc * (18 * s + 4) + 4 * (nc + 1) + 6 * nb * s + 3 * nc * s
    return 2 * hx


def measured_traffic(key: str, units: float):
    """HBM traffic per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes.

    profiles/traffic.json holds, per kernel and workload, the calibrated bytes of
    one profiled launch and its work units (operator applications); the figure is
    scaled to this run's units. None when no pass was recorded for the configuration.
    """
    try:
        rec = json.loads((ROOT / "profiles" / "traffic.json").read_text())[key]
    except (OSError, KeyError, ValueError):
        return None, None
    return rec["traffic_bytes"] * units / rec["units"], rec["summary"]


def host_cpu_info() -> dict:
    """CPU model, physical cores (distinct core/socket pairs of lscpu) and the cores this
    process may run on."""
    info = {"model": "", "physical_cores": 0, "logical_cpus": os.cpu_count() or 0,
            "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                info["model"] = ln.split(":", 1)[1].strip()
        pairs = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        info["physical_cores"] = len({ln for ln in pairs.splitlines() if ln and not ln.startswith("#")})
    except (OSError, subprocess.SubprocessError):
        pass
    if not info["model"]:
        try:
            with open("/proc/cpuinfo") as f:
                info["model"] = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
        except OSError:
            pass
    return info


def cpu_share() -> int:
    """Host threads for the CPU legs: one per physical core available to this process,
    at most 16 (this machine's CPU share per GPU)."""
    info = host_cpu_info()
    n = min(info["affinity_cpus"], info["physical_cores"] or info["affinity_cpus"])
    return max(1, min(16, n))


def cpu_baseline_child(spec: dict) -> dict:
    """Runs in a child process whose OpenMP runtime was configured by its environment
    (OMP_NUM_THREADS, OMP_PROC_BIND=close, OMP_PLACES=cores): the oracle's FP64 APGD
    (matrix-free, Coulomb, diagonal scaling, cold start) on every `stride`-th problem of
    the batch, one problem after the other; one warm-up pass, then `reps` timed passes
    (median reported); then one pass on one thread over every `stride1`-th problem."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    _load_package()
    from sobogus_amd import generator as G
    sizes = G.batch_sizes(512, seed0=spec["seed0"])
    sample = [G.generate(nc, nb, seed) for nc, nb, seed in sizes[::spec["stride"]]]
    tol, max_iters = spec["tol"], spec["max_iters"]

    def run(probs):
        t0 = time.perf_counter()
        solved = nconv = 0
        for p in probs:
            res, _, _ = O.OracleProblem(p, assemble=False).pg_solve(np.zeros(3 * p.nc), variant=O.PG_APGD, tol=tol,
                                                                   max_iters=max_iters, coulomb=True,
                                                                   assembled=False, diag_scaling=True)
            if res < tol:
                solved += p.nc
                nconv += 1
        return time.perf_counter() - t0, solved, nconv

    threads = O.threads()
    run(sample)  # warm-up
    passes = [run(sample) for _ in range(spec["reps"])]
    walls = sorted(p[0] for p in passes)
    med = walls[len(walls) // 2]
    solved, nconv = passes[0][1], passes[0][2]
    O.set_threads(1)
    small = [G.generate(nc, nb, seed) for nc, nb, seed in sizes[::spec["stride1"]]]
    dt1, solved1, _ = run(small)
    return {"value": solved / med, "unit": "contact-solves/s", "cores": threads, "kind": "port",
            "sample": f"problems 0, {spec['stride']}, {2 * spec['stride']}, ... of the batch ({len(sample)} problems, "
                      f"{sum(p.nc for p in sample)} contacts, {nconv} converged), oracle FP64 APGD matrix-free to FB "
                      f"{tol:g}; median of {spec['reps']} passes after a warm-up pass: {med:.2f} s "
                      f"(passes {', '.join(f'{w:.2f}' for w in walls)} s)",
            "omp": {k: os.environ.get(k, "") for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
            "one_thread": {"value": solved1 / dt1, "unit": "contact-solves/s",
                           "sample": f"problems 0, {spec['stride1']}, {2 * spec['stride1']}, ... ({len(small)} problems, "
                                     f"{sum(p.nc for p in small)} contacts), one pass, {dt1:.1f} s"}}


def cpu_baseline_batch(tol: float, max_iters: int, threads: int, seed0: int = 4000) -> dict:
    """The CPU baseline in a child process (the OpenMP binding must be set before the
    runtime starts): `threads` threads bound one per physical core."""
    info = host_cpu_info()
    env = dict(os.environ)
    env.update(OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    spec = {"seed0": seed0, "stride": 12, "stride1": 32, "reps": 5, "tol": tol, "max_iters": max_iters}
    out = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-baseline-child", json.dumps(spec)],
                         env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError(f"CPU baseline child failed: {out.stderr[-2000:]}")
    res = json.loads(out.stdout.strip().splitlines()[-1])
    res["cpu_model"] = info["model"]
    res["physical_cores"] = info["physical_cores"]
    res["affinity_cpus"] = info["affinity_cpus"]
    res["note"] = (f"{threads} threads: this GPU's share of the host cores (at most 16 per GPU on this machine; "
                   f"{info['physical_cores']} physical cores in the node, {info['affinity_cpus']} CPUs in this "
                   f"process's affinity), one per physical core")
    return res


def tight_50k(ctx, R, G, cpu_threads: int) -> dict:
    """The 50,000-contact problem to FB 1e-8: APGD (FP64 and mixed, matrix-free) and
    Gauss-Seidel (SOC law; assembled W and matrix-free) on the GPU; with cpu_threads > 0
    the oracle's APGD and GS (same coloring, each color's contacts swept in parallel as the
    reference's OpenMP loop) on the host."""
    p = G.workload("contacts-50k")
    dp = R.DualFrictionProblem(ctx, p, assemble=True)
    r = R.DeviceArray(ctx, 3 * p.nc)
    out = {}
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-8,
              max_iters=20000)
    for name, prec in (("apgd_fp64", R.PRECISION_FP64), ("apgd_mixed", R.PRECISION_MIXED)):
        o = R.pg_options(precision=prec, **kw)
        ts = []
        for _ in range(3):
            r.zero()
            s = dp.solve_pg(r, o, on_device=True)
            ts.append(s.device_ms)
        out[name] = {"ms": float(np.median(ts[1:])), "iterations": s.iterations, "residual": s.residual,
                     "converged": bool(s.converged)}
    # colored Gauss-Seidel, SOC law: with the Coulomb law GS stalls near FB 1e-3 on this
    # problem (the oracle's too), so the tight tolerance is measured on the SOC problem
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    for name, fn, prec in (("gs_assembled_w", dp.solve_gs, R.PRECISION_FP64),
                           ("gs_assembled_w_mixed", dp.solve_gs, R.PRECISION_MIXED),
                           ("gs_matrix_free", dp.solve_pgs, R.PRECISION_FP64)):
        ts = []
        for _ in range(3):
            r.zero()
            s = fn(r, law=R.LAW_SOC, tol=1e-8, max_iters=3000, colors=colors, on_device=True, precision=prec)
            ts.append(s.device_ms)
        out[name] = {"ms": float(np.median(ts[1:])), "sweeps": s.iterations, "residual": s.residual,
                     "converged": bool(s.converged), "colors": len(colors) - 1,
                     "us_per_sweep": 1e3 * float(np.median(ts[1:])) / max(s.iterations, 1)}
        if prec == R.PRECISION_MIXED:
            out[name].update(fp32_sweeps=s.iterations - s.refine_iterations, fp32_residual=s.mixed_residual)
    log("contacts-50k to 1e-8: " + ", ".join(f"{k} {v['ms']:.1f} ms" for k, v in out.items()))
    if cpu_threads > 0:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle as O
        O.set_threads(cpu_threads)
        op = O.OracleProblem(p, assemble=True)
        z = np.zeros(3 * p.nc)
        t0 = time.perf_counter()
        res, _, d = op.pg_solve(z, variant=O.PG_APGD, tol=1e-8, max_iters=20000, coulomb=True, assembled=False,
                                diag_scaling=True)
        out["cpu_apgd_fp64"] = {"ms": 1e3 * (time.perf_counter() - t0), "iterations": d["iterations"],
                                "residual": res, "threads": cpu_threads}
        op.permute(perm.astype(np.int32))
        t0 = time.perf_counter()
        res, _, d = op.gs_solve(z, tol=1e-8, max_iters=3000, coulomb=False, colors=colors.astype(np.int32),
                                ndx_ref_per_color=True)
        out["cpu_gs_assembled_w"] = {"ms": 1e3 * (time.perf_counter() - t0), "sweeps": d["iterations"],
                                     "residual": res, "threads": cpu_threads}
        log(f"oracle on the host: APGD {out['cpu_apgd_fp64']['ms']:.0f} ms ({cpu_threads} threads), "
            f"GS {out['cpu_gs_assembled_w']['ms']:.0f} ms (each color's contacts on {cpu_threads} threads)")
    del r, dp
    return out


def adjoint_100k(ctx, R, G, n_dirs: int = 64) -> dict:
    """The 100,000-contact derivative workload (Coulomb law, L = g . r, g ~ N(0,1) seed 55):
    mixed-precision solve (FP32 bulk, FP64 residual and verification), then
      * the gradient of a 200-step checkpointed unroll from the solution in b, mu and
        Minv, its cost against the forward steps, and central differences (h = 1e-6,
        FP64) on n_dirs random directions of the map replayed on the recorded piece
        (projection branches and De Saxce directions of the unperturbed run);
      * the implicit gradient at the solution (GMRES on the adjoint system);
      * at an FP64 solution to FB 1e-10, the implicit gradient against unrolls of
        40,000 and 80,000 steps from it (the unroll's limit)."""
    prob = G.workload("adjoint-100k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL,
              max_iters=100000)
    sol = dp.solve_pg(np.zeros(3 * prob.nc), tol=1e-6, precision=R.PRECISION_MIXED, **kw)
    g = np.random.default_rng(55).standard_normal(3 * prob.nc)
    K = 200
    law = R.LAW_COULOMB
    ad = dp.adjoint(sol.r, g, iterations=K, law=law, record_branches=True)
    ad = dp.adjoint(sol.r, g, iterations=K, law=law, record_branches=True)
    fwd = dp.adjoint(sol.r, None, iterations=K, rho=ad["rho"], law=law)
    fwd = dp.adjoint(sol.r, None, iterations=K, rho=ad["rho"], law=law)
    rng = np.random.default_rng(56)
    h, errs = 1e-6, []
    for _ in range(n_dirs):
        db = rng.standard_normal(3 * prob.nc)
        dmu = 0.1 * rng.standard_normal(prob.nc)
        A = rng.standard_normal((prob.nb, 6, 6)) * 0.05
        dM = (A + A.transpose(0, 2, 1)).reshape(prob.nb, 36)
        an = ad["grad_b"] @ db + ad["grad_mu"] @ dmu + np.sum(ad["grad_minv"] * dM)
        lv = []
        for t in (h, -h):
            dp.update(b=prob.b + t * db, mu=prob.mu + t * dmu, minv=prob.minv.reshape(prob.nb, 36) + t * dM)
            lv.append(float(g @ dp.adjoint(sol.r, None, iterations=K, rho=ad["rho"], law=law,
                                           branches=ad["branches"])["r"]))
        errs.append(abs((lv[0] - lv[1]) / (2 * h) - an) / max(abs(an), 1e-12))
    dp.update(b=prob.b, mu=prob.mu, minv=prob.minv.reshape(prob.nb, 36))
    imp = dp.adjoint(sol.r, g, mode=R.ADJOINT_IMPLICIT, rho=ad["rho"], law=law, tol=1e-10, max_iters=5000)
    # the implicit gradient as the limit of the unroll, at a tight FP64 fixed point
    s64 = dp.solve_pg(np.zeros(3 * prob.nc), tol=1e-10, precision=R.PRECISION_FP64, **kw)
    imp64 = dp.adjoint(s64.r, g, mode=R.ADJOINT_IMPLICIT, rho=ad["rho"], law=law, tol=1e-10, max_iters=5000)
    # unrolls from the fixed point approach the implicit gradient as their length grows
    limit = {}
    for K_long in (40000, 80000):
        unr = dp.adjoint(s64.r, g, iterations=K_long, rho=ad["rho"], law=law)
        limit[K_long] = ({k: float(np.linalg.norm(imp64[k] - unr[k]) / np.linalg.norm(imp64[k]))
                          for k in ("grad_b", "grad_mu", "grad_minv")}, unr["device_ms"])
    agree, unroll_ms = limit[K_long]
    out = {"contacts": prob.nc, "bodies": prob.nb, "law": "coulomb",
           "forward_solve": {"precision": "mixed", "ms": sol.device_ms, "iterations": sol.iterations,
                             "residual": sol.residual},
           "unroll": {"steps": K, "gradient_ms": ad["device_ms"], "forward_steps_ms": fwd["device_ms"],
                      "gradient_vs_forward": ad["device_ms"] / fwd["device_ms"], "fd_directions": n_dirs,
                      "fd_h": h, "fd_max_rel_error": float(max(errs)), "fd_median_rel_error": float(np.median(errs)),
                      "fd_map": "the map replayed on the recorded piece (projection branches and De Saxce "
                                "directions of the unperturbed run)"},
           "implicit": {"solver": "restarted GMRES(40) on (I - dT/dr^T) lambda = g",
                        "at": f"the FP64 fixed point (FB 1e-10, {s64.iterations} iterations)",
                        "ms": imp64["device_ms"], "arnoldi_steps": imp64["iterations"],
                        "relative_residual": imp64["change"], "converged": bool(imp64["converged"])},
           "implicit_at_tolerance_solution": {
               "at": f"the mixed-precision solution (FB {sol.residual:.1e})", "ms": imp["device_ms"],
               "arnoldi_steps": imp["iterations"], "relative_residual": imp["change"],
               "converged": bool(imp["converged"]),
               "note": "the adjoint system is taken at an iterate that meets the tolerance, not at the fixed "
                       "point; its conditioning depends on that iterate's branches"},
           "implicit_vs_unroll_limit": {"fixed_point": f"FP64 APGD to FB 1e-10 ({s64.iterations} iterations)",
                                        "unroll_steps": K_long, "relative_difference": agree,
                                        "implicit_ms": imp64["device_ms"], "unroll_ms": unroll_ms,
                                        "by_unroll_length": {str(k): v[0] for k, v in limit.items()}}}
    log(f"adjoint-100k (Coulomb): gradient of {K} steps {ad['device_ms']:.1f} ms ({out['unroll']['gradient_vs_forward']:.2f}x "
        f"the forward steps), FD max rel error {max(errs):.1e} on {n_dirs} directions; implicit (GMRES) at the "
        f"FB 1e-10 point {imp64['device_ms']:.0f} ms, {imp64['iterations']} steps (at the 1e-6 solution: "
        f"{imp['iterations']} steps, {'converged' if imp['converged'] else 'not converged'}); implicit vs "
        f"{K_long}-step unroll: "
        + ", ".join(f"{k} {v:.1e}" for k, v in agree.items()))
    return out


def setup_times(ctx, R, G) -> dict:
    """Set-up cost of a contact problem. A new contact set: creating the problem with
    the set-up on the device (default: tiles, transpose, W pattern and values built by
    kernels; the arrays in host memory, so the PCIe upload is included, or already in
    HBM) against the set-up on the host. An unchanged pattern: refreshing the values on
    the device (sb_problem_update_values), from host memory or from HBM."""
    out = {}
    for name, assemble in (("contacts-50k", True), ("zipf-1m", False)):
        p = G.workload(name)
        t0 = time.perf_counter()
        dp = R.DualFrictionProblem(ctx, p, assemble=assemble, host_setup=True)
        ctx.synchronize()
        create_host = time.perf_counter() - t0
        del dp
        t0 = time.perf_counter()
        dp = R.DualFrictionProblem(ctx, p, assemble=assemble)
        ctx.synchronize()
        create = time.perf_counter() - t0
        arrs = []
        for a in (p.hvals, p.minv, p.b, p.mu):
            d = R.DeviceArray(ctx, np.asarray(a).size)
            d.upload(np.asarray(a, dtype=np.float64).reshape(-1))
            arrs.append(d)
        irp = R.DeviceArray(ctx, p.nc + 1, np.int32).upload(p.rowptr)
        ici = R.DeviceArray(ctx, np.asarray(p.colidx).size, np.int32).upload(p.colidx)
        ctx.synchronize()
        t0 = time.perf_counter()
        dd = R.DualFrictionProblem.from_device(ctx, p.nc, p.nb, irp, ici, *arrs, assemble=assemble)
        ctx.synchronize()
        create_dev = time.perf_counter() - t0
        del dd
        host, dev = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            dp.update_values(p.hvals, p.minv, p.b, p.mu)
            host.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            dp.update_values(*arrs)
            dev.append(time.perf_counter() - t0)
        out[name] = {"assembled_w": assemble, "create_host_setup_ms": 1e3 * create_host,
                     "create_device_setup_ms": 1e3 * create, "create_from_device_arrays_ms": 1e3 * create_dev,
                     "update_from_host_ms": 1e3 * float(np.median(host)),
                     "update_on_device_ms": 1e3 * float(np.median(dev))}
        log(f"set-up {name}: create with the host set-up {1e3 * create_host:.1f} ms, device set-up "
            f"{1e3 * create:.1f} ms ({1e3 * create_dev:.1f} ms from arrays in HBM); value refresh "
            f"{out[name]['update_on_device_ms']:.2f} ms ({out[name]['update_from_host_ms']:.2f} ms from host memory)")
        del dp, arrs, irp, ici
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iters", type=int, default=5000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-large", action="store_true", help="skip the 1M-contact SpMV leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile-run", action="store_true",
                    help="short run for rocprofv3: tuned mixed batch solved 3 times + calibration copies, no JSON")
    ap.add_argument("--profile-spmv", action="store_true",
                    help="short run for rocprofv3: the stand-alone products at 1M contacts (zipf-1m: H x, "
                         "H^T r + Minv, matrix-free W x; 20 launches each after one warm-up), no JSON")
    ap.add_argument("--flush", action="store_true",
                    help="with --profile-spmv: evict L2 and the Infinity Cache before every launch (768 MB copy)")
    ap.add_argument("--profile-grid", action="store_true",
                    help="short run for rocprofv3: the single-problem solver on zipf-1m, 100 APGD iterations in "
                         "FP64 and in mixed precision (FP32 stage, then the FP64 stage), twice each, no JSON")
    ap.add_argument("--profile-50k", action="store_true",
                    help="short run for rocprofv3: the 50k-contact problem to FB 1e-8, FP64 APGD (matrix-free) and "
                         "Gauss-Seidel (SOC law, assembled W, greedy colors), twice each, no JSON")
    ap.add_argument("--cpu-baseline-child", type=str, default="", help=argparse.SUPPRESS)
    ap.add_argument("--profile-config", type=str, default="",
                    help="with --profile-run: solve with this launch shape (threads,blocks_per_cu,contacts_per_block) "
                         "after the autotuner instead of the tuned one, so that separate profiler passes see the "
                         "same kernels")
    args = ap.parse_args()
    if args.cpu_baseline_child:
        print(json.dumps(cpu_baseline_child(json.loads(args.cpu_baseline_child))), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this program runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is not implemented)", file=sys.stderr)
        return 2

    _load_package()
    from sobogus_amd import generator as G
    from sobogus_amd import runtime as R

    ctx = R.Context(0)
    log(f"{ctx.arch} {ctx.n_cu} CUs")
    if args.profile_spmv:
        p1m = G.workload("zipf-1m")
        dp1m = R.DualFrictionProblem(ctx, p1m, assemble=False)
        s = 8
        hx = 2 * p1m.nc * (18 * s + 4) + 4 * (p1m.nc + 1) + 6 * p1m.nb * s + 3 * p1m.nc * s
        flush = R.BENCH_FLUSH_CACHES if args.flush else 0
        for kind, name, nbytes in ((1, "H x", hx), (2, "H^T r + Minv", hx + 48 * s * p1m.nb),
                                   (3, "matrix-free W x", wx_matfree_bytes(p1m.nc, p1m.nb))):
            ms = dp1m.bench_spmv(kind | flush, reps=20)
            log(f"profile spmv {name}{' (caches flushed)' if flush else ''}: {1e3 * ms:.1f} us per launch, "
                f"{nbytes} compulsory bytes, {nbytes / (ms * 1e-3) / 1e9:.0f} GB/s")
        ctx.bench_copy(COPY_BYTES, reps=2)
        del dp1m
        ctx.close()
        return 0
    if args.profile_grid:
        p1m = G.workload("zipf-1m")
        dp1m = R.DualFrictionProblem(ctx, p1m, assemble=False)
        r1m = R.DeviceArray(ctx, 3 * p1m.nc)
        kwg = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL,
                   tol=args.tol, max_iters=100)
        for name, prec in (("FP64", R.PRECISION_FP64), ("mixed", R.PRECISION_MIXED)):
            o = R.pg_options(precision=prec, mixed_stall_iters=0, **kwg)
            for _ in range(2):
                r1m.zero()
                s1 = dp1m.solve_pg(r1m, o, on_device=True)
                n32 = s1.spmvs - s1.refine_spmvs
                log(f"profile grid {name}: {s1.device_ms:.1f} ms, {s1.iterations} iterations, {s1.spmvs} operator "
                    f"applications ({n32 if name == 'mixed' else 0} in the FP32 stage), "
                    f"{1e3 * s1.device_ms / s1.spmvs:.0f} us per application")
        ctx.bench_copy(COPY_BYTES, reps=2)
        del r1m, dp1m
        ctx.close()
        return 0
    if args.profile_50k:
        p50 = G.workload("contacts-50k")
        dp50 = R.DualFrictionProblem(ctx, p50, assemble=True)
        r50 = R.DeviceArray(ctx, 3 * p50.nc)
        o = R.pg_options(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL,
                         tol=1e-8, max_iters=20000, precision=R.PRECISION_FP64)
        for _ in range(2):
            r50.zero()
            s = dp50.solve_pg(r50, o, on_device=True)
            log(f"profile 50k APGD FP64: {s.device_ms:.2f} ms, {s.iterations} iterations, {s.spmvs} operator "
                f"applications, {1e3 * s.device_ms / s.spmvs:.1f} us per application, residual {s.residual:.2e}")
        perm, colors = dp50.coloring(R.COLORING_GREEDY)
        dp50.apply_permutation(perm)
        for _ in range(2):
            r50.zero()
            s = dp50.solve_gs(r50, law=R.LAW_SOC, tol=1e-8, max_iters=3000, colors=colors, on_device=True,
                              precision=R.PRECISION_FP64)
            log(f"profile 50k GS (SOC, assembled W, {len(colors) - 1} colors): {s.device_ms:.1f} ms, "
                f"{s.iterations} sweeps, {1e3 * s.device_ms / max(s.iterations, 1):.0f} us per sweep, "
                f"residual {s.residual:.2e}")
        ctx.bench_copy(COPY_BYTES, reps=2)
        del r50, dp50
        ctx.close()
        return 0
    t0 = time.perf_counter()
    probs = G.batch(512)
    ncs = np.array([p.nc for p in probs])
    nbs = np.array([p.nb for p in probs])
    B = R.Batch(ctx, probs)
    log(f"batch-512: {B.total_contacts} contacts; generate + upload {time.perf_counter() - t0:.1f} s")
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=args.tol,
              max_iters=args.max_iters)
    opts64 = R.pg_options(precision=R.PRECISION_FP64, **kw)
    optsmx = R.pg_options(precision=R.PRECISION_MIXED, **kw)
    r = R.DeviceArray(ctx, 3 * B.total_contacts)

    def solve(opts, cfg):
        r.zero()
        return B.solve(r, opts, config=cfg, on_device=True)

    def tune_elsewhere():
        """Launch shapes tuned on another batch of the generator (seeds 5000 + k): the timed
        batch is solved with a shape chosen without its inputs and with no per-problem
        iteration counts (its planner weights every problem alike)."""
        t0 = time.perf_counter()
        B_t = R.Batch(ctx, G.batch(512, seed0=5000))
        c64, _ = B_t.autotune(opts64, probe_iters=40)
        cmx, _ = B_t.autotune(optsmx, probe_iters=40)
        log(f"autotuned on the batch of seeds 5000+k: FP64 {c64}, mixed {cmx} ({time.perf_counter() - t0:.1f} s)")
        del B_t
        return c64, cmx

    if args.profile_run:
        if args.profile_config:
            cfg = tuple(int(v) for v in args.profile_config.split(","))
        else:
            cfg = tune_elsewhere()[1]
        for _ in range(3):
            res = solve(optsmx, cfg)
            st = res.stage_spmvs
            algo = sum(int(st[k, 0]) * wx_matfree_bytes(ncs[k], nbs[k], 4) +
                       int(st[k, 1]) * wx_matfree_bytes(ncs[k], nbs[k], 8) for k in range(len(probs)))
            log(f"profile solve: {res.device_ms:.1f} ms, {res.n_converged} converged, {res.total_spmvs} W x, "
                f"config {res.config}, algorithmic bytes {algo:.6e}")
        ctx.bench_copy(COPY_BYTES, reps=2)
        del r, B
        ctx.close()
        return 0

    def timed(opts, cfg, label):
        for _ in range(args.warmup):
            solve(opts, cfg)
        ctx.synchronize()
        results = []
        t_start = time.perf_counter()
        for _ in range(args.steps):
            results.append(solve(opts, cfg))
        ctx.synchronize()
        wall = time.perf_counter() - t_start
        last = results[-1]
        solved = int(ncs[last.converged].sum())
        rate = solved * args.steps / wall
        log(f"{label}: {1e3 * wall / args.steps:.1f} ms/step (graph {np.mean([x.device_ms for x in results]):.1f} ms), "
            f"{last.n_converged}/{len(probs)} converged, {rate:.4g} contact-solves/s, config {last.config}")
        return last, wall, rate, results

    # headline: FP64 and mixed batches with the shapes tuned on the other batch
    cfg64, cfgmx = tune_elsewhere()
    res64, wall64, rate64, _ = timed(opts64, cfg64, "FP64 batch")
    resmx, wallmx, ratemx, runsmx = timed(optsmx, cfgmx, "mixed batch")
    # FP64-evaluated residual of every returned r (independent of the solver's own report)
    rmx = r.download()
    check = np.array([R.DualFrictionProblem(ctx, p, assemble=False).eval_with(rk, law=R.LAW_COULOMB,
                                                                                op=R.OP_MATRIX_FREE,
                                                                                scaling=R.SCALING_DIAGONAL)
                      for p, rk in zip(probs, B.split(rmx))])
    n_ok = int((check < args.tol).sum())
    log(f"re-evaluated FP64 residuals of the mixed results: {n_ok} below tol, max {check.max():.3e}")
    # the same mixed solve from and to host memory (r crosses PCIe both ways)
    host_walls = []
    for _ in range(3):
        t0 = time.perf_counter()
        resh = B.solve(None, optsmx, config=cfgmx, on_device=False)
        host_walls.append(time.perf_counter() - t0)
    host_rate = int(ncs[resh.converged].sum()) / float(np.median(host_walls))
    log(f"mixed batch with host r: {1e3 * np.median(host_walls):.1f} ms, {host_rate:.4g} contact-solves/s")
    # tuned on this batch itself: the autotuner also records every problem's iteration
    # count in a full solve, and the planner sizes teams and orders queues with them
    same = {}
    for name, opts in (("fp64", opts64), ("mixed", optsmx)):
        t0 = time.perf_counter()
        cfg_s, _ = B.autotune(opts, probe_iters=40)
        rs, ws, rate_s, _ = timed(opts, cfg_s, f"{name} batch, tuned on itself")
        same[name] = {"contact_solves_per_s": rate_s, "ms_per_step": 1e3 * ws / args.steps, "config": list(cfg_s),
                      "converged": int(rs.n_converged), "autotune_s": time.perf_counter() - t0}
    same["speedup_mixed_vs_fp64"] = same["mixed"]["contact_solves_per_s"] / same["fp64"]["contact_solves_per_s"]

    # plain FP64 path: the single-problem solver on each problem in turn
    t0 = time.perf_counter()
    dev_ms = 0.0
    seq_solved = 0
    for p in probs:
        dp = R.DualFrictionProblem(ctx, p, assemble=False)
        rr = R.DeviceArray(ctx, 3 * p.nc).zero()
        s1 = dp.solve_pg(rr, opts64, on_device=True)
        dev_ms += s1.device_ms
        seq_solved += p.nc if s1.converged else 0
        del rr, dp
    seq_wall = time.perf_counter() - t0
    seq_rate = seq_solved / (dev_ms * 1e-3)
    log(f"plain FP64 path (512 single-problem solves): {dev_ms:.0f} ms on the device, {seq_rate:.4g} contact-solves/s")

    # roofline of the dominant kernel: the batch solver graph (FP32 stage + FP64 verification)
    g_ms = float(np.mean([x.device_ms for x in runsmx]))
    st = resmx.stage_spmvs
    algo_bytes = float(sum(int(st[k, 0]) * wx_matfree_bytes(ncs[k], nbs[k], 4) +
                           int(st[k, 1]) * wx_matfree_bytes(ncs[k], nbs[k], 8) for k in range(len(probs))))
    achieved = algo_bytes / (g_ms * 1e-3) / 1e9
    own_bytes = float(sum(int(st[k, 0]) * wx_factored_bytes(ncs[k], nbs[k], 4) +
                          int(st[k, 1]) * wx_factored_bytes(ncs[k], nbs[k], 8) for k in range(len(probs))))
    traffic, traffic_src = measured_traffic("pg_batch_kernel|batch-512|matrix-free|mixed", float(st.sum()))

    # 50,000-contact single problem, FP64 and mixed
    p50 = G.workload("contacts-50k")
    dp50 = R.DualFrictionProblem(ctx, p50, assemble=True)
    r50 = R.DeviceArray(ctx, 3 * p50.nc)
    single = {}
    # APGD's trajectory on this problem turns on rounding-level differences (line-search
    # decisions near ties), the FP32 stage's most of all: the solve is timed from r0 = 0
    # and from four starts perturbed at 1e-8, and the median reported
    starts = [np.zeros(3 * p50.nc)] + [1e-8 * np.random.default_rng(k).standard_normal(3 * p50.nc)
                                        for k in range(1, 5)]
    for name, prec in (("fp64", R.PRECISION_FP64), ("mixed", R.PRECISION_MIXED)):
        o = R.pg_options(precision=prec, **{**kw, "max_iters": 20000})
        r50.zero()
        dp50.solve_pg(r50, o, on_device=True)  # warm-up
        times, its = [], []
        for x0 in starts:
            r50.upload(x0)
            s1 = dp50.solve_pg(r50, o, on_device=True)
            times.append(s1.device_ms)
            its.append(s1.iterations)
        single[name] = {"ms": float(np.median(times)), "ms_per_start": times, "iterations_per_start": its,
                        "residual": s1.residual, "refine_iterations": s1.refine_iterations,
                        "contact_solves_per_s": p50.nc / (np.median(times) * 1e-3),
                        "starts": "r0 = 0 and four starts perturbed by 1e-8 N(0, 1); median"}
    log(f"contacts-50k: FP64 {single['fp64']['ms']:.2f} ms (iterations {single['fp64']['iterations_per_start']}), "
        f"mixed {single['mixed']['ms']:.2f} ms (iterations {single['mixed']['iterations_per_start']}) per solve, median")
    cpu_threads = args.cpu_threads or cpu_share()
    tight = tight_50k(ctx, R, G, 0 if args.no_cpu else cpu_threads)
    adj = adjoint_100k(ctx, R, G)
    setup = setup_times(ctx, R, G)
    nc50, nb50, _, wb50 = dp50.sizes()
    spmv = {}
    for kind, name, nbytes in ((0, "assembled_wx_50k", wx_assembled_bytes(wb50, nc50)),
                               (3, "matrix_free_wx_50k", wx_matfree_bytes(nc50, nb50))):
        ms = dp50.bench_spmv(kind, reps=50)
        spmv[name] = {"ms": ms, "bytes": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9}
    del r50, dp50
    zipf = None
    if not args.no_large:
        p1m = G.workload("zipf-1m")
        dp1m = R.DualFrictionProblem(ctx, p1m, assemble=False)
        s = 8
        hx = 2 * p1m.nc * (18 * s + 4) + 4 * (p1m.nc + 1) + 6 * p1m.nb * s + 3 * p1m.nc * s
        for kind, name, nbytes in ((1, "h_x_1m", hx), (2, "ht_minv_1m", hx + 48 * s * p1m.nb),
                                   (3, "matrix_free_wx_1m", wx_matfree_bytes(p1m.nc, p1m.nb))):
            ms = dp1m.bench_spmv(kind, reps=20)
            msf = dp1m.bench_spmv(kind | R.BENCH_FLUSH_CACHES, reps=10)
            spmv[name] = {"ms": ms, "bytes": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9,
                          "frac_of_peak": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "caches_flushed": {"ms": msf, "GBs": nbytes / (msf * 1e-3) / 1e9,
                                             "frac_of_peak": nbytes / (msf * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                             "note": "each launch timed alone after a 768 MB copy (L2 and "
                                                     "Infinity Cache evicted)"}}
        # the same products with the contacts listed body by body (as a simulator that
        # collects contacts per body passes them): the H^T gathers of contact vectors then
        # run along memory for each contact's first body instead of at random
        p1b = G.body_order(p1m)
        dp1b = R.DualFrictionProblem(ctx, p1b, assemble=False)
        for kind, name in ((1, "h_x_1m"), (2, "ht_minv_1m"), (3, "matrix_free_wx_1m")):
            nbytes = spmv[name]["bytes"]
            ms = dp1b.bench_spmv(kind, reps=20)
            msf = dp1b.bench_spmv(kind | R.BENCH_FLUSH_CACHES, reps=10)
            spmv[name + "_body_order"] = {
                "ms": ms, "bytes": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9,
                "frac_of_peak": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "caches_flushed": {"ms": msf, "GBs": nbytes / (msf * 1e-3) / 1e9,
                                   "frac_of_peak": nbytes / (msf * 1e-3) / 1e9 / HBM_PEAK_GBS},
                "note": "zipf-1m with its contacts in body order (generator.body_order)"}
        del dp1b, p1b
        # the 1M-contact Zipf problem has no solution at FB 1e-6 (its hubs over-constrain the
        # bodies); FP64 vs mixed is compared on a fixed number of APGD iterations
        zipf = {}
        r1m = R.DeviceArray(ctx, 3 * p1m.nc)
        for name, prec in (("fp64", R.PRECISION_FP64), ("mixed", R.PRECISION_MIXED)):
            o = R.pg_options(precision=prec, mixed_stall_iters=0, **{**kw, "max_iters": 100})
            best = None
            for _ in range(2):
                r1m.zero()
                s1 = dp1m.solve_pg(r1m, o, on_device=True)
                if best is None or s1.device_ms < best.device_ms:
                    best = s1
            if name == "fp64":
                zipf[name] = {"ms": best.device_ms, "iterations": best.iterations, "spmvs": best.spmvs,
                              "us_per_operator_application": 1e3 * best.device_ms / best.spmvs,
                              "residual": best.residual}
            else:
                n32 = best.spmvs - best.refine_spmvs
                zipf["fp32_stage"] = {"ms": best.mixed_device_ms, "iterations": best.iterations - best.refine_iterations,
                                      "spmvs": n32, "us_per_operator_application": 1e3 * best.mixed_device_ms / n32}
                zipf["mixed_fp64_stage"] = {"ms": best.device_ms - best.mixed_device_ms,
                                            "iterations": best.refine_iterations, "spmvs": best.refine_spmvs,
                                            "us_per_operator_application":
                                                1e3 * (best.device_ms - best.mixed_device_ms) / best.refine_spmvs}
        zipf["fp32_vs_fp64_per_application"] = (zipf["fp64"]["us_per_operator_application"] /
                                                zipf["fp32_stage"]["us_per_operator_application"])
        zipf["stand_alone_matrix_free_wx_us"] = 1e3 * spmv["matrix_free_wx_1m"]["ms"]
        zipf["fp64_in_solver_vs_stand_alone"] = (zipf["fp64"]["us_per_operator_application"] /
                                                 zipf["stand_alone_matrix_free_wx_us"])
        del r1m, dp1m
        log("1M-contact SpMV: " + ", ".join(f"{k} {v['GBs']:.0f} GB/s" for k, v in spmv.items() if "_1m" in k))
        log(f"zipf-1m, 100 APGD iterations: FP64 {zipf['fp64']['us_per_operator_application']:.0f} us, FP32 stage "
            f"{zipf['fp32_stage']['us_per_operator_application']:.0f} us per operator application "
            f"(stand-alone FP64 W x {zipf['stand_alone_matrix_free_wx_us']:.0f} us)")
    ms_copy = ctx.bench_copy(COPY_BYTES, reps=10)
    copy_gbs = 2 * COPY_BYTES / (ms_copy * 1e-3) / 1e9

    value = ratemx
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "contact-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wallmx / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 bulk + f64 residual/decisions/verification",
        "data": "synthetic",
        "config": {"workload": "batch-512", "problems": len(probs), "total_contacts": int(B.total_contacts),
                   "contacts_per_problem": "floor(exp(U[ln 64, ln 20000])), bodies = contacts/2",
                   "solver": "APGD", "law": "coulomb", "operator": "matrix-free", "tol": args.tol,
                   "max_iters": args.max_iters, "precision": "mixed", "launch": "one hipGraph per step",
                   "seeds": "4000 + k", "launch_shape": list(cfgmx),
                   "launch_shape_tuned_on": "another batch of the generator (seeds 5000 + k); no iteration "
                                            "counts of the timed problems"},
        "mixed": {"contact_solves_per_s": ratemx, "ms_per_step": 1e3 * wallmx / args.steps,
                  "graph_ms": g_ms, "converged": int(resmx.n_converged),
                  "fp64_checked_below_tol": n_ok, "refined_problems": int(resmx.refined_problems),
                  "total_iterations": int(resmx.total_iterations), "max_iterations": int(resmx.max_iterations),
                  "with_host_r": {"contact_solves_per_s": host_rate, "ms": 1e3 * float(np.median(host_walls)),
                                  "note": "r from and to pageable host memory (PCIe both ways), not the value"}},
        "fp64": {"contact_solves_per_s": rate64, "ms_per_step": 1e3 * wall64 / args.steps,
                 "converged": int(res64.n_converged), "total_iterations": int(res64.total_iterations),
                 "max_iterations": int(res64.max_iterations), "launch_shape": list(cfg64)},
        "plain_fp64_sequential": {"contact_solves_per_s": seq_rate, "device_ms": dev_ms, "wall_s": seq_wall,
                                  "note": "single-problem persistent solver, 512 launches one after the other"},
        "speedup_mixed_vs_fp64_batch": ratemx / rate64,
        "tuned_on_the_timed_batch": same,
        "speedup_mixed_batch_vs_plain_fp64": ratemx / seq_rate,
        "contacts_50k": single,
        "contacts_50k_tol_1e-8": tight,
        "adjoint_100k": adj,
        "setup": setup,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "pg_batch_kernel (FP32 stage + FP64 verification stage, one graph)",
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "model": "sum over problems of operator applications x compulsory bytes of one matrix-free "
                              "W x (4-byte values in the FP32 stage, 8-byte in the FP64 stage)",
                     "kernel_compulsory_bytes_per_launch": own_bytes,
                     "kernel_note": "the solver applies the factored product F (F^T x), F = H L, Minv = L L^T: "
                                    "it reads no Minv blocks, so its own compulsory bytes are below the model's"},
        "spmv": spmv,
        "zipf_1m_fixed_iterations": zipf,
        "stream_copy_GBs": copy_gbs,
    }
    if not args.no_cpu:
        log(f"CPU baseline on {cpu_threads} threads ...")
        out["cpu_baseline"] = cpu_baseline_batch(args.tol, args.max_iters, cpu_threads)
    print(json.dumps(out), flush=True)
    del r, B
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
