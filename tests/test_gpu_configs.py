"""The benchmark workloads at their full sizes on the GPU, against the oracle.

* batch-512 (512 problems, n_c log-uniform in [64, 20000], seeds 4000 + k,
  1,839,956 contacts): mixed and FP64 at the benchmark's launch shape (256 threads,
  2 workgroups per CU, 1000 contacts per workgroup: teams of up to 16 workgroups),
  with teams on one XCD and spread over the XCDs; every returned r re-evaluated by
  the oracle (Coulomb law, diagonal scaling) below 1.05 tol, and every problem
  that did not converge also stalls in the oracle's APGD.
* one 20,000-contact problem on a 16-workgroup team against the grid solver, 15
  FP64 iterations: the same iterates up to summation order (1e-6 relative).
* zipf-1m (1,000,000 contacts, 300,000 bodies, Zipf(1.2) body degrees): the
  matrix-free W x against the oracle (1e-12 relative), and the first 20 FP64 APGD
  residuals against the oracle's same-algorithm run (1e-6 relative).
* contacts-50k (50,000 contacts, 20,000 bodies): FP64 APGD (matrix-free, Coulomb)
  and colored Gauss-Seidel on the assembled W (SOC law, greedy coloring) to FB 1e-8,
  each result re-evaluated by the oracle below 1.05e-8.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BENCH_SHAPE = (256, 2, 1000)  # threads, workgroups per CU, contacts per workgroup


@pytest.fixture(scope="module")
def batch512(G):
    return G.batch(512)


def _team_sizes(probs, cpb, grid):
    """Team size of each problem under the batch plan without iteration weights:
    pow2ceil(n_c / cpb), at most min(16, grid / 2)."""
    kcap = 1
    while kcap * 2 <= min(16, grid // 2):
        kcap *= 2
    out = []
    for p in probs:
        need = -(-p.nc // cpb)
        k = 1
        while k < need:
            k *= 2
        out.append(min(k, kcap))
    return np.array(out)


def _oracle_check(O, probs, res, out, tol, max_iters):
    worst = 0.0
    for k, (p, rk) in enumerate(zip(probs, out)):
        op = O.OracleProblem(p, assemble=False)
        if res.converged[k]:
            chk = op.eval(rk, coulomb=True, assembled=False, diag_scaling=True)
            worst = max(worst, chk)
            assert chk <= 1.05 * tol, f"problem {k}: nc={p.nc} oracle residual {chk:.3e}"
        else:
            ores, _, _ = op.pg_solve(np.zeros(3 * p.nc), tol=tol, max_iters=max_iters, coulomb=True,
                                     assembled=False, diag_scaling=True)
            assert ores >= tol, f"problem {k}: the oracle converges ({ores:.3e}) but the batch did not"
    return worst


@pytest.mark.parametrize("spread", [False, True], ids=["one-xcd", "spread"])
@pytest.mark.parametrize("precision", ["PRECISION_MIXED", "PRECISION_FP64"])
def test_batch512_full_size(R, O, ctx, batch512, precision, spread):
    print(f"[configs] batch-512 {precision} spread={spread} shape {BENCH_SHAPE}", flush=True)
    O.set_threads(16)
    ks = _team_sizes(batch512, BENCH_SHAPE[2], BENCH_SHAPE[1] * ctx.n_cu)
    assert (ks == 16).sum() > 0 and (ks == 8).sum() > 0  # the shape makes teams of 8 and 16
    B = R.Batch(ctx, batch512, spread_teams=spread)
    res = B.solve(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-6,
                  max_iters=5000, precision=getattr(R, precision), config=BENCH_SHAPE)
    print(f"   {res.n_converged}/512 converged, {res.device_ms:.1f} ms, config {res.config}", flush=True)
    assert res.config == BENCH_SHAPE
    assert res.n_converged >= 510
    worst = _oracle_check(O, batch512, res, B.split(res.r), 1e-6, 5000)
    print(f"   worst oracle residual of the converged problems {worst:.3e}", flush=True)
    # the batch's own FP64 residuals agree with the oracle's re-evaluation
    assert np.all(res.residuals[res.converged] < 1e-6)


def test_batch_team_of_16_matches_grid_solver(R, G, ctx):
    """A 20,000-contact problem solved by a team of 16 workgroups (both placements) and by
    the single-problem grid solver: 15 FP64 iterations give the same iterates up to
    summation order."""
    print("[configs] 20k problem on a 16-workgroup team", flush=True)
    p = G.generate(20000, 10000, seed=4242)
    assert _team_sizes([p], 1000, 2 * ctx.n_cu)[0] == 16
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-30,
              max_iters=15, precision=R.PRECISION_FP64)
    g = R.DualFrictionProblem(ctx, p, assemble=False).solve_pg(np.zeros(3 * p.nc), **kw)
    for spread in (False, True):
        B = R.Batch(ctx, [p], spread_teams=spread)
        t = B.solve(config=BENCH_SHAPE, **kw)
        assert t.status == R.SB_NOT_CONVERGED and t.iterations[0] == 15
        rel = np.linalg.norm(t.r - g.r) / np.linalg.norm(g.r)
        print(f"   spread={spread}: |r_team - r_grid| / |r_grid| = {rel:.2e}", flush=True)
        assert rel <= 1e-6
    # and the mixed-precision team solve converges to an oracle-verified result
    t = B.solve(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-6,
                max_iters=5000, precision=R.PRECISION_MIXED, config=BENCH_SHAPE)
    assert t.converged[0]


@pytest.fixture(scope="module")
def zipf1m(G):
    return G.workload("zipf-1m")


def test_zipf_1m_matrix_free_product(R, O, ctx, zipf1m):
    print("[configs] zipf-1m matrix-free W x vs the oracle", flush=True)
    O.set_threads(16)
    p = zipf1m
    dp = R.DualFrictionProblem(ctx, p, assemble=False)
    op = O.OracleProblem(p, assemble=False)
    x = np.random.default_rng(31).standard_normal(3 * p.nc)
    y = dp.apply_w(x, op=R.OP_MATRIX_FREE)
    ref = op.apply_w(x, assembled=False)
    rel = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"   relative difference {rel:.2e} (max body degree {np.bincount(p.colidx).max()})", flush=True)
    assert rel <= 1e-12


def test_zipf_1m_apgd_first_residuals(R, O, ctx, zipf1m):
    """The first 20 FP64 APGD residuals (matrix-free, Coulomb, diagonal scaling, cold
    start) equal the oracle's to 1e-6: same algorithm at 1M contacts with hub rows."""
    print("[configs] zipf-1m 20 APGD iterations vs the oracle", flush=True)
    O.set_threads(16)
    p = zipf1m
    dp = R.DualFrictionProblem(ctx, p, assemble=False)
    g = dp.solve_pg(np.zeros(3 * p.nc), variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE,
                    scaling=R.SCALING_DIAGONAL, tol=1e-30, max_iters=20, hist_cap=64)
    op = O.OracleProblem(p, assemble=False)
    ores, ox, oinfo = op.pg_solve(np.zeros(3 * p.nc), tol=1e-30, max_iters=20, coulomb=True, assembled=False,
                                  diag_scaling=True)
    k = min(20, len(g.history), len(oinfo["history"]))
    print(f"   gpu {g.history[:k][-1]:.6e} oracle {oinfo['history'][k - 1]:.6e} after {k} residuals", flush=True)
    assert k == 20
    assert np.allclose(g.history[:k], oinfo["history"][:k], rtol=1e-6)


@pytest.fixture(scope="module")
def c250k(G):
    # 250,000 contacts / 100,000 bodies (the oracle's APGD passes FB 1e-6 at iteration 1220)
    return G.generate(250_000, 100_000, seed=12)


@pytest.mark.parametrize("variant,precision", [("PG_APGD", "PRECISION_FP64"), ("PG_APGD", "PRECISION_MIXED"),
                                               ("PG_SPG", "PRECISION_FP64")])
def test_candidate_phase_solver_250k(R, O, ctx, c250k, variant, precision):
    """From 200,000 contacts on, matrix-free single-problem solves form each line-search
    candidate in a contact-major phase of its own (the H^T pass then gathers one
    vector). At 250,000 contacts: the first residuals follow the oracle's same-algorithm
    run (1e-6), APGD reaches FB 1e-6 with a result the oracle re-evaluates below
    1.05e-6, and an SPG run of 60 iterations returns an iterate whose residual the
    oracle confirms."""
    print(f"[configs] 250k contacts, candidate phase, {variant} {precision}", flush=True)
    O.set_threads(16)
    p = c250k
    dp = R.DualFrictionProblem(ctx, p, assemble=False)
    spg = variant == "PG_SPG"
    tol, iters = (1e-30, 60) if spg else (1e-6, 10000)
    g = dp.solve_pg(np.zeros(3 * p.nc), variant=getattr(R, variant), law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE,
                    scaling=R.SCALING_DIAGONAL, tol=tol, max_iters=iters, precision=getattr(R, precision),
                    hist_cap=64)
    op = O.OracleProblem(p, assemble=False)
    chk = op.eval(g.r, coulomb=True, assembled=False, diag_scaling=True)
    print(f"   {g.iterations} iterations, residual {g.residual:.3e}, oracle {chk:.3e}", flush=True)
    assert chk == pytest.approx(g.residual, rel=1e-6)
    if not spg:
        assert g.converged and chk <= 1.05 * tol
    if precision == "PRECISION_FP64":
        ores, ox, oinfo = op.pg_solve(np.zeros(3 * p.nc), tol=1e-30, max_iters=10, coulomb=True, assembled=False,
                                      diag_scaling=True, variant=getattr(O, variant))
        assert np.allclose(g.history[:10], oinfo["history"][:10], rtol=1e-6)


@pytest.fixture(scope="module")
def c50k(G):
    return G.workload("contacts-50k")


def test_contacts_50k_apgd_fp64_to_1e8(R, O, ctx, c50k):
    print("[configs] contacts-50k FP64 APGD to 1e-8", flush=True)
    O.set_threads(16)
    p = c50k
    dp = R.DualFrictionProblem(ctx, p, assemble=False)
    g = dp.solve_pg(np.zeros(3 * p.nc), variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE,
                    scaling=R.SCALING_DIAGONAL, tol=1e-8, max_iters=20000)
    chk = O.OracleProblem(p, assemble=False).eval(g.r, coulomb=True, assembled=False, diag_scaling=True)
    print(f"   {g.iterations} iterations, residual {g.residual:.3e}, oracle {chk:.3e}", flush=True)
    assert g.converged and chk <= 1.05e-8
    assert chk == pytest.approx(g.residual, rel=1e-6)


def test_contacts_50k_gs_assembled_to_1e8(R, O, ctx, c50k):
    print("[configs] contacts-50k colored GS (assembled W, SOC) to 1e-8", flush=True)
    O.set_threads(16)
    p = c50k
    dp = R.DualFrictionProblem(ctx, p, assemble=True)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    g = dp.solve_gs(np.zeros(3 * p.nc), law=R.LAW_SOC, tol=1e-8, max_iters=3000, colors=colors)
    op = O.OracleProblem(p, assemble=True)
    op.permute(perm.astype(np.int32))
    chk = op.eval(g.r, coulomb=False)
    print(f"   {len(colors) - 1} colors, {g.iterations} sweeps, residual {g.residual:.3e}, oracle {chk:.3e}",
          flush=True)
    assert g.converged and chk <= 1.05e-8
