"""Batched projected gradient on the GPU: one launch (captured in a hipGraph)
in which every workgroup solves whole problems taken from a queue.

Parity bar, per problem: the returned r re-evaluated by the FP64 oracle meets the
tolerance (x1.05 for the residual's summation order) and the solve converged;
FP64 batch iteration counts within 25% of the oracle's same-algorithm solve
(reductions run in a different fixed order); graph replays reproduce the same
bits; the autotuned geometry gives results of the same quality.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch(G):
    # 24 problems, n_c log-uniform in [64, 2000], n_b = n_c / 2 (the batch generator at a smaller range)
    return G.batch(24, seed0=4000, lo=64, hi=2000)


def _check_each(O, probs, res, out, tol, max_iters):
    """Converged problems: the oracle re-evaluates r below the tolerance. Problems
    that did not converge: APGD stalls on them in the oracle as well (some
    generated problems do not reach the tolerance with APGD at all)."""
    for k, (p, rk) in enumerate(zip(probs, out)):
        op = O.OracleProblem(p)
        if res.converged[k]:
            chk = op.eval(rk, coulomb=True)
            assert chk <= 1.05 * tol, f"problem {k}: nc={p.nc} oracle residual {chk:.3e}"
        else:
            ores, _, _ = op.pg_solve(np.zeros(3 * p.nc), tol=tol, max_iters=max_iters, coulomb=True,
                                     diag_scaling=True)
            assert ores >= tol, f"problem {k}: the oracle converges ({ores:.3e}) but the batch did not"


@pytest.mark.parametrize("op", ["OP_MATRIX_FREE", "OP_ASSEMBLED"])
def test_batch_fp64_matches_oracle(R, O, ctx, small_batch, op):
    print(f"[batch] fp64 {op}", flush=True)
    B = R.Batch(ctx, small_batch, assemble=(op == "OP_ASSEMBLED"))
    res = B.solve(variant=R.PG_APGD, law=R.LAW_COULOMB, op=getattr(R, op), scaling=R.SCALING_DIAGONAL, tol=1e-8,
                  max_iters=5000, precision=R.PRECISION_FP64)
    assert res.n_converged >= len(small_batch) - 2
    out = B.split(res.r)
    _check_each(O, small_batch, res, out, 1e-8, 5000)
    for k in (0, 5, 11):
        p = small_batch[k]
        ores, ox, oinfo = O.OracleProblem(p).pg_solve(np.zeros(3 * p.nc), tol=1e-8, max_iters=5000, coulomb=True,
                                                     diag_scaling=True)
        assert abs(res.iterations[k] - oinfo["iterations"]) <= 0.25 * oinfo["iterations"] + 5


def test_batch_mixed_converges_fp64_verified(R, O, ctx, small_batch):
    print("[batch] mixed matrix-free", flush=True)
    B = R.Batch(ctx, small_batch)
    res = B.solve(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-6,
                  max_iters=5000, precision=R.PRECISION_MIXED)
    assert res.n_converged >= len(small_batch) - 2
    _check_each(O, small_batch, res, B.split(res.r), 1e-6, 5000)
    assert np.all(res.residuals[res.converged] < 1e-6)


def test_batch_graph_replay_is_deterministic(R, ctx, small_batch):
    B = R.Batch(ctx, small_batch)
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-6,
              max_iters=5000, precision=R.PRECISION_MIXED)
    first = B.solve(**kw)
    # eager work on the same stream between replays
    dp = R.DualFrictionProblem(ctx, small_batch[0], assemble=True)
    dp.solve_pg(np.zeros(3 * small_batch[0].nc), law=R.LAW_COULOMB, tol=1e-8, max_iters=5000)
    for _ in range(3):
        again = B.solve(**kw)
        assert np.array_equal(again.r, first.r)
        assert np.array_equal(again.iterations, first.iterations)
    # another grid size with the same workgroup size and one workgroup per problem:
    # a problem's arithmetic does not depend on which workgroup solves it
    one = B.solve(config=(512, 1, -1), **kw)
    other = B.solve(config=(512, 2, -1), **kw)
    assert np.array_equal(other.r, one.r)


def test_batch_warm_start_and_autotune(R, O, ctx, small_batch):
    B = R.Batch(ctx, small_batch)
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-6,
              max_iters=5000, precision=R.PRECISION_MIXED)
    cfg, ms = B.autotune(probe_iters=20, **kw)
    assert cfg[0] in (256, 512, 1024) and 1 <= cfg[1] <= 8 and cfg[2] != 0 and ms > 0
    res = B.solve(**kw)
    assert res.config == cfg
    warm = B.solve(r0=res.r, **kw)
    # converged results stay converged; the FP32 stage may take a few iterations to
    # get from the FP64-verified residual (< tol) below its own target (tol * 0.95)
    assert warm.n_converged >= res.n_converged
    assert np.all(warm.converged[res.converged])
    assert warm.iterations[res.converged].sum() < 0.5 * res.iterations[res.converged].sum()


def test_batch_hub_rows_match_grid_solver(G, R, ctx):
    """Problems with bodies of degree > 64 (long rows in H^T, summed through the
    carry buffer inside one workgroup): a fixed number of iterations gives the same
    iterates as the grid solver up to summation order."""
    probs = [G.generate(3000, 60, seed=77 + k) for k in range(3)]
    B = R.Batch(ctx, probs)
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-30,
              max_iters=15, precision=R.PRECISION_FP64)
    single = B.solve(config=(512, 1, -1), **kw)   # one workgroup per problem
    teams = B.solve(config=(512, 1, 1000), **kw)  # teams of 4 workgroups
    assert single.status == R.SB_NOT_CONVERGED and teams.status == R.SB_NOT_CONVERGED
    for p, r1, r4 in zip(probs, B.split(single.r), B.split(teams.r)):
        assert np.bincount(p.colidx, minlength=p.nb).max() > 64
        dp = R.DualFrictionProblem(ctx, p, assemble=False)
        g = dp.solve_pg(np.zeros(3 * p.nc), **kw)
        assert np.linalg.norm(r1 - g.r) <= 1e-6 * np.linalg.norm(g.r)
        assert np.linalg.norm(r4 - g.r) <= 1e-6 * np.linalg.norm(g.r)


def test_batch_multi_workgroup_teams(R, O, ctx, small_batch):
    """Teams of 2-4 workgroups per problem (contacts_per_block 500) give
    oracle-verified results, and replays of the same plan are bit-identical."""
    B = R.Batch(ctx, small_batch)
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-6,
              max_iters=5000, precision=R.PRECISION_FP64, config=(256, 2, 500))
    res = B.solve(**kw)
    assert res.n_converged >= len(small_batch) - 2
    _check_each(O, small_batch, res, B.split(res.r), 1e-6, 5000)
    again = B.solve(**kw)
    assert np.array_equal(again.r, res.r)


def test_batch_static_body_problem(R, O, ctx, small_batch):
    """A problem with a static body (Minv_k = 0, not positive definite) keeps the
    unfactored matrix-free product inside a batch whose other problems run the
    factored one; every result is oracle-verified."""
    print("[batch] static body", flush=True)
    probs = list(small_batch[:8])
    probs[3] = dataclasses.replace(probs[3], minv=probs[3].minv.copy())  # the fixture stays unchanged
    probs[3].minv[0] = 0.0
    B = R.Batch(ctx, probs)
    res = B.solve(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-8,
                  max_iters=5000, precision=R.PRECISION_FP64)
    assert res.n_converged >= len(probs) - 2
    _check_each(O, probs, res, B.split(res.r), 1e-8, 5000)


def test_batch_team_placement_is_transparent(R, ctx, small_batch):
    """Teams of several workgroups run on one XCD by default (L2 kept, L1-only acquire);
    spread over the XCDs they write back and invalidate L2 at every barrier. Either
    way a problem's arithmetic is the same: results are bit-identical, both stages."""
    print("[batch] team placement", flush=True)
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-6,
              max_iters=5000, precision=R.PRECISION_MIXED, config=(256, 2, 500))
    B = R.Batch(ctx, small_batch)
    S = R.Batch(ctx, small_batch, spread_teams=True)
    rb = B.solve(**kw)
    rs = S.solve(**kw)
    assert rb.n_converged >= len(small_batch) - 2
    assert np.array_equal(rb.iterations, rs.iterations)
    assert np.array_equal(rb.r, rs.r)


def test_batch_locality_order_is_transparent(R, O, ctx, small_batch):
    """The batch stores each problem in its locality order by default; r goes in and
    comes out in the caller's order (host and device r), the same as a batch kept in
    the caller's order up to summation order, and every result is oracle-verified."""
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-6,
              max_iters=5000, precision=R.PRECISION_FP64)
    B = R.Batch(ctx, small_batch)
    C = R.Batch(ctx, small_batch, caller_order=True)
    rb = B.solve(config=(512, 1, -1), **kw)
    rc = C.solve(config=(512, 1, -1), **kw)
    _check_each(O, small_batch, rb, B.split(rb.r), 1e-6, 5000)
    assert rb.n_converged >= len(small_batch) - 2
    # (APGD trajectories depend on summation order: a problem that stalls in one order
    # may converge in the other)
    for k, (a, c) in enumerate(zip(B.split(rb.r), C.split(rc.r))):
        if rb.converged[k] and rc.converged[k]:
            assert np.linalg.norm(a - c) <= 5e-2 * max(np.linalg.norm(c), 1e-12), k
    # a warm start from the caller-order result, device-resident: converges at once
    d = R.DeviceArray(ctx, 3 * B.total_contacts)
    d.upload(rc.r)
    warm = B.solve(d, R.pg_options(**kw), config=(512, 1, -1), on_device=True)
    assert all(warm.iterations[k] == 0 for k in range(len(small_batch)) if rc.converged[k])
    back = d.download()
    for k, (a, c) in enumerate(zip(B.split(back), C.split(rc.r))):
        if rc.converged[k]:
            assert np.linalg.norm(a - c) <= 1e-3 * max(np.linalg.norm(c), 1e-12), k


def test_batch_iteration_hints(R, O, ctx, small_batch):
    """Caller-provided iteration counts change the schedule (team sizes, queue order),
    not the quality: every result still meets the tolerance in the oracle. They read
    back as set; None forgets them; a negative count is refused."""
    print("[batch] iteration hints", flush=True)
    B = R.Batch(ctx, small_batch)
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-6,
              max_iters=5000, precision=R.PRECISION_MIXED)
    cfg = (256, 1, 250)  # teams of several workgroups for the larger problems
    assert np.all(B.iteration_hints() == 0)
    base = B.solve(config=cfg, **kw)
    # the largest problem expected to iterate longest, the others short: its team grows
    hints = np.full(B.n, 50, dtype=np.int32)
    hints[int(np.argmax(B.ncs))] = 2000
    B.set_iteration_hints(hints)
    assert np.array_equal(B.iteration_hints(), hints)
    res = B.solve(config=cfg, **kw)
    assert res.n_converged == base.n_converged
    _check_each(O, small_batch, res, B.split(res.r), 1e-6, 5000)
    # the counts of a previous solve (what a simulation passes from its last step)
    B.set_iteration_hints(base.iterations)
    res2 = B.solve(config=cfg, **kw)
    _check_each(O, small_batch, res2, B.split(res2.r), 1e-6, 5000)
    B.set_iteration_hints(None)
    assert np.all(B.iteration_hints() == 0)
    again = B.solve(config=cfg, **kw)
    assert np.array_equal(again.r, base.r)
    with pytest.raises(R.SbError):
        B.set_iteration_hints(np.full(B.n, -1, dtype=np.int32))
