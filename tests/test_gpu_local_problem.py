"""A dual problem given by W itself (sb_problem_create_from_w) and FCLib's merit1 (sb_merit1),
the path of the reference's FCLib application (apps/FCLibLoader.cpp:40-73), against the
oracle: the converted and pruned blocks bit for bit, solutions by the oracle's residual,
merit1 against its restatement."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def full_from_lower(nc, rp, ci, bv):
    """The 3n x 3n matrix of a lower-triangle block CSR: blocks as stored, the strictly lower
    ones mirrored (a diagonal block is kept as it is)."""
    W = np.zeros((3 * nc, 3 * nc))
    for i in range(nc):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            B = np.asarray(bv[p]).reshape(3, 3)
            W[3 * i:3 * i + 3, 3 * j:3 * j + 3] = B
            if j != i:
                W[3 * j:3 * j + 3, 3 * i:3 * i + 3] = B.T
    return W


def dense_to_csr(A):
    rp, ci, v = [0], [], []
    for row in A:
        nz = np.nonzero(row)[0]
        ci.extend(nz.tolist())
        v.extend(row[nz].tolist())
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)


@pytest.fixture(scope="module")
def pair(R, O, G, ctx):
    """small-1k from H (assembled W) and the same W given as a scalar matrix, with a few tiny
    blocks added to the lower triangle that prune removes."""
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    W = full_from_lower(prob.nc, *dp.w_lower())
    rng = np.random.default_rng(11)
    for _ in range(25):
        i, j = sorted(rng.integers(0, prob.nc, 2))[::-1]
        if i != j and not W[3 * i:3 * i + 3, 3 * j:3 * j + 3].any():
            tiny = 1e-13 * rng.standard_normal((3, 3))
            W[3 * i:3 * i + 3, 3 * j:3 * j + 3] = tiny
            W[3 * j:3 * j + 3, 3 * i:3 * i + 3] = tiny.T
    csr = dense_to_csr(W)
    dw = R.DualFrictionProblem.from_w(ctx, prob.nc, *csr, prob.b, prob.mu, prune_tol=1e-12)
    return prob, dp, dw, W, csr


def test_from_w_blocks_equal_the_oracle_conversion(O, pair):
    prob, dp, dw, W, csr = pair
    want = O.prune_blocks(*O.convert_symmetric_3x3(prob.nc, *csr), 1e-12)
    got = dw.w_lower()
    print(f"case: from_w 1k, {len(want[1])} blocks after prune (H-built W: {dp.w_lower()[0][-1]})", flush=True)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert np.array_equal(got[2], want[2].reshape(-1, 9))
    # the tiny blocks are gone: the pattern is the H-built one
    assert np.array_equal(got[0], dp.w_lower()[0]) and np.array_equal(got[1], dp.w_lower()[1])


def test_from_w_solves_the_problem(R, O, pair):
    prob, dp, dw, W, csr = pair
    op = O.OracleProblem(prob)
    r = dw.solve_pg(np.zeros(3 * prob.nc), variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_ASSEMBLED, tol=1e-8,
                    max_iters=3000)
    ref = dp.solve_pg(np.zeros(3 * prob.nc), variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_ASSEMBLED, tol=1e-8,
                      max_iters=3000)
    print(f"case: from_w APGD 1k: {r.iterations} it res {r.residual:.2e}; from H {ref.iterations} it", flush=True)
    assert r.converged and op.eval(r.r) <= 1.05e-8
    perm, colors = dw.coloring(R.COLORING_GREEDY)
    dw2 = R.DualFrictionProblem.from_w(dw.ctx, prob.nc, *csr, prob.b, prob.mu, prune_tol=1e-12)
    dw2.apply_permutation(perm)
    g = dw2.solve_gs(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, tol=1e-8, max_iters=1000, colors=colors)
    rg = np.zeros(3 * prob.nc)
    rg.reshape(-1, 3)[perm] = g.r.reshape(-1, 3)  # back to the caller's contact order
    print(f"case: from_w GS 1k: {len(colors) - 1} colors, {g.iterations} sweeps res {g.residual:.2e}", flush=True)
    assert g.converged and op.eval(rg) <= 1.05e-8


def test_merit1_against_the_oracle(R, O, pair):
    prob, dp, dw, W, csr = pair
    rng = np.random.default_rng(12)
    r = rng.standard_normal(3 * prob.nc)
    Wc = full_from_lower(prob.nc, *dw.w_lower())
    want = O.merit1(r, Wc @ r + prob.b, prob.b, prob.mu)
    got = dw.merit1(r)
    print(f"case: merit1 random r: {got:.12e} oracle {want:.12e}", flush=True)
    assert got == pytest.approx(want, rel=1e-12)
    Wh = full_from_lower(prob.nc, *dp.w_lower())
    want_h = O.merit1(r, Wh @ r + prob.b, prob.b, prob.mu)
    assert dp.merit1(r, op=R.OP_ASSEMBLED) == pytest.approx(want_h, rel=1e-12)
    assert dp.merit1(r, op=R.OP_MATRIX_FREE) == pytest.approx(want_h, rel=1e-10)
    # at a solution the merit is small and equal to the oracle's to rounding
    s = dw.solve_pg(np.zeros(3 * prob.nc), variant=R.PG_SPG, law=R.LAW_COULOMB, op=R.OP_ASSEMBLED, tol=1e-8,
                    max_iters=5000)
    m = dw.merit1(s.r)
    assert m == pytest.approx(O.merit1(s.r, Wc @ s.r + prob.b, prob.b, prob.mu), rel=1e-6, abs=1e-13)
    assert m < 1e-3 * dw.merit1(np.zeros(3 * prob.nc))


def test_from_w_refuses_what_needs_h(R, pair):
    prob, dp, dw, W, csr = pair
    z = np.zeros(3 * prob.nc)
    for call in (lambda: dw.solve_pg(z, op=R.OP_MATRIX_FREE, tol=1e-6, max_iters=10),
                 lambda: dw.apply_w(z, op=R.OP_MATRIX_FREE),
                 lambda: dw.merit1(z, op=R.OP_MATRIX_FREE),
                 lambda: dw.solve_pgs(z, tol=1e-6, max_iters=10)):
        with pytest.raises(R.SbError):
            call()


def test_local_problem_file_to_solution(R, O, ctx, pair, tmp_path):
    # the FCLib application's flow with this library's file layout: read, convert, solve, merit1
    prob, dp, dw, W, csr = pair
    path = tmp_path / "small-1k.sblp"
    R.write_local_problem(path, *csr, prob.b, prob.mu)
    d = R.read_local_problem(path)
    dl = R.DualFrictionProblem.from_w(ctx, d["n"], d["rowptr"], d["colidx"], d["vals"], d["q"], d["mu"],
                                      prune_tol=1e-12)
    assert all(np.array_equal(a, b) for a, b in zip(dl.w_lower(), dw.w_lower()))
    s = dl.solve_pg(np.zeros(3 * d["n"]), variant=R.PG_SPG, law=R.LAW_COULOMB, op=R.OP_ASSEMBLED, tol=1e-8,
                    max_iters=5000)
    print(f"case: file -> SPG 1k: {s.iterations} it res {s.residual:.2e} merit1 {dl.merit1(s.r):.2e}", flush=True)
    assert s.converged and O.OracleProblem(prob).eval(s.r) <= 1.05e-8


def test_from_w_rejects_bad_input(R, ctx):
    n = 2
    W = np.eye(3 * n) * 2.0
    rp, ci, v = dense_to_csr(W)
    b, mu = np.zeros(3 * n), np.full(n, 0.5)
    R.DualFrictionProblem.from_w(ctx, n, rp, ci, v, b, mu)  # the valid case
    bad_col = ci.copy()
    bad_col[2] = 3 * n
    bad_rp = rp.copy()
    bad_rp[3] = bad_rp[2] - 1
    for args, tol in (((rp, bad_col, v), 0.0), ((bad_rp, ci, v), 0.0), ((rp, ci, v), -1.0)):
        with pytest.raises(R.SbError):
            R.DualFrictionProblem.from_w(ctx, n, *args, b, mu, prune_tol=tol)
    with pytest.raises(ValueError):
        R.DualFrictionProblem.from_w(ctx, n, rp[:-1], ci, v, b, mu)
    # every block pruned: W = 0 (each empty row holds one zero block in the tiles), and with
    # b = 0 the residual of r = 0 is 0
    dz = R.DualFrictionProblem.from_w(ctx, n, rp, ci, v * 1e-20, b, mu, prune_tol=1e-12)
    assert dz.w_lower()[0][-1] == 0
    assert dz.eval_with(np.zeros(3 * n)) == 0.0
