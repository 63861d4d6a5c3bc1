"""The oracle against the reference's own known-answer tests.

Each test restates an assertion of the reference test suite (bogus tests/,
file:line cited) on the oracle, so the CPU restatement is pinned before it is
used to check the GPU path.
"""
import numpy as np
import pytest


# ---------------------------------------------------------------- tests/Polynomial.cpp:16-38
def test_quadratic_roots(O):
    assert len(O.real_roots_monic([-1, 0])) == 2
    pos = O.real_roots_monic([-1, 0], O.STRICTLY_POSITIVE)
    assert len(pos) == 1 and pos[0] == pytest.approx(1.0, rel=4 * np.finfo(float).eps)
    neg = O.real_roots_monic([-1, 0], O.STRICTLY_NEGATIVE)
    assert len(neg) == 1 and neg[0] == pytest.approx(-1.0, rel=4 * np.finfo(float).eps)
    assert len(O.real_roots_monic([1, 0])) == 0
    r = O.real_roots_monic([1, -2])
    assert len(r) == 2 and np.allclose(r, [1.0, 1.0], rtol=4 * np.finfo(float).eps, atol=0)


# ---------------------------------------------------------------- tests/Polynomial.cpp:40-94
SOQP_CASES = [
    (np.array([[0.01344, -9.421e-07, 0.001486], [-9.421e-07, 0.1061, 0.0001733], [0.001486, 0.0001733, 0.001442]]),
     np.array([-0.1458, -0.2484, -0.1515]), 0.6),
    (np.diag([0.4, 0.9, 0.9]), np.array([0, 0, 0.542629]), 0.8),
    (np.diag([0.4, 1.4, 1.4]), np.array([-0.128762, 3.21747e-16, 14.5114]), 0.8),
]


@pytest.mark.parametrize("W,b,mu", SOQP_CASES)
def test_soc_quartic_local_solver(O, W, b, mu):
    res, r = O.local_solve(W, b, np.zeros(3), mu, coulomb=False, strategy=O.PURE_ENUMERATIVE, tol=1e-12)
    u = W @ r + b
    assert res < 1e-24
    assert r[0] > 0 and u[0] > 0
    assert np.float32(u[0]) == pytest.approx(np.float32(np.linalg.norm(u[1:]) * mu), rel=4 * 1.2e-7)
    assert u @ r < 1e-12


# ---------------------------------------------------------------- tests/ProjectedGradient.cpp:85-108
def test_projection_known_answers(O):
    assert np.array_equal(O.project(0.5, [2, 0, 1]), [2, 0, 1])
    x = O.project(3, [1, 3, 0])
    assert np.array_equal(x, [1, 3, 0])
    assert np.allclose(O.project(0.5, x), [2, 1, 0], rtol=1e-12)
    assert np.allclose(O.project(0.5, [-1, 7, 0]), [2, 1, 0], rtol=1e-12)
    assert np.all(O.project(3, [-1, .3, 0]) == 0)


# ---------------------------------------------------------------- tests/SmallFrictionPb.hpp + GaussSeidel.cpp:19-65
@pytest.fixture(scope="module")
def small(G, O):
    prob, sol = G.small_friction_pb()
    return prob, sol, O.OracleProblem(prob)


def _is_approx(a, b, prec):
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


@pytest.mark.parametrize("strategy", ["HYBRID", "PURE_NEWTON", "PURE_ENUMERATIVE", "REV_HYBRID"])
def test_small_gs_coulomb_known_solution(O, small, strategy):
    prob, sol, op = small
    res, x, _ = op.gs_solve(np.ones(6), coulomb=True, strategy=getattr(O, strategy))
    assert res < 1e-8
    assert _is_approx(sol, x, 1e-4)


def test_small_gs_soc(O, small):
    prob, sol, op = small
    res, x, _ = op.gs_solve(np.ones(6), coulomb=False, tol=1e-8)
    assert res < 1e-8


# ---------------------------------------------------------------- tests/ProductGaussSeidel.cpp:16-62
def test_small_pgs_known_solution(O, small):
    """Product_GS_D / Product_GS_D_precomp: D = the inverse mass, x0 = ones."""
    prob, sol, op = small
    res, x, _ = op.pgs_solve(np.ones(6), coulomb=True, tol=1e-8)
    assert res < 1e-8
    assert _is_approx(sol, x, 1e-4)


def test_small_pgs_trivial_diagonal(G, O, small):
    """Product_GS: D = identity, so the operator is H H^T; the residual of the
    result, re-evaluated with y = H H^T x + b (the oracle's W of this problem), is
    below 1e-8."""
    import dataclasses
    prob, sol, _ = small
    eye = dataclasses.replace(prob, minv=np.tile(np.eye(6).reshape(36), (prob.nb, 1)))
    op = O.OracleProblem(eye)
    res, x, _ = op.pgs_solve(np.ones(6), coulomb=True, tol=1e-8)
    assert res < 1e-8
    H = eye.dense_h()
    assert np.allclose(eye.dense_w(), H @ H.T)
    assert op.eval(x, coulomb=True) < 1e-8


@pytest.mark.parametrize("coulomb", [True, False])
def test_pgs_equals_gs_on_w(G, O, coulomb):
    """The product form (lb from H (D H^T x), no W) and the assembled-W Gauss-Seidel
    are the same iteration: with the same coloring they take the same number of
    sweeps and their residual histories agree to rounding."""
    prob = G.workload("small-1k")
    op = O.OracleProblem(prob)
    perm, colors = op.coloring()
    op.permute(perm)
    r1, x1, d1 = op.gs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=1000, colors=colors, coulomb=coulomb)
    r2, x2, d2 = op.pgs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=1000, colors=colors, coulomb=coulomb)
    assert d1["iterations"] == d2["iterations"]
    assert np.allclose(d1["history"], d2["history"], rtol=1e-9, atol=0)
    assert np.linalg.norm(x1 - x2) <= 1e-12 * np.linalg.norm(x1)


# ---------------------------------------------------------------- Cadoux.hpp:44-155 (no reference test)
@pytest.mark.parametrize("use_gs", [True, False])
def test_small_cadoux_reaches_coulomb_solution(O, small, use_gs):
    """The fixed point of the SOC solves with the De Saxce-corrected affine term is the
    Coulomb solution: on the two-contact fixture it reaches the known Coulomb `sol`
    (tests/GaussSeidel.cpp:38-64) to 1e-4, and the returned residual is the oracle's
    independent Coulomb evaluation of the result."""
    prob, sol, op = small
    res, x, d = op.cadoux_solve(np.zeros(6), use_gs=use_gs, tol=1e-8, iterations=50)
    assert res < 1e-8 and d["iterations"] < 50
    assert _is_approx(sol, x, 1e-4)
    assert op.eval(x, coulomb=True) == pytest.approx(res, rel=1e-12, abs=1e-300)
    assert np.all(np.diff(d["history"][:3]) < 0)


@pytest.mark.parametrize("use_gs", [True, False])
def test_small_cadoux_velocity_form(O, small, use_gs):
    """solveCadouxVel: the problem read as r = W u + b. The returned residual equals an
    independent numpy evaluation from u: u* = u + (|u_T| / mu, 0, 0), r = W u + b
    (dense W), SOC Fischer-Burmeister of (s u*, r) averaged over n + 1."""
    prob, sol, op = small
    res, u, d = op.cadoux_solve(np.zeros(6), use_gs=use_gs, tol=1e-8, iterations=100, velocity=True)
    assert res < 1e-8
    W = prob.dense_w()
    ustar = u.copy()
    for i in range(2):
        ustar[3 * i] += np.hypot(u[3 * i + 1], u[3 * i + 2]) / prob.mu[i]
    r = W @ u + prob.b
    scal = [max(1.0, np.trace(W[3 * i:3 * i + 3, 3 * i:3 * i + 3]) / 3) for i in range(2)]
    acc = 0.0
    for i in range(2):
        fb = O.fb(prob.mu[i], scal[i] * ustar[3 * i:3 * i + 3], r[3 * i:3 * i + 3], coulomb=False)
        acc += float(np.dot(fb, fb))
    assert acc / 3 == pytest.approx(res, rel=1e-9, abs=1e-20)


# ---------------------------------------------------------------- tests/ProjectedGradient.cpp:15-83
@pytest.mark.parametrize("variant", ["PG_APGD", "PG_DESCENT", "PG_STANDARD", "PG_CONJUGATED", "PG_SPG"])
def test_small_pg_soc(O, small, variant):
    prob, sol, op = small
    res, x, _ = op.pg_solve(np.ones(6), variant=getattr(O, variant), tol=1e-8, coulomb=False)
    assert res < 1e-8


def test_small_pg_matrix_free_product(O, small):
    prob, sol, op = small
    res, x, _ = op.pg_solve(np.ones(6), tol=1e-8, coulomb=False, assembled=False)
    assert res < 1e-8


def test_small_w_dense(G, small):
    prob, sol, op = small
    rp, ci, v = op.w_lower()
    W = np.zeros((6, 6))
    for i in range(2):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            W[3 * i:3 * i + 3, 3 * j:3 * j + 3] = v[p].reshape(3, 3)
            W[3 * j:3 * j + 3, 3 * i:3 * i + 3] = v[p].reshape(3, 3).T
    assert np.allclose(W, prob.dense_w(), rtol=1e-14, atol=1e-12)


# ---------------------------------------------------------------- tests/BlockMV.cpp:14-125 (3x4 blocks, 5x2)
def _small_sbm():
    rowptr = [0, 1, 1, 1, 3, 3]
    colidx = [1, 0, 1]
    ones = np.ones((3, 4))
    vals = np.stack([ones, 2 * ones, 3 * ones]).reshape(3, 12)
    return rowptr, colidx, vals


def test_blockmv_known_vectors(O):
    rowptr, colidx, vals = _small_sbm()
    e1 = np.array([4, 4, 4, 0, 0, 0, 0, 0, 0, 20, 20, 20, 0, 0, 0], float)
    e2 = np.array([120, 120, 120, 120, 192, 192, 192, 192], float)
    e3 = np.array([768, 768, 768, 0, 0, 0, 0, 0, 0, 3264, 3264, 3264, 0, 0, 0], float)
    res = O.bsr_multiply(5, 2, 3, 4, rowptr, colidx, vals, np.ones(8))
    assert np.array_equal(res, e1)
    rhs = O.bsr_multiply(5, 2, 3, 4, rowptr, colidx, vals, res, transpose=True)
    assert np.array_equal(rhs, e2)
    assert np.array_equal(O.bsr_multiply(5, 2, 3, 4, rowptr, colidx, vals, rhs), e3)
    # alpha/beta semantics: res = 2 * (.5 * A x)
    half = O.bsr_multiply(5, 2, 3, 4, rowptr, colidx, vals, np.ones(8), alpha=0.5)
    assert np.array_equal(2 * half, e1)
    acc = O.bsr_multiply(5, 2, 3, 4, rowptr, colidx, vals, np.ones(8), res=e1, alpha=1.0, beta=1.0)
    assert np.array_equal(acc, 2 * e1)


# ---------------------------------------------------------------- tests/BlockMV.cpp:192-245 (symmetric)
def _sym_sbm():
    sec = np.array([[2, 0, 0], [2, 2, 0], [2, 2, 2]], float)
    vals = np.stack([np.ones((3, 3)), sec, 3 * np.ones((3, 3))]).reshape(3, 9)
    return [0, 1, 2, 3], [0, 0, 2], vals


def test_blockmv_symmetric_and_split_row(O):
    rowptr, colidx, vals = _sym_sbm()
    e1 = np.array([9, 7, 5, 2, 4, 6, 9, 9, 9], float)
    e2 = np.array([6, 4, 2, 2, 4, 6, 0, 0, 0], float)
    assert np.array_equal(O.bsr_multiply(3, 3, 3, 3, rowptr, colidx, vals, np.ones(9), sym=True), e1)
    assert np.array_equal(O.bsr_multiply(3, 3, 3, 3, rowptr, colidx, vals, np.ones(9), sym=True, transpose=True), e1)
    res = np.concatenate([O.bsr_split_row(3, rowptr, colidx, vals, k, np.ones(9)) for k in range(3)])
    assert np.array_equal(res, e2)


# ---------------------------------------------------------------- tests/SparseBlock.cpp:16-49 (W = A A^T diagonal)
def test_product_diagonal_block_136(G, O):
    """sbm * sbm^T has diagonal block 2 equal to Constant(136); embedded as H (3x4 padded
    to 3x6) with identity inverse mass, W = H H^T."""
    sample = np.array([[1, 2, 3, 4], [3, 2, 1, 0], [1, 2, 3, 4]], float)
    blocks = {(0, 1): np.ones((3, 4)), (1, 0): sample, (2, 0): 3 * np.ones((3, 4)), (2, 1): 5 * np.ones((3, 4))}
    rowptr, colidx, hv = [0], [], []
    for i in range(4):
        for j in range(2):
            if (i, j) in blocks:
                pad = np.zeros((3, 6))
                pad[:, :4] = blocks[(i, j)]
                hv.append(pad.reshape(-1))
                colidx.append(j)
        rowptr.append(len(colidx))
    prob = G.FrictionProblem(nc=4, nb=2, rowptr=np.array(rowptr, np.int32), colidx=np.array(colidx, np.int32),
                             hvals=np.array(hv), minv=np.tile(np.eye(6).reshape(-1), (2, 1)), b=np.zeros(12),
                             mu=np.full(4, 0.5))
    op = O.OracleProblem(prob)
    d = op.diag()
    assert np.array_equal(d[2], np.full(9, 136.0))
    rp, ci, v = op.w_lower()
    W = np.zeros((12, 12))
    for i in range(4):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            W[3 * i:3 * i + 3, 3 * j:3 * j + 3] = v[p].reshape(3, 3)
            W[3 * j:3 * j + 3, 3 * i:3 * i + 3] = v[p].reshape(3, 3).T
    Hd = prob.dense_h()
    assert np.array_equal(W, Hd @ Hd.T)
