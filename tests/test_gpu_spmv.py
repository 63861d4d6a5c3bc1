"""Block SpMV on the GPU (through the C ABI) against the oracle and the reference's
known answers. Integer-valued inputs make every sum exact in FP64, so those
comparisons are bit-exact; random inputs are compared at 1e-14 relative
(different but fixed summation order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(99)


def test_blockmv_known_vectors_gpu(R, ctx):
    print("case: BlockMV SmallSBM 3x4 blocks", flush=True)
    rowptr, colidx = [0, 1, 1, 1, 3, 3], [1, 0, 1]
    ones = np.ones((3, 4))
    vals = np.stack([ones, 2 * ones, 3 * ones]).reshape(3, 12)
    A = R.SparseBlockMatrix(ctx, 5, 2, 3, 4, rowptr, colidx, vals)
    e1 = np.array([4, 4, 4, 0, 0, 0, 0, 0, 0, 20, 20, 20, 0, 0, 0], float)
    e2 = np.array([120, 120, 120, 120, 192, 192, 192, 192], float)
    e3 = np.array([768, 768, 768, 0, 0, 0, 0, 0, 0, 3264, 3264, 3264, 0, 0, 0], float)
    res = A.multiply(np.ones(8))
    assert np.array_equal(res, e1)
    rhs = A.multiply(res, transpose=True)
    assert np.array_equal(rhs, e2)
    assert np.array_equal(A.multiply(rhs), e3)
    assert np.array_equal(2 * A.multiply(np.ones(8), alpha=0.5), e1)
    assert np.array_equal(A.multiply(np.ones(8), beta=1.0, res=e1), 2 * e1)


def test_blockmv_symmetric_gpu(R, ctx):
    print("case: BlockMV symmetric 3x3", flush=True)
    sec = np.array([[2, 0, 0], [2, 2, 0], [2, 2, 2]], float)
    vals = np.stack([np.ones((3, 3)), sec, 3 * np.ones((3, 3))]).reshape(3, 9)
    A = R.SparseBlockMatrix(ctx, 3, 3, 3, 3, [0, 1, 2, 3], [0, 0, 2], vals, symmetric=True)
    e1 = np.array([9, 7, 5, 2, 4, 6, 9, 9, 9], float)
    e2 = np.array([6, 4, 2, 2, 4, 6, 0, 0, 0], float)
    assert np.array_equal(A.multiply(np.ones(9)), e1)
    assert np.array_equal(A.multiply(np.ones(9), transpose=True), e1)
    assert np.array_equal(np.concatenate([A.split_row_multiply(k, np.ones(9)) for k in range(3)]), e2)


def test_generic_bsr_random_against_oracle(R, O, ctx):
    print("case: generic BSR 6x3 random, with a long row", flush=True)
    nbr, nbc, br, bc = 40, 30, 6, 3
    rows = [np.sort(rng.choice(nbc, size=int(rng.integers(0, 6)), replace=False)) for _ in range(nbr)]
    rows[7] = np.arange(nbc)  # a full row
    rowptr = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int32)
    colidx = np.concatenate(rows).astype(np.int32)
    vals = rng.integers(-5, 6, size=(len(colidx), br * bc)).astype(float)
    A = R.SparseBlockMatrix(ctx, nbr, nbc, br, bc, rowptr, colidx, vals)
    x = rng.integers(-3, 4, size=nbc * bc).astype(float)
    ref = O.bsr_multiply(nbr, nbc, br, bc, rowptr, colidx, vals, x)
    assert np.array_equal(A.multiply(x), ref)
    xt = rng.integers(-3, 4, size=nbr * br).astype(float)
    reft = O.bsr_multiply(nbr, nbc, br, bc, rowptr, colidx, vals, xt, transpose=True)
    assert np.array_equal(A.multiply(xt, transpose=True), reft)
    # a row of 200 blocks (3 long tiles + carry fix-up)
    nbc2 = 200
    rowptr2 = np.array([0, 200, 203], np.int32)
    colidx2 = np.concatenate([np.arange(200), [1, 5, 9]]).astype(np.int32)
    vals2 = rng.integers(-4, 5, size=(203, 9)).astype(float)
    A2 = R.SparseBlockMatrix(ctx, 2, nbc2, 3, 3, rowptr2, colidx2, vals2)
    x2 = rng.integers(-3, 4, size=nbc2 * 3).astype(float)
    assert np.array_equal(A2.multiply(x2), O.bsr_multiply(2, nbc2, 3, 3, rowptr2, colidx2, vals2, x2))
    y2 = rng.normal(size=6)
    assert np.allclose(A2.multiply(y2, transpose=True),
                       O.bsr_multiply(2, nbc2, 3, 3, rowptr2, colidx2, vals2, y2, transpose=True), rtol=1e-14)


def _int_problem(G, nc, nb, seed):
    """Generator structure with small integer H / Minv entries: all products exact in FP64."""
    p = G.generate(nc=nc, nb=nb, seed=seed)
    r = np.random.default_rng(seed)
    p.hvals = r.integers(-3, 4, size=p.hvals.shape).astype(float)
    m = r.integers(-2, 3, size=(nb, 6, 6)).astype(float)
    p.minv = (m + np.transpose(m, (0, 2, 1))).reshape(nb, 36)
    return p


def test_friction_operators_bit_exact_on_integers(R, O, G, ctx):
    print("case: W x / matrix-free on integer data (nc=3000)", flush=True)
    p = _int_problem(G, 3000, 1200, 5)
    dp = R.DualFrictionProblem(ctx, p)
    op = O.OracleProblem(p)
    x = rng.integers(-4, 5, size=3 * p.nc).astype(float)
    ref = op.apply_w(x, assembled=True)
    assert np.array_equal(dp.apply_w(x, R.OP_ASSEMBLED), ref)
    assert np.array_equal(dp.apply_w(x, R.OP_MATRIX_FREE), ref)
    # the assembled W itself
    rp, ci, v = dp.w_lower()
    orp, oci, ov = op.w_lower()
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(v, ov)


def test_friction_operators_ragged_contact_rows(R, O, G, ctx):
    """Contacts with one body (every fifth row keeps only its first block) take the
    tiled H kernel; rows of two blocks take the one-lane-per-contact kernel. Both
    bit-exact on integer data."""
    print("case: W x / matrix-free on integer data with one-body contacts (nc=3000)", flush=True)
    p = _int_problem(G, 3000, 1200, 6)
    keep = np.ones(2 * p.nc, bool)
    keep[2 * np.arange(0, p.nc, 5) + 1] = False
    p.colidx = np.ascontiguousarray(p.colidx[keep])
    p.hvals = np.ascontiguousarray(p.hvals[keep])
    p.rowptr = np.concatenate([[0], np.cumsum(keep.reshape(-1, 2).sum(axis=1))]).astype(np.int32)
    dp = R.DualFrictionProblem(ctx, p)
    op = O.OracleProblem(p)
    x = rng.integers(-4, 5, size=3 * p.nc).astype(float)
    ref = op.apply_w(x, assembled=True)
    assert np.array_equal(dp.apply_w(x, R.OP_ASSEMBLED), ref)
    assert np.array_equal(dp.apply_w(x, R.OP_MATRIX_FREE), ref)


def test_friction_operators_random(R, O, G, ctx):
    print("case: W x / matrix-free random (nc=20000)", flush=True)
    p = G.generate(nc=20000, nb=8000, seed=12)
    dp = R.DualFrictionProblem(ctx, p)
    op = O.OracleProblem(p)
    x = rng.normal(size=3 * p.nc)
    ref = op.apply_w(x, assembled=True)
    scale = np.abs(ref).max()
    assert np.abs(dp.apply_w(x, R.OP_ASSEMBLED) - ref).max() <= 1e-13 * scale
    assert np.abs(dp.apply_w(x, R.OP_MATRIX_FREE) - ref).max() <= 1e-13 * scale


def test_hub_bodies_matrix_free(R, O, G, ctx):
    print("case: Zipf hub bodies (long H^T rows)", flush=True)
    p = G.generate(nc=20000, nb=3000, seed=3, zipf=1.2, mu_range=(0.1, 1.0))
    deg = np.bincount(p.colidx, minlength=p.nb)
    assert deg.max() > 1000  # several long rows with carries
    dp = R.DualFrictionProblem(ctx, p, assemble=False)
    op = O.OracleProblem(p, assemble=False)
    x = rng.normal(size=3 * p.nc)
    ref = op.apply_w(x, assembled=False)
    assert np.abs(dp.apply_w(x, R.OP_MATRIX_FREE) - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("law", ["LAW_COULOMB", "LAW_SOC"])
def test_fb_residual_against_oracle(R, O, G, ctx, law):
    print(f"case: FB residual {law}", flush=True)
    p = G.generate(nc=5000, nb=2000, seed=31)
    dp = R.DualFrictionProblem(ctx, p)
    op = O.OracleProblem(p)
    r = rng.normal(size=3 * p.nc)
    coul = law == "LAW_COULOMB"
    for inf in (False, True):
        g = dp.eval_with(r, law=getattr(R, law), inf_norm=inf)
        o = op.eval(r, coulomb=coul, inf_norm=inf)
        assert g == pytest.approx(o, rel=1e-12)
    # matrix-free, reference scaling (all ones) vs diagonal scaling
    g1 = dp.eval_with(r, law=getattr(R, law), op=R.OP_MATRIX_FREE)
    assert g1 == pytest.approx(op.eval(r, coulomb=coul, diag_scaling=False), rel=1e-12)
    g2 = dp.eval_with(r, law=getattr(R, law), op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL)
    assert g2 == pytest.approx(op.eval(r, coulomb=coul, diag_scaling=True), rel=1e-12)
