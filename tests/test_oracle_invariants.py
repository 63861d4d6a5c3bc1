"""The oracle against mathematics it must satisfy (independent numpy restatements,
closed forms, invariants, brute force). These pin the parts the reference's
known-answer tests leave open."""
import numpy as np
import pytest


def jordan_sqrt(w):
    """sqrt of w = (w0, wT) in the Jordan algebra of the second-order cone (spectral form)."""
    w0, wt = w[..., 0], w[..., 1:]
    n = np.linalg.norm(wt, axis=-1)
    safe = np.where(n > 0, n, 1.0)
    d = np.where((n > 0)[..., None], wt / safe[..., None], np.array([1.0, 0.0]))
    l1 = np.sqrt(np.maximum(w0 - n, 0.0))
    l2 = np.sqrt(w0 + n)
    out = np.empty_like(w)
    out[..., 0] = 0.5 * (l1 + l2)
    out[..., 1:] = 0.5 * (l2 - l1)[..., None] * d
    return out


def jordan_sq(a):
    out = np.empty_like(a)
    out[..., 0] = np.sum(a * a, axis=-1)
    out[..., 1:] = 2.0 * a[..., :1] * a[..., 1:]
    return out


def np_law_eval(mu, x, y, coulomb):
    """|FB_mu(x, y)|^2 per contact, from the Jordan-algebra definition."""
    x = np.atleast_2d(x).astype(float)
    y = np.atleast_2d(y).astype(float).copy()
    mu = np.broadcast_to(np.asarray(mu, float), x.shape[:1])
    if coulomb:
        y[:, 0] += mu * np.linalg.norm(y[:, 1:], axis=1)
    out = np.empty(x.shape[0])
    for i in range(x.shape[0]):
        m = mu[i]
        if m < 0:
            out[i] = np.sum(y[i] ** 2)  # disabled contact
            continue
        if abs(m) < np.finfo(float).eps:
            f0 = x[i, 0] + y[i, 0] - np.hypot(x[i, 0], y[i, 0])
            out[i] = f0 * f0 + x[i, 1] ** 2 + x[i, 2] ** 2
            continue
        xh = np.array([m * x[i, 0], x[i, 1], x[i, 2]])
        yh = np.array([y[i, 0], m * y[i, 1], m * y[i, 2]])
        fb = xh + yh - jordan_sqrt(jordan_sq(xh) + jordan_sq(yh))
        out[i] = fb @ fb
    return out


rng = np.random.default_rng(1234)


@pytest.mark.parametrize("coulomb", [False, True])
def test_fb_matches_jordan_algebra(O, coulomb):
    for _ in range(300):
        mu = rng.uniform(0.05, 1.5)
        x, y = rng.normal(size=3), rng.normal(size=3)
        assert O.law_eval(mu, x, y, coulomb) == pytest.approx(np_law_eval(mu, x, y, coulomb)[0], rel=1e-10,
                                                              abs=1e-14)


def test_fb_special_mu(O):
    x, y = rng.normal(size=3), rng.normal(size=3)
    # mu = 0: scalar FB on the normal part, x_T on the tangent part
    f0 = x[0] + y[0] - np.hypot(x[0], y[0])
    assert O.law_eval(0.0, x, y, False) == pytest.approx(f0 * f0 + x[1] ** 2 + x[2] ** 2, rel=1e-14)
    # mu < 0 (disabled): |y|^2
    assert O.law_eval(-1.0, x, y, True) == pytest.approx(y @ y, rel=1e-15)


@pytest.mark.parametrize("coulomb", [False, True])
def test_fb_zero_on_complementarity(O, coulomb):
    for _ in range(200):
        mu = rng.uniform(0.1, 1.5)
        th = rng.uniform(0, 2 * np.pi)
        a, c = rng.uniform(0.1, 3), rng.uniform(0.1, 3)
        x = a * np.array([1.0, mu * np.cos(th), mu * np.sin(th)])   # boundary of K_mu
        yd = c * np.array([mu, -np.cos(th), -np.sin(th)])          # boundary of the dual cone, x . y = 0
        if coulomb:
            # u with u + mu |u_T| e_N = yd
            y = yd.copy()
            y[0] -= mu * np.linalg.norm(yd[1:])
        else:
            y = yd
        assert O.law_eval(mu, x, y, coulomb) < 1e-20
        # and positive away from it
        assert O.law_eval(mu, x + np.array([0.3, 0, 0]), y, coulomb) > 1e-6


def test_projection_invariants(O):
    for _ in range(500):
        mu = rng.uniform(0.0, 2.0)
        x = rng.normal(size=3) * rng.uniform(0.1, 10)
        p = O.project(mu, x)
        # membership
        assert np.linalg.norm(p[1:]) <= mu * p[0] + 1e-12 * (1 + abs(p[0]))
        # idempotence
        assert np.allclose(O.project(mu, p), p, rtol=1e-12, atol=1e-14)
        # Moreau: x - p in the polar cone -K_{1/mu}, orthogonal to p
        q = x - p
        assert abs(p @ q) <= 1e-10 * (1 + np.linalg.norm(x) ** 2)
        assert mu * np.linalg.norm(q[1:]) <= -q[0] + 1e-10 * (1 + np.linalg.norm(x))
    assert np.all(O.project(-0.5, [1.0, 0.2, 0.1]) == 0)


def test_cov(O):
    y = np.array([0.3, -1.2, 0.5])
    assert np.allclose(O.cov(0.4, y, True), [0.4 * np.linalg.norm(y[1:]), 0, 0], rtol=1e-15)
    assert np.all(O.cov(0.4, y, False) == 0)


def test_roots_against_numpy(O):
    for _ in range(400):
        c = rng.normal(size=4) * rng.uniform(0.1, 10)
        ours = np.sort(O.real_roots_monic(c))
        ref = np.roots(np.concatenate([[1.0], c[::-1]]))
        ref = np.sort(ref[np.abs(ref.imag) < 1e-7].real)
        if len(ours) != len(ref):
            # near-double roots may legitimately flip between real and complex
            cc = np.roots(np.concatenate([[1.0], c[::-1]]))
            assert np.min(np.abs(cc.imag[np.abs(cc.imag) > 0])) < 1e-5
            continue
        assert np.allclose(ours, ref, rtol=1e-7, atol=1e-7)
    # distinct known roots (x-1)(x-2)(x+3)(x-0.5)
    poly = np.poly([1, 2, -3, 0.5])
    got = np.sort(O.real_roots_monic(poly[1:][::-1]))
    assert np.allclose(got, [-3, 0.5, 1, 2], rtol=1e-12)
    assert np.allclose(np.sort(O.real_roots_monic(poly[1:][::-1], O.STRICTLY_POSITIVE)), [0.5, 1, 2], rtol=1e-12)


def test_degenerate_roots(O):
    # leading coefficient vanishing: 0 x^4 + 0 x^3 + 1 x^2 + 0 x - 4 -> roots +-2
    assert np.allclose(np.sort(O.real_roots_degenerate([-4, 0, 1, 0, 0])), [-2, 2], rtol=1e-14)
    # all zero: one root 0 with AllRoots
    assert len(O.real_roots_degenerate([0, 0, 0, 0, 0])) == 1


def test_hessenberg_eigenvalues(O):
    for _ in range(200):
        H = np.triu(rng.normal(size=(4, 4)), -1)
        wr, wi = O.eig_hessenberg(H)
        ours = np.sort_complex(wr + 1j * wi)
        ref = np.sort_complex(np.linalg.eigvals(H))
        assert np.allclose(ours, ref, rtol=1e-9, atol=1e-9)


def test_lu_solve(O):
    for _ in range(200):
        A = rng.normal(size=(3, 3))
        b = rng.normal(size=3)
        x, rank = O.lu_solve(A, b)
        assert rank == 3
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-12)
    # rank-deficient: free unknowns are 0 and the consistent part is solved
    A = np.array([[1.0, 2, 0], [2, 4, 0], [0, 0, 3]])
    x, rank = O.lu_solve(A, np.array([1.0, 2, 3]))
    assert rank == 2 and np.allclose(A @ x, [1, 2, 3])


def test_fb_jacobian_finite_differences(O):
    for _ in range(100):
        mu = rng.uniform(0.1, 1.2)
        x, y = rng.normal(size=3), rng.normal(size=3)
        fb, dx, dy = O.fb_base_jacobian(mu, x, y)
        h = 1e-7
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            fx = (O.fb_base_jacobian(mu, x + e, y)[0] - O.fb_base_jacobian(mu, x - e, y)[0]) / (2 * h)
            fy = (O.fb_base_jacobian(mu, x, y + e)[0] - O.fb_base_jacobian(mu, x, y - e)[0]) / (2 * h)
            assert np.allclose(dx[:, k], fx, rtol=1e-5, atol=1e-6)
            assert np.allclose(dy[:, k], fy, rtol=1e-5, atol=1e-6)


def _spd(n=3):
    A = rng.normal(size=(n, n))
    return A @ A.T + 0.5 * np.eye(n)


@pytest.mark.parametrize("coulomb", [True, False])
def test_local_solver_reaches_fb_zero(O, coulomb):
    """Brute-force check of the analytic+Newton local solver: the returned point has
    a vanishing FB residual (independent numpy FB) for random one-contact problems."""
    nsolved = 0
    for _ in range(300):
        A = _spd()
        b = rng.normal(size=3)
        mu = rng.uniform(0.2, 1.0)
        res, x = O.local_solve(A, b, np.zeros(3), mu, coulomb=coulomb)
        if res < O.LOCAL_TOL:
            nsolved += 1
            assert np_law_eval(mu, x, A @ x + b, coulomb)[0] < 1e-10
    assert nsolved >= 270


def test_local_solver_closed_forms(O):
    A = _spd()
    # take-off: b_N >= 0 -> r = 0
    res, x = O.local_solve(A, np.array([0.5, 0.2, -0.1]), np.ones(3), 0.5)
    assert res == 0 and np.all(x == 0)
    # sticking: r = -A^{-1} b inside the cone
    b = -A @ np.array([1.0, 0.1, -0.05])
    res, x = O.local_solve(A, b, np.zeros(3), 0.8)
    assert res == 0 and np.allclose(x, [1.0, 0.1, -0.05], rtol=1e-10)
    # frictionless
    res, x = O.local_solve(A, np.array([-2.0, 1, 1]), np.zeros(3), 0.0)
    assert res == 0 and np.allclose(x, [2.0 / A[0, 0], 0, 0])


def test_problem_w_and_matrix_free_against_dense(G, O):
    prob = G.generate(nc=60, nb=30, seed=7)
    op = O.OracleProblem(prob)
    Wd = prob.dense_w()
    x = rng.normal(size=3 * prob.nc)
    assert np.allclose(op.apply_w(x, assembled=True), Wd @ x, rtol=1e-12, atol=1e-11)
    assert np.allclose(op.apply_w(x, assembled=False), Wd @ x, rtol=1e-12, atol=1e-11)
    d = op.diag()
    for i in range(prob.nc):
        assert np.allclose(d[i].reshape(3, 3), Wd[3 * i:3 * i + 3, 3 * i:3 * i + 3], rtol=1e-13, atol=1e-13)


def test_generator_structure(G):
    prob = G.generate(nc=500, nb=200, seed=3)
    assert prob.rowptr[-1] == 2 * prob.nc
    a, b = prob.colidx[0::2], prob.colidx[1::2]
    assert np.all(a < b)
    for i in range(0, prob.nc, 50):
        Ha = prob.hvals[2 * i].reshape(3, 6)
        E = Ha[:, :3]
        assert np.allclose(E @ E.T, np.eye(3), atol=1e-12) or np.allclose(E @ E.T, np.eye(3), atol=1e-12)
    m = prob.minv.reshape(-1, 6, 6)
    assert np.allclose(m, np.transpose(m, (0, 2, 1)))
    assert np.all(np.linalg.eigvalsh(m) > 0)
    assert np.all((prob.mu >= 0.3) & (prob.mu <= 0.8))


def test_eval_against_numpy(G, O):
    prob = G.generate(nc=200, nb=100, seed=9)
    op = O.OracleProblem(prob)
    r = rng.normal(size=3 * prob.nc)
    u = prob.dense_w() @ r + prob.b
    d = op.diag()
    s = np.maximum(1.0, (d[:, 0] + d[:, 4] + d[:, 8]) / 3)
    terms = np_law_eval(prob.mu, (s[:, None] * r.reshape(-1, 3)), u.reshape(-1, 3), True)
    assert op.eval(r, coulomb=True) == pytest.approx(terms.sum() / (prob.nc + 1), rel=1e-10)
    assert op.eval(r, coulomb=True, inf_norm=True) == pytest.approx(terms.max(), rel=1e-10)


@pytest.mark.parametrize("variant", ["PG_APGD", "PG_SPG"])
def test_pg_solution_checked_by_numpy(G, O, variant):
    """APGD/SPG on a 300-contact problem: the returned r has FB residual <= tol by an
    independent numpy evaluation, and lies in the friction cones up to the APGD
    extrapolation (the residual is evaluated at the extrapolated x, as in the reference)."""
    prob = G.generate(nc=300, nb=150, seed=21)
    op = O.OracleProblem(prob)
    res, r, info = op.pg_solve(np.zeros(3 * prob.nc), variant=getattr(O, variant), tol=1e-8, max_iters=5000,
                               coulomb=True)
    u = prob.dense_w() @ r + prob.b
    d = op.diag()
    s = np.maximum(1.0, (d[:, 0] + d[:, 4] + d[:, 8]) / 3)
    terms = np_law_eval(prob.mu, s[:, None] * r.reshape(-1, 3), u.reshape(-1, 3), True)
    assert terms.sum() / (prob.nc + 1) == pytest.approx(res, rel=1e-8)
    assert res < 1e-8
    # best residual history is what test_residual saw
    assert np.min(info["history"]) == pytest.approx(res)


def test_coloring_literal_semantics(G, O):
    """Coloring::compute as written: a color is rejected when std::lower_bound(color, j)
    != end for a lower neighbour j (bisect restatement on a small problem), and every
    color is conflict-free."""
    import bisect
    prob = G.generate(nc=150, nb=60, seed=5)
    op = O.OracleProblem(prob)
    rp, ci, _ = op.w_lower()
    colors_lists = []
    for i in range(prob.nc):
        placed = False
        for col in colors_lists:
            ok = True
            for p in range(rp[i], rp[i + 1]):
                j = ci[p]
                if j >= i:
                    continue
                if bisect.bisect_left(col, j) != len(col):
                    ok = False
                    break
            if ok:
                col.append(i)
                placed = True
                break
        if not placed:
            colors_lists.append([i])
    perm, colors = op.coloring()
    assert len(colors) - 1 == len(colors_lists)
    assert np.array_equal(perm, np.concatenate(colors_lists))
    # conflict-free: no lower W block inside a color
    color_of = np.empty(prob.nc, int)
    for c in range(len(colors) - 1):
        color_of[perm[colors[c]:colors[c + 1]]] = c
    for i in range(prob.nc):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            if j != i:
                assert color_of[i] != color_of[j]


def test_permutation_preserves_solution(G, O):
    prob = G.generate(nc=120, nb=60, seed=8)
    op = O.OracleProblem(prob)
    r = rng.normal(size=3 * prob.nc)
    e0 = op.eval(r, coulomb=True)
    perm = rng.permutation(prob.nc).astype(np.int32)
    op.permute(perm)
    rp = r.reshape(-1, 3)[perm].reshape(-1)
    assert op.eval(rp, coulomb=True) == pytest.approx(e0, rel=1e-12)


def test_gs_per_color_ndx_matches_sequential_on_one_color(G, O):
    """With a single color the per-color ndxRef snapshot and the reference's running
    maximum differ only in when later contacts see updates; both must converge."""
    prob = G.generate(nc=100, nb=80, seed=4)
    op = O.OracleProblem(prob)
    res1, r1, _ = op.gs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=2000)
    perm, colors = op.coloring()
    op.permute(perm)
    res2, r2, _ = op.gs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=2000, ndx_ref_per_color=True,
                              colors=colors)
    assert res1 < 1e-8 and res2 < 1e-8


def test_gs_parallel_colors_equal_one_thread(G, O):
    """The oracle sweeps a color's contacts on several threads when the colors hold no
    coupled pair (the reference's OpenMP loop over a color): the iterates and sweep
    count equal a one-thread run bit for bit. A single color holding
    coupled contacts keeps the sequential sweep, whatever the thread count (the residual
    evaluations are OpenMP sums whose order follows the thread count)."""
    prob = G.generate(nc=3000, nb=1500, seed=21)
    op = O.OracleProblem(prob)
    perm, colors = op.coloring()
    op.permute(perm)
    saved = O.threads()
    try:
        out = []
        for t in (1, 8):
            O.set_threads(t)
            out.append(op.gs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=400, ndx_ref_per_color=True,
                                   colors=colors, coulomb=False))
            out.append(op.gs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=50, ndx_ref_per_color=True,
                                   colors=np.array([0, prob.nc]), coulomb=False))
        (res1, r1, d1), (ress1, rs1, _), (res8, r8, d8), (ress8, rs8, _) = out
        # the iterates are identical; the residual's own OpenMP sum follows the thread count
        assert d1["iterations"] == d8["iterations"] and np.array_equal(r1, r8)
        assert np.allclose(d1["history"], d8["history"], rtol=1e-12, atol=0)
        assert np.array_equal(rs1, rs8) and ress1 == pytest.approx(ress8, rel=1e-12)
    finally:
        O.set_threads(saved)
