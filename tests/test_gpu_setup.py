"""Problem set-up on the GPU (sb_problem_update_values, sb_mass_inverse) against the
host set-up of a freshly created problem with the same values.

The reference recomputes W = H Minv H^T, the diagonal blocks and the scalings on the
host at every computeFrom / setMatrix (FrictionProblem.cpp:30-40,
SparseMatrixMatrixProduct.impl.hpp:295-349) and inverts the masses with per-body LU
(PrimalFrictionProblem.cpp:28-38). Here a problem keeps its pattern and the values
are refreshed by kernels; the bar: every product, residual and solve of the updated
problem equals the freshly created problem's (products to 1e-14 relative: the same
sums, possibly contracted differently; solves to the same iteration count where the
arithmetic is the same, else at tolerance).
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _perturbed(G, prob, seed):
    """Same pattern, new values: H blocks, Minv (symmetric), b and mu perturbed."""
    rng = np.random.default_rng(seed)
    scale = 1.0 + 0.1 * rng.random(prob.nb)  # keeps every Minv block symmetric positive definite
    return dataclasses.replace(
        prob,
        hvals=prob.hvals * (1.0 + 0.1 * rng.standard_normal(prob.hvals.shape)),
        minv=(prob.minv.reshape(prob.nb, 36) * scale[:, None]).reshape(prob.minv.shape),
        b=prob.b + 0.05 * rng.standard_normal(prob.b.shape),
        mu=np.clip(prob.mu + 0.05 * rng.standard_normal(prob.mu.shape), 0.05, 1.5))


@pytest.mark.parametrize("on_device", [False, True])
def test_update_values_matches_fresh_problem(R, G, ctx, on_device):
    print(f"case: device set-up, on_device={on_device}", flush=True)
    p0 = G.workload("small-1k")
    p1 = _perturbed(G, p0, 11)
    dp = R.DualFrictionProblem(ctx, p0, assemble=True)
    fresh = R.DualFrictionProblem(ctx, p1, assemble=True)
    kw = dict(variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-8,
              max_iters=5000)
    # a matrix-free solve before the update factors the operator tiles of the old values
    assert dp.solve_pg(np.zeros(3 * p0.nc), **kw).converged
    if on_device:
        arrs = []
        for a in (p1.hvals, p1.minv, p1.b, p1.mu):
            d = R.DeviceArray(ctx, np.asarray(a).size)
            d.upload(np.asarray(a, dtype=np.float64).reshape(-1))
            arrs.append(d)
        dp.update_values(*arrs)
    else:
        dp.update_values(p1.hvals, p1.minv, p1.b, p1.mu)
    x = np.random.default_rng(5).standard_normal(3 * p0.nc)
    for op in (R.OP_ASSEMBLED, R.OP_MATRIX_FREE):
        a, b = dp.apply_w(x, op=op), fresh.apply_w(x, op=op)
        assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(b), op
        for law in (R.LAW_COULOMB, R.LAW_SOC):
            for sc in (R.SCALING_REFERENCE, R.SCALING_DIAGONAL):
                ra = dp.eval_with(x, law=law, op=op, scaling=sc)
                rb = fresh.eval_with(x, law=law, op=op, scaling=sc)
                assert ra == pytest.approx(rb, rel=1e-12), (op, law, sc)
    # the assembled W read back (host copy refreshed from the device values)
    wa, wb = dp.w_lower(), fresh.w_lower()
    assert np.array_equal(wa[0], wb[0]) and np.array_equal(wa[1], wb[1])
    assert np.allclose(wa[2], wb[2], rtol=1e-13, atol=1e-13 * np.abs(wb[2]).max())
    # solves: APGD (device data only) and GS (local matrices from the refreshed diagonal)
    s1, s2 = dp.solve_pg(np.zeros(3 * p0.nc), **kw), fresh.solve_pg(np.zeros(3 * p0.nc), **kw)
    assert s1.converged and s2.converged
    # solved with the refactored tiles of the new values: the fresh problem agrees
    assert fresh.eval_with(s1.r, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL) <= 1.05e-8
    assert abs(s1.iterations - s2.iterations) <= max(5, s2.iterations // 4)
    g1 = dp.solve_gs(np.zeros(3 * p0.nc), law=R.LAW_SOC, tol=1e-8, max_iters=2000)
    g2 = fresh.solve_gs(np.zeros(3 * p0.nc), law=R.LAW_SOC, tol=1e-8, max_iters=2000)
    assert g1.converged and g2.converged and abs(g1.iterations - g2.iterations) <= 25


def test_update_values_then_permutation_and_pgs(R, G, O, ctx):
    """After a device update, the host paths (coloring + permutation, the matrix-free
    GS's per-block products) see the new values."""
    print("case: device set-up, then permutation and product GS", flush=True)
    p0 = G.workload("small-1k")
    p1 = _perturbed(G, p0, 12)
    dp = R.DualFrictionProblem(ctx, p0, assemble=False)
    dp.update_values(p1.hvals, p1.minv, p1.b, p1.mu)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    g = dp.solve_pgs(np.zeros(3 * p0.nc), law=R.LAW_SOC, tol=1e-8, max_iters=2000, colors=colors)
    op = O.OracleProblem(p1, assemble=False)
    op.permute(perm.astype(np.int32))
    ores, ox, od = op.pgs_solve(np.zeros(3 * p0.nc), tol=1e-8, max_iters=2000, coulomb=False, colors=colors,
                                ndx_ref_per_color=True)
    assert g.converged and g.iterations == od["iterations"]
    assert np.linalg.norm(g.r - ox) <= 1e-6 * np.linalg.norm(ox)


def test_partial_updates(R, G, ctx):
    """b / mu alone leave the matrices; Minv alone refreshes W and the scalings."""
    p0 = G.workload("small-1k")
    p1 = _perturbed(G, p0, 13)
    x = np.random.default_rng(6).standard_normal(3 * p0.nc)
    dp = R.DualFrictionProblem(ctx, p0, assemble=True)
    dp.update_values(b=p1.b, mu=p1.mu)
    ref = R.DualFrictionProblem(ctx, dataclasses.replace(p0, b=p1.b, mu=p1.mu), assemble=True)
    assert np.array_equal(dp.apply_w(x), ref.apply_w(x))
    assert dp.eval_with(x, law=R.LAW_COULOMB) == pytest.approx(ref.eval_with(x, law=R.LAW_COULOMB), rel=1e-14)
    dp.update_values(minv=p1.minv)
    ref = R.DualFrictionProblem(ctx, dataclasses.replace(p0, b=p1.b, mu=p1.mu, minv=p1.minv), assemble=True)
    for op in (R.OP_ASSEMBLED, R.OP_MATRIX_FREE):
        a, b = dp.apply_w(x, op=op), ref.apply_w(x, op=op)
        assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(b)


def test_mass_inverse(R, ctx):
    rng = np.random.default_rng(7)
    nb = 5000
    A = rng.standard_normal((nb, 6, 6))
    M = A @ A.transpose(0, 2, 1) + 6.0 * np.eye(6)  # symmetric positive definite mass blocks
    Minv = R.mass_inverse(ctx, M.reshape(nb, 36)).reshape(nb, 6, 6)
    ref = np.linalg.inv(M)
    assert np.abs(Minv - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(M @ Minv - np.eye(6)).max() <= 1e-12
    # a zero leading pivot: the elimination must swap rows
    Mz = 2.0 * np.eye(6)[[1, 0, 2, 3, 4, 5]]
    assert np.array_equal(R.mass_inverse(ctx, Mz.reshape(1, 36)).reshape(6, 6), Mz.T / 4.0)
    with pytest.raises(RuntimeError):
        R.mass_inverse(ctx, np.zeros((1, 36)))


def _ragged_problem(G, seed=3):
    """Contacts with no body, one body, three bodies, the same body twice, and one
    contact with 100 blocks (a long row of H); the last 100 bodies touch nothing
    (empty rows of H^T)."""
    base = G.generate(nc=3000, nb=1500, seed=seed)
    rng = np.random.default_rng(seed)
    nbu = base.nb - 100
    rows = []
    for i in range(base.nc):
        k = i % 7
        if k == 0:
            cols = []
        elif k == 1:
            cols = [int(rng.integers(nbu))]
        elif k == 2:
            cols = sorted(rng.choice(nbu, size=3, replace=False).tolist())
        elif k == 3:
            a = int(rng.integers(nbu))
            cols = [a, a]
        else:
            cols = sorted(rng.choice(nbu, size=2, replace=False).tolist())
        rows.append(cols)
    rows[5] = list(range(0, 200, 2))
    rowptr = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int32)
    colidx = np.concatenate([np.asarray(c, np.int32) for c in rows]).astype(np.int32)
    hvals = rng.standard_normal((len(colidx), 18))
    return dataclasses.replace(base, rowptr=rowptr, colidx=colidx, hvals=hvals)


def _same_layout(a, b, which):
    la, lb = a.layout(which), b.layout(which)
    for k in la:
        assert np.array_equal(la[k], lb[k]), (which, k)
    return la


def _set_up_cases(G):
    return {
        "uniform-5k": (G.generate(nc=5000, nb=2000, seed=21), True),
        "zipf-20k": (G.generate(nc=20000, nb=3000, seed=3, zipf=1.2, mu_range=(0.1, 1.0)), True),
        "ragged": (_ragged_problem(G), True),
    }


@pytest.mark.parametrize("case", ["uniform-5k", "zipf-20k", "ragged"])
def test_device_setup_equals_host_setup(R, G, ctx, case):
    """The set-up on the device (default) builds every array the host set-up builds:
    tiles of H, H^T and the two-triangle W (indices, long rows, source blocks and
    values, bit for bit), W's lower triangle, the diagonal blocks and scalings (so
    every product and residual is bit-identical), and the colorings."""
    print(f"case: device set-up vs host set-up, {case}", flush=True)
    p, assemble = _set_up_cases(G)[case]
    host = R.DualFrictionProblem(ctx, p, assemble=assemble, host_setup=True)
    dev = R.DualFrictionProblem(ctx, p, assemble=assemble)
    lh = _same_layout(host, dev, 0)
    lt = _same_layout(host, dev, 1)
    if assemble:
        _same_layout(host, dev, 2)
    if case == "ragged":
        assert lh["nlong"] == 1 and (lh["src"] < 0).sum() == (np.diff(p.rowptr) == 0).sum()
        untouched = (np.bincount(p.colidx, minlength=p.nb) == 0).sum()
        assert untouched >= 100 and (lt["src"] < 0).sum() == untouched
    if case == "zipf-20k":
        assert lt["nlong"] > 0
    assert host.sizes() == dev.sizes()
    x = np.random.default_rng(4).standard_normal(3 * p.nc)
    for op in ((R.OP_ASSEMBLED, R.OP_MATRIX_FREE) if assemble else (R.OP_MATRIX_FREE,)):
        assert np.array_equal(host.apply_w(x, op=op), dev.apply_w(x, op=op)), op
        for sc in (R.SCALING_REFERENCE, R.SCALING_DIAGONAL):
            assert host.eval_with(x, law=R.LAW_COULOMB, op=op, scaling=sc) == \
                dev.eval_with(x, law=R.LAW_COULOMB, op=op, scaling=sc)
    if assemble:
        wh, wd = host.w_lower(), dev.w_lower()
        for a, b in zip(wh, wd):
            assert np.array_equal(a, b)
    for mode in (R.COLORING_REFERENCE, R.COLORING_GREEDY):
        ph, ch = host.coloring(mode | R.COLORING_HOST)
        for prob in (host, dev):
            pd, cd = prob.coloring(mode)
            assert np.array_equal(ph, pd) and np.array_equal(ch, cd)


def test_device_setup_from_device_arrays(R, G, ctx):
    """sb_problem_create_device: the arrays already in HBM; same layout and products as
    the host set-up, and invalid patterns are refused."""
    print("case: set-up from device arrays (uniform-5k, ragged)", flush=True)
    for p in (G.generate(nc=5000, nb=2000, seed=22), _ragged_problem(G, seed=4)):
        d = {k: R.DeviceArray(ctx, np.asarray(v).size, np.int32 if k in ("rowptr", "colidx") else np.float64)
             .upload(np.asarray(v).reshape(-1))
             for k, v in (("rowptr", p.rowptr), ("colidx", p.colidx), ("hvals", p.hvals), ("minv", p.minv),
                          ("b", p.b), ("mu", p.mu))}
        dev = R.DualFrictionProblem.from_device(ctx, p.nc, p.nb, d["rowptr"], d["colidx"], d["hvals"], d["minv"],
                                                d["b"], d["mu"], assemble=True)
        host = R.DualFrictionProblem(ctx, p, assemble=True, host_setup=True)
        for which in (0, 1, 2):
            _same_layout(host, dev, which)
        x = np.random.default_rng(8).standard_normal(3 * p.nc)
        assert np.array_equal(host.apply_w(x, op=R.OP_MATRIX_FREE), dev.apply_w(x, op=R.OP_MATRIX_FREE))
    p = G.generate(nc=100, nb=50, seed=23)
    bad_cols = np.array(p.colidx, np.int32)
    bad_cols[7] = p.nb  # a body index out of range
    bad_rows = np.array(p.rowptr, np.int32)
    bad_rows[10] = bad_rows[12]  # a decreasing row pointer
    for rp, ci in ((p.rowptr, bad_cols), (bad_rows, p.colidx)):
        arrs = [R.DeviceArray(ctx, np.asarray(v).size, t).upload(np.asarray(v).reshape(-1))
                for v, t in ((rp, np.int32), (ci, np.int32), (p.hvals, np.float64), (p.minv, np.float64),
                             (p.b, np.float64), (p.mu, np.float64))]
        with pytest.raises(RuntimeError):
            R.DualFrictionProblem.from_device(ctx, p.nc, p.nb, *arrs)


def test_device_setup_zipf_1m(R, G, ctx):
    """The 1M-contact Zipf problem (matrix-free): the device set-up's layouts equal the
    host's; its creation time is printed beside the host's."""
    import time
    print("case: device set-up vs host set-up, zipf-1m", flush=True)
    p = G.workload("zipf-1m")
    t0 = time.perf_counter()
    host = R.DualFrictionProblem(ctx, p, assemble=False, host_setup=True)
    ctx.synchronize()
    t1 = time.perf_counter()
    dev = R.DualFrictionProblem(ctx, p, assemble=False)
    ctx.synchronize()
    t2 = time.perf_counter()
    print(f"   create: host set-up {1e3 * (t1 - t0):.1f} ms, device set-up {1e3 * (t2 - t1):.1f} ms", flush=True)
    _same_layout(host, dev, 0)
    lt = _same_layout(host, dev, 1)
    assert lt["nlong"] > 0
    x = np.random.default_rng(9).standard_normal(3 * p.nc)
    assert np.array_equal(host.apply_w(x, op=R.OP_MATRIX_FREE), dev.apply_w(x, op=R.OP_MATRIX_FREE))


@pytest.mark.parametrize("case", ["uniform-5k", "contacts-50k", "ragged", "zipf-1m"])
def test_device_coloring_equals_host(R, G, ctx, case):
    """The greedy coloring on the device (rounds that reproduce the sequential rule)
    gives the host's permutation and color offsets, also for the ragged problem's
    contacts with no, one, three, twice the same, or a hundred bodies; where the device
    rule's limits are exceeded (the 1M Zipf problem's hubs: more contacts at a body than
    colors) the device-only call reports SB_ERR_LIMIT; the reference rule stays on the
    host."""
    import time
    print(f"case: device coloring vs host coloring, {case}", flush=True)
    if case == "uniform-5k":
        p, assemble = G.generate(nc=5000, nb=2000, seed=21), True
    elif case == "contacts-50k":
        p, assemble = G.workload("contacts-50k"), True
    elif case == "ragged":
        p, assemble = _ragged_problem(G), True
    else:
        p, assemble = G.workload("zipf-1m"), False
    dp = R.DualFrictionProblem(ctx, p, assemble=assemble)
    mode = R.COLORING_GREEDY
    if case == "zipf-1m":
        with pytest.raises(R.SbError) as e:
            dp.coloring(mode | R.COLORING_DEVICE)
        assert e.value.code == R.SB_ERR_LIMIT
        return
    t0 = time.perf_counter()
    pd, cd = dp.coloring(mode | R.COLORING_DEVICE)
    t1 = time.perf_counter()
    ph, ch = dp.coloring(mode | R.COLORING_HOST)
    t2 = time.perf_counter()
    print(f"   {len(cd) - 1} colors, device {1e3 * (t1 - t0):.1f} ms, host {1e3 * (t2 - t1):.1f} ms", flush=True)
    assert np.array_equal(ph, pd) and np.array_equal(ch, cd)
    with pytest.raises(R.SbError) as e:
        dp.coloring(R.COLORING_REFERENCE | R.COLORING_DEVICE)
    assert e.value.code == R.SB_ERR_UNSUPPORTED


@pytest.mark.parametrize("assemble", [True, False])
def test_device_permutation_equals_host(R, G, ctx, assemble):
    """sb_problem_apply_permutation on a device set-up permutes H, b and mu on the device
    and sets the problem up again there; the result equals the host path's (a problem
    created with the host set-up) array for array, and the colored GS agrees."""
    print(f"case: device permutation vs host permutation, assemble={assemble}", flush=True)
    p = _ragged_problem(G, seed=5)
    host = R.DualFrictionProblem(ctx, p, assemble=assemble, host_setup=True)
    dev = R.DualFrictionProblem(ctx, p, assemble=assemble)
    perm, colors = dev.coloring(R.COLORING_GREEDY)
    for prob in (host, dev):
        prob.apply_permutation(perm)
    for which in ((0, 1, 2) if assemble else (0, 1)):
        _same_layout(host, dev, which)
    assert host.sizes() == dev.sizes()
    x = np.random.default_rng(10).standard_normal(3 * p.nc)
    for op in ((R.OP_ASSEMBLED, R.OP_MATRIX_FREE) if assemble else (R.OP_MATRIX_FREE,)):
        assert np.array_equal(host.apply_w(x, op=op), dev.apply_w(x, op=op))
        assert host.eval_with(x, law=R.LAW_SOC, op=op, scaling=R.SCALING_DIAGONAL) == \
            dev.eval_with(x, law=R.LAW_SOC, op=op, scaling=R.SCALING_DIAGONAL)
    if assemble:
        for a, b in zip(host.w_lower(), dev.w_lower()):
            assert np.array_equal(a, b)
        gh = host.solve_gs(np.zeros(3 * p.nc), law=R.LAW_SOC, tol=1e-8, max_iters=500, colors=colors)
        gd = dev.solve_gs(np.zeros(3 * p.nc), law=R.LAW_SOC, tol=1e-8, max_iters=500, colors=colors)
        assert gh.iterations == gd.iterations and np.array_equal(gh.r, gd.r)
