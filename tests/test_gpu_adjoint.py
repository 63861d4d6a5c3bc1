"""Reverse-mode derivatives of the projected fixed-point map on the GPU.

T(r) = P_mu(r - rho (W r + b + s_mu(W r + b))), L = g . r. Checks:
  * the forward map against a dense numpy restatement (1e-12 relative);
  * the unrolled gradient (dL/db, dL/dmu, dL/dMinv) against central finite
    differences of the same map, h = 1e-6 in FP64, on random directions: 1e-5
    relative (directions whose difference quotient crosses a kink of the
    projection - a contact changing branch inside [-h, h] - are the only
    admissible misses; at most one of eight, each still within 1e-3);
  * the implicit gradient at a fixed point (GMRES on the adjoint system) against
    finite differences of the fixed point itself and against a long unroll started
    at the fixed point, on a small full-rank problem and on the 100,000-contact
    Coulomb problem (rank-deficient W, ill-conditioned adjoint system);
  * at 100,000 contacts, both laws, the unrolled gradient against central
    differences on 64 directions of the map replayed on the recorded piece
    (projection branches and, Coulomb law, the De Saxce directions);
  * the same gradients through a torch.autograd.Function built on the C ABI with
    ctypes (a test harness only), with torch.autograd.gradcheck.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def project_np(Z, mu):
    Z = Z.copy()
    t = np.linalg.norm(Z[:, 1:], axis=1)
    out = np.zeros_like(Z)
    inside = (t <= mu * Z[:, 0]) & (mu >= 0)
    polar = (mu * t <= -Z[:, 0]) | (mu < 0)
    rest = ~inside & ~polar
    out[inside] = Z[inside]
    a = (Z[rest, 0] + mu[rest] * t[rest]) / (1 + mu[rest] ** 2)
    out[rest, 0] = a
    f = a * mu[rest] / t[rest]
    out[rest, 1] = Z[rest, 1] * f
    out[rest, 2] = Z[rest, 2] * f
    return out


def map_np(prob, r, rho, coulomb, K):
    W = prob.dense_w()
    for _ in range(K):
        U = (W @ r + prob.b).reshape(-1, 3)
        R = r.reshape(-1, 3)
        ut = np.linalg.norm(U[:, 1:], axis=1)
        S = np.zeros_like(U)
        if coulomb:
            S[:, 0] = prob.mu * ut
        r = project_np(R - rho * (U + S), prob.mu).reshape(-1)
    return r


def _directions(prob, rng, n):
    out = []
    for _ in range(n):
        db = rng.standard_normal(3 * prob.nc)
        dmu = 0.1 * rng.standard_normal(prob.nc)
        A = rng.standard_normal((prob.nb, 6, 6)) * 0.05
        dM = (A + A.transpose(0, 2, 1)).reshape(prob.nb, 36)
        out.append((db, dmu, dM))
    return out


def _loss(dp, prob, r0, g, rho, K, db, dmu, dM, t, law):
    dp.update(b=prob.b + t * db, mu=prob.mu + t * dmu, minv=prob.minv.reshape(prob.nb, 36) + t * dM)
    rK = dp.adjoint(r0, None, iterations=K, rho=rho, law=law)["r"]
    return float(g @ rK)


def test_forward_map_matches_numpy(G, R, ctx):
    prob = G.generate(200, 100, seed=21)
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    r0 = np.zeros(3 * prob.nc)
    for law, coul in ((R.LAW_COULOMB, True), (R.LAW_SOC, False)):
        out = dp.adjoint(r0, None, iterations=50, law=law)
        ref = map_np(prob, r0, out["rho"], coul, 50)
        assert np.linalg.norm(out["r"] - ref) <= 1e-12 * np.linalg.norm(ref)
    # the step is 1 / lambda_max(W)
    lam = np.linalg.eigvalsh(prob.dense_w()).max()
    assert out["rho"] == pytest.approx(1.0 / lam, rel=1e-6)


@pytest.mark.parametrize("law", ["LAW_COULOMB", "LAW_SOC"])
def test_unroll_gradient_matches_finite_differences(G, R, ctx, law):
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    rng = np.random.default_rng(55)
    g = rng.standard_normal(3 * prob.nc)
    r0 = np.zeros(3 * prob.nc)
    K = 200
    lw = getattr(R, law)
    print(f"[adjoint] unroll K={K} law={law}", flush=True)
    ad = dp.adjoint(r0, g, iterations=K, law=lw)
    rho = ad["rho"]
    assert ad["loss"] == pytest.approx(float(g @ ad["r"]), rel=1e-12)
    h = 1e-6
    rel = []
    for db, dmu, dM in _directions(prob, rng, 8):
        lp = _loss(dp, prob, r0, g, rho, K, db, dmu, dM, h, lw)
        lm = _loss(dp, prob, r0, g, rho, K, db, dmu, dM, -h, lw)
        fd = (lp - lm) / (2 * h)
        an = ad["grad_b"] @ db + ad["grad_mu"] @ dmu + np.sum(ad["grad_minv"] * dM)
        rel.append(abs(fd - an) / max(abs(an), 1e-12))
    dp.update(b=prob.b, mu=prob.mu, minv=prob.minv.reshape(prob.nb, 36))
    rel = np.array(rel)
    print("[adjoint] relative FD errors", rel, flush=True)
    assert (rel <= 1e-5).sum() >= 7 and rel.max() <= 1e-3


def test_implicit_gradient_at_fixed_point(G, R, ctx):
    """SOC law on a problem with full-rank W (3 n_c = 6 n_b): T is a contraction,
    its fixed point r* is reached to rounding by iterating the map, and the
    implicit gradient equals finite differences of r*(theta) and a long unroll."""
    prob = G.generate(60, 30, seed=31)
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    rng = np.random.default_rng(7)
    g = rng.standard_normal(3 * prob.nc)
    Kfix = 20000
    fwd = dp.adjoint(np.zeros(3 * prob.nc), None, iterations=Kfix, law=R.LAW_SOC)
    rho, rstar = fwd["rho"], fwd["r"]
    again = dp.adjoint(rstar, None, iterations=1, rho=rho, law=R.LAW_SOC)["r"]
    assert np.linalg.norm(again - rstar) <= 1e-13 * max(np.linalg.norm(rstar), 1.0)
    imp = dp.adjoint(rstar, g, mode=R.ADJOINT_IMPLICIT, rho=rho, law=R.LAW_SOC, tol=1e-13, max_iters=200000)
    assert imp["converged"]
    unr = dp.adjoint(rstar, g, iterations=Kfix, rho=rho, law=R.LAW_SOC)
    for key in ("grad_b", "grad_mu", "grad_minv"):
        a, b = imp[key], unr[key]
        assert np.linalg.norm(a - b) <= 1e-6 * max(np.linalg.norm(a), 1e-12), key
    h = 1e-6
    for db, dmu, dM in _directions(prob, rng, 4):
        lp = _loss(dp, prob, np.zeros(3 * prob.nc), g, rho, Kfix, db, dmu, dM, h, R.LAW_SOC)
        lm = _loss(dp, prob, np.zeros(3 * prob.nc), g, rho, Kfix, db, dmu, dM, -h, R.LAW_SOC)
        fd = (lp - lm) / (2 * h)
        an = imp["grad_b"] @ db + imp["grad_mu"] @ dmu + np.sum(imp["grad_minv"] * dM)
        assert abs(fd - an) <= 1e-5 * max(abs(an), 1e-12)
    dp.update(b=prob.b, mu=prob.mu, minv=prob.minv.reshape(prob.nb, 36))


def test_torch_autograd_function_gradcheck(G, R, ctx):
    torch = pytest.importorskip("torch")
    prob = G.generate(16, 8, seed=41)
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    K = 30
    rho = dp.adjoint(np.zeros(3 * prob.nc), None, iterations=1)["rho"]
    r0 = np.zeros(3 * prob.nc)

    class UnrolledSolve(torch.autograd.Function):
        """r_K = T^K(r0; b, mu, Minv) through the C ABI (test harness)."""

        @staticmethod
        def forward(fctx, b, mu, minv):
            dp.update(b=b.detach().numpy(), mu=mu.detach().numpy(), minv=minv.detach().numpy())
            fctx.save_for_backward(b, mu, minv)
            return torch.from_numpy(dp.adjoint(r0, None, iterations=K, rho=rho)["r"].copy())

        @staticmethod
        def backward(fctx, gout):
            b, mu, minv = fctx.saved_tensors
            dp.update(b=b.detach().numpy(), mu=mu.detach().numpy(), minv=minv.detach().numpy())
            out = dp.adjoint(r0, gout.detach().numpy().astype(np.float64), iterations=K, rho=rho)
            return (torch.from_numpy(out["grad_b"].copy()), torch.from_numpy(out["grad_mu"].copy()),
                    torch.from_numpy(out["grad_minv"].reshape(-1).copy()).reshape(minv.shape))

    b = torch.tensor(prob.b, dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(prob.mu, dtype=torch.float64, requires_grad=True)
    minv = torch.tensor(prob.minv.reshape(prob.nb, 36), dtype=torch.float64, requires_grad=True)
    w = torch.tensor(np.random.default_rng(3).standard_normal(3 * prob.nc))
    f = lambda b_, mu_, m_: (UnrolledSolve.apply(b_, mu_, m_) * w).sum()
    assert torch.autograd.gradcheck(f, (b, mu, minv), eps=1e-6, atol=1e-6, rtol=1e-4)
    dp.update(b=prob.b, mu=prob.mu, minv=prob.minv.reshape(prob.nb, 36))


@pytest.mark.parametrize("law", ["LAW_SOC", "LAW_COULOMB"])
def test_adjoint_100k_against_64_fd_directions(G, R, ctx, law):
    """100,000 contacts / 40,000 bodies: the solution of the mixed-precision solve
    (FP32 bulk, FP64 residual and verification) starts a 200-step unroll whose
    gradient is checked against central differences (h = 1e-6, FP64) on 64 random
    directions in (b, mu, Minv).

    The map is piecewise smooth: some (contact, step) pairs change projection branch
    within +-h in every direction at this size, and the De Saxce term mu |u_T| has a
    kink at u_T = 0, which sticking contacts approach at the solution (|u_T| of 1e-9
    and below). The perturbed runs therefore replay the unperturbed run's recorded
    piece: its projection branches and its De Saxce directions d = u_T / |u_T|
    (mu (d . u_T) in place of mu |u_T|: the same value and gradient at the recorded
    point). Both laws: 1e-5 on all 64 directions. Plain differences are reported with
    the number of branch changes they cross."""
    prob = G.workload("adjoint-100k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    print(f"[adjoint] 100k {law}: mixed solve", flush=True)
    sol = dp.solve_pg(np.zeros(3 * prob.nc), variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE,
                      scaling=R.SCALING_DIAGONAL, tol=1e-6, max_iters=20000, precision=R.PRECISION_MIXED)
    assert sol.converged
    lw = getattr(R, law)
    rng = np.random.default_rng(56)
    g = np.random.default_rng(55).standard_normal(3 * prob.nc)
    K = 200
    ad = dp.adjoint(sol.r, g, iterations=K, law=lw, record_branches=True)
    rho, br = ad["rho"], ad["branches"]
    h = 1e-6
    frozen, plain, flips = [], [], []
    for db, dmu, dM in _directions(prob, rng, 64):
        an = ad["grad_b"] @ db + ad["grad_mu"] @ dmu + np.sum(ad["grad_minv"] * dM)
        lf, lp, nfl = [], [], 0
        for t in (h, -h):
            dp.update(b=prob.b + t * db, mu=prob.mu + t * dmu, minv=prob.minv.reshape(prob.nb, 36) + t * dM)
            lf.append(float(g @ dp.adjoint(sol.r, None, iterations=K, rho=rho, law=lw, branches=br)["r"]))
            free = dp.adjoint(sol.r, None, iterations=K, rho=rho, law=lw)
            lp.append(float(g @ free["r"]))
        frozen.append(abs((lf[0] - lf[1]) / (2 * h) - an) / max(abs(an), 1e-12))
        plain.append(abs((lp[0] - lp[1]) / (2 * h) - an) / max(abs(an), 1e-12))
    dp.update(b=prob.b, mu=prob.mu, minv=prob.minv.reshape(prob.nb, 36))
    frozen, plain = np.array(frozen), np.array(plain)
    print(f"[adjoint] 100k {law} replayed-piece FD: max rel {frozen.max():.2e}, median {np.median(frozen):.2e}; "
          f"plain FD: median {np.median(plain):.2e}, max {plain.max():.2e}", flush=True)
    assert frozen.max() <= 1e-5


def test_adjoint_100k_implicit_coulomb(G, R, ctx):
    """The implicit derivative of the 100,000-contact Coulomb solution. W is
    rank-deficient (3 n_c = 300k > 6 n_b = 240k) and I - dT/dr^T is ill-conditioned
    (a Neumann series contracts by about 0.9997 per term), so the adjoint system is
    solved by restarted GMRES; the result must reach its residual tolerance and agree
    with long unrolls started at the fixed point, which converge to it as K grows."""
    prob = G.workload("adjoint-100k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    print("[adjoint] 100k Coulomb implicit: FP64 solve to 1e-10", flush=True)
    s = dp.solve_pg(np.zeros(3 * prob.nc), variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE,
                    scaling=R.SCALING_DIAGONAL, tol=1e-10, max_iters=100000)
    assert s.converged
    g = np.random.default_rng(55).standard_normal(3 * prob.nc)
    rho = dp.adjoint(s.r, None, iterations=1, law=R.LAW_COULOMB)["rho"]
    imp = dp.adjoint(s.r, g, mode=R.ADJOINT_IMPLICIT, rho=rho, law=R.LAW_COULOMB, tol=1e-10, max_iters=5000)
    print(f"[adjoint] GMRES: {imp['iterations']} steps, residual {imp['change']:.2e}, {imp['device_ms']:.0f} ms",
          flush=True)
    assert imp["converged"] and imp["change"] < 1e-10
    assert imp["loss"] == pytest.approx(float(g @ s.r), rel=1e-12)
    errs = {}
    for K in (10000, 40000):
        u = dp.adjoint(s.r, g, iterations=K, rho=rho, law=R.LAW_COULOMB)
        errs[K] = [np.linalg.norm(imp[k] - u[k]) / np.linalg.norm(imp[k]) for k in ("grad_b", "grad_mu", "grad_minv")]
        print(f"[adjoint] unroll K={K}: relative difference to implicit (b, mu, Minv) "
              f"{', '.join(f'{e:.2e}' for e in errs[K])}", flush=True)
    assert max(errs[40000]) <= 5e-4
    assert all(e4 < 0.1 * e1 for e1, e4 in zip(errs[10000], errs[40000]))
