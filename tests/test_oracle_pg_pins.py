"""Pins of the oracle's projected-gradient step sequences (no GPU).

Each variant of ProjectedGradient (src/Core/BlockSolvers/ProjectedGradient.impl.hpp:
APGD and Descent :74-155, Standard :157-209, Conjugated :211-286, SPG :288-351, with
guess_alpha / nesterov_inertia / armijo_ls at :22-72) is restated here in numpy on a
dense W, line by line, and the oracle's residual history must follow it for the
first iterations to 1e-9 relative. A wrong line-search constant, a wrong Nesterov
coefficient, a wrong Barzilai-Borwein step or a missing projection shows up within
a few iterations. Standard's first line-search trial is the Cauchy step along the
projected gradient, where the Armijo test holds with equality in exact arithmetic
when the projection is inactive; its pin therefore uses an Armijo coefficient of 0.4
(set in both codes), which takes that tie away.
"""
from __future__ import annotations

import numpy as np
import pytest


def _proj(X, mu):
    X = X.reshape(-1, 3).copy()
    t = np.linalg.norm(X[:, 1:], axis=1)
    out = X.copy()
    neg = mu < 0
    polar = (mu * t <= -X[:, 0]) | neg
    inside = (t <= mu * X[:, 0]) & ~neg
    rest = ~inside & ~polar
    a = (X[rest, 0] + mu[rest] * t[rest]) / (1 + mu[rest] ** 2)
    f = a * mu[rest] / t[rest]
    out[rest, 0] = a
    out[rest, 1] *= f
    out[rest, 2] *= f
    out[polar] = 0.0
    return out.reshape(-1)


class _Np:
    def __init__(self, O, prob, coulomb):
        self.W, self.b, self.mu, self.coul, self.nc = prob.dense_w(), prob.b, prob.mu, coulomb, prob.nc
        d = O.OracleProblem(prob).diag()
        self.scal = np.maximum(1.0, (d[:, 0] + d[:, 4] + d[:, 8]) / 3.0)
        self.O = O

    def cov(self, y):
        Y = y.reshape(-1, 3)
        s = np.zeros_like(Y)
        if self.coul:
            s[:, 0] = self.mu * np.linalg.norm(Y[:, 1:], axis=1)
        return s.reshape(-1)

    def res(self, x, y):
        return sum(self.O.law_eval(self.mu[i], self.scal[i] * x[3 * i:3 * i + 3], y[3 * i:3 * i + 3],
                                   coulomb=self.coul) for i in range(self.nc)) / (self.nc + 1)

    def guess_alpha(self, x):
        d = x - 1.0
        nMx = np.sum((self.W @ d) ** 2)
        return min(1.0, d @ d / nMx) if nMx > 1e-16 else 1.0

    @staticmethod
    def nesterov(theta, q=0.0):
        tp2 = theta * theta
        bb = tp2 - q
        delta = np.sqrt(bb * bb + 4 * tp2)
        th = 0.5 * (delta - bb)
        return (theta - tp2) / (tp2 + th), th

    def armijo(self, x, bs, d, grad, J, alpha, armijo):
        for _ in range(8):
            xs = _proj(x + alpha * d, self.mu)
            Mx = self.W @ xs
            Js = xs @ (0.5 * Mx + bs)
            if Js < J + armijo * (grad @ (xs - x)):
                break
            alpha *= 0.5
        return xs, Mx, alpha

    def apgd(self, x, K, nesterov_on=True):
        W, b, mu = self.W, self.b, self.mu
        x = _proj(x, mu)
        Mx = W @ x
        theta, prev, alpha = 1.0, x.copy(), self.guess_alpha(x)
        hist = []
        for _ in range(K):
            y = Mx + b
            hist.append(self.res(x, y))
            s = self.cov(y)
            y = y + s
            J = x @ (0.5 * Mx + b + s)
            alpha *= 1.25
            for _ in range(8):
                xs = _proj(x - alpha * y, mu)
                Mx = W @ xs
                Js = xs @ (0.5 * Mx + b + s)
                if Js < J + y @ (xs - x) + 0.5 / alpha * np.sum((xs - x) ** 2):
                    break
                alpha *= 0.5
            if nesterov_on and (xs - prev) @ y < 0:
                beta, theta = self.nesterov(theta)
                x = xs + beta * (xs - prev)
                Mx = W @ x
            else:
                x = xs
                theta = 1.0
            prev = xs
        return np.array(hist)

    def spg(self, x, K):
        W, b, mu = self.W, self.b, self.mu
        x = _proj(x, mu)
        prev = x.copy()
        Mx = W @ x
        theta, alpha = 1.0, self.guess_alpha(x)
        hist = []
        for _ in range(K):
            y = Mx + b
            hist.append(self.res(x, y))
            y = y + self.cov(y)
            xs = _proj(x - alpha * y, mu)
            if (xs - prev) @ y < 0.0:
                beta, theta = self.nesterov(theta)
                w = (xs + beta * (xs - prev)) - x
                x = x + w
            else:
                w = xs - x
                x = xs
                theta = 1.0
            z = -Mx
            Mx = W @ x
            prev = xs
            z = z + Mx
            alpha = min(1e6, max(1e-6, (w @ z) / (z @ z)))
        return np.array(hist)

    def standard(self, x, K, armijo):
        W, b, mu = self.W, self.b, self.mu
        x = _proj(x, mu)
        Mx = W @ x
        alpha = 1.0
        hist = []
        for _ in range(K):
            y = Mx + b
            hist.append(self.res(x, y))
            s = self.cov(y)
            y = y + s
            J = x @ (0.5 * Mx + b + s)
            pg = x - _proj(x - y, mu)
            beta = -(y @ pg) / (pg @ (W @ pg))
            alpha = min(1.0, alpha * 1.25)
            x, Mx, alpha = self.armijo(x, b + s, beta * pg, y, J, alpha, armijo)
        return np.array(hist)

    def conjugated(self, x, K, armijo):
        W, b, mu = self.W, self.b, self.mu
        x = _proj(x, mu)
        Mx = W @ x
        prev_pg, prev_n2, prev_dir, alpha = x.copy(), 0.0, None, 1.0
        hist = []
        for _ in range(K):
            y = Mx + b
            hist.append(self.res(x, y))
            s = self.cov(y)
            y = y + s
            J = x @ (0.5 * Mx + b + s)
            pg = x - _proj(x - y, mu)
            ng2 = pg @ pg
            if prev_n2 == 0.0:
                d = -y
            else:
                den = prev_dir @ (pg - prev_pg)
                beta = max(0.0, ng2 - pg @ prev_pg) / (prev_n2 if den < 1e-12 else den)
                d = beta * prev_dir - y
                if d @ y >= 0.0:
                    d = -y
            prev_pg, prev_n2 = pg, ng2
            alpha *= 1.25
            xs, Mx, alpha = self.armijo(x, b + s, d, y, J, alpha, armijo)
            prev_dir = (xs - x) / alpha
            x = xs
        return np.array(hist)


@pytest.mark.parametrize("coulomb", [False, True])
@pytest.mark.parametrize("variant", ["PG_APGD", "PG_DESCENT", "PG_SPG", "PG_STANDARD", "PG_CONJUGATED"])
def test_oracle_pg_follows_numpy_restatement(G, O, variant, coulomb):
    prob = G.generate(60, 30, seed=5)
    K = 30
    armijo = 0.4 if variant == "PG_STANDARD" else 0.5
    ref = _Np(O, prob, coulomb)
    x0 = np.zeros(3 * prob.nc)
    hist = {"PG_APGD": lambda: ref.apgd(x0, K), "PG_DESCENT": lambda: ref.apgd(x0, K, nesterov_on=False),
            "PG_SPG": lambda: ref.spg(x0, K), "PG_STANDARD": lambda: ref.standard(x0, K, armijo),
            "PG_CONJUGATED": lambda: ref.conjugated(x0, K, armijo)}[variant]()
    O.set_pg_line_search(armijo=armijo)
    try:
        _, _, info = O.OracleProblem(prob).pg_solve(x0, variant=getattr(O, variant), tol=1e-30, max_iters=K,
                                                    coulomb=coulomb)
    finally:
        O.set_pg_line_search()
    assert len(info["history"]) == K
    assert np.allclose(info["history"], hist, rtol=1e-9, atol=0), (info["history"][:8], hist[:8])


def test_oracle_pg_pin_detects_a_wrong_constant(G, O):
    """The pin is sensitive: an optimistic factor of 1.3 instead of 1.25 in the oracle
    separates its APGD history from the restatement within the compared iterations."""
    prob = G.generate(60, 30, seed=5)
    hist = _Np(O, prob, True).apgd(np.zeros(3 * prob.nc), 30)
    O.set_pg_line_search(optimistic=1.3)
    try:
        _, _, info = O.OracleProblem(prob).pg_solve(np.zeros(3 * prob.nc), variant=O.PG_APGD, tol=1e-30,
                                                    max_iters=30, coulomb=True)
    finally:
        O.set_pg_line_search()
    assert not np.allclose(info["history"], hist, rtol=1e-9, atol=0)
