"""The synthetic generator's body-ordered variant of a problem (generator.body_order):
the same problem with contacts listed body by body. CPU only."""
from __future__ import annotations

import numpy as np


def test_body_order_is_a_symmetric_permutation(G):
    p = G.generate(300, 120, seed=3, zipf=1.2)
    q = G.body_order(p)
    c, bmap = q.meta["contact_perm"], q.meta["body_map"]
    assert sorted(c.tolist()) == list(range(p.nc)) and sorted(bmap.tolist()) == list(range(p.nb))
    rows = (3 * c[:, None] + np.arange(3)).reshape(-1)
    W, Wq = p.dense_w(), q.dense_w()
    assert np.abs(Wq - W[np.ix_(rows, rows)]).max() <= 1e-12 * np.abs(W).max()
    assert np.array_equal(q.b.reshape(-1, 3), p.b.reshape(-1, 3)[c])
    assert np.array_equal(q.mu, p.mu[c])
    cols = q.colidx.reshape(-1, 2)
    assert np.all(cols[:, 0] < cols[:, 1])
    # contacts sorted by their (lower, higher) renumbered body
    key = cols[:, 0].astype(np.int64) * p.nb + cols[:, 1]
    assert np.all(np.diff(key) >= 0)
    # bodies renumbered by first touch: the first contact's bodies are 0 and 1
    assert cols[0, 0] == 0 and cols[0, 1] == 1
    assert np.array_equal(q.minv[bmap], p.minv)
