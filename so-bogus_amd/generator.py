"""Synthetic 3D Coulomb-friction problems (host side, deterministic).

Every contact touches two distinct rigid bodies. For contact i with bodies (a, b):

* ``E_i`` in SO(3) is the Q factor of a 3x3 N(0,1) matrix with the signs fixed
  so that diag(R) > 0 and det(E_i) = +1; row 0 is the contact normal.
* lever arms ``p_a, p_b ~ U[-1,1]^3``;
* ``H_ia = E_i [I | -[p_a]x]`` and ``H_ib = -E_i [I | -[p_b]x]`` (3x6 blocks: the
  relative velocity of the contact point, v + w x p).

For body k: ``m_k ~ U[1,10]``, ``A_k ~ N(0,1)^{3x3}``,
``M_k = diag(m_k I, A_k A_k^T + 0.1 I)`` and ``Minv_k`` is its FP64 inverse
(Cholesky). ``mu_i ~ U[0.3, 0.8]`` (``U[0.1, 1.0]`` for the Zipf workload) and
``b_i ~ N(0,1)^3``.

Random numbers come from a counter-based SplitMix64 stream per (seed, field),
so every array is a pure function of (seed, index) and independent of numpy's
generator versions.

The dual problem is ``u = W r + b`` with ``W = H Minv H^T``; this is the input of
``DualFrictionProblem<3>`` (reference: src/Interfaces/FrictionProblem.hpp:127-214).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


_FIELDS = {
    "body_a": 1, "body_b": 2, "frame": 3, "lever": 4, "mass": 5, "inertia": 6,
    "mu": 7, "b": 8, "extra": 9,
}


class Stream:
    """Counter-based uniform/normal stream for one (seed, field)."""

    def __init__(self, seed: int, name: str):
        key = np.uint64((int(seed) * 0x9E3779B97F4A7C15 + _FIELDS[name] * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        self.key = _splitmix64(np.array([key], dtype=np.uint64))[0]
        self.pos = 0

    def u64(self, n: int) -> np.ndarray:
        idx = np.arange(self.pos, self.pos + n, dtype=np.uint64)
        self.pos += n
        with np.errstate(over="ignore"):
            return _splitmix64(idx ^ self.key)

    def uniform(self, n: int) -> np.ndarray:
        """Uniform doubles in [0, 1) with 53 random bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """Standard normals by Box-Muller (one normal per pair of uniforms)."""
        u1 = 1.0 - self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


@dataclass
class FrictionProblem:
    """Host arrays of one dual friction problem (contact-major BSR for H)."""

    nc: int
    nb: int
    rowptr: np.ndarray  # int32 [nc+1]
    colidx: np.ndarray  # int32 [nnzb]
    hvals: np.ndarray   # float64 [nnzb, 18], row-major 3x6 blocks
    minv: np.ndarray    # float64 [nb, 36], row-major 6x6 blocks
    b: np.ndarray       # float64 [3 nc]
    mu: np.ndarray      # float64 [nc]
    seed: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def nnzb(self) -> int:
        return int(self.rowptr[-1])

    def dense_h(self) -> np.ndarray:
        H = np.zeros((3 * self.nc, 6 * self.nb))
        for i in range(self.nc):
            for p in range(self.rowptr[i], self.rowptr[i + 1]):
                k = self.colidx[p]
                H[3 * i:3 * i + 3, 6 * k:6 * k + 6] = self.hvals[p].reshape(3, 6)
        return H

    def dense_minv(self) -> np.ndarray:
        M = np.zeros((6 * self.nb, 6 * self.nb))
        for k in range(self.nb):
            M[6 * k:6 * k + 6, 6 * k:6 * k + 6] = self.minv[k].reshape(6, 6)
        return M

    def dense_w(self) -> np.ndarray:
        H = self.dense_h()
        return H @ self.dense_minv() @ H.T


def _skew(p: np.ndarray) -> np.ndarray:
    """[p]x for a batch of 3-vectors: [p]x v = p x v."""
    z = np.zeros(p.shape[0])
    return np.stack([
        np.stack([z, -p[:, 2], p[:, 1]], axis=-1),
        np.stack([p[:, 2], z, -p[:, 0]], axis=-1),
        np.stack([-p[:, 1], p[:, 0], z], axis=-1),
    ], axis=1)


def _frames(st: Stream, n: int) -> np.ndarray:
    G = st.normal(9 * n).reshape(n, 3, 3)
    Q, R = np.linalg.qr(G)
    s = np.sign(np.diagonal(R, axis1=1, axis2=2))
    s[s == 0] = 1.0
    Q = Q * s[:, None, :]
    det = np.linalg.det(Q)
    Q[det < 0, :, 2] *= -1.0
    return Q


def _zipf_bodies(st: Stream, n: int, nb: int, s: float) -> np.ndarray:
    w = (np.arange(nb, dtype=np.float64) + 1.0) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    out = np.searchsorted(cdf, st.uniform(n), side="right")
    return np.minimum(out, nb - 1).astype(np.int64)


def generate(nc: int, nb: int, seed: int, zipf: float | None = None,
             mu_range: tuple[float, float] = (0.3, 0.8)) -> FrictionProblem:
    """Generate one problem. ``zipf`` is the body-degree exponent (None: uniform)."""
    if nb < 2:
        raise ValueError("need at least two bodies")
    sa, sb = Stream(seed, "body_a"), Stream(seed, "body_b")
    if zipf is None:
        a = np.minimum((sa.uniform(nc) * nb).astype(np.int64), nb - 1)
        bb = np.minimum((sb.uniform(nc) * (nb - 1)).astype(np.int64), nb - 2)
        bb += (bb >= a)
    else:
        a = _zipf_bodies(sa, nc, nb, zipf)
        bb = _zipf_bodies(sb, nc, nb, zipf)
        bad = np.nonzero(bb == a)[0]
        while bad.size:
            bb[bad] = _zipf_bodies(sb, bad.size, nb, zipf)
            bad = bad[bb[bad] == a[bad]]
    E = _frames(Stream(seed, "frame"), nc)
    lev = Stream(seed, "lever").uniform(6 * nc).reshape(nc, 2, 3) * 2.0 - 1.0
    eye = np.broadcast_to(np.eye(3), (nc, 3, 3))
    Ja = np.concatenate([eye, -_skew(lev[:, 0])], axis=2)  # 3x6
    Jb = np.concatenate([eye, -_skew(lev[:, 1])], axis=2)
    Ha = E @ Ja
    Hb = -(E @ Jb)
    # rows sorted by body index
    lo = np.minimum(a, bb)
    swap = bb < a
    H_lo = np.where(swap[:, None, None], Hb, Ha)
    H_hi = np.where(swap[:, None, None], Ha, Hb)
    hi = np.maximum(a, bb)
    rowptr = (2 * np.arange(nc + 1)).astype(np.int32)
    colidx = np.stack([lo, hi], axis=1).reshape(-1).astype(np.int32)
    hvals = np.stack([H_lo, H_hi], axis=1).reshape(2 * nc, 18)

    m = 1.0 + 9.0 * Stream(seed, "mass").uniform(nb)
    A = Stream(seed, "inertia").normal(9 * nb).reshape(nb, 3, 3)
    inertia = A @ np.transpose(A, (0, 2, 1)) + 0.1 * np.eye(3)
    L = np.linalg.cholesky(inertia)
    Linv = np.linalg.inv(L)
    inv_inertia = np.transpose(Linv, (0, 2, 1)) @ Linv
    inv_inertia = 0.5 * (inv_inertia + np.transpose(inv_inertia, (0, 2, 1)))
    minv = np.zeros((nb, 6, 6))
    minv[:, 0, 0] = minv[:, 1, 1] = minv[:, 2, 2] = 1.0 / m
    minv[:, 3:, 3:] = inv_inertia
    mu = mu_range[0] + (mu_range[1] - mu_range[0]) * Stream(seed, "mu").uniform(nc)
    b = Stream(seed, "b").normal(3 * nc)
    return FrictionProblem(nc=nc, nb=nb, rowptr=rowptr, colidx=colidx, hvals=np.ascontiguousarray(hvals),
                           minv=minv.reshape(nb, 36).copy(), b=b, mu=mu, seed=seed,
                           meta={"zipf": zipf, "mu_range": mu_range})


def body_order(p: FrictionProblem) -> FrictionProblem:
    """The same problem with its contacts listed in body order, as a simulator that
    collects contacts body by body passes them: bodies renumbered by first touch along
    the contacts sorted by their (lower, higher) body, then contacts sorted by their
    renumbered bodies (each row's blocks kept in increasing body order). W becomes
    P W P^T for the contact permutation P; the solution is permuted alike. Returns the
    reordered problem; its meta holds 'contact_perm' (new contact -> old contact) and
    'body_map' (old body -> new body)."""
    nc, nb = p.nc, p.nb
    if not np.all(np.diff(p.rowptr) == 2):
        raise ValueError("body_order expects two blocks per contact")
    col = p.colidx.reshape(nc, 2).astype(np.int64)
    lo, hi = col.min(1), col.max(1)
    perm0 = np.lexsort((hi, lo))
    seq = col[perm0].reshape(-1)
    _, first = np.unique(seq, return_index=True)
    touched = seq[np.sort(first)]
    bmap = np.full(nb, -1, dtype=np.int64)
    bmap[touched] = np.arange(touched.size)
    rest = np.nonzero(bmap < 0)[0]
    bmap[rest] = touched.size + np.arange(rest.size)
    c2 = bmap[col]
    cperm = np.lexsort((c2.max(1), c2.min(1)))
    newcol = c2[cperm]
    hv = p.hvals.reshape(nc, 2, 18)[cperm].copy()
    swap = newcol[:, 0] > newcol[:, 1]
    newcol[swap] = newcol[swap][:, ::-1]
    hv[swap] = hv[swap][:, ::-1]
    minv = np.empty_like(p.minv)
    minv[bmap] = p.minv
    meta = dict(p.meta, contact_perm=cperm, body_map=bmap)
    return FrictionProblem(nc=nc, nb=nb, rowptr=p.rowptr.copy(), colidx=newcol.reshape(-1).astype(np.int32),
                           hvals=np.ascontiguousarray(hv.reshape(2 * nc, 18)), minv=minv,
                           b=p.b.reshape(nc, 3)[cperm].reshape(-1).copy(), mu=p.mu[cperm].copy(), seed=p.seed,
                           meta=meta)


# Named workloads (sizes and seeds of the synthetic benchmark set)
WORKLOADS = {
    "small-1k": dict(nc=1000, nb=500, seed=1),
    "contacts-50k": dict(nc=50_000, nb=20_000, seed=2),
    "zipf-1m": dict(nc=1_000_000, nb=300_000, seed=3, zipf=1.2, mu_range=(0.1, 1.0)),
    "adjoint-100k": dict(nc=100_000, nb=40_000, seed=5),
}


def workload(name: str) -> FrictionProblem:
    return generate(**WORKLOADS[name])


def batch_sizes(n_problems: int = 512, seed0: int = 4000, lo: int = 64, hi: int = 20000) -> list[tuple[int, int, int]]:
    """(nc, nb, seed) of the ragged batch: nc = floor(exp(U[ln lo, ln hi])), nb = floor(nc/2)."""
    out = []
    for k in range(n_problems):
        u = Stream(seed0 + k, "extra").uniform(1)[0]
        nc = int(math.floor(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))
        nc = max(nc, 2)
        out.append((nc, max(nc // 2, 2), seed0 + k))
    return out


def small_friction_pb():
    """The reference's 2-contact fixture (tests/SmallFrictionPb.hpp:16-85), embedded
    into the fixed 3x6 / 6x6 layout by zero-padding H columns and identity-padding
    Minv (W is unchanged). Returns (problem, sol) with b = w - H Minv f."""
    dofs = [4, 2]
    H00 = np.array([[3, 5, 7, -1], [2, 4, 5, 4], [1, 6, 9, 1]], dtype=float)
    H01 = np.array([[-3, -6], [-6, -5], [-5, -8]], dtype=float)
    H10 = np.array([[3, 7, 1, -1], [3, 6, 3, 2], [4, 4, 7, -1]], dtype=float)
    H2B = np.array([[6, 7, 10, 9], [5, 8, 11, 8], [4, 9, 12, 7]], dtype=float)
    H10 = H10 - H2B
    E = np.array([[0, 0, 1], [1, 0, 0], [0, 1, 0]], dtype=float)
    H10 = E.T @ H10
    M0 = np.eye(4)
    M1 = 2 * np.eye(2)
    f = np.array([1, 2, 3, 4, 5, 6], dtype=float)
    w = np.array([2, 1, 2, 1, 3, 3], dtype=float)
    mu = np.array([0.5, 0.7])
    sol = np.array([0.0152695, 0.0073010, 0.0022325, 0.0, 0.0, 0.0])
    Hd = np.zeros((6, 6))
    Hd[0:3, 0:4] = H00
    Hd[0:3, 4:6] = H01
    Hd[3:6, 0:4] = H10
    Minv_d = np.zeros((6, 6))
    Minv_d[0:4, 0:4] = np.linalg.inv(M0)
    Minv_d[4:6, 4:6] = np.linalg.inv(M1)
    b = w - Hd @ (Minv_d @ f)

    def pad(block, nd):
        out = np.zeros((3, 6))
        out[:, :nd] = block
        return out

    rowptr = np.array([0, 2, 3], dtype=np.int32)
    colidx = np.array([0, 1, 0], dtype=np.int32)
    hvals = np.stack([pad(H00, 4), pad(H01, 2), pad(H10, 4)]).reshape(3, 18)
    minv = np.zeros((2, 6, 6))
    minv[0] = np.eye(6)
    minv[0, :4, :4] = np.linalg.inv(M0)
    minv[1] = np.eye(6)
    minv[1, :2, :2] = np.linalg.inv(M1)
    P = FrictionProblem(nc=2, nb=2, rowptr=rowptr, colidx=colidx, hvals=hvals, minv=minv.reshape(2, 36),
                        b=b, mu=mu, seed=-1, meta={"dofs": dofs})
    return P, sol


def batch(n_problems: int = 512, seed0: int = 4000, lo: int = 64, hi: int = 20000) -> list[FrictionProblem]:
    """The ragged batch: problem k from generate(nc_k, nb_k, seed0 + k)."""
    return [generate(nc, nb, seed) for nc, nb, seed in batch_sizes(n_problems, seed0, lo, hi)]
