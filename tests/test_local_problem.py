"""Local (dual) problems as the reference's FCLib application handles them
(apps/FCLibLoader.cpp:40-73): the oracle's restatements of the symmetric 3x3 block
conversion, prune and merit1, pinned by dense numpy constructions and closed forms;
and the library's file layout (sb_local_problem_write / _read: host code, no GPU)."""
from __future__ import annotations

import numpy as np
import pytest


def dense_to_csr(A):
    rp, ci, v = [0], [], []
    for row in A:
        nz = np.nonzero(row)[0]
        ci.extend(nz.tolist())
        v.extend(row[nz].tolist())
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(v)


def block_pattern_matrix(rng, n, density=0.3, symmetric=True):
    """A 3n x 3n matrix whose nonzeros come in 3x3 blocks (random block pattern), with a
    diagonal block in every block row."""
    A = np.zeros((3 * n, 3 * n))
    for I in range(n):
        for J in range(I + 1):
            if I == J or rng.random() < density:
                B = rng.standard_normal((3, 3))
                A[3 * I:3 * I + 3, 3 * J:3 * J + 3] = B
                if symmetric:
                    A[3 * J:3 * J + 3, 3 * I:3 * I + 3] = B.T
                else:
                    A[3 * J:3 * J + 3, 3 * I:3 * I + 3] = rng.standard_normal((3, 3))
    if symmetric:
        A = 0.5 * (A + A.T)
    return A


def blocks_to_dense_lower(n, rp, ci, bv):
    L = np.zeros((3 * n, 3 * n))
    for I in range(n):
        for p in range(rp[I], rp[I + 1]):
            J = ci[p]
            L[3 * I:3 * I + 3, 3 * J:3 * J + 3] = bv[p]
    return L


def test_convert_symmetric_round_trip(O):
    # a symmetric W comes back exactly from its lower blocks (the strict lower part transposed)
    rng = np.random.default_rng(3)
    n = 9
    W = block_pattern_matrix(rng, n)
    rp, ci, bv = O.convert_symmetric_3x3(n, *dense_to_csr(W))
    L = blocks_to_dense_lower(n, rp, ci, bv)
    D = np.zeros_like(W)
    for I in range(n):
        D[3 * I:3 * I + 3, 3 * I:3 * I + 3] = L[3 * I:3 * I + 3, 3 * I:3 * I + 3]
    assert np.array_equal(L + (L - D).T, W)
    assert all(np.all(np.diff(ci[rp[i]:rp[i + 1]]) > 0) and ci[rp[i + 1] - 1] == i for i in range(n))


def test_convert_unsymmetric_keeps_lower_blocks_and_symmetrises_diagonal(O):
    # upper blocks are ignored; the diagonal block is (D + D^T) / 2: a dense construction
    rng = np.random.default_rng(4)
    n = 7
    A = block_pattern_matrix(rng, n, symmetric=False)
    rp, ci, bv = O.convert_symmetric_3x3(n, *dense_to_csr(A))
    mask = np.kron(np.tril(np.ones((n, n))), np.ones((3, 3)))
    want = A * mask
    for I in range(n):
        Dg = A[3 * I:3 * I + 3, 3 * I:3 * I + 3]
        want[3 * I:3 * I + 3, 3 * I:3 * I + 3] = 0.5 * (Dg + Dg.T)
    assert np.array_equal(blocks_to_dense_lower(n, rp, ci, bv), want)


def test_convert_row_read_stops_at_first_upper_entry(O):
    # an unsorted scalar row is read up to its first entry right of the diagonal block (the
    # reference's `break`): row 0 lists column 4 (block 1) before column 1, so (0, 1) is dropped
    rp = np.array([0, 2, 3, 3, 5, 5, 5], np.int32)
    ci = np.array([4, 1, 0, 0, 3], np.int32)
    v = np.array([9.0, 2.0, 3.0, 5.0, 6.0])
    brp, bci, bv = O.convert_symmetric_3x3(2, rp, ci, v)
    assert list(brp) == [0, 1, 3] and list(bci) == [0, 0, 1]
    D0 = np.zeros((3, 3))
    D0[1, 0] = D0[0, 1] = 1.5  # (3 + 0) / 2 both ways: the dropped 2.0 is not there
    B10 = np.zeros((3, 3))
    B10[0, 0] = 5.0
    D1 = np.zeros((3, 3))
    D1[0, 0] = 6.0
    assert np.array_equal(bv[0], D0) and np.array_equal(bv[1], B10) and np.array_equal(bv[2], D1)


def test_prune_drops_blocks_with_all_entries_at_most_precision(O):
    rng = np.random.default_rng(5)
    n = 6
    W = block_pattern_matrix(rng, n)
    rp, ci, bv0 = O.convert_symmetric_3x3(n, *dense_to_csr(W))
    bv = bv0 * np.where(rng.random(len(bv0)) < 0.4, 1e-9, 1.0)[:, None, None]
    prp, pci, pbv = O.prune_blocks(rp, ci, bv, 1e-8)
    keep = np.abs(bv).max(axis=(1, 2)) > 1e-8
    assert 0 < keep.sum() < len(bv)
    assert prp[-1] == keep.sum()
    assert np.array_equal(pbv, bv[keep]) and np.array_equal(pci, ci[keep])
    # the boundary: a block whose largest entry equals the precision is dropped
    b2 = bv0.copy()
    b2[0] = 0.0
    b2[0, 1, 2] = -1e-8
    _, _, q = O.prune_blocks(rp, ci, b2, 1e-8)
    assert len(q) == len(b2) - 1


@pytest.mark.parametrize("mu", [0.3, 0.8])
def test_merit1_closed_forms(O, mu):
    b = np.array([0.5, -0.2, 0.1])
    nb = 1.0 + np.linalg.norm(b)
    # sticking: r strictly inside the cone, u = 0
    r = np.array([1.0, 0.2 * mu, -0.1 * mu])
    assert O.merit1(r, np.zeros(3), b, [mu]) == 0.0
    # separation: r = 0, u_N >= 0, any u_T
    assert O.merit1(np.zeros(3), np.array([0.7, 0.3, -0.4]), b, [mu]) == 0.0
    # sliding: r on the boundary, u_N = 0, u_T opposite to r_T
    t = np.array([0.6, 0.8])
    r = np.array([2.0, *(mu * 2.0 * t)])
    u = np.array([0.0, *(-0.35 * t)])
    assert O.merit1(r, u, b, [mu]) == pytest.approx(0.0, abs=1e-15)
    # penetration with r = 0 and u_T = 0: x = (-u_N, 0, 0) is inside the cone, error |u_N|
    assert O.merit1(np.zeros(3), np.array([-1.5, 0.0, 0.0]), b, [mu]) == pytest.approx(1.5 / nb, rel=1e-15)
    # a disabled contact (mu < 0) projects to 0: error |r|
    r = np.array([0.3, 0.4, 0.0])
    assert O.merit1(r, np.array([0.1, 0.2, 0.3]), b, [-1.0]) == pytest.approx(0.5 / nb, rel=1e-15)


def test_merit1_sums_contacts(O):
    rng = np.random.default_rng(6)
    r, u, b = rng.standard_normal(12), rng.standard_normal(12), rng.standard_normal(12)
    mu = np.array([0.2, 0.5, 0.9, 0.4])
    nb = 1.0 + np.linalg.norm(b)
    parts = [(O.merit1(r[3 * i:3 * i + 3], u[3 * i:3 * i + 3], np.zeros(3), mu[i:i + 1])) ** 2 for i in range(4)]
    assert O.merit1(r, u, b, mu) == pytest.approx(np.sqrt(sum(parts)) / nb, rel=1e-14)


def test_local_problem_file_round_trip(R, tmp_path):
    rng = np.random.default_rng(7)
    n = 11
    W = block_pattern_matrix(rng, n)
    rp, ci, v = dense_to_csr(W)
    q, mu, r0 = rng.standard_normal(3 * n), rng.uniform(0.1, 0.9, n), rng.standard_normal(3 * n)
    for with_r0 in (False, True):
        path = tmp_path / f"p{int(with_r0)}.sblp"
        R.write_local_problem(path, rp, ci, v, q, mu, r0 if with_r0 else None)
        d = R.read_local_problem(path)
        assert d["n"] == n
        for k, ref in (("rowptr", rp), ("colidx", ci), ("vals", v), ("q", q), ("mu", mu)):
            assert np.array_equal(d[k], ref), k
        assert (d["r0"] is None) == (not with_r0)
        if with_r0:
            assert np.array_equal(d["r0"], r0)
        size = path.stat().st_size
        assert size == 32 + 4 * (3 * n + 1) + 12 * len(v) + 8 * 3 * n + 8 * n + (8 * 3 * n if with_r0 else 0)


def test_local_problem_file_rejects_damage(R, tmp_path):
    rng = np.random.default_rng(8)
    n = 4
    rp, ci, v = dense_to_csr(block_pattern_matrix(rng, n))
    path = tmp_path / "p.sblp"
    R.write_local_problem(path, rp, ci, v, np.zeros(3 * n), np.full(n, 0.5))
    raw = path.read_bytes()
    bad = tmp_path / "bad.sblp"
    for damaged in (raw[:-8], raw + b"\0", b"XBLOCAL1" + raw[8:],
                    raw[:32] + np.int32(1).tobytes() + raw[36:],           # rowptr[0] != 0
                    raw[:32 + 4 * (3 * n + 1)] + np.int32(3 * n).tobytes() + raw[32 + 4 * (3 * n + 1) + 4:]):
        bad.write_bytes(damaged)
        with pytest.raises(R.SbError):
            R.read_local_problem(bad)
    with pytest.raises(R.SbError):
        R.read_local_problem(tmp_path / "missing.sblp")
    with pytest.raises(R.SbError):  # a column out of range is refused on writing too
        R.write_local_problem(tmp_path / "w.sblp", rp, np.full_like(ci, 3 * n), v, np.zeros(3 * n), np.ones(n))
