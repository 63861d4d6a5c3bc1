"""ctypes binding of the FP64 CPU oracle (oracle/_build/libcpuref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to check
results, and by bench.py's cpu_baseline leg. The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libcpuref.so"

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

PG_STANDARD, PG_DESCENT, PG_CONJUGATED, PG_APGD, PG_SPG = range(5)
PURE_NEWTON, PURE_ENUMERATIVE, HYBRID, REV_HYBRID = range(4)
STRICTLY_POSITIVE, STRICTLY_NEGATIVE, ALL_ROOTS = range(3)
LOCAL_TOL = np.finfo(np.float64).eps ** 0.75


def build(force: bool = False) -> Path:
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE), "-s"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        sig = {
            "orc_set_threads": (None, [C.c_int]),
            "orc_set_pg_line_search": (None, [C.c_int, C.c_double, C.c_double, C.c_double]),
            "orc_get_threads": (C.c_int, []),
            "orc_problem_create": (C.c_void_p, [C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]),
            "orc_problem_destroy": (None, [C.c_void_p]),
            "orc_problem_assemble_w": (C.c_int, [C.c_void_p]),
            "orc_w_nnzb": (C.c_int, [C.c_void_p]),
            "orc_w_export": (None, [C.c_void_p, _ip, _ip, _dp]),
            "orc_diag_export": (None, [C.c_void_p, _dp]),
            "orc_problem_export": (None, [C.c_void_p, _ip, _ip, _dp, _dp, _dp]),
            "orc_apply_w": (None, [C.c_void_p, C.c_int, _dp, _dp]),
            "orc_eval": (C.c_double, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp]),
            "orc_eval_raw": (C.c_double, [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
            "orc_pg_solve": (C.c_double, [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, _dp, _ip, _ip, _dp, C.c_int, _ip]),
            "orc_gs_solve": (C.c_double, [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                          C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                                          _dp, _ip, _dp, C.c_int, _ip]),
            "orc_pgs_solve": (C.c_double, [C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                           C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int,
                                           _dp, _ip, _dp, C.c_int, _ip]),
            "orc_cadoux_solve": (C.c_double, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                              C.c_int, C.c_double, C.c_int, _ip, C.c_int, _dp, _ip, _ip, _dp,
                                              C.c_int, _ip]),
            "orc_coloring": (C.c_int, [C.c_void_p, _ip, _ip]),
            "orc_problem_permute": (None, [C.c_void_p, _ip]),
            "orc_fb": (None, [C.c_int, C.c_double, _dp, _dp, _dp]),
            "orc_fb_base_jacobian": (None, [C.c_double, _dp, _dp, _dp, _dp, _dp]),
            "orc_law_eval": (C.c_double, [C.c_int, C.c_double, _dp, _dp]),
            "orc_project": (None, [C.c_double, _dp]),
            "orc_cov": (None, [C.c_int, C.c_double, _dp, _dp]),
            "orc_local_solve": (C.c_double, [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double]),
            "orc_solve_orthogonality": (C.c_int, [C.c_int, _dp, _dp, _dp, C.c_double]),
            "orc_newton": (C.c_double, [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double, C.c_int, _dp]),
            "orc_real_roots_monic": (C.c_uint, [C.c_int, _dp, _dp, C.c_int]),
            "orc_real_roots_degenerate": (C.c_uint, [C.c_int, _dp, _dp, C.c_int]),
            "orc_eig_hessenberg": (None, [C.c_int, _dp, _dp, _dp]),
            "orc_lu_solve": (C.c_int, [C.c_int, _dp, _dp, _dp]),
            "orc_bsr_multiply": (None, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, C.c_int, _dp,
                                        _dp, C.c_double, C.c_double]),
            "orc_bsr_split_row": (None, [C.c_int, _ip, _ip, _dp, C.c_int, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def set_pg_line_search(ls_iters: int = 8, optimistic: float = 1.25, pessimistic: float = 0.5,
                       armijo: float = 0.5) -> None:
    """Line-search parameters of later pg_solve calls (defaults: the reference's)."""
    lib().orc_set_pg_line_search(int(ls_iters), float(optimistic), float(pessimistic), float(armijo))


class OracleProblem:
    """A dual friction problem held by the oracle (W = H Minv H^T, b, mu)."""

    def __init__(self, prob, assemble: bool = True):
        self.nc, self.nb = prob.nc, prob.nb
        self._keep = [i32(prob.rowptr), i32(prob.colidx), f64(prob.hvals), f64(prob.minv), f64(prob.b),
                      f64(prob.mu)]
        rp, ci, hv, mi, b, mu = self._keep
        self.h = lib().orc_problem_create(self.nc, self.nb, _i(rp), _i(ci), _d(hv), _d(mi), _d(b), _d(mu))
        self.assembled = False
        if assemble:
            self.assemble()

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_problem_destroy(h)
            self.h = None

    def assemble(self) -> int:
        n = lib().orc_problem_assemble_w(self.h)
        self.assembled = True
        return n

    def w_lower(self):
        nnz = lib().orc_w_nnzb(self.h)
        rp = np.zeros(self.nc + 1, np.int32)
        ci = np.zeros(nnz, np.int32)
        v = np.zeros(nnz * 9)
        lib().orc_w_export(self.h, _i(rp), _i(ci), _d(v))
        return rp, ci, v.reshape(nnz, 9)

    def diag(self) -> np.ndarray:
        d = np.zeros(self.nc * 9)
        lib().orc_diag_export(self.h, _d(d))
        return d.reshape(self.nc, 9)

    def export(self):
        rp = np.zeros(self.nc + 1, np.int32)
        lib_ = lib()
        nnz = int(self._keep[0][-1])
        ci = np.zeros(nnz, np.int32)
        hv = np.zeros(nnz * 18)
        b = np.zeros(3 * self.nc)
        mu = np.zeros(self.nc)
        lib_.orc_problem_export(self.h, _i(rp), _i(ci), _d(hv), _d(b), _d(mu))
        return rp, ci, hv.reshape(nnz, 18), b, mu

    def apply_w(self, x, assembled: bool = True) -> np.ndarray:
        x = f64(x)
        y = np.zeros(3 * self.nc)
        lib().orc_apply_w(self.h, int(assembled), _d(x), _d(y))
        return y

    def eval(self, r, coulomb: bool = True, inf_norm: bool = False, diag_scaling: bool = True,
             assembled: bool = True) -> float:
        r = f64(r)
        return lib().orc_eval(self.h, int(coulomb), int(inf_norm), int(diag_scaling), int(assembled), _d(r))

    def pg_solve(self, r0, variant=PG_APGD, tol=1e-6, max_iters=300, coulomb=False, assembled=True,
                 diag_scaling=False, inf_norm=False, hist_cap=100000):
        r = f64(r0).copy()
        it = C.c_int(0)
        sp = C.c_int(0)
        hist = np.zeros(hist_cap)
        hl = C.c_int(0)
        res = lib().orc_pg_solve(self.h, variant, tol, max_iters, int(coulomb), int(assembled), int(diag_scaling),
                                 int(inf_norm), _d(r), C.byref(it), C.byref(sp), _d(hist), hist_cap, C.byref(hl))
        return res, r, dict(iterations=it.value, spmvs=sp.value, history=hist[:min(hl.value, hist_cap)].copy())

    def _gs(self, fn, r0, tol=1e-6, max_iters=250, eval_every=25, skip_tol=1e-6, skip_iters=10, auto_reg=0.0,
                 try_zero=True, coulomb=True, strategy=REV_HYBRID, inf_norm=False, ndx_ref_per_color=False,
                 colors=None, hist_cap=10000):
        r = f64(r0).copy()
        it = C.c_int(0)
        hist = np.zeros(hist_cap)
        hl = C.c_int(0)
        if colors is None:
            cp, nco = None, 0
        else:
            colors = i32(colors)
            cp, nco = _i(colors), len(colors) - 1
        res = fn(self.h, tol, max_iters, eval_every, skip_tol, skip_iters, auto_reg, int(try_zero),
                                 int(coulomb), strategy, int(inf_norm), int(ndx_ref_per_color), cp, nco, _d(r),
                                 C.byref(it), _d(hist), hist_cap, C.byref(hl))
        return res, r, dict(iterations=it.value, history=hist[:min(hl.value, hist_cap)].copy())

    def gs_solve(self, r0, **kw):
        """GaussSeidel::solve on the assembled W (GaussSeidel.impl.hpp:24-205)."""
        return self._gs(lib().orc_gs_solve, r0, **kw)

    def pgs_solve(self, r0, **kw):
        """ProductGaussSeidel::solve with D = Minv (ProductGaussSeidel.impl.hpp:64-304): W is not used."""
        return self._gs(lib().orc_pgs_solve, r0, **kw)

    def cadoux_solve(self, r0, use_gs=False, variant=PG_APGD, assembled=True, diag_scaling=True, tol=1e-6,
                     inner_max_iters=300, iterations=10, tol_tighten=0.1, velocity=False, colors=None,
                     hist_cap=1000):
        """solveCadoux / solveCadouxVel (src/Interfaces/Cadoux.hpp:44-155) with the GS or PG inner solver."""
        r = f64(r0).copy()
        it, sp, hl = C.c_int(0), C.c_int(0), C.c_int(0)
        hist = np.zeros(hist_cap)
        if colors is None:
            cp, nco = None, 0
        else:
            colors = i32(colors)
            cp, nco = _i(colors), len(colors) - 1
        res = lib().orc_cadoux_solve(self.h, int(use_gs), variant, int(assembled), int(diag_scaling), tol,
                                     inner_max_iters, iterations, tol_tighten, int(velocity), cp, nco, _d(r),
                                     C.byref(it), C.byref(sp), _d(hist), hist_cap, C.byref(hl))
        return res, r, dict(iterations=it.value, spmvs=sp.value, history=hist[:min(hl.value, hist_cap)].copy())

    def coloring(self):
        perm = np.zeros(self.nc, np.int32)
        colors = np.zeros(self.nc + 1, np.int32)
        n = lib().orc_coloring(self.h, _i(perm), _i(colors))
        return perm, colors[:n + 1].copy()

    def permute(self, perm) -> None:
        perm = i32(perm)
        lib().orc_problem_permute(self.h, _i(perm))


# ---- per-contact functions
def fb(mu, x, y, coulomb=True):
    out = np.zeros(3)
    lib().orc_fb(int(coulomb), float(mu), _d(f64(x)), _d(f64(y)), _d(out))
    return out


def fb_base_jacobian(mu, x, y):
    fbv, dx, dy = np.zeros(3), np.zeros(9), np.zeros(9)
    lib().orc_fb_base_jacobian(float(mu), _d(f64(x)), _d(f64(y)), _d(fbv), _d(dx), _d(dy))
    return fbv, dx.reshape(3, 3), dy.reshape(3, 3)


def law_eval(mu, x, y, coulomb=True) -> float:
    return lib().orc_law_eval(int(coulomb), float(mu), _d(f64(x)), _d(f64(y)))


def project(mu, x) -> np.ndarray:
    x = f64(x).copy()
    lib().orc_project(float(mu), _d(x))
    return x


def cov(mu, y, coulomb=True) -> np.ndarray:
    s = np.zeros(3)
    lib().orc_cov(int(coulomb), float(mu), _d(f64(y)), _d(s))
    return s


def local_solve(A, b, x0, mu, coulomb=True, strategy=REV_HYBRID, tol=LOCAL_TOL, scaling=1.0):
    x = f64(x0).copy()
    res = lib().orc_local_solve(int(coulomb), strategy, _d(f64(A).reshape(9)), _d(f64(b)), _d(x), float(mu), tol,
                                scaling)
    return res, x


def solve_orthogonality(W, b, mu, coulomb=True):
    r = np.zeros(3)
    ok = lib().orc_solve_orthogonality(int(coulomb), _d(f64(W).reshape(9)), _d(f64(b)), _d(r), float(mu))
    return bool(ok), r


def newton(A, b, x0, mu, coulomb=True, scaling=1.0, tol=LOCAL_TOL, max_iters=20):
    x = f64(x0).copy()
    res = lib().orc_newton(int(coulomb), _d(f64(A).reshape(9)), _d(f64(b)), float(mu), scaling, tol, max_iters,
                           _d(x))
    return res, x


def real_roots_monic(coeffs, filt=ALL_ROOTS):
    c = f64(coeffs)
    roots = np.zeros(max(len(c), 1))
    n = lib().orc_real_roots_monic(len(c), _d(c), _d(roots), filt)
    return roots[:n].copy()


def real_roots_degenerate(coeffs, filt=ALL_ROOTS):
    c = f64(coeffs).copy()
    roots = np.zeros(max(len(c), 1))
    n = lib().orc_real_roots_degenerate(len(c) - 1, _d(c), _d(roots), filt)
    return roots[:n].copy()


def eig_hessenberg(H):
    H = f64(H)
    n = H.shape[0]
    wr, wi = np.zeros(n), np.zeros(n)
    lib().orc_eig_hessenberg(n, _d(H.reshape(-1)), _d(wr), _d(wi))
    return wr, wi


def lu_solve(A, b):
    A = f64(A)
    n = A.shape[0]
    x = np.zeros(n)
    rank = lib().orc_lu_solve(n, _d(A.reshape(-1)), _d(f64(b)), _d(x))
    return x, rank


def bsr_multiply(nbr, nbc, br, bc, rowptr, colidx, vals, x, transpose=False, sym=False, res=None, alpha=1.0,
                 beta=0.0):
    rowptr, colidx, vals, x = i32(rowptr), i32(colidx), f64(vals).reshape(-1), f64(x)
    n_out = (nbc * bc if (transpose and not sym) else nbr * br)
    out = np.zeros(n_out) if res is None else f64(res).copy()
    lib().orc_bsr_multiply(nbr, nbc, br, bc, int(sym), _i(rowptr), _i(colidx), _d(vals), int(transpose), _d(x),
                           _d(out), alpha, beta)
    return out


def bsr_split_row(nbr, rowptr, colidx, vals, row, x):
    res = np.zeros(3)
    lib().orc_bsr_split_row(nbr, _i(i32(rowptr)), _i(i32(colidx)), _d(f64(vals).reshape(-1)), row, _d(f64(x)),
                            _d(res))
    return res


def threads() -> int:
    return lib().orc_get_threads()


def cpu_count_physical() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


# ----------------------------------------------------------------- FCLib local problems
# Pure-Python restatements (small inputs only) of the dual set-up and error of the
# reference's FCLib application.

def convert_symmetric_3x3(n_contacts: int, rowptr, colidx, vals):
    """bogus::convert of a scalar row-major sparse matrix into a symmetric block matrix of
    3x3 blocks (Core/Eigen/SparseConversions.hpp:26-95): block row I reads scalar rows
    3I..3I+2, each up to its first entry right of the diagonal block (the `break` at :60
    and :77); blocks are ordered by column (the std::map of :55); a repeated entry
    overwrites (:82); the diagonal block becomes (D + D^T) / 2 (:87-93). Returns the
    lower-triangle block CSR (rowptr, colidx, vals[nblocks, 3, 3])."""
    rowptr, colidx, vals = np.asarray(rowptr), np.asarray(colidx), np.asarray(vals, dtype=np.float64)
    brp, bci, bv = [0], [], []
    for I in range(n_contacts):
        blocks = {}
        for i in range(3):
            for p in range(rowptr[3 * I + i], rowptr[3 * I + i + 1]):
                J = int(colidx[p]) // 3
                if J > I:
                    break
                blocks.setdefault(J, np.zeros((3, 3)))[i, int(colidx[p]) - 3 * J] = vals[p]
        if I in blocks:
            D = blocks[I].copy()
            blocks[I] = 0.5 * (D + D.T)
        for J in sorted(blocks):
            bci.append(J)
            bv.append(blocks[J])
        brp.append(len(bci))
    return (np.array(brp, np.int32), np.array(bci, np.int32),
            np.array(bv).reshape(-1, 3, 3) if bv else np.zeros((0, 3, 3)))


def prune_blocks(rowptr, colidx, vals, precision: float):
    """SparseBlockMatrixBase::prune (Core/Block/SparseBlockMatrixBase.impl.hpp:286-305): drops the
    blocks for which is_zero holds, i.e. Eigen's isZero(precision): every |entry| <= precision
    (Core/Eigen/BlockBindings.hpp:40-43)."""
    rp, out_ci, out_v = [0], [], []
    for i in range(len(rowptr) - 1):
        for p in range(rowptr[i], rowptr[i + 1]):
            if not np.all(np.abs(vals[p]) <= precision):
                out_ci.append(int(colidx[p]))
                out_v.append(vals[p])
        rp.append(len(out_ci))
    return (np.array(rp, np.int32), np.array(out_ci, np.int32),
            np.array(out_v).reshape(-1, 3, 3) if out_v else np.zeros((0, 3, 3)))


def merit1(r, u, b, mu) -> float:
    """FCLib's merit1 as the reference application computes it (apps/FCLibLoader.cpp:40-61):
    sqrt(sum_i |P_K(r_i - u_i - mu_i |u_iT| e_N) - r_i|^2) / (1 + |b|), P_K the projection on the
    cone of aperture mu_i (SOCLaw::projectOnConstraint, restated by project above)."""
    r, u, b, mu = f64(r), f64(u), f64(b), f64(mu)
    err = 0.0
    for i in range(mu.size):
        x = r[3 * i:3 * i + 3] - u[3 * i:3 * i + 3]
        x[0] -= mu[i] * np.linalg.norm(u[3 * i + 1:3 * i + 3])
        x = project(mu[i], x)
        err += float(np.sum((x - r[3 * i:3 * i + 3]) ** 2))
    return float(np.sqrt(err) / (1.0 + np.linalg.norm(b)))
