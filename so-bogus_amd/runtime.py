"""ctypes binding of libsbhip.so (the C ABI declared in include/sobogus_hip.h).

Python-side mirror of the reference's solve-path API for tests and benchmarks:
``SparseBlockMatrix`` (BSR multiply), ``DualFrictionProblem`` (solveWith /
evalWith / applyPermutation / coloring), ``ProjectedGradient`` and
``GaussSeidel`` parameter objects. Every product, residual and solve runs in the
HIP library; if the library is missing or no GPU is usable, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("SOBOGUS_LIB", PKG_DIR / "lib" / "libsbhip.so"))
HEADER = PKG_DIR.parent / "include" / "sobogus_hip.h"

SB_OK, SB_NOT_CONVERGED = 0, 1
SB_ERR = {-1: "invalid argument", -2: "out of memory", -3: "HIP error", -4: "no device", -5: "unsupported",
          -6: "barrier timeout", -7: "non-finite residual", -8: "fixed limit exceeded"}
SB_ERR_INVALID, SB_ERR_UNSUPPORTED, SB_ERR_LIMIT = -1, -5, -8
PG_STANDARD, PG_DESCENT, PG_CONJUGATED, PG_APGD, PG_SPG = range(5)
PURE_NEWTON, PURE_ENUMERATIVE, HYBRID, REV_HYBRID = range(4)
LAW_SOC, LAW_COULOMB = 0, 1
OP_ASSEMBLED, OP_MATRIX_FREE = 0, 1
SCALING_REFERENCE, SCALING_DIAGONAL, SCALING_NONE = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
ADJOINT_UNROLL, ADJOINT_IMPLICIT = 0, 1
COLORING_REFERENCE, COLORING_GREEDY, COLORING_HOST, COLORING_DEVICE = 0, 1, 16, 32
BATCH_CALLER_ORDER = 2
BATCH_SPREAD_TEAMS = 4
BSR_SYMMETRIC = 1
PROBLEM_ASSEMBLE_W = 1
PROBLEM_HOST_SETUP = 8
BENCH_FLUSH_CACHES = 16

# Every extern "C" symbol of include/sobogus_hip.h
EXPORTED = [
    "sb_status_string", "sb_pg_options_default", "sb_gs_options_default", "sb_ctx_create", "sb_ctx_destroy",
    "sb_ctx_info", "sb_ctx_launch_counts", "sb_bsr_create", "sb_bsr_destroy", "sb_bsr_multiply", "sb_bsr_split_row_multiply",
    "sb_problem_create", "sb_problem_create_device", "sb_problem_layout", "sb_problem_set_w", "sb_problem_destroy", "sb_problem_sizes", "sb_problem_get_w",
    "sb_problem_apply_permutation", "sb_problem_coloring", "sb_problem_apply_w", "sb_fb_residual", "sb_pg_solve",
    "sb_pg_solve_noalloc",
    "sb_gs_solve", "sb_pgs_solve", "sb_pg_solve_cadoux", "sb_gs_solve_cadoux", "sb_problem_update_values",
    "sb_mass_inverse", "sb_device_alloc", "sb_device_free", "sb_memcpy_h2d", "sb_memcpy_d2h", "sb_synchronize",
    "sb_bench_spmv", "sb_bench_copy", "sb_last_phase_times", "sb_batch_create", "sb_batch_destroy",
    "sb_batch_sizes", "sb_batch_solve", "sb_batch_autotune", "sb_batch_set_iteration_hints",
    "sb_batch_get_iteration_hints", "sb_memset_device", "sb_adjoint_options_default",
    "sb_adjoint", "sb_problem_update", "sb_problem_create_from_w", "sb_merit1", "sb_local_problem_write",
    "sb_local_problem_read",
]


class PGOptions(C.Structure):
    _fields_ = [("variant", C.c_int), ("law", C.c_int), ("op", C.c_int), ("scaling", C.c_int),
                ("inf_norm", C.c_int), ("tol", C.c_double), ("max_iters", C.c_int), ("ls_iters", C.c_int),
                ("ls_optimistic", C.c_double), ("ls_pessimistic", C.c_double), ("ls_armijo", C.c_double),
                ("grid_blocks", C.c_int), ("block_threads", C.c_int), ("precision", C.c_int),
                ("mixed_tol_factor", C.c_double), ("mixed_stall_iters", C.c_int)]


class GSOptions(C.Structure):
    _fields_ = [("law", C.c_int), ("strategy", C.c_int), ("inf_norm", C.c_int), ("try_zero", C.c_int),
                ("tol", C.c_double), ("max_iters", C.c_int), ("eval_every", C.c_int), ("skip_tol", C.c_double),
                ("skip_iters", C.c_int), ("auto_regularization", C.c_double), ("grid_blocks", C.c_int),
                ("block_threads", C.c_int), ("precision", C.c_int), ("mixed_tol_factor", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("spmvs", C.c_int), ("ls_trials", C.c_int), ("nesterov_steps", C.c_int),
                ("converged", C.c_int), ("local_failures", C.c_int), ("residual", C.c_double),
                ("device_ms", C.c_double), ("history", C.POINTER(C.c_double)), ("history_capacity", C.c_int),
                ("history_length", C.c_int), ("refine_iterations", C.c_int), ("refine_spmvs", C.c_int),
                ("mixed_residual", C.c_double), ("plain_launch", C.c_int), ("mixed_device_ms", C.c_double)]


class AdjointOptions(C.Structure):
    _fields_ = [("mode", C.c_int), ("law", C.c_int), ("iterations", C.c_int), ("checkpoint_every", C.c_int),
                ("rho", C.c_double), ("tol", C.c_double), ("max_iters", C.c_int), ("grid_blocks", C.c_int),
                ("branch_mode", C.c_int), ("krylov_dim", C.c_int)]


class AdjointStats(C.Structure):
    _fields_ = [("rho", C.c_double), ("loss", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("change", C.c_double), ("device_ms", C.c_double)]


class BatchConfig(C.Structure):
    _fields_ = [("block_threads", C.c_int), ("blocks_per_cu", C.c_int), ("contacts_per_block", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("n_converged", C.c_int), ("total_iterations", C.c_int), ("max_iterations", C.c_int),
                ("refined_problems", C.c_int), ("total_spmvs", C.c_int64), ("device_ms", C.c_double),
                ("config", BatchConfig), ("iterations", C.POINTER(C.c_int)), ("residuals", C.POINTER(C.c_double)),
                ("converged", C.POINTER(C.c_int)), ("stage_spmvs", C.POINTER(C.c_int)),
                ("phase_us", C.POINTER(C.c_double))]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)
_vp = C.c_void_p

_lib = None


class SbError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: status {code} ({_status_string(code)})")
        self.code = code


def _status_string(code: int) -> str:
    try:
        return lib().sb_status_string(code).decode()
    except Exception:  # pragma: no cover - library absent
        return SB_ERR.get(code, "unknown")


def lib():
    """Load libsbhip.so; raises if it has not been built (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C so-bogus_amd)")
    L = C.CDLL(str(LIB_PATH))
    sig = {
        "sb_status_string": (C.c_char_p, [C.c_int]),
        "sb_pg_options_default": (None, [C.POINTER(PGOptions)]),
        "sb_gs_options_default": (None, [C.POINTER(GSOptions)]),
        "sb_ctx_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
        "sb_ctx_destroy": (C.c_int, [_vp]),
        "sb_ctx_info": (C.c_int, [_vp, C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
        "sb_ctx_launch_counts": (C.c_int, [_vp, _lp, _lp]),
        "sb_bsr_create": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _ip, _ip, _dp]),
        "sb_bsr_destroy": (C.c_int, [_vp]),
        "sb_bsr_multiply": (C.c_int, [_vp, _vp, C.c_int, C.c_double, _dp, C.c_double, _dp, C.c_int]),
        "sb_bsr_split_row_multiply": (C.c_int, [_vp, _vp, C.c_int, _dp, _dp]),
        "sb_problem_create_device": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp,
                                               C.c_int]),
        "sb_problem_layout": (C.c_int, [_vp, _vp, C.c_int, _lp, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _dp]),
        "sb_problem_create": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp,
                                        C.c_int]),
        "sb_problem_set_w": (C.c_int, [_vp, _vp, _ip, _ip, _dp]),
        "sb_problem_destroy": (C.c_int, [_vp]),
        "sb_problem_sizes": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), _lp, _lp]),
        "sb_problem_get_w": (C.c_int, [_vp, _vp, _ip, _ip, _dp]),
        "sb_problem_apply_permutation": (C.c_int, [_vp, _vp, _lp]),
        "sb_problem_coloring": (C.c_int, [_vp, _vp, C.c_int, _lp, _lp, C.POINTER(C.c_int)]),
        "sb_problem_apply_w": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, C.c_int]),
        "sb_fb_residual": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _dp, C.c_int]),
        "sb_pg_solve": (C.c_int, [_vp, _vp, C.POINTER(PGOptions), _vp, C.c_int, _dp, C.POINTER(Stats)]),
        "sb_pg_solve_noalloc": (C.c_int, [_vp, _vp, C.POINTER(PGOptions), _vp, _dp, C.POINTER(Stats)]),
        "sb_gs_solve": (C.c_int, [_vp, _vp, C.POINTER(GSOptions), _lp, C.c_int, _vp, C.c_int, _dp,
                                  C.POINTER(Stats)]),
        "sb_pgs_solve": (C.c_int, [_vp, _vp, C.POINTER(GSOptions), _lp, C.c_int, _vp, C.c_int, _dp,
                                   C.POINTER(Stats)]),
        "sb_problem_update_values": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int]),
        "sb_mass_inverse": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int]),
        "sb_pg_solve_cadoux": (C.c_int, [_vp, _vp, C.POINTER(PGOptions), C.c_int, C.c_double, C.c_int, _vp,
                                         C.c_int, _dp, C.POINTER(Stats)]),
        "sb_gs_solve_cadoux": (C.c_int, [_vp, _vp, C.POINTER(GSOptions), _lp, C.c_int, C.c_int, C.c_double,
                                         C.c_int, _vp, C.c_int, _dp, C.POINTER(Stats)]),
        "sb_device_alloc": (C.c_int, [_vp, C.POINTER(_vp), C.c_size_t]),
        "sb_device_free": (C.c_int, [_vp, _vp]),
        "sb_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
        "sb_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
        "sb_synchronize": (C.c_int, [_vp]),
        "sb_memset_device": (C.c_int, [_vp, _vp, C.c_int, C.c_size_t]),
        "sb_adjoint_options_default": (None, [C.POINTER(AdjointOptions)]),
        "sb_adjoint": (C.c_int, [_vp, _vp, C.POINTER(AdjointOptions), _dp, _dp, _dp, _dp, _dp, _dp, _vp,
                                 C.POINTER(AdjointStats)]),
        "sb_problem_update": (C.c_int, [_vp, _vp, _dp, _dp, _dp]),
        "sb_problem_create_from_w": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, _ip, _ip, _dp, _dp, _dp, C.c_double]),
        "sb_merit1": (C.c_int, [_vp, _vp, C.c_int, _vp, C.c_int, _dp]),
        "sb_local_problem_write": (C.c_int, [C.c_char_p, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]),
        "sb_local_problem_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                            C.POINTER(C.c_int), _ip, _ip, _dp, _dp, _dp, _dp]),
        "sb_bench_spmv": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _dp]),
        "sb_last_phase_times": (C.c_int, [_vp, _dp]),
        "sb_bench_copy": (C.c_int, [_vp, C.c_size_t, C.c_int, _dp]),
        "sb_batch_create": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(_ip), C.POINTER(_ip), C.POINTER(_dp), C.POINTER(_dp),
                                      C.POINTER(_dp), C.POINTER(_dp), C.c_int]),
        "sb_batch_destroy": (C.c_int, [_vp]),
        "sb_batch_sizes": (C.c_int, [_vp, C.POINTER(C.c_int), _lp]),
        "sb_batch_solve": (C.c_int, [_vp, _vp, C.POINTER(PGOptions), C.POINTER(BatchConfig), _vp, C.c_int,
                                     C.POINTER(BatchStats)]),
        "sb_batch_autotune": (C.c_int, [_vp, _vp, C.POINTER(PGOptions), C.c_int, C.POINTER(BatchConfig), _dp]),
        "sb_batch_set_iteration_hints": (C.c_int, [_vp, C.POINTER(C.c_int)]),
        "sb_batch_get_iteration_hints": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(code: int, what: str, allow_not_converged: bool = False) -> int:
    if code == SB_OK or (allow_not_converged and code == SB_NOT_CONVERGED):
        return code
    raise SbError(code, what)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _l(a):
    return a.ctypes.data_as(_lp)


class Context:
    """One HIP device and stream (sb_ctx)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().sb_ctx_create(C.byref(h), device), "sb_ctx_create")
        self.h = h
        name = C.create_string_buffer(128)
        ncu = C.c_int(0)
        lib().sb_ctx_info(self.h, name, 128, C.byref(ncu))
        self.arch = name.value.decode()
        self.n_cu = ncu.value

    def close(self):
        if getattr(self, "h", None):
            lib().sb_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().sb_synchronize(self.h), "sb_synchronize")

    def bench_copy(self, nbytes: int, reps: int = 10) -> float:
        """ms per launch of a plain device copy of nbytes (2 * nbytes of traffic)."""
        ms = C.c_double()
        _check(lib().sb_bench_copy(self.h, int(nbytes), reps, C.byref(ms)), "sb_bench_copy")
        return ms.value

    def launch_counts(self) -> dict:
        """Persistent-kernel launches since the context was created, by launch mode."""
        co, pl = C.c_int64(), C.c_int64()
        _check(lib().sb_ctx_launch_counts(self.h, C.byref(co), C.byref(pl)), "sb_ctx_launch_counts")
        return {"cooperative": co.value, "plain": pl.value}

    def last_phase_times(self) -> dict:
        """Phase-time breakdown (us) of the last projected-gradient solve."""
        out = np.zeros(4)
        _check(lib().sb_last_phase_times(self.h, _d(out)), "sb_last_phase_times")
        return dict(op_work=out[0], op_wait=out[1], update_work=out[2], update_wait=out[3])


class DeviceArray:
    """A vector resident in HBM (float64 by default; int32 for index arrays), for
    solves and set-ups without host round trips."""

    def __init__(self, ctx: Context, n: int, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.dtype.itemsize * self.n
        p = _vp()
        _check(lib().sb_device_alloc(ctx.h, C.byref(p), max(self.nbytes, 8)), "sb_device_alloc")
        self.ptr = p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(-1)
        assert a.size == self.n
        _check(lib().sb_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(_vp), self.nbytes), "h2d")
        return self

    def zero(self):
        _check(lib().sb_memset_device(self.ctx.h, self.ptr, 0, self.nbytes), "sb_memset_device")
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        _check(lib().sb_memcpy_d2h(self.ctx.h, out.ctypes.data_as(_vp), self.ptr, self.nbytes), "d2h")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().sb_device_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SparseBlockMatrix:
    """Device BSR matrix (bogus::SparseBlockMatrix) with multiply<Transpose>."""

    def __init__(self, ctx: Context, nbr, nbc, br, bc, rowptr, colidx, vals, symmetric=False):
        self.ctx, self.nbr, self.nbc, self.br, self.bc, self.symmetric = ctx, nbr, nbc, br, bc, symmetric
        rp, ci, v = _i32(rowptr), _i32(colidx), _f64(vals).reshape(-1)
        h = _vp()
        _check(lib().sb_bsr_create(ctx.h, C.byref(h), nbr, nbc, br, bc, BSR_SYMMETRIC if symmetric else 0, _i(rp),
                                   _i(ci), _d(v)), "sb_bsr_create")
        self.h = h

    def multiply(self, x, transpose=False, alpha=1.0, beta=0.0, res=None) -> np.ndarray:
        x = _f64(x)
        tr = transpose and not self.symmetric
        n_out = self.nbc * self.bc if tr else self.nbr * self.br
        out = np.zeros(n_out) if res is None else _f64(res).copy()
        _check(lib().sb_bsr_multiply(self.ctx.h, self.h, int(transpose), alpha, _d(x), beta, _d(out), 0),
               "sb_bsr_multiply")
        return out

    def split_row_multiply(self, row: int, x, res3=None) -> np.ndarray:
        out = np.zeros(3) if res3 is None else _f64(res3).copy()
        _check(lib().sb_bsr_split_row_multiply(self.ctx.h, self.h, row, _d(_f64(x)), _d(out)),
               "sb_bsr_split_row_multiply")
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().sb_bsr_destroy(self.h)
                self.h = None
        except Exception:
            pass


def pg_options(**kw) -> PGOptions:
    o = PGOptions()
    lib().sb_pg_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def gs_options(**kw) -> GSOptions:
    o = GSOptions()
    lib().sb_gs_options_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


@dataclass
class SolveResult:
    status: int
    residual: float
    r: np.ndarray | None
    iterations: int
    spmvs: int
    ls_trials: int
    nesterov_steps: int
    converged: bool
    local_failures: int
    device_ms: float
    history: np.ndarray = field(default_factory=lambda: np.zeros(0))
    refine_iterations: int = 0
    refine_spmvs: int = 0
    mixed_residual: float = 0.0
    plain_launch: bool = False
    mixed_device_ms: float = 0.0


def write_local_problem(path, rowptr, colidx, vals, q, mu, r0=None) -> None:
    """A local problem (W as the scalar CSR of a 3n x 3n matrix, q, mu, optional r0) in the
    library's file layout (sb_local_problem_write)."""
    rp, ci, v, qq, m = _i32(rowptr), _i32(colidx), _f64(vals), _f64(q), _f64(mu)
    n = m.size
    if rp.size != 3 * n + 1 or qq.size != 3 * n or ci.size < rp[-1] or v.size < rp[-1]:
        raise ValueError("write_local_problem: rowptr 3n + 1, q 3n, colidx and vals rowptr[-1] entries")
    r0p = None
    if r0 is not None:
        r0a = _f64(r0)
        if r0a.size != 3 * n:
            raise ValueError("write_local_problem: r0 has 3n entries")
        r0p = _d(r0a)
    _check(lib().sb_local_problem_write(str(path).encode(), n, _i(rp), _i(ci), _d(v), _d(qq), _d(m), r0p),
           "sb_local_problem_write")


def read_local_problem(path) -> dict:
    """A local problem written by write_local_problem (sb_local_problem_read): dict with n,
    rowptr, colidx, vals, q, mu and r0 (None when the file has none)."""
    n, nnz, has = C.c_int(), C.c_int64(), C.c_int()
    p = str(path).encode()
    _check(lib().sb_local_problem_read(p, C.byref(n), C.byref(nnz), C.byref(has), None, None, None, None, None,
                                       None), "sb_local_problem_read")
    nc, nz = n.value, nnz.value
    out = dict(n=nc, rowptr=np.zeros(3 * nc + 1, np.int32), colidx=np.zeros(nz, np.int32), vals=np.zeros(nz),
               q=np.zeros(3 * nc), mu=np.zeros(nc), r0=np.zeros(3 * nc) if has.value else None)
    _check(lib().sb_local_problem_read(p, C.byref(n), C.byref(nnz), C.byref(has), _i(out["rowptr"]),
                                       _i(out["colidx"]), _d(out["vals"]), _d(out["q"]), _d(out["mu"]),
                                       _d(out["r0"]) if has.value else None), "sb_local_problem_read")
    return out


class DualFrictionProblem:
    """u = W r + b with W = H Minv H^T on the device (sb_problem)."""

    def __init__(self, ctx: Context, prob, assemble: bool = True, host_setup: bool = False):
        self.ctx, self.nc, self.nb = ctx, prob.nc, prob.nb
        rp, ci, hv, mi, b, mu = (_i32(prob.rowptr), _i32(prob.colidx), _f64(prob.hvals).reshape(-1),
                                 _f64(prob.minv).reshape(-1), _f64(prob.b), _f64(prob.mu))
        h = _vp()
        flags = (PROBLEM_ASSEMBLE_W if assemble else 0) | (PROBLEM_HOST_SETUP if host_setup else 0)
        _check(lib().sb_problem_create(ctx.h, C.byref(h), prob.nc, prob.nb, _i(rp), _i(ci), _d(hv), _d(mi), _d(b),
                                       _d(mu), flags), "sb_problem_create")
        self.h = h
        self.assembled = assemble

    @classmethod
    def from_device(cls, ctx: Context, nc: int, nb: int, rowptr: "DeviceArray", colidx: "DeviceArray",
                    hvals: "DeviceArray", minv: "DeviceArray", b: "DeviceArray", mu: "DeviceArray",
                    assemble: bool = True) -> "DualFrictionProblem":
        """A problem from arrays already in HBM (sb_problem_create_device): int32 row
        pointer and columns, float64 values."""
        self = cls.__new__(cls)
        self.ctx, self.nc, self.nb = ctx, int(nc), int(nb)
        h = _vp()
        _check(lib().sb_problem_create_device(ctx.h, C.byref(h), self.nc, self.nb, rowptr.ptr, colidx.ptr, hvals.ptr,
                                              minv.ptr, b.ptr, mu.ptr, PROBLEM_ASSEMBLE_W if assemble else 0),
               "sb_problem_create_device")
        self.h = h
        self.assembled = assemble
        return self

    @classmethod
    def from_w(cls, ctx: Context, nc: int, rowptr, colidx, vals, b, mu,
               prune_tol: float = 0.0) -> "DualFrictionProblem":
        """A problem given by W itself (sb_problem_create_from_w): the scalar CSR of the
        3 nc x 3 nc matrix (int32 row pointer and columns, float64 values), converted to
        the 3x3 blocks of its lower triangle and pruned as the reference's FCLib path does;
        only the assembled operator exists for it."""
        self = cls.__new__(cls)
        self.ctx, self.nc, self.nb = ctx, int(nc), 0
        rp, ci, v, bb, m = _i32(rowptr), _i32(colidx), _f64(vals), _f64(b), _f64(mu)
        if rp.size != 3 * self.nc + 1 or bb.size != 3 * self.nc or m.size != self.nc:
            raise ValueError("from_w: rowptr has 3 nc + 1 entries, b 3 nc, mu nc")
        h = _vp()
        _check(lib().sb_problem_create_from_w(ctx.h, C.byref(h), self.nc, _i(rp), _i(ci), _d(v), _d(bb), _d(m),
                                              float(prune_tol)), "sb_problem_create_from_w")
        self.h = h
        self.assembled = True
        return self

    def merit1(self, r, op=OP_ASSEMBLED) -> float:
        """FCLib's merit1 of r (sb_merit1)."""
        r = _f64(r)
        out = C.c_double()
        _check(lib().sb_merit1(self.ctx.h, self.h, op, r.ctypes.data_as(_vp), 0, C.byref(out)), "sb_merit1")
        return out.value

    def layout(self, which: int) -> dict:
        """The tile layout of H (0), H^T (1) or W (2) as the solvers use it."""
        sz = np.zeros(6, np.int64)
        _check(lib().sb_problem_layout(self.ctx.h, self.h, which, sz.ctypes.data_as(_lp), None, None, None, None,
                                       None, None, None, None), "sb_problem_layout")
        nrows, nq, nt, nl, steps, pairs = (int(v) for v in sz)
        bs = 9 if which == 2 else 18
        out = dict(nrows=nrows, nblocks=nq, ntiles=nt, nlong=nl, scan_steps=steps, pairs=pairs,
                   tstart=np.zeros(nt + 1, np.int32), brow=np.zeros(nq, np.int32), bcol=np.zeros(nq, np.int32),
                   tlong=np.zeros(nt, np.int32), long_rows=np.zeros(max(3 * nl, 1), np.int32),
                   rowptr=np.zeros(nrows + 1, np.int32), src=np.zeros(nq, np.int32), vals=np.zeros(bs * nq))
        _check(lib().sb_problem_layout(self.ctx.h, self.h, which, sz.ctypes.data_as(_lp), _i(out["tstart"]),
                                       _i(out["brow"]), _i(out["bcol"]), _i(out["tlong"]), _i(out["long_rows"]),
                                       _i(out["rowptr"]), _i(out["src"]), _d(out["vals"])), "sb_problem_layout")
        out["long_rows"] = out["long_rows"][:3 * nl]
        return out

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().sb_problem_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def sizes(self):
        nc, nb = C.c_int(), C.c_int()
        hb, wb = C.c_int64(), C.c_int64()
        _check(lib().sb_problem_sizes(self.h, C.byref(nc), C.byref(nb), C.byref(hb), C.byref(wb)), "sizes")
        return nc.value, nb.value, hb.value, wb.value

    def w_lower(self):
        nc, _, _, wb = self.sizes()
        rp, ci, v = np.zeros(nc + 1, np.int32), np.zeros(wb, np.int32), np.zeros(wb * 9)
        _check(lib().sb_problem_get_w(self.ctx.h, self.h, _i(rp), _i(ci), _d(v)), "sb_problem_get_w")
        return rp, ci, v.reshape(wb, 9)

    def apply_w(self, x, op=OP_ASSEMBLED) -> np.ndarray:
        x = _f64(x)
        y = np.zeros(3 * self.nc)
        _check(lib().sb_problem_apply_w(self.ctx.h, self.h, op, x.ctypes.data_as(_vp), y.ctypes.data_as(_vp), 0),
               "sb_problem_apply_w")
        return y

    def eval_with(self, r, law=LAW_COULOMB, op=OP_ASSEMBLED, scaling=SCALING_REFERENCE, inf_norm=False) -> float:
        r = _f64(r)
        out = C.c_double()
        _check(lib().sb_fb_residual(self.ctx.h, self.h, law, op, scaling, int(inf_norm), r.ctypes.data_as(_vp),
                                    C.byref(out), 0), "sb_fb_residual")
        return out.value

    def coloring(self, mode=COLORING_GREEDY):
        perm = np.zeros(self.nc, np.int64)
        colors = np.zeros(self.nc + 1, np.int64)
        n = C.c_int()
        _check(lib().sb_problem_coloring(self.ctx.h, self.h, mode, _l(perm), _l(colors), C.byref(n)), "coloring")
        return perm, colors[:n.value + 1].copy()

    def apply_permutation(self, perm):
        perm = _i64(perm)
        _check(lib().sb_problem_apply_permutation(self.ctx.h, self.h, _l(perm)), "apply_permutation")

    def _run(self, fn, r, on_device, hist_cap):
        st = Stats()
        hist = np.zeros(max(hist_cap, 1))
        st.history = _d(hist)
        st.history_capacity = hist_cap
        res = C.c_double()
        if on_device:
            code = fn(r.ptr, 1, C.byref(res), C.byref(st))
            rout = None
        else:
            rout = _f64(r).copy()
            code = fn(rout.ctypes.data_as(_vp), 0, C.byref(res), C.byref(st))
        _check(code, "solve", allow_not_converged=True)
        return SolveResult(status=code, residual=res.value, r=rout, iterations=st.iterations, spmvs=st.spmvs,
                           ls_trials=st.ls_trials, nesterov_steps=st.nesterov_steps, converged=bool(st.converged),
                           local_failures=st.local_failures, device_ms=st.device_ms,
                           history=hist[:min(st.history_length, hist_cap)].copy(),
                           refine_iterations=st.refine_iterations, refine_spmvs=st.refine_spmvs,
                           mixed_residual=st.mixed_residual, plain_launch=bool(st.plain_launch),
                           mixed_device_ms=st.mixed_device_ms)

    def solve_pg(self, r0, opts: PGOptions | None = None, on_device=False, hist_cap=0, **kw) -> SolveResult:
        """ProjectedGradient::solve through DualFrictionProblem::solveWith(pg, r, staticProblem)."""
        o = opts if opts is not None else pg_options(**kw)
        fn = lambda rp, dev, res, st: lib().sb_pg_solve(self.ctx.h, self.h, C.byref(o), rp, dev, res, st)
        return self._run(fn, r0, on_device, hist_cap)

    def solve_pg_noalloc(self, r: "DeviceArray", opts: PGOptions | None = None, **kw) -> SolveResult:
        """sb_pg_solve_noalloc: the same solve on a device-resident r, allocating nothing
        (grid state in the library's fixed per-GPU array; no residual history)."""
        o = opts if opts is not None else pg_options(**kw)
        fn = lambda rp, dev, res, st: lib().sb_pg_solve_noalloc(self.ctx.h, self.h, C.byref(o), rp, res, st)
        return self._run(fn, r, True, 0)

    def solve_gs(self, r0, opts: GSOptions | None = None, colors=None, on_device=False, hist_cap=0,
                 **kw) -> SolveResult:
        """GaussSeidel::solve on color-contiguous contacts (DualFrictionProblem::solveWith(gs, r))."""
        o = opts if opts is not None else gs_options(**kw)
        if colors is None:
            cp, nco = None, 0
        else:
            colors = _i64(colors)
            cp, nco = _l(colors), len(colors) - 1
        fn = lambda rp, dev, res, st: lib().sb_gs_solve(self.ctx.h, self.h, C.byref(o), cp, nco, rp, dev, res, st)
        return self._run(fn, r0, on_device, hist_cap)

    def solve_pgs(self, r0, opts: GSOptions | None = None, colors=None, on_device=False, hist_cap=0,
                  **kw) -> SolveResult:
        """ProductGaussSeidel::solve, the matrix-free Gauss-Seidel (PrimalFrictionProblem::solveWith(pgs, r)):
        no assembled W needed; contacts color-contiguous as for solve_gs."""
        o = opts if opts is not None else gs_options(**kw)
        if colors is None:
            cp, nco = None, 0
        else:
            colors = _i64(colors)
            cp, nco = _l(colors), len(colors) - 1
        fn = lambda rp, dev, res, st: lib().sb_pgs_solve(self.ctx.h, self.h, C.byref(o), cp, nco, rp, dev, res, st)
        return self._run(fn, r0, on_device, hist_cap)

    def solve_cadoux(self, r0, solver="pg", opts=None, colors=None, iterations=10, tol_tighten=0.1,
                     velocity=False, on_device=False, hist_cap=0, **kw) -> SolveResult:
        """DualFrictionProblem::solveCadoux (velocity=False) / solveCadouxVel (True) with the
        ProjectedGradient (solver="pg") or colored GaussSeidel ("gs") inner solver. The
        result's iterations are fixed-point iterations and its history the residual
        before and after each."""
        vel = int(bool(velocity))
        if solver == "gs":
            o = opts if opts is not None else gs_options(**kw)
            if colors is None:
                cp, nco = None, 0
            else:
                colors = _i64(colors)
                cp, nco = _l(colors), len(colors) - 1
            fn = lambda rp, dev, res, st: lib().sb_gs_solve_cadoux(self.ctx.h, self.h, C.byref(o), cp, nco,
                                                                   iterations, tol_tighten, vel, rp, dev, res, st)
        else:
            o = opts if opts is not None else pg_options(**kw)
            fn = lambda rp, dev, res, st: lib().sb_pg_solve_cadoux(self.ctx.h, self.h, C.byref(o), iterations,
                                                                   tol_tighten, vel, rp, dev, res, st)
        return self._run(fn, r0, on_device, hist_cap)

    def update_values(self, hvals=None, minv=None, b=None, mu=None):
        """Device-side refresh for an unchanged pattern (sb_problem_update_values): new H
        block values (BSR order of this problem, 18 per block), Minv, b, mu; numpy arrays
        (host) or DeviceArray (device), all of one kind; None keeps a part."""
        parts = [hvals, minv, b, mu]
        dev = any(isinstance(v, DeviceArray) for v in parts if v is not None)
        if dev:
            ptrs = [None if v is None else v.ptr for v in parts]
        else:
            keep = [None if v is None else _f64(v).reshape(-1) for v in parts]
            ptrs = [None if v is None else v.ctypes.data_as(_vp) for v in keep]
        _check(lib().sb_problem_update_values(self.ctx.h, self.h, *ptrs, 1 if dev else 0), "sb_problem_update_values")

    def update(self, b=None, mu=None, minv=None):
        """Replace b, mu and/or Minv in place (sb_problem_update)."""
        keep = [None if v is None else _f64(v).reshape(-1) for v in (b, mu, minv)]
        ptr = [None if v is None else _d(v) for v in keep]
        _check(lib().sb_problem_update(self.ctx.h, self.h, *ptr), "sb_problem_update")

    def adjoint(self, r0, g=None, mode=ADJOINT_UNROLL, law=LAW_COULOMB, iterations=200, checkpoint_every=0, rho=0.0,
                tol=1e-10, max_iters=3000, record_branches=False, branches=None, krylov_dim=40) -> dict:
        """Forward map T^K(r0) (g None) or the gradients of L = g . r (sb_adjoint).
        record_branches: also return the projection branch of every contact and step;
        branches: replay them (the map then stays on one smooth piece)."""
        o = AdjointOptions()
        lib().sb_adjoint_options_default(C.byref(o))
        o.mode, o.law, o.iterations, o.checkpoint_every = mode, law, iterations, checkpoint_every
        o.rho, o.tol, o.max_iters, o.krylov_dim = rho, tol, max_iters, krylov_dim
        br = None
        if branches is not None:
            br = np.ascontiguousarray(branches, dtype=np.uint8)
            assert br.size == iterations * self.nc
            o.branch_mode = 2
        elif record_branches:
            br = np.zeros(iterations * self.nc, dtype=np.uint8)
            o.branch_mode = 1
        brp = None if br is None else br.ctypes.data_as(_vp)
        r0 = _f64(r0)
        rout = np.zeros(3 * self.nc)
        st = AdjointStats()
        out = {"r": rout}
        if g is None:
            code = lib().sb_adjoint(self.ctx.h, self.h, C.byref(o), _d(r0), None, _d(rout), None, None, None, brp,
                                    C.byref(st))
        else:
            g = _f64(g)
            gb, gm, gM = np.zeros(3 * self.nc), np.zeros(self.nc), np.zeros(36 * self.nb)
            code = lib().sb_adjoint(self.ctx.h, self.h, C.byref(o), _d(r0), _d(g), _d(rout), _d(gb), _d(gm), _d(gM),
                                    brp, C.byref(st))
            out.update(grad_b=gb, grad_mu=gm, grad_minv=gM.reshape(self.nb, 36))
        if record_branches and branches is None:
            out["branches"] = br.reshape(iterations, self.nc)
        _check(code, "sb_adjoint", allow_not_converged=True)
        out.update(status=code, rho=st.rho, loss=st.loss, iterations=st.iterations, converged=bool(st.converged),
                   change=st.change, device_ms=st.device_ms)
        return out

    def bench_spmv(self, kind: int, reps: int = 20) -> float:
        ms = C.c_double()
        _check(lib().sb_bench_spmv(self.ctx.h, self.h, kind, reps, C.byref(ms)), "sb_bench_spmv")
        return ms.value


@dataclass
class BatchResult:
    status: int
    r: np.ndarray | None            # concatenated results (None when solved on device)
    residuals: np.ndarray           # FP64 residual of each problem's result
    iterations: np.ndarray
    converged: np.ndarray
    n_converged: int
    total_iterations: int
    max_iterations: int
    refined_problems: int
    total_spmvs: int
    device_ms: float
    stage_spmvs: np.ndarray         # [n][2] operator applications: FP32 stage, FP64 stage
    config: tuple[int, int, int]
    phase_us: np.ndarray = field(default_factory=lambda: np.zeros(0))  # [n][stage][op work, op wait, update work, update wait]


class Batch:
    """Many independent dual friction problems solved in one launch (sb_batch)."""

    def __init__(self, ctx: Context, probs, assemble: bool = False, caller_order: bool = False,
                 spread_teams: bool = False):
        """caller_order: keep each problem in the given contact/body order instead of the
        library's locality order (results are returned in the caller's order either way).
        spread_teams: place each team's workgroups on consecutive blocks, over several XCDs
        (barriers with L2 write-back and invalidation; same results, slower)."""
        self.ctx = ctx
        self.n = len(probs)
        self.ncs = np.array([p.nc for p in probs], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(3 * self.ncs)])
        keep = []

        def arr(ctype, items):
            a = (ctype * self.n)(*items)
            keep.append(a)
            return a
        rps = [_i32(p.rowptr) for p in probs]
        cis = [_i32(p.colidx) for p in probs]
        hvs = [_f64(p.hvals).reshape(-1) for p in probs]
        mis = [_f64(p.minv).reshape(-1) for p in probs]
        bs = [_f64(p.b) for p in probs]
        mus = [_f64(p.mu) for p in probs]
        h = _vp()
        _check(lib().sb_batch_create(
            ctx.h, C.byref(h), self.n, arr(C.c_int, [p.nc for p in probs]), arr(C.c_int, [p.nb for p in probs]),
            arr(_ip, [_i(a) for a in rps]), arr(_ip, [_i(a) for a in cis]), arr(_dp, [_d(a) for a in hvs]),
            arr(_dp, [_d(a) for a in mis]), arr(_dp, [_d(a) for a in bs]), arr(_dp, [_d(a) for a in mus]),
            (PROBLEM_ASSEMBLE_W if assemble else 0) | (BATCH_CALLER_ORDER if caller_order else 0) |
            (BATCH_SPREAD_TEAMS if spread_teams else 0)),
            "sb_batch_create")
        self.h = h
        self.total_contacts = int(self.ncs.sum())

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().sb_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def split(self, r) -> list[np.ndarray]:
        return [r[self.offsets[k]:self.offsets[k + 1]] for k in range(self.n)]

    def solve(self, r0=None, opts: PGOptions | None = None, config=None, on_device=False, **kw) -> BatchResult:
        o = opts if opts is not None else pg_options(**kw)
        cfg = BatchConfig(*(config or (0, 0, 0)))
        its = np.zeros(self.n, dtype=np.int32)
        res = np.zeros(self.n)
        conv = np.zeros(self.n, dtype=np.int32)
        stage = np.zeros((self.n, 2), dtype=np.int32)
        phase = np.zeros((self.n, 2, 4))
        st = BatchStats()
        st.stage_spmvs = stage.ctypes.data_as(C.POINTER(C.c_int))
        st.phase_us = _d(phase)
        st.iterations = its.ctypes.data_as(C.POINTER(C.c_int))
        st.residuals = _d(res)
        st.converged = conv.ctypes.data_as(C.POINTER(C.c_int))
        if on_device:
            rp, rout = r0.ptr, None
        else:
            rout = np.zeros(3 * self.total_contacts) if r0 is None else _f64(r0).copy()
            rp = rout.ctypes.data_as(_vp)
        code = lib().sb_batch_solve(self.ctx.h, self.h, C.byref(o), C.byref(cfg), rp, 1 if on_device else 0,
                                    C.byref(st))
        _check(code, "sb_batch_solve", allow_not_converged=True)
        return BatchResult(status=code, r=rout, residuals=res, iterations=its, converged=conv.astype(bool),
                           n_converged=st.n_converged, total_iterations=st.total_iterations,
                           max_iterations=st.max_iterations, refined_problems=st.refined_problems,
                           total_spmvs=st.total_spmvs, device_ms=st.device_ms, stage_spmvs=stage, phase_us=phase,
                           config=(st.config.block_threads, st.config.blocks_per_cu,
                                   st.config.contacts_per_block))

    def set_iteration_hints(self, hints=None):
        """Expected iteration count of every problem (0 = unknown; None forgets them):
        the planner sizes teams and orders queues with them."""
        if hints is None:
            _check(lib().sb_batch_set_iteration_hints(self.h, None), "sb_batch_set_iteration_hints")
            return
        h = np.ascontiguousarray(hints, dtype=np.int32)
        if h.shape != (self.n,):
            raise ValueError(f"need {self.n} hints")
        _check(lib().sb_batch_set_iteration_hints(self.h, h.ctypes.data_as(C.POINTER(C.c_int))),
               "sb_batch_set_iteration_hints")

    def iteration_hints(self) -> np.ndarray:
        h = np.zeros(self.n, dtype=np.int32)
        _check(lib().sb_batch_get_iteration_hints(self.h, h.ctypes.data_as(C.POINTER(C.c_int))),
               "sb_batch_get_iteration_hints")
        return h

    def autotune(self, opts: PGOptions | None = None, probe_iters: int = 50, **kw):
        o = opts if opts is not None else pg_options(**kw)
        best = BatchConfig()
        ms = C.c_double()
        _check(lib().sb_batch_autotune(self.ctx.h, self.h, C.byref(o), probe_iters, C.byref(best), C.byref(ms)),
               "sb_batch_autotune")
        return (best.block_threads, best.blocks_per_cu, best.contacts_per_block), ms.value


def mass_inverse(ctx: Context, M) -> np.ndarray:
    """Minv_k = M_k^-1 for [n_bodies, 36] row-major 6x6 blocks, on the GPU (sb_mass_inverse)."""
    M = _f64(M).reshape(-1, 36)
    out = np.zeros_like(M)
    _check(lib().sb_mass_inverse(ctx.h, M.shape[0], M.ctypes.data_as(_vp), out.ctypes.data_as(_vp), 0),
           "sb_mass_inverse")
    return out


def gpu_available() -> bool:
    """True when the library loads and a device context can be created."""
    try:
        ctx = Context(0)
        ctx.close()
        return True
    except Exception:
        return False


def exported_symbols() -> list[str]:
    return list(EXPORTED)


def header_symbols() -> list[str]:
    import re
    text = HEADER.read_text()
    decl = re.findall(r"^\s*(?:int|void|const char\s*\*)\s*\*?\s*(sb_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(decl))


os.environ.setdefault("HSA_ENABLE_IPC_MODE_LEGACY", "0")
