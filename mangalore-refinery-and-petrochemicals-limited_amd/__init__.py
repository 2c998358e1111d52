"""MI355X PDHG engine for BharatOpt's first-order LP path (ctypes bindings).

The compute lives in ``libbo_hip.so`` (HIP kernels for gfx950 + the C ABI of
``include/bo_capi.h`` + the C++ drop-in classes).  This module only marshals
numpy arrays across that ABI.  There is no CPU fallback: if the library is
missing, importing ``lib()`` raises, and on a machine without a GPU every
compute entry point returns ``BO_GPU_UNAVAILABLE``, which is raised here.

The package directory name is not a Python identifier; load it with
``importlib`` (see ``__graft_entry__.load_package``).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import lpdata  # noqa: F401  (re-exported for callers)
from . import warmstart  # noqa: F401
from .lpdata import FlatModel, StdLP

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbo_hip.so")

STATUS_NAMES = ["SUCCESS", "GPU_UNAVAILABLE", "GPU_UNINITIALIZED", "OUT_OF_MEMORY",
                "INVALID_ARGUMENT", "EXECUTION_FAILED"]
FO_STATUS = ["OPTIMAL", "MAX_ITERATIONS", "NUMERICAL_FAILURE", "INFEASIBLE_DETECTED",
             "UNBOUNDED_DETECTED", "INTERRUPTED"]

P = C.c_void_p
_LIB = None


class BoError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = lib().bo_last_error().decode(errors="replace") if _LIB is not None else ""
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"{where}: {name} {msg}".strip())


class DeviceInfo(C.Structure):
    _fields_ = [("is_available", C.c_int), ("device_name", C.c_char * 256),
                ("total_memory_bytes", C.c_size_t), ("free_memory_bytes", C.c_size_t),
                ("compute_capability_major", C.c_int), ("compute_capability_minor", C.c_int),
                ("multiprocessor_count", C.c_int), ("status_message", C.c_char * 256)]


class Timing(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("h2d_time_ms", "kernel_time_ms", "d2h_time_ms",
                                          "sync_time_ms", "total_gpu_time_ms")]


class BoLp(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("nnz", C.c_longlong), ("row_ptr", P),
                ("col_idx", P), ("val", P), ("b", P), ("c", P), ("lb", P), ("ub", P), ("is_eq", P)]


class RefOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_longlong), ("feasibility_tolerance", C.c_double),
                ("primal_step_size", C.c_double), ("dual_step_size", C.c_double),
                ("adaptive_step_size", C.c_int), ("check_frequency", C.c_longlong)]


class RefResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_longlong), ("loop_iterations", C.c_longlong),
                ("constraint_violation", C.c_double), ("bound_violation", C.c_double),
                ("used_average", C.c_int), ("kernel_time_ms", C.c_double),
                ("transfer_time_ms", C.c_double), ("loop", C.c_int)]


class PdlpParams(C.Structure):
    _fields_ = [("max_iterations", C.c_longlong), ("eps_rel_kkt", C.c_double),
                ("ruiz_iterations", C.c_int), ("pock_chambolle", C.c_int), ("check_every", C.c_int),
                ("lanes_rows", C.c_int), ("lanes_cols", C.c_int), ("fp64_iterates", C.c_int),
                ("use_graph", C.c_int), ("time_limit_s", C.c_double), ("verbose", C.c_int),
                ("hot_columns", C.c_int), ("stall_iterations", C.c_longlong), ("fp64_polish", C.c_int),
                ("compact_tail", C.c_int), ("single_workgroup", C.c_int), ("eps_objective", C.c_double)]


class PdlpStats(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_longlong), ("restarts", C.c_int),
                ("used_average", C.c_int)] + [(k, C.c_double) for k in (
                    "rel_kkt", "rel_primal", "rel_dual", "rel_gap", "primal_obj", "dual_obj",
                    "primal_weight", "step_size", "setup_ms", "solve_ms", "spmv_csr_bytes",
                    "spmv_csc_bytes")] + [("lanes_rows", C.c_int), ("lanes_cols", C.c_int),
                                          ("hot_columns", C.c_int), ("fp64_polish_iterations", C.c_longlong),
                                          ("check_every", C.c_int), ("single_workgroup", C.c_int)]


class PdlpPassInfo(C.Structure):
    _fields_ = [("schedule", C.c_int), ("kernels", C.c_int), ("blocks", C.c_int), ("block_entries", C.c_int),
                ("tiles", C.c_int), ("slices", C.c_int), ("pairs", C.c_longlong), ("padded", C.c_longlong),
                ("gather_bytes", C.c_double), ("reduce_bytes", C.c_double), ("reduce_loads", C.c_int)]


class PdlpBatchStats(C.Structure):
    _fields_ = [("count", C.c_int), ("solved", C.c_int), ("max_iterations", C.c_longlong),
                ("sum_iterations", C.c_longlong), ("loop_steps", C.c_longlong), ("solve_ms", C.c_double),
                ("tiled", C.c_int), ("polished", C.c_int), ("tail_lps", C.c_int), ("tail_ms", C.c_double)]


class BoModel(C.Structure):
    _fields_ = [("num_vars", C.c_int), ("num_cons", C.c_int), ("sense", C.c_int),
                ("obj_offset", C.c_double), ("lb", P), ("ub", P), ("obj", P), ("var_type", P),
                ("con_sense", P), ("rhs", P), ("range_upper", P), ("term_ptr", P),
                ("term_var", P), ("term_val", P)]


class VerifyOptions(C.Structure):
    _fields_ = [("feasibility_tolerance", C.c_double), ("integrality_tolerance", C.c_double),
                ("objective_tolerance", C.c_double), ("check_objective", C.c_int)]


class VerifyResult(C.Structure):
    _fields_ = [("verified", C.c_int), ("has_reported_objective", C.c_int)] + [(k, C.c_double) for k in (
        "objective_reported", "objective_recomputed", "objective_difference", "maximum_constraint_violation",
        "maximum_lower_bound_violation", "maximum_upper_bound_violation", "maximum_integrality_violation")] + [
        (k, C.c_longlong) for k in ("violated_constraint_count", "violated_bound_count", "violated_integrality_count",
                                    "variable_count", "constraint_count")] + [("status", C.c_int), ("diagnostics", C.c_int)]


class CostEstimate(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("create_ms", C.c_double), ("fixed_ms", C.c_double),
                ("per_iteration_ms", C.c_double), ("iterations", C.c_longlong), ("loop", C.c_int),
                ("in_domain", C.c_int)]


class MpsInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("num_vars", C.c_int), ("num_cons", C.c_int), ("nonzeros", C.c_longlong),
                ("integer_vars", C.c_int), ("binary_vars", C.c_int), ("parse_time_ms", C.c_double),
                ("problem_name", C.c_char * 128), ("message", C.c_char * 256)]


MPS_STATUS = ["SUCCESS", "FILE_NOT_FOUND", "MISSING_ENDATA", "UNKNOWN_SECTION", "MALFORMED_ROW",
              "MALFORMED_COLUMN", "MALFORMED_BOUND", "UNCLOSED_INTEGER_MARKER", "INVALID_NUMERIC_VALUE"]


class FoOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_longlong), ("feasibility_tolerance", C.c_double),
                ("optimality_tolerance", C.c_double), ("primal_step_size", C.c_double),
                ("dual_step_size", C.c_double), ("adaptive_step_size", C.c_int),
                ("check_frequency", C.c_longlong), ("engine", C.c_int), ("eps_rel_kkt", C.c_double),
                ("cpu_backend", C.c_int), ("eps_objective", C.c_double)]


class FoResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_longlong)] + [(k, C.c_double) for k in (
        "objective_value", "primal_objective", "dual_objective", "primal_residual", "dual_residual",
        "constraint_violation", "bound_violation", "solve_time_ms", "transfer_time_ms",
        "kernel_time_ms")] + [("primal_len", C.c_longlong), ("dual_len", C.c_longlong),
                              ("message", C.c_char * 256), ("verify_status", C.c_int),
                              ("verify_max_violation", C.c_double)]


# every symbol include/bo_capi.h declares
EXPORTS = [
    "bo_last_error", "bo_device_init", "bo_buf_alloc", "bo_buf_free", "bo_buf_h2d", "bo_buf_d2h",
    "bo_spmv_csr_f64", "bo_vec_add_f64", "bo_vec_sub_f64", "bo_vec_scale_f64", "bo_axpy_f64",
    "bo_dot_f64", "bo_norm2_f64", "bo_pdhg_ref_solve", "bo_pdhg_ref_solve_noalloc", "bo_pdlp_default_params",
    "bo_pdlp_create",
    "bo_pdlp_set_warm_start", "bo_pdlp_solve", "bo_pdlp_get_solution", "bo_pdlp_destroy",
    "bo_pdlp_time_spmv", "bo_presolve_stats", "bo_first_order_solve", "bo_copy_kernel_calibrate",
    "bo_pdlp_get_pass_info", "bo_pdlp_time_kernels", "bo_pdlp_batch_create", "bo_pdlp_batch_set_warm_start",
    "bo_pdlp_batch_set_primal_weights",
    "bo_pdlp_batch_set_objective_rhs", "bo_pdlp_batch_solve", "bo_pdlp_batch_get_solution",
    "bo_pdlp_batch_time_kernels", "bo_pdlp_batch_destroy",
    "bo_pdlp_family_create", "bo_pdlp_family_set_lps", "bo_pdlp_family_set_warm_start",
    "bo_pdlp_family_set_primal_weights", "bo_pdlp_family_solve",
    "bo_pdlp_family_get_solutions", "bo_pdlp_family_time_passes", "bo_pdlp_family_destroy",
    "bo_mlp_create", "bo_mlp_num_params", "bo_mlp_get_params", "bo_mlp_set_params", "bo_mlp_forward",
    "bo_mlp_gradients", "bo_mlp_train_step", "bo_mlp_time", "bo_mlp_destroy", "bo_gemm_bf16",
    "bo_mps_read", "bo_mps_model", "bo_mps_destroy", "bo_verify_solution", "bo_pdhg_cost_estimate",
]


def lib():
    """Load libbo_hip.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.bo_last_error.restype = C.c_char_p
        L.bo_device_init.argtypes = [P]
        L.bo_buf_alloc.argtypes = [C.c_size_t, C.POINTER(P)]
        L.bo_buf_free.argtypes = [P]
        L.bo_buf_free.restype = None
        L.bo_buf_h2d.argtypes = [P, P, C.c_size_t]
        L.bo_buf_d2h.argtypes = [P, P, C.c_size_t]
        L.bo_spmv_csr_f64.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t, P, P, P, P, P, P]
        for f in ("bo_vec_add_f64", "bo_vec_sub_f64"):
            getattr(L, f).argtypes = [P, P, P, C.c_size_t, P]
        L.bo_vec_scale_f64.argtypes = [P, C.c_double, P, C.c_size_t, P]
        L.bo_axpy_f64.argtypes = [C.c_double, P, P, P, C.c_size_t, P]
        L.bo_dot_f64.argtypes = [P, P, P, C.c_size_t, P]
        L.bo_norm2_f64.argtypes = [P, P, C.c_size_t, P]
        L.bo_pdhg_ref_solve.argtypes = [P, P, P, P, P]
        L.bo_pdhg_ref_solve_noalloc.argtypes = [P, P, P, P, P]
        L.bo_pdlp_default_params.argtypes = [P]
        L.bo_pdlp_default_params.restype = None
        L.bo_pdlp_create.argtypes = [P, P, C.POINTER(P)]
        L.bo_pdlp_set_warm_start.argtypes = [P, P, P]
        L.bo_pdlp_solve.argtypes = [P, P]
        L.bo_pdlp_get_solution.argtypes = [P, P, P]
        L.bo_pdlp_destroy.argtypes = [P]
        L.bo_pdlp_destroy.restype = None
        L.bo_pdlp_time_spmv.argtypes = [P, C.c_int, P, P]
        L.bo_first_order_solve.argtypes = [P, P, P, P, P]
        L.bo_mps_read.argtypes = [C.c_char_p, C.c_char_p, P, C.POINTER(P)]
        L.bo_mps_model.argtypes = [P, P]
        L.bo_mps_destroy.argtypes = [P]
        L.bo_mps_destroy.restype = None
        L.bo_presolve_stats.argtypes = [P, P, P, P, P, P]
        L.bo_copy_kernel_calibrate.argtypes = [C.c_size_t, C.c_int, P]
        L.bo_verify_solution.argtypes = [P, P, C.c_longlong, P, P, P]
        L.bo_pdhg_cost_estimate.argtypes = [C.c_int, C.c_int, C.c_longlong, C.c_int, C.c_longlong, P]
        for name in EXPORTS:
            if name not in ("bo_last_error", "bo_buf_free", "bo_pdlp_destroy", "bo_pdlp_default_params",
                            "bo_pdlp_batch_destroy", "bo_pdlp_family_destroy", "bo_mlp_destroy",
                            "bo_mlp_num_params"):
                getattr(L, name).restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def _check(status, where):
    if status != 0:
        raise BoError(status, where)


def device_info():
    d = DeviceInfo()
    st = lib().bo_device_init(C.byref(d))
    return st, d


# ---------------------------------------------------------------- buffers
class DeviceBuffer:
    """Owning device allocation (float64 unless dtype says otherwise)."""

    def __init__(self, count, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.ptr = P()
        _check(lib().bo_buf_alloc(self.count * self.dtype.itemsize, C.byref(self.ptr)), "bo_buf_alloc")

    @classmethod
    def from_host(cls, arr, dtype=None):
        arr = np.ascontiguousarray(arr, dtype=dtype or np.asarray(arr).dtype)
        b = cls(arr.size, arr.dtype)
        b.upload(arr)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        _check(lib().bo_buf_h2d(self.ptr, _p(arr), arr.nbytes), "bo_buf_h2d")

    def download(self, count=None):
        out = np.empty(self.count if count is None else count, self.dtype)
        _check(lib().bo_buf_d2h(_p(out), self.ptr, out.nbytes), "bo_buf_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().bo_buf_free(self.ptr)
            self.ptr = P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def spmv_f64(lp: StdLP, x):
    """GpuBackend::spmv on a standard-form matrix (FP64, reference order)."""
    rp = DeviceBuffer.from_host(lp.row_ptr, np.int32)
    ci = DeviceBuffer.from_host(lp.col_idx if lp.nnz else np.zeros(1, np.int32), np.int32)
    v = DeviceBuffer.from_host(lp.val if lp.nnz else np.zeros(1), np.float64)
    dx = DeviceBuffer.from_host(np.asarray(x, np.float64))
    dy = DeviceBuffer(max(lp.m, 1))
    t = Timing()
    _check(lib().bo_spmv_csr_f64(lp.m, lp.n, lp.nnz, rp.ptr, ci.ptr, v.ptr, dx.ptr, dy.ptr,
                                 C.byref(t)), "bo_spmv_csr_f64")
    return dy.download(lp.m), t


# ---------------------------------------------------------------- LP marshalling
class _LpHolder:
    def __init__(self, lp: StdLP):
        self.a = dict(rp=np.ascontiguousarray(lp.row_ptr, np.int32),
                      ci=np.ascontiguousarray(lp.col_idx, np.int32),
                      v=np.ascontiguousarray(lp.val, np.float64),
                      b=np.ascontiguousarray(lp.b, np.float64),
                      c=np.ascontiguousarray(lp.c, np.float64),
                      lb=np.ascontiguousarray(lp.lb, np.float64),
                      ub=np.ascontiguousarray(lp.ub, np.float64),
                      eq=np.ascontiguousarray(lp.is_eq, np.uint8))
        a = self.a
        self.s = BoLp(lp.m, lp.n, lp.nnz, _p(a["rp"]), _p(a["ci"]), _p(a["v"]), _p(a["b"]),
                      _p(a["c"]), _p(a["lb"]), _p(a["ub"]), _p(a["eq"]))


@dataclass
class RefSolve:
    status: int
    iterations: int
    loop_iterations: int
    constraint_violation: float
    bound_violation: float
    used_average: bool
    kernel_time_ms: float
    x: np.ndarray
    y: np.ndarray
    loop: int = 0  # 0 kernel per pass, 1 one workgroup, 2 one workgroup with the matrix in LDS
    transfer_time_ms: float = 0.0  # host-side CSC build and upload


def pdhg_ref_solve(lp: StdLP, max_iterations=50000, feasibility_tolerance=1e-7,
                   primal_step_size=0.01, dual_step_size=0.01, adaptive_step_size=True,
                   check_frequency=100, noalloc=False) -> RefSolve:
    """The reference's PDHG loop on the GPU (FP64, bit-equal).  noalloc:
    through bo_pdhg_ref_solve_noalloc (no allocation; small LPs only)."""
    h = _LpHolder(lp)
    o = RefOptions(max_iterations, feasibility_tolerance, primal_step_size, dual_step_size,
                   1 if adaptive_step_size else 0, check_frequency)
    r = RefResult()
    x = np.zeros(max(lp.n, 1))
    y = np.zeros(max(lp.m, 1))
    name = "bo_pdhg_ref_solve_noalloc" if noalloc else "bo_pdhg_ref_solve"
    _check(getattr(lib(), name)(C.byref(h.s), C.byref(o), _p(x), _p(y), C.byref(r)), name)
    return RefSolve(r.status, r.iterations, r.loop_iterations, r.constraint_violation,
                    r.bound_violation, bool(r.used_average), r.kernel_time_ms, x[:lp.n], y[:lp.m], r.loop,
                    r.transfer_time_ms)


def default_pdlp_params(**kw) -> PdlpParams:
    p = PdlpParams()
    lib().bo_pdlp_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Pdlp:
    """Handle of one LP resident on the GPU (bo_pdlp_*)."""

    def __init__(self, lp: StdLP, **params):
        self.lp = lp
        self._h = _LpHolder(lp)
        self.params = default_pdlp_params(**params)
        self.handle = P()
        _check(lib().bo_pdlp_create(C.byref(self._h.s), C.byref(self.params), C.byref(self.handle)),
               "bo_pdlp_create")

    def set_warm_start(self, x0=None, y0=None):
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        y0 = None if y0 is None else np.ascontiguousarray(y0, np.float64)
        _check(lib().bo_pdlp_set_warm_start(self.handle, _p(x0), _p(y0)), "bo_pdlp_set_warm_start")

    def solve(self) -> dict:
        s = PdlpStats()
        _check(lib().bo_pdlp_solve(self.handle, C.byref(s)), "bo_pdlp_solve")
        return {k: getattr(s, k) for k, _ in PdlpStats._fields_}

    def solution(self):
        x = np.zeros(max(self.lp.n, 1))
        y = np.zeros(max(self.lp.m, 1))
        _check(lib().bo_pdlp_get_solution(self.handle, _p(x), _p(y)), "bo_pdlp_get_solution")
        return x[:self.lp.n], y[:self.lp.m]

    def pass_info(self, pass_):
        """Schedule and algorithmic bytes of pass 0 (CSR, dual step) or 1 (CSC, primal step)."""
        info = PdlpPassInfo()
        _check(lib().bo_pdlp_get_pass_info(self.handle, int(pass_), C.byref(info)), "bo_pdlp_get_pass_info")
        return {k: getattr(info, k) for k, _ in PdlpPassInfo._fields_}

    def time_kernels4(self, reps=20):
        """Per-launch ms of [CSC gather, CSC reduce, CSR gather, CSR reduce]."""
        ms = (C.c_double * 4)()
        _check(lib().bo_pdlp_time_kernels(self.handle, int(reps), ms), "bo_pdlp_time_kernels")
        return list(ms)

    def time_kernels(self, reps=20):
        a, b = C.c_double(), C.c_double()
        _check(lib().bo_pdlp_time_spmv(self.handle, reps, C.byref(a), C.byref(b)), "bo_pdlp_time_spmv")
        return a.value, b.value

    def close(self):
        if self.handle:
            lib().bo_pdlp_destroy(self.handle)
            self.handle = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PdlpBatch:
    """Handle of many LPs solved together on the GPU (bo_pdlp_batch_*)."""

    def __init__(self, lps, **params):
        self.lps = list(lps)
        self._h = [_LpHolder(lp) for lp in self.lps]
        arr = (BoLp * len(self.lps))(*[h.s for h in self._h])
        self._arr = arr
        self.params = default_pdlp_params(**params)
        self.handle = P()
        _check(lib().bo_pdlp_batch_create(arr, len(self.lps), C.byref(self.params), C.byref(self.handle)),
               "bo_pdlp_batch_create")

    def set_warm_start(self, xs=None, ys=None):
        """xs, ys: lists of per-LP vectors (original space), or None to clear."""
        if xs is None and ys is None:
            _check(lib().bo_pdlp_batch_set_warm_start(self.handle, None, None), "bo_pdlp_batch_set_warm_start")
            return
        x = np.ascontiguousarray(np.concatenate([np.asarray(v, np.float64) for v in xs]))
        y = np.ascontiguousarray(np.concatenate([np.asarray(v, np.float64) for v in ys]))
        _check(lib().bo_pdlp_batch_set_warm_start(self.handle, _p(x), _p(y)), "bo_pdlp_batch_set_warm_start")

    def time_kernels(self, reps=20):
        """ms and algorithmic bytes of [CSC gather, CSC reduce, CSR gather, CSR reduce], all LPs running."""
        ms, nb = np.zeros(4), np.zeros(4)
        _check(lib().bo_pdlp_batch_time_kernels(self.handle, int(reps), _p(ms), _p(nb)), "bo_pdlp_batch_time_kernels")
        return ms, nb

    def set_objective_rhs(self, cs=None, bs=None):
        """New c and / or b for every LP (lists in batch order; None keeps them)."""
        c = None if cs is None else np.ascontiguousarray(np.concatenate([np.asarray(v, np.float64) for v in cs]))
        b = None if bs is None else np.ascontiguousarray(np.concatenate([np.asarray(v, np.float64) for v in bs]))
        assert c is None or c.size == sum(lp.n for lp in self.lps)
        assert b is None or b.size == sum(lp.m for lp in self.lps)
        _check(lib().bo_pdlp_batch_set_objective_rhs(self.handle, _p(c), _p(b)), "bo_pdlp_batch_set_objective_rhs")

    def set_primal_weights(self, w=None):
        """Starting primal weight per LP (0 = default); None clears them."""
        if w is None:
            _check(lib().bo_pdlp_batch_set_primal_weights(self.handle, None), "bo_pdlp_batch_set_primal_weights")
            return
        w = np.ascontiguousarray(w, np.float64)
        assert w.size == len(self.lps)
        _check(lib().bo_pdlp_batch_set_primal_weights(self.handle, _p(w)), "bo_pdlp_batch_set_primal_weights")

    def solve(self):
        per = (PdlpStats * len(self.lps))()
        tot = PdlpBatchStats()
        _check(lib().bo_pdlp_batch_solve(self.handle, per, C.byref(tot)), "bo_pdlp_batch_solve")
        return ([{k: getattr(p, k) for k, _ in PdlpStats._fields_} for p in per],
                {k: getattr(tot, k) for k, _ in PdlpBatchStats._fields_})

    def solution(self, k):
        lp = self.lps[k]
        x = np.zeros(max(lp.n, 1))
        y = np.zeros(max(lp.m, 1))
        _check(lib().bo_pdlp_batch_get_solution(self.handle, int(k), _p(x), _p(y)), "bo_pdlp_batch_get_solution")
        return x[:lp.n], y[:lp.m]

    def close(self):
        if self.handle:
            lib().bo_pdlp_batch_destroy(self.handle)
            self.handle = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PdlpFamily:
    """Up to 64 LPs over one matrix (bo_pdlp_family_*): `base` gives A, bounds
    and row senses; set_lps swaps in the LPs' costs and right-hand sides."""

    def __init__(self, base: StdLP, capacity=64, **params):
        self.base = base
        self._h = _LpHolder(base)
        self.params = default_pdlp_params(**params)
        self.handle = P()
        _check(lib().bo_pdlp_family_create(C.byref(self._h.s), int(capacity), C.byref(self.params),
                                           C.byref(self.handle)), "bo_pdlp_family_create")
        self.count = capacity

    def set_lps(self, cs, bs):
        """Costs and right-hand sides of the next solves (sequences of rows)."""
        c = np.ascontiguousarray(np.asarray(cs, np.float64).reshape(len(cs), self.base.n))
        b = np.ascontiguousarray(np.asarray(bs, np.float64).reshape(len(bs), self.base.m))
        assert c.shape[0] == b.shape[0]
        _check(lib().bo_pdlp_family_set_lps(self.handle, c.shape[0], _p(c), _p(b)), "bo_pdlp_family_set_lps")
        self.count = c.shape[0]

    def set_warm_start(self, xs=None, ys=None):
        if xs is None and ys is None:
            _check(lib().bo_pdlp_family_set_warm_start(self.handle, None, None), "bo_pdlp_family_set_warm_start")
            return
        x = np.ascontiguousarray(np.asarray(xs, np.float64).reshape(self.count, self.base.n))
        y = np.ascontiguousarray(np.asarray(ys, np.float64).reshape(self.count, self.base.m))
        _check(lib().bo_pdlp_family_set_warm_start(self.handle, _p(x), _p(y)), "bo_pdlp_family_set_warm_start")

    def set_primal_weights(self, w=None):
        """Starting primal weights of the next solves (one per LP); None: each
        LP's own ||c_hat|| / ||b_hat||."""
        if w is None:
            _check(lib().bo_pdlp_family_set_primal_weights(self.handle, None), "bo_pdlp_family_set_primal_weights")
            return
        a = np.ascontiguousarray(np.asarray(w, np.float64).reshape(self.count))
        _check(lib().bo_pdlp_family_set_primal_weights(self.handle, _p(a)), "bo_pdlp_family_set_primal_weights")

    def solve(self):
        per = (PdlpStats * self.count)()
        tot = PdlpBatchStats()
        _check(lib().bo_pdlp_family_solve(self.handle, per, C.byref(tot)), "bo_pdlp_family_solve")
        return ([{k: getattr(p, k) for k, _ in PdlpStats._fields_} for p in per],
                {k: getattr(tot, k) for k, _ in PdlpBatchStats._fields_})

    def solutions(self):
        x = np.zeros((self.count, self.base.n))
        y = np.zeros((self.count, self.base.m))
        _check(lib().bo_pdlp_family_get_solutions(self.handle, _p(x), _p(y)), "bo_pdlp_family_get_solutions")
        return x, y

    def time_passes(self, reps=20):
        ms = np.zeros(2)
        _check(lib().bo_pdlp_family_time_passes(self.handle, int(reps), _p(ms)), "bo_pdlp_family_time_passes")
        return float(ms[0]), float(ms[1])

    def close(self):
        if self.handle:
            lib().bo_pdlp_family_destroy(self.handle)
            self.handle = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def f32_to_bf16_bits(a):
    """float32 -> bfloat16 bit patterns (round to nearest even), as the kernels do."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    nan = np.isnan(np.asarray(a, np.float32))
    if nan.any():
        r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r


def bf16_bits_to_f32(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def gemm_bf16(a_bits, b_bits, trans_a=False, trans_b=False):
    """C = op(A) op(B) on the matrix cores; A, B are 2-D uint16 BF16 bit arrays."""
    a_bits = np.ascontiguousarray(a_bits, np.uint16)
    b_bits = np.ascontiguousarray(b_bits, np.uint16)
    if trans_a:
        K, M = a_bits.shape
        sam, sak = 1, M
    else:
        M, K = a_bits.shape
        sam, sak = K, 1
    if trans_b:
        N, K2 = b_bits.shape
        sbk, sbn = 1, K2
    else:
        K2, N = b_bits.shape
        sbk, sbn = N, 1
    assert K2 == K
    c = np.zeros((M, N), np.float32)
    ll = C.c_longlong
    _check(lib().bo_gemm_bf16(M, N, K, a_bits.ctypes.data_as(P), ll(sam), ll(sak), b_bits.ctypes.data_as(P),
                              ll(sbk), ll(sbn), _p(c)), "bo_gemm_bf16")
    return c


class Mlp:
    """The warm-start network (bo_mlp_*)."""

    def __init__(self, n_in, width, n_out, layers=4, max_batch=64, seed=1):
        self.n_in, self.width, self.n_out, self.layers, self.max_batch = n_in, width, n_out, layers, max_batch
        self.handle = P()
        _check(lib().bo_mlp_create(n_in, width, n_out, layers, max_batch, C.c_ulonglong(seed),
                                   C.byref(self.handle)), "bo_mlp_create")
        lib().bo_mlp_num_params.restype = C.c_longlong
        self.nparam = int(lib().bo_mlp_num_params(self.handle))

    def shapes(self):
        dims = [self.n_in] + [self.width] * (self.layers - 1) + [self.n_out]
        return [(dims[l], dims[l + 1]) for l in range(self.layers)]

    def params(self):
        p = np.zeros(self.nparam, np.float32)
        _check(lib().bo_mlp_get_params(self.handle, _p(p)), "bo_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparam
        _check(lib().bo_mlp_set_params(self.handle, _p(p)), "bo_mlp_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        _check(lib().bo_mlp_forward(self.handle, _p(x), x.shape[0], _p(y)), "bo_mlp_forward")
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_double()
        _check(lib().bo_mlp_gradients(self.handle, _p(x), _p(t), x.shape[0], C.byref(loss), _p(g)),
               "bo_mlp_gradients")
        return loss.value, g

    def loss(self, x, t):
        """Mean squared error of forward(x) against t (no gradient copy)."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        _check(lib().bo_mlp_gradients(self.handle, _p(x), _p(t), x.shape[0], C.byref(loss), None),
               "bo_mlp_gradients")
        return loss.value

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        _check(lib().bo_mlp_train_step(self.handle, _p(x), _p(t), x.shape[0], C.c_float(lr), C.byref(loss)),
               "bo_mlp_train_step")
        return loss.value

    def time(self, what, batch, reps=10):
        """ms per forward (what 0) or forward + backward (what 1) pass on the device."""
        ms = C.c_double()
        _check(lib().bo_mlp_time(self.handle, int(what), int(batch), int(reps), C.byref(ms)), "bo_mlp_time")
        return ms.value

    def close(self):
        if self.handle:
            lib().bo_mlp_destroy(self.handle)
            self.handle = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _ModelHolder:
    def __init__(self, fm: FlatModel):
        f64 = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
        self.a = dict(lb=f64(fm.lb), ub=f64(fm.ub), obj=f64(fm.obj),
                      vt=None if fm.var_type is None else np.ascontiguousarray(fm.var_type, np.int32),
                      cs=np.ascontiguousarray(fm.con_sense, np.int32), rhs=f64(fm.rhs),
                      ru=None if fm.range_upper is None else f64(fm.range_upper),
                      tp=np.ascontiguousarray(fm.term_ptr, np.int64),
                      tv=np.ascontiguousarray(fm.term_var, np.int32), tw=f64(fm.term_val))
        a = self.a
        self.s = BoModel(fm.num_vars, fm.num_cons, int(fm.sense), float(fm.obj_offset),
                         _p(a["lb"]), _p(a["ub"]), _p(a["obj"]), _p(a["vt"]), _p(a["cs"]),
                         _p(a["rhs"]), _p(a["ru"]), _p(a["tp"]), _p(a["tv"]), _p(a["tw"]))


def verify_solution(fm: FlatModel, x, reported_objective=None, feasibility_tolerance=1e-7,
                    integrality_tolerance=1e-5, objective_tolerance=1e-4, check_objective=True):
    """SolutionVerifier::verify on the host (no GPU): a dict of the
    VerificationResult fields (status 0 verified, 1 failed, 2 objective
    mismatch, 3 dimension mismatch)."""
    h = _ModelHolder(fm)
    xx = np.ascontiguousarray(x, np.float64)
    o = VerifyOptions(feasibility_tolerance, integrality_tolerance, objective_tolerance, 1 if check_objective else 0)
    r = VerifyResult()
    rep = None if reported_objective is None else C.c_double(reported_objective)
    _check(lib().bo_verify_solution(C.byref(h.s), _p(xx) if xx.size else None, int(xx.size),
                                    None if rep is None else C.byref(rep), C.byref(o), C.byref(r)),
           "bo_verify_solution")
    return {k: getattr(r, k) for k, _ in VerifyResult._fields_}


def cost_estimate(m, n, nnz, engine=1, iterations=0):
    """Predicted ms of a GPU first-order solve (bo_pdhg_cost_estimate; host only)."""
    out = CostEstimate()
    _check(lib().bo_pdhg_cost_estimate(int(m), int(n), int(nnz), int(engine), int(iterations), C.byref(out)),
           "bo_pdhg_cost_estimate")
    return {k: getattr(out, k) for k, _ in CostEstimate._fields_}


def copy_kernel_ms(nbytes, reps=20):
    """Mean time of a plain copy kernel over nbytes (read + write)."""
    ms = C.c_double()
    _check(lib().bo_copy_kernel_calibrate(int(nbytes), int(reps), C.byref(ms)), "bo_copy_kernel_calibrate")
    return ms.value


def read_mps(text=None, path=None):
    """MpsParser::parse_string (text) or parse_file (path) through the C ABI.
    Returns (info dict, FlatModel or None when the text is not a model)."""
    L = lib()
    info = MpsInfo()
    h = P()
    st = L.bo_mps_read(None if text is None else text.encode(), None if path is None else path.encode(),
                       C.byref(info), C.byref(h))
    out = {k: getattr(info, k) for k, _ in MpsInfo._fields_}
    out["problem_name"] = info.problem_name.decode()
    out["message"] = info.message.decode()
    out["status_name"] = MPS_STATUS[info.status] if 0 <= info.status < len(MPS_STATUS) else str(info.status)
    if st != 0:
        raise BoError(st, "bo_mps_read: " + out["message"])
    if not h:
        return out, None
    try:
        v = BoModel()
        _check(L.bo_mps_model(h, C.byref(v)), "bo_mps_model")
        nv, nc = v.num_vars, v.num_cons

        def arr(ptr, ctype, count, dtype):
            if count == 0:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).astype(dtype, copy=True)

        tp = arr(v.term_ptr, C.c_longlong, nc + 1, np.int64)
        nnz = int(tp[-1]) if nc > 0 else 0
        fm = FlatModel(num_vars=nv, num_cons=nc, sense=v.sense, obj_offset=v.obj_offset,
                       lb=arr(v.lb, C.c_double, nv, np.float64), ub=arr(v.ub, C.c_double, nv, np.float64),
                       obj=arr(v.obj, C.c_double, nv, np.float64), var_type=arr(v.var_type, C.c_int, nv, np.int32),
                       con_sense=arr(v.con_sense, C.c_int, nc, np.int32), rhs=arr(v.rhs, C.c_double, nc, np.float64),
                       range_upper=arr(v.range_upper, C.c_double, nc, np.float64), term_ptr=tp,
                       term_var=arr(v.term_var, C.c_int, nnz, np.int32),
                       term_val=arr(v.term_val, C.c_double, nnz, np.float64))
    finally:
        L.bo_mps_destroy(h)
    return out, fm


def presolve_stats(fm: FlatModel):
    """PresolveEngine::presolve on the host (no GPU needed)."""
    h = _ModelHolder(fm)
    st, nv, nc, off, nz = C.c_int(), C.c_int(), C.c_int(), C.c_double(), C.c_longlong()
    _check(lib().bo_presolve_stats(C.byref(h.s), C.byref(st), C.byref(nv), C.byref(nc), C.byref(off),
                                   C.byref(nz)), "bo_presolve_stats")
    return dict(status=st.value, vars=nv.value, cons=nc.value, offset=off.value, nnz=nz.value)


def first_order_solve(fm: FlatModel, engine=0, max_iterations=50000, feasibility_tolerance=1e-7,
                      optimality_tolerance=1e-7, primal_step_size=0.01, dual_step_size=0.01,
                      adaptive_step_size=True, check_frequency=100, eps_rel_kkt=1e-6, backend="gpu",
                      eps_objective=0.0):
    """FirstOrderLPSolver::solve on a flattened LPModel.  backend "gpu":
    GPUFirstOrderSolver (engine 0 = reference PDHG semantics, 1 = PDLP);
    "cpu": CPUFirstOrderSolver, the reference's loop on the host (needs no
    GPU).  Returns a dict of the result fields."""
    if backend not in ("gpu", "cpu"):
        raise ValueError("backend must be 'gpu' or 'cpu'")
    h = _ModelHolder(fm)
    o = FoOptions(max_iterations, feasibility_tolerance, optimality_tolerance, primal_step_size,
                  dual_step_size, 1 if adaptive_step_size else 0, check_frequency, engine, eps_rel_kkt,
                  1 if backend == "cpu" else 0, eps_objective)
    r = FoResult()
    x = np.zeros(max(fm.num_vars, 1))
    y = np.zeros(max(fm.num_cons * 2, 1))
    st = lib().bo_first_order_solve(C.byref(h.s), C.byref(o), C.byref(r), _p(x), _p(y))
    out = {k: getattr(r, k) for k, _ in FoResult._fields_ if k != "message"}
    out["message"] = r.message.decode()
    out["x"] = x[:r.primal_len].copy()
    out["y"] = y[:r.dual_len].copy()
    out["capi_status"] = st
    if st not in (0,):
        raise BoError(st, "bo_first_order_solve: " + out["message"])
    return out
