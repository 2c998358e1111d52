"""ctypes bindings for the CPU checker libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` -- as the checker or the timed CPU baseline,
never as the measured or shipped path.

* ``liboracle.so``            -- FP64 C restatement (see oracle.h for the
  reference file:line each function follows).
* ``_ref/libbharatopt_ref.so`` -- the reference's own sources compiled by
  oracle/Makefile plus ref_shim.cpp; optional (absent when the reference
  sources were never available to build it).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_ORACLE = None
_REF = None

P = C.c_void_p


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def oracle_lib():
    global _ORACLE
    if _ORACLE is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        lib = C.CDLL(path)
        lib.orc_csr_from_coo.restype = C.c_long
        lib.orc_csr_from_coo.argtypes = [C.c_int, C.c_int, C.c_long, P, P, P, P, P, P]
        lib.orc_csr_spmv.argtypes = [C.c_int, P, P, P, P, P]
        lib.orc_csr_spmv_t.argtypes = [C.c_int, C.c_int, P, P, P, P, P]
        lib.orc_pdhg_ref.restype = C.c_int
        lib.orc_pdhg_ref.argtypes = [C.c_int, C.c_int] + [P] * 8 + [P, P, P, P]
        lib.orc_kkt.argtypes = [C.c_int, C.c_int] + [P] * 8 + [P, P, P]
        lib.orc_max_violation.argtypes = [C.c_int, C.c_int] + [P] * 7 + [P, P, P]
        _ORACLE = lib
    return _ORACLE


class OrcPdlpOpts(C.Structure):
    _fields_ = [("max_iterations", C.c_longlong), ("eps_rel_kkt", C.c_double), ("check_every", C.c_int),
                ("ruiz_iterations", C.c_int), ("pock_chambolle", C.c_int), ("threads", C.c_int),
                ("eps_objective", C.c_double), ("x_start", P), ("y_start", P), ("primal_weight_start", C.c_double),
                ("checkpoint_path", C.c_char_p), ("checkpoint_every", C.c_longlong)]


class OrcPdlpOut(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_longlong), ("restarts", C.c_int),
                ("rel_kkt", C.c_double), ("rel_primal", C.c_double), ("rel_dual", C.c_double),
                ("rel_gap", C.c_double), ("primal_obj", C.c_double), ("dual_obj", C.c_double),
                ("threads", C.c_int), ("step_size", C.c_double), ("primal_weight", C.c_double),
                ("prep_s", C.c_double), ("setup_s", C.c_double), ("solve_s", C.c_double)]


_PDLP_OMP = None


def pdlp_omp_lib():
    global _PDLP_OMP
    if _PDLP_OMP is None:
        path = os.path.join(HERE, "libpdlp_omp.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run make -C oracle)")
        lib = C.CDLL(path)
        lib.orc_pdlp_omp.restype = C.c_int
        lib.orc_pdlp_omp.argtypes = [C.c_int, C.c_int] + [P] * 8 + [P, P, P, P]
        _PDLP_OMP = lib
    return _PDLP_OMP


def pdlp_omp(lp, max_iterations=100000, eps_rel_kkt=1e-6, check_every=64, ruiz_iterations=10,
             pock_chambolle=True, threads=0, eps_objective=0.0, x_start=None, y_start=None, primal_weight_start=0.0,
             checkpoint_path=None, checkpoint_every=0):
    """The restarted Halpern PDHG in FP64 on `threads` host cores (0: all),
    from zero or from (x_start, y_start) with an optional starting primal
    weight.  checkpoint_path / checkpoint_every: the iteration state is
    written there every so many iterations, and a call that finds a matching
    file continues that run.  Returns (stats dict, x, y)."""
    lib = pdlp_omp_lib()
    xs = None if x_start is None else _f64(x_start)
    ys = None if y_start is None else _f64(y_start)
    o = OrcPdlpOpts(max_iterations, eps_rel_kkt, check_every, ruiz_iterations, 1 if pock_chambolle else 0,
                    threads, eps_objective, _p(xs) if xs is not None else None, _p(ys) if ys is not None else None,
                    float(primal_weight_start), checkpoint_path.encode() if checkpoint_path else None,
                    int(checkpoint_every))
    out = OrcPdlpOut()
    x = np.zeros(max(lp.n, 1))
    y = np.zeros(max(lp.m, 1))
    eq = np.ascontiguousarray(lp.is_eq, dtype=np.uint8)
    lib.orc_pdlp_omp(lp.m, lp.n, _p(_i32(lp.row_ptr)), _p(_i32(lp.col_idx)), _p(_f64(lp.val)),
                     _p(_f64(lp.b)), _p(_f64(lp.c)), _p(_f64(lp.lb)), _p(_f64(lp.ub)), _p(eq),
                     C.byref(o), _p(x), _p(y), C.byref(out))
    return {k: getattr(out, k) for k, _ in OrcPdlpOut._fields_}, x[:lp.n], y[:lp.m]


def ref_mps(text):
    """The reference's MpsParser::parse_string on `text`, flattened like
    bo_model: dict(status, sense, obj_offset, lb, ub, obj, var_type, con_sense,
    rhs, range_upper, term_ptr, term_var, term_val); None without oracle/_ref."""
    lib = ref_lib()
    if lib is None:
        return None
    raw = text.encode()
    dims = np.zeros(5, np.int64)
    off = C.c_double()
    if lib.ref_mps_parse(raw, _p(dims), C.byref(off), -1, *([None] * 10)) != 0:
        raise RuntimeError("reference MpsParser raised")
    nv, nc, nt = int(dims[1]), int(dims[2]), int(dims[3])
    a = dict(lb=np.zeros(max(nv, 1)), ub=np.zeros(max(nv, 1)), obj=np.zeros(max(nv, 1)),
             var_type=np.zeros(max(nv, 1), np.int32), con_sense=np.zeros(max(nc, 1), np.int32),
             rhs=np.zeros(max(nc, 1)), range_upper=np.zeros(max(nc, 1)), term_ptr=np.zeros(nc + 1, np.int64),
             term_var=np.zeros(max(nt, 1), np.int32), term_val=np.zeros(max(nt, 1)))
    rc = lib.ref_mps_parse(raw, _p(dims), C.byref(off), nt, _p(a["lb"]), _p(a["ub"]), _p(a["obj"]),
                           _p(a["var_type"]), _p(a["con_sense"]), _p(a["rhs"]), _p(a["range_upper"]),
                           _p(a["term_ptr"]), _p(a["term_var"]), _p(a["term_val"]))
    if rc != 0:
        raise RuntimeError("reference MpsParser raised")
    out = {k: v[:{"lb": nv, "ub": nv, "obj": nv, "var_type": nv, "con_sense": nc, "rhs": nc,
                  "range_upper": nc, "term_ptr": nc + 1, "term_var": nt, "term_val": nt}[k]] for k, v in a.items()}
    out.update(status=int(dims[0]), sense=int(dims[4]), obj_offset=off.value)
    return out


def ref_lib():
    """The reference build, or None when it is not present."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libbharatopt_ref.so")
        if not os.path.exists(path):
            return None
        lib = C.CDLL(path)
        lib.ref_fo_solve.restype = C.c_int
        lib.ref_fo_solve.argtypes = [P, P, P, P, P]
        lib.ref_presolve_dims.restype = C.c_int
        lib.ref_presolve_dims.argtypes = [P, P, P, P, P, P]
        lib.ref_simplex_solve.restype = C.c_int
        lib.ref_simplex_solve.argtypes = [P, P, P, P, P]
        lib.ref_verify.restype = C.c_int
        lib.ref_verify.argtypes = [P, P, C.c_double, P, P, P, P]
        lib.ref_verify_full.restype = C.c_int
        lib.ref_verify_full.argtypes = [P, P, C.c_longlong, P, C.c_double, C.c_double, C.c_double, C.c_int, P]
        lib.ref_csr_spmv.restype = C.c_longlong
        lib.ref_csr_spmv.argtypes = [C.c_int, C.c_int, C.c_longlong, P, P, P, P, P, P, P]
        lib.ref_mps_parse.restype = C.c_int
        lib.ref_mps_parse.argtypes = [C.c_char_p, P, P, C.c_longlong] + [P] * 10
        _REF = lib
    return _REF


# ---------------------------------------------------------------- structs
class OrcPdhgOpts(C.Structure):
    _fields_ = [("max_iterations", C.c_long), ("feasibility_tol", C.c_double),
                ("primal_step", C.c_double), ("dual_step", C.c_double),
                ("adaptive", C.c_int), ("check_frequency", C.c_long)]


class OrcPdhgOut(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_long), ("loop_iterations", C.c_long),
                ("constraint_violation", C.c_double), ("bound_violation", C.c_double),
                ("used_average", C.c_int)]


class OrcKkt(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "r_primal", "r_dual", "primal_obj", "dual_obj", "gap_rel", "b_norm", "c_norm",
        "rel_primal", "rel_dual", "rel_kkt", "max_row_viol", "max_bound_viol")]


class RefModel(C.Structure):
    _fields_ = [("num_vars", C.c_int), ("num_cons", C.c_int), ("sense", C.c_int),
                ("obj_offset", C.c_double), ("lb", P), ("ub", P), ("obj", P), ("var_type", P),
                ("con_sense", P), ("rhs", P), ("range_upper", P), ("term_ptr", P),
                ("term_var", P), ("term_val", P)]


class RefFoOptions(C.Structure):
    _fields_ = [("max_iterations", C.c_longlong), ("feasibility_tolerance", C.c_double),
                ("optimality_tolerance", C.c_double), ("primal_step_size", C.c_double),
                ("dual_step_size", C.c_double), ("adaptive_step_size", C.c_int),
                ("backend_gpu", C.c_int), ("check_frequency", C.c_longlong)]


class RefFoResult(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_longlong),
                ("objective_value", C.c_double), ("constraint_violation", C.c_double),
                ("bound_violation", C.c_double), ("solve_time_ms", C.c_double),
                ("primal_len", C.c_longlong), ("dual_len", C.c_longlong),
                ("message", C.c_char * 256)]


# ---------------------------------------------------------------- helpers
def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def csr_from_coo(rows, cols, ri, ci, v):
    lib = oracle_lib()
    ri, ci, v = _i32(ri), _i32(ci), _f64(v)
    rp = np.zeros(rows + 1, np.int32)
    oc = np.zeros(max(len(v), 1), np.int32)
    ov = np.zeros(max(len(v), 1), np.float64)
    nnz = lib.orc_csr_from_coo(rows, cols, len(v), _p(ri), _p(ci), _p(v), _p(rp), _p(oc), _p(ov))
    if nnz < 0:
        raise IndexError("COO index out of range")
    return rp, oc[:nnz].copy(), ov[:nnz].copy()


def spmv(rp, ci, v, x):
    lib = oracle_lib()
    rows = len(rp) - 1
    y = np.zeros(max(rows, 1))
    x = _f64(x)
    lib.orc_csr_spmv(rows, _p(_i32(rp)), _p(_i32(ci)), _p(_f64(v)), _p(x), _p(y))
    return y[:rows]


def spmv_t(rp, ci, v, x, cols):
    lib = oracle_lib()
    rows = len(rp) - 1
    y = np.zeros(max(cols, 1))
    x = _f64(x)
    lib.orc_csr_spmv_t(rows, cols, _p(_i32(rp)), _p(_i32(ci)), _p(_f64(v)), _p(x), _p(y))
    return y[:cols]


@dataclass
class PdhgRefResult:
    status: int
    iterations: int
    loop_iterations: int
    constraint_violation: float
    bound_violation: float
    used_average: bool
    x: np.ndarray
    y: np.ndarray


def pdhg_ref(lp, max_iterations=50000, feasibility_tol=1e-7, primal_step=0.01,
             dual_step=0.01, adaptive=True, check_frequency=100):
    """Reference PDHG on a standard-form LP object (attributes m, n, row_ptr,
    col_idx, val, b, c, lb, ub, is_eq)."""
    lib = oracle_lib()
    o = OrcPdhgOpts(max_iterations, feasibility_tol, primal_step, dual_step,
                    1 if adaptive else 0, check_frequency)
    out = OrcPdhgOut()
    x = np.zeros(max(lp.n, 1))
    y = np.zeros(max(lp.m, 1))
    eq = np.ascontiguousarray(lp.is_eq, dtype=np.uint8)
    lib.orc_pdhg_ref(lp.m, lp.n, _p(_i32(lp.row_ptr)), _p(_i32(lp.col_idx)), _p(_f64(lp.val)),
                     _p(_f64(lp.b)), _p(_f64(lp.c)), _p(_f64(lp.lb)), _p(_f64(lp.ub)), _p(eq),
                     C.byref(o), _p(x), _p(y), C.byref(out))
    return PdhgRefResult(out.status, out.iterations, out.loop_iterations,
                         out.constraint_violation, out.bound_violation, bool(out.used_average),
                         x[:lp.n], y[:lp.m])


def kkt(lp, x, y):
    lib = oracle_lib()
    out = OrcKkt()
    eq = np.ascontiguousarray(lp.is_eq, dtype=np.uint8)
    xx, yy = _f64(x), _f64(y)
    lib.orc_kkt(lp.m, lp.n, _p(_i32(lp.row_ptr)), _p(_i32(lp.col_idx)), _p(_f64(lp.val)),
                _p(_f64(lp.b)), _p(_f64(lp.c)), _p(_f64(lp.lb)), _p(_f64(lp.ub)), _p(eq),
                _p(xx), _p(yy), C.byref(out))
    return {k: getattr(out, k) for k, _ in OrcKkt._fields_}


# ---------------------------------------------------------------- reference
class _RefModelHolder:
    """Keeps the numpy arrays behind a RefModel alive."""

    def __init__(self, fm):
        self.arrs = dict(
            lb=_f64(fm.lb), ub=_f64(fm.ub), obj=_f64(fm.obj),
            var_type=None if getattr(fm, "var_type", None) is None else _i32(fm.var_type),
            con_sense=_i32(fm.con_sense), rhs=_f64(fm.rhs),
            range_upper=None if getattr(fm, "range_upper", None) is None else _f64(fm.range_upper),
            term_ptr=np.ascontiguousarray(fm.term_ptr, dtype=np.int64),
            term_var=_i32(fm.term_var), term_val=_f64(fm.term_val))
        a = self.arrs
        self.s = RefModel(fm.num_vars, fm.num_cons, int(fm.sense), float(fm.obj_offset),
                          _p(a["lb"]), _p(a["ub"]), _p(a["obj"]), _p(a["var_type"]),
                          _p(a["con_sense"]), _p(a["rhs"]), _p(a["range_upper"]),
                          _p(a["term_ptr"]), _p(a["term_var"]), _p(a["term_val"]))


def ref_fo_solve(fm, max_iterations=50000, feasibility_tolerance=1e-7, optimality_tolerance=1e-7,
                 primal_step_size=0.01, dual_step_size=0.01, adaptive_step_size=True,
                 backend_gpu=False, check_frequency=100):
    lib = ref_lib()
    if lib is None:
        return None
    h = _RefModelHolder(fm)
    o = RefFoOptions(max_iterations, feasibility_tolerance, optimality_tolerance,
                     primal_step_size, dual_step_size, 1 if adaptive_step_size else 0,
                     1 if backend_gpu else 0, check_frequency)
    r = RefFoResult()
    x = np.zeros(max(fm.num_vars, 1))
    y = np.zeros(max(fm.num_cons, 1))
    rc = lib.ref_fo_solve(C.byref(h.s), C.byref(o), C.byref(r), _p(x), _p(y))
    if rc != 0:
        raise RuntimeError(r.message.decode())
    return dict(status=r.status, iterations=r.iterations, objective=r.objective_value,
                constraint_violation=r.constraint_violation, bound_violation=r.bound_violation,
                solve_time_ms=r.solve_time_ms, x=x[:r.primal_len].copy(),
                y=y[:r.dual_len].copy(), message=r.message.decode())


def ref_presolve_dims(fm):
    lib = ref_lib()
    if lib is None:
        return None
    h = _RefModelHolder(fm)
    st, nv, nc, off, nz = C.c_int(), C.c_int(), C.c_int(), C.c_double(), C.c_longlong()
    rc = lib.ref_presolve_dims(C.byref(h.s), C.byref(st), C.byref(nv), C.byref(nc), C.byref(off),
                               C.byref(nz))
    if rc != 0:
        raise RuntimeError("reference presolve raised")
    return dict(status=st.value, vars=nv.value, cons=nc.value, offset=off.value, nnz=nz.value)


def ref_simplex(fm):
    lib = ref_lib()
    if lib is None:
        return None
    h = _RefModelHolder(fm)
    x = np.zeros(max(fm.num_vars, 1))
    obj, stage, it = C.c_double(), C.c_int(), C.c_longlong()
    st = lib.ref_simplex_solve(C.byref(h.s), _p(x), C.byref(obj), C.byref(stage), C.byref(it))
    return dict(status=st, objective=obj.value, x=x[:fm.num_vars], stage=stage.value,
                iterations=it.value)


def ref_verify(fm, x, tol=1e-7):
    lib = ref_lib()
    if lib is None:
        return None
    h = _RefModelHolder(fm)
    xx = _f64(x)
    a, b, c, d = C.c_double(), C.c_double(), C.c_double(), C.c_double()
    ok = lib.ref_verify(C.byref(h.s), _p(xx), tol, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
    return dict(verified=ok == 1, max_cons_viol=a.value, max_lb_viol=b.value,
                max_ub_viol=c.value, objective=d.value)


class RefVerifyOut(C.Structure):
    _fields_ = [("verified", C.c_int), ("has_reported_objective", C.c_int)] + [(k, C.c_double) for k in (
        "objective_reported", "objective_recomputed", "objective_difference", "maximum_constraint_violation",
        "maximum_lower_bound_violation", "maximum_upper_bound_violation", "maximum_integrality_violation")] + [
        (k, C.c_longlong) for k in ("violated_constraint_count", "violated_bound_count", "violated_integrality_count",
                                    "variable_count", "constraint_count")] + [("status", C.c_int), ("diagnostics", C.c_int)]


def ref_verify_full(fm, x, reported_objective=None, feasibility_tolerance=1e-7, integrality_tolerance=1e-5,
                    objective_tolerance=1e-4, check_objective=True):
    """The reference's SolutionVerifier::verify, every VerificationResult
    field (same keys as the package's verify_solution); None without
    oracle/_ref."""
    lib = ref_lib()
    if lib is None:
        return None
    h = _RefModelHolder(fm)
    xx = _f64(x)
    out = RefVerifyOut()
    rep = None if reported_objective is None else C.c_double(reported_objective)
    rc = lib.ref_verify_full(C.byref(h.s), _p(xx) if xx.size else None, int(xx.size),
                             None if rep is None else C.byref(rep), feasibility_tolerance, integrality_tolerance,
                             objective_tolerance, 1 if check_objective else 0, C.byref(out))
    if rc != 0:
        raise RuntimeError("ref_verify_full failed")
    return {k: getattr(out, k) for k, _ in RefVerifyOut._fields_}


def ref_spmv(rows, cols, ri, ci, v, x=None, xt=None):
    lib = ref_lib()
    if lib is None:
        return None
    ri, ci, v = _i32(ri), _i32(ci), _f64(v)
    y = np.zeros(max(rows, 1)) if x is not None else None
    yt = np.zeros(max(cols, 1)) if xt is not None else None
    xx = None if x is None else _f64(x)
    xxt = None if xt is None else _f64(xt)
    nnz = lib.ref_csr_spmv(rows, cols, len(v), _p(ri), _p(ci), _p(v), _p(xx), _p(xxt), _p(y), _p(yt))
    if nnz < 0:
        raise IndexError("reference raised")
    return nnz, (None if y is None else y[:rows]), (None if yt is None else yt[:cols])
