"""Writes the golden fixtures of tests/golden/ (run in the build container,
where oracle/_ref -- the reference's own sources compiled by oracle/Makefile --
and scipy's HiGHS are available):

* ref_pdhg_kat.json   CPUFirstOrderSolver results of the reference build on
                      seeded models: status, iteration count, objective,
                      constraint violation and SHA-256 of the raw FP64 bytes of
                      the primal and dual vectors (bit-exact pins), plus a
                      SHA-256 of the model arrays so a changed generator is
                      detected instead of silently compared.
* lp_optima.json      optimal objectives of the PDLP test instances from two
                      independent exact solvers: the reference's simplex
                      chain (dual -> primal Dantzig -> primal Bland) where it
                      is tractable, and HiGHS (scipy.optimize.linprog).
* mps_golden.json     for every MPS file under mps/ (the reference's
                      conformance set, example1 and the refinery demo, copied
                      as data): the reference MpsParser's model (status,
                      sizes, SHA-256 of the flattened arrays) and the LP
                      optimum of that model by HiGHS (integrality relaxed,
                      objective in the model's sense with its offset).

Usage:  python tests/golden/make_golden.py        (ref_pdhg_kat, lp_optima)
        python tests/golden/make_golden.py kat    (ref_pdhg_kat only)
        python tests/golden/make_golden.py verifier  (verifier_kat: the reference SolutionVerifier's
                                                      results on candidates that break each rule)
        python tests/golden/make_golden.py mps    (mps_golden)
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

pkg = __graft_entry__.load_package()
lpd = pkg.lpdata


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        if a is None:
            h.update(b"none")
        else:
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def model_sha(fm):
    return sha(fm.lb, fm.ub, fm.obj, fm.con_sense, fm.rhs, fm.range_upper, fm.term_ptr,
               fm.term_var, fm.term_val, np.array([fm.sense, fm.num_vars, fm.num_cons]))


def kat_models():
    """name -> (FlatModel, solver kwargs)"""
    out = {}
    out["benchmark10"] = (lpd.benchmark10_model(), {})
    for s in (42, 43, 44):
        out[f"random_300x1000_s{s}_it2000"] = (lpd.random_lp(300, 1000, 5, s).to_flat(),
                                              dict(max_iterations=2000))
    out["random_60x200_s7_converge"] = (lpd.random_lp(60, 200, 5, 7).to_flat(), {})
    out["staircase_2x60x150_s5"] = (lpd.staircase_lp(2, 60, 150, 10, seed=5).to_flat(),
                                    dict(max_iterations=3000))
    mb = lpd.ModelBuilder(sense=1)  # max 3x + 5y, 2x + y <= 8, x + 2y <= 8
    x = mb.add_variable(0.0, np.inf, 3.0)
    y = mb.add_variable(0.0, np.inf, 5.0)
    mb.add_constraint([(x, 2.0), (y, 1.0)], lpd.LE, 8.0)
    mb.add_constraint([(x, 1.0), (y, 2.0)], lpd.LE, 8.0)
    out["hand_lp_64_3"] = (mb.build(), {})
    mb = lpd.ModelBuilder(sense=1)
    x = mb.add_variable(0.0, 100.0, 5.0)
    y = mb.add_variable(0.0, 100.0, 6.0)
    mb.add_constraint([(x, 1.0), (y, 1.0)], lpd.LE, 50.0)
    out["iteration_limit_5"] = (mb.build(), dict(max_iterations=5))
    out["presolve_mix"] = (presolve_mix_model(), {})
    out["nonadaptive_steps"] = (lpd.random_lp(40, 120, 4, 11).to_flat(),
                                dict(adaptive_step_size=False, max_iterations=1500))
    out["check_frequency_7"] = (lpd.random_lp(40, 120, 4, 12).to_flat(),
                                dict(check_frequency=7, max_iterations=1500))
    for seed in range(6):  # LE / EQ / GE rows mixed, both objective senses
        fm = lpd.random_lp(30, 80, 3, 100 + seed).to_flat()
        fm.con_sense = np.random.default_rng(1000 + seed).choice([0, 1, 2], size=fm.num_cons).astype(np.int32)
        fm.sense = int(seed % 2)
        out[f"mixed_senses_30x80_s{100 + seed}"] = (fm, dict(max_iterations=3000))
    # MPS models without RANGED rows (the drop-in splits those into two <=
    # rows where the reference keeps only the lower side)
    for fname in ("afiro.free.mps", "example1.mps", "bharatopt_refinery_demo.mps", "edge_obj_rhs.free.mps",
                  "sc50a.free.mps", "blend2.free.mps"):
        info, fm = pkg.read_mps(text=open(os.path.join(HERE, "mps", fname)).read())
        assert info["status"] == 0, fname
        out["mps_" + fname] = (fm, dict(max_iterations=5000))
    return out


def presolve_mix_model():
    """Exercises every presolve pass before PDHG: a fixed variable, an empty
    column, a singleton row, GE and EQ rows, a free variable, maximisation."""
    mb = lpd.ModelBuilder(sense=1)
    v = [mb.add_variable(0.0, 10.0, 1.0 + 0.5 * j) for j in range(6)]
    fixed = mb.add_variable(2.0, 2.0, 3.0)
    empty = mb.add_variable(-1.0, 4.0, -2.0)
    free = mb.add_variable(-np.inf, np.inf, 0.5)
    mb.add_constraint([(v[0], 1.0), (v[1], 2.0), (fixed, 1.0)], lpd.LE, 12.0)
    mb.add_constraint([(v[1], 1.0), (v[2], 1.0), (v[3], 1.0)], lpd.LE, 9.0)
    mb.add_constraint([(v[2], 1.0), (v[4], -1.0)], lpd.GE, -3.0)
    mb.add_constraint([(v[3], 1.0), (v[4], 1.0), (v[5], 1.0), (free, 1.0)], lpd.EQ, 10.0)
    mb.add_constraint([(v[5], 2.0)], lpd.LE, 7.0)          # singleton row -> bound
    mb.add_constraint([(free, 1.0), (v[0], 1.0)], lpd.LE, 6.0)
    mb.add_constraint([(free, -1.0), (v[1], 1.0)], lpd.LE, 5.0)
    _ = empty
    return mb.build()


def verifier_model():
    """Every variable kind and row sense: max x0 - 2x1 + 3x2 + 0.5x3 + x5 + 2.5,
    x0 in [0, 10], x1 >= -5, x2 integer in [0, 8], x3 binary, x4 free,
    x5 <= 3; rows x0 + 2x1 <= 10, x0 - x2 >= -2, x1 + x3 + x4 = 4,
    1 <= x2 + x5 <= 6 (ranged)."""
    mb = lpd.ModelBuilder(sense=1)
    mb.offset = 2.5
    x0 = mb.add_variable(0.0, 10.0, 1.0)
    x1 = mb.add_variable(-5.0, np.inf, -2.0)
    x2 = mb.add_variable(0.0, 8.0, 3.0, vtype=1)
    x3 = mb.add_variable(0.0, 1.0, 0.5, vtype=2)
    x4 = mb.add_variable(-np.inf, np.inf, 0.0)
    x5 = mb.add_variable(-np.inf, 3.0, 1.0)
    mb.add_constraint([(x0, 1.0), (x1, 2.0)], lpd.LE, 10.0)
    mb.add_constraint([(x0, 1.0), (x2, -1.0)], lpd.GE, -2.0)
    mb.add_constraint([(x1, 1.0), (x3, 1.0), (x4, 1.0)], lpd.EQ, 4.0)
    mb.add_constraint([(x2, 1.0), (x5, 1.0)], lpd.RANGED, 1.0, 6.0)
    return mb.build()


def verifier_cases():
    """name -> (FlatModel, candidate x, kwargs of verify_solution)."""
    fm = verifier_model()
    good = np.array([2.0, 1.0, 3.0, 1.0, 2.0, 1.0])  # feasible; objective 2 - 2 + 9 + 0.5 + 1 + 2.5 = 13

    def at(**kv):
        x = good.copy()
        for k, v in kv.items():
            x[int(k[1:])] = v
        return x

    out = {
        "feasible": (fm, good, {}),
        "lower_bound": (fm, at(x0=-0.5), {}),
        "upper_bound": (fm, at(x0=10.2), {}),
        "upper_bound_within_tol": (fm, at(x0=10.0 + 5e-8), {}),
        "integer": (fm, at(x2=3.4), {}),
        "integer_within_tol": (fm, at(x2=3.0 + 5e-6), {}),
        "binary": (fm, at(x3=0.5), {}),
        "nan": (fm, at(x4=np.nan), {}),
        "plus_inf": (fm, at(x1=np.inf), {}),
        "minus_inf": (fm, at(x5=-np.inf), {}),
        "le_and_eq_rows": (fm, at(x1=5.0), {}),
        "ge_row": (fm, at(x2=5.0, x5=0.5), {}),
        "eq_row": (fm, at(x4=2.5), {}),
        "ranged_low": (fm, at(x5=-3.0), {}),
        "ranged_high": (fm, at(x2=4.0, x5=3.0), {}),
        "several_rules": (fm, at(x0=-1.0, x2=2.5, x4=7.0), {}),
        "short_candidate": (fm, good[:5], {}),
        "long_candidate": (fm, np.append(good, 0.0), {}),
        "empty_candidate": (fm, good[:0], {}),
        "reported_exact": (fm, good, dict(reported_objective=13.0)),
        "reported_off_1e-3": (fm, good, dict(reported_objective=13.001)),
        "reported_off_1e-5": (fm, good, dict(reported_objective=13.00001)),
        "reported_off_unchecked": (fm, good, dict(reported_objective=14.0, check_objective=False)),
        "reported_with_violation": (fm, at(x0=-0.5), dict(reported_objective=1.0)),
        "loose_feasibility": (fm, at(x4=2.005), dict(feasibility_tolerance=1e-2)),
        "loose_integrality": (fm, at(x2=3.2), dict(integrality_tolerance=0.25)),
    }
    rng = np.random.default_rng(77)
    for t in range(40):  # seeded random models: every sense, both objective senses, mixed types
        m_, n_ = int(rng.integers(3, 12)), int(rng.integers(3, 12))
        mb = lpd.ModelBuilder(sense=int(t % 2))
        mb.offset = float(rng.normal())
        for j in range(n_):
            lo = -np.inf if rng.random() < 0.2 else float(rng.integers(-5, 3))
            hi = np.inf if rng.random() < 0.2 else max(lo, 0.0) + float(rng.integers(0, 8))
            mb.add_variable(lo, hi, float(rng.normal()), vtype=int(rng.integers(0, 3)) if rng.random() < 0.3 else 0)
        for i in range(m_):
            k = int(rng.integers(1, n_ + 1))
            cols = rng.choice(n_, size=k, replace=False)
            terms = [(int(j), float(np.round(rng.normal() * 3, 3))) for j in cols]
            sense = int(rng.integers(0, 4))
            rhs = float(rng.integers(-4, 8))
            mb.add_constraint(terms, sense, rhs, rhs + float(rng.integers(0, 6)))
        fmr = mb.build()
        x = rng.integers(-3, 6, size=n_).astype(float) + np.where(rng.random(n_) < 0.3, rng.normal(size=n_) * 1e-6, 0.0)
        kw = dict(reported_objective=float(rng.normal() * 10)) if t % 3 == 0 else {}
        out[f"random_{t}"] = (fmr, x, kw)
    return out


def verifier_main():
    """verifier_kat.json: the reference SolutionVerifier's full result on every case."""
    if po.ref_lib() is None:
        raise SystemExit("oracle/_ref is not built; run make -C oracle first")
    out = {}
    for name, (fm, x, kw) in verifier_cases().items():
        r = po.ref_verify_full(fm, x, **kw)
        out[name] = {k: (float.hex(float(v)) if isinstance(v, float) else v) for k, v in r.items()}
        print(name, r["status"], r["verified"], r["violated_constraint_count"], r["violated_bound_count"],
              r["violated_integrality_count"], flush=True)
    with open(os.path.join(HERE, "verifier_kat.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


def pdlp_instances():
    out = {}
    for s in (42, 43, 44, 45, 46):
        out[f"random_300x1000_s{s}"] = lpd.random_lp(300, 1000, 5, s)
    out["random_60x200_s7"] = lpd.random_lp(60, 200, 5, 7)
    out["staircase_3x200x500_s3"] = lpd.staircase_lp(3, 200, 500, 20, seed=3)
    out["zipf_2000x5000_s9"] = lpd.zipf_lp(2000, 5000, 1.5, 200, seed=9)
    return out


def highs(lp):
    from scipy.optimize import linprog
    from scipy.sparse import csr_matrix
    A = csr_matrix((lp.val, lp.col_idx, lp.row_ptr), shape=(lp.m, lp.n))
    eq = lp.is_eq.astype(bool)
    res = linprog(lp.c, A_ub=A[~eq] if (~eq).any() else None, b_ub=lp.b[~eq] if (~eq).any() else None,
                  A_eq=A[eq] if eq.any() else None, b_eq=lp.b[eq] if eq.any() else None,
                  bounds=list(zip(lp.lb, lp.ub)), method="highs",
                  options=dict(primal_feasibility_tolerance=1e-10, dual_feasibility_tolerance=1e-10))
    assert res.status == 0, res.message
    return float(res.fun)


def main(kat_only=False):
    if po.ref_lib() is None:
        raise SystemExit("oracle/_ref is not built; run make -C oracle first")
    kat = {}
    for name, (fm, kw) in kat_models().items():
        r = po.ref_fo_solve(fm, **kw)
        kat[name] = dict(kwargs=kw, model_sha=model_sha(fm), status=r["status"],
                         iterations=r["iterations"], objective=float.hex(r["objective"]),
                         constraint_violation=float.hex(r["constraint_violation"]),
                         x_sha=sha(r["x"]), y_sha=sha(r["y"]), x_len=len(r["x"]), y_len=len(r["y"]))
        print(name, r["status"], r["iterations"], r["objective"], flush=True)
    with open(os.path.join(HERE, "ref_pdhg_kat.json"), "w") as f:
        json.dump(kat, f, indent=1, sort_keys=True)
    if kat_only:
        return

    opt = {}
    for name, lp in pdlp_instances().items():
        entry = dict(model_sha=sha(lp.row_ptr, lp.col_idx, lp.val, lp.b, lp.c, lp.lb, lp.ub, lp.is_eq),
                     highs=highs(lp))
        if lp.m * lp.n <= 400_000:
            s = po.ref_simplex(lp.to_flat())
            if s["status"] == 0:
                entry["reference_simplex"] = s["objective"]
                entry["reference_simplex_stage"] = s["stage"]
        opt[name] = entry
        print(name, entry, flush=True)
    with open(os.path.join(HERE, "lp_optima.json"), "w") as f:
        json.dump(opt, f, indent=1, sort_keys=True)


MPS_KEYS = ("lb", "ub", "obj", "var_type", "con_sense", "rhs", "range_upper", "term_ptr", "term_var", "term_val")


def flat_sha(d):
    """SHA-256 over the flattened model arrays in a fixed dtype per field."""
    h = hashlib.sha256()
    for k in MPS_KEYS:
        dt = np.int64 if k in ("var_type", "con_sense", "term_ptr", "term_var") else np.float64
        h.update(np.ascontiguousarray(np.asarray(d[k]), dt).tobytes())
    h.update(np.array([d["sense"]], np.int64).tobytes())
    h.update(np.array([d["obj_offset"]], np.float64).tobytes())
    return h.hexdigest()


def highs_flat(d):
    """LP optimum of a flattened model by HiGHS: (status name, objective)."""
    from scipy.optimize import linprog
    from scipy.sparse import csr_matrix
    nv, nc = len(d["lb"]), len(d["rhs"])
    A = csr_matrix((d["term_val"], d["term_var"], d["term_ptr"]), shape=(nc, nv))
    cs, rhs, ru = d["con_sense"], d["rhs"], d["range_upper"]
    ub_rows, ub_rhs, eq_rows, eq_rhs = [], [], [], []
    for i in range(nc):
        if cs[i] == 0:
            ub_rows.append(A[i]); ub_rhs.append(rhs[i])
        elif cs[i] == 2:
            ub_rows.append(-A[i]); ub_rhs.append(-rhs[i])
        elif cs[i] == 1:
            eq_rows.append(A[i]); eq_rhs.append(rhs[i])
        else:  # ranged: rhs <= a'x <= range_upper
            ub_rows.append(A[i]); ub_rhs.append(ru[i])
            ub_rows.append(-A[i]); ub_rhs.append(-rhs[i])
    from scipy.sparse import vstack
    sign = -1.0 if d["sense"] == 1 else 1.0
    res = linprog(sign * np.asarray(d["obj"]), A_ub=vstack(ub_rows) if ub_rows else None,
                  b_ub=ub_rhs if ub_rows else None, A_eq=vstack(eq_rows) if eq_rows else None,
                  b_eq=eq_rhs if eq_rows else None, bounds=list(zip(d["lb"], d["ub"])), method="highs",
                  options=dict(primal_feasibility_tolerance=1e-10, dual_feasibility_tolerance=1e-10))
    name = {0: "optimal", 2: "infeasible", 3: "unbounded"}.get(res.status, "other")
    return name, (float(sign * res.fun + d["obj_offset"]) if res.status == 0 else None)


def mps_main():
    if po.ref_lib() is None:
        raise SystemExit("oracle/_ref is not built; run make -C oracle first")
    out = {}
    folder = os.path.join(HERE, "mps")
    for fname in sorted(os.listdir(folder)):
        if not fname.endswith(".mps"):
            continue
        d = po.ref_mps(open(os.path.join(folder, fname)).read())
        entry = dict(status=d["status"], num_vars=len(d["lb"]), num_cons=len(d["rhs"]),
                     terms=int(d["term_ptr"][-1]) if len(d["term_ptr"]) else 0, sense=d["sense"],
                     obj_offset=d["obj_offset"], model_sha=flat_sha(d),
                     integer_vars=int(np.sum(np.asarray(d["var_type"]) > 0)))
        if d["status"] == 0:
            entry["highs_status"], entry["highs_objective"] = highs_flat(d)
        out[fname] = entry
        print(fname, entry, flush=True)
    with open(os.path.join(HERE, "mps_golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "mps":
        mps_main()
    elif len(sys.argv) > 1 and sys.argv[1] == "verifier":
        verifier_main()
    else:
        main(kat_only=len(sys.argv) > 1 and sys.argv[1] == "kat")
