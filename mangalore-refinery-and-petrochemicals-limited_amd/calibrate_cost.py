"""Measures the GPU first-order solve's cost on this device, the data the
routing cost model of bo_pdhg_cost_estimate is fitted to (it replaces the
reference's analytic GPU estimate 24.5 + 0.0008 nnz ms,
src/cost_estimator.cpp:193).

For a ladder of LP sizes (the random generator: m rows, 3m columns, 5
nonzeros per column, so nnz = 15 m) and the benchmark's staircase and Zipf
LPs it times, with the LP's arrays on the host as a drop-in caller holds
them:

* PDLP engine (bo_pdlp_*): handle creation (upload, layout, schedule
  tuning), then solves capped at two iteration counts with an unreachable
  tolerance, whose difference gives the cost per iteration and whose
  remainder the per-solve device set-up;
* reference engine (bo_pdhg_ref_solve): two iteration caps, per-iteration
  cost and fixed cost, for each loop tier it picks, and its host-side
  preparation and upload (transfer_time_ms).

It prints one JSON line per size and then the tables of
host/src/cost_model.cpp (nnz, ms), which bo_pdhg_cost_estimate interpolates
in log-log: PDLP handle creation, per-solve set-up and per-iteration cost;
reference-engine host preparation and per-iteration cost per loop tier.

Usage:  python calibrate_cost.py [--sizes 100,300,...]   (needs a GPU)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def _pkg():
    import __graft_entry__
    return __graft_entry__.load_package()


def measure(pkg, lp, k1, k2, ref_k1, ref_k2, name):
    out = {"lp": name, "m": lp.m, "n": lp.n, "nnz": lp.nnz}
    t0 = time.perf_counter()
    h = pkg.Pdlp(lp, eps_rel_kkt=1e-300, max_iterations=k1)
    out["create_ms"] = (time.perf_counter() - t0) * 1e3
    h.solve()  # warm
    s1 = h.solve()
    t1 = s1["solve_ms"]
    h.close()
    h = pkg.Pdlp(lp, eps_rel_kkt=1e-300, max_iterations=k2)
    h.solve()
    s2 = h.solve()
    t2 = s2["solve_ms"]
    out["single_workgroup"] = int(s2["single_workgroup"])
    h.close()
    out["pdlp_iter_ms"] = (t2 - t1) / (s2["iterations"] - s1["iterations"])
    out["pdlp_fixed_ms"] = t1 - s1["iterations"] * out["pdlp_iter_ms"]
    r1 = pkg.pdhg_ref_solve(lp, max_iterations=ref_k1, feasibility_tolerance=0.0)
    r2 = pkg.pdhg_ref_solve(lp, max_iterations=ref_k2, feasibility_tolerance=0.0)
    out["ref_loop"] = int(r2.loop)
    out["ref_transfer_ms"] = r1.transfer_time_ms
    out["ref_iter_ms"] = (r2.kernel_time_ms - r1.kernel_time_ms) / (r2.iterations - r1.iterations)
    out["ref_fixed_ms"] = r1.kernel_time_ms - r1.iterations * out["ref_iter_ms"]
    return out


def tables(rows):
    """The measured points as the (nnz, ms) tables of cost_model.cpp."""
    def tab(key, pick=lambda r: True):
        return [[r["nnz"], round(r[key], 6)] for r in rows if pick(r)]
    out = {"pdlp_iter": tab("pdlp_iter_ms"), "pdlp_create": tab("create_ms"), "pdlp_fixed": tab("pdlp_fixed_ms"),
           "ref_transfer": tab("ref_transfer_ms")}
    for tier in sorted({r["ref_loop"] for r in rows}):
        out["ref_iter_loop%d" % tier] = tab("ref_iter_ms", lambda r, t=tier: r["ref_loop"] == t)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="100,300,1000,3000,10000,30000,100000,300000")
    ap.add_argument("--workloads", default="c2,c1", help="benchmark LPs measured too (c2: staircase, c1: Zipf)")
    args = ap.parse_args()
    pkg = _pkg()
    st, info = pkg.device_info()
    if st != 0:
        raise SystemExit("no GPU: " + info.status_message.decode())
    rows = []
    warm = pkg.Pdlp(pkg.lpdata.random_lp_sampled(50, 150, 5, 1), max_iterations=10)  # one-time start-up
    warm.solve()
    warm.close()
    lps = [("random %d x %d" % (m, 3 * m), pkg.lpdata.random_lp_sampled(m, 3 * m, 5, 1000 + m))
           for m in [int(v) for v in args.sizes.split(",") if v]]
    for w in [v for v in args.workloads.split(",") if v]:
        if w == "c1":
            lps.append(("zipf 200000 x 500000", pkg.lpdata.zipf_lp(200_000, 500_000, 1.5, 2000, seed=42)))
        elif w == "c2":
            lps.append(("staircase 24 periods", pkg.lpdata.staircase_lp(24, 2000, 5000, 200, seed=42)))
    for name, lp in lps:
        small = lp.nnz <= (1 << 18)
        r = measure(pkg, lp, 1024 if small else 512, 3072 if small else 1536, 500,
                    1500 if lp.nnz <= 500_000 else 700, name)
        rows.append(r)
        print(json.dumps(r), flush=True)
    rows.sort(key=lambda r: r["nnz"])
    print(json.dumps({"tables": tables(rows), "device": info.device_name.decode()}), flush=True)


if __name__ == "__main__":
    main()
