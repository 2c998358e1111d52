"""Learned warm starts for a parametric family of LPs over one matrix.

No counterpart exists in the reference: its first-order solver always starts
from zero (src/first_order_solver.cpp:133-134 sets x = 0, y = 0) and its only
warm start is a dual-simplex basis for branch and bound
(include/bharatopt/dual_revised_simplex.hpp:70).

Pipeline (all heavy work on the GPU, through libbo_hip.so):

* ``FamilySolver`` solves draws of a family (``lpdata.StaircaseFamily``) in
  batches of up to 64: by default with the family engine (``PdlpFamily``:
  one matrix, one LP per lane), or with the block-diagonal batched engine
  (``PdlpBatch``, one per batch size, built once over copies of the
  family's matrix).  Every batch swaps in its draws' costs and right-hand
  sides, optionally a warm start, and solves.  A draw the warm-started solve
  leaves unsolved is solved again from the cold start (the fallback); the
  engine's FP64 rel-KKT of the unscaled LP decides every result.
* ``WarmStartNet`` is the BF16 multilayer perceptron of bo_mlp_* (forward,
  backward and Adam kernels on the matrix cores).  It maps the normalised
  perturbations of a draw (n + m features) to its primal / dual solution
  and, optionally, the log of the primal weight its solve ended with,
  standardised per coordinate with the training set's mean and spread; the
  output layer starts at zero, so the untrained network predicts the
  training mean.  Training keeps the parameters of the best validation loss.
  A predicted primal weight starts the solve in place of ||c|| / ||b||: on
  the staircase family that shortens the slowest solves of a batch.
"""
from __future__ import annotations

import time

import numpy as np


class FamilySolver:
    """Batched solves of family draws, reusing one device layout per batch size."""

    def __init__(self, pkg, family, batch=64, engine="family", **params):
        assert engine in ("family", "batch") and (engine == "batch" or batch <= 64)
        self.pkg, self.family, self.batch, self.engine, self.params = pkg, family, batch, engine, params
        self._engines = {}

    def _engine(self, k):
        key = "family" if self.engine == "family" else k
        if key not in self._engines:
            if self.engine == "family":
                self._engines[key] = self.pkg.PdlpFamily(self.family.base, capacity=self.batch, **self.params)
            else:
                self._engines[key] = self.pkg.PdlpBatch([self.family.base] * k, **self.params)
        return self._engines[key]

    def _run_chunk(self, chunk, warm_rows, want_solutions):
        eng = self._engine(len(chunk))
        draws = [self.family.draw(k) for k in chunk]
        if self.engine == "family":
            eng.set_lps([d.c for d in draws], [d.b for d in draws])
        else:
            eng.set_objective_rhs([d.c for d in draws], [d.b for d in draws])
        if warm_rows is None:
            eng.set_warm_start(None, None)
            eng.set_primal_weights(None)
        else:
            eng.set_warm_start(warm_rows[0], warm_rows[1])
            eng.set_primal_weights(warm_rows[2] if len(warm_rows) > 2 else None)
        t0 = time.perf_counter()
        p, _ = eng.solve()
        dt = time.perf_counter() - t0
        xs, ys = [], []
        if want_solutions:
            if self.engine == "family":
                X, Y = eng.solutions()
                xs, ys = list(X), list(Y)
            else:
                for i in range(len(chunk)):
                    x, y = eng.solution(i)
                    xs.append(x)
                    ys.append(y)
        return p, dt, xs, ys

    def solve(self, ids, warm=None, want_solutions=True, progress=None):
        """Solve draws `ids` (cold, or from warm = (X, Y) rows aligned with ids).

        warm may carry a third member, starting primal weights.  progress
        (optional): called with a line of text about every 64 batches.  Returns
        dict(per=[stats], x=[...], y=[...], solve_s=wall time of the engine's
        solves, iterations, status, omega = the primal weights the solves
        ended with).  Draw generation and the copy of the costs / right-hand
        sides to the device are outside solve_s."""
        ids = list(ids)
        per, xs, ys = [], [], []
        solve_s = 0.0
        for s in range(0, len(ids), self.batch):
            chunk = ids[s:s + self.batch]
            rows = None if warm is None else tuple([part[s + i] for i in range(len(chunk))] for part in warm)
            p, dt, x, y = self._run_chunk(chunk, rows, want_solutions)
            solve_s += dt
            per.extend(p)
            xs.extend(x)
            ys.extend(y)
            if progress and (s // self.batch) % 64 == 63:
                progress(f"{len(per)}/{len(ids)} draws solved, {solve_s:.1f}s of solves")
        return dict(per=per, x=xs, y=ys, solve_s=solve_s,
                    iterations=np.array([q["iterations"] for q in per]),
                    status=np.array([q["status"] for q in per]),
                    omega=np.array([q["primal_weight"] for q in per]))

    def solve_warm_with_fallback(self, ids, warm, fallback=None):
        """Warm-started solve; draws it leaves unsolved are solved again cold
        (by `fallback`, another FamilySolver, or this one).  solve_s covers
        both; `fallback` lists the positions that fell back."""
        r = self.solve(ids, warm)
        bad = [i for i, st in enumerate(r["status"]) if st != 0]
        r["fallback"] = bad
        if bad:
            c = (fallback or self).solve([ids[i] for i in bad], None)
            r["solve_s"] += c["solve_s"]
            for j, i in enumerate(bad):
                r["per"][i] = c["per"][j]
                if r["x"]:
                    r["x"][i], r["y"][i] = c["x"][j], c["y"][j]
                r["iterations"][i] += c["iterations"][j]
                r["status"][i] = c["status"][j]
        return r

    def close(self):
        for e in self._engines.values():
            e.close()
        self._engines = {}


class WarmStartNet:
    """features (n + m) -> standardised (x*, y*) with the bo_mlp_* network."""

    def __init__(self, pkg, n, m, width=1024, layers=4, max_batch=64, seed=1, primal_weight=False):
        self.pkg, self.n, self.m = pkg, n, m
        self.with_w = bool(primal_weight)
        dout = n + m + (1 if self.with_w else 0)
        self.net = pkg.Mlp(n + m, width, dout, layers=layers, max_batch=max_batch, seed=seed)
        self.max_batch = max_batch
        self.mu = np.zeros(dout)
        self.sd = np.ones(dout)
        p = self.net.params()
        din, dout = self.net.shapes()[-1]
        p[p.size - din * dout - dout:] = 0.0  # output layer: weights and bias
        self.net.set_params(p)

    def _standardise(self, X, Y, W=None):
        parts = [np.asarray(X, np.float64), np.asarray(Y, np.float64)]
        if self.with_w:
            parts.append(np.log(np.asarray(W, np.float64)).reshape(-1, 1))
        Z = np.hstack(parts)
        self.mu = Z.mean(0)
        sd = Z.std(0)
        pos = sd[sd > 0]
        floor = 1e-3 * float(np.median(pos)) if pos.size else 1.0
        self.sd = np.where(sd > 0, np.maximum(sd, floor), 1.0)
        return ((Z - self.mu) / self.sd).astype(np.float32)

    def fit(self, F, X, Y, W=None, epochs=20, lr=1e-4, batch=None, val_frac=0.1, seed=0, log=None):
        """Adam on the mean squared error of the standardised solutions (and
        log primal weights W, for a network built with primal_weight=True).
        Returns the history (train / validation loss per epoch, seconds)."""
        assert (W is not None) == self.with_w
        batch = batch or self.max_batch
        F = np.ascontiguousarray(F, np.float32)
        T = self._standardise(X, Y, W)
        rng = np.random.default_rng(seed)
        order = rng.permutation(len(F))
        nval = max(1, int(round(val_frac * len(F)))) if len(F) > 1 else 0
        val, tr = order[:nval], order[nval:]
        best = (self.val_loss(F[val], T[val]) if nval else np.inf, self.net.params(), -1)
        hist = dict(train=[], val=[best[0]], epochs=epochs, lr=lr, batch=batch, train_draws=len(tr),
                    val_draws=nval)
        t0 = time.perf_counter()
        for ep in range(epochs):
            perm = rng.permutation(tr)
            losses = []
            for s in range(0, len(perm) - batch + 1, batch):
                idx = np.sort(perm[s:s + batch])
                losses.append(self.net.train_step(F[idx], T[idx], lr))
            vl = self.val_loss(F[val], T[val]) if nval else float(np.mean(losses))
            hist["train"].append(float(np.mean(losses)) if losses else float("nan"))
            hist["val"].append(vl)
            if vl < best[0]:
                best = (vl, self.net.params(), ep)
            if log:
                log(f"epoch {ep}: train {hist['train'][-1]:.4f} val {vl:.4f}")
        hist["seconds"] = time.perf_counter() - t0
        hist["best_epoch"] = best[2]
        hist["best_val"] = best[0]
        # keep the best parameters (set_params also resets Adam)
        self.net.set_params(best[1])
        return hist

    def val_loss(self, F, T):
        tot, cnt = 0.0, 0
        for s in range(0, len(F), self.max_batch):
            loss = self.net.loss(F[s:s + self.max_batch], T[s:s + self.max_batch])
            k = min(self.max_batch, len(F) - s)
            tot += loss * k
            cnt += k
        return tot / max(cnt, 1)

    def _split(self, Z):
        nm = self.n + self.m
        if self.with_w:
            return Z[:, :self.n], Z[:, self.n:nm], np.exp(Z[:, nm])
        return Z[:, :self.n], Z[:, self.n:nm]

    def predict(self, F):
        """Warm starts (X rows, Y rows[, primal weights]) for feature rows F."""
        out = []
        for s in range(0, len(F), self.max_batch):
            out.append(self.net.forward(np.ascontiguousarray(F[s:s + self.max_batch], np.float32)))
        return self._split(self.mu + self.sd * np.vstack(out).astype(np.float64))

    def mean_start(self, count):
        return self._split(np.tile(self.mu, (count, 1)))

    def close(self):
        self.net.close()


def start_sensitivity(solver, ids, X, Y, lb, ub, seed=0):
    """Iterations of warm-started solves of draws `ids` from their own
    solutions (X, Y rows, e.g. from a cold solve), disturbed two ways: every
    entry by a relative error r (x (1 + r N(0,1)): zeros and bound-active
    entries at 0 stay put, each entry keeps r of its own precision), and
    every entry by an absolute error r times the vector's root-mean-square
    entry (small entries and zeros are lost); plus the mean solution with
    the true active set (x at a bound, zero duals) kept.  How close, and in
    which sense, a predicted start must be before it saves iterations.
    Returns {label: mean iterations}."""
    rng = np.random.default_rng(seed)
    X, Y = np.asarray(X, np.float64), np.asarray(Y, np.float64)

    def run(WX, WY):
        return float(solver.solve(ids, (WX, WY), want_solutions=False)["iterations"].mean())

    out = {"exact": run(X, Y)}
    for r in (1e-3, 1e-2, 1e-1):
        out[f"relative {r:g} per entry"] = run(np.clip(X * (1 + r * rng.standard_normal(X.shape)), lb, ub),
                                               Y * (1 + r * rng.standard_normal(Y.shape)))
    nx = np.linalg.norm(X, axis=1, keepdims=True) / np.sqrt(X.shape[1])
    ny = np.linalg.norm(Y, axis=1, keepdims=True) / np.sqrt(Y.shape[1])
    for r in (1e-4, 1e-2):
        out[f"absolute {r:g} x rms"] = run(X + r * nx * rng.standard_normal(X.shape),
                                            Y + r * ny * rng.standard_normal(Y.shape))
    atb = (np.abs(X - lb) <= 1e-9 * (1 + np.abs(lb))) | (np.abs(X - ub) <= 1e-9 * (1 + np.abs(ub)))
    yz = np.abs(Y) <= 1e-12
    out["mean values, true active set"] = run(np.where(atb, X, np.clip(np.tile(X.mean(0), (len(X), 1)), lb, ub)),
                                              np.where(yz, 0.0, np.tile(Y.mean(0), (len(Y), 1))))
    return out


def relative_distance(Xp, Yp, X, Y):
    """Median over draws of ||(xp, yp) - (x*, y*)|| / ||(x*, y*)||."""
    d = [np.sqrt(np.sum((xp - x) ** 2) + np.sum((yp - y) ** 2)) / max(np.sqrt(np.sum(x ** 2) + np.sum(y ** 2)), 1e-300)
         for xp, yp, x, y in zip(Xp, Yp, X, Y)]
    return float(np.median(d))
