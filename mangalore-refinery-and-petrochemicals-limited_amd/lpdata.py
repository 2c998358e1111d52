"""LP containers and seeded synthetic generators.

``FlatModel`` mirrors BharatOpt's ``LPModel`` (include/bharatopt/lp_model.hpp:60-127)
as arrays: row-wise terms in insertion order, constraint senses
0 LE / 1 EQ / 2 GE / 3 RANGED, objective sense 0 minimise / 1 maximise.
``StdLP`` is the standard form the solver builds from it
(src/first_order_solver.cpp:68-101): min c'x, rows ``a_i x <= b_i`` (``== b_i``
when ``is_eq``), ``lb <= x <= ub``, CSR with ascending columns.

Generators follow the workload descriptions of the benchmark configurations:

* ``random_lp``      m x n, ``k`` distinct uniform rows per column, N(0,1) values,
                     x0 ~ U[0,10], u = 20, b = A x0 + U[0,1], c ~ N(0,1)
* ``zipf_lp``        the same recipe with column degrees ~ Zipf(s) truncated to
                     [1, cap] (renormalised), 200k x 500k at full size
* ``staircase_lp``   multi-period refinery-planning structure: per period a
                     blending block (equality material balances, <= yield rows),
                     inventory columns linking consecutive periods, column scales
                     log-uniform in [1e-3, 1e3]; b from a feasible x0
* ``batch_sizes``    row counts of the ragged batch (m ~ U{200..20000}, n = 3m)

All draws use numpy's PCG64 ``default_rng(seed)``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

LE, EQ, GE, RANGED = 0, 1, 2, 3


@dataclass
class StdLP:
    m: int
    n: int
    row_ptr: np.ndarray  # int32[m+1]
    col_idx: np.ndarray  # int32[nnz]
    val: np.ndarray      # float64[nnz]
    b: np.ndarray
    c: np.ndarray
    lb: np.ndarray
    ub: np.ndarray
    is_eq: np.ndarray    # uint8[m]
    x0: np.ndarray | None = None  # a known feasible point, when the generator has one
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    def to_flat(self) -> "FlatModel":
        """As an LPModel: every row LE (or EQ), minimise."""
        return FlatModel(
            num_vars=self.n, num_cons=self.m, sense=0, obj_offset=0.0,
            lb=self.lb.copy(), ub=self.ub.copy(), obj=self.c.copy(), var_type=None,
            con_sense=np.where(self.is_eq.astype(bool), EQ, LE).astype(np.int32),
            rhs=self.b.copy(), range_upper=None,
            term_ptr=self.row_ptr.astype(np.int64), term_var=self.col_idx.copy(),
            term_val=self.val.copy())

    def spmv(self, x):
        rows = np.repeat(np.arange(self.m), np.diff(self.row_ptr))
        return np.bincount(rows, weights=self.val * np.asarray(x)[self.col_idx], minlength=self.m)

    def spmv_t(self, y):
        rows = np.repeat(np.arange(self.m), np.diff(self.row_ptr))
        return np.bincount(self.col_idx, weights=self.val * np.asarray(y)[rows], minlength=self.n)


@dataclass
class FlatModel:
    num_vars: int
    num_cons: int
    sense: int
    obj_offset: float
    lb: np.ndarray
    ub: np.ndarray
    obj: np.ndarray
    var_type: np.ndarray | None
    con_sense: np.ndarray
    rhs: np.ndarray
    range_upper: np.ndarray | None
    term_ptr: np.ndarray
    term_var: np.ndarray
    term_val: np.ndarray


class ModelBuilder:
    """Small LPModel-style builder used to restate the reference's test models."""

    def __init__(self, sense=0):
        self.sense = sense
        self.vars = []   # (lb, ub, obj, type)
        self.cons = []   # (terms, sense, rhs, range_upper)
        self.offset = 0.0

    def add_variable(self, lb=0.0, ub=np.inf, obj=0.0, vtype=0):
        self.vars.append((lb, ub, obj, vtype))
        return len(self.vars) - 1

    def add_constraint(self, terms, sense, rhs, range_upper=0.0):
        self.cons.append((list(terms), sense, rhs, range_upper))
        return len(self.cons) - 1

    def build(self) -> FlatModel:
        ptr = [0]
        tv, tw = [], []
        for terms, _, _, _ in self.cons:
            for j, a in terms:
                tv.append(j)
                tw.append(a)
            ptr.append(len(tv))
        return FlatModel(
            num_vars=len(self.vars), num_cons=len(self.cons), sense=self.sense,
            obj_offset=self.offset,
            lb=np.array([v[0] for v in self.vars], float), ub=np.array([v[1] for v in self.vars], float),
            obj=np.array([v[2] for v in self.vars], float),
            var_type=np.array([v[3] for v in self.vars], np.int32),
            con_sense=np.array([c[1] for c in self.cons], np.int32),
            rhs=np.array([c[2] for c in self.cons], float),
            range_upper=np.array([c[3] for c in self.cons], float),
            term_ptr=np.array(ptr, np.int64), term_var=np.array(tv, np.int32),
            term_val=np.array(tw, float))


def _csr_from_cols(m, n, col_of_entry, row_of_entry, val):
    """CSR (rows ascending, columns ascending within a row) from entry lists."""
    order = np.lexsort((col_of_entry, row_of_entry))
    r = row_of_entry[order]
    c = col_of_entry[order].astype(np.int32)
    v = val[order]
    row_ptr = np.zeros(m + 1, np.int64)
    np.add.at(row_ptr, r + 1, 1)
    row_ptr = np.cumsum(row_ptr).astype(np.int32)
    return row_ptr, c, v


def _finish(m, n, row_ptr, col_idx, val, rng, x0, slack_lo=0.0, slack_hi=1.0, eq_mask=None,
            c=None, lb=None, ub=None, meta=None):
    lp = StdLP(m, n, row_ptr, col_idx, val, np.zeros(m), np.zeros(n), np.zeros(n), np.zeros(n),
               np.zeros(m, np.uint8), x0=x0, meta=meta or {})
    ax = lp.spmv(x0)
    slack = rng.uniform(slack_lo, slack_hi, m)
    if eq_mask is not None:
        slack = np.where(eq_mask, 0.0, slack)
        lp.is_eq = eq_mask.astype(np.uint8)
    lp.b = ax + slack
    lp.c = c if c is not None else rng.standard_normal(n)
    lp.lb = lb if lb is not None else np.zeros(n)
    lp.ub = ub if ub is not None else np.full(n, 20.0)
    return lp


def random_lp(m=300, n=1000, k=5, seed=42):
    """m x n, k distinct uniform rows per column, N(0,1) values; x0 ~ U[0,10],
    u = 20, b = A x0 + U[0,1], c ~ N(0,1), minimise, all rows <=."""
    rng = np.random.default_rng(seed)
    k = min(k, m)
    if m <= 4096:
        rows = np.argsort(rng.random((n, m)), axis=1)[:, :k]
    else:  # sample with replacement, then redraw collisions
        rows = rng.integers(0, m, size=(n, k))
        for _ in range(64):
            srt = np.sort(rows, axis=1)
            dup = (np.diff(srt, axis=1) == 0).any(axis=1)
            if not dup.any():
                break
            rows[dup] = rng.integers(0, m, size=(int(dup.sum()), k))
    cols = np.repeat(np.arange(n), k)
    vals = rng.standard_normal(n * k)
    row_ptr, col_idx, val = _csr_from_cols(m, n, cols, rows.reshape(-1), vals)
    x0 = rng.uniform(0.0, 10.0, n)
    return _finish(m, n, row_ptr, col_idx, val, rng, x0, meta=dict(kind="random", seed=seed, k=k))


def zipf_degrees(n, s=1.5, cap=2000, rng=None):
    """Column degrees ~ Zipf(s) truncated to [1, cap] and renormalised."""
    ks = np.arange(1, cap + 1, dtype=np.float64)
    p = ks ** (-s)
    p /= p.sum()
    return rng.choice(np.arange(1, cap + 1), size=n, p=p)


def zipf_lp(m=200_000, n=500_000, s=1.5, cap=2000, seed=42):
    """The random_lp recipe with Zipf(s)-distributed column degrees (<= cap)."""
    rng = np.random.default_rng(seed)
    deg = zipf_degrees(n, s, cap, rng)
    cols = np.repeat(np.arange(n, dtype=np.int64), deg)
    rows = rng.integers(0, m, size=cols.size)
    # drop repeated (row, col) pairs: a column keeps distinct rows
    key = cols * m + rows
    key, first = np.unique(key, return_index=True)
    cols = key // m
    rows = key % m
    vals = rng.standard_normal(cols.size)
    row_ptr, col_idx, val = _csr_from_cols(m, n, cols, rows, vals)
    x0 = rng.uniform(0.0, 10.0, n)
    return _finish(m, n, row_ptr, col_idx, val, rng, x0,
                   meta=dict(kind="zipf", seed=seed, s=s, cap=cap, realised_nnz=int(val.size)))


def staircase_lp(periods=24, blk_rows=2000, blk_cols=5000, inv_per_link=200, seed=42,
                 eq_fraction=0.5, scale_lo=1e-3, scale_hi=1e3):
    """Refinery-planning staircase.

    Period t owns rows [t*R, (t+1)*R) and columns [t*C, (t+1)*C).  The first
    ``eq_fraction`` of a block's rows are equality material balances with
    coefficients +-U[0.1,1]; the rest are <= yield rows with coefficients
    U[0.1,1].  Each block column touches U{6..17} rows of its block.
    ``inv_per_link`` inventory columns link period t to t+1: +U[0.1,1] in a
    balance row of t, -U[0.1,1] in the same balance row of t+1.  Every
    column j is then scaled by s_j ~ log-U[scale_lo, scale_hi] (A[:,j] *= s_j,
    c_j *= s_j, bounds / s_j).  b = A x0 (+ U[0,1] slack on <= rows)."""
    rng = np.random.default_rng(seed)
    R, C = blk_rows, blk_cols
    n_eq = int(round(R * eq_fraction))
    m = periods * R
    n_blk = periods * C
    n_inv = (periods - 1) * inv_per_link
    n = n_blk + n_inv
    deg = rng.integers(6, 18, size=n_blk)
    cols = np.repeat(np.arange(n_blk, dtype=np.int64), deg)
    period = cols // C
    local = rng.integers(0, R, size=cols.size)
    key = np.unique(cols * R + local)
    cols = key // R
    local = key % R
    rows = (cols // C) * R + local
    is_bal = local < n_eq
    mag = rng.uniform(0.1, 1.0, cols.size)
    sign = np.where(is_bal, rng.choice([-1.0, 1.0], size=cols.size), 1.0)
    vals = mag * sign
    # inventory links
    t = np.repeat(np.arange(periods - 1), inv_per_link)
    brow = rng.integers(0, n_eq, size=n_inv)
    inv_cols = n_blk + np.arange(n_inv)
    cols = np.concatenate([cols, inv_cols, inv_cols])
    rows = np.concatenate([rows, t * R + brow, (t + 1) * R + brow])
    vals = np.concatenate([vals, rng.uniform(0.1, 1.0, n_inv), -rng.uniform(0.1, 1.0, n_inv)])
    # natural units: x0 in [0, 10], u = 20
    x0n = rng.uniform(0.0, 10.0, n)
    cn = rng.standard_normal(n)
    scale = np.exp(rng.uniform(np.log(scale_lo), np.log(scale_hi), n))
    vals = vals * scale[cols]
    x0 = x0n / scale
    c = cn * scale
    ub = 20.0 / scale
    row_ptr, col_idx, val = _csr_from_cols(m, n, cols, rows, vals)
    eq_mask = (np.arange(m) % R) < n_eq
    return _finish(m, n, row_ptr, col_idx, val, rng, x0, eq_mask=eq_mask, c=c,
                   lb=np.zeros(n), ub=ub,
                   meta=dict(kind="staircase", seed=seed, periods=periods, blk_rows=R, blk_cols=C,
                             realised_nnz=int(val.size)))


def batch_sizes(count=512, lo=200, hi=20000, seed=42):
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=count)


def random_lp_sampled(m, n, k=5, seed=42):
    """The random_lp recipe (k distinct uniform rows per column, N(0,1) values,
    x0 ~ U[0,10], u = 20, b = A x0 + U[0,1], c ~ N(0,1)) with the rows drawn
    by sampling and redrawing collisions at every size -- linear in n*k, for
    generating many LPs quickly (a different random stream from random_lp)."""
    rng = np.random.default_rng(seed)
    k = min(k, m)
    rows = rng.integers(0, m, size=(n, k))
    for _ in range(256):
        srt = np.sort(rows, axis=1)
        dup = (np.diff(srt, axis=1) == 0).any(axis=1)
        if not dup.any():
            break
        rows[dup] = rng.integers(0, m, size=(int(dup.sum()), k))
    cols = np.repeat(np.arange(n), k)
    vals = rng.standard_normal(n * k)
    row_ptr, col_idx, val = _csr_from_cols(m, n, cols, rows.reshape(-1), vals)
    x0 = rng.uniform(0.0, 10.0, n)
    return _finish(m, n, row_ptr, col_idx, val, rng, x0, meta=dict(kind="random_sampled", seed=seed, k=k))


def batch_lps(count=512, lo=200, hi=20000, seed=42, k=5):
    """The batched workload: count LPs, LP i of size m_i ~ U{lo..hi} rows,
    n_i = 3 m_i columns, k nonzeros per column, seed + 1 + i."""
    sizes = batch_sizes(count, lo, hi, seed)
    return [random_lp_sampled(int(mi), 3 * int(mi), k, seed + 1 + i) for i, mi in enumerate(sizes)]


class StaircaseFamily:
    """Parametric family over one fixed staircase matrix (the learned
    warm-start workload): draw k has

        c = c0 * (1 + 0.2 N(0,1)),   b = b0 * (1 + U[-0.1, 0.1])

    elementwise, from ``default_rng(k)``, with A, bounds and row senses of
    ``staircase_lp(periods, blk_rows, blk_cols, inv_per_link, seed)``.  The
    features of a draw are its normalised perturbations
    (c / c0 - 1) / 0.2 ~ N(0,1) and (b / b0 - 1) / 0.1 ~ U[-1,1] (0 where b0 is 0),
    n + m values.  With rank > 0 the perturbations of draw k are instead
    fixed mixtures of `rank` latent N(0,1) factors z from default_rng(k)
    (c: U_c z, each entry ~N(0,1); b: tanh(U_b z); U from default_rng(seed + 1)):
    a family whose solutions move along a few directions only."""

    def __init__(self, periods=12, blk_rows=2000, blk_cols=5000, inv_per_link=200, seed=42, rank=0):
        self.base = staircase_lp(periods, blk_rows, blk_cols, inv_per_link, seed)
        self.m, self.n = self.base.m, self.base.n
        self.rank = rank
        self.meta = dict(kind="staircase_family", periods=periods, blk_rows=blk_rows, blk_cols=blk_cols,
                         inv_per_link=inv_per_link, seed=seed, nnz=self.base.nnz, rank=rank)
        if rank:
            g = np.random.default_rng(seed + 1)
            self._uc = g.standard_normal((self.n, rank)) / np.sqrt(rank)
            self._ub = g.standard_normal((self.m, rank)) / np.sqrt(rank)

    def perturbation(self, k):
        rng = np.random.default_rng(k)
        if self.rank:
            z = rng.standard_normal(self.rank)
            return self._uc @ z, np.tanh(self._ub @ z)
        ec = rng.standard_normal(self.n)
        eb = rng.uniform(-1.0, 1.0, self.m)
        return ec, eb

    def features(self, k):
        ec, eb = self.perturbation(k)
        eb = np.where(self.base.b != 0.0, eb, 0.0)
        return np.concatenate([ec, eb]).astype(np.float32)

    def draw(self, k):
        ec, eb = self.perturbation(k)
        b0 = self.base
        return StdLP(self.m, self.n, b0.row_ptr, b0.col_idx, b0.val, b0.b * (1.0 + 0.1 * eb),
                     b0.c * (1.0 + 0.2 * ec), b0.lb, b0.ub, b0.is_eq, meta=dict(self.meta, draw=int(k)))


def benchmark10_model():
    """Deterministic 100 x 50 model of the reference's Benchmark 5-10
    (benchmarks/benchmark_main.cpp:139-151): maximise sum ((j % 7) + 1) x_j,
    0 <= x_j <= 100, rows i: sum_k (k+1) x_{(3i+k) % 100} <= 50 + i."""
    mb = ModelBuilder(sense=1)
    xs = [mb.add_variable(0.0, 100.0, float((j % 7) + 1)) for j in range(100)]
    for i in range(50):
        mb.add_constraint([(xs[(i * 3 + k) % 100], float(k + 1)) for k in range(5)], LE, 50.0 + i)
    return mb.build()


def flat_to_std(fm: FlatModel) -> StdLP:
    """Standard form of a model without presolve reductions (all rows kept):
    c scaled by the objective sense, GE rows negated, EQ flagged.  RANGED rows
    are split into a <= upper row and an appended -a x <= -lower row."""
    sm = -1.0 if fm.sense == 1 else 1.0
    rows, cols, vals, b, eq = [], [], [], [], []
    extra = []
    for i in range(fm.num_cons):
        s = int(fm.con_sense[i])
        lo, hi = fm.term_ptr[i], fm.term_ptr[i + 1]
        mult = -1.0 if s == GE else 1.0
        rhs = fm.rhs[i]
        if s == GE:
            rhs = -rhs
        if s == RANGED:
            rhs = fm.range_upper[i]
            extra.append(i)
        for k in range(lo, hi):
            rows.append(i)
            cols.append(int(fm.term_var[k]))
            vals.append(float(fm.term_val[k]) * mult)
        b.append(rhs)
        eq.append(1 if s == EQ else 0)
    for e, i in enumerate(extra):
        r = fm.num_cons + e
        for k in range(fm.term_ptr[i], fm.term_ptr[i + 1]):
            rows.append(r)
            cols.append(int(fm.term_var[k]))
            vals.append(-float(fm.term_val[k]))
        b.append(-fm.rhs[i])
        eq.append(0)
    m = fm.num_cons + len(extra)
    rows = np.array(rows, np.int64)
    cols = np.array(cols, np.int64)
    vals = np.array(vals, float)
    # sum duplicates, drop zeros (the CSR construction rule)
    key = rows * max(fm.num_vars, 1) + cols
    uk, inv = np.unique(key, return_inverse=True)
    sv = np.zeros(uk.size)
    np.add.at(sv, inv, vals)
    keep = sv != 0.0
    uk, sv = uk[keep], sv[keep]
    r = uk // max(fm.num_vars, 1)
    c = uk % max(fm.num_vars, 1)
    row_ptr, col_idx, val = _csr_from_cols(m, fm.num_vars, c, r, sv)
    return StdLP(m, fm.num_vars, row_ptr, col_idx, val, np.array(b, float), fm.obj * sm,
                 np.asarray(fm.lb, float), np.asarray(fm.ub, float), np.array(eq, np.uint8))
