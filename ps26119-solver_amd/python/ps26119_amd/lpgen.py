"""Seeded LP generators with a known optimum, for the workloads this backend is measured on.

Every generator returns an :class:`Lp` in the reference's model contract (CSC matrix,
``inf`` for missing bounds, ``sense`` +1 min / -1 max) together with the optimal objective
value known by construction (a KKT point is planted). All draws come from
``numpy.random.default_rng(seed)``, so a (generator, sizes, seed) triple always yields the
same bits.

random_lp            m rows, n = 2m columns, exactly k distinct uniformly drawn columns per
                     row, values N(0,1); columns in [0, 10]; 40% of x* at a bound, the rest
                     U(1, 9); 70% equality rows b = A x*, 30% rows <= with slack U(0, 1);
                     y* N(0,1) on equality rows, 0 on slack rows; z* U(0.1, 1) signed by the
                     active bound, 0 on interior columns; c = Aᵀy* + z*.
refinery_lp          T periods of a 49 x 60 block with yield/balance coefficients
                     log-uniform in [1e-3, 1e3] and +-1 inventory links between periods t
                     and t+1 (block staircase); capacities U(10, 1000), demands U(1, 100)
                     times a daily sinusoid, costs U(-5, 20) (see the function docstring).
scenario_set         perturbs b, c and upper bounds of a base LP by lognormal(0, sigma).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

INF = np.inf


@dataclass
class Lp:
    m: int
    n: int
    c: np.ndarray
    col_lower: np.ndarray
    col_upper: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    col_start: np.ndarray  # int32, n+1
    row_index: np.ndarray  # int32, nnz
    value: np.ndarray      # float64, nnz
    sense: int = 1
    offset: float = 0.0
    optimum: float = float("nan")  # known optimal objective (c·x* + offset), NaN if unknown
    name: str = ""
    x_star: np.ndarray | None = field(default=None, repr=False)
    y_star: np.ndarray | None = field(default=None, repr=False)

    @property
    def nnz(self) -> int:
        return int(self.col_start[-1])

    def csr(self):
        """(row_ptr, col, val) of A with columns sorted inside each row."""
        cols = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.col_start))
        order = np.lexsort((cols, self.row_index))
        row_ptr = np.zeros(self.m + 1, dtype=np.int64)
        np.cumsum(np.bincount(self.row_index, minlength=self.m), out=row_ptr[1:])
        return row_ptr, cols[order].astype(np.int32), self.value[order]

    def activity(self, x: np.ndarray) -> np.ndarray:
        cols = np.repeat(np.arange(self.n), np.diff(self.col_start))
        return np.bincount(self.row_index, weights=self.value * x[cols], minlength=self.m)


def _csc_from_triplets(m: int, n: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray):
    key = cols.astype(np.int64) * m + rows.astype(np.int64)
    order = np.argsort(key, kind="stable")
    col_start = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(cols, minlength=n), out=col_start[1:])
    return (col_start.astype(np.int32), rows[order].astype(np.int32), vals[order].astype(np.float64))


def random_lp(m: int, n: int | None = None, per_row: int = 6, seed: int = 1) -> Lp:
    """The random LP with a planted optimum (see the module docstring)."""
    rng = np.random.default_rng(seed)
    n = 2 * m if n is None else n
    cols = rng.integers(0, n, size=(m, per_row))
    while True:  # exactly `per_row` DISTINCT columns per row: redraw rows with a repeat
        s = np.sort(cols, axis=1)
        dup = (s[:, 1:] == s[:, :-1]).any(axis=1)
        if not dup.any():
            break
        cols[dup] = rng.integers(0, n, size=(int(dup.sum()), per_row))
    vals = rng.standard_normal((m, per_row))
    at_bound = rng.random(n) < 0.4
    upper = rng.random(n) < 0.5
    x = rng.uniform(1.0, 9.0, n)
    x[at_bound & ~upper] = 0.0
    x[at_bound & upper] = 10.0
    ax = np.zeros(m)
    for k in range(per_row):  # fixed left-to-right order per row
        ax = ax + vals[:, k] * x[cols[:, k]]
    eq = rng.random(m) < 0.7
    slack = rng.uniform(0.0, 1.0, m)
    rl = np.where(eq, ax, -INF)
    ru = np.where(eq, ax, ax + slack)
    y = np.where(eq, rng.standard_normal(m), 0.0)
    z = rng.uniform(0.1, 1.0, n)
    z[~at_bound] = 0.0
    z[at_bound & upper] *= -1.0
    rows = np.repeat(np.arange(m, dtype=np.int64), per_row)
    col_start, row_index, value = _csc_from_triplets(m, n, rows, cols.ravel(), vals.ravel())
    colidx = np.repeat(np.arange(n), np.diff(col_start))
    aty = np.bincount(colidx, weights=value * y[row_index], minlength=n)
    c = aty + z
    opt = float(np.dot(c, x))
    return Lp(m=m, n=n, c=c, col_lower=np.zeros(n), col_upper=np.full(n, 10.0), row_lower=rl, row_upper=ru,
              col_start=col_start, row_index=row_index, value=value, optimum=opt,
              name=f"random-{m}x{n}-k{per_row}-s{seed}", x_star=x, y_star=y)


def scenario_set(base: Lp, k: int, sigma: float = 0.05, seed: int = 7):
    """k scenarios of `base`: b (finite row bounds), c and finite column upper bounds each
    multiplied by an independent lognormal(0, sigma) factor. Returns a list of dicts with
    keys obj, col_upper, row_lower, row_upper."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        fb = rng.lognormal(0.0, sigma, base.m)
        fc = rng.lognormal(0.0, sigma, base.n)
        fu = rng.lognormal(0.0, sigma, base.n)
        rl = np.where(np.isfinite(base.row_lower), base.row_lower * fb, base.row_lower)
        ru = np.where(np.isfinite(base.row_upper), base.row_upper * fb, base.row_upper)
        cu = np.where(np.isfinite(base.col_upper), base.col_upper * fu, base.col_upper)
        out.append({"obj": base.c * fc, "col_upper": cu, "row_lower": rl, "row_upper": ru})
    return out


def with_scenario(base: Lp, sc: dict) -> Lp:
    """The model a scenario describes: `base` with any of obj, col_lower, col_upper, row_lower,
    row_upper replaced."""
    import copy
    m = copy.copy(base)
    for key, attr in (("obj", "c"), ("col_lower", "col_lower"), ("col_upper", "col_upper"),
                      ("row_lower", "row_lower"), ("row_upper", "row_upper")):
        if key in sc:
            setattr(m, attr, np.asarray(sc[key], dtype=np.float64))
    return m


def perturbation_scenarios(base: Lp):
    """The four scenarios of the reference's batch test (tests/unit/test_batch.cpp:28-42 there):
    unchanged; objective scaled by 1 + 0.05·((j mod 3) − 1); finite, non-equality row upper
    bounds loosened by 0.02·(1 + |u|); both changes."""
    j = np.arange(base.n)
    obj = base.c * (1.0 + 0.05 * ((j % 3) - 1.0))
    ru = base.row_upper.copy()
    loosen = np.isfinite(ru) & (base.row_lower != ru)
    ru[loosen] += 0.02 * (1 + np.abs(ru[loosen]))
    rows = {"row_lower": base.row_lower.copy(), "row_upper": ru}
    return [{}, {"obj": obj}, dict(rows), dict(rows, obj=obj)]


def read_lpm(path: str) -> Lp:
    """Reader for the reference's `.lpm` text format, version 1 (sections as in its
    tools/lpm.py specification: header keys, then OBJ, COL_LOWER, COL_UPPER, ROW_LOWER,
    ROW_UPPER, COL_START, ROW_INDEX, VALUE; INTEGER and name sections are skipped)."""
    with open(path) as f:
        lines = [ln.rstrip("\n") for ln in f]
    hdr = {}
    sections = {}
    cur = None
    skip_names = False
    for ln in lines:
        s = ln.strip()
        if not s or s.startswith("#"):
            continue
        head = s.split()[0]
        if head in ("ROW_NAMES", "COL_NAMES"):
            skip_names = True
            cur = None
            continue
        if head == "END":
            break
        if skip_names:
            continue
        if head in ("LPM", "NAME", "SENSE", "ROWS", "COLS", "NNZ", "OFFSET"):
            hdr[head] = s[len(head):].strip()
            cur = None
            continue
        if head in ("OBJ", "COL_LOWER", "COL_UPPER", "ROW_LOWER", "ROW_UPPER", "COL_START", "ROW_INDEX", "VALUE",
                    "INTEGER"):
            cur = head
            sections[cur] = []
            rest = s[len(head):].split()
            sections[cur].extend(rest)
            continue
        if cur is not None:
            sections[cur].extend(s.split())
    m, n = int(hdr["ROWS"]), int(hdr["COLS"])

    def f(key):
        return np.array([float(t) for t in sections.get(key, [])], dtype=np.float64)

    def i(key):
        return np.array([int(t) for t in sections.get(key, [])], dtype=np.int32)

    return Lp(m=m, n=n, c=f("OBJ"), col_lower=f("COL_LOWER"), col_upper=f("COL_UPPER"),
              row_lower=f("ROW_LOWER"), row_upper=f("ROW_UPPER"), col_start=i("COL_START"),
              row_index=i("ROW_INDEX"), value=f("VALUE"), sense=-1 if hdr.get("SENSE", "MIN") == "MAX" else 1,
              offset=float(hdr.get("OFFSET", "0") or 0), name=hdr.get("NAME", ""))


def write_lpm(lp: Lp, path: str) -> None:
    """Writes `lp` in the `.lpm` text format, version 1, as the reference's tools/lpm.py lays it
    out (tools/lpm.py:186-222 of the reference): each section tag on its own line, then its numbers
    eight to a line, doubles as repr (so they round-trip) and infinities as inf / -inf."""
    def block(f, tag, values, ints=False):
        f.write(tag + "\n")
        text = [str(int(x)) if ints else (repr(float(x)) if np.isfinite(x) else ("inf" if x > 0 else "-inf"))
                for x in values]
        for k in range(0, len(text), 8):
            f.write(" ".join(text[k:k + 8]) + "\n")

    with open(path, "w") as f:
        f.write(f"LPM 1\nNAME {lp.name or 'model'}\nSENSE {'MAX' if lp.sense < 0 else 'MIN'}\n")
        f.write(f"ROWS {lp.m}\nCOLS {lp.n}\nNNZ {lp.nnz}\nOFFSET {float(lp.offset)!r}\n")
        for tag, v in (("OBJ", lp.c), ("COL_LOWER", lp.col_lower), ("COL_UPPER", lp.col_upper),
                       ("ROW_LOWER", lp.row_lower), ("ROW_UPPER", lp.row_upper)):
            block(f, tag, v)
        block(f, "COL_START", lp.col_start, ints=True)
        block(f, "ROW_INDEX", lp.row_index, ints=True)
        block(f, "VALUE", lp.value)
        f.write("END\n")


def from_dense(c, A, rl, ru, cl, cu, sense: int = 1, name: str = "") -> Lp:
    """Small models from dense rows (test fixtures)."""
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    rows, cols = np.nonzero(A.T)  # column-major order: cols of A are rows of A.T
    col_start = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=col_start[1:])
    return Lp(m=m, n=n, c=np.asarray(c, float), col_lower=np.asarray(cl, float), col_upper=np.asarray(cu, float),
              row_lower=np.asarray(rl, float), row_upper=np.asarray(ru, float), col_start=col_start,
              row_index=cols.astype(np.int32), value=A.T[rows, cols].astype(np.float64), sense=sense, name=name)


def refinery_lp(T: int = 8760, seed: int = 3) -> Lp:
    """Multi-period planning LP in block-staircase form, feasible by construction.

    Each period t has 60 columns and 49 rows:
      columns  0..10  inventory of product q at the end of period t (bounds [0, U(10, 1000)])
               11..21 purchase of product q, 22..32 sale of product q
               33..59 process activities
      rows     0..10  balance of product q (equality):
                      yields of the activities + inv[t-1, q] − inv[t, q] + purchase − sale = demand
               11..48 capacity rows (<=) over 2–4 activities
    Yield / capacity coefficients are log-uniform in [1e-3, 1e3] with random signs on the yields;
    the inventory columns link periods t and t+1 with +1 / −1. Demands are U(1, 100) times a daily
    sinusoid (period 24); column upper bounds U(10, 1000); costs U(−5, 20). A point x0 inside the
    bounds is drawn first and every capacity row gets rhs = (A x0)_i + slack, purchases and sales
    absorb the balance residual (their bounds widened to contain it), so x0 is feasible; all
    columns are bounded, so the LP has an optimum (not known in closed form)."""
    rng = np.random.default_rng(seed)
    P, R = 60, 49
    n, m = T * P, T * R
    rows, cols, vals = [], [], []

    def logu(k):
        return np.exp(rng.uniform(np.log(1e-3), np.log(1e3), k))

    cap = rng.uniform(10.0, 1000.0, n)
    x0 = rng.uniform(0.0, 1.0, n) * cap * 0.5
    t_idx = np.arange(T)
    phase = rng.uniform(0, 2 * np.pi, 11)
    demand = rng.uniform(1.0, 100.0, (T, 11)) * (1.0 + 0.5 * np.sin(2 * np.pi * t_idx[:, None] / 24.0 + phase[None, :]))
    # structure of one period (same pattern every period, fresh coefficients)
    bal_acts = [rng.choice(np.arange(33, 60), size=rng.integers(2, 5), replace=False) for _ in range(11)]
    cap_acts = [rng.choice(np.arange(33, 60), size=rng.integers(2, 5), replace=False) for _ in range(38)]
    rl = np.empty(m)
    ru = np.empty(m)
    for t in range(T):
        c0, r0 = t * P, t * R
        for q in range(11):
            r = r0 + q
            acts = bal_acts[q]
            coef = logu(len(acts)) * rng.choice([-1.0, 1.0], len(acts))
            s = float(np.dot(coef, x0[c0 + acts]))
            rows.extend([r] * len(acts)), cols.extend((c0 + acts).tolist()), vals.extend(coef.tolist())
            rows.append(r), cols.append(c0 + q), vals.append(-1.0)          # inventory out of period t
            s -= x0[c0 + q]
            if t > 0:
                rows.append(r), cols.append(c0 - P + q), vals.append(1.0)   # inventory from period t-1
                s += x0[c0 - P + q]
            rows.append(r), cols.append(c0 + 11 + q), vals.append(1.0)      # purchase
            rows.append(r), cols.append(c0 + 22 + q), vals.append(-1.0)     # sale
            gap = demand[t, q] - s
            buy, sell = (gap, 0.0) if gap >= 0 else (0.0, -gap)
            x0[c0 + 11 + q], x0[c0 + 22 + q] = buy, sell
            cap[c0 + 11 + q] = max(cap[c0 + 11 + q], 2.0 * buy)
            cap[c0 + 22 + q] = max(cap[c0 + 22 + q], 2.0 * sell)
            rl[r] = ru[r] = demand[t, q]
        for k in range(38):
            r = r0 + 11 + k
            acts = cap_acts[k]
            coef = logu(len(acts))
            rows.extend([r] * len(acts)), cols.extend((c0 + acts).tolist()), vals.extend(coef.tolist())
            rl[r] = -INF
            ru[r] = float(np.dot(coef, x0[c0 + acts])) * rng.uniform(1.05, 2.0)
    cs, ri, v = _csc_from_triplets(m, n, np.asarray(rows), np.asarray(cols), np.asarray(vals))
    c = rng.uniform(-5.0, 20.0, n)
    return Lp(m=m, n=n, c=c, col_lower=np.zeros(n), col_upper=cap, row_lower=rl, row_upper=ru, col_start=cs,
              row_index=ri, value=v, name=f"refinery-T{T}-s{seed}", x_star=x0)
