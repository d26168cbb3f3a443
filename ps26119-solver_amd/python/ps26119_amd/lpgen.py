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
reference_bench_lp   the reference's own benchmark family rand-<m>-s<seed> (its
                     bench/generate_lp.py), restated draw for draw so that the same seed gives
                     the same instance as the reference's published scale runs.
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


def _plant_duals(rng, m, n, col_start, row_index, value, x, status, eq, ge, le, ranged, active, slack):
    """Row bounds, duals and costs around a primal point x (the KKT construction of the
    reference's bench/lpgen.py:67-118, draw for draw): an equality row's bounds are its activity
    (snapped to 0 below 1e-12) with a dual ±U(0.5, 5); an active ≥ / ≤ row is tight with dual
    +/−U(0.5, 5); an inactive one, and a ranged row, keeps a slack U(slack)·(1 + |activity|);
    columns at a bound get reduced costs ±U(0.5, 5) (sign by the bound), others 0, and
    c = Aᵀy + z. Returns c, rl, ru, y and the optimum cᵀx summed exactly."""
    import math
    cidx = np.repeat(np.arange(n), np.diff(col_start))
    ax = np.bincount(row_index, weights=value * x[cidx], minlength=m)
    slack_lo = rng.uniform(slack[0], slack[1], m) * (1 + np.abs(ax))
    slack_up = rng.uniform(slack[0], slack[1], m) * (1 + np.abs(ax))
    dual_mag = rng.uniform(0.5, 5.0, m)
    y = np.zeros(m)
    rl = np.full(m, -INF)
    ru = np.full(m, INF)
    snapped = np.where(np.abs(ax) < 1e-12, 0.0, ax)
    rl[eq] = snapped[eq]
    ru[eq] = snapped[eq]
    y[eq] = dual_mag[eq] * rng.choice([-1.0, 1.0], int(eq.sum()))
    rl[ge & active] = ax[ge & active]
    y[ge & active] = dual_mag[ge & active]
    rl[ge & ~active] = (ax - slack_lo)[ge & ~active]
    ru[le & active] = ax[le & active]
    y[le & active] = -dual_mag[le & active]
    ru[le & ~active] = (ax + slack_up)[le & ~active]
    rl[ranged] = (ax - slack_lo)[ranged]
    ru[ranged] = (ax + slack_up)[ranged]
    z_mag = rng.uniform(0.5, 5.0, n)
    z = np.where(status < 0, z_mag, np.where(status > 0, -z_mag, 0.0))
    c = np.bincount(cidx, weights=value * y[row_index], minlength=n) + z
    return c, rl, ru, y, math.fsum((c * x).tolist())


def _summed_csc(m, n, r_idx, c_idx, vals):
    """CSC of triplets with duplicates summed in input order and exact zeros dropped (the
    reference's bench/lpgen.py:37-54)."""
    key = c_idx.astype(np.int64) * m + r_idx.astype(np.int64)
    order = np.argsort(key, kind="stable")
    key, vals = key[order], vals[order]
    uniq, first = np.unique(key, return_index=True)
    summed = np.add.reduceat(vals, first) if len(vals) else vals
    nz = summed != 0.0
    uniq, summed = uniq[nz], summed[nz]
    col_start = np.zeros(n + 1, dtype=np.int64)
    np.add.at(col_start, uniq // m + 1, 1)
    return np.cumsum(col_start), uniq % m, summed


def reference_bench_lp(m: int, seed: int = 1, cols_per_row: float = 1.5, nnz_per_col: int = 4) -> Lp:
    """The reference's benchmark LP rand-<m>-s<seed> (reference bench/generate_lp.py:25-57 with
    the KKT construction of bench/lpgen.py:37-118), restated here: the random draws must come in
    the reference's order from numpy's default_rng(seed) for the instance — and its planted
    optimum — to be the reference's. n = round(1.5 m) columns, each with 4 random rows (duplicates
    summed), then one entry for every row left empty; values ±U(0.5, 2). Columns: 40% at the lower
    bound 0, 10% at an upper bound, 50% between (a tenth of those free); rows: 20% equalities,
    40% active one-sided (half ≥, half ≤), 40% inactive (ranged or one-sided). Duals of active
    rows are ±U(0.5, 5), reduced costs of bound columns ±U(0.5, 5), and c = Aᵀy + z, so the
    planted point is optimal; its objective is summed exactly (math.fsum), as there."""
    rng = np.random.default_rng(seed)
    n = int(round(cols_per_row * m))
    k = min(nnz_per_col, m)
    # the matrix: k draws per column, then one random column for each row no draw hit
    r_idx = rng.integers(0, m, size=(n, k)).ravel()
    c_idx = np.repeat(np.arange(n), k)
    empty = np.setdiff1d(np.arange(m), r_idx)
    r_idx = np.concatenate([r_idx, empty])
    c_idx = np.concatenate([c_idx, rng.integers(0, n, len(empty))])
    magnitude = rng.uniform(0.5, 2.0, len(r_idx))
    vals = magnitude * rng.choice([-1.0, 1.0], len(r_idx))
    col_start, row_index, value = _summed_csc(m, n, r_idx, c_idx, vals)
    # column statuses and bounds around x*
    draw = rng.random(n)
    status = np.where(draw < 0.4, -1, np.where(draw < 0.5, 1, 0))  # -1 at lower, +1 at upper
    free = (status == 0) & (rng.random(n) < 0.1)
    lower = np.where(free, -INF, 0.0)
    x = np.where(status < 0, 0.0, rng.uniform(0.5, 5.0, n))
    finite_up = rng.random(n) < 0.4
    upper = np.where(status > 0, x, np.where(finite_up & ~free, x + rng.uniform(0.5, 5.0, n), INF))
    upper = np.where(status < 0, np.where(finite_up, rng.uniform(0.5, 5.0, n), INF), upper)
    x = np.where(free, rng.uniform(-5.0, 5.0, n), x)
    upper = np.where(free, INF, upper)
    # row types: 20% eq, 20% ≥ and 20% ≤ active, 20% ranged, 10% ≥ and 10% ≤ inactive
    t = rng.random(m)
    eq, ge, le, ranged = t < 0.2, ((t >= 0.2) & (t < 0.4)) | ((t >= 0.8) & (t < 0.9)), \
        ((t >= 0.4) & (t < 0.6)) | (t >= 0.9), (t >= 0.6) & (t < 0.8)
    active = (t >= 0.2) & (t < 0.6)
    c, rl, ru, y, optimum = _plant_duals(rng, m, n, col_start, row_index, value, x, status, eq, ge, le, ranged, active,
                                         slack=(0.1, 2.0))
    return Lp(m=m, n=n, c=c, col_lower=lower, col_upper=upper, row_lower=rl, row_upper=ru,
              col_start=col_start.astype(np.int32), row_index=row_index.astype(np.int32), value=value,
              optimum=optimum, name=f"rand-{m}-s{seed}", x_star=x, y_star=y)


# The reference's refinery model (bench/generate_refinery_lp.py:45-54 there): six crudes bought,
# run through the crude unit and stored; six cuts; an FCC unit with four products; ten blend
# components routed over fifteen arcs into five products with six quality specs; product tanks
# and demands. Each period has 59 columns and 49 rows in this order.
_RF_CUTS, _RF_FCC, _RF_PRODS = 6, 4, 5
_RF_COMPS = ("lpg", "nap", "kero", "dsl", "vgo", "res", "fcc_lpg", "fcc_gas", "lco", "slurry")
_RF_ARCS = ((0, 0), (6, 0), (1, 1), (7, 1), (0, 1), (2, 2), (2, 3), (3, 3), (8, 3), (4, 4), (5, 4), (9, 4), (8, 4),
            (3, 4), (1, 4))  # (component, product)
_RF_SPECS = ((1, 0, "le"), (1, 1, "ge"), (2, 0, "le"), (3, 0, "le"), (4, 0, "le"), (4, 1, "ge"))  # (product, quality, sense)
_RF_FY = np.array([0.15, 0.50, 0.25, 0.10])  # FCC yields


def _rf_layout():
    """Offsets of the column and row groups of one period."""
    cols, rows, at = {}, {}, 0
    for name, size in (("buy", 6), ("cdu", 6), ("cinv", 6), ("prod", _RF_CUTS), ("fcc", 1), ("fout", _RF_FCC),
                       ("blend", len(_RF_ARCS)), ("out", _RF_PRODS), ("pinv", _RF_PRODS), ("sell", _RF_PRODS)):
        cols[name], at = np.arange(at, at + size), at + size
    ncols, at = at, 0
    for name, size in (("cbal", 6), ("cdu", 1), ("yield", _RF_CUTS), ("fcccap", 1), ("fy", _RF_FCC),
                       ("comp", len(_RF_COMPS)), ("prod", _RF_PRODS), ("spec", len(_RF_SPECS)), ("pbal", _RF_PRODS),
                       ("dem", _RF_PRODS)):
        rows[name], at = np.arange(at, at + size), at + size
    return cols, rows, ncols, at


def reference_refinery_lp(T: int, seed: int = 1) -> Lp:
    """The reference's refinery benchmark LP refinery-T<T>-s<seed> (reference
    bench/generate_refinery_lp.py:88-276 with the KKT construction of bench/lpgen.py), restated
    draw for draw so that a seed gives the reference's instance and planted optimum: a physically
    consistent flow x* simulated period by period (crude unit at capacity in 30% of periods, FCC
    cap binding in 30%, some crude purchases at their availability limit, tanks at capacity when
    full), then duals and prices planted around it with slacks U(0.05, 0.5); a MAX problem."""
    rng = np.random.default_rng(seed)
    C, R, ncols, nrows = _rf_layout()
    K, J, P = 6, _RF_CUTS, _RF_PRODS
    yields = rng.dirichlet(np.ones(J) * 3, K)  # crude k → cut j, rows sum to 1
    quality = np.stack([rng.uniform(0.1, 3.0, len(_RF_COMPS)), rng.uniform(60, 100, len(_RF_COMPS))], 1)
    cdu_cap, crude_tank, prod_tank = 1000.0, 3000.0, 2000.0
    # which period columns feed each component's balance row
    supply = [[(C["prod"][0], 1.0)], [(C["prod"][1], 1.0)], [(C["prod"][2], 1.0)], [(C["prod"][3], 1.0)],
              [(C["prod"][4], 1.0), (C["fcc"][0], -1.0)], [(C["prod"][5], 1.0)], [(C["fout"][0], 1.0)],
              [(C["fout"][1], 1.0)], [(C["fout"][2], 1.0)], [(C["fout"][3], 1.0)]]
    # the period template: (row, column, value, period offset of the column)
    tpl = []
    for k in range(K):
        tpl += [(R["cbal"][k], C["cinv"][k], 1.0, 0), (R["cbal"][k], C["cinv"][k], -1.0, -1),
                (R["cbal"][k], C["buy"][k], -1.0, 0), (R["cbal"][k], C["cdu"][k], 1.0, 0), (R["cdu"][0], C["cdu"][k], 1.0, 0)]
    for j in range(J):
        tpl.append((R["yield"][j], C["prod"][j], 1.0, 0))
        tpl += [(R["yield"][j], C["cdu"][k], -yields[k, j], 0) for k in range(K)]
    tpl.append((R["fcccap"][0], C["fcc"][0], 1.0, 0))
    for o in range(_RF_FCC):
        tpl += [(R["fy"][o], C["fout"][o], 1.0, 0), (R["fy"][o], C["fcc"][0], -_RF_FY[o], 0)]
    for ci, terms in enumerate(supply):
        tpl += [(R["comp"][ci], col, coef, 0) for col, coef in terms]
    for a, (ci, pi) in enumerate(_RF_ARCS):
        tpl += [(R["comp"][ci], C["blend"][a], -1.0, 0), (R["prod"][pi], C["blend"][a], -1.0, 0)]
    tpl += [(R["prod"][p], C["out"][p], 1.0, 0) for p in range(P)]
    for s_, (pi, qi, _) in enumerate(_RF_SPECS):
        arcs = [a for a, (_, pp) in enumerate(_RF_ARCS) if pp == pi]
        limit = float(np.mean(quality[[_RF_ARCS[a][0] for a in arcs], qi]))
        tpl += [(R["spec"][s_], C["blend"][a], quality[_RF_ARCS[a][0], qi] - limit, 0) for a in arcs]
    for p in range(P):
        tpl += [(R["pbal"][p], C["pinv"][p], 1.0, 0), (R["pbal"][p], C["pinv"][p], -1.0, -1),
                (R["pbal"][p], C["out"][p], -1.0, 0), (R["pbal"][p], C["sell"][p], 1.0, 0), (R["dem"][p], C["sell"][p], 1.0, 0)]
    tr, tc, tv, ts = (np.asarray(v) for v in zip(*tpl))
    period = np.repeat(np.arange(T), len(tr))
    col_period = period + np.tile(ts, T)
    kept = col_period >= 0  # period 0 has no previous inventory column
    m, n = T * nrows, T * ncols
    col_start, row_index, value = _summed_csc(m, n, (np.tile(tr, T) + period * nrows)[kept],
                                              (np.tile(tc, T) + col_period * ncols)[kept], np.tile(tv, T)[kept])
    # the flows x*, period by period
    x = np.zeros((T, ncols))
    share = rng.dirichlet(np.ones(K), T)
    load = rng.uniform(0.6, 1.0, T)
    cdu_full = rng.random(T) < 0.3
    load[cdu_full] = 1.0
    x[:, C["cdu"]] = share * (cdu_cap * load)[:, None]
    x[:, C["prod"]] = x[:, C["cdu"]] @ yields
    x[:, C["fcc"][0]] = rng.uniform(0.5, 0.9, T) * x[:, C["prod"][4]]
    fcc_full = rng.random(T) < 0.3
    x[:, C["fout"]] = x[:, [C["fcc"][0]]] * _RF_FY[None, :]
    avail = np.zeros((T, len(_RF_COMPS)))
    for ci, terms in enumerate(supply):
        for col, coef in terms:
            avail[:, ci] += coef * x[:, col]
    for ci in range(len(_RF_COMPS)):  # each component split over its arcs, the last arc takes the rest
        arcs = [a for a, (cc, _) in enumerate(_RF_ARCS) if cc == ci]
        split = rng.dirichlet(np.ones(len(arcs)), T) * avail[:, [ci]]
        split[:, -1] = avail[:, ci] - split[:, :-1].sum(1)
        x[:, C["blend"][arcs]] = split
    for p in range(P):
        x[:, C["out"][p]] = x[:, C["blend"][[a for a, (_, pp) in enumerate(_RF_ARCS) if pp == p]]].sum(1)
    crude, stock = np.full(K, 500.0), np.full(P, 300.0)
    for t in range(T):  # purchases around consumption and sales around output, inventories within tanks
        buy = np.clip(x[t, C["cdu"]] * rng.uniform(0.7, 1.3, K), 0, None)
        buy = np.minimum(buy, crude_tank - crude + x[t, C["cdu"]])
        crude = crude + buy - x[t, C["cdu"]]
        short = crude < 0
        buy[short] -= crude[short]
        crude[short] = 0.0
        x[t, C["buy"]], x[t, C["cinv"]] = buy, crude
        sell = x[t, C["out"]] * rng.uniform(0.7, 1.3, P)
        nxt = stock + x[t, C["out"]] - sell
        under = nxt < 0
        sell[under] += nxt[under]
        nxt[under] = 0.0
        over = nxt > prod_tank
        sell[over] += nxt[over] - prod_tank
        nxt[over] = prod_tank
        stock = nxt
        x[t, C["sell"]], x[t, C["pinv"]] = sell, stock
    xs = np.maximum(x, 0.0).ravel()
    # bounds and the columns that rest on one
    lower, upper = np.zeros(n), np.full(n, INF)
    status = np.zeros(n, dtype=np.int8)
    status[xs == 0.0] = -1
    base = np.arange(T)[:, None] * ncols
    buys = (base + C["buy"][None, :]).ravel()
    at_limit = (rng.random(buys.size) < 0.2) & (xs[buys] > 0)
    upper[buys] = np.where(at_limit, xs[buys], xs[buys] * 1.5 + 50.0)
    status[buys[at_limit]] = 1
    cinvs, pinvs = (base + C["cinv"][None, :]).ravel(), (base + C["pinv"][None, :]).ravel()
    upper[cinvs], upper[pinvs] = crude_tank, prod_tank
    status[pinvs[xs[pinvs] >= prod_tank]] = 1
    status[cinvs[xs[cinvs] >= crude_tank]] = 1
    # row senses and which inequalities are tight at x*
    eq, ge, le, active = (np.zeros(m, dtype=bool) for _ in range(4))
    rbase = np.arange(T)[:, None] * nrows
    for g in ("cbal", "yield", "fy", "comp", "prod", "pbal"):
        eq[(rbase + R[g][None, :]).ravel()] = True
    cdu_rows, fcc_rows = np.arange(T) * nrows + R["cdu"][0], np.arange(T) * nrows + R["fcccap"][0]
    le[cdu_rows], le[fcc_rows] = True, True
    active[cdu_rows], active[fcc_rows] = cdu_full, fcc_full
    dem_rows = (rbase + R["dem"][None, :]).ravel()
    le[dem_rows] = True
    active[dem_rows] = rng.random(dem_rows.size) < 0.6
    spec_rows = (rbase + R["spec"][None, :]).ravel()
    spec_ge = np.tile([sense == "ge" for _, _, sense in _RF_SPECS], T)
    ge[spec_rows[spec_ge]], le[spec_rows[~spec_ge]] = True, True
    active[spec_rows] = rng.random(spec_rows.size) < 0.5
    c, rl, ru, y, optimum = _plant_duals(rng, m, n, col_start, row_index, value, xs, status, eq, ge, le,
                                         np.zeros(m, dtype=bool), active, slack=(0.05, 0.5))
    return Lp(m=m, n=n, c=-c, col_lower=lower, col_upper=upper, row_lower=rl, row_upper=ru,
              col_start=col_start.astype(np.int32), row_index=row_index.astype(np.int32), value=value, sense=-1,
              optimum=-optimum, name=f"refinery-T{T}-s{seed}", x_star=xs, y_star=y)


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
