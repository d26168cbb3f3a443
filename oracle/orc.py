"""ctypes access to the two CPU checkers (TEST INFRASTRUCTURE — see orc_abi.h).

    lib = orc.load("oracle")      # oracle/liboracle.so   — our restatement (orc_* symbols)
    lib = orc.load("ref")         # oracle/_ref/libps26119_ref.so — reference sources (ref_*)
    lib = orc.load("ref_hip")     # the same with the MI355X backend behind its make_backend(true)
    res, vecs = lib.solve(lp, presolve=False)

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATHS = {"oracle": os.path.join(HERE, "liboracle.so"), "ref": os.path.join(HERE, "_ref", "libps26119_ref.so"),
         "ref_hip": os.path.join(HERE, "_ref", "libps26119_ref_hip.so")}
PREFIX = {"oracle": "orc_", "ref": "ref_", "ref_hip": "ref_"}

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


class OrcLp(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int), ("n", ctypes.c_int), ("sense", ctypes.c_int), ("offset", ctypes.c_double),
                ("c", D), ("col_lower", D), ("col_upper", D), ("row_lower", D), ("row_upper", D),
                ("col_start", I), ("row_index", I), ("value", D)]


class OrcOpts(ctypes.Structure):
    _fields_ = [("algorithm", ctypes.c_int), ("precision", ctypes.c_int), ("presolve", ctypes.c_int),
                ("threads", ctypes.c_int), ("check_every", ctypes.c_int), ("tolerance", ctypes.c_double),
                ("time_limit", ctypes.c_double), ("iteration_limit", ctypes.c_longlong), ("warm_x", D),
                ("warm_y", D), ("warm_primal_weight", ctypes.c_double), ("n_params", ctypes.c_int),
                ("param_names", ctypes.POINTER(ctypes.c_char_p)), ("param_values", D), ("use_gpu", ctypes.c_int)]


class OrcResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_longlong), ("iterations_to_fast", ctypes.c_longlong),
                ("objective", ctypes.c_double), ("dual_objective", ctypes.c_double),
                ("primal_residual", ctypes.c_double), ("dual_residual", ctypes.c_double), ("gap", ctypes.c_double),
                ("certified_bound", ctypes.c_double), ("primal_weight", ctypes.c_double),
                ("seconds", ctypes.c_double), ("setup_seconds", ctypes.c_double), ("check_primal", ctypes.c_double),
                ("check_dual", ctypes.c_double), ("check_gap", ctypes.c_double), ("check", ctypes.c_int),
                ("engine", ctypes.c_char * 32), ("message", ctypes.c_char * 512)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["engine"] = self.engine.decode()
        d["message"] = self.message.decode()
        return d


def _p(a, t=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(t))


class LpHandle:
    """Keeps the numpy arrays alive for as long as the struct is used."""

    def __init__(self, lp):
        self.arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in
                       (lp.c, lp.col_lower, lp.col_upper, lp.row_lower, lp.row_upper, lp.value)]
        self.ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (lp.col_start, lp.row_index)]
        c, cl, cu, rl, ru, v = self.arrays
        cs, ri = self.ints
        self.s = OrcLp(lp.m, lp.n, lp.sense, lp.offset, _p(c), _p(cl), _p(cu), _p(rl), _p(ru),
                       _p(cs, ctypes.c_int), _p(ri, ctypes.c_int), _p(v))


class OrcSolutionFile(ctypes.Structure):
    _fields_ = [(k, ctypes.c_char_p) for k in ("status", "engine", "precision", "message", "check")] + \
               [(k, ctypes.c_double) for k in ("objective", "dual_objective", "certified_bound", "primal_residual",
                                               "dual_residual", "gap")] + \
               [("iterations", ctypes.c_longlong), ("seconds", ctypes.c_double), ("setup_seconds", ctypes.c_double),
                ("primal_weight", ctypes.c_double), ("iterations_to_fast", ctypes.c_longlong)] + \
               [(k, ctypes.c_double) for k in ("seconds_to_fast", "check_primal", "check_dual", "check_gap")] + \
               [(k, D) for k in ("x", "z", "activity", "y")] + \
               [("dual_ray_len", ctypes.c_longlong), ("primal_ray_len", ctypes.c_longlong), ("dual_ray", D),
                ("primal_ray", D)]


class OrcScenario(ctypes.Structure):
    _fields_ = [("obj", D), ("col_lower", D), ("col_upper", D), ("row_lower", D), ("row_upper", D)]


class Checker:
    def __init__(self, kind: str):
        path = PATHS[kind]
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.kind = kind
        self.lib = ctypes.CDLL(path)
        pre = PREFIX[kind]
        L = ctypes.POINTER(OrcLp)

        def fn(name, argtypes, restype=ctypes.c_int):
            f = getattr(self.lib, pre + name)
            f.argtypes = argtypes
            f.restype = restype
            return f

        self._solve = fn("solve", [L, ctypes.POINTER(OrcOpts), ctypes.POINTER(OrcResult), D, D, D, D, D])
        self._scale = fn("scale", [L, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   D, D, D, D, D, D, D])
        self._norm2 = fn("norm2", [L], ctypes.c_double)
        self._kkt = fn("kkt", [L, D, D, D])
        self._ops = fn("backend_ops", [L, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, D, D, D, D, D, D, D])
        self._bound = fn("certified_bound", [L, D, ctypes.c_int], ctypes.c_double)
        self._batch = fn("solve_batch", [L, ctypes.c_int, ctypes.POINTER(OrcScenario), ctypes.POINTER(OrcOpts),
                                         ctypes.POINTER(OrcResult), D, D])
        self._check = fn("check_solution", [L, D, D, D])
        self._write_sol = (fn("write_solution", [ctypes.c_char_p, ctypes.POINTER(OrcSolutionFile), ctypes.c_char_p])
                           if hasattr(self.lib, PREFIX[kind] + "write_solution") else None)
        I = ctypes.POINTER(ctypes.c_int)
        self._presolve = (fn("presolve", [L, D, I, I, D, D, D, D, I, I, D])
                          if hasattr(self.lib, PREFIX[kind] + "presolve") else None)

    def solve(self, lp, precision=0, presolve=True, tolerance=1e-8, iteration_limit=0, time_limit=0.0,
              threads=1, check_every=0, warm_x=None, warm_y=None, warm_primal_weight=0.0, params=None, algorithm=3,
              use_gpu=0):
        """algorithm: 3 = r²HPDHG, 2 = PDLP (the reference build only). use_gpu = 1 (the ref_hip build
        only) runs the reference's drivers on the MI355X backend."""
        h = LpHandle(lp)
        names = [k.encode() for k in (params or {})]
        vals = np.array(list((params or {}).values()), dtype=np.float64)
        name_arr = (ctypes.c_char_p * max(1, len(names)))(*names)
        wx = None if warm_x is None else np.ascontiguousarray(warm_x, dtype=np.float64)
        wy = None if warm_y is None else np.ascontiguousarray(warm_y, dtype=np.float64)
        o = OrcOpts(algorithm, precision, int(presolve), threads, check_every, tolerance, time_limit, iteration_limit,
                    _p(wx), _p(wy), warm_primal_weight, len(names), name_arr, _p(vals) if len(names) else None,
                    int(use_gpu))
        r = OrcResult()
        x, y, z = np.full(lp.n, np.nan), np.full(lp.m, np.nan), np.full(lp.n, np.nan)
        dr, pr = np.full(lp.m, np.nan), np.full(lp.n, np.nan)
        self._solve(ctypes.byref(h.s), ctypes.byref(o), ctypes.byref(r), _p(x), _p(y), _p(z), _p(dr), _p(pr))
        return r.as_dict(), {"x": x, "y": y, "z": z, "dual_ray": dr, "primal_ray": pr}

    def solve_batch(self, lp, scenarios, tolerance=1e-8, iteration_limit=0, check_every=0, params=None):
        """scenarios: list of dicts with optional keys obj, col_lower, col_upper, row_lower, row_upper."""
        h = LpHandle(lp)
        K = len(scenarios)
        keep = []
        arr = (OrcScenario * max(1, K))()
        for k, sc in enumerate(scenarios):
            vals = []
            for key, length in (("obj", lp.n), ("col_lower", lp.n), ("col_upper", lp.n), ("row_lower", lp.m),
                                ("row_upper", lp.m)):
                a = sc.get(key)
                a = None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()
                if a is not None and a.size != length:  # same verdict as a wrong size: "non-finite data"
                    a = np.full(length, np.nan)
                keep.append(a)
                vals.append(_p(a))
            arr[k] = OrcScenario(*vals)
        names = [k.encode() for k in (params or {})]
        pv = np.array(list((params or {}).values()), dtype=np.float64)
        name_arr = (ctypes.c_char_p * max(1, len(names)))(*names)
        o = OrcOpts(3, 0, 0, 1, check_every, tolerance, 0.0, iteration_limit, None, None, 0.0, len(names), name_arr,
                    _p(pv) if len(names) else None)
        res = (OrcResult * max(1, K))()
        xs, ys = np.full(K * lp.n, np.nan), np.full(K * lp.m, np.nan)
        self._batch(ctypes.byref(h.s), K, arr, ctypes.byref(o), res, _p(xs), _p(ys))
        return [res[k].as_dict() for k in range(K)], xs.reshape(K, lp.n), ys.reshape(K, lp.m)

    def scale(self, lp, geo_iters=12, geo_min_range=4.5, ruiz=10, pock_chambolle=True, bound_objective_rescaling=True):
        h = LpHandle(lp)
        rs, cs, sc = np.zeros(lp.m), np.zeros(lp.n), np.zeros(4)
        av, atv, cc = np.zeros(lp.nnz), np.zeros(lp.nnz), np.zeros(lp.n)
        bnd = np.zeros(2 * lp.n + 2 * lp.m)
        self._scale(ctypes.byref(h.s), geo_iters, geo_min_range, ruiz, int(pock_chambolle),
                    int(bound_objective_rescaling), _p(rs), _p(cs), _p(sc), _p(av), _p(atv), _p(cc), _p(bnd))
        return {"row_scale": rs, "col_scale": cs, "bound_scale": sc[0], "obj_scale": sc[1], "log10_range": sc[2],
                "geometric_mean_applied": bool(sc[3]), "a_vals": av, "at_vals": atv, "c": cc,
                "col_lower": bnd[:lp.n], "col_upper": bnd[lp.n:2 * lp.n], "row_lower": bnd[2 * lp.n:2 * lp.n + lp.m],
                "row_upper": bnd[2 * lp.n + lp.m:]}

    def presolve(self, lp):
        """The reduced model as arrays (the oracle build only): info = {outcome (0 reduced,
        1 infeasible, 2 unchanged), rows removed, cols removed, rows, cols, nnz, offset}."""
        h = LpHandle(lp)
        info = np.zeros(7)
        I = ctypes.c_int
        out = {"kept_rows": np.zeros(lp.m, np.int32), "kept_cols": np.zeros(lp.n, np.int32), "rl": np.zeros(lp.m),
               "ru": np.zeros(lp.m), "cl": np.zeros(lp.n), "cu": np.zeros(lp.n),
               "col_start": np.zeros(lp.n + 1, np.int32), "row_index": np.zeros(max(1, lp.nnz), np.int32),
               "value": np.zeros(max(1, lp.nnz))}
        self._presolve(ctypes.byref(h.s), _p(info), _p(out["kept_rows"], I), _p(out["kept_cols"], I), _p(out["rl"]),
                       _p(out["ru"]), _p(out["cl"]), _p(out["cu"]), _p(out["col_start"], I), _p(out["row_index"], I),
                       _p(out["value"]))
        return info, _trim(info, out)

    def norm2(self, lp) -> float:
        h = LpHandle(lp)
        return self._norm2(ctypes.byref(h.s))

    def kkt(self, lp, x, ymin):
        h = LpHandle(lp)
        out = np.zeros(10)
        x = np.ascontiguousarray(x, dtype=np.float64)
        ymin = np.ascontiguousarray(ymin, dtype=np.float64)
        self._kkt(ctypes.byref(h.s), _p(x), _p(ymin), _p(out))
        keys = ["primal_residual", "dual_residual", "primal_obj", "dual_obj", "b_norm", "c_norm", "primal_max_rel",
                "dual_max_rel", "rel_kkt", "converged"]
        return dict(zip(keys, out))

    def backend_ops(self, lp, precision=0, iters=20, tau=0.3, sigma=0.4, rho=1.0):
        h = LpHandle(lp)
        outs = [np.zeros(lp.n) for _ in range(3)] + [np.zeros(lp.m) for _ in range(3)]
        red = np.zeros(11)
        self._ops(ctypes.byref(h.s), precision, iters, tau, sigma, rho, *[_p(o) for o in outs], _p(red))
        return dict(zip(["x", "xh", "xb", "y", "yh", "yb"], outs)), red

    def certified_bound(self, lp, y, use_implied=True) -> float:
        h = LpHandle(lp)
        y = np.ascontiguousarray(y, dtype=np.float64)
        return self._bound(ctypes.byref(h.s), _p(y), int(use_implied))

    def write_solution(self, lpm_path, fields, out_path):
        """The reference's solution writer (the reference build only) on the model its .lpm reader
        reads from lpm_path and `fields` (the keys of OrcSolutionFile; vectors as sequences, absent
        keys = no such block). Returns the writer's code (0 = written)."""
        keep = {k: np.ascontiguousarray(fields[k], dtype=np.float64)
                for k in ("x", "z", "activity", "y", "dual_ray", "primal_ray") if fields.get(k) is not None}
        f = OrcSolutionFile()
        for k, t in OrcSolutionFile._fields_:
            if k in keep:
                setattr(f, k, _p(keep[k]))
            elif k.endswith("_ray_len"):
                ray = keep.get(k[:-4])
                setattr(f, k, 0 if ray is None else ray.size)
            elif t is ctypes.c_char_p:
                setattr(f, k, fields.get(k, "").encode())
            elif k in fields:
                setattr(f, k, fields[k])
        return self._write_sol(lpm_path.encode(), ctypes.byref(f), out_path.encode())

    def check_solution(self, lp, x, y):
        h = LpHandle(lp)
        out = np.zeros(6)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        self._check(ctypes.byref(h.s), _p(x), _p(y), _p(out))
        return dict(zip(["max_primal_violation", "max_dual_violation", "relative_gap", "primal_objective",
                         "dual_objective", "passed"], out))


def _trim(info, out):
    """The reduced model's arrays cut to its sizes (nothing when presolve did not reduce)."""
    if info[0] != 0:
        return {}
    m, n, nz = int(info[3]), int(info[4]), int(info[5])
    return {"kept_rows": out["kept_rows"][:m], "kept_cols": out["kept_cols"][:n], "rl": out["rl"][:m],
            "ru": out["ru"][:m], "cl": out["cl"][:n], "cu": out["cu"][:n], "col_start": out["col_start"][:n + 1],
            "row_index": out["row_index"][:nz], "value": out["value"][:nz]}


def load(kind: str) -> Checker:
    return Checker(kind)


def available(kind: str) -> bool:
    return os.path.exists(PATHS[kind])
