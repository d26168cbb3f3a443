"""ps26119 — Python interface to the MI355X r²HPDHG backend (ctypes over include/ps26119/ps26119.h).

Same entry points as the reference's binding (its python/ps26119/__init__.py): ``solve_lp``,
``Result``, ``version`` and the status constants, so scripts written against the reference keep
working; the library loaded is this repository's ``libps26119.so``. Because this library serves
the GPU path, ``gpu`` defaults to True here.

    import numpy as np, ps26119
    res = ps26119.solve_lp(c, A, row_lower, row_upper, col_lower, col_upper, sense=-1)
    res.status, res.objective, res.x, res.y, res.z

Extensions: ``presolve``, ``check_every``, ``params`` (named engine knobs, e.g.
``{"hip_cpu_order": 1}``) and ``stats=True`` (setup time, kernel timing) go through
``ps26119_amd_solve_lp``; ``solve_batch`` (scenarios sharing one matrix) goes through
``ps26119_amd_solve_batch``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

__all__ = ["solve_lp", "solve_batch", "Server", "Result", "version", "device_count", "library_path", "OPTIMAL", "INFEASIBLE", "UNBOUNDED",
           "ITERATION_LIMIT", "TIME_LIMIT", "NUMERICAL_ERROR", "NOT_SOLVED", "INVALID_ARGUMENT"]

OPTIMAL, INFEASIBLE, UNBOUNDED, ITERATION_LIMIT, TIME_LIMIT, NUMERICAL_ERROR, NOT_SOLVED, INVALID_ARGUMENT = range(8)
_STATUS = ["Optimal", "Infeasible", "Unbounded", "IterationLimit", "TimeLimit", "NumericalError", "NotSolved",
           "InvalidArgument"]
_ALG = {"auto": 0, "oracle": 1, "pdlp": 2, "r2hpdhg": 3, "simplex": 4}
_PREC = {"fp64": 0, "mixed": 1}
_CHECK = {1: "PASS", 0: "FAIL", -1: ""}

_HERE = os.path.dirname(os.path.abspath(__file__))
_CSRC = os.path.normpath(os.path.join(_HERE, "..", "..", "csrc"))
library_path = os.path.join(_CSRC, "libps26119.so")
hip_library_path = os.path.join(_CSRC, "hip", "libps26119_hip.so")

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


class Options(ctypes.Structure):  # ps26119_options
    _fields_ = [("algorithm", ctypes.c_int), ("precision", ctypes.c_int), ("use_gpu", ctypes.c_int),
                ("tolerance", ctypes.c_double), ("time_limit", ctypes.c_double),
                ("iteration_limit", ctypes.c_longlong), ("verbosity", ctypes.c_int), ("threads", ctypes.c_int),
                ("warm_x", D), ("warm_y", D)]


class CResult(ctypes.Structure):  # ps26119_result
    _fields_ = [("status", ctypes.c_int), ("objective", ctypes.c_double), ("dual_objective", ctypes.c_double),
                ("primal_residual", ctypes.c_double), ("dual_residual", ctypes.c_double), ("gap", ctypes.c_double),
                ("certified_bound", ctypes.c_double), ("iterations", ctypes.c_longlong),
                ("seconds", ctypes.c_double), ("engine", ctypes.c_char * 16), ("message", ctypes.c_char * 160),
                ("check", ctypes.c_int)]


class AmdOptions(ctypes.Structure):  # ps26119_amd_options
    _fields_ = [("presolve", ctypes.c_int), ("check_every", ctypes.c_int), ("n_params", ctypes.c_int),
                ("param_names", ctypes.POINTER(ctypes.c_char_p)), ("param_values", D)]


class AmdStats(ctypes.Structure):  # ps26119_amd_stats
    _fields_ = [("setup_seconds", ctypes.c_double), ("iterations_to_fast", ctypes.c_longlong),
                ("seconds_to_fast", ctypes.c_double), ("primal_weight", ctypes.c_double),
                ("check_primal", ctypes.c_double), ("check_dual", ctypes.c_double), ("check_gap", ctypes.c_double),
                ("kernel_launches", ctypes.c_longlong * 3), ("kernel_ms", ctypes.c_double * 3),
                ("device_bytes", ctypes.c_double), ("iterate_seconds", ctypes.c_double)]

    def as_dict(self):
        return {"setup_seconds": self.setup_seconds, "iterations_to_fast": self.iterations_to_fast,
                "seconds_to_fast": self.seconds_to_fast, "primal_weight": self.primal_weight,
                "check_primal": self.check_primal, "check_dual": self.check_dual, "check_gap": self.check_gap,
                "kernel_launches": list(self.kernel_launches), "kernel_ms": list(self.kernel_ms),
                "device_bytes": self.device_bytes, "iterate_seconds": self.iterate_seconds}


class Scenario(ctypes.Structure):  # ps26119_amd_scenario
    _fields_ = [("obj", D), ("col_lower", D), ("col_upper", D), ("row_lower", D), ("row_upper", D)]


class ServerOptions(ctypes.Structure):  # ps26119_amd_server_options
    _fields_ = [("device", ctypes.c_int), ("budget_bytes", ctypes.c_double), ("max_instances", ctypes.c_int),
                ("check_every", ctypes.c_int), ("setup_threads", ctypes.c_int), ("eager_launches", ctypes.c_int),
                ("timing", ctypes.c_int)]


def _load():
    path = os.environ.get("PS26119_LIB", library_path)
    if not os.path.exists(path):
        raise OSError(f"{path} not found: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    model = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, D, D, D, D, D, I, I, D]
    lib.ps26119_solve_lp.argtypes = model + [ctypes.POINTER(Options), ctypes.POINTER(CResult), D, D, D]
    lib.ps26119_solve_lp.restype = ctypes.c_int
    lib.ps26119_solve_lp_ex.argtypes = lib.ps26119_solve_lp.argtypes + [D, D]
    lib.ps26119_solve_lp_ex.restype = ctypes.c_int
    lib.ps26119_amd_solve_lp.argtypes = model + [ctypes.POINTER(Options), ctypes.POINTER(AmdOptions),
                                                 ctypes.POINTER(CResult), ctypes.POINTER(AmdStats), D, D, D, D, D]
    lib.ps26119_amd_solve_lp.restype = ctypes.c_int
    lib.ps26119_amd_solve_batch.argtypes = model + [ctypes.c_int, ctypes.POINTER(Scenario), ctypes.POINTER(Options),
                                                    ctypes.POINTER(AmdOptions), ctypes.POINTER(CResult), D, D]
    lib.ps26119_amd_solve_batch.restype = ctypes.c_int
    lib.ps26119_amd_server_create.argtypes = [ctypes.POINTER(ServerOptions), ctypes.c_char_p, ctypes.c_int]
    lib.ps26119_amd_server_create.restype = ctypes.c_void_p
    lib.ps26119_amd_server_destroy.argtypes = [ctypes.c_void_p]
    lib.ps26119_amd_server_submit.argtypes = [ctypes.c_void_p] + model + [ctypes.POINTER(Options),
                                                                          ctypes.POINTER(AmdOptions)]
    lib.ps26119_amd_server_submit.restype = ctypes.c_longlong
    lib.ps26119_amd_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_double,
                                            ctypes.POINTER(CResult), D, D, D, D]
    lib.ps26119_amd_server_wait.restype = ctypes.c_int
    lib.ps26119_amd_server_in_flight.argtypes = [ctypes.c_void_p]
    lib.ps26119_amd_server_in_flight.restype = ctypes.c_int
    lib.ps26119_amd_server_kernel_time.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong),
                                                   D, D]
    lib.ps26119_amd_server_kernel_time.restype = ctypes.c_int
    lib.ps26119_default_options.argtypes = [ctypes.POINTER(Options)]
    lib.ps26119_amd_default_options.argtypes = [ctypes.POINTER(AmdOptions)]
    lib.ps26119_solve_mps.argtypes = [ctypes.c_char_p, ctypes.POINTER(Options), ctypes.POINTER(CResult),
                                      ctypes.c_char_p]
    lib.ps26119_solve_mps.restype = ctypes.c_int
    lib.ps26119_version.restype = ctypes.c_char_p
    lib.ps26119_amd_device_count.restype = ctypes.c_int
    nomodel = [ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D, D, D, D, I, I, D]
    lib.ps26119_amd_scale.argtypes = nomodel + [D, D, D, D, D]
    lib.ps26119_amd_scale.restype = ctypes.c_int
    lib.ps26119_amd_certified_bound.argtypes = model + [D, ctypes.c_int]
    lib.ps26119_amd_certified_bound.restype = ctypes.c_double
    lib.ps26119_amd_check_solution.argtypes = model + [D, D, D]
    lib.ps26119_amd_check_solution.restype = ctypes.c_int
    lib.ps26119_amd_presolve_info.argtypes = model + [D]
    lib.ps26119_amd_presolve_info.restype = ctypes.c_int
    IP = ctypes.POINTER(ctypes.c_int)
    lib.ps26119_amd_presolve_model.argtypes = model + [D, IP, IP, D, D, D, D, IP, IP, D]
    lib.ps26119_amd_presolve_model.restype = ctypes.c_int
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def version() -> str:
    return lib().ps26119_version().decode()


def device_count() -> int:
    return lib().ps26119_amd_device_count()


@dataclass
class Result:
    status: int
    status_name: str
    objective: float
    dual_objective: float
    primal_residual: float
    dual_residual: float
    gap: float
    certified_bound: float
    iterations: int
    seconds: float
    engine: str
    message: str
    check: str = ""
    x: np.ndarray = field(repr=False, default=None)
    y: np.ndarray = field(repr=False, default=None)
    z: np.ndarray = field(repr=False, default=None)
    dual_ray: np.ndarray = field(repr=False, default=None)
    primal_ray: np.ndarray = field(repr=False, default=None)
    stats: dict = field(repr=False, default=None)


def _lookup(table, key, what):
    if key not in table:
        raise ValueError(f"unknown {what} {key!r}; expected one of {sorted(table)}")
    return table[key]


def csc(A, m, n):
    """(col_start, row_index, value) from a CSC triple, a scipy.sparse matrix or a dense array."""
    if isinstance(A, tuple):
        cs, ri, v = A
        return (np.ascontiguousarray(cs, dtype=np.int32), np.ascontiguousarray(ri, dtype=np.int32),
                np.ascontiguousarray(v, dtype=np.float64))
    if hasattr(A, "tocsc"):
        A = A.tocsc()
        A.sort_indices()
        return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)
    A = np.asarray(A, dtype=np.float64)
    if A.shape != (m, n):
        raise ValueError(f"A has shape {A.shape}, expected {(m, n)}")
    rows, cols = np.nonzero(A.T)
    cs = np.zeros(n + 1, dtype=np.int32)
    np.cumsum(np.bincount(rows, minlength=n), out=cs[1:])
    return cs, cols.astype(np.int32), A.T[rows, cols].astype(np.float64)


def _ptr(a, ctype=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ctype)) if a is not None else None


def solve_lp(c, A, row_lower, row_upper, col_lower=None, col_upper=None, sense=1, offset=0.0, algorithm="auto",
             precision="fp64", gpu=True, tolerance=1e-8, time_limit=3600.0, iteration_limit=0, warm_x=None,
             warm_y=None, verbosity=0, threads=1, presolve=True, check_every=0, params=None, stats=False) -> Result:
    """min (sense=+1) or max (sense=-1) cᵀx + offset  s.t. row_lower <= A x <= row_upper, bounds."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    n = c.size
    rl = np.ascontiguousarray(row_lower, dtype=np.float64)
    ru = np.ascontiguousarray(row_upper, dtype=np.float64)
    m = rl.size
    cl = np.zeros(n) if col_lower is None else np.ascontiguousarray(col_lower, dtype=np.float64)
    cu = np.full(n, np.inf) if col_upper is None else np.ascontiguousarray(col_upper, dtype=np.float64)
    cs, ri, v = csc(A, m, n)
    L = lib()
    opt = Options()
    L.ps26119_default_options(ctypes.byref(opt))
    opt.algorithm = _lookup(_ALG, algorithm, "algorithm")
    opt.precision = _lookup(_PREC, precision, "precision")
    opt.use_gpu = int(bool(gpu))
    opt.tolerance = tolerance
    opt.time_limit = time_limit
    opt.iteration_limit = iteration_limit
    opt.verbosity = verbosity
    opt.threads = threads
    wx = None if warm_x is None else np.ascontiguousarray(warm_x, dtype=np.float64)
    wy = None if warm_y is None else np.ascontiguousarray(warm_y, dtype=np.float64)
    opt.warm_x = _ptr(wx)
    opt.warm_y = _ptr(wy)
    ext = AmdOptions()
    L.ps26119_amd_default_options(ctypes.byref(ext))
    ext.presolve = int(bool(presolve))
    ext.check_every = int(check_every)
    names = [k.encode() for k in (params or {})]
    vals = np.array([float(x) for x in (params or {}).values()], dtype=np.float64)
    name_arr = (ctypes.c_char_p * max(1, len(names)))(*names)
    ext.n_params = len(names)
    ext.param_names = name_arr
    ext.param_values = _ptr(vals) if len(names) else None
    st_out = AmdStats()
    res = CResult()
    x, y, z = np.full(n, np.nan), np.empty(m), np.empty(n)
    dr, pr = np.full(m, np.nan), np.full(n, np.nan)
    st = L.ps26119_amd_solve_lp(m, n, int(sense), float(offset), _ptr(c), _ptr(cl), _ptr(cu), _ptr(rl), _ptr(ru),
                                _ptr(cs, ctypes.c_int), _ptr(ri, ctypes.c_int), _ptr(v), ctypes.byref(opt),
                                ctypes.byref(ext), ctypes.byref(res), ctypes.byref(st_out), _ptr(x), _ptr(y), _ptr(z),
                                _ptr(dr), _ptr(pr))
    have = st in (OPTIMAL, ITERATION_LIMIT, TIME_LIMIT)
    ray_d = dr if st == INFEASIBLE and m > 0 and not np.isnan(dr).any() else None
    ray_p = pr if st == UNBOUNDED and n > 0 and not np.isnan(pr).any() else None
    return Result(st, _STATUS[st] if 0 <= st < len(_STATUS) else str(st), res.objective, res.dual_objective,
                  res.primal_residual, res.dual_residual, res.gap, res.certified_bound, res.iterations, res.seconds,
                  res.engine.decode(), res.message.decode(), _CHECK.get(res.check, ""),
                  x if have or ray_p is not None else None, y if have else None, z if have else None, ray_d, ray_p,
                  st_out.as_dict() if stats else None)


def _options(L, algorithm, precision, gpu, tolerance, time_limit, iteration_limit, verbosity, threads, wx, wy):
    opt = Options()
    L.ps26119_default_options(ctypes.byref(opt))
    opt.algorithm = _lookup(_ALG, algorithm, "algorithm")
    opt.precision = _lookup(_PREC, precision, "precision")
    opt.use_gpu = int(bool(gpu))
    opt.tolerance = tolerance
    opt.time_limit = time_limit
    opt.iteration_limit = iteration_limit
    opt.verbosity = verbosity
    opt.threads = threads
    opt.warm_x = _ptr(wx)
    opt.warm_y = _ptr(wy)
    return opt


def _ext(L, presolve, check_every, params, keep):
    ext = AmdOptions()
    L.ps26119_amd_default_options(ctypes.byref(ext))
    ext.presolve = int(bool(presolve))
    ext.check_every = int(check_every)
    names = [k.encode() for k in (params or {})]
    vals = np.array([float(x) for x in (params or {}).values()], dtype=np.float64)
    name_arr = (ctypes.c_char_p * max(1, len(names)))(*names)
    keep += [names, vals, name_arr]
    ext.n_params = len(names)
    ext.param_names = name_arr
    ext.param_values = _ptr(vals) if len(names) else None
    return ext


def solve_batch(c, A, row_lower, row_upper, col_lower=None, col_upper=None, scenarios=(), sense=1, offset=0.0,
                gpu=True, tolerance=1e-8, time_limit=3600.0, iteration_limit=0, warm_x=None, warm_y=None, verbosity=0,
                threads=1, check_every=0, params=None, precision="fp64") -> list:
    """Scenarios of one LP sharing A (the reference's ps26119::solve_batch, include/ps26119/batch.h).

    ``scenarios`` is a list of dicts with any of ``obj``, ``col_lower``, ``col_upper``,
    ``row_lower``, ``row_upper`` (missing = the base's). Returns one Result per scenario. A vector
    of the wrong length makes only that scenario NotSolved, as in the reference. ``precision`` =
    "mixed" runs every scenario in increment form (fp32 increments on fp64 anchors), an addition:
    the reference's batch is fp64 only."""
    c = np.ascontiguousarray(c, dtype=np.float64)
    n = c.size
    rl = np.ascontiguousarray(row_lower, dtype=np.float64)
    ru = np.ascontiguousarray(row_upper, dtype=np.float64)
    m = rl.size
    cl = np.zeros(n) if col_lower is None else np.ascontiguousarray(col_lower, dtype=np.float64)
    cu = np.full(n, np.inf) if col_upper is None else np.ascontiguousarray(col_upper, dtype=np.float64)
    cs, ri, v = csc(A, m, n)
    L = lib()
    keep = []
    wx = None if warm_x is None else np.ascontiguousarray(warm_x, dtype=np.float64)
    wy = None if warm_y is None else np.ascontiguousarray(warm_y, dtype=np.float64)
    opt = _options(L, "r2hpdhg", precision, gpu, tolerance, time_limit, iteration_limit, verbosity, threads, wx, wy)
    ext = _ext(L, False, check_every, params, keep)
    K = len(scenarios)
    arr = (Scenario * max(1, K))()
    for k, sc in enumerate(scenarios):
        ptrs = []
        for key, length in (("obj", n), ("col_lower", n), ("col_upper", n), ("row_lower", m), ("row_upper", m)):
            a = sc.get(key)
            if a is not None:
                a = np.ascontiguousarray(a, dtype=np.float64).ravel()
                if a.size != length:  # NaNs: the library refuses a scenario with a non-finite cost or NaN bound
                    a = np.full(length, np.nan)
                keep.append(a)
            ptrs.append(_ptr(a))
        arr[k] = Scenario(*ptrs)
    res = (CResult * max(1, K))()
    xs, ys = np.full(K * n, np.nan), np.full(K * m, np.nan)
    L.ps26119_amd_solve_batch(m, n, int(sense), float(offset), _ptr(c), _ptr(cl), _ptr(cu), _ptr(rl), _ptr(ru),
                              _ptr(cs, ctypes.c_int), _ptr(ri, ctypes.c_int), _ptr(v), K, arr, ctypes.byref(opt),
                              ctypes.byref(ext), res, _ptr(xs), _ptr(ys))
    out = []
    for k in range(K):
        r = res[k]
        st = r.status
        have = st in (OPTIMAL, ITERATION_LIMIT, TIME_LIMIT)
        out.append(Result(st, _STATUS[st] if 0 <= st < len(_STATUS) else str(st), r.objective, r.dual_objective,
                          r.primal_residual, r.dual_residual, r.gap, r.certified_bound, r.iterations, r.seconds,
                          r.engine.decode(), r.message.decode(), _CHECK.get(r.check, ""),
                          xs[k * n:(k + 1) * n].copy() if have else None,
                          ys[k * m:(k + 1) * m].copy() if have else None))
    return out


def solve_model_batch(lp, scenarios, **kw) -> list:
    """solve_batch for an object with the model fields (see solve_model)."""
    return solve_batch(lp.c, (lp.col_start, lp.row_index, lp.value), lp.row_lower, lp.row_upper, lp.col_lower,
                       lp.col_upper, scenarios=scenarios, sense=lp.sense, offset=lp.offset, **kw)


def solve_model(lp, **kw) -> Result:
    """solve_lp for an object with the model fields (m, n, c, col_lower, ..., value, sense, offset)."""
    return solve_lp(lp.c, (lp.col_start, lp.row_index, lp.value), lp.row_lower, lp.row_upper, lp.col_lower,
                    lp.col_upper, sense=lp.sense, offset=lp.offset, **kw)


def check_solution(lp, x, y) -> dict:
    """In-process verifier-grade check of (x, y) on an lp-like object (ps26119_amd_check_solution)."""
    L = lib()
    c, cl, cu, rl, ru, v = (np.ascontiguousarray(a, dtype=np.float64) for a in
                            (lp.c, lp.col_lower, lp.col_upper, lp.row_lower, lp.row_upper, lp.value))
    cs, ri = (np.ascontiguousarray(a, dtype=np.int32) for a in (lp.col_start, lp.row_index))
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(6)
    rc = L.ps26119_amd_check_solution(lp.m, lp.n, int(lp.sense), float(lp.offset), _ptr(c), _ptr(cl), _ptr(cu),
                                      _ptr(rl), _ptr(ru), _ptr(cs, ctypes.c_int), _ptr(ri, ctypes.c_int), _ptr(v),
                                      _ptr(x), _ptr(y), _ptr(out))
    if rc != OPTIMAL:
        raise ValueError(f"check_solution: invalid model or vectors (status {rc})")
    keys = ["max_primal_violation", "max_dual_violation", "relative_gap", "primal_objective", "dual_objective"]
    d = dict(zip(keys, out[:5].tolist()))
    d["passed"] = bool(out[5])
    return d


class Server:
    """Continuous batching of LP solves on one GPU (ps26119_amd_server_*).

        with ps26119.Server(budget_bytes=8 << 30) as srv:
            ids = [srv.submit(lp, tolerance=1e-6) for lp in models]     # returns at once
            results = [srv.wait(i) for i in ids]                          # Result with .latency

    Requests are prepared on host threads, admitted while the HBM budget has room, advanced
    together as one packed block-diagonal system and finished on their own schedule."""

    def __init__(self, budget_bytes=8 << 30, max_instances=256, check_every=0, setup_threads=4, device=0,
                 eager=False, timing=False):
        """eager: launch every chunk's kernels directly instead of replaying captured graphs;
        timing: HIP events around every chunk (kernel_time())."""
        L = lib()
        o = ServerOptions(device, float(budget_bytes), max_instances, check_every, setup_threads, int(bool(eager)),
                          int(bool(timing)))
        err = ctypes.create_string_buffer(512)
        self._h = L.ps26119_amd_server_create(ctypes.byref(o), err, 512)
        if not self._h:
            raise RuntimeError(err.value.decode() or "server could not be created")
        self._dims = {}
        self._keep = {}

    def submit(self, lp, tolerance=1e-8, iteration_limit=0, time_limit=3600.0, verbosity=0, params=None,
               precision="fp64") -> int:
        """Queues one LP; ``precision`` = "mixed" serves it in increment form (fp32 increments on
        fp64 anchors, fp64 takeover on a stall below 1e-5)."""
        L = lib()
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (lp.c, lp.col_lower, lp.col_upper, lp.row_lower, lp.row_upper, lp.value)]
        ints = [np.ascontiguousarray(a, dtype=np.int32) for a in (lp.col_start, lp.row_index)]
        c, cl, cu, rl, ru, v = arrs
        cs, ri = ints
        keep = arrs + ints
        opt = _options(L, "r2hpdhg", precision, True, tolerance, time_limit, iteration_limit, verbosity, 1, None,
                       None)
        ext = _ext(L, False, 0, params, keep)
        rid = L.ps26119_amd_server_submit(self._h, lp.m, lp.n, int(lp.sense), float(lp.offset), _ptr(c), _ptr(cl),
                                          _ptr(cu), _ptr(rl), _ptr(ru), _ptr(cs, ctypes.c_int),
                                          _ptr(ri, ctypes.c_int), _ptr(v), ctypes.byref(opt), ctypes.byref(ext))
        if rid < 0:
            raise ValueError("server rejected the request arguments")
        self._dims[rid] = (lp.m, lp.n)
        return rid

    def wait(self, rid, timeout=-1.0):
        """The request's Result (with .stats = {submitted, admitted, done, latency}); None on timeout."""
        L = lib()
        m, n = self._dims[rid]
        res = CResult()
        x, y, z = np.full(n, np.nan), np.full(m, np.nan), np.full(n, np.nan)
        times = np.zeros(4)
        rc = L.ps26119_amd_server_wait(self._h, rid, float(timeout), ctypes.byref(res), _ptr(x), _ptr(y), _ptr(z),
                                       _ptr(times))
        if rc == 1:
            return None
        if rc != 0:
            raise KeyError(rid)
        del self._dims[rid]
        st = res.status
        have = st in (OPTIMAL, ITERATION_LIMIT, TIME_LIMIT)
        return Result(st, _STATUS[st] if 0 <= st < len(_STATUS) else str(st), res.objective, res.dual_objective,
                      res.primal_residual, res.dual_residual, res.gap, res.certified_bound, res.iterations,
                      res.seconds, res.engine.decode(), res.message.decode(), _CHECK.get(res.check, ""),
                      x if have else None, y if have else None, z if have else None, None, None,
                      {"submitted": times[0], "admitted": times[1], "done": times[2], "latency": times[3]})

    def in_flight(self) -> int:
        return lib().ps26119_amd_server_in_flight(self._h)

    def kernel_time(self):
        """With timing=True: {iterations, ms, bytes} of the served iterations so far (device time
        of both half steps over the live instances, and their algorithmic bytes)."""
        n = ctypes.c_longlong(0)
        ms, b = ctypes.c_double(0), ctypes.c_double(0)
        if lib().ps26119_amd_server_kernel_time(self._h, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)) != 0:
            return None
        return {"iterations": n.value, "ms": ms.value, "bytes": b.value}

    def close(self):
        if self._h:
            lib().ps26119_amd_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
