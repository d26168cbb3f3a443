"""The reference's C API on the GPU, through its own symbols (ps26119_solve_lp, ps26119_solve_lp_ex).

These mirror the reference's C-API tests (its tests/unit/test_c_api.cpp: SolvesWyndorWithEveryAlgorithm
at :33-52, NullOptionsAndOutputsAreAllowed, InvalidArgumentsNeverCrash, WarmStartFromOptimum,
SolveLpExReturnsCertificates at :125-156) and its time-limit test (tests/unit/test_certificates.cpp:
337-382), with use_gpu = 1: this library serves the first-order engines on the MI355X and answers a
request for a CPU engine (oracle, simplex, or use_gpu = 0) with NotSolved and the reason.
"""
import ctypes
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INF = math.inf
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)
ALG_AUTO, ALG_ORACLE, ALG_PDLP, ALG_R2HPDHG, ALG_SIMPLEX = range(5)

# Wyndor (max 3x + 5y; x <= 4, 2y <= 12, 3x + 2y <= 18) in CSC, as in the reference test
WYNDOR = dict(m=3, n=2, sense=-1, c=[3, 5], cl=[0, 0], cu=[INF, INF], rl=[-INF, -INF, -INF], ru=[4, 12, 18],
              cs=[0, 2, 4], ri=[0, 2, 1, 2], v=[1, 3, 2, 2])


def _d(a):
    return (ctypes.c_double * len(a))(*a)


def _i(a):
    return (ctypes.c_int * len(a))(*a)


def _opts(alg=ALG_R2HPDHG, **kw):
    import ps26119
    o = ps26119.Options()
    ps26119.lib().ps26119_default_options(ctypes.byref(o))
    o.algorithm = alg
    o.use_gpu = 1
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _solve(model, opt, offset=0.0, x=None, y=None, z=None, ray=None, pray=None, ex=False, result=True):
    import ps26119
    L = ps26119.lib()
    r = ps26119.CResult()
    args = [model["m"], model["n"], model["sense"], offset, _d(model["c"]), _d(model["cl"]), _d(model["cu"]),
            _d(model["rl"]), _d(model["ru"]), _i(model["cs"]), _i(model["ri"]), _d(model["v"]),
            ctypes.byref(opt) if opt is not None else None, ctypes.byref(r) if result else None,
            x, y, z]
    st = L.ps26119_solve_lp_ex(*args, ray, pray) if ex else L.ps26119_solve_lp(*args)
    return st, r


@pytest.mark.parametrize("alg", [ALG_AUTO, ALG_PDLP, ALG_R2HPDHG])
def test_solves_wyndor_through_the_c_api(gpu_available, alg):
    import ps26119
    x, y, z = (ctypes.c_double * 2)(), (ctypes.c_double * 3)(), (ctypes.c_double * 2)()
    st, r = _solve(WYNDOR, _opts(alg), x=x, y=y, z=z)
    print(f"alg {alg}: {r.engine.decode()} it={r.iterations} obj={r.objective!r} {r.message.decode()[:100]}", flush=True)
    assert st == ps26119.OPTIMAL, r.message
    assert abs(r.objective - 36) <= 1e-6
    assert abs(x[0] - 2) <= 1e-5 and abs(x[1] - 6) <= 1e-5
    assert abs(y[1] - 1.5) <= 1e-5 and abs(y[2] - 1.0) <= 1e-5
    assert 36 - 1e-9 <= r.certified_bound < 36 + 1e-4  # MAX model: certified upper bound on the optimum
    assert len(r.engine) > 0 and r.check == 1


@pytest.mark.parametrize("alg", [ALG_ORACLE, ALG_SIMPLEX])
def test_cpu_engines_are_refused_with_a_reason(gpu_available, alg):
    import ps26119
    st, r = _solve(WYNDOR, _opts(alg))
    assert st == ps26119.NOT_SOLVED and b"reference CPU library" in r.message
    st, r = _solve(WYNDOR, _opts(ALG_R2HPDHG, use_gpu=0))
    assert st == ps26119.NOT_SOLVED and b"use_gpu = 1" in r.message


def test_offset_and_null_outputs(gpu_available):
    import ps26119
    st, r = _solve(WYNDOR, _opts(), offset=1.0)
    assert st == ps26119.OPTIMAL and abs(r.objective - 37) <= 1e-6  # offset included
    assert _solve(WYNDOR, _opts(), result=False)[0] == ps26119.INVALID_ARGUMENT


def test_invalid_arguments_never_crash(gpu_available):
    import ps26119
    bad = dict(WYNDOR, ri=[0, 7, 1, 2])  # row index out of range
    st, r = _solve(bad, _opts())
    assert st == ps26119.INVALID_ARGUMENT and b"invalid model" in r.message
    st, _ = _solve(dict(WYNDOR, m=-1), _opts())
    assert st == ps26119.INVALID_ARGUMENT
    st, _ = _solve(WYNDOR, _opts(42))
    assert st == ps26119.INVALID_ARGUMENT


def test_warm_start_from_optimum(gpu_available):
    import ps26119
    x, y = (ctypes.c_double * 2)(), (ctypes.c_double * 3)()
    st, cold = _solve(WYNDOR, _opts(), x=x, y=y)
    assert st == ps26119.OPTIMAL
    st, warm = _solve(WYNDOR, _opts(warm_x=ctypes.cast(x, D), warm_y=ctypes.cast(y, D)))
    assert st == ps26119.OPTIMAL and warm.iterations <= cold.iterations and abs(warm.objective - 36) <= 1e-6


@pytest.mark.parametrize("alg", [ALG_R2HPDHG, ALG_PDLP])
def test_solve_lp_ex_returns_certificates(gpu_available, alg):
    """x + y <= 1, x + y >= 3, x, y >= 0 is infeasible, with the Farkas vector (-1, 1) up to scale;
    min -x, x - y <= 0 is unbounded along a ray with d_x > 0, from a feasible point. The reference
    test asks this of its simplex, whose ray is exact (1e-12); a first-order engine's ray comes from
    the iterates' drift, and the reference's own r²HPDHG returns (-3947.11272, 3947.11124) here, so
    the direction is held to 1e-6."""
    import ps26119
    nan2 = lambda: (ctypes.c_double * 2)(math.nan, math.nan)  # noqa: E731
    infeasible = dict(m=2, n=2, sense=1, c=[1, 1], cl=[0, 0], cu=[INF, INF], rl=[-INF, 3], ru=[1, INF],
                      cs=[0, 2, 4], ri=[0, 1, 0, 1], v=[1, 1, 1, 1])
    ray, pray = nan2(), nan2()
    st, r = _solve(infeasible, _opts(alg), ray=ray, pray=pray, ex=True)
    print(f"infeasible: {r.message.decode()[:120]} ray=({ray[0]!r}, {ray[1]!r})", flush=True)
    assert st == ps26119.INFEASIBLE and r.check == 1, r.message
    assert ray[0] < 0 < ray[1]
    assert abs(ray[0] + ray[1]) <= 1e-6 * abs(ray[1])
    assert math.isnan(pray[0])  # no primal ray for an Infeasible verdict
    unbounded = dict(m=1, n=2, sense=1, c=[-1, 0], cl=[0, 0], cu=[INF, INF], rl=[-INF], ru=[0],
                     cs=[0, 1, 2], ri=[0, 0], v=[1, -1])
    x, pray = nan2(), nan2()
    st, r = _solve(unbounded, _opts(alg), x=x, pray=pray, ex=True)
    print(f"unbounded: {r.message.decode()[:120]}", flush=True)
    assert st == ps26119.UNBOUNDED and r.check == 1, r.message
    assert pray[0] > 0 and not math.isnan(x[0])


def covering_lp(n=60000, seed=3):
    """min cᵀx, A x >= 1, x >= 0 with 5 entries per column in [1, 3), after the reference's
    time-limit instance. There one of the 5 rows of column j is random too, which leaves about
    n·e⁻⁵ rows empty and the model infeasible (a CPU run reaches its time limit before it proves
    that; the MI355X proves it in 0.1 s): here column j always touches row j."""
    from ps26119_amd import lpgen
    rng = np.random.default_rng(seed)
    rows = np.sort(np.concatenate([np.arange(n)[:, None], rng.integers(0, n, (n, 4))], axis=1), axis=1)
    keep = np.ones_like(rows, dtype=bool)
    keep[:, 1:] = rows[:, 1:] != rows[:, :-1]  # no duplicate rows within a column
    cols = np.repeat(np.arange(n), 5).reshape(n, 5)
    vals = 1.0 + rng.integers(0, 100, (n, 5)) / 50.0
    cs, ri, v = lpgen._csc_from_triplets(n, n, rows[keep], cols[keep], vals[keep])
    c = 1.0 + rng.integers(0, 1000, n) / 100.0
    return lpgen.Lp(m=n, n=n, c=c, col_lower=np.zeros(n), col_upper=np.full(n, INF), row_lower=np.ones(n),
                    row_upper=np.full(n, INF), col_start=cs, row_index=ri, value=v, name="covering60k")


@pytest.mark.parametrize("case", ["covering60k-r2hpdhg", "covering60k-pdlp", "random1M-r2hpdhg"])
def test_first_order_engines_stop_near_the_time_limit(gpu_available, case):
    """A 0.3 s limit returns TimeLimit (or Optimal) within the limit + setup + 3 s, the reference's
    own bound: the loop reads the clock every 64 iterations (a few ms here) and the certified bound
    skips its implied-bound refinement once the call is past its time."""
    import ps26119
    from ps26119_amd import lpgen
    name, alg = case.split("-")
    lp = covering_lp() if name == "covering60k" else lpgen.random_lp(1_000_000, seed=1)
    t0 = time.time()
    r = ps26119.solve_model(lp, algorithm=alg, time_limit=0.3, stats=True)
    wall = time.time() - t0
    setup = r.stats["setup_seconds"]
    print(f"{case}: {r.status_name} it={r.iterations} wall={wall:.2f}s setup={setup:.2f}s", flush=True)
    assert r.status in (ps26119.TIME_LIMIT, ps26119.OPTIMAL)
    assert wall < 0.3 + setup + 3.0
