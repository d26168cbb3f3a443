"""Host side of the MI355X backend, on the CPU: the C-ABI libraries load and export every
declared entry point, the C API layout matches the reference header, argument validation and
error statuses follow the reference's C API, and the host stages of the path (scaling,
presolve, certified bound, in-process check) reproduce the oracle bit for bit."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import NETLIB, ROOT, hand_models, netlib

HEADERS = [os.path.join(ROOT, "include", "ps26119", "ps26119.h"), os.path.join(ROOT, "include", "ps26119", "ps26119_amd.h"),
           os.path.join(ROOT, "include", "ps26119_hip.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:ps26119|psh)_\w+)\s*\(", text)))


def test_libraries_export_every_declared_symbol():
    import ps26119
    from ps26119_amd import psh
    main = ctypes.CDLL(ps26119.library_path)
    hip = ctypes.CDLL(psh.HIP_LIB)
    for h in HEADERS:
        lib = hip if h.endswith("ps26119_hip.h") else main
        names = declared_functions(h)
        assert names, h
        missing = [n for n in names if not hasattr(lib, n)]
        assert not missing, (h, missing)


def test_c_header_compiles_as_c_and_matches_reference_layout(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "abi.c"
    src.write_text(r"""
#include <stddef.h>
#include <stdio.h>
#include "ps26119/ps26119.h"
#define F(T, f) printf(#T "." #f " %zu %zu\n", offsetof(T, f), sizeof(((T*)0)->f));
int main(void) {
  printf("options %zu result %zu\n", sizeof(ps26119_options), sizeof(ps26119_result));
  F(ps26119_options, algorithm) F(ps26119_options, precision) F(ps26119_options, use_gpu) F(ps26119_options, tolerance)
  F(ps26119_options, time_limit) F(ps26119_options, iteration_limit) F(ps26119_options, verbosity)
  F(ps26119_options, threads) F(ps26119_options, warm_x) F(ps26119_options, warm_y)
  F(ps26119_result, status) F(ps26119_result, objective) F(ps26119_result, dual_objective)
  F(ps26119_result, primal_residual) F(ps26119_result, dual_residual) F(ps26119_result, gap)
  F(ps26119_result, certified_bound) F(ps26119_result, iterations) F(ps26119_result, seconds)
  F(ps26119_result, engine) F(ps26119_result, message) F(ps26119_result, check)
  printf("enums %d %d %d %d %d %d %d %d %d %d\n", PS26119_OPTIMAL, PS26119_INFEASIBLE, PS26119_UNBOUNDED,
         PS26119_ITERATION_LIMIT, PS26119_TIME_LIMIT, PS26119_NUMERICAL_ERROR, PS26119_NOT_SOLVED,
         PS26119_INVALID_ARGUMENT, PS26119_ALG_R2HPDHG, PS26119_PREC_MIXED);
  return 0;
}
""")

    def layout(include):
        exe = tmp_path / ("abi_" + str(abs(hash(include))))
        subprocess.run([gcc, "-std=c99", "-Wall", "-Werror", f"-I{include}", str(src), "-o", str(exe)], check=True)
        return subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout

    ours = layout(os.path.join(ROOT, "include"))
    assert "options 64 result 256" in ours
    if os.path.isdir("/root/reference/include"):
        assert ours == layout("/root/reference/include")


def _arrays(lp):
    return dict(c=np.ascontiguousarray(lp.c, float), cl=np.ascontiguousarray(lp.col_lower, float),
                cu=np.ascontiguousarray(lp.col_upper, float), rl=np.ascontiguousarray(lp.row_lower, float),
                ru=np.ascontiguousarray(lp.row_upper, float), cs=np.ascontiguousarray(lp.col_start, np.int32),
                ri=np.ascontiguousarray(lp.row_index, np.int32), v=np.ascontiguousarray(lp.value, float))


def _p(a, t=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _model_args(lp, a, offset=True):
    head = [lp.m, lp.n, lp.sense] + ([lp.offset] if offset else [])
    return head + [_p(a["c"]), _p(a["cl"]), _p(a["cu"]), _p(a["rl"]), _p(a["ru"]), _p(a["cs"], ctypes.c_int),
                   _p(a["ri"], ctypes.c_int), _p(a["v"])]


def _host_scale(lp):
    import ps26119
    a = _arrays(lp)
    rs, cs, sc = np.zeros(lp.m), np.zeros(lp.n), np.zeros(4)
    av, atv = np.zeros(lp.nnz), np.zeros(lp.nnz)
    assert ps26119.lib().ps26119_amd_scale(*_model_args(lp, a, offset=False), _p(rs), _p(cs), _p(sc), _p(av), _p(atv)) == 0
    return rs, cs, sc, av, atv


@pytest.mark.parametrize("name", ["afiro", "adlittle", "random", "random200k"])
def test_host_scaling_is_bit_identical_to_oracle(oracle, name):
    """random200k runs every host sweep on several pieces of the worker pool."""
    from ps26119_amd import lpgen
    lp = {"random": lambda: lpgen.random_lp(3000, seed=8), "random200k": lambda: lpgen.random_lp(200_000, seed=8)}.get(
        name, lambda: netlib(name))()
    rs, cs, sc, av, atv = _host_scale(lp)
    e = oracle.scale(lp)
    assert np.array_equal(rs, e["row_scale"]) and np.array_equal(cs, e["col_scale"])
    assert np.array_equal(av, e["a_vals"]) and np.array_equal(atv, e["at_vals"])
    assert sc[0] == e["bound_scale"] and sc[1] == e["obj_scale"] and bool(sc[3]) == e["geometric_mean_applied"]


def test_host_loops_from_several_threads_give_the_same_bits():
    """Host loops called from several threads at once (the serving finish workers do this): one
    owns the worker pool, the others run on their own thread, and every result is the same."""
    from concurrent.futures import ThreadPoolExecutor
    from ps26119_amd import lpgen
    lp = lpgen.random_lp(150_000, seed=2)
    first = _host_scale(lp)
    with ThreadPoolExecutor(4) as ex:
        for got in ex.map(lambda _: _host_scale(lp), range(8)):
            assert all(np.array_equal(a, b) for a, b in zip(got, first))


@pytest.mark.parametrize("name", ["stocfor1", "wyndor", "random"])
def test_host_verification_is_bit_identical_to_oracle(oracle, name):
    import ps26119
    from ps26119_amd import lpgen
    lp = lpgen.random_lp(2000, seed=3) if name == "random" else (hand_models()[name] if name == "wyndor" else netlib(name))
    r, v = oracle.solve(lp, presolve=False)
    a = _arrays(lp)
    rng = np.random.default_rng(1)
    for y in (v["y"], v["y"] + 1e-3 * rng.standard_normal(lp.m)):
        y = np.ascontiguousarray(y)
        got = ps26119.lib().ps26119_amd_certified_bound(*_model_args(lp, a), _p(y), 1)
        assert got == oracle.certified_bound(lp, y)
    out = np.zeros(6)
    assert ps26119.lib().ps26119_amd_check_solution(*_model_args(lp, a), _p(v["x"]), _p(np.ascontiguousarray(v["y"])), _p(out)) == 0
    assert list(out) == list(oracle.check_solution(lp, v["x"], v["y"]).values())


def test_host_presolve_counts(oracle):
    import ps26119
    from ps26119_amd import lpgen
    lp = lpgen.random_lp(2000, seed=4)
    a = _arrays(lp)
    out = np.zeros(7)
    assert ps26119.lib().ps26119_amd_presolve_info(*_model_args(lp, a), _p(out)) == 0
    empty_cols = int(np.sum(np.diff(lp.col_start) == 0))
    assert out[0] == 0 and out[1] == 0 and out[2] == empty_cols  # random LP: presolve drops the empty columns
    assert out[4] == lp.n - empty_cols and out[5] == lp.nnz


def _presolve_cases():
    from ps26119_amd import lpgen
    inf = np.inf
    cases = {name: netlib(name) for name in NETLIB}
    cases.update(hand_models())
    cases["random300_s3"] = lpgen.random_lp(300, seed=3)
    cases["random2000_s4"] = lpgen.random_lp(2000, seed=4)
    cases["refinery_T6_s3"] = lpgen.refinery_lp(T=6, seed=3)
    # a singleton row that tightens both bounds, a fixed column whose substitution leaves a row
    # empty (dropped) and another a singleton (next round), an entry-free column at its best bound
    cases["rounds"] = lpgen.from_dense([1, -2, 3, 1, 0.5], [[2, 0, 0, 0, 0], [1, 1, 0, 0, 0], [0, 1, 1, 1, 0],
                                                            [0, 0, 0, 3, 0]],
                                       [1, -inf, 2, 6], [4, 5, 8, 6], [0, 0, 0, 2, -1], [10, 10, 10, 2, 4])
    cases["singleton_conflict"] = lpgen.from_dense([1, 1], [[1, 0], [1, 1]], [20, 0], [30, 5], [0, 0], [10, 10])
    cases["empty_row_excludes_0"] = lpgen.from_dense([1, 1], [[0, 0], [1, 1]], [1, 0], [2, 5], [0, 0], [10, 10])
    return cases


@pytest.mark.parametrize("name", sorted(_presolve_cases()))
def test_host_presolve_is_bit_identical_to_oracle(oracle, name):
    """The reduced model of the product's presolve equals the oracle's (whose presolve = 1 solves
    are pinned to the reference's fixtures) bit for bit: outcome, counts, offset, kept rows and
    columns, reduced bounds and the reduced matrix."""
    import ctypes
    import ps26119
    lp = _presolve_cases()[name]
    a = _arrays(lp)
    info = np.zeros(7)
    I = ctypes.c_int
    got = {"kept_rows": np.zeros(lp.m, np.int32), "kept_cols": np.zeros(lp.n, np.int32), "rl": np.zeros(lp.m),
           "ru": np.zeros(lp.m), "cl": np.zeros(lp.n), "cu": np.zeros(lp.n), "col_start": np.zeros(lp.n + 1, np.int32),
           "row_index": np.zeros(max(1, lp.nnz), np.int32), "value": np.zeros(max(1, lp.nnz))}
    rc = ps26119.lib().ps26119_amd_presolve_model(
        *_model_args(lp, a), _p(info), _p(got["kept_rows"], I), _p(got["kept_cols"], I), _p(got["rl"]), _p(got["ru"]),
        _p(got["cl"]), _p(got["cu"]), _p(got["col_start"], I), _p(got["row_index"], I), _p(got["value"]))
    assert rc == 0
    einfo, exp = oracle.presolve(lp)
    print(f"{name}: outcome {int(info[0])}, {int(info[1])} rows and {int(info[2])} columns removed", flush=True)
    assert [float(v).hex() for v in info] == [float(v).hex() for v in einfo]
    import orc
    got = orc._trim(info, got)
    assert got.keys() == exp.keys()
    for k in exp:
        assert np.array_equal(got[k], exp[k]) and got[k].tobytes() == exp[k].tobytes(), k
    if name == "rounds":
        assert info[0] == 0 and info[1] >= 2 and info[2] >= 2
    if name in ("singleton_conflict", "empty_row_excludes_0"):
        assert info[0] == 1


def _solve_raw(lp, opt=None, result=True, **over):
    import ps26119
    L = ps26119.lib()
    a = _arrays(lp)
    for k, v in over.items():
        a[k] = v
    res = ps26119.CResult() if result else None
    st = L.ps26119_solve_lp(*_model_args(lp, a), ctypes.byref(opt) if opt is not None else None,
                            ctypes.byref(res) if res is not None else None, None, None, None)
    return st, res


def test_invalid_arguments_never_crash():
    import ps26119
    lp = hand_models()["wyndor"]
    st, _ = _solve_raw(lp, result=False)
    assert st == ps26119.INVALID_ARGUMENT
    st, res = _solve_raw(lp, ri=np.array([0, 7, 1, 2], np.int32))
    assert st == ps26119.INVALID_ARGUMENT and b"invalid model" in res.message
    for field, value in [("time_limit", -1.0), ("time_limit", float("nan")), ("tolerance", -1e-8),
                         ("tolerance", float("nan")), ("iteration_limit", -5), ("threads", -3), ("precision", 7),
                         ("algorithm", 99)]:
        opt = ps26119.Options()
        ps26119.lib().ps26119_default_options(ctypes.byref(opt))
        setattr(opt, field, value)
        st, res = _solve_raw(lp, opt)
        assert st == ps26119.INVALID_ARGUMENT, field
        assert len(res.message) > 0


def test_warm_primal_weight_is_checked():
    """The named parameter "warm_primal_weight" sets the reference's Options::warm_primal_weight
    (the C struct has no field for it); a negative or non-finite value is an invalid argument,
    in the single solve and in the batch."""
    import ps26119
    lp = hand_models()["wyndor"]
    for bad in (-1.0, float("nan"), float("inf")):
        r = ps26119.solve_model(lp, params={"warm_primal_weight": bad})
        assert r.status == ps26119.INVALID_ARGUMENT and "warm_primal_weight" in r.message, (bad, r.message)
        res = ps26119.solve_model_batch(lp, [{}], params={"warm_primal_weight": bad})
        assert res[0].status == ps26119.INVALID_ARGUMENT and "warm_primal_weight" in res[0].message


def test_cpu_requests_and_other_engines_are_not_solved():
    import ps26119
    lp = hand_models()["wyndor"]
    r = ps26119.solve_model(lp, gpu=False)
    assert r.status == ps26119.NOT_SOLVED and "use_gpu" in r.message
    r = ps26119.solve_model(lp, algorithm="simplex")
    assert r.status == ps26119.NOT_SOLVED and "r2hpdhg" in r.message
    assert ps26119.lib().ps26119_solve_mps(b"x.mps", None, ctypes.byref(ps26119.CResult()), None) == ps26119.INVALID_ARGUMENT


def test_no_device_is_not_solved_with_reason():
    import ps26119
    if ps26119.device_count() > 0:
        pytest.skip("a GPU is visible")
    r = ps26119.solve_model(hand_models()["wyndor"])
    assert r.status == ps26119.NOT_SOLVED and "HIP backend unavailable" in r.message


def test_reference_gpu_hook_reaches_the_hip_backend():
    """The reference core built with its GPU hook answered by the pdhg::Backend adapter
    (oracle/_ref/libps26119_ref_hip.so) runs its CPU backend unchanged for use_gpu = 0 and reaches
    the adapter for use_gpu = 1, which without a device reports why (make_backend's NotSolved)."""
    import orc
    import ps26119
    if not orc.available("ref_hip"):
        pytest.skip("oracle/_ref/libps26119_ref_hip.so not built (needs the reference sources)")
    if ps26119.device_count() > 0:
        pytest.skip("a GPU is visible")
    L = orc.load("ref_hip")
    lp = hand_models()["wyndor"]
    cpu, v = L.solve(lp, use_gpu=0)
    assert cpu["status"] == 0 and abs(cpu["objective"] - 36) <= 1e-6
    gpu, _ = L.solve(lp, use_gpu=1, presolve=False)
    assert gpu["status"] == 6 and "no usable MI355X" in gpu["message"], gpu["message"]


def test_engine_parameters_are_checked_before_the_device():
    """Every engine knob the extension header documents is accepted by the parser, which runs
    before the device is opened (so this holds without a GPU); an unknown name is refused with
    that name in the message, as the reference's solve() does."""
    import ps26119
    text = open(os.path.join(ROOT, "include", "ps26119", "ps26119_amd.h")).read()
    block = text[text.index("Named engine knobs"):text.index("typedef struct")]
    names = sorted(set(re.findall(r'"([a-z_0-9]+)"', block)))
    assert {"hip_max_epoch", "hip_max_epoch_gap_only", "hip_refine_max_epoch", "hip_cpu_order"} <= set(names)
    lp = hand_models()["wyndor"]
    for name in names:
        value = 0.0 if name == "hip_device" else 1.0
        r = ps26119.solve_model(lp, params={name: value}, iteration_limit=64)
        assert "unknown engine parameter" not in r.message, (name, r.message)
    r = ps26119.solve_model(lp, params={"hip_max_epochs": 8192.0})
    assert r.status == ps26119.NOT_SOLVED and "hip_max_epochs" in r.message


def test_version_and_defaults():
    import ps26119
    assert ps26119.version()
    opt = ps26119.Options()
    ps26119.lib().ps26119_default_options(ctypes.byref(opt))
    assert (opt.algorithm, opt.precision, opt.use_gpu, opt.tolerance, opt.threads) == (0, 0, 0, 1e-8, 1)


def test_bench_refuses_more_than_one_gpu_and_lists_its_workloads():
    """bench.py measures one GPU: --gpus 2 exits with a message before touching a device; --help
    names the workloads and the lp1m knobs."""
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "one GPU" in (r.stdout + r.stderr)
    h = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120)
    assert h.returncode == 0
    for name in ("lp1m", "refinery", "scenarios", "serving", "copy", "stream", "--epoch-cap", "--compare-seconds"):
        assert name in h.stdout, name
