"""Writes the golden fixtures of tests/golden/ by running the REFERENCE solver core, built from its
own sources by oracle/ref.mk into oracle/_ref/ (run in a container that has the reference):

    make -f oracle/ref.mk -j8 && python tests/golden/make_golden.py [--big]

Fixtures are data only: per case the status, iteration count, objective (exact, as a hex
float), certified bound, in-process check verdict, final primal weight, iterations to 1e-4
and FNV-1a hashes of the bits of x and y; per kernel-level case the hash of every output vector
of a fixed backend op sequence and the reductions. --big adds the 10,000-row random LP (about
five minutes of reference CPU time). --refinery only adds the refinery planning LPs (T = 6 and 24
periods, seed 3; fp64 and mixed, with and without presolve; a few minutes) to the existing file.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.normpath(os.path.join(HERE, "..", ".."))
sys.path[:0] = [os.path.join(ROOT, "ps26119-solver_amd", "python"), os.path.join(ROOT, "oracle")]

import orc  # noqa: E402
from ps26119_amd import lpgen  # noqa: E402

NETLIB = ["afiro", "sc50a", "sc50b", "kb2", "adlittle", "blend", "share2b", "sc105", "stocfor1", "recipe"]


def fnv(a: np.ndarray) -> str:
    h = 0xcbf29ce484222325
    for b in np.ascontiguousarray(a, dtype=np.float64).tobytes():
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"


def hexf(v: float) -> str:
    return float(v).hex()


def hand_models():
    inf = np.inf
    return {
        "wyndor": lpgen.from_dense([3, 5], [[1, 0], [0, 2], [3, 2]], [-inf] * 3, [4, 12, 18], [0, 0], [inf, inf], sense=-1),
        "mixed_rows": lpgen.from_dense([2, 3, -1, 0.5], [[1, 1, 0, 0], [0, 1, 1, 0], [1, 0, 0, 1], [1, -1, 1, 1]],
                                       [2, 1, 1, -3], [inf, 1, 4, 3], [0, 0, -1, -inf], [10, 10, 2, inf]),
        "infeasible": lpgen.from_dense([1, 1], [[1, 1], [1, 1]], [-inf, 3], [1, inf], [0, 0], [inf, inf]),
        "unbounded": lpgen.from_dense([-1, 0], [[1, -1]], [-inf], [1], [0, 0], [inf, inf]),
        "max_ranged": lpgen.read_lpm(os.path.join(HERE, "max_ranged.lpm")),
    }


def solve_case(ref, lp, **kw):
    r, v = ref.solve(lp, **kw)
    out = {k: r[k] for k in ("status", "iterations", "iterations_to_fast", "check")}
    for k in ("objective", "certified_bound", "primal_weight", "dual_objective"):
        out[k] = hexf(r[k])
    out["x_hash"] = fnv(v["x"]) if r["status"] in (0, 3, 4) else None
    out["y_hash"] = fnv(v["y"]) if r["status"] in (0, 3, 4) else None
    out["message"] = r["message"][:120]
    return out


def batch_case(ref, lp, scenarios):
    res, xs, ys = ref.solve_batch(lp, scenarios, iteration_limit=3_000_000)
    out = []
    for k, r in enumerate(res):
        d = {key: r[key] for key in ("status", "iterations", "iterations_to_fast")}
        for key in ("objective", "primal_weight", "dual_objective"):
            d[key] = hexf(r[key])
        have = r["status"] in (0, 3, 4)
        d["x_hash"] = fnv(xs[k]) if have else None
        d["y_hash"] = fnv(ys[k]) if have else None
        d["message"] = r["message"][:120]
        out.append(d)
    return out


REFINERY = {"refinery_T6_s3": (6, 3), "refinery_T24_s3": (24, 3)}


def refinery_only(ref, out_path):
    with open(out_path) as f:
        data = json.load(f)
    for name, (T, seed) in REFINERY.items():
        lp = lpgen.refinery_lp(T=T, seed=seed)
        for pre in (True, False):
            for prec in (0, 1):
                key = f"{name}|presolve={int(pre)}|{'mixed' if prec else 'fp64'}"
                data["solves"][key] = solve_case(ref, lp, presolve=pre, precision=prec, iteration_limit=3_000_000)
                print(key, data["solves"][key]["status"], data["solves"][key]["iterations"], flush=True)
        sc = ref.scale(lp)
        data["kernels"][name] = {"norm2": hexf(ref.norm2(lp)), "row_scale": fnv(sc["row_scale"]),
                                 "col_scale": fnv(sc["col_scale"]), "a_vals": fnv(sc["a_vals"]),
                                 "bound_scale": hexf(sc["bound_scale"]), "obj_scale": hexf(sc["obj_scale"]),
                                 "geometric_mean_applied": int(sc["geometric_mean_applied"])}
    with open(out_path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    print("wrote", out_path)


def main():
    big = "--big" in sys.argv
    ref = orc.load("ref")
    if "--refinery" in sys.argv:
        refinery_only(ref, os.path.join(HERE, "reference_runs.json"))
        return
    models = {name: lpgen.read_lpm(os.path.join(HERE, "netlib_small", name + ".lpm")) for name in NETLIB}
    models.update(hand_models())
    models["random300_s3"] = lpgen.random_lp(300, seed=3)
    models["random2000_s4"] = lpgen.random_lp(2000, seed=4)
    solves = {}
    for name, lp in models.items():
        for pre in (True, False):
            for prec in (0, 1):
                key = f"{name}|presolve={int(pre)}|{'mixed' if prec else 'fp64'}"
                solves[key] = solve_case(ref, lp, presolve=pre, precision=prec, iteration_limit=3_000_000)
                print(key, solves[key]["status"], solves[key]["iterations"], flush=True)
    if big:
        lp = lpgen.random_lp(10000, seed=1)
        for pre in (True, False):
            key = f"random10000_s1|presolve={int(pre)}|fp64"
            solves[key] = solve_case(ref, lp, presolve=pre, precision=0)
            print(key, solves[key]["status"], solves[key]["iterations"], flush=True)
    pdlp = {}
    for name in ("afiro", "sc50a", "sc105", "stocfor1", "wyndor", "random300_s3"):
        for pre in (True, False):
            for prec in (0, 1):
                key = f"{name}|presolve={int(pre)}|{'mixed' if prec else 'fp64'}"
                pdlp[key] = solve_case(ref, models[name], presolve=pre, precision=prec, iteration_limit=3_000_000,
                                       algorithm=2)
                print("pdlp", key, pdlp[key]["status"], pdlp[key]["iterations"], flush=True)
    batches = {}
    for name in ("afiro", "sc50a", "stocfor1", "share2b", "random300_s3"):
        lp = models[name]
        batches[name] = batch_case(ref, lp, lpgen.perturbation_scenarios(lp))
        print("batch", name, [b["iterations"] for b in batches[name]], flush=True)
    mr = models["max_ranged"]
    cu = mr.col_upper.copy()
    cu[0] = 2.0
    batches["max_ranged"] = batch_case(ref, mr, [{}, {"obj": [1.0]}, {"col_upper": cu}])
    batches["random300_s3|scenario_set"] = batch_case(ref, models["random300_s3"],
                                                     lpgen.scenario_set(models["random300_s3"], 6))
    kernels = {}
    for name in ("afiro", "stocfor1", "random300_s3"):
        lp = models[name]
        sc = ref.scale(lp)
        kernels[name] = {"norm2": hexf(ref.norm2(lp)), "row_scale": fnv(sc["row_scale"]), "col_scale": fnv(sc["col_scale"]),
                         "a_vals": fnv(sc["a_vals"]), "bound_scale": hexf(sc["bound_scale"]),
                         "obj_scale": hexf(sc["obj_scale"])}
        for prec in (0, 1):
            vecs, red = ref.backend_ops(lp, precision=prec, iters=20, tau=0.3, sigma=0.4, rho=1.0)
            kernels[name][f"ops_{'mixed' if prec else 'fp64'}"] = {"vectors": {k: fnv(v) for k, v in vecs.items()},
                                                                  "reductions": [hexf(x) for x in red]}
    out_path = os.path.join(HERE, "reference_runs.json")
    old = {}
    if os.path.exists(out_path):
        with open(out_path) as f:
            old = json.load(f)
    if not big:  # keep previously computed big cases
        for k, v in old.get("solves", {}).items():
            if k.startswith("random10000"):
                solves[k] = v
    for k, v in old.get("solves", {}).items():  # and the refinery cases (--refinery)
        if k.startswith("refinery"):
            solves[k] = v
    for k, v in old.get("kernels", {}).items():
        if k.startswith("refinery"):
            kernels[k] = v
    with open(out_path, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "solves": solves, "kernels": kernels,
                   "batches": batches, "pdlp": pdlp}, f, indent=1,
                  sort_keys=True)
    print("wrote", out_path)


if __name__ == "__main__":
    main()
