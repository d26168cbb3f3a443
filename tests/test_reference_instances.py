"""The reference's own benchmark instances and its published results on them.

The reference publishes time-to-1e-8 runs on instances of its own generators
(bench/results/scale-rtx4050-laptop-183c59c.csv there): rand-<m>-s1 (bench/generate_lp.py) and
refinery-T<T>-s1 (bench/generate_refinery_lp.py). `lpgen.reference_bench_lp` and
`lpgen.reference_refinery_lp` restate both generators draw for draw; the published shape and planted
optimum of every instance pin them (the optimum is a sum over the whole instance, so one wrong draw
or a wrong coefficient changes it). The published CPU rows then pin what a solve must give: the
reference build reproduces their iteration counts and objectives here on the CPU, and the MI355X
path reproduces them on the GPU (its default reductions follow the reference's decisions, so fp64
solves take the same iterations to the same bits; "mixed" with the reference's precision policy too).
"""
import math

import numpy as np
import pytest

# (name, rows, cols, nnz, planted optimum, fp64: iterations / objective, mixed: iterations / objective)
# from the published CPU rows (CSV lines 2-3, 8-9, 20-21, 26-27, 32-33; lines 14-19 for rand-1e6:
# CPU 12 threads and GPU fp64 both 17,600 iterations, objective -1138593.929543431 on the CPU)
PUBLISHED = {
    "rand-10000-s1": (10000, 15000, 60024, -13805.501052078082, (1536, -13805.501060830773), (1536, -13805.501054477845)),
    "rand-100000-s1": (100000, 150000, 600228, -109311.46887857676, (2496, -109311.46890526047),
                       (2944, -109311.46888417374)),
    "rand-1000000-s1": (1000000, 1500000, 6002453, -1138593.9295595926, (17600, -1138593.929543431),
                        (16576, -1138593.9295456805)),
    "refinery-T12-s1": (588, 708, 2065, 12118.440547364007, (1088, 12118.440546976286), (2240, 12118.440556847085)),
    "refinery-T365-s1": (17885, 21535, 63134, 843641.450693402, (1984, 843641.4507003281), (2432, 843641.4507331152)),
    "refinery-T8760-s1": (429240, 516840, 1515469, 21158525.625104126, (2880, 21158525.625100967),
                          (2752, 21158525.625120834)),
}


def instance(name):
    from ps26119_amd import lpgen
    kind, size, _ = name.split("-")
    return lpgen.reference_bench_lp(int(size), seed=1) if kind == "rand" else lpgen.reference_refinery_lp(int(size[1:]), 1)


@pytest.mark.parametrize("name", ["rand-10000-s1", "rand-100000-s1", "refinery-T12-s1", "refinery-T365-s1",
                                  "refinery-T8760-s1"])
def test_generators_give_the_published_instances(name):
    m, n, nnz, optimum = PUBLISHED[name][:4]
    lp = instance(name)
    assert (lp.m, lp.n, lp.nnz) == (m, n, nnz)
    assert lp.optimum == optimum  # bit for bit
    # the planted point is optimal: feasible, and its objective is the planted optimum
    act = lp.activity(lp.x_star)
    scale = 1e-9 * (1 + np.abs(act))
    assert np.all(act >= lp.row_lower - scale) and np.all(act <= lp.row_upper + scale)
    assert np.all(lp.x_star >= lp.col_lower) and np.all(lp.x_star <= lp.col_upper)
    assert math.isclose(float(lp.c @ lp.x_star), lp.optimum, rel_tol=1e-12)


@pytest.mark.parametrize("name", ["rand-10000-s1", "refinery-T12-s1", "refinery-T365-s1"])
def test_reference_build_reproduces_its_published_rows(ref, name):
    lp = instance(name)
    for prec, (iters, obj) in ((0, PUBLISHED[name][4]), (1, PUBLISHED[name][5])):
        r, _ = ref.solve(lp, precision=prec, presolve=True, threads=8)
        assert r["status"] == 0 and r["iterations"] == iters and r["objective"] == obj, (name, prec, r)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["rand-10000-s1", "rand-100000-s1", "refinery-T12-s1", "refinery-T365-s1",
                                  "refinery-T8760-s1", "rand-1000000-s1"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_mi355x_solves_follow_the_published_cpu_rows(gpu_available, name, prec):
    """Default options through the C API (presolve, tolerance 1e-8, the reference's restart
    schedule); "mixed" with the reference's precision policy (hip_refine = 0)."""
    import ps26119
    lp = instance(name)
    iters, obj = PUBLISHED[name][4 if prec == "fp64" else 5]
    params = {"hip_refine": 0} if prec == "mixed" else None
    r = ps26119.solve_model(lp, precision=prec, params=params)
    print(f"{name} {prec}: {r.status_name} it={r.iterations} (published {iters}) {r.seconds:.2f}s "
          f"obj={r.objective!r} (published {obj!r}) check={r.check}", flush=True)
    assert r.status == ps26119.OPTIMAL and r.check == "PASS", r.message
    assert r.iterations == iters
    assert r.objective == obj
    assert abs(r.objective - lp.optimum) <= 1e-8 * (1 + abs(lp.optimum))
