"""The reference's own drivers on the MI355X backend, through its pdhg::Backend interface.

oracle/_ref/libps26119_ref_hip.so is the reference core built from its sources with its GPU hook
switched on (make_backend(true, …) → make_cuda_backend(), reference src/pdhg/cpu_backend.cpp:
293-310) and that hook answered by ps26119-solver_amd/csrc/adapter/hip_backend.cpp, a
pdhg::Backend over the psh_* C-ABI. The reference's r²HPDHG and PDLP drivers (src/pdhg/r2hpdhg.cpp,
src/pdhg/pdlp.cpp) then run unchanged on either backend. With every reduction in CPU order the two
runs must agree bit for bit: status, iteration count, objective, x, y, z and the certificates.
Without a GPU the hooked build answers use_gpu = 1 with NotSolved, so an Optimal here is the
MI355X backend's.
"""
import numpy as np
import pytest

from conftest import hand_models, netlib

pytestmark = pytest.mark.gpu

R2HPDHG, PDLP = 3, 2


@pytest.fixture(scope="module")
def ref_hip():
    import orc
    if not orc.available("ref_hip"):
        pytest.skip("oracle/_ref/libps26119_ref_hip.so not built (needs the reference sources)")
    return orc.load("ref_hip")


def model(name):
    from ps26119_amd import lpgen
    if name in ("wyndor", "mixed_rows", "infeasible", "unbounded", "max_ranged"):
        return hand_models()[name]
    if name.startswith("rand"):
        return lpgen.random_lp(int(name[4:]), seed=3)
    return netlib(name)


CASES = [
    ("afiro", R2HPDHG, 0), ("stocfor1", R2HPDHG, 0), ("sc105", R2HPDHG, 0), ("rand300", R2HPDHG, 0),
    ("rand300", R2HPDHG, 1), ("stocfor1", R2HPDHG, 1), ("rand2000", R2HPDHG, 0),
    ("afiro", PDLP, 0), ("rand300", PDLP, 0), ("kb2", PDLP, 1),
    ("wyndor", R2HPDHG, 0), ("max_ranged", R2HPDHG, 0), ("infeasible", R2HPDHG, 0), ("unbounded", R2HPDHG, 0),
]


@pytest.mark.parametrize("name,alg,prec", CASES, ids=[f"{n}-{'r2h' if a == R2HPDHG else 'pdlp'}-{'mixed' if p else 'fp64'}"
                                                       for n, a, p in CASES])
def test_reference_drivers_on_the_hip_backend_are_bit_identical(gpu_available, ref_hip, name, alg, prec):
    lp = model(name)
    kw = dict(algorithm=alg, precision=prec, presolve=True, iteration_limit=200000)
    cpu, vc = ref_hip.solve(lp, use_gpu=0, **kw)
    gpu, vg = ref_hip.solve(lp, use_gpu=1, **kw)
    print(f"{name}: cpu status={cpu['status']} it={cpu['iterations']} | hip status={gpu['status']} "
          f"it={gpu['iterations']} {gpu['message'][:100]}", flush=True)
    assert gpu["status"] == cpu["status"], gpu["message"]
    assert gpu["iterations"] == cpu["iterations"]
    assert float(gpu["objective"]).hex() == float(cpu["objective"]).hex()
    for key in ("x", "y", "z", "dual_ray", "primal_ray"):
        assert np.array_equal(vg[key], vc[key], equal_nan=True), key
    assert gpu["check"] == cpu["check"]
