"""Shared fixtures. GPU tests are marked ``gpu``; everything else runs on a CPU-only machine.

Checkers (oracle/liboracle.so — our CPU restatement; oracle/_ref/libps26119_ref.so — the
reference sources built from source) are loaded only here, in tests, as the checker.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[:0] = [os.path.join(ROOT, "ps26119-solver_amd", "python"), os.path.join(ROOT, "oracle"), ROOT]

GOLDEN = os.path.join(ROOT, "tests", "golden")
NETLIB = ["afiro", "sc50a", "sc50b", "kb2", "adlittle", "blend", "share2b", "sc105", "stocfor1", "recipe"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    import orc
    if not orc.available("oracle"):
        pytest.skip("oracle/liboracle.so not built")
    return orc.load("oracle")


@pytest.fixture(scope="session")
def ref():
    import orc
    if not orc.available("ref"):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return orc.load("ref")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(GOLDEN, "reference_runs.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def gpu_available():
    import ps26119
    if ps26119.device_count() <= 0:
        pytest.fail("GPU test on a machine without a visible GPU: the HIP path must run here")
    return True


def netlib(name):
    from ps26119_amd import lpgen
    return lpgen.read_lpm(os.path.join(GOLDEN, "netlib_small", name + ".lpm"))


def hand_models():
    from ps26119_amd import lpgen
    inf = np.inf
    return {
        "wyndor": lpgen.from_dense([3, 5], [[1, 0], [0, 2], [3, 2]], [-inf] * 3, [4, 12, 18], [0, 0], [inf, inf], sense=-1),
        "mixed_rows": lpgen.from_dense([2, 3, -1, 0.5], [[1, 1, 0, 0], [0, 1, 1, 0], [1, 0, 0, 1], [1, -1, 1, 1]],
                                       [2, 1, 1, -3], [inf, 1, 4, 3], [0, 0, -1, -inf], [10, 10, 2, inf]),
        "infeasible": lpgen.from_dense([1, 1], [[1, 1], [1, 1]], [-inf, 3], [1, inf], [0, 0], [inf, inf]),
        "unbounded": lpgen.from_dense([-1, 0], [[1, -1]], [-inf], [1], [0, 0], [inf, inf]),
        "max_ranged": lpgen.read_lpm(os.path.join(GOLDEN, "max_ranged.lpm")),
    }


def fnv(a) -> str:
    h = 0xcbf29ce484222325
    for b in np.ascontiguousarray(a, dtype=np.float64).tobytes():
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return f"{h:016x}"
