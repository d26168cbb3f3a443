"""The command-line front end `ps26119-hip` (the reference CLI's `solve --gpu` and `batch`
commands, reference apps/cli/main.cpp): argument handling, exit codes, the .lpm reader and the
solution file, on the CPU; solves on the GPU (marked gpu)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, netlib

CLI = os.path.join(ROOT, "ps26119-solver_amd", "csrc", "ps26119-hip")
AFIRO = os.path.join(GOLDEN, "netlib_small", "afiro.lpm")


def run(*args, timeout=120):
    if not os.path.exists(CLI):
        pytest.skip("ps26119-hip not built")
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)


def read_solution(path):
    out = {"x": [], "y": []}
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    while k < len(lines):
        parts = lines[k].split()
        if parts and parts[0] in ("COLUMNS", "ROWS"):
            cnt = int(parts[1])
            vals = [lines[k + 1 + i].split() for i in range(cnt)]
            out["x" if parts[0] == "COLUMNS" else "y"] = [float(v[1 if parts[0] == "COLUMNS" else 2]) for v in vals]
            k += cnt + 1
            continue
        if len(parts) >= 2 and parts[0] in ("status", "objective", "model", "engine", "check"):
            out[parts[0]] = " ".join(parts[1:])
        k += 1
    return out


def solution_file_fields(path):
    """Every field of a solution file, keyed as the reference writer's inputs (the "model" line
    is left out: the reference writer computes the fingerprint from the model itself)."""
    nan = float("nan")
    f = {"check": "", "message": "", "certified_bound": nan, "primal_weight": nan, "iterations_to_fast": -1,
         "seconds_to_fast": -1.0, "check_primal": nan, "check_dual": nan, "check_gap": nan}
    lines = open(path).read().split("\n")
    assert lines[0] == "PS26119-SOLUTION 1"
    k = 1
    while lines[k] != "END":
        key, _, rest = lines[k].partition(" ")
        if key in ("COLUMNS", "ROWS", "DUAL_RAY", "PRIMAL_RAY"):
            block = [lines[k + 1 + i].split(" ") for i in range(int(rest))]
            if key == "COLUMNS":
                f["x"], f["z"] = [float(b[1]) for b in block], [float(b[2]) for b in block]
            elif key == "ROWS":
                f["activity"], f["y"] = [float(b[1]) for b in block], [float(b[2]) for b in block]
            else:
                f[key.lower()] = [float(b[1]) for b in block]
            k += len(block) + 1
            continue
        if key in ("status", "engine", "precision", "message"):
            f[key] = rest
        elif key == "check":
            parts = rest.split(" ")
            f["check"], f["check_primal"], f["check_dual"], f["check_gap"] = parts[0], *map(float, parts[2::2])
        elif key in ("iterations", "iterations_to_fast"):
            f[key] = int(rest)
        elif key not in ("model", "name"):
            f[key] = float(rest)
        k += 1
    return f


def assert_reference_writer_agrees(lpm, ours, tmp_path):
    """The reference's own writer (src/io/solution_writer.cpp:8-67 of the reference, built from
    source in oracle/_ref), given this file's contents and the model its own .lpm reader reads,
    writes the same bytes: header keys and their order, %.17g / %.6g / %.3g numbers, the model
    fingerprint, names, the COLUMNS / ROWS blocks and the certificates."""
    import orc
    if not orc.available("ref"):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    theirs = tmp_path / "reference.sol"
    assert orc.load("ref").write_solution(str(lpm), solution_file_fields(ours), str(theirs)) == 0
    assert open(ours, "rb").read() == open(theirs, "rb").read()


def test_lpm_files_written_here_read_in_the_reference_reader(tmp_path):
    """lpgen.write_lpm lays files out as the reference's tools/lpm.py does: the reference's own
    .lpm reader (src/io/lpm_reader.cpp of the reference, built in oracle/_ref) accepts them, and
    reading them back here gives the same arrays."""
    import orc
    from conftest import hand_models
    from ps26119_amd import lpgen
    if not orc.available("ref"):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    ref = orc.load("ref")
    for name, lp in list(hand_models().items()) + [("random300", lpgen.random_lp(300, seed=3))]:
        path = str(tmp_path / f"{name}.lpm")
        lpgen.write_lpm(lp, path)
        assert ref.write_solution(path, {"status": "NotSolved"}, str(tmp_path / "x.sol")) == 0, name
        back = lpgen.read_lpm(path)
        for key in ("c", "col_lower", "col_upper", "row_lower", "row_upper", "col_start", "row_index", "value"):
            assert np.array_equal(getattr(lp, key), getattr(back, key)), (name, key)


def test_version_and_usage():
    assert run("--version").returncode == 0
    assert run().returncode == 2
    assert run("solve").returncode == 2
    assert run("solve", AFIRO, "--tol", "abc").returncode == 2
    assert run("solve", AFIRO, "--bogus").returncode == 2
    assert run("frobnicate").returncode == 2


def test_read_errors(tmp_path):
    assert run("solve", str(tmp_path / "missing.lpm")).returncode == 3
    bad = tmp_path / "bad.lpm"
    bad.write_text("LPM 1\nROWS 1\nCOLS 1\nNNZ 1\nOBJ x\nEND\n")
    r = run("solve", str(bad))
    assert r.returncode == 3 and "bad or missing number" in r.stderr
    assert run("solve", str(tmp_path / "model.mps")).returncode == 3


def test_unwritable_output_fails_before_solving(tmp_path):
    r = run("solve", AFIRO, "--out", str(tmp_path / "no" / "such" / "dir" / "a.sol"))
    assert r.returncode == 4


def test_reader_matches_the_python_reader_and_writer_round_trip(tmp_path):
    from ps26119_amd import lpgen
    lp = lpgen.random_lp(50, seed=4)
    p = tmp_path / "r.lpm"
    lpgen.write_lpm(lp, str(p))
    back = lpgen.read_lpm(str(p))
    for a, b in ((lp.c, back.c), (lp.value, back.value), (lp.row_upper, back.row_upper), (lp.col_start, back.col_start)):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    r = run("solve", str(p), "--iteration-limit", "64")
    assert f"rows {lp.m}  cols {lp.n}  nnz {lp.nnz}" in r.stdout


def test_batch_rejects_a_different_matrix(tmp_path):
    from ps26119_amd import lpgen
    a, b = lpgen.random_lp(30, seed=1), lpgen.random_lp(30, seed=2)
    pa, pb = tmp_path / "a.lpm", tmp_path / "b.lpm"
    lpgen.write_lpm(a, str(pa))
    lpgen.write_lpm(b, str(pb))
    r = run("batch", str(pa), str(pb))
    assert r.returncode == 2 and "differs from the base model" in r.stderr


def test_without_a_gpu_the_solve_is_not_solved(tmp_path):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    out = tmp_path / "a.sol"
    r = run("solve", AFIRO, "--out", str(out))
    assert r.returncode == 5 and "HIP backend unavailable" in r.stdout
    sol = read_solution(out)
    assert sol["status"] == "NotSolved"
    assert_reference_writer_agrees(AFIRO, out, tmp_path)


@pytest.mark.gpu
def test_gpu_solve_writes_a_verified_solution_and_warm_starts(gpu_available, golden, tmp_path):
    out = tmp_path / "afiro.sol"
    r = run("solve", AFIRO, "--gpu", "--out", str(out))
    assert r.returncode == 0, r.stdout + r.stderr
    sol = read_solution(out)
    ref = float.fromhex(golden["solves"]["afiro|presolve=1|fp64"]["objective"])
    assert sol["status"] == "Optimal" and sol["check"].startswith("PASS")
    assert abs(float(sol["objective"]) - ref) <= 1e-9 * (1 + abs(ref))
    import ps26119
    assert ps26119.check_solution(netlib("afiro"), sol["x"], sol["y"])["passed"]
    r2 = run("solve", AFIRO, "--warm", str(out), "--warm-weight")
    assert r2.returncode == 0 and "warm start" in r2.stdout
    r3 = run("solve", AFIRO, "--algorithm", "pdlp", "--precision", "mixed")
    assert r3.returncode == 0 and "engine     pdlp" in r3.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["afiro", "max_ranged", "infeasible", "unbounded"])
def test_gpu_solution_files_are_the_reference_writers(gpu_available, name, tmp_path):
    """Optimal (with row and column names, a maximisation with ranged rows) and the two
    certificates (DUAL_RAY, PRIMAL_RAY blocks), written by `solve --gpu --out`."""
    from conftest import hand_models
    from ps26119_amd import lpgen
    if name == "afiro":
        lpm = AFIRO
    elif name == "max_ranged":
        lpm = os.path.join(GOLDEN, "max_ranged.lpm")
    else:
        lpm = str(tmp_path / f"{name}.lpm")
        lpgen.write_lpm(hand_models()[name], lpm)
    out = tmp_path / f"{name}.sol"
    r = run("solve", lpm, "--gpu", "--out", str(out))
    assert r.returncode in (0, 1), r.stdout + r.stderr
    print(open(out).read()[:400], flush=True)
    assert_reference_writer_agrees(lpm, out, tmp_path)


@pytest.mark.gpu
def test_gpu_batch_command(gpu_available, tmp_path):
    from ps26119_amd import lpgen
    base = netlib("stocfor1")
    files = []
    for k, sc in enumerate(lpgen.perturbation_scenarios(base)):
        p = tmp_path / f"s{k}.lpm"
        lpgen.write_lpm(lpgen.with_scenario(base, sc), str(p))
        files.append(str(p))
    bp = tmp_path / "base.lpm"
    lpgen.write_lpm(base, str(bp))
    outdir = tmp_path / "sols"
    outdir.mkdir()
    r = run("batch", str(bp), *files, "--out-dir", str(outdir))
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Optimal") == 4
    assert len(list(outdir.glob("*.sol"))) == 4
    rm = run("batch", str(bp), *files, "--precision", "mixed")
    assert rm.returncode == 0 and rm.stdout.count("Optimal") == 4, rm.stdout + rm.stderr
    assert "fp32 increments" in rm.stdout


def test_batch_precision_must_be_known(tmp_path):
    from ps26119_amd import lpgen
    a = lpgen.random_lp(30, seed=1)
    pa = tmp_path / "a.lpm"
    lpgen.write_lpm(a, str(pa))
    r = run("batch", str(pa), str(pa), "--precision", "fp16")
    assert r.returncode == 2 and "unknown precision" in r.stderr
