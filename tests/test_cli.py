"""The command-line front end rama_text_input (src/rama_text_input.cu:8-50 and
the options of include/multicut_solver_options.h:82-105 of RAMA).

CPU tests cover the option parser (no solve runs); the GPU test solves a
file end to end and reads the printed objective, bound and solution file
back."""
import os
import subprocess

import numpy as np
import pytest

from tests.instances_small import named

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "rama_amd", "rama_text_input")


def run(*args):
    return subprocess.run([CLI] + [str(a) for a in args], capture_output=True, text=True, timeout=120)


def write_instance(path, i, j, c):
    with open(path, "w") as f:
        f.write("MULTICUT\n")
        for a, b, x in zip(i, j, c):
            f.write("%d %d %.9g\n" % (a, b, x))


def test_help_exits_zero():
    r = run("--help")
    assert r.returncode == 0
    assert "--out_sol_file" in r.stdout and "max_cycle_dual" in r.stdout


@pytest.mark.parametrize("args, code, text", [
    ((), 106, "--file is required"),
    (("/nonexistent/instance.txt",), 105, "Path does not exist"),
    (("{f}", "7"), 105, "max_cycle_dual"),
    (("{f}", "5", "-3"), 105, "dual_itr"),
    (("{f}", "5", "10", "3", "5", "1", "0.05", "0.05", "0"), 105, "tri_memory_factor"),
    (("{f}", "five"), 104, "Could not convert"),
    (("{f}", "5", "10", "3", "5", "1", "0.05", "0.05", "2.0", "9"), 109, "not expected"),
    (("--bogus", "{f}"), 109, "not expected"),
    (("-f",), 114, "requires a value"),
])
def test_parse_errors(tmp_path, args, code, text):
    f = tmp_path / "g.txt"
    write_instance(f, [0], [1], [1.0])
    r = run(*[a.format(f=f) for a in args])
    assert r.returncode == code, (r.returncode, r.stderr)
    assert text in r.stderr


def _parse_output(out):
    vals = {}
    for line in out.splitlines():
        line = line.strip()
        if line.startswith("cost w.r.t original objective:"):
            vals["objective"] = float(line.split(":")[1])
        elif line.startswith("final lower bound:"):
            vals["lb"] = float(line.split(":")[1])
        elif line.startswith("GPU compute time:"):
            vals["ms"] = int(line.split(":")[1].strip().rstrip("ms"))
    return vals


def _read_solution(path, n):
    labels = np.full(n, -1, np.int64)
    with open(path) as f:
        for c, line in enumerate(f):
            assert line.endswith("\n")
            for tok in line.rstrip("\n").split("\t"):
                if tok:
                    labels[int(tok)] = c
    return labels


def _same_partition(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return len(set(zip(a.tolist(), b.tolist()))) == len(set(a.tolist())) == len(set(b.tolist()))


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["grid64", "er3000"])
def test_cli_round_trip(tmp_path, oracle, name):
    i, j, c = named(name)
    f, sol = tmp_path / "inst.txt", tmp_path / "sol.txt"
    write_instance(f, i, j, c)
    # values as the file holds them (9 significant digits, exact for fp32)
    c = np.float32([float("%.9g" % x) for x in c])
    for preset, extra in (("PD", []), ("PD+", ["5", "10", "5", "10", "1"])):
        r = run("-f", f, "-o", sol, *extra)
        assert r.returncode == 0, r.stderr
        vals = _parse_output(r.stdout)
        assert "Writing solution to file" in r.stdout
        ref, ref_lb, _ = oracle.solve(i, j, c, oracle.options(preset))
        lab = _read_solution(sol, len(ref))
        assert np.all(lab >= 0)
        assert _same_partition(lab, ref), preset
        ref_obj = oracle.objective(ref, i, j, c)
        # std::cout prints 6 significant digits
        assert float("%.6g" % ref_obj) == vals["objective"]
        assert float("%.6g" % ref_lb) == vals["lb"]
        assert vals["ms"] >= 0
    r = run(f, "--only_lb")
    assert r.returncode == 0, r.stderr
    vals = _parse_output(r.stdout)
    assert "objective" not in vals
    _, d_lb, _ = oracle.solve(i, j, c, oracle.options("PD", only_compute_lb=1))
    assert float("%.6g" % d_lb) == vals["lb"]
