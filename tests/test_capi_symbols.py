"""CPU checks of the drop-in boundary: librama_hip.so loads and exports every
function include/rama_hip.h declares, and the host-only option helpers behave
like the reference's multicut_solver_options.  No kernel runs here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "rama_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(rama_[a-z_0-9]+)\s*\(", src)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    from rama_amd import _capi

    L = _capi.lib()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python binding declares argument types for every one of them
    assert set(names) <= set(_capi.SIGNATURES)


def test_options_presets_match_reference():
    from rama_amd import _capi

    d = _capi.options(None)
    assert (d.max_cycle_length_lb, d.num_dual_itr_lb, d.max_cycle_length_primal, d.num_dual_itr_primal,
            d.num_outer_itr_dual) == (5, 10, 3, 5, 1)
    assert abs(d.mean_multiplier_mm - 0.05) < 1e-7 and abs(d.tri_memory_factor - 2.0) < 1e-7
    assert d.max_time_sec == -1 and d.verbose == 1 and d.only_compute_lb == 0
    p = _capi.options("P")
    assert (p.max_cycle_length_lb, p.num_dual_itr_lb, p.max_cycle_length_primal, p.num_dual_itr_primal) == (0, 0, 0, 0)
    pp = _capi.options("PD+")
    assert (pp.max_cycle_length_primal, pp.num_dual_itr_primal) == (5, 10)
    dd = _capi.options("D")
    assert dd.only_compute_lb == 1 and dd.num_outer_itr_dual == 5
    o = _capi.Options()
    # unknown preset: PD defaults, status 1 (the reference never throws its error)
    assert _capi.lib().rama_options_preset(b"XYZ", ctypes.byref(o)) == 1
    assert o.max_cycle_length_lb == 5


def test_rama_py_options_surface(capfd):
    from rama_amd import rama_py

    o = rama_py.multicut_solver_options("PD")
    assert "best compute time" in capfd.readouterr().out
    assert repr(o).startswith("<multicut_solver_options>:max_cycle_length_lb: 5, num_dual_itr_lb: 10")
    o.verbose = False
    o.dump_timeline = True
    assert o.dump_timeline and not o.verbose
    o11 = rama_py.multicut_solver_options(4, 7, 3, 2, 1, 0.1, 0.2, 1.5, True, 9, True)
    assert (o11.max_cycle_length_lb, o11.num_dual_itr_lb, o11.max_time_sec) == (4, 7, 9)
    assert o11.only_compute_lb and o11.dump_timeline and not o11.sanitize_graph
    for name in ("rama_cuda", "rama_cuda_gpu_pointers", "read_multicut_file", "rama_torch"):
        assert hasattr(rama_py, name)


def test_read_multicut_file(tmp_path):
    from rama_amd import rama_py

    p = tmp_path / "g.txt"
    p.write_text("MULTICUT\n0 1 1.5\n1 2 -2\n0 2 0.25\n")
    i, j, c = rama_py.read_multicut_file(str(p))
    assert i == [0, 1, 0] and j == [1, 2, 2] and c == pytest.approx([1.5, -2.0, 0.25])
    bad = tmp_path / "b.txt"
    bad.write_text("GRAPH\n0 1 1\n")
    with pytest.raises(RuntimeError):
        rama_py.read_multicut_file(str(bad))


def test_objective_host_helper():
    from rama_amd import _capi

    assert _capi.objective([0, 1, 0], [0, 1, 2], [1, 2, 0], [1.1, -2, 3]) == pytest.approx(-0.9, abs=1e-6)
