"""The verbose solve (src/rama_cuda.cu prints every message-passing
iteration's bound, so the loop is driven from the host: bound from the
per-block partials plus the edges in no triangle, stall test on the host)
against the device-driven loop of a quiet solve: the same partition and the
same bound to 1e-9 relative (the two sum the partials in different orders)."""
import numpy as np
import pytest

from tests.instances_small import named

pytestmark = pytest.mark.gpu

CASES = ["grid64", "grid96x128_lr", "grid3d_24", "er800_dense", "random_dup"]


@pytest.mark.parametrize("preset", ["PD", "PD+", "D"])
def test_verbose_loop_equals_device_loop(gpu, preset, capfd):
    for name in CASES:
        i, j, c = named(name)
        lab, lb = gpu.solve(i, j, c, gpu.options(preset, verbose=0))[:2]
        vlab, vlb = gpu.solve(i, j, c, gpu.options(preset, verbose=1))[:2]
        capfd.readouterr()  # the per-iteration lines
        assert abs(vlb - lb) <= 1e-9 * max(1.0, abs(lb)), (name, vlb, lb)
        assert np.array_equal(np.asarray(vlab), np.asarray(lab)), name
