"""The allocation-free solve (rama_solve_device_noalloc) against the device
entry it is a variant of (rama_solve_device), through the C ABI: two calls
in a row on one GPU, the second on a smaller graph (the static device array
keeps the first call's contents), then the size limit, which the host
rejects before any device work."""
import numpy as np
import pytest

from rama_amd import instances as I

pytestmark = pytest.mark.gpu


def _device_solve(gpu, torch, entry, inst, opts):
    i, j, c = inst
    di, dj, dc = (torch.from_numpy(np.ascontiguousarray(x)).to("cuda:0") for x in (i, j, c))
    cap = int(max(i.max(), j.max())) + 1
    labels = torch.full((cap,), -7, dtype=torch.int32, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    if entry == "device":
        n, lb, _, _ = gpu.solve_device(di.data_ptr(), dj.data_ptr(), dc.data_ptr(), len(c), labels.data_ptr(), cap,
                                       opts, device=0, stream=stream)
        peak = None
    else:
        n, lb, _, peak = gpu.solve_device_noalloc(di.data_ptr(), dj.data_ptr(), dc.data_ptr(), len(c),
                                                  labels.data_ptr(), cap, opts, device=0, stream=stream)
    torch.cuda.synchronize()
    # inputs are not modified
    assert np.array_equal(di.cpu().numpy(), i) and np.array_equal(dc.cpu().numpy(), c)
    return labels[:n].cpu().numpy(), lb, peak


@pytest.mark.parametrize("preset", ["PD", "PD+", "D"])
def test_noalloc_equals_device_entry_twice(gpu, preset):
    import torch

    opts = gpu.options(preset, verbose=0)
    big = I.grid2d_voronoi(H=160, W=640, n_seeds=100, seed=4)[:3]   # 584,800 edges
    small = I.grid2d_rectangles(seed=2)[:3]                          # 8,064 edges
    for name, inst in (("big", big), ("small", small)):
        print("noalloc case: preset %s, %s, %d edges" % (preset, name, len(inst[2])), flush=True)
        ref_lab, ref_lb, _ = _device_solve(gpu, torch, "device", inst, opts)
        lab, lb, peak = _device_solve(gpu, torch, "noalloc", inst, opts)
        assert lb == ref_lb
        assert np.array_equal(lab, ref_lab)
        assert 0 < peak <= 256 << 20  # RAMA_NOALLOC_SCRATCH_BYTES
        print("  scratch peak %d bytes (%.1f per edge)" % (peak, peak / len(inst[2])), flush=True)


def test_noalloc_rejects_more_edges_than_its_limit(gpu):
    import torch

    m = gpu.NOALLOC_MAX_EDGES + 1
    i = torch.arange(m, dtype=torch.int32, device="cuda:0")
    j = i + 1
    c = torch.ones(m, dtype=torch.float32, device="cuda:0")
    labels = torch.empty(m + 1, dtype=torch.int32, device="cuda:0")
    with pytest.raises(gpu.RamaError) as e:
        gpu.solve_device_noalloc(i.data_ptr(), j.data_ptr(), c.data_ptr(), m, labels.data_ptr(), m + 1,
                                 gpu.options("PD", verbose=0), device=0,
                                 stream=torch.cuda.current_stream().cuda_stream)
    assert "exceed the limit" in str(e.value)
    assert e.value.status == 1  # RAMA_ERR_INVALID
