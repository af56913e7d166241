"""fip_pipe_device_noalloc (the entry point that allocates nothing) against
fip_pipe_device on the same device inputs, through the C ABI.

Two calls in a row on the shared static array, the second with a smaller
input (fewer baselines, smaller image), each compared with the allocating
entry point; the small one also against the FP64 oracle.  Sizes beyond the
array are refused before anything is launched."""
import numpy as np
import pytest

from fastimagingpipe_amd import hip
from tests.inputs import snapshots

pytestmark = pytest.mark.gpu


def _device_inputs(vr, vi, B, V):
    return [hip.DeviceArray(np.ascontiguousarray(a)) for a in (vr, vi, B, V)]


def _run(fip, entry, dev, nbl, image, cell, tile):
    un = image // tile
    out = hip.DeviceArray(nbytes=un * un * 4)
    stream = hip.Stream()
    entry(dev[0].addr, dev[1].addr, dev[2].addr, dev[3].addr, out.addr, nbl, image, 3, cell, tile, stream.addr)
    stream.synchronize()
    return out.to_host(np.float32, (un, un))


# (image, baselines, cell, tile): the largest shape the 64 MiB array holds
# (1024^2 with 130816 baselines, 66.1 MB of work), then a smaller one.
CASES = [(1024, 130816, 2.9e-5, 32), (256, 2016, 5.8e-5, 16)]


def test_noalloc_matches_pipe_device_two_calls(fip, oracle):
    for k, (image, nbl, cell, tile) in enumerate(CASES):
        print(f"noalloc call {k}: image {image} baselines {nbl} tile {tile}", flush=True)
        vr, vi, B, V = snapshots(700 + k, 3, nbl, image, cell)
        dev = _device_inputs(vr, vi, B, V)
        got = _run(fip, fip.pipe_device_noalloc, dev, nbl, image, cell, tile)
        ref = _run(fip, fip.pipe_device, dev, nbl, image, cell, tile)
        err = float(np.abs(got - ref).max())
        same = np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        print(f"  max |noalloc - pipe_device| = {err:.3e} (bit-identical: {same})")
        assert np.isfinite(got).all()
        assert err <= 1e-6
        if image == 256:
            orc = oracle.pipe(vr, vi, B, V, nbl, image, 3, cell, tile)
            e_orc = float(np.abs(got - orc).max())
            print(f"  max |noalloc - fp64 oracle| = {e_orc:.3e}")
            assert e_orc < 2e-5


def test_noalloc_refuses_shapes_beyond_its_array(fip):
    vr, vi, B, V = snapshots(9, 3, 64, 256, 5.8e-5)
    dev = _device_inputs(vr, vi, B, V)
    out = hip.DeviceArray(nbytes=64 * 64 * 4)
    with pytest.raises(fip.FipError) as e:
        fip.pipe_device_noalloc(dev[0].addr, dev[1].addr, dev[2].addr, dev[3].addr, out.addr, 64, 2048, 3,
                                1e-5, 32)
    assert e.value.status == -3
    # 1024^2 with 200000 baselines needs more than 64 MiB of work buffers
    # (inputs of that size, so that nothing could be read out of bounds)
    n = 200000
    big = [hip.DeviceArray(np.zeros(a, np.float32)) for a in ((3, n), (3, n), (3, n, 2), (3, 3, 3))]
    out = hip.DeviceArray(nbytes=32 * 32 * 4)
    with pytest.raises(fip.FipError) as e:
        fip.pipe_device_noalloc(big[0].addr, big[1].addr, big[2].addr, big[3].addr, out.addr, n, 1024, 3,
                                2.9e-5, 32)
    assert e.value.status == -3
