"""The reference's C++ entry point called from C++: tests/abi/fipipe_caller is
compiled against the FIpipe prototype of pipe_interface.cpp:8 (not fip.h) and
linked with -lfip, as an unmodified pipe_interface.o would be.  Its result
must be bit-identical to fip_pipe's on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from tests.inputs import snapshots

pytestmark = pytest.mark.gpu

CALLER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi", "fipipe_caller")


@pytest.mark.parametrize("image,nbl,cell,tile", [(256, 2016, 5.8e-5, 16), (1024, 32640, 2.9e-5, 32)])
def test_cxx_fipipe_caller_matches_fip_pipe(fip, tmp_path, image, nbl, cell, tile):
    assert os.access(CALLER, os.X_OK), "build the caller first (__graft_entry__.build())"
    vr, vi, B, V = snapshots(31 + image, 3, nbl, image, cell)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        for a in (vr, vi, B, V):
            np.ascontiguousarray(a, np.float32).tofile(f)
    cell32 = repr(float(np.float32(cell)))
    r = subprocess.run([CALLER, str(inp), str(out), str(nbl), str(image), "3", cell32, str(tile)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    un = image // tile
    got = np.fromfile(out, np.float32).reshape(un, un)
    ref = fip.fipipe(vr, vi, B, V, nbl, image, 3, cell, tile)
    print(f"C++ FIpipe caller at {image}^2: bit-identical to fip_pipe = "
          f"{np.array_equal(got.view(np.uint32), ref.view(np.uint32))}")
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
