"""C-ABI surface on the CPU: the library loads, exports every declared entry
point, and rejects malformed calls before touching a device."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fip_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def F():
    import fastimagingpipe_amd as F
    return F


def test_every_declared_symbol_is_exported(F):
    names = declared_functions()
    assert len(names) >= 30
    lib = ctypes.CDLL(F.library_path)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # C++ drop-in with the reference's signature (pipe_interface.cpp:8).
    assert hasattr(lib, "_Z6FIpipePfS_S_S_S_mmmfm")


def test_status_strings_and_version(F):
    assert F.version()
    texts = {F.lib.fip_status_string(code).decode() for code in F.STATUS}
    assert len(texts) == len(F.STATUS)          # one distinct message per status
    assert F.lib.fip_status_string(0) == b"ok"
    assert F.lib.fip_status_string(12345) == b"unknown status"


@pytest.mark.parametrize("image,g", [(64, 96), (2048, 3072), (4096, 6144), (8192, 12288)])
def test_grid_size_supported(F, image, g):
    assert F.grid_size(image) == g


@pytest.mark.parametrize("image", [0, 48, 100, 200, 16384])
def test_grid_size_unsupported(F, image):
    assert F.grid_size(image) == 0


def _arrays(nbl=16, nsnap=3, image=64):
    return (np.zeros((nsnap, nbl), np.float32), np.zeros((nsnap, nbl), np.float32),
            np.zeros((nsnap, nbl, 2), np.float32), np.tile(np.eye(3, dtype=np.float32), (nsnap, 1, 1)))


def _expect(F, status, fn, *args):
    with pytest.raises(F.FipError) as ei:
        fn(*args)
    assert ei.value.status == status, str(ei.value)


def test_pipe_argument_errors(F):
    vr, vi, b, v = _arrays()
    _expect(F, -2, F.fipipe, vr[:2], vi[:2], b[:2], v[:2], 16, 64, 2, 1e-3, 16)   # NSNAP
    _expect(F, -3, F.fipipe, vr, vi, b, v, 16, 100, 3, 1e-3, 16)                  # SIZE: image
    _expect(F, -3, F.fipipe, vr, vi, b, v, 16, 64, 3, 1e-3, 0)                    # SIZE: tile 0
    _expect(F, -3, F.fipipe, vr, vi, b, v, 16, 64, 3, 1e-3, 128)                  # SIZE: tile > image
    _expect(F, -1, F.fipipe, vr, vi, b, v, 16, 64, 3, float("nan"), 16)          # ARG: cell
    _expect(F, -1, F.fipipe, vr, vi, b, v, 0, 64, 3, 1e-3, 16)                    # ARG: no baselines
    out = np.zeros(16, np.float32)
    rc = F.lib.fip_pipe(None, vi.ctypes.data, b.ctypes.data, v.ctypes.data, out.ctypes.data,
                        16, 64, 3, 1e-3, 16)
    assert rc == -1
    assert b"null" in F.lib.fip_last_error()


def test_lowp_rejects_unknown_vis_type(F):
    vr, vi, b, v = _arrays()
    _expect(F, -1, F.fipipe_lowp, vr.view(np.uint16), vi.view(np.uint16), 7, b, v, 16, 64, 3,
            1e-3, 16)


def test_batch_host_needs_three_snapshots(F):
    vr, vi, b, v = _arrays(nsnap=2)
    _expect(F, -2, F.batch_host, vr, vi, b, v, None, 0, 64, 1e-3, 16)


def test_valid_call_without_device_reports_nodevice(F):
    if F.device_count() > 0:
        pytest.skip("a device is visible: covered by the GPU parity tests")
    vr, vi, b, v = _arrays()
    _expect(F, -6, F.fipipe, vr, vi, b, v, 16, 64, 3, 1e-3, 16)
    _expect(F, -6, F.project, np.ones((4, 3)), [1e8])


def test_synth_is_deterministic_and_projected(F):
    kw = dict(nant=12, radius_m=2000.0, nsnap=3, dt_s=10.0, freq_hz=1.4e8, dec_deg=-30.0,
              lat_deg=-26.7, image_size=256, cell=5e-4, nsrc=5, smin=0.1, smax=1.0, index=-0.8,
              noise_sigma=0.1, transients=[1], transient_flux=2.0)
    a = F.synth(seed=3, **kw)
    b = F.synth(seed=3, **kw)
    c = F.synth(seed=4, **kw)
    for k in ("vis_re", "vis_im", "bin", "vin"):
        assert np.array_equal(a[k], b[k]), k
    assert not np.array_equal(a["vis_re"], c["vis_re"])
    assert a["nbl"] == 66 and np.all(a["nvis"] == 66)
    for s in range(3):
        V = a["vin"][s].astype(np.float64)
        assert np.allclose(V @ V.T, np.eye(3), atol=1e-6)
        assert np.allclose(V[2], np.cross(V[0], V[1]), atol=1e-6)
    # Baselines are at most 2 * radius long: |B| <= 2 R f / c wavelengths.
    assert np.abs(a["bin"]).max() <= 2 * 2000.0 * 1.4e8 / 299792458.0 * (1 + 1e-5)


def test_synth_flagging_reports_counts(F):
    d = F.synth(nant=10, radius_m=1000.0, nsnap=4, dt_s=10.0, freq_hz=1.4e8, dec_deg=-30.0,
                lat_deg=-26.7, image_size=128, cell=1e-3, nsrc=2, smin=0.1, smax=1.0, index=-0.8,
                noise_sigma=0.0, flag_frac_max=0.5, seed=9)
    nv = d["nvis"]
    assert np.all(nv <= 45) and np.all(nv >= 22)
    for s in range(4):   # flagged rows are compacted away and padded with zeros
        assert np.all(d["vis_re"][s, nv[s]:] == 0) and np.all(d["bin"][s, nv[s]:] == 0)


def test_synth_rejects_bad_arguments(F):
    with pytest.raises(F.FipError):
        F.synth(nant=1, radius_m=1.0, nsnap=3, dt_s=1.0, freq_hz=1e8, dec_deg=0.0, lat_deg=0.0,
                image_size=64, cell=1e-3, nsrc=1, smin=0.1, smax=1.0, index=-0.8,
                noise_sigma=0.0)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in (r.stdout + r.stderr)


def test_cxx_caller_links_fipipe_by_its_reference_prototype():
    """tests/abi/fipipe_caller declares FIpipe as pipe_interface.cpp:8 does;
    its undefined symbol must be the mangled name libfip.so exports."""
    import os
    import subprocess
    caller = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi", "fipipe_caller")
    assert os.access(caller, os.X_OK), "build the caller first (__graft_entry__.build())"
    syms = subprocess.run(["nm", "-D", "--undefined-only", caller], capture_output=True, text=True,
                          check=True).stdout
    assert "_Z6FIpipePfS_S_S_S_mmmfm" in syms
    deps = subprocess.run(["ldd", caller], capture_output=True, text=True, check=True).stdout
    assert "libfip.so" in deps and "not found" not in deps.split("libfip.so")[1].splitlines()[0]


def test_noalloc_argument_errors_before_any_device_use(F):
    """fip_pipe_device_noalloc refuses what its static array cannot hold before
    it touches a device or a pointer (the pointers here are never read)."""
    fake = 256
    f = F.lib.fip_pipe_device_noalloc
    assert f(None, fake, fake, fake, fake, 2016, 256, 3, 5.8e-5, 16, None) == -1
    assert f(fake, fake, fake, fake, fake, 2016, 256, 2, 5.8e-5, 16, None) == -2
    assert f(fake, fake, fake, fake, fake, 2016, 2048, 3, 5.8e-5, 32, None) == -3   # beyond 1024^2
    assert b"64 MiB" in F.lib.fip_last_error()
    assert f(fake, fake, fake, fake, fake, 200000, 1024, 3, 2.9e-5, 32, None) == -3   # work > 64 MiB
    assert b"bytes of scratch" in F.lib.fip_last_error()


def test_raw_chain_argument_errors_before_any_device_use(F):
    """fip_pipe_raw_device validates its view and shape first (fake, never-read
    device pointers)."""
    fake = 256
    keys = ("uvw", "freqs", "xx_re", "xx_im", "yy_re", "yy_im", "flag0", "flag3", "w0", "w3")
    ptrs = {k: fake for k in keys}
    _expect(F, -2, F.pipe_raw_device, ptrs, 2016, 1, 2, 256, 5.8e-5, 16, fake)     # NSNAP
    _expect(F, -1, F.pipe_raw_device, ptrs, 0, 1, 3, 256, 5.8e-5, 16, fake)        # no rows
    _expect(F, -3, F.pipe_raw_device, ptrs, 2016, 1, 3, 100, 5.8e-5, 16, fake)     # image size
    _expect(F, -3, F.pipe_raw_device, ptrs, 1 << 30, 1 << 20, 3, 256, 5.8e-5, 16, fake)   # nrows*nchan
    _expect(F, -1, F.pipe_raw_device, dict(ptrs, w3=0), 2016, 1, 3, 256, 5.8e-5, 16, fake)


def test_server_argument_errors_before_any_device_use(F):
    """The serving entry points reject null and malformed arguments before they
    create a context or touch a device."""
    lib = F.lib
    n = ctypes.c_long(7)
    assert lib.fip_server_sequences(None, ctypes.byref(n)) == -1
    assert n.value == 7
    assert lib.fip_server_drain(None) == -1
    assert lib.fip_server_stats(None, None, None, None, None, None) == -1
    out = ctypes.c_void_p()
    cfg = (F._StreamConfig * 1)()
    cfg[0].num_baselines, cfg[0].image_size, cfg[0].cell_size, cfg[0].unit_size = 2016, 256, 5.8e-5, 16
    assert lib.fip_server_create(cfg, 1, 2, 1 << 30, ctypes.byref(out)) == -1     # ring < 3
    assert lib.fip_server_create(cfg, 0, 8, 1 << 30, ctypes.byref(out)) == -1     # no stream
    cfg[0].vis_type = 9
    assert lib.fip_server_create(cfg, 1, 8, 1 << 30, ctypes.byref(out)) == -1     # bad vis_type
    cfg[0].vis_type = 0
    cfg[0].image_size = 100
    assert lib.fip_server_create(cfg, 1, 8, 1 << 30, ctypes.byref(out)) != 0      # unsupported size
    cfg[0].image_size = 256
    assert lib.fip_server_create(cfg, 1, 8, 1 << 16, ctypes.byref(out)) == -7     # budget, before any device use
    assert not out.value
