"""The device chain from raw visibilities to trigger (fip_pipe_raw_device)
against the oracle's collapse, projection and pipe (SVD_MFS.py:38-110;
SVDPreprocessKernels.cu:208-424; FIPkernels.cu:554-862), through the C ABI."""
import numpy as np
import pytest

from fastimagingpipe_amd import hip
from tests.inputs import raw_snapshots, triggered

pytestmark = pytest.mark.gpu

# (name, antennas, radius m, channels, image, cell, tile): the 64-antenna
# array at 256^2 (2016 baselines), the 512-antenna array at 4096^2 (130816
# baselines), and four channels of a 96-antenna array (MFS, 18240 samples).
CASES = [("64ant_256px", 64, 5000.0, [140e6], 256, 5.8e-5, 16),
         ("512ant_4096px", 512, 40000.0, [140e6], 4096, 7.2722e-6, 32),
         ("96ant_4chan_512px", 96, 8000.0, [138e6, 139e6, 140e6, 141e6], 512, 2.9e-5, 32)]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_raw_device_chain_matches_oracle(fip, oracle, case):
    name, nant, radius, freqs, image, cell, tile = case
    uvw, fq, xx, yy, f0, f3, w0, w3 = raw_snapshots(1000 + nant, nant, radius, 3, freqs, image, cell)
    nrows, nchan = uvw.shape[1], fq.size
    ns = nrows * nchan
    print(f"{name}: {nrows} rows x {nchan} channels, image {image}", flush=True)
    dev = {"uvw": hip.DeviceArray(uvw), "freqs": hip.DeviceArray(fq),
           "xx_re": hip.DeviceArray(np.ascontiguousarray(xx.real, np.float32)),
           "xx_im": hip.DeviceArray(np.ascontiguousarray(xx.imag, np.float32)),
           "yy_re": hip.DeviceArray(np.ascontiguousarray(yy.real, np.float32)),
           "yy_im": hip.DeviceArray(np.ascontiguousarray(yy.imag, np.float32)),
           "flag0": hip.DeviceArray(f0), "flag3": hip.DeviceArray(f3),
           "w0": hip.DeviceArray(w0), "w3": hip.DeviceArray(w3)}
    un = image // tile
    out = hip.DeviceArray(nbytes=un * un * 4)
    bout = hip.DeviceArray(nbytes=3 * ns * 8)
    vout = hip.DeviceArray(nbytes=3 * 9 * 4)
    st = hip.Stream()
    fip.pipe_raw_device({k: v.addr for k, v in dev.items()}, nrows, nchan, 3, image, cell, tile, out.addr,
                        bout.addr, vout.addr, stream=st.addr)
    st.synchronize()
    got = out.to_host(np.float32, (un, un))
    gb = bout.to_host(np.float32, (3, ns, 2))
    gv = vout.to_host(np.float32, (3, 3, 3))
    vr = np.zeros((3, ns), np.float32)
    vi = np.zeros((3, ns), np.float32)
    ob = np.zeros((3, ns, 2), np.float32)
    ov = np.zeros((3, 3, 3), np.float32)
    for s in range(3):
        vis = oracle.collapse(xx[s], yy[s], f0[s], f3[s], w0[s], w3[s])
        vr[s], vi[s] = vis.real, vis.imag
        b, v = oracle.project(uvw[s], fq)
        eb = np.abs(gb[s] - b).max() / np.abs(b).max()
        ev = np.abs(gv[s] - v).max()
        print(f"  snapshot {s}: B rel err {eb:.2e}, V err {ev:.2e}")
        assert eb <= 2e-7 and ev <= 2e-7
        ob[s], ov[s] = b, v
    ref = oracle.pipe(vr, vi, ob, ov, ns, image, 3, cell, tile)
    err = float(np.abs(got - ref).max())
    print(f"  tLISI max abs err vs oracle collapse+project+pipe {err:.2e}, min {ref.min():.6f}")
    assert err < 1e-6
    for tau in (0.9, 0.99, 0.999):
        assert triggered(got, tau) == triggered(ref, tau), tau


def test_project_and_collapse_reuse_context_buffers(fip, oracle):
    """The host-array projection and collapse agree with the oracle call
    after call at changing sizes (their device buffers are the context's)."""
    rng = np.random.default_rng(12)
    for n in (5000, 130816, 777):
        uvw = rng.normal(size=(n, 3)) * np.array([2e4, 2e4, 1e3])
        B, V = fip.project(uvw, [140e6, 150e6])
        ob, ov = oracle.project(uvw, [140e6, 150e6])
        assert np.abs(V - ov).max() < 2e-7
        assert np.abs(B - ob).max() <= 2e-7 * np.abs(ob).max()
        xx = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        yy = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
        f0 = (rng.uniform(size=n) < 0.2).astype(np.uint8)
        f3 = (rng.uniform(size=n) < 0.2).astype(np.uint8)
        w0 = rng.uniform(0, 2, n).astype(np.float32)
        w3 = rng.uniform(0, 2, n).astype(np.float32)
        got = fip.collapse(xx, yy, f0, f3, w0, w3)
        ref = oracle.collapse(xx, yy, f0, f3, w0, w3)
        assert np.abs(got - ref).max() < 1e-5 * np.abs(ref).max()
