"""Pins the FP64 oracle's imaging, map, splat, trigger and preprocessing
functions against closed forms, brute force and scikit-learn fixtures (CPU)."""
import json
import os

import numpy as np
import pytest

from tests.inputs import rotation

GOLD = os.path.join(os.path.dirname(__file__), "golden")
NORM = json.load(open(os.path.join(GOLD, "reference_quadrature.json")))["norm"]


def _taps(p, g):
    lo = max(int(np.ceil(p - 4)), -g // 2)
    hi = min(int(np.floor(p + 4)), (g - 1) // 2)
    return np.arange(lo, hi + 1)


def direct_tap_dft(oracle, vis, B, V, image, cell):
    """Re sum_n sum_taps v_n ES ES e^{+2 pi i (r1 y + r2 x)/g} / J, corrected and |.|,
    evaluated directly (no FFT, no modulation)."""
    g = oracle.grid_size(image)
    h = image // 2
    Vd = V.astype(np.float64)
    J = abs(Vd[0, 0] * Vd[1, 1] - Vd[0, 1] * Vd[1, 0])
    scale = np.float64(np.float32(cell)) * g
    y = np.arange(-h, h)
    acc = np.zeros((image, image), complex)
    for v, (b0, b1) in zip(vis, B):
        p1, p2 = np.float64(b0) * scale, np.float64(b1) * scale
        r1, r2 = _taps(p1, g), _taps(p2, g)
        w1 = np.array([oracle.es((r - p1) / 4) for r in r1])
        w2 = np.array([oracle.es((r - p2) / 4) for r in r2])
        e1 = (w1[:, None] * np.exp(2j * np.pi * np.outer(r1, y) / g)).sum(0)   # over r1 -> y
        e2 = (w2[:, None] * np.exp(2j * np.pi * np.outer(r2, y) / g)).sum(0)   # over r2 -> x
        acc += (v / J) * np.outer(e1, e2)
    c = oracle.conv_corr(image)
    cy = c[np.abs(y)]
    return np.abs(acc.real) / (cy[:, None] * cy[None, :] * NORM * NORM)


def _one_snapshot(seed, nbl, image, cell, tilt=0.2):
    rng = np.random.default_rng(seed)
    g = int(np.ceil(1.5 * image))
    p = rng.uniform(-0.3 * g, 0.3 * g, (nbl, 2))
    B = (p / (np.float64(np.float32(cell)) * g)).astype(np.float32)
    vis = (rng.standard_normal(nbl) + 1j * rng.standard_normal(nbl)).astype(np.complex64)
    V = rotation(tilt, rng.uniform(0, 6.3), rng.uniform(0, 6.3)).astype(np.float32)
    return vis, B, V


@pytest.mark.parametrize("image,nbl", [(64, 1), (64, 7), (128, 25)])
def test_grid_fft_crop_equals_direct_tap_dft(oracle, image, nbl):
    """FIPkernels.cu:197-287: modulation + inverse FFT + crop + flip == direct DFT."""
    cell = 1e-4
    vis, B, V = _one_snapshot(image + nbl, nbl, image, cell)
    g = oracle.grid_size(image)
    G = oracle.grid(vis.real, vis.imag, B, V, g, cell)
    pre = oracle.crop_correct(oracle.fft2_inverse(G), image)
    ref = direct_tap_dft(oracle, vis, B, V, image, cell)
    # 1e-8: the norm constant above is the reference literal (about 10 digits).
    assert np.abs(pre - ref).max() <= 1e-8 * max(1.0, np.abs(ref).max())


def test_single_visibility_known_answer(oracle):
    """One visibility images to |Re(v e^{2 pi i (p1 y + p2 x)/g})|/J up to the
    gridding kernel's correction error."""
    image, cell = 256, 5.8e-5
    g = oracle.grid_size(image)
    h = image // 2
    p1, p2 = 37.3, -12.6
    B = np.array([[p1 / (np.float64(np.float32(cell)) * g), p2 / (np.float64(np.float32(cell)) * g)]],
                 dtype=np.float32)
    V = np.eye(3, dtype=np.float32)
    v = 0.8 - 0.35j
    G = oracle.grid([v.real], [v.imag], B, V, g, cell)
    pre = oracle.crop_correct(oracle.fft2_inverse(G), image)
    P1 = np.float64(B[0, 0]) * np.float64(np.float32(cell)) * g
    P2 = np.float64(B[0, 1]) * np.float64(np.float32(cell)) * g
    y = np.arange(-h, h)
    kat = np.abs((v * np.exp(2j * np.pi * (P1 * y[:, None] + P2 * y[None, :]) / g)).real)
    assert np.abs(pre - kat).max() < 3e-5 * abs(v)


def test_gridding_is_linear(oracle):
    image, cell = 64, 2e-4
    g = oracle.grid_size(image)
    a, B, V = _one_snapshot(1, 9, image, cell)
    b = (np.arange(9) * (0.1 + 0.2j)).astype(np.complex64)
    ga = oracle.grid(a.real, a.imag, B, V, g, cell)
    gb = oracle.grid(b.real, b.imag, B, V, g, cell)
    s = (a + b).astype(np.complex64)
    gs = oracle.grid(s.real, s.imag, B, V, g, cell)
    assert np.abs(gs - ga - gb).max() < 1e-6 * np.abs(gs).max()


def test_identity_geometry_is_column_mirror(oracle):
    """V = I: xi = eta = 0, p2p is the identity and coordschange maps (i, j) -> (2h - i, j)."""
    image = 64
    h = image // 2
    idx = oracle.coord_map(np.eye(3, dtype=np.float32), image, 1e-4)
    i, j = np.meshgrid(np.arange(image), np.arange(image), indexing="ij")
    # The degree-trig chain leaves ~1e-12 of round-off at the identity.
    assert np.abs(idx[..., 0] - (2 * h - i)).max() < 1e-9
    assert np.abs(idx[..., 1] - j).max() < 1e-9


@pytest.mark.parametrize("tilt,cell", [(0.1, 7.27e-6), (0.4, 5.8e-5), (0.8, 2e-4)])
def test_sin_degree_chain_equals_closed_form(oracle, tilt, cell):
    """FIPkernels.cu:301-376 == P1 - xi n'/dc, P2 + eta n'/dc with n' = 1 - sqrt(1 - r^2)."""
    image = 256
    V = rotation(tilt, 0.7, 1.9).astype(np.float32)
    aff = oracle.coord_map(V, image, cell, sin_projection=False)
    got = oracle.coord_map(V, image, cell, sin_projection=True)
    c0 = image / 2 + 1
    d = np.float64(np.float32(cell))
    l = -d * (aff[..., 0] - c0)
    m = d * (aff[..., 1] - c0)
    r2 = l * l + m * m
    n1 = r2 / (1 + np.sqrt(1 - r2))
    xi = np.float64(V[2, 0]) / np.float64(V[2, 2])
    eta = np.float64(V[2, 1]) / np.float64(V[2, 2])
    assert np.abs(got[..., 0] - (aff[..., 0] - xi * n1 / d)).max() < 1e-7
    assert np.abs(got[..., 1] - (aff[..., 1] + eta * n1 / d)).max() < 1e-7


def test_splat_integer_positions_copy_exactly(oracle):
    image, nbl = 64, 10
    h = image // 2
    rng = np.random.default_rng(5)
    pre = rng.uniform(0, 1, (image, image))
    idx = oracle.coord_map(np.eye(3, dtype=np.float32), image, 1e-4, sin_projection=False)
    out = oracle.splat(idx, pre, nbl)
    expect = np.zeros_like(pre)
    for j in range(image):
        for i in range(image):
            LL, MM = h - i, j - h
            if abs(LL) < h - 1 and abs(MM) < h - 1:
                expect[MM + h, LL + h] += pre[j, i] / nbl
    assert np.abs(out - expect).max() < 1e-15


def test_splat_conserves_mass(oracle):
    image, nbl = 128, 3
    h = image // 2
    rng = np.random.default_rng(6)
    pre = rng.uniform(0, 1, (image, image))
    V = rotation(0.3, 1.1, 0.4).astype(np.float32)
    idx = oracle.coord_map(V, image, 7e-5)
    out = oracle.splat(idx, pre, nbl)
    LL = idx[..., 0] - h   # [i, j]
    MM = idx[..., 1] - h
    inside = (np.abs(LL) < h - 1) & (np.abs(MM) < h - 1)
    assert out.sum() == pytest.approx((pre.T[inside]).sum() / nbl, rel=1e-12)
    assert out.min() >= 0.0


def test_max_starts_from_zero(oracle):
    assert oracle.image_max(-np.ones((4, 4))) == 0.0
    assert oracle.image_max(np.array([[0.5, 2.0], [1.0, 0.0]])) == 2.0


def test_trigger_constant_images_give_one(oracle):
    d = np.full((64, 64), 0.7)
    assert np.array_equal(oracle.trigger(d, d, d, [0.7] * 3, 16), np.ones((4, 4)))


def test_trigger_single_pixel_hand_value(oracle):
    I, T, a, dlt = 64, 16, 0.5, 0.25
    d1 = np.full((I, I), a)
    d2 = d1.copy()
    d3 = d1.copy()
    d3[20, 37] += dlt
    out = oracle.trigger(d1, d2, d3, [a, a, a + dlt], T)
    expect = np.ones((4, 4))
    expect[1, 2] = 1 - (dlt / T / T) * dlt * (min(dlt / a, 1) / T / T) / (a + dlt) ** 2
    assert np.abs(out - expect).max() < 1e-15
    # A change in the middle snapshot alone cancels (| |d1-d2| - |d2-d3| | = 0).
    d2m = d1.copy()
    d2m[20, 37] += dlt
    assert np.array_equal(oracle.trigger(d1, d2m, d1, [a, a + dlt, a], T), np.ones((4, 4)))


def test_trigger_nonpositive_guard(oracle):
    I, T = 32, 16
    z = np.zeros((I, I))
    d3 = z.copy()
    d3[3, 4] = 2e-7     # ratio 2e-7 / 1e-6f < 1 is kept
    d3[20, 25] = 5.0    # ratio capped at 1
    out = oracle.trigger(z, z, d3, [0.0, 0.0, 5.0], T)
    C = float(np.float32(1e-6))
    e00 = 1 - (2e-7 / T / T) * 2e-7 * ((2e-7 / C) / T / T) / 25.0
    e11 = 1 - (5.0 / T / T) * 5.0 * (1.0 / T / T) / 25.0
    assert out[0, 0] == pytest.approx(e00, abs=1e-15)
    assert out[1, 1] == pytest.approx(e11, abs=1e-15)


def test_trigger_matches_numpy_brute_force(oracle):
    rng = np.random.default_rng(7)
    I, T = 64, 8
    d1, d2, d3 = (rng.uniform(-0.1, 1, (I, I)) for _ in range(3))
    mx = [max(0.0, x.max()) for x in (d1, d2, d3)]
    out = oracle.trigger(d1, d2, d3, mx, T)
    D = np.abs(np.abs(d1 - d2) - np.abs(d2 - d3))
    S = np.where(d2 <= 0, float(np.float32(1e-6)), d2)
    R = np.minimum(D / S, 1.0)
    tiles = lambda A: A.reshape(I // T, T, I // T, T).swapaxes(1, 2).reshape(I // T, I // T, -1)
    mu, mxd, rho = tiles(D).mean(-1), tiles(D).max(-1), tiles(R).mean(-1)
    assert np.allclose(out, 1 - mu * mxd * rho / max(mx) ** 2, atol=1e-14)


def test_projection_matches_sklearn_fixtures(oracle):
    fx = np.load(os.path.join(GOLD, "pca_fixtures.npz"))
    for k in range(4):
        B, V = oracle.project(fx[f"uvw{k}"], fx[f"freqs{k}"])
        assert np.allclose(V, fx[f"V{k}"], atol=1e-12), k
        assert np.abs(B - fx[f"B{k}"]).max() < 1e-9 * np.abs(fx[f"B{k}"]).max(), k


def test_collapse_special_cases(oracle):
    xx = np.array([1 + 1j, 2 - 1j, 3 + 0j, 4 + 4j])
    yy = np.array([3 - 1j, 0 + 0j, 1 + 1j, 2 + 2j])
    f0 = np.array([0, 0, 1, 1], np.uint8)
    f3 = np.array([0, 1, 0, 1], np.uint8)
    w0 = np.array([1.0, 2.0, 1.0, 1.0])
    w3 = np.array([1.0, 1.0, 3.0, 1.0])
    out = oracle.collapse(xx, yy, f0, f3, w0, w3)
    assert out[0] == pytest.approx((xx[0] + yy[0]) / 2)      # equal weights: mean
    assert out[1] == pytest.approx(xx[1])                     # YY flagged: XX
    assert out[2] == pytest.approx(yy[2])                     # XX flagged: YY
    assert out[3] == 0                                        # all flagged: 0
    assert oracle.collapse([1j], [2j], [0], [0], [0.0], [0.0])[0] == 0   # zero weights
