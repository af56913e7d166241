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


# ----------------------------------------------------------------------------
# Geometry known answers: where a point source of the sky lands.
#
# A flat array seen from a tilted uvw frame has coplanar baselines D, so the
# PCA normal V2 = V0 x V1 is the plane's normal and D . V2 = 0.  A source at
# direction cosines (l, m) has visibilities exp(-2 pi i D . s'), s' =
# (l, m, n - 1), and D . s' = B0 (V0 . s') + B1 (V1 . s').  The gridding
# phase e^{+2 pi i (B0 cell y + B1 cell x)} therefore peaks at row
# y = V0 . s' / cell, column x = V1 . s' / cell of the pre-image, whatever V.
# Re-projection undoes V: with V orthonormal, the cofactors of V22 give
# -V00 (V1 . s') + V10 (V0 . s') = -V22 m + V21 (n - 1) and
# -V01 (V1 . s') + V11 (V0 . s') = V22 l - V20 (n - 1), so after the division
# by |V22| the source sits at column -sign(V22) m / cell + h and row
# sign(V22) l / cell + h, up to the w terms (n - 1)/cell, which are below
# 0.2 px for the fields used here.  A V with V22 < 0 (the PCA's sign rule
# decides it) gives the point-mirrored sky.

C_LIGHT = 299792458.0


def _coplanar_array(rng, nant, cell, tilt):
    """Baselines (wavelengths) of a flat array in a plane tilted by `tilt`."""
    # antenna radius: longest baseline |B| cell <= 0.5, inside the half grid
    rad = 0.25 / cell
    r = rad * np.sqrt(rng.uniform(0, 1, nant))
    a = rng.uniform(0, 2 * np.pi, nant)
    pos = np.stack([r * np.cos(a), r * np.sin(a), np.zeros(nant)], axis=1)
    i, j = np.triu_indices(nant, 1)
    R = rotation(tilt, rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi))
    return (pos[i] - pos[j]) @ R


def _peak(img):
    """Sub-pixel peak: argmax refined by the centroid of its 3x3 neighbourhood."""
    r, c = np.unravel_index(np.argmax(img), img.shape)
    w = img[r - 1:r + 2, c - 1:c + 2]
    yy, xx = np.mgrid[-1:2, -1:2]
    return r + (w * yy).sum() / w.sum(), c + (w * xx).sum() / w.sum()


GEOM_SEEDS = [11, 12, 13, 14, 15, 16, 17, 18]


def _geometry_case(oracle, seed, image=256, cell=5.8e-5, nant=64):
    rng = np.random.default_rng(seed)
    tilt = rng.uniform(0.35, 0.5)
    uvw = _coplanar_array(rng, nant, cell, tilt)
    # 60..100 px from the centre: a 1/|V22| error moves it by >= 3 px
    rad = rng.uniform(60, 100) * cell
    ang = rng.uniform(0, 2 * np.pi)
    l, m = rad * np.cos(ang), rad * np.sin(ang)
    n = np.sqrt(1 - l * l - m * m)
    vis = np.exp(-2j * np.pi * (uvw[:, 0] * l + uvw[:, 1] * m + uvw[:, 2] * (n - 1)))
    B, V = oracle.project(uvw, np.array([C_LIGHT]))   # D = uvw (already wavelengths)
    pre, dirty, _ = oracle.image(vis.real.astype(np.float32), vis.imag.astype(np.float32),
                                 B.astype(np.float32), V.astype(np.float32), image, cell,
                                 want_pre=True)
    return dict(V=V, l=l, m=m, n=n, pre=pre, dirty=dirty, tilt=tilt)


def _source_on_pixel(V, y, x, cell):
    """(l, m, n) with V0 . s' = y cell and V1 . s' = x cell, s' = (l, m, n - 1)."""
    l = m = 0.0
    for _ in range(50):
        n1 = np.sqrt(1 - l * l - m * m) - 1
        A = V[:2, :2]
        rhs = np.array([y * cell - V[0, 2] * n1, x * cell - V[1, 2] * n1])
        l, m = np.linalg.solve(A, rhs)
    return l, m, np.sqrt(1 - l * l - m * m)


@pytest.mark.parametrize("seed", GEOM_SEEDS[:4])
def test_point_source_on_pixel_known_amplitude(oracle, seed):
    """A unit source placed so that its PCA-frame position falls on pre-image
    pixel (y, x): every baseline's phase there is exactly zero (D . V2 = 0),
    so the corrected pre-image holds N / J, J = |V00 V11 - V01 V10| = |V22|,
    at that pixel (FIPkernels.cu:188-287; B un-centred, SVD_MFS.py:105-110)."""
    image, cell, nant = 256, 5.8e-5, 64
    h = image // 2
    rng = np.random.default_rng(seed)
    uvw = _coplanar_array(rng, nant, cell, rng.uniform(0.35, 0.5))
    B, V = oracle.project(uvw, np.array([C_LIGHT]))
    y, x = (int(v) for v in rng.integers(-90, 90, 2))
    l, m, n = _source_on_pixel(V, y, x, cell)
    vis = np.exp(-2j * np.pi * (uvw[:, 0] * l + uvw[:, 1] * m + uvw[:, 2] * (n - 1)))
    pre, _, _ = oracle.image(vis.real.astype(np.float32), vis.imag.astype(np.float32),
                             B.astype(np.float32), V.astype(np.float32), image, cell, want_pre=True)
    N = uvw.shape[0]
    J = abs(V[0, 0] * V[1, 1] - V[0, 1] * V[1, 0])
    assert J == pytest.approx(abs(V[2, 2]), rel=1e-12)
    assert np.unravel_index(np.argmax(pre), pre.shape) == (y + h, x + h)
    # float32 inputs (B, vis) and the gridding kernel's correction error
    assert pre[y + h, x + h] == pytest.approx(N / J, rel=2e-4)


def test_point_source_geometry_cases_cover_both_v22_signs(oracle):
    signs = {np.sign(_geometry_case(oracle, s)["V"][2, 2]) for s in GEOM_SEEDS}
    assert signs == {-1.0, 1.0}


@pytest.mark.parametrize("seed", GEOM_SEEDS)
def test_point_source_pre_image_peak_in_pca_frame(oracle, seed):
    """fo_project + fo_grid + FFT + crop: the pre-image peak is at
    (V0 . s', V1 . s') / cell + h for a random tilted frame."""
    c = _geometry_case(oracle, seed)
    image, cell = 256, 5.8e-5
    h = image // 2
    s = np.array([c["l"], c["m"], c["n"] - 1])
    V = c["V"].astype(np.float64)
    r, col = _peak(c["pre"])
    assert abs(r - (V[0] @ s / cell + h)) < 0.5
    assert abs(col - (V[1] @ s / cell + h)) < 0.5


@pytest.mark.parametrize("seed", GEOM_SEEDS)
def test_point_source_final_peak_on_sky_position(oracle, seed):
    """fo_coordschange + fo_p2p + fo_splat: the re-projected peak sits at row
    sign(V22) l / cell + h, column -sign(V22) m / cell + h for every tilted V
    (FIPkernels.cu:289-299, 378-413).  A swapped V01/V10, a dropped / |V22|
    or a signed V22 in the divisor moves it by >= 3 px."""
    c = _geometry_case(oracle, seed)
    image, cell = 256, 5.8e-5
    h = image // 2
    sg = np.sign(c["V"][2, 2])
    r, col = _peak(c["dirty"])
    assert abs(r - (sg * c["l"] / cell + h)) < 0.6, (r, sg * c["l"] / cell + h)
    assert abs(col - (-sg * c["m"] / cell + h)) < 0.6, (col, -sg * c["m"] / cell + h)


def _splat_one(oracle, LL, MM, image=64, nbl=4, val=3.0, src=(5, 9)):
    """fo_splat of one source (column i, row j = src) placed at (LL, MM)."""
    h = image // 2
    idx = np.full((image, image, 2), 1e6)          # every other source outside
    idx[src[0], src[1]] = (LL + h, MM + h)
    pre = np.zeros((image, image))
    pre[src[1], src[0]] = val                      # pre[row j, column i]
    return oracle.splat(idx, pre, nbl), h, val / nbl


@pytest.mark.parametrize("LL,MM", [(10.25, -7.625), (-3.625, 4.25), (0.75, 0.125), (-20.5, -0.875)])
def test_splat_subpixel_bilinear_corners(oracle, LL, MM):
    """finalinterp (FIPkernels.cu:385-410): a source at fractional (LL, MM)
    puts (1-a)(1-b), (1-a)b, a(1-b), ab of val/N on (fM,fL), (cM,fL),
    (fM,cL), (cM,cL) with a = {LL}, b = {MM}; rows follow MM, columns LL.
    Equivalent invariants: the mass is val/N, the centroid is exactly
    (row MM + h, column LL + h), and the 2 x 2 footprint is rank one."""
    out, h, v = _splat_one(oracle, LL, MM)
    a, b = LL - np.floor(LL), MM - np.floor(MM)
    fL, fM = int(np.floor(LL)) + h, int(np.floor(MM)) + h
    blk = out[fM:fM + 2, fL:fL + 2]
    assert out.sum() == pytest.approx(v, rel=1e-14)
    assert blk.sum() == pytest.approx(v, rel=1e-14)          # nothing outside the 2 x 2 block
    rows, cols = np.mgrid[0:out.shape[0], 0:out.shape[1]]
    assert (out * rows).sum() / v == pytest.approx(MM + h, abs=1e-12)
    assert (out * cols).sum() / v == pytest.approx(LL + h, abs=1e-12)
    assert abs(np.linalg.det(blk)) < 1e-14 * v * v
    expect = v * np.array([[(1 - a) * (1 - b), a * (1 - b)],
                           [(1 - a) * b, a * b]])
    assert np.abs(blk - expect).max() < 1e-15


def test_splat_first_moments_are_linear(oracle):
    """Many sources at random sub-pixel positions: the output's first moments
    equal the value-weighted mean target position (bilinear splatting keeps
    the centroid of every source)."""
    image, nbl = 64, 2
    rng = np.random.default_rng(9)
    idx = rng.uniform(6, image - 6, (image, image, 2))
    pre = rng.uniform(0, 1, (image, image))
    out = oracle.splat(idx, pre, nbl)
    w = pre.T / nbl                                # source (i, j) carries pre[j, i]
    rows, cols = np.mgrid[0:image, 0:image]
    assert out.sum() == pytest.approx(w.sum(), rel=1e-13)
    assert (out * cols).sum() == pytest.approx((w * idx[..., 0]).sum(), rel=1e-13)
    assert (out * rows).sum() == pytest.approx((w * idx[..., 1]).sum(), rel=1e-13)


@pytest.mark.parametrize("image,nbl,cell", [(128, 900, 1e-4), (256, 2016, 5.8e-5)])
def test_gpu_image_cases_resolve_the_splat_group_curvature(oracle, image, nbl, cell):
    """The splat places the 8 sources of a row group (columns i0..i0+7) from a
    quadratic model in k = i - i0 of the SIN displacement D(i) = SIN position -
    affine position.  Its k^2 term is at most ~5e-4 px at these sizes.  Shifting
    every source by minus that term (D''/2 k^2, from second differences of the
    oracle's own map) moves the oracle image of each snapshot of
    test_image_matches_oracle by more than ten times that test's 1e-5 rel-rms
    tolerance (1.4e-4 to 4.1e-4), so the GPU test would catch a dropped or wrong
    k^2 term.  Those snapshots take the series form at 128^2 and 256^2 (model
    remainder bound 1.2e-8 to 4.4e-8 px, within the 5e-8 px limit) and the
    closed form per source at 64^2 (bound ~1.9e-7 px)."""
    from tests.inputs import rel_rms, snapshots

    vr, vi, B, V = snapshots(11 + image, 3, nbl, image, cell)
    k = (np.arange(image) % 8)[:, None, None]
    for s in range(3):
        pre, _, _ = oracle.image(vr[s], vi[s], B[s], V[s], image, cell, want_pre=True)
        sin = oracle.coord_map(V[s], image, cell, sin_projection=True)
        disp = sin - oracle.coord_map(V[s], image, cell, sin_projection=False)
        o = disp[0::8]                                   # [group, j, 2] at i0, i0+1, i0+2
        half_d2 = 0.5 * (disp[2::8] - 2.0 * disp[1::8] + o)
        shifted = sin - np.repeat(half_d2, 8, axis=0) * k ** 2
        base = oracle.splat(sin, pre, nbl)
        moved = rel_rms(oracle.splat(shifted, pre, nbl), base)
        assert moved > 1e-4, (s, moved)
