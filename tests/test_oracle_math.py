"""Pins the FP64 oracle's building blocks: ES kernel and quadrature tables,
grid-correction function, FFT (CPU only)."""
import json
import os

import numpy as np
import pytest
from scipy.integrate import quad

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ref_tables():
    with open(os.path.join(GOLD, "reference_quadrature.json")) as f:
        return json.load(f)


def test_quadrature_tables_match_reference_literals(oracle, ref_tables):
    n, w, k = oracle.quadrature_tables()
    # The reference tabulates 8-9 significant digits (FIPkernels.cu:14-29).
    assert np.allclose(n, ref_tables["nodes"], rtol=0, atol=2e-8)
    assert np.allclose(w, ref_tables["weights"], rtol=0, atol=2e-8)
    assert np.allclose(k, ref_tables["kernel"], rtol=5e-8, atol=0)


def test_quadrature_is_gauss_legendre_28(oracle):
    n, w, _ = oracle.quadrature_tables()
    xg, wg = np.polynomial.legendre.leggauss(28)
    pos = xg > 0
    assert np.allclose(np.sort(n), np.sort(xg[pos]), atol=1e-14)
    assert np.allclose(w[np.argsort(n)], wg[pos][np.argsort(xg[pos])], atol=1e-14)


def test_es_kernel_closed_form(oracle, ref_tables):
    beta = ref_tables["beta"]
    for x in [0.0, 0.3, -0.7, 0.999, 1.0]:
        assert oracle.es(x) == pytest.approx(np.exp(beta * (np.sqrt(1 - x * x) - 1)), rel=1e-15)
    assert oracle.es(1.0000001) == 0.0
    assert oracle.es(-2.0) == 0.0


def test_norm_constant_is_four_times_es_integral(oracle, ref_tables):
    integral = quad(oracle.es, -1, 1, epsabs=1e-15, epsrel=1e-15, limit=400)[0]
    assert 4 * integral == pytest.approx(ref_tables["norm"], rel=3e-10)
    # c(0) = 8 sum w ES / norm = 1 (FIPkernels.cu:173-186 at t = 0)
    c = oracle.conv_corr(256)
    assert c[0] == pytest.approx(1.0, abs=1e-9)


@pytest.mark.parametrize("image", [64, 256, 4096])
def test_conv_corr_is_cosine_transform_of_es(oracle, ref_tables, image):
    g = oracle.grid_size(image)
    c = oracle.conv_corr(image)
    norm = ref_tables["norm"]
    for t in [0, 1, image // 7, image // 2]:
        ref = 4 * quad(lambda x: oracle.es(x) * np.cos(8 * np.pi * t * x / g), -1, 1,
                       epsabs=1e-14, limit=400)[0] / norm
        assert c[t] == pytest.approx(ref, abs=2e-7)


@pytest.mark.parametrize("image,g", [(64, 96), (256, 384), (2048, 3072), (4096, 6144), (8192, 12288), (100, 150)])
def test_grid_size(oracle, image, g):
    assert oracle.grid_size(image) == g


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 48, 96, 100, 384, 1536, 6144])
def test_fft_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(oracle.fft(x, 1), np.fft.ifft(x) * n, atol=1e-12 * max(1, n))
    assert np.allclose(oracle.fft(x, -1), np.fft.fft(x), atol=1e-12 * max(1, n))


@pytest.mark.parametrize("n", [3, 8, 12, 15])
def test_fft_matches_brute_force_dft(oracle, n):
    rng = np.random.default_rng(100 + n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    k = np.arange(n)
    dft = np.exp(2j * np.pi * np.outer(k, k) / n) @ x
    assert np.allclose(oracle.fft(x, 1), dft, atol=1e-12)


@pytest.mark.parametrize("g", [6, 96, 384])
def test_fft2_inverse_matches_numpy(oracle, g):
    rng = np.random.default_rng(g)
    G = rng.standard_normal((g, g)) + 1j * rng.standard_normal((g, g))
    assert np.allclose(oracle.fft2_inverse(G), np.fft.ifft2(G) * g * g, atol=1e-10 * g)
