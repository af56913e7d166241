"""FITS codec and the sharedlibrary_gpu command line on the CPU.

Hand-built FITS files (numpy, written card by card here) check the reader
against the standard's layout independently of the library's own writer.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.environ.get("FIP_CLI") or os.path.join(ROOT, "sharedlibrary_gpu")


@pytest.fixture(scope="module")
def F():
    import fastimagingpipe_amd as F
    return F


def _card(key, value=None):
    if value is None:
        s = key
    elif isinstance(value, str):
        s = f"{key:<8}= '{value:<8}'"
    elif isinstance(value, bool):
        s = f"{key:<8}= {'T' if value else 'F':>20}"
    else:
        s = f"{key:<8}= {value!s:>20}"
    return s.ljust(80)[:80]


def _pad(b, fill=b" "):
    return b + fill * ((-len(b)) % 2880)


def _hdu(data, bitpix, primary=True, extname=None, extra=()):
    cards = [_card("SIMPLE", True) if primary else _card("XTENSION", "IMAGE"),
             _card("BITPIX", bitpix), _card("NAXIS", data.ndim)]
    for k, n in enumerate(data.shape[::-1]):
        cards.append(_card(f"NAXIS{k + 1}", n))
    if not primary:
        cards += [_card("PCOUNT", 0), _card("GCOUNT", 1)]
    if primary:
        cards.append(_card("EXTEND", True))
    if extname:
        cards.append(_card("EXTNAME", extname))
    cards += [_card(k, v) for k, v in extra]
    cards.append(_card("END"))
    header = _pad("".join(cards).encode("ascii"))
    dt = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[bitpix]
    return header + _pad(np.ascontiguousarray(data, dtype=dt).tobytes(), b"\0")


def test_write_read_round_trip_2d_3d(F, tmp_path):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((5, 7)).astype(np.float32)
    p = str(tmp_path / "a.fits")
    F.fits_write(p, a)
    assert os.path.getsize(p) % 2880 == 0
    assert np.array_equal(F.fits_read(p).reshape(5, 7), a)
    F.fits_write(p, a[:2])     # overwrites an existing file
    assert np.array_equal(F.fits_read(p).reshape(2, 7), a[:2])
    c = rng.standard_normal((3, 4, 6)).astype(np.float32)
    F.fits_write3(p, c)
    assert np.array_equal(F.fits_read(p), c)


@pytest.mark.parametrize("bitpix", [8, 16, 32, 64, -32, -64])
def test_reads_every_bitpix(F, tmp_path, bitpix):
    rng = np.random.default_rng(bitpix + 100)
    if bitpix > 0:
        hi = {8: 255, 16: 30000, 32: 1 << 20, 64: 1 << 20}[bitpix]
        d = rng.integers(0 if bitpix == 8 else -hi, hi, (6, 9))
    else:
        d = rng.standard_normal((6, 9))
    p = tmp_path / "b.fits"
    p.write_bytes(_hdu(d, bitpix))
    got = F.fits_read(str(p)).reshape(6, 9)
    assert np.array_equal(got, d.astype(np.float32))


def test_bscale_bzero(F, tmp_path):
    d = np.arange(-6, 6, dtype=np.int16).reshape(3, 4)
    p = tmp_path / "s.fits"
    p.write_bytes(_hdu(d, 16, extra=[("BSCALE", 0.5), ("BZERO", 32768.0)]))
    assert np.array_equal(F.fits_read(str(p)).reshape(3, 4), (d * 0.5 + 32768.0).astype(np.float32))


def test_named_extensions(F, tmp_path):
    a = np.arange(9, dtype=np.float32).reshape(1, 3, 3)
    b = np.arange(2 * 5 * 1, dtype=np.float64).reshape(1, 5, 2)
    blob = (_hdu(np.zeros((1,), np.float32), -32) + _hdu(a, -32, False, "TRANSFORMATION MATRICES")
            + _hdu(b, -64, False, "VISIBILITIES"))
    p = tmp_path / "c.fits"
    p.write_bytes(blob)
    assert np.array_equal(F.fits_read(str(p), "TRANSFORMATION MATRICES"), a)
    assert np.array_equal(F.fits_read(str(p), "VISIBILITIES"), b.astype(np.float32))
    with pytest.raises(F.FipError) as ei:
        F.fits_read(str(p), "R-COORDINATES")
    assert ei.value.status == -8


def test_truncated_and_missing_files(F, tmp_path):
    d = np.ones((40, 40), np.float32)
    full = _hdu(d, -32)
    p = tmp_path / "t.fits"
    p.write_bytes(full[:2880 + 100])
    with pytest.raises(F.FipError) as ei:
        F.fits_read(str(p))
    assert ei.value.status == -8
    with pytest.raises(F.FipError):
        F.fits_read(str(tmp_path / "nope.fits"))
    p.write_bytes(b"NOT A FITS FILE".ljust(2880))
    with pytest.raises(F.FipError):
        F.fits_read(str(p))


def _raw_header(cards):
    return _pad("".join([_card(k, v) for k, v in cards] + [_card("END")]).encode("ascii"))


@pytest.mark.parametrize("cards", [
    # negative PCOUNT: |BITPIX|/8 * (PCOUNT + n) would wrap to 0 bytes
    [("BITPIX", -32), ("NAXIS", 1), ("NAXIS1", 1000), ("PCOUNT", -1000)],
    # 2^31 x 2^31 x 4 bytes: the unchecked product wraps
    [("BITPIX", -32), ("NAXIS", 2), ("NAXIS1", 2 ** 31), ("NAXIS2", 2 ** 31)],
    # 2^62 elements: past the element limit
    [("BITPIX", 8), ("NAXIS", 1), ("NAXIS1", 2 ** 62)],
    [("BITPIX", -32), ("NAXIS", 3), ("NAXIS1", 2 ** 21), ("NAXIS2", 2 ** 21), ("NAXIS3", 2 ** 21)],
    [("BITPIX", 16), ("NAXIS", 1), ("NAXIS1", 10), ("GCOUNT", 0)],
    [("BITPIX", 16), ("NAXIS", 1), ("NAXIS1", 10), ("GCOUNT", -5)],
    [("BITPIX", 16), ("NAXIS", 1), ("NAXIS1", 10), ("PCOUNT", "'x'")],
], ids=["neg_pcount", "wrap_2d", "huge_1d", "wrap_3d", "gcount0", "neg_gcount", "bad_pcount"])
def test_crafted_headers_are_rejected(F, tmp_path, cards):
    """Size fields that would overflow or underflow the data size are errors,
    not short reads or heap overruns."""
    p = tmp_path / "c.fits"
    p.write_bytes(_raw_header([("SIMPLE", True)] + cards) + b"\0" * 2880)
    with pytest.raises(F.FipError) as ei:
        F.fits_read(str(p))
    assert ei.value.status == -8


def test_crafted_extension_header_is_rejected(F, tmp_path):
    good = _hdu(np.ones((2, 2), np.float32), -32)
    bad = _raw_header([("XTENSION", "IMAGE"), ("BITPIX", -32), ("NAXIS", 2), ("NAXIS1", 2 ** 33),
                       ("NAXIS2", 2 ** 33), ("PCOUNT", 0), ("GCOUNT", 1), ("EXTNAME", "BIG")])
    p = tmp_path / "e.fits"
    p.write_bytes(good + bad + b"\0" * 2880)
    with pytest.raises(F.FipError) as ei:
        F.fits_read(str(p), "BIG")
    assert ei.value.status == -8


def _run(*args):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=120)


@pytest.fixture(scope="module")
def cli():
    if not os.path.exists(CLI):
        pytest.fail("sharedlibrary_gpu is not built (python -m fastimagingpipe_amd.build)")
    return CLI


def test_cli_usage(cli):
    r = _run()
    assert r.returncode == 1 and "usage" in r.stderr
    r = _run("--combined", "x.fits")
    assert r.returncode == 1
    r = _run("a", "b", "--tau", "x")
    assert r.returncode == 1 and "--tau" in r.stderr
    r = _run("a", "b", "--tau", "-1")
    assert r.returncode == 1 and "--tau" in r.stderr


def _write_inputs(F, d, nbl=16, nsnap=3, naxis3=False):
    F.fits_write(str(d / "vr.fits"), np.zeros((nsnap, nbl), np.float32))
    F.fits_write(str(d / "vi.fits"), np.zeros((nsnap, nbl), np.float32))
    F.fits_write(str(d / "b.fits"), np.zeros((nsnap, 2 * nbl), np.float32))
    if naxis3:
        F.fits_write3(str(d / "v.fits"), np.zeros((nsnap, 3, 3), np.float32))
    else:
        F.fits_write(str(d / "v.fits"), np.tile(np.eye(3, dtype=np.float32).reshape(1, 9), (nsnap, 1)))


def test_cli_rejects_bad_numbers_and_sizes(F, cli, tmp_path):
    _write_inputs(F, tmp_path)
    base = [tmp_path / "vr.fits", tmp_path / "vi.fits", tmp_path / "b.fits", tmp_path / "v.fits"]
    r = _run(*base, "sixty-four", 16, 1e-3, 3, 16, tmp_path / "o.fits")
    assert r.returncode == 1 and "malformed" in r.stderr
    r = _run(*base, 64, 17, 1e-3, 3, 16, tmp_path / "o.fits")      # arrays hold 16 baselines
    assert r.returncode == 1 and "do not hold" in r.stderr
    r = _run(*base, 64, 16, 1e-3, 3, 0, tmp_path / "o.fits")
    assert r.returncode == 1
    r = _run(tmp_path / "missing.fits", *base[1:], 64, 16, 1e-3, 3, 16, tmp_path / "o.fits")
    assert r.returncode == 1 and "ERROR" in r.stderr


def test_cli_rejects_three_axis_inputs(F, cli, tmp_path):
    _write_inputs(F, tmp_path, naxis3=True)
    base = [tmp_path / "vr.fits", tmp_path / "vi.fits", tmp_path / "b.fits", tmp_path / "v.fits"]
    r = _run(*base, 64, 16, 1e-3, 3, 16, tmp_path / "o.fits")
    assert r.returncode == 1 and "naxis" in r.stderr


def test_cli_without_device_fails_cleanly(F, cli, tmp_path):
    if F.device_count() > 0:
        pytest.skip("a device is visible: the GPU CLI test covers this")
    _write_inputs(F, tmp_path)
    base = [tmp_path / "vr.fits", tmp_path / "vi.fits", tmp_path / "b.fits", tmp_path / "v.fits"]
    r = _run(*base, 64, 16, 1e-3, 3, 16, tmp_path / "o.fits")
    assert r.returncode == 1 and "no GPU" in r.stderr
    r = _run(*base, 64, 16, 1.4e8, 1e-3, 3, 16, tmp_path / "o.fits")   # README form
    assert r.returncode == 1 and "no GPU" in r.stderr
