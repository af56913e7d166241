"""ctypes binding of the FP64 CPU oracle (oracle/fip_oracle.c).

Test infrastructure: only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this module, and only to check or
time against it.  The product (``fastimagingpipe_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FIP_ORACLE_LIB names another build of fip_oracle.c (the mutation tests load
# deliberately broken copies through it); by default the in-tree library.
_LIB_OVERRIDE = os.environ.get("FIP_ORACLE_LIB")
_LIB_PATH = _LIB_OVERRIDE or os.path.join(_HERE, "libfip_oracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_sz = ctypes.c_size_t


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if _LIB_OVERRIDE:
        return _LIB_PATH
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []),
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        L.fo_es.argtypes = [ctypes.c_double]
        L.fo_es.restype = ctypes.c_double
        L.fo_quadrature_tables.argtypes = [_f64p, _f64p, _f64p]
        L.fo_grid_size.argtypes = [_sz]
        L.fo_grid_size.restype = _sz
        L.fo_conv_corr.argtypes = [_sz, _sz, _f64p]
        L.fo_fft.argtypes = [_f64p, _sz, ctypes.c_int]
        L.fo_fft.restype = ctypes.c_int
        L.fo_fft2_inverse.argtypes = [_f64p, _sz, ctypes.c_int]
        L.fo_fft2_inverse.restype = ctypes.c_int
        L.fo_grid.argtypes = [_f32p, _f32p, _f32p, _f32p, _sz, _sz, ctypes.c_float, _f64p]
        L.fo_crop_correct.argtypes = [_f64p, _f64p, _sz, _sz, _f64p]
        L.fo_coordschange.argtypes = [_f32p, _sz, _f64p]
        L.fo_p2p.argtypes = [_f64p, _f32p, ctypes.c_float, _sz]
        L.fo_splat.argtypes = [_f64p, _f64p, _sz, _sz, _f64p]
        L.fo_max.argtypes = [_f64p, _sz]
        L.fo_max.restype = ctypes.c_double
        L.fo_trigger.argtypes = [_f64p, _f64p, _f64p, _f64p, _sz, _sz, _f64p]
        L.fo_image.argtypes = [_f32p, _f32p, _f32p, _f32p, _sz, _sz, ctypes.c_float,
                               ctypes.c_void_p, _f64p, ctypes.POINTER(ctypes.c_double),
                               ctypes.c_int]
        L.fo_image.restype = ctypes.c_int
        L.fo_pipe.argtypes = [_f32p, _f32p, _f32p, _f32p, _f64p, _sz, _sz, _sz,
                              ctypes.c_float, _sz, ctypes.c_int]
        L.fo_pipe.restype = ctypes.c_int
        L.fo_project.argtypes = [_f64p, _sz, _f64p, _sz, _f64p, _f64p]
        L.fo_project.restype = ctypes.c_int
        L.fo_collapse.argtypes = [_f64p, _f64p, _f64p, _f64p, _u8p, _u8p, _f64p, _f64p,
                                  _sz, _f64p, _f64p]
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def es(x: float) -> float:
    return lib().fo_es(float(x))


def quadrature_tables():
    n, w, k = (np.zeros(14) for _ in range(3))
    lib().fo_quadrature_tables(n, w, k)
    return n, w, k


def grid_size(image: int) -> int:
    return int(lib().fo_grid_size(image))


def conv_corr(image: int, grid: int | None = None) -> np.ndarray:
    g = grid_size(image) if grid is None else grid
    c = np.zeros(image // 2 + 1)
    lib().fo_conv_corr(image, g, c)
    return c


def fft(x: np.ndarray, sign: int = 1) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64).copy()
    rc = lib().fo_fft(a, len(x), sign)
    assert rc == 0
    return a.view(np.complex128)


def fft2_inverse(grid: np.ndarray, nthreads: int = 0) -> np.ndarray:
    g = grid.shape[0]
    a = np.ascontiguousarray(grid, dtype=np.complex128).view(np.float64).reshape(-1).copy()
    rc = lib().fo_fft2_inverse(a, g, nthreads)
    assert rc == 0
    return a.view(np.complex128).reshape(g, g)


def grid(vis_re, vis_im, bin_, vin, g: int, cell: float) -> np.ndarray:
    nbl = len(vis_re)
    out = np.zeros(2 * g * g)
    lib().fo_grid(_f32(vis_re), _f32(vis_im), _f32(bin_).reshape(-1), _f32(vin).reshape(-1),
                  nbl, g, cell, out)
    return out.view(np.complex128).reshape(g, g)


def crop_correct(fgrid: np.ndarray, image: int) -> np.ndarray:
    g = fgrid.shape[0]
    c = conv_corr(image, g)
    out = np.zeros(image * image)
    lib().fo_crop_correct(np.ascontiguousarray(fgrid).view(np.float64).reshape(-1), c,
                          image, g, out)
    return out.reshape(image, image)


def coord_map(vin, image: int, cell: float, sin_projection: bool = True) -> np.ndarray:
    """coordschange (+ p2p): returns out_index[i, j, 2] (i = column, j = row)."""
    idx = np.zeros(2 * image * image)
    v = _f32(vin).reshape(-1)
    lib().fo_coordschange(v, image, idx)
    if sin_projection:
        lib().fo_p2p(idx, v, cell, image)
    return idx.reshape(image, image, 2)


def splat(out_index: np.ndarray, dirty_pre: np.ndarray, nbl: int) -> np.ndarray:
    image = dirty_pre.shape[0]
    out = np.zeros(image * image)
    lib().fo_splat(_f64(out_index).reshape(-1), _f64(dirty_pre).reshape(-1), image, nbl, out)
    return out.reshape(image, image)


def image_max(img: np.ndarray) -> float:
    a = _f64(img).reshape(-1)
    return float(lib().fo_max(a, a.size))


def trigger(d1, d2, d3, maxall, tile: int) -> np.ndarray:
    image = d1.shape[0]
    un = image // tile
    out = np.zeros(un * un)
    lib().fo_trigger(_f64(d1).reshape(-1), _f64(d2).reshape(-1), _f64(d3).reshape(-1),
                     _f64(maxall), image, tile, out)
    return out.reshape(un, un)


def image(vis_re, vis_im, bin_, vin, image: int, cell: float, nthreads: int = 0,
          want_pre: bool = False):
    """One snapshot through the whole imaging stage.

    Returns (dirty, max) or (dirty_pre, dirty, max) when ``want_pre``.
    """
    nbl = len(vis_re)
    pre = np.zeros(image * image) if want_pre else None
    dirty = np.zeros(image * image)
    mx = ctypes.c_double(0.0)
    rc = lib().fo_image(_f32(vis_re), _f32(vis_im), _f32(bin_).reshape(-1),
                        _f32(vin).reshape(-1), nbl, image, cell,
                        pre.ctypes.data if want_pre else None, dirty, ctypes.byref(mx),
                        nthreads)
    if rc != 0:
        raise RuntimeError(f"fo_image failed: {rc}")
    if want_pre:
        return pre.reshape(image, image), dirty.reshape(image, image), mx.value
    return dirty.reshape(image, image), mx.value


def pipe(vis_re, vis_im, bin_, vin, nbl: int, image: int, nsnap: int, cell: float,
         tile: int, nthreads: int = 0) -> np.ndarray:
    """FIpipe in FP64: the first three snapshots -> tLISI [I/T, I/T]."""
    un = image // tile
    out = np.zeros(un * un)
    rc = lib().fo_pipe(_f32(vis_re).reshape(-1), _f32(vis_im).reshape(-1),
                       _f32(bin_).reshape(-1), _f32(vin).reshape(-1), out, nbl, image,
                       nsnap, cell, tile, nthreads)
    if rc != 0:
        raise RuntimeError(f"fo_pipe failed: {rc}")
    return out.reshape(un, un)


def project(uvw, freqs):
    """SVD_MFS.py projection: returns (bin [nsamp,2], vin [3,3]) in float64."""
    uvw = _f64(uvw).reshape(-1, 3)
    freqs = _f64(np.atleast_1d(freqs))
    ns = uvw.shape[0] * freqs.size
    b = np.zeros(2 * ns)
    v = np.zeros(9)
    rc = lib().fo_project(uvw.reshape(-1), uvw.shape[0], freqs, freqs.size, b, v)
    assert rc == 0
    return b.reshape(ns, 2), v.reshape(3, 3)


def collapse(xx, yy, flag0, flag3, w0, w3):
    xx = np.asarray(xx, dtype=np.complex128)
    yy = np.asarray(yy, dtype=np.complex128)
    n = xx.size
    ore, oim = np.zeros(n), np.zeros(n)
    lib().fo_collapse(_f64(xx.real), _f64(xx.imag), _f64(yy.real), _f64(yy.imag),
                      np.ascontiguousarray(flag0, dtype=np.uint8),
                      np.ascontiguousarray(flag3, dtype=np.uint8), _f64(w0), _f64(w3),
                      n, ore, oim)
    return ore + 1j * oim
