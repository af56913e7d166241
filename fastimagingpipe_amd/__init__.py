"""MI355X-native FIP-TOI imaging-and-trigger path.

Python mirror of the reference's interface (``FIpipe``, FIPkernels.cu:554-557)
over the C ABI in ``include/fip.h``, implemented by the in-tree HIP library
``libfip.so``.  There is no CPU fallback: every call goes through the HIP
library, and importing the package raises if the library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

__all__ = [
    "FipError", "lib", "library_path", "grid_size", "fipipe", "fipipe_lowp", "image",
    "trigger", "pipe_device", "pipe_device_noalloc", "pipe_raw_device", "batch_device", "batch_host", "project", "collapse", "ingest_table",
    "synth",
    "set_profiling", "stage_times", "reset_stage_times", "autotune", "set_tuning",
    "debug_fft", "fits_read", "fits_write", "fits_write3", "Server", "STATUS",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
# FIP_LIBRARY names another build of the same sources (the host-sanitizer
# build of tests/sanitize); by default the in-tree library.
library_path = os.environ.get("FIP_LIBRARY") or os.path.join(_HERE, "libfip.so")

STATUS = {
    0: "FIP_OK", -1: "FIP_ERR_ARG", -2: "FIP_ERR_NSNAP", -3: "FIP_ERR_SIZE",
    -4: "FIP_ERR_ALLOC", -5: "FIP_ERR_HIP", -6: "FIP_ERR_NODEVICE", -7: "FIP_ERR_BUDGET",
    -8: "FIP_ERR_IO", -9: "FIP_ERR_DEGENERATE",
}
VIS_F32, VIS_BF16, VIS_F16 = 0, 1, 2


class FipError(RuntimeError):
    """A non-zero fip_status returned through the C ABI."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


if not os.path.exists(library_path):
    raise ImportError(
        f"{library_path} is missing: build it with `python -m fastimagingpipe_amd.build` "
        "(hipcc, --offload-arch=gfx950); there is no CPU fallback")

lib = ctypes.CDLL(library_path)

_sz = ctypes.c_size_t
_vp = ctypes.c_void_p
_f = ctypes.c_float
_i = ctypes.c_int
_fp = ctypes.POINTER(ctypes.c_float)


def _sig(name, argtypes, restype=ctypes.c_int):
    fn = getattr(lib, name)
    fn.argtypes = argtypes
    fn.restype = restype
    return fn


_sig("fip_version", [], ctypes.c_char_p)
_sig("fip_last_error", [], ctypes.c_char_p)
_sig("fip_status_string", [_i], ctypes.c_char_p)
_sig("fip_device_count", [])
_sig("fip_grid_size", [_sz], _sz)
_sig("fip_pipe", [_vp, _vp, _vp, _vp, _vp, _sz, _sz, _sz, _f, _sz])
_sig("fip_pipe_device", [_vp, _vp, _vp, _vp, _vp, _sz, _sz, _sz, _f, _sz, _vp])
_sig("fip_pipe_device_noalloc", [_vp, _vp, _vp, _vp, _vp, _sz, _sz, _sz, _f, _sz, _vp])
_sig("fip_pipe_lowp", [_vp, _vp, _i, _vp, _vp, _vp, _sz, _sz, _sz, _f, _sz])
_sig("fip_image", [_vp, _vp, _vp, _vp, _sz, _sz, _sz, _f, _vp, _vp, _vp])
_sig("fip_trigger", [_vp, _vp, _vp, _vp, _sz, _sz, _vp])
_sig("fip_batch_device", [_vp, _vp, _i, _vp, _vp, _vp, _sz, _sz, _sz, _sz, _f, _sz, _vp, _vp])
_sig("fip_project", [_vp, _sz, _vp, _sz, _vp, _vp])
_sig("fip_collapse", [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp])
_sig("fip_ingest_table", [_vp, _vp, _vp, _sz, _vp, _vp, _vp, _sz, _sz, _vp, _sz, _i, _i, ctypes.c_char_p,
                          ctypes.c_char_p, _vp, _vp])
_sig("fip_set_profiling", [_i])
_sig("fip_stage_count", [])
_sig("fip_stage_name", [_i], ctypes.c_char_p)
_sig("fip_stage_times", [_vp, _vp, _i])
_sig("fip_stage_times_reset", [])
_sig("fip_autotune", [_sz, _sz, _f, _i, _vp])
_sig("fip_set_tuning", [_sz, _vp])
_sig("fip_debug_fft", [_i, _vp, _i])
_sig("fip_synth", [_i, ctypes.c_double, _sz, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                   ctypes.c_double, _sz, _f, _i, ctypes.c_double, ctypes.c_double,
                   ctypes.c_double, ctypes.c_double, _i, _vp, ctypes.c_double,
                   ctypes.c_double, ctypes.c_ulonglong, _vp, _vp, _vp, _vp, _vp])
_sig("fip_fits_read", [ctypes.c_char_p, ctypes.POINTER(_fp), ctypes.POINTER(ctypes.c_long)])
_sig("fip_fits_read_ext", [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_fp),
                           ctypes.POINTER(ctypes.c_long)])
_sig("fip_fits_write", [ctypes.c_char_p, _vp, ctypes.c_long, ctypes.c_long])
_sig("fip_fits_write3", [ctypes.c_char_p, _vp, ctypes.c_long, ctypes.c_long, ctypes.c_long])
_sig("fip_batch_host", [_vp, _vp, _i, _vp, _vp, _vp, _sz, _sz, _sz, _sz, _f, _sz, _vp])
_sig("fip_free", [_vp], None)


class _RawView(ctypes.Structure):
    _fields_ = [("uvw", _vp), ("freqs", _vp), ("xx_re", _vp), ("xx_im", _vp), ("yy_re", _vp),
                ("yy_im", _vp), ("flag0", _vp), ("flag3", _vp), ("w0", _vp), ("w3", _vp),
                ("nrows", _sz), ("nchan", _sz), ("nsnap", _sz)]


_sig("fip_pipe_raw_device", [ctypes.POINTER(_RawView), _sz, _f, _sz, _vp, _vp, _vp, _vp])


def pipe_raw_device(ptrs: dict, nrows: int, nchan: int, nsnap: int, image_size: int, cell_size: float,
                    unit_size: int, out_ptr: int, bin_out_ptr: int = 0, vin_out_ptr: int = 0,
                    stream: int = 0) -> None:
    """fip_pipe_raw_device: ``ptrs`` maps uvw, freqs, xx_re, xx_im, yy_re, yy_im, flag0,
    flag3, w0, w3 to device addresses."""
    rv = _RawView(**{k: ptrs[k] for k in ("uvw", "freqs", "xx_re", "xx_im", "yy_re", "yy_im", "flag0",
                                           "flag3", "w0", "w3")}, nrows=nrows, nchan=nchan, nsnap=nsnap)
    _check(lib.fip_pipe_raw_device(ctypes.byref(rv), image_size, cell_size, unit_size, out_ptr,
                                   bin_out_ptr or None, vin_out_ptr or None, stream or None))


class _StreamConfig(ctypes.Structure):
    _fields_ = [("num_baselines", _sz), ("image_size", _sz), ("cell_size", _f),
                ("unit_size", _sz), ("vis_type", _i)]


_sig("fip_server_create", [ctypes.POINTER(_StreamConfig), _i, _i, _sz, ctypes.POINTER(_vp)])
_sig("fip_server_submit", [_vp, _i, _vp, _vp, _vp, _vp, ctypes.c_double])
_sig("fip_server_drain", [_vp])
_sig("fip_server_stats", [_vp, _vp, _vp, _vp, _vp, _vp])
_sig("fip_server_latest", [_vp, _i, _vp, _vp])
_sig("fip_server_sequences", [_vp, _vp])
_sig("fip_server_destroy", [_vp], None)


def _check(rc: int) -> int:
    if rc != 0:
        msg = lib.fip_last_error()
        raise FipError(rc, msg.decode() if msg else "")
    return rc


def _c(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return a.ctypes.data_as(_vp)


def version() -> str:
    return lib.fip_version().decode()


def device_count() -> int:
    return int(lib.fip_device_count())


def grid_size(image_size: int) -> int:
    """ceil(1.5 * image) (FIPkernels.cu:583); 0 for unsupported sizes."""
    return int(lib.fip_grid_size(image_size))


def fipipe(vis_re, vis_im, bin_, vin, num_baselines: int, image_size: int,
           num_snapshots: int, cell_size: float, unit_size: int) -> np.ndarray:
    """FIpipe (FIPkernels.cu:554-862): host arrays in, tLISI [I/T, I/T] out."""
    vr, vi, b, v = _c(vis_re), _c(vis_im), _c(bin_), _c(vin)
    un = image_size // unit_size if unit_size else 0
    out = np.zeros((un, un), dtype=np.float32)
    _check(lib.fip_pipe(_ptr(vr), _ptr(vi), _ptr(b), _ptr(v), _ptr(out), num_baselines,
                        image_size, num_snapshots, cell_size, unit_size))
    return out


def fipipe_lowp(vis_re, vis_im, vis_type: int, bin_, vin, num_baselines: int, image_size: int,
                num_snapshots: int, cell_size: float, unit_size: int) -> np.ndarray:
    """Precision form: vis_re/vis_im given as uint16 bit patterns of bf16 or fp16."""
    vr = np.ascontiguousarray(vis_re).view(np.uint16)
    vi = np.ascontiguousarray(vis_im).view(np.uint16)
    b, v = _c(bin_), _c(vin)
    un = image_size // unit_size
    out = np.zeros((un, un), dtype=np.float32)
    _check(lib.fip_pipe_lowp(_ptr(vr), _ptr(vi), vis_type, _ptr(b), _ptr(v), _ptr(out),
                             num_baselines, image_size, num_snapshots, cell_size, unit_size))
    return out


def image(vis_re, vis_im, bin_, vin, num_baselines: int, image_size: int, num_snapshots: int,
          cell_size: float):
    """Imaging half: returns (pre [S,I,I], img [S,I,I], max [S])."""
    vr, vi, b, v = _c(vis_re), _c(vis_im), _c(bin_), _c(vin)
    S, I = num_snapshots, image_size
    pre = np.zeros((S, I, I), dtype=np.float32)
    img = np.zeros((S, I, I), dtype=np.float32)
    mx = np.zeros(S, dtype=np.float32)
    _check(lib.fip_image(_ptr(vr), _ptr(vi), _ptr(b), _ptr(v), num_baselines, I, S, cell_size,
                         _ptr(pre), _ptr(img), _ptr(mx)))
    return pre, img, mx


def trigger(d1, d2, d3, maxall, unit_size: int, allow_degenerate: bool = False) -> np.ndarray:
    """Trigger half (FIPkernels.cu:806-851): tLISI of three images."""
    a, b, c = _c(d1), _c(d2), _c(d3)
    m = _c(maxall)
    I = a.shape[0]
    un = I // unit_size
    out = np.zeros((un, un), dtype=np.float32)
    rc = lib.fip_trigger(_ptr(a), _ptr(b), _ptr(c), _ptr(m), I, unit_size, _ptr(out))
    if not (allow_degenerate and rc == -9):
        _check(rc)
    return out


def pipe_device(vr_ptr: int, vi_ptr: int, b_ptr: int, v_ptr: int, out_ptr: int,
                num_baselines: int, image_size: int, num_snapshots: int, cell_size: float,
                unit_size: int, stream: int = 0) -> None:
    """fip_pipe_device on raw device pointers (e.g. torch ``data_ptr()``)."""
    _check(lib.fip_pipe_device(vr_ptr, vi_ptr, b_ptr, v_ptr, out_ptr, num_baselines, image_size,
                               num_snapshots, cell_size, unit_size, stream or None))


def pipe_device_noalloc(vr_ptr: int, vi_ptr: int, b_ptr: int, v_ptr: int, out_ptr: int,
                        num_baselines: int, image_size: int, num_snapshots: int, cell_size: float,
                        unit_size: int, stream: int = 0) -> None:
    """fip_pipe_device_noalloc: as pipe_device, with every intermediate in the
    library's static device array (no allocation; image sizes up to 1024)."""
    _check(lib.fip_pipe_device_noalloc(vr_ptr, vi_ptr, b_ptr, v_ptr, out_ptr, num_baselines, image_size,
                                       num_snapshots, cell_size, unit_size, stream or None))


def batch_device(vr_ptr: int, vi_ptr: int, vis_type: int, b_ptr: int, v_ptr: int, nvis,
                 stride: int, nsnap: int, norm_n: int, image_size: int, cell_size: float,
                 unit_size: int, out_ptr: int, stream: int = 0) -> None:
    """fip_batch_device: images nsnap device-resident snapshots and triggers every window."""
    nv = _c(nvis, np.int32)
    _check(lib.fip_batch_device(vr_ptr, vi_ptr, vis_type, b_ptr, v_ptr, _ptr(nv), stride, nsnap,
                                norm_n, image_size, cell_size, unit_size, out_ptr, stream or None))


def batch_host(vis_re, vis_im, bin_, vin, nvis=None, norm_n: int = 0, image_size: int = 0,
               cell_size: float = 0.0, unit_size: int = 0, vis_type: int = VIS_F32) -> np.ndarray:
    """fip_batch_host: vis [nsnap][stride], bin [nsnap][stride][2], vin [nsnap][3][3] on the
    host -> tLISI of every window (s, s+1, s+2), shape (nsnap-2, I/T, I/T)."""
    if vis_type == VIS_F32:
        vr, vi = _c(vis_re), _c(vis_im)
    else:
        vr = np.ascontiguousarray(vis_re).view(np.uint16)
        vi = np.ascontiguousarray(vis_im).view(np.uint16)
    nsnap, stride = vr.shape
    b, v = _c(bin_), _c(vin)
    nv = _c(np.full(nsnap, stride) if nvis is None else nvis, np.int32)
    un = image_size // unit_size if unit_size else 0
    out = np.zeros((max(nsnap - 2, 0), un, un), np.float32)
    _check(lib.fip_batch_host(_ptr(vr), _ptr(vi), vis_type, _ptr(b), _ptr(v), _ptr(nv), stride,
                              nsnap, norm_n, image_size, cell_size, unit_size, _ptr(out)))
    return out


def project(uvw, freqs):
    """SVD-plane projection: (bin [nsamp,2] float32, vin [3,3] float32)."""
    u = _c(uvw, np.float64).reshape(-1, 3)
    f = _c(np.atleast_1d(freqs), np.float64)
    ns = u.shape[0] * f.size
    b = np.zeros((ns, 2), dtype=np.float32)
    v = np.zeros((3, 3), dtype=np.float32)
    _check(lib.fip_project(_ptr(u), u.shape[0], _ptr(f), f.size, _ptr(b), _ptr(v)))
    return b, v


def collapse(xx, yy, flag0, flag3, w0, w3) -> np.ndarray:
    xx = np.asarray(xx, dtype=np.complex64)
    yy = np.asarray(yy, dtype=np.complex64)
    n = xx.size
    a = [_c(xx.real), _c(xx.imag), _c(yy.real), _c(yy.imag)]
    f0, f3 = _c(flag0, np.uint8), _c(flag3, np.uint8)
    ww0, ww3 = _c(w0), _c(w3)
    ore, oim = np.zeros(n, np.float32), np.zeros(n, np.float32)
    _check(lib.fip_collapse(*[_ptr(x) for x in a], _ptr(f0), _ptr(f3), _ptr(ww0), _ptr(ww3), n,
                            _ptr(ore), _ptr(oim)))
    return ore + 1j * oim


def ingest_table(uvw, data, flag, freqs, time=None, file_index=None, weight=None,
                 group_by_time: bool = False, workers: int = 0, prefix: str = "",
                 combined: str | None = None):
    """Measurement Set ingest from exported columns (fip_ingest_table).

    uvw [nrows, 3] metres; data [nrows, nchan_data, npol] complex; flag of the
    same shape (bool); freqs [nchan] Hz; time [nrows]; file_index [nrows]
    (table of each row); weight [nrows, nchan_data, npol] or None.  Writes the
    preprocessor's files; returns (snapshots, padded samples per snapshot).
    """
    u = _c(uvw, np.float64).reshape(-1, 3)
    n = u.shape[0]
    d = np.asarray(data)
    if d.ndim != 3 or d.shape[0] != n:
        raise ValueError("data must be [nrows, nchan, npol]")
    nchan_data, npol = d.shape[1], d.shape[2]
    dd = _c(np.stack([d.real, d.imag], axis=-1), np.float32)
    fl = _c(np.asarray(flag).reshape(d.shape), np.uint8)
    f = _c(np.atleast_1d(freqs), np.float64)
    t = None if time is None else _c(np.asarray(time).reshape(n), np.float64)
    fi = None if file_index is None else _c(np.asarray(file_index).reshape(n), np.int32)
    w = None if weight is None else _c(np.asarray(weight).reshape(d.shape), np.float32)
    ns, np_ = ctypes.c_size_t(0), ctypes.c_size_t(0)
    _check(lib.fip_ingest_table(None if fi is None else _ptr(fi), None if t is None else _ptr(t), _ptr(u), n,
                                _ptr(dd), _ptr(fl), None if w is None else _ptr(w), nchan_data, npol, _ptr(f),
                                f.size, int(bool(group_by_time)), int(workers), prefix.encode(),
                                None if combined is None else combined.encode(), ctypes.byref(ns),
                                ctypes.byref(np_)))
    return ns.value, np_.value


def synth(nant: int, radius_m: float, nsnap: int, dt_s: float, freq_hz: float, dec_deg: float,
          lat_deg: float, image_size: int, cell: float, nsrc: int, smin: float, smax: float,
          index: float, noise_sigma: float, transients=(), transient_flux: float = 0.5,
          flag_frac_max: float = 0.0, seed: int = 0):
    """Synthetic workload: returns dict(vis_re, vis_im, bin, vin, nvis, nbl)."""
    nbl = nant * (nant - 1) // 2
    vr = np.zeros((nsnap, nbl), np.float32)
    vi = np.zeros((nsnap, nbl), np.float32)
    b = np.zeros((nsnap, nbl, 2), np.float32)
    v = np.zeros((nsnap, 3, 3), np.float32)
    nv = np.zeros(nsnap, np.int32)
    tr = _c(list(transients) or [0], np.int32)
    _check(lib.fip_synth(nant, radius_m, nsnap, dt_s, freq_hz, dec_deg, lat_deg, image_size, cell,
                         nsrc, smin, smax, index, noise_sigma, len(transients), _ptr(tr),
                         transient_flux, flag_frac_max, seed, _ptr(vr), _ptr(vi), _ptr(b),
                         _ptr(v), _ptr(nv)))
    return {"vis_re": vr, "vis_im": vi, "bin": b, "vin": v, "nvis": nv, "nbl": nbl}


def set_profiling(on: bool) -> None:
    _check(lib.fip_set_profiling(1 if on else 0))


def stage_times() -> dict:
    n = lib.fip_stage_count()
    ms = np.zeros(n, np.float64)
    cnt = np.zeros(n, np.int64)
    _check(lib.fip_stage_times(_ptr(ms), _ptr(cnt), n))
    return {lib.fip_stage_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(n)}


def reset_stage_times() -> None:
    _check(lib.fip_stage_times_reset())


def autotune(image_size: int, num_baselines: int, cell_size: float, reps: int = 3):
    ch = np.zeros(4, np.int32)
    _check(lib.fip_autotune(image_size, num_baselines, cell_size, reps, _ptr(ch)))
    return {"stage1_records": int(ch[0]), "stage2_rows": int(ch[1]), "splat_tile": int(ch[2]),
            "persistent": int(ch[3])}


def set_tuning(image_size: int, stage1_records: int, stage2_rows: int, splat_tile: int,
               persistent: int) -> None:
    ch = np.array([stage1_records, stage2_rows, splat_tile, persistent], np.int32)
    _check(lib.fip_set_tuning(image_size, _ptr(ch)))


def debug_fft(x: np.ndarray) -> np.ndarray:
    """Inverse unnormalised FFT of each row of x (complex) with the LDS kernel."""
    a = np.ascontiguousarray(np.atleast_2d(x), dtype=np.complex64).copy()
    _check(lib.fip_debug_fft(a.shape[1], _ptr(a), a.shape[0]))
    return a


def fits_read(path: str, extname: str | None = None):
    """Read a FITS image (primary or named extension) as float32, numpy axis order."""
    data = _fp()
    dims = (ctypes.c_long * 3)()
    if extname:
        _check(lib.fip_fits_read_ext(path.encode(), extname.encode(), ctypes.byref(data), dims))
    else:
        _check(lib.fip_fits_read(path.encode(), ctypes.byref(data), dims))
    n = dims[0] * dims[1] * dims[2]
    try:
        arr = np.ctypeslib.as_array(data, shape=(n,)).copy()
    finally:
        lib.fip_free(ctypes.cast(data, _vp))
    shape = [d for d in (dims[2], dims[1], dims[0])]
    return arr.reshape(shape)


def fits_write(path: str, data2d) -> None:
    a = _c(data2d)
    if a.ndim != 2:
        raise ValueError("fits_write takes a 2-D array")
    _check(lib.fip_fits_write(path.encode(), _ptr(a), a.shape[1], a.shape[0]))


def fits_write3(path: str, data3d) -> None:
    a = _c(data3d)
    if a.ndim != 3:
        raise ValueError("fits_write3 takes a 3-D array")
    _check(lib.fip_fits_write3(path.encode(), _ptr(a), a.shape[2], a.shape[1], a.shape[0]))


class Server:
    """Serving form: device ring buffers per sub-array stream, HBM-budget admission."""

    def __init__(self, streams, ring: int = 8, hbm_budget_bytes: int = 64 << 30):
        cfg = (_StreamConfig * len(streams))()
        for k, s in enumerate(streams):
            cfg[k].num_baselines = s["num_baselines"]
            cfg[k].image_size = s["image_size"]
            cfg[k].cell_size = s["cell_size"]
            cfg[k].unit_size = s["unit_size"]
            cfg[k].vis_type = s.get("vis_type", VIS_F32)
        self._h = _vp()
        _check(lib.fip_server_create(cfg, len(streams), ring, hbm_budget_bytes,
                                     ctypes.byref(self._h)))
        self._streams = list(streams)

    def submit(self, sid: int, vis_re, vis_im, bin_, vin, arrival_s: float = -1.0) -> bool:
        """arrival_s on time.monotonic()'s clock (negative: now).  Returns False when the
        snapshot is refused (its inputs do not fit the free part of the input pool
        that the HBM budget leaves)."""
        vt = self._streams[sid].get("vis_type", VIS_F32)
        if vt == VIS_F32:
            vr, vi = _c(vis_re), _c(vis_im)
        else:
            vr = np.ascontiguousarray(vis_re).view(np.uint16)
            vi = np.ascontiguousarray(vis_im).view(np.uint16)
        b, v = _c(bin_), _c(vin)
        rc = lib.fip_server_submit(self._h, sid, _ptr(vr), _ptr(vi), _ptr(b), _ptr(v), arrival_s)
        if rc == -7:
            return False
        _check(rc)
        return True

    def drain(self) -> None:
        _check(lib.fip_server_drain(self._h))

    def stats(self) -> dict:
        p50, p99, mean = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        done, refused = ctypes.c_long(), ctypes.c_long()
        _check(lib.fip_server_stats(self._h, ctypes.byref(p50), ctypes.byref(p99),
                                    ctypes.byref(mean), ctypes.byref(done), ctypes.byref(refused)))
        return {"p50_ms": p50.value, "p99_ms": p99.value, "mean_ms": mean.value,
                "processed": done.value, "refused": refused.value}

    def sequences(self) -> int:
        """Launch sequences enqueued so far: snapshots that queued up behind a
        running sequence of their stream were imaged together."""
        n = ctypes.c_long()
        _check(lib.fip_server_sequences(self._h, ctypes.byref(n)))
        return n.value

    def latest(self, sid: int):
        s = self._streams[sid]
        un = s["image_size"] // s["unit_size"]
        out = np.zeros((un, un), np.float32)
        win = ctypes.c_long()
        _check(lib.fip_server_latest(self._h, sid, _ptr(out), ctypes.byref(win)))
        return out, win.value

    def close(self) -> None:
        if self._h:
            lib.fip_server_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
