"""Seeded synthetic inputs for parity tests (numpy only; no reference code)."""
from __future__ import annotations

import numpy as np


def rotation(tilt: float, azimuth: float, spin: float) -> np.ndarray:
    """Orthonormal V with rows V0, V1, V0 x V1: the uv plane tilted by `tilt`."""
    def rz(a):
        c, s = np.cos(a), np.sin(a)
        return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])

    def rx(a):
        c, s = np.cos(a), np.sin(a)
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])

    R = rz(azimuth) @ rx(tilt) @ rz(spin)
    V = R.T.copy()
    V[2] = np.cross(V[0], V[1])
    return V


def snapshot(seed: int, nbl: int, image: int, cell: float, radius_frac: float = 0.27,
             nsrc: int = 5, tilt: float = 0.25, noise: float = 0.05, extra=None):
    """One snapshot: B inside a disc of radius_frac*g grid cells, point-source vis."""
    rng = np.random.default_rng(seed)
    g = int(np.ceil(np.float32(1.5 * np.float32(image))))
    R = radius_frac * g
    r = R * np.sqrt(rng.uniform(0, 1, nbl)) ** 1.5
    a = rng.uniform(0, 2 * np.pi, nbl)
    p = np.stack([r * np.cos(a), r * np.sin(a)], axis=1)
    B = (p / (np.float64(cell) * g)).astype(np.float32)
    V = rotation(tilt, rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi)).astype(np.float32)
    fov = 0.35 * image * cell
    l = rng.uniform(-fov, fov, nsrc)
    m = rng.uniform(-fov, fov, nsrc)
    S = rng.uniform(0.1, 1.0, nsrc)
    if extra is not None:
        l = np.append(l, extra[0]); m = np.append(m, extra[1]); S = np.append(S, extra[2])
    Bd = B.astype(np.float64)
    ph = -2j * np.pi * (Bd[:, :1] * l[None, :] + Bd[:, 1:] * m[None, :])
    vis = (S[None, :] * np.exp(ph)).sum(axis=1)
    vis = vis + noise * (rng.standard_normal(nbl) + 1j * rng.standard_normal(nbl))
    return (vis.real.astype(np.float32), vis.imag.astype(np.float32), B, V)


def snapshots(seed: int, nsnap: int, nbl: int, image: int, cell: float, transient_snap=1, **kw):
    """nsnap snapshots of one array (same B per snapshot up to a small rotation)."""
    out = [snapshot(seed + 100 * s, nbl, image, cell, **kw) for s in range(nsnap)]
    if transient_snap is not None:
        rng = np.random.default_rng(seed + 7)
        fov = 0.3 * image * cell
        l, m = rng.uniform(-fov, fov), rng.uniform(-fov, fov)
        vr, vi, B, V = out[transient_snap]
        Bd = B.astype(np.float64)
        t = 2.0 * np.exp(-2j * np.pi * (Bd[:, 0] * l + Bd[:, 1] * m))
        out[transient_snap] = ((vr + t.real).astype(np.float32), (vi + t.imag).astype(np.float32), B, V)
    vr = np.stack([o[0] for o in out])
    vi = np.stack([o[1] for o in out])
    B = np.stack([o[2] for o in out])
    V = np.stack([o[3] for o in out])
    return vr, vi, B, V


def rel_rms(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-300))


# Synthetic workloads from the library's own generator (fip_synth): a
# 64-antenna array within 5 km imaged at 256^2 (2016 baselines, tile 16) and a
# 512-antenna array within 40 km imaged at 4096^2 (130816 baselines, tile 32),
# 140 MHz, dec -30 deg, power-law sky plus a transient in snapshot 2.
SMALL = dict(nant=64, radius_m=5000.0, image=256, cell=5.8e-5, tile=16, nsrc=20, noise=0.05,
             seed=20261202)
LARGE = dict(nant=512, radius_m=40000.0, image=4096, cell=7.2722e-6, tile=32, nsrc=200, noise=0.1,
             seed=20261203)


def synth_workload(F, w, nsnap=3, transients=(2,), flag=0.0, seed=None, image=None):
    return F.synth(w["nant"], w["radius_m"], nsnap, 1.0, 140e6, -30.0, -26.7,
                   image or w["image"], w["cell"], w["nsrc"], 0.01, 1.0, -1.8, w["noise"],
                   transients=list(transients), transient_flux=0.5, flag_frac_max=flag,
                   seed=w["seed"] if seed is None else seed)


def to_bf16_bits(x):
    """float32 -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


def triggered(tlisi, tau):
    return set(zip(*np.nonzero(np.asarray(tlisi) < tau)))


def raw_snapshots(seed: int, nant: int, radius_m: float, nsnap: int, freqs, image: int, cell: float,
                  nsrc: int = 30, noise: float = 0.05, flag_frac: float = 0.1, dec_deg: float = -30.0,
                  lat_deg: float = -26.7, dt_s: float = 1.0):
    """Raw per-snapshot inputs of the preprocessing chain (numpy only): uvw of
    every antenna pair of a log-radial array under earth rotation, and XX/YY
    from an FP64 DFT of point sources over D = uvw nu/c including the w term,
    with noise, random flags and weights.  Samples are channel-major
    (n = row + chan*nrows).  Returns uvw [S][nrows][3] (float64), freqs,
    xx, yy [S][ns] complex64, flag0, flag3 [S][ns] uint8, w0, w3 [S][ns] f32."""
    rng = np.random.default_rng(seed)
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    r = np.exp(rng.uniform(np.log(20.0), np.log(radius_m), nant))
    az = rng.uniform(0, 2 * np.pi, nant)
    east, north = r * np.cos(az), r * np.sin(az)
    lat, dec = np.radians(lat_deg), np.radians(dec_deg)
    X, Y, Z = -np.sin(lat) * north, east, np.cos(lat) * north
    i, j = np.triu_indices(nant, 1)
    bx, by, bz = X[j] - X[i], Y[j] - Y[i], Z[j] - Z[i]
    nrows = bx.size
    fov = 0.3 * image * cell
    l = rng.uniform(-fov, fov, nsrc)
    m = rng.uniform(-fov, fov, nsrc)
    n1 = np.sqrt(1.0 - l * l - m * m) - 1.0
    flux = rng.uniform(0.05, 1.0, nsrc)
    c = 299792458.0
    ns = nrows * freqs.size
    uvw = np.zeros((nsnap, nrows, 3))
    xx = np.zeros((nsnap, ns), np.complex64)
    yy = np.zeros((nsnap, ns), np.complex64)
    for s in range(nsnap):
        H = 2 * np.pi * s * dt_s / 86164.0
        u = np.sin(H) * bx + np.cos(H) * by
        v = -np.sin(dec) * np.cos(H) * bx + np.sin(dec) * np.sin(H) * by + np.cos(dec) * bz
        w = np.cos(dec) * np.cos(H) * bx - np.cos(dec) * np.sin(H) * by + np.sin(dec) * bz
        uvw[s] = np.stack([u, v, w], axis=1)
        vis = np.zeros(ns, np.complex128)
        for ch, f in enumerate(freqs):
            D = uvw[s] * (f / c)
            ph = -2j * np.pi * (D[:, :1] * l[None, :] + D[:, 1:2] * m[None, :] + D[:, 2:3] * n1[None, :])
            vis[ch * nrows:(ch + 1) * nrows] = (flux[None, :] * np.exp(ph)).sum(axis=1)
        xx[s] = vis + noise * (rng.standard_normal(ns) + 1j * rng.standard_normal(ns))
        yy[s] = vis + noise * (rng.standard_normal(ns) + 1j * rng.standard_normal(ns))
    flag0 = (rng.uniform(size=(nsnap, ns)) < flag_frac).astype(np.uint8)
    flag3 = (rng.uniform(size=(nsnap, ns)) < flag_frac).astype(np.uint8)
    w0 = rng.uniform(0.5, 2.0, (nsnap, ns)).astype(np.float32)
    w3 = rng.uniform(0.5, 2.0, (nsnap, ns)).astype(np.float32)
    return uvw, freqs, xx, yy, flag0, flag3, w0, w3
