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
