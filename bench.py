#!/usr/bin/env python3
"""Benchmark of the MI355X FIP-TOI imaging-and-trigger path (one JSON line on stdout).

Headline (``value``): snapshots/s of the plain FP32 drop-in path
(fip_pipe_device: three snapshots imaged + one tLISI trigger per step) for a
512-antenna array (130816 baselines) imaged at 4096^2 with tile 32, inputs
resident in HBM before the timed region starts.  Beside it: per-stage kernel
times (hipEvents on the launching stream), the roofline of the dominant
kernel, the batched (hipGraph) form in FP32 and with BF16-stored
visibilities over 64 unevenly flagged snapshots at 4096^2 and 2048^2, the
serving form's p50/p99 arrival->tLISI latency for three sub-arrays, and the
FP64 CPU oracle timed on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# 512 antennas log-radially within 40 km, 140 MHz, dec -30 deg, 4096^2 pixels of
# 7.2722e-6 rad, tile 32; 200 power-law sources (0.01-1 Jy, index -1.8), a 0.5 Jy
# transient in snapshot 2, complex noise sigma 0.1 Jy.
ARRAY = dict(nant=512, radius_m=40000.0, image=4096, cell=7.2722e-6, tile=32, nsrc=200, noise=0.1,
             freq=140e6, dec=-30.0, lat=-26.7, seed=20261203)
WORKLOAD = "FIpipe: 3 snapshots, 512 antennas (130816 baselines), 4096^2, tile 32, cell 7.2722e-6 rad"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_4096px_130816bl.json")
# Serving sub-arrays: (antennas, radius m, image, cell): the cell scales with
# the longest baseline so every stream fills the same fraction of its grid.
SUBARRAYS = [(256, 20000.0, 2048, 2 * 7.2722e-6), (512, 40000.0, 4096, 7.2722e-6),
             (1024, 80000.0, 8192, 7.2722e-6 / 2)]


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def synth(F, nsnap=3, transients=(2,), flag=0.0, seed=None, nant=None, radius=None, image=None,
          cell=None):
    a = ARRAY
    return F.synth(nant or a["nant"], radius or a["radius_m"], nsnap, 1.0, a["freq"], a["dec"], a["lat"],
                   image or a["image"], cell or a["cell"], a["nsrc"], 0.01, 1.0, -1.8, a["noise"],
                   transients=list(transients), transient_flux=0.5, flag_frac_max=flag,
                   seed=a["seed"] if seed is None else seed)


def to_bf16_bits(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def rmax_of(B, cell, g):
    """Largest |column tap| over visibilities (the column extent gridded and transformed)."""
    m = g // 2
    p2 = B[..., 1].astype(np.float64) * (np.float64(np.float32(cell)) * g)
    lo = np.maximum(np.ceil(p2 - 4), -m)
    hi = np.minimum(np.floor(p2 + 4), m - 1)
    return int(max(hi.max(), (-lo).max()))


def stage_bytes(nbl, image, g, rmax, tile, nsnap=3, nwin=1, fused=False):
    """Algorithmic HBM bytes per launch of each stage: every input read once and
    every output written once, for the nsnap snapshots (nwin windows) one launch
    processes.  Visibility records are 24 B (two bucket/offset pairs + complex
    value), the column-transformed grid holds rmax+1 complex columns of I rows.
    fused: the last snapshot's splat carries the trigger (splat_trigger reads
    its corrected image and the window's first two images, and writes three
    terms per trigger tile; the trigger stage turns the terms into the tLISI)."""
    I = image
    kcols = rmax + 1
    un2 = (I // tile) ** 2
    b = {
        "reset": nsnap * (g * 4 + 64),
        "bin": nsnap * nbl * (4 + 4 + 8 + 24 + 8),
        "scan": nsnap * (g * 4 + (g + 1) * 4),
        "scatter": nsnap * nbl * (8 + 24 + 24),
        "grid_colfft": nsnap * (nbl * 24 + kcols * I * 8),
        "row_c2r": nsnap * (kcols * I * 8 + I * I * 4),
        "splat": nsnap * (I * I * 4 * 2),
        "trigger": nwin * (3 * I * I * 4 + un2 * 4),
    }
    if fused:
        b["splat"] = (nsnap - 1) * (I * I * 4 * 2)
        b["splat_trigger"] = I * I * 4 + 2 * I * I * 4 + un2 * 12
        b["trigger"] = un2 * (12 + 4)
    return b


def load_traffic(stage: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (copy-calibrated), if present."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    try:
        with open(TRAFFIC_FILE) as f:
            k = json.load(f)["kernels"].get(stage)
        return None if k is None else int(round(k["bytes"]))
    except Exception:
        return None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(d, nbl):
    """FP64 oracle on one snapshot (projection + imaging) + 1/3 of a trigger window."""
    from oracle import oracle as O
    a = ARRAY
    nthreads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(1)
    uvw = rng.normal(size=(nbl, 3)) * np.array([2e4, 2e4, 1e3])   # metres, one snapshot's rows
    t0 = time.perf_counter()
    O.project(uvw, [a["freq"]])
    t1 = time.perf_counter()
    img, mx = O.image(d["vis_re"][0], d["vis_im"][0], d["bin"][0], d["vin"][0], a["image"], a["cell"],
                      nthreads=nthreads)
    t2 = time.perf_counter()
    O.trigger(img, img * 0.5, img, np.array([mx, mx, mx]), a["tile"])
    t3 = time.perf_counter()
    per_snap = (t1 - t0) + (t2 - t1) + (t3 - t2) / 3.0
    return {"value": round(1.0 / per_snap, 4), "unit": "snapshots/s", "cores": nthreads, "kind": "port",
            "sample": "FP64 oracle (oracle/fip_oracle.c, OpenMP) on 1 of the 3 snapshots of the 4096^2 "
                      "workload: SVD_MFS.py projection of 130816 uvw rows + dense 6144^2 gridding, FFT, "
                      "correction, SIN map, splat + 1/3 of one tLISI window; "
                      f"{per_snap:.2f} s per snapshot; {nthreads} OpenMP threads of {os.cpu_count()} "
                      f"logical CPUs ({cpu_model()})"}


def bench_batched(F, H, d, image, cell, vis_type, reps):
    """64 snapshots resident in HBM through fip_batch_device (hipGraph replay)."""
    S, nbl = d["vis_re"].shape
    un = image // ARRAY["tile"]
    if vis_type == F.VIS_BF16:
        vr, vi = H.DeviceArray(to_bf16_bits(d["vis_re"])), H.DeviceArray(to_bf16_bits(d["vis_im"]))
    else:
        vr, vi = H.DeviceArray(d["vis_re"]), H.DeviceArray(d["vis_im"])
    b, v = H.DeviceArray(d["bin"]), H.DeviceArray(d["vin"])
    out = H.DeviceArray(nbytes=(S - 2) * un * un * 4)
    st = H.Stream()

    def call():
        F.batch_device(vr.addr, vi.addr, vis_type, b.addr, v.addr, d["nvis"], nbl, S, nbl, image, cell,
                       ARRAY["tile"], out.addr, stream=st.addr)

    call()                        # capture
    call()
    st.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    st.synchronize()
    wall = time.perf_counter() - t0
    res = out.to_host(np.float32, (S - 2, un, un))
    for a in (vr, vi, b, v, out):
        a.free()
    return S * reps / wall, res


def bench_serving(F, seconds=2.0, rate=200.0, saturate=120):
    """Poisson arrivals over three sub-arrays (BF16 visibilities, ring of 8)."""
    streams, data = [], []
    for k, (nant, radius, image, cell) in enumerate(SUBARRAYS):
        d = synth(F, nsnap=4, transients=(2,), seed=20261210 + k, nant=nant, radius=radius, image=image,
                  cell=cell)
        streams.append(dict(num_baselines=d["nbl"], image_size=image, cell_size=cell, unit_size=32,
                            vis_type=F.VIS_BF16))
        data.append((to_bf16_bits(d["vis_re"]), to_bf16_bits(d["vis_im"]), d["bin"], d["vin"]))
    srv = F.Server(streams, ring=8, hbm_budget_bytes=48 << 30)
    try:
        rng = np.random.default_rng(7)
        n = int(seconds * rate)
        gaps = rng.exponential(1.0 / rate, n)
        which = rng.integers(0, len(streams), n)
        t0 = time.monotonic()
        t_arr = t0
        refused = 0
        for i in range(n):
            t_arr += gaps[i]
            now = time.monotonic()
            if t_arr > now:
                time.sleep(t_arr - now)
            k = int(which[i])
            vr, vi, b, v = data[k]
            s = i % 4
            if not srv.submit(k, vr[s], vi[s], b[s], v[s], t_arr):
                refused += 1
        srv.drain()
        elapsed = time.monotonic() - t0
        st = srv.stats()
        poisson = {"offered_per_s": rate, "submitted": n, "processed": st["processed"], "refused": refused,
                   "p50_ms": round(st["p50_ms"], 4), "p99_ms": round(st["p99_ms"], 4),
                   "mean_ms": round(st["mean_ms"], 4), "sustained_per_s": round(st["processed"] / elapsed, 1)}
    finally:
        srv.close()
    # Saturation: the 4096^2 stream alone, submitted back to back, after an
    # untimed pass on a server of its own (first launches of the kernels that
    # only batched sequences use, first use of a server's buffers).
    vr, vi, b, v = data[1]
    warm = F.Server([streams[1]], ring=8, hbm_budget_bytes=48 << 30)
    try:
        for k in range(24):
            while not warm.submit(0, vr[k % 4], vi[k % 4], b[k % 4], v[k % 4], -1.0):
                pass
        warm.drain()
    finally:
        warm.close()
    srv = F.Server([streams[1]], ring=8, hbm_budget_bytes=48 << 30)
    try:
        t0 = time.monotonic()
        done = 0
        in_submit = 0.0   # host time inside accepted submits (staging copy included)
        while done < saturate:
            ts = time.perf_counter()
            ok = srv.submit(0, vr[done % 4], vi[done % 4], b[done % 4], v[done % 4], -1.0)
            if ok:
                in_submit += time.perf_counter() - ts
                done += 1
        srv.drain()
        elapsed = time.monotonic() - t0
        st = srv.stats()
        sat = {"stream": "512 antennas, 4096^2", "snapshots": saturate,
               "snapshots_per_s": round(saturate / elapsed, 1), "p50_ms": round(st["p50_ms"], 4),
               "p99_ms": round(st["p99_ms"], 4), "launch_sequences": srv.sequences(),
               "submit_ms_per_snapshot": round(in_submit / saturate * 1e3, 4)}
    finally:
        srv.close()
    return {"poisson": poisson, "saturated_4096": sat,
            "subarrays": [f"{n} ant ({s['num_baselines']} bl) {s['image_size']}^2"
                          for (n, *_), s in zip(SUBARRAYS, streams)]}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extra", action="store_true", help="skip batched/serving/latency extras")
    ap.add_argument("--profile-only", action="store_true",
                    help="only run warmup+steps of the plain path and a calibration copy (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(sharding across GPUs is a design note, not a launch path)"}))
        return 2

    import fastimagingpipe_amd as F
    from fastimagingpipe_amd import hip as H

    a = ARRAY
    I, T, cell = a["image"], a["tile"], a["cell"]
    g = F.grid_size(I)
    log("generating the 4096^2 synthetic snapshots ...")
    d = synth(F)
    nbl = d["nbl"]
    rmax = rmax_of(d["bin"], cell, g)
    # The profiled run keeps the default launch configuration (what the
    # autotuner picks at this shape) so rocprof sees only the timed kernels.
    tuning = None if args.profile_only else F.autotune(I, nbl, cell, reps=3)
    log(f"autotuned launch configuration at {I}^2: {tuning}")
    stream = H.Stream()
    dvr, dvi = H.DeviceArray(d["vis_re"]), H.DeviceArray(d["vis_im"])
    db, dv = H.DeviceArray(d["bin"]), H.DeviceArray(d["vin"])
    un = I // T
    dout = H.DeviceArray(nbytes=un * un * 4)

    def step():
        F.pipe_device(dvr.addr, dvi.addr, db.addr, dv.addr, dout.addr, nbl, I, 3, cell, T, stream.addr)

    log("warmup ...")
    for _ in range(args.warmup):
        step()
    stream.synchronize()
    H.synchronize()
    e0, e1 = H.Event(), H.Event()
    log(f"timing {args.steps} steps ...")
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    stream.synchronize()
    H.synchronize()
    wall = time.perf_counter() - t0
    ms_step = e0.elapsed_ms(e1) / args.steps
    value = 3.0 * args.steps / wall
    tlisi = dout.to_host(np.float32, (un, un))
    log(f"plain path: {value:.1f} snapshots/s, {ms_step:.4f} ms/step")
    if args.profile_only:
        # Calibration for the rocprofv3 counters: three 256 MiB device-to-device
        # copies (a copy kernel of known size).
        nb = 256 << 20
        src, dst = H.DeviceArray(nbytes=nb), H.DeviceArray(nbytes=nb)
        for _ in range(3):
            H._check(H._hip.hipMemcpy(dst.ptr, src.ptr, nb, H.D2D), "hipMemcpy D2D")
        H.synchronize()
        print(json.dumps({"metric": "snapshots/s", "value": round(value, 2), "ms_per_step": ms_step,
                          "copy_bytes": nb}))
        return 0

    # Per-stage durations: hipEvents around every launch on the launching stream, K steps.
    F.reset_stage_times()
    F.set_profiling(True)
    for _ in range(args.steps):
        step()
    F.set_profiling(False)
    st = F.stage_times()
    per = {k: v[0] / v[1] for k, v in st.items() if v[1] > 0}
    sb = stage_bytes(nbl, I, g, rmax, T, fused="splat_trigger" in per)
    stages = {k: {"ms": round(per[k], 5), "GB_s": round(sb[k] / (per[k] * 1e-3) / 1e9, 1) if per[k] > 0 else None,
                  "traffic_bytes": load_traffic(k)} for k in per}
    dom = max(per, key=lambda k: per[k])
    achieved = sb[dom] / (per[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(dom), "kernel": dom,
                "algorithmic_bytes": sb[dom], "avg_ms": round(per[dom], 5)}

    extra = {}
    if not args.no_extra:
        lat = []
        for _ in range(max(20, args.steps // 2)):
            t = time.perf_counter()
            step()
            stream.synchronize()
            lat.append((time.perf_counter() - t) * 1e3)
        extra["step_latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 4),
                                    "p99": round(float(np.percentile(lat, 99)), 4),
                                    "what": "enqueue of 3 snapshots -> tLISI in HBM, one step at a time"}
        n_host = 50
        host_lat = []
        for _ in range(3):   # warm the host path's staging and copy stream
            F.fipipe(d["vis_re"], d["vis_im"], d["bin"], d["vin"], nbl, I, 3, cell, T)
        for _ in range(n_host):
            t = time.perf_counter()
            F.fipipe(d["vis_re"], d["vis_im"], d["bin"], d["vin"], nbl, I, 3, cell, T)
            host_lat.append((time.perf_counter() - t) * 1e3)
        extra["host_pipe_snapshots_per_s"] = round(3 * n_host / (sum(host_lat) * 1e-3), 1)
        extra["host_pipe_latency_ms"] = {"p50": round(float(np.percentile(host_lat, 50)), 4),
                                         "p99": round(float(np.percentile(host_lat, 99)), 4),
                                         "what": "fip_pipe: host arrays in -> tLISI in host memory"}
        extra["tlisi_min"] = float(tlisi.min())
        for a_ in (dvr, dvi, db, dv):
            a_.free()

        log("batched form: generating 64 flagged snapshots ...")
        db64 = synth(F, nsnap=64, transients=(9, 31, 50), flag=0.3, seed=20261204)
        batched = {"snapshots": 64, "nvis_min": int(db64["nvis"].min()), "nvis_max": int(db64["nvis"].max())}
        for image, cell_b in ((4096, cell), (2048, 2 * cell)):
            for name, vt in (("f32", F.VIS_F32), ("bf16", F.VIS_BF16)):
                rate, res = bench_batched(F, H, db64, image, cell_b, vt, reps=10 if image == 4096 else 20)
                batched[f"{image}_{name}_snapshots_per_s"] = round(rate, 1)
                if image == 4096:
                    batched[f"{image}_{name}_vs_plain"] = round(rate / value, 3)
                    # trigger windows per second: 62 per batch of 64, 1 per plain 3-snapshot step
                    batched[f"{image}_{name}_windows_vs_plain"] = round((rate * 62 / 64) / (value / 3), 3)
                log(f"batched {image}^2 {name}: {rate:.1f} snapshots/s")
        extra["batched"] = batched
        log("serving form ...")
        extra["serving"] = bench_serving(F)

    cpu = None
    if not args.no_cpu:
        log("CPU oracle baseline ...")
        cpu = cpu_baseline(d, nbl)

    out = {
        "metric": "snapshots/s at 4096^2, 130816 bl (plain FP32 path)",
        "value": round(value, 2),
        "unit": "snapshots/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "latency_ms": extra.get("step_latency_ms"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (fip_synth: 512-antenna log-radial array, 200-source power-law sky + 0.5 Jy "
                "transient, sigma 0.1 Jy)",
        "config": {"workload": WORKLOAD, "image": I, "grid": g, "baselines": nbl, "snapshots_per_step": 3,
                   "tile": T, "uv_columns": rmax + 1, "launch_tuning": tuning},
        "stages": stages,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
