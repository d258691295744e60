#!/usr/bin/env python3
"""Benchmark: batched 2D elastic RTM on one MI355X.

Workload (default): 64 shots on a seeded layered + anomaly model of
nz=800 x nx=2400 cells, dx=10 m, dt=0.8 ms, 6000 time levels, 15 Hz Ricker
at z=20 m, receivers at every column of the shot's aperture at z=20 m;
aperture widths drawn uniformly from 600..1200 columns (seeded), source x
uniform along the model. Each shot's grid is its aperture padded by the 45-cell
sponge on every side, so grids are uneven within one launch.

A "step" is one full RTM of the whole batch: forward source propagation with
snapshots, adjoint propagation of the observed data and the dlambda / dmu /
drho imaging condition, stacked over shots. Observed data (true model minus
smooth model, both modelled on the GPU) and every input are resident in HBM
before the timed region.

Prints ONE JSON line. ``value`` is Mcell-updates/s on the physical grid
(forward + adjoint cell-updates of every shot per second); shots/hour and the
computational-grid rate are reported beside it. The run is valid only when
(a) shot 0 of the timed batch, run with the batch's own kernel form and strip
widths, equals the same shot run alone on the tile kernels bit for bit, and
(b) the images of a 200 x 400 sub-run on those tile kernels are within 1e-3
relative L2 of the FP64 oracle; both checks run in the same process.

After the timed steps the same survey and observed data run once through the
plain FP32 GPU form (one shot per launch, no graphs) and once through the
BF16 form (snapshots every 4th level with per-strip scales, compensated
images): ``variants`` reports their shots/hour, the ratios, and the stacked
images' relative L2 against the batched FP32 run.

Other workloads (``--workload``): acoustic3d (3D acoustic subdomain),
elastic3d (3D elastic 9-field adjoint step), probe (the library's streaming
kernel: attainable HBM rate and counter calibration).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

VARIANTS = {
    # plain FP32 GPU form: one shot at a time, fp32 Vx/Vz history of every
    # level, one launch per level
    "plain": dict(snap_every=1, snap_prec="fp32", comp_image=0, max_group=1, use_graph=0),
    # batched FP32: many shots per launch (memory-bound group size), hipGraph
    "batched": dict(snap_every=1, snap_prec="fp32", comp_image=0, max_group=0, use_graph=1),
    # mixed precision: bf16 / fp16 snapshots every 4th level, compensated images
    "bf16": dict(snap_every=4, snap_prec="bf16", comp_image=1, max_group=0, use_graph=1),
    "fp16": dict(snap_every=4, snap_prec="fp16", comp_image=1, max_group=0, use_graph=1),
}

# algorithmic bytes per cell-update of each kernel (fp32 fields, 5 fields of
# the previous level + 5 of the level before + 5 written + 5 coefficients =
# 80 B, plus the snapshot traffic of the variant)
def algo_bytes(variant):
    v = VARIANTS[variant]
    j = v["snap_every"]
    if v["snap_prec"] == "fp32" and j == 1:
        # forward: Vx, Vz history written (8 B). Adjoint: the history read at
        # levels L and L+1 (S1 = (V(L+1) - V(L)) / dt, 16 B; a kernel that
        # advances one level per launch reads each history level twice) and
        # 3 images read and written (24 B)
        fwd, adj = 80 + 8, 80 + 16 + 24
    else:
        b = 2 if v["snap_prec"] in ("bf16", "fp16") else 4
        img = 48 if v["comp_image"] else 24
        fwd = 80 + 4 * b / j
        adj = 80 + (4 * b + img) / j
    return fwd, adj


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_batch(lib, synth, nz, nx, nshots, nt, dx, dt, fp, abc, variant, seed, txr=0, graph=True):
    vp, vps, layout = synth.survey(nz, nx, nshots, seed, abc, dx, dt)
    specs, ctrue, cmig = [], [], []
    for xs, x0, w in layout:
        sp, ct, cm = synth.window_shot(vp, vps, x0, w, xs, abc, dx, dt)
        specs.append(sp)
        ctrue.append(ct)
        cmig.append(cm)
    wav = lib.ricker(nt, dt, fp)
    opts = dict(VARIANTS[variant])
    if not graph:
        opts["use_graph"] = 0
    b = lib.Batch(nt, dx, dt, specs, cmig, wav, abc=abc, adj_abc=1, stack_shape=(nz, nx), txr=txr, **opts)
    b.set_true_model(ctrue)
    return b, specs, cmig, wav, layout


class env_pinned:
    """Temporarily set RTM_* environment variables (the library reads them
    at each call)."""

    def __init__(self, **kv):
        self.kv = {k: str(v) for k, v in kv.items() if v is not None}
        self.old = {}

    def __enter__(self):
        for k, v in self.kv.items():
            self.old[k] = os.environ.get(k)
            os.environ[k] = v
        return self

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def validate(lib, synth, variant, seed, txr):
    """Sub-run on a 200 x 400 model (2 shots, 1000 levels) through the same
    engine options, images against the FP64 oracle. It runs the tile kernels
    at the timed batch's tile width; shot_matches_single() ties the timed
    batch's own kernel form to that path bit for bit."""
    import oracle as O

    nz, nx, nt, dx, dt, fp, abc = 200, 400, 1000, 10.0, 8e-4, 15.0, 40
    vp, vps, layout = synth.survey(nz, nx, 2, seed + 7, abc, dx, dt, wmin=300, wmax=400)
    specs, ctrue, cmig = [], [], []
    for xs, x0, w in layout:
        sp, ct, cm = synth.window_shot(vp, vps, x0, w, xs, abc, dx, dt)
        specs.append(sp)
        ctrue.append(ct)
        cmig.append(cm)
    wav = lib.ricker(nt, dt, fp)
    opts = VARIANTS[variant]
    b = lib.Batch(nt, dx, dt, specs, cmig, wav, abc=abc, adj_abc=1, stack_shape=(nz, nx), txr=txr, **opts)
    b.set_true_model(ctrue)
    b.synthesize()
    b.migrate()
    errs = []
    nth = cpu_threads()
    for i, sp in enumerate(specs):
        rec = b.get_data(i)
        img = b.image(i)
        p = O.make_params(sp["nz"], sp["nx"], nt, dx, dt, sp["xs"], sp["zs"], sp["zr"], sp["rx0"], sp["nrx"], abc,
                          adj_abc=1, snap_every=opts["snap_every"], nthreads=nth)
        ref = O.rtm(p, wav, cmig[i], rec, np.float64)
        z0, x0 = sp["img_z0"], sp["img_x0"]
        sl = (slice(z0, z0 + sp["img_nz"]), slice(x0, x0 + sp["img_nx"]))
        errs.append([O.rel_l2(a[sl], r[sl]) for a, r in zip(img, ref)])
    b.close()
    return errs


def shot_matches_single(lib, sample, txr, specs, cmig, wav, nt, dx, dt, abc, variant, k=0):
    """Shot k of a timed batch (its kernel form and strip widths; sample =
    its observed data and images, taken before the batch released its device
    memory) against the same shot run alone through the single-shot entry
    point on the tile kernels of the batch's tile width txr: the images must
    agree bit for bit."""
    v = VARIANTS[variant]
    sp = specs[k]
    rec, img_b = sample
    s = lib.shot(sp["nz"], sp["nx"], nt, dx, dt, sp["xs"], sp["zs"], sp["zr"], sp["rx0"], sp["nrx"], abc, adj_abc=1,
                 snap_every=v["snap_every"], snap_prec=lib.SNAP_PREC[v["snap_prec"]], comp_image=v["comp_image"])
    with env_pinned(RTM_TBLOCK=0, RTM_TXR=txr, RTM_GRAPH=0):
        img_s = lib.migrate(s, cmig[k], wav, rec)
    return all(np.array_equal(a, c) for a, c in zip(img_b, img_s)) and all(np.isfinite(a).all() for a in img_b)


def cpu_threads():
    """Threads of the CPU baseline: every CPU this process may run on, within
    the CPU share the job is granted (cgroup quota; at most 16 on the
    one-GPU hosts this runs on)."""
    return host_info()["threads"]


_HOST = None


def host_info():
    global _HOST
    if _HOST is None:
        import platform
        import subprocess

        aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        quota = None
        try:
            q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
            if q != "max":
                quota = int(q) / int(per)
        except (OSError, ValueError):
            pass
        model = platform.processor()
        try:
            model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        try:
            gxx = subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
        except (OSError, IndexError):
            gxx = "unknown"
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = aff
        if quota:
            threads = min(threads, max(1, int(quota)))
        if share > 0:
            threads = min(threads, share)
        _HOST = dict(nproc=os.cpu_count(), affinity=aff, cgroup_cpus=quota, omp_num_threads=share or None,
                     cpu_model=model, compiler=gxx, threads=threads)
    return _HOST


def cpu_baseline(lib, synth, seed, dx, dt, fp, abc, nt_sample, nthreads):
    """Oracle FP32 port (reference operation order, -O3, no FMA) on one
    config-shaped shot (800 x 900 physical, padded), nt_sample levels of
    forward + adjoint + imaging."""
    import oracle as O

    vp, vps, layout = synth.survey(800, 900, 1, seed, abc, dx, dt, wmin=900, wmax=900)
    xs, x0, w = layout[0]
    sp, ct, cm = synth.window_shot(vp, vps, x0, w, xs, abc, dx, dt)
    wav = lib.ricker(nt_sample, dt, fp)
    p = O.make_params(sp["nz"], sp["nx"], nt_sample, dx, dt, sp["xs"], sp["zs"], sp["zr"], sp["rx0"], sp["nrx"], abc,
                      adj_abc=1, nthreads=nthreads)
    rec = np.random.default_rng(0).standard_normal((5, sp["nrx"], nt_sample)).astype(np.float32) * 1e-6
    t0 = time.perf_counter()
    O.rtm(p, wav, cm, rec, np.float32)
    dtm = time.perf_counter() - t0
    phys = 2.0 * 800 * 900 * nt_sample
    return phys / dtm / 1e6, dtm, sp


def load_traffic(variant, kernel):
    """Measured HBM bytes per cell-update of a kernel (profiles/traffic_<variant>.json:
    one entry per kernel form, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{variant}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get(kernel)


def acoustic3d(args):
    """3D acoustic 10th-order subdomain of one GPU (n = 512 interior + 5-cell
    halo per side), Born imaging with BF16 snapshots every 5th step: prints
    its own JSON line (not the headline metric)."""
    from elastic_amd import lib

    n = args.n3d + 10
    nt, dx, dt, fp, abc = args.nt3d, 10.0, 1e-3, 12.0, 40
    rng = np.random.default_rng(args.seed)
    z = np.arange(n, dtype=np.float32)[:, None, None]
    x = np.arange(n, dtype=np.float32)[None, :, None]
    y = np.arange(n, dtype=np.float32)[None, None, :]
    v = 1500.0 + 2500.0 * z / (n - 1)
    for _ in range(3):  # seeded dipping layers
        z0, sx, sy, dv = rng.uniform(0.2, 0.9) * n, rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2), rng.uniform(-150, 150)
        v = v + dv * (z > z0 + sx * (x - n / 2) + sy * (y - n / 2))
    v = np.clip(v, 1500.0, 4000.0).astype(np.float32)
    c_true = np.asfortranarray((v * dt / dx) ** 2)
    del v
    c_mig = np.asfortranarray(np.full_like(c_true, float(c_true.mean())))
    t = np.arange(1, nt + 1) * dt - 1.2 / fp
    amp = ((1 - 2 * (np.pi * fp * t) ** 2) * np.exp(-(np.pi * fp * t) ** 2)).astype(np.float32)
    src = (5 + 20, n // 2, n // 2)
    g = lib.Acoustic3D(c_mig, amp, src, 5 + 20, abc=abc, snap_every=5, snap_prec="bf16")
    del c_mig
    g.model_data(c_true)
    del c_true
    g.migrate()  # warm-up
    lib.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.migrate()
    lib.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    tm = g.timings()
    g.close()
    cells = float(n) ** 3 * nt
    fwd_gbs = 16.0 * cells / (tm[0] * 1e-3) / 1e9
    out = {"metric": "Mcell-updates/s, 3D acoustic 10th-order Born imaging (one GPU subdomain)",
           "value": round(2 * cells / (ms * 1e-3) / 1e6, 1), "unit": "Mcell-updates/s", "n_gpus": 1,
           "steps": args.steps, "ms_per_step": round(ms, 1), "higher_is_better": True, "dtype": "fp32",
           "data": "synthetic (seeded v(z) + 3 dipping layers; observed data modelled on the GPU)",
           "config": {"workload": f"3D acoustic {n}^3 ({args.n3d}^3 + 5-cell halos), dx=10 m, dt=1 ms, {nt} steps, "
                                  "12 Hz Ricker, 40-pt sponge, BF16 snapshots every 5th step"},
           "roofline": {"bound": "hbm", "kernel": "k_ac3_step (source pass)", "achieved": round(fwd_gbs, 1),
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(fwd_gbs / HBM_PEAK_GBS, 4),
                        "algo_bytes_per_cell": 16,
                        "other": {"source_pass_ms": round(tm[0], 1), "receiver_pass_ms": round(tm[2], 1)}}}
    print(json.dumps(out), flush=True)


def elastic3d(args):
    """9-field 3D elastic adjoint step (include/elastic3d.h) on the reference's
    200^3 test shape (K3/gpu.h) with its data distribution, nt3e steps with the
    reference's buffer rotation, device-resident: prints its own JSON line
    (not the headline metric)."""
    from elastic_amd import lib
    import oracle as O

    n = args.n3e
    # parity gate: one step of a 40 x 70 x 30 grid against FP64
    v = [O.el3_random(40, 70, 30, s) for s in (1, 2, 3)]
    vc = O.el3_random(40, 70, 30, 4, k=6)
    dv = [lib.DeviceArray(a) for a in v]
    got = lib.el3_steps_dev(dv, lib.DeviceArray(vc), 1)[2].get()
    err = O.rel_l2(got, O.el3_step(v[0], v[1], v[2], vc, nthreads=8))
    if not err <= 1e-5:
        raise SystemExit(f"bench: 3D elastic step differs from the FP64 oracle (rel-L2 {err:.3e})")
    del dv, got
    log(f"bench: 3D elastic parity vs FP64 rel-L2 {err:.2e}; building {n}^3 data")
    blocks = [O.el3_random(n, n, n, args.seed + s) for s in range(3)]
    coef = O.el3_random(n, n, n, args.seed + 3, k=6)
    dev = [lib.DeviceArray(b) for b in blocks]
    dc = lib.DeviceArray(coef)
    cpu = None
    if not args.no_cpu:
        nthr = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        O.el3_step(blocks[0], blocks[1], blocks[2], coef, dtype=np.float32, nthreads=1)
        t1 = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.el3_step(blocks[0], blocks[1], blocks[2], coef, dtype=np.float32, nthreads=nthr)
        tn = time.perf_counter() - t0
        cells = float(n) ** 3
        cpu = {"value": round(cells / tn / 1e6, 2), "unit": "Mcell-updates/s", "cores": nthr, "kind": "port",
               "sample": f"one step of the {n}^3 workload, oracle FP32 in the reference's term order "
                         f"(-O3, no FMA), OpenMP over y/x; 1 thread: {round(cells / t1 / 1e6, 2)} Mcell-updates/s"}
    del blocks, coef
    for _ in range(args.warmup):
        lib.el3_steps_dev(dev, dc, min(args.nt3e, 20))
    lib.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev = list(lib.el3_steps_dev(dev, dc, args.nt3e))
    lib.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    tm = lib.el3_timings()
    per_launch_ms = tm[0] / tm[1]
    cells = tm[2]
    gbs = 168.0 * cells / (per_launch_ms * 1e-3) / 1e9
    tr = load_traffic("elastic3d", "k_el3_adj")
    out = {"metric": "Mcell-updates/s, 3D elastic 9-field adjoint step (one GPU)",
           "value": round(cells * args.nt3e / (ms * 1e-3) / 1e6, 1), "unit": "Mcell-updates/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic (the reference's test distribution: (r % 1000)/500 - 1 inside a zero 5-cell rim)",
           "config": {"workload": f"3D elastic adjoint step, {n}^3 interior + 5-cell padding, 9 fields + 6 "
                                  f"coefficients, {args.nt3e} steps per run with the reference's buffer rotation",
                      "parity_rel_l2_fp64": float(f"{err:.3e}")},
           "roofline": {"bound": "hbm", "kernel": "k_el3_adj", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "traffic": None if tr is None else round(tr * cells),
                        "algo_bytes_per_cell": 168, "cells_per_launch": int(cells),
                        "ms_per_launch": round(per_launch_ms, 5), "planes_per_chunk": int(tm[3])},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def probe(args):
    """Streaming copy / read / write kernel of the library over a buffer of
    --probe-bytes, timed with HIP events: the attainable HBM rate, and the
    kernel of known traffic that calibrates the FETCH_SIZE / WRITE_SIZE
    counters (run it under rocprofv3 --pmc). Prints its own JSON line."""
    from elastic_amd import lib

    modes = {0: "copy", 1: "copy, non-temporal loads and stores", 2: "read only", 3: "write only"}
    gbs, ms = lib.bandwidth_probe(args.probe_bytes, args.probe_reps, args.probe_mode)
    moved = args.probe_bytes * (2 if args.probe_mode in (0, 1) else 1)
    out = {"metric": "streaming kernel bandwidth (bytes read + written per second)", "value": round(gbs, 1),
           "unit": "GB/s", "n_gpus": 1, "steps": args.probe_reps, "ms_per_step": round(ms, 4),
           "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"{modes[args.probe_mode]} of {args.probe_bytes} bytes",
                      "bytes_per_launch": moved},
           "roofline": {"bound": "hbm", "kernel": "k_copy", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": moved}}
    print(json.dumps(out), flush=True)


def run_variant(lib, synth, args, variant, steps, warmup, data=None, graph=True):
    """Build the survey's batch for one variant, put the observed data in
    place (given, or modelled on the GPU), run warmup + timed migrations.
    Returns (batch, specs, cmig, wav, seconds per migration, info)."""
    dx, dt, fp, abc = 10.0, 8e-4, 15.0, 40
    t0 = time.perf_counter()
    b, specs, cmig, wav, _ = build_batch(lib, synth, args.nz, args.nx, args.shots, args.nt, dx, dt, fp, abc,
                                         variant, args.seed, args.txr, graph=graph)
    log(f"bench: {variant}: setup {time.perf_counter() - t0:.1f}s (form {b.tblock}, tile width {b.txr}, "
        f"strip width {b.march_w})")
    t1 = time.perf_counter()
    if data is None:
        b.synthesize()
        log(f"bench: {variant}: data synthesis on the GPU {time.perf_counter() - t1:.1f}s")
    else:
        for i, d in enumerate(data):
            b.set_data(i, d)
        log(f"bench: {variant}: data upload {time.perf_counter() - t1:.1f}s")
    for w in range(warmup):
        t1 = time.perf_counter()
        b.migrate()
        log(f"bench: {variant}: warmup {w} {time.perf_counter() - t1:.2f}s")
    lib.synchronize()
    ts = time.perf_counter()
    info = None
    for k in range(steps):
        b.migrate()
        info = b.info()
        log(f"bench: {variant}: step {k} done")
    lib.synchronize()
    return b, specs, cmig, wav, (time.perf_counter() - ts) / max(steps, 1), info


def survey_with_layer(args, net):
    """The batched BF16 survey (64 shots, 800 x 2400, 6000 levels, BF16
    snapshots every 4th level, compensated images) timed with four
    boundaries: the 40-point sponge (45 padding cells), the 8-point band
    sponge alone, with the learned layer in the forward runs, and with the
    layer in the forward and the adjoint runs (13 cells; the network is the
    one trained at 400 x 800, used for both directions here: this measures
    cost, not image quality). One timed migration each, after the data
    synthesis."""
    from elastic_amd import lib, synth

    nz, nx, nshots, nt, dx, dt, fp = args.nz, args.nx, args.shots, args.nt, 10.0, 8e-4, 15.0
    res = {}
    for key, abc, use_net in (("sponge40", 40, 0), ("band_sponge", 8, 0), ("band_sponge_net", 8, 1),
                              ("band_sponge_net_both", 8, 2)):
        b, specs, _, _, _ = build_batch(lib, synth, nz, nx, nshots, nt, dx, dt, fp, abc, "bf16", args.seed,
                                        graph=not args.no_graph)
        try:
            if use_net:
                b.set_learned(net)
            if use_net == 2:
                b.set_learned_adjoint(net)
            b.synthesize()
            lib.synchronize()
            t0 = time.perf_counter()
            b.migrate()
            lib.synchronize()
            sec = time.perf_counter() - t0
            info = b.info()
            cells = sum(s["nz"] * s["nx"] for s in specs) * 2.0 * nt
            res[key] = {"seconds": round(sec, 2), "shots_per_hour": round(3600.0 * nshots / sec, 1),
                        "padding_cells": abc + 5, "cell_updates": cells, "fwd_ms": round(info["fwd_ms"], 1),
                        "adj_ms": round(info["adj_ms"], 1), "kernel_form": b.tblock}
            log(f"learned survey {key}: {res[key]}")
        finally:
            b.close()
    res["net_over_sponge40"] = round(res["band_sponge_net"]["shots_per_hour"] / res["sponge40"]["shots_per_hour"], 3)
    res["net_both_over_sponge40"] = round(res["band_sponge_net_both"]["shots_per_hour"] /
                                          res["sponge40"]["shots_per_hour"], 3)
    return res


def learned_cost(args):
    """Inference cost of the learned layer: the first --shots shots of the
    survey (one BF16 group, 13-cell band grids, 8-point band sponge, --nt
    levels; a network with random weights, which costs what a trained one
    does), one migration after the data synthesis without the layer, with
    it in the forward runs and with it in both directions, host-launched
    levels. Prints its own JSON line."""
    from elastic_amd import learned as LRN
    from elastic_amd import lib, synth

    net = LRN.Net(seed=123)
    res = {}
    try:
        for key, use_net in (("band_sponge", 0), ("band_sponge_net", 1), ("band_sponge_net_both", 2)):
            b, specs, _, _, _ = build_batch(lib, synth, args.nz, args.nx, args.shots, args.nt, 10.0, 8e-4, 15.0, 8,
                                            "bf16", args.seed, graph=False)
            try:
                if use_net:
                    b.set_learned(net)
                if use_net == 2:
                    b.set_learned_adjoint(net)
                b.synthesize()
                lib.synchronize()
                b.migrate()
                lib.synchronize()
                info = b.info()
                res[key] = {"fwd_ms": round(info["fwd_ms"], 2), "adj_ms": round(info["adj_ms"], 2),
                            "fwd_launches": int(info["fwd_launches"]), "adj_launches": int(info["adj_launches"]),
                            "group_size": int(info["group_size"])}
            finally:
                b.close()
    finally:
        net.close()
    per = (res["band_sponge_net"]["fwd_ms"] - res["band_sponge"]["fwd_ms"]) / res["band_sponge_net"]["fwd_launches"]
    res["adjoint_ms_per_launch"] = round((res["band_sponge_net_both"]["adj_ms"] - res["band_sponge"]["adj_ms"]) /
                                         res["band_sponge_net_both"]["adj_launches"], 4)
    line = {"metric": "learned-layer inference, ms per launch (forward-phase time with the layer minus without, "
                      "per level launch)", "value": round(per, 4), "unit": "ms", "n_gpus": 1, "steps": 1,
            "higher_is_better": False, "dtype": "fp32", "data": "synthetic (the survey's first shots)",
            "config": {"workload": f"learned-layer cost, {args.shots} shots of the {args.nz}x{args.nx} survey, "
                                   f"{args.nt} levels, BF16 snapshots every 4th level, 13-cell band grids"},
            "results": res}
    print(json.dumps(line), flush=True)


def learned(args):
    """The learned absorbing layer on the 400 x 800 configuration: training
    on wide-sponge (120-point) teachers of seeded random layered models,
    then boundary-reflection energy, receiver error and RTM image error on
    held-out models for the 8-point band sponge alone, with the network, and
    for the 40-point sponge, all against the wide-sponge runs; forward cost
    per level with and without the network. Prints its own JSON line and
    saves the trained parameters (--labc-out)."""
    from elastic_amd import learned as LRN

    t0 = time.perf_counter()
    out, net = LRN.experiment(nz=args.n3l[0], nx=args.n3l[1], nt=args.nt3l, steps=tuple(args.labc_steps),
                              levels_per_shot=args.labc_levels, log=log)
    if args.labc_out:
        np.save(args.labc_out, net.params)
    out["survey_bf16"] = survey_with_layer(args, net)
    net.close()
    img = out["image_rel_l2_vs_wide"]
    line = {"metric": "learned absorbing layer: reflection and image rel-L2 vs a 120-point sponge (value: the "
                      "image with the networks in both directions)",
            "value": img.get("band_sponge_net_both", img["band_sponge_net"]), "unit": "rel-L2", "n_gpus": 1,
            "steps": 1,
            "ms_per_step": round((time.perf_counter() - t0) * 1e3, 1), "higher_is_better": False, "dtype": "fp32",
            "data": "synthetic (seeded random layered models, Vp 1500-4500 m/s, 3-10 layers, Ricker 8-25 Hz)",
            "config": {"workload": f"learned layer, {args.n3l[0]}x{args.n3l[1]} grid, {args.nt3l} levels, 8-cell band "
                                   "+ 8-point band sponge vs 40-point sponge, teachers with a 120-point sponge"},
            "results": out}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--variant", default="batched", choices=sorted(VARIANTS))
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--nt", type=int, default=6000)
    ap.add_argument("--nz", type=int, default=800)
    ap.add_argument("--nx", type=int, default=2400)
    ap.add_argument("--txr", type=int, default=0)
    ap.add_argument("--seed", type=int, default=20260102)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-validate", action="store_true")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the plain-FP32 and BF16 runs of the same survey (one timed migration each)")
    ap.add_argument("--cpu-steps", type=int, default=400)
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every level from the host instead of replaying HIP graphs (same kernels; "
                         "for tracing tools that do not follow graph launches)")
    ap.add_argument("--workload", default="elastic2d",
                    choices=("elastic2d", "acoustic3d", "elastic3d", "probe", "learned", "learned-cost"),
                    help="acoustic3d: the 3D acoustic subdomain; elastic3d: the 9-field 3D elastic adjoint "
                         "step; probe: the streaming-kernel bandwidth probe; learned: training and evaluation "
                         "of the learned absorbing layer; learned-cost: the layer's inference cost on one BF16 "
                         "group of the survey; all instead of the headline 2D workload")
    ap.add_argument("--n3l", type=int, nargs=2, default=(400, 800))
    ap.add_argument("--nt3l", type=int, default=2500)
    ap.add_argument("--labc-steps", type=int, nargs=3, default=(4000, 2000, 2000))
    ap.add_argument("--labc-levels", type=int, default=200)
    ap.add_argument("--labc-out", default="")
    ap.add_argument("--n3d", type=int, default=512)
    ap.add_argument("--nt3d", type=int, default=1500)
    ap.add_argument("--n3e", type=int, default=200)
    ap.add_argument("--nt3e", type=int, default=500)
    ap.add_argument("--probe-bytes", type=int, default=1 << 30)
    ap.add_argument("--probe-reps", type=int, default=3)
    ap.add_argument("--probe-mode", type=int, default=0, choices=(0, 1, 2, 3))
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(shots are independent: a multi-GPU run would split them and all-reduce the images)", file=sys.stderr)
        sys.exit(2)

    from elastic_amd import lib, synth

    lib.load()
    if args.workload == "acoustic3d":
        acoustic3d(args)
        return
    if args.workload == "elastic3d":
        elastic3d(args)
        return
    if args.workload == "probe":
        probe(args)
        return
    if args.workload == "learned":
        learned(args)
        return
    if args.workload == "learned-cost":
        learned_cost(args)
        return
    dx, dt, fp, abc = 10.0, 8e-4, 15.0, 40
    log(f"bench: {args.shots}-shot survey {args.nz}x{args.nx}, nt={args.nt}, variant={args.variant}")
    b, specs, cmig, wav, sec_step, info = run_variant(lib, synth, args, args.variant, args.steps, args.warmup,
                                                      graph=not args.no_graph)
    ms_step = sec_step * 1e3
    phys = sum(2.0 * (s["img_nz"] * s["img_nx"]) * args.nt for s in specs)
    comp = info["fwd_cells"] + info["adj_cells"]
    value = phys / (ms_step * 1e-3) / 1e6
    shots_per_hour = 3600.0 * args.shots / (ms_step * 1e-3)
    fb, abytes = algo_bytes(args.variant)
    fwd_gbs = fb * info["fwd_cells"] / (info["fwd_ms"] * 1e-3) / 1e9
    adj_gbs = abytes * info["adj_cells"] / (info["adj_ms"] * 1e-3) / 1e9
    dom = "adj" if info["adj_ms"] >= info["fwd_ms"] else "fwd"
    achieved = adj_gbs if dom == "adj" else fwd_gbs
    launches = info["adj_launches"] if dom == "adj" else info["fwd_launches"]
    ms_launch = (info["adj_ms"] if dom == "adj" else info["fwd_ms"]) / launches
    # kernel form the batch runs (autotuned unless --txr / RTM_TXR / RTM_TBLOCK fix it)
    form, txr = b.tblock, b.txr
    mw = b.march_w if form >= 2 else 0
    kname = ("k_" + dom) + {0: "", 1: "2", 2: "m", 3: ("m" if dom == "fwd" else "")}[form]
    cells_per_launch = (info["adj_cells"] / info["adj_launches"]) if dom == "adj" else \
        (info["fwd_cells"] / info["fwd_launches"])
    # HBM bytes per launch of the dominant kernel: measured bytes per
    # cell-update (rocprofv3 FETCH_SIZE / WRITE_SIZE passes, calibrated on a
    # copy kernel) times this run's cell-updates per launch; marching figures
    # are kept per strip width, and a width without figures of its own
    # reports null
    tr = load_traffic(args.variant, kname + (f"_w{mw}" if kname.endswith("m") and mw else ""))
    traffic = round(tr["bytes_per_cell_update"] * cells_per_launch) if tr else None
    same_as_single = None
    stack_fp32 = None
    data = None
    sample = (b.get_data(0), b.image(0))
    if not args.no_compare and args.variant == "batched":
        stack_fp32 = b.stack((args.nz, args.nx))
        t1 = time.perf_counter()
        data = [b.get_data(i) for i in range(len(specs))]
        log(f"bench: observed data to host {time.perf_counter() - t1:.1f}s")
    b.close()  # the single-shot check needs the device memory the batch held
    if not args.no_validate:
        same_as_single = shot_matches_single(lib, sample, txr, specs, cmig, wav, args.nt, dx, dt, abc, args.variant)
        log(f"bench: shot 0 of the timed batch equals the single-shot tile path bit for bit: {same_as_single}")

    img_errs = None
    if not args.no_validate:
        t1 = time.perf_counter()
        img_errs = validate(lib, synth, args.variant, args.seed, txr)
        log(f"bench: validation {time.perf_counter() - t1:.1f}s, image rel-L2 vs FP64 {img_errs}")

    compare = None
    if data is not None:
        # the same survey and observed data through the plain FP32 GPU form
        # (one shot per launch, no graphs) and through the BF16 form
        # (snapshots every 4th level, compensated images), one timed
        # migration each
        import oracle as O

        compare = {}
        for v in ("plain", "bf16"):
            # the plain form replays no graph: the autotuner's short run has
            # loaded its kernels, so it needs no warm-up migration
            bv, _, _, _, sv, _ = run_variant(lib, synth, args, v, 1, 0 if v == "plain" else 1, data=data)
            st = bv.stack((args.nz, args.nx))
            ent = {"shots_per_hour": round(3600.0 * args.shots / sv, 1), "seconds": round(sv, 2),
                   "mcell_updates_per_s": round(phys / sv / 1e6, 1),
                   "stack_rel_l2_vs_batched_fp32": [float(f"{O.rel_l2(a, r):.3e}") for a, r in zip(st, stack_fp32)],
                   "kernel_form": bv.tblock, "tile_width": bv.txr}
            vtxr, vsample = bv.txr, (data[0], bv.image(0))
            bv.close()
            if not args.no_validate:
                ent["single_shot_bitwise"] = shot_matches_single(lib, vsample, vtxr, specs, cmig, wav, args.nt, dx, dt,
                                                                 abc, v)
                if v == "bf16":
                    ev = validate(lib, synth, v, args.seed, vtxr)
                    ent["subrun_rel_l2_vs_fp64"] = [[float(f"{x:.3e}") for x in e] for e in ev]
            compare[v] = ent
            log(f"bench: {v}: {ent}")
        del data
        compare["batched"] = {"shots_per_hour": round(shots_per_hour, 1), "seconds": round(sec_step, 2)}
        compare["bf16_over_plain"] = round(compare["bf16"]["shots_per_hour"] / compare["plain"]["shots_per_hour"], 3)
        compare["batched_over_plain"] = round(shots_per_hour / compare["plain"]["shots_per_hour"], 3)

    cpu = None
    if not args.no_cpu:
        hi = host_info()
        nth = hi["threads"]
        t1 = time.perf_counter()
        mc, secs, _ = cpu_baseline(lib, synth, args.seed, dx, dt, fp, abc, args.cpu_steps, nth)
        mc1, secs1, _ = cpu_baseline(lib, synth, args.seed, dx, dt, fp, abc, max(4, args.cpu_steps // 4), 1)
        log(f"bench: cpu baseline {time.perf_counter() - t1:.1f}s")
        cpu = {"value": round(mc, 3), "unit": "Mcell-updates/s", "cores": nth, "kind": "port",
               "sample": f"oracle FP32 RTM (forward+adjoint+imaging, reference operation order, -O3, no FMA, "
                         f"OpenMP over x) of one 800x900 shot (890x990 padded), {args.cpu_steps} levels, "
                         f"{nth} threads, {secs:.1f}s",
               "single_thread": {"value": round(mc1, 3), "cores": 1, "seconds": round(secs1, 2),
                                 "levels": max(4, args.cpu_steps // 4)},
               "host": hi}
    max_err = max(max(e) for e in img_errs) if img_errs else None
    out = {
        "metric": "Mcell-updates/s & shots/hour, 2D elastic RTM at <=1e-3 image rel-L2; % HBM BW",
        "value": round(value, 1),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded layered+anomaly model; observed data modelled on the GPU)",
        "config": {"workload": f"batched 2D elastic RTM, {args.shots} shots, model {args.nz}x{args.nx}, dx=10 m, "
                               f"dt=0.8 ms, {args.nt} levels, 15 Hz Ricker, apertures 600-1200 columns, 40-pt sponge",
                   "variant": args.variant, "shots": args.shots, "nt": args.nt,
                   "shots_per_group": int(info["group_size"]),
                   "kernel_form": {0: "tiles, one level per launch", 1: "tiles, two levels per launch",
                                   2: "strips marching along x",
                                   3: "strips marching along x (forward), tiles (adjoint)"}[form],
                   "tile_width": txr, "strip_width": mw or None},
        "shots_per_hour": round(shots_per_hour, 1),
        "mcells_computational": round(comp / (ms_step * 1e-3) / 1e6, 1),
        "image_rel_l2_vs_fp64": max_err,
        "timed_shot_equals_validated_path": same_as_single,
        "valid": bool(max_err is not None and max_err <= 1e-3 and same_as_single),
        "variants": compare,
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_per_cell_update": tr["bytes_per_cell_update"] if traffic else None,
                     "ms_per_launch": round(ms_launch, 5),
                     "cells_per_launch": round(cells_per_launch),
                     "algo_bytes_per_cell": fb if dom == "fwd" else abytes,
                     "other": {"k_fwd_GBs": round(fwd_gbs, 1), "k_adj_GBs": round(adj_gbs, 1),
                               "fwd_ms": round(info["fwd_ms"], 2), "adj_ms": round(info["adj_ms"], 2)}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
