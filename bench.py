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
computational-grid rate are reported beside it. The run is valid only when the
images of a 200 x 400 sub-run through the same engine are within 1e-3
relative L2 of the FP64 oracle; that check runs in the same process.
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
        fwd, adj = 80 + 8, 80 + 8 + 24  # Vx, Vz history write / read; 3 images RMW
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


def validate(lib, synth, variant, seed):
    """Sub-run on a 200 x 400 model (2 shots, 1000 levels) through the same
    engine configuration, images against the FP64 oracle."""
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
    b = lib.Batch(nt, dx, dt, specs, cmig, wav, abc=abc, adj_abc=1, stack_shape=(nz, nx), **opts)
    b.set_true_model(ctrue)
    b.synthesize()
    b.migrate()
    errs = []
    nth = min(16, os.cpu_count() or 1)
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
    ap.add_argument("--cpu-steps", type=int, default=600)
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every level from the host instead of replaying HIP graphs (same kernels; "
                         "for tracing tools that do not follow graph launches)")
    ap.add_argument("--workload", default="elastic2d", choices=("elastic2d", "acoustic3d", "elastic3d"),
                    help="acoustic3d: the 3D acoustic subdomain; elastic3d: the 9-field 3D elastic adjoint "
                         "step; both instead of the headline 2D workload")
    ap.add_argument("--n3d", type=int, default=512)
    ap.add_argument("--nt3d", type=int, default=1500)
    ap.add_argument("--n3e", type=int, default=200)
    ap.add_argument("--nt3e", type=int, default=500)
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
    dx, dt, fp, abc = 10.0, 8e-4, 15.0, 40
    t0 = time.perf_counter()
    log(f"bench: building {args.shots}-shot survey {args.nz}x{args.nx}, nt={args.nt}, variant={args.variant}")
    b, specs, cmig, wav, layout = build_batch(lib, synth, args.nz, args.nx, args.shots, args.nt, dx, dt, fp, abc,
                                              args.variant, args.seed, args.txr, graph=not args.no_graph)
    log(f"bench: setup {time.perf_counter() - t0:.1f}s; synthesizing observed data on the GPU")
    t1 = time.perf_counter()
    b.synthesize()
    log(f"bench: data synthesis {time.perf_counter() - t1:.1f}s")
    for w in range(args.warmup):
        t1 = time.perf_counter()
        b.migrate()
        log(f"bench: warmup {w} {time.perf_counter() - t1:.2f}s")
    lib.synchronize()
    ts = time.perf_counter()
    infos = []
    for k in range(args.steps):
        b.migrate()
        infos.append(b.info())
        log(f"bench: step {k} done")
    lib.synchronize()
    elapsed = time.perf_counter() - ts
    ms_step = elapsed * 1e3 / args.steps
    info = infos[-1]
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
    # HBM bytes per launch of the dominant kernel: measured bytes per
    # cell-update (rocprofv3 FETCH_SIZE / WRITE_SIZE passes, calibrated on a
    # copy kernel) times this run's cell-updates per launch
    # kernel form the batch runs (autotuned unless --txr / RTM_TXR / RTM_TBLOCK fix it)
    form, txr = b.tblock, b.txr
    mw = b.march_w if form >= 2 else 0
    kname = ("k_" + dom) + {0: "", 1: "2", 2: "m", 3: ("m" if dom == "fwd" else "")}[form]
    cells_per_launch = (info["adj_cells"] / info["adj_launches"]) if dom == "adj" else \
        (info["fwd_cells"] / info["fwd_launches"])
    # marching figures are measured per strip width (the 16-column ones
    # unsuffixed); a width without its own figures takes the nearest measured
    tr = None
    if kname.endswith("m") and mw:
        for w in sorted((16, 24, 26, 28, 32), key=lambda w: abs(w - mw)):
            tr = load_traffic(args.variant, kname + ("" if w == 16 else f"_w{w}"))
            if tr:
                tr = dict(tr, strip_width=w)
                break
    else:
        tr = load_traffic(args.variant, kname)
    traffic = round(tr["bytes_per_cell_update"] * cells_per_launch) if tr else None
    b.close()

    img_errs = None
    if not args.no_validate:
        t1 = time.perf_counter()
        img_errs = validate(lib, synth, args.variant, args.seed)
        log(f"bench: validation {time.perf_counter() - t1:.1f}s, image rel-L2 vs FP64 {img_errs}")
    cpu = None
    if not args.no_cpu:
        nth = min(16, os.cpu_count() or 1)
        t1 = time.perf_counter()
        mc, secs, _ = cpu_baseline(lib, synth, args.seed, dx, dt, fp, abc, args.cpu_steps, nth)
        mc1, secs1, _ = cpu_baseline(lib, synth, args.seed, dx, dt, fp, abc, max(4, args.cpu_steps // 4), 1)
        log(f"bench: cpu baseline {time.perf_counter() - t1:.1f}s")
        cpu = {"value": round(mc, 3), "unit": "Mcell-updates/s", "cores": nth, "kind": "port",
               "sample": f"oracle FP32 RTM (forward+adjoint+imaging) of one 800x900 shot (890x990 padded), "
                         f"{args.cpu_steps} levels, {nth} threads, {secs:.1f}s",
               "single_thread": {"value": round(mc1, 3), "cores": 1, "seconds": round(secs1, 2)}}
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
        "valid": (max_err is not None and max_err <= 1e-3),
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_per_cell_update": tr["bytes_per_cell_update"] if traffic else None,
                     "traffic_strip_width": tr.get("strip_width") if traffic else None,
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
