"""The benchmark configurations at their own sizes on the GPU.

* Plain single-shot RTM on the 300 x 600 model (390 x 690 padded), 2000
  levels, 40-point sponge, the drop-in's adjoint without sponge: observed
  data (true minus smooth model) and the dlambda / dmu / drho images against
  the FP64 oracle (data <= 1e-5, images <= 1e-4).
* Batched RTM of 64 shots on the 800 x 2400 model, 6000 levels, apertures of
  600-1200 columns, strips marching along x with widths fitted per group and
  kernel: size-independent properties (finite and nonzero images, two shots
  bit for bit equal to the single-shot tile path, the stack equal to the sum
  of the shot windows, linearity in the data).
* Mixed precision on the 200 x 400 sub-run: BF16 and FP16 snapshots every
  4th level with compensated images against the FP64 oracle of the same
  decimation (<= 1e-3), and against the FP32 plain run (reported).
* 3D acoustic subdomain at 522^3: receiver data of 40 steps against FP64
  (<= 1e-5); the full 1500-step Born imaging with BF16 snapshots: finite,
  nonzero, and linear in the data.
"""
import os

import numpy as np
import pytest

import oracle as O
from elastic_amd import synth

pytestmark = pytest.mark.gpu


def _threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, n))


# --------------------------------------------------------------------------- plain single shot
def test_plain_rtm_300x600_2000_levels_vs_fp64(hip):
    c = synth.config0()
    geo, nt, dx, dt, abc = c["geo"], c["nt"], c["dx"], c["dt"], c["abc"]
    ct = [np.asfortranarray(a) for a in O.model_op(c["true"].rho, c["true"].lam, c["true"].mu, dx, dt, np.float32)]
    cm = [np.asfortranarray(a) for a in O.model_op(c["mig"].rho, c["mig"].lam, c["mig"].mu, dx, dt, np.float32)]
    w = O.ricker(nt, dt, c["fp"], dtype=np.float32)
    s = hip.shot(geo.nz, geo.nx, nt, dx, dt, geo.xs, geo.zs, geo.zr, geo.rx0, geo.nrx, abc)
    print("plain RTM: GPU data and images", flush=True)
    d = hip.forward(s, ct, w) - hip.forward(s, cm, w)
    img = hip.migrate(s, cm, w, d)
    p = O.make_params(geo.nz, geo.nx, nt, dx, dt, geo.xs, geo.zs, geo.zr, geo.rx0, geo.nrx, abc, nthreads=_threads())
    print("plain RTM: FP64 oracle", flush=True)
    d64 = O.forward(p, w, ct, np.float64)["rec"] - O.forward(p, w, cm, np.float64)["rec"]
    assert O.rel_l2(d, d64) <= 1e-5
    ref = O.rtm(p, w, cm, d, np.float64)
    sl = (slice(geo.pad, geo.pad + 300), slice(geo.pad, geo.pad + 600))
    errs = [O.rel_l2(a[sl], r[sl]) for a, r in zip(img, ref)]
    print("plain RTM image rel-L2 vs FP64 (drho, dlambda, dmu):", errs, flush=True)
    for r in ref:
        assert np.abs(r[sl]).max() > 0
    assert max(errs) <= 1e-4


# --------------------------------------------------------------------------- 64-shot batch
def _survey(nz, nx, nshots, nt, dx, dt, abc, seed=20260102):
    vp, vps, layout = synth.survey(nz, nx, nshots, seed, abc, dx, dt)
    specs, ctrue, cmig = [], [], []
    for xs, x0, w in layout:
        sp, ct, cm = synth.window_shot(vp, vps, x0, w, xs, abc, dx, dt)
        specs.append(sp)
        ctrue.append(ct)
        cmig.append(cm)
    return specs, ctrue, cmig


def _single(hip, sp, cm, w, rec, nt, dx, dt, abc, monkeypatch, **kw):
    s = hip.shot(sp["nz"], sp["nx"], nt, dx, dt, sp["xs"], sp["zs"], sp["zr"], sp["rx0"], sp["nrx"], abc, adj_abc=1,
                 **kw)
    with monkeypatch.context() as m:
        m.setenv("RTM_TBLOCK", "0")
        m.setenv("RTM_TXR", "2")
        return hip.migrate(s, cm, w, rec)


def test_batched_64_shots_800x2400_6000_levels_properties(hip, monkeypatch):
    nz, nx, nshots, nt, dx, dt, fp, abc = 800, 2400, 64, 6000, 10.0, 8e-4, 15.0, 40
    monkeypatch.setenv("RTM_TBLOCK", "2")  # marching strips, widths fitted per group and kernel
    monkeypatch.delenv("RTM_MARCH_W", raising=False)
    monkeypatch.delenv("RTM_MARCH_FIT", raising=False)
    specs, ctrue, cmig = _survey(nz, nx, nshots, nt, dx, dt, abc)
    w = hip.ricker(nt, dt, fp)
    b = hip.Batch(nt, dx, dt, specs, cmig, w, abc=abc, adj_abc=1, stack_shape=(nz, nx), txr=2, use_graph=1)
    try:
        print(f"batch: form {b.tblock}, first group's strip width {b.march_w}", flush=True)
        assert b.tblock == 2 and b.march_w >= 16
        b.set_true_model(ctrue)
        b.synthesize()
        print("batch: migrate", flush=True)
        b.migrate()
        stack = b.stack((nz, nx))
        acc = [np.zeros((nz, nx)) for _ in range(3)]
        for i, sp in enumerate(specs):
            img = b.image(i)
            z0, x0, x1 = sp["img_z0"], sp["img_x0"], sp["img_x0"] + sp["img_nx"]
            for k in range(3):
                a = img[k][z0:z0 + sp["img_nz"], x0:x1]
                assert np.isfinite(a).all() and np.abs(a).max() > 0, (i, k)
                acc[k][:, sp["stack_x0"]:sp["stack_x0"] + sp["img_nx"]] += a
        for k in range(3):
            assert O.rel_l2(stack[k], acc[k]) <= 1e-7, k
        sample = {i: (b.get_data(i), b.image(i)) for i in (0, 37)}
    finally:
        b.close()  # the single-shot runs below need the device memory the batch held
    for i, (rec, got) in sample.items():
        print(f"batch: shot {i} against the single-shot tile path", flush=True)
        one = _single(hip, specs[i], cmig[i], w, rec, nt, dx, dt, abc, monkeypatch)
        for k in range(3):
            assert np.array_equal(got[k], one[k]), (i, k)
    # linearity in the data: doubling every sample doubles the image
    rec, got = sample[0]
    two = _single(hip, specs[0], cmig[0], w, 2.0 * rec, nt, dx, dt, abc, monkeypatch)
    for k in range(3):
        # exact in binary FP except where the far-field precursor is subnormal
        assert O.rel_l2(two[k], 2.0 * got[k]) <= 1e-6, k


# --------------------------------------------------------------------------- mixed precision
def test_mixed_precision_200x400_subrun(hip):
    nz, nx, nt, dx, dt, fp, abc = 200, 400, 1000, 10.0, 8e-4, 15.0, 40
    specs, ctrue, cmig = _survey(nz, nx, 2, nt, dx, dt, abc, seed=20260109)
    w = hip.ricker(nt, dt, fp)
    runs = {}
    data = None
    for prec, j, comp in (("fp32", 1, 0), ("bf16", 4, 1), ("fp16", 4, 1)):
        b = hip.Batch(nt, dx, dt, specs, cmig, w, abc=abc, adj_abc=1, stack_shape=(nz, nx), txr=2, use_graph=1,
                      snap_every=j, snap_prec=prec, comp_image=comp)
        if data is None:
            b.set_true_model(ctrue)
            b.synthesize()
            data = [b.get_data(i) for i in range(len(specs))]
        else:
            for i, d in enumerate(data):
                b.set_data(i, d)
        b.migrate()
        runs[prec] = [b.image(i) for i in range(len(specs))]
        b.close()
    for i, sp in enumerate(specs):
        z0, x0 = sp["img_z0"], sp["img_x0"]
        sl = (slice(z0, z0 + sp["img_nz"]), slice(x0, x0 + sp["img_nx"]))
        p4 = O.make_params(sp["nz"], sp["nx"], nt, dx, dt, sp["xs"], sp["zs"], sp["zr"], sp["rx0"], sp["nrx"], abc,
                           adj_abc=1, snap_every=4, nthreads=_threads())
        ref4 = O.rtm(p4, w, cmig[i], data[i], np.float64)
        for prec in ("bf16", "fp16"):
            e64 = [O.rel_l2(a[sl], r[sl]) for a, r in zip(runs[prec][i], ref4)]
            e32 = [O.rel_l2(a[sl], r[sl]) for a, r in zip(runs[prec][i], runs["fp32"][i])]
            print(f"shot {i} {prec}: rel-L2 vs FP64 (j=4) {e64}, vs FP32 plain (j=1) {e32}", flush=True)
            assert max(e64) <= 1e-3
            # the FP32 run images every level: the difference is the
            # decimation error (about 1e-3 here) plus the storage precision
            assert max(e32) <= 3e-3


# --------------------------------------------------------------------------- 3D acoustic
def _ac3_model(n, dt=1e-3, dx=10.0, seed=20260102):
    rng = np.random.default_rng(seed)
    z = np.arange(n, dtype=np.float32)[:, None, None]
    x = np.arange(n, dtype=np.float32)[None, :, None]
    y = np.arange(n, dtype=np.float32)[None, None, :]
    v = 1500.0 + 2500.0 * z / (n - 1)
    for _ in range(3):
        z0, sx, sy, dv = rng.uniform(0.2, 0.9) * n, rng.uniform(-0.2, 0.2), rng.uniform(-0.2, 0.2), rng.uniform(-150, 150)
        v = v + dv * (z > z0 + sx * (x - n / 2) + sy * (y - n / 2))
    v = np.clip(v, 1500.0, 4000.0).astype(np.float32)
    return np.asfortranarray((v * dt / dx) ** 2)


def _ricker(nt, fp=12.0, dt=1e-3):
    t = np.arange(1, nt + 1) * dt - 1.2 / fp
    return ((1 - 2 * (np.pi * fp * t) ** 2) * np.exp(-(np.pi * fp * t) ** 2)).astype(np.float32)


def test_acoustic3d_522_cubed(hip):
    n, abc = 522, 40
    c = _ac3_model(n)
    src, zr = (25, n // 2, n // 2), 25
    nt_short = 40
    print("3D: 40 steps against FP64", flush=True)
    amp = _ricker(nt_short)
    g = hip.Acoustic3D(c, amp, src, zr, abc=abc, snap_every=5, snap_prec="bf16")
    g.model_data()
    rec = g.data()
    g.close()
    ref, _ = O.ac3_run(c, amp, src, zr, abc=abc, nthreads=_threads())
    assert np.abs(ref).max() > 0
    assert O.rel_l2(rec, ref) <= 1e-5
    del ref
    print("3D: 1500-step Born imaging, BF16 snapshots", flush=True)
    nt = 1500
    cm = np.asfortranarray(np.full_like(c, float(c.mean())))
    g = hip.Acoustic3D(cm, _ricker(nt), src, zr, abc=abc, snap_every=5, snap_prec="bf16")
    try:
        g.model_data(c_true=c)
        d = g.data()
        assert np.isfinite(d).all() and np.abs(d).max() > 0
        g.migrate()
        img = g.image()
        assert np.isfinite(img).all() and np.abs(img).max() > 0
        g.set_data(2.0 * d)
        g.migrate()
        img2 = g.image()
        assert O.rel_l2(img2, 2.0 * img) <= 1e-6
    finally:
        g.close()
