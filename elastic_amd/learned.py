"""Learned absorbing layer: ctypes binding of include/learned_abc.h and the
host-side training / evaluation driver.

The layer replaces the reference's 40-point sponge (W/abc-mod.f90:59-78,
applied after every field update in W/rtm-mod-adj.f90:225-252) by an 8-cell
band inside the 5-cell rim. For every band cell a small network predicts the
part of the next level that the plain stencil cannot see (the wavefield that
would continue outside the grid). Training pairs come from runs of the same
model with a wide (120-point) sponge far outside the band: at a level L the
band grid's levels L-1 and L-2 are windows of the wide run, and the target is
the wide run's level L minus the band grid's plain update. Everything runs on
the GPU (collection, training with Adam, inference during propagation).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _lib
from .lib import _chk, _p, pack_fields, unpack_fields

RIM, WIDTH, NFEAT, KPAD, HID, NOUT = 5, 8, 245, 248, 32, 5
NPARAM = NFEAT * HID + HID + HID * HID + HID + HID * NOUT + NOUT
Z0 = 6.0e6
PAD = RIM + WIDTH  # padding of a band grid around the physical region


def _L():
    lib = _lib.load()
    if not getattr(lib, "_labc_typed", False):
        lib.labc_last_forward_ms.restype = C.c_double
        lib._labc_typed = True
    return lib


def cell_count(nz, nx):
    return int(_L().labc_cell_count(C.c_int(nz), C.c_int(nx)))


def cell_coords(nz, nx, c):
    iz, ix, e, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _chk(_L().labc_cell_coords(C.c_int(nz), C.c_int(nx), C.c_int(c), C.byref(iz), C.byref(ix), C.byref(e), C.byref(d)))
    return iz.value, ix.value, e.value, d.value


def features(lev_m1, lev_m2, coef5, adjoint=False):
    """Normalised bf16 features (ncell, KPAD) and scales of the band cells
    of one grid from levels L-1 and L-2 and the grid's coefficients (five
    (nz, nx) arrays each); with adjoint=True from levels L+1 and L+2 of an
    adjoint state."""
    nz, nx = np.asarray(lev_m1[0]).shape
    nc = cell_count(nz, nx)
    a, b, c = pack_fields(lev_m1), pack_fields(lev_m2), pack_fields(coef5)
    f = np.zeros((nc, KPAD), np.float32)
    s = np.zeros(nc, np.float32)
    fn = _L().labc_features_adj if adjoint else _L().labc_features
    _chk(fn(C.c_int(nz), C.c_int(nx), _p(a), _p(b), _p(c), _p(f), _p(s)))
    return f, s


class Net:
    """The 245 -> 32 -> 32 -> 5 GELU network (FP32 master weights, BF16 use)."""

    def __init__(self, seed=1):
        self.h = C.c_void_p()
        _chk(_L().labc_net_create(C.byref(self.h), C.c_uint(seed)))

    def close(self):
        if self.h:
            _L().labc_net_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(NPARAM, np.float32)
        _chk(_L().labc_net_get_params(self.h, _p(p)))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == NPARAM
        _chk(_L().labc_net_set_params(self.h, _p(p)))

    def forward(self, feat, mode=0):
        feat = np.ascontiguousarray(feat, np.float32)
        out = np.zeros((feat.shape[0], NOUT), np.float32)
        _chk(_L().labc_net_forward(self.h, _p(feat), C.c_int(feat.shape[0]), _p(out), C.c_int(mode)))
        return out

    def grad(self, feat, target):
        feat = np.ascontiguousarray(feat, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        g = np.zeros(NPARAM, np.float32)
        loss = C.c_float()
        _chk(_L().labc_net_grad(self.h, _p(feat), _p(t), C.c_int(feat.shape[0]), _p(g), C.byref(loss)))
        return loss.value, g

    def train(self, data: "Data", steps, batch=4096, lr=1e-3, seed=1):
        losses = np.zeros(steps, np.float32)
        _chk(_L().labc_train(self.h, data.h, C.c_int(steps), C.c_int(batch), C.c_float(lr), C.c_uint(seed),
                             _p(losses)))
        return losses

    def eval(self, data: "Data"):
        loss = C.c_float()
        _chk(_L().labc_eval(self.h, data.h, C.byref(loss)))
        return loss.value


class Data:
    """Training pairs resident on the device."""

    def __init__(self, capacity):
        self.h = C.c_void_p()
        _chk(_L().labc_data_create(C.byref(self.h), C.c_int(capacity)))

    def close(self):
        if self.h:
            _L().labc_data_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return int(_L().labc_data_size(self.h))

    def get(self, row0=0, n=None):
        n = len(self) - row0 if n is None else n
        f = np.zeros((n, KPAD), np.float32)
        t = np.zeros((n, NOUT), np.float32)
        _chk(_L().labc_data_get(self.h, C.c_int(row0), C.c_int(n), _p(f), _p(t)))
        return f, t

    def collect(self, nz_phys, nx_phys, pad_wide, abc_wide, nt, dt, zs_phys, xs_phys, coefs_wide, coefs_small,
                wavelet, levels, noise=0.0, seed=0, abc_band=0):
        """Teacher-forced pairs of one wide-sponge run (see labc_collect)."""
        cw, cs = pack_fields(coefs_wide), pack_fields(coefs_small)
        w = np.ascontiguousarray(wavelet, np.float32)
        lv = np.ascontiguousarray(levels, np.int32)
        _chk(_L().labc_collect(self.h, C.c_int(nz_phys), C.c_int(nx_phys), C.c_int(pad_wide), C.c_int(abc_wide),
                               C.c_int(nt), C.c_float(dt), C.c_int(zs_phys), C.c_int(xs_phys), _p(cw), _p(cs),
                               _p(w), _p(lv), C.c_int(lv.size), C.c_float(noise), C.c_uint(seed),
                               C.c_int(abc_band)))

    def collect_adj(self, nz_phys, nx_phys, pad_wide, abc_wide, nt, dt, zs_phys, xs_phys, zr_phys, coefs_wide,
                    coefs_small, wavelet, levels, noise=0.0, seed=0, abc_band=0, coefs_wide_obs=None):
        """Teacher-forced pairs of the adjoint propagation of one wide-sponge
        run's receiver data, or with coefs_wide_obs of the data of that model
        minus those of coefs_wide (see labc_collect_adj)."""
        cw, cs = pack_fields(coefs_wide), pack_fields(coefs_small)
        co = None if coefs_wide_obs is None else pack_fields(coefs_wide_obs)
        w = np.ascontiguousarray(wavelet, np.float32)
        lv = np.ascontiguousarray(levels, np.int32)
        _chk(_L().labc_collect_adj(self.h, C.c_int(nz_phys), C.c_int(nx_phys), C.c_int(pad_wide),
                                   C.c_int(abc_wide), C.c_int(nt), C.c_float(dt), C.c_int(zs_phys),
                                   C.c_int(xs_phys), C.c_int(zr_phys), _p(cw), _p(cs), _p(w), _p(lv),
                                   C.c_int(lv.size), C.c_float(noise), C.c_uint(seed), C.c_int(abc_band),
                                   _p(co)))


def propagate(net, s: _lib.RtmShot, coef5, wavelet, want_last=False):
    """Forward propagation of a band grid with the learned layer (net, and
    the sponge s.abc when > 0) or, with net None, with the sponge s.abc alone.
    Returns receivers (5, nrx, nt) [and the last two levels] and the device
    time of the time loop (ms)."""
    lib = _L()
    cb = pack_fields(coef5)
    w = np.ascontiguousarray(wavelet, np.float32)
    rec = np.zeros((5, max(s.nrx, 0), s.nt), np.float32)
    last = np.zeros((10, s.nx, s.nz), np.float32) if want_last else None
    _chk(lib.labc_forward(net.h if net is not None else None, C.byref(s), _p(cb), _p(w), _p(rec), _p(last)))
    ms = float(lib.labc_last_forward_ms())
    if want_last:
        return rec, unpack_fields(last), ms
    return rec, ms


# --------------------------------------------------------------------------- experiment
def random_layered_vp(nz, nx, rng, vmin=1500.0, vmax=4500.0, nmin=3, nmax=10):
    """Seeded random layered Vp model: nmin..nmax interfaces at random depths
    with small random dips, each layer's velocity uniform in [vmin, vmax]."""
    n = int(rng.integers(nmin, nmax + 1))
    z = np.arange(nz, dtype=np.float64)[:, None]
    x = np.arange(nx, dtype=np.float64)[None, :]
    depths = np.sort(rng.uniform(0.05 * nz, 0.95 * nz, n))
    dips = rng.uniform(-0.08, 0.08, n)
    vels = rng.uniform(vmin, vmax, n + 1)
    vp = np.full((nz, nx), vels[0])
    for d, s, v in zip(depths, dips, vels[1:]):
        vp = np.where(z >= d + s * (x - nx / 2), v, vp)
    return np.asfortranarray(vp)


def _coefs(vp, pad, dx, dt):
    from . import synth

    m = synth.elastic_from_vp(synth.pad_edge(vp, pad))
    return _lib.model_op(m.rho, m.lam, m.mu, dx, dt)


def _rel(a, b):
    den = float(np.sqrt(sum(np.sum(np.asarray(y, np.float64) ** 2) for y in b)))
    num = float(np.sqrt(sum(np.sum((np.asarray(x, np.float64) - np.asarray(y, np.float64)) ** 2)
                            for x, y in zip(a, b))))
    return num / den if den > 0 else 0.0


def experiment(nz=400, nx=800, dx=10.0, dt=1e-3, nt=2500, abc_wide=120, band_abc=8, train_models=5,
               shots_per_model=2, levels_per_shot=200, test_models=2, steps=(4000, 2000, 2000), batch=8192,
               noise=0.03, seed=20260301, adjoint_layer=True, adjoint_residual=False, log=None):
    """The learned layer end to end on nz x nx grids (all on the GPU):

    * training: `train_models` seeded random layered models (Vp 1500-4500
      m/s, 3-10 layers) with `shots_per_model` shots each (Ricker 8-25 Hz,
      random positions), teacher-forced pairs from runs with an `abc_wide`
      sponge placed outside the band grid (`levels_per_shot` levels each:
      train_models * shots_per_model * levels_per_shot snapshot pairs),
      noise augmentation, Adam in three learning-rate stages; with
      adjoint_layer, a second network trained the same way on the adjoint
      runs of the same shots' receiver data (labc_collect_adj; with
      adjoint_residual, through the smoothed model injecting the residual
      data of a migration instead);
    * reflection energy: on held-out models, the physical-region wavefield
      of the last two levels and the receiver data of an 8-cell band grid
      with the band sponge alone, with the sponge and the network, and of a
      40-point-sponge grid, each relative to the wide-sponge run;
    * images: RTM (exact-adjoint sponge) of the held-out shots through the
      batch engine with each boundary treatment, the observed data the wide
      grid's (true minus smoothed model), relative L2 of the physical-region
      images against the wide-sponge RTM (with adjoint_layer also with the
      adjoint network alone and with both networks);
    * cost: device time per forward level with and without the layer.
    Returns (metrics dict, trained Net)."""
    from . import synth

    say = log or (lambda *a: None)
    rng = np.random.default_rng(seed)
    pad_w = abc_wide + 5 + PAD
    nzb, nxb = nz + 2 * PAD, nx + 2 * PAD
    ncell = cell_count(nzb, nxb)
    nshot_train = train_models * shots_per_model
    d = Data(ncell * levels_per_shot * nshot_train)
    t0 = _time()
    zr = 2  # receiver depth below the band, in the training and held-out shots
    train_shots = []
    for mi in range(train_models):
        vp = random_layered_vp(nz, nx, rng)
        cw, cs = _coefs(vp, pad_w, dx, dt), _coefs(vp, PAD, dx, dt)
        # the migration model of the adjoint teachers (smoothed, as in the
        # held-out RTM below)
        vps = synth.smooth(vp, 10.0) if adjoint_residual else None
        cws = _coefs(vps, pad_w, dx, dt) if adjoint_residual else cw
        css = _coefs(vps, PAD, dx, dt) if adjoint_residual else cs
        for si in range(shots_per_model):
            fp = float(rng.uniform(8.0, 25.0))
            zs, xs = int(rng.integers(2, nz // 3)), int(rng.integers(0, nx))
            wav = _lib.ricker(nt, dt, fp)
            lv = np.linspace(150, nt, levels_per_shot).round().astype(int)
            d.collect(nz, nx, pad_w, abc_wide, nt, dt, zs, xs, cw, cs, wav, lv, noise=noise,
                      seed=int(seed + 31 * mi + si), abc_band=band_abc)
            if adjoint_layer:
                train_shots.append((cw, cws, css, zs, xs, wav, int(seed + 31 * mi + si)))
    say(f"learned: {len(d)} training rows from {nshot_train * levels_per_shot} snapshot pairs, "
        f"{_time() - t0:.1f}s")

    def fit(data, net_seed):
        n = Net(seed=net_seed)
        a = n.eval(data)
        t = _time()
        for lr, k in zip((2e-3, 5e-4, 1e-4), steps):
            n.train(data, k, batch=batch, lr=lr, seed=int(lr * 1e6) + seed)
        return n, a, n.eval(data), _time() - t

    net, l0, l1, train_s = fit(d, int(seed % 100000))
    say(f"learned: loss {l0:.3e} -> {l1:.3e} in {sum(steps)} Adam steps, {train_s:.1f}s")
    d.close()
    net_adj, adj_train = None, None
    if adjoint_layer:
        # the adjoint network: pairs of the adjoint runs of the same shots
        # (the recorded data, or with adjoint_residual the residual data a
        # migration injects through the smoothed model), levels spread over
        # the adjoint's range, the set built after the forward one is freed
        da = Data(ncell * levels_per_shot * nshot_train)
        t0 = _time()
        for cw, cws, css, zs, xs, wav, sd in train_shots:
            lv = np.linspace(1, nt - 150, levels_per_shot).round().astype(int)
            da.collect_adj(nz, nx, pad_w, abc_wide, nt, dt, zs, xs, zr, cws, css, wav, lv, noise=noise, seed=sd,
                           abc_band=band_abc, coefs_wide_obs=cw if adjoint_residual else None)
        collect_s = _time() - t0
        net_adj, a0, a1, a_s = fit(da, int(seed % 100000) + 1)
        da.close()
        adj_train = {"rows": ncell * levels_per_shot * nshot_train, "loss_initial": a0, "loss_final": a1,
                     "collect_seconds": collect_s, "seconds": a_s}
        say(f"learned: adjoint network loss {a0:.3e} -> {a1:.3e}, {collect_s:.1f}s collection, {a_s:.1f}s training")
    train_shots = None

    # held-out models: reflection energy and images
    refl = {"band_sponge": [], "band_sponge_net": [], "sponge40": []}
    rec_err = {k: [] for k in refl}
    img_err = {k: [] for k in refl}
    if net_adj is not None:
        img_err["band_sponge_net_adj"] = []
        img_err["band_sponge_net_both"] = []
    ms_level = {}
    for mi in range(test_models):
        vp = random_layered_vp(nz, nx, rng)
        vps = synth.smooth(vp, 10.0)
        shots = [(int(rng.integers(2, nz // 3)), int(rng.integers(nx // 8, nx - nx // 8)), float(rng.uniform(8, 25)))
                 for _ in range(shots_per_model)]
        for zs, xs, fp in shots:
            wav = _lib.ricker(nt, dt, fp)
            # forward runs of the true model: the wide reference and the three boundaries
            sw = _lib.shot(nz + 2 * pad_w, nx + 2 * pad_w, nt, dx, dt, pad_w + xs, pad_w + zs, pad_w + zr, pad_w, nx,
                           abc_wide)
            rw, lw = _lib.forward(sw, _coefs(vp, pad_w, dx, dt), wav, want_last=True)
            ref = [a[pad_w:pad_w + nz, pad_w:pad_w + nx] for a in lw]
            sb = _lib.shot(nzb, nxb, nt, dx, dt, PAD + xs, PAD + zs, PAD + zr, PAD, nx, band_abc)
            cb = _coefs(vp, PAD, dx, dt)
            for key, n in (("band_sponge", None), ("band_sponge_net", net)):
                r, last, ms = propagate(n, sb, cb, wav, want_last=True)
                refl[key].append(_rel([a[PAD:PAD + nz, PAD:PAD + nx] for a in last], ref))
                rec_err[key].append(_rel(r, rw))
                ms_level.setdefault(key, []).append(ms / (nt - 1))
            p40 = 45
            s40 = _lib.shot(nz + 2 * p40, nx + 2 * p40, nt, dx, dt, p40 + xs, p40 + zs, p40 + zr, p40, nx, 40)
            r40, l40 = _lib.forward(s40, _coefs(vp, p40, dx, dt), wav, want_last=True)
            refl["sponge40"].append(_rel([a[p40:p40 + nz, p40:p40 + nx] for a in l40], ref))
            rec_err["sponge40"].append(_rel(r40, rw))
            # RTM: observed data of the wide grid (true minus smooth model)
            rws = _lib.forward(sw, _coefs(vps, pad_w, dx, dt), wav)
            dobs = rw - rws
            imgs = {}
            cases = [("wide", pad_w, abc_wide, None, None), ("sponge40", p40, 40, None, None),
                     ("band_sponge", PAD, band_abc, None, None), ("band_sponge_net", PAD, band_abc, net, None)]
            if net_adj is not None:
                cases += [("band_sponge_net_adj", PAD, band_abc, None, net_adj),
                          ("band_sponge_net_both", PAD, band_abc, net, net_adj)]
            for key, pad, abc, fwd_net, adj_net in cases:
                spec = dict(nz=nz + 2 * pad, nx=nx + 2 * pad, xs=pad + xs, zs=pad + zs, zr=pad + zr, rx0=pad, nrx=nx,
                            img_z0=pad, img_x0=pad, img_nz=nz, img_nx=nx, stack_z0=0, stack_x0=0)
                b = _lib.Batch(nt, dx, dt, [spec], [_coefs(vps, pad, dx, dt)], wav, abc=abc, adj_abc=1, txr=2)
                if fwd_net is not None:
                    b.set_learned(fwd_net)
                if adj_net is not None:
                    b.set_learned_adjoint(adj_net)
                b.set_data(0, dobs)
                b.migrate()
                img = b.image(0)
                imgs[key] = [a[pad:pad + nz, pad:pad + nx] for a in img]
                b.close()
            for key in img_err:
                img_err[key].append(_rel(imgs[key], imgs["wide"]))
            say(f"learned: model {mi} shot ({zs},{xs}) {fp:.1f} Hz: reflection "
                f"{[round(refl[k][-1], 4) for k in refl]}, image {[round(img_err[k][-1], 4) for k in img_err]}")

    def mean(v):
        return float(np.mean(v)) if v else None

    if net_adj is not None:
        net_adj.close()
    out = {"training": {"snapshot_pairs": nshot_train * levels_per_shot, "rows": ncell * levels_per_shot * nshot_train,
                        "loss_initial": l0, "loss_final": l1, "adam_steps": int(sum(steps)), "seconds": train_s},
           "adjoint_training": adj_train,
           "reflection_rel_l2": {k: mean(v) for k, v in refl.items()},
           "receiver_rel_l2": {k: mean(v) for k, v in rec_err.items()},
           "image_rel_l2_vs_wide": {k: mean(v) for k, v in img_err.items()},
           "per_case": {"reflection": refl, "image": img_err},
           "ms_per_forward_level": {k: mean(v) for k, v in ms_level.items()}}
    return out, net


def _time():
    import time

    return time.perf_counter()
