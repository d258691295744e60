"""GPU tests of the learned absorbing layer (elastic_amd/csrc/learned_abc.hip)
against the numpy oracle (oracle/learned_abc.py): band geometry, features
(forward and adjoint states), network forward and gradient, teacher-forced
pairs from wide-sponge runs (forward and adjoint), training, propagation
with the layer, and the layer in the batch engine's forward and adjoint
runs."""
import numpy as np
import pytest

import oracle as O
from oracle import learned_abc as LA
from elastic_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L(hip):
    from elastic_amd import learned

    return learned


def test_cell_geometry(L):
    for nz, nx in ((40, 57), (64, 200)):
        iz, ix, edge, depth = LA.band_cells(nz, nx)
        assert L.cell_count(nz, nx) == iz.size
        for c in list(range(0, iz.size, 7)) + [iz.size - 1]:
            assert L.cell_coords(nz, nx, c) == (iz[c], ix[c], edge[c], depth[c])
    assert L.cell_count(20, 20) == 0  # too small for a band on both sides


def test_features_bit_exact(L):
    rng = np.random.default_rng(5)
    nz, nx = 48, 70
    m1 = [(rng.standard_normal((nz, nx)) * (1e-3 if f < 2 else 1e4)).astype(np.float32) for f in range(5)]
    m2 = [(rng.standard_normal((nz, nx)) * (1e-3 if f < 2 else 1e4)).astype(np.float32) for f in range(5)]
    m1[0][:, 30:] = 0  # some cells with an all-zero patch of Vx
    vp = synth.layered_vp(nz, nx, 2, vp_top=1800.0, vp_bot=3500.0, nlayers=2, anom_sigma=5.0)
    m = synth.elastic_from_vp(vp)
    coefs = O.model_op(m.rho, m.lam, m.mu, 10.0, 1e-3, np.float32)
    f, s = L.features(m1, m2, coefs)
    fo, so = LA.features(m1, m2, coefs, np.float32)
    assert np.array_equal(s, so.astype(np.float32))
    assert np.array_equal(f, fo)


def test_adjoint_features_bit_exact(L):
    """Features of an adjoint state (levels L+1, L+2; the velocities divided
    by Z0) equal the oracle's float32 emulation bit for bit."""
    rng = np.random.default_rng(15)
    nz, nx = 48, 70
    p1 = [(rng.standard_normal((nz, nx)) * (1e4 if f < 2 else 1e-3)).astype(np.float32) for f in range(5)]
    p2 = [(rng.standard_normal((nz, nx)) * (1e4 if f < 2 else 1e-3)).astype(np.float32) for f in range(5)]
    p1[3][:, 40:] = 0
    vp = synth.layered_vp(nz, nx, 2, vp_top=1800.0, vp_bot=3500.0, nlayers=2, anom_sigma=5.0)
    m = synth.elastic_from_vp(vp)
    coefs = O.model_op(m.rho, m.lam, m.mu, 10.0, 1e-3, np.float32)
    f, s = L.features(p1, p2, coefs, adjoint=True)
    fo, so = LA.features(p1, p2, coefs, np.float32, adjoint=True)
    assert np.array_equal(s, so.astype(np.float32))
    assert np.array_equal(f, fo)
    fw, _ = L.features(p1, p2, coefs)
    assert not np.array_equal(f, fw)


def _params(seed, scale=0.2):
    return (np.random.default_rng(seed).standard_normal(LA.NPARAM) * scale).astype(np.float32)


def _feat(n, seed):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, LA.KP), np.float32)
    f[:, :LA.NRAW] = rng.uniform(-1, 1, (n, LA.NRAW))
    f[np.arange(n), LA.NRAW + rng.integers(0, 4, n)] = 1
    f[np.arange(n), LA.NRAW + 4 + rng.integers(0, 8, n)] = 1
    return LA.bf16(f)


def test_forward_and_gradient_vs_oracle(L):
    net = L.Net(seed=3)
    p = _params(1)
    net.params = p
    assert np.array_equal(net.params, p)
    for n in (1, 37, 1000):  # partial blocks of 32 samples
        f = _feat(n, n)
        y = net.forward(f)
        yo = LA.mlp(p, f)
        assert O.rel_l2(y, yo) <= 1e-5, n
        t = np.random.default_rng(n + 1).standard_normal((n, 5)).astype(np.float32)
        loss, g = net.grad(f, t)
        lo, go = LA.loss_grad(p, f, t)
        assert abs(loss - lo) <= 1e-5 * abs(lo), n
        assert O.rel_l2(g, go) <= 1e-4, n
        for a, b in ((LA.OW1, LA.OB1), (LA.OW2, LA.OB2), (LA.OW3, LA.OB3), (LA.OB3, LA.NPARAM)):
            assert O.rel_l2(g[a:b], go[a:b]) <= 1e-4, (n, a)
    net.close()


def _case(nt=160):
    nzp, nxp, dx, dt, fp = 30, 40, 10.0, 1e-3, 30.0
    abc_w = 20
    pad_w = abc_w + 5 + LA.R + LA.W
    vp = synth.layered_vp(nzp, nxp, 4, vp_top=2000.0, vp_bot=2600.0, nlayers=2, jump=100.0, anom_sigma=4.0)
    mw = synth.elastic_from_vp(synth.pad_edge(vp, pad_w))
    ms = synth.elastic_from_vp(synth.pad_edge(vp, LA.R + LA.W))
    cw = O.model_op(mw.rho, mw.lam, mw.mu, dx, dt, np.float32)
    cs = O.model_op(ms.rho, ms.lam, ms.mu, dx, dt, np.float32)
    wav = O.ricker(nt, dt, fp, nsource=60, dtype=np.float32)
    return dict(nz_phys=nzp, nx_phys=nxp, pad_wide=pad_w, abc_wide=abc_w, nt=nt, dt=dt, zs_phys=14, xs_phys=9,
                coefs_wide=cw, coefs_small=cs, wavelet=wav)


@pytest.mark.parametrize("band", [0, 8])
def test_collect_vs_oracle(L, band):
    case = _case()
    levels = [2, 70, 100, 131, 160]
    d = L.Data(100000)
    d.collect(levels=levels, abc_band=band, **case)
    fo, to = LA.teacher_pairs(levels=levels, abc_band=band, **case)
    f, t = d.get()
    assert f.shape == fo.shape  # level 2 has no pair
    # features: FP32 run vs FP64 run, then bf16 rounding: within one bf16 step
    # (plus a floor for values near cancellation)
    assert np.all(np.abs(f - fo) <= 2.0 ** -7 * np.abs(fo) + 1e-4)
    assert np.abs(f - fo).mean() < 1e-4
    # targets: FP32 differences of nearly equal levels; normalised units
    assert np.abs(t - to).max() <= 1e-4 * max(1.0, np.abs(to).max())
    assert np.abs(to).max() > 1e-3  # the wave reached the rim
    d.close()


@pytest.mark.parametrize("band,resid", [(0, False), (8, False), (8, True)])
def test_collect_adj_vs_oracle(L, band, resid):
    """Adjoint teacher pairs (wide forward recording receivers, its adjoint
    run with the exact-adjoint sponge, windows of levels L+1, L+2, the band
    grid's adjoint update) against the FP64 oracle; with resid, the data are
    those of an observed model (an added anomaly) minus the adjoint's."""
    case = _case()
    obs = {}
    if resid:
        vp = synth.layered_vp(30, 40, 4, vp_top=2000.0, vp_bot=2600.0, nlayers=2, jump=100.0, anom_sigma=4.0)
        vp = vp + 300.0 * np.exp(-((np.arange(30)[:, None] - 18.0) ** 2 + (np.arange(40)[None, :] - 20.0) ** 2) / 20.0)
        mo = synth.elastic_from_vp(synth.pad_edge(np.asfortranarray(vp), case["pad_wide"]))
        obs = dict(coefs_wide_obs=O.model_op(mo.rho, mo.lam, mo.mu, 10.0, case["dt"], np.float32))
    levels = [20, 60, 95, 130, 159]
    d = L.Data(100000)
    d.collect_adj(levels=levels, abc_band=band, zr_phys=2, **case, **obs)
    fo, to = LA.teacher_pairs_adj(levels=levels, abc_band=band, zr_phys=2, **case, **obs)
    f, t = d.get()
    assert f.shape == fo.shape
    # residual data are the difference of two runs about 50x larger, so the
    # FP32 runs' rounding reaches them 50x larger than in full data: a floor
    # of 5e-3 of the patch scale instead of 1e-4 (measured: 3.9e-3 at most,
    # 0.03% of the values above 1e-4; the full data instead of the residual
    # would miss by up to 2.0 on 57% of them)
    assert np.all(np.abs(f - fo) <= 2.0 ** -7 * np.abs(fo) + (5e-3 if resid else 1e-4))
    assert np.abs(f - fo).mean() < 1e-4
    assert np.abs(t - to).max() <= 1e-4 * max(1.0, np.abs(to).max())
    assert np.abs(to).max() > 1e-3  # the adjoint wavefield reached the rim
    d.close()


def test_collect_rejects_bad_level_lists(L):
    """A level outside the run or listed twice is an error, not a silently
    shorter training set."""
    from elastic_amd.lib import RtmError

    case = _case(nt=60)
    d = L.Data(100000)
    for levels in ([30, 30], [1, 30], [30, 61]):
        with pytest.raises(RtmError):
            d.collect(levels=levels, **case)
    for levels in ([0, 30], [30, 61], [20, 20]):
        with pytest.raises(RtmError):
            d.collect_adj(levels=levels, zr_phys=2, **case)
    assert len(d) == 0
    d.collect(levels=[2, 30, 60], **case)
    d.collect_adj(levels=[1, 30, 60], zr_phys=2, **case)
    n = L.cell_count(case["nz_phys"] + 2 * L.PAD, case["nx_phys"] + 2 * L.PAD)
    assert len(d) == 2 * n + 3 * n  # level 2 gives no forward pair
    d.close()


def test_training_reduces_loss(L):
    case = _case(nt=200)
    d = L.Data(200000)
    d.collect(levels=list(range(40, 201, 4)), **case)
    net = L.Net(seed=7)
    l0 = net.eval(d)
    losses = net.train(d, steps=300, batch=2048, lr=2e-3, seed=1)
    l1 = net.eval(d)
    assert np.all(np.isfinite(losses))
    assert l1 < 0.2 * l0, (l0, l1)
    # the full-set loss agrees with the oracle on the trained weights
    f, t = d.get()
    lo, _ = LA.loss_grad(net.params, f, t)
    assert abs(l1 - lo) <= 1e-3 * lo
    net.close()
    d.close()


def test_propagation_without_net_equals_forward(L, hip):
    """labc_forward with no network is the plain sponge forward."""
    case = _case(nt=150)
    nz, nx = case["nz_phys"] + 2 * 13, case["nx_phys"] + 2 * 13
    s = hip.shot(nz, nx, 150, 10.0, case["dt"], 13 + case["xs_phys"], 13 + case["zs_phys"], 15, 13, 40, abc=8)
    rec, last, _ = L.propagate(None, s, case["coefs_small"], case["wavelet"], want_last=True)
    rec2, last2 = hip.forward(s, case["coefs_small"], case["wavelet"], want_last=True)
    assert np.array_equal(rec, rec2)
    for a, b in zip(last, last2):
        assert np.array_equal(a, b)


def test_learned_propagation_vs_oracle(L, hip):
    """Free run with the network in the loop: at every level L >= 3 the
    correction s * NN(features(L-1, L-2)) is added to the band cells of the
    plain update and the rim is cleared; GPU (FP32 state, parallel inference
    kernel) against the oracle (FP64 state, FP64 network on the same bf16
    weights)."""
    nzp, nxp, dt, nt = 20, 26, 1e-3, 60
    pad = LA.R + LA.W
    vp = synth.layered_vp(nzp, nxp, 5, vp_top=2000.0, vp_bot=2600.0, nlayers=2, jump=100.0, anom_sigma=4.0)
    m = synth.elastic_from_vp(synth.pad_edge(vp, pad))
    coefs = O.model_op(m.rho, m.lam, m.mu, 10.0, dt, np.float32)
    wav = O.ricker(nt, dt, 40.0, nsource=40, dtype=np.float32)
    nz, nx = nzp + 2 * pad, nxp + 2 * pad
    zs, xs = pad + 4, pad + 5
    params = (np.random.default_rng(9).standard_normal(LA.NPARAM) * 0.05).astype(np.float32)
    net = L.Net(seed=1)
    net.params = params
    s = hip.shot(nz, nx, nt, 10.0, dt, xs, zs, pad + 1, pad, nxp, abc=0)
    rec, last, _ = L.propagate(net, s, coefs, wav, want_last=True)
    net.close()
    # oracle free run
    p = O.make_params(nz, nx, nt, 10.0, dt, xs, zs, 0, 0, 0, 0)
    iz, ix, _, _ = LA.band_cells(nz, nx)
    z = [np.zeros((nz, nx)) for _ in range(5)]
    levs = {0: z, 1: [a.copy() for a in z]}
    for lv in range(2, nt + 1):
        new = O.fwd_level(p, levs[lv - 2], levs[lv - 1], coefs, srcamp=float(dt) * float(wav[lv - 2]))
        if lv >= 3:
            f, sc = LA.features(levs[lv - 1], levs[lv - 2], coefs)
            y = LA.mlp(params, f)
            for fi in range(5):
                new[fi][iz, ix] += y[:, fi] * sc * (LA.Z0 if fi >= 2 else 1.0)
                new[fi][:LA.R, :] = 0
                new[fi][nz - LA.R:, :] = 0
                new[fi][:, :LA.R] = 0
                new[fi][:, nx - LA.R:] = 0
        levs[lv] = new
        levs.pop(lv - 2)
    # the correction was active: the run differs from the plain one
    plain, _ = L.propagate(None, hip.shot(nz, nx, nt, 10.0, dt, xs, zs, pad + 1, pad, nxp, abc=0), coefs, wav)
    assert O.rel_l2(rec, plain) > 1e-3
    for fi in range(5):
        assert O.rel_l2(last[5 + fi], levs[nt][fi]) <= 1e-3, fi


def test_learned_layer_gain_over_same_width_sponge(L, hip):
    """The layer trained on the GPU (band-sponge teachers from 120-point wide
    runs of two seeded layered models) reduces the boundary-reflection
    energy of an 8-cell sponge: on shots of a held-out model, the wavefield
    error in the physical region at the last level against the wide-sponge
    reference is smaller with the network than with the 8-cell sponge alone
    (measured gain on the full experiment: 1.6-3.7x on 4 of 5 shots)."""
    dx, dt, nt = 10.0, 1e-3, 700
    nzp, nxp = 80, 160
    abc_w = 120
    pad_w = abc_w + 5 + L.PAD

    def model(seed):
        vp = synth.layered_vp(nzp, nxp, seed, vp_top=1800.0, vp_bot=3500.0, nlayers=4, anom_sigma=8.0)
        return np.clip(vp, 1500, 4500)

    def coefs(vp, pad):
        m = synth.elastic_from_vp(synth.pad_edge(vp, pad))
        return hip.model_op(m.rho, m.lam, m.mu, dx, dt)

    wav = hip.ricker(nt, dt, 15.0)
    levels = list(range(100, nt + 1, 4))
    ncell = L.cell_count(nzp + 2 * L.PAD, nxp + 2 * L.PAD)
    d = L.Data(ncell * len(levels) * 4)
    for seed in (31, 32):
        vp = model(seed)
        cw, cs = coefs(vp, pad_w), coefs(vp, L.PAD)
        for zs, xs in ((10, 20), (60, 130)):
            d.collect(nzp, nxp, pad_w, abc_w, nt, dt, zs, xs, cw, cs, wav, levels, noise=0.03, seed=seed + zs,
                      abc_band=8)
    net = L.Net(seed=2)
    l0 = net.eval(d)
    for lr, n in ((2e-3, 1500), (5e-4, 700), (1e-4, 700)):
        net.train(d, n, batch=8192, lr=lr, seed=int(lr * 1e6))
    assert net.eval(d) < 0.2 * l0
    d.close()
    vp = model(77)
    cref, csm = coefs(vp, pad_w), coefs(vp, L.PAD)
    ratios = []
    for zs, xs in ((15, 30), (50, 120)):
        sref = hip.shot(nzp + 2 * pad_w, nxp + 2 * pad_w, nt, dx, dt, pad_w + xs, pad_w + zs, pad_w + 2, pad_w,
                        nxp, abc_w)
        _, last_ref = hip.forward(sref, cref, wav, want_last=True)
        ref = [a[pad_w:pad_w + nzp, pad_w:pad_w + nxp] for a in last_ref[5:7]]
        s8 = hip.shot(nzp + 2 * L.PAD, nxp + 2 * L.PAD, nt, dx, dt, L.PAD + xs, L.PAD + zs, L.PAD + 2, L.PAD, nxp, 8)
        errs = []
        for n in (None, net):
            _, last, _ = L.propagate(n, s8, csm, wav, want_last=True)
            win = [a[L.PAD:L.PAD + nzp, L.PAD:L.PAD + nxp] for a in last[5:7]]
            assert all(np.all(np.isfinite(w)) for w in win)
            errs.append(np.sqrt(sum(np.sum((a - b) ** 2) for a, b in zip(win, ref)) / sum(np.sum(b ** 2) for b in ref)))
        ratios.append(errs[1] / errs[0])
    net.close()
    assert np.mean(ratios) < 0.85, ratios


def test_batch_engine_learned_layer_equals_single_grid_path(L, hip):
    """The batch engine with the learned layer (per-level MFMA inference over
    all shots of a group, uneven grids) gives the single-grid propagation's
    receivers bit for bit: synthesized data = labc_forward(true) -
    labc_forward(migration model), on every shot."""
    nt, dx, dt = 300, 10.0, 1e-3
    pad = L.PAD
    rng = np.random.default_rng(21)
    params = (rng.standard_normal(LA.NPARAM) * 0.05).astype(np.float32)
    net = L.Net(seed=1)
    net.params = params
    specs, ct, cm = [], [], []
    for k, (nzp, nxp) in enumerate(((40, 70), (52, 61), (40, 90))):
        vt = synth.layered_vp(nzp, nxp, 30 + k, vp_top=1800.0, vp_bot=3000.0, nlayers=3, anom_sigma=5.0)
        vm = synth.smooth(vt, 3.0)
        specs.append(dict(nz=nzp + 2 * pad, nx=nxp + 2 * pad, xs=pad + nxp // 2, zs=pad + 3, zr=pad + 2, rx0=pad,
                          nrx=nxp, img_z0=pad, img_x0=pad, img_nz=nzp, img_nx=nxp, stack_z0=0, stack_x0=0))
        for v, lst in ((vt, ct), (vm, cm)):
            m = synth.elastic_from_vp(synth.pad_edge(v, pad))
            lst.append(hip.model_op(m.rho, m.lam, m.mu, dx, dt))
    w = hip.ricker(nt, dt, 25.0, nsource=60)
    b = hip.Batch(nt, dx, dt, specs, cm, w, abc=8, adj_abc=1, txr=2, use_graph=1)
    b.set_true_model(ct)
    b.set_learned(net)
    b.synthesize()
    for i, sp in enumerate(specs):
        s = hip.shot(sp["nz"], sp["nx"], nt, dx, dt, sp["xs"], sp["zs"], sp["zr"], sp["rx0"], sp["nrx"], abc=8)
        rt, _ = L.propagate(net, s, ct[i], w)
        rm, _ = L.propagate(net, s, cm[i], w)
        assert np.array_equal(b.get_data(i), rt - rm), i
        plain, _ = L.propagate(None, s, ct[i], w)
        assert O.rel_l2(rt, plain) > 1e-4  # the layer acted
    b.migrate()
    with_net = b.image(0)
    b.set_learned(None)
    b.migrate()
    without = b.image(0)
    b.close()
    net.close()
    for a, c in zip(with_net, without):
        assert np.isfinite(a).all() and np.abs(a).max() > 0
    assert max(O.rel_l2(a, c) for a, c in zip(with_net, without)) > 1e-6


def test_batch_engine_adjoint_layer_vs_oracle(L, hip):
    """The learned layer in the batch engine's adjoint runs: after every
    adjoint level (formed and imaged in its step) the band cells get the
    network's correction from levels L+1, L+2 and the rim is cleared. The
    images (FP32 state, the parallel inference kernel) match the oracle's
    RTM with the same layer (FP64 state, FP64 network on the same bf16
    weights); the layer acted; dropping it restores the plain images."""
    nzp, nxp, nt, dx, dt = 30, 44, 120, 10.0, 1e-3
    pad = L.PAD
    nz, nx = nzp + 2 * pad, nxp + 2 * pad
    vp = synth.layered_vp(nzp, nxp, 8, vp_top=2000.0, vp_bot=2800.0, nlayers=2, jump=100.0, anom_sigma=4.0)
    m = synth.elastic_from_vp(synth.pad_edge(vp, pad))
    coefs = O.model_op(m.rho, m.lam, m.mu, dx, dt, np.float32)
    wav = O.ricker(nt, dt, 30.0, nsource=50, dtype=np.float32)
    spec = dict(nz=nz, nx=nx, xs=pad + 20, zs=pad + 4, zr=pad + 2, rx0=pad, nrx=nxp)
    p = O.make_params(nz, nx, nt, dx, dt, spec["xs"], spec["zs"], spec["zr"], pad, nxp, 8, adj_abc=1)
    rec = O.forward(p, wav, coefs, np.float32)["rec"]
    params = (np.random.default_rng(19).standard_normal(LA.NPARAM) * 0.05).astype(np.float32)
    net = L.Net(seed=1)
    net.params = params
    b = hip.Batch(nt, dx, dt, [spec], [coefs], wav, abc=8, adj_abc=1, txr=2)
    b.set_data(0, rec)
    b.migrate()
    plain = b.image(0)
    b.set_learned_adjoint(net)
    b.migrate()
    got = b.image(0)
    b.set_learned_adjoint(None)
    b.migrate()
    again = b.image(0)
    b.close()
    net.close()
    ref = LA.rtm_adjoint_layer(p, wav, coefs, rec, params)
    ref_plain = LA.rtm_adjoint_layer(p, wav, coefs, rec)
    for a, r, q, r0 in zip(got, ref, plain, ref_plain):
        assert O.rel_l2(a, r) <= 1e-3
        assert O.rel_l2(q, r0) <= 1e-4
        assert O.rel_l2(r, r0) > 1e-3  # the layer acted
    for a, c in zip(again, plain):
        assert np.array_equal(a, c)


def test_batch_refuses_layer_on_small_grids(L, hip):
    """Grids below 27 cells a side have no room for the band's patches: the
    batch refuses either network and keeps running without a layer."""
    nt, dx, dt = 80, 10.0, 1e-3
    vp = synth.layered_vp(20, 24, 3, vp_top=2000.0, vp_bot=2600.0, nlayers=2, anom_sigma=4.0)
    m = synth.elastic_from_vp(vp)
    coefs = hip.model_op(m.rho, m.lam, m.mu, dx, dt)
    spec = dict(nz=20, nx=24, xs=12, zs=6, zr=3, rx0=0, nrx=24)
    w = hip.ricker(nt, dt, 30.0, nsource=40)
    b = hip.Batch(nt, dx, dt, [spec], [coefs], w, abc=0, adj_abc=0, txr=2)
    b.set_data(0, np.ones((5, 24, nt), np.float32))
    b.migrate()
    plain = b.image(0)
    net = L.Net(seed=1)
    for setter in (b.set_learned, b.set_learned_adjoint):
        with pytest.raises(hip.RtmError):
            setter(net)
    b.migrate()
    for a, c in zip(b.image(0), plain):
        assert np.array_equal(a, c)
    b.close()
    net.close()


def test_learned_experiment_small(L, hip):
    """The end-to-end experiment (training on wide-sponge teachers, held-out
    reflection and RTM image errors for the three boundaries) at a small
    size: every metric is finite and the training lowers the loss."""
    out, net = L.experiment(nz=60, nx=120, nt=500, train_models=2, shots_per_model=1, levels_per_shot=40,
                            test_models=1, steps=(300, 150, 150), batch=4096, seed=7)
    net.close()
    print(out["reflection_rel_l2"], out["image_rel_l2_vs_wide"], out["ms_per_forward_level"])
    assert out["training"]["loss_final"] < out["training"]["loss_initial"]
    for grp in ("reflection_rel_l2", "receiver_rel_l2", "image_rel_l2_vs_wide"):
        for k, v in out[grp].items():
            assert v is not None and np.isfinite(v), (grp, k)


def test_learned_layer_400x800_against_both_sponges(L, hip):
    """The learned layer at the 400 x 800 size: 2000 teacher-forced snapshot
    pairs from runs with a 120-point sponge, held-out models, every boundary
    against the 120-point run. The layer must cut the forward reflection of
    its own 8-point band sponge by at least 1.3x (1.57x measured) and leave
    every image finite; the 40-point sponge (45 cells against 13) stays
    ahead of it on reflection, receiver data and images, and the check
    states so, so that a layer that closed the gap would show up here."""
    out, net = L.experiment(nz=400, nx=800, nt=2500)
    net.close()
    r, d, im = out["reflection_rel_l2"], out["receiver_rel_l2"], out["image_rel_l2_vs_wide"]
    print("reflection", r, "receiver", d, "image", im, "ms/level", out["ms_per_forward_level"], flush=True)
    assert out["training"]["snapshot_pairs"] == 2000
    assert r["band_sponge_net"] * 1.3 <= r["band_sponge"]
    assert d["band_sponge_net"] < d["band_sponge"]
    for grp in (r, d, im):
        for v in grp.values():
            assert v is not None and np.isfinite(v) and v > 0
        assert grp["sponge40"] < grp["band_sponge_net"]
