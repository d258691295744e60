"""Pins of the learned-absorbing-layer oracle (oracle/learned_abc.py), CPU only.

The learned layer has no counterpart in the reference, so the oracle is pinned
by independent routes: the band against a brute-force distance mask, bf16
rounding against torch's bfloat16 conversion, the network and its backward
pass against torch float64 autograd and central finite differences, the
feature normalisation by its exact scaling law, and the teacher-forced
targets by the locality of the 10th-order stencil (a target is nonzero only
within the stencil radius of the rim) and by linearity (the target equals the
plain update of the rim values alone)."""
import numpy as np
import pytest
import torch

import oracle as O
from oracle import learned_abc as LA
from elastic_amd import synth


def test_band_matches_distance_mask():
    for nz, nx in ((40, 57), (31, 31), (64, 200)):
        iz, ix, edge, depth = LA.band_cells(nz, nx)
        z, x = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
        dist = np.minimum(np.minimum(z, nz - 1 - z), np.minimum(x, nx - 1 - x))
        mask = (dist >= LA.R) & (dist < LA.R + LA.W)
        got = np.zeros((nz, nx), int)
        np.add.at(got, (iz, ix), 1)
        assert got.max() == 1  # each cell once
        assert np.array_equal(got.astype(bool), mask)
        # depth is the distance to the cell's own edge, minus the rim
        own = np.select([edge == 0, edge == 1, edge == 2], [iz, nz - 1 - iz, ix], nx - 1 - ix)
        assert np.array_equal(own - LA.R, depth)
        # corners belong to the top/bottom rows
        corner = (dist[iz, ix] < LA.R + LA.W) & ((iz < LA.R + LA.W) | (iz >= nz - LA.R - LA.W))
        assert np.all(edge[corner] < 2)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-20, 20, 10000),
                        np.float32([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8, 0.0, 65504.0])])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(LA.bf16(x), ref)
    assert LA.bf16(np.float32([1.0 + 2.0 ** -8]))[0] == 1.0  # tie to even
    assert LA.bf16(np.float32([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6


def _rand_params(seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(LA.NPARAM) * 0.2).astype(np.float32)


def _rand_feat(n, seed):
    rng = np.random.default_rng(seed)
    f = np.zeros((n, LA.KP), np.float32)
    f[:, :LA.NRAW] = rng.uniform(-1, 1, (n, LA.NRAW))
    f[np.arange(n), LA.NRAW + rng.integers(0, 4, n)] = 1
    f[np.arange(n), LA.NRAW + 4 + rng.integers(0, 8, n)] = 1
    return LA.bf16(f)


def _torch_net(params, feat, tgt):
    W1, b1, W2, b2, W3, b3 = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in LA.unpack(params)]
    x = torch.tensor(feat[:, :LA.NF], dtype=torch.float64)
    h1 = torch.nn.functional.gelu(x @ W1 + b1)
    h2 = torch.nn.functional.gelu(h1 @ W2 + b2)
    y = h2 @ W3 + b3
    loss = 0.5 * torch.mean((y - torch.tensor(tgt, dtype=torch.float64)) ** 2)
    loss.backward()
    g = np.concatenate([t.grad.numpy().ravel() for t in (W1, b1, W2, b2, W3, b3)])
    return y.detach().numpy(), loss.item(), g


def test_mlp_and_gradient_match_torch_autograd():
    p = _rand_params(1)
    f = _rand_feat(64, 2)
    t = np.random.default_rng(3).standard_normal((64, LA.NO))
    y_t, loss_t, g_t = _torch_net(p, f, t)
    assert np.allclose(LA.mlp(p, f), y_t, rtol=1e-12, atol=1e-12)
    loss, g = LA.loss_grad(p, f, t)
    assert abs(loss - loss_t) <= 1e-12 * abs(loss_t)
    assert np.allclose(g, g_t, rtol=1e-10, atol=1e-14)


def test_gradient_central_differences():
    p = _rand_params(4)
    f = _rand_feat(16, 5)
    t = np.random.default_rng(6).standard_normal((16, LA.NO))
    _, g = LA.loss_grad(p, f, t)
    # differentiate in the rounded weight space: perturb the unpacked arrays
    W = [a.copy() for a in LA.unpack(p)]
    x = f[:, :LA.NF].astype(np.float64)

    def loss_of(ws):
        h1 = LA.gelu(x @ ws[0] + ws[1])
        h2 = LA.gelu(h1 @ ws[2] + ws[3])
        y = h2 @ ws[4] + ws[5]
        return 0.5 * np.mean((y - t) ** 2)

    offs = [LA.OW1, LA.OB1, LA.OW2, LA.OB2, LA.OW3, LA.OB3]
    rng = np.random.default_rng(7)
    for a in range(6):
        for _ in range(6):
            i = rng.integers(0, W[a].size)
            h = 1e-6
            wp = [w.copy() for w in W]
            wm = [w.copy() for w in W]
            wp[a].flat[i] += h
            wm[a].flat[i] -= h
            fd = (loss_of(wp) - loss_of(wm)) / (2 * h)
            assert abs(fd - g[offs[a] + i]) <= 1e-7 + 1e-5 * abs(fd), (a, i)


def _rand_levels(nz, nx, seed):
    rng = np.random.default_rng(seed)
    m1 = [rng.standard_normal((nz, nx)) * (1e-3 if f < 2 else 1e4) for f in range(5)]
    m2 = [rng.standard_normal((nz, nx)) * (1e-3 if f < 2 else 1e4) for f in range(5)]
    return m1, m2


def _coefs(nz, nx, seed, dx=10.0, dt=1e-3):
    """Piecewise-constant blocks (so that most cells have a constant 2x2
    neighbourhood) with seeded velocities."""
    rng = np.random.default_rng(seed)
    z, x = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    vp = np.asfortranarray(rng.uniform(1800, 3500, (4,))[(z >= nz // 2) * 2 + (x >= nx // 3)])
    m = synth.elastic_from_vp(vp)
    return O.model_op(m.rho, m.lam, m.mu, dx, dt), m


def test_features_scaling_and_layout():
    nz, nx = 36, 44
    m1, m2 = _rand_levels(nz, nx, 8)
    coefs, model = _coefs(nz, nx, 9)
    f, s = LA.features(m1, m2, coefs)
    cells = LA.band_cells(nz, nx)
    iz, ix, edge, depth = cells
    # exact scaling law: features are invariant, scales double
    f2, s2 = LA.features([2 * a for a in m1], [2 * a for a in m2], coefs)
    assert np.array_equal(f, f2) and np.array_equal(s2, 2 * s)
    # max-abs normalisation: the largest raw feature is exactly +-1
    assert np.all(np.abs(f[:, :LA.NRAW]).max(axis=1) == 1.0)
    # one-hots and zero padding
    assert np.array_equal(np.argmax(f[:, 230:234], 1), edge) and np.all(f[:, 230:234].sum(1) == 1)
    assert np.array_equal(np.argmax(f[:, 234:242], 1), depth) and np.all(f[:, 234:242].sum(1) == 1)
    assert np.all(f[:, 245:] == 0)
    # material: Courant numbers and impedance of the cell. The coefficients
    # use averaged rho / harmonic lambda, mu, so compare where the model is
    # locally constant (its 2x2 neighbourhood) to the closed form
    vp, vs, rho = model.vp[iz, ix], model.vs[iz, ix], model.rho[iz, ix]
    flat = np.ones(iz.size, bool)
    for a in (model.vp,):
        for dz, dx in ((1, 0), (0, 1), (1, 1)):
            flat &= a[np.minimum(iz + dz, nz - 1), np.minimum(ix + dx, nx - 1)] == a[iz, ix]
    assert flat.sum() > 100
    for k, want in ((242, vp * 1e-3 / 10.0), (243, vs * 1e-3 / 10.0), (244, rho * vp / LA.Z0)):
        got = f[flat, k]
        assert np.all(np.abs(got - want[flat]) <= 2.0 ** -8 * np.abs(want[flat]) * 1.01), k
    # spot check: patch slot (kn, m, f) of a left-edge cell reads (iz+m, ix+kn)
    c = np.nonzero(edge == 2)[0][5]
    for kn, m, fi in ((0, 0, 0), (8, -2, 3), (3, 2, 4)):
        v = m1[fi][iz[c] + m, ix[c] + kn] / (LA.Z0 if fi >= 2 else 1.0)
        assert f[c, (fi * 5 + m + 2) * 9 + kn] == LA.bf16(np.float32([v / s[c]]))[0]
    # own cell of level L-2
    assert f[c, 227] == LA.bf16(np.float32([m2[2][iz[c], ix[c]] / LA.Z0 / s[c]]))[0]
    # float32 emulation agrees with the exact definition within one bf16 step
    f32, s32 = LA.features(m1, m2, coefs, np.float32)
    assert np.all(np.abs(f32 - f) <= 2.0 ** -7 * np.abs(f) + 1e-30)
    assert np.allclose(s32, s, rtol=1e-6)


def test_features_outside_grid_are_zero():
    nz, nx = 30, 30
    m1 = [np.ones((nz, nx)) for _ in range(5)]
    m2 = [np.ones((nz, nx)) for _ in range(5)]
    coefs, _ = _coefs(nz, nx, 1)
    f, s = LA.features(m1, m2, coefs)
    # every patch of this grid stays inside it (9 inward, 2 tangential from
    # a cell at least 5 from every edge): all raw features are in-grid
    iz, ix, edge, _ = LA.band_cells(nz, nx)
    assert np.all(s > 0)
    # a grid too small for the 9-deep patch: the far slots fall outside
    m1[0][:, :] = 0
    m1[0][5, 5] = 1.0
    fz, _ = LA.features(m1, m2, coefs)
    assert fz.shape[0] == iz.size


def _small_case():
    nzp, nxp, dx, dt, fp = 24, 30, 10.0, 1e-3, 30.0
    abc_w = 20
    pad_w = abc_w + 5 + LA.R + LA.W
    vp = np.full((nzp, nxp), 2500.0)
    mw = synth.elastic_from_vp(synth.pad_edge(vp, pad_w))
    ms = synth.elastic_from_vp(synth.pad_edge(vp, LA.R + LA.W))
    cw = O.model_op(mw.rho, mw.lam, mw.mu, dx, dt)
    cs = O.model_op(ms.rho, ms.lam, ms.mu, dx, dt)
    nt = 120
    wav = O.ricker(nt, dt, fp, nsource=60)
    return dict(nz_phys=nzp, nx_phys=nxp, pad_wide=pad_w, abc_wide=abc_w, nt=nt, dt=dt, zs_phys=12, xs_phys=6,
                coefs_wide=cw, coefs_small=cs, wavelet=wav)


def test_teacher_targets_local_to_the_rim():
    case = _small_case()
    levels = [60, 90, 119]
    f, t = LA.teacher_pairs(levels=levels, **case)
    nzs, nxs = case["nz_phys"] + 2 * (LA.R + LA.W), case["nx_phys"] + 2 * (LA.R + LA.W)
    iz, ix, _, _ = LA.band_cells(nzs, nxs)
    dist = np.minimum(np.minimum(iz, nzs - 1 - iz), np.minimum(ix, nxs - 1 - ix)) - LA.R
    d = np.tile(dist, len(levels))
    assert np.all(t[d >= 5] == 0.0)  # beyond the stencil radius the band sees the same cells
    assert np.abs(t[d < 5]).max() > 1e-3  # the wave has reached the rim


def test_teacher_targets_are_the_rim_contribution():
    case = _small_case()
    L = 90
    _, t = LA.teacher_pairs(levels=[L], **case)
    pad_s = LA.R + LA.W
    nzs, nxs = case["nz_phys"] + 2 * pad_s, case["nx_phys"] + 2 * pad_s
    nzw, nxw = case["nz_phys"] + 2 * case["pad_wide"], case["nx_phys"] + 2 * case["pad_wide"]
    off = case["pad_wide"] - pad_s
    pw = O.make_params(nzw, nxw, case["nt"], 1.0, case["dt"], case["pad_wide"] + case["xs_phys"],
                       case["pad_wide"] + case["zs_phys"], 0, 0, 0, case["abc_wide"])
    hist = O.forward(pw, case["wavelet"], case["coefs_wide"], want_hist5=True)["hist5"]
    # rim values of level L-1 alone (the band grid zeroes them)
    rim = []
    for h in hist:
        a = np.array(h[off:off + nzs, off:off + nxs, L - 2])
        a[LA.R:nzs - LA.R, LA.R:nxs - LA.R] = 0
        rim.append(a)
    ps = O.make_params(nzs, nxs, case["nt"], 1.0, case["dt"], 0, 0, 0, 0, 0, 0)
    zero = [np.zeros((nzs, nxs)) for _ in range(5)]
    contrib = O.fwd_level(ps, zero, rim, case["coefs_small"])
    _, s = LA.features(LA.window([h[:, :, L - 2] for h in hist], off, nzs, nxs),
                       LA.window([h[:, :, L - 3] for h in hist], off, nzs, nxs), case["coefs_small"])
    iz, ix, _, _ = LA.band_cells(nzs, nxs)
    for fi in range(5):
        u = LA.Z0 if fi >= 2 else 1.0
        want = np.where(s > 0, contrib[fi][iz, ix] / np.where(s > 0, s * u, 1.0), 0.0)
        assert np.allclose(t[:, fi], want, rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_custom_damping_forward_matches_oracle_sponge():
    """The separable-damping forward used for band-sponge teachers equals the
    oracle's own sponge forward when the damping is that sponge."""
    case = _small_case()
    nzw = case["nz_phys"] + 2 * case["pad_wide"]
    nxw = case["nx_phys"] + 2 * case["pad_wide"]
    p = O.make_params(nzw, nxw, 80, 1.0, case["dt"], case["pad_wide"] + 6, case["pad_wide"] + 12, 0, 0, 0,
                      case["abc_wide"])
    kz, kx = LA.sponge_axes(nzw, nxw, case["abc_wide"])
    # the factors reproduce K on a field
    f = np.random.default_rng(1).standard_normal((nzw, nxw))
    assert np.allclose(O.abc_K(f, case["abc_wide"]), f * kz[:, None] * kx[None, :], rtol=1e-13, atol=0)
    h = LA._forward_custom_k(p, np.asarray(case["wavelet"]), case["coefs_wide"], kz, kx, 80)
    ref = O.forward(p, case["wavelet"], case["coefs_wide"], want_hist5=True)["hist5"]
    for a, b in zip(h, ref):
        assert O.rel_l2(a, b) <= 1e-12


def test_band_sponge_targets_local_to_the_rim():
    """With the band sponge in both runs the targets are still the exterior
    contribution only: zero beyond the stencil radius of the rim."""
    case = _small_case()
    levels = [60, 90, 119]
    f, t = LA.teacher_pairs(levels=levels, abc_band=8, **case)
    nzs, nxs = case["nz_phys"] + 2 * (LA.R + LA.W), case["nx_phys"] + 2 * (LA.R + LA.W)
    iz, ix, _, _ = LA.band_cells(nzs, nxs)
    dist = np.minimum(np.minimum(iz, nzs - 1 - iz), np.minimum(ix, nxs - 1 - ix)) - LA.R
    d = np.tile(dist, len(levels))
    assert np.abs(t[d >= 5]).max() <= 1e-12 * np.abs(t).max()
    assert np.abs(t[d < 5]).max() > 1e-3


def test_adjoint_features_are_forward_features_in_swapped_units():
    """An adjoint state's features divide the velocities (not the stresses)
    by Z0: an adjoint state (v Z0, s / Z0) has the features of the forward
    state (v, s), and the scales differ by nothing."""
    nz, nx = 36, 44
    m1, m2 = _rand_levels(nz, nx, 11)
    coefs, _ = _coefs(nz, nx, 12)
    f, s = LA.features(m1, m2, coefs)

    def swap(lev):
        return [a * LA.Z0 if fi < 2 else a / LA.Z0 for fi, a in enumerate(lev)]

    fa, sa = LA.features(swap(m1), swap(m2), coefs, adjoint=True)
    assert np.allclose(sa, s, rtol=1e-14, atol=0)
    assert np.all(np.abs(fa - f) <= 2.0 ** -8 * np.abs(f))
    # and the unit is not applied to a forward state's velocities
    fw, _ = LA.features(swap(m1), swap(m2), coefs)
    assert not np.allclose(fw, f)


def test_adjoint_rtm_loop_matches_oracle_rtm():
    """The per-level adjoint recursion of the learned-layer oracle (the
    gather-form rtm_op3_c_ on three-level histories, then the separable
    exact-adjoint sponge), with its imaging, reproduces the oracle's own RTM
    images with and without the adjoint sponge."""
    nz, nx, nt, dx, dt = 40, 52, 90, 10.0, 1e-3
    vp = synth.layered_vp(nz - 16, nx - 16, 3, vp_top=2000.0, vp_bot=3000.0, nlayers=2, anom_sigma=4.0)
    m = synth.elastic_from_vp(synth.pad_edge(vp, 8))
    coefs = O.model_op(m.rho, m.lam, m.mu, dx, dt)
    wav = O.ricker(nt, dt, 30.0, nsource=50)
    rec = np.random.default_rng(4).standard_normal((5, nx - 16, nt))
    for adj_abc in (0, 1):
        p = O.make_params(nz, nx, nt, dx, dt, 20, 12, 10, 8, nx - 16, 8, adj_abc=adj_abc)
        got = LA.rtm_adjoint_layer(p, wav, coefs, rec)
        ref = O.rtm(p, wav, coefs, rec)
        for a, b in zip(got, ref):
            assert O.rel_l2(a, b) <= 1e-12, adj_abc


def test_adjoint_teacher_targets_are_the_rim_contribution():
    """Adjoint teacher targets: the wide and band-grid updates share the
    injection and level L+2, so with no band sponge the target is the wide
    grid's adjoint stencil applied to the rim values of level L+1 alone
    (zero beyond the stencil radius of the rim), in the adjoint's units.
    (The band grid's own adjoint stencil reads nothing from its rim: the
    gather form masks the coefficients of cells that the forward operators
    never update.)"""
    case = _small_case()
    case["nt"] = 100
    case["wavelet"] = case["wavelet"][:100]
    levels = [30, 55]
    f, t = LA.teacher_pairs_adj(levels=levels, zr_phys=2, **case)
    pad_s = LA.R + LA.W
    nzs, nxs = case["nz_phys"] + 2 * pad_s, case["nx_phys"] + 2 * pad_s
    nzw, nxw = case["nz_phys"] + 2 * case["pad_wide"], case["nx_phys"] + 2 * case["pad_wide"]
    off = case["pad_wide"] - pad_s
    pw = O.make_params(nzw, nxw, case["nt"], 1.0, case["dt"], case["pad_wide"] + case["xs_phys"],
                       case["pad_wide"] + case["zs_phys"], case["pad_wide"] + 2, case["pad_wide"],
                       case["nx_phys"], case["abc_wide"])
    rec = O.forward(pw, case["wavelet"], case["coefs_wide"])["rec"]
    kz, kx = LA.sponge_axes(nzw, nxw, case["abc_wide"])
    wide = LA.adjoint_run(case["coefs_wide"], kz, kx, rec, pw.zr, pw.rx0, case["nt"])
    iz, ix, _, _ = LA.band_cells(nzs, nxs)
    ones_z, ones_x = np.ones(nzw), np.ones(nxw)
    zero = [np.zeros((nzw, nxw)) for _ in range(5)]
    rows = []
    for L in sorted(levels, reverse=True):  # the stacking order
        rim = []
        for a in wide[L + 1]:
            w = np.zeros((nzw, nxw))
            w[off:off + nzs, off:off + nxs] = a[off:off + nzs, off:off + nxs]
            w[off + LA.R:off + nzs - LA.R, off + LA.R:off + nxs - LA.R] = 0
            rim.append(w)
        cw = LA.adjoint_level(zero, rim, zero, case["coefs_wide"], ones_z, ones_x)
        contrib = [a[off:off + nzs, off:off + nxs] for a in cw]
        _, s = LA.features(LA.window(wide[L + 1], off, nzs, nxs), LA.window(wide[L + 2], off, nzs, nxs),
                           case["coefs_small"], adjoint=True)
        rows.append(np.stack([np.where(s > 0, contrib[fi][iz, ix] / np.where(s > 0, s * (LA.Z0 if fi < 2 else 1.0),
                                                                             1.0), 0.0) for fi in range(5)], 1))
    want = np.concatenate(rows)
    assert np.abs(want).max() > 1e-3  # the adjoint wavefield reached the rim
    assert np.allclose(t, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())
    dist = np.minimum(np.minimum(iz, nzs - 1 - iz), np.minimum(ix, nxs - 1 - ix)) - LA.R
    assert np.all(t[np.tile(dist, len(levels)) >= 5] == 0.0)


def test_adjoint_teacher_residual_data():
    """With an observed model equal to the adjoint's model the residual data
    vanish, and with it every feature and target; the residual of a model
    with an anomaly equals the difference of the two models' recorded data
    (the injected data enter linearly: pairs of obs-data minus pairs of the
    model's own data, before normalisation, cancel to the residual's)."""
    case = _small_case()
    case["nt"] = 80
    case["wavelet"] = case["wavelet"][:80]
    levels = [20, 50]
    f0, t0 = LA.teacher_pairs_adj(levels=levels, zr_phys=2, coefs_wide_obs=case["coefs_wide"], **case)
    assert np.all(f0[:, :LA.NRAW] == 0) and np.all(t0 == 0)
    # the injected data: obs minus the model's own data, as recorded
    nzw, nxw = case["nz_phys"] + 2 * case["pad_wide"], case["nx_phys"] + 2 * case["pad_wide"]
    vp = np.full((case["nz_phys"], case["nx_phys"]), 2500.0)
    vp[8:14, 10:18] = 3000.0
    mo = synth.elastic_from_vp(synth.pad_edge(vp, case["pad_wide"]))
    co = O.model_op(mo.rho, mo.lam, mo.mu, 10.0, case["dt"])
    f1, t1 = LA.teacher_pairs_adj(levels=levels, zr_phys=2, coefs_wide_obs=co, **case)
    assert np.abs(f1[:, :LA.NRAW]).max() > 0
    pw = O.make_params(nzw, nxw, case["nt"], 1.0, case["dt"], case["pad_wide"] + case["xs_phys"],
                       case["pad_wide"] + case["zs_phys"], case["pad_wide"] + 2, case["pad_wide"], case["nx_phys"],
                       case["abc_wide"])
    r_own = O.forward(pw, case["wavelet"], case["coefs_wide"])["rec"]
    r_obs = O.forward(pw, case["wavelet"], co)["rec"]
    kz, kx = LA.sponge_axes(nzw, nxw, case["abc_wide"])
    lev = LA.adjoint_run(case["coefs_wide"], kz, kx, r_obs - r_own, pw.zr, pw.rx0, case["nt"])
    off = case["pad_wide"] - (LA.R + LA.W)
    nzs, nxs = case["nz_phys"] + 2 * (LA.R + LA.W), case["nx_phys"] + 2 * (LA.R + LA.W)
    fa, _ = LA.features(LA.window(lev[51], off, nzs, nxs), LA.window(lev[52], off, nzs, nxs), case["coefs_small"],
                        adjoint=True)
    assert np.array_equal(f1[:fa.shape[0]], fa)  # level 50 comes first (decreasing order)
