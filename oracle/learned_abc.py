"""Learned absorbing layer oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` import this module. It restates, in numpy, what the HIP
learned-layer kernels compute (elastic_amd/csrc/learned_abc.hip), so that the
GPU results can be checked:

* band geometry: the 8-cell band inside the 5-cell rim of a grid, cell order
  top rows, bottom rows (both over all inner columns), left columns, right
  columns (over the remaining rows);
* features of a band cell from levels L-1 and L-2 (a 9 x 5 patch of the five
  fields of L-1 along the inward normal, the cell's own fields of L-2, divided
  by their max-abs scale s; one-hot edge and depth; the cell's P and S Courant
  numbers and P impedance from the coefficients), rounded to bf16; ``dtype=np.float32`` reproduces the kernel's float32 operation
  order bit for bit, ``np.float64`` is the exact definition;
* the network 245 -> 32 -> 32 -> 5 with exact GELU on bf16-rounded weights,
  evaluated in FP64, with the loss 0.5 * mean((y - t)^2) and its gradient
  (hand-written backward pass);
* teacher-forced training pairs: the band grid's levels L-1, L-2 are windows
  of a wide-sponge run, the target is (wide(L) - plain band-grid update(L)) / s
  per field unit, using the 2D oracle's forward and single-level update;
* the same for the adjoint propagation of the wide run's receiver data
  (levels L+1, L+2, the adjoint's field units), each adjoint level the
  oracle's gather-form rtm_op3_c_ on a three-level history followed by the
  exact adjoint of the sponge.

The reference has no learned boundary: it only has the sponge of
W/abc-mod.f90:59-78. The network is this project's addition, so there is no
reference vector for it; the pins are torch's float64 autograd (an
independent implementation of the same network), central finite differences
of the loss, and a brute-force construction of the band from its distance
definition (tests/test_learned_abc_pins.py).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

from . import abc_K, forward, fwd_level, image_S, image_adj_level, make_params, rtm_op3_c

R, W, NF, KP, H, NO = 5, 8, 245, 248, 32, 5
NRAW = 230
Z0 = 6.0e6
OW1 = 0
OB1 = NF * H
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H * NO
NPARAM = OB3 + NO


def band_cells(nz, nx):
    """(iz, ix, edge, depth) of every band cell, in kernel cell order."""
    cols = np.arange(R, nx - R)
    rows = np.arange(R + W, nz - R - W)
    out = []
    for e in range(4):
        for d in range(W):
            if e == 0:
                out.append((np.full(cols.size, R + d), cols))
            elif e == 1:
                out.append((np.full(cols.size, nz - R - 1 - d), cols))
            elif e == 2:
                out.append((rows, np.full(rows.size, R + d)))
            else:
                out.append((rows, np.full(rows.size, nx - R - 1 - d)))
    iz = np.concatenate([a for a, _ in out])
    ix = np.concatenate([b for _, b in out])
    per = [cols.size] * (2 * W) + [rows.size] * (2 * W)
    edge = np.concatenate([np.full(n, e) for (e, n) in zip(np.repeat(np.arange(4), W), per)])
    depth = np.concatenate([np.full(n, d) for (d, n) in zip(np.tile(np.arange(W), 4), per)])
    return iz.astype(np.int64), ix.astype(np.int64), edge.astype(np.int64), depth.astype(np.int64)


def bf16(x):
    """Round float32 values to the nearest bf16 (ties to even), as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def _unit(f, dtype):
    if f < 2:
        return None
    return dtype(1.0) / dtype(Z0) if dtype == np.float32 else 1.0 / Z0


def _field_unit(f, dtype, adjoint):
    """Scale of field f in the features: the stresses of a forward state,
    the velocities of an adjoint state (the larger field pair of each, by
    about the impedance Z0) are divided by Z0."""
    return _unit(2, dtype) if (f < 2 if adjoint else f >= 2) else None


def raw_patch(lev_m1, lev_m2, cells, dtype=np.float64, adjoint=False):
    """Raw features 0..229 of the band cells: (ncell, 230) in dtype.
    lev_m1 / lev_m2: five (nz, nx) arrays each (levels L-1, L-2 of a forward
    state; L+1, L+2 of an adjoint state with adjoint=True)."""
    iz, ix, edge, _ = cells
    nz, nx = lev_m1[0].shape
    dz = np.select([edge == 0, edge == 1], [1, -1], 0)
    dx = np.select([edge == 2, edge == 3], [1, -1], 0)
    tz = (edge >= 2).astype(np.int64)
    tx = (edge < 2).astype(np.int64)
    out = np.zeros((iz.size, NRAW), dtype)
    for kn in range(9):
        for m in range(-2, 3):
            z = iz + kn * dz + m * tz
            x = ix + kn * dx + m * tx
            ok = (z >= 0) & (z < nz) & (x >= 0) & (x < nx)
            zc, xc = np.clip(z, 0, nz - 1), np.clip(x, 0, nx - 1)
            for f in range(5):
                v = np.where(ok, np.asarray(lev_m1[f], dtype)[zc, xc], dtype(0))
                u = _field_unit(f, dtype, adjoint)
                if u is not None:
                    v = (v * u).astype(dtype)
                out[:, (f * 5 + m + 2) * 9 + kn] = v
    for f in range(5):
        v = np.asarray(lev_m2[f], dtype)[iz, ix]
        u = _field_unit(f, dtype, adjoint)
        out[:, 225 + f] = v if u is None else (v * u).astype(dtype)
    return out


def material(coefs, cells, dtype=np.float64):
    """(vp dt/dx, vs dt/dx, rho vp / Z0) of the band cells from the dt/dx-scaled
    coefficients m1_x, aux_m3, aux_m2m3."""
    iz, ix = cells[0], cells[1]
    m1 = np.asarray(coefs[0], dtype)[iz, ix]
    m3 = np.asarray(coefs[3], dtype)[iz, ix]
    m23 = np.asarray(coefs[4], dtype)[iz, ix]
    cp = np.sqrt((m1 * m23).astype(dtype))
    cs = np.sqrt((m1 * m3).astype(dtype))
    with np.errstate(divide="ignore", invalid="ignore"):
        zp = np.where(m1 > 0, np.sqrt((m23 / np.where(m1 > 0, m1, 1)).astype(dtype)) * _unit(2, dtype), 0)
    return cp, cs, zp.astype(dtype)


def features(lev_m1, lev_m2, coefs, dtype=np.float64, cells=None, adjoint=False):
    """Normalised, bf16-rounded features (ncell, KP) and scales (ncell,)."""
    if cells is None:
        cells = band_cells(*lev_m1[0].shape)
    raw = raw_patch(lev_m1, lev_m2, cells, dtype, adjoint)
    s = np.abs(raw).max(axis=1)
    live = s > 1e-30
    inv = np.zeros_like(s)
    inv[live] = dtype(1.0) / s[live]
    x = np.zeros((raw.shape[0], KP), dtype)
    x[:, :NRAW] = (raw * inv[:, None]).astype(dtype)
    _, _, edge, depth = cells
    rows = np.arange(raw.shape[0])
    x[rows, NRAW + edge] = 1.0
    x[rows, NRAW + 4 + depth] = 1.0
    cp, cs, zp = material(coefs, cells, dtype)
    x[:, 242], x[:, 243], x[:, 244] = cp, cs, zp
    feat = bf16(x.astype(np.float32))
    return feat, np.where(live, s, 0).astype(dtype)


def unpack(params):
    p = bf16(np.asarray(params, np.float32)).astype(np.float64)
    return (p[OW1:OB1].reshape(NF, H), p[OB1:OW2], p[OW2:OB2].reshape(H, H), p[OB2:OW3],
            p[OW3:OB3].reshape(H, NO), p[OB3:NPARAM])


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def mlp(params, feat):
    """FP64 network output (n, 5) on bf16-rounded weights; feat (n, KP)."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    x = np.asarray(feat, np.float64)[:, :NF]
    h1 = gelu(x @ W1 + b1)
    h2 = gelu(h1 @ W2 + b2)
    return h2 @ W3 + b3


def loss_grad(params, feat, tgt):
    """0.5 * mean((y - t)^2) over n*5 values and its gradient w.r.t. the
    (rounded) weights, in the header's parameter order."""
    W1, b1, W2, b2, W3, b3 = unpack(params)
    x = np.asarray(feat, np.float64)[:, :NF]
    t = np.asarray(tgt, np.float64)
    z1 = x @ W1 + b1
    h1 = gelu(z1)
    z2 = h1 @ W2 + b2
    h2 = gelu(z2)
    y = h2 @ W3 + b3
    cnt = y.size
    d = (y - t) / cnt
    loss = 0.5 * np.sum((y - t) ** 2) / cnt
    g = np.zeros(NPARAM)
    g[OW3:OB3] = (h2.T @ d).ravel()
    g[OB3:NPARAM] = d.sum(0)
    dz2 = (d @ W3.T) * gelu_d(z2)
    g[OW2:OB2] = (h1.T @ dz2).ravel()
    g[OB2:OW3] = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * gelu_d(z1)
    g[OW1:OB1] = (x.T @ dz1).ravel()
    g[OB1:OW2] = dz1.sum(0)
    return loss, g


def window(lev, off, nz, nx):
    """Band-grid window of a wide level with the 5-cell rim zeroed."""
    out = []
    for a in lev:
        w = np.array(a[off:off + nz, off:off + nx], dtype=np.float64)
        w[:R, :] = 0
        w[nz - R:, :] = 0
        w[:, :R] = 0
        w[:, nx - R:] = 0
        out.append(w)
    return out


def sponge_axes(n1, n2, abc):
    """Per-row and per-column factors of the sponge K on an (n1, n2) grid
    (K(field) = field * kz[iz] * kx[ix], W/abc-mod.f90:59-78)."""
    if abc <= 0:
        return np.ones(n1), np.ones(n2)
    k = abc_K(np.ones((n1, n2)), abc)
    kx = k[n1 // 2, :].copy()
    kz = k[:, n2 // 2] / kx[n2 // 2]
    return kz, kx


def _forward_custom_k(p, wavelet, coefs, kz, kx, nt):
    """Forward levels 1..nt (W/rtm-mod-adj.f90:178-260) with a separable
    damping kz[iz]*kx[ix] after every update from level 3 on; returns the
    five (nz, nx, nt) histories."""
    nz, nx = p.nz, p.nx
    hist = [np.zeros((nz, nx, nt), order="F") for _ in range(5)]
    lev = {0: [np.zeros((nz, nx)) for _ in range(5)], 1: [np.zeros((nz, nx)) for _ in range(5)]}
    for L in range(2, nt + 1):
        new = fwd_level(p, lev[L - 2], lev[L - 1], coefs, apply_k=False, srcamp=p.dt * wavelet[L - 2])
        if L >= 3:
            new = [(a * kz[:, None]) * kx[None, :] for a in new]
        for f in range(5):
            hist[f][:, :, L - 1] = new[f]
        lev[L] = new
        lev.pop(L - 2)
    return hist


def teacher_pairs(nz_phys, nx_phys, pad_wide, abc_wide, nt, dt, zs_phys, xs_phys, coefs_wide,
                  coefs_small, wavelet, levels, abc_band=0):
    """FP64 (features, targets) of the band cells at each level in levels
    (features of levels L-1, L-2 windowed from the wide run), stacked in
    level order like labc_collect. abc_band > 0: the band grid's sponge of
    that width acts in the band grid's update, and the wide run is damped by
    the same factors at the band's position (clamped outside the band grid)
    on top of its own sponge."""
    pad_s = R + W
    nzw, nxw = nz_phys + 2 * pad_wide, nx_phys + 2 * pad_wide
    nzs, nxs = nz_phys + 2 * pad_s, nx_phys + 2 * pad_s
    off = pad_wide - pad_s
    pw = make_params(nzw, nxw, nt, 1.0, dt, pad_wide + xs_phys, pad_wide + zs_phys, 0, 0, 0, abc_wide)
    if abc_band <= 0:
        hist = forward(pw, wavelet, coefs_wide, want_hist5=True)["hist5"]
    else:
        kzw, kxw = sponge_axes(nzw, nxw, abc_wide)
        kzb, kxb = sponge_axes(nzs, nxs, abc_band)
        kz = kzw * kzb[np.clip(np.arange(nzw) - off, 0, nzs - 1)]
        kx = kxw * kxb[np.clip(np.arange(nxw) - off, 0, nxs - 1)]
        hist = _forward_custom_k(pw, np.asarray(wavelet, np.float64), coefs_wide, kz, kx, nt)
    ps = make_params(nzs, nxs, nt, 1.0, dt, pad_s + xs_phys, pad_s + zs_phys, 0, 0, 0, abc_band)
    cells = band_cells(nzs, nxs)
    iz, ix = cells[0], cells[1]
    feats, tgts = [], []
    for L in sorted(levels):
        if L < 3:
            continue
        m1 = window([h[:, :, L - 2] for h in hist], off, nzs, nxs)
        m2 = window([h[:, :, L - 3] for h in hist], off, nzs, nxs)
        f, s = features(m1, m2, coefs_small, np.float64, cells)
        plain = fwd_level(ps, m2, m1, coefs_small, apply_k=abc_band > 0, srcamp=dt * wavelet[L - 2])
        t = np.zeros((iz.size, NO))
        for fi in range(5):
            wide = hist[fi][iz + off, ix + off, L - 1]
            u = Z0 if fi >= 2 else 1.0
            with np.errstate(divide="ignore", invalid="ignore"):
                t[:, fi] = np.where(s > 0, (wide - plain[fi][iz, ix]) / (s * u), 0.0)
        feats.append(f)
        tgts.append(t)
    return np.concatenate(feats), np.concatenate(tgts)


def adjoint_level(inj, p1, p2, coefs, kz, kx):
    """One adjoint level K (inj + X(L+2) + Op^T X(L+1)) (the step of
    W/rtmlib-modC.c:120-253 in gather form, followed by the exact adjoint of
    a separable sponge kz[iz]*kx[ix]): the oracle's rtm_op3_c_ on the
    three-level history (inj, X(L+1), X(L+2)), which computes level 0 only."""
    nz, nx = p1[0].shape
    h = []
    for f in range(5):
        a = np.zeros((nz, nx, 3), order="F")
        a[:, :, 0], a[:, :, 1], a[:, :, 2] = inj[f], p1[f], p2[f]
        h.append(a)
    out = rtm_op3_c(h, coefs, form="gather", dtype=np.float64)
    return [(o[:, :, 0] * kz[:, None]) * kx[None, :] for o in out]


def adjoint_run(coefs, kz, kx, rec, zr, rx0, nt, on_level=None):
    """Adjoint levels nt..1 of receiver data rec (5, nrx, nt) injected on row
    zr, columns rx0.. at the level of the same index (level L gets
    rec[:, :, L-1]), levels nt+1, nt+2 zero, each level
    K (inj + X(L+2) + Op^T X(L+1)) (the recursion of the oracle's RTM,
    oracle/rtm_oracle.cpp rtm). on_level(L, level, levels), when given, sees
    each new level before it is stored and may change it in place (imaging,
    a learned correction). Returns {L: five (nz, nx)}."""
    nz, nx = np.asarray(coefs[0]).shape
    nrx = rec.shape[1]
    lev = {nt + 1: [np.zeros((nz, nx)) for _ in range(5)], nt + 2: [np.zeros((nz, nx)) for _ in range(5)]}
    for L in range(nt, 0, -1):
        inj = [np.zeros((nz, nx)) for _ in range(5)]
        for f in range(5):
            inj[f][zr, rx0:rx0 + nrx] = rec[f, :, L - 1]
        cur = adjoint_level(inj, lev[L + 1], lev[L + 2], coefs, kz, kx)
        if on_level is not None:
            on_level(L, cur, lev)
        lev[L] = cur
    return lev


def rtm_adjoint_layer(p, wavelet, coefs, rec, params=None):
    """Single-shot RTM images (drho, dlambda, dmu) as the oracle's rtm
    (forward with the sponge p.abc and a full Vx/Vz history, the adjoint
    recursion with the exact-adjoint sponge when p.adj_abc, the Lame imaging
    condition at every level) with, when params is given, the learned layer
    in the adjoint: after a level is formed and imaged, its band cells get
    s * NN(features of levels L+1, L+2 of the adjoint state) (field units
    applied) and its rim is cleared."""
    nz, nx, nt = p.nz, p.nx, p.nt
    fw = forward(p, wavelet, coefs, want_hist=True)
    hvx, hvz = fw["hist_vx"], fw["hist_vz"]
    kz, kx = sponge_axes(nz, nx, p.abc) if p.adj_abc else (np.ones(nz), np.ones(nx))
    cells = band_cells(nz, nx)
    iz, ix = cells[0], cells[1]
    dm = [None]

    def on_level(L, cur, lev):
        S = image_S(hvx[:, :, L - 1], hvz[:, :, L - 1], hvx[:, :, L] if L < nt else None,
                    hvz[:, :, L] if L < nt else None, p.dx, p.dt)
        dm[0] = image_adj_level(cur, S, 1.0, dm[0])
        if params is None:
            return
        f, sc = features(lev[L + 1], lev[L + 2], coefs, np.float64, cells, adjoint=True)
        y = mlp(params, f)
        for fi in range(5):
            cur[fi][iz, ix] += y[:, fi] * sc * (Z0 if fi < 2 else 1.0)
            cur[fi][:R, :] = 0
            cur[fi][nz - R:, :] = 0
            cur[fi][:, :R] = 0
            cur[fi][:, nx - R:] = 0
        lev.pop(L + 3, None)

    adjoint_run(coefs, kz, kx, np.asarray(rec, np.float64), p.zr, p.rx0, nt, on_level)
    return tuple(dm[0])


def teacher_pairs_adj(nz_phys, nx_phys, pad_wide, abc_wide, nt, dt, zs_phys, xs_phys, zr_phys, coefs_wide,
                      coefs_small, wavelet, levels, abc_band=0, coefs_wide_obs=None):
    """FP64 (features, targets) of the adjoint propagation, stacked in
    decreasing level order like labc_collect_adj: the wide grid's forward run
    (sponges as in teacher_pairs) records receivers on row zr_phys of every
    physical column; its adjoint run (the same separable sponge, exact
    adjoint) back-propagates them; at each listed level L the band grid's
    levels L+1, L+2 are windows of it, the features those of an adjoint
    state, the target (wide(L) - band-grid adjoint update(L)) / (s * unit)
    with the velocities' unit Z0. With coefs_wide_obs the injected data
    are that model's wide run minus the run of coefs_wide (the model of the
    adjoint runs), as in a migration."""
    pad_s = R + W
    nzw, nxw = nz_phys + 2 * pad_wide, nx_phys + 2 * pad_wide
    nzs, nxs = nz_phys + 2 * pad_s, nx_phys + 2 * pad_s
    off = pad_wide - pad_s
    pw = make_params(nzw, nxw, nt, 1.0, dt, pad_wide + xs_phys, pad_wide + zs_phys, 0, 0, 0, abc_wide)
    kzw, kxw = sponge_axes(nzw, nxw, abc_wide)
    kzb, kxb = sponge_axes(nzs, nxs, abc_band)
    if abc_band > 0:
        kzw = kzw * kzb[np.clip(np.arange(nzw) - off, 0, nzs - 1)]
        kxw = kxw * kxb[np.clip(np.arange(nxw) - off, 0, nxs - 1)]
    hist = _forward_custom_k(pw, np.asarray(wavelet, np.float64), coefs_wide, kzw, kxw, nt)
    zrw = pad_wide + zr_phys
    rec = np.stack([h[zrw, pad_wide:pad_wide + nx_phys, :] for h in hist])
    if coefs_wide_obs is not None:
        ho = _forward_custom_k(pw, np.asarray(wavelet, np.float64), coefs_wide_obs, kzw, kxw, nt)
        rec = np.stack([h[zrw, pad_wide:pad_wide + nx_phys, :] for h in ho]) - rec
    wide = adjoint_run(coefs_wide, kzw, kxw, rec, zrw, pad_wide, nt)
    cells = band_cells(nzs, nxs)
    iz, ix = cells[0], cells[1]
    zrs = pad_s + zr_phys
    feats, tgts = [], []
    for L in sorted(levels, reverse=True):
        p1 = window(wide[L + 1], off, nzs, nxs)
        p2 = window(wide[L + 2], off, nzs, nxs)
        f, s = features(p1, p2, coefs_small, np.float64, cells, adjoint=True)
        inj = [np.zeros((nzs, nxs)) for _ in range(5)]
        for fi in range(5):
            inj[fi][zrs, pad_s:pad_s + nx_phys] = rec[fi, :, L - 1]
        band = adjoint_level(inj, p1, p2, coefs_small, kzb, kxb)
        t = np.zeros((iz.size, NO))
        for fi in range(5):
            w = wide[L][fi][iz + off, ix + off]
            u = Z0 if fi < 2 else 1.0
            with np.errstate(divide="ignore", invalid="ignore"):
                t[:, fi] = np.where(s > 0, (w - band[fi][iz, ix]) / (s * u), 0.0)
        feats.append(f)
        tgts.append(t)
    return np.concatenate(feats), np.concatenate(tgts)
