"""Independent pins of three oracle functions the whole-chain dot test alone
cannot pin (it uses the same S1..S6 in both directions):

* ``image_S`` (S1..S6, W/image-mod-rtm.f90:26-67, T :268-299, vxdx :301-344,
  vzdz :346-390, vxdz :392-435, vzdx :437-480): exact half-cell derivatives
  of polynomials of degree <= 9 on each operator's own index range, divided
  by dx; S1 of a field linear in the level is slope/dt and S1 = S2 = 0 at
  the last level;
* ``image_adj_level`` (W/image-mod-rtm.f90:154-162): its response to every
  one-hot pairing of an adjoint field with one of S1..S6, read off those
  lines;
* the forward loop's first levels and source timing (W/rtm-mod-adj.f90:
  185-252, explosive_source W/source-mod.f90:47-53): level 1 is zero, level
  2 is dt*s(1) at the source without K, level 3's single taps read off Ax
  and Bz (W/rtmlib-mod.f90:229-237, :344-352), and with m1 = 0 the stresses
  at the source follow sigma(L) = K(sigma(L-2) + dt*s(L-1)) in closed form.

It also pins the oracle's restatement of the R/Src adjoint with the sponge
(R/Src/rtm-mod-adj.f90:365-475, ``rtm_rsrc`` / ``rsrc_adjoint``): without a
sponge it is the drop-in path's linear map, and with every operator zero the
powers of K on each stored level are counted off the loop.

Each pin is shown to reject a planted slip of the kind it guards against (a
wrong S6 tap shape, a missing 1/dx, a shrunken range, a dropped factor 2, a
one-level source shift, K applied at level 2, shifted taps, K once per
level).
"""
import numpy as np
import pytest

import oracle as O
from elastic_amd import synth

DX, DT = 10.0, 1e-3
NZ, NX = 23, 27
TOL = 1e-10


def _p(deg, t):
    """Test polynomial of degree deg in metres, and its derivative."""
    u = (t - 130.0) / 100.0
    return u ** deg + 0.3 * u ** max(deg - 2, 0)


def _dp(deg, t):
    u = (t - 130.0) / 100.0
    e = max(deg - 2, 0)
    d = deg * u ** (deg - 1) if deg >= 1 else 0.0 * u
    d2 = e * u ** (e - 1) if e >= 1 else 0.0 * u
    return (d + 0.3 * d2) / 100.0


def _mask(lo, nz=NZ, nx=NX):
    """Cells updated by an imaging derivative whose 1-based loops run
    lo..n-5 along both axes (0-based lo-1..n-6)."""
    m = np.zeros((nz, nx), bool)
    m[lo - 1:nz - 5, lo - 1:nx - 5] = True
    return m


MASK_PLUS = _mask(5)   # vxdx, vzdz: do ix=5,nx-5; do iz=5,nz-5
MASK_MINUS = _mask(6)  # vxdz, vzdx: do ix=6,nx-5; do iz=6,nz-5
ZC = np.arange(NZ, dtype=np.float64)[:, None] * DX + 0.0 * np.arange(NX)[None, :]
XC = 0.0 * np.arange(NZ)[:, None] + np.arange(NX, dtype=np.float64)[None, :] * DX


def _close(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b)))) <= TOL * max(1.0, float(np.max(np.abs(b))))


# --------------------------------------------------------------------------- image_S
def _check_S4(S4, deg):
    """vxdx of Vx = p(x): dp/dx at the half cell ix + 1/2, zero off range."""
    return _close(S4, np.where(MASK_PLUS, _dp(deg, XC + DX / 2), 0.0))


def _check_S5(S5, deg):
    return _close(S5, np.where(MASK_PLUS, _dp(deg, ZC + DX / 2), 0.0))


def _check_S6(S6, dz, dx_):
    """vxdz of Vx = r(z) plus vzdx of Vz = s(x): derivatives at iz - 1/2 and
    ix - 1/2 on the H/J range."""
    return _close(S6, np.where(MASK_MINUS, _dp(dz, ZC - DX / 2) + _dp(dx_, XC - DX / 2), 0.0))


@pytest.mark.parametrize("deg", [1, 2, 5, 9])
def test_image_S_x_derivative_of_vx(deg):
    vx = np.asfortranarray(_p(deg, XC))
    S = O.image_S(vx, np.zeros_like(vx), vx, np.zeros_like(vx), DX, DT)
    assert _check_S4(S[3], deg)
    assert np.all(S[4] == 0) and _close(S[2], S[3])
    assert _close(S[5], 0.0 * vx)  # no z variation: vxdz = 0
    assert np.all(S[0] == 0) and np.all(S[1] == 0)  # level-constant: S1 = S2 = 0


@pytest.mark.parametrize("deg", [1, 3, 9])
def test_image_S_z_derivative_of_vz(deg):
    vz = np.asfortranarray(_p(deg, ZC))
    S = O.image_S(np.zeros_like(vz), vz, None, None, DX, DT)
    assert _check_S5(S[4], deg)
    assert np.all(S[3] == 0) and _close(S[2], S[4])
    assert _close(S[5], 0.0 * vz)


@pytest.mark.parametrize("dz,dx_", [(1, 1), (4, 7), (9, 9)])
def test_image_S_shear_term(dz, dx_):
    vx = np.asfortranarray(_p(dz, ZC))
    vz = np.asfortranarray(_p(dx_, XC))
    S = O.image_S(vx, vz, None, None, DX, DT)
    assert _check_S6(S[5], dz, dx_)
    assert _close(S[3], 0.0 * vx) and _close(S[4], 0.0 * vx)


def test_image_S_time_derivative_and_last_level():
    rng = np.random.default_rng(4)
    vx = np.asfortranarray(rng.uniform(-1, 1, (NZ, NX)))
    vz = np.asfortranarray(rng.uniform(-1, 1, (NZ, NX)))
    bx, bz = 0.375, -1.25  # slopes per level
    S = O.image_S(vx, vz, vx + bx, vz + bz, DX, DT)
    assert _close(S[0], np.full((NZ, NX), bx / DT)) and _close(S[1], np.full((NZ, NX), bz / DT))
    last = O.image_S(vx, vz, None, None, DX, DT)
    assert np.all(last[0] == 0) and np.all(last[1] == 0)
    # the spatial terms do not depend on the next level
    for k in range(2, 6):
        assert np.array_equal(S[k], last[k])


def test_image_S_pins_reject_planted_slips():
    deg = 5
    vx = np.asfortranarray(_p(deg, XC))
    S4 = O.image_S(vx, 0 * vx, None, None, DX, DT)[3]
    assert _check_S4(S4, deg)
    assert not _check_S4(S4 * DX, deg)  # derivative not divided by dx
    shrunk = S4.copy()
    shrunk[:, 4] = 0.0  # loop starting one column late
    assert not _check_S4(shrunk, deg)
    S6 = O.image_S(np.asfortranarray(_p(4, ZC)), np.asfortranarray(_p(7, XC)), None, None, DX, DT)[5]
    assert _check_S6(S6, 4, 7)
    # S+ tap shape for vxdz (derivative at iz + 1/2 instead of iz - 1/2)
    plus = np.where(MASK_MINUS, np.roll(S6, -1, axis=0), 0.0)
    assert not _check_S6(plus, 4, 7)
    plus_x = np.where(MASK_MINUS, np.roll(S6, -1, axis=1), 0.0)
    assert not _check_S6(plus_x, 4, 7)


# --------------------------------------------------------------------------- image_adj_level
ADJ = ["Vx", "Vz", "sxx", "szz", "sxz"]
# (adjoint field, S index 1..6) -> (d drho, d dlambda, d dmu) per unit product,
# read off W/image-mod-rtm.f90:154-162; every other pairing contributes 0.
TABLE = {
    ("Vx", 1): (1.0, 0.0, 0.0),
    ("Vz", 2): (1.0, 0.0, 0.0),
    ("sxx", 3): (0.0, -1.0, 0.0),
    ("szz", 3): (0.0, -1.0, 0.0),
    ("sxx", 4): (0.0, 0.0, -2.0),
    ("szz", 5): (0.0, 0.0, -2.0),
    ("sxz", 6): (0.0, 0.0, -1.0),
}


def _one_hot_response(fn, a=1.5, s=0.25, w=3.0):
    """Response of an imaging-level function to each (field, S) pair at one
    cell of a 3 x 4 grid: {(field, k): (drho, dlambda, dmu)} / (w a s)."""
    out = {}
    for fi, f in enumerate(ADJ):
        for k in range(1, 7):
            adj = [np.zeros((3, 4)) for _ in range(5)]
            S = [np.zeros((3, 4)) for _ in range(6)]
            adj[fi][1, 2] = a
            S[k - 1][1, 2] = s
            dm = fn(adj, S, w)
            for d in dm:
                d_off = d.copy()
                d_off[1, 2] = 0
                assert np.all(d_off == 0)
            out[(f, k)] = tuple(float(d[1, 2]) / (w * a * s) for d in dm)
    return out


def _matches_table(resp):
    return all(resp[key] == TABLE.get(key, (0.0, 0.0, 0.0)) for key in resp)


def test_image_adj_level_one_hot_table():
    resp = _one_hot_response(lambda adj, S, w: O.image_adj_level(adj, S, w))
    assert _matches_table(resp), resp


def test_image_adj_level_accumulates():
    rng = np.random.default_rng(2)
    adj = [rng.uniform(-1, 1, (5, 6)) for _ in range(5)]
    S = [rng.uniform(-1, 1, (5, 6)) for _ in range(6)]
    dm0 = [rng.uniform(-1, 1, (5, 6)) for _ in range(3)]
    a = O.image_adj_level(adj, S, 2.0, dm=dm0)
    b = O.image_adj_level(adj, S, 2.0)
    for x, y, z in zip(a, b, dm0):
        assert np.allclose(x, y + z, rtol=0, atol=1e-14)


def test_image_adj_level_table_rejects_planted_slips():
    good = _one_hot_response(lambda adj, S, w: O.image_adj_level(adj, S, w))

    def drop_factor_2(adj, S, w):
        d = O.image_adj_level(adj, S, w)
        # dmu -= sxx * S4 (factor 2 dropped): undo half of that term
        d[2] = d[2] + w * adj[2] * S[3]
        return d

    def sign_flip_dlambda(adj, S, w):
        d = O.image_adj_level(adj, S, w)
        d[1] = -d[1]
        return d

    def s4_s5_swapped(adj, S, w):
        return O.image_adj_level(adj, [S[0], S[1], S[2], S[4], S[3], S[5]], w)

    assert _matches_table(good)
    for bad in (drop_factor_2, sign_flip_dlambda, s4_s5_swapped):
        assert not _matches_table(_one_hot_response(bad)), bad.__name__


# --------------------------------------------------------------------------- forward levels
NZF, NXF, ZS, XS = 31, 33, 15, 16


def _wavelet(nt, seed=3):
    """Arbitrary nonzero wavelet (the Ricker is zero for its first levels)."""
    return np.random.default_rng(seed).uniform(0.5, 1.5, nt) * np.where(np.arange(nt) % 2, 1.0, -1.0)


def _coefs(m1x, m1z, m2, m3, m2m3, nz=NZF, nx=NXF):
    return [np.full((nz, nx), v, np.float64, order="F") for v in (m1x, m1z, m2, m3, m2m3)]


def _levels(coefs, w, nt, abc=0, zs=ZS, xs=XS, nz=NZF, nx=NXF):
    p = O.make_params(nz, nx, nt, DX, DT, xs, zs, zs, 0, nx, abc)
    return O.forward(p, w, coefs, np.float64, want_hist5=True)["hist5"]  # hist5[f][:, :, L-1] = level L


def _sigma_closed_form(w, nt, kf=1.0):
    """Stress at the source with m1 = 0: sigma(1) = 0, sigma(2) = dt*s(1)
    (no K), sigma(L) = kf*(sigma(L-2) + dt*s(L-1)) for L >= 3; w[i] = s(i+1)."""
    sig = np.zeros(nt + 1)
    if nt >= 2:
        sig[2] = DT * w[0]
    for L in range(3, nt + 1):
        sig[L] = kf * (sig[L - 2] + DT * w[L - 2])
    return sig[1:]


def _check_source_levels(h, sig, zs=ZS, xs=XS):
    for f in (0, 1, 4):
        if np.any(h[f] != 0):
            return False
    for f in (2, 3):
        at = h[f][zs, xs, :]
        rest = h[f].copy()
        rest[zs, xs, :] = 0
        if np.any(rest != 0) or not _close(at, sig):
            return False
    return True


def test_forward_source_timing_closed_form():
    """With m1_x = m1_z = 0 the velocities stay zero and the stresses at the
    source accumulate dt*s(L-1) on the two leapfrog chains."""
    nt = 40
    w = _wavelet(nt)
    h = _levels(_coefs(0.0, 0.0, 2.0, 3.0, 5.0), w, nt)
    sig = _sigma_closed_form(w, nt)
    assert _check_source_levels(h, sig)
    # closed form of the chain sum: sigma(L) = dt * sum_m s(L-1-2m)
    for L in (2, 7, 8, 40):
        assert abs(sig[L - 1] - DT * sum(w[L - 2 - 2 * m] for m in range((L - 1 + 1) // 2) if L - 2 - 2 * m >= 0)) < 1e-15


def test_forward_source_timing_inside_sponge():
    """Source inside the sponge: K multiplies every level from 3 on, never
    level 2 (W/rtm-mod-adj.f90:193-198)."""
    nt, abc, zs = 30, 8, 3
    w = _wavelet(nt, 5)
    ones = np.ones((NZF, NXF))
    kf = float(O.abc_K(ones, abc)[zs, XS])
    assert 0.9 < kf < 1.0
    h = _levels(_coefs(0.0, 0.0, 2.0, 3.0, 5.0), w, nt, abc=abc, zs=zs)
    assert _check_source_levels(h, _sigma_closed_form(w, nt, kf), zs=zs)


def test_forward_source_timing_rejects_planted_slips():
    nt = 30
    w = _wavelet(nt, 7)
    c0 = _coefs(0.0, 0.0, 2.0, 3.0, 5.0)
    sig = _sigma_closed_form(w, nt)
    assert _check_source_levels(_levels(c0, w, nt), sig)
    # one-level source shift: dt*s(L) added at level L instead of dt*s(L-1)
    shifted = np.concatenate([w[1:], [0.0]])
    assert not _check_source_levels(_levels(c0, shifted, nt), sig)
    # K applied at level 2 as well (source inside the sponge)
    abc, zs = 8, 3
    kf = float(O.abc_K(np.ones((NZF, NXF)), abc)[zs, XS])
    h = _levels(c0, w, nt, abc=abc, zs=zs)
    good = _sigma_closed_form(w, nt, kf)
    assert _check_source_levels(h, good, zs=zs)
    bad = np.zeros(nt + 1)
    bad[2] = kf * DT * w[0]
    for L in range(3, nt + 1):
        bad[L] = kf * (bad[L - 2] + DT * w[L - 2])
    assert not _close(bad[1:], good)


# taps of Ax along x (and Bz along z) applied to a unit impulse at offset 0:
# data(i) = c1 (m(i+4) - m(i-5)) + c2 (m(i+3) - m(i-4)) + c3 (m(i+2) - m(i-3))
#         + c4 (m(i+1) - m(i-2)) + c5 (m(i) - m(i-1))     (W/rtmlib-mod.f90:229-237)
def _impulse_taps():
    c1, c2, c3, c4, c5 = O.fd_coefs(np.float64)
    return {-4: c1, -3: c2, -2: c3, -1: c4, 0: c5, 1: -c5, 2: -c4, 3: -c3, 4: -c2, 5: -c1}


def _check_level3(h, a1, a2, m1x, m1z, taps):
    lv = lambda f: h[f][:, :, 2]  # noqa: E731  level 3
    ex = np.zeros((NZF, NXF))
    for d, t in taps.items():
        ex[ZS, XS + d] = m1x * DT * a1 * t
    ez = np.zeros((NZF, NXF))
    for d, t in taps.items():
        ez[ZS + d, XS] = m1z * DT * a1 * t
    es = np.zeros((NZF, NXF))
    es[ZS, XS] = DT * a2
    return (_close(lv(0), ex) and _close(lv(1), ez) and _close(lv(2), es) and _close(lv(3), es)
            and np.all(lv(4) == 0))


def test_forward_first_three_levels_by_hand():
    nt = 4
    w = _wavelet(nt, 9)
    m1x, m1z = 0.7, 1.3
    h = _levels(_coefs(m1x, m1z, 2.0, 3.0, 5.0), w, nt)
    for f in range(5):
        assert np.all(h[f][:, :, 0] == 0)  # level 1
    lv2 = np.zeros((NZF, NXF))
    lv2[ZS, XS] = DT * w[0]
    for f in range(5):
        assert _close(h[f][:, :, 1], lv2 if f in (2, 3) else 0 * lv2)  # level 2
    taps = _impulse_taps()
    assert _check_level3(h, w[0], w[1], m1x, m1z, taps)
    # planted: S+ tap shape (offsets shifted by one) and the source of level 3
    # taken one level late
    shifted = {d + 1: t for d, t in taps.items()}
    assert not _check_level3(h, w[0], w[1], m1x, m1z, shifted)
    assert not _check_level3(h, w[0], w[2], m1x, m1z, taps)


# --------------------------------------------------------------------------- R/Src adjoint
def _small_shot(abc, nt=60, seed=5):
    nzp, nxp, dt = 24, 34, 1e-3
    vp = synth.layered_vp(nzp, nxp, seed, vp_top=1800.0, vp_bot=2600.0, nlayers=2, anom_sigma=4.0)
    pad = abc + 5 if abc > 0 else 8
    m = synth.elastic_from_vp(synth.pad_edge(vp, pad))
    coefs = O.model_op(m.rho, m.lam, m.mu, DX, dt, np.float64)
    nz, nx = nzp + 2 * pad, nxp + 2 * pad
    p = O.make_params(nz, nx, nt, DX, dt, pad + nxp // 2, pad + 3, pad + 2, pad, nxp, abc)
    w = O.ricker(nt, dt, 40.0, nsource=30)
    rec = np.random.default_rng(seed).uniform(-1, 1, (5, nxp, nt))
    return p, w, coefs, rec


def test_rsrc_adjoint_without_sponge_equals_the_drop_in_path():
    """With no sponge the R/Src loop (scatter adjoints, full histories) and
    the gather-form recursion of the drop-in path are the same linear map."""
    p, w, coefs, rec = _small_shot(0)
    a = O.rtm_rsrc(p, w, coefs, rec)
    b = O.rtm(p, w, coefs, rec)
    for x, y in zip(a, b):
        assert np.abs(y).max() > 0
        assert O.rel_l2(x, y) <= 1e-12


def test_rsrc_adjoint_with_sponge_differs_from_the_exact_adjoint():
    p, w, coefs, rec = _small_shot(8)
    p.adj_abc = 1
    a = O.rtm_rsrc(p, w, coefs, rec)
    b = O.rtm(p, w, coefs, rec)
    assert min(O.rel_l2(x, y) for x, y in zip(a, b)) > 1e-8  # far above FP64 rounding


def _rsrc_powers(nt, l0):
    """K powers of the stored levels when every operator is zero and data
    d sits at level l0 only, counted off R/Src/rtm-mod-adj.f90:380-462:
    level L receives A(L+2) after both of that level's K (computed(L) =
    K^2 computed(L+2)); a stored level L is damped twice (L >= 3), once
    (L = 2) or never (L = 1)."""
    count = {L: (2 if L >= 3 else 1 if L == 2 else 0) for L in range(1, nt + 1)}
    comp = {}
    L, e = l0, 0
    while L >= 1:
        comp[L] = e
        e += count[L]
        L -= 2
    return {L: comp[L] + count[L] for L in comp}


def _check_rsrc_powers(hist, nt, zr, xr, kf, powers, d=1.0):
    for f in range(5):
        for L in range(1, nt + 1):
            lv = hist[f][:, :, L - 1]
            want = d * kf ** powers[L] if L in powers else 0.0
            if abs(lv[zr, xr] - want) > 1e-14:
                return False
            rest = lv.copy()
            rest[zr, xr] = 0.0
            if np.any(rest != 0):
                return False
    return True


@pytest.mark.parametrize("l0", [12, 11])
def test_rsrc_sponge_powers_closed_form(l0):
    nt, abc, nz, nx, zr = 12, 8, 30, 32, 2
    p = O.make_params(nz, nx, nt, DX, DT, 15, 15, zr, 0, nx, abc)
    coefs = _coefs(0.0, 0.0, 0.0, 0.0, 0.0, nz, nx)
    xr = 3  # a receiver column inside the sponge, near the corner
    rec = np.zeros((5, nx, nt))
    rec[:, xr, l0 - 1] = 1.0
    kf = float(O.abc_K(np.ones((nz, nx)), abc)[zr, xr])
    assert kf < 0.95
    h = O.rsrc_adjoint(p, coefs, rec)
    good = _rsrc_powers(nt, l0)
    assert _check_rsrc_powers(h, nt, zr, xr, kf, good)
    # planted slips: K once per level (as the exact adjoint applies it), and
    # level 2 damped twice like the others
    once = {L: (good[L] + 1) // 2 for L in good}
    assert not _check_rsrc_powers(h, nt, zr, xr, kf, once)
    if 2 in good:
        two = dict(good)
        two[2] += 1
        assert not _check_rsrc_powers(h, nt, zr, xr, kf, two)
