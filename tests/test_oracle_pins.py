"""Pins of the CPU oracle (oracle/rtm_oracle.cpp) before it is trusted as the
checker of the HIP library. Each test is chosen so that a plausible slip in
the restatement (a dropped term, a wrong sign, index or operand) fails it:

* the reference's own C adjoint, compiled from its source into oracle/_ref,
  and the golden vectors made from it (a10 bit-for-bit in FP32);
* closed forms: FD weights from the staggered Taylor system (a1), polynomial
  exactness and the plane-wave symbol of the stencils (a5), sponge values
  with doubly damped corners (a7), Ricker peak and onset (a4), averaged
  coefficients (a3), the dt/nt known answer of R/Dat/seis-Vx-layer.H (a2);
* adjointness (the reference's own dot-product method, W/dptest_mod.f90:8-37):
  every forward operator against its scatter adjoint, the full-history
  adjoint against a forward propagation, and Born modelling against RTM;
* invariants: linearity, mirror symmetry, causality of the forward loop.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from elastic_amd import synth


def _dot(a, b):
    return float(np.vdot(np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()))


# --------------------------------------------------------------------------- a1
def _staggered_weights():
    """Solve sum_k a_k (2k-1)^(2m+1) = [m==0], m=0..4, exactly (a_1 nearest)."""
    n = 5
    A = [[Fraction((2 * k - 1) ** (2 * m + 1)) for k in range(1, n + 1)] for m in range(n)]
    b = [Fraction(1 if m == 0 else 0) for m in range(n)]
    # Gauss-Jordan in rationals
    M = [row[:] + [bb] for row, bb in zip(A, b)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        piv = M[c][c]
        M[c] = [v / piv for v in M[c]]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [vr - f * vc for vr, vc in zip(M[r], M[c])]
    return [M[k][n] for k in range(n)]


def test_fd_coefficients_exact():
    a = _staggered_weights()  # a[0] multiplies the nearest pair (reference c5)
    expect = [Fraction(19845, 16384), Fraction(-735, 8192), Fraction(567, 40960), Fraction(-405, 229376),
              Fraction(35, 294912)]
    assert a == expect
    c64 = O.fd_coefs(np.float64)
    c32 = O.fd_coefs(np.float32)
    ref = [expect[4], expect[3], expect[2], expect[1], expect[0]]  # c1..c5
    for k in range(5):
        assert c64[k] == float(ref[k])
        # single-precision literal quotient, rounded once
        assert c32[k] == np.float32(np.float32(ref[k].numerator) / np.float32(ref[k].denominator))
        assert abs(float(c32[k]) - float(ref[k])) <= abs(float(ref[k])) * 2 ** -24


# --------------------------------------------------------------------------- a5
NZ, NX = 30, 34


def _range_mask(op):
    ix0 = [5, 5, 4, 4, 4, 4, 4, 4, 5, 5][op]
    iz0 = [4, 4, 5, 5, 4, 4, 4, 4, 5, 5][op]
    m = np.zeros((NZ, NX), bool, order="F")
    m[iz0:NZ - 5, ix0:NX - 5] = True
    return m


@pytest.mark.parametrize("op", range(10))
@pytest.mark.parametrize("deg", range(0, 11))
def test_stencil_polynomial_exact(op, deg):
    """S+ is the derivative at i+1/2 and S- at i-1/2, exact up to degree 10."""
    plus = op in (1, 2, 4, 5, 6, 7)
    xaxis = op in (0, 2, 4, 6, 9)
    n = NX if xaxis else NZ
    u = (np.arange(n, dtype=np.float64) - n / 2) / n  # keep x^10 small
    h = 1.0 / n
    f1 = u ** deg
    g = (deg * (u + (0.5 if plus else -0.5) * h) ** (deg - 1) if deg > 0 else 0 * u) * h
    model = np.broadcast_to(f1[None, :] if xaxis else f1[:, None], (NZ, NX))
    expect = np.broadcast_to(g[None, :] if xaxis else g[:, None], (NZ, NX))
    coef = np.ones((NZ, NX))
    out = O.fwd_op(op, coef, model, np.zeros((NZ, NX)))
    m = _range_mask(op)
    assert np.allclose(out[m], expect[m], rtol=0, atol=1e-14)
    assert np.all(out[~m] == 0.0)  # outside the operator range nothing is added


@pytest.mark.parametrize("op", range(10))
def test_stencil_plane_wave_symbol(op):
    """S+- sin(k x)(i) = 2 cos(k(i +- 1/2)) sum_k c_k sin(k (11-2k)/2)."""
    plus = op in (1, 2, 4, 5, 6, 7)
    xaxis = op in (0, 2, 4, 6, 9)
    n = NX if xaxis else NZ
    kap = 0.7
    i = np.arange(n, dtype=np.float64)
    c = O.fd_coefs()
    sym = 2.0 * sum(c[k] * np.sin(kap * (11 - 2 * (k + 1)) / 2.0) for k in range(5))
    expect1 = np.cos(kap * (i + (0.5 if plus else -0.5))) * sym
    f = np.sin(kap * i)
    model = np.broadcast_to(f[None, :] if xaxis else f[:, None], (NZ, NX))
    expect = np.broadcast_to(expect1[None, :] if xaxis else expect1[:, None], (NZ, NX))
    out = O.fwd_op(op, np.ones((NZ, NX)), model, np.zeros((NZ, NX)))
    m = _range_mask(op)
    assert np.allclose(out[m], expect[m], rtol=0, atol=1e-13)


@pytest.mark.parametrize("op", range(10))
def test_operator_dot_product(op):
    rng = np.random.default_rng(100 + op)
    coef = rng.uniform(-1, 1, (NZ, NX))
    m = rng.uniform(-1, 1, (NZ, NX))
    d = rng.uniform(-1, 1, (NZ, NX))
    fm = O.fwd_op(op, coef, m, np.zeros((NZ, NX)))
    ad = O.adj_op(op, coef, np.zeros((NZ, NX)), d)
    lhs, rhs = _dot(fm, d), _dot(m, ad)
    assert abs(lhs - rhs) <= 1e-13 * max(abs(lhs), 1.0)


# --------------------------------------------------------------------------- a10
def _adjoint_inputs(seed, nz, nx, nt, zrec=3):
    """rtm_op3_c_ entry state: levels nt-1, nt-2 complete, lower levels hold an
    injected row. Coefficients with dt/dx-scaled magnitudes (stable)."""
    rng = np.random.default_rng(seed)
    hist = []
    for _ in range(5):
        h = np.zeros((nz, nx, nt), np.float32, order="F")
        h[:, :, nt - 2:] = rng.uniform(-1, 1, (nz, nx, 2))
        h[zrec, :, : nt - 2] = rng.uniform(-1, 1, (nx, nt - 2))
        hist.append(h)
    m1 = [rng.uniform(0.4e-7, 0.6e-7, (nz, nx)).astype(np.float32) for _ in range(2)]
    aux = [rng.uniform(0.5e6, 2.0e6, (nz, nx)).astype(np.float32) for _ in range(3)]
    return hist, [m1[0], m1[1], aux[0], aux[1], aux[2]]


def test_gather_equals_scatter_fp64():
    hist, coefs = _adjoint_inputs(7, 26, 30, 40)
    g = O.rtm_op3_c(hist, coefs, "gather", np.float64)
    s = O.rtm_op3_c(hist, coefs, "scatter", np.float64)
    for a, b in zip(g, s):
        assert O.rel_l2(a, b) <= 1e-14


def test_fp32_oracle_bitwise_equals_reference_c():
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (reference tree absent); golden vectors cover this")
    hist, coefs = _adjoint_inputs(11, 22, 26, 64)
    ref = O.ref_rtm_op3_c(hist, coefs)
    sc = O.rtm_op3_c(hist, coefs, "scatter", np.float32)
    ga = O.rtm_op3_c(hist, coefs, "gather", np.float32)
    g64 = O.rtm_op3_c(hist, coefs, "gather", np.float64)
    for f in range(5):
        assert np.array_equal(sc[f], ref[f])
        assert np.array_equal(ga[f], ref[f])
        assert O.rel_l2(ref[f], g64[f]) <= 1e-5


def test_golden_rtm_op3_reference_vectors():
    """Vectors written by tests/golden/make_golden_op3.py from the reference's
    own C adjoint; the FP32 oracle reproduces them bit for bit."""
    from tests.golden import golden_op3

    g = golden_op3.load()
    out = O.rtm_op3_c(g["hist"], g["coefs"], "gather", np.float32)
    golden_op3.check(g, out, exact=True)
    out64 = O.rtm_op3_c(g["hist"], g["coefs"], "gather", np.float64)
    golden_op3.check(g, out64, exact=False, tol=1e-5)


def _forward_full_nok(src, coefs):
    """X(it) = s(it) + X(it-2) + Op X(it-1), X(<0)=0, full five-field histories
    (the Born-loop structure of W/rtm-mod-adj.f90:323-348 without K)."""
    nz, nx, nt = src[0].shape
    X = [np.zeros((nz, nx, nt), order="F") for _ in range(5)]
    m1x, m1z, m2, m3, m2m3 = coefs
    ops = [(0, m1x, 2, 0), (1, m1x, 4, 0), (2, m1z, 4, 1), (3, m1z, 3, 1), (4, m2m3, 0, 2), (5, m2, 1, 2),
           (6, m2, 0, 3), (7, m2m3, 1, 3), (8, m3, 0, 4), (9, m3, 1, 4)]
    for it in range(nt):
        cur = [src[f][:, :, it] + (X[f][:, :, it - 2] if it >= 2 else 0.0) for f in range(5)]
        if it >= 1:
            for op, c, fin, fout in ops:
                cur[fout] = O.fwd_op(op, c, X[fin][:, :, it - 1], cur[fout])
        for f in range(5):
            X[f][:, :, it] = cur[f]
    return X


def test_full_history_adjoint_dot_product():
    """The gather adjoint (with the rtm_op3 pre-step) is the transpose of the
    sponge-free forward propagation."""
    nz, nx, nt = 22, 24, 12
    rng = np.random.default_rng(5)
    coefs = [rng.uniform(0.1, 0.3, (nz, nx)) for _ in range(5)]
    s = [rng.uniform(-1, 1, (nz, nx, nt)) for _ in range(5)]
    d = [rng.uniform(-1, 1, (nz, nx, nt)) for _ in range(5)]
    Ls = _forward_full_nok(s, coefs)
    # adjoint: level nt-1 = d; level nt-2 = d + Op^T(level nt-1); lower: gather loop
    A = [np.asfortranarray(x.copy()) for x in d]
    lvl = [A[f][:, :, nt - 1] for f in range(5)]
    m1x, m1z, m2, m3, m2m3 = coefs
    adj = [(4, m2m3, 0, 2), (6, m2, 0, 3), (8, m3, 0, 4), (5, m2, 1, 2), (7, m2m3, 1, 3), (9, m3, 1, 4),
           (0, m1x, 2, 0), (3, m1z, 3, 1), (1, m1x, 4, 0), (2, m1z, 4, 1)]
    for op, c, fout, fin in adj:
        A[fout][:, :, nt - 2] = O.adj_op(op, c, A[fout][:, :, nt - 2], lvl[fin])
    LtD = O.rtm_op3_c(A, coefs, "gather", np.float64)
    lhs = sum(_dot(Ls[f], d[f]) for f in range(5))
    rhs = sum(_dot(s[f], LtD[f]) for f in range(5))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_prestep_matches_manual():
    nz, nx, nt, zrec = 20, 22, 8, 6
    rng = np.random.default_rng(9)
    coefs = [rng.uniform(0.1, 0.3, (nz, nx)) for _ in range(5)]
    data = [rng.uniform(-1, 1, (nt, nx)) for _ in range(5)]
    h = O.rtm_op3_prestep(nt, zrec, data, coefs)
    for f in range(5):
        assert np.array_equal(h[f][zrec, :, nt - 1], data[f][nt - 1])
        for lev in range(nt - 2):
            assert np.array_equal(h[f][zrec, :, lev], data[f][lev])
            others = np.delete(h[f][:, :, lev], zrec, axis=0)
            assert not others.any()
    # level nt-2 of Vx = D^T sxx + F^T szz + H^T sxz of level nt-1 plus its row
    top = [np.zeros((nz, nx)) for _ in range(5)]
    for f in range(5):
        top[f][zrec] = data[f][nt - 1]
    vx = np.zeros((nz, nx))
    vx[zrec] = data[0][nt - 2]
    vx = O.adj_op("D", coefs[4], vx, top[2])
    vx = O.adj_op("F", coefs[2], vx, top[3])
    vx = O.adj_op("H", coefs[3], vx, top[4])
    assert np.array_equal(h[0][:, :, nt - 2], vx)


# --------------------------------------------------------------------------- a7, a4, a3, a2
def test_sponge_filter_known_answer_and_closed_form():
    f = O.abc_filter(40)
    assert f.shape == (45,)  # R/Dat/filter.H: n1=45 for abc=40
    i = np.arange(1, 41)
    assert np.allclose(f[:40], np.exp(-0.09 * (i / 40.0) ** 2), rtol=1e-15)
    assert np.all(f[40:] == f[39])
    f32 = O.abc_filter(40, np.float32)
    assert np.allclose(f32, f, rtol=1e-6)


def test_sponge_K_closed_form_double_corners():
    abc, nz, nx = 6, 40, 44
    abc2 = abc + 5
    f = O.abc_filter(abc)
    K = O.abc_K(np.ones((nz, nx)), abc)
    kz = np.ones(nz)
    kx = np.ones(nx)
    for j in range(1, abc2 + 1):
        kz[j - 1] *= f[abc2 - j]
        kz[nz - abc2 + j - 1] *= f[j - 1]
        kx[j - 1] *= f[abc2 - j]
        kx[nx - abc2 + j - 1] *= f[j - 1]
    assert np.allclose(K, kz[:, None] * kx[None, :], rtol=1e-15)
    assert K[0, 0] == pytest.approx(f[abc2 - 1] ** 2)  # corner damped twice
    assert K[nz // 2, nx // 2] == 1.0


def test_ricker_closed_form():
    nt, dt, fp = 400, 1e-3, 15.0
    s = O.ricker(nt, dt, fp)
    assert np.all(s[:24] == 0.0)  # first nonzero sample it = nsource/10 = 25
    assert s[124] == pytest.approx(1e8, rel=1e-12)  # peak at it = nsource/2
    it = np.arange(25, nt + 1)
    t = it * dt - 125 * dt
    a = (3.14159265 * fp * t) ** 2
    assert np.allclose(s[24:], 1e8 * (1 - 2 * a) * np.exp(-a), rtol=1e-12, atol=1e-6)
    # zero crossings at t = +-1/(pi sqrt 2 fp)
    tz = 1.0 / (np.pi * np.sqrt(2.0) * fp)
    k = int(np.floor(125 + tz / dt))
    assert s[k - 1] > 0 > s[k]


def test_model_op_closed_forms():
    nz, nx, dx, dt = 16, 18, 10.0, 1e-3
    rho = np.full((nz, nx), 2000.0)
    lam = np.full((nz, nx), 5e9)
    mu = np.full((nz, nx), 3e9)
    m1x, m1z, m2, m3, m2m3 = O.model_op(rho, lam, mu, dx, dt)
    r = dt / dx
    assert np.allclose(m1x, r / 2000.0, rtol=1e-15)
    assert np.allclose(m1z, r / 2000.0, rtol=1e-15)
    assert np.allclose(m2, r * 5e9, rtol=1e-15)
    assert np.allclose(m3, r * 3e9, rtol=1e-15)
    assert np.allclose(m2m3, r * 11e9, rtol=1e-15)
    # harmonic 2x2 mean and the mu=0 rule
    rng = np.random.default_rng(3)
    lam = rng.uniform(1e9, 9e9, (nz, nx))
    mu = rng.uniform(1e9, 4e9, (nz, nx))
    mu[8, 9] = 0.0
    rho = rng.uniform(1500, 2600, (nz, nx))
    m1x, m1z, m2, m3, m2m3 = O.model_op(rho, lam, mu, dx, dt)
    iz, ix = 5, 6
    h = 4.0 / (1 / lam[iz, ix] + 1 / lam[iz + 1, ix] + 1 / lam[iz, ix + 1] + 1 / lam[iz + 1, ix + 1])
    assert m2[iz, ix] == pytest.approx(r * h, rel=1e-14)
    for (a, b) in [(8, 9), (7, 9), (8, 8), (7, 8)]:
        assert m3[a, b] == 0.0
    assert m1x[iz, ix] == pytest.approx(r / (0.5 * (rho[iz + 1, ix] + rho[iz, ix])), rel=1e-14)
    assert m1z[iz, ix] == pytest.approx(r / (0.5 * (rho[iz, ix + 1] + rho[iz, ix])), rel=1e-14)
    assert m1x[nz - 1, 3] == pytest.approx(r / rho[nz - 1, 3], rel=1e-14)
    assert m1z[4, nx - 1] == pytest.approx(r / rho[4, nx - 1], rel=1e-14)


def test_reference_dt_nt_known_answer():
    dt, nt = synth.reference_dt_nt(dx=4.0, ttime=2.0)  # R/Dat/seis-Vx-layer.H
    assert nt == 3535
    assert abs(dt - 0.000565685) < 5e-10


# --------------------------------------------------------------------------- a6 invariants
def _small_shot(nt=160, abc=0, nz=60, nx=81, dt=1e-3):
    vp = np.full((nz, nx), 2000.0)
    vp[30:] = 2600.0  # x-invariant layered model
    m = synth.elastic_from_vp(vp)
    coefs = O.model_op(m.rho, m.lam, m.mu, 10.0, dt)
    p = O.make_params(nz, nx, nt, 10.0, dt, xs=nx // 2, zs=20, zr=15, rx0=0, nrx=nx, abc=abc)
    w = O.ricker(nt, dt, 15.0)
    return p, w, coefs


def test_forward_linearity_and_causality():
    p, w, coefs = _small_shot()
    r1 = O.forward(p, w, coefs)["rec"]
    r2 = O.forward(p, 2.0 * w, coefs)["rec"]
    assert O.rel_l2(r2, 2.0 * r1) <= 1e-14
    assert not r1[:, :, :25].any()  # nothing before the first nonzero source sample
    assert np.abs(r1).max() > 0


def test_forward_mirror_symmetry():
    """x-invariant model, centred source: Vx, sxz odd about the source (half-cell
    grids map i -> 2xs+1-i), Vz, sxx, szz even (i -> 2xs-i)."""
    p, w, coefs = _small_shot()
    r = O.forward(p, w, coefs)["rec"]
    xs = p.xs
    scale = np.abs(r).max(axis=(1, 2))
    for f in (1, 2, 3):
        for k in range(1, 25):
            assert np.allclose(r[f, xs + k], r[f, xs - k], atol=1e-12 * scale[f])
    for f in (0, 4):
        for k in range(1, 25):
            assert np.allclose(r[f, xs + k], -r[f, xs + 1 - k], atol=1e-12 * scale[f])


def _born_rtm_pair(abc, adj_abc):
    nz, nx, nt = 44, 48, 70
    rng = np.random.default_rng(21)
    vp = np.asfortranarray(rng.uniform(1800, 2600, (nz, nx)))
    m = synth.elastic_from_vp(vp)
    dt = 1e-3
    coefs = O.model_op(m.rho, m.lam, m.mu, 10.0, dt)
    p = O.make_params(nz, nx, nt, 10.0, dt, xs=24, zs=16, zr=14, rx0=10, nrx=28, abc=abc, adj_abc=adj_abc)
    w = O.ricker(nt, dt, 25.0, nsource=40)
    dm = [rng.uniform(-1, 1, (nz, nx)) * s for s in (1e-3, 1e-3, 1e-3)]
    d = rng.uniform(-1, 1, (5, p.nrx, nt))
    return p, w, coefs, dm, d


@pytest.mark.parametrize("abc,adj_abc", [(0, 0), (6, 1)])
def test_born_rtm_dot_product(abc, adj_abc):
    """Whole chain: RTM (adjoint + Lame imaging) is the transpose of Born
    modelling (W/rtm-mod-adj.f90:273-375), with the sponge applied exactly."""
    p, w, coefs, dm, d = _born_rtm_pair(abc, adj_abc)
    Bm = O.born(p, w, coefs, dm)
    img = O.rtm(p, w, coefs, d)
    lhs = _dot(Bm, d)
    rhs = sum(_dot(a, b) for a, b in zip(dm, img))
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def test_reference_adjoint_without_sponge_is_not_transpose_with_sponge():
    """The drop-in adjoint (no K, W/ tree) is only the transpose of a forward
    run without K: with K in the forward the dot test must fail."""
    p, w, coefs, dm, d = _born_rtm_pair(6, 0)
    Bm = O.born(p, w, coefs, dm)
    img = O.rtm(p, w, coefs, d)
    lhs = _dot(Bm, d)
    rhs = sum(_dot(a, b) for a, b in zip(dm, img))
    assert abs(lhs - rhs) > 1e-6 * abs(lhs)
