"""GPU parity of the 2D elastic kernels against the FP64 oracle (and the
reference's own FP32 adjoint through its golden vectors).

Tolerances (stated by the task for this path): wavefield relative L2 <= 1e-5
after N steps and images <= 1e-4 against FP64, plain FP32 form.
"""
import numpy as np
import pytest

import oracle as O
from elastic_amd import synth
from tests.golden import golden_op3

pytestmark = pytest.mark.gpu

WAVE_TOL = 1e-5
IMG_TOL = 1e-4


def _adjoint_inputs(seed, nz, nx, nt, zrec=3):
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


def _max_level_err(a, b):
    nt = a.shape[2]
    return max(O.rel_l2(a[:, :, k], b[:, :, k]) for k in range(nt) if np.any(b[:, :, k]))


def test_rtm_op3_golden_reference_vectors(hip):
    g = golden_op3.load()
    out = hip.rtm_op3(g["hist"], g["coefs"], g["zrec"])
    golden_op3.check(g, out, exact=False, tol=WAVE_TOL)
    ref64 = O.rtm_op3_c(g["hist"], g["coefs"], "gather", np.float64)
    for f in range(5):
        assert _max_level_err(out[f], ref64[f]) <= WAVE_TOL


@pytest.mark.parametrize("nz,nx,nt", [(37, 45, 30), (70, 130, 25), (130, 21, 20), (64, 16, 12), (12, 12, 5)])
def test_rtm_op3_ragged_sizes(hip, nz, nx, nt):
    hist, coefs = _adjoint_inputs(nz * 1000 + nx, nz, nx, nt)
    out = hip.rtm_op3(hist, coefs)
    ref = O.rtm_op3_c(hist, coefs, "gather", np.float64)
    for f in range(5):
        assert _max_level_err(out[f], ref[f]) <= WAVE_TOL, f
        # levels nt-1, nt-2 are inputs and must come back unchanged
        assert np.array_equal(out[f][:, :, nt - 2:], hist[f][:, :, nt - 2:])


def test_rtm_op3_void_symbol_and_config0_grid(hip):
    """The exported void rtm_op3_c_ on the config-0 grid (300 x 600), 40 levels."""
    hist, coefs = _adjoint_inputs(3, 300, 600, 40, zrec=2)
    out = hip.rtm_op3_c_void(hist, coefs, zrec=2)
    ref = O.rtm_op3_c(hist, coefs, "gather", np.float64, nthreads=8)
    for f in range(5):
        assert _max_level_err(out[f], ref[f]) <= WAVE_TOL, f


def _big_history(nz, nx, nt, zrec, seed):
    """Seeded drop-in input: complete top two levels, receiver rows below."""
    rng = np.random.default_rng(seed)
    hist = []
    for _ in range(5):
        h = np.zeros((nz, nx, nt), np.float32, order="F")
        h[:, :, nt - 2:] = rng.uniform(-1, 1, (nz, nx, 2))
        h[zrec, :, : nt - 2] = rng.uniform(-1e-2, 1e-2, (nx, nt - 2))
        hist.append(h)
    return hist


@pytest.mark.parametrize("form", ["tiles", "marching"])
def test_rtm_op3_history_past_2_31_elements(hip, monkeypatch, form):
    """Config-0 padded grid 390 x 690 with its 2000 levels: the sigma_xz
    history starts at element 4*nt*nz*nx = 2,152,800,000 > 2^31, so field
    offsets must be 64-bit. The top 12 levels are checked against the FP64
    oracle run on the top 14 levels (the recursion of a level reads only the
    two above it); the whole chain against the same history advanced in two
    chunks whose offsets stay below 2^31 (levels 998..1999, then 0..999),
    which must agree bit for bit."""
    if form == "marching":
        monkeypatch.setenv("RTM_TBLOCK", "2")
        monkeypatch.setenv("RTM_MARCH_W", "26")
    else:
        monkeypatch.setenv("RTM_TBLOCK", "0")
    c = synth.config0()
    nz, nx, nt, zrec = c["geo"].nz, c["geo"].nx, c["nt"], c["geo"].zr
    assert (nz, nx, nt) == (390, 690, 2000) and 4 * nt * nz * nx >= 2 ** 31
    m = c["mig"]
    coefs = [np.asfortranarray(a) for a in O.model_op(m.rho, m.lam, m.mu, c["dx"], c["dt"], np.float32)]
    print(f"op3 {form}: {nz}x{nx}x{nt}, one call", flush=True)
    full = _big_history(nz, nx, nt, zrec, 11)
    top_in = [h[:, :, nt - 14:].copy(order="F") for h in full]  # a copy: full is updated in place
    hip.rtm_op3(full, coefs, zrec, inplace=True)
    ref = O.rtm_op3_c(top_in, coefs, "gather", np.float64, nthreads=8)
    for f in range(5):
        assert _max_level_err(full[f][:, :, nt - 14:], ref[f]) <= WAVE_TOL, f
    del top_in, ref
    print(f"op3 {form}: two chunks", flush=True)
    chunk = _big_history(nz, nx, nt, zrec, 11)
    hip.rtm_op3([h[:, :, 998:] for h in chunk], coefs, zrec, inplace=True)
    hip.rtm_op3([h[:, :, :1000] for h in chunk], coefs, zrec, inplace=True)
    for f in range(5):
        assert np.array_equal(full[f], chunk[f]), f
        assert np.isfinite(full[f][:, :, 0]).all() and np.abs(full[f][:, :, 0]).max() > 0


def test_rtm_op3_trivial_nt(hip):
    hist, coefs = _adjoint_inputs(1, 20, 20, 2)
    out = hip.rtm_op3(hist, coefs)
    for f in range(5):
        assert np.array_equal(out[f], hist[f])


def test_rtm_op3_rejects_bad_dims(hip):
    hist, coefs = _adjoint_inputs(1, 11, 20, 8)
    with pytest.raises(hip.RtmError) as e:
        hip.rtm_op3(hist, coefs)
    assert e.value.code == -1


# --------------------------------------------------------------------------- forward
def _shot(nzp=60, nxp=90, nt=300, abc=10, dt=1e-3, fp=15.0, seed=5):
    vp = synth.layered_vp(nzp, nxp, seed)
    geo = synth.single_shot_geometry(nzp, nxp, abc, 10.0)
    m = synth.elastic_from_vp(synth.pad_edge(vp, geo.pad))
    coefs = O.model_op(m.rho, m.lam, m.mu, 10.0, dt, np.float32)
    p = O.make_params(geo.nz, geo.nx, nt, 10.0, dt, geo.xs, geo.zs, geo.zr, geo.rx0, geo.nrx, abc)
    w = O.ricker(nt, dt, fp, dtype=np.float32)
    return p, w, [np.asfortranarray(c) for c in coefs], m, geo


def _hip_shot(hip, p, **kw):
    return hip.shot(p.nz, p.nx, p.nt, p.dx, p.dt, p.xs, p.zs, p.zr, p.rx0, p.nrx, p.abc, **kw)


@pytest.mark.parametrize("nt", [60, 300])
def test_forward_receivers_and_last_levels(hip, nt):
    p, w, coefs, _, _ = _shot(nt=nt)
    rec, last = hip.forward(_hip_shot(hip, p), coefs, w, want_last=True)
    ref = O.forward(p, w, coefs, np.float64, want_hist5=True)
    assert O.rel_l2(rec, ref["rec"]) <= WAVE_TOL
    for f in range(5):
        assert O.rel_l2(last[f], ref["hist5"][f][:, :, nt - 2]) <= WAVE_TOL, f
        assert O.rel_l2(last[5 + f], ref["hist5"][f][:, :, nt - 1]) <= WAVE_TOL, f


def test_forward_config0_grid_reduced_steps(hip):
    """Config-0 padded grid (390 x 690), seeded model, 400 of its 2000 levels."""
    c = synth.config0()
    geo, nt = c["geo"], 400
    coefs = O.model_op(c["true"].rho, c["true"].lam, c["true"].mu, c["dx"], c["dt"], np.float32)
    coefs = [np.asfortranarray(a) for a in coefs]
    p = O.make_params(geo.nz, geo.nx, nt, c["dx"], c["dt"], geo.xs, geo.zs, geo.zr, geo.rx0, geo.nrx, c["abc"],
                      nthreads=8)
    w = O.ricker(nt, c["dt"], c["fp"], dtype=np.float32)
    rec, last = hip.forward(_hip_shot(hip, p), coefs, w, want_last=True)
    ref = O.forward(p, w, coefs, np.float64, want_hist5=False)
    assert O.rel_l2(rec, ref["rec"]) <= WAVE_TOL
    assert np.abs(rec).max() > 0


def test_library_setup_matches_oracle(hip):
    """model_op, Ricker and K of the library's host setup against the oracle."""
    _, _, _, m, _ = _shot()
    got = hip.model_op(m.rho, m.lam, m.mu, 10.0, 1e-3)
    exp = O.model_op(m.rho, m.lam, m.mu, 10.0, 1e-3, np.float32)
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)
    assert np.array_equal(hip.ricker(500, 1e-3, 15.0), O.ricker(500, 1e-3, 15.0, dtype=np.float32))
    f = np.random.default_rng(0).uniform(-1, 1, (60, 70)).astype(np.float32)
    assert np.array_equal(hip.abc_k(f, 12), O.abc_K(f, 12, np.float32))


# --------------------------------------------------------------------------- RTM / Born
def _observed(p, w, coefs, m, geo, dt):
    """Observed data = true - smooth (both from the FP64 oracle)."""
    vps = synth.smooth(m.vp, 4.0)
    ms = synth.elastic_from_vp(vps)
    cs = O.model_op(ms.rho, ms.lam, ms.mu, 10.0, dt, np.float32)
    cs = [np.asfortranarray(c) for c in cs]
    d = O.forward(p, w, coefs, np.float64)["rec"] - O.forward(p, w, cs, np.float64)["rec"]
    return d.astype(np.float32), cs


@pytest.mark.parametrize("adj_abc,comp", [(0, 0), (1, 0), (1, 1)])
def test_migrate_images_vs_fp64(hip, adj_abc, comp):
    p, w, coefs, m, geo = _shot(nt=260)
    d, cs = _observed(p, w, coefs, m, geo, 1e-3)
    p.adj_abc = adj_abc
    img = hip.migrate(_hip_shot(hip, p, adj_abc=adj_abc, comp_image=comp), cs, w, d)
    ref = O.rtm(p, w, cs, d, np.float64)
    for a, b in zip(img, ref):
        assert np.abs(b).max() > 0
        assert O.rel_l2(a, b) <= IMG_TOL


@pytest.mark.parametrize("tblock", [0, 2])
def test_migrate_rsrc_sponge_adjoint_vs_fp64(hip, monkeypatch, tblock):
    """adj_abc = 2: the R/Src adjoint (K applied in place to the stored
    levels it+1 and it+2 before every step) against the oracle's literal
    restatement of that loop, over the whole grid, sponge included; tile and
    marching kernels agree bit for bit."""
    p, w, coefs, m, geo = _shot(nt=260)
    d, cs = _observed(p, w, coefs, m, geo, 1e-3)
    monkeypatch.setenv("RTM_TBLOCK", str(tblock))
    monkeypatch.setenv("RTM_MARCH_W", "16")
    img = hip.migrate(_hip_shot(hip, p, adj_abc=2), cs, w, d)
    ref = O.rtm_rsrc(p, w, cs, d, np.float64)
    p.adj_abc = 1
    exact = O.rtm(p, w, cs, d, np.float64)
    for a, b, e in zip(img, ref, exact):
        assert np.abs(b).max() > 0
        assert O.rel_l2(a, b) <= IMG_TOL
        assert O.rel_l2(e, b) > 10 * O.rel_l2(a, b)  # the two sponge adjoints are told apart
    monkeypatch.setenv("RTM_TBLOCK", str(2 - tblock))
    other = hip.migrate(_hip_shot(hip, p, adj_abc=2), cs, w, d)
    for a, b in zip(img, other):
        assert np.array_equal(a, b)


def test_migrate_decimated_imaging(hip):
    p, w, coefs, m, geo = _shot(nt=240)
    d, cs = _observed(p, w, coefs, m, geo, 1e-3)
    p.snap_every = 4
    img = hip.migrate(_hip_shot(hip, p, snap_every=4), cs, w, d)
    ref = O.rtm(p, w, cs, d, np.float64)
    for a, b in zip(img, ref):
        assert O.rel_l2(a, b) <= IMG_TOL


def test_born_vs_fp64(hip):
    p, w, coefs, m, geo = _shot(nt=200)
    rng = np.random.default_rng(2)
    dm = [np.asfortranarray(rng.uniform(-1, 1, (p.nz, p.nx)).astype(np.float32) * s) for s in (1e-3, 1e-3, 1e-3)]
    rec = hip.born(_hip_shot(hip, p), coefs, w, dm)
    ref = O.born(p, w, coefs, dm, np.float64)
    assert O.rel_l2(rec, ref) <= WAVE_TOL


def test_born_rtm_dot_product_on_gpu(hip):
    """Size-independent check: the GPU RTM (exact sponge adjoint) is the
    transpose of the GPU Born modelling to FP32 accuracy."""
    p, w, coefs, m, geo = _shot(nt=200)
    rng = np.random.default_rng(4)
    dm = [np.asfortranarray(rng.uniform(-1, 1, (p.nz, p.nx)).astype(np.float32)) for _ in range(3)]
    d = rng.uniform(-1, 1, (5, p.nrx, p.nt)).astype(np.float32)
    s = _hip_shot(hip, p, adj_abc=1)
    Bm = hip.born(s, coefs, w, dm)
    img = hip.migrate(s, coefs, w, d)
    lhs = float(np.vdot(Bm.astype(np.float64).ravel(), d.astype(np.float64).ravel()))
    rhs = sum(float(np.vdot(a.astype(np.float64).ravel(), b.astype(np.float64).ravel())) for a, b in zip(dm, img))
    assert abs(lhs - rhs) <= 1e-4 * abs(lhs)


def test_migrate_linearity(hip):
    p, w, coefs, m, geo = _shot(nt=200)
    rng = np.random.default_rng(8)
    d = rng.uniform(-1, 1, (5, p.nrx, p.nt)).astype(np.float32)
    s = _hip_shot(hip, p)
    i1 = hip.migrate(s, coefs, w, d)
    i2 = hip.migrate(s, coefs, w, 2.0 * d)
    for a, b in zip(i1, i2):
        # exact in binary FP except where the far-field precursor is subnormal
        assert O.rel_l2(b, 2.0 * a) <= 1e-6


def test_fortran_module_mirror(hip):
    """RtmOperator mirrors rtm_init/rtm_op (1-based geometry, in-place arrays)."""
    p, w, coefs, m, geo = _shot(nt=150, fp=15.0)
    rho, lam, mu = (np.asfortranarray(a, np.float32) for a in (m.rho, m.lam, m.mu))
    op = hip.RtmOperator(p.nz, 10.0, p.nx, 10.0, 1e-3, p.xs + 1, p.zs + 1, p.zr + 1, 15.0, p.nt, p.abc, False,
                         rho, lam, mu)
    rng = np.random.default_rng(6)
    data = [np.asfortranarray(rng.uniform(-1, 1, (p.nt, p.nx)).astype(np.float32)) for _ in range(5)]
    dm = [np.zeros((p.nz, p.nx), np.float32, order="F") for _ in range(3)]
    op.op(True, *dm, *data)
    pr = O.make_params(p.nz, p.nx, p.nt, 10.0, 1e-3, p.xs, p.zs, p.zr, 0, p.nx, p.abc)
    cref = O.model_op(m.rho, m.lam, m.mu, 10.0, 1e-3, np.float32)
    wref = O.ricker(p.nt, 1e-3, 15.0, dtype=np.float32)
    recref = np.stack([d.T for d in data])
    ref = O.rtm(pr, wref, cref, recref, np.float64)
    for a, b in zip(dm, ref):
        assert O.rel_l2(a, b) <= IMG_TOL
    # Born direction writes the five data arrays
    out = [np.zeros((p.nt, p.nx), np.float32, order="F") for _ in range(5)]
    dmb = [np.asfortranarray(rng.uniform(-1, 1, (p.nz, p.nx)).astype(np.float32) * 1e-3) for _ in range(3)]
    op.op(False, *dmb, *out)
    bref = O.born(pr, wref, cref, dmb, np.float64)
    assert O.rel_l2(np.stack([o.T for o in out]), bref) <= WAVE_TOL
    op.close()


def test_shot_argument_errors(hip):
    p, w, coefs, _, _ = _shot(nt=20)
    bad = _hip_shot(hip, p)
    bad.xs = p.nx + 3
    with pytest.raises(hip.RtmError):
        hip.forward(bad, coefs, w)
    bad = _hip_shot(hip, p)
    bad.nrx = p.nx + 1
    with pytest.raises(hip.RtmError):
        hip.forward(bad, coefs, w)
    ok, cr, lim = hip.cfl_check(5000.0, 1e-3, 10.0)
    assert ok and abs(lim - 0.5370) < 1e-3
    ok, cr, lim = hip.cfl_check(6000.0, 1e-3, 10.0)
    assert not ok


def test_movie_frames_are_forward_levels(hip, tmp_path):
    """rtm_forward_movie_hip frames equal the FP64 oracle's levels k*dsnap in
    the window, and snap_wfld writes them as the reference's SEP cubes."""
    from elastic_amd import sep, snap

    nzp, nxp, abc, nt = 40, 50, 10, 130
    geo = synth.single_shot_geometry(nzp, nxp, abc, 10.0)
    m = synth.elastic_from_vp(synth.pad_edge(synth.layered_vp(nzp, nxp, 3), geo.pad))
    coefs = hip.model_op(m.rho, m.lam, m.mu, 10.0, 1e-3)
    wav = hip.ricker(nt, 1e-3, 25.0)
    s = hip.shot(geo.nz, geo.nx, nt, 10.0, 1e-3, geo.xs, geo.zs, geo.zr, geo.rx0, geo.nrx, abc)
    z0, x0, wnz, wnx = snap.window(geo.nz, geo.nx, abc)
    movie, dsnap = hip.forward_movie(s, coefs, wav, 6, z0, x0, wnz, wnx)
    assert dsnap == nt // 6 and movie.shape == (6, 5, wnz, wnx)
    p = O.make_params(geo.nz, geo.nx, nt, 10.0, 1e-3, geo.xs, geo.zs, geo.zr, geo.rx0, geo.nrx, abc)
    h5 = O.forward(p, wav, coefs, want_hist5=True)["hist5"]
    for k in range(6):
        L = (k + 1) * dsnap
        for f in range(5):
            ref = h5[f][z0:z0 + wnz, x0:x0 + wnx, L - 1]
            assert O.rel_l2(movie[k, f], ref) <= 1e-5, (k, f)
    paths = snap.snap_wfld(s, coefs, wav, str(tmp_path), nsnap=6)
    cube, h = sep.read(paths["Vz0"])
    assert cube.shape == (wnz, wnx, 6) and np.array_equal(cube[:, :, 2], movie[2, 1])
    assert sep.axes(h)[2][2] == pytest.approx(1e-3 * dsnap)


@pytest.mark.parametrize("tblock", ["0", "1", "2"])
def test_sponge_free_forward_and_migrate(hip, monkeypatch, tblock):
    """abc = 0 (no sponge anywhere; the kernels get no sponge factors):
    forward levels and images against the FP64 oracle, one and two levels per
    launch and the marching form."""
    monkeypatch.setenv("RTM_TBLOCK", tblock)
    p, w, coefs, m, geo = _shot(nt=180)
    p.abc = 0
    rec, last = hip.forward(_hip_shot(hip, p), coefs, w, want_last=True)
    ref = O.forward(p, w, coefs, np.float64, want_hist5=True)
    assert O.rel_l2(rec, ref["rec"]) <= WAVE_TOL
    for f in range(5):
        assert O.rel_l2(last[5 + f], ref["hist5"][f][:, :, p.nt - 1]) <= WAVE_TOL, f
    d, cs = _observed(p, w, coefs, m, geo, 1e-3)
    img = hip.migrate(_hip_shot(hip, p), cs, w, d)
    refi = O.rtm(p, w, cs, d, np.float64)
    for a, b in zip(img, refi):
        assert O.rel_l2(a, b) <= IMG_TOL


@pytest.mark.parametrize("mw,nw", [("2", "1"), ("6", "2"), ("32", "4")])
def test_marching_form_single_shot_paths_bitwise(hip, monkeypatch, mw, nw):
    """The marching step kernels (RTM_TBLOCK=2, strips of mw columns and nw
    waves) give bitwise the tile kernels' results on the drop-in adjoint
    (ragged grid), the forward with receivers, Born modelling and migration
    with the exact sponge adjoint and compensated images."""
    hist, coefs5 = _adjoint_inputs(7, 70, 130, 25)
    p, w, coefs, m, geo = _shot(nt=160)
    rng = np.random.default_rng(2)
    dm = [np.asfortranarray(rng.uniform(-1, 1, (p.nz, p.nx)).astype(np.float32) * 1e-3) for _ in range(3)]
    d, cs = _observed(p, w, coefs, m, geo, 1e-3)
    out = {}
    monkeypatch.setenv("RTM_MARCH_W", mw)
    monkeypatch.setenv("RTM_MARCH_NW", nw)
    for tb in ("0", "2"):
        monkeypatch.setenv("RTM_TBLOCK", tb)
        op3 = hip.rtm_op3(hist, coefs5)
        rec, last = hip.forward(_hip_shot(hip, p), coefs, w, want_last=True)
        born = hip.born(_hip_shot(hip, p), coefs, w, dm)
        img = hip.migrate(_hip_shot(hip, p, adj_abc=1, comp_image=1), cs, w, d)
        out[tb] = [*op3, rec, *last, born, *img]
    for k, (a, b) in enumerate(zip(out["0"], out["2"])):
        assert np.array_equal(a, b), k
