"""CPU oracle loader -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package. It wraps ``liboracle.so`` (the C++ restatement in
``rtm_oracle.cpp``) and, when present, ``_ref/librtm_ref.so`` (the reference's
own scatter-form CPU adjoint, built from its source by ``Makefile``).

Array conventions (reference layout, W/gpu.h:8-14):
  field     numpy (nz, nx)      Fortran order   -> index iz + nz*ix
  history   numpy (nz, nx, nt)  Fortran order   -> iz + nz*ix + nz*nx*it
  receivers numpy (5, nrx, nt)  C order         -> component f, trace ix, sample it
            (each component block is the reference's (nt, nx) layout it + nt*ix)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None


class Params(C.Structure):
    """Mirror of orc::Params in rtm_oracle.cpp."""

    _fields_ = [
        ("nz", C.c_int), ("nx", C.c_int), ("nt", C.c_int),
        ("dx", C.c_double), ("dt", C.c_double),
        ("xs", C.c_int), ("zs", C.c_int),
        ("zr", C.c_int), ("rx0", C.c_int), ("nrx", C.c_int),
        ("abc", C.c_int), ("adj_abc", C.c_int), ("snap_every", C.c_int),
        ("nthreads", C.c_int),
    ]


def build():
    """Compile liboracle.so (and _ref when the reference tree is mounted)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
    return _LIB


def ref_lib():
    """The reference's own CPU adjoint (R/Src_cuda_722-1/rtmlib-modC.c) or None."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "librtm_ref.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
    return _REF


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _suffix(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def _scal(dtype):
    return C.c_double if np.dtype(dtype) == np.float64 else C.c_float


def F(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


def fd_coefs(dtype=np.float64):
    out = np.zeros(5, dtype)
    getattr(lib(), "orc_fd_coefs_" + _suffix(dtype))(_p(out))
    return out


def model_op(rho, lam, mu, dx, dt, dtype=np.float64):
    """W/model-mod.f90:21-89 + rtmmod_init + aux_*_c -> (m1_x, m1_z, m2, m3, m2m3)."""
    nz, nx = rho.shape
    rho, lam, mu = F(rho, dtype), F(lam, dtype), F(mu, dtype)
    outs = [np.zeros((nz, nx), dtype, order="F") for _ in range(5)]
    s = _scal(dtype)
    getattr(lib(), "orc_model_op_" + _suffix(dtype))(
        C.c_int(nz), C.c_int(nx), s(dx), s(dt), _p(rho), _p(lam), _p(mu), *[_p(o) for o in outs])
    return tuple(outs)


def ricker(nt, dt, fp, nsource=250, dtype=np.float64):
    out = np.zeros(nt, dtype)
    s = _scal(dtype)
    getattr(lib(), "orc_ricker_" + _suffix(dtype))(C.c_int(nt), s(dt), s(fp), C.c_int(nsource), _p(out))
    return out


def abc_filter(abc, dtype=np.float64):
    out = np.zeros(abc + 5, dtype)
    getattr(lib(), "orc_abc_filter_" + _suffix(dtype))(C.c_int(abc), _p(out))
    return out


def abc_K(field, abc, dtype=np.float64):
    f = F(field, dtype).copy(order="F")
    nz, nx = f.shape
    getattr(lib(), "orc_abc_K_" + _suffix(dtype))(C.c_int(nz), C.c_int(nx), C.c_int(abc), _p(f))
    return f


OPS = ["Ax", "Bx", "Az", "Bz", "D", "E", "F", "G", "H", "J"]


def fwd_op(op, coef, model, data, dtype=np.float64):
    """data += op(model) (adj=.false., add=.true.). Returns a new array."""
    k = OPS.index(op) if isinstance(op, str) else op
    coef, model = F(coef, dtype), F(model, dtype)
    data = F(data, dtype).copy(order="F")
    nz, nx = data.shape
    getattr(lib(), "orc_fwd_op_" + _suffix(dtype))(C.c_int(k), C.c_int(nz), C.c_int(nx), _p(coef), _p(model), _p(data))
    return data


def adj_op(op, coef, model, data, dtype=np.float64):
    """model += op^T(data) in scatter form (adj=.true., add=.true.). Returns a new array."""
    k = OPS.index(op) if isinstance(op, str) else op
    coef, data = F(coef, dtype), F(data, dtype)
    model = F(model, dtype).copy(order="F")
    nz, nx = model.shape
    getattr(lib(), "orc_adj_op_scatter_" + _suffix(dtype))(C.c_int(k), C.c_int(nz), C.c_int(nx), _p(coef), _p(model), _p(data))
    return model


def rtm_op3_c(hist, coefs, form="gather", dtype=np.float64, nthreads=1):
    """In-place semantics of rtm_op3_c_ on copies. hist: 5 arrays (nz,nx,nt);
    coefs: (m1_x, m1_z, aux_m2, aux_m3, aux_m2m3). Returns 5 new histories."""
    h = [F(a, dtype).copy(order="F") for a in hist]
    c = [F(a, dtype) for a in coefs]
    nz, nx, nt = h[0].shape
    if form == "gather":
        getattr(lib(), "orc_rtm_op3_c_gather_" + _suffix(dtype))(
            C.c_int(nt), C.c_int(nz), C.c_int(nx), *[_p(a) for a in h], *[_p(a) for a in c], C.c_int(nthreads))
    else:
        getattr(lib(), "orc_rtm_op3_c_scatter_" + _suffix(dtype))(
            C.c_int(nt), C.c_int(nz), C.c_int(nx), *[_p(a) for a in h], *[_p(a) for a in c])
    return h


def ref_rtm_op3_c(hist, coefs, zrec=0):
    """Run the reference's own rtm_op3_c_ (19-argument scatter form) on copies.
    The five (nt,nx) arrays of that form are unused by it; zero arrays are passed."""
    r = ref_lib()
    if r is None:
        raise RuntimeError("oracle/_ref/librtm_ref.so is not built")
    h = [F(a, np.float32).copy(order="F") for a in hist]
    c = [F(a, np.float32) for a in coefs]
    nz, nx, nt = h[0].shape
    if nt < 59:
        raise ValueError("the reference adjoint prints progress with it%(nt/59): nt must be >= 59")
    dummy = [np.zeros((nt, nx), np.float32, order="F") for _ in range(5)]
    r.rtm_op3_c_(C.c_int(nt), C.c_int(nz), C.c_int(nx), C.c_int(zrec), *[_p(a) for a in h],
                 *[_p(a) for a in dummy], *[_p(a) for a in c])
    return h


def rtm_op3_prestep(nt, zrec, data, coefs, dtype=np.float64):
    """rtm_op3 pre-step and injection (W/rtmlib-mod.f90:91-146). data: 5 arrays (nt,nx)."""
    c = [F(a, dtype) for a in coefs]
    nz, nx = c[0].shape
    h = [np.zeros((nz, nx, nt), dtype, order="F") for _ in range(5)]
    d = [F(a, dtype) for a in data]
    getattr(lib(), "orc_rtm_op3_prestep_" + _suffix(dtype))(
        C.c_int(nt), C.c_int(nz), C.c_int(nx), C.c_int(zrec), *[_p(a) for a in h], *[_p(a) for a in d], *[_p(a) for a in c])
    return h


def make_params(nz, nx, nt, dx, dt, xs, zs, zr, rx0, nrx, abc, adj_abc=0, snap_every=1, nthreads=1):
    return Params(nz, nx, nt, dx, dt, xs, zs, zr, rx0, nrx, abc, adj_abc, snap_every, nthreads)


def forward(p: Params, wavelet, coefs, dtype=np.float64, want_hist=False, want_hist5=False):
    """Forward source propagation (W/rtm-mod-adj.f90:178-260) with receivers.
    Returns dict(rec=(5,nrx,nt), hist_vx, hist_vz, hist5)."""
    c = [F(a, dtype) for a in coefs]
    w = np.ascontiguousarray(wavelet, dtype)
    rec = np.zeros((5, p.nrx, p.nt), dtype)
    hvx = hvz = None
    if want_hist:
        hvx = np.zeros((p.nz, p.nx, p.nt), dtype, order="F")
        hvz = np.zeros((p.nz, p.nx, p.nt), dtype, order="F")
    h5 = [None] * 5
    if want_hist5:
        h5 = [np.zeros((p.nz, p.nx, p.nt), dtype, order="F") for _ in range(5)]
    getattr(lib(), "orc_forward_" + _suffix(dtype))(
        C.byref(p), _p(w), *[_p(a) for a in c], _p(rec), _p(hvx), _p(hvz), *[_p(a) for a in h5])
    return dict(rec=rec, hist_vx=hvx, hist_vz=hvz, hist5=h5 if want_hist5 else None)


def fwd_level(p: Params, cur, prev, coefs, apply_k=False, srcamp=None, dtype=np.float64):
    """One forward level X(L) = K(X(L-2) [+ src] + Op X(L-1)) (W/rtm-mod-adj.f90:225-252).
    cur: five (nz,nx) arrays of level L-2; prev: level L-1. Returns level L
    (new arrays). srcamp = dt*source(L-1) or None."""
    out = [np.array(F(a, dtype), order="F", copy=True) for a in cur]
    pv = [F(a, dtype) for a in prev]
    c = [F(a, dtype) for a in coefs]
    getattr(lib(), "orc_fwd_level_" + _suffix(dtype))(
        C.byref(p), *[_p(a) for a in out], *[_p(a) for a in pv], *[_p(a) for a in c],
        C.c_int(int(apply_k)), C.c_int(srcamp is not None), _scal(dtype)(0.0 if srcamp is None else srcamp))
    return out


def image_S(vx, vz, vx1, vz1, dx, dt, dtype=np.float64):
    """S1..S6 at one level (W/image-mod-rtm.f90:26-67). vx1/vz1 may be None (last level)."""
    vx, vz = F(vx, dtype), F(vz, dtype)
    vx1 = None if vx1 is None else F(vx1, dtype)
    vz1 = None if vz1 is None else F(vz1, dtype)
    nz, nx = vx.shape
    S = [np.zeros((nz, nx), dtype, order="F") for _ in range(6)]
    s = _scal(dtype)
    getattr(lib(), "orc_image_S_" + _suffix(dtype))(
        C.c_int(nz), C.c_int(nx), s(dx), s(dt), _p(vx), _p(vz), _p(vx1), _p(vz1), *[_p(a) for a in S])
    return S


def image_adj_level(adj5, S6, w=1.0, dm=None, dtype=np.float64):
    """One level of the Lame imaging condition, adjoint branch
    (W/image-mod-rtm.f90:154-162), accumulated into dm = (drho, dlambda,
    dmu) (zeros when None) with weight w. adj5: the five adjoint fields of
    the level; S6: S1..S6 of the level. Returns the three updated arrays."""
    a = [F(x, dtype) for x in adj5]
    nz, nx = a[0].shape
    S = np.concatenate([F(x, dtype).ravel(order="F") for x in S6])
    out = [np.zeros((nz, nx), dtype, order="F") if dm is None else np.array(F(x, dtype), order="F", copy=True)
           for x in (dm if dm is not None else range(3))]
    getattr(lib(), "orc_image_adj_level_" + _suffix(dtype))(
        C.c_int64(nz * nx), *[_p(x) for x in a], _p(S), _scal(dtype)(w), *[_p(x) for x in out])
    return out


def rtm(p: Params, wavelet, coefs, rec, dtype=np.float64):
    """Full single-shot RTM -> (drho, dlambda, dmu), each (nz, nx) Fortran order."""
    c = [F(a, dtype) for a in coefs]
    w = np.ascontiguousarray(wavelet, dtype)
    r = np.ascontiguousarray(rec, dtype)
    out = [np.zeros((p.nz, p.nx), dtype, order="F") for _ in range(3)]
    getattr(lib(), "orc_rtm_" + _suffix(dtype))(C.byref(p), _p(w), *[_p(a) for a in c], _p(r), *[_p(a) for a in out])
    return tuple(out)


def rtm_rsrc(p: Params, wavelet, coefs, rec, dtype=np.float64):
    """RTM with the R/Src adjoint (K applied in place to the stored levels
    it+1 and it+2 before every step, R/Src/rtm-mod-adj.f90:365-475), full
    histories, no decimation -> (drho, dlambda, dmu)."""
    c = [F(a, dtype) for a in coefs]
    w = np.ascontiguousarray(wavelet, dtype)
    r = np.ascontiguousarray(rec, dtype)
    out = [np.zeros((p.nz, p.nx), dtype, order="F") for _ in range(3)]
    getattr(lib(), "orc_rtm_rsrc_" + _suffix(dtype))(C.byref(p), _p(w), *[_p(a) for a in c], _p(r),
                                                     *[_p(a) for a in out])
    return tuple(out)


def rsrc_adjoint(p: Params, coefs, rec, dtype=np.float64):
    """The R/Src adjoint alone: the five stored adjoint histories (each
    (nz, nx, nt), Fortran order) after the whole loop."""
    c = [F(a, dtype) for a in coefs]
    r = np.ascontiguousarray(rec, dtype)
    h = np.zeros((5, p.nt, p.nx, p.nz), dtype)
    getattr(lib(), "orc_rsrc_adjoint_" + _suffix(dtype))(C.byref(p), *[_p(a) for a in c], _p(r), _p(h))
    return [np.asfortranarray(h[f].transpose(2, 1, 0)) for f in range(5)]


def born(p: Params, wavelet, coefs, dm, dtype=np.float64):
    """Born modelling (W/rtm-mod-adj.f90:273-375) -> receivers (5, nrx, nt)."""
    c = [F(a, dtype) for a in coefs]
    w = np.ascontiguousarray(wavelet, dtype)
    d = [F(a, dtype) for a in dm]
    rec = np.zeros((5, p.nrx, p.nt), dtype)
    getattr(lib(), "orc_born_" + _suffix(dtype))(C.byref(p), _p(w), *[_p(a) for a in c], *[_p(a) for a in d], _p(rec))
    return rec


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


# --------------------------------------------------------------------------- 3D acoustic
def ac3_weights():
    w = np.zeros(6)
    lib().ac3_weights_f64(_p(w))
    return w


def ac3_laplacian(p):
    p = np.asfortranarray(p, np.float64)
    out = np.zeros_like(p, order="F")
    n1, n2, n3 = p.shape
    lib().ac3_laplacian_f64(C.c_int(n1), C.c_int(n2), C.c_int(n3), _p(p), _p(out))
    return out


def ac3_taper(n, abc):
    """Per-axis sponge factors (W/abc-mod.f90 taper, 1 inside)."""
    k = np.ones(n)
    if abc <= 0:
        return k
    f = abc_filter(abc)
    abc2 = abc + 5
    for i1 in range(1, abc2 + 1):
        k[i1 - 1] *= f[abc2 - i1]
        k[n - abc2 + i1 - 1] *= f[i1 - 1]
    return k


def ac3_run(c, amp, src, zr, abc=0, mode=0, rec=None, j=5, snap=None, want_snap=False, want_img=False,
            dtype=np.float64, nthreads=1):
    """3D acoustic pass (acoustic3d_oracle.cpp). mode 0: source pass returning
    (rec (nt, n3, n2), snapshots (nsnap, n1, n2, n3) F-ordered per snapshot);
    mode 1: receiver pass with injected rec and imaging against snap."""
    c = np.asfortranarray(c, dtype)
    n1, n2, n3 = c.shape
    nt = len(amp)
    a = np.ascontiguousarray(amp, dtype)
    kz, kx, ky = (np.ascontiguousarray(ac3_taper(n, abc), dtype) for n in (n1, n2, n3))
    nsnap = nt // j
    if mode == 0:
        r = np.zeros((nt, n3, n2), dtype)
        s = np.zeros((max(nsnap, 1), n3, n2, n1), dtype) if want_snap else None
        img = None
    else:
        r = np.ascontiguousarray(rec, dtype)
        s = np.ascontiguousarray(snap, dtype)
        img = np.zeros((n3, n2, n1), dtype)
    getattr(lib(), "ac3_run_" + _suffix(dtype))(
        C.c_int(n1), C.c_int(n2), C.c_int(n3), C.c_int(nt), _p(c), _p(kz) if abc > 0 else None,
        _p(kx) if abc > 0 else None, _p(ky) if abc > 0 else None, C.c_int(mode), C.c_int(src[0]), C.c_int(src[1]),
        C.c_int(src[2]), _p(a), C.c_int(zr), _p(r), C.c_int(j), _p(s), _p(img), C.c_int(nthreads))
    if mode == 0:
        return r, s
    return img


# --- 3D elastic adjoint step (elastic3d_oracle.cpp) --------------------------
# Arrays are numpy (k, ny+10, nx+10, nz+10) C-ordered: k fields in the
# reference's order, each padded by 5 cells per face, z fastest.
EL3_FIELDS = ("Vy", "Vx", "Vz", "syy", "sxx", "szz", "sxy", "sxz", "syz")
EL3_COEFS = ("m1_y", "m1_x", "m1_z", "m2", "m3", "m2m3")


def _el3_dims(a):
    ny, nx, nz = (d - 10 for d in a.shape[1:])
    return ny, nz, nx


def el3_step(fin, fin1, fout, coef, dtype=np.float64, nthreads=1):
    """One step out = out + in1 + gather(in); returns the new out."""
    fin, fin1, coef = (np.ascontiguousarray(a, dtype) for a in (fin, fin1, coef))
    out = np.array(fout, dtype=dtype, order="C", copy=True)
    ny, nz, nx = _el3_dims(out)
    getattr(lib(), "el3_adj_step_" + _suffix(dtype))(
        C.c_int(ny), C.c_int(nz), C.c_int(nx), _p(fin), _p(fin1), _p(out), _p(coef), C.c_int(nthreads))
    return out


def el3_steps(b0, b1, b2, coef, nsteps, dtype=np.float64, nthreads=1):
    """nsteps with the reference's pointer rotation, buffers starting as
    (in, in1, out) = (b0, b1, b2). Returns (in, in1, out) after the run."""
    bufs = [np.array(b, dtype=dtype, order="C", copy=True) for b in (b0, b1, b2)]
    coef = np.ascontiguousarray(coef, dtype)
    ny, nz, nx = _el3_dims(bufs[0])
    slots = np.zeros(3, np.int32)
    getattr(lib(), "el3_adj_steps_" + _suffix(dtype))(
        C.c_int(ny), C.c_int(nz), C.c_int(nx), _p(bufs[0]), _p(bufs[1]), _p(bufs[2]), _p(coef), C.c_int(nsteps),
        C.c_int(nthreads), _p(slots))
    return tuple(bufs[s] for s in slots)


def el3_fwd(x, coef, nthreads=1):
    """The transpose of the step's gather, as the velocity-stress system (FP64)."""
    x = np.ascontiguousarray(x, np.float64)
    coef = np.ascontiguousarray(coef, np.float64)
    y = np.zeros_like(x)
    ny, nz, nx = _el3_dims(x)
    lib().el3_fwd_apply_f64(C.c_int(ny), C.c_int(nz), C.c_int(nx), _p(x), _p(y), _p(coef), C.c_int(nthreads))
    return y


def el3_random(ny, nz, nx, seed, k=9):
    """Reference-style test data (K3/rtmlib-modC.c:190-242): values
    (r % 1000) / 500 - 1 over the interior, a zero 5-cell rim."""
    rng = np.random.default_rng(seed)
    a = np.zeros((k, ny + 10, nx + 10, nz + 10), np.float32)
    a[:, 5:-5, 5:-5, 5:-5] = rng.integers(0, 1000, size=(k, ny, nx, nz)) / 500.0 - 1.0
    return a
