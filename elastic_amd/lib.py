"""ctypes binding of librtm_hip.so (the C ABI declared in include/rtmlib_modC.h).

Every function here runs on the GPU through the HIP library. There is no CPU
fallback: when the library is missing or no HIP device is present the calls
raise.

Array conventions follow the reference layout (W/gpu.h:8-14):
  field      numpy (nz, nx), Fortran order (index iz + nz*ix)
  history    numpy (nz, nx, nt), Fortran order
  receivers  numpy (5, nrx, nt), C order: component f, trace ix, sample it
             (each component block is the reference (nt, nx) layout it + nt*ix)
  coef5      list of the five (nz, nx) coefficient arrays m1_x, m1_z, aux_m2,
             aux_m3, aux_m2m3 (already dt/dx scaled)
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librtm_hip.so")

RTM_OPT_ADJ_ABC, RTM_OPT_SNAP_EVERY, RTM_OPT_SNAP_PREC, RTM_OPT_COMP_IMAGE, RTM_OPT_REC_X0, RTM_OPT_REC_NX = range(1, 7)

_LIB = None


class RtmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"librtm_hip error {code}: {msg}")
        self.code = code


class RtmShot(C.Structure):
    """Mirror of rtm_shot_t."""

    _fields_ = [
        ("nz", C.c_int), ("nx", C.c_int), ("nt", C.c_int),
        ("dx", C.c_float), ("dt", C.c_float),
        ("xs", C.c_int), ("zs", C.c_int),
        ("zr", C.c_int), ("rx0", C.c_int), ("nrx", C.c_int),
        ("abc", C.c_int), ("adj_abc", C.c_int), ("snap_every", C.c_int),
        ("snap_prec", C.c_int), ("comp_image", C.c_int),
    ]


def load():
    """Load the in-tree HIP library; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m elastic_amd.build` first")
        _LIB = C.CDLL(LIB_PATH)
        _LIB.rtm_hip_last_error.restype = C.c_char_p
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _chk(code):
    if code != 0:
        raise RtmError(code, load().rtm_hip_last_error().decode())


def _f32F(a):
    return np.asfortranarray(a, dtype=np.float32)


def pack_fields(arrs):
    """List of (nz, nx) arrays -> contiguous (k, nx, nz) float32 block whose
    k-th slab is the k-th field in the reference layout."""
    return np.ascontiguousarray(np.stack([np.asarray(a, np.float32).T for a in arrs]))


def unpack_fields(block):
    return [np.asfortranarray(block[k].T) for k in range(block.shape[0])]


def shot(nz, nx, nt, dx, dt, xs, zs, zr, rx0=0, nrx=None, abc=0, adj_abc=0, snap_every=1,
         snap_prec=0, comp_image=0):
    if nrx is None:
        nrx = nx - rx0
    return RtmShot(nz, nx, nt, dx, dt, xs, zs, zr, rx0, nrx, abc, adj_abc, snap_every, snap_prec, comp_image)


def rtm_op3(hist, coefs, zrec=0, inplace=False):
    """Drop-in adjoint rtm_op3_c_ semantics (status-returning form), on
    copies of the histories, or in place on Fortran-ordered float32 arrays
    (views along the level axis included) with inplace=True."""
    lib = load()
    if inplace:
        for a in hist:
            if a.dtype != np.float32 or not a.flags.f_contiguous:
                raise ValueError("in-place histories must be Fortran-contiguous float32")
        h = list(hist)
    else:
        h = [_f32F(a).copy(order="F") for a in hist]
    c = [_f32F(a) for a in coefs]
    nz, nx, nt = h[0].shape
    _chk(lib.rtm_op3_hip(C.c_int(nt), C.c_int(nz), C.c_int(nx), C.c_int(zrec), *[_p(a) for a in h], *[_p(a) for a in c]))
    return h


def rtm_op3_noalloc(hist, coefs, zrec=0):
    """rtm_op3_hip_noalloc (the allocation-free drop-in) on copies."""
    lib = load()
    h = [_f32F(a).copy(order="F") for a in hist]
    c = [_f32F(a) for a in coefs]
    nz, nx, nt = h[0].shape
    _chk(lib.rtm_op3_hip_noalloc(C.c_int(nt), C.c_int(nz), C.c_int(nx), C.c_int(zrec), *[_p(a) for a in h],
                                 *[_p(a) for a in c]))
    return h


def rtm_op3_c_void(hist, coefs, zrec=0):
    """Call the void drop-in symbol rtm_op3_c_ itself (in place on copies)."""
    lib = load()
    h = [_f32F(a).copy(order="F") for a in hist]
    c = [_f32F(a).copy(order="F") for a in coefs]
    nz, nx, nt = h[0].shape
    lib.rtm_op3_c_(C.c_int(nt), C.c_int(nz), C.c_int(nx), C.c_int(zrec), *[_p(a) for a in h], *[_p(a) for a in c])
    return h


def forward(s: RtmShot, coef5, wavelet, want_last=False):
    lib = load()
    cb = pack_fields(coef5)
    w = np.ascontiguousarray(wavelet, np.float32)
    rec = np.zeros((5, s.nrx, s.nt), np.float32)
    last = np.zeros((10, s.nx, s.nz), np.float32) if want_last else None
    _chk(lib.rtm_forward_hip(C.byref(s), _p(cb), _p(w), _p(rec), _p(last)))
    if want_last:
        return rec, unpack_fields(last)
    return rec


def forward_movie(s: RtmShot, coef5, wavelet, nsnap, z0, x0, wnz, wnx):
    """Frames of one forward run (rtm_forward_movie_hip): returns (movie,
    dsnap) with movie[k][f] the (wnz, wnx) window of field f at level
    (k+1)*dsnap, dsnap = nt // nsnap."""
    lib = load()
    cb = pack_fields(coef5)
    w = np.ascontiguousarray(wavelet, np.float32)
    out = np.zeros((nsnap, 5, wnx, wnz), np.float32)
    ds = C.c_int()
    _chk(lib.rtm_forward_movie_hip(C.byref(s), _p(cb), _p(w), C.c_int(nsnap), C.c_int(z0), C.c_int(x0), C.c_int(wnz),
                                   C.c_int(wnx), _p(out), C.byref(ds)))
    return np.ascontiguousarray(out.transpose(0, 1, 3, 2)), ds.value


def migrate(s: RtmShot, coef5, wavelet, rec):
    lib = load()
    cb = pack_fields(coef5)
    w = np.ascontiguousarray(wavelet, np.float32)
    r = np.ascontiguousarray(rec, np.float32)
    img = np.zeros((3, s.nx, s.nz), np.float32)
    _chk(lib.rtm_migrate_hip(C.byref(s), _p(cb), _p(w), _p(r), _p(img)))
    return tuple(unpack_fields(img))


def born(s: RtmShot, coef5, wavelet, dm):
    lib = load()
    cb = pack_fields(coef5)
    w = np.ascontiguousarray(wavelet, np.float32)
    d = pack_fields(dm)
    rec = np.zeros((5, s.nrx, s.nt), np.float32)
    _chk(lib.rtm_born_hip(C.byref(s), _p(cb), _p(w), _p(d), _p(rec)))
    return rec


def model_op(rho, lam, mu, dx, dt):
    """Staggered, dt/dx-scaled coefficients (host setup, FP32 reference order)."""
    lib = load()
    rho, lam, mu = _f32F(rho), _f32F(lam), _f32F(mu)
    nz, nx = rho.shape
    out = np.zeros((5, nx, nz), np.float32)
    _chk(lib.rtm_model_op(C.c_int(nz), C.c_int(nx), C.c_float(dx), C.c_float(dt), _p(rho), _p(lam), _p(mu), _p(out)))
    return unpack_fields(out)


def ricker(nt, dt, fp, nsource=250):
    lib = load()
    out = np.zeros(nt, np.float32)
    _chk(lib.rtm_ricker(C.c_int(nt), C.c_float(dt), C.c_float(fp), C.c_int(nsource), _p(out)))
    return out


def abc_k(field, abc):
    lib = load()
    f = _f32F(field).copy(order="F")
    nz, nx = f.shape
    _chk(lib.abc_k_hip(C.c_int(nz), C.c_int(nx), C.c_int(abc), _p(f)))
    return f


def cfl_check(vmax, dt, dx):
    lib = load()
    cr, lim = C.c_double(), C.c_double()
    code = lib.rtm_cfl_check(C.c_float(vmax), C.c_float(dt), C.c_float(dx), C.byref(cr), C.byref(lim))
    return code == 0, cr.value, lim.value


def timings():
    """[fwd ms, fwd launches, adj ms, adj launches, total ms] of the last call."""
    lib = load()
    out = np.zeros(5, np.float64)
    lib.rtm_last_timings(_p(out))
    return out


class RtmOperator:
    """Mirror of the Fortran module rtm_mod_adj (W/rtm-mod-adj.f90):
    ``RtmOperator(...)`` is rtm_init (:43-117) and ``op(adj, ...)`` is rtm_op
    (:120-129). xsource, zsource, zrec are Fortran 1-based indices."""

    def __init__(self, nz, dz, nx, dx, dt, xsource, zsource, zrec, fp, nt, abc, free_surface, m1, m2, m3, lame=True):
        lib = load()
        self.nz, self.nx, self.nt = nz, nx, nt
        self._ctx = C.c_void_p()
        m1, m2, m3 = _f32F(m1), _f32F(m2), _f32F(m3)
        _chk(lib.rtm_init_c_(C.byref(self._ctx), C.c_int(nz), C.c_float(dz), C.c_int(nx), C.c_float(dx), C.c_float(dt),
                             C.c_int(xsource), C.c_int(zsource), C.c_int(zrec), C.c_float(fp), C.c_int(nt), C.c_int(abc),
                             C.c_int(int(free_surface)), _p(m1), _p(m2), _p(m3), C.c_int(int(lame))))

    def set_option(self, key, value):
        _chk(load().rtm_set_option_c_(self._ctx, C.c_int(key), C.c_int(value)))

    def op(self, adj, dm1, dm2, dm3, Vx, Vz, sigmaxx, sigmazz, sigmaxz):
        """In place, like the Fortran rtm_op: adj=True writes dm1..dm3 (drho,
        dlambda, dmu) from the (nt, nrx) data; adj=False writes the data."""
        arrs = [dm1, dm2, dm3, Vx, Vz, sigmaxx, sigmazz, sigmaxz]
        for a in arrs:
            if a.dtype != np.float32 or not a.flags.f_contiguous:
                raise ValueError("rtm_op arrays must be float32 Fortran-ordered (written in place)")
        _chk(load().rtm_op_c_(self._ctx, C.c_int(int(adj)), *[_p(a) for a in arrs]))
        return 0

    def close(self):
        if self._ctx:
            load().rtm_final_c_(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------------------- batch
class RtmBatchShot(C.Structure):
    """Mirror of rtm_batch_shot_t."""

    _fields_ = [(n, C.c_int) for n in ("nz", "nx", "xs", "zs", "zr", "rx0", "nrx", "img_z0", "img_x0", "img_nz",
                                       "img_nx", "stack_z0", "stack_x0")]


class RtmBatchCfg(C.Structure):
    """Mirror of rtm_batch_cfg_t."""

    _fields_ = [
        ("nt", C.c_int), ("dx", C.c_float), ("dt", C.c_float), ("abc", C.c_int), ("adj_abc", C.c_int),
        ("snap_every", C.c_int), ("snap_prec", C.c_int), ("comp_image", C.c_int), ("txr", C.c_int),
        ("use_graph", C.c_int), ("max_group", C.c_int), ("stack_nz", C.c_int), ("stack_nx", C.c_int),
        ("mem_fraction", C.c_double),
    ]


SNAP_PREC = {"fp32": 0, "bf16": 1, "fp16": 2}


class Batch:
    """Multi-shot RTM on device-resident inputs (rtm_batch_* C API)."""

    def __init__(self, nt, dx, dt, shots, coefs, wavelet, abc=0, adj_abc=0, snap_every=1, snap_prec="fp32",
                 comp_image=0, txr=0, use_graph=0, max_group=0, stack_shape=(0, 0), mem_fraction=0.0):
        lib = load()
        self.nt = nt
        self.shots = [dict(s) for s in shots]
        n = len(shots)
        arr = (RtmBatchShot * n)()
        for i, s in enumerate(shots):
            for k, _ in RtmBatchShot._fields_:
                setattr(arr[i], k, int(s.get(k, 0)))
        cfg = RtmBatchCfg(nt, dx, dt, abc, adj_abc, snap_every, SNAP_PREC[snap_prec] if isinstance(snap_prec, str) else snap_prec,
                          comp_image, txr, use_graph, max_group, stack_shape[0], stack_shape[1], mem_fraction)
        self._keep = [pack_fields(c) for c in coefs]
        ptrs = (C.c_void_p * n)(*[k.ctypes.data for k in self._keep])
        w = np.ascontiguousarray(wavelet, np.float32)
        self._h = C.c_void_p()
        _chk(lib.rtm_batch_create(C.byref(self._h), C.byref(cfg), C.c_int(n), arr, ptrs, _p(w)))
        self._keep = None  # the library copied the coefficients

    def set_true_model(self, coefs):
        keep = [pack_fields(c) for c in coefs]
        ptrs = (C.c_void_p * len(keep))(*[k.ctypes.data for k in keep])
        _chk(load().rtm_batch_set_true_model(self._h, ptrs))

    def synthesize(self):
        _chk(load().rtm_batch_synthesize(self._h))

    def set_data(self, i, rec):
        r = np.ascontiguousarray(rec, np.float32)
        _chk(load().rtm_batch_set_data(self._h, C.c_int(i), _p(r)))

    def get_data(self, i):
        s = self.shots[i]
        rec = np.zeros((5, s["nrx"], self.nt), np.float32)
        _chk(load().rtm_batch_get_data(self._h, C.c_int(i), _p(rec)))
        return rec

    def migrate(self):
        _chk(load().rtm_batch_migrate(self._h))

    def image(self, i):
        s = self.shots[i]
        img = np.zeros((3, s["nx"], s["nz"]), np.float32)
        _chk(load().rtm_batch_get_image(self._h, C.c_int(i), _p(img)))
        return tuple(unpack_fields(img))

    def stack(self, shape):
        img = np.zeros((3, shape[1], shape[0]), np.float32)
        _chk(load().rtm_batch_get_stack(self._h, _p(img)))
        return tuple(unpack_fields(img))

    @property
    def txr(self):
        """Tile width the batch runs with (autotuned when created with txr=0)."""
        return int(load().rtm_batch_txr(self._h))

    @property
    def tblock(self):
        """Kernel form: 0 one level per launch on tiles, 1 two levels per
        launch (temporal blocking), 2 one level per launch marching along x,
        3 the marching forward with the tile adjoint."""
        return int(load().rtm_batch_tblock(self._h))

    @property
    def march_w(self):
        """Strip width (columns) of the marching forms."""
        return int(load().rtm_batch_march_w(self._h))

    def info(self):
        out = np.zeros(8, np.float64)
        _chk(load().rtm_batch_info(self._h, _p(out)))
        keys = ["fwd_ms", "fwd_launches", "adj_ms", "adj_launches", "total_ms", "fwd_cells", "adj_cells", "group_size"]
        return dict(zip(keys, out.tolist()))

    def set_learned(self, net):
        """Use the learned absorbing layer `net` (elastic_amd.learned.Net) in
        this batch's forward runs, or drop it with None; the network must stay
        alive while the batch runs."""
        self._net = net
        _chk(load().rtm_batch_set_learned(self._h, net.h if net is not None else None))

    def set_learned_adjoint(self, net):
        """Use the learned layer `net` (trained on adjoint states) in this
        batch's adjoint runs, or drop it with None; the network must stay
        alive while the batch runs."""
        self._net_adj = net
        _chk(load().rtm_batch_set_learned_adjoint(self._h, net.h if net is not None else None))

    def close(self):
        if self._h:
            load().rtm_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(nt, dx, dt, shots, coefs, wavelet, abc=0, adj_abc=0, snap_every=1, snap_prec="fp32", comp_image=0,
             use_graph=0):
    """Run the tile autotuner on a workload (rtm_batch_autotune): returns
    ((txr, form), {(txr, form): ms}); form 0 one level per launch on tiles,
    1 two levels per launch, 2 marching along x, 3 marching forward with
    the tile adjoint (the (2, 2) entry times the faster of forms 2 and 3)."""
    lib = load()
    n = len(shots)
    arr = (RtmBatchShot * n)()
    for i, s in enumerate(shots):
        for k, _ in RtmBatchShot._fields_:
            setattr(arr[i], k, int(s.get(k, 0)))
    cfg = RtmBatchCfg(nt, dx, dt, abc, adj_abc, snap_every, SNAP_PREC[snap_prec] if isinstance(snap_prec, str) else snap_prec,
                      comp_image, 0, use_graph, 0, 0, 0, 0.0)
    keep = [pack_fields(c) for c in coefs]
    ptrs = (C.c_void_p * n)(*[k.ctypes.data for k in keep])
    w = np.ascontiguousarray(wavelet, np.float32)
    best, tb = C.c_int(), C.c_int()
    ms = np.zeros(8, np.float64)
    _chk(lib.rtm_batch_autotune(C.byref(cfg), C.c_int(n), arr, ptrs, _p(w), C.byref(best), C.byref(tb), _p(ms)))
    cand = [(1, 0), (2, 0), (4, 0), (8, 0), (1, 1), (2, 1), (4, 1), (2, 2)]
    times = {c: ms[i] for i, c in enumerate(cand) if ms[i] >= 0}
    return (best.value, tb.value), times


def bandwidth_probe(nbytes=1 << 30, reps=20, mode=0):
    """Achieved GB/s (read + written) of a streaming kernel and ms per launch.
    mode 0 copy, 1 non-temporal copy, 2 read only, 3 write only."""
    g, ms = C.c_double(), C.c_double()
    _chk(load().rtm_bandwidth_probe(C.c_size_t(nbytes), C.c_int(reps), C.c_int(mode), C.byref(g), C.byref(ms)))
    return g.value, ms.value


def synchronize():
    _chk(load().rtm_hip_finalize())


# --------------------------------------------------------------------------- 3D acoustic
class Ac3dCfg(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("n1", "n2", "n3", "nt", "sz", "sx", "sy", "zr", "abc", "snap_every",
                                       "snap_prec")]


class Acoustic3D:
    """3D acoustic subdomain propagator + imaging (include/acoustic3d.h).
    Arrays: c (n1, n2, n3) Fortran order (z fastest); data (nt, n3, n2)."""

    def __init__(self, c, amp, src, zr, abc=40, snap_every=5, snap_prec="bf16"):
        lib = load()
        c = np.asfortranarray(c, np.float32)
        self.shape = c.shape
        self.nt = len(amp)
        cfg = Ac3dCfg(*c.shape, self.nt, *src, zr, abc, snap_every, 1 if snap_prec == "bf16" else 0)
        a = np.ascontiguousarray(amp, np.float32)
        self._h = C.c_void_p()
        _chk(lib.ac3d_create(C.byref(self._h), C.byref(cfg), _p(c), _p(a)))

    def model_data(self, c_true=None):
        ct = None if c_true is None else np.asfortranarray(c_true, np.float32)
        _chk(load().ac3d_model_data(self._h, _p(ct)))

    def migrate(self):
        _chk(load().ac3d_migrate(self._h))

    def data(self):
        n1, n2, n3 = self.shape
        out = np.zeros((self.nt, n3, n2), np.float32)
        _chk(load().ac3d_get_data(self._h, _p(out)))
        return out

    def set_data(self, rec):
        r = np.ascontiguousarray(rec, np.float32)
        _chk(load().ac3d_set_data(self._h, _p(r)))

    def image(self):
        n1, n2, n3 = self.shape
        out = np.zeros((n3, n2, n1), np.float32)
        _chk(load().ac3d_get_image(self._h, _p(out)))
        return np.asfortranarray(out.transpose(2, 1, 0))

    def timings(self):
        out = np.zeros(6, np.float64)
        _chk(load().ac3d_timings(self._h, _p(out)))
        return out

    def close(self):
        if self._h:
            load().ac3d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --- 3D elastic adjoint step (include/elastic3d.h) ---------------------------
# Blocks are (k, ny+10, nx+10, nz+10) float32, C-ordered: k fields in the
# order Vy, Vx, Vz, syy, sxx, szz, sxy, sxz, syz (or the 6 coefficients
# m1_y, m1_x, m1_z, m2, m3, m2m3), each padded by 5 cells per face, z fastest.

def _el3_dims(block):
    ny, nx, nz = (d - 10 for d in block.shape[1:])
    return ny, nz, nx


def el3_gpu_func(fin, fin1, fout, coef):
    """The reference's host interface (rtm3d_gpu_init / func / final): one
    step on host arrays. Returns (new out block, [copy-in, compute, copy-out] s)."""
    lib = load()
    fin, fin1, coef = (np.ascontiguousarray(a, np.float32) for a in (fin, fin1, coef))
    out = np.array(fout, dtype=np.float32, order="C", copy=True)
    ny, nz, nx = _el3_dims(out)
    t = np.zeros(3, np.float32)
    _chk(lib.rtm3d_gpu_init(ny, nz, nx))
    try:
        ptrs = [_p(a[k]) for a in (fin, fin1, out) for k in range(9)] + [_p(coef[k]) for k in range(6)]
        _chk(lib.rtm3d_gpu_func(C.c_int(ny), C.c_int(nz), C.c_int(nx), *ptrs, None, _p(t)))
    finally:
        lib.rtm3d_gpu_final()
    return out, t


class DeviceArray:
    """A float32 array in device memory, allocated through the HIP runtime
    the library itself links (torch bundles a runtime of its own, and two
    runtimes in one process do not share a device)."""

    def __init__(self, host):
        h = np.ascontiguousarray(host, np.float32)
        self.shape = h.shape
        self.nbytes = h.nbytes
        rt = _hiprt()
        p = C.c_void_p()
        if rt.hipMalloc(C.byref(p), C.c_size_t(self.nbytes)) != 0:
            raise RtmError(-2, f"hipMalloc({self.nbytes}) failed")
        self.ptr = p.value
        self.put(h)

    def put(self, host):
        h = np.ascontiguousarray(host, np.float32)
        assert h.nbytes == self.nbytes
        if _hiprt().hipMemcpy(C.c_void_p(self.ptr), _p(h), C.c_size_t(self.nbytes), 1) != 0:
            raise RtmError(-3, "hipMemcpy to device failed")

    def get(self):
        h = np.empty(self.shape, np.float32)
        if _hiprt().hipMemcpy(_p(h), C.c_void_p(self.ptr), C.c_size_t(self.nbytes), 2) != 0:
            raise RtmError(-3, "hipMemcpy to host failed")
        return h

    def slab(self, k):
        """Device address of the k-th slab along the first axis."""
        return self.ptr + k * (self.nbytes // self.shape[0])

    def free(self):
        if self.ptr:
            _hiprt().hipFree(C.c_void_p(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_HIPRT = None


def _hiprt():
    global _HIPRT
    if _HIPRT is None:
        load()  # pulls in the runtime librtm_hip.so was linked with
        _HIPRT = C.CDLL("libamdhip64.so.7")
    return _HIPRT


def el3_steps_dev(bufs, coef, nsteps):
    """nsteps on device blocks (rtm3d_adj_steps_dev): bufs = [in, in1, out]
    DeviceArrays of shape (9, ny+10, nx+10, nz+10), coef (6, ...). Returns
    the three blocks in their rotated roles (in, in1, out)."""
    lib = load()
    ny, nz, nx = _el3_dims(bufs[0])
    P9 = C.c_void_p * 9
    arrs = [P9(*[C.c_void_p(b.slab(k)) for k in range(9)]) for b in bufs]
    cp = (C.c_void_p * 6)(*[C.c_void_p(coef.slab(k)) for k in range(6)])
    _chk(lib.rtm3d_adj_steps_dev(C.c_int(ny), C.c_int(nz), C.c_int(nx), arrs[0], arrs[1], arrs[2], cp,
                                 C.c_int(nsteps)))
    by_ptr = {b.ptr: b for b in bufs}
    return tuple(by_ptr[arrs[i][0]] for i in range(3))


def el3_timings():
    out = np.zeros(4, np.float64)
    _chk(load().rtm3d_timings(_p(out)))
    return out
