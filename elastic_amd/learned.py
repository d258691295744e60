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


def features(lev_m1, lev_m2, coef5):
    """Normalised bf16 features (ncell, KPAD) and scales of the band cells
    of one grid from levels L-1 and L-2 and the grid's coefficients (five
    (nz, nx) arrays each)."""
    nz, nx = np.asarray(lev_m1[0]).shape
    nc = cell_count(nz, nx)
    a, b, c = pack_fields(lev_m1), pack_fields(lev_m2), pack_fields(coef5)
    f = np.zeros((nc, KPAD), np.float32)
    s = np.zeros(nc, np.float32)
    _chk(_L().labc_features(C.c_int(nz), C.c_int(nx), _p(a), _p(b), _p(c), _p(f), _p(s)))
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
