"""SEP header (.H) / data (.H@) reader and writer.

The reference reads and writes every model, seismogram and image through the
SEP library (``sep_read``/``sep_write``/``from_history``/``to_history`` in
W/rtm.f90:38-56, W/seis-mod-rtm.f90:21-142, W/create-model-lame.f90:88-112).
A SEP dataset is a text header that is an append-only history -- each program
that touched the data appends its lines -- plus a binary payload:

* parameters are ``key=value`` tokens anywhere in the text (also inside
  ``From par:`` and ``sets next:`` lines); a later token overrides an earlier
  one, so the last ``n1=`` of the history is the current size;
* ``in=path`` names the payload (written as ``header_path@`` by convention);
* ``data_format`` is ``xdr_float`` (big-endian IEEE float32, the default),
  ``native_float`` (host order), ``xdr_double``/``native_double`` or
  ``xdr_int``/``native_int``;
* axes are ``n1, o1, d1, label1`` (fastest) to ``nK...``; arrays map to
  numpy in Fortran order, shape (n1, n2, ...).

Parameter files (``par=file.p``, R/Par/*.p) use the same ``key=value`` syntax.
This module is host-side input/output; nothing here runs on the GPU.
"""
from __future__ import annotations

import datetime
import getpass
import os
import re
import shlex
import socket

import numpy as np

_TOKEN = re.compile(r'([A-Za-z_][A-Za-z0-9_.]*)=("(?:[^"]*)"|\'(?:[^\']*)\'|[^\s]*)')

_FORMATS = {
    "xdr_float": ">f4", "native_float": "=f4",
    "xdr_double": ">f8", "native_double": "=f8",
    "xdr_int": ">i4", "native_int": "=i4",
}


def parse_text(text):
    """key=value tokens of a SEP header or parameter file; later tokens win.
    Quoted values lose their quotes. Lines starting with '#' are comments."""
    out = {}
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("#"):
            continue
        for k, v in _TOKEN.findall(line):
            if len(v) >= 2 and v[0] == v[-1] and v[0] in "\"'":
                v = v[1:-1]
            out[k] = v
    return out


def read_header(path):
    with open(path, "r", errors="replace") as f:
        return parse_text(f.read())


def _num(v, typ):
    if typ is bool:
        return str(v).strip().lower() in ("1", "y", "yes", "t", "true", ".true.")
    return typ(float(v)) if typ is int else typ(v)


def get(pars, key, typ=float, default=None):
    """Typed parameter lookup (``from_param``/``from_history`` semantics):
    raises KeyError when the key is missing and no default is given."""
    if key in pars:
        return _num(pars[key], typ)
    if default is None:
        raise KeyError(f"missing parameter {key}")
    return default


def axes(h):
    """[(n, o, d, label)] of the header's axes, trailing n=1 axes dropped."""
    out = []
    k = 1
    while f"n{k}" in h:
        out.append((int(float(h[f"n{k}"])), float(h.get(f"o{k}", 0.0)), float(h.get(f"d{k}", 1.0)),
                    h.get(f"label{k}", "")))
        k += 1
    while len(out) > 1 and out[-1][0] == 1:
        out.pop()
    if not out:
        raise ValueError("SEP header has no n1")
    return out


def data_path(header_path, h):
    """Resolve in= of a header: as written, else next to the header."""
    p = h.get("in")
    if p is None:
        raise ValueError(f"{header_path}: no in= in the header")
    if os.path.isabs(p) and os.path.exists(p):
        return p
    cand = os.path.join(os.path.dirname(os.path.abspath(header_path)), p)
    if os.path.exists(cand):
        return cand
    cand = os.path.join(os.path.dirname(os.path.abspath(header_path)), os.path.basename(p))
    if os.path.exists(cand):
        return cand
    raise FileNotFoundError(f"{header_path}: payload {p} not found")


def read(path, dtype=np.float32, mmap=False):
    """(array in Fortran order with shape (n1, n2, ...), header dict)."""
    h = read_header(path)
    ax = axes(h)
    shape = tuple(a[0] for a in ax)
    fmt = _FORMATS.get(h.get("data_format", "xdr_float"))
    if fmt is None:
        raise ValueError(f"{path}: unsupported data_format {h.get('data_format')}")
    dp = data_path(path, h)
    count = int(np.prod(shape))
    if mmap:
        raw = np.memmap(dp, dtype=fmt, mode="r", shape=(count,))
    else:
        raw = np.fromfile(dp, dtype=fmt, count=count)
    if raw.size != count:
        raise ValueError(f"{path}: payload holds {raw.size} values, header says {count}")
    a = np.asarray(raw, dtype=dtype).reshape(shape, order="F")
    return a, h


def write(path, array, axes_=None, data_format="xdr_float", program="elastic_amd", extra=None, data=None):
    """Write a SEP header at path and its payload at data (default path + '@').
    axes_: [(o, d, label)] per axis (defaults o=0, d=1). extra: more key=value
    pairs for the history block."""
    a = np.asarray(array)
    if a.ndim == 0:
        a = a.reshape(1)
    fmt = _FORMATS[data_format]
    data = data or path + "@"
    np.asfortranarray(a).ravel(order="F").astype(fmt).tofile(data)
    try:
        who = f"{getpass.getuser()}@{socket.gethostname()}"
    except Exception:
        who = "unknown"
    lines = ["", "", f"{program}:   {who}   {datetime.datetime.now().ctime()}"]
    for k, n in enumerate(a.shape, start=1):
        o, d, lab = (axes_[k - 1] if axes_ and k - 1 < len(axes_) else (0.0, 1.0, ""))
        lines += [f"\t\tn{k}={n}", f"\t\to{k}={o:g}", f"\t\td{k}={d:g}"]
        if lab:
            lines.append(f'\t\tlabel{k}="{lab}"')
    for k, v in (extra or {}).items():
        lines.append(f"\t\t{k}={shlex.quote(str(v)) if ' ' in str(v) else v}")
    lines += [f'\t\tsets next: in="{os.path.abspath(data)}"', f'\t\tdata_format="{data_format}"', ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))
    return path


def parse_args(argv):
    """Command line in SEP style: key=value tokens plus ``par=file`` files,
    later tokens overriding earlier ones (files expand in place)."""
    pars = {}
    for tok in argv:
        if "=" not in tok:
            continue
        k, v = tok.split("=", 1)
        if k == "par":
            with open(v) as f:
                pars.update(parse_text(f.read()))
        else:
            pars[k] = v
    return pars
