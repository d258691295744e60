"""The C ABI library loads and exports every entry point include/*.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from elastic_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M):
            if m.group(1) not in ("if", "for", "while", "return", "sizeof"):
                names.add(m.group(1))
    return names


def test_header_declares_the_drop_in():
    names = _declared()
    assert "rtm_op3_c_" in names
    assert "rtm_op3_hip" in names


def test_library_exports_every_declared_symbol():
    so = ctypes.CDLL(lib.LIB_PATH)
    missing = [n for n in sorted(_declared()) if not hasattr(so, n)]
    assert not missing, missing


def test_rtm_op3_c_abi_signature_matches_reference():
    """14 arguments: four ints by value then ten float pointers (the Fortran
    interface at W/rtmlib-modF.f90:29-34)."""
    text = open(os.path.join(ROOT, "include", "rtmlib_modC.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"void rtm_op3_c_\((.*?)\);", text, re.S)
    args = [a.strip() for a in m.group(1).split(",")]
    assert len(args) == 14
    assert all(a.startswith("int ") for a in args[:4])
    assert all(a.startswith("float *") for a in args[4:])


def test_rtm_op3_rejects_grids_past_32_bit_cell_offsets():
    """Cells inside a level and inside the five coefficient arrays are
    addressed with 32-bit offsets; the entry point refuses 5*nz*nx >= 2^31
    with RTM_E_ARG before it touches a device (no compute call). Level and
    field offsets are 64-bit: the config-0 drop-in history 390 x 690 x 2000
    puts sigma_xz at element 4*nt*nz*nx, past 2^31, and is accepted."""
    assert 4 * 2000 * 390 * 690 >= 2 ** 31 > 5 * 390 * 690
    so = lib.load()
    dummy = (ctypes.c_float * 4)()
    ptrs = [ctypes.cast(dummy, ctypes.c_void_p)] * 10
    nz, nx = 20000, 21475  # 5*nz*nx = 2,147,500,000 >= 2^31
    rc = so.rtm_op3_hip(ctypes.c_int(8), ctypes.c_int(nz), ctypes.c_int(nx), ctypes.c_int(0), *ptrs)
    assert rc == -1
    assert "2^31" in so.rtm_hip_last_error().decode()
