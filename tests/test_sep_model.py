"""SEP .H/.H@ input/output and the Lame model builder (host side, CPU).

Fixtures under tests/golden/sep/ are headers and parameter files the
reference ships (R/Dat/*.H, R/Dat_check/*.H, R/Par/*.p); their payloads are
not in the reference, so the data checks use payloads written here."""
import os

import numpy as np
import pytest

from elastic_amd import model_builder as MB
from elastic_amd import sep, synth

GOLD = os.path.join(os.path.dirname(__file__), "golden", "sep")


def test_reference_headers_parse():
    h = sep.read_header(os.path.join(GOLD, "filter.H"))
    assert h["n1"] == "45" and h["in"] == "/data/lin7/filter.H@" and h["data_format"] == "xdr_float"
    assert sep.axes(h) == [(45, 0.0, 1.0, "")]
    h = sep.read_header(os.path.join(GOLD, "seis-Vx-layer.H"))
    assert sep.axes(h) == [(3535, 0.0, 0.000565685, ""), (300, 0.0, 4.0, "")]
    # the header's time axis is the reference's dt / nt rule for dx = 4, Ttime = 2
    dt, nt = synth.reference_dt_nt(4.0, 2.0)
    assert nt == 3535 and abs(dt - 0.000565685) < 1e-9
    # a history: Smooth's later 'sets next: in=' wins over the first program's
    h = sep.read_header(os.path.join(GOLD, "lambdaS-layer.H"))
    assert h["in"] == "/data/lin7/Smooth@Wz2BOs"
    assert sep.axes(h) == [(300, 0.0, 4.0, "z"), (300, 0.0, 4.0, "x")]
    assert h["rect1"] == "20" and h["lambda2"] == "6e+09" and h["ref"] == "yes"
    h = sep.read_header(os.path.join(GOLD, "drho-layer-check.H"))
    assert h["in"] == "/data/lin7/drho-layer-check.H@" and h["copy_intag"] == "Dat/drho-layer.H"


def test_xdr_roundtrip_and_byte_order(tmp_path):
    a = np.asfortranarray(np.arange(24, dtype=np.float32).reshape(4, 6, order="F") * 0.5 - 3)
    p = str(tmp_path / "a.H")
    sep.write(p, a, [(0.0, 4.0, "z"), (10.0, 4.0, "x")])
    b, h = sep.read(p)
    assert np.array_equal(a, b) and b.flags.f_contiguous
    assert sep.axes(h) == [(4, 0.0, 4.0, "z"), (6, 10.0, 4.0, "x")]
    raw = open(p + "@", "rb").read()
    assert raw[:4] == np.array([-3.0], ">f4").tobytes() == b"\xc0\x40\x00\x00"
    assert len(raw) == 4 * a.size
    # native and double formats, memory-mapped read
    for fmt in ("native_float", "xdr_double"):
        q = str(tmp_path / f"{fmt}.H")
        sep.write(q, a, data_format=fmt)
        c, _ = sep.read(q, mmap=True)
        assert np.array_equal(a, c)


def test_payload_found_next_to_header_and_size_checked(tmp_path):
    a = np.ones((3, 5), np.float32)
    d = tmp_path / "sub"
    d.mkdir()
    p = str(d / "x.H")
    sep.write(p, a)
    # move the pair elsewhere: in= is stale, the payload sits next to the header
    new = tmp_path / "moved"
    new.mkdir()
    os.replace(p, new / "x.H")
    os.replace(p + "@", new / "x.H@")
    b, _ = sep.read(str(new / "x.H"))
    assert np.array_equal(a, b)
    with open(new / "x.H", "a") as f:
        f.write("\n\t\tn2=6\n")
    with pytest.raises(ValueError):
        sep.read(str(new / "x.H"))


def test_parse_args_par_files_and_override(tmp_path):
    pars = sep.parse_args([f"par={os.path.join(GOLD, 'rtm-layer.p')}", "abc=10", "junk"])
    assert pars["xsource"] == "150" and pars["zsource"] == "50" and pars["abc"] == "10"
    assert sep.get(pars, "Ttime") == 2.0 and sep.get(pars, "free_surface", bool) is False
    assert sep.get(pars, "missing", int, 7) == 7
    with pytest.raises(KeyError):
        sep.get(pars, "missing")


def test_layer_model_matches_reference_parameters():
    p = MB.params_from(sep.parse_args([f"par={os.path.join(GOLD, 'layer-lambda.p')}"]))
    lam, mu, rho = MB.build(p)
    assert lam.shape == (300, 300) and lam.dtype == np.float32
    # rows 1..199 background, 200..300 the reflector (z1 .. z1+thick inclusive)
    assert np.all(lam[:199] == np.float32(5.12e9)) and np.all(mu[:199] == np.float32(2.88e9))
    assert np.all(rho[:199] == 2000.0)
    assert np.all(lam[199:] == np.float32(6e9)) and np.all(mu[199:] == np.float32(4e9)) and np.all(rho[199:] == 2500)
    # the parameters echoed in the reference's own header build the same model
    h = sep.read_header(os.path.join(GOLD, "lambdaS-layer.H"))
    lam2, mu2, rho2 = MB.build(MB.params_from(h))
    assert np.array_equal(lam, lam2) and np.array_equal(mu, mu2) and np.array_equal(rho, rho2)


def test_model_builder_features():
    base = dict(nz=40, dz=5.0, nx=30, dx=5.0)
    # water + Poisson low-velocity layer + gradient background with mu = lambda/2
    p = MB.params_from({**{k: str(v) for k, v in base.items()}, "water": "y", "waterz": "5", "lvl": "1",
                        "nlvl": "4", "lvlmu": "1e9", "lvlvz1": "1e8", "lvlrho": "1500", "lambda1": "4e9",
                        "vz": "1e7", "rho1": "2100"})
    lam, mu, rho = MB.build(p)
    assert np.all(lam[:5] == np.float32(2.25e9)) and np.all(mu[:5] == 0) and np.all(rho[:5] == 1000)
    for k in range(4):  # rows 6..9: mu = lvlmu + k*lvlvz1, lambda = mu
        assert np.all(mu[5 + k] == np.float32(1e9 + k * 1e8)) and np.all(lam[5 + k] == mu[5 + k])
    for iz in range(10, 41):  # background from row waterz+nlvl+1 = 10
        want = np.float32(np.float32(4e9) + np.float32(1e7) * np.float32(iz - 9))
        assert np.all(lam[iz - 1] == want) and np.all(mu[iz - 1] == np.float32(0.5) * want)
    # reflector with mu2 = -1 halves the background mu; rows beyond nz clipped
    p2 = MB.params_from({**{k: str(v) for k, v in base.items()}, "lambda1": "4e9", "mu1": "2e9", "ref": "1",
                         "z1": "35", "thick": "10", "lambda2": "5e9", "rho2": "2400"})
    lam, mu, rho = MB.build(p2)
    assert np.all(lam[34:] == np.float32(5e9)) and np.all(mu[34:] == np.float32(1e9)) and np.all(lam[:34] == 4e9)
    # Gaussian anomaly, relative form: exactly (1+anomstr) at the centre
    p3 = MB.params_from({**{k: str(v) for k, v in base.items()}, "lambda1": "4e9", "mu1": "2e9", "rho1": "2000",
                         "anom": "1", "anomz": "20", "anomx": "15", "anomsigma": "3", "anomstr": "0.1"})
    lam, mu, rho = MB.build(p3)
    assert lam[19, 14] == np.float32(np.float32(4e9) * np.float32(1.1))
    assert abs(lam[19, 14 + 3] / 4e9 - (1 + 0.1 * np.exp(-0.5))) < 1e-6
    # blend form: the centre takes the anomaly values
    p4 = dict(p3, anomstr=0.0, anomlambda=1e9, anommu=5e8, anomrho=1200.0)
    lam, mu, rho = MB.build(p4)
    assert lam[19, 14] == np.float32(1e9) and rho[19, 14] == np.float32(1200)
    # disc: squared distance compared with the radius value (reference rule)
    p5 = MB.params_from({**{k: str(v) for k, v in base.items()}, "lambda1": "4e9", "circle": "1", "circlez": "20",
                         "circlex": "15", "circleradius": "9", "circlelambda": "7e9", "circlemu": "1e9",
                         "circlerho": "3000"})
    lam, _, _ = MB.build(p5)
    z, x = np.meshgrid(np.arange(1, 41), np.arange(1, 31), indexing="ij")
    inside = (x - 15) ** 2 + (z - 20) ** 2 <= 9
    assert np.all(lam[inside] == np.float32(7e9)) and np.all(lam[~inside] != np.float32(7e9))
    assert inside.sum() == 29  # lattice points within distance 3


def test_model_builder_cli_writes_sep(tmp_path):
    out = str(tmp_path / "lambda.H")
    MB.main([f"par={os.path.join(GOLD, 'layer-lambda.p')}", f"out={out}", f"mu={tmp_path / 'mu.H'}",
             f"rho={tmp_path / 'rho.H'}"])
    lam, h = sep.read(out)
    assert lam.shape == (300, 300) and sep.axes(h)[0] == (300, 0.0, 4.0, "z")
    rho, _ = sep.read(str(tmp_path / "rho.H"))
    assert rho[0, 0] == 2000 and rho[-1, 0] == 2500
