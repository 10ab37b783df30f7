"""The C-ABI library: loads, exports every entry point of include/pwgpu.h, and
its host-side PLANEWAVE set-up (grid rule, G order, MINUSG, G vectors) equals
the oracle.  No GPU compute is issued here."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import pwgpu
from pwgpu import _lib, model
from oracle import planewave as opw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pwgpu.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^int\s+(pw_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 30
    lib = ctypes.CDLL(_lib.LIBPATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIBPATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (pw_\w+)", out))
    assert set(names) <= exported


@pytest.mark.parametrize("rbas,epsi,erho", [
    (np.diag([10.26] * 3), 15.0, 30.0),
    (np.diag([20.52] * 3), 15.0, 60.0),
    (np.diag([19.7] * 3), 20.0, 80.0),
    (model.tiny().rbas, 3.0, 12.0),
])
def test_plan_matches_oracle(rbas, epsi, erho):
    p = pwgpu.PlaneWave(rbas, epsi, erho)
    o = opw.Plane(rbas, epsi, erho)
    assert p.ngl == o.ngl and p.nr == (o.nr1, o.nr2, o.nr3)
    assert np.array_equal(p.geti4a("IGVEC"), o.igvec)
    assert np.array_equal(p.geti4a("MINUSG") - 1, o.minusg)
    assert np.max(np.abs(p.getr8a("G2") - o.g2)) <= 1e-12 * np.max(o.g2)
    assert np.max(np.abs(p.getr8a("GVEC") - o.gvec)) <= 1e-12
    assert p.geti4("NGAMMA") == o.ngamma + 1
    assert abs(p.gweight - o.vol) < 1e-9 and abs(p.rweight - o.rweight) < 1e-12


def test_error_paths_return_status():
    lib = _lib.load()
    rb = np.ascontiguousarray(np.diag([10.0] * 3).T).ravel()
    h = ctypes.c_void_p()
    k = np.array([0.25, 0.0, 0.0])
    rc = lib.pw_plan_create(_lib.dptr(rb), _lib.dptr(k), 1, 5.0, 20.0, ctypes.byref(h))
    assert rc == 2 and "Gamma" in _lib.last_error()
    rc = lib.pw_plan_create(_lib.dptr(rb), None, 1, 5.0, 1.0, ctypes.byref(h))
    assert rc == 1
    p = pwgpu.PlaneWave(np.diag([10.0] * 3), 5.0, 20.0)
    with pytest.raises(pwgpu.PwError) as e:
        p.geti4("NOPE")
    assert e.value.code == 1
    rc = lib.pw_fft(p.handle, b"GTOR", 1, p.ngl + 1, ctypes.c_void_p(8), p.nrl, ctypes.c_void_p(8), 0)
    assert rc == 1 and "F(G)" in _lib.last_error()
    rc = lib.pw_fft(p.handle, b"XTOY", 1, p.ngl, ctypes.c_void_p(8), p.nrl, ctypes.c_void_p(8), 0)
    assert rc == 1


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(OSError):
        _lib.load(str(tmp_path / "absent.so"))


def test_bench_rejects_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stdout


def test_vofrho_requires_tables():
    # POTENTIAL$VOFRHO before its setup tables: a status, no GPU work
    lib = _lib.load()
    p = pwgpu.PlaneWave(np.diag([10.0] * 3), 5.0, 5.0)
    isp = np.array([1], dtype=np.int32)
    pos = np.zeros(3)
    qlm = np.zeros(1)
    en = np.zeros(3)
    fake = ctypes.c_void_p(256)
    rc = lib.pw_vofrho(p.handle, 1, 1, isp.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                       pos.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                       qlm.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), fake, fake,
                       en.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None, None)
    assert rc == 4
    assert "tables" in _lib.last_error()
