"""PLANEWAVE$FFT without allocation (pw_fft_noalloc) against pw_fft.

pw_fft_noalloc runs the same kernels as pw_fft with its work buffers in the
library's static device scratch, so both must give bit-identical results.
Two calls in a row on two plane-wave sets (the second smaller) check that the
shared scratch keeps no state that leaks from one call into the next; the
error paths (scratch too small, plan not yet on the device) are host-side
checks that issue no GPU work."""
import ctypes
import sys

import numpy as np
import pytest

from pwgpu import _lib

from harness import SYSTEMS


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _both(plane, key, nfft, prec, rng):
    """Outputs of pw_fft and of pw_fft_noalloc for the same seeded input."""
    import pwgpu
    rdt = np.complex128 if prec == 0 else np.complex64
    lib = _lib.load()
    if key == "GTOR":
        x = rng.normal(size=(nfft, plane.ngl)) + 1j * rng.normal(size=(nfft, plane.ngl))
        src = pwgpu.DeviceArray.from_host(x)
        shape = (nfft, plane.nr[2], plane.nr[1], plane.nr[0])
        a, b = pwgpu.DeviceArray(shape, rdt), pwgpu.DeviceArray(shape, rdt)
        plane.fft("GTOR", nfft, src, a, prec)
        plane.fft_noalloc("GTOR", nfft, src, b, prec)
    else:
        shape = (nfft, plane.nr[2], plane.nr[1], plane.nr[0])
        x = (rng.normal(size=shape) + 1j * rng.normal(size=shape)).astype(rdt)
        src = pwgpu.DeviceArray.from_host(x)
        a = pwgpu.DeviceArray((nfft, plane.ngl), np.complex128)
        b = pwgpu.DeviceArray((nfft, plane.ngl), np.complex128)
        plane.fft("RTOG", nfft, a, src, prec)
        plane.fft_noalloc("RTOG", nfft, b, src, prec)
    _lib.check(lib.pw_device_synchronize())
    return a.to_host(), b.to_host()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_fft_equals_pw_fft(prec):
    import pwgpu
    rng = np.random.default_rng(5)
    # first the 72^3 Si64 set with 5 transforms, then the 28^3 Si8 set with 2
    for name, nfft in (("si64", 5), ("si8", 2)):
        s = SYSTEMS[name]()
        plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
        for key in ("GTOR", "RTOG"):
            say(f"case noalloc {name} {key} nfft={nfft} prec={prec}")
            a, b = _both(plane, key, nfft, prec, rng)
            assert np.array_equal(a, b), f"{name} {key}: max diff {np.max(np.abs(a - b))}"
            assert np.max(np.abs(a)) > 0.0
        plane.close()


@pytest.mark.gpu
def test_noalloc_fft_batches_beyond_scratch():
    """More transforms than the scratch holds at once run in batches (a 72^3
    FP64 transform needs about 12 MB of work space; 9 transforms exceed the
    64 MiB), still bit-identical to pw_fft."""
    import pwgpu
    rng = np.random.default_rng(6)
    s = SYSTEMS["si64"]()
    plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
    for key in ("GTOR", "RTOG"):
        say(f"case noalloc batches si64 {key} nfft=9")
        a, b = _both(plane, key, 9, 0, rng)
        assert np.array_equal(a, b)
    plane.close()


def test_noalloc_fft_rejects_plan_not_on_device():
    """A plan whose tables were never uploaded is refused before any GPU work."""
    import pwgpu
    s = SYSTEMS["si8"]()
    plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
    lib = _lib.load()
    rc = lib.pw_fft_noalloc(plane.handle, b"GTOR", 1, plane.ngl, ctypes.c_void_p(16), plane.nrl,
                            ctypes.c_void_p(16), 0)
    assert rc == 4          # PW_ERR_STATE
    assert "not on the device" in _lib.last_error()
    plane.close()


def test_noalloc_fft_rejects_transform_beyond_scratch():
    """One transform larger than the static scratch is refused with a clear
    message, before any GPU work (a 50 bohr cubic cell at 50 Ha, a 160^3 grid)."""
    import pwgpu
    plane = pwgpu.PlaneWave(np.diag([50.0] * 3), 50.0, 50.0)
    lib = _lib.load()
    rc = lib.pw_fft_noalloc(plane.handle, b"GTOR", 1, plane.ngl, ctypes.c_void_p(16), plane.nrl,
                            ctypes.c_void_p(16), 0)
    assert rc == 1          # PW_ERR_ARG
    assert "static scratch" in _lib.last_error()
    plane.close()
