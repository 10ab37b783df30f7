"""Host mirror of the PLANEWAVE object (src/paw_fft.f90) over the C ABI.

Names follow the reference interface: PlaneWave(...) stands for
PLANEWAVE$INITIALIZE, .geti4/.getr8/.getr8a for PLANEWAVE$GETxx and .fft for
PLANEWAVE$FFT.  Errors surface as PwError, where the reference stops with
ERROR$STOP.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr


class DeviceArray:
    """Device buffer owned by the caller (pw_malloc/pw_free)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctypes.c_void_p()
        check(_lib.load().pw_malloc(ctypes.byref(self.ptr), self.nbytes))

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(_lib.load().pw_memcpy_htod(d.ptr, a.ctypes.data_as(ctypes.c_void_p), d.nbytes))
        return d

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check(_lib.load().pw_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr:
            _lib.load().pw_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PlaneWave:
    """PLANEWAVE instance for a Gamma-point wave set with super wave functions.

    rbas: (3,3) lattice vectors as columns (RBAS); epw_psi / epw_rho in Hartree.
    """

    def __init__(self, rbas, epw_psi, epw_rho, kvec=(0.0, 0.0, 0.0), tinv=True):
        self.lib = _lib.load()
        rb = np.asfortranarray(np.asarray(rbas, dtype=np.float64))
        rb = np.ascontiguousarray(rb.T).ravel()          # column-major RBAS(3,3)
        kv = np.asarray(kvec, dtype=np.float64)
        h = ctypes.c_void_p()
        check(self.lib.pw_plan_create(dptr(rb), dptr(kv), int(bool(tinv)), float(epw_psi),
                                      float(epw_rho), ctypes.byref(h)))
        self.handle = h
        self.rbas = np.array(rbas, dtype=np.float64)
        self.kvec = np.array(kvec, dtype=np.float64)
        self.tinv = bool(tinv)
        self.ngl = self.geti4("NGL")
        self.nrl = self.geti4("NRL")
        self.nr = (self.geti4("NR1"), self.geti4("NR2"), self.geti4("NR3"))
        self.gweight = self.getr8("GWEIGHT")
        self.rweight = self.getr8("RWEIGHT")

    def close(self):
        if self.handle:
            self.lib.pw_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geti4(self, key):
        v = ctypes.c_int()
        check(self.lib.pw_plan_geti4(self.handle, key.encode(), ctypes.byref(v)))
        return v.value

    def getr8(self, key):
        v = ctypes.c_double()
        check(self.lib.pw_plan_getr8(self.handle, key.encode(), ctypes.byref(v)))
        return v.value

    def getr8a(self, key):
        n = 3 * self.ngl if key == "GVEC" else self.ngl
        out = np.empty(n, dtype=np.float64)
        check(self.lib.pw_plan_getr8a(self.handle, key.encode(), n, dptr(out)))
        return out.reshape(self.ngl, 3) if key == "GVEC" else out

    def geti4a(self, key):
        n = 3 * self.ngl if key == "IGVEC" else self.ngl
        out = np.empty(n, dtype=np.int32)
        check(self.lib.pw_plan_geti4a(self.handle, key.encode(), n, iptr(out)))
        return out.reshape(self.ngl, 3) if key == "IGVEC" else out

    def fft(self, key, nfft, fofg_dev, fofr_dev, prec=_lib.PW_PREC_FP64):
        """PLANEWAVE$FFT(ID,NFFT,NGL,FOFG,NRL,FOFR) on device buffers."""
        check(self.lib.pw_fft(self.handle, key.encode(), int(nfft), self.ngl, fofg_dev.ptr,
                              self.nrl, fofr_dev.ptr, int(prec)))

    def fft_noalloc(self, key, nfft, fofg_dev, fofr_dev, prec=_lib.PW_PREC_FP64):
        """PLANEWAVE$FFT on device buffers with the library's static work space
        (pw_fft_noalloc: no device allocation; the plan must be on the device)."""
        check(self.lib.pw_fft_noalloc(self.handle, key.encode(), int(nfft), self.ngl, fofg_dev.ptr,
                                      self.nrl, fofr_dev.ptr, int(prec)))

    def fft_host(self, key, fofg=None, fofr=None, prec=_lib.PW_PREC_FP64):
        """Convenience: PLANEWAVE$FFT on host arrays (copies in and out).

        fofg: (NFFT, NGL) complex128; fofr: (NFFT, NR3, NR2, NR1) complex.
        """
        rdt = np.complex128 if prec == _lib.PW_PREC_FP64 else np.complex64
        if key == "GTOR":
            g = DeviceArray.from_host(np.asarray(fofg, dtype=np.complex128))
            nfft = g.shape[0] if g.shape and len(g.shape) > 1 else 1
            r = DeviceArray((nfft, self.nr[2], self.nr[1], self.nr[0]), rdt)
            self.fft("GTOR", nfft, g, r, prec)
            check(self.lib.pw_device_synchronize())
            return r.to_host()
        r = DeviceArray.from_host(np.asarray(fofr, dtype=rdt))
        nfft = r.shape[0]
        g = DeviceArray((nfft, self.ngl), np.complex128)
        self.fft("RTOG", nfft, g, r, prec)
        check(self.lib.pw_device_synchronize())
        return g.to_host()
