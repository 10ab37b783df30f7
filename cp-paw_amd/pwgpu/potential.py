"""Host mirror of the density -> potential step POTENTIAL$VOFRHO
(src/paw_potential.f90:324-896) over the C ABI (pw_potential_tables,
pw_vofrho): one spin, the local-density functional of type 1, compensation
charges of the moments QLM(LMRXX,NAT) with l <= 4, no pair potential, no
isolation.

Potential(dplane, tables, lmrx) stands for the POTENTIAL module data of one
density plane-wave object (PLANEWAVE$SELECT('DENSITY')); .vofrho(...) for
POTENTIAL$VOFRHO.  Errors surface as PwError.
"""
import numpy as np

from . import _lib
from ._lib import check, dptr, iptr
from .planewave import DeviceArray


class Potential:
    """dplane: PlaneWave of the density cutoff; tables: (4, NSP, NGL) real
    PSCORG, VBARG, G0, V0 on its sphere; lmrx: LMRX per species (default 1);
    functional: DFT type, 1 (local exchange + Perdew-Zunger) or 10 (PBE)."""

    def __init__(self, dplane, tables, lmrx=None, functional=1):
        self.lib = _lib.load()
        self.plane = dplane
        t = np.ascontiguousarray(np.asarray(tables, dtype=np.float64))
        if t.ndim != 3 or t.shape[0] != 4 or t.shape[2] != dplane.ngl:
            raise ValueError("tables must be (4, NSP, NGL) on the density sphere")
        self.nsp = t.shape[1]
        self.lmrx = np.ones(self.nsp, dtype=np.int32) if lmrx is None else \
            np.ascontiguousarray(np.asarray(lmrx, dtype=np.int32))
        check(self.lib.pw_potential_tables(dplane.handle, self.nsp, iptr(self.lmrx), dptr(t)))
        self.functional = int(functional)
        check(self.lib.pw_potential_functional(dplane.handle, self.functional))

    def vofrho_device(self, isp, pos, qlm, rho_dev, vofr_dev):
        """POTENTIAL$VOFRHO on device arrays (NRL doubles); isp 0-based; qlm
        QLM(LMRXX,NAT) as (NAT, LMRXX), or (NAT,) for LMRXX = 1.  Returns
        (energies{ehartree, exc, rhob}, vqlm (shape of qlm), force(NAT,3))."""
        isp1 = np.ascontiguousarray(np.asarray(isp, dtype=np.int32) + 1)
        pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64))
        flat = np.ndim(qlm) == 1
        nat = len(isp1)
        q = np.ascontiguousarray(np.asarray(qlm, dtype=np.float64).reshape(max(nat, 1), -1) if nat else
                                 np.zeros((1, 1)))
        lmrxx = q.shape[1]
        en = np.zeros(3)
        vq = np.zeros((max(nat, 1), lmrxx))
        fo = np.zeros((max(nat, 1), 3))
        check(self.lib.pw_vofrho(self.plane.handle, nat, lmrxx, iptr(isp1), dptr(pos), dptr(q), rho_dev.ptr,
                                 vofr_dev.ptr, dptr(en), dptr(vq), dptr(fo)))
        vq = vq[:nat, 0] if flat else vq[:nat]
        return {"ehartree": en[0], "exc": en[1], "rhob": en[2]}, vq, fo[:nat]

    def vofrho(self, isp, pos, qlm, rho_r):
        """Convenience on host arrays: rho_r (NR3, NR2, NR1) real.  Returns the
        same dict as the oracle: v, ehartree, exc, rhob, vqlm, force."""
        nr1, nr2, nr3 = self.plane.nr
        r = DeviceArray.from_host(np.ascontiguousarray(np.asarray(rho_r, dtype=np.float64)).reshape(-1))
        v = DeviceArray((self.plane.nrl,), np.float64)
        en, vq, fo = self.vofrho_device(isp, pos, qlm, r, v)
        out = dict(en)
        out["v"] = v.to_host().reshape(nr3, nr2, nr1)
        out["vqlm"] = vq
        out["force"] = fo
        return out
