"""Host mirror of the WAVES object (src/paw_waves1.f90, src/paw_waves2.f90).

Method names follow the reference entry points: etot (WAVES$ETOT), propagate
(WAVES$PROPAGATE), orthogonalize (WAVES$ORTHOGONALIZE), switch (WAVES$SWITCH),
gramschmidt (WAVES$GRAMMSCHMIDT), projections (WAVES$PROJECTIONS('PSI0')) and
the keyword setters/getters WAVES$SETR8/GETR8 etc.  All state lives on the
device inside the library; the arrays below are host copies in the
reference's layouts (numpy C order with reversed Fortran dimensions).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr

FP64 = _lib.PW_PREC_FP64
MIXED = _lib.PW_PREC_MIXED


def dh_array(nat, lmnxx, dh_list):
    """DH(LMNXX,LMNXX,NAT) (numpy [atom][j][i]) from per-atom (LMNX,LMNX) matrices."""
    a = np.zeros((nat, lmnxx, lmnxx), dtype=np.complex128)
    for i, d in enumerate(dh_list):
        n = d.shape[0]
        a[i, :n, :n] = np.asarray(d).T
    return a


def do_array(nsp, lnxx, do_list):
    """DO(LNXX,LNXX,NSP) (numpy [species][j][i]) from per-species (LNX,LNX) matrices."""
    a = np.zeros((nsp, lnxx, lnxx))
    for s, d in enumerate(do_list):
        n = d.shape[0]
        a[s, :n, :n] = np.asarray(d).T
    return a


class Waves:
    def __init__(self, plane, nb, isp, lox_list, prec=FP64):
        self.lib = _lib.load()
        self.plane = plane
        self.nb = int(nb)
        self.nbh = (self.nb + 1) // 2
        self.isp = np.asarray(isp, dtype=np.int64)
        self.nat = len(self.isp)
        self.lox_list = [list(map(int, l)) for l in lox_list]
        self.nsp = len(self.lox_list)
        self.lnx = np.array([len(l) for l in self.lox_list], dtype=np.int32)
        self.lnxx = int(self.lnx.max())
        lox = np.zeros((self.nsp, self.lnxx), dtype=np.int32)      # LOX(LNXX,NSP)
        for s, l in enumerate(self.lox_list):
            lox[s, :len(l)] = l
        self.lmnx = [sum(2 * l + 1 for l in lo) for lo in self.lox_list]
        self.lmnxx = max(self.lmnx)
        isp1 = (self.isp + 1).astype(np.int32)
        h = ctypes.c_void_p()
        check(self.lib.pw_waves_create(plane.handle, self.nb, self.nat, self.nsp, iptr(isp1),
                                       iptr(self.lnx), self.lnxx, iptr(np.ascontiguousarray(lox)),
                                       int(prec), ctypes.byref(h)))
        self.handle = h
        self.prec = prec
        self.npro = self.geti4("NPRO")
        self.qlm_lmrxx = 1

    def close(self):
        if self.handle:
            self.lib.pw_waves_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- keyword interface ----------------------------------------------------
    def setr8(self, key, val):
        check(self.lib.pw_waves_setr8(self.handle, key.encode(), float(val)))

    def setr8a(self, key, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        check(self.lib.pw_waves_setr8a(self.handle, key.encode(), a.size, dptr(a)))

    def setc8a(self, key, arr):
        a = np.ascontiguousarray(arr, dtype=np.complex128).ravel()
        check(self.lib.pw_waves_setc8a(self.handle, key.encode(), a.size, dptr(a)))

    def getc8a(self, key):
        shapes = {"PSI0": (self.nbh, self.plane.ngl), "PSIM": (self.nbh, self.plane.ngl),
                  "HPSI": (self.nbh, self.plane.ngl), "OPSI": (self.nbh, self.plane.ngl),
                  "PROJ": (self.npro, self.nbh), "OPROJ": (self.npro, self.nbh),
                  "RLAM0": (self.nb, self.nb), "RLAMM": (self.nb, self.nb)}
        out = np.empty(shapes[key], dtype=np.complex128)
        check(self.lib.pw_waves_getc8a(self.handle, key.encode(), out.size, dptr(out)))
        return out.T.copy() if key in ("RLAM0", "RLAMM") else out

    def getr8a(self, key):
        L = self.lmnxx
        shapes = {"RHO": (self.plane.nr[2], self.plane.nr[1], self.plane.nr[0]),
                  "FORCE": (self.nat, 3), "EIG": (self.nb,), "MAT": (self.nb, self.nb),
                  "OMAT": (self.nb, self.nb), "OOMAT": (self.nb, self.nb),
                  "DENMAT": (self.nat, L, L), "VQLM": (self.nat, self.qlm_lmrxx),
                  "QLM": (self.nat, self.qlm_lmrxx)}
        out = np.empty(shapes[key], dtype=np.float64)
        check(self.lib.pw_waves_getr8a(self.handle, key.encode(), out.size, dptr(out)))
        if key in ("MAT", "OMAT", "OOMAT"):
            return out.T.copy()
        if key == "DENMAT":
            return np.transpose(out, (0, 2, 1)).copy()
        return out

    def getr8(self, key):
        v = ctypes.c_double()
        check(self.lib.pw_waves_getr8(self.handle, key.encode(), ctypes.byref(v)))
        return v.value

    def geti4(self, key):
        v = ctypes.c_int()
        check(self.lib.pw_waves_geti4(self.handle, key.encode(), ctypes.byref(v)))
        return v.value

    def devptr(self, key):
        p = ctypes.c_void_p()
        check(self.lib.pw_waves_devptr(self.handle, key.encode(), ctypes.byref(p)))
        return p

    # -- inputs in reference layouts ------------------------------------------
    def dh_array(self, dh_list):
        """DH(LMNXX,LMNXX,NAT) from a list of per-atom (LMNX,LMNX) matrices."""
        return dh_array(self.nat, self.lmnxx, dh_list)

    def do_array(self, do_list):
        """DO(LNXX,LNXX,NSP) from a list of per-species (LNX,LNX) matrices."""
        return do_array(self.nsp, self.lnxx, do_list)

    def set_dh(self, dh_list):
        self.setc8a("DH", self.dh_array(dh_list))

    def set_do(self, do_list):
        self.setr8a("DO", self.do_array(do_list))

    def set_matrix_c(self, key, m):
        self.setc8a(key, np.asarray(m, dtype=np.complex128).T)

    # -- the step ----------------------------------------------------------------
    def gramschmidt(self):
        check(self.lib.pw_waves_gramschmidt(self.handle))

    def projections(self):
        check(self.lib.pw_waves_projections(self.handle))

    def etot(self):
        check(self.lib.pw_waves_etot(self.handle))
        return {k.lower(): self.getr8(k) for k in ("EKIN", "EPOT", "E1C", "ETOT")}

    def autotune(self):
        """Launch shapes for this system shape, timed on the device
        (pw_waves_autotune; PSI0 and R0 must be set); cached per shape."""
        check(self.lib.pw_waves_autotune(self.handle))

    def set_potential(self, potential, qlm=None):
        """Link the density -> potential step (a pwgpu.Potential on the
        density object of this cell) into etot; qlm = QLM(LMRXX,NAT) as (NAT,
        LMRXX) or (NAT,), or None when the one-centre terms produce it."""
        if qlm is not None:
            q = np.asarray(qlm, dtype=np.float64).reshape(self.nat, -1)
            self.qlm_lmrxx = q.shape[1]
            self.setr8a("QLM", q)
        check(self.lib.pw_waves_set_potential(self.handle, potential.plane.handle))
        self._potential = potential            # keeps the density object alive

    def sphere_setup(self, isp, d):
        """Setup data of the one-centre terms for species isp (0-based), from
        a dict with the keys of pwgpu.model.Species.sphere_setup():
        r1, dex, nr, aez, rnuc, rcsm, lmrx, aephi/psphi (LNX, NR), aecore,
        pscore, vadd (NR), dekin (LNX, LNX) (pw_waves_sphere_setup)."""
        f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64))
        check(self.lib.pw_waves_sphere_setup(
            self.handle, int(isp) + 1, int(d["nr"]), float(d["r1"]), float(d["dex"]), float(d["aez"]),
            float(d["rnuc"]), float(d["rcsm"]), int(d["lmrx"]), dptr(f(d["aephi"])), dptr(f(d["psphi"])),
            dptr(f(d["aecore"])), dptr(f(d["pscore"])), dptr(f(d["vadd"])), dptr(f(np.asarray(d["dekin"]).T))))
        self.qlm_lmrxx = max(self.qlm_lmrxx, int(d["lmrx"]))

    def getc8a_dh(self):
        """DH(LMNXX,LMNXX,NAT) as per-atom (LMNXX, LMNXX) complex matrices [atom][i][j]."""
        L = self.lmnxx
        out = np.empty((self.nat, L, L), dtype=np.complex128)
        check(self.lib.pw_waves_getc8a(self.handle, b"DH", out.size, dptr(out)))
        return np.transpose(out, (0, 2, 1)).copy()

    def write_restart(self, path, static=False):
        """WAVES$WRITE into a restart file of the reference's format
        (pwgpu.restart.save)."""
        from . import restart
        restart.save(self, path, static=static)

    def read_restart(self, path):
        """WAVES$READ from a restart file of the reference's format
        (pwgpu.restart.load); False when it holds no WAVES section."""
        from . import restart
        return restart.load(self, path)

    def propagate(self):
        check(self.lib.pw_waves_propagate(self.handle))

    def orthogonalize(self):
        check(self.lib.pw_waves_orthogonalize(self.handle))

    def switch(self):
        check(self.lib.pw_waves_switch(self.handle))

    def step(self):
        check(self.lib.pw_waves_step(self.handle))

    def profile(self, enable=True):
        check(self.lib.pw_waves_profile(self.handle, int(bool(enable))))

    def kernel_time(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_int()
        check(self.lib.pw_waves_kernel_time(self.handle, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # -- set-up from a model system ----------------------------------------------
    def system_inputs(self, system, psi=None):
        """The keyword inputs of the step for a pwgpu.model.System, in call order:
        [(setter, ID, value)] with setter "r8", "r8a" or "c8a" and values in the
        reference layouts (PSIM = PSI0, R0 = RP = RM)."""
        g2 = self.plane.getr8a("G2")
        c = system.initial_psi(g2) if psi is None else psi
        return [("r8a", "PRO", system.pro_table(g2)),
                ("r8a", "DO", self.do_array(system.do_list())),
                ("c8a", "DH", self.dh_array(system.dh_matrices())),
                ("r8a", "OCC", np.asarray(system.occ, dtype=np.float64)),
                ("r8a", "VOFR", system.potential(self.plane.nr)),
                ("r8", "TIMESTEP", system.dt),
                ("r8", "EMASS", system.emass),
                ("r8", "EMASSCG2", system.emasscg2),
                ("r8", "FRICTION", system.annee),
                ("r8a", "R0", system.pos),
                ("r8a", "RM", system.pos),
                ("r8a", "RP", system.pos),
                ("c8a", "PSI0", c),
                ("c8a", "PSIM", c)]

    @classmethod
    def from_system(cls, plane, system, prec=FP64, psi=None):
        """Waves for a pwgpu.model.System: all inputs of the step set, PSIM=PSI0."""
        w = cls(plane, system.nb, system.isp, system.lox_list, prec=prec)
        for kind, key, val in w.system_inputs(system, psi):
            getattr(w, "set" + kind)(key, val)
        return w


def ortho_x(occ, chichi, psipsi, chipsi, lam):
    """WAVES_ORTHO_X on the GPU (FP64); numpy (NB,NB) matrices, returns (lambda, iter)."""
    lib = _lib.load()
    nb = len(occ)
    f = lambda m: np.ascontiguousarray(np.asarray(m, dtype=np.float64).T)
    occ = np.ascontiguousarray(occ, dtype=np.float64)
    a, b, c = f(chichi), f(psipsi), f(chipsi)
    l = f(lam)
    it = ctypes.c_int()
    check(lib.pw_ortho_x(nb, dptr(occ), dptr(a), dptr(b), dptr(c), dptr(l), ctypes.byref(it)))
    return l.T.copy(), it.value
