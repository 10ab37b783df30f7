"""Request loop of independent wave-function relaxations (pw_server_* of
include/pwgpu.h).

Each request is a cell with its own plane-wave set; the server admits
requests first-fit by device footprint and slot, runs each on its own
stream and returns the relaxed energies with the request's submit, start and
end times (seconds on the server's clock).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr
from .planewave import PlaneWave
from .waves import do_array, dh_array


def request_arrays(system, psi=None):
    """The request's arrays in the layouts of the cell's plane-wave set."""
    plane = PlaneWave(system.rbas, system.epw_psi, system.epw_rho)
    try:
        g2 = plane.getr8a("G2")
        lox_list = system.lox_list
        lnx = np.array([len(l) for l in lox_list], dtype=np.int32)
        lnxx = int(lnx.max())
        lox = np.zeros((len(lox_list), lnxx), dtype=np.int32)
        for s, l in enumerate(lox_list):
            lox[s, :len(l)] = l
        lmnxx = max(sum(2 * l + 1 for l in lo) for lo in lox_list)
        pro = np.ascontiguousarray(system.pro_table(g2), dtype=np.float64)
        arr = {
            "isp": (np.asarray(system.isp) + 1).astype(np.int32),
            "lnx": lnx, "lox": np.ascontiguousarray(lox), "lnxx": lnxx,
            "pro": pro, "nbarepro": pro.size // plane.ngl,
            "occ": np.asarray(system.occ, dtype=np.float64),
            "dov": np.ascontiguousarray(do_array(len(lox_list), lnxx, system.do_list())),
            "dh": np.ascontiguousarray(dh_array(system.nat, lmnxx, system.dh_matrices())),
            "vofr": np.ascontiguousarray(system.potential(plane.nr), dtype=np.float64),
            "r": np.ascontiguousarray(system.pos, dtype=np.float64),
            "psi": np.ascontiguousarray(system.initial_psi(g2) if psi is None else psi, dtype=np.complex128),
        }
    finally:
        plane.close()
    return arr


class Server:
    def __init__(self, mem_budget_bytes, max_active):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.pw_server_create(int(mem_budget_bytes), int(max_active), ctypes.byref(h)))
        self.handle = h

    def submit(self, system, prec=1, friction=0.1, etol=1.0e-5, nconv=3, max_steps=400, autotune=True,
               arrays=None, rtol=0.0):
        """rtol > 0: converged when the residual (the largest super-band norm
        of HPSI - O psi Lambda, Ha) stays below rtol for nconv steps; else
        when |dE| < etol for nconv steps."""
        a = request_arrays(system) if arrays is None else arrays
        q = _lib.Request()
        q.rbas[:] = list(np.asarray(system.rbas, dtype=np.float64).T.ravel())
        q.epw_psi, q.epw_rho = system.epw_psi, system.epw_rho
        q.nb, q.nat, q.nsp = system.nb, system.nat, len(system.lox_list)
        q.lnxx, q.nbarepro, q.prec = a["lnxx"], a["nbarepro"], int(prec)
        q.isp, q.lnx, q.lox = iptr(a["isp"]), iptr(a["lnx"]), iptr(a["lox"])
        q.pro, q.occ, q.dov, q.dh = dptr(a["pro"]), dptr(a["occ"]), dptr(a["dov"]), dptr(a["dh"])
        q.vofr, q.r, q.psi = dptr(a["vofr"]), dptr(a["r"]), dptr(a["psi"])
        q.dt, q.emass, q.emasscg2, q.friction = system.dt, system.emass, system.emasscg2, friction
        q.etol, q.nconv, q.max_steps, q.autotune = etol, nconv, max_steps, int(bool(autotune))
        q.rtol = float(rtol)
        rid = ctypes.c_longlong()
        check(self.lib.pw_server_submit(self.handle, ctypes.byref(q), ctypes.byref(rid)))
        return rid.value

    def wait(self, rid):
        r = _lib.Result()
        check(self.lib.pw_server_wait(self.handle, ctypes.c_longlong(rid), ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in _lib.Result._fields_}

    def stats(self):
        b, a, p = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int()
        check(self.lib.pw_server_stats(self.handle, ctypes.byref(b), ctypes.byref(a), ctypes.byref(p)))
        r, c, i = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        check(self.lib.pw_server_pool_stats(self.handle, ctypes.byref(r), ctypes.byref(c), ctypes.byref(i)))
        return {"peak_bytes": b.value, "peak_active": a.value, "pending": p.value,
                "pool_reused": r.value, "pool_created": c.value, "pool_idle_bytes": i.value}

    def close(self):
        if self.handle:
            self.lib.pw_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
