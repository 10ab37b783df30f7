"""Shared set-up for parity tests: the same seeded system for the oracle and
the HIP library."""
import numpy as np

from oracle import planewave as opw
from oracle import waves as owv
from pwgpu import model


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def oracle_waveset(system):
    pl = opw.Plane(system.rbas, system.epw_psi, system.epw_rho)
    st = owv.Setup(system.isp, system.lox_list, system.do_list())
    pro = system.pro_table(pl.g2)
    v = system.potential((pl.nr1, pl.nr2, pl.nr3))
    ws = owv.WaveSet(pl, st, pro, system.nb, system.occ, v, system.dh_matrices(),
                     system.dt, system.emass, system.emasscg2, system.annee)
    c = system.initial_psi(pl.g2)
    ws.psi0 = c.copy()
    ws.psim = c.copy()
    return ws


def gpu_waveset(system, prec=0):
    import pwgpu
    plane = pwgpu.PlaneWave(system.rbas, system.epw_psi, system.epw_rho)
    w = pwgpu.Waves.from_system(plane, system, prec=prec)
    return plane, w


SYSTEMS = {
    "tiny": model.tiny,
    "si8": model.si8,
    # Si8 compressed to a = 8.7 bohr: a 22^3 grid, so the y and z transforms take
    # a radix-11 stage (the in-place kernels' second instantiation)
    "si8c": lambda: model.si8(a=8.7),
    # two species with different projector sets (O: 2s2p, H: 2s), 12 atoms
    "h2o4": lambda: model.water(nmol=4, a=9.85, seed=3),
    # 256 bands, 64^3 grid, 512 projectors (the oracle takes ~11 s per etot)
    "synthetic": model.synthetic,
    # 32 H2O, 40 Ry, 80^3 grid, 128 bands, 384 projectors (~13 s per etot)
    "h2o32": model.water,
    # Si8 with the si2 deck's projector set 2 s + 2 p + 1 d (LMNX = 13)
    "si8d": lambda: model.si8(species=model.si_species_d()),
    # the Si8 and Si16 cells of the request mix (36^3 and 36 x 36 x 72 grids:
    # the x, y and z lengths of the compile-time FFT kernels besides 72)
    "si8s": lambda: model.si_supercell(8, seed=5),
    "si16": lambda: model.si_supercell(16, seed=6),
    # the benchmarked cell: 64 atoms, 144 bands, 23 847 G, 72^3, 512 projectors
    # (the oracle takes ~2 s per step)
    "si64": model.si64,
}
