"""Input deck of the Fortran driver cp-paw_amd/fortran/pawgpu_cpstep.f90.

The deck is a native-endian stream of the driver's READ list (see the header
of pawgpu_cpstep.f90); the arrays are the keyword inputs of Waves.from_system
in the reference's Fortran layouts, so the Fortran host and the Python host
set up identical wave-function steps.
"""
import numpy as np

DECK_VERSION = 1


def write_deck(path, waves, system, nstep, prec=0, psi=None):
    """Write the deck for `system` on the plane-wave set of `waves` (a
    pwgpu.Waves whose plane defines NGL/NRL and the G order)."""
    plane = waves.plane
    inp = {key: (kind, val) for kind, key, val in waves.system_inputs(system, psi)}
    pro = np.ascontiguousarray(inp["PRO"][1], dtype=np.float64)
    nbarepro = pro.size // plane.ngl
    lox = np.zeros((waves.nsp, waves.lnxx), dtype=np.int32)        # LOX(LNXX,NSP)
    for s, lo in enumerate(waves.lox_list):
        lox[s, :len(lo)] = lo
    head = np.array([DECK_VERSION, waves.nat, waves.nsp, waves.lnxx, waves.nb, nbarepro, waves.lmnxx,
                     plane.ngl, plane.nrl, int(nstep), int(prec)], dtype=np.int32)
    reals = np.concatenate([np.asarray(system.rbas, dtype=np.float64).T.ravel(),
                            [system.epw_psi, system.epw_rho, system.dt, system.emass, system.emasscg2,
                             system.annee]]).astype(np.float64)
    with open(path, "wb") as f:
        head.tofile(f)
        reals.tofile(f)
        (np.asarray(waves.isp) + 1).astype(np.int32).tofile(f)
        np.asarray(waves.lnx, dtype=np.int32).tofile(f)
        lox.tofile(f)
        np.asarray(inp["R0"][1], dtype=np.float64).ravel().tofile(f)
        pro.ravel().tofile(f)
        np.asarray(inp["OCC"][1], dtype=np.float64).ravel().tofile(f)
        np.asarray(inp["DO"][1], dtype=np.float64).ravel().tofile(f)
        np.asarray(inp["DH"][1], dtype=np.complex128).ravel().tofile(f)
        np.asarray(inp["VOFR"][1], dtype=np.float64).ravel().tofile(f)
        np.asarray(inp["PSI0"][1], dtype=np.complex128).ravel().tofile(f)


def read_output(path, nat):
    """(steps, forces) from the driver's OUT file: steps rows are
    (istep, etot, ekin, epot, e1c, efict, orthoiter); forces (NAT,3)."""
    lines = [l.split() for l in open(path) if l.strip()]
    steps = [[float(x) for x in l] for l in lines[:-nat]]
    forces = np.array([[float(x) for x in l] for l in lines[-nat:]])
    return np.array(steps), forces
