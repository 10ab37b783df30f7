"""GPU parity of the one-centre terms linked into the CP step: WAVES$MOMENTS ->
POTENTIAL$VOFRHO -> WAVES$SPHERE -> WAVES$ADDCONSTANTPOT on the device
(pw_waves_sphere_setup, src/paw_waves1.f90:1843-1915) against the FP64 oracle
with the same coupling (oracle/augmentation.py).

Tolerances (FP64 on both sides; radial sums differ in summation order):
energies within 1e-10 relative per step, QLM/VQLM/DH within 1e-10 of their
largest magnitude, forces within 1e-9, Lambda within 1e-9.
"""
import sys

import numpy as np
import pytest

from harness import gpu_waveset, relerr
from oracle import planewave as opw
from pwgpu import model
from test_oracle_augmentation import coupled_waveset

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def neutral_tiny():
    s = model.tiny()
    s.occ = np.array([2.0, 2.0, 2.0, 1.0, 1.0, 0.0, 0.0, 0.0])
    return s


CASES = {
    "tiny": neutral_tiny,
    "si8": model.si8,
    "h2o4": lambda: model.water(nmol=4, a=9.85, seed=3),
    "si8d": lambda: model.si8(species=model.si_species_d()),
}


def gpu_coupled(s, prec=0):
    import pwgpu
    plane, w = gpu_waveset(s, prec=prec)
    dplane = pwgpu.PlaneWave(s.rbas, s.epw_rho, s.epw_rho)
    tabs = s.potential_tables(opw.Plane(s.rbas, s.epw_rho, s.epw_rho).g2)
    pot = pwgpu.Potential(dplane, tabs, lmrx=[sp.lmrx for sp in s.species])
    for k, sp in enumerate(s.species):
        w.sphere_setup(k, sp.sphere_setup())
    w.set_potential(pot)
    return plane, w


@pytest.mark.parametrize("name,nsteps", [("tiny", 4), ("si8", 3), ("h2o4", 3), ("si8d", 2)])
def test_sphere_cp_steps_match_oracle(name, nsteps):
    s = model.paw_system(CASES[name]())
    say(f"case one-centre cp {name} x{nsteps}")
    o = coupled_waveset(s)
    plane, w = gpu_coupled(s)
    for it in range(nsteps):
        eo = o.etot(s.pos)
        eg = w.etot()
        for k in ("ekin", "epot", "e1c", "etot"):
            assert abs(eg[k] - eo[k]) <= 1e-10 * max(1.0, abs(eo[k])), (it, k, eg[k], eo[k])
        assert abs(w.getr8("POTB") - o.potb) <= 1e-10 * max(1.0, abs(o.potb)), it
        assert relerr(w.getr8a("QLM"), o.qlm) < 1e-10, it
        lmrx = [sp.lmrx for sp in s.species]
        vq = np.array([np.pad(o.pot["vqlm"][a, :lmrx[s.isp[a]]], (0, o.qlm.shape[1] - lmrx[s.isp[a]]))
                       for a in range(s.nat)])
        assert relerr(w.getr8a("VQLM"), vq) < 1e-10, it
        dh = w.getc8a_dh()
        for a in range(s.nat):
            n = o.dh[a].shape[0]
            assert relerr(dh[a, :n, :n], o.dh[a]) < 1e-10, (it, a)
        assert relerr(w.getr8a("FORCE"), o.force) < 1e-9, it
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        w.propagate()
        w.orthogonalize()
        assert w.geti4("ORTHOITER") == o.ortho_iter, it
        assert relerr(w.getc8a("RLAM0").real, o.rlam0) < 1e-9, it
        o.switch()
        w.switch()
    assert relerr(w.getc8a("PSI0"), o.psi0) < 1e-9
    say(f"  {name}: etot={eg['etot']:.12f} e1c={eg['e1c']:.10f} potb={w.getr8('POTB'):.3e}")


def test_sphere_mixed_form_step_gates():
    """The mixed form with the one-centre terms (FP64 on the device in both
    forms) keeps the per-step gates of the mixed form on Si8."""
    s = model.paw_system(model.si8())
    say("case one-centre mixed si8")
    o = coupled_waveset(s)
    plane, w = gpu_coupled(s, prec=1)
    w.etot()
    for it in range(2):
        eo = o.etot(s.pos)
        w.setc8a("PSI0", o.psi0)
        w.setc8a("PSIM", o.psim)
        w.set_matrix_c("RLAM0", np.asarray(o.rlam0, dtype=np.complex128))
        w.projections()
        e = w.etot()
        assert abs(e["etot"] - eo["etot"]) / s.nat < 1e-6, it
        assert float(np.max(np.abs(w.getr8a("FORCE") - o.force))) < 1e-4, it
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        o.switch()
