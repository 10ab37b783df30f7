"""GPU parity of POTENTIAL$VOFRHO (pw_vofrho, src/paw_potential.f90:324-896)
against the FP64 oracle (oracle/potential.py) on identical seeded inputs:
the model valence density of the cell and the model setup tables on the
density sphere, with the local functional (type 1) and PBE (type 10).

Tolerances (FP64 on both sides; the transforms and sums differ only in
rounding order): EHARTREE, EXC, RHOB and dE/dQLM within 1e-10 relative, the
potential within 1e-10 of its largest magnitude, forces within 1e-9 of the
largest force.
"""
import sys

import numpy as np
import pytest

from harness import relerr
from oracle import planewave as opw
from oracle import potential as opot
from pwgpu import model

pytestmark = pytest.mark.gpu

CASES = {
    "tiny": model.tiny,
    "si8": model.si8,
    "h2o4": lambda: model.water(nmol=4, a=9.85, seed=3),
    "si64": model.si64,
}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def inputs(s):
    pl = opw.Plane(s.rbas, s.epw_rho, s.epw_rho)
    rho = s.valence_density((pl.nr1, pl.nr2, pl.nr3))
    tabs = s.potential_tables(pl.g2)
    return pl, rho, tabs


def run_oracle(pl, s, rho, tabs, functional=1):
    return opot.vofrho(pl, rho, s.pos, s.isp, s.qlm(),
                       {"pscorg": tabs[0], "vbarg": tabs[1], "g0": tabs[2], "v0": tabs[3]}, functional=functional)


@pytest.mark.parametrize("functional", [1, 10])
@pytest.mark.parametrize("name", list(CASES))
def test_vofrho_matches_oracle(name, functional):
    import pwgpu
    s = CASES[name]()
    say(f"case vofrho {name} functional={functional}")
    pl, rho, tabs = inputs(s)
    o = run_oracle(pl, s, rho, tabs, functional)
    dplane = pwgpu.PlaneWave(s.rbas, s.epw_rho, s.epw_rho)
    assert dplane.ngl == pl.ngl and dplane.nr == (pl.nr1, pl.nr2, pl.nr3)
    pot = pwgpu.Potential(dplane, tabs, functional=functional)
    g = pot.vofrho(s.isp, s.pos, s.qlm(), rho)
    for k in ("ehartree", "exc", "rhob"):
        assert abs(g[k] - o[k]) <= 1e-10 * max(abs(o[k]), 1e-12), (k, g[k], o[k])
    assert relerr(g["v"], o["v"]) < 1e-10
    assert relerr(g["vqlm"], o["vqlm"]) < 1e-10
    assert relerr(g["force"], o["force"]) < 1e-9
    say(f"  {name}: ngl={pl.ngl} E_H={g['ehartree']:.12f} Exc={g['exc']:.12f} "
        f"dv={relerr(g['v'], o['v']):.1e} dF={relerr(g['force'], o['force']):.1e}")


@pytest.mark.parametrize("name", ["si8", "h2o4", "si64"])
def test_vofrho_higher_moments_matches_oracle(name):
    """Compensation charges of moments to l = 2 (LMRXX = 9), with the LMRX of
    each species (O and Si 9, H 1)."""
    import pwgpu
    s = CASES[name]()
    say(f"case vofrho LMRXX=9 {name}")
    pl, rho, tabs = inputs(s)
    lmrx = np.array([sp.lmrx for sp in s.species])
    qlm = s.qlm(9)
    for a, k in enumerate(s.isp):
        qlm[a, lmrx[k]:] = 0.0
    o = opot.vofrho(pl, rho, s.pos, s.isp, qlm, {"pscorg": tabs[0], "vbarg": tabs[1], "g0": tabs[2], "v0": tabs[3]},
                    lmrx=lmrx)
    dplane = pwgpu.PlaneWave(s.rbas, s.epw_rho, s.epw_rho)
    pot = pwgpu.Potential(dplane, tabs, lmrx=lmrx)
    g = pot.vofrho(s.isp, s.pos, qlm, rho)
    for k in ("ehartree", "exc", "rhob"):
        assert abs(g[k] - o[k]) <= 1e-10 * max(abs(o[k]), 1e-12), (k, g[k], o[k])
    assert relerr(g["v"], o["v"]) < 1e-10
    assert g["vqlm"].shape == (s.nat, 9)
    assert relerr(g["vqlm"], o["vqlm"]) < 1e-10
    assert relerr(g["force"], o["force"]) < 1e-9


def test_vofrho_in_place_and_repeatable():
    # vofr may alias rho; two calls give bit-identical results (ordered sums)
    import pwgpu
    s = CASES["h2o4"]()
    say("case vofrho in-place h2o4")
    pl, rho, tabs = inputs(s)
    dplane = pwgpu.PlaneWave(s.rbas, s.epw_rho, s.epw_rho)
    pot = pwgpu.Potential(dplane, tabs)
    a = pot.vofrho(s.isp, s.pos, s.qlm(), rho)
    buf = pwgpu.DeviceArray.from_host(np.ascontiguousarray(rho).reshape(-1))
    en, vq, fo = pot.vofrho_device(s.isp, s.pos, s.qlm(), buf, buf)
    v = buf.to_host().reshape(rho.shape)
    assert en["ehartree"] == a["ehartree"] and en["exc"] == a["exc"]
    assert np.array_equal(v, a["v"]) and np.array_equal(vq, a["vqlm"]) and np.array_equal(fo, a["force"])


@pytest.mark.parametrize("name,nsteps,functional", [("tiny", 4, 1), ("si8", 3, 1), ("h2o4", 3, 1),
                                                     ("si8", 3, 10), ("h2o4", 2, 10)])
def test_scf_cp_steps_match_oracle(name, nsteps, functional):
    # CP steps with the density -> potential step linked into ETOT
    # (pw_waves_set_potential) against the oracle with the same coupling
    import pwgpu
    from harness import gpu_waveset, oracle_waveset
    s = CASES[name]()
    say(f"case scf cp {name} steps={nsteps} functional={functional}")
    o = oracle_waveset(s)
    dpl = opw.Plane(s.rbas, s.epw_rho, s.epw_rho)
    tabs = s.potential_tables(dpl.g2)
    o.potential = (dpl, {"pscorg": tabs[0], "vbarg": tabs[1], "g0": tabs[2], "v0": tabs[3]}, s.isp, s.qlm(),
                   functional)
    plane, w = gpu_waveset(s)
    dplane = pwgpu.PlaneWave(s.rbas, s.epw_rho, s.epw_rho)
    pot = pwgpu.Potential(dplane, tabs, functional=functional)
    w.set_potential(pot, s.qlm())
    for it in range(nsteps):
        eo = o.etot(s.pos)
        eg = w.etot()
        assert abs(eg["etot"] - eo["etot"]) <= 1e-10 * abs(eo["etot"]), it
        assert abs(w.getr8("EHARTREE") - o.pot["ehartree"]) <= 1e-10 * abs(o.pot["ehartree"]), it
        assert abs(w.getr8("EXC") - o.pot["exc"]) <= 1e-10 * abs(o.pot["exc"]), it
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
    say(f"  {name}: etot={eg['etot']:.12f} E_H={w.getr8('EHARTREE'):.10f} Exc={w.getr8('EXC'):.10f}")


def test_pbe_with_one_centre_terms_is_refused():
    """The one-centre terms take the local functional only: an ETOT that would
    couple them to PBE fails with PW_ERR_UNSUPPORTED instead of mixing the two."""
    import pwgpu
    from pwgpu import _lib
    from harness import gpu_waveset
    s = model.paw_system(model.tiny())
    say("case pbe with one-centre terms refused")
    plane, w = gpu_waveset(s)
    dplane = pwgpu.PlaneWave(s.rbas, s.epw_rho, s.epw_rho)
    tabs = s.potential_tables(opw.Plane(s.rbas, s.epw_rho, s.epw_rho).g2)
    pot = pwgpu.Potential(dplane, tabs, lmrx=[sp.lmrx for sp in s.species], functional=10)
    for k, sp in enumerate(s.species):
        w.sphere_setup(k, sp.sphere_setup())
    w.set_potential(pot)
    with pytest.raises(_lib.PwError) as e:
        w.etot()
    assert "type 1" in str(e.value)
