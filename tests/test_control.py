"""The reader of the reference's control/structure decks (pwgpu.control) on
the reference's own test decks (tests/fulltests/si2/si2.cntl and si2.strc,
copied as fixtures under tests/golden/decks/), on the defaults of its reader
routines and on its error messages; the deck's cell through the plane-wave
set-up (against the oracle) and, on the GPU, through CP steps against the
oracle."""
import math
import os

import numpy as np
import pytest

from pwgpu import control, model

DECKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "decks")


def si2():
    return control.read(os.path.join(DECKS, "si2.cntl"), os.path.join(DECKS, "si2.strc"))


def test_si2_deck_values():
    d = si2()
    # !FOURIER EPWPSI=20 CDUAL=2 (Ry): EPWRHO = CDUAL EPWPSI (src/paw_ioroutines.f90:1270-1295)
    assert d.epw_psi == 10.0 and d.epw_rho == 20.0
    # !GENERIC DT=10.0 NSTEP=180; !PSIDYN FRIC=0.005, MPSI default 10 DT^2
    assert d.dt == 10.0 and d.nstep == 180 and d.friction == 0.005
    assert d.emass == 1000.0
    assert abs(d.emasscg2 - 0.5 * (10.0 / (2 * math.pi)) ** 2 * 100.0 / 1000.0) < 1e-15
    assert d.safeortho is True
    # LUNIT=10.26, fcc lattice vectors as the columns of RBAS
    a = 10.26
    want = a * np.array([[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]).T
    assert np.allclose(d.rbas, want, atol=1e-14)
    assert abs(abs(np.linalg.det(d.rbas)) - a ** 3 / 4) < 1e-9
    names = [n for n, _, _ in d.atoms]
    assert names == ["Si_1", "Si_2"]
    assert np.allclose(d.positions, [[0, 0, 0], [0.25 * a] * 3], atol=1e-14)
    assert d.species["Si"]["zv"] == 4.0 and d.species["Si"]["npro"] == [2, 2, 1]
    # 8 electrons, EMPTY=5: NOCC = 4, NB = 9 -> 10 (even) (:4730-4760)
    assert d.nocc == 4 and d.nb == 10
    assert np.array_equal(d.occupations(), [2, 2, 2, 2, 0, 0, 0, 0, 0, 0])
    assert d.kdiv == [1, 1, 1]


MIN_CNTL = "!CONTROL !GENERIC !END !END !EOB"
MIN_STRC = """!STRUCTURE
 !LATTICE T= 1 0 0  0 1 0  0 0 1 !END
 !SPECIES NAME='O_' ZV=6. !END
 !ATOM NAME='O_1' R= 0.1 0.2 0.3 !END
 !OCCUPATIONS !END
!END
!EOB"""


def test_defaults():
    d = control.CPInput(MIN_CNTL, MIN_STRC)
    assert d.epw_psi == 15.0 and d.epw_rho == 60.0        # 30 Ry, CDUAL 4
    assert d.dt == 5.0 and d.nstep == 100 and d.friction == 0.0
    assert d.emass == 250.0                               # 10 DT^2
    assert d.lunit == 1.0 and np.allclose(d.rbas, np.eye(3))
    assert d.atoms[0][1] == "O_"                          # SP defaults to NAME(1:2) (:4508-4512)
    # 6 electrons, EMPTY default 1: NOCC 3, NB 4
    assert d.nocc == 3 and d.nb == 4


def test_lunit_in_angstrom():
    s = MIN_STRC.replace("!LATTICE", "!GENERIC LUNIT[AA]=2.0 !END\n !LATTICE")
    d = control.CPInput(MIN_CNTL, s)
    # the reference's Angstrom, from c, hbar, m_e, alpha (src/paw_constants.f90:83-101)
    assert abs(d.lunit - 2.0 * 1.8897259925730) < 1e-10


@pytest.mark.parametrize("cntl,strc,msg", [
    (MIN_CNTL, MIN_STRC.replace(" !LATTICE T= 1 0 0  0 1 0  0 0 1 !END\n", ""), "LATTICE IS MANDATORY"),
    (MIN_CNTL, MIN_STRC.replace("!LATTICE", "!GENERIC LUNIT=1 LUNIT[AA]=1 !END\n !LATTICE"), "MUTUALLY EXCLUSIVE"),
    ("!CONTROL !PSIDYN DIAG=T !END !END", MIN_STRC, "DIAG"),
    (MIN_CNTL, MIN_STRC.replace("!END\n!EOB", "!EOB"), "not closed"),
    (MIN_CNTL, MIN_STRC.replace("NAME='O_1'", "NAME='O_1' SP='C'"), "not defined"),
])
def test_errors(cntl, strc, msg):
    with pytest.raises(ValueError, match=msg):
        control.CPInput(cntl, strc)


def test_si2_plane_matches_oracle():
    import pwgpu
    from oracle import planewave as opw
    d = si2()
    p = pwgpu.PlaneWave(d.rbas, d.epw_psi, d.epw_rho)
    o = opw.Plane(d.rbas, d.epw_psi, d.epw_rho)
    assert p.ngl == o.ngl and p.nr == (o.nr1, o.nr2, o.nr3)
    assert np.array_equal(p.geti4a("IGVEC"), o.igvec)


@pytest.mark.gpu
def test_si2_deck_cp_steps_match_oracle():
    # the deck's cell, bands, cutoffs and dynamics, with a model Si setup that
    # has the deck's projector set (NPRO= 2 2 1: 2 s, 2 p and 1 d projector)
    import pwgpu
    from harness import relerr
    from oracle import planewave as opw
    from oracle import waves as owv
    d = si2()
    assert d.species["Si"]["npro"] == [2, 2, 1]
    s = d.to_system({"Si": model.species_from_npro(d.species["Si"]["npro"])})
    assert s.species[0].lox == [0, 0, 1, 1, 2]
    print(f"case si2 deck: nb={s.nb} epw={s.epw_psi}/{s.epw_rho} dt={s.dt}", flush=True)
    pl = opw.Plane(s.rbas, s.epw_psi, s.epw_rho)
    o = owv.WaveSet(pl, owv.Setup(s.isp, s.lox_list, s.do_list()), s.pro_table(pl.g2), s.nb, s.occ,
                    s.potential((pl.nr1, pl.nr2, pl.nr3)), s.dh_matrices(), s.dt, s.emass, s.emasscg2, s.annee)
    o.psi0 = s.initial_psi(pl.g2)
    o.psim = o.psi0.copy()
    plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
    w = pwgpu.Waves.from_system(plane, s)
    for it in range(3):
        eo = o.etot(s.pos)
        eg = w.etot()
        assert abs(eg["etot"] - eo["etot"]) <= 1e-10 * abs(eo["etot"]), it
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        w.propagate()
        w.orthogonalize()
        assert relerr(w.getc8a("RLAM0").real, o.rlam0) < 1e-9, it
        o.switch()
        w.switch()
