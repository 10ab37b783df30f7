"""The mixed-precision form (FP32 FFT passes, split-BF16x3 MFMA products,
FP64 constraint solve, energies and forces) against the FP64 oracle.

Gates, per step of the oracle's FP64 trajectory, with the mixed form handed
the oracle's state (PSI0, PSIM, RLAM0): |dE| < 1e-6 Ha/atom and
|dF| < 1e-4 Ha/bohr of WAVES$ETOT, and the orthonormality residual
||psi^H S psi - 1|| of the wave functions that one mixed PROPAGATE +
ORTHOGONALIZE makes from that state, measured in the oracle's FP64 metric.

Free-running trajectories of the two forms separate over the steps: the
mixed form keeps the constraint to FP32-level accuracy (~1e-7), and the
CP dynamics carries such differences forward.  That drift is measured and
bounded loosely (it catches a broken step, not rounding)."""
import numpy as np
import pytest

from harness import SYSTEMS, gpu_waveset, oracle_waveset, relerr
from oracle import waves as owv

pytestmark = pytest.mark.gpu

DE_ATOM = 1.0e-6      # Ha/atom
DF = 1.0e-4           # Ha/bohr
ORTHO_RES = 1.0e-5    # max |psi^H S psi - 1| in FP64, mixed-form wave functions
DRIFT_DE_ATOM = 1.0e-4
DRIFT_DF = 1.0e-2


def say(msg):
    print(msg, flush=True)


def ortho_residual(o, psi, r=None):
    """max |<psi_i|S|psi_j> - delta_ij| with the oracle's FP64 overlaps at
    positions r (default R(0))."""
    pl, st = o.plane, o.setup
    proj = owv.projections(pl, st, o.pro, o.ylm, o.r0 if r is None else r, psi)
    s = owv.overlap(pl, o.nb, psi, psi) + owv.overlap_1c(st, o.nb, proj, proj)
    return float(np.max(np.abs(s - np.eye(o.nb))))


@pytest.mark.parametrize("name,nsteps,moving", [("tiny", 4, False), ("si8", 5, False), ("si8c", 3, False),
                                                 ("h2o4", 4, False), ("synthetic", 2, False), ("h2o32", 2, False),
                                                 ("si8d", 3, False), ("si8s", 2, False), ("si16", 2, False),
                                                 ("si64", 2, False),
                                                 ("si8", 3, True), ("si64", 2, True)])
def test_mixed_form_error_gates(name, nsteps, moving):
    s = SYSTEMS[name]()
    say(f"case mixed gates {name} x{nsteps}{' moving ions' if moving else ''}")
    o = oracle_waveset(s)
    o.r0 = s.pos
    vel = np.random.default_rng(12).normal(0.0, 0.03, s.pos.shape) if moving else np.zeros_like(s.pos)
    plane, w = gpu_waveset(s, prec=1)
    w.etot()                          # first call: Gram-Schmidt of its own start
    worst = [0.0, 0.0, 0.0]
    for it in range(nsteps):
        r0, rp = s.pos + it * vel, s.pos + (it + 1) * vel
        eo = o.etot(r0)
        w.setr8a("R0", r0)
        w.setr8a("RP", rp)
        w.setc8a("PSI0", o.psi0)
        w.setc8a("PSIM", o.psim)
        w.set_matrix_c("RLAM0", np.asarray(o.rlam0, dtype=np.complex128))
        w.projections()
        e = w.etot()
        de = abs(e["etot"] - eo["etot"]) / s.nat
        df = float(np.max(np.abs(w.getr8a("FORCE") - o.force)))
        o.propagate()
        o.orthogonalize(r0, rp)
        w.propagate()
        w.orthogonalize()
        res = ortho_residual(o, w.getc8a("PSIM"), rp)
        worst = [max(worst[0], de), max(worst[1], df), max(worst[2], res)]
        assert de < DE_ATOM, (it, de)
        assert df < DF, (it, df)
        assert res < ORTHO_RES, (it, res)
        o.switch()
    say(f"  worst |dE|/atom {worst[0]:.2e}  |dF| {worst[1]:.2e}  ortho residual {worst[2]:.2e}")


@pytest.mark.parametrize("name,nsteps", [("si8", 5), ("h2o4", 4)])
def test_mixed_trajectory_drift(name, nsteps):
    s = SYSTEMS[name]()
    say(f"case mixed drift {name} x{nsteps}")
    o = oracle_waveset(s)
    plane, w = gpu_waveset(s, prec=1)
    for it in range(nsteps):
        eo = o.etot(s.pos)
        e = w.etot()
        de = abs(e["etot"] - eo["etot"]) / s.nat
        df = float(np.max(np.abs(w.getr8a("FORCE") - o.force)))
        say(f"  step {it}: |dE|/atom {de:.2e}  |dF| {df:.2e}")
        assert de < DRIFT_DE_ATOM, (it, de)
        assert df < DRIFT_DF, (it, df)
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        w.propagate()
        w.orthogonalize()
        o.switch()
        w.switch()


def test_mixed_projections_accuracy():
    s = SYSTEMS["si8"]()
    say("case mixed projections si8")
    o = oracle_waveset(s)
    plane, w = gpu_waveset(s, prec=1)
    w.gramschmidt()
    w.projections()
    psi = w.getc8a("PSI0")
    want = owv.projections(o.plane, o.setup, o.pro, o.ylm, s.pos, psi)
    err = relerr(w.getc8a("PROJ"), want)
    say(f"  projections rel err {err:.2e}")
    assert err < 1e-5
