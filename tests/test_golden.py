"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the
pinned oracle): the oracle must keep reproducing them (CPU) and the HIP
library must reproduce them within the FP64 gate (GPU)."""
import os

import numpy as np
import pytest

from harness import gpu_waveset, oracle_waveset, relerr
from pwgpu import model

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    with np.load(os.path.join(HERE, "golden", name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_reproduces_tiny_fixture():
    g = load("tiny_cp3.npz")
    s = model.tiny()
    o = oracle_waveset(s)
    assert np.array_equal(o.plane.igvec, g["igvec"])
    for it in range(3):
        e = o.etot(s.pos)
        if it == 0:
            assert relerr(o.psi0, g["psi0_gs"]) < 1e-12
            assert relerr(o.rho, g["rho0"]) < 1e-12
            assert relerr(o.force, g["force0"]) < 1e-11
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        assert np.allclose([e["ekin"], e["epot"], e["e1c"], e["etot"], o.fictitious_ekin()],
                           g[f"etot{it}"], rtol=1e-12, atol=0)
        assert o.ortho_iter == int(g[f"ortho_iter{it}"])
        o.switch()
    assert relerr(o.psi0, g["psi0_final"]) < 1e-11


@pytest.mark.gpu
@pytest.mark.parametrize("fixture,factory,nsteps", [("tiny_cp3.npz", model.tiny, 3),
                                                     ("si8_cp2.npz", model.si8, 2)])
def test_gpu_reproduces_fixture(fixture, factory, nsteps):
    print(f"case golden {fixture}", flush=True)
    g = load(fixture)
    s = factory()
    plane, w = gpu_waveset(s)
    assert np.array_equal(plane.geti4a("IGVEC"), g["igvec"])
    assert np.array_equal(plane.geti4a("MINUSG") - 1, g["minusg"])
    for it in range(nsteps):
        e = w.etot()
        if it == 0:
            if "rho0" in g:
                assert relerr(w.getr8a("RHO"), g["rho0"]) < 1e-11
                assert relerr(w.getc8a("HPSI"), g["hpsi0"]) < 1e-11
            assert relerr(w.getr8a("FORCE"), g["force0"]) < 1e-10
            assert relerr(w.getr8a("EIG"), g["eig0"]) < 1e-11
        w.propagate()
        w.orthogonalize()
        got = np.array([e["ekin"], e["epot"], e["e1c"], e["etot"], w.getr8("EFICT")])
        assert np.allclose(got, g[f"etot{it}"], rtol=1e-10, atol=1e-12)
        assert w.geti4("ORTHOITER") == int(g[f"ortho_iter{it}"])
        w.switch()
    if "psi0_final" in g:
        assert relerr(w.getc8a("PSI0"), g["psi0_final"]) < 1e-9
