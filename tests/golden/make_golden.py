"""Writes the golden fixtures of the CP step from the FP64 oracle.

The reference's own tests hold no vectors for this path (only the end-to-end
si2 energy of tests/fulltests/si2), so these fixtures record the oracle's
outputs on seeded inputs after the oracle has been pinned by the checks in
tests/test_oracle_*.py.  Inputs are regenerated from pwgpu.model by seed.

python tests/golden/make_golden.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "cp-paw_amd"), os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

from harness import oracle_waveset  # noqa: E402
from pwgpu import model  # noqa: E402


def run(system, nsteps):
    o = oracle_waveset(system)
    r0 = system.pos
    out = {"igvec": o.plane.igvec, "minusg": o.plane.minusg, "g2": o.plane.g2}
    for it in range(nsteps):
        e = o.etot(r0)
        if it == 0:
            out.update(psi0_gs=o.psi0, proj0=o.proj, rho0=o.rho, hpsi0=o.hpsi, eig0=o.eig,
                       force0=o.force, denmat0=np.array(o.denmat))
        o.propagate()
        o.orthogonalize(r0, r0)
        out[f"etot{it}"] = np.array([e["ekin"], e["epot"], e["e1c"], e["etot"], o.fictitious_ekin()])
        out[f"ortho_iter{it}"] = np.array(o.ortho_iter)
        out[f"lambda{it}"] = o.rlam0
        o.switch()
    out["psi0_final"] = o.psi0
    out["force_last"] = o.force
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "tiny_cp3.npz"), **run(model.tiny(), 3))
    si8 = run(model.si8(), 2)
    keep = {k: v for k, v in si8.items() if k.startswith(("etot", "ortho_iter", "force", "eig0"))}
    keep["igvec"] = si8["igvec"]
    keep["minusg"] = si8["minusg"]
    np.savez_compressed(os.path.join(HERE, "si8_cp2.npz"), **keep)
    print("written")
