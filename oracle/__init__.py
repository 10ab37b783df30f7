"""FP64 CPU oracle for the CP-PAW Car-Parrinello wavefunction step.

TEST INFRASTRUCTURE ONLY.  This package is a plain numpy restatement of the
reference algorithm (cp-paw ``src/paw_fft.f90``, ``src/paw_waves1.f90``,
``src/paw_waves2.f90``, ``src/paw_library.f90``, ``src/paw_generalpurpose.f90``).
It is used by ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` as the checker.  The product path (the HIP library behind
``include/pwgpu.h`` and the ``pwgpu`` host package) never imports it.

Conventions (pinned from the reference):
  * arrays use numpy C order with the Fortran dimensions reversed, so the memory
    layout equals the reference's column-major layout:
    PSI(NGL,NDIM=1,NBH) -> psi[nbh, ngl];  PROJ(1,NBH,NPRO) -> proj[npro, nbh];
    FOFR(NR1*NR2*NR3) with ir1 fastest -> f[ir3, ir2, ir1].
  * only NDIM=1, NSPIN=1 and the Gamma point with super wave functions
    (TINV=.TRUE.) are restated; these are the cases the CP step configs use.

Parity status per function (each is pinned by tests under ``tests/``):
  planewave.*     pinned: brute-force enumeration, closed-form grid rule,
                  direct DFT, plane-wave delta, round trip.
  waves.ekin/density/vpsi/projections/addpro/overlap/1coverlap/addopsi/
  addoproj/denmat/dedproj/proforce/propagate/wavekinetic:
                  pinned against unpacked real-band restatements (independent
                  formulas without super wave functions), closed forms and
                  finite differences.
  waves.ortho_x:  pinned by the constraint invariants (overlap residual <1e-8,
                  occupation symmetry of Lambda) and by an independent
                  Newton solve of the same fixed point.
"""
