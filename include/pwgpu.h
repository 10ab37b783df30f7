/*
 * pwgpu.h - C ABI of the MI355X-native Car-Parrinello wavefunction step.
 *
 * The entry points replace, one for one, the routines of CP-PAW's PLANEWAVE
 * and WAVES objects that form the CP wavefunction step.  Each declaration
 * names the reference routine (file:line, relative to the cp-paw checkout) it
 * stands in for.  A Fortran host binds them through ISO_C_BINDING (see
 * cp-paw_amd/fortran/pwgpu_c.f90); a Python host binds them through ctypes
 * (cp-paw_amd/pwgpu/_lib.py).
 *
 * Conventions
 *  - Every function returns 0 on success and a non-zero PW_ERR_* code on
 *    failure; pw_last_error() returns the message of the last failure of the
 *    calling thread.  The Fortran shim turns a non-zero status into
 *    ERROR$MSG + ERROR$STOP, the reference's error behaviour
 *    (src/paw_error.f90:221-281).
 *  - Arrays are Fortran column-major with the reference's dimensions:
 *    PSI(NGL,NBH) complex, PROJ(NBH,NPRO) complex (NDIM=1 dropped),
 *    FOFR(NR1*NR2*NR3) with ir1 fastest, square matrices (NB,NB).
 *    Complex values are interleaved (re, im) pairs of double (complex128)
 *    or float (complex64).
 *  - Pointers named *_dev are device pointers (from pw_malloc or any HIP
 *    allocation on the current device); all other array pointers are host
 *    pointers that the library copies from/to.
 *  - A plan or waveset is not thread safe; callers serialise access.  Work is
 *    queued on the object's stream; functions that return host data
 *    synchronise that stream.
 *  - Only the Gamma point with super wave functions (TINV=.TRUE.), NDIM=1 and
 *    NSPIN=1 are supported, the case of the CP step configurations; other
 *    requests fail with PW_ERR_UNSUPPORTED.
 */
#ifndef PWGPU_H
#define PWGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define PW_OK 0
#define PW_ERR_ARG 1          /* inconsistent size or bad argument          */
#define PW_ERR_UNSUPPORTED 2  /* case outside the supported set             */
#define PW_ERR_HIP 3          /* HIP runtime error                          */
#define PW_ERR_STATE 4        /* call order violated (e.g. no HPSI yet)     */
#define PW_ERR_NOCONV 5       /* orthogonalisation loop did not converge    */

/* Precision forms of a waveset. */
#define PW_PREC_FP64 0        /* every kernel in FP64                       */
#define PW_PREC_MIXED 1       /* FP32 FFT/V.psi, split-BF16x3 MFMA products,
                                 FP64 constraint solve and reductions       */

typedef struct pw_plan pw_plan;       /* PLANEWAVE instance (PWPARALLEL_TYPE) */
typedef struct pw_waveset pw_waveset; /* WAVES state (MAP/GSET/WVSET_TYPE)    */

/* ---- library / runtime ------------------------------------------------ */
int pw_version(void);
/* Copies the message of the last failure into buf (NUL terminated). */
int pw_last_error(char *buf, int len);
int pw_malloc(void **ptr_dev, long long nbytes);
int pw_free(void *ptr_dev);
int pw_memcpy_htod(void *dst_dev, const void *src, long long nbytes);
int pw_memcpy_dtoh(void *dst, const void *src_dev, long long nbytes);
int pw_device_synchronize(void);

/* ---- PLANEWAVE object ----------------------------------------------------
 * pw_plan_create replaces PLANEWAVE$DIVIDERGRIDONTASKS (src/paw_fft.f90:823)
 * followed by PLANEWAVE$INITIALIZE (src/paw_fft.f90:872) for one task:
 * the real-space grid follows BOXSPH+LIB$FFTADJUSTGRD for epw_rho, the
 * G sphere 0.5|G|^2<=epw_psi follows PLANEWAVE_GETIG's order and MINUSG.
 * rbas: RBAS(3,3) column-major (lattice vectors as columns), Hartree units. */
int pw_plan_create(const double rbas[9], const double kvec[3], int tinv,
                   double epw_psi, double epw_rho, pw_plan **plan);
int pw_plan_destroy(pw_plan *plan);
/* PLANEWAVE$GETI4 (src/paw_fft.f90:286): "NGL","NRL","NR1","NR2","NR3",
 * "NGAMMA" (1-based), "NSTRIPE", "NZPLANE". */
int pw_plan_geti4(const pw_plan *plan, const char *id, int *val);
/* PLANEWAVE$GETR8 (src/paw_fft.f90:334): "GWEIGHT","RWEIGHT". */
int pw_plan_getr8(const pw_plan *plan, const char *id, double *val);
/* PLANEWAVE$GETR8A (src/paw_fft.f90:377): "GVEC" (3*NGL), "G2" (NGL). */
int pw_plan_getr8a(const pw_plan *plan, const char *id, int len, double *val);
/* Integer maps: "IGVEC" (3*NGL), "MINUSG" (NGL, 1-based as the reference). */
int pw_plan_geti4a(const pw_plan *plan, const char *id, int len, int *val);

/* PLANEWAVE$FFT(ID,NFFT,NGL,FOFG,NRL,FOFR) (src/paw_fft.f90:1948).
 * id "GTOR": FOFR(r)=sum_G FOFG(G) exp(+iGr); "RTOG": FOFG(G)=1/N sum_r
 * FOFR(r) exp(-iGr).  fofg_dev: NGL x NFFT complex128 (FP64 storage of the
 * coefficients in both precisions); fofr_dev: NRL x NFFT complex128 for
 * prec=PW_PREC_FP64, complex64 for prec=PW_PREC_MIXED.  Asynchronous on the
 * plan's stream. */
int pw_fft(pw_plan *plan, const char *id, int nfft, int ngl, void *fofg_dev,
           int nrl, void *fofr_dev, int prec);
/* PLANEWAVE$FFT as pw_fft, for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The work buffers come from
 * a static device array of 64 MiB per GPU instead of the plan's lazily
 * allocated ones, so the call makes no hipMalloc/hipFree; one transform must
 * fit in it (es*(NSTRIPE*NR1 + NZPLANE*NR2*NR1) + 256 bytes, es = 16 for
 * PW_PREC_FP64, 8 for PW_PREC_MIXED), else PW_ERR_ARG; larger NFFT run in
 * batches that fit.  The plan's tables must already be on the device (after
 * any pw_fft or waveset call on the plan), else PW_ERR_STATE.  All calls on
 * one GPU share the array: calls that overlap on one GPU race.  Results are
 * bit-identical to pw_fft's. */
int pw_fft_noalloc(pw_plan *plan, const char *id, int nfft, int ngl, void *fofg_dev,
                   int nrl, void *fofr_dev, int prec);
/* Stream used by the plan (hipStream_t), or 0 for the default stream. */
int pw_plan_set_stream(pw_plan *plan, void *stream);

/* ---- WAVES object --------------------------------------------------------
 * pw_waves_create replaces the MAP/GSET/WVSET parts of WAVES$INITIALIZE
 * (src/paw_waves1.f90:1187-1420) for one k-point and spin.
 * isp(NAT): 1-based species of each atom; lnx(NSP); lox(LNXX,NSP).       */
int pw_waves_create(pw_plan *plan, int nb, int nat, int nsp, const int *isp,
                    const int *lnx, int lnxx, const int *lox, int prec,
                    pw_waveset **ws);
int pw_waves_destroy(pw_waveset *ws);
/* WAVES$SETR8 (src/paw_waves1.f90:322): "TIMESTEP" (DELT), "EMASS",
 * "EMASSCG2", "FRICTION" (ANNEE); and "RESIDUALCHECK" (no counterpart in the
 * reference): nonzero makes each ORTHOGONALIZE evaluate the residual that
 * pw_waves_getr8("RESIDUAL") returns. */
int pw_waves_setr8(pw_waveset *ws, const char *id, double val);
/* Real arrays: "PRO" (NGL*NBAREPRO, GSET%PRO, already / cell volume),
 * "OCC" (NB), "DO" (LNXX*LNXX*NSP), "VOFR" (NRL, local potential),
 * "R0" / "RP" / "RM" (3*NAT atomic positions R(0), R(+), R(-)),
 * "ORTHOBOUNDS" (2: spectral bounds of sym(CHIPSI) for the next
 * orthogonalisation; checked on the device and recomputed if rejected --
 * no counterpart in the reference, whose DSYEV needs none). */
int pw_waves_setr8a(pw_waveset *ws, const char *id, int len, const double *val);
/* Complex arrays (interleaved re,im): "DH" (LMNXX*LMNXX*NAT),
 * "PSI0" / "PSIM" (NGL*NBH), "RLAM0" / "RLAMM" (NB*NB). */
int pw_waves_setc8a(pw_waveset *ws, const char *id, int len, const double *val);
/* Complex results: "PSI0","PSIM","HPSI","OPSI","PROJ","OPROJ" (NBH*NPRO),
 * "RLAM0" / "RLAMM" (NB*NB; RLAMM fails with PW_ERR_STATE while NLAMBDA < 2).
 * PSI0, PSIM, RLAM0 and RLAMM are what WAVES_WRITEPSI / WAVES_WRITELAMBDA
 * (src/paw_waves2.f90:3168, 3920) put into the restart file; the setters take
 * them back as WAVES_READPSI / WAVES_READLAMBDA (:5147, 4041) do. */
int pw_waves_getc8a(pw_waveset *ws, const char *id, int len, double *val);
/* Real results: "RHO" (NRL), "FORCE" (3*NAT), "EIG" (NB, <psi|H|psi>),
 * "MAT","OMAT","OOMAT" (NB*NB, last orthogonalisation), "DENMAT"
 * (LMNXX*LMNXX*NAT). */
int pw_waves_getr8a(pw_waveset *ws, const char *id, int len, double *val);
/* Scalars: "EKIN","EPOT","E1C","ETOT","WAVEEKIN1","WAVEEKIN2","EFICT".
 * "RESIDUAL" (with RESIDUALCHECK set): of the last ORTHOGONALIZE, the largest
 * norm over super bands of the constrained force HPSI - O PSI0 LAMBDA =
 * HPSI - (1+ANNEE) m(G)/DELT^2 (PSI(+) - PSIBAR), which vanishes at the fixed
 * point of the damped dynamics (a super band's norm is that of its two real
 * bands' residuals together).
 * "ORTHOERR": the orthonormality test of WAVES$ORTHOGONALIZE's TTEST branch
 * (src/paw_waves2.f90:425-464) on PSI0 at R(0), max_ij |<psi_i|S|psi_j> -
 * delta_ij|, always evaluated with the FP64 kernels (for a mixed-form set it
 * measures the mixed wave functions in the FP64 metric); overwrites OPROJ. */
int pw_waves_getr8(pw_waveset *ws, const char *id, double *val);
/* "ORTHOITER" (iterations of the last WAVES_ORTHO_X), "NPRO", "NBH",
 * "LMNXX", "ORTHOLEVELS" (squared-Smith levels of the last solve), "ORTHONS"
 * (Newton-Schulz steps of the last solve), "ORTHOLANCZOS" (Lanczos runs for
 * the spectral bounds in the last solve: 0 when the previous bounds were
 * accepted), "ORTHOEIGEN" (solves so far through the eigenvectors of
 * sym(CHIPSI), taken when it is not positive definite), "ORTHOGRAPHMISSES"
 * (solves issued without a captured hipGraph: captures are made only while
 * one plan and one waveset exist in the process), "NLAMBDA" (multiplier history held: 0 none, 1 RLAM0, 2 RLAM0
 * and RLAMM; counted as in WAVES_WRITELAMBDA, src/paw_waves2.f90:3948),
 * and the launch shapes of pw_waves_autotune: "SPLITWAVES" (waves of a
 * G-split row-product launch), "DENSBLOCKS" (workgroups of the density pass),
 * "COLWAVES" (waves of a column-product workgroup, mixed form). */
int pw_waves_geti4(pw_waveset *ws, const char *id, int *val);
/* Device pointer of a resident array ("PSI0","PSIM","HPSI","PROJ","VOFR",
 * "RHO") for zero-copy access by the caller. */
int pw_waves_devptr(pw_waveset *ws, const char *id, void **ptr_dev);

/* WAVES$GRAMMSCHMIDT (src/paw_waves1.f90:2638): orthonormalise PSI0 at R(0)
 * and PSIM at R(-) (WAVES_GRAMSCHMIDT, src/paw_waves2.f90:2025). */
int pw_waves_gramschmidt(pw_waveset *ws);
/* WAVES$PROJECTIONS('PSI0') (src/paw_waves1.f90:2681): PROJ=<p(R0)|PSI0>. */
int pw_waves_projections(pw_waveset *ws);
/* WAVES$ETOT hot path (src/paw_waves1.f90:1610-2233): WAVES_EKIN, WAVES$DENMAT,
 * WAVES_DENSITY, the energy of the fixed potential and one-centre terms,
 * WAVES$FORCE, WAVES$HPSI (VPSI+HPROJ+ADDPRO), <psi|H|psi>, WAVEKINETIC. */
int pw_waves_etot(pw_waveset *ws);
/* WAVES$PROPAGATE (src/paw_waves1.f90:6479). */
int pw_waves_propagate(pw_waveset *ws);
/* WAVES$ORTHOGONALIZE (src/paw_waves2.f90:11), TSAFEORTHO=.TRUE. */
int pw_waves_orthogonalize(pw_waveset *ws);
/* WAVES$SWITCH (src/paw_waves2.f90:2180), NLAMBDAX=2. */
int pw_waves_switch(pw_waveset *ws);
/* One full CP step: etot, propagate, orthogonalize, switch (TIMESTEP order,
 * src/paw_driver.f90:333-561 restricted to the wave functions).
 * pw_waves_etot (without a density object) and pw_waves_orthogonalize return
 * once their work is queued: the energies, forces and EFICT they produce are
 * completed by the next call that reads or changes the waveset (any getter or
 * setter), which waits for them; an error of their kernels is reported
 * there. */
int pw_waves_step(pw_waveset *ws);
/* Set the stream (hipStream_t) all waveset work is queued on. */
int pw_waves_set_stream(pw_waveset *ws, void *stream);
/* Per-kernel GPU timing: enable=1 records HIP events around each launch on
 * the waveset's stream; pw_waves_kernel_time returns the summed duration (ms)
 * and launch count of the named kernel since the last reset. */
int pw_waves_profile(pw_waveset *ws, int enable);
int pw_waves_kernel_time(pw_waveset *ws, const char *name, double *ms, int *count);

/* WAVES_VPSI (src/paw_waves1.f90:5099) on caller arrays:
 * hpsi = FFT^-1[V FFT psi] + G^2/2 psi for NBH super bands (device). */
int pw_waves_vpsi(pw_waveset *ws, const void *psi_dev, void *hpsi_dev);
/* WAVES_DENSITY (src/paw_waves1.f90:5377) of PSI0 into rho_dev (NRL). */
int pw_waves_density(pw_waveset *ws, const void *psi_dev, double *rho_dev);

/* WAVES_ORTHO_X (src/paw_waves2.f90:1844) on host matrices (NB,NB):
 * lambda in/out; returns the iteration count in *iter.  Runs on the GPU in
 * FP64; any symmetric sym(CHIPSI) with e_i + e_j != 0 (a positive definite
 * one by a Cayley series of GEMMs, others through its eigenvectors). */
int pw_ortho_x(int nb, const double *occ, const double *chichi,
               const double *psipsi, const double *chipsi, double *lambda,
               int *iter);

/* ---- density -> potential (POTENTIAL$VOFRHO) --------------------------------
 * Replaces the core of POTENTIAL$VOFRHO (src/paw_potential.f90:324-896:
 * SUPFFT of the density, POTENTIAL_ADDCORE :1038, POTENTIAL_HARTREE :1069,
 * POTENTIAL_XC :1339, POTENTIAL_FPSCORE :897) for one spin, the local-density
 * functional of type 1 (local exchange + Perdew-Zunger, src/paw_dft.f90:
 * 124-130), compensation charges of the multipole moments QLM(LMRXX,NAT)
 * with l <= 4 (LMRXX <= 25; POTENTIAL_HARTREE :1145-1173, POTENTIAL_YLMOFG
 * :1548), no pair potential of the compensation charges and no isolation;
 * functional type 10 (PBE) through pw_potential_functional.
 * The plan is the density plane-wave object (PLANEWAVE$SELECT('DENSITY')):
 * its sphere is that of the density cutoff.  FP64.
 * pw_potential_tables sets the setup tables the reference keeps as module
 * data, on the sphere and resident on the device until replaced:
 *   tabs[4][nsp][NGL] = PSCORG, VBARG, G0, V0  (each G fastest),
 *   lmrx(NSP) = LMRX of each species (SETUP$GETI4('LMRX')), NULL for all 1.
 * pw_vofrho: rho_dev (NRL, x fastest) is the pseudo valence density;
 * vofr_dev (NRL) receives the total local potential dE/drho (may alias
 * rho_dev).  isp(NAT) 1-based species, pos R(3,NAT), qlm QLM(LMRXX,NAT).
 * Host outputs: energies = {EHARTREE, EXC, RHOB}, vqlm VQLM(LMRXX,NAT) =
 * dE/dQLM (zero beyond a species' LMRX), force(3,NAT) = the Hartree, V-bar
 * and pseudo-core forces (each may be NULL except energies). */
int pw_potential_tables(pw_plan *dplan, int nsp, const int *lmrx, const double *tabs);
/* DFT$SETI4('TYPE',type) (src/paw_dft.f90:511) for the density object:
 * 1 = local exchange + Perdew-Zunger (the default), 10 = PBE (exchange type 3
 * and PBE correlation over Perdew-Wang 1992, src/paw_dft.f90:248-262), whose
 * gradient of rho + core is formed from its Fourier coefficients and whose
 * divergence term joins the potential in G space (POTENTIAL_GRADIENT,
 * src/paw_potential.f90:977-1034).  Other types: PW_ERR_UNSUPPORTED.  The
 * one-centre terms (pw_waves_sphere_setup) take type 1 only. */
int pw_potential_functional(pw_plan *dplan, int type);
int pw_vofrho(pw_plan *dplan, int nat, int lmrxx, const int *isp, const double *pos,
              const double *qlm, const double *rho_dev, double *vofr_dev,
              double *energies, double *vqlm, double *force);
/* Links the density -> potential step into WAVES$ETOT, as the reference's
 * energy evaluation calls POTENTIAL$VOFRHO between WAVES_DENSITY and
 * WAVES_VPSI: each ETOT then forms V(r) from the density of PSI0 on the
 * device (positions R0, QLM(LMRXX,NAT) from WAVES$SETR8A('QLM'), or from
 * the one-centre terms when pw_waves_sphere_setup has given every species
 * its setup data), uses it for HPSI,
 * takes EPOT = EHARTREE + EXC and adds the potential forces to FORCE.
 * dplan must have the waveset's cell and r-space grid and its setup tables;
 * NULL detaches (VOFR is then the caller's again).  Calls on one density
 * object must not overlap. */
int pw_waves_set_potential(pw_waveset *ws, pw_plan *dplan);

/* ---- one-centre terms (AUGMENTATION$SPHERE) ---------------------------------
 * The setup data of species isp (1-based) that AUGMENTATION$SPHERE and
 * AUGMENTATION$MOMENTS read from the SETUP object (src/paw_augmentation.f90:
 * 265-305, 171-193): the shifted logarithmic radial grid r_i = r1 (exp(dex
 * (i-1)) - 1), i = 1..nr (SHLOGRADIAL, src/paw_radial.f90:1626-1655), the
 * atomic number aez and nuclear radius rnuc (RADIAL$NUCPOT :1464), the width
 * rcsm of the compensation Gaussians, LMRX, AEPHI(NR,LNX), PSPHI(NR,LNX),
 * AECORE(NR), PSCORE(NR), VADD(NR) and DEKIN(LNX,LNX); the one-centre
 * overlap is the waveset's DO.  Limits: NR <= 512, LNX <= 8, LMNX <= 32,
 * LMRX <= 25.  Once every species has its setup data and a density object is
 * linked (pw_waves_set_potential, tables with the same LMRX), each
 * WAVES$ETOT runs on the device, as src/paw_waves1.f90:1843-1915 does:
 * WAVES$MOMENTS (QLM from the density matrix), POTENTIAL$VOFRHO with these
 * QLM, WAVES$SPHERE (DH, the one-centre energy E1C, POTB) and
 * WAVES$ADDCONSTANTPOT (V += POTB, DH += POTB DO); DH is then no input.
 * One spin (NDIMD=1), local functional of type 1, no soft core, no external
 * one-centre potential.  Results: WAVES$GETR8 "E1C", "POTB"; GETR8A "QLM",
 * "VQLM" (LMRXX*NAT); GETC8A "DH" (LMNXX*LMNXX*NAT). */
int pw_waves_sphere_setup(pw_waveset *ws, int isp, int nr, double r1, double dex,
                          double aez, double rnuc, double rcsm, int lmrx,
                          const double *aephi, const double *psphi,
                          const double *aecore, const double *pscore,
                          const double *vadd, const double *dekin);

/* ---- launch-shape autotuner and device footprint ---------------------------
 * No counterpart in the reference (FFTW plans with FFTW_MEASURE,
 * src/paw_library.f90:2688-2707, are its nearest analogue).
 * pw_waves_autotune times the G-split width of the projection GEMMs and the
 * band grouping of the density pass on the waveset's stream (PSI0 and R0
 * must be set) and keeps the fastest; results are cached per system shape
 * for the process.  pw_waves_footprint estimates the device bytes a waveset
 * of this shape holds during a CP step (host-only, no GPU use). */
int pw_waves_autotune(pw_waveset *ws);
int pw_waves_footprint(pw_plan *plan, int nb, int nat, int nsp, const int *isp,
                       const int *lnx, int lnxx, const int *lox, int prec,
                       long long *bytes);

/* ---- request loop (serving) --------------------------------------------------
 * No counterpart in the reference, which relaxes one system per run with
 * damped CP dynamics and the autopilot's energy window
 * (src/paw_assist.f90:297-400).  The server takes independent relaxation
 * requests as they arrive, admits them first-fit while their device
 * footprint fits mem_budget and one of max_active slots is free, and runs
 * each admitted request's CP loop on its own stream, so requests overlap on
 * the device.  A request converges when |E(step) - E(step-1)| < etol on
 * nconv consecutive steps (at most max_steps steps).  All arrays are copied
 * at submit; layouts as for pw_waves_setr8a/setc8a on the plane-wave set of
 * the cell (rbas, epw_psi, epw_rho). */
typedef struct pw_server pw_server;
typedef struct pw_request {
  double rbas[9];          /* RBAS(3,3), lattice vectors as columns */
  double epw_psi, epw_rho;
  int nb, nat, nsp, lnxx, nbarepro, prec;
  const int *isp;          /* NAT, 1-based species */
  const int *lnx;          /* NSP */
  const int *lox;          /* LNXX*NSP */
  const double *pro;       /* NGL*NBAREPRO */
  const double *occ;       /* NB */
  const double *dov;       /* DO(LNXX,LNXX,NSP) */
  const double *dh;        /* DH(LMNXX,LMNXX,NAT), complex interleaved */
  const double *vofr;      /* NRL */
  const double *r;         /* R(3,NAT) */
  const double *psi;       /* PSI(NGL,NBH), complex interleaved; PSI0 = PSIM */
  double dt, emass, emasscg2, friction;
  double etol;
  int nconv, max_steps;
  int autotune;            /* 1: run pw_waves_autotune on admission */
  double rtol;             /* > 0: converged when RESIDUAL < rtol (Ha) on
                              nconv consecutive steps, instead of |dE| < etol */
} pw_request;
typedef struct pw_result {
  int status, steps, converged;
  double etot, ekin, epot, e1c;
  double t_submit, t_start, t_end;   /* seconds since pw_server_create */
  long long footprint;               /* bytes reserved while running */
  double residual;                   /* RESIDUAL of the last step (rtol > 0) */
} pw_result;
int pw_server_create(long long mem_budget, int max_active, pw_server **srv);
/* Waits for every submitted request, then stops the workers. */
int pw_server_destroy(pw_server *srv);
int pw_server_submit(pw_server *srv, const pw_request *req, long long *id);
/* Blocks until request id has finished; its record is then released. */
int pw_server_wait(pw_server *srv, long long id, pw_result *res);
/* peak_bytes: device memory reserved at the peak (running requests and
 * pooled sets). */
int pw_server_stats(pw_server *srv, long long *peak_bytes, int *peak_active,
                    int *pending);
/* The pool of finished plans and wavesets (kept per system shape, reset and
 * handed to the next request of that shape): requests served from the pool
 * (reused), requests that built a new set (created), and the device memory
 * the idle pooled sets reserve now. */
int pw_server_pool_stats(pw_server *srv, long long *reused, long long *created,
                         long long *idle_bytes);

#ifdef __cplusplus
}
#endif
#endif /* PWGPU_H */
