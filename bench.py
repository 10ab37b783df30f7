"""Benchmark of the CP wavefunction step on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--system si64]
                [--prec fp64|mixed|both] [--no-cpu] [--profile]

A step is one Car-Parrinello time step of the wave functions (WAVES$ETOT,
WAVES$PROPAGATE, WAVES$ORTHOGONALIZE, WAVES$SWITCH) of the Si64 diamond
supercell (a=20.52 bohr, Gamma point, 128 occupied + 16 empty bands, 30 Ry
wave cutoff, 72^3 grid, 512 projectors) with all state resident in HBM.
Inputs are synthetic (seeded displacements, seeded initial coefficients,
model potential and one-centre matrices).  Prints one JSON line whose value
is the mixed form's steps/s; the plain FP64 form is timed in the same run and
the error of the mixed form against it is measured step by step.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "cp-paw_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_counts(system, plane, nb, prec):
    """Algorithmic work per launch of the step's kernels.

    Complex GEMM M x N x K: 8 M N K flops (split-BF16x3 credited at the same
    count).  Grid passes at the element sizes of the form: the FP64 form
    transforms complex128 with float64 V and density, the mixed form complex64
    with float32 V and density.  The coefficients are complex128 in both."""
    nbh = (nb + 1) // 2
    npro = system.npro
    ngl = plane.ngl
    nr = plane.nrl
    c, r = (16.0, 8.0) if prec == 0 else (8.0, 4.0)
    return {
        "projections": ("flop", 8.0 * npro * nbh * ngl),
        "addpro": ("flop", 8.0 * npro * nbh * ngl),
        "force": ("flop", 8.0 * npro * nbh * ngl),
        "overlap": ("flop", 2 * 8.0 * nbh * nbh * ngl),
        "addopsi": ("flop", 2 * 8.0 * nbh * nbh * ngl),
        # V.psi for all super bands: N (2 sizeof(c) + sizeof(r)) per band
        "vpsi": ("byte", nbh * nr * (2 * c + r)),
        # density: N (sizeof(c) + 2 sizeof(r)) per band
        "density": ("byte", nbh * nr * (c + 2 * r)),
        "propagate": ("byte", 4.0 * nbh * ngl * 16.0),
    }


# peak of each counted kernel's arithmetic in each form.  FP64 form: the
# projector and overlap GEMMs run on FP64 MFMA.  Mixed form: projections,
# ADDPRO, overlaps and ADDOPSI run as split-BF16x3 on BF16 MFMA (six piece
# products per real product, credited at the plain 8MNK count against the
# BF16 dense peak); the force GEMM stays FP64.  The rest is HBM-bound.
GEMM_KERNELS = ("projections", "addpro", "force", "overlap", "addopsi")
BF16_PEAK_TFLOPS = 2500.0
SPLIT_PRODUCTS = 6


def kernel_peak(name, prec):
    if name in GEMM_KERNELS:
        if prec == 1:
            return BF16_PEAK_TFLOPS, "mfma"
        return FP64_PEAK_TFLOPS, "mfma"
    return HBM_PEAK_GBS, "hbm"


def _copy_state(src, dst):
    """PSI0, PSIM and the multiplier history of one set into another."""
    dst.setc8a("PSI0", src.getc8a("PSI0"))
    dst.setc8a("PSIM", src.getc8a("PSIM"))
    dst.set_matrix_c("RLAM0", src.getc8a("RLAM0"))
    if src.geti4("NLAMBDA") >= 2:
        dst.set_matrix_c("RLAMM", src.getc8a("RLAMM"))


def _trajectory(s, plane, start, nsteps, per_step):
    """Both forms from the state of `start` (an FP64 set) for nsteps CP steps.

    trajectory: the two forms run free; the largest |dE|/atom and |dF| along
    the way, their values at a few steps (how the difference grows), and the
    orthonormality residual of each form's wave functions in the FP64 metric
    (WAVES$GETR8('ORTHOERR')).  per step (per_step=True): at every step the
    mixed form is also handed the FP64 state and evaluates WAVES$ETOT; that
    difference is one step's error without the divergence of trajectories."""
    import numpy as np
    import pwgpu
    w64, wtr = [pwgpu.Waves.from_system(plane, s, prec=p) for p in (0, 1)]
    wmx = pwgpu.Waves.from_system(plane, s, prec=1) if per_step else None
    for w in (w64, wtr) + ((wmx,) if wmx else ()):
        w.etot()                     # first call: Gram-Schmidt of its own start, then replaced
        _copy_state(start, w)
        w.projections()
    de = df = tde = tdf = 0.0
    ortho_mx = ortho_64 = 0.0
    grow = {}
    marks = {1, 2, 5, 10, 20, 50, 100, 150, 200, nsteps}
    for it in range(nsteps):
        e0 = w64.etot()["etot"]
        f0 = w64.getr8a("FORCE")
        if wmx is not None:
            _copy_state(w64, wmx)
            wmx.projections()
            e1 = wmx.etot()["etot"]
            de = max(de, abs(e1 - e0) / s.nat)
            df = max(df, float(np.max(np.abs(wmx.getr8a("FORCE") - f0))))
        et = wtr.etot()["etot"]
        ste = abs(et - e0) / s.nat
        stf = float(np.max(np.abs(wtr.getr8a("FORCE") - f0)))
        tde, tdf = max(tde, ste), max(tdf, stf)
        for w in (w64, wtr):
            w.propagate()
            w.orthogonalize()
            w.switch()
        om, o6 = wtr.getr8("ORTHOERR"), w64.getr8("ORTHOERR")
        ortho_mx, ortho_64 = max(ortho_mx, om), max(ortho_64, o6)
        if it + 1 in marks:
            grow[str(it + 1)] = {"dE_per_atom": ste, "dF": stf, "ortho_mixed": om}
    for w in (w64, wtr) + ((wmx,) if wmx else ()):
        w.close()
    out = {"trajectory_max_dE_per_atom_Ha": tde, "trajectory_max_dF_Ha_per_bohr": tdf,
           "max_ortho_residual_mixed": ortho_mx, "max_ortho_residual_fp64": ortho_64, "growth": grow}
    if per_step:
        out.update({"per_step_max_dE_per_atom_Ha": de, "per_step_max_dF_Ha_per_bohr": df})
    return out


def accuracy(system_name, nsteps, relax_tol=1.0e-8, relax_max=3000, friction=0.1, random_steps=20):
    """Error of the mixed form against the FP64 form on the Si64 cell.

    The CP run starts, as a CP simulation does, from relaxed wave functions:
    the FP64 form runs damped CP (FRICTION) from the seeded start until the
    energy changes by less than relax_tol per step for 5 steps; both forms
    then continue without friction for nsteps steps from that state (see
    _trajectory).  The same comparison from the raw seeded start (far from
    the ground state, where the energy is first order in any deviation of the
    wave functions) is reported beside it over random_steps steps."""
    import pwgpu
    from pwgpu import model
    s = model.SYSTEMS[system_name]()
    plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
    w = pwgpu.Waves.from_system(plane, s, prec=0)
    w.setr8("FRICTION", friction)
    e_old, calm, nrelax, de_last = None, 0, 0, None
    while nrelax < relax_max and calm < 5:
        e = w.etot()["etot"]
        w.propagate()
        w.orthogonalize()
        w.switch()
        nrelax += 1
        if e_old is not None:
            de_last = abs(e - e_old)
        calm = calm + 1 if (de_last is not None and de_last < relax_tol) else 0
        e_old = e
    w.setr8("FRICTION", s.annee)
    t0 = time.perf_counter()
    res = _trajectory(s, plane, w, nsteps, True)
    w.close()
    log(f"accuracy: relaxed start ({nrelax} damped steps), {nsteps} steps in {time.perf_counter() - t0:.1f}s")
    w = pwgpu.Waves.from_system(plane, s, prec=0)
    w.etot()
    rnd = _trajectory(s, plane, w, random_steps, False)
    w.close()
    plane.close()
    gates = {"dE_per_atom": 1e-6, "dF": 1e-4}
    p_step = bool(res["per_step_max_dE_per_atom_Ha"] < 1e-6 and res["per_step_max_dF_Ha_per_bohr"] < 1e-4)
    p_traj = bool(res["trajectory_max_dE_per_atom_Ha"] < 1e-6 and res["trajectory_max_dF_Ha_per_bohr"] < 1e-4)
    reached = calm >= 5
    res.update({"steps": nsteps, "start": f"relaxed: damped FP64 CP (friction {friction}) for {nrelax} steps, "
                                          + (f"|dE| < {relax_tol:g} Ha per step reached" if reached else
                                             f"stopped at the step limit with |dE| = {de_last:.2e} Ha per step"),
                "relax_converged": bool(reached), "relax_last_dE_Ha": de_last,
                "random_start": dict(rnd, steps=random_steps),
                "gates": gates, "pass_per_step": p_step, "pass_trajectory": p_traj, "pass": p_step and p_traj})
    return res


def fft_pair_us(system_name, prec, reps=10):
    """PLANEWAVE$FFT of all super bands on device arrays: one GTOR and one RTOG
    of NBH transforms (the full-grid FFT pair), microseconds per pair."""
    import numpy as np
    import pwgpu
    from pwgpu import model
    s = model.SYSTEMS[system_name]()
    plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
    nbh = (s.nb + 1) // 2
    rng = np.random.default_rng(3)
    g = pwgpu.DeviceArray.from_host(rng.normal(size=(nbh, plane.ngl)) + 1j * rng.normal(size=(nbh, plane.ngl)))
    r = pwgpu.DeviceArray((nbh, plane.nrl), np.complex128 if prec == 0 else np.complex64)
    lib = plane.lib
    for _ in range(2):
        plane.fft("GTOR", nbh, g, r, prec)
        plane.fft("RTOG", nbh, g, r, prec)
    lib.pw_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plane.fft("GTOR", nbh, g, r, prec)
        plane.fft("RTOG", nbh, g, r, prec)
    lib.pw_device_synchronize()
    us = 1e6 * (time.perf_counter() - t0) / reps
    g.free()
    r.free()
    plane.close()
    return us


def synthetic_components(steps=10, warmup=2):
    """The kernel-level synthetic step (256 bands, 64^3, 512 projectors):
    per-component times of both forms, microseconds per call."""
    out = {}
    for prec, form in ((0, "fp64"), (1, "mixed")):
        r = run_gpu("synthetic", prec, steps, warmup, True)
        kt = r["kernels"]
        avg = lambda n: 1e3 * kt[n]["ms_avg"] if n in kt else None
        per_step = lambda n: 1e3 * kt[n]["ms_total"] / steps if n in kt else 0.0
        out[form] = {"steps_per_s": r["steps_per_s"], "ms_per_step": r["ms_per_step"],
                     "fft_pair_us": fft_pair_us("synthetic", prec),
                     "vpsi_us": avg("vpsi"), "density_us": avg("density"),
                     "projections_us": avg("projections"), "overlap_us": avg("overlap"),
                     "verlet_plus_ortho_us_per_step": per_step("propagate") + per_step("ortho_x")
                     + per_step("addopsi") + per_step("addoproj"),
                     "desc": r["desc"]}
    return out


def energy_drift(system_name, nsteps=500, relax_max=400, friction=0.1):
    """Energy conservation of the undamped CP run in both forms: the FP64 form
    relaxes the seeded start with friction (relax_max steps), then both forms
    run nsteps steps without friction from that state.  The conserved quantity
    is ETOT + EFICT (fixed ions); its drift is the slope of a straight-line fit,
    in Ha per ps and per atom.  Beside it the per-step error of the mixed form
    handed the FP64 state each step (as in accuracy())."""
    import numpy as np
    import pwgpu
    from pwgpu import model
    s = model.SYSTEMS[system_name]()
    plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
    w = pwgpu.Waves.from_system(plane, s, prec=0)
    w.setr8("FRICTION", friction)
    for _ in range(relax_max):
        w.step()
    w.setr8("FRICTION", 0.0)
    forms = {}
    wmx = pwgpu.Waves.from_system(plane, s, prec=1)
    wmx.etot()
    de = df = 0.0
    for prec, form in ((0, "fp64"), (1, "mixed")):
        x = pwgpu.Waves.from_system(plane, s, prec=prec)
        x.etot()
        _copy_state(w, x)
        x.projections()
        x.setr8("FRICTION", 0.0)
        econs = []
        t0 = time.perf_counter()
        for it in range(nsteps):
            e = x.etot()["etot"]
            if prec == 0:
                f0 = x.getr8a("FORCE")
                _copy_state(x, wmx)
                wmx.projections()
                e1 = wmx.etot()["etot"]
                de = max(de, abs(e1 - e) / s.nat)
                df = max(df, float(np.max(np.abs(wmx.getr8a("FORCE") - f0))))
            x.propagate()
            x.orthogonalize()
            econs.append(e + x.getr8("EFICT"))
            x.switch()
        secs = time.perf_counter() - t0
        t_ps = np.arange(nsteps) * s.dt * 2.4188843265857e-5     # atomic time unit in ps
        slope = float(np.polyfit(t_ps, np.array(econs), 1)[0])
        forms[form] = {"drift_Ha_per_ps": slope, "drift_Ha_per_ps_per_atom": slope / s.nat,
                       "max_dev_Ha": float(np.max(np.abs(np.array(econs) - econs[0]))),
                       "wall_s": secs}
        x.close()
    wmx.close()
    w.close()
    plane.close()
    return {"steps": nsteps, "dt_au": s.dt, "relax_steps": relax_max, "forms": forms,
            "per_step_max_dE_per_atom_Ha": de, "per_step_max_dF_Ha_per_bohr": df,
            "conserved": "ETOT + EFICT (ions fixed), undamped, from a damped FP64 relaxation"}


def serving(nreq, rate, slots, seed=7, reference=True):
    """Poisson stream of independent relaxations (Si8/16/32/64, seeded
    displacements) through the request loop in the mixed form: req/s,
    p50/p99 latency (submit to result) and, per request (reference=True),
    |dE| per atom against the FP64 form run to the same step."""
    import numpy as np
    from pwgpu import model, serve
    rng = np.random.default_rng(seed)
    sizes = rng.choice([8, 16, 32, 64], size=nreq)
    gaps = rng.exponential(1.0 / rate, size=nreq)
    cells = [model.si_supercell(int(n), seed=100 + i, dt=8.0, emass=300.0,
                                disp_sigma=float(rng.uniform(0.05, 0.1))) for i, n in enumerate(sizes)]
    opts = dict(friction=0.05, etol=1.0e-5, nconv=3, max_steps=1500)
    arrays = [serve.request_arrays(c) for c in cells]          # host set-up outside the clock
    srv = serve.Server(200 << 30, slots)
    t0 = time.perf_counter()
    ids = []
    for c, a, g in zip(cells, arrays, gaps):
        time.sleep(g)
        ids.append(srv.submit(c, prec=1, arrays=a, **opts))
    res = [srv.wait(i) for i in ids]
    wall = time.perf_counter() - t0
    st = srv.stats()
    lat = np.array([r["t_end"] - r["t_submit"] for r in res])
    if not reference:
        srv.close()
        return {"arrival_rate_per_s": rate, "requests": nreq, "req_per_s": nreq / wall,
                "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
                "converged": int(sum(r["converged"] for r in res)), "peak_active": st["peak_active"],
                "pool_reused": st["pool_reused"], "pool_created": st["pool_created"]}
    # FP64 form of every request, to the step the mixed form stopped at
    ref = [srv.submit(c, prec=0, arrays=a, friction=opts["friction"], etol=0.0, nconv=1, max_steps=r["steps"])
           for c, a, r in zip(cells, arrays, res)]
    ref = [srv.wait(i) for i in ref]
    srv.close()
    de = [abs(r["etot"] - f["etot"]) / c.nat for r, f, c in zip(res, ref, cells)]
    return {"requests": nreq, "arrival_rate_per_s": rate, "slots": slots, "form": "mixed",
            "req_per_s": nreq / wall, "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_steps": float(np.mean([r["steps"] for r in res])),
            "converged": int(sum(r["converged"] for r in res)),
            "peak_active": st["peak_active"], "peak_device_GiB": st["peak_bytes"] / 2**30,
            "pool_reused": st["pool_reused"], "pool_created": st["pool_created"],
            "max_dE_per_atom_vs_fp64": float(max(de)),
            "mix": {f"Si{n}": int(sum(sizes == n)) for n in (8, 16, 32, 64)},
            "relaxation": "damped CP, emass 300, dt 8, friction 0.05, |dE|<1e-5 Ha on 3 steps"}


def run_gpu(system_name, prec, steps, warmup, profile):
    import pwgpu
    from pwgpu import model
    s = model.SYSTEMS[system_name]()
    plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
    w = pwgpu.Waves.from_system(plane, s, prec=prec)
    w.autotune()                      # launch shapes of this system shape (the product's autotuner)
    lib = w.lib
    t0 = time.time()
    for i in range(warmup):
        w.step()
    lib.pw_device_synchronize()
    log(f"[{system_name} prec={prec}] warmup {warmup} steps {time.time() - t0:.2f}s")
    if profile:
        w.profile(True)
    lib.pw_device_synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        w.step()
    lib.pw_device_synchronize()
    dt = time.perf_counter() - t0
    e = {k: w.getr8(k) for k in ("ETOT", "EKIN", "EFICT")}
    res = {"steps_per_s": steps / dt, "ms_per_step": 1e3 * dt / steps, "energies": e,
           "ortho_iter": w.geti4("ORTHOITER"), "ortho_levels": w.geti4("ORTHOLEVELS"),
           "ortho_ns": w.geti4("ORTHONS"), "ortho_series": w.geti4("ORTHOSERIES"),
           "launch_shapes": {"split_waves": w.geti4("SPLITWAVES"), "density_blocks": w.geti4("DENSBLOCKS"),
                             "col_waves": w.geti4("COLWAVES")},
           "desc": f"Gamma, {s.nat} atoms, {s.nb} bands, {plane.ngl} G, {plane.nr[0]}x{plane.nr[1]}x{plane.nr[2]}, "
                   f"{s.npro} projectors"}
    if profile:
        names = ["projections", "addpro", "force", "overlap", "overlap1c", "addopsi", "addoproj",
                 "vpsi", "density", "propagate", "opsi_scale", "ortho_x", "etot_sums", "wavekinetic",
                 "denmat", "atom_mat", "phase", "reduce", "vrho", "split", "pgen"]
        kt = {}
        for n in names:
            ms, cnt = w.kernel_time(n)
            if cnt:
                kt[n] = {"ms_total": ms, "count": cnt, "ms_avg": ms / cnt}
        res["kernels"] = kt
        res["counts"] = kernel_counts(s, plane, s.nb, prec)
        # per-section times under the reference's clock names (TIMING$CLOCKON
        # in src/paw_waves1.f90 / paw_waves2.f90), microseconds per step
        sec = {}
        for n in ("WAVES$ETOT", "W:FORCE", "W:HPSI", "WAVES$PROPAGATE", "WAVES$ORTHOGONALIZE", "WAVES_OPSI",
                  "WAVES_OVERLAP", "WAVES_ADDOPSI"):
            ms, cnt = w.kernel_time(n)
            if cnt:
                sec[n] = 1e3 * ms / steps
        for ours, ref in (("density", "W:RHO"), ("vpsi", "W:HPSI.VPSI"), ("ortho_x", "WAVES_ORTHO_X")):
            if ours in kt:
                sec[ref] = 1e3 * kt[ours]["ms_total"] / steps
        res["sections_us_per_step"] = sec
    w.close()
    plane.close()
    return res


def potential_measure(system_name, calls, steps, with_cpu):
    """POTENTIAL$VOFRHO on the density object of the cell (model valence
    density, model setup tables): ms per call (host results included), the
    mixed-form CP step rate with the step linked into ETOT, and the FP64 oracle
    on the host for one call."""
    import numpy as np
    import pwgpu
    from pwgpu import model
    s = model.SYSTEMS[system_name]()
    dplane = pwgpu.PlaneWave(s.rbas, s.epw_rho, s.epw_rho)
    g2 = dplane.getr8a("G2")
    tabs = s.potential_tables(g2)
    pot = pwgpu.Potential(dplane, tabs)
    rho = s.valence_density(dplane.nr)
    r = pwgpu.DeviceArray.from_host(np.ascontiguousarray(rho).reshape(-1))
    v = pwgpu.DeviceArray((dplane.nrl,), np.float64)
    qlm = s.qlm()
    for _ in range(3):
        pot.vofrho_device(s.isp, s.pos, qlm, r, v)
    t0 = time.perf_counter()
    for _ in range(calls):
        pot.vofrho_device(s.isp, s.pos, qlm, r, v)
    ms = 1e3 * (time.perf_counter() - t0) / calls
    plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
    w = pwgpu.Waves.from_system(plane, s, prec=pwgpu.MIXED)
    w.set_potential(pot, qlm)
    for _ in range(3):
        w.step()
    w.lib.pw_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        w.step()
    w.lib.pw_device_synchronize()
    sps = steps / (time.perf_counter() - t0)
    out = {"vofrho_ms": ms, "density_G": dplane.ngl, "grid": list(dplane.nr), "atoms": s.nat,
           "scf_steps_per_s": sps, "scf_form": "mixed (VOFRHO in FP64)",
           "scope": "one spin, LDA (PZ), l=0 compensation, model setup tables"}
    w.close()
    plane.close()
    if with_cpu:
        from oracle import planewave as opw
        from oracle import potential as opot
        pl = opw.Plane(s.rbas, s.epw_rho, s.epw_rho)
        t0 = time.perf_counter()
        opot.vofrho(pl, rho, s.pos, s.isp, qlm,
                    {"pscorg": tabs[0], "vbarg": tabs[1], "g0": tabs[2], "v0": tabs[3]})
        out["cpu_oracle_vofrho_s"] = time.perf_counter() - t0
    return out


def cpu_baseline(system_name, nsteps=20, nsteps_1t=20):
    """The FP64 oracle (numpy) on the host, as the reference's clock takes it
    (first step excluded, src/paw_driver.f90:156-160): on all the BLAS
    threads of the process (the host's share of cores) the median of nsteps
    steps after one untimed step, and the same on one thread over nsteps_1t."""
    from threadpoolctl import threadpool_info, threadpool_limits
    threads = max([i.get("num_threads", 1) for i in threadpool_info()] + [1])
    out = _cpu_oracle_steps(system_name, nsteps)
    out["cores"] = int(threads)
    out["host_cpus"] = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    with threadpool_limits(limits=1):
        one = _cpu_oracle_steps(system_name, nsteps_1t)
    one["cores"] = 1
    out["single_thread"] = one
    out["sample"] += f" on {threads} threads; single_thread: the same on one thread"
    return out


def _cpu_oracle_steps(system_name, nsteps):
    import numpy as np
    from pwgpu import model
    from oracle import planewave as opw
    from oracle import waves as owv
    s = model.SYSTEMS[system_name]()
    pl = opw.Plane(s.rbas, s.epw_psi, s.epw_rho)
    st = owv.Setup(s.isp, s.lox_list, s.do_list())
    ws = owv.WaveSet(pl, st, s.pro_table(pl.g2), s.nb, s.occ, s.potential((pl.nr1, pl.nr2, pl.nr3)),
                     s.dh_matrices(), s.dt, s.emass, s.emasscg2, s.annee)
    ws.psi0 = s.initial_psi(pl.g2)
    ws.psim = ws.psi0.copy()

    def step():
        ws.etot(s.pos)
        ws.propagate()
        ws.orthogonalize(s.pos, s.pos)
        ws.switch()

    step()
    times = []
    for _ in range(nsteps):
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": 1.0 / med, "unit": "CP steps/s", "kind": "port",
            "sample": f"median of {nsteps} CP steps of {s.name} (FP64 numpy oracle, one band per "
                      f"FFT as the reference) after one untimed step"}


def profiled_traffic(kernel, system, prec):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the
    copy-kernel calibration), or None when that kernel/workload was not measured."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    v = tab.get(f"{system}/{'fp64' if prec == 0 else 'mixed'}/{kernel}")
    return None if v is None else v.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--system", default="si64")
    ap.add_argument("--prec", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=20,
                    help="timed oracle steps of the CPU baseline on all cores (median)")
    ap.add_argument("--cpu-steps-1t", type=int, default=20,
                    help="timed oracle steps of the CPU baseline on one thread (median)")
    ap.add_argument("--acc-steps", type=int, default=200,
                    help="steps of the mixed-vs-FP64 error measurement (0: skip)")
    ap.add_argument("--serve-requests", type=int, default=12,
                    help="requests of the serving measurement (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-slots", type=int, default=8, help="concurrent requests on the GPU")
    ap.add_argument("--serve-sweep", default="1,4,16,50",
                    help="arrival rates (req/s) of the sweep; empty for none")
    ap.add_argument("--serve-sweep-requests", type=int, default=40, help="requests per arrival rate of the sweep")
    ap.add_argument("--serve-p99-bound", type=float, default=2.0,
                    help="p99 latency (s) within which a swept arrival rate counts as sustained")
    ap.add_argument("--profile", action="store_true", help="print per-kernel times to stderr")
    ap.add_argument("--other-configs", type=int, default=1,
                    help="1: also measure the synthetic step's components and the 32 H2O run (throughput, "
                         "500-step energy drift, per-step error)")
    ap.add_argument("--pot-calls", type=int, default=20,
                    help="calls of the density -> potential measurement (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (calibrates the HBM "
                         "byte counters of a profiler run) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20)
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(the decomposition over several GPUs is a design only)"}))
        sys.exit(2)

    import pwgpu
    pwgpu.load()
    precs = {"fp64": [0], "mixed": [1], "both": [0, 1]}[args.prec]
    results = {}
    for pr in precs:
        results[pr] = run_gpu(args.system, pr, args.steps, args.warmup, False)
        log(f"prec={pr}: {results[pr]['steps_per_s']:.2f} steps/s "
            f"({results[pr]['ms_per_step']:.3f} ms/step) E={results[pr]['energies']} "
            f"ortho iter/levels/ns={results[pr]['ortho_iter']}/{results[pr]['ortho_levels']}/{results[pr]['ortho_ns']}")
    head = precs[-1]
    # kernel timing in a separate short run (events add host overhead)
    prof = run_gpu(args.system, head, max(5, min(20, args.steps)), 2, True)
    kt = prof["kernels"]
    counts = prof["counts"]
    for n, v in sorted(kt.items(), key=lambda kv: -kv[1]["ms_total"]):
        log(f"  {n:12s} {v['ms_avg']*1e3:9.1f} us x{v['count']:4d}  total {v['ms_total']:.2f} ms")
    # dominant kernel among those with an algorithmic count
    cand = [(v["ms_total"], n) for n, v in kt.items() if n in counts]
    dom = max(cand)[1]
    kind, work = counts[dom]
    avg_s = kt[dom]["ms_avg"] * 1e-3
    peak, bound = kernel_peak(dom, head)
    traffic = profiled_traffic(dom, args.system, head)
    if kind == "flop":
        achieved = work / avg_s / 1e12
        unit = "TFLOP/s"
    else:
        achieved = work / avg_s / 1e9
        unit = "GB/s"
    roof = {"kernel": dom, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": traffic,
            "traffic_rate_GBs": None if traffic is None else traffic / avg_s / 1e9,
            "work_per_launch": work, "avg_us": avg_s * 1e6}
    if peak == BF16_PEAK_TFLOPS:
        roof["split_products"] = SPLIT_PRODUCTS
    acc = accuracy(args.system, args.acc_steps) if (args.acc_steps > 0 and head == 1) else None
    srv = None
    if args.serve_requests > 0:
        log(f"serving: {args.serve_requests} requests at {args.serve_rate}/s, {args.serve_slots} slots ...")
        srv = serving(args.serve_requests, args.serve_rate, args.serve_slots)
        log(f"serving: {srv['req_per_s']:.2f} req/s p50 {srv['p50_latency_s']:.2f}s p99 {srv['p99_latency_s']:.2f}s")
        rates = [float(x) for x in args.serve_sweep.split(",") if x.strip()]
        if rates:
            # the arrival-rate sweep
            srv["sweep"] = []
            for k, rate in enumerate(rates):
                pt = serving(args.serve_sweep_requests, rate, args.serve_slots, seed=11 + k, reference=False)
                srv["sweep"].append(pt)
                log(f"serving sweep {rate:g}/s: {pt['req_per_s']:.2f} req/s p50 {pt['p50_latency_s']:.2f}s "
                    f"p99 {pt['p99_latency_s']:.2f}s")
            # sustained rate: the highest swept arrival rate served without a
            # queue building up (p99 latency within the bound), beside the
            # saturation throughput
            kept = [pt for pt in srv["sweep"] if pt["p99_latency_s"] <= args.serve_p99_bound]
            top = max(srv["sweep"], key=lambda pt: pt["req_per_s"])
            best = max(kept, key=lambda pt: pt["arrival_rate_per_s"]) if kept else None
            srv["sustained"] = {
                "arrival_rate_per_s": best["arrival_rate_per_s"] if best else None,
                "req_per_s": best["req_per_s"] if best else None,
                "p99_latency_s": best["p99_latency_s"] if best else None,
                "saturation_req_per_s": top["req_per_s"], "saturation_p99_latency_s": top["p99_latency_s"],
                "rule": f"highest swept arrival rate with p99 latency <= {args.serve_p99_bound:g} s"}
    potm = None
    if args.pot_calls > 0:
        potm = potential_measure(args.system, args.pot_calls, max(5, min(20, args.steps)), not args.no_cpu)
        log(f"vofrho: {potm['vofrho_ms']:.3f} ms/call; self-consistent mixed step "
            f"{potm['scf_steps_per_s']:.1f} steps/s")
    other = None
    if args.other_configs:
        log("synthetic step (256 bands, 64^3, 512 projectors): components ...")
        syn = synthetic_components()
        log("  " + ", ".join(f"{f}: {v['ms_per_step']:.2f} ms/step, FFT pair {v['fft_pair_us']:.0f} us"
                             for f, v in syn.items()))
        log("32 H2O: throughput and 500-step energy drift ...")
        h2o = {f: run_gpu("h2o32", p, 20, 3, False) for p, f in ((0, "fp64"), (1, "mixed"))}
        drift = energy_drift("h2o32")
        log(f"  fp64 {h2o['fp64']['steps_per_s']:.1f}, mixed {h2o['mixed']['steps_per_s']:.1f} steps/s; drift "
            + ", ".join(f"{f} {v['drift_Ha_per_ps']:.2e} Ha/ps" for f, v in drift["forms"].items()))
        other = {"synthetic": syn,
                 "h2o32": {"steps_per_s": {f: v["steps_per_s"] for f, v in h2o.items()},
                           "desc": h2o["mixed"]["desc"], "energy_drift": drift}}
    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(args.system, args.cpu_steps, args.cpu_steps_1t)
        log(f"cpu baseline: {cpu['value']:.4f} steps/s on {cpu['cores']} threads, "
            f"{cpu['single_thread']['value']:.4f} steps/s on 1 thread")
    r = results[head]
    out = {
        "metric": "CP steps/s Si64 mixed vs FP64 @|dE|<1e-6 Ha/atom; serving req/s + p50/p99 lat",
        "value": r["steps_per_s"],
        "unit": "CP steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == 0 else "mixed(f32 fft+V.psi, split-bf16x3 mfma gemms, f64 constraint/energy/force)",
        "data": f"synthetic (seeded {args.system} positions and coefficients, model potential and projectors)",
        "config": {"workload": f"{args.system} CP wavefunction step ({r['desc']})",
                   "form": "fp64" if head == 0 else "mixed"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if acc is not None:
        out["mixed_error_vs_fp64"] = acc
    if srv is not None:
        out["serving"] = srv
    if potm is not None:
        out["potential"] = potm
    if other is not None:
        out["other_configs"] = other
    if 0 in results and 1 in results:
        out["fp64_steps_per_s"] = results[0]["steps_per_s"]
        out["mixed_over_fp64"] = results[1]["steps_per_s"] / results[0]["steps_per_s"]
    out["sections_us_per_step"] = prof["sections_us_per_step"]
    out["ortho_iter_last"] = r["ortho_iter"]
    out["ortho_series_terms"] = r["ortho_series"]
    out["launch_shapes"] = r["launch_shapes"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
