"""Request loop (pw_server_*): admission by device footprint, concurrent
relaxations on separate streams, and the same result for a request alone and
in a mix of concurrent requests."""
import numpy as np
import pytest

from pwgpu import model

GIB = 1 << 30


def test_footprint_is_host_only_and_scales():
    import pwgpu
    from pwgpu import _lib
    import ctypes
    lib = _lib.load()
    sizes = []
    for nat in (8, 16, 32, 64):
        s = model.si_supercell(nat)
        plane = pwgpu.PlaneWave(s.rbas, s.epw_psi, s.epw_rho)
        isp = (np.asarray(s.isp) + 1).astype(np.int32)
        lnx = np.array([len(l) for l in s.lox_list], dtype=np.int32)
        lox = np.array(s.lox_list[0], dtype=np.int32)
        b = ctypes.c_longlong()
        for prec in (0, 1):
            _lib.check(lib.pw_waves_footprint(plane.handle, s.nb, s.nat, 1, _lib.iptr(isp), _lib.iptr(lnx),
                                              int(lnx.max()), _lib.iptr(lox), prec, ctypes.byref(b)))
            nbh = (s.nb + 1) // 2
            assert b.value > 5 * nbh * plane.ngl * 16          # at least the coefficient arrays
            if prec == 1:
                sizes.append(b.value)
        plane.close()
    assert sizes == sorted(sizes)


def test_request_larger_than_budget_is_rejected():
    from pwgpu import serve, _lib
    srv = serve.Server(16 << 20, 2)         # 16 MiB: too small for Si8
    try:
        with pytest.raises(_lib.PwError) as e:
            srv.submit(model.si8())
        assert e.value.code == _lib.PW_OK + 2   # PW_ERR_UNSUPPORTED
    finally:
        srv.close()


@pytest.mark.gpu
def test_alone_equals_in_mix():
    """Each of Si8, Si16, Si32 and Si64 relaxed alone, then all four at once
    with one more Si8: every result in the mix equals its result alone (same
    step count, energy to 1e-12), including those that run on a pooled set
    that an earlier request of the same shape used."""
    from pwgpu import serve
    print("case serving alone vs mix", flush=True)
    # damped CP relaxation: fictitious mass 300, time step 8, friction 0.05
    cell = lambda nat, seed: model.si_supercell(nat, seed=seed, dt=8.0, emass=300.0)
    opts = dict(friction=0.05, etol=1.0e-5, nconv=3, max_steps=800)
    targets = [cell(8, 11), cell(16, 3), cell(32, 5), cell(64, 9)]
    srv = serve.Server(64 * GIB, 5)
    try:
        alone = [srv.wait(srv.submit(t, **opts)) for t in targets]
        mix = [targets[1], targets[0], cell(8, 12), targets[3], targets[2]]
        ids = [srv.submit(s, **opts) for s in mix]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    for t, a in zip(targets, alone):
        print(f"  alone {t.name}: E={a['etot']:.12f} steps={a['steps']} converged={a['converged']}", flush=True)
    for s, r in zip(mix, res):
        print(f"  {s.name}: E={r['etot']:.10f} steps={r['steps']} converged={r['converged']} "
              f"wait={r['t_start'] - r['t_submit']:.3f}s run={r['t_end'] - r['t_start']:.3f}s", flush=True)
    print(f"  peak active {st['peak_active']}, peak bytes {st['peak_bytes'] / 2**20:.0f} MiB, "
          f"pool reused {st['pool_reused']} created {st['pool_created']}", flush=True)
    assert all(a["converged"] == 1 for a in alone)
    assert all(r["converged"] == 1 for r in res)
    for k, t in enumerate(targets):
        mine = res[next(i for i, m in enumerate(mix) if m is t)]
        assert mine["steps"] == alone[k]["steps"], t.name
        assert abs(mine["etot"] - alone[k]["etot"]) <= 1e-12 * abs(alone[k]["etot"]), t.name
    assert st["peak_active"] >= 2
    # the four shapes of the alone runs were pooled and served the mix again
    assert st["pool_reused"] >= 4 and st["pool_created"] >= 4


@pytest.mark.gpu
def test_residual_convergence():
    """A request converged on the residual norm (rtol) instead of the energy
    change: it stops with the residual below rtol, at an energy within the
    energy rule's tolerance of the same cell relaxed on |dE|."""
    from pwgpu import serve
    print("case serving residual convergence", flush=True)
    cell = model.si_supercell(8, seed=11, dt=8.0, emass=300.0)
    srv = serve.Server(16 * GIB, 2)
    try:
        by_r = srv.wait(srv.submit(cell, friction=0.05, rtol=1.0e-5, nconv=3, max_steps=4000))
        by_e = srv.wait(srv.submit(cell, friction=0.05, etol=1.0e-7, nconv=3, max_steps=4000))
    finally:
        srv.close()
    print(f"  residual rule: steps={by_r['steps']} residual={by_r['residual']:.2e} E={by_r['etot']:.10f}", flush=True)
    print(f"  energy rule:   steps={by_e['steps']} E={by_e['etot']:.10f}", flush=True)
    assert by_r["converged"] == 1 and by_e["converged"] == 1
    assert 0.0 < by_r["residual"] < 1.0e-5
    assert abs(by_r["etot"] - by_e["etot"]) < 1.0e-5
