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
    from pwgpu import serve
    print("case serving alone vs mix", flush=True)
    # damped CP relaxation: fictitious mass 300, time step 8, friction 0.05
    cell = lambda nat, seed: model.si_supercell(nat, seed=seed, dt=8.0, emass=300.0)
    opts = dict(friction=0.05, etol=1.0e-5, nconv=3, max_steps=800)
    target = cell(8, 11)
    srv = serve.Server(64 * GIB, 4)
    try:
        alone = srv.wait(srv.submit(target, **opts))
        mix = [cell(16, 3), target, cell(32, 5), cell(8, 12)]
        ids = [srv.submit(s, **opts) for s in mix]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    print(f"  alone: E={alone['etot']:.12f} steps={alone['steps']} converged={alone['converged']}", flush=True)
    for s, r in zip(mix, res):
        print(f"  {s.name}: E={r['etot']:.10f} steps={r['steps']} converged={r['converged']} "
              f"wait={r['t_start'] - r['t_submit']:.3f}s run={r['t_end'] - r['t_start']:.3f}s", flush=True)
    print(f"  peak active {st['peak_active']}, peak bytes {st['peak_bytes'] / 2**20:.0f} MiB", flush=True)
    assert alone["converged"] == 1
    assert all(r["converged"] == 1 for r in res)
    mine = res[1]
    assert mine["steps"] == alone["steps"]
    assert abs(mine["etot"] - alone["etot"]) <= 1e-12 * abs(alone["etot"])
    assert st["peak_active"] >= 2
