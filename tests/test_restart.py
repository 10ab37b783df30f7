"""Restart files in the reference's format (pwgpu.restart and the Fortran
module PAWGPU_RESTART).

The record layout is pinned against a real Fortran compiler: the test
program tests/fortran/restart_probe.f90 writes the WAVES section with
Fortran WRITE statements of the reference's record lists (WAVES_WRITEPSI /
WAVES_WRITELAMBDA, src/paw_waves2.f90:3168-3350, 3920-4040; separators
src/paw_report.f90:107-240) and reads one written by pwgpu.restart; both
directions must agree value for value, and the WAVES section pwgpu.restart
writes must be byte-identical to the compiler's.  The G-vector mapping
(WAVES_MAPG, :4268-4399) is pinned by permutations, a strained cell and
vectors absent from the file.  No restart file of the reference ships with
its tests, so agreement with files the reference itself wrote is pinned
only through this record layout.

On the GPU: a restart written after CP steps and read into a fresh set
continues the run as the oracle does, from Python and from the Fortran
driver (WAVESGPU$WRITE / WAVESGPU$READ)."""
import os
import subprocess

import numpy as np
import pytest

from pwgpu import restart as rs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FMOD = os.path.join(ROOT, "cp-paw_amd", "fortran", "pawgpu_restart.f90")
FPROBE = os.path.join(ROOT, "tests", "fortran", "restart_probe.f90")


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    from pwgpu import build
    fc = build.flang()
    if fc is None:
        pytest.skip("no Fortran compiler")
    d = str(tmp_path_factory.mktemp("probe"))
    obj = os.path.join(d, "pawgpu_restart.o")
    exe = os.path.join(d, "restart_probe")
    for cmd in ([fc, "-O1", "-c", FMOD, "-o", obj, "-module-dir", d],
                [fc, "-O1", FPROBE, obj, "-o", exe, "-I", d, "-module-dir", d]):
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
    return exe


def run_probe(exe, mode, path):
    r = subprocess.run([exe, mode, path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


# the fixed data of restart_probe.f90, restated
RBAS = np.array([[6.0, 0.0, 0.0], [1.0, 7.0, 0.0], [0.5, 1.0, 8.0]]).T
IGVEC = np.array([(k, j - k, i + 2 * j) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)],
                 dtype=np.int32)
NB, NBH, NGG = 5, 3, 27


def probe_psi():
    ig = np.arange(1, NGG + 1)[None, :].astype(float)
    ib = np.arange(1, NBH + 1)[:, None].astype(float)
    psi0 = (ig + 0.25 * ib) - 0.5j * ig * ib
    psim = 2.0 * ig + 1j * ib + 0.0 * ig
    return psi0, psim


def probe_lam():
    i = np.arange(1, NB + 1)[:, None].astype(float)
    j = np.arange(1, NB + 1)[None, :].astype(float)
    return (i + 10.0 * j) + 0.5j * i + 0.0 * j, -i + 1j * j


def probe_kpoint():
    psi0, psim = probe_psi()
    l0, lm = probe_lam()
    return rs.KPointWaves(NB, True, np.zeros(3), IGVEC, np.stack([psi0, psim])[:, None, :, None, :],
                          np.stack([l0, lm])[:, None])


def record_spans(data):
    """(start, end) byte offsets of the records of a sequential file."""
    spans, pos = [], 0
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "little", signed=True)
        spans.append((pos, pos + 8 + n))
        pos += 8 + n
    return spans


def test_separator_layout():
    assert rs.SEPARATOR.size == 4 + 32 + 32 + 64 + 256
    sep = rs.unpack_separator(rs.pack_separator(17, "WAVES"))
    assert sep == {"nrec": 17, "id": "WAVES", "name": "NONE", "version": "AUG1996", "notes": "NONE"}


def test_fortran_written_file_is_read(probe, tmp_path):
    path = str(tmp_path / "f.rstrt")
    run_probe(probe, "write", path)
    data = open(path, "rb").read()
    assert int.from_bytes(data[:4], "little") == rs.SEPARATOR.size
    r = rs.read_file(path)
    assert r.skipped == ["ATOMS"]
    assert r.nspin == 1 and r.nwave == 2 and len(r.kpoints) == 1
    assert np.array_equal(r.rbas, RBAS)
    k = r.kpoints[0]
    assert (k.key, k.nb, k.ndim, k.tsuper, k.nbh, k.ngg) == ("PSI", NB, 1, True, NBH, NGG)
    assert np.array_equal(k.xk, np.zeros(3)) and np.array_equal(k.igvec, IGVEC)
    psi0, psim = probe_psi()
    assert np.array_equal(k.psi[0, 0, :, 0], psi0) and np.array_equal(k.psi[1, 0, :, 0], psim)
    l0, lm = probe_lam()
    assert k.nlambda == 2 and np.array_equal(k.lam[0, 0], l0) and np.array_equal(k.lam[1, 0], lm)


def test_python_waves_section_is_byte_identical(probe, tmp_path):
    fpath, ppath = str(tmp_path / "f.rstrt"), str(tmp_path / "p.rstrt")
    run_probe(probe, "write", fpath)
    rs.write_file(ppath, RBAS, [probe_kpoint()])
    fdata, pdata = open(fpath, "rb").read(), open(ppath, "rb").read()
    fs, ps = record_spans(fdata), record_spans(pdata)
    # Fortran: HEADER, ATOMS + 2 records, WAVES section, EOF; Python: no ATOMS
    fwaves = fdata[fs[4][0]:fs[-1][0]]
    pwaves = pdata[ps[1][0]:ps[-1][0]]
    assert fwaves == pwaves
    assert fdata[:fs[1][0]] == pdata[:ps[1][0]]              # HEADER
    assert fdata[fs[-1][0]:] == pdata[ps[-1][0]:]            # ENDOFFILE


def present_set():
    gbas = rs.gbass(RBAS)
    ig = np.concatenate([IGVEC[::-1], [[9, 9, 9]]]).astype(float)
    return ig @ gbas.T


def test_fortran_reads_python_file(probe, tmp_path):
    path = str(tmp_path / "p.rstrt")
    rs.write_file(path, RBAS, [probe_kpoint()])
    out = run_probe(probe, "read", path)
    rows = [l.split() for l in out.splitlines()]
    counts = [r for r in rows if r[0] == "counts"][0]
    assert (int(counts[1]), int(counts[2])) == (2, 2)
    ngp, nbp = NGG + 1, 4
    fp0 = np.zeros((2, ngp), complex)
    fpm = np.zeros((2, ngp), complex)
    fl0 = np.zeros((nbp, nbp), complex)
    flm = np.zeros((nbp, nbp), complex)
    for r in rows:
        if r[0] == "psi":
            b, g = int(r[1]) - 1, int(r[2]) - 1
            v = [float(x) for x in r[3:7]]
            fp0[b, g], fpm[b, g] = complex(v[0], v[1]), complex(v[2], v[3])
        elif r[0] == "lam":
            i, j = int(r[1]) - 1, int(r[2]) - 1
            v = [float(x) for x in r[3:7]]
            fl0[i, j], flm[i, j] = complex(v[0], v[1]), complex(v[2], v[3])
    cur = np.full((2, ngp), -1 - 1j)
    p0, pm, lams = rs.restore(rs.read_file(path), present_set(), cur, cur, nbp)
    assert np.array_equal(fp0, p0) and np.array_equal(fpm, pm)
    assert np.array_equal(fl0, lams[0]) and np.array_equal(flm, lams[1])
    # reversed order; the vector absent from the file is zero; 2 of 3 bands
    psi0, _ = probe_psi()
    assert np.array_equal(p0[:, :NGG], psi0[:2, ::-1]) and np.all(p0[:, NGG] == 0)
    l0, _ = probe_lam()
    assert np.array_equal(lams[0], l0[:4, :4])


def test_nint_rounds_halves_away_from_zero():
    assert rs._nint(np.array([0.5, -0.5, 1.5, -1.5, 2.5, 0.49])).tolist() == [1, -1, 2, -2, 3, 0]


def test_map_g_permutation_strain_and_absent():
    rng = np.random.default_rng(4)
    gbas = rs.gbass(RBAS)
    ga = IGVEC.astype(float) @ gbas.T
    perm = rng.permutation(NGG)
    assert np.array_equal(rs.map_g(gbas, ga, ga), np.arange(1, NGG + 1))
    assert np.array_equal(rs.map_g(gbas, ga, ga[perm]), perm + 1)
    # the present cell strained by 1 %: the same integer coordinates map
    gs = IGVEC.astype(float) @ rs.gbass(RBAS * 1.01).T
    assert np.array_equal(rs.map_g(gbas, ga, gs[perm]), perm + 1)
    # vectors outside the file's box or inside the box but absent map to 0
    extra = np.array([[5, 0, 0], [0, 2, 0]], float) @ gbas.T
    m = rs.map_g(gbas, ga[:10], np.concatenate([ga[:10], extra, ga[20:]]))
    assert m[:10].tolist() == list(range(1, 11)) and np.all(m[10:] == 0)


def test_map_g_errors():
    gbas = rs.gbass(RBAS)
    ga = IGVEC.astype(float) @ gbas.T
    with pytest.raises(rs.RestartError, match="SAME POINT"):
        rs.map_g(gbas, np.concatenate([ga, ga[:1]]), ga)
    bad = ga.copy()
    bad[3] += 0.3 * gbas[:, 0]
    with pytest.raises(rs.RestartError, match="CANNOT BE REPRESENTED"):
        rs.map_g(gbas, bad, ga)


def test_static_file_and_band_folding(tmp_path):
    path = str(tmp_path / "s.rstrt")
    k = probe_kpoint()
    st = rs.KPointWaves(NB, True, np.zeros(3), IGVEC, k.psi[:1], k.lam[:1])
    rs.write_file(path, RBAS, [st])
    r = rs.read_file(path)
    assert r.nwave == 1 and r.kpoints[0].nlambda == 1
    gvec = IGVEC.astype(float) @ rs.gbass(RBAS).T
    nb = 8                                   # more bands than the file: NBH 4 > 3
    cur0 = np.full((4, NGG), 3.0 + 0j)
    curm = np.full((4, NGG), 5.0 + 0j)
    p0, pm, lams = rs.restore(r, gvec, cur0, curm, nb)
    psi0, _ = probe_psi()
    assert np.array_equal(p0[:3], psi0) and np.array_equal(pm[:3], psi0)    # PSIM = PSI0
    assert np.all(p0[3] == 3.0) and np.all(pm[3] == 5.0)                      # kept
    assert len(lams) == 1
    l0, _ = probe_lam()
    assert np.array_equal(lams[0][:5, :5], l0) and np.all(lams[0][5:] == 0) and np.all(lams[0][:, 5:] == 0)


def corrupt(tmp_path, edit):
    path = str(tmp_path / "c.rstrt")
    rs.write_file(path, RBAS, [probe_kpoint()])
    data = bytearray(open(path, "rb").read())
    edit(data)
    open(path, "wb").write(bytes(data))
    return path


@pytest.mark.parametrize("what,edit,msg", [
    ("header id", lambda d: d.__setitem__(slice(8, 14), b"HEADX "), "NOT RECOGNIZED"),
    ("version", lambda d: d.__setitem__(slice(4 + 4 + 64, 4 + 4 + 64 + 7), b"JAN2000"), "version"),
    ("truncated", lambda d: d.__delitem__(slice(len(d) - 600, len(d))), "truncated|corrupted|end"),
])
def test_corrupted_files(tmp_path, what, edit, msg):
    with pytest.raises(rs.RestartError, match=msg):
        rs.read_file(corrupt(tmp_path, edit))


def test_psi_key_checked(tmp_path):
    path = str(tmp_path / "k.rstrt")
    rs.write_file(path, RBAS, [probe_kpoint()])
    data = open(path, "rb").read()
    spans = record_spans(data)
    s0 = spans[3][0] + 4                          # HEADER, WAVES, sizes, then KEY
    data = data[:s0] + b"PSX     " + data[s0 + 8:]
    open(path, "wb").write(data)
    with pytest.raises(rs.RestartError, match='NOT "PSI"'):
        rs.read_file(path)


def test_restore_rejects_mismatches():
    k = probe_kpoint()
    gvec = IGVEC.astype(float) @ rs.gbass(RBAS).T
    cur = np.zeros((3, NGG), complex)
    r2 = rs.RestartWaves(1, RBAS, [k, k])
    with pytest.raises(rs.RestartError, match="KPOINT SETS ARE FIXED"):
        rs.restore(r2, gvec, cur, cur, NB)
    r1 = rs.RestartWaves(1, RBAS, [k])
    with pytest.raises(rs.RestartError, match="INCONSISTENT K-POINTS"):
        rs.restore(r1, gvec, cur, cur, NB, xk=(0.5, 0.0, 0.0))
    with pytest.raises(rs.RestartError, match="SUPER"):
        rs.restore(r1, gvec, cur, cur, NB, tsuper=False)


# ----------------------------------------------------------------------------
# GPU
# ----------------------------------------------------------------------------
def cp_steps(o, w, s, n, check=True):
    from harness import relerr
    for it in range(n):
        eo = o.etot(s.pos)
        eg = w.etot()
        if check:
            assert abs(eg["etot"] - eo["etot"]) <= 1e-10 * abs(eo["etot"]), it
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        w.propagate()
        w.orthogonalize()
        if check:
            assert relerr(w.getc8a("RLAM0").real, o.rlam0) < 1e-9, it
        o.switch()
        w.switch()


def oracle_from_restart(s, path):
    """The oracle started from a restart file: a fresh set whose first ETOT
    re-orthogonalises (Gram-Schmidt) the wave functions read, as WAVES$ETOT
    does after a restart (src/paw_waves1.f90:1738-1740)."""
    from harness import oracle_waveset
    k = rs.read_file(path).kpoints[0]
    o = oracle_waveset(s)
    o.psi0 = k.psi[0, 0, :, 0].copy()
    o.psim = k.psi[-1, 0, :, 0].copy()
    o.rlam0 = k.lam[0, 0].real.copy()
    o.rlamm = k.lam[1, 0].real.copy() if k.nlambda > 1 else None
    return o


@pytest.mark.gpu
def test_gpu_restart_continues_run(tmp_path):
    from harness import SYSTEMS, gpu_waveset, oracle_waveset, relerr
    s = SYSTEMS["si8"]()
    print("case restart python si8 3+2", flush=True)
    o = oracle_waveset(s)
    plane, w = gpu_waveset(s)
    cp_steps(o, w, s, 3)
    path = str(tmp_path / "si8.rstrt")
    assert w.geti4("NLAMBDA") == 2
    w.write_restart(path)
    saved = {k: w.getc8a(k) for k in ("PSI0", "PSIM", "RLAM0", "RLAMM")}
    w.close()
    # a fresh set (Gram-Schmidt start wave functions) takes the state back
    plane2, w2 = gpu_waveset(s)
    assert w2.geti4("NLAMBDA") == 0
    assert w2.read_restart(path)
    assert w2.geti4("NLAMBDA") == 2
    for k, v in saved.items():
        assert np.array_equal(w2.getc8a(k), v), k
    assert np.array_equal(w2.restart_epsilon, np.diag(saved["RLAM0"]).real)
    # the file holds the state of the run (against the oracle's)
    assert relerr(saved["PSI0"], o.psi0) < 1e-9 and relerr(saved["PSIM"], o.psim) < 1e-9
    assert relerr(saved["RLAM0"].real, o.rlam0) < 1e-9 and relerr(saved["RLAMM"].real, o.rlamm) < 1e-9
    o2 = oracle_from_restart(s, path)
    cp_steps(o2, w2, s, 2)
    assert relerr(w2.getc8a("PSI0"), o2.psi0) < 1e-9


@pytest.mark.gpu
def test_gpu_static_restart(tmp_path):
    from harness import SYSTEMS, gpu_waveset
    s = SYSTEMS["tiny"]()
    print("case restart static tiny", flush=True)
    plane, w = gpu_waveset(s)
    w.etot()
    w.propagate()
    w.orthogonalize()
    w.switch()
    path = str(tmp_path / "tiny.rstrt")
    w.write_restart(path, static=True)
    r = rs.read_file(path)
    assert r.nwave == 1 and r.kpoints[0].nlambda == 1
    psi0 = w.getc8a("PSI0")
    w.close()
    plane2, w2 = gpu_waveset(s)
    assert w2.read_restart(path)
    assert np.array_equal(w2.getc8a("PSI0"), psi0) and np.array_equal(w2.getc8a("PSIM"), psi0)
    assert w2.geti4("NLAMBDA") == 1


@pytest.mark.gpu
def test_fortran_driver_restart(tmp_path):
    from harness import SYSTEMS, gpu_waveset, oracle_waveset, relerr
    from pwgpu import deck
    from pwgpu.build import FDRIVER
    if not os.path.exists(FDRIVER):
        pytest.skip("Fortran driver not built (no Fortran compiler)")
    s = SYSTEMS["si8"]()
    print("case restart fortran si8 3+2", flush=True)
    plane, w = gpu_waveset(s)
    d3, d2 = str(tmp_path / "a.deck"), str(tmp_path / "b.deck")
    deck.write_deck(d3, w, s, 3)
    deck.write_deck(d2, w, s, 2)
    w.close()
    plane.close()
    rst = str(tmp_path / "si8.rstrt")
    r = subprocess.run([FDRIVER, d3, str(tmp_path / "a.out"), "--write-restart", rst],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    o = oracle_waveset(s)
    for _ in range(3):
        o.etot(s.pos)
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        o.switch()
    k = rs.read_file(rst).kpoints[0]
    assert k.nwave == 2 and k.nlambda == 2
    assert relerr(k.psi[0, 0, :, 0], o.psi0) < 1e-9 and relerr(k.psi[1, 0, :, 0], o.psim) < 1e-9
    assert relerr(k.lam[0, 0].real, o.rlam0) < 1e-9
    r = subprocess.run([FDRIVER, d2, str(tmp_path / "b.out"), "--read-restart", rst],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "READING: WAVES" in r.stdout
    steps, _ = deck.read_output(str(tmp_path / "b.out"), s.nat)
    o = oracle_from_restart(s, rst)
    for it in range(2):
        e = o.etot(s.pos)
        o.propagate()
        o.orthogonalize(s.pos, s.pos)
        assert abs(steps[it, 1] - e["etot"]) <= 1e-10 * abs(e["etot"]), it
        assert int(steps[it, 6]) == o.ortho_iter, it
        o.switch()


def test_map_g_between_real_sphere_sets():
    # the G spheres of two cutoffs and of a strained cell (the reference's
    # sphere and order, oracle/planewave.py): every present vector inside the
    # file's sphere maps to the file's vector with the same integer coordinates
    from oracle import planewave as opw
    from pwgpu import model
    s = model.si8()
    lo = opw.Plane(s.rbas, 6.0, 24.0)
    hi = opw.Plane(s.rbas, 8.0, 32.0)
    m = rs.map_g(lo.gbas, lo.gvec, hi.gvec)
    key = {tuple(v): i + 1 for i, v in enumerate(lo.igvec.tolist())}
    want = np.array([key.get(tuple(v), 0) for v in hi.igvec.tolist()])
    assert np.array_equal(m, want) and np.count_nonzero(m) == lo.ngl
    # down to the smaller sphere: a permutation-free subset
    m2 = rs.map_g(hi.gbas, hi.gvec, lo.gvec)
    assert np.array_equal(hi.igvec[m2 - 1], lo.igvec)
    # a cell strained by 0.5 %: same integer set at the same cutoff maps fully
    st = opw.Plane(s.rbas * 1.005, 6.0, 24.0)
    m3 = rs.map_g(lo.gbas, lo.gvec, st.gvec)
    assert np.array_equal(lo.igvec[m3[m3 > 0] - 1], st.igvec[m3 > 0])


@pytest.mark.gpu
def test_gpu_restart_across_cutoffs(tmp_path):
    # a file written by a set at a lower cutoff, read into a set at a higher
    # one: the common plane waves carry over, the new ones are zero
    import pwgpu
    from harness import SYSTEMS
    s = SYSTEMS["si8"]()
    print("case restart si8 cutoff 6 -> 8 Ha", flush=True)
    plo = pwgpu.PlaneWave(s.rbas, 6.0, 24.0)
    wlo = pwgpu.Waves.from_system(plo, s)
    path = str(tmp_path / "lo.rstrt")
    wlo.write_restart(path, static=True)
    psi_lo = wlo.getc8a("PSI0")
    ig_lo = plo.geti4a("IGVEC")
    wlo.close()
    phi = pwgpu.PlaneWave(s.rbas, 8.0, 32.0)
    whi = pwgpu.Waves.from_system(phi, s)
    assert whi.read_restart(path)
    got = whi.getc8a("PSI0")
    key = {tuple(v): i for i, v in enumerate(ig_lo.tolist())}
    idx = np.array([key.get(tuple(v), -1) for v in phi.geti4a("IGVEC").tolist()])
    assert np.array_equal(got[:, idx >= 0], psi_lo[:, idx[idx >= 0]])
    assert np.all(got[:, idx < 0] == 0)
    assert np.array_equal(whi.getc8a("PSIM"), got)
