"""Restart files of the reference for the wave-function step.

The reference writes its restart file as a Fortran unformatted sequential
file (WRITERESTART / READRESTART, src/paw_ionew.f90:12-275): a HEADER
separator, one section per object (TIMESTEP, THERMOSTAT, ATOMS, CELL, DYNOCC,
WAVES, QMMM, COSMO), each opened by a separator record that carries the number
of records that follow, and an ENDOFFILE separator.  This module writes and
reads that file with the WAVES section only (WAVES$WRITE / WAVES$READ,
src/paw_waves2.f90:3058-3166); sections of other objects found on reading are
skipped by their record count, as RESTART$SKIP does (src/paw_report.f90:183-207).

Records (Fortran sequential: a 4-byte little-endian length before and after
each record's bytes; arrays in column-major order):

  separator  INTEGER(4) NREC, CHARACTER(32) ID, CHARACTER(32) NAME,
             CHARACTER(64) VERSION, CHARACTER(256) NOTES     (SEPARATOR_TYPE,
             a SEQUENCE type: 388 bytes, src/paw_report.f90:107-125)
  WAVES section (WAVES_WRITEPSI, src/paw_waves2.f90:3168-3350), NREC records:
    NKPT, NSPIN (int32), RBAS(3,3) (real64), NWAVE (int32)
    per k-point:
      KEY='PSI' (char 8), NGG, NDIM, NB, ILOGICAL (int32; bit 0 = super waves)
      XK(3) (real64, relative), IGVEC(3,NGG) (int32)
      for IWAVE = 1..NWAVE (PSI0, PSIM), ISPIN, IB = 1..NBH: PSI(NGG,NDIM) (complex128)
      KEY='LAMBDA' (char 8), NB, NDIM, NSPIN, NLAMBDA (int32)    (WAVES_WRITELAMBDA,
      for I = 1..NLAMBDA (RLAM0, RLAMM, ...), ISPIN: LAMBDA(NB,NB) (complex128)   :3920-4040)

NWAVE is 1 for RSTRTTYPE='STATIC' (PSI0 only, NLAMBDA=1) and 2 for 'DYNAMIC'.
Reading (WAVES_READPSI :5147-5513, WAVES_READLAMBDA :4041-4267) maps the
file's plane waves onto the present G set by their integer coordinates in the
file's reciprocal basis (WAVES_MAPG :4268-4399; absent ones are zero), copies
min(NBH, NBH') super bands and min(NSPIN, NSPIN') spins over the present
wave functions, sets PSIM = PSI0 when NWAVE = 1, and folds the multipliers
into the leading min(NB, NB') block of zeroed matrices.

Record lengths above 2**31 - 9 bytes would need the compiler's subrecord
convention; such a record (a single band of over 134 million plane waves) is
refused on writing and on reading.
"""
import struct

import numpy as np

SEPARATOR = struct.Struct("<i32s32s64s256s")
VERSION = "AUG1996"
_MAXREC = 2 ** 31 - 9


class RestartError(ValueError):
    """A restart file that does not fit the reference's format or the present
    wave-function set; the message is the reference's where it has one."""


# ----------------------------------------------------------------------------
# Fortran sequential unformatted records
# ----------------------------------------------------------------------------
class RecordWriter:
    def __init__(self, f):
        self.f = f

    def write(self, *parts):
        data = b"".join(parts)
        if len(data) > _MAXREC:
            raise RestartError(f"record of {len(data)} bytes exceeds the 4-byte record marker")
        m = struct.pack("<i", len(data))
        self.f.write(m + data + m)


class RecordReader:
    def __init__(self, f):
        self.f = f

    def read(self):
        """The next record's bytes, or None at the end of the file."""
        head = self.f.read(4)
        if len(head) == 0:
            return None
        if len(head) != 4:
            raise RestartError("truncated record marker")
        (n,) = struct.unpack("<i", head)
        if n < 0:
            raise RestartError("continued (sub)records are not supported")
        data = self.f.read(n)
        tail = self.f.read(4)
        if len(data) != n or len(tail) != 4 or struct.unpack("<i", tail)[0] != n:
            raise RestartError("record markers inconsistent: file is truncated or corrupted")
        return data

    def need(self):
        d = self.read()
        if d is None:
            raise RestartError("unexpected end of the restart file")
        return d


def _chars(s, n):
    b = s.encode("ascii")
    if len(b) > n:
        raise RestartError(f"{s!r} longer than CHARACTER({n})")
    return b.ljust(n, b" ")


def _str(b):
    return b.decode("ascii").rstrip(" ")


def pack_separator(nrec, ident, name="NONE", version=VERSION, notes="NONE"):
    return SEPARATOR.pack(int(nrec), _chars(ident, 32), _chars(name, 32), _chars(version, 64),
                          _chars(notes, 256))


def unpack_separator(data):
    if len(data) != SEPARATOR.size:
        raise RestartError("FORMAT OF RESTART FILE NOT RECOGNIZED")
    nrec, ident, name, version, notes = SEPARATOR.unpack(data)
    return {"nrec": nrec, "id": _str(ident), "name": _str(name), "version": _str(version),
            "notes": _str(notes)}


# ----------------------------------------------------------------------------
# WAVES section
# ----------------------------------------------------------------------------
class KPointWaves:
    """One k-point of the WAVES section.

    psi: (NWAVE, NSPIN, NBH, NDIM, NGG) complex128 -- record (IWAVE,ISPIN,IB)
         is PSI(NGG,NDIM); lam: (NLAMBDA, NSPIN, NB, NB) complex128 with
         lam[i, s, a, b] = LAMBDA(a+1, b+1).
    """

    def __init__(self, nb, tsuper, xk, igvec, psi, lam, ndim=1, key="PSI"):
        self.key = key
        self.nb = int(nb)
        self.ndim = int(ndim)
        self.tsuper = bool(tsuper)
        self.xk = np.asarray(xk, dtype=np.float64).reshape(3)
        self.igvec = np.ascontiguousarray(igvec, dtype=np.int32).reshape(-1, 3)
        self.psi = np.asarray(psi, dtype=np.complex128)
        self.lam = np.asarray(lam, dtype=np.complex128)
        self.ngg = self.igvec.shape[0]
        self.nbh = (self.nb + 1) // 2 if self.tsuper else self.nb
        if self.psi.ndim != 5 or self.psi.shape[2:] != (self.nbh, self.ndim, self.ngg):
            raise RestartError(f"psi shape {self.psi.shape} does not fit NBH={self.nbh} "
                               f"NDIM={self.ndim} NGG={self.ngg}")
        if self.lam.ndim != 4 or self.lam.shape[2:] != (self.nb, self.nb):
            raise RestartError(f"lambda shape {self.lam.shape} does not fit NB={self.nb}")

    @property
    def nwave(self):
        return self.psi.shape[0]

    @property
    def nlambda(self):
        return self.lam.shape[0]


def waves_nrec(kpoints, nspin):
    """Records of the WAVES section after its separator (WAVES_WRITEPSI with
    NREC=-1, src/paw_waves2.f90:3210-3232)."""
    n = 1
    for k in kpoints:
        n += 2 + k.nwave * k.nbh * nspin + 1 + nspin * k.nlambda
    return n


def write_waves(rec, rbas, kpoints, nspin):
    nwave = kpoints[0].nwave
    if any(k.nwave != nwave for k in kpoints):
        raise RestartError("all k-points must hold the same NWAVE")
    rb = np.asarray(rbas, dtype=np.float64).reshape(3, 3)
    rec.write(struct.pack("<ii", len(kpoints), nspin), rb.T.astype("<f8").tobytes(),
              struct.pack("<i", nwave))
    for k in kpoints:
        if k.psi.shape[1] != nspin or k.lam.shape[1] != nspin:
            raise RestartError("spin count of a k-point differs from NSPIN")
        rec.write(_chars(k.key, 8), struct.pack("<iiii", k.ngg, k.ndim, k.nb, 1 if k.tsuper else 0))
        rec.write(k.xk.astype("<f8").tobytes(), k.igvec.astype("<i4").tobytes())
        for iw in range(nwave):
            for s in range(nspin):
                for b in range(k.nbh):
                    rec.write(np.ascontiguousarray(k.psi[iw, s, b]).astype("<c16").tobytes())
        rec.write(_chars("LAMBDA", 8), struct.pack("<iiii", k.nb, k.ndim, nspin, k.nlambda))
        for i in range(k.nlambda):
            for s in range(nspin):
                rec.write(np.ascontiguousarray(k.lam[i, s].T).astype("<c16").tobytes())


def read_waves(rec):
    """The WAVES section after its separator: (NSPIN, RBAS, [KPointWaves])."""
    d = rec.need()
    if len(d) != 8 + 72 + 4:
        raise RestartError("WAVES section: first record has the wrong length")
    nkpt, nspin = struct.unpack("<ii", d[:8])
    rbas = np.frombuffer(d[8:80], dtype="<f8").reshape(3, 3).T.copy()
    (nwave,) = struct.unpack("<i", d[80:84])
    if nwave not in (1, 2):
        raise RestartError(f"NWAVE={nwave} on file: neither STATIC (1) nor DYNAMIC (2)")
    kpts = []
    for _ in range(nkpt):
        d = rec.need()
        key = _str(d[:8])
        if key != "PSI" or len(d) != 24:
            raise RestartError('ID IS NOT "PSI": FILE IS CORRUPTED')
        ngg, ndim, nb, ilogical = struct.unpack("<iiii", d[8:24])
        tsuper = bool(ilogical & 1)
        d = rec.need()
        if len(d) != 24 + 12 * ngg:
            raise RestartError("k-point / G-vector record has the wrong length")
        xk = np.frombuffer(d[:24], dtype="<f8").copy()
        igvec = np.frombuffer(d[24:], dtype="<i4").reshape(ngg, 3).copy()
        nbh = (nb + 1) // 2 if tsuper else nb
        psi = np.empty((nwave, nspin, nbh, ndim, ngg), dtype=np.complex128)
        for iw in range(nwave):
            for s in range(nspin):
                for b in range(nbh):
                    d = rec.need()
                    if len(d) != 16 * ngg * ndim:
                        raise RestartError("ERROR WHILE READING FROM FILE: wave-function record length")
                    psi[iw, s, b] = np.frombuffer(d, dtype="<c16").reshape(ndim, ngg)
        d = rec.need()
        if _str(d[:8]) != "LAMBDA" or len(d) != 24:
            raise RestartError('ID IS NOT "LAMBDA": FILE IS CORRUPTED')
        nb_l, ndim_l, nspin_l, nlambda = struct.unpack("<iiii", d[8:24])
        lam = np.empty((nlambda, nspin_l, nb_l, nb_l), dtype=np.complex128)
        for i in range(nlambda):
            for s in range(nspin_l):
                d = rec.need()
                if len(d) != 16 * nb_l * nb_l:
                    raise RestartError("ERROR WHILE READING FROM FILE: multiplier record length")
                lam[i, s] = np.frombuffer(d, dtype="<c16").reshape(nb_l, nb_l).T
        if nspin_l != nspin or nb_l != nb:
            raise RestartError("LAMBDA record does not match the wave functions of its k-point")
        kpts.append(KPointWaves(nb, tsuper, xk, igvec, psi, lam, ndim=ndim, key=key))
    return nspin, rbas, kpts


# ----------------------------------------------------------------------------
# whole file
# ----------------------------------------------------------------------------
class RestartWaves:
    """What a restart file holds for the wave-function step."""

    def __init__(self, nspin, rbas, kpoints, skipped=()):
        self.nspin = nspin
        self.rbas = rbas
        self.kpoints = kpoints
        self.skipped = list(skipped)

    @property
    def nwave(self):
        return self.kpoints[0].nwave if self.kpoints else 0


def write_file(path, rbas, kpoints, nspin=1):
    """HEADER, the WAVES section and ENDOFFILE (WRITERESTART with only the
    wave functions, src/paw_ionew.f90:154-275)."""
    with open(path, "wb") as f:
        rec = RecordWriter(f)
        rec.write(pack_separator(0, "HEADER", version=VERSION, notes="DEFAULT HEADER"))
        rec.write(pack_separator(waves_nrec(kpoints, nspin), "WAVES"))
        write_waves(rec, rbas, kpoints, nspin)
        rec.write(pack_separator(0, "ENDOFFILE", notes="DEFAULT EOF"))


def read_file(path):
    """The WAVES section of a restart file (READRESTART, src/paw_ionew.f90:12-152):
    the file must open with the HEADER separator of version AUG1996; other
    sections are skipped by their record count.  Returns a RestartWaves, or
    None when the file holds no WAVES section."""
    with open(path, "rb") as f:
        rec = RecordReader(f)
        d = rec.read()
        if d is None or len(d) != SEPARATOR.size:
            raise RestartError("FORMAT OF RESTART FILE NOT RECOGNIZED")
        head = unpack_separator(d)
        if head["id"] != "HEADER":
            raise RestartError("FORMAT OF RESTART FILE NOT RECOGNIZED")
        if head["version"] != VERSION:
            raise RestartError(f"restart file version {head['version']!r} is not {VERSION!r}")
        out, skipped = None, []
        while True:
            d = rec.read()
            if d is None:
                raise RestartError("restart file ends without the ENDOFFILE separator")
            sep = unpack_separator(d)
            if sep["id"] == "ENDOFFILE":
                break
            if sep["id"] == "WAVES" and sep["name"] == "NONE" and out is None:
                nspin, rbas, kpts = read_waves(rec)
                if waves_nrec(kpts, nspin) != sep["nrec"]:
                    raise RestartError("#(RECORDS) OF THE WAVES SECTION DIFFERS FROM ITS SEPARATOR")
                out = RestartWaves(nspin, rbas, kpts)
                continue
            if sep["nrec"] < 0 or sep["nrec"] > 10000000:
                raise RestartError(f"FILE STRUCTURE CORRUPTED: {sep['id']} NREC={sep['nrec']}")
            for _ in range(sep["nrec"]):
                rec.need()
            skipped.append(sep["id"])
        if out is not None:
            out.skipped = skipped
        return out


# ----------------------------------------------------------------------------
# G-vector mapping
# ----------------------------------------------------------------------------
def gbass(rbas):
    """Reciprocal lattice vectors (columns) of RBAS: GBAS^T RBAS = 2 pi."""
    return 2.0 * np.pi * np.linalg.inv(np.asarray(rbas, dtype=np.float64)).T


def _nint(x):
    # Fortran NINT: halves round away from zero
    return (np.sign(x) * np.floor(np.abs(x) + 0.5)).astype(np.int64)


def map_g(gbas_a, gvec_a, gvec_b):
    """WAVES_MAPG (src/paw_waves2.f90:4268-4399): for each present G vector
    (rows of gvec_b) the 1-based index of the file's vector (rows of gvec_a,
    reciprocal basis gbas_a) on the same point of the file's integer grid,
    0 where the file has none."""
    ginv = np.linalg.inv(np.asarray(gbas_a, dtype=np.float64))
    xa = np.asarray(gvec_a, dtype=np.float64) @ ginv.T
    xk = xa[0] - _nint(xa[0])
    iv = _nint(xa - xk)
    if np.any(np.sum((xa - xk - iv) ** 2, axis=1) > 1e-6):
        raise RestartError("G-VECTORS CANNOT BE REPRESENTED AS SUM OF K-VECTOR AND "
                           "RECIPROCAL LATTICE VECTORS: FAILED TO MAP G-VECTORS")
    lo, hi = iv.min(axis=0), iv.max(axis=0)
    dims = hi - lo + 1
    grid = np.zeros(int(np.prod(dims)), dtype=np.int64)
    flat = np.ravel_multi_index((iv - lo).T, dims)
    if np.unique(flat).size != flat.size:
        raise RestartError("TWO G-VECTORS ARE MAPPED ONTO THE SAME POINT")
    grid[flat] = np.arange(1, len(iv) + 1)
    xb = np.asarray(gvec_b, dtype=np.float64) @ ginv.T - xk
    ib = _nint(xb)
    inside = np.all((ib >= lo) & (ib <= hi), axis=1)
    out = np.zeros(len(ib), dtype=np.int64)
    out[inside] = grid[np.ravel_multi_index((ib[inside] - lo).T, dims)]
    return out


# ----------------------------------------------------------------------------
# a GPU wave-function set <-> restart file
# ----------------------------------------------------------------------------
def relative_k(plane):
    return np.linalg.solve(gbass(plane.rbas), plane.kvec)


def save(waves, path, static=False):
    """WAVES$WRITE of a pwgpu.Waves (RSTRTTYPE 'STATIC' when static): PSI0 (and
    PSIM), RLAM0 (and RLAMM) through the getters of the C ABI."""
    plane = waves.plane
    psi = [waves.getc8a("PSI0")]
    if not static:
        psi.append(waves.getc8a("PSIM"))
    nlambda = 1 if static else waves.geti4("NLAMBDA")
    lam = [waves.getc8a(key) for key in ("RLAM0", "RLAMM")[:nlambda]]
    k = KPointWaves(waves.nb, plane.tinv, relative_k(plane), plane.geti4a("IGVEC"),
                    np.stack(psi)[:, None, :, None, :],
                    np.array(lam, dtype=np.complex128).reshape(nlambda, 1, waves.nb, waves.nb))
    write_file(path, plane.rbas, [k], nspin=1)


def restore(restart, gvec, psi0, psim, nb, tsuper=True, xk=(0.0, 0.0, 0.0)):
    """The wave functions and multipliers of a restart file on the present
    set (WAVES_READPSI / WAVES_READLAMBDA): gvec (NGL,3) present G vectors,
    psi0 / psim (NBH,NGL) present wave functions (kept for bands the file does
    not hold), nb present bands.  Returns (psi0, psim, [RLAM0, RLAMM][:NLAMBDA])."""
    r = restart
    if len(r.kpoints) != 1:
        raise RestartError(f"KPOINT SETS ARE FIXED: NKPT ON FILE {len(r.kpoints)}, EXPECTED 1")
    k = r.kpoints[0]
    xk = np.asarray(xk, dtype=np.float64)
    if np.sum(np.abs(k.xk - xk)) > 1e-8:
        raise RestartError(f"INCONSISTENT K-POINTS: {k.xk} ON FILE, {xk} EXPECTED")
    if k.tsuper != bool(tsuper):
        raise RestartError("TRANSFORMATION BETWEEN REGULAR AND SUPER WAVE FUNCTIONS NOT IMPLEMENTED")
    if k.ndim != 1:
        raise RestartError("TRANSFORMATION BETWEEN SCALAR AND SPINOR WAVE FUNCTIONS NOT IMPLEMENTED")
    gbas_f = gbass(r.rbas)
    gvec_f = (k.igvec.astype(np.float64) + k.xk) @ gbas_f.T
    mapg = map_g(gbas_f, gvec_f, gvec)
    have = mapg > 0
    out = [np.array(psi0, dtype=np.complex128, copy=True), np.array(psim, dtype=np.complex128, copy=True)]
    nbh = min(out[0].shape[0], k.nbh)
    for iw in range(k.nwave):
        cur = out[iw]
        cur[:nbh] = 0.0
        cur[:nbh, have] = k.psi[iw, 0, :nbh, 0][:, mapg[have] - 1]
    if k.nwave == 1:
        out[1][:nbh] = out[0][:nbh]
    nba = min(nb, k.nb)
    lams = []
    for i in range(min(k.nlambda, 2)):
        lam = np.zeros((nb, nb), dtype=np.complex128)
        lam[:nba, :nba] = k.lam[i, 0, :nba, :nba]
        lams.append(lam)
    return out[0], out[1], lams


def load(waves, path):
    """WAVES$READ into a pwgpu.Waves: returns False when the file has no WAVES
    section, else True after PSI0, PSIM and the multipliers have been set.
    The diagonal of RLAM0 read (the reference hands it to DYNOCC as
    EPSILON) is left in waves.restart_epsilon."""
    r = read_file(path)
    if r is None:
        return False
    plane = waves.plane
    psi0, psim, lams = restore(r, plane.getr8a("GVEC"), waves.getc8a("PSI0"), waves.getc8a("PSIM"),
                               waves.nb, plane.tinv, relative_k(plane))
    waves.setc8a("PSI0", psi0)
    waves.setc8a("PSIM", psim)
    for key, lam in zip(("RLAM0", "RLAMM"), lams):
        waves.set_matrix_c(key, lam)
    waves.restart_epsilon = np.real(np.diag(lams[0])) if lams else None
    return True
