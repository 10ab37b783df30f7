"""Oracle restatement of the PLANEWAVE object (cp-paw src/paw_fft.f90) and of
the grid helpers it uses (src/paw_generalpurpose.f90, src/paw_library.f90).

Test infrastructure only; see oracle/__init__.py.
"""
import numpy as np

PI = 4.0 * np.arctan(1.0)


def gbass(rbas):
    """Reciprocal lattice vectors and cell volume.

    Follows src/paw_generalpurpose.f90:72-110 (GBASS).  ``rbas`` is the Fortran
    RBAS(3,3) with lattice vectors as columns, given here as a (3,3) array
    indexed rbas[i, j] = RBAS(i+1, j+1).  Returns (gbas, vol) with gbas[i, j] =
    GBAS(i+1, j+1), i.e. reciprocal vectors as columns, gbas^T rbas = 2 pi 1.
    """
    r = np.asarray(rbas, dtype=np.float64)
    g = np.empty((3, 3))
    g[0, 0] = r[1, 1] * r[2, 2] - r[2, 1] * r[1, 2]
    g[1, 0] = r[2, 1] * r[0, 2] - r[0, 1] * r[2, 2]
    g[2, 0] = r[0, 1] * r[1, 2] - r[1, 1] * r[0, 2]
    g[0, 1] = r[1, 2] * r[2, 0] - r[2, 2] * r[1, 0]
    g[1, 1] = r[2, 2] * r[0, 0] - r[0, 2] * r[2, 0]
    g[2, 1] = r[0, 2] * r[1, 0] - r[1, 2] * r[0, 0]
    g[0, 2] = r[1, 0] * r[2, 1] - r[2, 0] * r[1, 1]
    g[1, 2] = r[2, 0] * r[0, 1] - r[0, 0] * r[2, 1]
    g[2, 2] = r[0, 0] * r[1, 1] - r[1, 0] * r[0, 1]
    det = r[0, 0] * g[0, 0] + r[1, 0] * g[1, 0] + r[2, 0] * g[2, 0]
    if det == 0.0:
        raise ValueError("GBASS: lattice vectors linearly dependent")
    g *= 2.0 * PI / det
    return g, abs(det)


def boxsph(rbas, x0, y0, z0, rmax):
    """Index box around a sphere; src/paw_generalpurpose.f90:115-201 (BOXSPH)."""
    t1, t2, t3 = (np.asarray(rbas, dtype=np.float64)[:, k] for k in range(3))
    g1 = np.cross(t2, t3)
    g2 = np.cross(t3, t1)
    g3 = np.cross(t1, t2)
    det = float(np.dot(t1, g1))
    if det == 0.0:
        raise ValueError("BOXSPH: displacement vectors linearly dependent")
    g1, g2, g3 = g1 / det, g2 / det, g3 / det
    xn1 = np.sqrt(np.dot(g1, g1)) * rmax
    xn2 = np.sqrt(np.dot(g2, g2)) * rmax
    xn3 = np.sqrt(np.dot(g3, g3)) * rmax
    r0 = np.array([x0, y0, z0])
    xp0, yp0, zp0 = np.dot(g1, r0), np.dot(g2, r0), np.dot(g3, r0)
    ibig = 1000000
    # Fortran INT truncates toward zero; the +IBIG shift keeps arguments positive
    mn = [int(v - xn + 1.0 + ibig) - ibig for v, xn in ((xp0, xn1), (yp0, xn2), (zp0, xn3))]
    mx = [int(v + xn + ibig) - ibig for v, xn in ((xp0, xn1), (yp0, xn2), (zp0, xn3))]
    return mn[0], mx[0], mn[1], mx[1], mn[2], mx[2]


def _allowed_fft_lengths():
    """Allowed FFT lengths 2^h 3^i 5^j 7^k 11^m; src/paw_library.f90:2336-2377."""
    out = []
    for h in range(1, 26):
        for i in range(0, 3):
            for j in range(0, 2):
                for k in range(0, 2):
                    for m in range(0, 2):
                        if len(out) >= 300:
                            return sorted(out)
                        v = 2.0 ** h * 3.0 ** i * 5.0 ** j * 7.0 ** k * 11.0 ** m
                        if v > 37748736.0:
                            continue
                        out.append(int(round(v)))
    return sorted(out)


_IFR = _allowed_fft_lengths()


def fftadjustgrd(nr):
    """Smallest allowed length >= nr, skipping the first list entry.

    src/paw_library.f90:2379-2393 (the search starts at I=2).
    """
    for v in _IFR[1:]:
        if v >= nr:
            return v
    raise ValueError("FFTADJUSTGRD: requested gridpoint out of range")


def divider_grid(epw, rbas):
    """Real-space grid for cutoff epw (Hartree); src/paw_fft.f90:823-869.

    One task: NR1START=1, NR1L=NR1.  Returns (nr1, nr2, nr3).
    """
    gbas, _ = gbass(rbas)
    rad = np.sqrt(2.0 * epw) + 1.0e-6
    mn1, mx1, mn2, mx2, mn3, mx3 = boxsph(gbas, 0.0, 0.0, 0.0, rad)
    return (fftadjustgrd(mx1 - mn1 + 1), fftadjustgrd(mx2 - mn2 + 1),
            fftadjustgrd(mx3 - mn3 + 1))


def getig(gbas, epw, kvec=(0.0, 0.0, 0.0)):
    """G vectors in the sphere 0.5|G|^2 <= epw in the reference order.

    src/paw_fft.f90:1211-1325 (PLANEWAVE_GETIG): enumeration i,j,k nested,
    sorted by |G|^2 (heap sort, order of ties irrelevant because of the next
    step), then inside each run of |G|^2 within 1e-5 of the run's first
    element, sorted lexicographically by (i1,i2,i3).  The reference closes the
    last run at NG even if its last element differs (src/paw_fft.f90:1298-1301);
    that quirk is reproduced.  Returns igvec as an (NG,3) int32 array.
    """
    gbas = np.asarray(gbas, dtype=np.float64)
    kartk = gbas @ np.asarray(kvec, dtype=np.float64)
    rad = np.sqrt(2.0 * epw) + 1.0e-6
    mn1, mx1, mn2, mx2, mn3, mx3 = boxsph(gbas, -kartk[0], -kartk[1], -kartk[2], rad)
    i, j, k = np.meshgrid(np.arange(mn1, mx1 + 1), np.arange(mn2, mx2 + 1),
                          np.arange(mn3, mx3 + 1), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()      # k fastest, as the DO loops
    t1, t2, t3 = i.astype(np.float64), j.astype(np.float64), k.astype(np.float64)
    g1 = kartk[0] + gbas[0, 0] * t1 + gbas[0, 1] * t2 + gbas[0, 2] * t3
    g2 = kartk[1] + gbas[1, 0] * t1 + gbas[1, 1] * t2 + gbas[1, 2] * t3
    g3 = kartk[2] + gbas[2, 0] * t1 + gbas[2, 1] * t2 + gbas[2, 2] * t3
    gsq = g1 ** 2 + g2 ** 2 + g3 ** 2
    keep = ~(0.5 * gsq > epw)
    ig = np.stack([i[keep], j[keep], k[keep]], axis=1)
    g2a = gsq[keep]
    order = np.argsort(g2a, kind="stable")
    ig, g2a = ig[order], g2a[order]
    ng = len(g2a)
    out = ig.copy()
    istart = 0
    while istart < ng:
        svar = g2a[istart]
        stop = istart + 1
        while stop < ng and abs(g2a[stop] - svar) < 1.0e-5:
            stop += 1
        # reference quirk: a run that reaches the second-to-last element also
        # swallows the last one
        if stop == ng - 1:
            stop = ng
        run = out[istart:stop]
        lex = np.lexsort((run[:, 2], run[:, 1], run[:, 0]))
        out[istart:stop] = run[lex]
        istart = stop
    return out.astype(np.int32)


def minusg(igvec, kvec=(0.0, 0.0, 0.0)):
    """Index of -G for each G (0-based); src/paw_fft.f90:1157-1209.

    The reference searches IG2 >= IG1 for IK - IG(:,IG1) with IK=-NINT(2K).
    """
    ik = -np.rint(2.0 * np.asarray(kvec)).astype(np.int64)
    ng = len(igvec)
    key = {tuple(v): n for n, v in enumerate(igvec.tolist())}
    out = np.empty(ng, dtype=np.int32)
    for n in range(ng):
        m = key.get((ik[0] - int(igvec[n, 0]), ik[1] - int(igvec[n, 1]),
                     ik[2] - int(igvec[n, 2])))
        if m is None:
            raise ValueError("MINUSG: inversion symmetric partner not found")
        out[n] = m
    return out


def gvectors(gbas, igvec, kvec=(0.0, 0.0, 0.0)):
    """Cartesian G+k; src/paw_fft.f90:1475-1509.  Returns (NG,3)."""
    gbas = np.asarray(gbas, dtype=np.float64)
    kartk = gbas @ np.asarray(kvec, dtype=np.float64)
    t = igvec.astype(np.float64)
    return kartk[None, :] + t @ gbas.T


def structurefactor(r, gbas, igvec):
    """exp(-i G.R) at the Gamma point; src/paw_fft.f90:1597-1696.

    The reference builds it from separable per-axis power recurrences; this
    restatement evaluates the same product of per-axis phases directly.
    """
    gbas = np.asarray(gbas, dtype=np.float64)
    gr = gbas.T @ np.asarray(r, dtype=np.float64)     # GR_k = GBAS(:,k).R
    ph = igvec.astype(np.float64) @ gr
    return np.exp(-1j * ph)


def invertg(minus, fofg):
    """conj(f(-G)); src/paw_fft.f90:1740-1773.  Works along the last axis."""
    return np.conj(fofg[..., minus])


class Plane:
    """The PLANEWAVE instance for one wave set (src/paw_fft.f90:872-1110).

    One task, Gamma point, TINV=.TRUE. (super wave functions), no TSUPER.
    """

    def __init__(self, rbas, epw_psi, epw_rho):
        self.rbas = np.asarray(rbas, dtype=np.float64)
        self.gbas, self.vol = gbass(self.rbas)
        self.nr1, self.nr2, self.nr3 = divider_grid(epw_rho, self.rbas)
        self.nrl = self.nr1 * self.nr2 * self.nr3
        self.igvec = getig(self.gbas, epw_psi)
        self.ngl = len(self.igvec)
        self.minusg = minusg(self.igvec)
        self.gvec = gvectors(self.gbas, self.igvec)
        self.g2 = np.sum(self.gvec ** 2, axis=1)
        self.gweight = self.vol
        self.rweight = self.vol / float(self.nrl)
        g0 = np.nonzero(self.minusg == np.arange(self.ngl))[0]
        self.ngamma = int(g0[0]) if len(g0) else -1
        # grid index of each sphere vector (src/paw_fft.f90:1063-1073, 2065-2069)
        self.i1 = np.mod(self.igvec[:, 0], self.nr1)
        self.i2 = np.mod(self.igvec[:, 1], self.nr2)
        self.i3 = np.mod(self.igvec[:, 2], self.nr3)

    # -- FFT ------------------------------------------------------------------
    def fft_gtor(self, fofg):
        """PLANEWAVE$FFT('GTOR'): f(r)=sum_G f(G) exp(+iGr), unnormalised.

        src/paw_fft.f90:1948-2177; sign/normalisation src/paw_library.f90:2398.
        The reference's pruned passes skip only lines that are identically
        zero, so the result equals the full 3D DFT of the scattered sphere.
        fofg: (..., NGL) -> (..., NR3, NR2, NR1).
        """
        fofg = np.asarray(fofg, dtype=np.complex128)
        lead = fofg.shape[:-1]
        grid = np.zeros(lead + (self.nr3, self.nr2, self.nr1), dtype=np.complex128)
        grid[..., self.i3, self.i2, self.i1] = fofg
        return np.fft.ifftn(grid, axes=(-3, -2, -1)) * float(self.nrl)

    def fft_rtog(self, fofr):
        """PLANEWAVE$FFT('RTOG'): f(G)=(1/N) sum_r f(r) exp(-iGr) on the sphere.

        src/paw_fft.f90:2180-2341; 1/N from src/paw_library.f90:2723-2726.
        """
        fofr = np.asarray(fofr, dtype=np.complex128)
        spec = np.fft.fftn(fofr, axes=(-3, -2, -1)) / float(self.nrl)
        return spec[..., self.i3, self.i2, self.i1]

    # -- scalar products -----------------------------------------------------
    def scalarproduct(self, f1, f2, mode=" "):
        """PLANEWAVE$SCALARPRODUCT, src/paw_fft.f90:2344-2422.

        f1: (N1, NGL), f2: (N2, NGL) -> MAT (N1, N2) as numpy (row=i1).
        ' ' / '=' : Omega * sum_G conj(f1) f2.
        '-'       : Omega * sum_G f1(G) f2(-G)  (conj of <f1|invertg(f2)>).
        """
        if mode in (" ", "="):
            mat = np.conj(f1) @ f2.T
        elif mode == "-":
            f2m = invertg(self.minusg, f2)
            mat = np.conj(np.conj(f1) @ f2m.T)
        else:
            raise ValueError(mode)
        return mat * self.gweight

    def addproduct(self, f1, f2, mat, mode=" "):
        """PLANEWAVE$ADDPRODUCT, src/paw_fft.f90:2425-2489.

        F1(:,i1) += sum_i2 F2(:,i2) MAT(i2,i1); f1: (N1,NGL), f2: (N2,NGL),
        mat: (N2, N1) numpy (row=i2).  '-' uses invertg(f2).
        """
        src = f2 if mode in (" ", "=") else invertg(self.minusg, f2)
        return f1 + mat.T @ src
