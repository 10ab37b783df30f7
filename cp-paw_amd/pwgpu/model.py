"""Seeded model systems for the CP wavefunction step.

The reference obtains the local potential V(r) from POTENTIAL$VOFRHO
(src/paw_potential.f90:324) and the one-centre matrices DH/DO from
AUGMENTATION$SPHERE (src/paw_augmentation.f90:223) and SETUP (src/paw_setups.f90).
Those producers sit outside the wavefunction step.  The systems below feed the
step with a fixed local potential (a periodic sum of Gaussian wells), analytic
Gaussian projector functions (2 s + 2 p per site, as the built-in Si setup's
NPRO 2 2 in tests/fulltests/si2) and fixed DO/DH matrices.  With these inputs the
total energy is the linear functional E = sum_n f_n <psi_n|H|psi_n>, which is
what the step's ETOT/FORCE/HPSI compute.

All arrays follow the reference's column-major layouts expressed as numpy C
order with reversed dimensions (PSI(NGL,NBH) -> psi[nbh, ngl]).
"""
import math
from dataclasses import dataclass, field

import numpy as np

PI = math.pi
RY = 0.5           # Hartree per Rydberg, src/paw_constants.f90:238


@dataclass
class Species:
    lox: list                    # angular momentum of each radial projector
    sigma: list                  # Gaussian width of each radial projector (bohr)
    do_ln: np.ndarray            # (LNX,LNX) one-centre overlap DO
    dh_ln: np.ndarray            # (LNX,LNX) one-centre Hamiltonian (spherical part)
    # model setup data of the density -> potential step (Gaussian forms)
    zval: float = 4.0            # valence charge
    rcsm: float = 1.0            # width of the compensation Gaussian G0
    rcore: float = 0.9           # width of the pseudo core density
    qcore: float = 0.3           # charge of the pseudo core density
    vbar: float = -0.4           # depth of V-bar (Hartree)
    rvbar: float = 1.1           # width of V-bar
    # model setup data of the one-centre terms (AUGMENTATION$SPHERE)
    aez: float = 14.0            # atomic number
    mass: float = 28.0855        # atomic mass (u), fixes the nuclear radius
    rcore_ae: float = 0.45       # width of the all-electron core density
    ae_bump: float = 0.6         # size of the all-electron partial waves' inner part
    lrhox: int = None            # LRHOX of the structure file (default 2 max(LOX))

    @property
    def lnx(self):
        return len(self.lox)

    @property
    def lmnx(self):
        return sum(2 * l + 1 for l in self.lox)

    @property
    def lmrx(self):
        """LMRX = (LRHOX+1)^2 with LRHOX = min(2 max(LOX), LRHOX of the
        structure file) (src/paw_setups.f90:1223-1230)."""
        lr = 2 * max(self.lox)
        if self.lrhox is not None:
            lr = min(lr, self.lrhox)
        return (lr + 1) ** 2

    def sphere_setup(self, r1=1.0e-4, dex=0.05, nr=250):
        """Setup data of AUGMENTATION$SPHERE on the shifted logarithmic grid
        r_i = r1 (exp(dex (i-1)) - 1), as a dict of host arrays (Gaussian
        model forms; the reference builds them in its setup construction):
          PSPHI_ln = N r^l exp(-r^2/(2 s^2)) for the projector widths s,
          AEPHI_ln = PSPHI_ln (1 + ae_bump exp(-r^2/(0.35 s)^2)),
          AECORE Y0 = (aez - zval) Gaussian of width rcore_ae,
          PSCORE Y0 = qcore Gaussian of width rcore   (as PSCORG),
          VADD Y0   = vbar exp(-r^2/rvbar^2)           (as VBARG),
          DEKIN     = 0.8 DH (the spherical one-centre Hamiltonian), DO,
          RNUC      = mass^(1/3) 1.85635065215e-6 (src/paw_periodictable.f90:95-97,274)."""
        y0 = 1.0 / math.sqrt(4.0 * PI)
        xexp = math.exp(dex)
        ri = np.empty(nr)
        v = r1 / xexp
        for i in range(nr):
            v *= xexp
            ri[i] = v
        r = ri - r1
        r[0] = 1.0e-9
        phi_ps, phi_ae = [], []
        for l, s in zip(self.lox, self.sigma):
            norm = 1.0 / math.sqrt(math.gamma(l + 1.5) * s ** (2 * l + 3) / 2.0)
            p = norm * r ** l * np.exp(-r * r / (2.0 * s * s))
            phi_ps.append(p)
            phi_ae.append(p * (1.0 + self.ae_bump * np.exp(-r * r / (0.35 * s) ** 2)))
        zc = self.aez - self.zval

        def gauss(q, w):
            return q * np.exp(-r * r / (w * w)) / (PI * w * w) ** 1.5 / y0

        return {"r1": r1, "dex": dex, "nr": nr, "aez": self.aez, "rnuc": self.mass ** (1.0 / 3.0) * 1.85635065215e-6,
                "lox": list(self.lox), "aephi": np.array(phi_ae), "psphi": np.array(phi_ps),
                "aecore": gauss(zc, self.rcore_ae), "pscore": gauss(self.qcore, self.rcore),
                "rcsm": self.rcsm, "vadd": self.vbar * np.exp(-r * r / self.rvbar ** 2) / y0,
                "dover": np.asarray(self.do_ln, dtype=np.float64),
                "dekin": 0.8 * np.asarray(self.dh_ln, dtype=np.float64), "lmrx": self.lmrx}

    def partial_wave_overlaps(self):
        """DO(LN,LN') = <AEPHI|AEPHI'> - <PSPHI|PSPHI'> for equal l (zero
        otherwise) of the model partial waves, by a fine trapezoid rule; with
        it as DO the one-centre charge makes the cell neutral (RHOB ~ 0)."""
        r = np.linspace(0.0, 30.0, 300001)
        n = len(self.lox)
        out = np.zeros((n, n))
        ph = []
        for l, s in zip(self.lox, self.sigma):
            norm = 1.0 / math.sqrt(math.gamma(l + 1.5) * s ** (2 * l + 3) / 2.0)
            p = norm * r ** l * np.exp(-r * r / (2.0 * s * s))
            ph.append((p, p * (1.0 + self.ae_bump * np.exp(-r * r / (0.35 * s) ** 2))))
        for i in range(n):
            for j in range(n):
                if self.lox[i] == self.lox[j]:
                    f = (ph[i][1] * ph[j][1] - ph[i][0] * ph[j][0]) * r * r
                    out[i, j] = float(np.sum(0.5 * (f[1:] + f[:-1])) * (r[1] - r[0]))
        return out


def paw_system(system):
    """A copy of a system whose species' DO are their partial-wave overlap
    differences (Species.partial_wave_overlaps), for runs with the one-centre
    terms on the device."""
    import copy
    s = copy.deepcopy(system)
    for sp in s.species:
        sp.do_ln = sp.partial_wave_overlaps()
    return s


@dataclass
class System:
    name: str
    rbas: np.ndarray             # (3,3), lattice vectors as columns (RBAS)
    pos: np.ndarray              # (NAT,3) cartesian positions R(0)
    isp: np.ndarray              # (NAT,) 0-based species index
    species: list
    nb: int
    occ: np.ndarray              # (NB,)
    epw_psi: float               # Hartree
    epw_rho: float               # Hartree
    wells: list = field(default_factory=list)   # (center(3), depth, width)
    dt: float = 5.0
    emass: float = 1000.0
    emasscg2: float = None
    annee: float = 0.0
    seed: int = 0
    dh_noise: float = 0.02

    def __post_init__(self):
        if self.emasscg2 is None:
            # reference default, src/paw_ioroutines.f90:1405-1414
            self.emasscg2 = 0.5 * (10.0 / (2.0 * PI)) ** 2 * self.dt ** 2 / self.emass

    # -- derived sizes --------------------------------------------------------
    @property
    def nat(self):
        return len(self.pos)

    @property
    def nbh(self):
        return (self.nb + 1) // 2

    @property
    def vol(self):
        return abs(float(np.linalg.det(self.rbas)))

    @property
    def lox_list(self):
        return [s.lox for s in self.species]

    @property
    def npro(self):
        return int(sum(self.species[s].lmnx for s in self.isp))

    def do_list(self):
        return [np.asarray(s.do_ln, dtype=np.float64) for s in self.species]

    # -- tables the wave-function step consumes ------------------------------
    def pro_table(self, g2):
        """Bare projector tables PRO(G,LN)/Omega as (NBAREPRO, NG).

        Radial projector p(r)=N r^l exp(-r^2/(2 s^2)), normalised to
        int p^2 r^2 dr = 1; its Bessel transform gives
        4 pi N sqrt(pi/2) s^(2l+3) |G|^l exp(-G^2 s^2/2) (table / Omega, as
        SETUP$GETFOFG, src/paw_setups.f90:1080).
        """
        g = np.sqrt(np.asarray(g2, dtype=np.float64))
        rows = []
        for sp in self.species:
            for l, s in zip(sp.lox, sp.sigma):
                norm = 1.0 / math.sqrt(math.gamma(l + 1.5) * s ** (2 * l + 3) / 2.0)
                rows.append(4.0 * PI * norm * math.sqrt(PI / 2.0) * s ** (2 * l + 3)
                            * g ** l * np.exp(-0.5 * g * g * s * s) / self.vol)
        return np.array(rows)

    def dh_matrices(self):
        """Per-atom DH (LMNX,LMNX) complex: spherical part + seeded symmetric noise."""
        rng = np.random.default_rng(self.seed + 7)
        out = []
        for iat in range(self.nat):
            sp = self.species[self.isp[iat]]
            n = sp.lmnx
            d = np.zeros((n, n))
            o1 = 0
            for ln1, l1 in enumerate(sp.lox):
                o2 = 0
                for ln2, l2 in enumerate(sp.lox):
                    if l1 == l2:
                        for m in range(2 * l1 + 1):
                            d[o1 + m, o2 + m] = sp.dh_ln[ln1, ln2]
                    o2 += 2 * l2 + 1
                o1 += 2 * l1 + 1
            noise = rng.normal(0.0, self.dh_noise, (n, n))
            d = d + 0.5 * (noise + noise.T)
            out.append(d.astype(np.complex128))
        return out

    def potential(self, nr):
        """V(r) on the (NR1,NR2,NR3) grid, returned as (NR3,NR2,NR1) float64.

        Periodic sum of wells depth*exp(-|r-c|^2/w^2), summed in reciprocal
        space: V(G) = sum_c depth (pi w^2)^1.5/Omega exp(-G^2 w^2/4) exp(-iGc).
        """
        nr1, nr2, nr3 = nr
        gb = 2.0 * PI * np.linalg.inv(self.rbas).T      # columns = b_k
        f1 = np.fft.fftfreq(nr1, 1.0 / nr1)
        f2 = np.fft.fftfreq(nr2, 1.0 / nr2)
        f3 = np.fft.fftfreq(nr3, 1.0 / nr3)
        g3, g2_, g1 = np.meshgrid(f3, f2, f1, indexing="ij")
        gx = g1 * gb[0, 0] + g2_ * gb[0, 1] + g3 * gb[0, 2]
        gy = g1 * gb[1, 0] + g2_ * gb[1, 1] + g3 * gb[1, 2]
        gz = g1 * gb[2, 0] + g2_ * gb[2, 1] + g3 * gb[2, 2]
        gsq = gx * gx + gy * gy + gz * gz
        vg = np.zeros((nr3, nr2, nr1), dtype=np.complex128)
        for c, depth, w in self.wells:
            c = np.asarray(c, dtype=np.float64)
            ph1 = np.exp(-1j * f1 * (gb[:, 0] @ c))
            ph2 = np.exp(-1j * f2 * (gb[:, 1] @ c))
            ph3 = np.exp(-1j * f3 * (gb[:, 2] @ c))
            amp = depth * (PI * w * w) ** 1.5 / self.vol * np.exp(-gsq * w * w / 4.0)
            vg += amp * ph3[:, None, None] * ph2[None, :, None] * ph1[None, None, :]
        v = np.fft.ifftn(vg) * float(nr1 * nr2 * nr3)
        return np.ascontiguousarray(v.real)

    # -- setup data of the density -> potential step (POTENTIAL$VOFRHO) --------
    def potential_tables(self, g2):
        """(4, NSP, NG) PSCORG, VBARG, G0, V0 on a sphere with |G|^2 = g2.

        Gaussian model forms (the reference obtains them from its setup
        construction, src/paw_setups.f90): with Y0 = 1/sqrt(4 pi),
          PSCORG Y0 = qcore exp(-G^2 rcore^2/4)/Omega   (core charge qcore),
          VBARG Y0  = vbar (pi rvbar^2)^1.5 exp(-G^2 rvbar^2/4)/Omega,
          G0        = 4 pi exp(-G^2 rcsm^2/4)/Omega      (the Gaussian of unit moment
                      C_0 exp(-r^2/rcsm^2) Y0 of GAUSSN, in the reference's normalisation:
                      a moment QLM(1) carries the charge QLM(1)/Y0),
          V0        = (4 pi)^2/G^2 (exp(-G^2 r1^2/4) - exp(-G^2 rcsm^2/4))/Omega,
                      r1 = 0.6 rcsm, with its limit 4 pi^2 (rcsm^2 - r1^2)/Omega at G = 0.
        """
        g2 = np.asarray(g2, dtype=np.float64)
        y0 = 1.0 / math.sqrt(4.0 * PI)
        om = self.vol
        out = np.zeros((4, len(self.species), len(g2)))
        nz = g2 > 1.0e-12
        for k, sp in enumerate(self.species):
            out[0, k] = sp.qcore * np.exp(-g2 * sp.rcore ** 2 / 4.0) / om / y0
            out[1, k] = sp.vbar * (PI * sp.rvbar ** 2) ** 1.5 * np.exp(-g2 * sp.rvbar ** 2 / 4.0) / om / y0
            out[2, k] = 4.0 * PI * np.exp(-g2 * sp.rcsm ** 2 / 4.0) / om
            r1 = 0.6 * sp.rcsm
            v0 = np.full(len(g2), 4.0 * PI * PI * (sp.rcsm ** 2 - r1 ** 2) / om)
            v0[nz] = (4.0 * PI) ** 2 / g2[nz] * (np.exp(-g2[nz] * r1 ** 2 / 4.0)
                                                 - np.exp(-g2[nz] * sp.rcsm ** 2 / 4.0)) / om
            out[3, k] = v0
        return out

    def qlm(self, lmrxx=1):
        """QLM(LMRXX,NAT) as (NAT,) for LMRXX = 1, else (NAT, LMRXX): the
        monopole -zval Y0 (charge -zval) plus a seeded deviation of 0.05 Y0,
        and seeded higher moments of 0.02 (so that every dE/dQLM is exercised)."""
        rng = np.random.default_rng(self.seed + 23)
        y0 = 1.0 / math.sqrt(4.0 * PI)
        z = np.array([self.species[s].zval for s in self.isp], dtype=np.float64)
        q0 = (-z + rng.normal(0.0, 0.05, self.nat)) * y0
        if lmrxx == 1:
            return q0
        out = rng.normal(0.0, 0.02, (self.nat, lmrxx))
        out[:, 0] = q0
        return out

    def valence_density(self, nr, width=1.2, background=2.0e-3):
        """Model pseudo valence density on the (NR1,NR2,NR3) grid as
        (NR3,NR2,NR1): periodic sum of zval-normalised Gaussians of the given
        width on the sites plus a constant background (keeps it above the
        functional's density floor)."""
        nr1, nr2, nr3 = nr
        gb = 2.0 * PI * np.linalg.inv(self.rbas).T
        f1 = np.fft.fftfreq(nr1, 1.0 / nr1)
        f2 = np.fft.fftfreq(nr2, 1.0 / nr2)
        f3 = np.fft.fftfreq(nr3, 1.0 / nr3)
        g3, g2_, g1 = np.meshgrid(f3, f2, f1, indexing="ij")
        gsq = sum((g1 * gb[k, 0] + g2_ * gb[k, 1] + g3 * gb[k, 2]) ** 2 for k in range(3))
        rg = np.zeros((nr3, nr2, nr1), dtype=np.complex128)
        for p, s in zip(self.pos, self.isp):
            ph1 = np.exp(-1j * f1 * (gb[:, 0] @ p))
            ph2 = np.exp(-1j * f2 * (gb[:, 1] @ p))
            ph3 = np.exp(-1j * f3 * (gb[:, 2] @ p))
            rg += self.species[s].zval / self.vol * np.exp(-gsq * width * width / 4.0) * \
                ph3[:, None, None] * ph2[None, :, None] * ph1[None, None, :]
        rho = np.fft.ifftn(rg).real * float(nr1 * nr2 * nr3)
        return np.ascontiguousarray(rho + background)

    def initial_psi(self, g2):
        """c(G) = (N(0,1)+iN(0,1))/(1+|G|^2)^2, seeded; (NBH, NG) complex128."""
        rng = np.random.default_rng(self.seed + 11)
        ng = len(g2)
        c = rng.normal(size=(self.nbh, ng)) + 1j * rng.normal(size=(self.nbh, ng))
        return (c / (1.0 + np.asarray(g2)[None, :]) ** 2).astype(np.complex128)


# ----------------------------------------------------------------------------
# concrete systems
# ----------------------------------------------------------------------------
def si_species():
    """Model Si site: 2 s + 2 p Gaussian projectors (LNX=4, LMNX=8)."""
    do = np.array([[0.35, 0.08, 0, 0], [0.08, 0.20, 0, 0],
                   [0, 0, 0.30, 0.05], [0, 0, 0.05, 0.15]], dtype=np.float64)
    dh = np.array([[-0.45, 0.10, 0, 0], [0.10, 0.25, 0, 0],
                   [0, 0, -0.25, 0.06], [0, 0, 0.06, 0.20]], dtype=np.float64)
    return Species(lox=[0, 0, 1, 1], sigma=[0.9, 1.3, 1.0, 1.4], do_ln=do, dh_ln=dh)


def si_species_d():
    """Model Si site with the projector set of the reference's si2 deck
    (NPRO= 2 2 1: 2 s + 2 p + 1 d, LNX=5, LMNX=13)."""
    sp = si_species()
    do = np.zeros((5, 5))
    dh = np.zeros((5, 5))
    do[:4, :4] = sp.do_ln
    dh[:4, :4] = sp.dh_ln
    do[4, 4] = 0.12
    dh[4, 4] = 0.35
    return Species(lox=[0, 0, 1, 1, 2], sigma=[0.9, 1.3, 1.0, 1.4, 1.1], do_ln=do, dh_ln=dh)


def species_from_npro(npro, base=None):
    """Model species with NPRO(l) radial projectors per angular momentum l
    (the !STRUCTURE!SPECIES NPRO keyword of the reference's decks): Gaussian
    radial functions of widths 0.9, 1.3, ... and block-diagonal DO/DH."""
    base = si_species() if base is None else base
    lox, sigma = [], []
    for l, n in enumerate(npro):
        for k in range(int(n)):
            lox.append(l)
            sigma.append(0.9 + 0.4 * k + 0.1 * l)
    nl = len(lox)
    do = np.zeros((nl, nl))
    dh = np.zeros((nl, nl))
    for i in range(nl):
        for j in range(nl):
            if lox[i] == lox[j]:
                k = abs(i - j)
                do[i, j] = 0.3 / (1 + lox[i]) if k == 0 else 0.05
                dh[i, j] = (-0.45 + 0.25 * lox[i] + 0.3 * (i - lox.index(lox[i]))) if k == 0 else 0.08
    return Species(lox=lox, sigma=sigma, do_ln=do, dh_ln=dh, zval=base.zval, rcsm=base.rcsm,
                   rcore=base.rcore, qcore=base.qcore, vbar=base.vbar, rvbar=base.rvbar)


def diamond_si(ncell=(1, 1, 1), a=10.26, nempty_per_cell=None, seed=1,
               disp_sigma=0.05, epw_psi_ry=30.0, epw_rho_ry=None, cdual=4.0,
               dt=5.0, emass=1000.0, nempty=None, species=None):
    """Diamond Si supercell of the 8-atom simple-cubic cell, Gamma point.

    16 doubly occupied bands per 8 atoms; ``nempty`` empty bands.
    Atoms displaced by seeded N(0, disp_sigma^2).
    """
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    frac = np.concatenate([base, base + 0.25])
    n1, n2, n3 = ncell
    rbas = np.diag([n1 * a, n2 * a, n3 * a]).astype(np.float64)
    pos = []
    for i in range(n1):
        for j in range(n2):
            for k in range(n3):
                for f in frac:
                    pos.append((np.array([i, j, k]) + f) * a)
    pos = np.array(pos, dtype=np.float64)
    rng = np.random.default_rng(seed)
    pos = pos + rng.normal(0.0, disp_sigma, pos.shape)
    nat = len(pos)
    nocc = 2 * nat              # 4 valence electrons per Si, doubly occupied
    if nempty is None:
        nempty = 4 if nat == 8 else nat // 4
    nb = nocc + nempty
    occ = np.array([2.0] * nocc + [0.0] * nempty)
    epw_psi = epw_psi_ry * RY
    epw_rho = (epw_rho_ry * RY) if epw_rho_ry is not None else cdual * epw_psi
    wells = [(p, -1.6, 1.5) for p in pos]
    return System(name=f"Si{nat}", rbas=rbas, pos=pos, isp=np.zeros(nat, dtype=np.int64),
                  species=[si_species() if species is None else species], nb=nb, occ=occ,
                  epw_psi=epw_psi, epw_rho=epw_rho, wells=wells, dt=dt, emass=emass, seed=seed)


def si8(**kw):
    """Si8 simple-cubic cell, a=10.26, 16 occupied + 4 empty, 30/60 Ry."""
    kw.setdefault("epw_rho_ry", 60.0)
    return diamond_si((1, 1, 1), nempty=4, **kw)


def si64(**kw):
    """Si64 = 2x2x2 supercell a=20.52, 128 occupied + 16 empty, 30 Ry, CDUAL=4."""
    return diamond_si((2, 2, 2), nempty=16, **kw)


def si_supercell(nat, **kw):
    """Si8/Si16/Si32/Si64 for the serving mix."""
    cells = {8: (1, 1, 1), 16: (1, 1, 2), 32: (1, 2, 2), 64: (2, 2, 2)}
    return diamond_si(cells[nat], nempty=max(4, nat // 4), **kw)


def synthetic(seed=3, nb=256, ncentre=64, nwell=32, a=20.0,
              epw_psi=9.45, epw_rho=45.0, dt=5.0, emass=1000.0):
    """Kernel-level synthetic step: ~11k plane waves on a 64^3 grid,
    256 bands, 512 projectors (64 centres x 2 s + 2 p), 32 Gaussian wells.
    All bands singly weighted f=2 for the lower half, 0 above."""
    rng = np.random.default_rng(seed)
    rbas = np.diag([a, a, a]).astype(np.float64)
    pos = rng.uniform(0.0, a, (ncentre, 3))
    sig = rng.uniform(0.8, 1.5, (ncentre, 4))
    # one species per centre would make NSP large; use 4 width classes
    species = []
    for k in range(4):
        sp = si_species()
        sp.sigma = [0.8 + 0.15 * k, 1.2 + 0.1 * k, 0.9 + 0.15 * k, 1.3 + 0.1 * k]
        species.append(sp)
    isp = np.asarray(np.floor(sig[:, 0] * 0 + rng.integers(0, 4, ncentre)), dtype=np.int64)
    wells = [(rng.uniform(0.0, a, 3), -rng.uniform(1.0, 3.0), rng.uniform(1.0, 2.0))
             for _ in range(nwell)]
    occ = np.array([2.0] * (nb // 2) + [0.0] * (nb - nb // 2))
    return System(name="synthetic", rbas=rbas, pos=pos, isp=isp, species=species,
                  nb=nb, occ=occ, epw_psi=epw_psi, epw_rho=epw_rho, wells=wells,
                  dt=dt, emass=emass, seed=seed)


def o_species():
    """Model O site: 2 s + 2 p Gaussian projectors (LNX=4, LMNX=8), more compact than Si."""
    do = np.array([[0.40, 0.10, 0, 0], [0.10, 0.22, 0, 0],
                   [0, 0, 0.34, 0.06], [0, 0, 0.06, 0.18]], dtype=np.float64)
    dh = np.array([[-0.80, 0.12, 0, 0], [0.12, 0.30, 0, 0],
                   [0, 0, -0.45, 0.08], [0, 0, 0.08, 0.25]], dtype=np.float64)
    return Species(lox=[0, 0, 1, 1], sigma=[0.55, 0.85, 0.6, 0.9], do_ln=do, dh_ln=dh,
                   zval=6.0, rcsm=0.7, rcore=0.5, qcore=0.2, vbar=-0.8, rvbar=0.7,
                   aez=8.0, mass=15.9994, rcore_ae=0.25)


def h_species():
    """Model H site: 2 s Gaussian projectors (LNX=2, LMNX=2)."""
    do = np.array([[0.25, 0.05], [0.05, 0.12]], dtype=np.float64)
    dh = np.array([[-0.35, 0.06], [0.06, 0.18]], dtype=np.float64)
    return Species(lox=[0, 0], sigma=[0.45, 0.75], do_ln=do, dh_ln=dh,
                   zval=1.0, rcsm=0.5, rcore=0.4, qcore=0.0, vbar=-0.3, rvbar=0.5,
                   aez=1.0, mass=1.00794, rcore_ae=0.3)


def water(nmol=32, a=19.7, seed=9, epw_psi_ry=40.0, cdual=4.0, dt=4.0, emass=1000.0,
          doo=4.5, doh=1.81, angle=104.5):
    """32 H2O in a cubic box (1 g/cm^3 at a=19.7 bohr): seeded random O positions
    with O-O >= doo (minimum image), O-H = doh, H-O-H = angle, random
    orientations; 4 doubly occupied bands per molecule, no empty bands."""
    rng = np.random.default_rng(seed)
    opos = []
    tries = 0
    while len(opos) < nmol:
        tries += 1
        if tries > 200000:
            raise RuntimeError("water: could not place the molecules")
        c = rng.uniform(0.0, a, 3)
        ok = True
        for o in opos:
            d = c - o
            d -= a * np.round(d / a)
            if d @ d < doo * doo:
                ok = False
                break
        if ok:
            opos.append(c)
    half = math.radians(angle) / 2.0
    pos, isp = [], []
    for o in opos:
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        v = rng.normal(size=3)
        v -= (v @ u) * u
        v /= np.linalg.norm(v)
        pos.append(o)
        isp.append(0)
        for sgn in (1.0, -1.0):
            pos.append(o + doh * (math.cos(half) * u + sgn * math.sin(half) * v))
            isp.append(1)
    pos = np.array(pos) % a
    nb = 4 * nmol
    wells = [(p, -3.0 if k == 0 else -1.0, 1.0 if k == 0 else 0.8) for p, k in zip(pos, isp)]
    epw_psi = epw_psi_ry * RY
    return System(name=f"H2O{nmol}", rbas=np.diag([a, a, a]).astype(np.float64), pos=pos,
                  isp=np.array(isp, dtype=np.int64), species=[o_species(), h_species()], nb=nb,
                  occ=np.full(nb, 2.0), epw_psi=epw_psi, epw_rho=cdual * epw_psi, wells=wells, dt=dt,
                  emass=emass, seed=seed)


def tiny(seed=5, nb=8, a=7.0, epw_psi=3.0, epw_rho=12.0, nat=2):
    """Tiny cell for fast CPU parity tests (two sites, small grid)."""
    rng = np.random.default_rng(seed)
    rbas = np.array([[a, 0.3, 0.0], [0.0, a * 0.95, 0.2], [0.1, 0.0, a * 1.05]]).T
    pos = rng.uniform(0.0, a, (nat, 3))
    wells = [(p, -1.2, 1.2) for p in pos]
    occ = np.array([2.0] * (nb - 2) + [1.0, 0.0])
    return System(name="tiny", rbas=rbas, pos=pos, isp=np.zeros(nat, dtype=np.int64),
                  species=[si_species()], nb=nb, occ=occ, epw_psi=epw_psi,
                  epw_rho=epw_rho, wells=wells, seed=seed)


SYSTEMS = {
    "si8": si8,
    "si64": si64,
    "synthetic": synthetic,
    "h2o32": water,
    "tiny": tiny,
}
