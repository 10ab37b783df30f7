"""Reader of the reference's control (.cntl) and structure (.strc) input decks
for the quantities of the Car-Parrinello wave-function step.

The file format is the reference's linked-list block format
(src/paw_linkedlist.f90, read by READIN/STRCIN in src/paw_ioroutines.f90):
blocks open with ``!NAME`` and close with ``!END``, ``!EOB`` ends the
buffer; inside a block ``KEY= v1 v2 ...`` assigns one or more values
(numbers, 'quoted strings', logicals T/F); keys may repeat across sibling
blocks (several ``!ATOM``).

What is taken from the decks, with the reference's units and defaults:
  * !CONTROL!FOURIER EPWPSI [Ry] (default 30), EPWRHO [Ry] or CDUAL (default
    4, EPWRHO = CDUAL EPWPSI) -> Hartree (READIN_FOURIER,
    src/paw_ioroutines.f90:1247-1300);
  * !CONTROL!GENERIC DT (default 5), NSTEP (default 100)
    (READIN_GENERIC, :873-960);
  * !CONTROL!PSIDYN FRIC (default 0), MPSI (default 10 DT^2), MPSICG2
    (default 0.5 (10/(2 pi))^2 DT^2/MPSI), SAFEORTHO (default T)
    (READIN_PSIDYN, :1319-1420);
  * !STRUCTURE!GENERIC LUNIT (default 1; LUNIT[AA] in Angstrom)
    (:3910-3925); !LATTICE T (columns are the lattice vectors, times LUNIT,
    STRCIN_LATTICE :4018-4073); !ATOM NAME, SP (default: first two letters
    of NAME), R (times LUNIT, STRCIN_ATOM :4460-4571); !SPECIES NAME, ZV,
    NPRO; !OCCUPATIONS NSPIN (default 1), EMPTY (default 1), NBAND,
    CHARGE[E] (default 0), and the band count
    NB = max(NBAND, NOCC + EMPTY), NOCC = 1 + int((sum ZV - CHARGE)/2 - 1e-6),
    rounded up to even (STRCIN_OCCUP :4670-4780); !KPOINTS DIV.
Errors raise ValueError with the reference's message where it has one.
"""
import math
import re

import numpy as np

RY = 0.5                     # Hartree per Rydberg, src/paw_constants.f90:238
# bohr per Angstrom as the reference derives it from c, hbar, m_e and the
# fine-structure constant (src/paw_constants.f90:83-101, 240):
# 1e-10 METER, METER = c[m/s] m_e[kg] / (hbar[J s] c[a.u.])
_ALPHA = 7.2973530833e-3
ANGSTROM = 1.0e-10 * 299792458.0 * 9.109389754e-31 / (1.0545726663e-34 / _ALPHA)


class Block:
    """One !NAME ... !END block: ordered (key, values) pairs and sub-blocks."""

    def __init__(self, name):
        self.name = name
        self.items = []
        self.blocks = []

    def has(self, key):
        return any(k == key for k, _ in self.items)

    def get(self, key, default=None):
        for k, v in self.items:
            if k == key:
                return v
        return default

    def sub(self, name):
        return [b for b in self.blocks if b.name == name]

    def one(self, name):
        b = self.sub(name)
        return b[0] if b else None


_TOKEN = re.compile(r"'[^']*'|\"[^\"]*\"|\S+")


def _value(tok):
    if tok[0] in "'\"":
        return tok[1:-1]
    low = tok.lower()
    if low in ("t", ".true.", "true"):
        return True
    if low in ("f", ".false.", "false"):
        return False
    try:
        if re.fullmatch(r"[+-]?\d+", tok):
            return int(tok)
        return float(tok.replace("d", "e").replace("D", "e"))
    except ValueError:
        return tok


def parse(text):
    """Linked-list tree of a deck; the root Block is named '~'."""
    root = Block("~")
    stack = [root]
    key = None
    for line in text.splitlines():
        for tok in _TOKEN.findall(line):
            if tok.upper() == "!EOB":
                if len(stack) != 1:
                    raise ValueError(f"block !{stack[-1].name} not closed")
                return root
            if tok.startswith("!"):
                key = None
                name = tok[1:].upper()
                if name == "END":
                    if len(stack) == 1:
                        raise ValueError("!END without an open block")
                    stack.pop()
                else:
                    b = Block(name)
                    stack[-1].blocks.append(b)
                    stack.append(b)
                continue
            if "=" in tok:
                k, _, rest = tok.partition("=")
                key = k.upper()
                stack[-1].items.append((key, []))
                if rest:
                    stack[-1].items[-1][1].append(_value(rest))
                continue
            if key is None:
                raise ValueError(f"value {tok!r} outside a keyword in block !{stack[-1].name}")
            stack[-1].items[-1][1].append(_value(tok))
    if len(stack) != 1:
        raise ValueError(f"block !{stack[-1].name} not closed")
    return root


def _scalar(block, key, default=None):
    v = None if block is None else block.get(key)
    if v is None or len(v) == 0:
        return default
    return v[0]


class CPInput:
    """The step's parameters from a control and a structure deck (Hartree
    atomic units); see the module docstring for the rules."""

    def __init__(self, cntl_text, strc_text):
        c = parse(cntl_text).one("CONTROL")
        s = parse(strc_text).one("STRUCTURE")
        if c is None:
            raise ValueError("block !CONTROL is mandatory")
        if s is None:
            raise ValueError("block !STRUCTURE is mandatory")
        fourier, generic, psidyn = c.one("FOURIER"), c.one("GENERIC"), c.one("PSIDYN")
        self.epw_psi = float(_scalar(fourier, "EPWPSI", 30.0)) * RY
        if fourier is not None and fourier.has("EPWRHO"):
            self.epw_rho = float(_scalar(fourier, "EPWRHO")) * RY
        else:
            self.cdual = float(_scalar(fourier, "CDUAL", 4.0))
            self.epw_rho = self.epw_psi * self.cdual
        self.dt = float(_scalar(generic, "DT", 5.0))
        self.nstep = int(_scalar(generic, "NSTEP", 100))
        if psidyn is not None and psidyn.has("DIAG"):
            raise ValueError("SYNTAX HAS CHANGED: OPTION DIAG HAS BEEN REMOVED")
        self.friction = float(_scalar(psidyn, "FRIC", 0.0))
        self.emass = float(_scalar(psidyn, "MPSI", 10.0 * self.dt ** 2))
        self.emasscg2 = float(_scalar(psidyn, "MPSICG2",
                                      0.5 * (10.0 / (2.0 * math.pi)) ** 2 * self.dt ** 2 / self.emass))
        self.safeortho = bool(_scalar(psidyn, "SAFEORTHO", True))
        # structure
        sg = s.one("GENERIC")
        if sg is not None and sg.has("LUNIT[AA]"):
            if sg.has("LUNIT"):
                raise ValueError("LUNIT AND LUNIT[AA] ARE MUTUALLY EXCLUSIVE")
            unit = float(_scalar(sg, "LUNIT[AA]")) * ANGSTROM
        else:
            unit = float(_scalar(sg, "LUNIT", 1.0))
        self.lunit = unit
        lat = s.one("LATTICE")
        if lat is None:
            raise ValueError("BLOCK !STRUCTURE!LATTICE IS MANDATORY")
        t = lat.get("T")
        if not t or len(t) != 9:
            raise ValueError("KEYWORD T= IS MANDATORY")
        # T(3,3) in Fortran order: the first three numbers are the first lattice vector
        self.rbas = np.array(t, dtype=np.float64).reshape(3, 3).T * unit
        self.species = {}
        for sp in s.sub("SPECIES"):
            name = _scalar(sp, "NAME")
            self.species[name] = {"zv": float(_scalar(sp, "ZV", 0.0)),
                                  "npro": [int(x) for x in (sp.get("NPRO") or [])],
                                  "id": _scalar(sp, "ID")}
        self.atoms = []
        for at in s.sub("ATOM"):
            name = _scalar(at, "NAME")
            sp = _scalar(at, "SP", name[:2])
            r = at.get("R")
            if not r or len(r) != 3:
                raise ValueError(f"!STRUCTURE!ATOM {name}: R= needs three values")
            self.atoms.append((name, sp, np.array(r, dtype=np.float64) * unit))
        for _, sp, _ in self.atoms:
            if sp not in self.species:
                raise ValueError(f"species {sp} of an atom is not defined in !STRUCTURE!SPECIES")
        occ = s.one("OCCUPATIONS")
        self.nspin = int(_scalar(occ, "NSPIN", 1))
        if self.nspin not in (1, 2, 3):
            raise ValueError("NSPIN MUST BE 1, 2 OR 3")
        if self.nspin != 1:
            raise ValueError("only NSPIN=1 is supported by the wave-function step here")
        charge = -float(_scalar(occ, "CHARGE[E]", 0.0))
        nband = int(_scalar(occ, "NBAND", 0))
        self.nempty = int(_scalar(occ, "EMPTY", 1))
        self.sumofz = sum(self.species[sp]["zv"] for _, sp, _ in self.atoms)
        fmax = 2.0
        self.nocc = 1 + int((self.sumofz + charge) / fmax - 1e-6)
        nb = max(nband, self.nocc + self.nempty)
        self.nb = 2 * ((nb + 1) // 2)
        self.nelectrons = self.sumofz + charge
        kp = s.one("KPOINTS")
        self.kdiv = [int(x) for x in (kp.get("DIV") if kp is not None and kp.has("DIV") else [1, 1, 1])]

    @property
    def positions(self):
        return np.array([r for _, _, r in self.atoms])

    def occupations(self):
        """Ground-state filling of the NB bands with FMAX = 2 per band."""
        occ = np.zeros(self.nb)
        left = self.nelectrons
        for b in range(self.nb):
            occ[b] = min(2.0, max(0.0, left))
            left -= occ[b]
        return occ

    def to_system(self, species_models, seed=1, wells_depth=-1.6, wells_width=1.5):
        """A model System (pwgpu.model) on this cell: the deck's lattice,
        atoms, band count, cutoffs and wave-function dynamics, with the
        model setups given by species_models {species name: model.Species}
        and a model local potential (Gaussian wells on the sites)."""
        from . import model
        names = list(self.species)
        isp = np.array([names.index(sp) for _, sp, _ in self.atoms], dtype=np.int64)
        missing = [n for n in names if n not in species_models]
        if missing:
            raise ValueError(f"no model setup for species {missing}")
        pos = self.positions
        return model.System(name="deck", rbas=self.rbas, pos=pos, isp=isp,
                            species=[species_models[n] for n in names], nb=self.nb,
                            occ=self.occupations(), epw_psi=self.epw_psi, epw_rho=self.epw_rho,
                            wells=[(p, wells_depth, wells_width) for p in pos], dt=self.dt,
                            emass=self.emass, emasscg2=self.emasscg2, annee=self.friction, seed=seed)


def read(cntl_path, strc_path):
    with open(cntl_path) as f:
        c = f.read()
    with open(strc_path) as f:
        s = f.read()
    return CPInput(c, s)
