"""basis.in / nadvibs.in readers and writers, and the problem description.

basis.in is the Fortran namelist &NADVIBS (keys and defaults of
Source/secondary.f90:835-870); nadvibs.in is list-directed with one skipped
record before every data group (Source/top.f90:122-135).
"""
import math
import os
import re
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

AU2WAVE = 219474.6313705   # Source/progdata.f90:7
AU2EV = 27.211384523       # Source/progdata.f90:8

NAMELIST_KEYS = ("niter", "natoms", "nmodes", "nstates", "basis", "restart", "bconv", "idroots", "soroots",
                 "reorthog", "chkorthog", "nseg", "ztoler", "maxdisk", "weights", "shiftref", "nirreps",
                 "npirr", "ordr", "initstate")


def noterms(nmodes, p):
    """C(nmodes+p-1, p): monomials of order p (Source/function.f90:253-258)."""
    return math.comb(nmodes + p - 1, p)


def monomials(nmodes, p):
    """Mode tuples (0-based) of order p in nadvibs.in order: non-increasing
    tuples with the first entry running fastest (Source/top.f90:153-168)."""
    from itertools import combinations_with_replacement
    return [tuple(reversed(c)) for c in combinations_with_replacement(range(nmodes), p)]


@dataclass
class Problem:
    """The contents of basis.in and nadvibs.in."""
    nstates: int
    nmodes: int
    ordr: int
    nfunc: List[int]
    aomega: np.ndarray                 # hartree
    concoef: np.ndarray                # [nstblks]
    potterms: np.ndarray               # [nstblks, ordr, noterms(ordr)], zero padded
    ztoler: float = 1e-20
    quirks: int = 3
    niter: int = 1
    reorthog: int = 0
    chkorthog: int = 100
    bconv: int = 0
    restart: int = 0
    natoms: int = 2
    nseg: int = 1
    maxdisk: float = 1000.0
    idroots: int = 0
    soroots: int = 0
    shiftref: int = 0
    statew: Optional[np.ndarray] = None
    nirreps: int = 1
    npirrep: Optional[List[int]] = None
    initstate: Optional[List[int]] = None
    bomega: Optional[np.ndarray] = None
    dvec: Optional[np.ndarray] = None
    tmat: Optional[np.ndarray] = None  # [nmodes*nmodes] column-major
    name: str = ""

    @property
    def nstblks(self):
        return self.nstates * (self.nstates + 1) // 2

    @property
    def dimen(self):
        return int(np.prod(self.nfunc))

    @property
    def n(self):
        return self.dimen * self.nstates

    def blocks(self):
        """(i, j) 0-based, i >= j, in nadvibs.in block order (Source/top.f90:80-85)."""
        return [(i, j) for i in range(self.nstates) for j in range(i + 1)]


# ---------------------------------------------------------------------------
# namelist
# ---------------------------------------------------------------------------
def _fnum(tok):
    t = tok.strip().lower().replace("d", "e")
    if t in (".true.", "t", ".t."):
        return True
    if t in (".false.", "f", ".f."):
        return False
    try:
        return int(t)
    except ValueError:
        return float(t)


def _values(text):
    out = []
    for tok in re.split(r"[,\s]+", text.strip()):
        if not tok:
            continue
        m = re.fullmatch(r"(\d+)\*(.+)", tok)
        if m:
            out.extend([_fnum(m.group(2))] * int(m.group(1)))
        else:
            out.append(_fnum(tok))
    return out


def read_namelist(path, group="nadvibs"):
    """Parses one &group ... / namelist into {key: value or list}."""
    with open(path) as f:
        text = "\n".join(line.split("!")[0] for line in f)
    m = re.search(r"&" + group + r"\b(.*?)(?:^|[\s,])/", text, re.I | re.S)
    if not m:
        raise ValueError(f"no &{group} group in {path}")
    body = m.group(1)
    parts = re.split(r"([A-Za-z_][A-Za-z0-9_]*)\s*(?:\(\s*\d*\s*:?\s*\d*\s*\))?\s*=", body)
    out = {}
    for k in range(1, len(parts) - 1, 2):
        vals = _values(parts[k + 1])
        out[parts[k].lower()] = vals if len(vals) != 1 else vals[0]
    return out


def _as_list(v):
    return list(v) if isinstance(v, list) else [v]


def read_basis(path):
    """basis.in with the reference's defaults (Source/secondary.f90:835-843)."""
    nl = read_namelist(path)
    for k in nl:
        if k not in NAMELIST_KEYS:
            raise ValueError(f"unknown basis.in key {k!r}")
    d = dict(niter=1, natoms=2, nmodes=1, nstates=1, restart=0, bconv=0, idroots=0, soroots=0, reorthog=0,
             chkorthog=100, nseg=1, ztoler=1e-20, maxdisk=1000.0, shiftref=0, nirreps=1, ordr=2)
    d.update({k: v for k, v in nl.items() if k not in ("basis", "weights", "npirr", "initstate")})
    nm, ns = int(d["nmodes"]), int(d["nstates"])
    basis = _as_list(nl.get("basis", [1]))
    basis = (basis + [1] * 100)[:nm]
    weights = (_as_list(nl.get("weights", [])) + [1.0] * 10)[:ns]
    initstate = (_as_list(nl.get("initstate", [])) + [0] * 100)[:nm]
    npirr = _as_list(nl.get("npirr", [nm]))
    nirreps = int(d["nirreps"])
    npirr = (npirr + [0] * 10)[:nirreps]
    if sum(npirr) != nm:  # symmetry disabled (Source/secondary.f90:848-852)
        nirreps, npirr = 1, [nm]
    d["basis"], d["weights"], d["initstate"], d["npirr"], d["nirreps"] = basis, weights, initstate, npirr, nirreps
    return d


def write_basis(path, prob, niter=None):
    p = prob
    lines = [" &NADVIBS",
             f"  niter={niter if niter is not None else p.niter}, natoms={p.natoms}, nmodes={p.nmodes}, "
             f"nstates={p.nstates}, ordr={p.ordr},",
             "  basis=" + ",".join(str(int(b)) for b in p.nfunc) + ",",
             f"  restart={p.restart}, bconv={p.bconv}, idroots={p.idroots}, soroots={p.soroots},",
             f"  reorthog={p.reorthog}, chkorthog={p.chkorthog}, nseg={p.nseg}, ztoler={p.ztoler!r},"
             .replace("e-", "d-"),
             f"  maxdisk={p.maxdisk!r}, shiftref={p.shiftref}, nirreps={p.nirreps},",
             "  npirr=" + ",".join(str(int(v)) for v in (p.npirrep or [p.nmodes])) + ",",
             "  weights=" + ",".join(repr(float(w)) for w in (p.statew if p.statew is not None
                                                              else [1.0] * p.nstates)) + ",",
             "  initstate=" + ",".join(str(int(v)) for v in (p.initstate or [0] * p.nmodes)),
             " /", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


# ---------------------------------------------------------------------------
# nadvibs.in (list-directed)
# ---------------------------------------------------------------------------
class _ListReader:
    """Fortran list-directed reads: every read starts on a new record and
    consumes records until it has its values; the rest of the last record is
    dropped."""

    def __init__(self, lines):
        self.lines = lines
        self.pos = 0

    def skip(self):
        if self.pos >= len(self.lines):
            raise ValueError("unexpected end of nadvibs.in")
        self.pos += 1

    def read(self, n):
        vals = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise ValueError("unexpected end of nadvibs.in")
            line = self.lines[self.pos]
            self.pos += 1
            line = line.split("/")[0] if "/" in line else line
            vals.extend(float(v) if not isinstance(v, bool) else float(v) for v in _values(line))
        return np.array(vals[:n], dtype=np.float64)


def read_potential(path, nstates, nmodes, ordr, neworigin=False):
    """Returns (aomega, concoef, potterms, bomega, dvec, tmat)."""
    with open(path) as f:
        rd = _ListReader(f.read().splitlines())
    nb = nstates * (nstates + 1) // 2
    ntmax = noterms(nmodes, ordr)
    rd.skip()
    aomega = rd.read(nmodes)
    concoef = np.zeros(nb)
    pot = np.zeros((nb, ordr, ntmax))
    for b in range(nb):
        rd.skip()
        concoef[b] = rd.read(1)[0]
        for p in range(1, ordr + 1):
            rd.skip()
            pot[b, p - 1, :noterms(nmodes, p)] = rd.read(noterms(nmodes, p))
    bomega = dvec = tmat = None
    if neworigin:
        rd.skip(); bomega = rd.read(nmodes)
        rd.skip(); dvec = rd.read(nmodes)
        rd.skip(); tmat = rd.read(nmodes * nmodes)
    return aomega, concoef, pot, bomega, dvec, tmat


def write_potential(path, prob):
    p = prob
    fmt = lambda a: " ".join(repr(float(v)) for v in a)
    out = ["Frequencies (hartree)", fmt(p.aomega)]
    for b, (i, j) in enumerate(p.blocks()):
        out += [f"Block ({i + 1},{j + 1}) constant", fmt([p.concoef[b]])]
        for q in range(1, p.ordr + 1):
            out += [f"Block ({i + 1},{j + 1}) order {q}", fmt(p.potterms[b, q - 1, :noterms(p.nmodes, q)])]
    if p.shiftref:
        out += ["Reference-state frequencies", fmt(p.bomega), "Displacements", fmt(p.dvec),
                "T matrix (column major)", fmt(p.tmat)]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def read_problem(directory, quirks=3):
    """Reads basis.in and nadvibs.in from `directory`."""
    b = read_basis(os.path.join(directory, "basis.in"))
    neworigin = int(b["shiftref"]) != 0
    ao, cc, pot, bo, dv, tm = read_potential(os.path.join(directory, "nadvibs.in"), int(b["nstates"]),
                                             int(b["nmodes"]), int(b["ordr"]), neworigin)
    return Problem(nstates=int(b["nstates"]), nmodes=int(b["nmodes"]), ordr=int(b["ordr"]),
                   nfunc=[int(v) for v in b["basis"]], aomega=ao, concoef=cc, potterms=pot,
                   ztoler=float(b["ztoler"]), quirks=quirks, niter=int(b["niter"]), reorthog=int(b["reorthog"]),
                   chkorthog=int(b["chkorthog"]), bconv=int(b["bconv"]), restart=int(b["restart"]),
                   natoms=int(b["natoms"]), nseg=int(b["nseg"]), maxdisk=float(b["maxdisk"]),
                   idroots=int(b["idroots"]), soroots=int(b["soroots"]), shiftref=int(b["shiftref"]),
                   statew=np.array(b["weights"], dtype=np.float64), nirreps=int(b["nirreps"]),
                   npirrep=[int(v) for v in b["npirr"]], initstate=[int(v) for v in b["initstate"]],
                   bomega=bo, dvec=dv, tmat=tm)


def write_problem(directory, prob):
    os.makedirs(directory, exist_ok=True)
    write_basis(os.path.join(directory, "basis.in"), prob)
    write_potential(os.path.join(directory, "nadvibs.in"), prob)
