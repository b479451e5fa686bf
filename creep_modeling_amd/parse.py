"""Readers for the LS-EVPFFT input contract -> :class:`Problem`.

Restates ``load_input`` (``IO_functions.f90:6-365``), ``data_crystal``
(``:367-637``), ``data_crystal_elast`` (``:640-719``) and ``data_grain``
(``:722-812``) up to the point where the files have been read and the
crystal tables assembled.  Per-voxel derived state (rotation matrices, rotated
stiffness, Hall-Petch initial strengths, ...) is computed by the solver that
consumes the :class:`Problem` (``cg_create`` in the HIP library).

Fortran list-directed reads (``read(u,*)``) are emulated: values are taken
from the current record and following records as needed, the remainder of the
last record is discarded, ``d`` exponents are accepted.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np


class InputError(ValueError):
    """The input files violate a check that stops the reference solver."""


class _Reader:
    def __init__(self, path):
        self.path = str(path)
        self.lines = Path(path).read_text().splitlines()
        self.pos = 0

    def line(self):
        if self.pos >= len(self.lines):
            raise InputError("%s: unexpected end of file" % self.path)
        s = self.lines[self.pos]
        self.pos += 1
        return s

    def skip(self):
        self.line()

    def values(self, n, conv=float):
        """List-directed read of ``n`` values (one or more records)."""
        out = []
        while len(out) < n:
            toks = self.line().replace(",", " ").split()
            for t in toks:
                if t == "/":
                    break
                if len(out) == n:
                    break
                out.append(t)
        return [_conv(t, conv, self.path) for t in out]


def _conv(tok, conv, path):
    try:
        if conv is int:
            return int(tok)
        return float(tok.replace("d", "e").replace("D", "E"))
    except ValueError as exc:
        raise InputError("%s: cannot read %r as %s" % (path, tok, conv.__name__)) from exc


@dataclass
class Phase:
    igas: bool
    nsyst: int = 0
    nmodes: int = 0
    nsm: list = field(default_factory=list)
    dnca: np.ndarray = None     # (ns, 3) unit plane normals, crystal axes
    dbca: np.ndarray = None     # (ns, 3) unit slip directions, crystal axes
    tau: np.ndarray = None      # (ns, 3): tau0f, tau0b, tau1
    thet: np.ndarray = None     # (ns, 2): thet0, thet1
    nrs: np.ndarray = None
    gamd0: np.ndarray = None
    hpfac: np.ndarray = None
    hard: np.ndarray = None     # (ns, ns)
    C: np.ndarray = None        # (3,3,3,3) crystal stiffness


@dataclass
class Process:
    iudot: np.ndarray
    udot: np.ndarray
    idsim: np.ndarray
    iscau: np.ndarray
    scauchy: np.ndarray
    dsim: np.ndarray
    tdot: float
    tdotref: float
    tdotmin: float
    devmmx: float
    ictrl: int
    devmapp: float
    nsteps: int
    error: float
    erroral: float
    itmax: int
    irecover: int
    isave: int
    iupdate: int
    iuphard: int
    iwtex: int
    iwfields: int
    iwstep: int
    xc0: float
    igamma: int
    ithermo: int
    eth_grain: np.ndarray = None  # (ngrains + 1, 3, 3) eigenstrain rates of ithermo = 1 (row 0 unused)


@dataclass
class Problem:
    n: tuple
    phases: list
    grain: np.ndarray     # (V,) int32 1-based, voxel index v = i + n1*(j + n2*k)
    phase: np.ndarray     # (V,) int32 1-based
    euler: np.ndarray     # (V, 3) degrees
    complgas: float
    idispini: int
    processes: list
    eth: np.ndarray = None  # (V, 3, 3) per-voxel eigenstrain rate (some process has ithermo = 1)
    dispini: np.ndarray = None  # idispini 1: (V, 3) cell displacements; 2: (nodes, 3), i fastest

    @property
    def nvox(self):
        return int(self.n[0] * self.n[1] * self.n[2])

    @property
    def nsyst_max(self):
        return max([p.nsyst for p in self.phases if not p.igas] + [0])


def _sym33(a):
    return 0.5 * (a + a.T)


def _vm_strain(d):
    h = (d[0, 0] + d[1, 1] + d[2, 2]) / 3.0
    s = ((d[0, 0] - h) ** 2 + (d[1, 1] - h) ** 2 + (d[2, 2] - h) ** 2 + d[0, 1] ** 2 + d[0, 2] ** 2
         + d[1, 2] ** 2 + d[1, 0] ** 2 + d[2, 0] ** 2 + d[2, 1] ** 2)
    return float(np.sqrt(2.0 / 3.0 * s))


def read_crystal(path):
    """``data_crystal`` (``IO_functions.f90:367-637``) for cubic/orthorhombic
    and hexagonal/trigonal files."""
    r = _Reader(path)
    r.skip()
    icryst = r.line()[:3]
    cdim = r.values(3)
    covera = cdim[2] / cdim[0]
    nmodesx = r.values(1, int)[0]
    nmodes = r.values(1, int)[0]
    mode = r.values(nmodes, int)
    if nmodes > 10:
        raise InputError("nmodes=%d exceeds supported modes" % nmodes)
    hself = {}
    hlat = {}
    rows = []
    kount = 0
    for nm in range(1, nmodesx + 1):
        r.skip()
        vals = r.values(6)
        modex, nsmx, nrsx, gamd0x, twshx, _isectw = int(vals[0]), int(vals[1]), vals[2], vals[3], vals[4], vals[5]
        tau0f, tau0b, tau1, thet0, thet1, hpfac = r.values(6)
        hv = r.values(1 + nmodesx)
        hself[nm] = hv[0]
        hlat[nm] = hv[1:]
        if kount >= nmodes or modex != mode[kount]:
            for _ in range(nsmx):
                r.skip()
            continue
        if thet0 < thet1:
            raise InputError("initial hardening lower than final hardening for mode %d" % (kount + 1))
        if tau1 <= 1.0e-6:
            tau1 = 1.0e-6
            thet0 = thet1
        if twshx > 1.0e-15:
            raise InputError("twinning modes are not supported (ntwmmx=0)")
        hex_ = icryst.upper() in ("HEX", "TRI")
        cub = icryst.upper() in ("CUB", "ORT")
        if not (hex_ or cub):
            raise InputError("cannot identify the crystal symmetry %r" % icryst)
        for _ in range(nsmx):
            if hex_:
                v = r.values(8, int)
                isn, isb = v[:4], v[4:]
                sn = np.array([isn[0], (isn[0] + 2.0 * isn[1]) / np.sqrt(3.0), isn[3] / covera])
                sb = np.array([1.5 * isb[0], (isb[0] / 2.0 + isb[1]) * np.sqrt(3.0), isb[3] * covera])
            else:
                v = r.values(6, int)
                sn = np.array([v[m] / cdim[m] for m in range(3)], dtype=np.float64)
                sb = np.array([v[3 + m] * cdim[m] for m in range(3)], dtype=np.float64)
            dn = sn / np.sqrt(sn @ sn)
            db = sb / np.sqrt(sb @ sb)
            dn[np.abs(dn) < 1.0e-3] = 0.0
            db[np.abs(db) < 1.0e-3] = 0.0
            if dn @ db >= 1.0e-3:
                raise InputError("slip system in mode %d is not orthogonal" % nm)
            rows.append((kount, dn, db, nrsx, gamd0x, tau0f, tau0b, tau1, thet0, thet1, hpfac))
        kount += 1
    ns = len(rows)
    ph = Phase(igas=False, nsyst=ns, nmodes=nmodes)
    ph.nsm = [sum(1 for rw in rows if rw[0] == m) for m in range(nmodes)]
    ph.dnca = np.array([rw[1] for rw in rows])
    ph.dbca = np.array([rw[2] for rw in rows])
    ph.nrs = np.array([rw[3] for rw in rows])
    ph.gamd0 = np.array([rw[4] for rw in rows])
    ph.tau = np.array([[rw[5], rw[6], rw[7]] for rw in rows])
    ph.thet = np.array([[rw[8], rw[9]] for rw in rows])
    ph.hpfac = np.array([rw[10] for rw in rows])
    # hard(i, j) = hlatex(mode_i, mode_j) with hselfx(mode_i) on the diagonal;
    # modes are relabelled to the active list order (IO_functions.f90:468-471, 578-591).
    act = [mode[m] for m in range(nmodes)]
    hard = np.zeros((ns, ns))
    im_of = [rw[0] for rw in rows]
    for i in range(ns):
        for j in range(ns):
            hard[i, j] = hlat[act[im_of[i]]][act[im_of[j]] - 1]
        hard[i, i] = hself[act[im_of[i]]]
    ph.hard = hard
    return ph


def read_elastic(path):
    """``data_crystal_elast`` (``IO_functions.f90:640-719``) -> C_ijkl."""
    r = _Reader(path)
    iso = r.values(1, int)[0]
    C = np.zeros((3, 3, 3, 3))
    ijv = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    if iso == 0:
        c66 = np.array([r.values(6) for _ in range(6)])
        for a, (i1, i2) in enumerate(ijv):
            for b, (j1, j2) in enumerate(ijv):
                for (p, q) in ((i1, i2), (i2, i1)):
                    for (s, t) in ((j1, j2), (j2, j1)):
                        C[p, q, s, t] = c66[a, b]
    else:
        young, nu = r.values(2)
        mu = young / (2.0 * (1.0 + nu))
        la = 2.0 * mu * nu / (1.0 - 2.0 * nu)
        d = np.eye(3)
        C = (la * np.einsum("ij,kl->ijkl", d, d)
             + mu * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d)))
    return C


def read_microstructure(path, n):
    """``data_grain`` (``IO_functions.f90:759-809``): explicit indices, any order."""
    n1, n2, n3 = n
    V = n1 * n2 * n3
    try:
        import pandas as pd
        arr = pd.read_csv(path, sep=r"\s+", header=None, nrows=V, engine="c").to_numpy()
    except ImportError:  # pragma: no cover
        arr = np.loadtxt(path, max_rows=V)
    if arr.shape[0] < V:
        raise InputError("%s: %d voxel lines, expected %d" % (path, arr.shape[0], V))
    i = arr[:, 3].astype(np.int64) - 1
    j = arr[:, 4].astype(np.int64) - 1
    k = arr[:, 5].astype(np.int64) - 1
    if (i.min() < 0 or j.min() < 0 or k.min() < 0 or i.max() >= n1 or j.max() >= n2
            or k.max() >= n3):
        raise InputError("%s: voxel index out of range" % path)
    v = i + n1 * (j + n2 * k)
    euler = np.zeros((V, 3))
    grain = np.zeros(V, np.int32)
    phase = np.zeros(V, np.int32)
    euler[v] = arr[:, 0:3]
    grain[v] = arr[:, 6].astype(np.int32)
    phase[v] = arr[:, 7].astype(np.int32)
    return euler, grain, phase


def _read_process(r, resolve):
    r.skip()
    r.skip()
    iudot = np.array([r.values(3, int) for _ in range(3)], dtype=np.int32)
    if iudot[0, 0] + iudot[1, 1] + iudot[2, 2] == 2:
        raise InputError("check diagonal boundary conditions iudot")
    for i in range(3):
        for j in range(3):
            if i != j and iudot[i, j] + iudot[j, i] == 0:
                raise InputError("check off-diagonal boundary conditions iudot")
    r.skip()
    udot = np.array([r.values(3) for _ in range(3)])
    dsim = _sym33(udot)
    dvm = _vm_strain(dsim)
    idsim = np.zeros(6, np.int32)
    idsim[0:3] = [iudot[0, 0], iudot[1, 1], iudot[2, 2]]
    idsim[3] = 1 if (iudot[1, 2] == 1 and iudot[2, 1] == 1) else 0
    idsim[4] = 1 if (iudot[0, 2] == 1 and iudot[2, 0] == 1) else 0
    idsim[5] = 1 if (iudot[0, 1] == 1 and iudot[1, 0] == 1) else 0
    r.skip()
    iscau = np.zeros(6, np.int32)
    iscau[0], iscau[5], iscau[4] = r.values(3, int)
    iscau[1], iscau[3] = r.values(2, int)
    iscau[2] = r.values(1, int)[0]
    for i in range(6):
        if iscau[i] * idsim[i] != 0 or iscau[i] + idsim[i] != 1:
            raise InputError("check boundary condits on strain-rate and stress")
    r.skip()
    sc = np.zeros((3, 3))
    sc[0, 0], sc[0, 1], sc[0, 2] = r.values(3)
    sc[1, 1], sc[1, 2] = r.values(2)
    sc[2, 2] = r.values(1)[0]
    sc[2, 1], sc[2, 0], sc[1, 0] = sc[1, 2], sc[0, 2], sc[0, 1]
    r.skip()
    tdot, tdotref, tdotmin, devmmx = r.values(4)
    ictrl = r.values(1, int)[0]
    devmapp = 0.0
    if ictrl == 0:
        devmapp = tdot
        tdot = tdot / dvm
        tdotref = tdot
    elif ictrl > 0:
        raise InputError("ictrl>0 not implemented")
    r.skip()
    nsteps = r.values(1, int)[0]
    error, erroral = r.values(2)
    itmax = r.values(1, int)[0]
    irecover = r.values(1, int)[0]
    isave = r.values(1, int)[0]
    iupdate = r.values(1, int)[0]
    iuphard = r.values(1, int)[0]
    iwtex = r.values(1, int)[0]
    iwfields, iwstep = r.values(2, int)
    xc0v = r.values(2)
    xc0, igamma = xc0v[0], int(xc0v[1])
    if not nsteps > 0:
        raise InputError("nsteps must be positive")
    if not (error > 0 or erroral > 0):
        raise InputError("error and erroral tolerances must be positive")
    if not itmax > 0:
        raise InputError("itmax must be positive")
    for name, val in (("irecover", irecover), ("isave", isave), ("iupdate", iupdate),
                      ("iuphard", iuphard), ("iwtex", iwtex), ("iwfields", iwfields)):
        if val not in (0, 1):
            raise InputError("%s must be 0 or 1" % name)
    if igamma not in (0, 1, 2):
        raise InputError("igamma must be 0 (continuous), 1 (discrete), or 2 (modified)")
    ithermo = r.values(1, int)[0]
    eth_grain = read_thermo(resolve(r.line())) if ithermo == 1 else None
    return Process(iudot, udot, idsim, iscau, sc, dsim, tdot, tdotref, tdotmin, devmmx, ictrl,
                   devmapp, nsteps, error, erroral, itmax, irecover, isave, iupdate, iuphard, iwtex,
                   iwfields, iwstep, xc0, igamma, ithermo, eth_grain)


def read_thermo(path):
    """The eigenstrain file of ithermo = 1 (IO_functions.f90:316-332): the number of grains,
    then eth11 eth22 eth33 eth23 eth31 eth12 per grain.  Returns (ngrains + 1, 3, 3), row 0
    unused (grain ids are 1-based)."""
    r = _Reader(path)
    ngr = r.values(1, int)[0]
    eth = np.zeros((ngr + 1, 3, 3))
    for g in range(1, ngr + 1):
        e = r.values(6)
        eth[g] = [[e[0], e[5], e[4]], [e[5], e[1], e[3]], [e[4], e[3], e[2]]]
    return eth


def voxel_eth(procs, grain):
    """Per-voxel eigenstrain rate: every process assigns the grain table to the voxels
    (IO_functions.f90:334-341) and the table holds the file read last, so the last process
    with ithermo = 1 decides.  None if no process has ithermo = 1."""
    tabs = [p.eth_grain for p in procs if p.eth_grain is not None]
    if not tabs:
        return None
    tab = tabs[-1]
    if grain.max() >= tab.shape[0]:
        raise InputError("grain id %d has no eigenstrain rate in the ithermo file" % grain.max())
    return tab[grain]


def read_fft_in(path):
    """Parse ``fft.in`` and every file it names into a :class:`Problem`."""
    base = Path(path).parent
    r = _Reader(path)

    def resolve(name):
        p = Path(name.strip())
        return p if p.is_absolute() or p.exists() else (base / p)

    nph = r.values(1, int)[0]
    n = tuple(r.values(3, int))
    r.values(3)          # delt (overwritten to 4,4,4 by the reference)
    r.skip()
    filetext = r.line()
    phases = []
    for _ in range(nph):
        r.skip()
        igas = r.values(1, int)[0] != 0
        r.skip()
        fpl = r.line()
        fel = r.line()
        if igas:
            phases.append(Phase(igas=True))
        else:
            ph = read_crystal(resolve(fpl))
            ph.C = read_elastic(resolve(fel))
            phases.append(ph)
    euler, grain, phase = read_microstructure(resolve(filetext), n)
    if phase.min() < 1 or phase.max() > nph:
        raise InputError("phase id out of range")
    r.skip()
    ires = r.values(2, int)
    del ires
    complgas = r.values(1)[0]
    r.skip()
    idispini = r.values(1, int)[0]
    dispini = None
    if idispini > 0:
        if idispini not in (1, 2):
            raise InputError("idispini must be 0, 1 or 2")
        dispini = read_displacement(resolve(r.line()), n, idispini)
    nproc = r.values(1, int)[0]
    if nproc < 1:
        raise InputError("cannot have less than one processes")
    procs = [_read_process(r, resolve) for _ in range(nproc)]
    return Problem(n, phases, grain, phase, euler, complgas, idispini, procs, voxel_eth(procs, grain), dispini)


def read_displacement(path, n, idispini):
    """The initial displacement file (init_defgrad, IO_functions.f90:850-880): one line
    "i j k u1 u2 u3" per cell (idispini 1, n1 n2 n3 lines) or per node (idispini 2,
    (n1+1)(n2+1)(n3+1) lines, one z layer for a 2D grid), i fastest; the indices are not used."""
    n1, n2, n3 = n
    rows = n1 * n2 * n3 if idispini == 1 else (n1 + 1) * (n2 + 1) * (1 if n3 == 1 else n3 + 1)
    try:
        txt = Path(path).read_text().replace("D", "E").replace("d", "e")
    except OSError as e:
        raise InputError("cannot read the initial displacement file %s: %s" % (path, e))
    vals = np.array(txt.split(), dtype=float)
    if vals.size < 6 * rows:
        raise InputError("initial displacement file %s: %d values, expected %d lines of 6" % (path, vals.size, rows))
    return vals[:6 * rows].reshape(rows, 6)[:, 3:].copy()


def problem_from_case(ms, mat, processes, complgas=0.01):
    """Build a :class:`Problem` in memory (no files) from the generator's
    objects, going through the same text formats as the files would."""
    import tempfile

    from . import inputs

    with tempfile.TemporaryDirectory() as d:
        d = Path(d)
        inputs.write_cupl2(d / "cupl2.sx", mat)
        inputs.write_cuel(d / "cuel.sx", mat)
        ph = read_crystal(d / "cupl2.sx")
        ph.C = read_elastic(d / "cuel.sx")
        (d / "micro.dat").write_text("")
        inputs.write_fft_in(d / "fft.in", ms.n, d / "micro.dat", d / "cupl2.sx", d / "cuel.sx",
                            processes, complgas)
        r = _Reader(d / "fft.in")
        for _ in range(15):
            r.skip()
        nproc = r.values(1, int)[0]
        procs = [_read_process(r, lambda name: Path(name.strip())) for _ in range(nproc)]
    grain = ms.grain.astype(np.int32)
    return Problem(tuple(ms.n), [ph], grain, ms.phase.astype(np.int32), np.asarray(ms.euler, np.float64), complgas, 0,
                   procs, voxel_eth(procs, grain))
