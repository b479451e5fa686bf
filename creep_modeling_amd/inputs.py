"""Synthetic RVE generator and writers for the LS-EVPFFT input files.

File formats follow the reference readers:

* microstructure ``.dat``: one line per voxel ``phi1 Phi phi2 i j k grain phase``
  (Bunge angles in degrees, 1-based indices) -- ``IO_functions.f90:759-809``.
* ``cupl2.sx``: crystal / slip-system / Voce hardening file --
  ``IO_functions.f90:423-578`` (writer in ``utils/io_utils.py:77-120``).
* ``cuel.sx``: elastic constants, ``iso`` flag then a 6x6 VPSC-Voigt matrix or
  ``young nu`` -- ``IO_functions.f90:682-717`` (writer ``utils/io_utils.py:36-74``).
* ``fft.in``: the process file -- ``IO_functions.f90:51-346`` (creep template
  ``utils/sim_cases/creep_calibration.py:112-179``).
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

# ---------------------------------------------------------------------------
# Slip-system tables
# ---------------------------------------------------------------------------

#: FCC {111}<110>: the 12 (n, b) pairs in the order ``utils/io_utils.py:108-119``
#: lists them (the order fixes the slip-system numbering of every output).
FCC_111_110 = (
    ((1, 1, -1), (0, 1, 1)), ((1, 1, -1), (1, 0, 1)), ((1, 1, -1), (1, -1, 0)),
    ((1, -1, -1), (0, 1, -1)), ((1, -1, -1), (1, 0, 1)), ((1, -1, -1), (1, 1, 0)),
    ((1, -1, 1), (0, 1, 1)), ((1, -1, 1), (1, 0, -1)), ((1, -1, 1), (1, 1, 0)),
    ((1, 1, 1), (0, 1, -1)), ((1, 1, 1), (1, 0, -1)), ((1, 1, 1), (1, -1, 0)),
)


def _canon(v):
    """Sign-canonical integer vector (first non-zero component positive)."""
    for c in v:
        if c != 0:
            return tuple(v) if c > 0 else tuple(-x for x in v)
    return tuple(v)


def _family(base):
    """All sign/permutation variants of ``base``, one per +-pair, sorted."""
    out = set()
    for perm in itertools.permutations(base):
        for signs in itertools.product((1, -1), repeat=3):
            v = tuple(s * p for s, p in zip(signs, perm))
            if any(v):
                out.add(_canon(v))
    return sorted(out, reverse=True)


def bcc_slip_families():
    """BCC <111> slip on {110}, {112} and {123}: 12 + 12 + 24 = 48 systems.

    Each family is a list of (n, b) integer pairs with n.b = 0; systems are
    enumerated deterministically (planes in descending lexicographic order of
    their sign-canonical Miller indices, then directions likewise).
    """
    dirs = _family((1, 1, 1))
    fams = []
    for base in ((1, 1, 0), (1, 1, 2), (1, 2, 3)):
        systems = []
        for n in _family(base):
            for b in dirs:
                if sum(a * c for a, c in zip(n, b)) == 0:
                    systems.append((n, b))
        fams.append(systems)
    return fams


# ---------------------------------------------------------------------------
# Material and process descriptions
# ---------------------------------------------------------------------------


@dataclass
class SlipMode:
    label: str
    systems: list  # [(n, b)], integer Miller indices
    nrs: float = 10.0
    gamd0: float = 1.0e-3
    tau0f: float = 100.0
    tau0b: float = 100.0
    tau1: float = 40.0
    thet0: float = 500.0
    thet1: float = 50.0
    hpfac: float = 0.0
    hself: float = 1.0
    hlat: float = 1.0


@dataclass
class Material:
    modes: list
    c11: float = 247000.0
    c12: float = 147000.0
    c44: float = 125000.0


def fcc_material(n=10.0, gamd0=1.0e-3, tau0=100.0, tau1=40.0, thet0=500.0, thet1=50.0,
                 c11=247000.0, c12=147000.0, c44=125000.0, hself=1.0, hlat=1.0):
    mode = SlipMode("{111}<110> SLIP", list(FCC_111_110), n, gamd0, tau0, tau0, tau1, thet0,
                    thet1, 0.0, hself, hlat)
    return Material([mode], c11, c12, c44)


def bcc_material(n=20.0, gamd0=1.0e-3, tau0=120.0, tau1=50.0, thet0=600.0, thet1=40.0,
                 c11=231000.0, c12=135000.0, c44=116000.0, hself=1.0, hlat=1.0):
    labels = ("{110}<111> SLIP", "{112}<111> SLIP", "{123}<111> SLIP")
    modes = [SlipMode(lab, fam, n, gamd0, tau0, tau0, tau1, thet0, thet1, 0.0, hself, hlat)
             for lab, fam in zip(labels, bcc_slip_families())]
    return Material(modes, c11, c12, c44)


@dataclass
class CreepProcess:
    """One process block of ``fft.in`` (creep template semantics)."""
    s11: float = 0.0
    s22: float = 0.0
    s33: float = 300.0
    udot33: float = 1.0e-4
    tdot: float = 1.0
    tdotref: float = 1.0
    tdotmin: float = 1.0
    devmmx: float = 1.0
    ictrl: int = -1
    nsteps: int = 50
    error: float = 1.0e-6
    erroral: float = 1.0e-9
    itmax: int = 200
    iupdate: int = 1
    iuphard: int = 1
    iwfields: int = 0
    iwstep: int = -1
    xc0: float = 2.0
    igamma: int = 0
    #: "creep": sigma33 imposed, the other stresses zero (the creep template,
    #: ``utils/sim_cases/creep_calibration.py``); "tension": the velocity gradient L33 = udot33
    #: imposed with sigma33 free (the tension template, ``utils/sim_cases/tension.py:33-47``)
    control: str = "creep"
    #: 1: an eigenstrain rate per grain acts in the first increment (IO_functions.f90:316-341);
    #: thermo_file names the file (written by write_thermo)
    ithermo: int = 0
    thermo_file: str = ""


# ---------------------------------------------------------------------------
# Microstructure generation
# ---------------------------------------------------------------------------


@dataclass
class Microstructure:
    n: tuple
    grain: np.ndarray          # (V,) int32, 1-based, x-fastest voxel order
    phase: np.ndarray          # (V,) int32, 1-based
    euler: np.ndarray          # (V, 3) float64 Bunge degrees (phi1, Phi, phi2)
    grain_euler: np.ndarray = field(default=None)  # (G, 3)


def voronoi_rve(n, ngrains, seed, phase=1):
    """Periodic Voronoi RVE with uniformly random (Haar on SO(3)) orientations.

    Seeds are i.i.d. uniform in [0, n)^3; voxel centres sit at integer
    indices; distances use the minimum image.  Orientations: phi1 = 360 u1,
    Phi = acos(2 u2 - 1), phi2 = 360 u3 (degrees).
    """
    from scipy.spatial import cKDTree

    if isinstance(n, int):
        n = (n, n, n)
    n1, n2, n3 = n
    rng = np.random.default_rng(seed)
    seeds = rng.uniform(0.0, 1.0, size=(ngrains, 3)) * np.array([n1, n2, n3], dtype=np.float64)
    u = rng.uniform(0.0, 1.0, size=(ngrains, 3))
    geul = np.empty((ngrains, 3))
    geul[:, 0] = 360.0 * u[:, 0]
    geul[:, 1] = np.degrees(np.arccos(2.0 * u[:, 1] - 1.0))
    geul[:, 2] = 360.0 * u[:, 2]
    box = np.array([n1, n2, n3], dtype=np.float64)
    tree = cKDTree(np.mod(seeds, box), boxsize=box)
    V = n1 * n2 * n3
    grain = np.empty(V, dtype=np.int32)
    # x fastest: v = i + n1*(j + n2*k); query slab by slab to bound memory.
    ii, jj = np.meshgrid(np.arange(n1), np.arange(n2), indexing="xy")
    plane = np.stack([ii.ravel(), jj.ravel(), np.zeros(n1 * n2)], axis=1).astype(np.float64)
    for k in range(n3):
        plane[:, 2] = k
        _, idx = tree.query(plane, k=1, workers=-1)  # independent queries: same result on any thread count
        grain[k * n1 * n2:(k + 1) * n1 * n2] = idx.astype(np.int32) + 1
    euler = geul[grain - 1]
    return Microstructure((n1, n2, n3), grain, np.full(V, phase, np.int32), euler, geul)


# ---------------------------------------------------------------------------
# Writers
# ---------------------------------------------------------------------------


def write_microstructure(path, ms: Microstructure):
    n1, n2, n3 = ms.n
    V = n1 * n2 * n3
    v = np.arange(V)
    i = v % n1 + 1
    j = (v // n1) % n2 + 1
    k = v // (n1 * n2) + 1
    e = ms.euler
    with open(path, "w") as f:
        for a in range(V):
            f.write("%.6f %.6f %.6f %d %d %d %d %d\n" % (e[a, 0], e[a, 1], e[a, 2], i[a], j[a], k[a],
                                                       ms.grain[a], ms.phase[a]))


def write_cuel(path, mat: Material):
    c11, c12, c44 = mat.c11, mat.c12, mat.c44
    rows = [
        (c11, c12, c12, 0, 0, 0), (c12, c11, c12, 0, 0, 0), (c12, c12, c11, 0, 0, 0),
        (0, 0, 0, c44, 0, 0), (0, 0, 0, 0, c44, 0), (0, 0, 0, 0, 0, c44),
    ]
    lines = ["0                    ISO"] + [" ".join(repr(float(x)) for x in r) for r in rows]
    Path(path).write_text("\n".join(lines) + "\n")


def write_cupl2(path, mat: Material):
    nm = len(mat.modes)
    out = [
        "SLIP SYSTEMS FOR CUBIC CRYSTAL",
        "CUBIC          icryst",
        "1.   1.   1.   crystal axes (cdim(i))",
        "%d              nmodesx (total # of modes listed in the file)" % nm,
        "%d              nmodes  (# of modes to be used in the calculation)" % nm,
        " ".join(str(i + 1) for i in range(nm)) + "              mode(i) (label of the modes to be used)",
    ]
    for im, m in enumerate(mat.modes):
        out.append(m.label)
        out.append("%d   %d   %r   %r   0.0  0                 modex,nsmx,nrsx,gamd0x,twshx,isectwx"
                   % (im + 1, len(m.systems), float(m.nrs), float(m.gamd0)))
        out.append("%r  %r   %r   %r  %r  %r  tau0xf,tau0xb,tau1x,thet0,thet1,hpfac"
                   % (float(m.tau0f), float(m.tau0b), float(m.tau1), float(m.thet0), float(m.thet1),
                      float(m.hpfac)))
        lat = ["%r" % float(m.hlat)] * nm
        out.append("%r   %s      hselfx,hlatex" % (float(m.hself), "  ".join(lat)))
        for (nv, bv) in m.systems:
            out.append("%3d %3d %3d    %3d %3d %3d" % (*nv, *bv))
    Path(path).write_text("\n".join(out) + "\n")


def write_fft_in(path, n, microstructure_path, cupl2_path, cuel_path, processes,
                 complgas=0.01, idispini=0, displacement_path=None):
    if isinstance(n, int):
        n = (n, n, n)
    lines = [
        "1                      number of phases (nph)",
        "%d %d %d               nx,ny,nz - number of points" % n,
        "1.  1.  1.             RVE dimensions (delt)",
        "* name and path of microstructure file (filetext)",
        str(microstructure_path),
        "*INFORMATION ABOUT PHASE #1",
        "0                          igas(iph)",
        "* name and path of single crystal files (dummy if igas(iph)=1)",
        str(cupl2_path),
        str(cuel_path),
        "* restart analysis, remeshing and gas phase compliance",
        "0 200                  ires,imicrosave",
        "%r                   complgas" % float(complgas),
        "* initial displacement field",
        "%d                      IDISPINI" % idispini,
    ] + ([str(displacement_path)] if idispini > 0 else []) + [
        "%d                      number of processes (nproc)" % len(processes),
    ]
    for p in processes:
        tension = p.control == "tension"
        lines += [
            "*INFORMATION ABOUT TEST CONDITIONS",
            "* boundary conditions",
            "    0       1       1           iudot",
            "    1       0       1",
            "    1       1       %d" % (1 if tension else 0),
            "",
            "    0.0     0.0      0.0          udot",
            "    0.0     0.0      0.0",
            "    0.0     0.0      %r" % float(p.udot33),
            "",
            "    1       0        0           iscau",
            "            1        0",
            "                     %d" % (0 if tension else 1),
            "",
            "    %r      0.       0.          scauchy" % float(p.s11),
            "            %r       0." % float(p.s22),
            "                     %r" % float(p.s33),
            "* other",
            "%r %r %r %r  eqincr or tdot, tdotref, tdotmin, devmmx"
            % (float(p.tdot), float(p.tdotref), float(p.tdotmin), float(p.devmmx)),
            "%d                     ictrl" % p.ictrl,
            "*INFORMATION ABOUT RUN CONDITIONS",
            "%d               nsteps" % p.nsteps,
            "%r %r  err,erral" % (float(p.error), float(p.erroral)),
            "%d             itmax" % p.itmax,
            "0               IRECOVER",
            "0               ISAVE",
            "%d               IUPDATE" % p.iupdate,
            "%d               IUPHARD" % p.iuphard,
            "0               IWTEX",
            "%d %d            IWFIELDS,IWSTEP" % (p.iwfields, p.iwstep),
            "%r %d            XC0, IGAMMA" % (float(p.xc0), p.igamma),
            "%d               ITHERMO" % p.ithermo,
        ]
        if p.ithermo == 1:
            lines.append(str(p.thermo_file))
    Path(path).write_text("\n".join(lines) + "\n")


def write_thermo(path, eth):
    """Eigenstrain file of ithermo = 1: the number of grains, then eth11 eth22 eth33 eth23 eth31
    eth12 per grain; eth is (ngrains, 3, 3) (symmetric part used)."""
    eth = np.asarray(eth, float)
    lines = ["%d" % eth.shape[0]]
    for e in eth:
        lines.append(" ".join(repr(float(x)) for x in (e[0, 0], e[1, 1], e[2, 2], e[1, 2], e[2, 0], e[0, 1])))
    Path(path).write_text("\n".join(lines) + "\n")


def write_displacement(path, n, disp, idispini):
    """Initial displacement file of idispini = 1 (one line "i j k u1 u2 u3" per cell, i fastest)
    or 2 (one line per node: (n1+1)(n2+1)(n3+1) nodes, n3 + 1 -> 1 for a 2D grid), as
    init_defgrad reads it (IO_functions.f90:850-880).  disp: (lines, 3)."""
    n1, n2, n3 = n
    if idispini == 1:
        dims = (n1, n2, n3)
    else:
        dims = (n1 + 1, n2 + 1, 1 if n3 == 1 else n3 + 1)
    disp = np.asarray(disp, float).reshape(-1, 3)
    if disp.shape[0] != dims[0] * dims[1] * dims[2]:
        raise ValueError("displacement has %d rows, expected %d" % (disp.shape[0], dims[0] * dims[1] * dims[2]))
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    with open(path, "w") as fh:
        for (a, b, c), u in zip(zip(i.ravel() + 1, j.ravel() + 1, k.ravel() + 1), disp):
            fh.write("%d %d %d %r %r %r\n" % (a, b, c, float(u[0]), float(u[1]), float(u[2])))


def write_case(directory, ms: Microstructure, mat: Material, processes, complgas=0.01, displacement=None):
    """Write a complete input set; returns the ``fft.in`` path.  displacement: (idispini, values)
    for an initial displacement file (see write_displacement)."""
    d = Path(directory).resolve()
    d.mkdir(parents=True, exist_ok=True)
    write_microstructure(d / "micro.dat", ms)
    write_cupl2(d / "cupl2.sx", mat)
    write_cuel(d / "cuel.sx", mat)
    idispini, dpath = 0, None
    if displacement is not None:
        idispini, dpath = int(displacement[0]), d / "dispini.txt"
        write_displacement(dpath, tuple(ms.n), displacement[1], idispini)
    write_fft_in(d / "fft.in", ms.n, d / "micro.dat", d / "cupl2.sx", d / "cuel.sx", processes,
                 complgas, idispini, dpath)
    return d / "fft.in"


# ---------------------------------------------------------------------------
# Named workloads (seeded; harness choices recorded in the docs)
# ---------------------------------------------------------------------------


def hall_petch_shift(nvox):
    """Hidden Hall-Petch shift of ``IO_functions.f90:107-169`` for a grain of
    ``nvox`` voxels: 252.8 / sqrt(max((64 nvox)^(1/3), 12))."""
    d = max((64.0 * nvox) ** (1.0 / 3.0), 12.0)
    return (632.0 / 2.5) / math.sqrt(d)


# Arrhenius scaling of the reference slip rate for the batched sweep: Q = 280 kJ/mol, the
# factor is 1 at 525 C
SWEEP_Q = 280.0e3
SWEEP_TREF = 525.0 + 273.15
R_GAS = 8.314462618


def sweep_cases(count=64, nsteps=100, seed=4000, sizes=(32, 48, 64, 96)):
    """The batched sweep of independent FCC RVEs: grid edge uniform over `sizes`, grain count
    uniform in 32..512, sigma33 uniform in 200-400 MPa, exponent n uniform in 8-14, and
    gamma0_dot scaled by exp(-Q/R (1/T - 1/Tref)) for T uniform over {475, 525, 575} C.
    RVE i uses microstructure seed seed + i.  Returns a list of (params, (ms, mat, [proc]))."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice(sizes))
        ng = int(rng.integers(32, 513))
        s33 = float(rng.uniform(200.0, 400.0))
        nrs = float(rng.uniform(8.0, 14.0))
        tc = float(rng.choice((475.0, 525.0, 575.0)))
        fac = float(np.exp(-SWEEP_Q / R_GAS * (1.0 / (tc + 273.15) - 1.0 / SWEEP_TREF)))
        mat = fcc_material(n=nrs, gamd0=1.0e-3 * fac)
        proc = CreepProcess(s33=s33, tdot=1.0, tdotref=1.0, tdotmin=1.0, nsteps=nsteps)
        ms = voronoi_rve(n, ng, seed + i)
        out.append((dict(n=n, grains=ng, s33=s33, nrs=nrs, temp_c=tc, gamd0=1.0e-3 * fac), (ms, mat, [proc])))
    return out


def config_case(config, n=None, ngrains=None, nsteps=None, s33=None, nrs=None):
    """(Microstructure, Material, [CreepProcess]) for the named workloads.

    config 0: 32^3 / 64 grains FCC, 50 x 1 s, sigma33 = 300 MPa.
    config 1: 128^3 / 512 grains FCC, 200 x 1 s.
    config 2: as 1 with n = 15 (also 256^3 / 4096 grains), sigma33 in {250,300,350}.
    config 3: BCC-48, 256^3 / 2048 grains, sigma11 = 150, sigma33 = 300, 100 x 0.5 s.
    Keyword overrides shrink a workload for tests.
    """
    seed = 1000 + config
    if config == 2 and n in (None, 128) and ngrains in (None, 512):
        seed = 1001  # the 128^3 input of the n = 15 workload is the 128^3 / 512-grain RVE of config 1
    if config in (0, 1, 2):
        dn, dg, ds = {0: (32, 64, 50), 1: (128, 512, 200), 2: (128, 512, 200)}[config]
        mat = fcc_material(n=15.0 if config == 2 else 10.0)
        proc = CreepProcess(s33=300.0, tdot=1.0, tdotref=1.0, tdotmin=1.0, nsteps=ds)
    elif config == 3:
        dn, dg, ds = 256, 2048, 100
        mat = bcc_material()
        proc = CreepProcess(s11=150.0, s33=300.0, tdot=0.5, tdotref=0.5, tdotmin=0.5, nsteps=ds)
    else:
        raise ValueError("config must be 0..3")
    if nrs is not None:
        for m in mat.modes:
            m.nrs = float(nrs)
    if s33 is not None:
        proc.s33 = float(s33)
    n = n or dn
    ngrains = ngrains or dg
    if nsteps is not None:
        proc.nsteps = nsteps
    ms = voronoi_rve(n, ngrains, seed)
    return ms, mat, [proc]
