"""ctypes binding of ``libcreep_hip.so`` (C-ABI declared in ``include/creep_hip.h``).

The library is built in-tree by ``__graft_entry__.build()`` (hipcc, gfx950).
There is no fallback: if the shared library is missing or no GPU is visible the
calls raise.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIBPATH = Path(__file__).resolve().parent / "libcreep_hip.so"

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)

#: symbols exported by the library (one per declaration in include/creep_hip.h)
EXPORTS = (
    "cg_create", "cg_destroy", "cg_last_error", "cg_set_phase", "cg_set_voxels", "cg_set_displacement", "cg_init",
    "cg_begin_increment", "cg_iterate", "cg_rollback", "cg_kinematic_substep", "cg_end_increment",
    "cg_calc_c0", "cg_field_ncomp", "cg_get_field", "cg_set_field", "cg_get_macro", "cg_get_stats",
    "cg_set_timing", "cg_set_tuning_file", "cg_get_fft_tuning", "cg_set_fft_tuning", "cg_set_graphs",
    "cg_set_overlap",
    "cg_kinematic_substep_noalloc",
    "cg_set_thermo", "cg_batch_create", "cg_batch_destroy", "cg_batch_iterate", "cg_batch_launch", "cg_batch_wait",
)


class CgProcess(C.Structure):
    _fields_ = [
        ("iudot", C.c_int * 9), ("udot", C.c_double * 9), ("idsim", C.c_int * 6), ("iscau", C.c_int * 6),
        ("scauchy", C.c_double * 9), ("dsim", C.c_double * 9),
        ("tdot", C.c_double), ("tdotref", C.c_double), ("tdotmin", C.c_double), ("devmmx", C.c_double),
        ("devmapp", C.c_double), ("ictrl", C.c_int), ("nsteps", C.c_int), ("error", C.c_double),
        ("erroral", C.c_double), ("itmax", C.c_int), ("iupdate", C.c_int), ("iuphard", C.c_int),
        ("iwfields", C.c_int), ("iwstep", C.c_int), ("xc0", C.c_double), ("igamma", C.c_int),
        ("ithermo", C.c_int),
    ]


class CgIterState(C.Structure):
    _fields_ = [("iter", C.c_int), ("mix", C.c_double), ("imicro", C.c_int)]


class CgIterOut(C.Structure):
    _fields_ = [
        ("errs", C.c_double), ("erre", C.c_double), ("errsbc", C.c_double),
        ("velgradavg", C.c_double * 9), ("sgavg", C.c_double * 9), ("sgavgsol", C.c_double * 9),
        ("dvelgradavg", C.c_double * 9), ("svmavg", C.c_double), ("dL_iter", C.c_double),
        ("nr_steps", C.c_int64), ("evals", C.c_int64), ("outer", C.c_int64),
    ]


class CgIncrOut(C.Structure):
    _fields_ = [
        ("eavg", C.c_double * 9), ("sgsolavg", C.c_double * 9), ("epavg", C.c_double * 9),
        ("eelavg", C.c_double * 9), ("edotpavg", C.c_double * 9), ("velgradavg", C.c_double * 9),
        ("sgavg", C.c_double * 9),
        ("evmavg", C.c_double), ("evmpavg", C.c_double), ("dvmavg", C.c_double), ("dvmpavg", C.c_double),
        ("svmavg", C.c_double), ("lvmavg", C.c_double), ("edotvmmx", C.c_double),
        ("gammaavg", C.c_double), ("gavg", C.c_double), ("geffavg", C.c_double),
        ("detF_min", C.c_double), ("detF_min_sol", C.c_double), ("detF_min_gas", C.c_double),
        ("volinc_max_sol", C.c_double), ("volinc_max_gas", C.c_double),
        ("stretchinc_max_sol", C.c_double), ("stretchinc_max_gas", C.c_double),
        ("rotang_max_sol", C.c_double), ("rotang_max_gas", C.c_double),
        ("dL_step_rms_sol", C.c_double), ("dL_step_rms_gas", C.c_double), ("dL_step_rms", C.c_double),
        ("dL_step_max_sol", C.c_double), ("dL_step_max_gas", C.c_double), ("dL_step_max", C.c_double),
        ("dvelgradavg_mag", C.c_double), ("cond_c066mod_max_sol", C.c_double),
        ("cond_c066mod_max_gas", C.c_double), ("cond_cgas_max", C.c_double),
        ("imin_sol", C.c_int * 3), ("imin_gas", C.c_int * 3),
    ]


class CgStats(C.Structure):
    _fields_ = [
        ("t_ms", C.c_double * 4), ("launches", C.c_int64 * 4), ("passes", C.c_int64),
        ("voxel_iters", C.c_int64), ("nr_steps", C.c_int64), ("evals", C.c_int64), ("outer", C.c_int64),
    ]


_lib = None


class HipError(RuntimeError):
    pass


def load():
    """Load the in-tree shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIBPATH.exists():
        raise HipError("%s is missing: run __graft_entry__.build() first" % LIBPATH)
    # PyTorch-ROCm ships its own HIP runtime under the same soname.  Loaded first, it is the
    # one the library binds to; loaded after the library's (/opt/rocm) copy, a second runtime
    # would enter the process and could not open the device.  One runtime per process either
    # way, so torch tensors and the library's contexts can be used together.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIBPATH))
    vp = C.c_void_p
    L.cg_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                            C.POINTER(vp)]
    L.cg_destroy.argtypes = [vp]
    L.cg_last_error.restype = C.c_char_p
    L.cg_last_error.argtypes = [vp]
    L.cg_set_phase.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, _I] + [_D] * 9
    L.cg_set_voxels.argtypes = [vp, _D, _I, _I]
    L.cg_init.argtypes = [vp, C.POINTER(CgProcess)]
    L.cg_begin_increment.argtypes = [vp, C.POINTER(CgProcess), C.c_int]
    L.cg_iterate.argtypes = [vp, C.POINTER(CgProcess), C.POINTER(CgIterState), C.POINTER(CgIterOut)]
    L.cg_rollback.argtypes = [vp, C.POINTER(CgProcess)]
    L.cg_kinematic_substep.argtypes = [vp, C.c_double, _D, _I]
    L.cg_end_increment.argtypes = [vp, C.POINTER(CgProcess), C.c_int, C.POINTER(CgIncrOut)]
    L.cg_calc_c0.argtypes = [vp, C.POINTER(CgProcess)]
    L.cg_field_ncomp.argtypes = [vp, C.c_char_p]
    L.cg_get_field.argtypes = [vp, C.c_char_p, _D, C.c_size_t]
    L.cg_set_field.argtypes = [vp, C.c_char_p, _D, C.c_size_t]
    L.cg_get_macro.argtypes = [vp, _D, _D, _D, _D, _D]
    L.cg_get_stats.argtypes = [vp, C.POINTER(CgStats), C.c_int]
    L.cg_set_timing.argtypes = [vp, C.c_int]
    L.cg_set_tuning_file.argtypes = [vp, C.c_char_p]
    L.cg_get_fft_tuning.argtypes = [vp, _I, _D]
    L.cg_set_fft_tuning.argtypes = [vp, _I]
    L.cg_set_graphs.argtypes = [vp, C.c_int]
    L.cg_set_overlap.argtypes = [vp, C.c_int]
    L.cg_set_displacement.argtypes = [vp, C.c_int, _D]
    L.cg_kinematic_substep_noalloc.argtypes = [C.c_long, vp, C.c_long, vp, C.c_double, vp, vp]
    L.cg_set_thermo.argtypes = [vp, _D]
    L.cg_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
    L.cg_batch_destroy.argtypes = [vp]
    L.cg_batch_iterate.argtypes = [vp, C.POINTER(C.POINTER(CgProcess)), C.POINTER(CgIterState), C.POINTER(CgIterOut),
                                   _I, _I]
    L.cg_batch_launch.argtypes = [vp, C.POINTER(C.POINTER(CgProcess)), C.POINTER(CgIterState), C.POINTER(CgIterOut),
                                  _I]
    L.cg_batch_wait.argtypes = [vp, _I]
    _lib = L
    return L


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I)


def process_struct(p):
    s = CgProcess()
    s.iudot[:] = [int(x) for x in np.asarray(p.iudot).ravel()]
    s.udot[:] = [float(x) for x in np.asarray(p.udot).ravel()]
    s.idsim[:] = [int(x) for x in p.idsim]
    s.iscau[:] = [int(x) for x in p.iscau]
    s.scauchy[:] = [float(x) for x in np.asarray(p.scauchy).ravel()]
    s.dsim[:] = [float(x) for x in np.asarray(p.dsim).ravel()]
    for name in ("tdot", "tdotref", "tdotmin", "devmmx", "devmapp", "error", "erroral", "xc0"):
        setattr(s, name, float(getattr(p, name)))
    for name in ("ictrl", "nsteps", "itmax", "iupdate", "iuphard", "iwfields", "iwstep", "igamma", "ithermo"):
        setattr(s, name, int(getattr(p, name)))
    return s


class Context:
    """Owns one ``cg_ctx`` (one RVE resident on the GPU)."""

    def __init__(self, problem, precision=0, modal=False, do_mix=True):
        L = load()
        self.L = L
        self.problem = problem
        n1, n2, n3 = problem.n
        h = C.c_void_p()
        rc = L.cg_create(n1, n2, n3, len(problem.phases), float(problem.complgas), int(modal), int(do_mix),
                         int(precision), C.byref(h))
        if rc != 0:
            raise HipError("cg_create failed (rc=%d): no usable GPU?" % rc)
        self.h = h
        for iph, ph in enumerate(problem.phases, start=1):
            if ph.igas:
                z = np.zeros(1)
                self._chk(L.cg_set_phase(h, iph, 1, 0, 0, ip(np.zeros(1, np.int32)), *([dp(z)] * 9)))
                continue
            arrs = [np.ascontiguousarray(x, np.float64) for x in
                    (ph.dnca, ph.dbca, ph.tau, ph.thet, ph.nrs, ph.gamd0, ph.hpfac, ph.hard, ph.C)]
            nsm = np.ascontiguousarray(ph.nsm, np.int32)
            self._chk(L.cg_set_phase(h, iph, 0, ph.nsyst, ph.nmodes, ip(nsm), *[dp(a) for a in arrs]))
        eul = np.ascontiguousarray(problem.euler, np.float64)
        gr = np.ascontiguousarray(problem.grain, np.int32)
        phv = np.ascontiguousarray(problem.phase, np.int32)
        self._chk(L.cg_set_voxels(h, dp(eul), ip(gr), ip(phv)))
        if getattr(problem, "idispini", 0) > 0:  # initial displacement field (init_defgrad)
            d = np.ascontiguousarray(problem.dispini, np.float64).reshape(-1)
            self._chk(L.cg_set_displacement(h, int(problem.idispini), dp(d)), "cg_set_displacement")
        self._p1 = process_struct(problem.processes[0])
        self._chk(L.cg_init(h, C.byref(self._p1)))
        eth = getattr(problem, "eth", None)
        if eth is not None:  # ithermo = 1 in some process: per-voxel eigenstrain rate
            self._chk(L.cg_set_thermo(h, dp(np.ascontiguousarray(eth, np.float64).reshape(-1))), "cg_set_thermo")

    def _chk(self, rc, what=""):
        if rc != 0:
            msg = self.L.cg_last_error(self.h).decode() if self.h else ""
            raise HipError("%s rc=%d: %s" % (what, rc, msg))
        return rc

    def field(self, name):
        k = self.L.cg_field_ncomp(self.h, name.encode())
        if k < 0:
            raise KeyError(name)
        if name == "xnode":
            n1, n2, n3 = self.problem.n
            out = np.zeros(3 * (n1 + 1) * (n2 + 1) * (1 if n3 == 1 else n3 + 1))
        else:
            out = np.zeros(k * self.problem.nvox)
        self._chk(self.L.cg_get_field(self.h, name.encode(), dp(out), out.size), "cg_get_field")
        return out

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        self._chk(self.L.cg_set_field(self.h, name.encode(), dp(a), a.size), "cg_set_field")

    def macro(self):
        c0, s0, de, g0 = np.zeros(81), np.zeros(81), np.zeros(81), np.zeros(81)
        c66 = np.zeros(36)
        self.L.cg_get_macro(self.h, dp(c0), dp(s0), dp(de), dp(c66), dp(g0))
        return dict(c0=c0, s0=s0, de_ds=de, c066=c66, goperr0=g0)

    def stats(self, reset=False):
        s = CgStats()
        self.L.cg_get_stats(self.h, C.byref(s), int(reset))
        return s

    def set_tuning_file(self, path):
        """FFT shape tuning cache (the reference's FFTW wisdom file); before the first pass."""
        self._chk(self.L.cg_set_tuning_file(self.h, str(path).encode()), "cg_set_tuning_file")

    def fft_tuning(self):
        """(tuned, shapes[8], us[5]) of the FFT passes (see cg_get_fft_tuning)."""
        sh = np.zeros(8, np.int32)
        us = np.zeros(5)
        rc = self.L.cg_get_fft_tuning(self.h, ip(sh), dp(us))
        if rc < 0:
            raise HipError("cg_get_fft_tuning rc=%d" % rc)
        return bool(rc), sh.tolist(), us.tolist()

    def set_graphs(self, on):
        """Replay each pass' fixed device work as HIP graphs, or launch it eagerly (default)."""
        self._chk(self.L.cg_set_graphs(self.h, int(bool(on))), "cg_set_graphs")

    def set_overlap(self, on):
        """Overlap each pass' host part with the next pass' transforms (default), or not."""
        self._chk(self.L.cg_set_overlap(self.h, int(bool(on))), "cg_set_overlap")

    def set_fft_tuning(self, shapes):
        """Fix the FFT shapes (8 entries of 0/1) and skip the tuning."""
        sh = np.ascontiguousarray(shapes, np.int32)
        self._chk(self.L.cg_set_fft_tuning(self.h, ip(sh)), "cg_set_fft_tuning")

    def close(self):
        if getattr(self, "h", None):
            self.L.cg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
