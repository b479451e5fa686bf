"""ctypes binding of the CPU oracle (``libevpfft_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libevpfft_oracle.so"

_DBL = C.POINTER(C.c_double)
_INT = C.POINTER(C.c_int)


class OrcRow(C.Structure):
    _fields_ = [
        ("iter", C.c_int), ("imacroloop", C.c_int), ("ipr", C.c_int), ("imicro", C.c_int),
        ("sim_time", C.c_double), ("wall_time", C.c_double), ("tdot", C.c_double), ("tdot_ref", C.c_double),
        ("eavg", C.c_double * 6), ("sgsolavg", C.c_double * 6), ("epavg", C.c_double * 6),
        ("eelavg", C.c_double * 6), ("velgradavg", C.c_double * 6), ("edotpavg", C.c_double * 6),
        ("evm", C.c_double), ("evmp", C.c_double), ("dvm", C.c_double), ("dvmp", C.c_double),
        ("svm", C.c_double), ("gamma", C.c_double), ("g", C.c_double), ("geff", C.c_double),
        ("devm_inc_max", C.c_double), ("diag_tdotref", C.c_double), ("erre", C.c_double),
        ("errs", C.c_double), ("errsbc", C.c_double), ("s33_target", C.c_double), ("s33", C.c_double),
        ("s33_sol", C.c_double),
        ("detF_min_sol", C.c_double), ("detF_min_gas", C.c_double), ("detF_min_sol_rel", C.c_double),
        ("detF_min_gas_rel", C.c_double),
        ("volinc_max_sol", C.c_double), ("volinc_max_gas", C.c_double), ("volinc_max_sol_rel", C.c_double),
        ("volinc_max_gas_rel", C.c_double),
        ("stretchinc_max_sol", C.c_double), ("stretchinc_max_gas", C.c_double),
        ("stretchinc_max_sol_rel", C.c_double), ("stretchinc_max_gas_rel", C.c_double),
        ("rotang_max_sol", C.c_double), ("rotang_max_gas", C.c_double),
        ("dL_step_rms_sol", C.c_double), ("dL_step_rms_gas", C.c_double), ("dL_step_rms", C.c_double),
        ("dL_step_max_sol", C.c_double), ("dL_step_max_gas", C.c_double), ("dL_step_max", C.c_double),
        ("dvelgradavg_mag", C.c_double), ("cond_c066mod_max_sol", C.c_double),
        ("cond_c066mod_max_gas", C.c_double), ("cond_cgas_max", C.c_double),
        ("imin_sol", C.c_int * 3), ("imin_gas", C.c_int * 3),
        ("passes", C.c_longlong), ("nr_steps", C.c_longlong), ("evals", C.c_longlong),
        ("outer", C.c_longlong), ("rollbacks", C.c_longlong),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if not isinstance(v, (int, float)) else v
        return out


class OrcProcess(C.Structure):
    _fields_ = [
        ("iudot", C.c_int * 9), ("udot", C.c_double * 9), ("idsim", C.c_int * 6), ("iscau", C.c_int * 6),
        ("scauchy", C.c_double * 9), ("dsim", C.c_double * 9),
        ("tdot", C.c_double), ("tdotref", C.c_double), ("tdotmin", C.c_double), ("devmmx", C.c_double),
        ("devmapp", C.c_double), ("ictrl", C.c_int), ("nsteps", C.c_int), ("error", C.c_double),
        ("erroral", C.c_double), ("itmax", C.c_int), ("iupdate", C.c_int), ("iuphard", C.c_int),
        ("iwfields", C.c_int), ("iwstep", C.c_int), ("xc0", C.c_double), ("igamma", C.c_int),
        ("ithermo", C.c_int),
    ]


class OracleStop(RuntimeError):
    """The restated reference stopped (status `code`, e.g. 1 for a singular system)."""

    def __init__(self, code, msg):
        super().__init__("stop %d: %s" % (code, msg))
        self.code = code


def build(force=False):
    if force or not LIB.exists() or LIB.stat().st_mtime < (HERE / "evpfft_oracle.c").stat().st_mtime:
        subprocess.run(["make", "-C", str(HERE), "-s"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_error.restype = C.c_char_p
        L.orc_error.argtypes = [C.c_void_p]
        L.orc_set_phase.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _INT] + [_DBL] * 9
        L.orc_set_voxels.argtypes = [C.c_void_p, _DBL, _INT, _INT]
        L.orc_add_process.argtypes = [C.c_void_p, C.POINTER(OrcProcess)]
        L.orc_init.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.POINTER(OrcRow)]
        L.orc_get_field.argtypes = [C.c_void_p, C.c_char_p, _DBL]
        L.orc_set_field.argtypes = [C.c_void_p, C.c_char_p, _DBL]
        L.orc_field_ncomp.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_chg_basis.argtypes = [_DBL, _DBL, _DBL, _DBL, C.c_int, C.c_int]
        L.orc_fft3d.argtypes = [_DBL, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_spatial_freq.argtypes = [C.c_int] * 6 + [_DBL]
        L.orc_green_at.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _DBL]
        L.orc_green_op.argtypes = [_DBL, _DBL, _DBL]
        L.orc_plastic_strain_rate.argtypes = [_DBL, _DBL, _DBL, _DBL, _DBL, C.c_int, _DBL, _DBL, _DBL]
        L.orc_expm_adaptive.argtypes = [_DBL, _DBL]
        L.orc_polar.argtypes = [_DBL, _DBL, _DBL]
        L.orc_cond1_est.restype = C.c_double
        L.orc_cond1_est.argtypes = [_DBL]
        L.orc_state_6x6.argtypes = [_INT, _DBL, _INT, _DBL, _DBL, _DBL, _DBL, _DBL]
        L.orc_solve_voxel.argtypes = [C.c_void_p, C.c_long, _DBL]
        L.orc_form_G.argtypes = [C.c_void_p, C.c_int]
        L.orc_convolve.argtypes = [C.c_void_p, _DBL, _DBL]
        L.orc_harden.argtypes = [C.c_void_p, C.c_long, _DBL, _DBL]
        L.orc_set_c066.argtypes = [C.c_void_p, _DBL]
        L.orc_set_ipr.argtypes = [C.c_void_p, C.c_int]
        L.orc_calc_c0.argtypes = [C.c_void_p]
        L.orc_kinematic.argtypes = [C.c_void_p, C.c_double, _DBL]
        L.orc_stage.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.orc_set_macro.argtypes = [C.c_void_p, C.c_char_p, _DBL]
        L.orc_get_lavg.argtypes = [C.c_void_p, _DBL]
        L.orc_set_displacement.argtypes = [C.c_void_p, C.c_int, _DBL]
        L.orc_gradient_tensor1.argtypes = [_DBL, _DBL, C.c_int, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def dp(a):
    return a.ctypes.data_as(_DBL)


def ip(a):
    return a.ctypes.data_as(_INT)


def process_struct(p):
    s = OrcProcess()
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


class Oracle:
    """The CPU restatement driven step by step (one accepted increment per step)."""

    def __init__(self, problem, modal=False, do_mix=True, nthreads=0):
        L = lib()
        self.L = L
        self.problem = problem
        n1, n2, n3 = problem.n
        self.h = L.orc_create(n1, n2, n3, len(problem.phases), float(problem.complgas), int(modal),
                              int(do_mix), int(nthreads))
        if not self.h:
            raise RuntimeError("orc_create failed")
        self._keep = []
        for iph, ph in enumerate(problem.phases, start=1):
            if ph.igas:
                z = np.zeros(1)
                L.orc_set_phase(self.h, iph, 1, 0, 0, ip(np.zeros(1, np.int32)), *([dp(z)] * 9))
                continue
            arrs = [np.ascontiguousarray(x, np.float64) for x in
                    (ph.dnca, ph.dbca, ph.tau, ph.thet, ph.nrs, ph.gamd0, ph.hpfac, ph.hard, ph.C)]
            nsm = np.ascontiguousarray(ph.nsm, np.int32)
            self._keep += arrs + [nsm]
            rc = L.orc_set_phase(self.h, iph, 0, ph.nsyst, ph.nmodes, ip(nsm), *[dp(a) for a in arrs])
            if rc:
                raise RuntimeError("orc_set_phase failed")
        eul = np.ascontiguousarray(problem.euler, np.float64)
        gr = np.ascontiguousarray(problem.grain, np.int32)
        phv = np.ascontiguousarray(problem.phase, np.int32)
        L.orc_set_voxels(self.h, dp(eul), ip(gr), ip(phv))
        if getattr(problem, "idispini", 0) > 0:  # initial displacement field (init_defgrad)
            d = np.ascontiguousarray(problem.dispini, np.float64).reshape(-1)
            if L.orc_set_displacement(self.h, int(problem.idispini), dp(d)) != 0:
                raise RuntimeError("orc_set_displacement: %s" % L.orc_error(self.h).decode())
        for p in problem.processes:
            s = process_struct(p)
            L.orc_add_process(self.h, C.byref(s))
        rc = L.orc_init(self.h)
        if rc < 0:
            raise RuntimeError("orc_init: %s" % L.orc_error(self.h).decode())
        if rc > 0:  # the reference's `stop 1` (singular system at initialisation)
            raise OracleStop(rc, L.orc_error(self.h).decode())
        eth = getattr(problem, "eth", None)
        if eth is not None:  # ithermo = 1: per-voxel eigenstrain rate
            self.set_field("eth", np.ascontiguousarray(eth, np.float64).reshape(-1))

    def step(self):
        row = OrcRow()
        rc = self.L.orc_step(self.h, C.byref(row))
        if rc == -7:  # singular system met during the increment (reference: stop 1)
            raise OracleStop(1, self.L.orc_error(self.h).decode())
        if rc < 0:
            raise RuntimeError("orc_step: %s" % self.L.orc_error(self.h).decode())
        return row if rc == 1 else None

    def field(self, name):
        k = self.L.orc_field_ncomp(self.h, name.encode())
        if k < 0:
            raise KeyError(name)
        if name == "xnode":
            n1, n2, n3 = self.problem.n
            n3n = 1 if n3 == 1 else n3 + 1
            out = np.zeros(3 * (n1 + 1) * (n2 + 1) * n3n)
        elif name == "macro":
            out = np.zeros(81 * 3 + 36 + 81 + 45 + 8)
        else:
            out = np.zeros(k * self.problem.nvox)
        rc = self.L.orc_get_field(self.h, name.encode(), dp(out))
        if rc:
            raise KeyError(name)
        return out

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, np.float64)
        if self.L.orc_set_field(self.h, name.encode(), dp(a)):
            raise KeyError(name)

    def stage(self, name, tdot=1.0):
        """Run one restated procedure of the increment on the current fields."""
        if self.L.orc_stage(self.h, name.encode(), float(tdot)):
            raise KeyError(name)

    def set_macro(self, name, arr):
        a = np.ascontiguousarray(arr, np.float64).ravel()
        if self.L.orc_set_macro(self.h, name.encode(), dp(a)):
            raise KeyError(name)

    def lavg(self):
        out = np.zeros(27)
        self.L.orc_get_lavg(self.h, dp(out))
        return out[:9].reshape(3, 3), out[9:18].reshape(3, 3), out[18:].reshape(3, 3)

    def macro(self):
        m = self.field("macro")
        keys = (("c0", 81), ("s0", 81), ("de_ds", 81), ("c066", 36), ("goperr0", 81), ("sgavg", 9),
                ("velgradavg", 9), ("dvelgradavg", 9), ("defgradinvavgc", 9), ("defgradavg", 9))
        out, o = {}, 0
        for k, n in keys:
            out[k] = m[o:o + n].copy()
            o += n
        out.update(wphcsol=m[o], wphcgas=m[o + 1], lvmavg=m[o + 2], svmavg=m[o + 3], wphsol=m[o + 7])
        return out

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def chg_basis(iopt, kdim=6, ce2=None, c2=None, ce4=None, c4=None):
    L = lib()
    ce2 = np.zeros(kdim) if ce2 is None else np.ascontiguousarray(ce2, np.float64).copy()
    c2 = np.zeros(9) if c2 is None else np.ascontiguousarray(c2, np.float64).ravel().copy()
    ce4 = np.zeros(kdim * kdim) if ce4 is None else np.ascontiguousarray(ce4, np.float64).ravel().copy()
    c4 = np.zeros(81) if c4 is None else np.ascontiguousarray(c4, np.float64).ravel().copy()
    L.orc_chg_basis(dp(ce2), dp(c2), dp(ce4), dp(c4), iopt, kdim)
    return {1: c2.reshape(3, 3), 2: ce2, 3: c4.reshape(3, 3, 3, 3), 4: ce4.reshape(kdim, kdim)}[iopt]


def fft3d(x, sign):
    """x: complex array (ncomp, n3, n2, n1) (C order, x fastest)."""
    a = np.ascontiguousarray(x, np.complex128).copy()
    nc, n3, n2, n1 = a.shape
    lib().orc_fft3d(a.view(np.float64).ctypes.data_as(_DBL), n1, n2, n3, nc, sign)
    return a
