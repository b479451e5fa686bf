"""The LS-EVPFFT increment loop driving the HIP kernels through the C-ABI.

Mirrors the control flow of ``lsevpfft/src/LS-EVPFFT.f90:180-590`` (process
loop, increment loop, fixed-point loop with convergence test, Delta-t control
with kinematic sub-stepping / rollback / ramp, accepted-step updates, CSV rows).
The Fortran host (``host/lsevpfft_host.f90``) implements the same loop for the
drop-in executable; this module is its Python twin for tests and benchmarks.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from pathlib import Path

import numpy as np

from . import lib as _lib

RESULTS_HEADER = (
    "iter,macrostep,ipr,sim_time,wall_time,tdot,tdot_ref,"
    "eav11,eav22,eav33,eav23,eav13,eav12,"
    "sav11,sav22,sav33,sav23,sav13,sav12,"
    "epav11,epav22,epav33,epav23,epav13,epav12,"
    "eelav11,eelav22,eelav33,eelav23,eelav13,eelav12,"
    "vgrad11,vgrad22,vgrad33,vgrad23,vgrad13,vgrad12,"
    "edotp11,edotp22,edotp33,edotp23,edotp13,edotp12,"
    "evm,evmp,dvm,dvmp,svm,gamma,g,geff"
)  # IO_functions.f90:1264-1271

DIAG_HEADER = (
    "iter,macrostep,ipr,sim_time,wall_time,tdot,vm_max_inc,tdotref,erre,errs,errsbc,"
    "s33_target,s33,s33_sol,detF_min_sol,detF_min_gas,detF_min_sol_rel,detF_min_gas_rel,"
    "volinc_max_sol,volinc_max_gas,volinc_max_sol_rel,volinc_max_gas_rel,"
    "stretchinc_max_sol,stretchinc_max_gas,stretchinc_max_sol_rel,stretchinc_max_gas_rel,"
    "rotang_max_sol,rotang_max_gas,dL_step_rms_sol,dL_step_rms_gas,dL_step_rms,"
    "dL_step_max_sol,dL_step_max_gas,dL_step_max,dvelgradavg_mag,cond_c066mod_max_sol,"
    "cond_c066mod_max_gas,cond_cgas_max,imin_sol_i,imin_sol_j,imin_sol_k,imin_gas_i,imin_gas_j,imin_gas_k"
)  # types.f90:787-798


def v6(m):
    """3x3 (row-major, 9) -> (11, 22, 33, 23, 13, 12)."""
    m = list(m)
    return [m[0], m[4], m[8], m[5], m[2], m[1]]


def _safe_rel(x, xp):
    huge = 1.7976931348623157e308
    if xp <= 2.2250738585072014e-308 or xp >= huge * 0.5:
        return 0.0
    return abs(x - xp) / xp


class Solver:
    """One RVE: fixed-point increments on the GPU, one accepted increment per :meth:`step`."""

    def __init__(self, problem, precision=0, modal=False, do_mix=True, scan_cond=True, max_pass_factor=100,
                 tune_file=None, fft_shapes=None, graphs=False, overlap=True):
        self.problem = problem
        self.ctx = _lib.Context(problem, precision=precision, modal=modal, do_mix=do_mix)
        if tune_file:  # FFT shape cache (--fftw_save)
            self.ctx.set_tuning_file(tune_file)
        if fft_shapes is not None:  # fixed FFT shapes, no tuning
            self.ctx.set_fft_tuning(fft_shapes)
        if graphs:  # replay each pass' fixed device work as HIP graphs
            self.ctx.set_graphs(True)
        if not overlap:  # no early transforms of the next pass during the host's part of a pass
            self.ctx.set_overlap(False)
        self.L = self.ctx.L
        self.h = self.ctx.h
        self.procs = [_lib.process_struct(p) for p in problem.processes]
        self.scan_cond = scan_cond
        self.max_pass_factor = max_pass_factor
        self.ipr = -1
        self.imicro = 0
        self.imacroloop = 1
        self.sim_time = 0.0
        self.proc_active = False
        self.total_macrosteps = sum(p.nsteps for p in problem.processes)
        self.prev = dict(sol=0.0, gas=0.0, vs=0.0, vg=0.0, ss=0.0, sg=0.0)
        self.start = time.time()
        self.passes_total = 0

    def _chk(self, rc, what):
        if rc != 0:
            raise _lib.HipError("%s rc=%d: %s" % (what, rc, self.L.cg_last_error(self.h).decode()))

    def _kin(self, tdot):
        out = (C.c_double * 4)()
        nsub = C.c_int()
        self._chk(self.L.cg_kinematic_substep(self.h, tdot, out, C.byref(nsub)), "cg_kinematic_substep")
        return nsub.value, list(out)

    def begin(self):
        """Start the next increment (LS-EVPFFT.f90:196-219: process advance, Delta-t plan,
        snapshot, update_schmid, form_G, c066mod); False when every process is done."""
        L, h = self.L, self.h
        while not self.proc_active or self.imicro >= self.procs[self.ipr].nsteps:
            if self.ipr + 1 >= len(self.procs):
                return False
            self.ipr += 1
            P = self.procs[self.ipr]
            self.nconverged = 0
            self.kin_ctrl = P.iscau[2] == 0
            self.tdot_plan = P.tdot
            self.proc_time_target = P.nsteps * self.tdot_plan
            self.proc_time_done = 0.0
            self.imicro = 0
            self.proc_active = True
        P = self.procs[self.ipr]
        self.imicro += 1
        if P.ictrl == -1:
            rem = max(0.0, self.proc_time_target - self.proc_time_done)
            P.tdot = min(self.tdot_plan, rem) if rem > 0.0 else self.tdot_plan
        do_schmid = int((self.imicro == 1 and self.ipr == 0) or P.iupdate == 1)
        self._chk(L.cg_begin_increment(h, C.byref(P), do_schmid), "cg_begin_increment")
        self.P = P
        self.st = _lib.CgIterState(0, 1.0, self.imicro)
        self.out = _lib.CgIterOut()
        self.did_converge = self.accept = False
        self.passes = self.rollbacks = self.nr = self.ev = self.ou = 0
        return True

    def step(self):
        """Run one accepted increment; returns a row dict (results + diagnostics) or None when done."""
        if not self.begin():
            return None
        while True:
            self._chk(self.L.cg_iterate(self.h, C.byref(self.P), C.byref(self.st), C.byref(self.out)), "cg_iterate")
            if self.pass_done():
                return self.finish()

    def pass_done(self):
        """The control of LS-EVPFFT.f90:358-429 after one pass (self.out): convergence, strain
        control, Delta-t cut-back with rollback.  True when the increment's loop has ended."""
        L, h, P, st, out = self.L, self.h, self.P, self.st, self.out
        did_converge = self.did_converge
        cap = self.max_pass_factor * max(P.itmax, 1) + 1000
        self.passes += 1
        passes = self.passes
        self.nr += out.nr_steps
        self.ev += out.evals
        self.ou += out.outer
        if passes > cap:
            raise _lib.HipError("fixed-point loop did not terminate")
        if P.ictrl == 0:
            vs = np.array(out.velgradavg).reshape(3, 3)
            vs = 0.5 * (vs + vs.T)
            hh = np.trace(vs) / 3.0
            d = vs - hh * np.eye(3)
            evm = math.sqrt(2.0 / 3.0 * float((d * d).sum()))
            P.tdot = P.devmapp / evm
            P.tdotref = P.tdot
        if out.errs <= P.error and out.erre <= P.error and out.errsbc <= P.error:
            did_converge = True
        if P.ictrl == -1:
            if P.tdot > P.tdotmin:
                if (not did_converge) and st.iter == P.itmax:
                    tdot = P.tdot
                    nsub, _ = self._kin(tdot)
                    tdot = tdot / nsub
                    tdot = max(P.tdotmin, min(tdot, 0.5 * P.tdot))
                    nsub = math.ceil(P.tdot / tdot)
                    if nsub > 1:
                        self._replan(P, tdot)
                        self.nconverged = 0
                        st.iter = 0
                        self._chk(L.cg_rollback(h, C.byref(P)), "cg_rollback")
                        self.rollbacks += 1
                elif did_converge:
                    tdot = P.tdot
                    nsub, kin = self._kin(tdot)
                    tdot = tdot / nsub
                    if kin[3] <= 1.0e-3:
                        tdot = max(P.tdotmin, min(tdot, 0.5 * P.tdot))
                    else:
                        tdot = max(P.tdotmin, tdot)
                    nsub = math.ceil(P.tdot / tdot)
                    if nsub > 1:
                        self._replan(P, tdot)
                        did_converge = False
                        self.nconverged = 0
                        st.iter = 0
                        st.mix = 1.0
                        self._chk(L.cg_rollback(h, C.byref(P)), "cg_rollback")
                        self.rollbacks += 1
            elif st.iter == P.itmax:
                self.accept = True
        self.did_converge = did_converge
        return did_converge or self.accept

    def finish(self):
        """The accepted increment: updates, averages, diagnostics, Delta-t ramp, calc_c0
        (LS-EVPFFT.f90:431-589); returns the row."""
        L, h, P, st, out = self.L, self.h, self.P, self.st, self.out
        did_converge = self.did_converge
        passes, rollbacks, nr, ev, ou = self.passes, self.rollbacks, self.nr, self.ev, self.ou
        last_iter = st.iter
        inc = _lib.CgIncrOut()
        self._chk(L.cg_end_increment(h, C.byref(P), int(self.scan_cond), C.byref(inc)), "cg_end_increment")
        tdot_used = P.tdot
        if P.ictrl == -1:
            self.proc_time_done += tdot_used
            P.tdot = self.tdot_plan
        self.sim_time += tdot_used
        row = dict(
            iter=last_iter, imacroloop=self.imacroloop, ipr=self.ipr + 1, imicro=self.imicro,
            sim_time=self.sim_time, wall_time=float(int(time.time() - self.start)), tdot=tdot_used,
            tdot_ref=P.tdotref,
            eavg=v6(inc.eavg), sgsolavg=v6(inc.sgsolavg), epavg=v6(inc.epavg), eelavg=v6(inc.eelavg),
            velgradavg=v6(inc.velgradavg), edotpavg=v6(inc.edotpavg),
            evm=inc.evmavg, evmp=inc.evmpavg, dvm=inc.dvmavg, dvmp=inc.dvmpavg, svm=inc.svmavg,
            gamma=inc.gammaavg, g=inc.gavg, geff=inc.geffavg,
            devm_inc_max=inc.edotvmmx * tdot_used, diag_tdotref=P.tdotref,
            erre=out.erre, errs=out.errs, errsbc=out.errsbc,
            s33_target=P.scauchy[8], s33=inc.sgavg[8], s33_sol=inc.sgsolavg[8],
            detF_min_sol=inc.detF_min_sol, detF_min_gas=inc.detF_min_gas,
            volinc_max_sol=inc.volinc_max_sol, volinc_max_gas=inc.volinc_max_gas,
            stretchinc_max_sol=inc.stretchinc_max_sol, stretchinc_max_gas=inc.stretchinc_max_gas,
            rotang_max_sol=inc.rotang_max_sol, rotang_max_gas=inc.rotang_max_gas,
            dL_step_rms_sol=inc.dL_step_rms_sol, dL_step_rms_gas=inc.dL_step_rms_gas,
            dL_step_rms=inc.dL_step_rms, dL_step_max_sol=inc.dL_step_max_sol,
            dL_step_max_gas=inc.dL_step_max_gas, dL_step_max=inc.dL_step_max,
            dvelgradavg_mag=inc.dvelgradavg_mag, cond_c066mod_max_sol=inc.cond_c066mod_max_sol,
            cond_c066mod_max_gas=inc.cond_c066mod_max_gas, cond_cgas_max=inc.cond_cgas_max,
            imin_sol=list(inc.imin_sol), imin_gas=list(inc.imin_gas),
            passes=passes, nr_steps=nr, evals=ev, outer=ou, rollbacks=rollbacks,
        )
        pv = self.prev
        row["detF_min_sol_rel"] = _safe_rel(inc.detF_min_sol, pv["sol"])
        row["detF_min_gas_rel"] = _safe_rel(inc.detF_min_gas, pv["gas"])
        row["volinc_max_sol_rel"] = _safe_rel(inc.volinc_max_sol, pv["vs"])
        row["volinc_max_gas_rel"] = _safe_rel(inc.volinc_max_gas, pv["vg"])
        row["stretchinc_max_sol_rel"] = _safe_rel(inc.stretchinc_max_sol, pv["ss"])
        row["stretchinc_max_gas_rel"] = _safe_rel(inc.stretchinc_max_gas, pv["sg"])
        pv.update(sol=inc.detF_min_sol, gas=inc.detF_min_gas, vs=inc.volinc_max_sol, vg=inc.volinc_max_gas,
                  ss=inc.stretchinc_max_sol, sg=inc.stretchinc_max_gas)
        # ramp tdot back up (LS-EVPFFT.f90:557-581)
        if P.ictrl == -1:
            if self.tdot_plan < P.tdotref and self.nconverged >= 5:
                self.nconverged = 0
                self.tdot_plan = min(2.0 * self.tdot_plan, P.tdotref)
                rem = max(0.0, self.proc_time_target - self.proc_time_done)
                left = math.ceil(rem / self.tdot_plan)
                nnew = self.imicro + left
                self.total_macrosteps += nnew - P.nsteps
                P.nsteps = nnew
                P.tdot = min(self.tdot_plan, rem) if rem > 0.0 else self.tdot_plan
        elif P.tdot < P.tdotref and self.nconverged >= 5:
            self.nconverged = 0
            t0 = P.tdot
            P.tdot = min(2.0 * P.tdot, P.tdotref)
            left = P.nsteps - self.imicro
            left = math.ceil(left * t0 / P.tdot)
            self.total_macrosteps += left - (P.nsteps - self.imicro)
            P.nsteps = self.imicro + left
        if did_converge:
            self.nconverged += 1
        self._chk(L.cg_calc_c0(h, C.byref(P)), "cg_calc_c0")
        self.imacroloop += 1
        self.passes_total += passes
        return row

    def _replan(self, P, tdot):
        rem = max(0.0, self.proc_time_target - self.proc_time_done)
        self.tdot_plan = tdot
        nnew = (self.imicro - 1) + math.ceil(rem / self.tdot_plan)
        self.total_macrosteps += nnew - P.nsteps
        P.nsteps = nnew
        P.tdot = min(self.tdot_plan, rem)

    def run(self, max_steps=None, out_dir=None):
        rows = []
        writer = CsvWriter(out_dir, self) if out_dir else None
        while max_steps is None or len(rows) < max_steps:
            r = self.step()
            if r is None:
                break
            rows.append(r)
            if writer:
                writer.write(r)
        return rows

    def close(self):
        self.ctx.close()


def _g0(x):
    """Compact list-directed-like number format for CSV cells."""
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    return repr(float(x))


class CsvWriter:
    """results.csv / diagnostics.csv with the reference's headers and row semantics
    (header + an initial row at step 0, then one row per accepted increment, flushed)."""

    def __init__(self, out_dir, solver):
        d = Path(out_dir)
        d.mkdir(parents=True, exist_ok=True)
        self.res = d / "results.csv"
        self.diag = d / "diagnostics.csv"
        self.res.write_text(RESULTS_HEADER + "\n")
        self.diag.write_text(DIAG_HEADER + "\n")
        zero = [0] * 3 + [0.0] * 3
        # initial row (LS-EVPFFT.f90:176-178): the averages of init_micro are zero except
        # dvmavg = vm_strain(dsim of process 1) (initial_guess, various_functions.f90:50)
        d = np.asarray(solver.procs[0].dsim[:], dtype=np.float64).reshape(3, 3)
        h = (d[0, 0] + d[1, 1] + d[2, 2]) / 3.0
        dvm0 = math.sqrt(2.0 / 3.0 * ((d[0, 0] - h) ** 2 + (d[1, 1] - h) ** 2 + (d[2, 2] - h) ** 2 + d[0, 1] ** 2
                                      + d[0, 2] ** 2 + d[1, 2] ** 2 + d[1, 0] ** 2 + d[2, 0] ** 2 + d[2, 1] ** 2))
        res0 = [0, 0, 0] + [0.0] * 4 + [0.0] * 36 + [0.0, 0.0, dvm0, 0.0, 0.0, 0.0, 0.0, 0.0]
        with open(self.res, "a") as f:
            f.write(",".join(_g0(x) for x in res0) + "\n")
        with open(self.diag, "a") as f:
            f.write(",".join(_g0(x) for x in zero + [0.0] * 32 + [0] * 6) + "\n")

    def write(self, r):
        vals = [r["iter"], r["imacroloop"], r["ipr"], r["sim_time"], r["wall_time"], r["tdot"], r["tdot_ref"]]
        for k in ("eavg", "sgsolavg", "epavg", "eelavg", "velgradavg", "edotpavg"):
            vals += r[k]
        vals += [r[k] for k in ("evm", "evmp", "dvm", "dvmp", "svm", "gamma", "g", "geff")]
        with open(self.res, "a") as f:
            f.write(",".join(_g0(x) for x in vals) + "\n")
        d = [r["iter"], r["imacroloop"], r["ipr"], r["sim_time"], r["wall_time"], r["tdot"], r["devm_inc_max"],
             r["diag_tdotref"], r["erre"], r["errs"], r["errsbc"], r["s33_target"], r["s33"], r["s33_sol"],
             r["detF_min_sol"], r["detF_min_gas"], r["detF_min_sol_rel"], r["detF_min_gas_rel"],
             r["volinc_max_sol"], r["volinc_max_gas"], r["volinc_max_sol_rel"], r["volinc_max_gas_rel"],
             r["stretchinc_max_sol"], r["stretchinc_max_gas"], r["stretchinc_max_sol_rel"],
             r["stretchinc_max_gas_rel"], r["rotang_max_sol"], r["rotang_max_gas"], r["dL_step_rms_sol"],
             r["dL_step_rms_gas"], r["dL_step_rms"], r["dL_step_max_sol"], r["dL_step_max_gas"],
             r["dL_step_max"], r["dvelgradavg_mag"], r["cond_c066mod_max_sol"], r["cond_c066mod_max_gas"],
             r["cond_cgas_max"]] + list(r["imin_sol"]) + list(r["imin_gas"])
        with open(self.diag, "a") as f:
            f.write(",".join(_g0(x) for x in d) + "\n")
