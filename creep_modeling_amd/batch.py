"""Independent RVEs on one GPU: a calibration or microstructure sweep in one process.

The reference runs its calibration and microstructure sweeps as independent solver
processes (``utils/calibrate.py:1054-1058``, ``utils/sweep_microstructures.py:24-29``,
``utils/run.py:183-196``).  Here every RVE is one context of the library and the RVEs advance
together, one fixed-point pass of every RVE per round (``cg_batch_launch`` / ``cg_batch_wait``):
each FFT pass, the sums and the correction + constitutive kernel run as one launch per kernel
over per-RVE work tables (one set of FFT launches per grid size), so small RVEs together fill
the GPU.  The RVEs are split into two batches launched in turn, so the host's increment
control of one batch overlaps the other's pass on the GPU.  Between passes each RVE runs its
own increment control (convergence, Delta-t cut-back, rollback); an RVE whose increment has
ended takes its accepted-increment updates and starts the next one, and an RVE whose run is
over retires from the work tables while the others continue.  A batched RVE computes exactly
what it computes alone.

``engine="threads"`` instead drives each RVE's ordinary loop from a pool of host threads
(one context and stream per RVE, ``workers`` in flight).
"""
from __future__ import annotations

import ctypes as C
import queue
import threading

from . import driver
from . import lib as _lib


def run_batch(problems, steps=None, workers=4, precision=0, solvers=None, on_done=None, engine="shared", groups=2):
    """Run `steps` accepted increments (all of them when None) of every problem.

    Returns the per-problem lists of row dicts, in input order.  `solvers`, if given, are
    already-created :class:`driver.Solver` objects (one per problem, closed here).
    `on_done(i, solver)`, if given, is called when RVE i has finished, before its context is
    closed (e.g. to read its fields).  The shared engine splits the RVEs into `groups` batches of
    about equal voxel counts, launched in turn: while the GPU runs one batch's pass the host
    does the increment control of the other's."""
    if engine == "threads":
        return _run_threads(problems, steps, workers, precision, solvers, on_done)
    n = len(problems)
    if solvers is None:
        solvers = [driver.Solver(p, precision=precision) for p in problems]
    L = _lib.load()
    # groups of about equal voxel counts (largest RVE first into the lightest group)
    ng = max(1, min(groups, n))
    members = [[] for _ in range(ng)]
    load = [0] * ng
    for i in sorted(range(n), key=lambda i: -problems[i].nvox):
        g = load.index(min(load))
        members[g].append(i)
        load[g] += problems[i].nvox
    batches = []
    try:
        for mem in members:
            k = len(mem)
            handles = (C.c_void_p * k)(*[solvers[i].h.value for i in mem])
            b = C.c_void_p()
            if L.cg_batch_create(handles, k, C.byref(b)) != 0:
                raise _lib.HipError("cg_batch_create: %s" % L.cg_last_error(None).decode())
            batches.append(dict(b=b, mem=mem, procs=(C.POINTER(_lib.CgProcess) * k)(), sts=(_lib.CgIterState * k)(),
                                outs=(_lib.CgIterOut * k)(), active=(C.c_int * k)(), rc=(C.c_int * k)(),
                                pending=False))
        results = [[] for _ in range(n)]
        live = [False] * n

        def start(i):
            s = solvers[i]
            live[i] = (steps is None or len(results[i]) < steps) and s.begin()
            if not live[i] and on_done is not None:
                on_done(i, s)

        def launch(bt):
            mem, any_live = bt["mem"], False
            for j, i in enumerate(mem):
                bt["active"][j] = int(live[i])
                if live[i]:
                    any_live = True
                    s = solvers[i]
                    bt["procs"][j] = C.pointer(s.P)
                    C.memmove(C.byref(bt["sts"], j * C.sizeof(_lib.CgIterState)), C.byref(s.st),
                              C.sizeof(_lib.CgIterState))
            if not any_live:
                return
            if L.cg_batch_launch(bt["b"], bt["procs"], bt["sts"], bt["outs"], bt["active"]) != 0:
                raise _lib.HipError("cg_batch_launch: %s" % L.cg_last_error(None).decode())
            bt["pending"] = True

        def finish(bt):
            bt["pending"] = False
            if L.cg_batch_wait(bt["b"], bt["rc"]) != 0:
                raise _lib.HipError("cg_batch_wait: %s" % L.cg_last_error(None).decode())
            for j, i in enumerate(bt["mem"]):
                if not bt["active"][j]:
                    continue
                s = solvers[i]
                if bt["rc"][j] != 0:
                    raise _lib.HipError("RVE %d: cg_iterate rc=%d: %s" % (i, bt["rc"][j], L.cg_last_error(s.h).decode()))
                C.memmove(C.byref(s.st), C.byref(bt["sts"], j * C.sizeof(_lib.CgIterState)),
                          C.sizeof(_lib.CgIterState))
                C.memmove(C.byref(s.out), C.byref(bt["outs"], j * C.sizeof(_lib.CgIterOut)), C.sizeof(_lib.CgIterOut))
                if s.pass_done():
                    results[i].append(s.finish())
                    start(i)

        for i in range(n):
            start(i)
        for bt in batches:
            launch(bt)
        while any(bt["pending"] for bt in batches):
            for bt in batches:
                if bt["pending"]:
                    finish(bt)
                    launch(bt)
        return results
    finally:
        for bt in batches:
            if bt["pending"]:
                L.cg_batch_wait(bt["b"], bt["rc"])
            L.cg_batch_destroy(bt["b"])
        for s in solvers:
            s.close()


def _run_threads(problems, steps, workers, precision, solvers, on_done):
    n = len(problems)
    results = [None] * n
    order = sorted(range(n), key=lambda i: -problems[i].nvox)
    work = queue.Queue()
    for i in order:
        work.put(i)
    errors = []

    def worker():
        while not errors:
            try:
                i = work.get_nowait()
            except queue.Empty:
                return
            s = None
            try:
                s = solvers[i] if solvers is not None else driver.Solver(problems[i], precision=precision)
                results[i] = s.run(max_steps=steps)
                if on_done is not None:
                    on_done(i, s)
            except BaseException as e:  # reported after the pool drains
                errors.append((i, e))
            finally:
                if s is not None:
                    s.close()

    threads = [threading.Thread(target=worker, daemon=True) for _ in range(max(1, min(workers, n)))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if solvers is not None:  # solvers the pool did not reach (after an error)
        for i in range(n):
            if results[i] is None:
                try:
                    solvers[i].close()
                except Exception:
                    pass
    if errors:
        i, e = errors[0]
        raise RuntimeError("RVE %d of the batch failed: %s" % (i, e)) from e
    return results
