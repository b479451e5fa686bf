"""Independent RVEs on one GPU: a calibration or microstructure sweep in one process.

The reference runs its calibration and microstructure sweeps as independent solver
processes (``utils/calibrate.py:1054-1058``, ``utils/sweep_microstructures.py:24-29``,
``utils/run.py:183-196``).  Here every RVE is one context of the library and the RVEs advance
together, one fixed-point pass of every RVE per round (``cg_batch_iterate``): each RVE's
spectral stage is enqueued on its own stream and the correction + constitutive kernel, most of
a pass, runs as one launch over a per-RVE work table, so small RVEs together fill the GPU.
Between passes each RVE runs its own increment control (convergence, Delta-t cut-back,
rollback); an RVE whose increment has ended takes its accepted-increment updates and starts the
next one, and an RVE whose run is over retires from the work table while the others continue.
A batched RVE computes exactly what it computes alone.

``engine="threads"`` instead drives each RVE's ordinary loop from a pool of host threads
(one context and stream per RVE, ``workers`` in flight).
"""
from __future__ import annotations

import ctypes as C
import queue
import threading

from . import driver
from . import lib as _lib


def run_batch(problems, steps=None, workers=4, precision=0, solvers=None, on_done=None, engine="shared"):
    """Run `steps` accepted increments (all of them when None) of every problem.

    Returns the per-problem lists of row dicts, in input order.  `solvers`, if given, are
    already-created :class:`driver.Solver` objects (one per problem, closed here).
    `on_done(i, solver)`, if given, is called when RVE i has finished, before its context is
    closed (e.g. to read its fields)."""
    if engine == "threads":
        return _run_threads(problems, steps, workers, precision, solvers, on_done)
    n = len(problems)
    if solvers is None:
        solvers = [driver.Solver(p, precision=precision) for p in problems]
    L = _lib.load()
    handles = (C.c_void_p * n)(*[s.h.value for s in solvers])
    b = C.c_void_p()
    try:
        if L.cg_batch_create(handles, n, C.byref(b)) != 0:
            raise _lib.HipError("cg_batch_create: %s" % L.cg_last_error(None).decode())
        results = [[] for _ in range(n)]
        live = [False] * n

        def start(i):
            s = solvers[i]
            live[i] = (steps is None or len(results[i]) < steps) and s.begin()
            if not live[i] and on_done is not None:
                on_done(i, s)

        for i in range(n):
            start(i)
        procs = (C.POINTER(_lib.CgProcess) * n)()
        sts = (_lib.CgIterState * n)()
        outs = (_lib.CgIterOut * n)()
        active = (C.c_int * n)()
        rc = (C.c_int * n)()
        while any(live):
            for i in range(n):
                active[i] = int(live[i])
                if live[i]:
                    s = solvers[i]
                    procs[i] = C.pointer(s.P)
                    C.memmove(C.byref(sts, i * C.sizeof(_lib.CgIterState)), C.byref(s.st), C.sizeof(_lib.CgIterState))
            if L.cg_batch_iterate(b, procs, sts, outs, active, rc) != 0:
                raise _lib.HipError("cg_batch_iterate: %s" % L.cg_last_error(None).decode())
            for i in range(n):
                if not live[i]:
                    continue
                s = solvers[i]
                if rc[i] != 0:
                    raise _lib.HipError("RVE %d: cg_iterate rc=%d: %s" % (i, rc[i], L.cg_last_error(s.h).decode()))
                C.memmove(C.byref(s.st), C.byref(sts, i * C.sizeof(_lib.CgIterState)), C.sizeof(_lib.CgIterState))
                C.memmove(C.byref(s.out), C.byref(outs, i * C.sizeof(_lib.CgIterOut)), C.sizeof(_lib.CgIterOut))
                if s.pass_done():
                    results[i].append(s.finish())
                    start(i)
        return results
    finally:
        if b.value:
            L.cg_batch_destroy(b)
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
