"""Independent RVEs on one GPU, several in flight at a time.

The reference runs its calibration and microstructure sweeps as independent solver
processes (``utils/calibrate.py:1054-1058``, ``utils/sweep_microstructures.py:24-29``,
``utils/run.py:183-196``).  Here each RVE is one context of the library with its own HIP
stream; a pool of host threads drives several contexts at once (the library's calls release
the interpreter lock), so RVEs too small to fill the GPU share it, and each RVE retires as
soon as its increments are done.  Contexts are independent: a batched RVE computes exactly
what it computes alone.
"""
from __future__ import annotations

import queue
import threading

from . import driver


def run_batch(problems, steps=None, workers=4, precision=0, solvers=None, on_done=None):
    """Run `steps` accepted increments (all of them when None) of every problem.

    Returns the per-problem lists of row dicts, in input order.  The work queue is ordered
    largest RVE first.  `solvers`, if given, are already-created :class:`driver.Solver`
    objects (one per problem, closed here).  `on_done(i, solver)`, if given, is called when
    RVE i has finished, before its context is closed (e.g. to read its fields)."""
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
