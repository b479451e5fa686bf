"""Independent RVEs driven together (creep_modeling_amd.batch) compute exactly what each
computes alone: in the batched rounds (one launch per FFT pass for all RVEs of a grid, one
constitutive launch and one set of sums over the per-RVE work tables, RVEs retiring as their
runs end; lengths without register FFT kernels on their own streams), in FP64 and in the mixed
form, and in the thread-pool engine."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("engine,precision", [("shared", 0), ("shared", 1), ("threads", 0)])
def test_batch_matches_single_runs(engine, precision):
    from creep_modeling_amd import batch, driver, inputs, parse

    cases = inputs.sweep_cases(5, nsteps=2, sizes=(16, 24, 32))
    problems = [parse.problem_from_case(*c) for _, c in cases]
    # uneven run lengths: RVE 1 stops after one increment and leaves the work table early
    steps = [2, 1, 2, 2, 2]
    print("batch of", [p.n for p in problems], engine, precision, flush=True)
    single = []
    for p, k in zip(problems, steps):
        s = driver.Solver(p, precision=precision)
        single.append(s.run(max_steps=k))
        s.close()
    fields = {}

    def keep(i, s):
        fields[i] = s.ctx.field("sg")

    if engine == "shared":
        solvers = [driver.Solver(p, precision=precision) for p in problems]
        for s, k in zip(solvers, steps):  # per-RVE run lengths through the process tables
            s.procs[0].nsteps = k
        batched = batch.run_batch(problems, solvers=solvers, on_done=keep)
    else:
        solvers = [driver.Solver(p) for p in problems]
        for s, k in zip(solvers, steps):
            s.procs[0].nsteps = k
        batched = batch.run_batch(problems, workers=4, solvers=solvers, on_done=keep, engine="threads")
    for a, b, k in zip(single, batched, steps):
        assert len(a) == len(b) == k
        for ra, rb in zip(a, b):
            assert ra["passes"] == rb["passes"]
            for key in ("sgsolavg", "edotpavg", "eavg", "velgradavg", "eelavg"):
                assert np.array_equal(np.asarray(ra[key]), np.asarray(rb[key])), key
    assert sorted(fields) == [0, 1, 2, 3, 4]
