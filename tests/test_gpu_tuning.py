"""FFT shape tuning (cg_set_tuning_file / cg_get_fft_tuning / cg_set_fft_tuning), the
counterpart of the reference's FFTW planning and wisdom file (fourier_functions.f90:35-194):
every shape computes the same values in the same order, so results are bitwise equal whatever
the tuner picks, and a tuning file written by one run is read by the next."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ("sgsolavg", "edotpavg", "eavg", "velgradavg", "eelavg")


def _problem(n=(32, 32, 48), ngrains=12, nsteps=2):
    from creep_modeling_amd import inputs, parse

    ms, mat, procs = inputs.config_case(0, n=n[0], ngrains=ngrains, nsteps=nsteps)
    if n[1:] != (n[0], n[0]):
        ms = inputs.voronoi_rve(n, ngrains, 77)
    return parse.problem_from_case(ms, mat, procs)


@pytest.mark.parametrize("precision", [0, 1])
def test_every_fft_shape_gives_bitwise_equal_results(precision):
    from creep_modeling_amd import driver

    p = _problem()
    runs = []
    for shapes in ([1, 1, 1, 1, 1, 0, 0, 0], [0, 0, 0, 0, 0, 1, 1, 1], None):
        print("shapes", shapes, "precision", precision, flush=True)
        s = driver.Solver(p, precision=precision, fft_shapes=shapes)
        rows = s.run(max_steps=2)
        tuned, got, us = s.ctx.fft_tuning()
        if shapes is None:
            assert tuned and all(x > 0 for x in us), (tuned, us)
        else:
            assert got == shapes
        runs.append((rows, s.ctx.field("sg"), s.ctx.field("velgrad")))
        s.close()
    for rows, sg, L in runs[1:]:
        assert [r["passes"] for r in rows] == [r["passes"] for r in runs[0][0]]
        for ra, rb in zip(runs[0][0], rows):
            for k in KEYS:
                assert np.array_equal(np.asarray(ra[k]), np.asarray(rb[k])), k
        assert np.array_equal(sg, runs[0][1]) and np.array_equal(L, runs[0][2])


def test_tuning_file_written_then_read(tmp_path):
    from creep_modeling_amd import driver

    p = _problem(n=(32, 32, 32), nsteps=1)
    wf = tmp_path / "sub" / "fftw_wisdom.save"
    wf.parent.mkdir()
    s = driver.Solver(p, tune_file=wf)
    r1 = s.run(max_steps=1)
    t1, shapes1, us1 = s.ctx.fft_tuning()
    s.close()
    lines = [ln for ln in wf.read_text().splitlines() if ln.startswith("creep_hip_fft_tune ")]
    assert len(lines) == 1 and " 32 32 32 0 0 : " in lines[0], lines
    assert t1 and all(u > 0 for u in us1)
    # a second run reads the line instead of timing (no times), with the same shapes and results
    s = driver.Solver(p, tune_file=wf)
    r2 = s.run(max_steps=1)
    t2, shapes2, us2 = s.ctx.fft_tuning()
    s.close()
    assert t2 and shapes2 == shapes1 and us2 == [0.0] * 5
    assert len([ln for ln in wf.read_text().splitlines() if ln.startswith("creep_hip_fft_tune ")]) == 1
    for k in KEYS:
        assert np.array_equal(np.asarray(r1[0][k]), np.asarray(r2[0][k])), k


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replays_match_eager_launches(precision):
    """Every pass replays the captured FFT-stage and reduction graphs between eager solve
    launches and host synchronisations (and the stage graph is captured again when C0 changes
    at the end of each increment); all rows and the final fields equal the eager run bitwise."""
    from creep_modeling_amd import driver

    p = _problem(nsteps=3)
    runs = []
    for graphs in (False, True):
        print("graphs", graphs, "precision", precision, flush=True)
        s = driver.Solver(p, precision=precision, graphs=graphs, fft_shapes=[1, 1, 1, 1, 1, 0, 0, 0])
        rows = s.run(max_steps=3)
        runs.append((rows, s.ctx.field("sg"), s.ctx.field("velgrad"), s.ctx.field("trialtau")))
        s.close()
    (ra, *fa), (rb, *fb) = runs
    assert [r["passes"] for r in ra] == [r["passes"] for r in rb]
    assert sum(r["passes"] for r in ra) > 10
    for x, y in zip(ra, rb):
        for k in KEYS:
            assert np.array_equal(np.asarray(x[k]), np.asarray(y[k])), k
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("precision", [0, 1])
def test_overlapped_passes_match_serial_passes(precision):
    """With overlap on (the default) each cg_iterate enqueues the next pass' x, y, z + Gamma and
    y transforms while the host takes the pass' sums, and the next pass runs only the x inverse;
    the early transforms are dropped at the end of each increment (C0 changes) and when a field
    is written between passes.  Rows and final fields equal the serial run bitwise; the stage
    timers still count the early transforms."""
    import ctypes as C

    from creep_modeling_amd import driver

    p = _problem(nsteps=3)
    runs = []
    for overlap in (False, True):
        print("overlap", overlap, "precision", precision, flush=True)
        s = driver.Solver(p, precision=precision, overlap=overlap, fft_shapes=[1, 1, 1, 1, 1, 0, 0, 0])
        s.L.cg_set_timing(s.h, 1)
        rows = [s.step()]
        # a state write between passes of the second increment (sigma rewritten with itself)
        assert s.begin()
        for _ in range(3):
            s._chk(s.L.cg_iterate(s.h, C.byref(s.P), C.byref(s.st), C.byref(s.out)), "cg_iterate")
            assert not s.pass_done()
        s.ctx.set_field("sg", s.ctx.field("sg"))
        while True:
            s._chk(s.L.cg_iterate(s.h, C.byref(s.P), C.byref(s.st), C.byref(s.out)), "cg_iterate")
            if s.pass_done():
                rows.append(s.finish())
                break
        rows.append(s.step())
        st = s.ctx.stats()
        assert st.t_ms[0] > 0 and st.t_ms[1] > 0
        runs.append((rows, s.ctx.field("sg"), s.ctx.field("velgrad"), s.ctx.field("trialtau")))
        s.close()
    (ra, *fa), (rb, *fb) = runs
    assert [r["passes"] for r in ra] == [r["passes"] for r in rb]
    for x, y in zip(ra, rb):
        for k in KEYS:
            assert np.array_equal(np.asarray(x[k]), np.asarray(y[k])), k
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y)
