"""The drop-in LS-EVPFFT executable (creep_modeling_amd/host): the reference's command
line (LS-EVPFFT.f90:41-139) on the CPU, and a whole run through the reference's input
files and CSV outputs on the GPU, checked against the oracle and the Python driver."""
import csv
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "creep_modeling_amd" / "host" / "LS-EVPFFT"


def run(args, timeout=60):
    assert EXE.exists(), "LS-EVPFFT is not built (make -C creep_modeling_amd/host)"
    return subprocess.run([str(EXE)] + [str(a) for a in args], capture_output=True, text=True, timeout=timeout)


def test_help_lists_the_reference_flags():
    r = run(["-h"])
    assert r.returncode == 0
    for flag in ("--in", "--out", "--nthreads", "--fftw_save", "--modal", "--mix", "--precision"):
        assert flag in r.stdout, flag


def test_unknown_flag_stops_with_status_zero():
    r = run(["--bogus"])
    assert r.returncode == 0
    assert "Unknown command-line option" in r.stdout


def read_csv(path):
    with open(path) as f:
        rows = list(csv.reader(f))
    return rows[0], [[float(x) for x in r] for r in rows[1:]]


@pytest.mark.gpu
def test_executable_run_matches_oracle_and_driver(tmp_path):
    from creep_modeling_amd import driver, inputs, parse
    from oracle import bind

    ms, mat, procs = inputs.config_case(0, n=8, ngrains=4, nsteps=2)
    fft_in = inputs.write_case(tmp_path / "case", ms, mat, procs)
    out = tmp_path / "out"
    print("LS-EVPFFT on", fft_in, flush=True)
    r = run(["--in", fft_in, "--out", out, "--nthreads", 4, "--mix", 1, "--modal", 0,
             "--fftw_save", tmp_path / "wisdom" / "w.dat"], timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    head, rows = read_csv(out / "results.csv")
    dhead, drows = read_csv(out / "diagnostics.csv")
    assert head == driver.RESULTS_HEADER.split(",")
    assert dhead == driver.DIAG_HEADER.split(",")
    assert len(rows) == 3 and len(drows) == 3  # initial row + one per increment
    assert (out / "microstr_cell_fields_stp_00000.vtk").exists()

    # the same case through the Python driver (same library) and the oracle
    problem = parse.read_fft_in(fft_in)
    g = driver.Solver(problem)
    gw = driver.CsvWriter(tmp_path / "py", g)
    o = bind.Oracle(problem)
    orows = []
    for _ in range(2):
        gw.write(g.step())
        orows.append(o.step().as_dict())
    g.close()
    o.close()
    _, prow = read_csv(tmp_path / "py" / "results.csv")
    col = {k: i for i, k in enumerate(head)}
    wall = col["wall_time"]
    for a, b in zip(rows, prow):
        a = np.array(a); b = np.array(b)
        a[wall] = b[wall] = 0.0
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-300)
    for a, ro in zip(rows[1:], orows):
        assert a[col["iter"]] == pytest.approx(ro["iter"], abs=1)
        for name, key, c in (("sav33", "sgsolavg", 2), ("edotp33", "edotpavg", 2), ("eav33", "eavg", 2)):
            assert a[col[name]] == pytest.approx(ro[key][c], rel=1e-8), name
