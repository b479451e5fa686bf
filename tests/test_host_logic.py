"""Host-side logic that needs no GPU: the batched-sweep generator and the batch runner's
work distribution and error handling (with stand-in solvers)."""
import math
import threading

import pytest


def test_sweep_cases_ranges_and_determinism():
    from creep_modeling_amd import inputs

    a = inputs.sweep_cases(12, nsteps=3, sizes=(8, 16))
    b = inputs.sweep_cases(12, nsteps=3, sizes=(8, 16))
    assert [p for p, _ in a] == [p for p, _ in b]
    for prm, (ms, mat, procs) in a:
        assert prm["n"] in (8, 16) and ms.n == (prm["n"],) * 3
        assert 32 <= prm["grains"] <= 512
        assert 200.0 <= prm["s33"] <= 400.0 and procs[0].s33 == prm["s33"]
        assert 8.0 <= prm["nrs"] <= 14.0 and mat.modes[0].nrs == prm["nrs"]
        assert prm["temp_c"] in (475.0, 525.0, 575.0)
        # Arrhenius factor: 1 at 525 C, exp(-Q/R (1/T - 1/Tref)) otherwise
        t = prm["temp_c"] + 273.15
        fac = math.exp(-280.0e3 / 8.314462618 * (1.0 / t - 1.0 / (525.0 + 273.15)))
        assert mat.modes[0].gamd0 == pytest.approx(1.0e-3 * fac, rel=1e-14)
        assert procs[0].nsteps == 3
    assert any(p["temp_c"] == 575.0 for p, _ in a) or any(p["temp_c"] == 475.0 for p, _ in a)


class _FakeSolver:
    """Stands in for driver.Solver: `steps` rows of which passes = problem id."""

    live = 0
    peak = 0
    lock = threading.Lock()

    def __init__(self, problem, precision=0):
        self.p = problem
        with _FakeSolver.lock:
            _FakeSolver.live += 1
            _FakeSolver.peak = max(_FakeSolver.peak, _FakeSolver.live)

    def run(self, max_steps=None):
        if self.p.fail:
            raise ValueError("boom %d" % self.p.id)
        import time

        time.sleep(0.01)
        return [dict(passes=self.p.id)] * (max_steps or 2)

    def close(self):
        with _FakeSolver.lock:
            _FakeSolver.live -= 1


class _P:
    def __init__(self, i, nvox, fail=False):
        self.id, self.nvox, self.fail = i, nvox, fail


def test_run_batch_order_and_pool(monkeypatch):
    from creep_modeling_amd import batch, driver

    monkeypatch.setattr(driver, "Solver", _FakeSolver)
    _FakeSolver.peak = 0
    probs = [_P(i, nvox=(i * 7) % 5 + 1) for i in range(9)]
    out = batch.run_batch(probs, steps=3, workers=3, engine="threads")
    assert [r[0]["passes"] for r in out] == list(range(9))  # input order
    assert all(len(r) == 3 for r in out)
    assert _FakeSolver.peak <= 3 and _FakeSolver.live == 0


def test_run_batch_reports_failure(monkeypatch):
    from creep_modeling_amd import batch, driver

    monkeypatch.setattr(driver, "Solver", _FakeSolver)
    probs = [_P(0, 1), _P(1, 1, fail=True), _P(2, 1)]
    with pytest.raises(RuntimeError, match="RVE 1"):
        batch.run_batch(probs, workers=2, engine="threads")
    assert _FakeSolver.live == 0
