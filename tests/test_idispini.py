"""idispini > 0: an initial displacement field (IO_functions.f90:147-154, init_defgrad :815-936).

CPU: the reader; the oracle's spectral gradient_tensor1 (various_functions.f90:56-112) against
the real-space stencils the multipliers stand for -- i sin(xi) is the central difference and
tan(xi/2) (i/4) prod(1 + e^{i xi}) the mean forward difference over a cell's four edges, because
e^{i xi} shifts by one voxel -- on random periodic fields (3D and one-layer grids); the oracle's
init_defgrad against F = grad(u0) + 4 I, J, w_c = J / sum J and the nodal shift recomputed here.
GPU: the HIP path (and the executable) against the oracle with cell and node displacements.
"""
import numpy as np
import pytest


def stencil_grad(u, ideriv):
    """grad[v, l, m] = d_m u_l of u[n1, n2, n3, 3] (x fastest in the flattened order) by the
    periodic real-space stencils."""
    sh = lambda a, ax, s: np.roll(a, -s, axis=ax)  # a(x + s e_ax)
    g = np.zeros(u.shape[:3] + (3, 3))
    for m in range(3):
        if ideriv == 1:
            g[..., :, m] = 0.5 * (sh(u, m, 1) - sh(u, m, -1))
        else:
            o = [a for a in range(3) if a != m]
            acc = 0.0
            for p in (0, 1):
                for q in (0, 1):
                    base = sh(sh(u, o[0], p), o[1], q)
                    acc = acc + sh(base, m, 1) - base
            g[..., :, m] = 0.25 * acc
    return g


def to_grid(flat, n):
    n1, n2, n3 = n
    return flat.reshape(n3, n2, n1, 3).transpose(2, 1, 0, 3)


def from_grid(g):
    return g.transpose(2, 1, 0, 3, 4).reshape(-1, 9)


@pytest.mark.parametrize("n", [(8, 12, 6), (12, 6, 1)])
@pytest.mark.parametrize("ideriv", [1, 2])
def test_oracle_spectral_gradient_is_the_stencil(n, ideriv):
    from oracle import bind

    L = bind.lib()
    rng = np.random.default_rng(11)
    V = n[0] * n[1] * n[2]
    u = rng.standard_normal((V, 3))
    grad = np.zeros((V, 9))
    L.orc_gradient_tensor1(bind.dp(u), bind.dp(grad), ideriv, *n)
    want = from_grid(stencil_grad(to_grid(u, n), ideriv))
    np.testing.assert_allclose(grad, want, rtol=0, atol=1e-13 * np.abs(want).max())
    # a plausible slip, the derivative index first, fails
    swapped = from_grid(stencil_grad(to_grid(u, n), ideriv).swapaxes(-1, -2))
    assert np.abs(grad - swapped).max() > 1e-3


def disp_case(tmp_path, idispini, n=(8, 8, 8), amp=0.05, seed=5, nsteps=2):
    from creep_modeling_amd import inputs

    ms, mat, procs = inputs.config_case(0, n=n[0], ngrains=4, nsteps=nsteps)
    if tuple(n) != (n[0],) * 3:
        ms = inputs.voronoi_rve(tuple(n), 4, 9)
    n1, n2, n3 = n
    rng = np.random.default_rng(seed)
    if idispini == 1:
        rows = n1 * n2 * n3
    else:
        rows = (n1 + 1) * (n2 + 1) * (1 if n3 == 1 else n3 + 1)
    # smooth periodic field: a few low modes
    d = np.zeros((rows, 3))
    if idispini == 1:
        k, j, i = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    else:
        k, j, i = np.meshgrid(np.arange(1 if n3 == 1 else n3 + 1), np.arange(n2 + 1), np.arange(n1 + 1),
                              indexing="ij")
    x = np.stack([i.ravel() / n1, j.ravel() / n2, k.ravel() / max(n3, 1)], 1)
    for l in range(3):
        for _ in range(3):
            kv = rng.integers(-2, 3, 3)
            d[:, l] += amp * rng.standard_normal() * np.sin(2 * np.pi * x @ kv + rng.uniform(0, 6.3))
    return inputs.write_case(tmp_path / ("case%d" % idispini), ms, mat, procs, displacement=(idispini, d)), d


@pytest.mark.parametrize("idispini", [1, 2])
def test_reader_returns_the_displacement(tmp_path, idispini):
    from creep_modeling_amd import parse

    fft_in, d = disp_case(tmp_path, idispini)
    p = parse.read_fft_in(fft_in)
    assert p.idispini == idispini
    np.testing.assert_array_equal(p.dispini, d)


@pytest.mark.parametrize("idispini", [1, 2])
def test_oracle_init_defgrad_closed_form(tmp_path, idispini):
    from creep_modeling_amd import parse
    from oracle import bind

    n = (8, 8, 8)
    fft_in, d = disp_case(tmp_path, idispini, n=n)
    p = parse.read_fft_in(fft_in)
    o = bind.Oracle(p)
    F = o.field("defgrad").reshape(-1, 9)
    F0 = o.field("defgradini").reshape(-1, 9)
    J = o.field("detF")
    w = o.field("wgtc")
    xn = o.field("xnode").reshape(-1, 3)
    o.close()
    n1, n2, n3 = n
    if idispini == 1:
        ucell = d
    else:
        node = d.reshape(n3 + 1, n2 + 1, n1 + 1, 3)
        ucell = node[:n3, :n2, :n1].reshape(-1, 3)
    want = from_grid(stencil_grad(to_grid(ucell, n), idispini)) + 4.0 * np.eye(3).ravel()
    np.testing.assert_allclose(F, want, rtol=0, atol=1e-13)
    np.testing.assert_array_equal(F0, F)
    np.testing.assert_allclose(J, np.linalg.det(want.reshape(-1, 3, 3)), rtol=1e-13)
    np.testing.assert_allclose(w, J / J.sum(), rtol=1e-12)
    # nodal coordinates: grid + nodal displacement (idispini 2) or the mean of the 8 cells (1)
    k, j, i = np.meshgrid(np.arange(n3 + 1), np.arange(n2 + 1), np.arange(n1 + 1), indexing="ij")
    grid = np.stack([i.ravel() + 0.5, j.ravel() + 0.5, k.ravel() + 0.5], 1)
    if idispini == 2:
        shift = d
    else:
        c = to_grid(d, n)
        shift = np.zeros(((n3 + 1) * (n2 + 1) * (n1 + 1), 3))
        q = 0
        for kk in range(n3 + 1):
            for jj in range(n2 + 1):
                for ii in range(n1 + 1):
                    s = 0.0
                    for a in (ii - 1, ii):
                        for b in (jj - 1, jj):
                            for e in (kk - 1, kk):
                                s = s + c[a % n1, b % n2, e % n3]
                    shift[q] = s / 8
                    q += 1
    np.testing.assert_allclose(xn, grid + shift, rtol=0, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("idispini", [1, 2])
def test_gpu_initial_displacement_matches_oracle(tmp_path, idispini):
    from creep_modeling_amd import parse
    from parity import check_fields, check_rows, relerr, run_pair

    fft_in, _ = disp_case(tmp_path, idispini, n=(16, 16, 16))
    p = parse.read_fft_in(fft_in)
    rows, fo, fg = run_pair(p, 2, fields=("sg", "velgrad", "edotp", "trialtau", "defgrad", "defgradini", "defgrade",
                                          "detF", "wgtc", "xnode"))
    check_rows(rows)
    check_fields(fo, fg)
    assert relerr(fo["defgradini"], fg["defgradini"]) < 1e-13


@pytest.mark.gpu
def test_exe_initial_displacement_matches_oracle(tmp_path):
    import subprocess
    from pathlib import Path

    from creep_modeling_amd import parse
    from oracle import bind
    from test_gpu_host_paths import compare, read_csv

    fft_in, _ = disp_case(tmp_path, 2)
    exe = Path(__file__).resolve().parents[1] / "creep_modeling_amd" / "host" / "LS-EVPFFT"
    print("LS-EVPFFT with idispini = 2", flush=True)
    r = subprocess.run([str(exe), "--in", str(fft_in), "--out", str(tmp_path / "out"), "--modal", "0"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    head, rows = read_csv(tmp_path / "out" / "results.csv")
    o = bind.Oracle(parse.read_fft_in(fft_in))
    orows = [o.step().as_dict() for _ in range(2)]
    o.close()
    compare(head, rows[1:], orows)
