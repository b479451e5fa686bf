"""The FFT + Green-operator stage against rocFFT (torch.fft on the ROCm build), at the
production transform lengths.

A random symmetric stress field is written into a fresh context (F = I, J = 1, so the
polarisation is P' = sigma - <sigma>), one pass runs, and the stage's output dvelgradref is
compared with the same convolution done by torch: rfftn of P', Gamma(xi):P^ per frequency from
the context's C0 (convolution_with_Goper, various_functions.f90:966-1059, with form_G's
operator :640-742: Gamma_pqij = -A^-1_pi xi_q xi_j, A = C0 : xi xi; zero at k = 0, -S0 when any
index is at Nyquist), irfftn (the reference's Re(IFFT)/V).  rocFFT is used here only as the
checker; the library's own transforms are hand-written.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def torch_convolution(P, c0, s0, n):
    """P: (n3, n2, n1, 9) float64 torch tensor on the GPU (x fastest) -> dL^ref, same shape."""
    import torch

    n1, n2, n3 = n
    Ph = torch.fft.rfftn(P, dim=(0, 1, 2))  # (n3, n2, n1//2+1, 9), forward e^{-i xi x}
    kx = torch.arange(n1 // 2 + 1, device=P.device, dtype=torch.float64)
    ky = torch.fft.fftfreq(n2, d=1.0 / n2, device=P.device).to(torch.float64)
    kz = torch.fft.fftfreq(n3, d=1.0 / n3, device=P.device).to(torch.float64)
    Z, Y, X = torch.meshgrid(kz, ky, kx, indexing="ij")
    xi = torch.stack([2 * np.pi * X / n1, 2 * np.pi * Y / n2, 2 * np.pi * Z / n3], dim=-1)
    C4 = torch.as_tensor(c0.reshape(3, 3, 3, 3), device=P.device)
    A = torch.einsum("ijkl,...j,...l->...ik", C4, xi, xi)
    # adjugate inverse of the 3x3 acoustic tensors
    a = A.reshape(-1, 9)
    cof = torch.stack([
        a[:, 4] * a[:, 8] - a[:, 5] * a[:, 7], a[:, 2] * a[:, 7] - a[:, 1] * a[:, 8], a[:, 1] * a[:, 5] - a[:, 2] * a[:, 4],
        a[:, 5] * a[:, 6] - a[:, 3] * a[:, 8], a[:, 0] * a[:, 8] - a[:, 2] * a[:, 6], a[:, 2] * a[:, 3] - a[:, 0] * a[:, 5],
        a[:, 3] * a[:, 7] - a[:, 4] * a[:, 6], a[:, 1] * a[:, 6] - a[:, 0] * a[:, 7], a[:, 0] * a[:, 4] - a[:, 1] * a[:, 3],
    ], dim=1)
    det = a[:, 0] * cof[:, 0] + a[:, 1] * cof[:, 3] + a[:, 2] * cof[:, 6]
    det = torch.where(det == 0, torch.ones_like(det), det)
    Ai = (cof / det[:, None]).reshape(A.shape)
    Pm = Ph.reshape(*Ph.shape[:3], 3, 3)
    u = torch.einsum("...ij,...j->...i", Pm, xi.to(Pm.dtype))
    w = torch.einsum("...pi,...i->...p", Ai.to(Pm.dtype), u)
    out = -(w[..., :, None] * xi.to(Pm.dtype)[..., None, :])
    # any index at Nyquist: Gamma = -S0 (igamma = 0); k = 0: Gamma = 0
    S = torch.as_tensor(s0.reshape(9, 9), device=P.device).to(Pm.dtype)
    nyq = torch.zeros(X.shape, dtype=torch.bool, device=P.device)
    if n1 % 2 == 0:
        nyq |= X == n1 // 2
    if n2 % 2 == 0:
        nyq |= Y == -(n2 // 2)
    if n3 % 2 == 0:
        nyq |= Z == -(n3 // 2)
    out = out.reshape(*out.shape[:3], 9)
    out[nyq] = -(Ph[nyq] @ S.T)
    out[0, 0, 0] = 0
    return torch.fft.irfftn(out, s=(n3, n2, n1), dim=(0, 1, 2))


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n", [128, 256])
def test_green_stage_matches_rocfft(n, precision):
    import torch

    from creep_modeling_amd import driver, inputs, lib, parse

    print("rocFFT cross-check n=%d precision=%d" % (n, precision), flush=True)
    torch.cuda.init()  # torch's HIP runtime first, then the library's contexts
    ms, mat, procs = inputs.config_case(1, n=n, ngrains=64, nsteps=1)
    p = parse.problem_from_case(ms, mat, procs)
    g = driver.Solver(p, precision=precision)
    mac = g.ctx.macro()
    V = p.nvox
    rng = np.random.default_rng(7 + n)
    sg = rng.standard_normal((V, 3, 3)) * 100.0
    sg = 0.5 * (sg + sg.transpose(0, 2, 1))
    g.ctx.set_field("sg", sg.reshape(V, 9))
    P = g.procs[0]
    st, out = lib.CgIterState(1, 1.0), lib.CgIterOut()  # iter != 0: no per-increment reset
    assert g.L.cg_begin_increment(g.h, C.byref(P), 1) == 0
    assert g.L.cg_iterate(g.h, C.byref(P), C.byref(st), C.byref(out)) == 0
    dl = g.ctx.field("dvelgradref").reshape(n, n, n, 9)
    g.close()
    Pt = torch.as_tensor(sg.reshape(n, n, n, 9), device="cuda")
    Pt = Pt - Pt.reshape(-1, 9).mean(dim=0)
    want = torch_convolution(Pt, mac["c0"], mac["s0"], (n, n, n)).cpu().numpy()
    err = float(np.abs(want - dl).max() / np.abs(want).max())
    print("max rel err vs rocFFT: %.2e" % err, flush=True)
    assert err < (1e-11 if precision == 0 else 2e-6)
