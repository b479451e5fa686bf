"""The constitutive Newton step of the HIP library is taken in a symmetric form.

The reference (various_functions.f90:1226-1254) solves J d = R with
J = I + cm (dedotp + s66/dt), R = sg - lambda + cm (edotp + s66 (sg - sgt)/dt - D).
The library factors cm out of both: J = cm K, R = cm R' with
K = cm^-1 + s66/dt + dedotp (symmetric positive definite) and
R' = (cm^-1 + s66/dt) sg - (cm^-1 lambda + s66 sgt/dt + D) + edotp,
and solves K d = R' by L D L^T, using |cm R'| for the line search.  These checks
pin that the two forms give the same step and the same residual norm, and that an
L D L^T written as in the library reproduces numpy's solve.  CPU only.
"""
import numpy as np


def spd(rng, n=6, scale=1.0):
    a = rng.standard_normal((n, n))
    return scale * (a @ a.T + n * np.eye(n))


def ldl_solve(K, r):
    """L D L^T without pivoting, column by column, as the library's ldl6/ldl6_solve."""
    n = len(r)
    L = np.eye(n)
    d = np.zeros(n)
    for j in range(n):
        d[j] = K[j, j] - sum(L[j, k] ** 2 * d[k] for k in range(j))
        for i in range(j + 1, n):
            L[i, j] = (K[i, j] - sum(L[i, k] * L[j, k] * d[k] for k in range(j))) / d[j]
    z = r.astype(float).copy()
    for i in range(n):
        z[i] -= sum(L[i, k] * z[k] for k in range(i))
    z /= d
    for i in reversed(range(n)):
        z[i] -= sum(L[k, i] * z[k] for k in range(i + 1, n))
    return z


def test_symmetric_newton_step_equals_reference_step():
    rng = np.random.default_rng(7)
    for _ in range(20):
        cm = spd(rng, scale=1e4)            # stiffness-like (MPa s)
        s66 = spd(rng, scale=1e-5)          # compliance-like
        dt = 0.5
        sch = rng.standard_normal((12, 5))
        w = rng.random(12) * 1e-3
        dd = np.zeros((6, 6))
        dd[:5, :5] = (sch.T * w) @ sch      # sum_a w_a S_a (x) S_a
        lam, sgt, sg, D = (rng.standard_normal(6) * s for s in (100, 100, 100, 1e-4))
        ep = np.zeros(6)
        ep[:5] = rng.standard_normal(5) * 1e-4
        # reference form
        R = sg - lam + cm @ (ep + s66 @ (sg - sgt) / dt - D)
        J = np.eye(6) + cm @ (dd + s66 / dt)
        d_ref = np.linalg.solve(J, R)
        # symmetric form
        ci = np.linalg.inv(cm)
        A = ci + s66 / dt
        b = ci @ lam + s66 @ sgt / dt + D
        Rp = A @ sg - b + ep
        K = A + dd
        d_sym = ldl_solve(K, Rp)
        assert np.allclose(cm @ Rp, R, rtol=1e-9, atol=1e-9 * np.abs(R).max())
        assert abs(np.linalg.norm(cm @ Rp) - np.linalg.norm(R)) <= 1e-9 * np.linalg.norm(R)
        assert np.allclose(d_sym, d_ref, rtol=1e-8, atol=1e-10 * np.abs(d_ref).max())
        assert np.allclose(K, K.T) and np.all(np.linalg.eigvalsh(K) > 0)


def test_ldl_matches_numpy_solve():
    rng = np.random.default_rng(3)
    for _ in range(20):
        K = spd(rng)
        r = rng.standard_normal(6)
        assert np.allclose(ldl_solve(K, r), np.linalg.solve(K, r), rtol=1e-12, atol=1e-12)
