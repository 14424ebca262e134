"""The moment storage form's coefficient algebra, on the CPU.

The library stores, per cell, the coefficients of the polynomial in c that
the regularized collision output w_q^-1 (f~eq + (1 - omega) f_neq + F_q / 2)
is (phasefieldlbm_amd/csrc/mom.hpp, mom_coeffs), and rebuilds each
population as w_q times that polynomial at c_q.  This test restates the
coefficient closed forms in FP64 and checks them against the reference's
per-direction formulas (src/velocitySet/D3Q19.cuh:143-187,
src/velocitySet/D3Q27.cuh:152-206, src/lbm.cu:206) for random states, and
that plausible slips in the closed forms fail the check.  The GPU tests
(test_gpu_moments.py) hold the HIP code to the oracle.
"""
import numpy as np
import pytest

CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1]
CY = [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1]
W = {19: [1 / 3, 1 / 18, 1 / 36], 27: [8 / 27, 2 / 27, 1 / 54, 1 / 216]}
CS2 = 1.0 / 3.0


def reference_post_collision(Q, rho, u, F, P, omco):
    """f~eq + (1 - omega) f_neq + F_q / 2, direction by direction, as the
    reference writes it (shifted f~ = f - w)."""
    uu = 1.5 * (u @ u)
    ux, uy, uz = u
    pxx, pyy, pzz, pxy, pxz, pyz = P[0, 0], P[1, 1], P[2, 2], P[0, 1], P[0, 2], P[1, 2]
    out = []
    for q in range(Q):
        c = np.array([CX[q], CY[q], CZ[q]], dtype=float)
        w = W[Q][int(np.abs(c).sum())]
        cu = 3.0 * (c @ u)
        feq = w * rho * (1 - uu + cu + 0.5 * cu * cu) - w
        if Q == 27:
            feq += w * rho * (cu ** 3 / 6 - uu * cu)
        s = sum((c[i] * c[j] - CS2 * (i == j)) * P[i, j] for i in range(3) for j in range(3))
        if Q == 27:
            cx, cy, cz = c
            k = [cx * cx * cy - CS2 * cy, cx * cx * cz - CS2 * cz, cx * cy * cy - CS2 * cx,
                 cy * cy * cz - CS2 * cz, cx * cz * cz - CS2 * cx, cy * cz * cz - CS2 * cy]
            X = [pxx * uy + 2 * ux * pxy, pxx * uz + 2 * ux * pxz, pyy * ux + 2 * uy * pxy,
                 pyy * uz + 2 * uy * pyz, pzz * ux + 2 * uz * pxz, pzz * uy + 2 * uz * pyz]
            s += 3 * sum(a * b for a, b in zip(k, X)) + 6 * cx * cy * cz * (ux * pyz + uy * pxz + uz * pxy)
        fneq = 4.5 * w * s
        frc = w * ((3 * (c - u) + 3 * cu * c) @ F)
        out.append(feq + omco * fneq + 0.5 * frc)
    return np.array(out)


def coefficients(Q, rho, u, F, P, omco, slip=None):
    """The closed forms of mom.hpp (mom_coeffs), in FP64."""
    ux, uy, uz = u
    fx, fy, fz = F
    usq, uF = u @ u, u @ F
    R3, R45, O45 = 3 * rho, 4.5 * rho, 4.5 * omco
    pxx, pyy, pzz, pxy, pxz, pyz = P[0, 0], P[1, 1], P[2, 2], P[0, 1], P[0, 2], P[1, 2]
    c = {"A": (rho - 1) - 1.5 * (rho * usq + uF)}
    for i, n in enumerate("xyz"):
        c["G" + n] = R45 * u[i] ** 2 + 4.5 * u[i] * F[i] + O45 * P[i, i]
    for (i, j) in ((0, 1), (0, 2), (1, 2)):
        c["D" + "xyz"[i] + "xyz"[j]] = 2 * R45 * u[i] * u[j] + 4.5 * (u[i] * F[j] + u[j] * F[i]) + O45 * 2 * P[i, j]
    if Q == 19:
        for i, n in enumerate("xyz"):
            c["B" + n] = R3 * u[i] + 1.5 * F[i]
    else:
        V = 1 - 1.5 * usq
        X1, X2 = pxx * uy + 2 * ux * pxy, pxx * uz + 2 * ux * pxz
        X3, X4 = pyy * ux + 2 * uy * pxy, pyy * uz + 2 * uy * pyz
        X5, X6 = pzz * ux + 2 * uz * pxz, pzz * uy + 2 * uz * pyz
        X7s = 6 * (ux * pyz + uy * pxz + uz * pxy)
        fterm = 0.0 if slip == "force" else 1.5
        c["Bx"] = R3 * (V * ux + 1.5 * ux ** 3) + fterm * fx - O45 * (X3 + X5)
        c["By"] = R3 * (V * uy + 1.5 * uy ** 3) + 1.5 * fy - O45 * (X1 + X6)
        c["Bz"] = R3 * (V * uz + 1.5 * uz ** 3) + 1.5 * fz - O45 * (X2 + X4)
        c["Exxy"] = 4.5 * R3 * ux * ux * uy + 3 * O45 * X1
        c["Exxz"] = 4.5 * R3 * ux * ux * uz + 3 * O45 * X2
        c["Eyyx"] = 4.5 * R3 * uy * uy * ux + 3 * O45 * X3
        c["Eyyz"] = 4.5 * R3 * uy * uy * uz + 3 * O45 * X4
        c["Ezzx"] = 4.5 * R3 * uz * uz * ux + 3 * O45 * X5
        c["Ezzy"] = 4.5 * R3 * uz * uz * uy + 3 * O45 * X6
        c["Z"] = (3.0 if slip == "zeta" else 9.0) * R3 * ux * uy * uz + O45 * X7s
    return c


def rebuild(Q, c):
    """w_q times the polynomial at c_q (mom_pop)."""
    out = []
    for q in range(Q):
        x, y, z = CX[q], CY[q], CZ[q]
        p = (c["A"] + x * c["Bx"] + y * c["By"] + z * c["Bz"] + x * x * c["Gx"] + y * y * c["Gy"] + z * z * c["Gz"]
             + x * y * c["Dxy"] + x * z * c["Dxz"] + y * z * c["Dyz"])
        if Q == 27:
            p += (x * x * y * c["Exxy"] + x * x * z * c["Exxz"] + y * y * x * c["Eyyx"] + y * y * z * c["Eyyz"]
                  + z * z * x * c["Ezzx"] + z * z * y * c["Ezzy"] + x * y * z * c["Z"])
        out.append(W[Q][abs(x) + abs(y) + abs(z)] * p)
    return np.array(out)


def _states(n=20, seed=3):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        rho = 1 + 0.01 * rng.standard_normal()
        u = 0.06 * rng.standard_normal(3)
        F = 1e-3 * rng.standard_normal(3)
        P = 1e-3 * rng.standard_normal((3, 3))
        P = P + P.T
        P -= np.trace(P) / 3 * np.eye(3)
        omco = 1 - rng.uniform(0.5, 1.99)
        yield rho, u, F, P, omco


@pytest.mark.parametrize("Q", [19, 27])
def test_coefficients_rebuild_reference_populations(Q):
    for st in _states():
        ref = reference_post_collision(Q, *st)
        got = rebuild(Q, coefficients(Q, *st))
        assert np.abs(got - ref).max() < 1e-15, np.abs(got - ref).max()


@pytest.mark.parametrize("slip", ["zeta", "force"])
def test_planted_slips_fail(slip):
    worst = 0.0
    for st in _states():
        ref = reference_post_collision(27, *st)
        worst = max(worst, np.abs(rebuild(27, coefficients(27, *st, slip=slip)) - ref).max())
    assert worst > 1e-7, (slip, worst)


@pytest.mark.parametrize("Q", [19, 27])
def test_rest_state_is_zero(Q):
    """rho = 1, u = 0, no stress, no force: every shifted population is 0."""
    c = coefficients(Q, 1.0, np.zeros(3), np.zeros(3), np.zeros((3, 3)), 0.3)
    assert all(v == 0.0 for v in c.values())
    assert np.all(rebuild(Q, c) == 0.0)


@pytest.mark.parametrize("Q", [19, 27])
def test_moments_of_rebuilt_populations(Q):
    """Mass and momentum of the rebuilt populations are those of the
    collision: sum f~ = rho - 1 and sum f~ c = rho u + F/2 (Guo)."""
    for rho, u, F, P, omco in _states(5):
        f = rebuild(Q, coefficients(Q, rho, u, F, P, omco))
        c = np.array([CX[:Q], CY[:Q], CZ[:Q]], dtype=float)
        assert abs(f.sum() - (rho - 1)) < 1e-14
        np.testing.assert_allclose(c @ f, rho * u + 0.5 * F, atol=1e-14)
