"""Closed-form checks of the oracle's lattice algebra (FP64).

They pin the restated D3Q7/D3Q19/D3Q27 tables and per-cell functions
(src/velocitySet/*.cuh) to the moment identities the Hermite construction
fixes, so a dropped term, a sign or an index slip fails here.
"""
import ctypes

import numpy as np
import pytest


def lattice(lib, Q):
    cx = np.zeros(Q, np.int32); cy = np.zeros(Q, np.int32); cz = np.zeros(Q, np.int32)
    w = np.zeros(Q)
    assert lib.pfo_lattice(Q, cx.ctypes.data, cy.ctypes.data, cz.ctypes.data, w.ctypes.data) == 0
    return np.stack([cx, cy, cz], 1).astype(float), w


CS2 = {7: 0.25, 19: 1 / 3, 27: 1 / 3}


@pytest.mark.parametrize("Q", [7, 19, 27])
def test_weights_and_isotropy(oracle_lib, Q):
    c, w = lattice(oracle_lib, Q)
    cs2 = CS2[Q]
    assert abs(w.sum() - 1) < 1e-15
    np.testing.assert_allclose(w @ c, 0, atol=1e-15)
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", w, c, c), cs2 * np.eye(3), atol=1e-15)
    np.testing.assert_allclose(np.einsum("q,qa,qb,qc->abc", w, c, c, c), 0, atol=1e-15)
    if Q != 7:  # fourth-order isotropy of the hydrodynamic sets
        m4 = np.einsum("q,qa,qb,qc,qd->abcd", w, c, c, c, c)
        d = np.eye(3)
        iso = cs2 ** 2 * (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
                          + np.einsum("ad,bc->abcd", d, d))
        np.testing.assert_allclose(m4, iso, atol=1e-15)
    # opposite pairs (2j-1, 2j)
    for q in range(1, Q, 2):
        np.testing.assert_array_equal(c[q], -c[q + 1])
    assert len({tuple(v) for v in c}) == Q


def feq(lib, Q, rho, u):
    out = np.zeros(Q)
    uu = np.asarray(u, float)
    assert lib.pfo_cell_feq(Q, rho, uu.ctypes.data, out.ctypes.data) == 0
    return out


@pytest.mark.parametrize("Q", [19, 27])
def test_feq_moments(oracle_lib, Q):
    c, w = lattice(oracle_lib, Q)
    rho, u = 1.07, np.array([0.031, -0.022, 0.047])
    f = feq(oracle_lib, Q, rho, u) + w  # unshift
    assert abs(f.sum() - rho) < 1e-14
    np.testing.assert_allclose(f @ c, rho * u, atol=1e-14)
    m2 = np.einsum("q,qa,qb->ab", f, c, c)
    np.testing.assert_allclose(m2, rho * (np.outer(u, u) + np.eye(3) / 3), atol=1e-14)
    # u = 0, rho = 1: the shifted equilibrium vanishes (frozen inlet wall)
    np.testing.assert_array_equal(feq(oracle_lib, Q, 1.0, [0, 0, 0]), 0)


@pytest.mark.parametrize("Q", [19, 27])
def test_fneq_projection(oracle_lib, Q):
    """Sum_q f_neq c c recovers a deviatoric P; mass and momentum vanish."""
    c, w = lattice(oracle_lib, Q)
    rng = np.random.default_rng(3)
    for _ in range(5):
        A = rng.normal(size=(3, 3)) * 1e-3
        P = A + A.T
        P -= np.trace(P) / 3 * np.eye(3)
        u = rng.normal(size=3) * 0.03
        Pv = np.array([P[0, 0], P[1, 1], P[2, 2], P[0, 1], P[0, 2], P[1, 2]])
        out = np.zeros(Q)
        assert oracle_lib.pfo_cell_fneq(Q, Pv.ctypes.data, u.ctypes.data, out.ctypes.data) == 0
        assert abs(out.sum()) < 1e-15
        np.testing.assert_allclose(out @ c, 0, atol=1e-15)
        np.testing.assert_allclose(np.einsum("q,qa,qb->ab", out, c, c), P, atol=1e-15)
        if Q == 27:
            # third-order Hermite moment: u_a P_bc + u_b P_ac + u_c P_ab
            m3 = np.einsum("q,qa,qb,qc->abc", out, c, c, c)
            want = (np.einsum("a,bc->abc", u, P) + np.einsum("b,ac->abc", u, P)
                    + np.einsum("c,ab->abc", u, P))
            # D3Q27 carries the off-diagonal third moments exactly (the
            # diagonal c_a^3 = c_a is aliased into the first moment)
            for a, b, cc in [(0, 0, 1), (0, 0, 2), (1, 1, 0), (1, 1, 2), (2, 2, 0), (2, 2, 1), (0, 1, 2)]:
                assert abs(m3[a, b, cc] - want[a, b, cc]) < 1e-15


@pytest.mark.parametrize("Q", [19, 27])
def test_guo_force_moments(oracle_lib, Q):
    c, w = lattice(oracle_lib, Q)
    u = np.array([0.02, -0.01, 0.04])
    F = np.array([1e-4, -3e-4, 2e-4])
    out = np.zeros(Q)
    assert oracle_lib.pfo_cell_force(Q, u.ctypes.data, F.ctypes.data, out.ctypes.data) == 0
    assert abs(out.sum()) < 1e-18
    np.testing.assert_allclose(out @ c, F, atol=1e-18)
    np.testing.assert_allclose(np.einsum("q,qa,qb->ab", out, c, c), np.outer(F, u) + np.outer(u, F),
                               atol=1e-18)


def test_phase_equilibrium(oracle_lib):
    c, w = lattice(oracle_lib, 7)
    u = np.array([0.03, -0.02, 0.05]); n = np.array([0.6, 0.0, 0.8])
    geq = np.zeros(7); h = np.zeros(7)
    phi, sharp = 0.37, 0.5 * 0.37 * 0.63
    assert oracle_lib.pfo_cell_geq(phi, u.ctypes.data, n.ctypes.data, sharp, geq.ctypes.data,
                                   h.ctypes.data) == 0
    assert abs(geq.sum() - phi) < 1e-15
    np.testing.assert_allclose(geq @ c, phi * u, atol=1e-15)
    assert abs(h.sum()) < 1e-17
    np.testing.assert_allclose(h @ c, 0.25 * sharp * n, atol=1e-17)


@pytest.mark.parametrize("Q", [19, 27])
def test_isotropic_gradient(oracle_lib, Q):
    """Exact on linear fields; O(h^2) on a quadratic (no bias at the origin)."""
    c, w = lattice(oracle_lib, Q)
    g = np.array([0.3, -1.1, 0.7])
    v = c @ g + 5.0
    out = np.zeros(3)
    assert oracle_lib.pfo_iso_grad(Q, v.ctypes.data, out.ctypes.data) == 0
    np.testing.assert_allclose(out, g, rtol=1e-14)
    v2 = v + 0.4 * (c ** 2).sum(1)  # even part does not bias the gradient
    assert oracle_lib.pfo_iso_grad(Q, v2.ctypes.data, out.ctypes.data) == 0
    np.testing.assert_allclose(out, g, rtol=1e-14)
    # a planted sign error in one direction would show up here
    v3 = v.copy(); v3[1] += 1.0
    assert oracle_lib.pfo_iso_grad(Q, v3.ctypes.data, out.ctypes.data) == 0
    assert abs(out[0] - g[0]) > 1e-3


@pytest.mark.parametrize("Q", [19, 27])
def test_sphere_divergence(oracle_lib, Q):
    """div(r/|r|) by the isotropic stencil at radius R approaches 2/R."""
    c, w = lattice(oracle_lib, Q)
    for R in (8.0, 16.0):
        p = np.array([R, 0.0, 0.0]) / np.sqrt(1.0)
        s = 0.0
        for a in range(3):
            vals = np.zeros(Q)
            for q in range(Q):
                r = p + c[q]
                vals[q] = r[a] / np.linalg.norm(r)
            out = np.zeros(3)
            oracle_lib.pfo_iso_grad(Q, vals.ctypes.data, out.ctypes.data)
            s += out[a]
        assert abs(s - 2 / R) / (2 / R) < 0.01
