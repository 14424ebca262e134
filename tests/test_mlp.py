"""Learned closure: reference network (CPU) and the matrix-core kernels (GPU)."""
import numpy as np
import pytest

import oracle.mlp_ref as ref


def test_reference_gradient_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(size=6017) * 0.2
    x = rng.uniform(-1, 1, (40, 27))
    t = rng.normal(size=40)
    _, g = ref.loss_and_grad(p, x, t, bf=False)
    for k in [0, 700, 1727, 1728, 1791, 1792, 5887, 5888, 5951, 5952, 6015, 6016]:
        e = np.zeros(6017)
        e[k] = 1e-6
        fd = (ref.loss_and_grad(p + e, x, t, bf=False)[0] - ref.loss_and_grad(p - e, x, t, bf=False)[0]) / 2e-6
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), k


def test_reference_bf16_rounding():
    # 1 + 2^-8 and 1 + 3*2^-8 are halfway cases: ties go to the even mantissa
    v = np.array([1.0, 1.00390625, 1.001953125, 3.14159265, -2.5e-3, 0.0, 1.01171875])
    b = ref.bf16(v)
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == 1.0 and b[6] == 1.015625
    assert abs(b[3] - 3.140625) < 1e-12
    np.testing.assert_array_equal(ref.bf16(b), b)


def test_cube_symmetry_labels_invariant():
    """The closure's training labels (curvature) do not depend on the
    orientation: a rotated ellipsoid gives the same label distribution."""
    from phasefieldlbm_amd import closure
    axes = np.array([14.0, 9.0, 7.0])
    c = np.full(3, 31.5)
    _, k1 = closure.ellipsoid_fields(axes, np.eye(3), c)
    _, k2 = closure.ellipsoid_fields(axes[[1, 2, 0]], np.eye(3)[[1, 2, 0]].T, c)
    np.testing.assert_allclose(np.sort(k1.ravel())[::997], np.sort(k2.ravel())[::997], rtol=1e-4, atol=1e-6)
    # sphere: kappa = -2/R everywhere
    _, ks = closure.ellipsoid_fields(np.full(3, 10.0), np.eye(3), c)
    np.testing.assert_allclose(ks, -0.2, rtol=1e-5)


@pytest.mark.gpu
def test_mfma_lane_layout():
    import phasefieldlbm_amd as pf
    rng = np.random.default_rng(1)
    a = rng.normal(size=(16, 32)).astype(np.float32)
    b = rng.normal(size=(32, 16)).astype(np.float32)
    c = pf.mfma_check(a, b)
    want = ref.bf16(a) @ ref.bf16(b)
    np.testing.assert_allclose(c, want, rtol=1e-5, atol=1e-5)


def _random_problem(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=6017) * 0.15
    x = rng.uniform(-1, 1, (n, 27))
    t = rng.normal(size=n)
    return p.astype(np.float32), x.astype(np.float32), t.astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 200])
def test_mlp_gradients_match_reference(n):
    import phasefieldlbm_amd as pf
    p, x, t = _random_problem(n, 2)
    net = pf.MLP(0, p)
    net.set_data(x, t)
    g, mse = net.gradients(n)
    y = net.predict(x)
    net.close()
    y_ref, _ = ref.forward(p, x, bf=True)
    np.testing.assert_allclose(y, y_ref, rtol=0, atol=2e-4 * max(1.0, np.abs(y_ref).max()))
    mse_ref, g_ref = ref.loss_and_grad(p, x, t, bf=True)
    assert abs(mse - mse_ref) <= 1e-4 * mse_ref
    blocks = np.cumsum([0, 64 * 27, 64, 64 * 64, 64, 64, 1])
    for a, b in zip(blocks[:-1], blocks[1:]):
        scale = np.abs(g_ref[a:b]).max()
        assert np.abs(g[a:b] - g_ref[a:b]).max() <= 1e-2 * scale, (a, b)
    # and against the exact FP64 network (bf16 operands cost ~1 %)
    _, g64 = ref.loss_and_grad(p, x, t, bf=False)
    assert np.linalg.norm(g - g64) <= 0.05 * np.linalg.norm(g64)


@pytest.mark.gpu
def test_mlp_training_fits():
    import phasefieldlbm_amd as pf
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, (8192, 27)).astype(np.float32)
    t = (np.tanh(x[:, 13] - x[:, 4]) + 0.5 * x[:, 22] * x[:, 0]).astype(np.float32)
    net = pf.MLP(0)
    net.set_data(x, t)
    losses = net.train(8, batch=256, lr=3e-3, seed=1, augment=False)
    assert losses[-1] < 0.1 * losses[0]
    assert np.all(np.isfinite(losses))
    net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("grid", [(32, 32, 32), (70, 30, 37)])
@pytest.mark.parametrize("storage", [0, 1])
def test_closure_in_timestep(storage, grid):
    """k_closure + the closure variant of the fused step: the curvature the
    step applies at interface cells equals the reference network evaluated
    on the same phi stencils, each in its canonical frame, averaged over the
    cell and its interface neighbours.  Both storage
    forms (binary16 uses the step's slim LDS layout).  The 70 x 30 x 37 box
    gives k_closure's 64 x 4 x 8 workgroups a second, partial x tile, a
    partial row block and a last z-chunk of 3 planes."""
    import phasefieldlbm_amd as pf
    from phasefieldlbm_amd import closure, presets
    rng = np.random.default_rng(5)
    p = (rng.normal(size=6017) * 0.1).astype(np.float32)
    cfg = presets.static_droplet(n=32, radius=8.0, storage=storage)
    cfg.nx, cfg.ny, cfg.nz = grid
    cfg.closure = pf.CLOSURE_MLP
    cfg.mlp_kmin, cfg.mlp_kmax = -100.0, 100.0  # network at every interface cell
    with pf.Simulation(cfg) as sim:
        sim.set_closure_params(p)
        sim.init().step(3, save=True)
        phi = sim.field("phi")
        ind = sim.field("ind").astype(np.float64)
        nv = np.stack([sim.field(k) for k in ("normx", "normy", "normz")]).astype(np.float64)
        fs = np.stack([sim.field(k) for k in ("fsx", "fsy", "fsz")]).astype(np.float64)
    inner = np.zeros(phi.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    sel = inner & (phi > cfg.mlp_lo) & (phi < cfg.mlp_hi) & (ind > 1e-3)
    cells = np.flatnonzero(sel.ravel())
    assert cells.size > 100
    # the closure evaluates the network once per interface cell, on its
    # stencil in the canonical frame under the cube symmetries, and the step
    # applies the mean over the cell and its interface neighbours
    band = inner & (phi > cfg.mlp_lo) & (phi < cfg.mlp_hi)
    bc = np.flatnonzero(band.ravel())
    kap = np.zeros(phi.size)
    kap[bc] = 0.25 * ref.forward(p, closure.stencils(phi, bc), bf=True)[0]
    kap = kap.reshape(phi.shape)
    ksum, cnt = np.zeros(phi.shape), np.zeros(phi.shape)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                sh = (-dz, -dy, -dx)  # value at x + d moved to x
                ksum += np.roll(kap * band, sh, axis=(0, 1, 2))
                cnt += np.roll(band.astype(float), sh, axis=(0, 1, 2))
    k_ref = (ksum / np.maximum(cnt, 1.0)).ravel()[cells]
    n2 = (nv ** 2).sum(0).ravel()[cells]
    k_gpu = -(fs * nv).sum(0).ravel()[cells] / (cfg.sigma * ind.ravel()[cells] * n2)
    np.testing.assert_allclose(k_gpu, k_ref, rtol=0, atol=1e-4 * max(1.0, np.abs(k_ref).max()))
