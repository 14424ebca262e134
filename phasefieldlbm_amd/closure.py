"""Learned curvature closure: data from the solver, training on the GPU.

Pipeline (no counterpart in the reference, whose curvature is always the
finite-difference div(n) of src/phaseField.cuh:142-219):

1. ``make_dataset``: static droplets and ellipsoids in periodic boxes
   (64^3, 96^3 for the largest), random semi-axes, orientation and centre,
   phase profile phi = 1/2 (1 + tanh(-2 d / W)) with d the signed distance to
   the surface (exact, by closest-point projection, within 12 cells of it),
   W = 4.  Each is relaxed for 500 timesteps by this library with surface
   tension off, so the velocity stays zero and the shape stays the analytic
   one while the phase profile settles into the scheme's own equilibrium.
   Samples are the normalised 3x3x3 phi stencils (``stencils``, in their
   canonical frame under the cube symmetries) of cells with
   0.001 < phi < 0.999, the whole diffuse interface, labelled 4 x the
   analytic curvature in the reference's convention, kappa = div(n) with
   n = grad(phi)/|grad phi| (= -2/R on a sphere): kappa = 2 H at the point of
   the ellipsoid closest to the cell centre, H the signed mean curvature.
   That is the curvature of the sharp interface, so sigma * kappa integrated
   across the diffuse one gives the Laplace jump 2 sigma / R exactly.
2. ``train``: Adam on the mean squared error with the device kernels
   (pf_mlp_train).  The stencils are in their canonical frame under the 48
   symmetries of the cube (``canonical_stencils``), as the closure pass
   evaluates them, so the network only ever sees one orientation of each
   orbit and needs no symmetry augmentation.
3. ``laplace``: static droplets run to rest (12000 steps) with the FD
   curvature and with the closure; Laplace pressure error and spurious
   velocity of both.

The semi-axes are drawn so the largest, a, lies in [4, 32] cells and the
others are 1 to 2 times smaller but never below sqrt(4 a): the smallest
radius of curvature of the ellipsoid, c^2 / a, then stays at 4 cells or more,
where a curvature is still resolved by the W = 4 diffuse profile.  Shapes with
a > 22 are relaxed in 96^3 boxes, which hold the interface and its sampled
tail.  Half the shapes are spheres.

Why the whole interface: a curvature that switches between the network's
value and the FD one inside the interface is not a function of phi along it,
so the surface force it gives has a rotational part that pressure cannot
balance.  With the exact curvature -2/R applied only where
0.05 < phi < 0.95 a droplet's spurious currents exceeded FD's; applied where
0.001 < phi < 0.999 they were 3-6 times smaller than FD's (R = 8-16).
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import CLOSURE_FD, CLOSURE_MLP, D3Q19, MLP, Simulation, presets, split_params

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "closure_weights.npz")
# The same parameters as 6017 little-endian float32 values, for the C++
# driver (pfsim --closure).
WEIGHTS_RAW = os.path.join(HERE, "closure_weights.f32")
BOX = 64
BOX_LARGE = 96
W = 4.0
# Phase window of the training samples: the whole diffuse interface out to
# where the force sigma kappa |grad phi| has fallen to ~0.4 % of its peak.
SAMPLE_LO, SAMPLE_HI = 0.001, 0.999
# The reference's phase relaxation time (src/constants.cuh:105, 134).  It sets
# the sharpening strength and with it the relaxed interface profile, so the
# closure is trained and evaluated at this value only.
TAU_G = 1.0


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def box_for(amax):
    """Periodic box edge for a shape whose largest semi-axis is amax: the
    shape, its sampled tail (phi down to SAMPLE_LO, ~7 cells out for W = 4)
    and the ghost shell fit in 64^3 up to amax = 22, in 96^3 beyond."""
    return BOX if amax <= 22.0 else BOX_LARGE


def draw_shape(rng, sphere=False):
    """Semi-axes, rotation, centre and box edge of one training shape: the
    largest semi-axis a uniform in [4, 32], the others a / U(1, 2), raised
    where needed to sqrt(4 a) so that no radius of curvature (c^2 / a at the
    tip) falls below 4 cells."""
    amax = rng.uniform(4.0, 32.0)
    if sphere:
        axes = np.full(3, amax)
    else:
        axes = np.maximum(amax / rng.uniform(1.0, 2.0, size=3), np.sqrt(4.0 * amax))
        axes[0] = amax
    n = box_for(amax)
    rot = random_rotation(rng)
    slack = (n - 2) / 2.0 - 1.0 - (axes.max() + 2 * W)
    centre = (n - 1) / 2.0 + rng.uniform(-1, 1, size=3) * max(slack, 0.0)
    return axes, rot, centre, n


def _body(points, rot, centre):
    return (points - centre) @ rot  # grid -> body frame


def closest_point(axes, p, iters=60):
    """Closest point on the ellipsoid sum (x_i / a_i)^2 = 1 to body-frame
    points p [..., 3], and the signed distance (negative inside).

    The closest point is x_i = a_i^2 p_i / (t + a_i^2) with t the root of
    F(t) = sum (a_i p_i / (t + a_i^2))^2 - 1 on t > -min a_i^2, where F is
    convex and decreasing; Newton from t0 = max_i (a_i |p_i| - a_i^2), where
    F(t0) >= 0, converges monotonically from the left."""
    e = axes * axes
    ap = np.abs(p) * axes
    t = np.max(ap - e, axis=-1)
    for _ in range(iters):
        r = ap / (t[..., None] + e)
        f = (r * r).sum(-1) - 1.0
        df = -2.0 * (r * r / (t[..., None] + e)).sum(-1)
        step = f / df
        t = t - step
        if np.all(np.abs(step) <= 1e-12 * (1.0 + np.abs(t))):
            break
    x = e * p / (t[..., None] + e)
    d = np.sqrt(((p - x) ** 2).sum(-1)) * np.sign(t)
    return x, d


def mean_curvature2(axes, x):
    """2 H of the ellipsoid at surface points x (body frame); H is negative
    on a convex surface (-1/R on a sphere)."""
    a2 = axes ** 2
    num = (x * x).sum(-1) - a2.sum()
    den = 2.0 * np.prod(a2) * ((x * x / (a2 * a2)).sum(-1)) ** 1.5
    return 2.0 * num / den


def phi_field(axes, rot, centre, n=BOX):
    """Initial phase field of an ellipsoid (ghost shell zero), computed in its
    bounding box only: exact signed distance within 12 cells of the surface,
    the first-order estimate (where tanh has saturated) elsewhere."""
    ext = axes.max() + 2 * W + 2
    lo = np.maximum(np.floor(centre - ext).astype(int), 1)
    hi = np.minimum(np.ceil(centre + ext).astype(int) + 1, n - 1)
    phi = np.zeros((n, n, n), dtype=np.float32)
    z, y, x = np.meshgrid(np.arange(lo[2], hi[2]), np.arange(lo[1], hi[1]), np.arange(lo[0], hi[0]),
                          indexing="ij")
    p = _body(np.stack([x, y, z], -1).astype(np.float64), rot, centre)
    s = p / axes
    rho = np.sqrt((s * s).sum(-1)) + 1e-12
    g = p / (axes * axes) / rho[..., None]
    d = (rho - 1.0) / (np.sqrt((g * g).sum(-1)) + 1e-12)
    band = np.abs(d) < 12.0
    d[band] = closest_point(axes, p[band])[1]
    phi[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] = 0.5 * (1.0 + np.tanh(-2.0 * d / W))
    return phi


def gauss_curvature(axes, x):
    """Gaussian curvature of the ellipsoid at surface points x (body frame)."""
    a2 = axes ** 2
    return 1.0 / (np.prod(a2) * ((x * x / (a2 * a2)).sum(-1)) ** 2)


def curvature_labels(axes, rot, centre, cells, n=BOX, label="surface"):
    """Curvature label of each cell, in the reference's sign (div n, -2/R on
    a sphere).  "surface": 2 H at the point of the ellipsoid closest to the
    cell centre.  "levelset": the curvature of the level set through the
    cell, at signed distance d from the surface (d > 0 outside): the sum of
    c_i / (1 - d c_i) over the principal curvatures c_i at the closest point
    (-2/r on a sphere), the quantity finite-difference div n approximates."""
    z, y, x = np.unravel_index(cells, (n, n, n))
    p = _body(np.stack([x, y, z], -1).astype(np.float64), rot, centre)
    xs, d = closest_point(axes, p)
    k2 = mean_curvature2(axes, xs)
    if label == "surface":
        return k2.astype(np.float32)
    K = gauss_curvature(axes, xs)
    disc = np.sqrt(np.maximum(k2 * k2 - 4.0 * K, 0.0))
    c1, c2 = 0.5 * (k2 + disc), 0.5 * (k2 - disc)
    return (c1 / (1.0 - d * c1) + c2 / (1.0 - d * c2)).astype(np.float32)


def ellipsoid_fields(axes, rot, centre, n=BOX):
    """phi on the grid and the curvature label of every cell."""
    phi = phi_field(axes, rot, centre, n)
    kap = curvature_labels(axes, rot, centre, np.arange(n ** 3), n).reshape(n, n, n)
    return phi, kap


def stencils(phi, cells, canonical=True):
    """27 normalised features of the 3x3x3 phi stencils at flat cell indices,
    computed in FP32 as the closure kernel does: the centre (feature 13) is
    2 phi_c - 1, the others (phi_f - phi_c) / (|g| + 1e-4) with g the
    central-difference gradient of phi at the centre.  canonical: each
    stencil in its canonical frame under the cube symmetries
    (``canonical_stencils``), as the closure pass evaluates it."""
    z, y, x = np.unravel_index(cells, phi.shape)
    phi = np.asarray(phi, dtype=np.float32)

    def at(dx, dy, dz):
        return phi[z + dz, y + dy, x + dx]

    pc = at(0, 0, 0)
    half = np.float32(0.5)
    gx = half * (at(1, 0, 0) - at(-1, 0, 0))
    gy = half * (at(0, 1, 0) - at(0, -1, 0))
    gz = half * (at(0, 0, 1) - at(0, 0, -1))
    inv = np.float32(1.0) / (np.sqrt(gx * gx + gy * gy + gz * gz) + np.float32(1e-4))
    out = np.empty((cells.size, 27), dtype=np.float32)
    for f in range(27):
        dx, dy, dz = f % 3 - 1, (f // 3) % 3 - 1, f // 9 - 1
        out[:, f] = (np.float32(2.0) * pc - np.float32(1.0)) if f == 13 else (at(dx, dy, dz) - pc) * inv
    if canonical:
        out = canonical_stencils(out, at(1, 0, 0) - at(-1, 0, 0), at(0, 1, 0) - at(0, -1, 0),
                                 at(0, 0, 1) - at(0, 0, -1))
    return out


def canonical_stencils(X, dx, dy, dz, order=True):
    """Stencils X [n, 27] in their canonical frame under the 48 symmetries of
    the cube, given the differences d = phi(+1) - phi(-1) across each centre
    along x, y, z: each axis reflected so that d >= 0, the axes ordered by
    |d|, largest first, the earlier axis first at a tie.  Canonical offset
    (u0, u1, u2) is original offset (s_x u[r_x], s_y u[r_y], s_z u[r_z]) with
    s the reflections and r the canonical positions of the original axes
    (order=False: the reflections only, r the identity)."""
    d = np.stack([dx, dy, dz], -1)
    a = np.abs(d)
    r = np.stack([(a[:, 1] > a[:, 0]).astype(int) + (a[:, 2] > a[:, 0]),
                  (a[:, 0] >= a[:, 1]).astype(int) + (a[:, 2] > a[:, 1]),
                  (a[:, 0] >= a[:, 2]).astype(int) + (a[:, 1] >= a[:, 2])], -1)
    if not order:
        r = np.broadcast_to(np.arange(3), r.shape)
    s = np.where(d < 0, -1, 1)
    rows = np.arange(X.shape[0])
    out = np.empty_like(X)
    for f in range(27):
        u = np.array([f % 3 - 1, (f // 3) % 3 - 1, f // 9 - 1])
        o = s * u[r]  # original offset of canonical feature f, per sample
        out[:, f] = X[rows, (o[:, 0] + 1) + 3 * (o[:, 1] + 1) + 9 * (o[:, 2] + 1)]
    return out


def relax_config(n=BOX):
    return presets.static_droplet(n=n, radius=8.0, velocity_set=D3Q19, sigma=0.0, tau_g=TAU_G)


def make_dataset(n_drops=8192, steps=500, per_drop=384, seed=7, concurrent=16, log=print, label="surface"):
    """(X [n,27], T [n]) from n_drops relaxed droplets/ellipsoids, stencils in
    canonical orientation.  Shapes are drawn up front; their phase fields are
    built on a thread pool while `concurrent` simulations per box size, each
    on its own stream, relax earlier ones."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    shapes = [draw_shape(rng, sphere=rng.uniform() < 0.5) for _ in range(n_drops)]
    shapes.sort(key=lambda sh: sh[3])  # by box size
    sims = {}
    xs, ts = [], []
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4)) as pool:
        phis = pool.map(lambda sh: phi_field(*sh), shapes)
        done = 0
        while done < n_drops:
            n = shapes[done][3]
            if n not in sims:
                for m in list(sims):
                    for sim in sims.pop(m):
                        sim.close()
                sims[n] = [Simulation(relax_config(n)) for _ in range(concurrent)]
            rho = np.ones((n, n, n), dtype=np.float32)
            zero = np.zeros((n, n, n), dtype=np.float32)
            batch = []
            for sim in sims[n]:
                if done + len(batch) >= n_drops or shapes[done + len(batch)][3] != n:
                    break
                shape = shapes[done + len(batch)]
                sim.set_field("rho", rho)
                for k in ("ux", "uy", "uz"):
                    sim.set_field(k, zero)
                sim.set_field("phi", next(phis))
                sim.init_distros()
                sim.step(steps)  # asynchronous on the simulation's stream
                batch.append((sim, shape))
            for sim, shape in batch:
                phi = sim.field("phi")
                inner = np.zeros(phi.shape, dtype=bool)
                inner[1:-1, 1:-1, 1:-1] = True
                cand = np.flatnonzero(((phi > SAMPLE_LO) & (phi < SAMPLE_HI) & inner).ravel())
                if cand.size > per_drop:
                    cand = rng.choice(cand, per_drop, replace=False)
                axes, rot, centre, _ = shape
                xs.append(stencils(phi, cand))
                ts.append(4.0 * curvature_labels(axes, rot, centre, cand, n, label=label))
            done += len(batch)
            if done % 512 < len(batch) or done == n_drops:
                log(f"[closure] {done}/{n_drops} droplets relaxed, {time.time() - t0:.1f} s")
    for group in sims.values():
        for sim in group:
            sim.close()
    return np.concatenate(xs), np.concatenate(ts)


def train(X, T, epochs=120, batch=512, lr=2e-3, seed=3, device=0, log=print):
    """Adam with the learning rate halved five times over the run."""
    net = MLP(device)
    net.set_data(X, T)
    losses = []
    period = max(1, epochs // 5)
    for e0 in range(0, epochs, 4):
        n = min(4, epochs - e0)
        lr_e = lr * (0.5 ** (e0 // period))
        losses += list(net.train(n, batch=batch, lr=lr_e, seed=seed + e0, augment=False))
        log(f"[closure] epoch {e0 + n}: mse {losses[-1]:.3e} (lr {lr_e:.1e})")
    p = net.params()
    net.close()
    return p, np.array(losses)


def save(params, path=WEIGHTS, **meta):
    np.savez(path, params=np.asarray(params, dtype=np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def save_raw(params, path=WEIGHTS_RAW):
    np.asarray(params, dtype="<f4").tofile(path)


def load(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


def predict_curvature(params, X, device=0):
    """Curvature the closure applies for canonical stencils X [n, 27]
    (``stencils``): one network evaluation, as the timestep's closure pass."""
    net = MLP(device, params)
    y = net.predict(X).astype(np.float64)
    net.close()
    return 0.25 * y


def curvature_error(params, radii=(6, 8, 12, 16, 20), steps=500, n=64, velocity_set=D3Q19, sigma=1e-3):
    """Pointwise curvature of relaxed static spheres against the exact -2/R:
    the finite-difference div(n) the timestep computes and the network on the
    same phi stencils, over the cells the closure replaces."""
    out = []
    for R in radii:
        cfg = presets.static_droplet(n=n, radius=float(R), velocity_set=velocity_set, sigma=sigma, tau_g=TAU_G)
        with Simulation(cfg) as sim:
            sim.init().step(steps, save=True)
            phi = sim.field("phi")
            ind = sim.field("ind").astype(np.float64)
            nv = np.stack([sim.field(k) for k in ("normx", "normy", "normz")]).astype(np.float64)
            fs = np.stack([sim.field(k) for k in ("fsx", "fsy", "fsz")]).astype(np.float64)
        inner = np.zeros(phi.shape, bool)
        inner[1:-1, 1:-1, 1:-1] = True
        sel = inner & (phi > cfg.mlp_lo) & (phi < cfg.mlp_hi) & (ind > 1e-3)
        cells = np.flatnonzero(sel.ravel())
        n2 = (nv ** 2).sum(0).ravel()[cells]
        k_fd = -(fs * nv).sum(0).ravel()[cells] / (sigma * ind.ravel()[cells] * n2)
        k_mlp = predict_curvature(params, stencils(phi, cells))
        want = -2.0 / R
        row = {"R": R, "cells": int(cells.size)}
        for name, k in (("fd", k_fd), ("mlp", k_mlp)):
            row[name] = {"mean_rel": float((k.mean() - want) / abs(want)),
                         "rms_rel": float(np.sqrt(np.mean((k - want) ** 2)) / abs(want))}
        out.append(row)
    return out


# Steps to rest of the 64^3 static droplets: the pressure jump of R >= 12
# is still settling at 3000 steps (FD, R = 12 / 16 / 20: +0.3 / -2.1 /
# -4.8 % at 3000, +1.6 / +0.5 / +0.1 % at 6000, +1.73 / +0.87 / +0.67 % at
# 12000 and +1.73 / +0.87 / +0.58 % at 24000 steps), so the closure is judged
# at 12000.
REST_STEPS = 12000


def laplace(params, radii=(8, 12, 16), steps=REST_STEPS, sigma=0.01, n=64, velocity_set=D3Q19, **window):
    """Laplace pressure error |dp - 2 sigma/R| / (2 sigma/R) (and its sign,
    signed_err) and max spurious |u| of static droplets run to rest, FD
    curvature against the learned closure.  window: config overrides of the
    closure's windows (mlp_lo, mlp_hi, mlp_kmin, mlp_kmax)."""
    out = []
    for R in radii:
        row = {"R": R}
        for name, clos in (("fd", CLOSURE_FD), ("mlp", CLOSURE_MLP)):
            cfg = presets.static_droplet(n=n, radius=float(R), velocity_set=velocity_set, sigma=sigma,
                                         tau_g=TAU_G)
            cfg.closure = clos
            for k, v in window.items():
                setattr(cfg, k, v)
            with Simulation(cfg) as sim:
                if clos == CLOSURE_MLP:
                    sim.set_closure(*split_params(params))
                sim.init().step(steps, save=True)
                rho = sim.field("rho").astype(np.float64)
                u = np.stack([sim.field(k) for k in ("ux", "uy", "uz")]).astype(np.float64)
            z, y, x = np.mgrid[0:n, 0:n, 0:n]
            c = (n - 1) / 2
            r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
            inner = (x > 0) & (x < n - 1) & (y > 0) & (y < n - 1) & (z > 0) & (z < n - 1)
            dp = (rho[r < R - 4].mean() - rho[(r > R + 4) & inner].mean()) / 3.0
            want = 2 * sigma / R
            row[name] = {"dp": dp, "rel_err": abs(dp - want) / want, "signed_err": (dp - want) / want,
                         "spurious_umax": float(np.sqrt((u ** 2).sum(0))[inner].max())}
        out.append(row)
    return out


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="generate data, train and evaluate the learned closure")
    ap.add_argument("--drops", type=int, default=8192)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--per-drop", type=int, default=384)
    ap.add_argument("--epochs", type=int, default=120)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--eval", metavar="NPZ", help="only evaluate saved weights")
    ap.add_argument("--label", choices=("surface", "levelset"), default="surface",
                    help="curvature label: sharp surface at the closest point, or the level set through the cell")
    args = ap.parse_args(argv)
    t0 = time.time()
    if args.eval:
        params = load(args.eval)
        print(json.dumps({"curvature": curvature_error(params)}), flush=True)
        print(json.dumps({"laplace": laplace(params), "seconds": time.time() - t0}), flush=True)
        return
    X, T = make_dataset(args.drops, args.steps, args.per_drop, label=args.label)
    rng = np.random.default_rng(11)
    perm = rng.permutation(T.size)
    nval = T.size // 20
    val, tr = perm[:nval], perm[nval:]
    print(f"[closure] {T.size} samples ({time.time() - t0:.1f} s)", flush=True)
    params, losses = train(X[tr], T[tr], epochs=args.epochs)
    net = MLP(0, params)
    pred = net.predict(X[val])
    net.close()
    val_rmse = float(np.sqrt(np.mean((pred - T[val]) ** 2)))
    base_rmse = float(np.sqrt(np.mean((T[val] - T[tr].mean()) ** 2)))
    save(params, args.out, losses=losses, val_rmse=val_rmse, samples=T.size)
    save_raw(params, os.path.splitext(args.out)[0] + ".f32")
    print(json.dumps({"samples": int(T.size), "train_mse": float(losses[-1]), "val_rmse": val_rmse,
                      "val_rmse_constant": base_rmse}), flush=True)
    print(json.dumps({"curvature": curvature_error(params)}), flush=True)
    print(json.dumps({"laplace": laplace(params), "seconds": time.time() - t0}), flush=True)


if __name__ == "__main__":
    main()
