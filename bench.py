"""Benchmark of the phase-field LBM timestep on one MI355X.

Workload (one "step" = one full timestep of the whole grid): round jet,
D3Q27 regularized hydrodynamics + D3Q7 phase field, 256x256x1024 cells,
nozzle D=32, u=0.05, Re=5000, We=500, FP32 populations.  Inputs are
resident in HBM before timing starts.  Also reported: the two binary16
storage forms and the FP32 moment storage form of the same workload, the same workload with the learned curvature closure
(its cost in MLUPS, and its Laplace pressure error against the FD curvature
for R = 6..20), the Laplace pressure error of static droplets run
to rest (R = 8, 12, 16, 12000 steps), the MLUPS of the D3Q19 128x128x512 jet, the 64^3
droplet and the reference's own default 64x64x128 jet, and the CPU oracle
timed on the host on a bounded sample.

MLUPS counts nx*ny*nz cells per step, as the reference does
(src/main.cu:203-205); the interior-cell rate is reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_bytes_per_cell(q, s):
    """Algorithmic bytes of the fused collide-stream kernel per interior cell:
    read and write Q f populations, write the 3-component phase flux (both
    in the storage type), read phi."""
    return 2 * q * s + 3 * s + 4


def run_form(pf, cfg, steps, warmup, closure_params=None, graph=False):
    sim = pf.Simulation(cfg)
    if closure_params is not None:
        sim.set_closure_params(closure_params)
    sim.init()
    sim.step(warmup, graph=graph)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(steps, graph=graph)
    sim.sync()
    wall = time.perf_counter() - t0
    tm = sim.time_steps(steps)  # HIP events on the simulation's stream
    sim.step(1, save=True)
    d = sim.diagnostics()
    sim.close()
    return wall, tm, d


def laplace(pf, presets, R, steps=12000):
    cfg = presets.static_droplet(n=64, radius=float(R), sigma=0.01)
    with pf.Simulation(cfg) as sim:
        sim.init().step(steps, save=True)
        rho = sim.field("rho").astype(np.float64)
        ux, uy, uz = (sim.field(k).astype(np.float64) for k in ("ux", "uy", "uz"))
    n = cfg.nx
    z, y, x = np.mgrid[0:n, 0:n, 0:n]
    c = (n - 1) / 2
    r = np.sqrt((x - c) ** 2 + (y - c) ** 2 + (z - c) ** 2)
    interior = (x > 0) & (x < n - 1) & (y > 0) & (y < n - 1) & (z > 0) & (z < n - 1)
    inside = rho[(r < R - 4)].mean()
    outside = rho[(r > R + 4) & interior].mean()
    dp = (inside - outside) / 3.0
    want = 2 * cfg.sigma / R
    umax = float(np.sqrt(ux ** 2 + uy ** 2 + uz ** 2)[interior].max())
    return {"R": R, "dp": dp, "dp_exact": want, "rel_err": abs(dp - want) / want, "spurious_umax": umax}


def cpu_baseline(cfg, steps):
    import oracle
    nthr = oracle.num_threads()
    o = oracle.Oracle(cfg, 32).init()
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    cells = cfg.nx * cfg.ny * cfg.nz
    o.close()
    return {"value": cells * steps / dt / 1e6, "unit": "MLUPS", "cores": nthr, "kind": "port",
            "sample": f"C++/OpenMP FP32 oracle, same {cfg.nx}x{cfg.ny}x{cfg.nz} D3Q{cfg.velocity_set} jet, "
                      f"{steps} timesteps after initialisation (wall clock of the steps only)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--closure-warmup", type=int, default=3000,
                    help="steps run before timing the learned-closure form against the FD form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip FP16 form and Laplace probes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the timestep "
              "would shard as z-slabs across GPUs, which this build does not run)", file=sys.stderr)
        sys.exit(2)

    import phasefieldlbm_amd as pf
    from phasefieldlbm_amd import presets

    cfg = presets.jet_d3q27_256x256x1024()
    cells = cfg.nx * cfg.ny * cfg.nz
    inner = (cfg.nx - 2) * (cfg.ny - 2) * (cfg.nz - 2)
    log(f"[bench] workload D3Q27 jet {cfg.nx}x{cfg.ny}x{cfg.nz} FP32, warmup {args.warmup}, steps {args.steps}")
    wall, tm, diag = run_form(pf, cfg, args.steps, args.warmup)
    ms = wall * 1e3 / args.steps
    mlups = cells * args.steps / wall / 1e6
    b_alg = presets.bytes_per_update(cfg)
    b_moved = presets.bytes_moved_per_update(cfg)
    step_ms = tm["step_ms"] / args.steps
    kbytes = kernel_bytes_per_cell(cfg.velocity_set, 4) * inner
    achieved = kbytes / (step_ms * 1e-3) / 1e9
    log(f"[bench] fp32: {mlups:.0f} MLUPS, {ms:.3f} ms/step, kernel {step_ms:.3f} ms, "
        f"phase {tm['phase_sum_ms'] / args.steps:.3f} ms, bc {tm['bc_ms'] / args.steps:.3f} ms")

    # HBM bytes per k_step launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE summary (a "traffic_bytes_per_launch:" line), if present.
    def profiled_traffic(form):
        tpath = os.path.join(ROOT, "profiles", f"traffic_k_step_jet_d3q27_256x256x1024_{form}.md")
        if os.path.exists(tpath):
            for line in open(tpath):
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
        return None

    traffic = profiled_traffic("fp32")
    bw = pf.stream_bandwidth(0, 4 << 30, 10)
    copy_gbs = bw["copy_gbs"]
    log(f"[bench] streaming bandwidth: read-only {bw['read_gbs']:.0f}, write-only {bw['write_gbs']:.0f}, "
        f"copy {copy_gbs:.0f} GB/s")

    out = {
        "metric": "MLUPS (D3Q27 round jet 256x256x1024, FP32 populations)",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "jet_d3q27_256x256x1024", "velocity_set": "D3Q27", "phase_set": "D3Q7",
                   "grid": [cfg.nx, cfg.ny, cfg.nz], "nozzle_diameter": cfg.diam, "u_inf": cfg.u_inf,
                   "reynolds": cfg.reynolds_water, "weber": cfg.weber, "inflow_noise": "uniform +-5% u_z",
                   "storage": "fp32", "parallelism": "single GPU"},
        "interior_mlups": round(inner * args.steps / wall / 1e6, 1),
        # whole step at the bytes this scheme moves per update (physical)
        "moved_roofline_frac": round(mlups * 1e6 * b_moved / (HBM_PEAK_GBS * 1e9), 4),
        "bytes_moved_per_update": b_moved,
        "step_frac_of_copy": round(mlups * 1e6 * b_moved / (copy_gbs * 1e9), 4),
        # the same MLUPS priced at the reference scheme's bytes per update
        # (f and the seven g read and written, phi once): an equivalence figure
        "reference_scheme_roofline_frac": round(mlups * 1e6 * b_alg / (HBM_PEAK_GBS * 1e9), 4),
        "bytes_per_update": b_alg,
        "measured_bandwidth_gbs": {k: round(v, 1) for k, v in bw.items()},
        "roofline": {"bound": "hbm", "kernel": "k_step", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": kbytes, "avg_launch_ms": round(step_ms, 4)},
        "kernel_ms": {"phase_sum": round(tm["phase_sum_ms"] / args.steps, 4), "step": round(step_ms, 4),
                      "jet_bc": round(tm["bc_ms"] / args.steps, 4)},
        "diagnostics": {"mass": diag["mass"], "phi_mass": diag["phi_mass"], "umax": diag["umax"]},
    }

    if not args.no_extras:
        c16 = presets.jet_d3q27_256x256x1024(pf.STORE_FP16)
        wall16, tm16, _ = run_form(pf, c16, args.steps, args.warmup)
        mlups16 = cells * args.steps / wall16 / 1e6
        b16 = presets.bytes_per_update(c16)
        b16m = presets.bytes_moved_per_update(c16)
        out["fp16_storage"] = {"value": round(mlups16, 1), "unit": "MLUPS",
                               "speedup_vs_fp32": round(mlups16 / mlups, 3),
                               "moved_roofline_frac": round(mlups16 * 1e6 * b16m / (HBM_PEAK_GBS * 1e9), 4),
                               "reference_scheme_roofline_frac": round(mlups16 * 1e6 * b16 / (HBM_PEAK_GBS * 1e9), 4),
                               "kernel_ms": round(tm16["step_ms"] / args.steps, 4), "bytes_per_update": b16,
                               "bytes_moved_per_update": b16m,
                               "kernel_bytes_per_launch": kernel_bytes_per_cell(c16.velocity_set, 2) * inner,
                               "kernel_traffic": profiled_traffic("fp16")}
        log(f"[bench] fp16 storage: {mlups16:.0f} MLUPS ({mlups16 / mlups:.2f}x)")
        # binary16 non-equilibrium storage (populations relative to the
        # writer's second-order equilibrium, 8 B of moments per cell more)
        cne = presets.jet_d3q27_256x256x1024(pf.STORE_FP16_NE)
        wallne, tmne, _ = run_form(pf, cne, args.steps, args.warmup)
        mlupsne = cells * args.steps / wallne / 1e6
        bne = presets.bytes_moved_per_update(cne)
        out["fp16_ne_storage"] = {"value": round(mlupsne, 1), "unit": "MLUPS",
                                  "speedup_vs_fp32": round(mlupsne / mlups, 3),
                                  "moved_roofline_frac": round(mlupsne * 1e6 * bne / (HBM_PEAK_GBS * 1e9), 4),
                                  "kernel_ms": round(tmne["step_ms"] / args.steps, 4),
                                  "bytes_moved_per_update": bne,
                                  "kernel_bytes_per_launch": (2 * cne.velocity_set * 2 + 3 * 2 + 4 + 16) * inner,
                                  "kernel_traffic": profiled_traffic("fp16_ne")}
        log(f"[bench] fp16 non-equilibrium storage: {mlupsne:.0f} MLUPS ({mlupsne / mlups:.2f}x)")
        # moment storage (FP32 coefficients of each cell's collision output,
        # 17 per D3Q27 cell, instead of 27 populations)
        cmo = presets.jet_d3q27_256x256x1024(pf.STORE_MOM)
        wallmo, tmmo, _ = run_form(pf, cmo, args.steps, args.warmup)
        mlupsmo = cells * args.steps / wallmo / 1e6
        bmo = presets.bytes_moved_per_update(cmo)
        kmo = (2 * 17 * 4 + 3 * 4 + 4) * inner
        mo_ms = tmmo["step_ms"] / args.steps
        out["moment_storage"] = {"value": round(mlupsmo, 1), "unit": "MLUPS", "dtype": "f32",
                                 "speedup_vs_fp32": round(mlupsmo / mlups, 3),
                                 "moved_roofline_frac": round(mlupsmo * 1e6 * bmo / (HBM_PEAK_GBS * 1e9), 4),
                                 "reference_scheme_roofline_frac": round(mlupsmo * 1e6 * b_alg / (HBM_PEAK_GBS * 1e9), 4),
                                 "kernel_ms": round(mo_ms, 4), "bytes_moved_per_update": bmo,
                                 "kernel_bytes_per_launch": kmo,
                                 "kernel_frac": round(kmo / (mo_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                 "phase_sum_ms": round(tmmo["phase_sum_ms"] / args.steps, 4),
                                 "kernel_traffic": profiled_traffic("mom")}
        log(f"[bench] moment storage: {mlupsmo:.0f} MLUPS ({mlupsmo / mlups:.2f}x), kernel {mo_ms:.3f} ms")
        # Learned closure against the FD curvature on a developed jet (its
        # interface has grown for `closure_warmup` steps), same step count.
        from phasefieldlbm_amd import closure
        cw = args.closure_warmup
        wallf, tmf, _ = run_form(pf, presets.jet_d3q27_256x256x1024(), args.steps, cw)
        cm = presets.jet_d3q27_256x256x1024()
        cm.closure = pf.CLOSURE_MLP
        wallm, tmm, diagm = run_form(pf, cm, args.steps, cw, closure_params=closure.load())
        mlupsf = cells * args.steps / wallf / 1e6
        mlupsm = cells * args.steps / wallm / 1e6
        out["learned_closure"] = {"value": round(mlupsm, 1), "unit": "MLUPS", "fd_value": round(mlupsf, 1),
                                  "cost_frac": round(1.0 - mlupsm / mlupsf, 4), "after_steps": cw,
                                  "phase_and_closure_ms": round(tmm["phase_sum_ms"] / args.steps, 4),
                                  "phase_ms_fd": round(tmf["phase_sum_ms"] / args.steps, 4),
                                  "step_ms": round(tmm["step_ms"] / args.steps, 4),
                                  "umax": diagm["umax"]}
        log(f"[bench] learned closure after {cw} steps: {mlupsm:.0f} MLUPS against {mlupsf:.0f} with FD "
            f"(cost {100 * (1 - mlupsm / mlupsf):.1f} %)")
        # Laplace pressure jump of static droplets run to rest (64^3, 12000
        # steps, tau_g = 1, the closure's training regime): FD curvature
        # against the closure
        rows = closure.laplace(closure.load(), radii=(6, 7, 8, 10, 12, 14, 16, 20))
        out["learned_closure"]["laplace"] = [
            {"R": r["R"], "fd_rel_err": round(r["fd"]["rel_err"], 5), "mlp_rel_err": round(r["mlp"]["rel_err"], 5),
             "fd_signed_err": round(r["fd"]["signed_err"], 5), "mlp_signed_err": round(r["mlp"]["signed_err"], 5),
             "fd_spurious_umax": r["fd"]["spurious_umax"], "mlp_spurious_umax": r["mlp"]["spurious_umax"]}
            for r in rows]
        out["learned_closure"]["laplace_steps"] = closure.REST_STEPS
        out["learned_closure"]["laplace_mean_rel_err"] = {
            "fd": round(float(np.mean([r["fd"]["rel_err"] for r in rows])), 5),
            "mlp": round(float(np.mean([r["mlp"]["rel_err"] for r in rows])), 5)}
        log(f"[bench] closure laplace mean |dP err|: {out['learned_closure']['laplace_mean_rel_err']}")
        out["laplace"] = [laplace(pf, presets, R) for R in (8, 12, 16)]
        # the other configurations of the workload set, FP32, for reference
        others = {}
        for name in ("jet_d3q19_128x128x512", "jet_d3q19_128x128x512_mom", "droplet_d3q19_64"):
            make, _ = presets.WORKLOADS[name]
            c = make()
            w, _, _ = run_form(pf, c, args.steps, args.warmup)
            others[name] = round(c.nx * c.ny * c.nz * args.steps / w / 1e6, 1)
        out["other_workloads_mlups"] = others
        log(f"[bench] other workloads: {others}")
        # the step replayed as a captured graph per parity (the reference's
        # execution model) against direct launches, on the small grids where
        # launch overhead could show; wall clock per step
        graph_ms = {}
        for name, make in (("droplet_d3q19_64", presets.WORKLOADS["droplet_d3q19_64"][0]),
                           ("reference_jet_d3q19_64x64x128", lambda: presets.reference_jet(pf.D3Q19))):
            row = {}
            for tag, g in (("direct", False), ("graph", True)):
                w, _, _ = run_form(pf, make(), args.steps, args.warmup, graph=g)
                row[tag] = round(1e3 * w / args.steps, 4)
            graph_ms[name] = row
        out["graph_vs_direct_ms_per_step"] = graph_ms
        log(f"[bench] graph vs direct ms/step: {graph_ms}")
        # the reference's own default jet (64x64x128, D = 10, Gaussian inflow
        # noise), at which its README quotes 470 (D3Q19) and 410 (D3Q27)
        # MLUPS on an RTX 3050, probably with binary16 f; context only
        ref_default = {}
        for vs in (pf.D3Q19, pf.D3Q27):
            for st, tag in ((pf.STORE_FP32, "fp32"), (pf.STORE_FP16, "fp16")):
                c = presets.reference_jet(vs)
                c.storage = st
                w, _, _ = run_form(pf, c, args.steps, args.warmup)
                ref_default[f"d3q{vs}_{tag}"] = round(c.nx * c.ny * c.nz * args.steps / w / 1e6, 1)
        out["reference_default_jet_mlups"] = ref_default
        log(f"[bench] reference default jet 64x64x128: {ref_default}")
        log(f"[bench] laplace: {[round(r['rel_err'], 4) for r in out['laplace']]}")

    if not args.no_cpu:
        log(f"[bench] cpu baseline: {args.cpu_steps} steps of the oracle")
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
