"""Benchmark of the pllcart3d time step on one MI355X.

metric: cell-steps/s at 256x128x128 (interior cells x time steps per second);
a step is the CFL reduction plus the three RK3-CN substages.  Inputs are
resident in HBM before the timed region; the timed region is K whole steps
bracketed by stream synchronisation, timed with HIP events on the solver's
stream.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config front_256x128x128] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s at 256x128x128, FP64 plain vs mixed/batched, L2 err vs FP64"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of each kernel class, in full-field touches of the
# interior cells per launch (reads + writes of FP64 fields; the viscosity
# and the A(mu) coefficients count as recomputed from c, scratch round trips
# of the line solvers do not count).
TOUCHES = {
    "rhs_c": 7, "rhs_u": 8, "rhs_v": 8, "rhs_w": 8,
    "sweep_y": None, "sweep_z": None, "sweep_x": None,   # per-field, see below
    "div": 4, "velo_corr": 7, "pg": 8, "corr_pg": 14, "cfl": 1.5, "visc": 5,
}
# sweeps per substage: c then u, v, w (one launch per direction for c and one
# for the three velocity components): y/z touch 2 + 3 + 3 + 3 fields, x
# 3 + 4 + 4 + 4; per launch that is the substage total over the launches
SWEEP_TOUCHES_PER_SUBSTAGE = {"sweep_y": 11.0, "sweep_z": 11.0, "sweep_x": 15.0}


# the learned closure's set-up: target horizon (coarse steps), the scale of
# the applied correction, its refresh period (steps; kept in between), and
# the coarse runs' Timestep caps (the fine run's 2e-3, and larger)
CLOSURE_HORIZON = 4
CLOSURE_ALPHA = 0.5
CLOSURE_EVERY = 16
CLOSURE_TIMESTEPS = (2e-3, 2.25e-3, 2.5e-3)

# initial-condition recipes, as the workload names them
IC_NAMES = {0: "erf front", 1: "noise front", 2: "16x16-mode front"}


def dct_is_fast(prm):
    """ny = nz = 128 or 256 run the fast transform (unless PLLC_DCT=gemm):
    every launch reads and writes the packed block once; other plane sizes
    run the dense contractions on the matrix cores."""
    return prm.ny == prm.nz and prm.ny in (128, 256) and os.environ.get("PLLC_DCT") != "gemm"


def touches_per_launch(k, launches_per_step, prm):
    if k in SWEEP_TOUCHES_PER_SUBSTAGE:
        return SWEEP_TOUCHES_PER_SUBSTAGE[k] * 3.0 / launches_per_step
    if k == "dct":
        return 2.0 if dct_is_fast(prm) else None
    return TOUCHES.get(k)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="front_256x128x128")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel timing")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble line")
    ap.add_argument("--no-large", action="store_true", help="skip the 512x256x256 plain/mixed line")
    ap.add_argument("--ensemble-members", type=int, default=64)
    ap.add_argument("--no-closure", action="store_true", help="skip the learned-closure line")
    return ap.parse_args()


def host_cpu():
    """The host's CPU model and the cores this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def _oracle_rate(prm, c0, cells, threads, max_steps, max_s):
    from oracle import oracle
    oracle.set_threads(threads)
    orc = oracle.Oracle(prm)
    orc.init(c0)
    t0 = time.perf_counter()
    n = 0
    while True:
        orc.run(1)
        n += 1
        el = time.perf_counter() - t0
        if el > max_s or n >= max_steps:
            break
    return cells * n / el, n, el


def cpu_baseline(prm, c0, cells, threads):
    """The FP64 oracle on the host cores: a bounded sample of the same
    workload on `threads` OpenMP threads (the reported value), and one step
    on a single thread."""
    v, n, el = _oracle_rate(prm, c0, cells, threads, 5, 10.0)
    v1, n1, el1 = _oracle_rate(prm, c0, cells, 1, 1, 0.0)
    return {"value": v, "unit": "cell-steps/s", "cores": threads, "kind": "port",
            "sample": "%d step(s) of the %dx%dx%d workload from the same initial state, %.1f s" %
                      (n, prm.nx, prm.ny, prm.nz, el),
            "one_thread": {"value": v1, "unit": "cell-steps/s", "cores": 1,
                           "sample": "%d step(s), %.1f s" % (n1, el1)},
            "host": host_cpu()}


def kernel_table(sim, ksteps):
    import pllcart3d_amd as P
    names = P.load().pllc_kernel_names().decode().split(";")
    out = {}
    for nm in names:
        avg, n = sim.time_kernel(nm, ksteps)
        if n:
            out[nm] = {"avg_ms": avg, "launches_per_step": n / float(ksteps), "ms_per_step": avg * n / ksteps}
    return out


# Peaks (MI355X datasheet): HBM3E 8 TB/s; dense FP64 78.6 TFLOP/s (vector and
# matrix); FP32 157.3 TFLOP/s.
PEAK_TFLOPS = {8.0: 78.6, 4.0: 157.3}


# HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated by the copy
# kernel) of the 256x128x128 workload, read from the committed rocprofv3
# summaries (profiles/step_fp64_256x128x128.md, step_mixed_256x128x128.md:
# their "HBM read MB | HBM write MB" table, per dispatch), per kernel class;
# the DCT is the mean of the forward and inverse plane transforms
PROFILES = os.path.join(ROOT, "profiles")
CLASS_KERNELS = {
    "sweep_x": ["k_sweep_part<{T}, 0>"], "sweep_y": ["k_sweep_part<{T}, 1>"],
    "sweep_z": ["k_sweep_part_z<{T}>", "k_sweep_part_zt<{T}>"], "corr_pg": ["k_corr_pg<{T}, false>"],
    "rhs_c": ["k_rhs_c<{T}, true, false>"], "rhs_u": ["k_rhs_mom<{T}, 1, true, false>"],
    "rhs_v": ["k_rhs_mom<{T}, 2, true, false>"], "rhs_w": ["k_rhs_mom<{T}, 3, true, false>"],
    "div": ["k_div<{T}, false>"], "visc": ["k_visc<{T}, false>"],
    "dct": ["k_dct_plane<{T}, false>", "k_dct_plane<{T}, true>"],
}


def measured_traffic_mb(bytes_per_value):
    """{kernel class: HBM MB read + written per launch} from the profile
    summary of the 256x128x128 workload in this precision ({} when absent)."""
    name = "step_fp64_256x128x128.md" if bytes_per_value == 8.0 else "step_mixed_256x128x128.md"
    T = "double" if bytes_per_value == 8.0 else "float"
    rows = {}
    try:
        with open(os.path.join(PROFILES, name)) as f:
            table = False
            for line in f:
                if "HBM read MB" in line:
                    table = True
                    continue
                if table and line.startswith("| `"):
                    cells = [c.strip() for c in line.strip().strip("|").split("|")]
                    rows[cells[0].strip("`")] = round(float(cells[2]) + float(cells[3]), 1)
                elif table and not line.startswith("|"):
                    table = False
    except OSError:
        return {}
    out = {}
    for cls, names in CLASS_KERNELS.items():
        vals = [rows[n.format(T=T)] for n in names if n.format(T=T) in rows]
        if cls == "dct" and len(vals) == 2:
            out[cls] = round(sum(vals) / 2, 1)
        elif cls != "dct" and vals:
            out[cls] = vals[0]
    return out


def dct_flops_per_launch(prm):
    """Dense REDFT contractions: the four launches of a substage are two
    along z (M = nx ny, N = K = nz) and two along y (nx batches of
    M = K = ny, N = nz), 2 M N K flops each; the mean per launch."""
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    return (2.0 * nx * ny * nz * nz + 2.0 * nx * ny * ny * nz) / 2.0


def roofline(kernels, cells, bytes_per_value, prm):
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    tmb = measured_traffic_mb(bytes_per_value).get(dom) if (prm.nx, prm.ny, prm.nz) == (256, 128, 128) else None
    traffic = {dom: round(tmb * 1e6)} if tmb else None
    avg_s = kernels[dom]["avg_ms"] * 1e-3
    if dom == "dct" and not dct_is_fast(prm):
        # the dense contractions are matrix-core work; the fast transform
        # counts as HBM
        fl = dct_flops_per_launch(prm)
        ach = fl / avg_s / 1e12
        peak = PEAK_TFLOPS[bytes_per_value]
        return {"bound": "mfma", "achieved": round(ach, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 4), "traffic": (traffic or {}).get(dom), "kernel": dom,
                "algorithmic_flops_per_launch": fl, "avg_launch_ms": round(kernels[dom]["avg_ms"], 5)}
    t = touches_per_launch(dom, kernels[dom]["launches_per_step"], prm)
    if t is None:
        return {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None,
                "traffic": None, "kernel": dom}
    ab = t * cells * bytes_per_value
    ach = ab / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": (traffic or {}).get(dom), "kernel": dom,
            "algorithmic_bytes_per_launch": ab, "avg_launch_ms": round(kernels[dom]["avg_ms"], 5)}


def hbm_table(kernels, cells, bytes_per_value, prm):
    """Achieved algorithmic GB/s of every kernel class with an AB model."""
    out = {}
    for k, v in kernels.items():
        t = touches_per_launch(k, v["launches_per_step"], prm)
        if t is None or not v["avg_ms"]:
            continue
        out[k] = round(t * cells * bytes_per_value / (v["avg_ms"] * 1e-3) / 1e9, 1)
    return out


def ensemble_line(a):
    """The batched ensemble (64 members of three grid sizes, random R and Pe)
    in the mixed precision, the members of each grid size stepped as one
    batch: aggregate interior cell-steps per second."""
    import pllcart3d_amd as P
    from pllcart3d_amd.configs import ensemble_members
    mem = ensemble_members(a.ensemble_members)
    ens = P.Ensemble([m[0] for m in mem], "mixed")
    for (prm, rec, seed, amp), sim in zip(mem, ens.members):
        sim.init(P.ic_fill(prm, rec, seed, amp))
    cells = sum(m[0].nx * m[0].ny * m[0].nz for m in mem)
    steps = max(2, min(a.steps, 10))
    ms = ens.time_steps(2, steps)
    batches = ens.batches()
    # the batched Poisson transform on the matrix cores (FP32 MFMA, after
    # the timed steps: each batch's members advance alone here): per
    # launch, every plane of every member contracted along z and y
    mfma = []
    for b in range(batches):
        nm, (nx, ny, nz) = ens.batch_info(b)
        avg, n = ens.time_kernel(b, "dct", 2)
        flops = nm * nx * 2.0 * (ny * nz * nz + ny * ny * nz)
        mfma.append({"members": nm, "grid": "%dx%dx%d" % (nx, ny, nz), "planes": nm * nx, "launches": n,
                     "avg_ms": round(avg, 5), "tflops": round(flops / (avg * 1e-3) / 1e12, 1),
                     "frac_fp32_mfma_peak": round(flops / (avg * 1e-3) / 157.3e12, 3)})
    ens.close()
    return {"value": cells * steps / (ms * 1e-3), "unit": "cell-steps/s", "members": len(mem),
            "batches": batches, "cells": cells, "steps": steps, "ms_per_step": ms / steps, "dtype": "f32/f64",
            "workload": "ensemble: %d members from {64x32x32, 128x48x48, 192x64x64}, R~U[1,4], "
                        "Pe log-uniform [200,4000], mixed precision" % len(mem),
            "dct_mfma": mfma}


def large_line(a):
    """512x256x256 (33.5 M cells, Pe = 4000): plain FP64 and mixed from the
    same state, and the mixed run's relative L2 error vs the FP64 run."""
    import numpy as np
    import pllcart3d_amd as P
    from pllcart3d_amd.configs import CONFIGS, config_params
    prm = config_params("front_512x256x256")
    cfg = CONFIGS["front_512x256x256"]
    c0 = P.ic_fill(prm, cfg["ic"], cfg["seed"], cfg["amp"])
    cells = prm.nx * prm.ny * prm.nz
    steps = max(2, min(a.steps, 6))
    out, sims = {}, {}
    for prec in ("plain", "mixed"):
        s = P.Sim(prm, prec)
        s.init(c0)
        ms = s.time_steps(2, steps)
        sims[prec] = s
        out[prec] = {"value": cells * steps / (ms * 1e-3), "ms_per_step": ms / steps}
    err = {}
    for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI"):
        ref = sims["plain"].download(f)
        err[f] = float(np.linalg.norm(sims["mixed"].download(f) - ref) / np.linalg.norm(ref))
    for s in sims.values():
        s.close()
    out["workload"] = "512x256x256 box 4x1x1, R=3 Pe=4000, 16x16-mode front"
    out["steps"] = steps
    out["speedup_mixed_vs_plain"] = out["plain"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    out["mixed_rel_l2_vs_fp64_after_%d_steps" % (steps + 2)] = err
    return out


def closure_line(a):
    """The learned closure (cfg: coarse 128x64x64 FP32 run with the 3-layer
    3x3x3 CNN, R=3, Pe=2000, compared with the fine 256x128x128 plain FP64
    run): training data from 4 box-filtered fine runs with R ~ U[2,4],
    Pe ~ U[1000,3000] (seeds 500..503; held-out run seed 600), training on
    the GPU with the library's BF16 kernels, then the fine run of the
    256x128x128 workload for the workload's 1000 steps against the coarse
    run without and with the closure (correction refreshed every
    CLOSURE_EVERY steps) up to the same time, at each Timestep cap of
    CLOSURE_TIMESTEPS: mixing-zone lengths, their error against the fine
    run's, and the cost ratio (fine GPU time over coarse wall time)."""
    import numpy as np
    import pllcart3d_amd as P
    from pllcart3d_amd import closure as CL
    from pllcart3d_amd.configs import CONFIGS, config_params
    t0 = time.perf_counter()
    fine_shape = (256, 128, 128)
    fp = config_params("front_256x128x128")
    cfg = CONFIGS["front_256x128x128"]
    cp = CL.coarse_params(fp)
    mask = CL.loss_mask(cp)
    quiet = lambda *x: None
    X, T = CL.collect(fine_shape, CL.draw_training_runs(4, 500), 600, 60, 60, quiet, CLOSURE_HORIZON)
    Xv, Tv = CL.collect(fine_shape, CL.draw_training_runs(1, 600), 600, 120, 120, quiet, CLOSURE_HORIZON)
    t1 = time.perf_counter()
    net, norm, losses = CL.train(X, T, mask, (cp.nx, cp.ny, cp.nz), 3000, 1, quiet)
    t2 = time.perf_counter()
    trained, none, gain = CL.validation_gain(net, norm, Xv, Tv, mask)
    shift, scale, osc = norm
    net.set_norm(shift, scale, osc * CLOSURE_ALPHA)
    c0 = P.ic_fill(fp, cfg["ic"], cfg["seed"], cfg["amp"])
    fine = P.Sim(fp, "plain")
    fine.init(c0)
    ms_f = fine.time_steps(0, cfg["steps"])
    t_end = CL.sim_time(fine)
    Lf = CL.mixing_length(fine.download("C_OLD"), fp)
    fine.close()
    out = {"workload": "coarse 128x64x64 mixed + 3-layer 3x3x3 CNN (16 ch, BF16 MFMA) vs fine 256x128x128 "
                       "plain FP64, R=3 Pe=2000, 16x16-mode front, to the fine run's time after %d steps"
                       % cfg["steps"],
           "training": {"fine_runs": 4, "samples": len(X), "horizon_steps": CLOSURE_HORIZON,
                        "adam_steps": len(losses), "loss_first": float(losses[:50].mean()),
                        "loss_last": float(losses[-50:].mean()), "data_s": round(t1 - t0, 1),
                        "train_s": round(t2 - t1, 1)},
           "heldout": {"loss_trained": trained, "loss_no_correction": none, "gain": gain},
           "fine": {"steps": cfg["steps"], "t": t_end, "gpu_ms": ms_f, "mixing_length": Lf},
           "correction_scale": CLOSURE_ALPHA}
    # the coarse run without and with the closure at each Timestep cap; the
    # cheapest of each whose mixing length is within 2% of the fine run's
    runs = []
    for dt in CLOSURE_TIMESTEPS:
        for name, n in (("plain", None), ("closure", net)):
            L, wall, steps, fb, worst = CL.coarse_run(fp, c0, t_end, n, dt, tol_bound=2e-2, every=CLOSURE_EVERY)
            runs.append({"coarse": name, "timestep_cap": dt, "steps": steps, "wall_ms": round(wall * 1e3, 1),
                         "cost_ratio": round(ms_f / (wall * 1e3), 2), "mixing_length": L,
                         "err_pct": round(100.0 * (L - Lf) / Lf, 2), "fallback_steps": fb,
                         "guard_worst_bound_excess": float(worst[0])})
    out["refresh_every_steps"] = CLOSURE_EVERY
    out["runs"] = runs
    for name in ("plain", "closure"):
        ok = [r for r in runs if r["coarse"] == name and abs(r["err_pct"]) <= 2.0]
        out["best_within_2pct_" + name] = max(ok, key=lambda r: r["cost_ratio"]) if ok else None
    net.close()
    out["seconds"] = round(time.perf_counter() - t0, 1)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import pllcart3d_amd as P
    from pllcart3d_amd.configs import CONFIGS, config_params
    cfg = CONFIGS[a.config]
    prm = config_params(a.config)
    cells = prm.nx * prm.ny * prm.nz
    c0 = P.ic_fill(prm, cfg["ic"], cfg["seed"], cfg["amp"])

    # the sweeps' block shapes are tuned on throwaway simulations first
    tuned = {}
    for prec in ("plain", "mixed"):
        t = P.Sim(prm, prec)
        t.init(c0)
        tuned[prec] = t.autotune(3)
        t.close()

    plain = P.Sim(prm, "plain")
    plain.init(c0)
    ms = plain.time_steps(a.warmup, a.steps)
    value = cells * a.steps / (ms * 1e-3)

    mixed = P.Sim(prm, "mixed")
    mixed.init(c0)
    ms_m = mixed.time_steps(a.warmup, a.steps)
    # both runs advanced warmup+steps steps from the same state: the plain
    # FP64 run is the reference trajectory of the mixed one
    err = {}
    for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI"):
        ref = plain.download(f)
        err[f] = float(np.linalg.norm(mixed.download(f) - ref) / np.linalg.norm(ref))

    roof = None
    kernels = {}
    kernels_m = {}
    if not a.no_kernels:
        ksteps = max(2, min(5, a.steps))
        kernels = kernel_table(plain, ksteps)
        kernels_m = kernel_table(mixed, ksteps)
        roof = roofline(kernels, cells, 8.0, prm)
    ens = None
    if not a.no_ensemble:
        ens = ensemble_line(a)
    large = None
    if not a.no_large:
        large = large_line(a)
    closure = None
    if not a.no_closure:
        closure = closure_line(a)
    step_ab = 2256.0 * cells   # whole-step algorithmic bytes (282 field touches x 8 B)
    cpu = None
    if not a.no_cpu:
        thr = a.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(prm, c0, cells, thr)
    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "%dx%dx%d box 4x1x1, R=%g Pe=%g, %s, RK3-CN step (plain FP64; mixed in 'mixed')" %
                   (prm.nx, prm.ny, prm.nz, prm.Vratio, prm.Pe, IC_NAMES[cfg["ic"]]), "cells": cells},
        "roofline": roof,
        "cpu_baseline": cpu,
        "step_hbm_fraction": round(step_ab / (ms / a.steps * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "mixed": {"value": cells * a.steps / (ms_m * 1e-3), "ms_per_step": ms_m / a.steps, "dtype": "f32/f64",
                  "speedup_vs_plain": ms / ms_m,
                  "rel_l2_vs_fp64_after_%d_steps" % (a.warmup + a.steps): err,
                  "roofline": roofline(kernels_m, cells, 4.0, prm) if kernels_m else None,
                  "kernel_gbs": hbm_table(kernels_m, cells, 4.0, prm) if kernels_m else None},
        "kernel_gbs": hbm_table(kernels, cells, 8.0, prm) if kernels else None,
        "ensemble": ens,
        "large": large,
        "closure": closure,
        "sweep_lines_per_block": {k: {"x": v[0], "y": v[1], "z": v[2]} for k, v in tuned.items()},
        "kernels": {k: {kk: round(vv, 5) for kk, vv in v.items()} for k, v in kernels.items()},
        "kernels_mixed": {k: {kk: round(vv, 5) for kk, vv in v.items()} for k, v in kernels_m.items()},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
