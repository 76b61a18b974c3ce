#!/usr/bin/env python3
"""F + dF/drho evaluations per second on one MI355X (classicalDFT hot path).

Workload (the metric's configuration): Lennard-Jones fluid with WCA split,
cutoff 2.5 sigma, kT = 0.8 eps, White-Bear II FMT with d = sigma, periodic
128^3 grid, dx = 0.1 sigma; a spherical droplet of radius 4 sigma (tanh width
0.5 sigma) between rho_l = 0.70 and rho_v = 0.02 plus seeded 1e-3 uniform
noise; mu = bulk mu(rho_v).  One step = one full evaluation of Omega, F and
delta(beta Omega)/delta rho * dV (functionals::total of the reference) on
device-resident inputs.

    python bench.py [--steps K] [--warmup W] [--precision mixed|fp64] [--no-cpu-baseline]

Prints one JSON line.  Only N=1 is supported (single-GPU build).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

METRIC = "F+δF/δρ evals/s at 128³: mixed vs FP64 on GPU, Ω rel err ≤1e-6"
W_BATCH_SIZES = (48, 64, 80, 96, 128, 160)  # the batched workload's grid sizes


def progress(msg):
    """A progress line on stderr (stdout carries only the JSON result)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)
PEAK_HBM_GBPS = 8000.0


def eval_bytes(shape, s_f, s_s):
    """Algorithmic bytes of one evaluation, stage-boundary model: N (46 s_f + 6 s_s) + 5 s_f M."""
    nx, ny, nz = shape
    N = nx * ny * nz
    M = nx * ny * (nz // 2 + 1)
    return N * (46 * s_f + 6 * s_s) + 5 * s_f * M


def cpu_baseline(rho, mu, shape, dx, budget_s=15.0):
    """Time the FP64 CPU oracle (a port of the reference algorithm) on the host cores."""
    from oracle import oracle as O
    from classicaldft_amd import workloads as W

    orc = O.Oracle(shape, dx, interactions=[W.LJ_WCA], kT=0.8, model="WhiteBearII")
    threads = O.max_threads()
    parts, forces = orc.total(rho, mu)  # warm-up, also the FP64 reference answer
    times = []
    t_start = time.perf_counter()
    while len(times) < 8 and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        orc.total(rho, mu)
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    # reference-faithful variant: the per-point Phi/dPhi loop on one thread (hard_sphere.hpp:185-221)
    orc.set_serial_pointwise(True)
    t0 = time.perf_counter()
    orc.total(rho, mu)
    t_faithful = time.perf_counter() - t0
    return {
        "value": 1.0 / med,
        "unit": "evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{len(times)} timed FP64 oracle evaluations of the same 128^3 workload (median), after 1 warm-up; "
                  f"FFTs and pointwise loop on {threads} OpenMP threads",
        "reference_faithful_evals_per_s": 1.0 / t_faithful,
        "nproc": os.cpu_count(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
    }, parts, forces


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_128cubed.json")


def measured_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (scaled by the plain-copy calibration recorded there), or None when it holds no entry for it."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    e = t.get("kernels", {}).get(precision, {}).get(kernel)
    return None if e is None else e["hbm_bytes_per_launch"]


def batched_fire(cd, W, precision, n=64, steps=500, sizes=W_BATCH_SIZES):
    """Batched form: n seeded LJ state points of uneven grid size (sizes drawn from 48^3 .. 160^3, kT ~
    U(0.7, 1.2), bulk rho ~ U(0.05, 0.80), droplet or slab), `steps` FIRE steps each, run one after another
    and then advanced together by cdft_minimize_grand_potential_batch.  Every context is autotuned first (the
    general-schedule sizes are timed once per size)."""
    import torch

    pts = W.batch_state_points(n, sizes=sizes)
    ctxs, mus = [], []
    t0 = time.perf_counter()
    tuned = {}
    for p in pts:
        ctx = cd.Context(p["shape"], p["dx"], interactions=[W.LJ_WCA], kT=p["kT"], model="WhiteBearII",
                         precision=precision)
        t_def, t_tun = ctx.autotune(reps=3)  # tuned once per grid size, cached for the others
        tuned.setdefault(p["shape"][0], (t_def, t_tun))
        ctxs.append(ctx)
        mus.append(ctx.bulk_mu(p["rho_bulk"]))
    t_tune = time.perf_counter() - t0
    fire = dict(force_tolerance=1e-30)  # every run takes all its steps (or stops at FIRE's uphill limit)
    rho0s = [p["rho0"] for p in pts]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq = []
    for i, (ctx, r0, mu) in enumerate(zip(ctxs, rho0s, mus)):
        if i % 16 == 0:
            progress(f"batched: sequential run {i}/{len(ctxs)}")
        try:
            _, res, _ = ctx.minimize_grand_potential(r0, mu, fire=fire, max_steps=steps, record_energies=False)
            seq.append(res["iterations"])
        except cd.CdftError:
            seq.append(-1)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = cd.minimize_grand_potential_batch(ctxs, rho0s, mus, fire=fire, max_steps=steps)
    torch.cuda.synchronize()
    t_bat = time.perf_counter() - t0
    done = sum(r["iterations"] for _, r, st in out if st == 0)
    return {
        "state_points": n,
        "grids": [p["shape"][0] for p in pts],
        "fire_steps_each": steps,
        "precision": precision,
        "uphill_stops": sum(1 for _, _, st in out if st != 0),
        "sequential_s": t_seq,
        "batched_s": t_bat,
        "batched_fire_steps_per_s": done / t_bat,
        "speedup_vs_sequential": t_seq / t_bat,
        "same_iterations": [r["iterations"] if st == 0 else -1 for _, r, st in out] == seq,
        "setup_and_autotune_s": t_tune,
        "eval_us_default_tuned": {str(k): [round(v[0], 1), round(v[1], 1)] for k, v in sorted(tuned.items())},
    }


def adjoint_ddft(cd, W, shape=(128, 128, 128), K=400, P=2, dt=1e-4, n_wells=16, fd_all=True):
    """The DDFT adjoint workload (FP64): hard-sphere (White-Bear II) + LJ fluid at kT = 0.9 on `shape`, dx = 0.1,
    n_wells seeded Gaussian wells, loss = sum (rho_K - target)^2 dV after K fixed-dt steps (P fixed-point
    iterations each), target from the same run with perturbed amplitudes; the gradient by the checkpointed
    adjoint, checked against central differences (h = 1e-5) of the forward map for the amplitude, width and
    z-centre of well 0 (of every well with fd_all)."""
    import numpy as np
    import torch

    dx = 0.1
    ctx = cd.Context(shape, dx, interactions=[W.LJ_WCA], kT=0.9, model="WhiteBearII", precision="fp64")
    rho0, prm, tprm = W.adjoint_problem(shape, dx, n_wells=n_wells, rho_bulk=0.3)
    mu = ctx.bulk_mu(0.3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    target, _ = ctx.ddft_fixed_run(rho0, mu, K, dt, P, vext=ctx.wells_potential(tprm))
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    grad, res = ctx.ddft_adjoint_wells(rho0, target, prm, mu, K, dt, P)
    t_adj = time.perf_counter() - t0

    def loss(p):
        rho, _ = ctx.ddft_fixed_run(rho0, mu, K, dt, P, vext=ctx.wells_potential(p))
        return float(np.sum((rho - target) ** 2) * dx ** 3)

    h = 1e-5
    checks = [(j, k) for j in (range(n_wells) if fd_all else [0]) for k in (0, 1, 4)]
    errs = []
    for i, (j, k) in enumerate(checks):
        if i % 12 == 0:
            progress(f"adjoint: finite differences {i}/{len(checks)}")
        pp, pm = prm.copy(), prm.copy()
        pp[j, k] += h
        pm[j, k] -= h
        fd = (loss(pp) - loss(pm)) / (2 * h)
        errs.append(abs(grad[j, k] - fd) / max(abs(fd), 1e-300))
    return {
        "grid": list(shape),
        "steps": K,
        "fp_iterations": P,
        "dt": dt,
        "wells": n_wells,
        "precision": "fp64",
        "loss": res["loss"],
        "forward_s": t_fwd,
        "gradient_s": t_adj,
        "gradient_over_forward": t_adj / t_fwd,
        "evaluations": res["evaluations"],
        "hessian_products": res["hessian_products"],
        "checkpoints": res["checkpoints"],
        "store_gb": res["store_bytes"] / 1e9,
        "fd_checked": len(checks),
        "fd_max_rel_err": max(errs),
    }


def ddft_slab(cd, W, steps=2000, shape=(256, 128, 128), dt=1e-4, precision="mixed"):
    """The DDFT slab workload: a planar LJ liquid slab (rho_l = 0.70 over 0.4 L along x, rho_v = 0.02, 1e-3 noise,
    kT = 0.8, WB-II + LJ WCA, dx = 0.1) relaxed by the integrating-factor scheme at fixed dt (dt_max = dt) for
    `steps` steps.  Reports steps/s, evaluations/s and the relative mass change."""
    import torch

    dx = 0.1
    ctx = cd.Context(shape, dx, interactions=[W.LJ_WCA], kT=0.8, model="WhiteBearII", precision=precision)
    rho0 = W.slab(shape, dx, fraction=0.4, width=0.5, rho_l=0.70, rho_v=0.02, noise=1e-3, axis=0)
    mu = ctx.bulk_mu(0.02)
    ctx.ddft_run(rho0, mu, 2, dt=dt, dt_max=dt, record_energies=False)  # warm-up (plans, graphs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, r, en = ctx.ddft_run(rho0, mu, steps, dt=dt, dt_max=dt)
    t = time.perf_counter() - t0
    return {
        "grid": list(shape),
        "schedule": ctx.schedule,
        "precision": precision,
        "steps": steps,
        "dt": dt,
        "dt_final": r["dt"],
        "seconds": t,
        "steps_per_s": steps / t,
        "evaluations": r["evaluations"],
        "evals_per_s": r["evaluations"] / t,
        "mass_rel_change": abs(r["mass_final"] - r["mass_initial"]) / r["mass_initial"],
        "energy_first": float(en[0]),
        "energy_last": float(en[-1]),
    }


def general_sizes(cd, W, sizes, precision="mixed", evals=20):
    """Per-stage times of an evaluation on cubic grids of the given sizes (general schedule for sizes the fused
    kernels do not cover): the config-5 LJ droplet settings, autotuned first, `evals` timed evaluations."""
    import torch

    out = {}
    for n in sizes:
        shape = (n, n, n)
        ctx = cd.Context(shape, 0.1, interactions=[W.LJ_WCA], kT=0.8, model="WhiteBearII", precision=precision)
        t_def, t_tun = ctx.autotune(reps=3)
        d = torch.tensor(W.droplet(shape, 0.1, radius=1.5), device="cuda")
        f = torch.empty_like(d)
        mu = ctx.bulk_mu(0.02)
        ctx.time_total_device(d.data_ptr(), mu, 2, d_forces=f.data_ptr())
        ms = ctx.time_total_device(d.data_ptr(), mu, evals, d_forces=f.data_ptr())
        stages = ctx.profile_stages(d.data_ptr(), mu, reps=5, d_forces=f.data_ptr())
        out[str(n)] = {"schedule": ctx.schedule, "precision": precision, "us_per_eval": 1000.0 * ms / evals,
                       "autotune_us_default_tuned": [t_def, t_tun],
                       "stages_us": [[name, round(1000.0 * t, 2)] for name, t, _ in stages]}
        progress(f"general sizes: {n}^3 {1000.0 * ms / evals:.1f} us per evaluation")
    return out


def stream_copy_gbps(reps=5):
    """Device-to-device copy of 1 GiB of doubles (reads 1 GiB, writes 1 GiB), timed with HIP events on
    torch's current stream; median of `reps` copies, in GB/s of bytes moved."""
    import torch

    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    rates = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        rates.append(2.0 * 8 * n / (e0.elapsed_time(e1) * 1e-3) / 1e9)
    del src, dst
    torch.cuda.empty_cache()
    return sorted(rates)[len(rates) // 2]


def converged_profile(cd, W, rho0, shape, dx, kT, tol=1e-8, cap=60000, dt_max=0.3):
    """The headline droplet relaxed by Minimizer::grand_potential (Unbounded alias, no boundary averaging:
    a periodic box; FIRE dt_max = 0.3) until max |delta(beta Omega)/delta rho| (no dV) < tol: FP64 on the
    GPU, then mixed precision with the same rule and a step cap of 1.5x the FP64 run (its FP32 rounding
    leaves a floor).  At mu = mu_bulk(rho_v) in this 12.8 sigma box the R = 4 droplet is supercritical and
    the box fills with the liquid whose bulk chemical potential is mu: the converged profile is that uniform
    liquid, which the bulk equation of state fixes exactly (its root is found by bisection).  Reports steps,
    step rates, the residual max force of each final profile under the FP64 functional, the FP64 profile's
    distance from the bulk root, and how far the mixed profile and Omega end from the FP64 ones."""
    import numpy as np
    import torch

    out = {"rule": f"max |delta(beta Omega)/delta rho| < {tol:g} (no dV), homogeneous_boundary off", "tol": tol}
    ctx64 = cd.Context(shape, dx, interactions=[W.LJ_WCA], kT=kT, model="WhiteBearII", precision="fp64")
    mu = ctx64.bulk_mu(0.02)
    dv = dx ** 3

    def max_force(r):
        _, f = ctx64.total(r, mu)
        return float(np.max(np.abs(f)) / dv)

    res = {}
    for prec in ("fp64", "mixed"):
        ctx = ctx64 if prec == "fp64" else cd.Context(shape, dx, interactions=[W.LJ_WCA], kT=kT,
                                                       model="WhiteBearII", precision=prec)
        steps = cap if prec == "fp64" else max(2000, int(1.5 * res["fp64"][1]["iterations"]))
        fire = dict(force_tolerance=0.0, force_max_tolerance=tol, dt_max=dt_max)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r, info, _ = ctx.minimize_grand_potential(rho0, mu, fire=fire, max_steps=steps, homogeneous_boundary=False,
                                                  record_energies=False)
        t = time.perf_counter() - t0
        res[prec] = (r, info)
        out[f"{prec}_iterations"] = info["iterations"]
        out[f"{prec}_reached_tol"] = bool(info["converged"])
        out[f"{prec}_fire_steps_per_s"] = info["iterations"] / t
        out[f"{prec}_max_force_fp64_functional"] = max_force(r)
        print(f"converged profile {prec}: {info['iterations']} steps in {t:.1f} s", file=sys.stderr, flush=True)
    (rd, idd), (rm, im) = res["fp64"], res["mixed"]
    # the uniform states at this mu: the vapour root rho_v and the liquid root in (0.4, 1.0) of
    # mu_bulk(rho) = mu (mu_bulk rises on the liquid branch)
    lo, hi = 0.4, 1.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if ctx64.bulk_mu(mid) < mu else (lo, mid)
    rho_l = 0.5 * (lo + hi)
    root = rho_l if float(np.mean(rd)) > 0.3 else 0.02
    out["fp64_final_phase"] = "liquid" if root == rho_l else "vapour"
    out["bulk_root_density"] = root
    out["fp64_max_abs_dev_from_bulk_root"] = float(np.max(np.abs(rd - root)))
    out["mixed_floor_max_force"] = out["mixed_max_force_fp64_functional"]
    out["profile_max_rel_err"] = float(np.max(np.abs(rm - rd)) / np.max(np.abs(rd)))
    out["omega_rel_err"] = abs(im["grand_potential"] - idd["grand_potential"]) / abs(idd["grand_potential"])
    out["omega_fp64"] = idd["grand_potential"]
    return out


def mixture(cd, W, shape=(128, 128, 128), evals=50, autotune=True):
    """Binary LJ mixture (diameters 1.0 and 0.8; self pairs LJ(1, 1) and LJ(0.8, 0.6), cross pair
    LJ(0.9, 0.8), WCA, r_c = 2.5; kT = 1.0, WB-II, dx = 0.1) on smooth noisy densities of mean 0.3 and 0.2:
    evaluations per second with forces, FP64 (after autotuning the context; a fused-schedule context is
    only timed) and mixed precision (FP32 real-space fields), timed with HIP events on the context
    stream; the mixed result is compared with the FP64 one."""
    import numpy as np
    import torch

    dx = 0.1
    inter = [dict(W.LJ_WCA, si=0, sj=0), dict(W.LJ_WCA, si=1, sj=1, sigma=0.8, epsilon=0.6),
             dict(W.LJ_WCA, si=0, sj=1, sigma=0.9, epsilon=0.8)]
    kw = dict(diameters=(1.0, 0.8), interactions=inter, kT=1.0, model="WhiteBearII")
    rho = np.stack([W.smooth_noise_fluid(shape, dx, rho0=0.3, amp=0.02, seed=1),
                    W.smooth_noise_fluid(shape, dx, rho0=0.2, amp=0.02, seed=2)])
    d_rho = torch.tensor(rho, device="cuda:0").contiguous()
    mu = (0.1, 0.2)
    out = {"grid": list(shape), "species": 2, "interactions": 3}
    forces = {}
    for prec in ("fp64", "mixed"):
        try:
            ctx = cd.Context(shape, dx, precision=prec, **kw)
        except cd.CdftError:  # mixed precision needs a fused-schedule grid
            continue
        tune = ctx.autotune(reps=5) if autotune else (None, None)
        d_f = torch.empty_like(d_rho)
        res = ctx.total_species_device(d_rho.data_ptr(), mu, d_forces=d_f.data_ptr())  # warm-up
        ms = ctx.time_total_species_device(d_rho.data_ptr(), mu, evals, d_forces=d_f.data_ptr())
        forces[prec] = d_f.cpu().numpy()
        key = "" if prec == "fp64" else "mixed_"
        out.update({f"{key}precision": prec, f"{key}schedule": ctx.schedule, f"{key}evals": evals,
                    f"{key}us_per_eval": 1000.0 * ms / evals, f"{key}evals_per_s": 1000.0 * evals / ms,
                    f"{key}grand_potential": res["grand_potential"]})
        if prec == "fp64":
            out["autotune_us_default_tuned"] = list(tune)
    if "mixed" in forces:
        out["mixed_omega_rel_err_vs_fp64"] = abs(out["mixed_grand_potential"] - out["grand_potential"]) / abs(
            out["grand_potential"])
        out["mixed_force_rel_err_vs_fp64"] = float(np.max(np.abs(forces["mixed"] - forces["fp64"])) /
                                                   np.max(np.abs(forces["fp64"])))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the same-GPU FP64 comparison run")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-sequential FIRE measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the DDFT adjoint measurement")
    ap.add_argument("--no-mixture", action="store_true", help="skip the binary-mixture evaluation measurement")
    ap.add_argument("--mixture-untuned", action="store_true",
                    help="time the mixture on the default launch shapes (no autotuning trials in a profile)")
    ap.add_argument("--no-converged", action="store_true",
                    help="skip relaxing the droplet to max |dOmega/drho| < 1e-8 (mixed vs FP64)")
    ap.add_argument("--ddft-steps", type=int, default=2000,
                    help="steps of the 256x128x128 DDFT slab run (0 skips it)")
    ap.add_argument("--adjoint-fd-sample", action="store_true",
                    help="check the adjoint gradient against differences for well 0 only (default: all 48)")
    ap.add_argument("--reps", type=int, default=5, help="timed repetitions of K evaluations (median reported)")
    ap.add_argument("--batch-points", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--general-sizes", default="",
                    help="comma-separated cube sizes whose per-stage evaluation times are added (e.g. 48,80,96,160)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only (there is no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import classicaldft_amd as cd
    from classicaldft_amd import workloads as W

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(3)

    if args.copy_calibration:
        n = 1 << 27  # 1 GiB of doubles
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 3}), flush=True)
        return

    shape, dx, kT = (128, 128, 128), 0.1, 0.8
    rho = W.droplet(shape, dx, radius=4.0, width=0.5, rho_l=0.70, rho_v=0.02, noise=1e-3, seed=2024)
    dev = torch.device("cuda:0")
    d_rho = torch.from_numpy(rho.copy()).to(dev)
    d_f = torch.empty_like(d_rho)

    def run(precision, steps, warmup, reps):
        """`reps` timed blocks of `steps` back-to-back evaluations after `warmup` untimed ones; each block is
        bracketed by device synchronisation and timed with HIP events on the context stream.  Returns the
        per-block milliseconds (the median is reported) and the wall time of all blocks."""
        ctx = cd.Context(shape, dx, interactions=[W.LJ_WCA], kT=kT, model="WhiteBearII", precision=precision)
        mu = ctx.bulk_mu(0.02)
        ctx.time_total_device(d_rho.data_ptr(), mu, max(warmup, 1), d_forces=d_f.data_ptr())
        blocks = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            torch.cuda.synchronize()
            blocks.append(ctx.time_total_device(d_rho.data_ptr(), mu, steps, d_forces=d_f.data_ptr()))
            torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        res = ctx.total_device(d_rho.data_ptr(), mu, d_forces=d_f.data_ptr())
        forces = d_f.cpu().numpy().reshape(shape)
        return ctx, mu, blocks, wall, res, forces

    def median(v):
        v = sorted(v)
        return v[len(v) // 2] if len(v) % 2 else 0.5 * (v[len(v) // 2 - 1] + v[len(v) // 2])

    reps = max(1, args.reps)
    progress(f"{args.precision} evaluation: {reps} x {args.steps} evaluations")
    ctx, mu, blocks, wall, res, forces = run(args.precision, args.steps, args.warmup, reps)
    ms = median(blocks)
    ms_per_step = ms / args.steps
    value = 1000.0 / ms_per_step

    stages = ctx.profile_stages(d_rho.data_ptr(), mu, reps=args.profile_reps, d_forces=d_f.data_ptr())
    timed = [s for s in stages if s[0] not in ("k_finalize",)]
    dom = max(timed, key=lambda s: s[1])
    achieved = dom[2] / (dom[1] * 1e-3) / 1e9

    s_f = 4 if args.precision == "mixed" else 8
    B = eval_bytes(shape, s_f, 8)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "dtype_detail": ("f32 fields and vector/tensor transforms; f64 scalar-channel transforms, pointwise "
                         "terms, reductions and state") if args.precision == "mixed" else "f64 throughout",
        "data": "synthetic",
        "config": {
            "workload": "LJ droplet F+dF/drho eval: 128^3, dx=0.1, WB-II FMT d=1, LJ WCA rc=2.5, kT=0.8, "
                        "R=4 tanh(0.5) rho_l=0.70 rho_v=0.02, 1e-3 noise, mu=mu_bulk(rho_v)",
            "grid": list(shape),
            "precision": args.precision,
            "field_bytes": s_f,
            "state_bytes": 8,
        },
        "wall_ms_per_step": wall * 1000.0 / (args.steps * reps),
        "timing": {"repetitions": reps, "evals_per_rep": args.steps,
                   "evals_per_s_each_rep": [1000.0 * args.steps / b for b in blocks],
                   "spread_rel": (max(blocks) - min(blocks)) / ms, "statistic": "median"},
        "eval_algorithmic_bytes": B,
        "eval_achieved_GBps": B * value / 1e9,
        "eval_roofline_frac": B * value / 1e9 / PEAK_HBM_GBPS,
        "roofline": {
            "kernel": dom[0],
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS,
            "traffic": measured_traffic(dom[0], args.precision),
            "kernel_ms": dom[1],
            "kernel_algorithmic_bytes": dom[2],
        },
        "stages": [{"kernel": n, "ms": round(t, 5), "bytes": b} for n, t, b in stages],
        "omega": res["grand_potential"],
    }

    if not args.no_fp64 and args.precision == "mixed":
        progress(f"fp64 evaluation: {reps} x {args.steps} evaluations")
        ctx64, mu64, blocks64, _, res64, f64 = run("fp64", args.steps, args.warmup, reps)
        out["fp64_value"] = 1000.0 * args.steps / median(blocks64)
        out["fp64_evals_per_s_each_rep"] = [1000.0 * args.steps / b for b in blocks64]
        out["speedup_vs_fp64"] = value / out["fp64_value"]
        out["omega_fp64_gpu"] = res64["grand_potential"]
        out["omega_rel_err_vs_fp64_gpu"] = abs(res["grand_potential"] - res64["grand_potential"]) / abs(
            res64["grand_potential"])
        out["force_rel_err_vs_fp64_gpu"] = float(np.max(np.abs(forces - f64)) / np.max(np.abs(f64)))
        del ctx64

    out["stream_copy_GBps"] = stream_copy_gbps()

    if not args.no_converged:
        progress("converged profile")
        out["converged_profile"] = converged_profile(cd, W, rho, shape, dx, kT)

    if not args.no_batch:
        progress(f"batched: {args.batch_points} state points x {args.batch_steps} FIRE steps")
        out["batched"] = batched_fire(cd, W, args.precision, n=args.batch_points, steps=args.batch_steps)

    if not args.no_adjoint:
        progress("adjoint")
        out["adjoint"] = adjoint_ddft(cd, W, fd_all=not args.adjoint_fd_sample)

    if args.ddft_steps > 0:
        progress("ddft slab")
        out["ddft_slab"] = ddft_slab(cd, W, steps=args.ddft_steps)

    if args.general_sizes:
        out["general_sizes"] = general_sizes(cd, W, [int(v) for v in args.general_sizes.split(",") if v])

    if not args.no_mixture:
        progress("mixture")
        out["mixture"] = mixture(cd, W, autotune=not args.mixture_untuned)

    if not args.no_cpu_baseline:
        progress("cpu baseline")
        cb, parts, fr = cpu_baseline(rho, mu, shape, dx)
        out["cpu_baseline"] = cb
        out["omega_oracle"] = parts[5]
        out["omega_rel_err_vs_oracle"] = abs(res["grand_potential"] - parts[5]) / abs(parts[5])
        out["force_rel_err_vs_oracle"] = float(np.max(np.abs(forces - fr)) / np.max(np.abs(fr)))
        out["gpu_over_cpu"] = value / cb["value"]
    else:
        out["cpu_baseline"] = None

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
