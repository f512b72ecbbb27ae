#!/usr/bin/env python3
"""Benchmark: SIP-steps/s of the super-droplet time step on one MI355X.

Default workload (--case icmw75; synthetic inputs from lacmopy_amd.setup_case): the 2-D ICMW
kinematic stratocumulus, 75x75 cells of 20 m, Poisson(128) SIPs per cell, bimodal AS aerosol,
Long E_col table, dt_adv = 1 s, 10 condensation substeps, 2 collision steps per advection step,
relaxation on.  --case stress150: 150x150 cells of 10 m, Poisson(256), Hall E_col, dt 0.5 s,
20 substeps, 16 seeds.
  * batch: 64 (16) independent seeds in one engine (the headline `value`)
  * plain: one run (reported next to it; ratio = batch / plain)
A "step" is one advection step of every SIP of every run.
SIP-steps/s = sum over runs of SIPs x timed steps / wall seconds, with all inputs resident in HBM;
the timed region is bracketed by torch.cuda.synchronize().

Also measured: the dominant kernel's duration with HIP events (roofline vs HBM 8 TB/s, algorithmic
bytes per SIP), and the CPU oracle on a bounded sample of the plain 75x75 case (cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_SUBSTEP = 104.0    # algorithmic bytes per SIP per particle substep (60 read + 44 write)
B_COLLISION = 64.0   # per SIP per collision step (40 read + 24 write)
B_REINDEX = 140.0    # per SIP per reindex

from lacmopy_amd.setup_case import WORKLOADS as CASES  # noqa: E402


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_icmw75_64seeds.md")
VALU_PROFILE = os.path.join(ROOT, "profiles", "valu_fp64_icmw75_64seeds.md")
VALU_ISSUE_PEAK = 256 * 2.4e9  # 64-lane VALU wave instructions per second: one per CU per cycle at 2.4 GHz
FP64_PEAK_TFLOPS = 78.6


def profile_value(path, tag, kernel):
    """A `TAG kernel value` line of a committed rocprofv3 summary, or None."""
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == tag and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the 64-seed workload, from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE scaled by a 1 GiB copy); None if absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "TRAFFIC" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_engine(n_runs, spc, seed0, case="icmw75", kernel=None):
    """kernel: 'Long_Bott' / 'Hall_Bott' table, or 'const' (Hydro_E_const, E = 0.5); default: the case's."""
    import torch  # noqa: F401
    from lacmopy_amd import load_kernel_data_Ecol
    from lacmopy_amd.engine import Engine
    from lacmopy_amd.setup_case import init_profiles, workload_case
    c = CASES[case]
    kernel = kernel or c["kernel"]
    if kernel == "const":
        E, lnR0, lnf = 0.5, None, None
    else:
        E, lnR0, lnf, nb = load_kernel_data_Ecol(kernel)
    cases = []
    for r in range(n_runs):
        grid, pos, vel, cells, m_w, m_s, xi, act, draws = workload_case(case, seed0 + 2 * r, spc=spc,
                                                                        return_draws=True)
        cases.append((grid, (pos, vel, cells, m_w, m_s, xi, act), draws))
    grid0 = cases[0][0]
    eng = Engine(c["nx"], c["nz"], [len(k[1][5]) for k in cases], grid0.ranges, grid0.steps, grid0.volume_cell)
    eng.set_params(solute="AS", no_iter_impl_mass=3, scale_dt_cond=c["sdc"], no_col_per_adv=2, act_collisions=True,
                   act_relaxation=True, dt=c["dt"], dt_col=c["dt"] / 2, g_set=9.80665, seed=seed0)
    eng.set_collision_kernel(E, lnR0, lnf)
    eng.set_flow(grid0.velocity, grid0.mass_flux_air_dry)
    for r, (g, _, _) in enumerate(cases):
        eng.set_relaxation(r, *init_profiles(g))
        eng.set_grid(r, g.scalar_fields(), g.mat_prop())
    eng.init_cells()
    eng.load_particles([k[1] for k in cases])
    n_sips = sum(len(k[1][5]) for k in cases)
    return eng, n_sips, cases


def time_steps(eng, steps, warmup, graph=False):
    """Wall time of `steps` advection steps after `warmup` untimed ones (graph: replays of a captured
    two-step HIP graph, captured during the warm-up)."""
    import torch
    run = eng.adv_steps_graph if graph else eng.adv_steps
    run(warmup)
    if graph and warmup < 2:
        run(2)  # capture outside the timed region
    torch.cuda.synchronize()
    eng.check()
    t0 = time.perf_counter()
    run(steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    eng.check()
    return t1 - t0


def kernel_times(eng, steps):
    """HIP-event durations of the step's kernels over `steps` profiled advection steps.

    The events are recorded by liblacmo_amd on the stream the kernels run on.
    Returns {kind: (mean ms per launch, launches)}.
    """
    import torch
    eng.prof_enable(max_events=64 * steps + 64)
    eng.adv_steps(steps)
    torch.cuda.synchronize()
    res = eng.prof_collect()
    eng.prof_disable()
    eng.check()
    return {k: (ms / max(n, 1), n) for k, (ms, n) in res.items()}


# adjoint window per case: (window steps, checkpoint interval, spin-up steps without collisions, event capacity).
# icmw75: 600 s after a 2 h spin-up, checkpoints every 60 steps; stress150: the last 300 s (600 steps of 0.5 s)
# of 1800 s, checkpoints every 10 steps (the tape of 10 steps of 5.8M SIPs x 20 substeps is 74 GB).
ADJ_DEFAULTS = {"icmw75": (600, 60, 7200, 1 << 20), "stress150": (600, 10, 3000, 1 << 23)}


def adjoint_cost(seed0, spc, window, every, spinup, check=0, case="icmw75"):
    """Adjoint of a `window`-step window of one run of `case` after `spinup` steps without collisions:
    wall time of the plain forward window, of the recording forward (events + checkpoints every `every`
    steps) and of one gradient (taped re-runs + reverse sweep), for the cloud-droplet-number objective,
    and the gradient with respect to the six aerosol-mode parameters through the SIP initialisation map
    (the window-start multiplicities and dry masses are still the map's: no collision before the window).
    check > 0: also compare the gradient with central finite differences on `check` random droplets
    (multiplicity, h = 1e-3 relative) and on the six aerosol-mode parameters (fourth order, h = 1e-3)."""
    import torch
    from lacmopy_amd.adjoint import AdjointWindow
    from lacmopy_amd.setup_case import aerosol_sips
    eng, n1, cases = build_engine(1, spc, seed0, case=case)
    c = CASES[case]
    e = eng.e
    e.act_collisions = 0
    t = time.perf_counter()
    eng.adv_steps(spinup)
    torch.cuda.synchronize()
    log(f"adjoint: spin-up {spinup} steps in {time.perf_counter() - t:.1f} s")
    e.act_collisions = 1
    eng.check()
    start = eng.snapshot()
    eng.adv_steps(2)  # warm-up of the collision path
    eng.restore(start)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.adv_steps(window)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    eng.check()
    eng.restore(start)
    w = AdjointWindow(eng, window, checkpoint_every=every, ev_cap=ADJ_DEFAULTS[case][3])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = w.run(record=True)
    torch.cuda.synchronize()
    t_rec = time.perf_counter() - t0
    w.gradient(w_nc=1.0, w_lwp=0.0)  # allocates the tape; the timed call below reuses it
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lam_xi, lam_ms = w.gradient(w_nc=1.0, w_lwp=0.0)
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    # aerosol-mode parameters (N, r_med, sigma of both modes) -> window-start xi, m_s -> J_nc, by the chain rule
    eng.restore(w.start)
    slots = w.slot_map()
    cur = int(eng.e.cur)
    theta = torch.tensor([60.0E6, 0.02, 1.4, 40.0E6, 0.075, 1.6], dtype=torch.float64, device="cuda",
                         requires_grad=True)
    xi_m, ms_m = aerosol_sips(theta, cases[0][2])
    # the map restated in torch rounds differently from make_case's numpy in the last bits
    map_dev = max(float(((xi_m.detach() - eng.buf[f"xi{cur}"][slots]).abs() / xi_m.detach().abs()).max()),
                  float(((ms_m.detach() - eng.buf[f"m_s{cur}"][slots]).abs() / ms_m.detach().abs()).max()))
    (g_theta,) = torch.autograd.grad((xi_m * lam_xi).sum() + (ms_m * lam_ms).sum(), theta)
    res = {"workload": f"one {case} run ({c['nx']}x{c['nz']} cells, {n1} SIPs), {window}-step window (dt "
                       f"{c['dt']:g} s, {2 * c['sdc']} substeps, 2 collision steps) after a {spinup}-step spin-up "
                       f"without collisions; fixed collision events; checkpoints every {every} steps",
           "sips": n1, "window_steps": window, "checkpoint_every": every, "events": w.n_events,
           "J_nc": float(J[0][0]), "J_lwp": float(J[1][0]),
           "forward_s": t_fwd, "record_s": t_rec, "gradient_s": t_grad,
           "cost_ratio": (t_rec + t_grad) / t_fwd, "gradient_over_forward": t_grad / t_fwd,
           "dJnc_dmodes": g_theta.tolist(), "mode_map_rel_dev_from_window_start": map_dev,
           "tape_gb": sum(t.numel() * t.element_size() for t in w.tape.values()) / 1e9}
    log(f"adjoint: {res}")
    if check:
        res["fd_check"] = adjoint_fd_check(w, eng, cases, lam_xi, check)
        log(f"adjoint fd check: {res['fd_check']}")
    del w, eng
    torch.cuda.empty_cache()
    return res


def adjoint_fd_check(w, eng, cases, lam_xi, n_check):
    import torch
    from lacmopy_amd.setup_case import aerosol_sips
    eps = np.finfo(np.float64).eps
    slots = w.slot_map()
    cur = int(eng.e.cur)
    xi0 = eng.buf[f"xi{cur}"][slots].clone()
    ms0 = eng.buf[f"m_s{cur}"][slots].clone()
    drop = ((eng.buf[f"cell{int(eng.e.ccur)}"][slots] >= 0) & (eng.buf[f"m_w{cur}"][slots] > 4.19E3)).cpu().numpy()
    J0 = float(np.sum(w.run(xi0, ms0)[0]))
    out = {"sips": [], "modes": []}
    rng = np.random.default_rng(11)
    for gid in rng.choice(np.flatnonzero(drop), n_check, replace=False):
        x, h = float(xi0[gid]), 1.0E-3
        Jpm = []
        for sgn in (1.0, -1.0):
            p = xi0.clone()
            p[gid] = x * (1.0 + sgn * h)
            Jpm.append(float(np.sum(w.run(p, ms0)[0])))
        fd = (Jpm[0] - Jpm[1]) / (2 * h * x)
        a = float(lam_xi[gid])
        out["sips"].append({"gid": int(gid), "adjoint": a, "fd": fd, "rel": abs(a - fd) / max(abs(fd), 1e-300),
                            "floor_rel": 256 * eps * abs(J0) / (2 * h * x) / max(abs(fd), 1e-300)})
    # mode parameters through the initialisation map (fourth-order quotient, h = 1e-3)
    draws = cases[0][2] if len(cases[0]) > 2 else None
    if draws is not None:
        theta0 = np.array([60.0E6, 0.02, 1.4, 40.0E6, 0.075, 1.6])
        theta = torch.tensor(theta0, dtype=torch.float64, device="cuda", requires_grad=True)
        xi_m, ms_m = aerosol_sips(theta, draws)
        w.run(xi_m.detach(), ms_m.detach())
        g_xi, g_ms = w.gradient(w_nc=1.0, w_lwp=0.0)
        (grad,) = torch.autograd.grad((xi_m * g_xi).sum() + (ms_m * g_ms).sum(), theta)
        for k in range(6):
            def J_at(f):
                th = theta0.copy()
                th[k] *= 1.0 + f
                xx, mm = aerosol_sips(torch.tensor(th, dtype=torch.float64, device="cuda"), draws)
                return float(np.sum(w.run(xx, mm)[0]))
            h = 1.0E-3
            fd = (8.0 * (J_at(h) - J_at(-h)) - (J_at(2 * h) - J_at(-2 * h))) / (12 * h * theta0[k])
            a = float(grad[k])
            out["modes"].append({"k": k, "adjoint": a, "fd": fd, "rel": abs(a - fd) / max(abs(fd), 1e-300)})
    return out


def _cpu_run(args):
    """One CPU worker of the all-cores baseline: its own seed of the plain case, SIP-steps/s of its steps."""
    seed0, spc, max_steps, max_s = args
    return cpu_baseline(seed0, spc, max_steps=max_steps, max_s=max_s)["value"]


def cpu_baseline_all_cores(seed0, spc, workers):
    """The reference's own pattern for many runs (one single-threaded process per run, run_lacmo.sh:7-16):
    `workers` processes, each the oracle on its own seed of the plain 75x75 case, started together; the sum
    of their SIP-steps/s."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        rates = pool.map(_cpu_run, [(seed0 + 2 * k, spc, 2, 8.0) for k in range(workers)])
    return float(np.sum(rates))


def cpu_baseline(seed0, spc, max_steps=3, max_s=10.0):
    """Oracle (C restatement, 1 thread) on the plain 75x75 case: bounded sample of adv steps."""
    from lacmopy_amd import load_kernel_data_Ecol
    from lacmopy_amd.setup_case import init_profiles, make_case
    from oracle import oracle as orc
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=75, nz=75, spc=spc, seed=seed0)
    N = len(xi)
    cfg = dict(solute_type="AS", dt_adv=1.0, dt_col=0.5, scale_dt_cond=5, no_col_per_adv=2, no_iter_impl_mass=3,
               g_set=9.80665, act_collisions=True, act_relaxation=True, no_kernel_bins=nb, R_kernel_low_log=lnR0,
               bin_factor_R_log=lnf)
    p = orc.adv_params(grid, N, cfg, seed=seed0)
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    prof = init_profiles(grid)
    rel = np.zeros((2, N))
    T_p = np.ones(N)
    wr = np.zeros(1)
    nco = np.zeros(2, dtype=np.int64)
    q = orc.adv_params(grid, N, cfg)
    orc.lib().or_update_pos_and_cells(orc.C.byref(q), 0.0, orc.P(pos), orc.P(vel), orc.P(cells), orc.P(rel),
                                      orc.P(m_w), orc.P(xi), orc.P(act.view(np.uint8)), orc.P(wr))
    steps = 0
    t0 = time.perf_counter()
    while True:
        orc.integrate_adv_steps(p, 1, pos, vel, cells, rel, m_w, m_s, xi, act, T_p, wr, nco, scal, mat,
                                grid.velocity, grid.mass_flux_air_dry, E, prof[0], prof[1], prof[2])
        steps += 1
        if time.perf_counter() - t0 > max_s or steps >= max_steps:
            break
    dt = time.perf_counter() - t0
    return {"value": N * steps / dt, "unit": "SIP-steps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} advection steps of the plain 75x75 case ({N} SIPs), oracle C restatement, 1 thread"}


def copy_probe(mib=1024, reps=3):
    """A plain device-to-device copy of known size: the calibration kernel for the HBM byte counters."""
    import torch
    n = mib * 1024 * 1024 // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "reps": reps}),
          flush=True)


def init_cost(n_seeds, cpu=True):
    """GPU initial states (SinSIP + saturation adjustment, lacmopy_amd.initialize) of n_seeds seeds of the
    reference's configuration on 75x75 cells, wall time after a warm-up; the oracle's time for the same
    initialisation measured on 4 of the 75 columns (all 75 levels) and scaled to one full seed."""
    import torch
    from lacmopy_amd.grid import Grid
    from lacmopy_amd.initialize import init_parameters, initialize_batch, reference_config
    cfg = reference_config()
    initialize_batch(reference_config(nx=4, nz=6), [1])
    torch.cuda.synchronize()
    t = time.perf_counter()
    states = initialize_batch(cfg, list(range(3711, 3711 + 2 * n_seeds, 2)))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    out = {"seeds": n_seeds, "cells": "75x75", "wall_s": wall,
           "sips_per_seed": float(np.mean([len(s.xi) for s in states])),
           "adjustment_iterations_per_seed": float(np.mean([s.iterations.sum() for s in states]))}
    del states
    if cpu:
        import ctypes
        from oracle import oracle as orc
        c4 = reference_config(nx=4)
        g = Grid(c4["grid_ranges"], c4["grid_steps"], c4["dy"])
        q = init_parameters(c4, g)
        q.n_runs = 1
        oq = orc.OrInitPar()
        ctypes.memmove(ctypes.addressof(oq), ctypes.addressof(q), ctypes.sizeof(q))
        t = time.perf_counter()
        orc.init_SinSIP(oq, 3711, g.corners[1][0], c4["p_0"], g.corners[0][:, 0])
        s4 = time.perf_counter() - t
        out["cpu_oracle_s_per_seed"] = s4 * 75 / 4
        out["cpu_sample"] = "oracle (1 thread) on 4 columns x 75 levels of one seed, scaled by 75/4"
        out["gpu_over_cpu"] = out["cpu_oracle_s_per_seed"] * n_seeds / wall
    log(f"init: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--case", default="icmw75", choices=sorted(CASES))
    ap.add_argument("--runs", type=int, default=None, help="seeds in the batch (default: the case's)")
    ap.add_argument("--spc", type=float, default=None, help="mean SIPs per cell (default: the case's)")
    ap.add_argument("--seed", type=int, default=1001)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=min(16, len(os.sched_getaffinity(0))),
                    help="processes of the all-cores CPU baseline (1: single core only)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-kernels", action="store_true")
    ap.add_argument("--kernel", default=None, choices=["Long_Bott", "Hall_Bott", "const"])
    ap.add_argument("--graph", action="store_true",
                    help="replay the step captured in a two-step HIP graph instead of launching kernel by kernel")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adj-window", type=int, default=None,
                    help="advection steps of the adjoint window (default 600 for both cases)")
    ap.add_argument("--adj-every", type=int, default=None,
                    help="checkpoint interval of the adjoint window (default 60; stress150: 10)")
    ap.add_argument("--adj-spinup", type=int, default=None,
                    help="spin-up steps before the adjoint window (default 7200; stress150: 3000)")
    ap.add_argument("--adj-check", type=int, default=0,
                    help="also check the adjoint against finite differences on this many droplets and the "
                         "six aerosol-mode parameters")
    ap.add_argument("--no-init", action="store_true", help="skip timing the GPU initial states")
    ap.add_argument("--init-seeds", type=int, default=64, help="seeds of the timed GPU initialisation")
    ap.add_argument("--init-only", action="store_true", help="only time the GPU initialisation (no CPU leg)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (calibration of the HBM byte counters)")
    args = ap.parse_args()
    case = CASES[args.case]
    args.runs = args.runs or case["runs"]
    args.spc = args.spc or case["spc"]
    args.kernel = args.kernel or case["kernel"]
    n_sub = 2 * case["sdc"]
    adj_w, adj_k, adj_s, _ = ADJ_DEFAULTS[args.case]
    args.adj_window = args.adj_window or adj_w
    args.adj_every = args.adj_every or adj_k
    args.adj_spinup = args.adj_spinup if args.adj_spinup is not None else adj_s
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): independent seeds shard "
              "trivially across GPUs and no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    if args.copy_probe:
        copy_probe()
        return
    if args.init_only:
        print(json.dumps({"init": init_cost(args.init_seeds, cpu=False)}), flush=True)
        return

    log(f"building batch of {args.runs} runs ...")
    t = time.perf_counter()
    eng, n_sips, _ = build_engine(args.runs, args.spc, args.seed, case=args.case, kernel=args.kernel)
    log(f"  {n_sips} SIPs loaded in {time.perf_counter() - t:.1f} s")
    tune = None
    if not args.no_tune:
        t = time.perf_counter()
        eng.adv_steps(1)  # tune on an evolved state
        tr = eng.autotune()
        ch = tr["chosen"]
        tune = {"chosen": ch,
                "particle_ms": {f"{b}x{sp}/w{wp}": round(ms, 4) for (b, sp, wp), ms in tr["particles"].items()},
                "dml_work_ms": {f"{w}/{c}": round(ms, 4) for (w, c), ms in tr["dml"].items()},
                "collision_ms": {str(c): round(ms, 4) for c, ms in tr["collision"].items()},
                "collision_work_ms": {f"{w}/{c}": round(ms, 4) for (w, c), ms in tr["collision_work"].items()},
                "seconds": time.perf_counter() - t}
        log(f"autotune: {tune}")
    secs = time_steps(eng, args.steps, args.warmup, graph=args.graph)
    batch_rate = n_sips * args.steps / secs
    log(f"batch: {secs / args.steps * 1e3:.2f} ms/step, {batch_rate:.4g} SIP-steps/s")
    ncols = eng.buf["no_cols"].sum(0).cpu().tolist()
    log(f"collision events so far (single, multiple): {ncols}")
    ktimes = {} if args.no_kernels else kernel_times(eng, max(1, min(args.steps, 3)))
    log(f"kernel times (ms): {ktimes}")
    n_batch_sips = n_sips
    del eng
    torch.cuda.empty_cache()

    plain = None
    if not args.no_plain:
        eng1, n1, _ = build_engine(1, args.spc, args.seed, case=args.case, kernel=args.kernel)
        shape1 = None
        if not args.no_tune:
            eng1.adv_steps(1)
            shape1 = eng1.autotune()["chosen"]
        s1 = time_steps(eng1, max(args.steps, 10), args.warmup, graph=args.graph)
        plain = {"sip_steps_per_s": n1 * max(args.steps, 10) / s1, "ms_per_step": s1 / max(args.steps, 10) * 1e3,
                 "sips": n1, "launch_shape": shape1}
        log(f"plain: {plain}")
        del eng1

    roof = None
    if ktimes:
        # dominant per-SIP kernel: the particle substep kernel (one thread per SIP)
        ms = ktimes["particles"][0]
        achieved = n_batch_sips * B_SUBSTEP / (ms * 1e-3) / 1e9
        default_workload = (args.case == "icmw75" and args.runs == 64 and args.spc == 128.0 and args.seed == 1001
                            and args.kernel == "Long_Bott")
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic("k_particles") if default_workload else None,
                "kernel": "k_particles", "bytes_per_launch": n_batch_sips * B_SUBSTEP, "launch_ms": ms}
        valu = profile_value(VALU_PROFILE, "VALU", "k_particles") if default_workload else None
        flops = profile_value(VALU_PROFILE, "FP64FLOPS", "k_particles") if default_workload else None
        if valu and flops:
            # the co-bound of this kernel: VALU instruction issue (counts per launch from the committed --pmc
            # summary of the same workload, over the launch time measured here)
            roof["valu"] = {"achieved": valu / (ms * 1e-3), "peak": VALU_ISSUE_PEAK, "unit": "wave-instructions/s",
                            "frac": valu / (ms * 1e-3) / VALU_ISSUE_PEAK,
                            "fp64_tflops": flops / (ms * 1e-3) / 1e12,
                            "fp64_frac": flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}

    adj = None
    if not args.no_adjoint:
        adj = adjoint_cost(args.seed, args.spc, args.adj_window, args.adj_every, args.adj_spinup, args.adj_check,
                           case=args.case)

    init = None if (args.no_init or args.case != "icmw75") else init_cost(args.init_seeds, cpu=not args.no_cpu)

    # the CPU baseline is the oracle on the plain 75x75 case (the metric's workload)
    cpu = None if (args.no_cpu or args.case != "icmw75") else cpu_baseline(args.seed, args.spc)
    if cpu and args.cpu_workers > 1:
        t = time.perf_counter()
        single = dict(cpu)
        agg = cpu_baseline_all_cores(args.seed, args.spc, args.cpu_workers)
        cpu = {"value": agg, "unit": "SIP-steps/s", "cores": args.cpu_workers, "kind": "port",
               "sample": f"{args.cpu_workers} processes started together, each the oracle C restatement (1 thread) "
                         f"on its own seed of the plain 75x75 case for 2 advection steps (the reference's one "
                         f"single-threaded process per run); sum of their SIP-steps/s",
               "single_core": single, "seconds": time.perf_counter() - t}
    if cpu:
        log(f"cpu baseline: {cpu}")
    res = {
        "metric": "SIP-steps/s, 2D ICMW 75x75: 64-seed batch vs plain GPU; adjoint/forward cost"
        if args.case == "icmw75"
        else f"SIP-steps/s, {args.case}",
        "value": batch_rate,
        "unit": "SIP-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ICMW 2D kinematic Sc ({args.case}), {case['nx']}x{case['nz']} cells of "
                               f"{case['dx']:g} m, Poisson({args.spc:g}) SIPs/cell, {args.runs} seeds batched, "
                               f"{args.kernel} E_col, dt={case['dt']:g} s, {n_sub} substeps, 2 collision steps",
                   "runs": args.runs, "sips_total": n_batch_sips, "cells": case["nx"] * case["nz"],
                   "algorithmic_bytes_per_sip_step": n_sub * B_SUBSTEP + 2 * B_COLLISION + B_REINDEX,
                   "launch": "two-step HIP graph replays" if args.graph else "eager",
                   "autotune": tune,
                   "plain_single_run": plain,
                   "batch_over_plain": (batch_rate / plain["sip_steps_per_s"]) if plain else None,
                   "kernel_ms": ktimes, "collision_events": ncols,
                   "adjoint": adj,
                   "initial_states": init},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
