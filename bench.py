#!/usr/bin/env python3
"""Benchmark: SIP-steps/s of the super-droplet time step on one MI355X.

Default workload (--case icmw75; synthetic inputs from lacmopy_amd.setup_case): the 2-D ICMW
kinematic stratocumulus, 75x75 cells of 20 m, Poisson(128) SIPs per cell, bimodal AS aerosol,
Long E_col table, dt_adv = 1 s, 10 condensation substeps, 2 collision steps per advection step,
relaxation on.  --case stress150: 150x150 cells of 10 m, Poisson(256), Hall E_col, dt 0.5 s,
20 substeps, 16 seeds.

A "step" is one advection step of every SIP of every run.  SIP-steps/s = sum over runs of SIPs x
timed steps / wall seconds, with all inputs resident in HBM; every timed region is bracketed by
torch.cuda.synchronize().  Legs (--legs, default all):
  * early:    the 64-seed batch a few steps after the synthetic start (eager and HIP-graph replays)
  * plain:    one run, same state, eager and graph (batch / plain ratio)
  * drizzle:  the collision-active state: one run spun up 7200 s without collisions, replicated to the
              64 runs, then 900 s with collisions (every run its own draws) before the timed steps;
              eager and graph, kernel times, and the plain run in the same state.  The headline `value`
              is this leg's graph-replay rate (the 64-seed batch in its stated regime, hipGraph step)
  * adjoint:  cost of the adjoint window after the spin-up and its finite-difference check (icmw75: from the
              reference's SinSIP initial state computed on the GPU, mode-parameter gradient through its map)
  * box:      the Golovin box model, 50 seeds x 3600 s (the oracle's time for one box beside it)
  * stress:   the 150x150 Poisson(256) 16-seed batch
  * stress_adjoint: the adjoint cost of one 150x150 run: 600-step window (the last 300 s of 1800 s) after a
              3000-step spin-up with collisions, checkpoints every 10 steps (no finite-difference check)
  * init:     GPU initial states of 64 seeds
  * cpu:      the oracle (C restatement) on the host cores: 5 warm-up, 3 x 60 timed advection steps
              (median of 3) per process, one process per core of the box's share and one alone
"""
import argparse
import json
import os
import platform
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B_SUBSTEP = 104.0    # algorithmic bytes per SIP per particle substep (60 read + 44 write)
B_COLLISION = 64.0   # per SIP per collision step (40 read + 24 write)
B_REINDEX = 140.0    # per SIP per reindex
HOST_THREADS = 16    # the box's CPU share for one GPU
LEGS = ("early", "plain", "drizzle", "adjoint", "box", "stress", "stress_adjoint", "init", "cpu")

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
    return profile_value(TRAFFIC_PROFILE, "TRAFFIC", kernel)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _kernel_data(kernel):
    from lacmopy_amd import load_kernel_data_Ecol
    if kernel == "const":
        return 0.5, None, None
    E, lnR0, lnf, _ = load_kernel_data_Ecol(kernel)
    return E, lnR0, lnf


def build_cases(n_runs, spc, seed0, case):
    """Host inputs of n_runs seeds (seed0 + 2 r) of `case`, built on the host threads."""
    from lacmopy_amd.setup_case import workload_case

    def one(r):
        grid, pos, vel, cells, m_w, m_s, xi, act, draws = workload_case(case, seed0 + 2 * r, spc=spc,
                                                                        return_draws=True)
        return grid, (pos, vel, cells, m_w, m_s, xi, act), draws

    with ThreadPoolExecutor(min(HOST_THREADS, max(n_runs, 1))) as ex:
        return list(ex.map(one, range(n_runs)))


def new_engine(case, kernel, seed0, n_per_run, grid0):
    from lacmopy_amd.engine import Engine
    c = CASES[case]
    E, lnR0, lnf = _kernel_data(kernel)
    eng = Engine(c["nx"], c["nz"], n_per_run, grid0.ranges, grid0.steps, grid0.volume_cell)
    eng.set_params(solute="AS", no_iter_impl_mass=3, scale_dt_cond=c["sdc"], no_col_per_adv=2, act_collisions=True,
                   act_relaxation=True, dt=c["dt"], dt_col=c["dt"] / 2, g_set=9.80665, seed=seed0)
    eng.set_collision_kernel(E, lnR0, lnf)
    eng.set_flow(grid0.velocity, grid0.mass_flux_air_dry)
    return eng


def build_engine(n_runs, spc, seed0, case="icmw75", kernel=None):
    """kernel: 'Long_Bott' / 'Hall_Bott' table, or 'const' (Hydro_E_const, E = 0.5); default: the case's."""
    import torch  # noqa: F401
    from lacmopy_amd.setup_case import init_profiles
    kernel = kernel or CASES[case]["kernel"]
    cases = build_cases(n_runs, spc, seed0, case)
    eng = new_engine(case, kernel, seed0, [len(k[1][5]) for k in cases], cases[0][0])
    for r, (g, _, _) in enumerate(cases):
        eng.set_relaxation(r, *init_profiles(g))
        eng.set_grid(r, g.scalar_fields(), g.mat_prop())
    eng.init_cells()
    eng.load_particles([k[1] for k in cases])
    n_sips = sum(len(k[1][5]) for k in cases)
    return eng, n_sips, cases


def replicate_engine(src, grid0, n_runs, case, kernel, seed0):
    """An n_runs engine whose runs all start from run 0 of `src` (particles, grid fields, step counter);
    the collision draws are keyed per run, so the runs evolve independently from there."""
    from lacmopy_amd.setup_case import init_profiles
    parts = src.export_particles(0)
    scal, mat = src.export_grid(0)
    eng = new_engine(case, kernel, seed0, [len(parts[5])] * n_runs, grid0)
    prof = init_profiles(grid0)
    for r in range(n_runs):
        eng.set_relaxation(r, *prof)
        eng.set_grid(r, scal, mat)
    eng.init_cells()
    eng.load_particles([parts] * n_runs)
    eng.col_step = src.col_step
    return eng


def tune(eng):
    t = time.perf_counter()
    tr = eng.autotune()
    out = {"chosen": tr["chosen"],
           "particle_ms": {f"{b}x{sp}/w{wp}": round(ms, 4) for (b, sp, wp), ms in tr["particles"].items()},
           "dml_work_ms": {f"{w}/{c}": round(ms, 4) for (w, c), ms in tr["dml"].items()},
           "collision_ms": {str(c): round(ms, 4) for c, ms in tr["collision"].items()},
           "collision_work_ms": {f"{w}/{c}": round(ms, 4) for (w, c), ms in tr["collision_work"].items()},
           "seconds": time.perf_counter() - t}
    log(f"autotune: {out}")
    return out


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


def rates(eng, n_sips, steps, warmup):
    """Eager and graph-replay SIP-steps/s over `steps` timed steps each."""
    s_e = time_steps(eng, steps, warmup, graph=False)
    s_g = time_steps(eng, steps, warmup, graph=True)
    eng.drop_graph()
    return {"eager_sip_steps_per_s": n_sips * steps / s_e, "eager_ms_per_step": s_e / steps * 1e3,
            "graph_sip_steps_per_s": n_sips * steps / s_g, "graph_ms_per_step": s_g / steps * 1e3,
            "steps": steps, "warmup": warmup}


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


def collision_events(eng):
    return eng.buf["no_cols"].sum(0).cpu().tolist()


# adjoint window per case: (window steps, checkpoint interval, spin-up steps, collisions in the spin-up,
# event capacity).  icmw75: 600 s after a 2 h spin-up without collisions, checkpoints every 60 steps;
# stress150: the last 300 s (600 steps of 0.5 s) of 1800 s with collisions, checkpoints every 10 steps (the
# tape of 10 steps of 5.8M SIPs x 20 substeps is 74 GB).
ADJ_DEFAULTS = {"icmw75": (600, 60, 7200, False, 1 << 20), "stress150": (600, 10, 3000, True, 1 << 23)}


def spin_up(eng, steps, collisions):
    """`steps` advection steps of the engine, with or without collisions (untimed preparation)."""
    import torch
    e = eng.e
    e.act_collisions = int(collisions)
    t = time.perf_counter()
    eng.adv_steps(steps)
    torch.cuda.synchronize()
    e.act_collisions = 1
    eng.check()
    log(f"spin-up: {steps} steps ({'with' if collisions else 'without'} collisions) in "
        f"{time.perf_counter() - t:.1f} s")


def sinsip_engine(seed, spc_modes=(52, 76), seed_sim=1001):
    """One run of the reference's own initial state on 75x75 cells, computed on the GPU (lacmopy_amd.initialize,
    init.py:439-1236) with SinSIP counts no_spcm = spc_modes (52, 76: about 128 SIPs per cell, twice the
    configuration's (26, 38)), in an engine with the icmw75 step (Long table, dt 1 s, 10 substeps, 2 collision
    steps).  Returns (engine, initial state, init config)."""
    from lacmopy_amd import load_kernel_data_Ecol
    from lacmopy_amd.engine import Engine
    from lacmopy_amd.initialize import initialize_batch, reference_config
    icfg = reference_config(no_spcm=np.array(spc_modes), seed_SIP_gen=seed)
    st = initialize_batch(icfg, [seed])[0]
    E, lnR0, lnf, _ = load_kernel_data_Ecol("Long_Bott")
    c = CASES["icmw75"]
    step = dict(solute_type="AS", dt_adv=c["dt"], dt_col=c["dt"] / 2, scale_dt_cond=c["sdc"], no_col_per_adv=2,
                no_iter_impl_mass=3, g_set=9.80665, act_collisions=True, act_relaxation=True, seed_sim=seed_sim,
                R_kernel_low_log=lnR0, bin_factor_R_log=lnf)
    return Engine.from_initial_states([st], step, E), st, icfg


def adjoint_cost(eng, start, window, every, mode_map=None, check=0, case="icmw75", state_desc=""):
    """Adjoint of a `window`-step window of one run of `case` from the state `start`: wall time of the plain
    forward window, of the recording forward (events + checkpoints every `every` steps) and of one gradient
    (taped re-runs + reverse sweep), for the cloud-droplet-number objective.  mode_map(theta) -> (xi, m_s) by
    SIP id: the initialisation map from the six aerosol-mode parameters to the window-start multiplicities and
    dry masses (valid when nothing collided before the window); the gradient with respect to the parameters
    follows by the chain rule through it.  check > 0: also compare the gradient with central differences of the
    window on `check` random droplets (multiplicity, h = 1e-3 relative) and on the six parameters (h = 1e-7),
    each run's cell membership fingerprinted after every step (adjoint.membership_fingerprint)."""
    import torch
    from lacmopy_amd.adjoint import AdjointWindow
    n1 = eng.n_tot
    eng.restore(start)
    eng.adv_steps(2)  # warm-up of the collision path
    eng.restore(start)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.adv_steps(window)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    eng.check()
    eng.restore(start)
    w = AdjointWindow(eng, window, checkpoint_every=every, ev_cap=ADJ_DEFAULTS[case][4])
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
    c = CASES[case]
    res = {"workload": f"one {case} run ({c['nx']}x{c['nz']} cells, {n1} SIPs), {window}-step window (dt "
                       f"{c['dt']:g} s, {2 * c['sdc']} substeps, 2 collision steps) from {state_desc}; fixed "
                       f"collision events; checkpoints every {every} steps",
           "sips": n1, "window_steps": window, "checkpoint_every": every, "events": w.n_events,
           "J_nc": float(J[0][0]), "J_lwp": float(J[1][0]),
           "forward_s": t_fwd, "record_s": t_rec, "gradient_s": t_grad,
           "cost_ratio": (t_rec + t_grad) / t_fwd, "gradient_over_forward": t_grad / t_fwd,
           "tape_gb": sum(t.numel() * t.element_size() for t in w.tape.values()) / 1e9}
    if mode_map is not None:
        eng.restore(w.start)
        slots = w.slot_map()
        cur = int(eng.e.cur)
        theta = torch.tensor(THETA0, dtype=torch.float64, device="cuda", requires_grad=True)
        xi_m, ms_m = mode_map(theta)
        res["mode_map_rel_dev_from_window_start"] = max(
            float(((xi_m.detach() - eng.buf[f"xi{cur}"][slots]).abs() / xi_m.detach().abs()).max()),
            float(((ms_m.detach() - eng.buf[f"m_s{cur}"][slots]).abs() / ms_m.detach().abs()).max()))
        (g_theta,) = torch.autograd.grad((xi_m * lam_xi).sum() + (ms_m * lam_ms).sum(), theta)
        res["dJnc_dmodes"] = g_theta.tolist()
    else:
        res["dJnc_dmodes"] = None
        res["dJnc_dmodes_note"] = ("collisions before the window change multiplicities and dry masses, so the "
                                   "window-start state is not the initialisation map's; per-SIP gradients only")
    log(f"adjoint: {res}")
    if check:
        res["fd_check"] = adjoint_fd_check(w, eng, mode_map, check)
        log(f"adjoint fd check: {res['fd_check']}")
    del w
    torch.cuda.empty_cache()
    return res


THETA0 = (60.0E6, 0.02, 1.4, 40.0E6, 0.075, 1.6)  # N [m^-3], r_med [um], sigma of both modes (config.py:106-114)


def adjoint_fd_check(w, eng, mode_map, n_check, h_sip=1.0E-3, h_mode=1.0E-7):
    """Central differences of the window's J_nc against the adjoint: `n_check` random droplets (multiplicity,
    h_sip relative) and, through mode_map, the six mode parameters (h_mode relative).  Each comparison carries
    the quotient's noise floor 256 eps |J| / (2 h |x| |fd|) (floor_rel: one perturbed trajectory rounds
    differently from the other) and the number of steps at whose end some SIP sat in another cell than in the
    unperturbed run (cell membership fingerprinted after every step: such a run carries piecewise-constant
    jumps that are no part of the derivative).  The maxima are taken over the comparisons resolved above their
    floor (floor_rel < 1e-3); for the droplets also without membership change."""
    import torch
    from lacmopy_amd.adjoint import membership_changes
    eps = np.finfo(np.float64).eps
    eng.restore(w.start)
    slots = w.slot_map()
    cur = int(eng.e.cur)
    xi0 = eng.buf[f"xi{cur}"][slots].clone()
    ms0 = eng.buf[f"m_s{cur}"][slots].clone()
    drop = ((eng.buf[f"cell{int(eng.e.ccur)}"][slots] >= 0) & (eng.buf[f"m_w{cur}"][slots] > 4.19E3)).cpu().numpy()
    fp0 = []
    J0 = float(np.sum(w.run(xi0, ms0, membership=fp0)[0]))
    lam_xi, _ = w.gradient(w_nc=1.0, w_lwp=0.0)
    out = {"h_sip": h_sip, "h_mode": h_mode, "sips": [], "modes": []}

    def J_run(x, m):
        fp = []
        J = float(np.sum(w.run(x, m, membership=fp)[0]))
        return J, membership_changes(fp, fp0)

    rng = np.random.default_rng(11)
    for gid in rng.choice(np.flatnonzero(drop), n_check, replace=False):
        x = float(xi0[gid])
        Jpm, ch = [], 0
        for sgn in (1.0, -1.0):
            p = xi0.clone()
            p[gid] = x * (1.0 + sgn * h_sip)
            J, c = J_run(p, ms0)
            Jpm.append(J)
            ch += c
        fd = (Jpm[0] - Jpm[1]) / (2 * h_sip * x)
        a = float(lam_xi[gid])
        scale = max(abs(fd), abs(a), 1e-300)
        out["sips"].append({"gid": int(gid), "adjoint": a, "fd": fd, "rel": abs(a - fd) / scale,
                            "floor_rel": 256 * eps * abs(J0) / (2 * h_sip * x) / scale,
                            "membership_changed_steps": ch})
    res_ = [s["rel"] for s in out["sips"] if s["membership_changed_steps"] == 0 and s["floor_rel"] < 1e-3]
    out["sips_resolved"] = len(res_)
    out["max_rel_sips_resolved"] = max(res_) if res_ else None
    out["sips_with_membership_change"] = sum(s["membership_changed_steps"] > 0 for s in out["sips"])
    if mode_map is not None:
        theta = torch.tensor(THETA0, dtype=torch.float64, device="cuda", requires_grad=True)
        xi_m, ms_m = mode_map(theta)
        w.run(xi_m.detach(), ms_m.detach())
        g_xi, g_ms = w.gradient(w_nc=1.0, w_lwp=0.0)
        (grad,) = torch.autograd.grad((xi_m * g_xi).sum() + (ms_m * g_ms).sum(), theta)
        fp0 = []
        Jm0 = float(np.sum(w.run(xi_m.detach(), ms_m.detach(), membership=fp0)[0]))
        for k in range(6):
            Jpm, ch = [], 0
            for sgn in (1.0, -1.0):
                th = np.array(THETA0)
                th[k] *= 1.0 + sgn * h_mode
                xx, mm = mode_map(torch.tensor(th, dtype=torch.float64, device="cuda"))
                J, c = J_run(xx.detach(), mm.detach())
                Jpm.append(J)
                ch += c
            fd = (Jpm[0] - Jpm[1]) / (2 * h_mode * THETA0[k])
            a = float(grad[k])
            scale = max(abs(fd), abs(a), 1e-300)
            out["modes"].append({"k": k, "adjoint": a, "fd": fd, "rel": abs(a - fd) / scale,
                                 "floor_rel": 256 * eps * abs(Jm0) / (2 * h_mode * THETA0[k]) / scale,
                                 "membership_changed_steps": ch})
        res_ = [m["rel"] for m in out["modes"] if m["floor_rel"] < 1e-3]
        out["modes_resolved"] = len(res_)
        out["max_rel_modes_resolved"] = max(res_) if res_ else None
        out["modes_with_membership_change"] = sum(m["membership_changed_steps"] > 0 for m in out["modes"])
    return out


def box_model_leg(seed0, n_seeds=50, t_end=3600.0, dt_save=150.0, b=1.5, LWC0=1.0E-3):
    """The Golovin box model of the reference (run_box_model.py:133-153, 306-328): n_seeds boxes of the
    expo SinSIP ensemble (kappa 40), dt 1 s, batched on the GPU (one launch per step for all boxes; wall
    time includes the snapshot copies), the ensemble moments against Golovin's closed forms, and the
    oracle's time for the first box on one host core."""
    import math
    import torch
    from lacmopy_amd.collision.box_model import golovin_expo_ensemble, simulate_collisions_batched
    seeds = [seed0 + 2 * k for k in range(n_seeds)]
    boxes, DNC0 = [], None
    for s in seeds:
        xi, m, DNC0 = golovin_expo_ensemble(s)
        boxes.append((xi, m))
    host = [(x.copy(), m.copy()) for x, m in boxes[:1]]
    n_sips = sum(len(x) for x, _ in boxes)
    simulate_collisions_batched([(x.copy(), m.copy()) for x, m in boxes], None, "Golovin", "analytic", 1.0, 1.0,
                                10.0, 10.0, seed0)  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    xs, ms, times, nc = simulate_collisions_batched(boxes, None, "Golovin", "analytic", 1.0, 1.0, t_end, dt_save,
                                                    seed0)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t
    mom = np.zeros((len(times), 3, n_seeds))
    for k in range(n_seeds):
        mk = 1.0E-18 * ms[k]
        for n in range(3):
            mom[:, n, k] = (xs[k] * mk ** n).sum(axis=1)
    avg = mom.mean(axis=2)
    an = np.array([[DNC0 * math.exp(-b * LWC0 * t), LWC0, 2 * LWC0 ** 2 / DNC0 * math.exp(2 * b * LWC0 * t)]
                   for t in times])
    rel = avg / an - 1.0
    out = {"workload": f"Golovin box model, {n_seeds} seeds (expo SinSIP, kappa 40), dt 1 s, {int(t_end)} s",
           "sips": n_sips, "steps": int(t_end), "wall_s": wall, "sip_steps_per_s": n_sips * t_end / wall,
           "collision_events": nc.sum(0).tolist(),
           "lambda0_max_rel_dev": float(np.max(np.abs(rel[:, 0]))),
           "lambda2_rel_dev_600s": float(rel[times == 600.0, 2][0]),
           "lambda2_rel_dev_end": float(rel[-1, 2])}
    try:
        from oracle import oracle as orc
        x, m = host[0]
        no = np.zeros(2, dtype=np.int64)
        t = time.perf_counter()
        for step in range(int(t_end)):
            orc.golovin_step(x, m, 1.0, no, seed=seed0, run=0, step=step)
        cpu_s = time.perf_counter() - t
        out["cpu_oracle_s_per_box"] = cpu_s
        out["gpu_over_cpu_1_core"] = cpu_s * n_seeds / wall
        # the oracle's first box against the GPU's (the same keyed draws)
        out["box0_bit_exact"] = bool(np.array_equal(x, boxes[0][0]) and np.array_equal(m, boxes[0][1]))
    except (ImportError, OSError) as ex:
        out["cpu_oracle_s_per_box"] = None
        out["cpu_note"] = f"oracle not loaded: {ex}"
    log(f"box model: {out}")
    return out


def _cpu_run(args):
    """One process of the CPU baseline: the protocol on its own seed of the reduced case."""
    seed0, spc, nx, warm, timed, reps = args
    return cpu_protocol(seed0, spc, nx, warm, timed, reps)


def cpu_protocol(seed0, spc, nx=2, warm=5, timed=60, reps=3):
    """Oracle (C restatement, 1 thread) on a reduced domain of the 75x75 case: `nx` columns x 75 levels of
    20 m cells (the whole vertical profile and its cloud layer), Poisson(spc) SIPs per cell, the same step
    (Long table, dt 1 s, 10 substeps, 2 collision steps, relaxation); `warm` untimed advection steps, then
    `reps` repetitions of `timed` steps; returns the median SIP-steps/s and all three."""
    from lacmopy_amd import load_kernel_data_Ecol
    from lacmopy_amd.setup_case import init_profiles, make_case, workload_config
    from oracle import oracle as orc
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=nx, nz=75, spc=spc, seed=seed0)
    N = len(xi)
    cfg = workload_config("icmw75", E, lnR0, lnf)
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

    def steps(n):
        orc.integrate_adv_steps(p, n, pos, vel, cells, rel, m_w, m_s, xi, act, T_p, wr, nco, scal, mat,
                                grid.velocity, grid.mass_flux_air_dry, E, prof[0], prof[1], prof[2])

    steps(warm)
    r = []
    for _ in range(reps):
        t0 = time.perf_counter()
        steps(timed)
        r.append(N * timed / (time.perf_counter() - t0))
    return {"median": float(np.median(r)), "reps": r, "sips": N, "collision_events": nco.tolist()}


def cpu_full_domain_sample(seed0, spc, n_steps=1):
    """The oracle on the full 75x75 case for n_steps advection steps (one thread): the per-SIP rate at the
    workload's own size, beside the reduced-domain protocol."""
    from lacmopy_amd import load_kernel_data_Ecol
    from lacmopy_amd.setup_case import init_profiles, make_case, workload_config
    from oracle import oracle as orc
    E, lnR0, lnf, nb = load_kernel_data_Ecol("Long_Bott")
    grid, pos, vel, cells, m_w, m_s, xi, act = make_case(nx=75, nz=75, spc=spc, seed=seed0)
    N = len(xi)
    cfg = workload_config("icmw75", E, lnR0, lnf)
    p = orc.adv_params(grid, N, cfg, seed=seed0)
    scal, mat = grid.scalar_fields(), grid.mat_prop()
    prof = init_profiles(grid)
    rel = np.zeros((2, N))
    q = orc.adv_params(grid, N, cfg)
    orc.lib().or_update_pos_and_cells(orc.C.byref(q), 0.0, orc.P(pos), orc.P(vel), orc.P(cells), orc.P(rel),
                                      orc.P(m_w), orc.P(xi), orc.P(act.view(np.uint8)), orc.P(np.zeros(1)))
    t0 = time.perf_counter()
    orc.integrate_adv_steps(p, n_steps, pos, vel, cells, rel, m_w, m_s, xi, act, np.ones(N), np.zeros(1),
                            np.zeros(2, dtype=np.int64), scal, mat, grid.velocity, grid.mass_flux_air_dry, E,
                            prof[0], prof[1], prof[2])
    return {"sips": N, "steps": n_steps, "sip_steps_per_s": N * n_steps / (time.perf_counter() - t0)}


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model or platform.processor(), "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(seed0, spc, workers, nx=2):
    """CPU timing protocol: single core (one process alone), then all cores of the share (one
    independent run per core, started together, the reference's one process per run, run_lacmo.sh:7-16);
    each process 5 warm-up + 3 x 60 timed advection steps, median of the 3; plus one step of the full 75x75
    case on one core."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    t = time.perf_counter()
    single = cpu_protocol(seed0, spc, nx)
    if workers > 1:
        with ctx.Pool(workers) as pool:
            per = pool.map(_cpu_run, [(seed0 + 2 * k, spc, nx, 5, 60, 3) for k in range(workers)])
        agg = float(np.sum([r["median"] for r in per]))
    else:
        per, agg = [single], single["median"]
    full = cpu_full_domain_sample(seed0, spc)
    out = {"value": agg, "unit": "SIP-steps/s", "cores": workers, "kind": "port",
           "sample": f"oracle C restatement, 1 thread per process, on {nx} columns x 75 levels of the 75x75 case "
                     f"(20 m cells, Poisson({spc:g}), the same step); per process 5 warm-up and 3 x 60 timed "
                     f"advection steps, median of 3; {workers} processes started together, one seed each; "
                     f"value = sum of their medians",
           "single_core": single["median"], "single_core_reps": single["reps"],
           "per_process_medians": [r["median"] for r in per], "sips_per_process": single["sips"],
           "full_domain_1_core": full, "host": host_info(), "seconds": time.perf_counter() - t}
    log(f"cpu baseline: {out}")
    return out


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


def roofline(ktimes, n_sips, default_workload):
    """The particle substep kernel (one thread per SIP, the step's dominant kernel) against HBM: algorithmic
    bytes of one launch over its HIP-event duration; its VALU co-bound from the committed counter summary."""
    if not ktimes or "particles" not in ktimes:
        return None
    ms = ktimes["particles"][0]
    achieved = n_sips * B_SUBSTEP / (ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic("k_particles") if default_workload else None,
            "kernel": "k_particles", "bytes_per_launch": n_sips * B_SUBSTEP, "launch_ms": ms}
    valu = profile_value(VALU_PROFILE, "VALU", "k_particles") if default_workload else None
    flops = profile_value(VALU_PROFILE, "FP64FLOPS", "k_particles") if default_workload else None
    if valu and flops:
        roof["valu"] = {"achieved": valu / (ms * 1e-3), "peak": VALU_ISSUE_PEAK, "unit": "wave-instructions/s",
                        "frac": valu / (ms * 1e-3) / VALU_ISSUE_PEAK,
                        "fp64_tflops": flops / (ms * 1e-3) / 1e12,
                        "fp64_frac": flops / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
    return roof


def roofline_collision(ktimes, n_sips, default_workload):
    """The collision step's main pass in the early state (where its counters were taken): algorithmic bytes of
    one launch (64 B per SIP) over its HIP-event duration against HBM, and its VALU issue rate from the
    committed counter summary (the pass is bound by instruction issue)."""
    if not ktimes or "collision" not in ktimes:
        return None
    ms = ktimes["collision"][0]
    achieved = n_sips * B_COLLISION / (ms * 1e-3) / 1e9
    roof = {"leg": "early", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_collision") if default_workload else None,
            "kernel": "k_collision<160>", "bytes_per_launch": n_sips * B_COLLISION, "launch_ms": ms}
    valu = profile_value(VALU_PROFILE, "VALU", "k_collision") if default_workload else None
    if valu:
        roof["valu"] = {"achieved": valu / (ms * 1e-3), "peak": VALU_ISSUE_PEAK, "unit": "wave-instructions/s",
                        "frac": valu / (ms * 1e-3) / VALU_ISSUE_PEAK}
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--case", default="icmw75", choices=sorted(CASES))
    ap.add_argument("--runs", type=int, default=None, help="seeds in the batch (default: the case's)")
    ap.add_argument("--spc", type=float, default=None, help="mean SIPs per cell (default: the case's)")
    ap.add_argument("--seed", type=int, default=1001)
    ap.add_argument("--legs", default=",".join(LEGS), help=f"comma-separated subset of {','.join(LEGS)}")
    ap.add_argument("--cpu-workers", type=int, default=min(HOST_THREADS, len(os.sched_getaffinity(0))),
                    help="processes of the all-cores CPU baseline (1: single core only)")
    ap.add_argument("--kernel", default=None, choices=["Long_Bott", "Hall_Bott", "const"])
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuning)")
    ap.add_argument("--spinup", type=int, default=None,
                    help="spin-up steps without collisions before the collision-active leg and the adjoint "
                         "(default 7200; stress150: 3000, with collisions)")
    ap.add_argument("--col-steps", type=int, default=900,
                    help="advection steps with collisions between the spin-up and the collision-active timing")
    ap.add_argument("--eager-prep", action="store_true",
                    help="run the collision-active leg's preparation steps as eager launches instead of graph "
                         "replays (same state; for kernel tracers that record graph launches poorly)")
    ap.add_argument("--adj-window", type=int, default=None,
                    help="advection steps of the adjoint window (default 600 for both cases)")
    ap.add_argument("--adj-every", type=int, default=None,
                    help="checkpoint interval of the adjoint window (default 60; stress150: 10)")
    ap.add_argument("--adj-check", type=int, default=20,
                    help="droplets of the finite-difference check of the adjoint (plus the six aerosol-mode "
                         "parameters); 0: no check")
    ap.add_argument("--adj-init", default=None, choices=["sinsip", "synthetic"],
                    help="start state of the icmw75 adjoint window: the reference's SinSIP initialisation computed "
                         "on the GPU (about 128 SIPs per cell; default) or the synthetic Poisson sampler; the "
                         "mode-parameter gradient follows the chosen state's initialisation map")
    ap.add_argument("--init-seeds", type=int, default=64, help="seeds of the timed GPU initialisation")
    ap.add_argument("--stress-runs", type=int, default=16, help="seeds of the stress150 batch leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (calibration of the HBM byte counters)")
    args = ap.parse_args()
    legs = [x for x in args.legs.split(",") if x]
    bad = [x for x in legs if x not in LEGS]
    if bad:
        ap.error(f"unknown legs {bad}")
    case = CASES[args.case]
    args.runs = args.runs or case["runs"]
    args.spc = args.spc or case["spc"]
    args.kernel = args.kernel or case["kernel"]
    n_sub = 2 * case["sdc"]
    adj_w, adj_k, adj_s, adj_col, _ = ADJ_DEFAULTS[args.case]
    args.adj_window = args.adj_window or adj_w
    args.adj_every = args.adj_every or adj_k
    args.spinup = args.spinup if args.spinup is not None else adj_s
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): independent seeds shard "
              "trivially across GPUs and no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "needs a GPU"
    if args.copy_probe:
        copy_probe()
        return
    icmw = args.case == "icmw75"
    default_workload = (icmw and args.runs == 64 and args.spc == 128.0 and args.seed == 1001
                        and args.kernel == "Long_Bott")
    out = {}

    # ---- early state: the batch a few steps after the synthetic start
    shape = None
    if "early" in legs:
        log(f"building batch of {args.runs} runs ...")
        t = time.perf_counter()
        eng, n_sips, _ = build_engine(args.runs, args.spc, args.seed, case=args.case, kernel=args.kernel)
        log(f"  {n_sips} SIPs loaded in {time.perf_counter() - t:.1f} s")
        tn = None
        if not args.no_tune:
            eng.adv_steps(1)  # tune on an evolved state
            tn = tune(eng)
            shape = tn["chosen"]
        r = rates(eng, n_sips, args.steps, args.warmup)
        r.update(sips=n_sips, collision_events=collision_events(eng), autotune=tn,
                 kernel_ms=kernel_times(eng, max(1, min(args.steps, 3))))
        out["early"] = r
        log(f"early batch: {r}")
        del eng
        torch.cuda.empty_cache()

    if "plain" in legs:
        eng1, n1, _ = build_engine(1, args.spc, args.seed, case=args.case, kernel=args.kernel)
        shape1 = None
        if not args.no_tune:
            eng1.adv_steps(1)
            shape1 = eng1.autotune()["chosen"]
        r = rates(eng1, n1, max(args.steps, 10), args.warmup)
        r.update(sips=n1, launch_shape=shape1)
        out["plain_early"] = r
        log(f"plain early: {r}")
        del eng1

    # ---- spin-up of one run: the collision-active state (and the synthetic adjoint window start)
    adj_init = args.adj_init or ("sinsip" if icmw else "synthetic")
    spin = None
    if "drizzle" in legs or ("adjoint" in legs and adj_init == "synthetic"):
        spin, n1, cases1 = build_engine(1, args.spc, args.seed, case=args.case, kernel=args.kernel)
        spin_up(spin, args.spinup, adj_col)
        spun = spin.snapshot()

    if "drizzle" in legs:
        spin.adv_steps(args.col_steps)
        torch.cuda.synchronize()
        spin.check()
        drz = {"state": f"seed {args.seed}: {args.spinup} steps {'with' if adj_col else 'without'} collisions, "
                        f"then {args.col_steps} steps with collisions",
               "plain_collision_events": collision_events(spin)}
        t = time.perf_counter()
        eng = replicate_engine(spin, cases1[0][0], args.runs, args.case, args.kernel, args.seed)
        log(f"collision-active batch: {args.runs} copies loaded in {time.perf_counter() - t:.1f} s")
        n_sips_d = eng.n_tot
        if not args.no_tune:
            eng.adv_steps(1)
            drz["autotune"] = tune(eng)
        # the runs diverge under their own draws for col_steps more steps before the timing
        if args.eager_prep:
            eng.adv_steps(args.col_steps)
        else:
            eng.adv_steps_graph(args.col_steps)
        torch.cuda.synchronize()
        eng.check()
        eng.drop_graph()
        c0 = collision_events(eng)
        r = rates(eng, n_sips_d, args.steps, args.warmup)
        drz.update(r)
        drz["sips"] = n_sips_d
        drz["collision_events_in_timed_steps"] = [a - b for a, b in zip(collision_events(eng), c0)]
        drz["kernel_ms"] = kernel_times(eng, max(1, min(args.steps, 3)))
        drz["state"] += (f"; replicated to {args.runs} runs, {args.col_steps} more steps with per-run draws "
                         f"before the timing")
        del eng
        torch.cuda.empty_cache()
        if not args.no_tune:
            spin.autotune()
        pr = rates(spin, spin.n_tot, max(args.steps, 10), args.warmup)
        drz["plain"] = pr
        drz["batch_over_plain_graph"] = drz["graph_sip_steps_per_s"] / pr["graph_sip_steps_per_s"]
        out["drizzle"] = drz
        log(f"collision-active: {drz}")

    if "adjoint" in legs:
        col_desc = "with" if adj_col else "without"
        if adj_init == "sinsip":
            from lacmopy_amd.initialize import sinsip_mode_map
            if spin is not None:
                del spin, spun
                spin = None
                torch.cuda.empty_cache()
            t = time.perf_counter()
            aeng, ast, aicfg = sinsip_engine(3711, seed_sim=args.seed)
            log(f"adjoint: SinSIP initial state ({aeng.n_tot} SIPs) in {time.perf_counter() - t:.1f} s")
            spin_up(aeng, args.spinup, False)
            out["adjoint"] = adjoint_cost(
                aeng, aeng.snapshot(), args.adj_window, args.adj_every,
                mode_map=lambda th: sinsip_mode_map(th, ast, aicfg), check=args.adj_check, case=args.case,
                state_desc=f"the reference's SinSIP initial state (computed on the GPU, seed_SIP_gen 3711, no_spcm "
                           f"(52, 76)) after a {args.spinup}-step spin-up without collisions")
            del aeng
        else:
            from lacmopy_amd.setup_case import aerosol_sips
            draws = cases1[0][2]
            out["adjoint"] = adjoint_cost(
                spin, spun, args.adj_window, args.adj_every,
                mode_map=None if adj_col else (lambda th: aerosol_sips(th, draws)), check=args.adj_check,
                case=args.case, state_desc=f"the synthetic start after a {args.spinup}-step spin-up {col_desc} "
                                           f"collisions")
    if spin is not None:
        del spin, spun
    torch.cuda.empty_cache()

    if "box" in legs and icmw:
        out["box_model"] = box_model_leg(args.seed)

    if "stress" in legs and icmw:
        t = time.perf_counter()
        eng, n_s, _ = build_engine(args.stress_runs, None, args.seed, case="stress150")
        log(f"stress150 batch: {n_s} SIPs loaded in {time.perf_counter() - t:.1f} s")
        st = {"workload": f"stress150: 150x150 cells of 10 m, Poisson(256), {args.stress_runs} seeds, Hall E_col, "
                          f"dt 0.5 s, 20 substeps, 2 collision steps; a few steps after the synthetic start"}
        if not args.no_tune:
            eng.adv_steps(1)
            st["launch_shape"] = eng.autotune()["chosen"]
        st.update(rates(eng, n_s, 5, 2))
        st["sips"] = n_s
        st["kernel_ms"] = kernel_times(eng, 1)
        out["stress"] = st
        log(f"stress: {st}")
        del eng
        torch.cuda.empty_cache()

    if "stress_adjoint" in legs and icmw:
        win, every, spin_steps, spin_col, _ = ADJ_DEFAULTS["stress150"]
        seng, _, _ = build_engine(1, None, args.seed, case="stress150")
        spin_up(seng, spin_steps, spin_col)
        out["stress_adjoint"] = adjoint_cost(
            seng, seng.snapshot(), win, every, mode_map=None, check=0, case="stress150",
            state_desc=f"the synthetic start after a {spin_steps}-step spin-up {'with' if spin_col else 'without'} "
                       f"collisions")
        del seng
        torch.cuda.empty_cache()

    if "init" in legs and icmw:
        out["initial_states"] = init_cost(args.init_seeds, cpu="cpu" in legs)

    cpu = cpu_baseline(args.seed, args.spc, args.cpu_workers) if ("cpu" in legs and icmw) else None

    # headline: the batch in the collision-active state with the hipGraph step; else the early state
    if "drizzle" in out:
        head, kt, n_head = out["drizzle"], out["drizzle"]["kernel_ms"], out["drizzle"]["sips"]
        head_desc = "collision-active state (" + out["drizzle"]["state"] + "), two-step HIP graph replays"
        value, ms_step = head["graph_sip_steps_per_s"], head["graph_ms_per_step"]
    elif "early" in out:
        head, kt, n_head = out["early"], out["early"]["kernel_ms"], out["early"]["sips"]
        head_desc = f"{args.warmup} steps after the synthetic start, eager launches"
        value, ms_step = head["eager_sip_steps_per_s"], head["eager_ms_per_step"]
    else:
        head, kt, n_head, head_desc, value, ms_step = None, None, None, None, None, None
    plain = out.get("drizzle", {}).get("plain") or out.get("plain_early")
    res = {
        "metric": "SIP-steps/s, 2D ICMW 75x75: 64-seed batch vs plain GPU; adjoint/forward cost"
        if icmw else f"SIP-steps/s, {args.case}",
        "value": value,
        "unit": "SIP-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"ICMW 2D kinematic Sc ({args.case}), {case['nx']}x{case['nz']} cells of "
                               f"{case['dx']:g} m, Poisson({args.spc:g}) SIPs/cell, {args.runs} seeds batched, "
                               f"{args.kernel} E_col, dt={case['dt']:g} s, {n_sub} substeps, 2 collision steps; "
                               f"{head_desc}",
                   "runs": args.runs, "sips_total": n_head, "cells": case["nx"] * case["nz"],
                   "algorithmic_bytes_per_sip_step": n_sub * B_SUBSTEP + 2 * B_COLLISION + B_REINDEX,
                   "batch_over_plain": (head["graph_sip_steps_per_s"] / plain["graph_sip_steps_per_s"])
                   if (head and plain) else None,
                   **out},
        "roofline": roofline(kt, n_head, default_workload) if kt else None,
        "roofline_collision": roofline_collision(out["early"]["kernel_ms"], out["early"]["sips"], default_workload)
        if "early" in out else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
