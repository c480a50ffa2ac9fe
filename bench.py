"""Benchmark: particle-substeps/s of the differentiable MLS-MPM substep on one
MI355X, forward + adjoint, on the single-rollout workload: 200k particles in
an ellipsoid (radii 0.2, 0.1, 0.15) on a 64^3 grid, Neo-Hookean + von Mises,
dt = 5e-5, 400 substeps, loss = mean squared distance of the final
positions to a target rollout made with other parameters, output the
gradient w.r.t. (log E, nu, log yield).

One step = one forward rollout recording its tape + loss + one reverse
sweep over the recorded substeps.  value = N * 400 / step time.

Prints one JSON line (see --help).  --gpus N > 1 exits: the library runs
the MPM hot path on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# reverse-mode tape: 1 = every substep recorded (state, Fe, grid; ~35 MB per
# substep here, 14 GB per rollout), so the backward recomputes nothing
CKPT_EVERY = 1
# the whole-step algorithmic byte count below is the stage-separated
# fwd+bwd minimum with checkpoint-recompute every 50 substeps; it is kept
# fixed as the unit of work whatever tape policy the run uses
B_FB_K = 50


def kernel_bytes(N, G, T=1):
    """Algorithmic HBM bytes per launch of each kernel (FP32): the particle
    and node fields that kernel must read or write, each once.  T: substeps
    of the reverse sweep whose launches are averaged (all but its last run
    the fused P2G + G2P adjoint)."""
    # P2G adjoint alone: reads state + constants (116), partial adjoints (48) and dL/d(mv, m) once per node;
    # writes the adjoint of the state (96) and read-modify-writes 5 parameter adjoints (40)
    p2g_adj = 300 * N + 16 * G
    # fused with the G2P adjoint of the substep before: the adjoints of x and v (24 B) stay in registers
    # instead of being written, and that G2P adjoint reads x, Fe (48) and v_out once per node and writes
    # dL/dx partial (12), dL/dFe (36) and dL/dv' (16); the adjoint of the next state comes in registers
    fused = (p2g_adj - 24 * N) + (112 * N + 12 * G)
    T = max(1, int(T))
    return {
        # reads x, v, C (60 B), the stress payload S (36 B) and mass (4 B); writes (mv, m) once per node
        "k_p2g": 100 * N + 16 * G,
        # reads (mv, m), writes v_out (12 B) per node
        "k_grid": 28 * G,
        # reads x, Fe (48 B), mu, lam, yield, vol (16 B) and v_out once per node; writes x, v, C, F_trial
        # (96 B), the next substep's Fe and S (72 B) and its binning's key and rank (8 B)
        "k_g2p": 240 * N + 12 * G,
        # reads F_trial (36 B) and mu, lam, yield, vol (16 B); writes Fe and S (72 B)
        "k_stress": 124 * N,
        # (the first substep of a sweep only) reads x, Fe (48), the adjoint of the next state (96) and v_out;
        # writes dL/dx partial (12), dL/dFe (36) and dL/dv' for the scatter (16)
        "k_g2p_adj": 208 * N + 12 * G,
        # reads x, dL/dv' (28 B), dL/dC', dL/dF' and Fe (108 B); writes dL/dv_out once per node (12 B)
        "k_g2p_adj_scatter": 136 * N + 12 * G,
        # reads (mv, m) and dL/dv_out, writes dL/d(mv, m)
        "k_grid_adj": 44 * G,
        # T - 1 fused launches and one alone per sweep
        "k_p2g_adj": (fused * (T - 1) + p2g_adj) / T,
        # the binning kernels of a substep whose counting the previous G2P did: key and rank (8 B) read,
        # perm (4 B) written per particle; per 2x2x2-cell block (G / 8): 8 counts read and re-zeroed,
        # 9 offsets written and the occupancy, near and list entries (about 116 B, 14.5 B per node)
        "k_bin": 12 * N + 14.5 * G,
    }


# HBM bytes per launch measured with FETCH_SIZE / WRITE_SIZE counter runs of
# this workload (calibrated on a plain copy); the file gives the commands
TRAFFIC_PROFILE = "profiles/c2_hbm_traffic_pmc.md"
TRAFFIC_PROFILE_LBS = "profiles/lbs_hbm_traffic_pmc.md"


def measured_traffic(kernel, profile=None):
    """Counter-measured HBM bytes per launch of `kernel`, or None."""
    try:
        with open(os.path.join(ROOT, profile or TRAFFIC_PROFILE)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return int(parts[3])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def emit(args, result):
    """Print the result line; with --metrics-jsonl also append it there."""
    print(json.dumps(result), flush=True)
    if args.metrics_jsonl:
        from empm_amd import metrics

        metrics.append_record(args.metrics_jsonl, result, argv=sys.argv[1:])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=200_000)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--substeps", type=int, default=400)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--only", choices=("all", "single", "population", "serving", "long_horizon", "loss", "lbs", "c1"),
                    default="all", help="run only the single-rollout, the population, the serving, the long-horizon, "
                                        "the frame-loss, the Gaussian interpolation (lbs) or the 32^3 cube (c1) "
                                        "workload")
    ap.add_argument("--serve-requests", type=int, default=60, help="requests of the serving workload")
    ap.add_argument("--serve-qps", type=float, default=10.0, help="Poisson arrival rate of the serving workload")
    ap.add_argument("--pop-steps", type=int, default=2, help="timed evaluations of the population workload")
    ap.add_argument("--pop-batched-only", action="store_true",
                    help="population workload: skip the sequential legs (for profiling the batched kernels)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (3 times) and a read-only sum of it "
                         "(3 times): the known-size kernels that calibrate FETCH_SIZE / WRITE_SIZE readings")
    ap.add_argument("--metrics-jsonl", default=None,
                    help="also append the result line to this JSON-lines file (time- and host-stamped)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the library runs the MPM hot path on one GPU only; --gpus must be 1 "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    from empm_amd import _lib
    from empm_amd import runtime as rt
    from empm_amd import scenes

    dev = torch.device("cuda:0")
    if args.copy_probe_mib > 0:
        src = torch.ones(args.copy_probe_mib * (1 << 18), device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {args.copy_probe_mib} MiB read + {args.copy_probe_mib} MiB written per launch")
        # a read-only probe of the same size (a sum: reads it, writes one
        # float per workgroup), which pins the FETCH_SIZE scale on its own
        for _ in range(3):
            src.sum()
        torch.cuda.synchronize()
        log(f"[bench] read probe: {args.copy_probe_mib} MiB read per launch (sum)")
        del src, dst
    if args.only == "population":
        emit(args, {"population": bench_population(args, dev)})
        return
    if args.only == "serving":
        emit(args, {"serving": bench_serving(args, dev)})
        return
    if args.only == "long_horizon":
        emit(args, {"long_horizon": bench_long_horizon(args, dev)})
        return
    if args.only == "loss":
        emit(args, {"loss": bench_loss(args, dev)})
        return
    if args.only == "lbs":
        emit(args, {"lbs": bench_lbs(args, dev)})
        return
    if args.only == "c1":
        emit(args, {"c1": bench_c1(args, dev)})
        return
    sc = scenes.ellipsoid_c2(n_particles=args.particles, n_grid=args.grid, n_steps=args.substeps, param_seed=3)
    tprm = scenes.sample_params_c2(2)
    N, n, T, dt = sc.n, sc.n_grid, sc.n_steps, sc.dt
    G = n ** 3
    log(f"[bench] workload {sc.name}: N={N} grid={n}^3 substeps={T} dt={dt} params={sc.params} target={tprm}")

    def t(a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()

    def model_for(prm):
        E = np.full(N, prm["E"])
        nu = np.full(N, prm["nu"])
        Et, nut = t(E), t(nu)
        mu, lam = torch.empty_like(Et), torch.empty_like(Et)
        rt.mu_lam(Et, nut, mu, lam)
        dx = float(np.float32(1.0 / n))
        return rt.ModelArrays(material=sc.material, n_grid=n, grid_lim=1.0, dx=dx, inv_dx=float(np.float32(n)),
                              mu=mu, lam=lam, yield_stress=t(np.full(N, prm["yield_stress"])),
                              mass=t(sc.vol * sc.rho), vol=t(sc.vol), gravity=sc.gravity), Et, nut

    def cols():
        from empm_amd.diff_simulator.warp_mpm.mpm_solver_diff import KIND_SURFACE
        out = []
        for c in sc.colliders:
            g = _lib.Collider()
            g.kind = KIND_SURFACE
            g.start_time, g.end_time = 0.0, 999.0
            for i in range(3):
                g.point[i] = c["point"][i]
                g.normal[i] = c["normal"][i]
            out.append(g)
        return out

    x0 = t(sc.x)
    s0 = rt.ParticleState(x0, t(sc.params["v0"]), torch.zeros((N, 3, 3), device=dev),
                          torch.eye(3, device=dev).repeat(N, 1, 1).contiguous())
    ctx = rt.Context(n, N, dev)

    # target rollout (forward only, untimed)
    tmodel, _, _ = model_for(tprm)
    tro = rt.Rollout(ctx, tmodel, N)
    tfinal, _, _, _ = tro.forward(s0, cols(), 0.0, dt, T)
    target = tfinal.x.clone()

    model, Et, nut = model_for(sc.params)
    ro = rt.Rollout(ctx, model, N)
    # launch parameters picked by the autotuner on this workload (every
    # candidate timed on 40 substeps of forward + reverse sweep; set-up, not
    # timed)
    tune = ro.autotune(s0, cols(), dt, n_steps=40, backward=True, cache={}, force=True)
    log(f"[bench] autotune: {tune['tuning']} after {tune['timings']} timings; 40 substeps fwd+bwd "
        f"{tune['seconds'] * 1e3:.2f} ms, default {tune['default_seconds'] * 1e3:.2f} ms")

    def step():
        final, _, _, tape = ro.forward(s0, cols(), 0.0, dt, T, tape=True, ckpt_every=CKPT_EVERY)
        diff = final.x - target
        loss = (diff * diff).sum(1).mean()
        d = rt.ParticleState.empty(N, dev)
        d.x.copy_(diff * (2.0 / N))
        _, grads = ro.backward(tape, d, flags=_lib.EMPM_ADJ_EXACT)
        dE = torch.zeros(N, device=dev)
        dnu = torch.zeros(N, device=dev)
        rt.mu_lam_adj(Et, nut, grads.mu, grads.lam, dE, dnu)
        g = torch.stack([(dE * Et).sum(), dnu.sum(), (grads.ys * model.yield_stress).sum()])
        return loss, g

    def fwd_only():
        final, _, _, _ = ro.forward(s0, cols(), 0.0, dt, T)
        return final

    lib = ctx.lib
    # profiling pass: time every kernel once to find the dominant one
    lib.empm_profile_enable(ctx.handle, (1 << len(_lib.KERNEL_NAMES)) - 1)
    step()
    torch.cuda.synchronize()
    per_kernel = {}
    for k, name in enumerate(_lib.KERNEL_NAMES):
        l, ms = C.c_long(), C.c_double()
        _lib.check(lib.empm_profile_read(ctx.handle, k, C.byref(l), C.byref(ms)))
        per_kernel[name] = (l.value, ms.value)
    lib.empm_profile_enable(ctx.handle, 0)
    _lib.check(lib.empm_profile_reset(ctx.handle))
    known = kernel_bytes(N, G, T)
    dom = max((k for k in per_kernel if k in known), key=lambda k: per_kernel[k][1])
    log("[bench] per-kernel (launches, ms) in one step: " +
        ", ".join(f"{k}={v[0]}/{v[1]:.2f}" for k, v in per_kernel.items()) + f" -> dominant {dom}")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # timed region; the dominant kernel is timed with events on its stream
    lib.empm_profile_enable(ctx.handle, 1 << _lib.KERNEL_NAMES.index(dom))
    torch.cuda.synchronize()
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    marks[0].record()
    for k in range(args.steps):
        loss, g = step()
        marks[k + 1].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    step_ms = sorted(marks[k].elapsed_time(marks[k + 1]) for k in range(args.steps))
    ms_median = step_ms[len(step_ms) // 2]
    l, ms = C.c_long(), C.c_double()
    _lib.check(lib.empm_profile_read(ctx.handle, _lib.KERNEL_NAMES.index(dom), C.byref(l), C.byref(ms)))
    lib.empm_profile_enable(ctx.handle, 0)
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    value = N * T / (ms_per_step / 1000.0)
    avg_ms = ms.value / max(l.value, 1)
    kb = kernel_bytes(N, G, T)[dom]
    achieved = kb / (avg_ms / 1000.0) / 1e9
    log(f"[bench] fwd+bwd: {ms_per_step:.2f} ms/step (median {ms_median:.2f}), {value:.4e} p-substeps/s; "
        f"{dom}: {l.value} launches, "
        f"avg {avg_ms * 1000:.1f} us, {achieved:.0f} GB/s; loss={float(loss):.6e} grad={g.tolist()}")

    # the same step with every rollout call replayed as a HIP graph (no
    # per-kernel events: a graph's launches are not timed one by one)
    gctx = rt.Context(n, N, dev).set_graphs(True).set_tuning(tune["tuning"])
    gro = rt.Rollout(gctx, model, N)

    def gstep():
        final, _, _, tape = gro.forward(s0, cols(), 0.0, dt, T, tape=True, ckpt_every=CKPT_EVERY)
        diff = final.x - target
        d = rt.ParticleState.empty(N, dev)
        d.x.copy_(diff * (2.0 / N))
        _, grads = gro.backward(tape, d, flags=_lib.EMPM_ADJ_EXACT)
        return grads

    for _ in range(max(2, args.warmup)):
        gstep()
    torch.cuda.synchronize()
    g0 = time.perf_counter()
    for _ in range(args.steps):
        gstep()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - g0) * 1000.0 / args.steps
    gcaps, greps = gctx.graph_stats()
    log(f"[bench] fwd+bwd with graphs: {graph_ms:.2f} ms/step ({N * T / (graph_ms / 1000.0):.4e} p-substeps/s), "
        f"{gcaps} captures / {greps} replays")
    graph_leg = {"ms_per_step": graph_ms, "value": N * T / (graph_ms / 1000.0), "captures": gcaps, "replays": greps}
    del gro, gctx
    torch.cuda.empty_cache()

    # the checkpoint-recompute policy the per-substep byte count B_fb below is
    # defined on: a state every 50 substeps, each segment replayed (recorded)
    # before its reverse sweep
    def kstep():
        final, _, _, tape = ro.forward(s0, cols(), 0.0, dt, T, tape=True, ckpt_every=B_FB_K)
        diff = final.x - target
        d = rt.ParticleState.empty(N, dev)
        d.x.copy_(diff * (2.0 / N))
        _, grads = ro.backward(tape, d, flags=_lib.EMPM_ADJ_EXACT)
        return grads

    kstep()
    torch.cuda.synchronize()
    k0 = time.perf_counter()
    for _ in range(args.steps):
        kstep()
    torch.cuda.synchronize()
    k50_ms = (time.perf_counter() - k0) * 1000.0 / args.steps
    k50_leg = {"ckpt_every": B_FB_K, "ms_per_step": k50_ms, "value": N * T / (k50_ms / 1000.0),
               "tape_gb": rt.tape_bytes(N, G, ro._bin_ints, T, B_FB_K) / 1e9,
               "tape_gb_recorded": rt.tape_bytes(N, G, ro._bin_ints, T, 1) / 1e9}
    log(f"[bench] fwd+bwd checkpoint every {B_FB_K}: {k50_ms:.2f} ms/step ({k50_leg['value']:.4e} p-substeps/s), "
        f"tape {k50_leg['tape_gb']:.2f} GB (recorded: {k50_leg['tape_gb_recorded']:.1f} GB)")

    # forward-only throughput (reported beside the headline)
    fwd_only()
    torch.cuda.synchronize()
    f0 = time.perf_counter()
    for _ in range(args.steps):
        fwd_only()
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - f0) * 1000.0 / args.steps
    fwd_value = N * T / (fwd_ms / 1000.0)
    log(f"[bench] fwd only: {fwd_ms:.2f} ms/rollout, {fwd_value:.4e} p-substeps/s")

    B_fb = 776 * N + 196 * G + 192 * N / B_FB_K
    B_f = 212 * N + 56 * G
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(sc, args.cpu_seconds)
    population = serving = long_horizon = frame_loss = lbs = c1 = None
    copy_gbs = copy_bandwidth(dev)
    log(f"[bench] device copy: {copy_gbs:.0f} GB/s read + write")
    if args.only == "all":
        del ro, tro, ctx
        torch.cuda.empty_cache()
        population = bench_population(args, dev)
        torch.cuda.empty_cache()
        serving = bench_serving(args, dev)
        torch.cuda.empty_cache()
        long_horizon = bench_long_horizon(args, dev)
        torch.cuda.empty_cache()
        frame_loss = bench_loss(args, dev)
        torch.cuda.empty_cache()
        lbs = bench_lbs(args, dev)
        torch.cuda.empty_cache()
        c1 = bench_c1(args, dev)

    out = {
        "metric": "particle-substeps/s (fwd+bwd) @200k pts/64^3",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "ms_per_step_median": ms_median,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded ellipsoid, random-init parameters)",
        "config": {"workload": "single rollout fwd+adjoint: ellipsoid 200k particles, 64^3 grid, "
                               "neo-hookean + von Mises, dt 5e-5, 400 substeps, every substep recorded",
                   "n_particles": N, "n_grid": n, "substeps": T, "ckpt_every": CKPT_EVERY},
        "fwd_value": fwd_value,
        "fwd_ms_per_rollout": fwd_ms,
        "fwd_bwd_graph": graph_leg,
        "fwd_bwd_ckpt50": k50_leg,
        "tuning": {"chosen": tune["tuning"], "timings": tune["timings"], "seconds_40_substeps": tune["seconds"],
                   "default_seconds_40_substeps": tune["default_seconds"]},
        # both against B_fb (checkpoint-recompute every 50 substeps): the
        # headline's recorded tape moves other bytes (no recompute, the
        # records instead), so its fraction is of the same unit of work
        "hbm_roofline_fraction_fwd_bwd": (B_fb * T / (ms_per_step / 1000.0)) / 1e9 / HBM_PEAK_GBS,
        "hbm_roofline_fraction_fwd_bwd_ckpt50": (B_fb * T / (k50_ms / 1000.0)) / 1e9 / HBM_PEAK_GBS,
        "hbm_roofline_fraction_fwd": (B_f * T / (fwd_ms / 1000.0)) / 1e9 / HBM_PEAK_GBS,
        "grad_logE_nu_logY": [float(v) for v in g.tolist()],
        "loss": float(loss),
        "per_kernel_ms_one_step": {k: v[1] for k, v in per_kernel.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "peak_measured_copy": copy_gbs,
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom) if (N, n) == (200_000, 64) else None,
                     "traffic_source": TRAFFIC_PROFILE, "algorithmic_bytes": kb,
                     "avg_launch_us": avg_ms * 1000.0, "launches": l.value},
        "cpu_baseline": cpu,
        "batched": population,
        "serving": serving,
        "long_horizon": long_horizon,
        "loss": frame_loss,
        "lbs": lbs,
        "c1": c1,
    }
    emit(args, out)


def bench_population(args, dev):
    """Batched vs sequential evaluation of a 32-candidate CMA-ES population
    (50k-150k particles each, cube / rod / ellipsoid, own 64^3 grids, 300
    substeps, dt 1e-4), sand and plasticine alternating: forward-only
    fitness and fitness + per-candidate gradients.
    Throughput = sum of particles * 300 / wall time of one evaluation of the
    whole population; the sequential leg runs the same candidates as plain
    rollouts one after another on the same GPU."""
    import torch

    from empm_amd import population as pop
    from empm_amd import scenes

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, r

    out = {"workload": "CMA-ES population: 32 candidates x 50k-150k particles (cube/rod/ellipsoid), own 64^3 grids, "
                       "300 substeps, dt 1e-4"}
    for leg, diff in (("fwd", False), ("fwd_bwd", True)):
        sc = scenes.population_c3()
        nom = scenes.population_c3(nominal=True)
        ev = pop.PopulationEvaluator(sc, dev)
        ev.set_targets([pop.Candidate.of_scene(s) for s in nom])
        cands = [pop.Candidate.of_scene(s) for s in sc]
        work = ev.n * ev.n_steps
        info = {"materials": sorted(set(s.material for s in sc)), "n_particles": ev.n, "n_cand": ev.n_cand}
        # launch parameters tuned on 20 substeps of each leg's own workload
        # (the batched evaluation, and the one-at-a-time rollouts it is
        # compared with); set-up, not timed
        tb_tune = ev.autotune(cands, backward=diff, n_steps=20, cache={}, force=True)
        info["tuning_batched"] = tb_tune["tuning"]
        if not args.pop_batched_only:
            info["tuning_sequential"] = ev.autotune(cands, backward=diff, n_steps=20, sequential=True, cache={},
                                                    force=True)["tuning"]
        log(f"[bench] population {leg} autotune: batched {info['tuning_batched']}, "
            f"one-at-a-time {info.get('tuning_sequential')}")
        if not diff:
            tb, fb = timed(lambda: ev.fitness(cands), args.pop_steps)
            ts, fs = (None, None) if args.pop_batched_only else timed(lambda: ev.fitness_sequential(cands),
                                                                       args.pop_steps)
        else:
            # the shortest checkpoint interval whose tape fits the free HBM
            # (every substep recorded when it fits: no replay)
            free, total = torch.cuda.mem_get_info(dev)
            Ks = [k for k in (1, 2, 3, 4, 5, 6, 10, 15, 20, 30, 50, 100, 300) if ev.tape_bytes(k) <= 0.9 * free]
            tb = None
            for K in Ks:
                try:
                    tb, (fb, gb) = timed(lambda: ev.fitness_and_grad(cands, ckpt_every=K), args.pop_steps)
                    break
                except torch.OutOfMemoryError:
                    log(f"[bench] population fwd_bwd: tape of {ev.tape_bytes(K) / 1e9:.0f} GB at ckpt_every={K} "
                        f"did not fit; next interval")
                    torch.cuda.empty_cache()
            if tb is None:
                raise RuntimeError("no checkpoint interval fits the device memory")
            info["ckpt_every_batched"] = K
            info["ckpt_every_sequential"] = 1
            info["tape_gb_batched"] = ev.tape_bytes(K) / 1e9
            info["hbm_free_gb_at_start"] = free / 1e9
            if not args.pop_batched_only:
                ts, (fs, gs) = timed(lambda: ev.fitness_and_grad_sequential(cands, ckpt_every=1), args.pop_steps)
                gerr = (gb - gs).abs().max(1).values / gs.abs().max(1).values
                info["grad_rel_err_batched_vs_sequential"] = float(gerr.max())
            else:
                ts = None
        info.update({"batched": work / tb, "ms_batched": tb * 1000.0})
        if ts is not None:
            ferr = ((fb - fs).abs() / fs.abs()).max()
            info.update({"sequential": work / ts, "speedup": ts / tb, "ms_sequential": ts * 1000.0,
                         "fitness_rel_err_batched_vs_sequential": float(ferr)})
        log(f"[bench] population {leg}: batched {tb * 1000:.1f} ms ({work / tb:.3e} p-substeps/s); {info}")
        out[leg] = info
        del ev
        torch.cuda.empty_cache()
    out["unit"] = "particle-substeps/s"
    return out


def bench_serving(args, dev):
    """Continuous batching under Poisson arrivals: a seeded mix of rollout
    requests (10k-500k particles, 32^3 / 64^3 / 128^3 grids, 100-2000
    substeps at dt 5e-5, 60% forward only and 40% forward + gradient, jelly /
    Neo-Hookean + von Mises / sand) submitted at --serve-qps into one
    engine; reports p50/p99 latency (arrival to result) and the aggregate
    particle-substeps/s (forward substeps of all requests over the wall
    time)."""
    import torch

    from empm_amd import scenes
    from empm_amd import serving as sv

    mix = scenes.serving_requests_c5(n_requests=args.serve_requests)
    reqs = [sv.RolloutRequest(**{k: v for k, v in m.items() if k != "shape"}) for m in mix]
    eng = sv.ServingEngine(dev, max_particles=8_000_000, max_cand=24, max_grid=128, dt=5e-5, chunk=50)
    # warm-up: one small request of each mode through the engine
    for r in reqs[:2]:
        w = sv.RolloutRequest(**{**r.__dict__, "n_steps": 60})
        eng.submit(w)
    eng.run_until_idle()
    eng.ticks = eng.calls = 0
    torch.cuda.synchronize()
    out = sv.serve_poisson(eng, reqs, args.serve_qps, log=log)
    out["workload"] = (f"{len(reqs)} rollout requests, Poisson arrivals at {args.serve_qps}/s: N log-uniform 10k-500k, "
                       f"grids 32/64/128, 100-2000 substeps (dt 5e-5), 40% with gradient")
    out["unit"] = "particle-substeps/s"
    out["hbm_budget_gb"] = eng.sched.budget / 1e9
    log(f"[bench] serving: {out}")
    eng.close()
    return out


def bench_long_horizon(args, dev):
    """Long-horizon embodied gradient: 8 parameter sets of a 300k-particle
    slab on 96^3 grids pressed and dragged by a box gripper, 3000 substeps
    (30 frames of 100, dt 5e-5), position loss every frame, checkpoint every
    50 substeps with recompute; one step = forward of all sets + reverse
    sweep; throughput = 8 * N * 3000 / step time."""
    import torch

    from empm_amd import long_horizon as lh
    from empm_amd import scenes
    from empm_amd.population import Candidate

    sc = scenes.slab_c4()
    sets = [Candidate(**{k: scenes.sample_params_c2(100 + i)[k] for k in ("E", "nu", "yield_stress")})
            for i in range(8)]
    b = lh.LongHorizonBatch(sc, len(sets), dev, ckpt_every=50)
    b.set_targets([Candidate(**{k: sc.params[k] for k in ("E", "nu", "yield_stress")})] * len(sets))
    b.loss_and_grad(sets)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss, grad = b.loss_and_grad(sets)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    T = sc.n_steps
    out = {"workload": f"8 parameter sets x slab of {sc.n} particles, 96^3 grids, box gripper, {T} substeps "
                       f"(dt 5e-5), position loss every 100 substeps, checkpoint every 50",
           "value": len(sets) * sc.n * T / el, "unit": "particle-substeps/s (fwd+bwd)", "seconds_per_step": el,
           "loss": loss.tolist(), "grad_logE_nu_logY": grad.tolist()}
    log(f"[bench] long horizon: {el:.2f} s per gradient batch, {out['value']:.3e} p-substeps/s")
    return out


def bench_c1(args, dev):
    """The reference's CPU-runnable case at full size: a 32^3 cube of 8
    particles per covered cell (about 7.1k), FCR, 3-cell sticky walls, dt
    1e-4, 200 forward substeps; the GPU rollout (median of 5 after 2
    warm-ups) and the FP32 CPU oracle on 1 and on all host threads (up to
    16)."""
    import torch

    from empm_amd import _lib
    from empm_amd import runtime as rt
    from empm_amd import scenes
    from empm_amd.diff_simulator.warp_mpm.mpm_solver_diff import KIND_SURFACE

    sc = scenes.cube_c1()
    N, n, T, dt = sc.n, sc.n_grid, sc.n_steps, sc.dt

    def t(a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev).clone()

    Et, nut = t(sc.E), t(sc.nu)
    mu, lam = torch.empty_like(Et), torch.empty_like(Et)
    rt.mu_lam(Et, nut, mu, lam)
    model = rt.ModelArrays(material=sc.material, n_grid=n, grid_lim=1.0, dx=float(np.float32(1.0 / n)),
                           inv_dx=float(np.float32(n)), mu=mu, lam=lam, yield_stress=t(sc.yield_stress),
                           mass=t(sc.vol * sc.rho), vol=t(sc.vol), gravity=sc.gravity)

    def cols():
        out = []
        for c in sc.colliders:
            g = _lib.Collider()
            g.kind = KIND_SURFACE
            g.start_time, g.end_time = 0.0, 999.0
            for i in range(3):
                g.point[i] = c["point"][i]
                g.normal[i] = c["normal"][i]
            out.append(g)
        return out

    s0 = rt.ParticleState(t(sc.x), torch.zeros((N, 3), device=dev), torch.zeros((N, 3, 3), device=dev),
                          torch.eye(3, device=dev).repeat(N, 1, 1).contiguous())
    ro = rt.Rollout(rt.Context(n, N, dev), model, N)
    tune = ro.autotune(s0, cols(), dt, n_steps=T, backward=False, cache={}, force=True)
    log(f"[bench] C1 autotune: {tune['tuning']}; {tune['seconds'] * 1e3:.2f} ms per rollout, default "
        f"{tune['default_seconds'] * 1e3:.2f} ms")
    for _ in range(2):
        ro.forward(s0, cols(), 0.0, dt, T)
    torch.cuda.synchronize()
    ms = []
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ro.forward(s0, cols(), 0.0, dt, T)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    med = sorted(ms)[len(ms) // 2]
    # the same rollouts replayed as a HIP graph (one launch per call instead
    # of 8 per substep): the launch-bound regime of a small scene
    gctx = rt.Context(n, N, dev).set_graphs(True).set_tuning(tune["tuning"])
    gro = rt.Rollout(gctx, model, N)
    for _ in range(2):
        gro.forward(s0, cols(), 0.0, dt, T)
    torch.cuda.synchronize()
    gms = []
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        gf, _, _, _ = gro.forward(s0, cols(), 0.0, dt, T)
        e1.record()
        torch.cuda.synchronize()
        gms.append(e0.elapsed_time(e1))
    gmed = sorted(gms)[len(gms) // 2]
    ef, _, _, _ = ro.forward(s0, cols(), 0.0, dt, T)
    gerr = float((gf.x - ef.x).abs().max() / ef.x.abs().max())
    caps, reps = gctx.graph_stats()
    log(f"[bench] C1 graph: {gmed:.2f} ms per rollout (eager {med:.2f}), {caps} captures / {reps} replays, "
        f"rel diff x {gerr:.1e}")
    out = {"workload": f"cube 32^3, {N} particles (8 per covered cell), FCR, 3-cell sticky walls, dt 1e-4, "
                       f"{T} forward substeps", "value": N * T / (med * 1e-3), "unit": "particle-substeps/s (forward)",
           "ms_per_rollout_median": med, "us_per_substep": med * 1e3 / T,
           "graph": {"value": N * T / (gmed * 1e-3), "ms_per_rollout_median": gmed, "us_per_substep": gmed * 1e3 / T,
                     "captures": caps, "replays": reps, "rel_diff_x_vs_eager": gerr},
           "tuning": {"chosen": tune["tuning"], "ms_per_rollout": tune["seconds"] * 1e3,
                      "default_ms_per_rollout": tune["default_seconds"] * 1e3}}
    if not args.no_cpu_baseline:
        from oracle import oracle as orc

        cpu = {}
        mu64, lam64 = orc.mu_lam(sc.E.astype(np.float64), sc.nu.astype(np.float64), "f32")
        P = orc.make_params(N, n, material=sc.material, gravity=sc.gravity)
        ocols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
        for threads in sorted({1, min(os.cpu_count() or 1, 16)}):
            orc.set_threads(threads)
            o = orc.OracleState(sc.x, np.zeros((N, 3)), np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu64,
                                lam64, sc.yield_stress, sc.vol * sc.rho, sc.vol, precision="f32")
            t0 = time.perf_counter()
            orc.rollout(P, ocols, 0.0, dt, T, o)
            cpu[threads] = N * T / (time.perf_counter() - t0)
        out["cpu_baseline"] = {"value": cpu[max(cpu)], "unit": "particle-substeps/s (forward)", "cores": max(cpu),
                               "kind": "port", "value_1_thread": cpu[1],
                               "sample": f"the whole workload ({T} substeps) on the FP32 CPU oracle"}
    log(f"[bench] C1 cube: {med:.2f} ms per {T}-substep rollout on the GPU, {out['value']:.3e} p-substeps/s"
        + (f"; CPU oracle {out['cpu_baseline']['value']:.3e} on {out['cpu_baseline']['cores']} threads, "
           f"{out['cpu_baseline']['value_1_thread']:.3e} on 1" if "cpu_baseline" in out else ""))
    return out


def copy_bandwidth(dev, mib=1024, reps=10):
    """Achieved HBM bandwidth of a plain device copy (read + write bytes per
    second), the measured counterpart of the 8 TB/s published peak."""
    import torch

    src = torch.ones(mib * (1 << 18), device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * mib * (1 << 20) * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def bench_loss(args, dev):
    """The training loss on frames with its gradient
    (phys_sim/engine/trainer_warp_mpm.py:727-766): 30 frames of 20k
    simulated points against 20k observed points, 80% of them visible and 90%
    with a valid motion per frame; one step = Chamfer + tracking summed over
    the frames and the backward to the points.  The nearest-neighbour kernel
    (k_nn_sqdist, brute force, two directions per frame) is VALU-bound:
    9 operations per pair (3 differences, 3 multiply-adds, compare, 2 selects)."""
    import torch

    from empm_amd import _lib
    from empm_amd import loss
    from empm_amd.runtime import stream_handle

    T, N = 30, 20_000
    rng = np.random.default_rng(7)
    pts = rng.random((T, N, 3)) * 0.2
    tgt = pts + 0.005 * rng.standard_normal((T, N, 3))
    vis = rng.random((T, N)) < 0.8
    mv = rng.random((T, N)) < 0.9
    tp = torch.tensor(pts, dtype=torch.float32, device=dev, requires_grad=True)
    tt = torch.tensor(tgt, dtype=torch.float32, device=dev)
    tv = torch.tensor(vis, device=dev)
    tm = torch.tensor(mv, device=dev)

    def step():
        tp.grad = None
        L = loss.frame_loss(tp, tt, tv, tm)
        L.backward()
        return L

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        L = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    # the kernel alone on one frame's visible points, HIP events on its stream
    a = tp[0][tv[0]].detach().contiguous()
    b = tt[0][tv[0]].contiguous()
    n = a.shape[0]
    dist = torch.empty(n, dtype=torch.float32, device=dev)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    lib = _lib.load()
    st = stream_handle(dev)
    ws_bytes = int(lib.empm_nn_sqdist_workspace(n, n))
    ws = torch.empty(max(ws_bytes // 4, 1), dtype=torch.int32, device=dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(50)]
    for e0, e1 in ev:
        e0.record()
        _lib.check(lib.empm_nn_sqdist(n, a.data_ptr(), n, b.data_ptr(), dist.data_ptr(), idx.data_ptr(),
                                      ws.data_ptr(), ws_bytes, st))
        e1.record()
    torch.cuda.synchronize()
    us = sum(e0.elapsed_time(e1) for e0, e1 in ev) / len(ev) * 1e3
    tflops = 9.0 * n * n / (us * 1e-6) / 1e12
    out = {"workload": f"frame loss (Chamfer + tracking) and its gradient: {T} frames of {N} simulated points vs "
                       f"{N} observed points, 80% visible, 90% with valid motion",
           "value": T / (ms * 1e-3), "unit": "frames/s (loss + gradient)", "ms_per_step": ms, "loss": float(L.detach()),
           "roofline": {"bound": "valu", "kernel": "k_nn_sqdist", "achieved": tflops, "peak": 157.3,
                        "unit": "TFLOP/s", "frac": tflops / 157.3, "traffic": None,
                        "note": f"{n} x {n} pairs, 9 FP32 operations per pair", "avg_launch_us": us}}
    if not args.no_cpu_baseline:
        from oracle import chamfer as och

        # a 4000-point subsample of each frame's visible points (the brute
        # force is quadratic: pairs/s scale to the full frame's 2 n^2 pairs)
        sub, f, pairs = 4000, 0, 0
        t0 = time.perf_counter()
        while f < T and time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            x, y = pts[f][vis[f]][:sub], tgt[f][vis[f]][:sub]
            och.chamfer(x, y)
            pairs += 2 * len(x) * len(y)
            f += 1
        el = time.perf_counter() - t0
        frame_pairs = float(np.mean([2 * int(v.sum()) ** 2 for v in vis]))
        out["cpu_baseline"] = {"value": pairs / el / frame_pairs * 1.0, "unit": "frames/s (Chamfer value)",
                               "cores": 1, "kind": "port",
                               "sample": f"{f} frames cut to {sub} visible points each ({pairs:.3g} point pairs), "
                                         f"float64 NumPy brute force, {el:.1f} s, scaled by pairs to full frames"}
    log(f"[bench] loss: {ms:.2f} ms per {T}-frame loss + gradient; k_nn_sqdist {us:.1f} us for {n}x{n} "
        f"({tflops:.1f} TFLOP/s)")
    return out


def _morton(p, lo, hi, bits=10):
    """Morton (Z-order) keys of points p in the box [lo, hi]."""
    q = np.clip(((p - lo) / (hi - lo) * (1 << bits)).astype(np.int64), 0, (1 << bits) - 1)
    key = np.zeros(len(p), np.int64)
    for b in range(bits):
        for a in range(3):
            key |= ((q[:, a] >> b) & 1) << (3 * b + a)
    return key


def bench_lbs(args, dev):
    """Per-frame motion interpolation of Gaussians on the simulated particles
    (linear blend skinning, third_party/gaussian_splatting/dynamic_utils.py:305,
    once per frame in scripts/interactive_viser.py:259-280): 50k particles as
    bones, 1M Gaussians, K = 16 bones per Gaussian and 16 neighbours per bone,
    a smooth non-rigid motion.  One frame = the bone transforms (k_lbs_bones)
    + the per-Gaussian blend (k_lbs_points) with the inverse-distance weights
    computed in the blend; the reference's two-call sequence (weights, then
    interpolation with the weights) is timed beside it.  Measured twice: with
    particles and Gaussians in random order (every bone-record gather a cache
    miss; `value`) and in Z-order (as a grid-filling particle sampler and a
    splat initialised from a scanned point cloud store them)."""
    import torch

    from empm_amd import _lib
    from empm_amd import gaussian_splatting as gs
    from empm_amd.runtime import stream_handle

    nb, ng, K = 50_000, 1_000_000, 16
    rng = np.random.default_rng(0)
    bones = rng.random((nb, 3)) * 0.2
    xyz = bones[rng.integers(0, nb, ng)] + 0.004 * rng.standard_normal((ng, 3))
    quat = rng.standard_normal((ng, 4))
    lib = _lib.load()
    st = stream_handle(dev)
    frames = 200

    def motions_of(b):
        c = np.full(3, 0.1)
        ang = 0.1
        Q = np.array([[np.cos(ang), -np.sin(ang), 0.0], [np.sin(ang), np.cos(ang), 0.0], [0.0, 0.0, 1.0]])
        x, y, z = (b - c).T
        return (b - c) @ Q.T + c - b + 0.01 * np.stack([np.sin(20 * y), np.sin(20 * z), np.sin(20 * x)], 1)

    def t(a):
        return torch.tensor(a, dtype=torch.float32, device=dev)

    def timed(fn):
        for _ in range(10):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / frames

    def case(bones, xyz, quat):
        tb, tm, tx, tq = t(bones), t(motions_of(bones)), t(xyz), t(quat)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rel = gs.get_topk_indices(tb, K=K)
        _, idx = gs.knn_weights_sparse(tb, tx, K=K)
        torch.cuda.synchronize()
        setup_ms = (time.perf_counter() - t0) * 1e3
        interp = gs.LBSInterpolator()
        xo, qo = torch.empty_like(tx), torch.empty_like(tq)

        def fused():
            interp(tb, tm, rel, tx, quat=tq, weights=None, weights_indices=idx, xyz_out=xo, quat_out=qo)

        def two_call():
            w = gs.calc_weights_vals_from_indices(tb, tx, idx)
            gs.interpolate_motions_speedup(bones=tb, motions=tm, relations=rel, weights=w, weights_indices=idx,
                                           xyz=tx, quat=tq)

        ms_fused = timed(fused)
        ms_two = timed(two_call)
        # the blend kernel alone, HIP events on its stream around each launch
        rec, degen = interp.bones(tb, tm, rel)
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(frames)]
        for e0, e1 in ev:
            e0.record()
            _lib.check(lib.empm_lbs_points(ng, tx.data_ptr(), tq.data_ptr(), K, idx.data_ptr(), 4, None, nb,
                                           rec.data_ptr(), degen.data_ptr(), xo.data_ptr(), qo.data_ptr(), st))
            e1.record()
        torch.cuda.synchronize()
        us = sorted(e0.elapsed_time(e1) * 1e3 for e0, e1 in ev)
        return {"ms_per_frame": ms_fused, "ms_per_frame_two_call": ms_two, "setup_ms_knn": setup_ms,
                "avg_launch_us": sum(us) / len(us), "median_launch_us": us[len(us) // 2]}, rel, idx

    rnd, rel, idx = case(bones, xyz, quat)
    ob = np.argsort(_morton(bones, 0.0, 0.2), kind="stable")
    og = np.argsort(_morton(xyz, -0.02, 0.22), kind="stable")
    zo, _, _ = case(bones[ob], xyz[og], quat[og])
    # per Gaussian: x (12) and quaternion (16) read, K int32 indices (64), x' (12) and q' (16) written;
    # per bone: its 64-byte record, read once
    algo = ng * (12 + 16 + 4 * K + 12 + 16) + nb * 64
    achieved = algo / (rnd["avg_launch_us"] * 1e-6) / 1e9
    out = {"workload": f"linear-blend-skinning of {ng} Gaussians on {nb} particles, K={K} bones per Gaussian, "
                       f"{K} neighbours per bone, per-frame bone transforms + blend; random storage order",
           "value": ng / (rnd["ms_per_frame"] * 1e-3), "unit": "Gaussians/s (per-frame interpolation)", **rnd,
           "roofline": {"bound": "hbm", "kernel": "k_lbs_points", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": measured_traffic("k_lbs_points", TRAFFIC_PROFILE_LBS),
                        "traffic_source": TRAFFIC_PROFILE_LBS, "algorithmic_bytes": algo,
                        "avg_launch_us": rnd["avg_launch_us"], "launches": frames},
           "z_order": {"value": ng / (zo["ms_per_frame"] * 1e-3), **zo,
                       "roofline_frac": algo / (zo["avg_launch_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS}}
    if not args.no_cpu_baseline:
        from oracle import lbs as olbs

        rel_np = rel.long().cpu().numpy()
        idx_np = idx.long().cpu().numpy()
        mo = motions_of(bones)
        n_cpu, t0 = 0, time.perf_counter()
        while n_cpu < ng and time.perf_counter() - t0 < args.cpu_seconds:
            sl = slice(n_cpu, n_cpu + 100_000)
            olbs.interpolate(bones, mo, rel_np, xyz[sl], quat[sl], None, idx_np[sl])
            n_cpu += len(xyz[sl])
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n_cpu / el, "unit": "Gaussians/s", "cores": 1, "kind": "port",
                               "sample": f"{n_cpu} of the Gaussians in chunks of 100000 (bone transforms of all "
                                         f"{nb} particles per chunk), float64 NumPy oracle, {el:.1f} s"}
    log(f"[bench] lbs: {rnd['ms_per_frame']:.3f} ms/frame fused ({zo['ms_per_frame']:.3f} in Z-order), "
        f"{rnd['ms_per_frame_two_call']:.3f} two-call, setup {rnd['setup_ms_knn']:.1f} ms; k_lbs_points avg "
        f"{rnd['avg_launch_us']:.1f} us ({achieved:.0f} GB/s), Z-order {zo['avg_launch_us']:.1f} us")
    return out


def cpu_baseline(sc, seconds):
    """FP32 CPU oracle (a C restatement of the reference substep, OpenMP) on
    the same workload's forward substeps, on the host cores."""
    from oracle import oracle as orc

    cores = min(os.cpu_count() or 1, 16)
    orc.set_threads(cores)
    N = sc.n
    E, nu = sc.params["E"], sc.params["nu"]
    mu = np.full(N, E / (2 * (1 + nu)))
    lam = np.full(N, E * nu / ((1 + nu) * (1 - 2 * nu)))
    P = orc.make_params(N, sc.n_grid, material=sc.material, gravity=sc.gravity)
    cols = [orc.surface_collider(c["point"], c["normal"]) for c in sc.colliders]
    s = orc.OracleState(sc.x, sc.params["v0"], np.zeros((N, 3, 3)), np.tile(np.eye(3), (N, 1, 1)), mu, lam,
                        sc.yield_stress, sc.vol * sc.rho, sc.vol, precision="f32")
    t0 = time.perf_counter()
    orc.rollout(P, cols, 0.0, sc.dt, 1, s)
    one = time.perf_counter() - t0
    k = int(max(1, min(200, seconds / max(one, 1e-6))))
    t0 = time.perf_counter()
    orc.rollout(P, cols, sc.dt, sc.dt, k, s)
    el = time.perf_counter() - t0
    return {"value": N * k / el, "unit": "particle-substeps/s (forward only)", "cores": cores, "kind": "port",
            "sample": f"{k} forward substeps of the {N}-particle 64^3 workload (substeps 2..{k + 1}) "
                      f"on the FP32 CPU oracle, {el:.1f} s"}


if __name__ == "__main__":
    main()
