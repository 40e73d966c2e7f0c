"""Benchmark: the forced-dissipative gSQG RK4 step with two passive scalars.

Workload (the configuration the headline metric is quoted on): N=1024,
alpha=1, FP64; random-phase k^-5/3 initial theta (seed 2, rms 1); ring
forcing kf=40, dk=2, eps=0.1 with per-step keys from seed 42; CombinedDamping
mu=0.1 (|k| < 20), nu_p = 1e-8*(512/N)^8, p=8; solver nu_p = 0; two Gaussian
passive scalars (sigma = 2pi/16); dt = 5e-4.  One "step" = one RK4 step of
theta plus both scalars.  The hyperviscous term is integrated with the
integrating-factor scheme: the explicit scheme is unstable at this nu (the
timed state would be non-finite within tens of steps).

Prints one JSON line; see the keys below.

``--config`` selects the other configurations of the reference's benchmark
set (each its own JSON line, same keys):
  plain  N=256, alpha=1, nu_p=1e-4, p=8, theta0 = sin(4x)cos(4y), dt=1e-3
         (integrating-factor scheme: explicit RK4 is unstable at this nu);
  large  N=8192, alpha=0.5, k^-5/3 theta0, ring kf=100 dk=2 eps=0.1,
         mu=0.1, nu_p=1e-8*(512/N)^8, p=8, 4 Gaussian scalars, dt=1e-4;
  batch  96 members, 32 each at N=128, 256, 512, alpha cycling
         {0, 0.5, 1, 1.5, 2}, seeded k^-5/3 theta0, nu_p=1e-4, p=8, dt=1e-3,
         against looping the single-realization solver over the members;
  closure  learned subgrid closure: forced DNS at N=4096 (alpha=1, ring
         kf=40 dk=2 eps=0.1, mu=0.1, nu_p=1e-8*(512/N)^8, p=8, dt=1e-4),
         snapshots sharp-filtered to N=512 with the filtered-minus-resolved
         Jacobian flux as target, the CNN trained on the GPU (batch 32, Adam
         lr=1e-3), then the closed N=512 run against the filtered DNS and the
         unclosed N=512 run: shell-spectrum error for k <= 170 and wall-clock
         per simulated time unit.  --nf/--nc/--snaps/--train-steps/--t-eval
         shrink it for quick checks;
  closure-decay  the learned closure on decaying turbulence, in about a
         minute: eight seeded k^-5/3 realizations at N=512 (alpha=1,
         nu_p=1e-15, p=8, dt=1e-3), 12 snapshots each filtered to N=256,
         training on six, then closed and unclosed N=256 runs from each
         held-out realization against the filtered DNS over one time unit
         (the gain the closure tests check, with wall-clocks; the coarse
         runs take dt=2e-3, the DNS's CFL number on the coarser grid).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12  # B/s, MI355X HBM3E


def step_bytes(N: int, n_s: int) -> float:
    """Algorithmic HBM bytes of one RK4 step (3-pass 2D-FFT model):
    60h + n_s (44h + 8r) + [n_s > 0] (5h + 2r), h = 16 N (N/2+1), r = 8 N^2."""
    h = 16.0 * N * (N // 2 + 1)
    r = 8.0 * N * N
    return 60 * h + n_s * (44 * h + 8 * r) + (5 * h + 2 * r if n_s > 0 else 0.0)


def launch_bytes(N: int, n_s: int) -> dict:
    """Algorithmic bytes per launch of each kernel class: the theta column
    kernel is pass 3 of one stage plus pass 1 of the next (5h + 5h), the row
    kernel pass 2 (5h); per scalar the column kernel moves 5h + 3h and the row
    kernel 3h + 2r."""
    h = 16.0 * N * (N // 2 + 1)
    r = 8.0 * N * N
    return {"theta_column": 10 * h, "theta_row": 5 * h, "scalar_column": 8 * h * n_s,
            "scalar_row": (3 * h + 2 * r) * n_s}


def read_traffic(kernel: str):
    """HBM bytes per launch measured with rocprofv3 PMC counters, from
    profiles/pmc_traffic.txt (lines 'kernel_class bytes_per_launch')."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 2 and parts[0] == kernel:
                    return float(parts[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=1024)
    ap.add_argument("--scheme", default="if", choices=["if", "explicit"])
    ap.add_argument("--cpu-steps", type=int, default=6)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-closure", action="store_true",
                    help="default configuration: skip the learned-closure measurement that the "
                         "JSON line carries as 'closure' (run after the timed region)")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tile shapes (forced configuration)")
    ap.add_argument("--shapes", default=None, metavar="COL,ROW",
                    help="set these column/row tile shapes instead of autotuning (forced, "
                         "plain and large configurations; profiles reuse the shapes a timed "
                         "run picked)")
    ap.add_argument("--copy-only", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates PMC byte counters)")
    ap.add_argument("--config", default="forced",
                    choices=["forced", "plain", "large", "batch", "closure", "closure-decay",
                             "cnn"])
    # closure workload sizes
    ap.add_argument("--nf", type=int, default=4096)
    ap.add_argument("--nc", type=int, default=512)
    ap.add_argument("--snaps", type=int, default=2000)
    ap.add_argument("--snap-every", type=int, default=5)
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--t-spin", type=float, default=0.5)
    ap.add_argument("--t-eval", type=float, default=10.0)
    ap.add_argument("--dt-coarse", default=None,
                    help="coarse time step (closure: 1e-4, closure-decay: 2e-3); a comma list "
                         "evaluates each (closure; the first is reported)")
    ap.add_argument("--nu", type=float, default=None,
                    help="hyperviscosity (default 1e-8 * (512 / nf)^8)")
    ap.add_argument("--coarse-seeds", default="43",
                    help="forcing seeds of the coarse runs; extra seeds add unclosed runs that "
                         "measure the spread between forcing realizations")
    ap.add_argument("--no-match-forcing", dest="match_forcing", action="store_false",
                    help="coarse forcing amplitude sqrt(eps) instead of sqrt(eps) (nc/nf)^2")
    ap.add_argument("--closure-every", default=None,
                    help="closure evaluated every M steps (0: every RK4 stage; closure: 8, "
                         "closure-decay: 4,0,8); a comma list evaluates each (the first is "
                         "reported)")
    args = ap.parse_args()
    if args.copy_only:
        import torch

        n = args.copy_only * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is "
              "not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.config == "batch":
        return bench_batch(args)
    if args.config == "closure":
        return bench_closure(args)
    if args.config == "closure-decay":
        return bench_closure_decay(args)
    if args.config == "cnn":
        return bench_cnn(args)
    if args.config in ("plain", "large"):
        return bench_single(args)

    import numpy as np
    import torch

    import pygsquig_amd as pg
    from pygsquig_amd import _lib, ics
    from pygsquig_amd.scalars import gSQGSolverWithScalars

    N, dt, n_s = args.N, 5e-4, 2
    L = 2 * np.pi
    dev = torch.device("cuda:0")
    g = pg.make_grid(N, L, dev)
    forcing = pg.RingForcing(kf=40, dk=2, epsilon=0.1, seed=42)
    damping = pg.CombinedDamping(mu=0.1, kf=40, nu_p=1e-8 * (512 / N) ** 8, p=8)
    solver = gSQGSolverWithScalars(g, 1.0, 0.0, 8, scheme=args.scheme,
                                   passive_scalars={f"s{i}": {"kappa": 0.0} for i in range(n_s)})
    theta0 = torch.from_numpy(ics.kolmogorov_53(N, seed=2)).to(dev)
    init = {f"s{i}": torch.from_numpy(ics.gaussian_blob(N, c, L / 16)).to(dev)
            for i, c in enumerate(ics.SCALAR_CENTRES[:n_s])}
    state = solver.initialize(theta0, scalar_init=init)
    tuned = None
    if args.shapes:
        solver._plan(forcing, damping).set_shape(*(int(x) for x in args.shapes.split(",")))
    elif not args.no_autotune:  # tile shapes picked by timing every pair (untimed here)
        tuned = solver.autotune(dt, forcing=forcing, damping=damping)
    # the replay graphs of this workload, captured before warm-up and timing
    solver.capture_graphs(dt, forcing=forcing, damping=damping)
    state = solver.multistep(state, args.warmup, dt, forcing=forcing, damping=damping)
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    state = solver.multistep(state, args.steps, dt, forcing=forcing, damping=damping)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = e0.elapsed_time(e1) / 1e3
    finite = bool(torch.isfinite(state["theta_hat"]).all())
    sec = max(wall, gpu_s)
    steps_per_s = args.steps / sec
    value = N * N * steps_per_s
    B = step_bytes(N, n_s)

    # Live per-kernel timing (HIP events around every launch, no graph).
    plan = solver._plan(forcing, damping)
    prof = plan.profile_advance(dt, args.profile_steps, _lib.KEY_DERIVED, forcing.seed,
                                int(state["step"]))
    lb = launch_bytes(N, n_s)
    dom = max((k for k in lb), key=lambda k: prof[k][0])
    dom_ms, dom_n = prof[dom]
    avg_s = dom_ms / 1e3 / max(dom_n, 1)
    achieved = lb[dom] / avg_s / 1e9
    kernel_table = {k: {"avg_us": round(prof[k][0] / max(prof[k][1], 1) * 1e3, 2),
                        "launches": prof[k][1],
                        "GBps": round(lb[k] / (prof[k][0] / 1e3 / max(prof[k][1], 1)) / 1e9, 1)
                        if prof[k][1] else None}
                    for k in lb}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import cpu_baseline

        r = cpu_baseline.time_forced_scalars(N=N, n_steps=args.cpu_steps,
                                             workers=cpu_baseline.host_cores(), n_scalars=n_s,
                                             scheme=args.scheme, budget_s=25.0)
        # the same path on one FFT worker (a shorter sample)
        r1 = cpu_baseline.time_forced_scalars(N=N, n_steps=2, workers=1, n_scalars=n_s,
                                              scheme=args.scheme, budget_s=10.0)
        cpu = {"value": N * N / r["sec_per_step"], "unit": "grid-point RK4 steps/s",
               "cores": r["workers"], "kind": "port",
               "sample": f"{r['steps']} timed steps (after 1 warm-up) of the same N={N} forced "
                         f"configuration with {n_s} scalars, NumPy/SciPy reference-structured "
                         f"c2c path, scipy.fft workers={r['workers']}, CPU {r['cpu']}",
               "sec_per_step": r["sec_per_step"],
               "single_worker": {"value": N * N / r1["sec_per_step"], "cores": 1,
                                 "steps": r1["steps"], "sec_per_step": r1["sec_per_step"]}}

    out = {
        "metric": "N^2*RK4-steps/s + HBM GB/s at N=1024 fp64",
        "value": value,
        "unit": "grid-point RK4 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"forced-dissipative gSQG N={N}, 2 passive scalars, dt=5e-4",
                   "N": N, "alpha": 1.0, "scheme": args.scheme, "forcing": "ring kf=40 dk=2 eps=0.1",
                   "damping": "mu=0.1, nu_p=1e-8*(512/N)^8, p=8", "scalars": n_s,
                   "steps_per_s": steps_per_s, "finite": finite},
        "hbm_GBps_algorithmic": B * steps_per_s / 1e9,
        "step_roofline": {"bound": "hbm", "achieved": B * steps_per_s / 1e9, "peak": PEAK_HBM / 1e9,
                          "unit": "GB/s", "frac": B * steps_per_s / PEAK_HBM,
                          "bytes_per_step": B},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved * 1e9 / PEAK_HBM,
                     "traffic": read_traffic(dom), "bytes_per_launch": lb[dom],
                     "avg_launch_us": avg_s * 1e6},
        "kernels": kernel_table,
        "autotune": None if tuned is None else {
            "shapes": list(tuned[0]),
            "ms_per_pair": {f"{c},{r}": round(ms, 4) for (c, r), ms in tuned[1].items()}},
        "shapes": list(plan.shape()[:2]),
        "shape_radix": [plan.shape_radix(k) for k in plan.shape()[:2]],
        "cpu_baseline": cpu,
    }
    if not args.no_closure:
        # the metric's second half, measured in the same run after the timed
        # region: the learned closure's wall-clock gain over the DNS at the
        # spectrum error it keeps (closure-decay workload, every 4 steps)
        cd = bench_closure_decay(args, emit=False)
        out["closure"] = {
            "metric": cd["metric"], "value": cd["value"], "unit": cd["unit"],
            "workload": cd["config"]["workload"],
            "speedup": cd["speedup"], "wall_clock_s": cd["wall_clock_s"],
            "mean_log_E_ratio_vs_dns": cd["mean_log_E_ratio_vs_dns"],
            "training_s": cd["training"]["seconds"],
            "full_config": "DNS N=4096 -> closed N=512: bench.py --config closure "
                           "(minutes; not part of this run)"}
    print(json.dumps(out))


def _timed(fn, steps):
    """Run fn() once between HIP events and wall clocks; seconds (the larger)."""
    import torch

    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    out = fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return out, max(time.perf_counter() - t0, e0.elapsed_time(e1) / 1e3)


def bench_single(args):
    """Configurations 'plain' (N=256) and 'large' (N=8192)."""
    import numpy as np
    import torch

    import pygsquig_amd as pg
    from pygsquig_amd import ics
    from pygsquig_amd.scalars import gSQGSolverWithScalars

    dev = torch.device("cuda:0")
    if args.config == "plain":
        N, alpha, dt, n_s = 256, 1.0, 1e-3, 0
        steps = args.steps if args.steps != 2000 else 1000
        g = pg.make_grid(N, 2 * np.pi, dev)
        solver = pg.gSQGSolverOptimized(g, alpha, 1e-4, 8, scheme="if")
        st = solver.initialize(torch.from_numpy(ics.sin4(N)).to(dev))
        run = lambda s, n: solver.multistep(s, n, dt)
        capture = lambda: solver.capture_graphs(dt)
        workload = "plain unforced gSQG N=256, theta0 = sin(4x)cos(4y), nu_p=1e-4, p=8, dt=1e-3"
    else:
        N, alpha, dt, n_s = 8192, 0.5, 1e-4, 4
        steps = args.steps if args.steps != 2000 else 20
        g = pg.make_grid(N, 2 * np.pi, dev)
        solver = gSQGSolverWithScalars(g, alpha, 0.0, 8, scheme="if",
                                       passive_scalars={f"s{i}": {"kappa": 0.0} for i in range(n_s)})
        f = pg.RingForcing(kf=100, dk=2, epsilon=0.1, seed=42)
        d = pg.CombinedDamping(mu=0.1, kf=100, nu_p=1e-8 * (512 / N) ** 8, p=8)
        init = {f"s{i}": torch.from_numpy(ics.gaussian_blob(N, c, 2 * np.pi / 16)).to(dev)
                for i, c in enumerate(ics.SCALAR_CENTRES[:n_s])}
        st = solver.initialize(torch.from_numpy(ics.kolmogorov_53(N, seed=2, alpha=alpha)).to(dev),
                               scalar_init=init)
        run = lambda s, n: solver.multistep(s, n, dt, forcing=f, damping=d)
        capture = lambda: solver.capture_graphs(dt, forcing=f, damping=d)
        workload = ("large forced gSQG N=8192, alpha=0.5, ring kf=100, 4 passive scalars, "
                    "dt=1e-4")
    warm = min(args.warmup, 5) if args.config == "large" else args.warmup
    tuned = None
    if args.shapes:
        plan0 = solver._plan(None, None) if args.config == "plain" else solver._plan(f, d)
        plan0.set_shape(*(int(x) for x in args.shapes.split(",")))
    elif not args.no_autotune:  # shapes picked by timing every pair (untimed here)
        if args.config == "plain":
            tuned = solver.autotune(dt)
        else:  # N=8192: radix 8 and radix 16, a few steps per pair
            tuned = solver.autotune(dt, forcing=f, damping=d, n_steps=4, repeats=1)
    capture()  # replay graphs captured before warm-up and timing
    st = run(st, warm)
    st, sec = _timed(lambda: run(st, steps), steps)
    finite = bool(torch.isfinite(st["theta_hat"]).all())
    sps = steps / sec
    B = step_bytes(N, n_s)
    # live per-kernel timing (HIP events at the kernel boundaries, no graph)
    from pygsquig_amd import _lib

    if args.config == "plain":
        plan = solver._plan(None, None)
        prof = plan.profile_advance(dt, 20, _lib.KEY_EXPLICIT, 0, 0)
    else:
        plan = solver._plan(f, d)
        prof = plan.profile_advance(dt, 2, _lib.KEY_DERIVED, f.seed, int(st["step"]))
    lb = launch_bytes(N, n_s)
    kernels = {k: {"avg_us": round(ms / cnt * 1e3, 2), "launches": cnt,
                   "GBps": round(lb[k] / (ms / 1e3 / cnt) / 1e9, 1) if k in lb else None}
               for k, (ms, cnt) in prof.items() if cnt}
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import cpu_baseline

        nc = cpu_baseline.host_cores()
        if args.config == "plain":
            r = cpu_baseline.time_plain(N=N, n_steps=100, workers=nc)
            sample = (f"{r['steps']} timed steps (after 1 warm-up) of the same N={N} plain "
                      f"configuration, NumPy/SciPy reference-structured c2c path, IF scheme, "
                      f"scipy.fft workers={nc}, CPU {r['cpu']}")
        else:
            r = cpu_baseline.time_large_rhs(N=N, workers=nc, n_scalars=n_s, alpha=alpha)
            sample = (f"one timed evaluation each of the theta right-hand side with ring "
                      f"forcing and damping ({r['t_theta_rhs']:.1f} s), a scalar right-hand side "
                      f"({r['t_scalar_rhs']:.1f} s) and the velocity ({r['t_velocity']:.1f} s) at "
                      f"N={N}, summed as 4 (theta + {n_s} scalars) + velocity per step; "
                      f"NumPy/SciPy reference-structured c2c path, scipy.fft workers={nc}, "
                      f"CPU {r['cpu']}")
        cpu = {"value": N * N / r["sec_per_step"], "unit": "grid-point RK4 steps/s",
               "cores": nc, "kind": "port", "sample": sample, "sec_per_step": r["sec_per_step"]}
    print(json.dumps({
        "metric": "N^2*RK4-steps/s + HBM GB/s fp64", "value": N * N * sps,
        "unit": "grid-point RK4 steps/s", "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": sec / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "N": N, "alpha": alpha, "scheme": "if", "scalars": n_s,
                   "steps_per_s": sps, "finite": finite},
        "step_roofline": {"bound": "hbm", "achieved": B * sps / 1e9, "peak": PEAK_HBM / 1e9,
                          "unit": "GB/s", "frac": B * sps / PEAK_HBM, "bytes_per_step": B},
        "autotune": None if tuned is None else {
            "shapes": list(tuned[0]),
            "ms_per_pair": {f"{c},{r}": round(ms, 4) for (c, r), ms in tuned[1].items()}},
        "shape_radix": [plan.shape_radix(k) for k in plan.shape()[:2]],
        "kernels": kernels,
        "cpu_baseline": cpu}))


def bench_batch(args):
    """96 uneven members in batched plans, against looping the
    single-realization solver over the same members on the same GPU."""
    import numpy as np
    import torch

    import pygsquig_amd as pg
    from pygsquig_amd import ics

    dev = torch.device("cuda:0")
    alphas = (0.0, 0.5, 1.0, 1.5, 2.0)
    members = [(N, alphas[m % 5]) for m, N in enumerate([128] * 32 + [256] * 32 + [512] * 32)]
    dt, nu = 1e-3, 1e-4
    steps = args.steps if args.steps != 2000 else 500
    thetas = [torch.from_numpy(ics.kolmogorov_53(N, seed=100 + m, alpha=a)).to(dev)
              for m, (N, a) in enumerate(members)]
    ens = pg.EnsembleSolver(members, nu_p=nu, p=8, scheme="if", device=dev)
    st0 = ens.initialize(thetas)
    # default tile shapes, then the autotuner's choice (tuning is untimed)
    st = ens.multistep(st0, args.warmup, dt)
    st_d, sec_d = _timed(lambda: ens.multistep(st, steps, dt), steps)
    tuned = ens.autotune(dt)
    st = ens.multistep(st0, args.warmup, dt)
    st, sec = _timed(lambda: ens.multistep(st, steps, dt), steps)
    same = all(bool(torch.equal(a, b)) for a, b in zip(st["theta_hat"], st_d["theta_hat"]))
    # one column and one row launch per RK4 stage for all three grid sizes
    # (default radix-8 tiles; equal to the default-shape streams run to rounding)
    ensb = pg.EnsembleSolver(members, nu_p=nu, p=8, scheme="if", device=dev, launch="batch")
    stb = ensb.multistep(ensb.initialize(thetas), args.warmup, dt)
    stb, sec_b = _timed(lambda: ensb.multistep(stb, steps, dt), steps)
    rel_b = max(float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
                for a, b in zip(stb["theta_hat"], st_d["theta_hat"]))
    del ensb
    pts = sum(N * N for N, _ in members)
    launch = "batch" if sec_b < sec else "streams"
    if launch == "batch":
        st, sec = stb, sec_b
    finite = all(bool(torch.isfinite(t).all()) for t in st["theta_hat"])
    value = pts * steps / sec
    # baseline: the single-realization solver looped over the members
    solvers = {}
    states = []
    for (N, a), th in zip(members, thetas):
        key = (N, a)
        if key not in solvers:
            solvers[key] = pg.gSQGSolverOptimized(pg.make_grid(N, 2 * np.pi, dev), a, nu, 8,
                                                  scheme="if")
        states.append(solvers[key].initialize(th))
    states = [solvers[k].multistep(s, args.warmup, dt) for k, s in zip(members, states)]
    loop_steps = max(steps // 5, 1)
    _, lsec = _timed(lambda: [solvers[k].multistep(s, loop_steps, dt)
                              for k, s in zip(members, states)], loop_steps)
    loop_value = pts * loop_steps / lsec
    B = sum(step_bytes(N, 0) for N, _ in members)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import cpu_baseline

        nc = cpu_baseline.host_cores()
        r = cpu_baseline.time_batch(members, n_steps=2, workers=nc, nu=nu, dt=dt)
        cpu = {"value": pts / r["sec_per_step"], "unit": "member-grid-point RK4 steps/s",
               "cores": nc, "kind": "port", "sec_per_step": r["sec_per_step"],
               "sample": f"{r['steps']} timed steps (after 1 warm-up) of the oracle solver "
                         f"looped over the same {len(members)} members, NumPy/SciPy "
                         f"reference-structured c2c path, IF scheme, scipy.fft workers={nc}, "
                         f"CPU {r['cpu']}"}
    print(json.dumps({
        "metric": "member-grid-point RK4 steps/s (batched uneven ensemble)", "value": value,
        "unit": "member-grid-point RK4 steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": sec / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "96 members: 32 each at N=128, 256, 512; alpha cycling "
                               "{0, 0.5, 1, 1.5, 2}; nu_p=1e-4, p=8, dt=1e-3, IF scheme",
                   "members": len(members), "finite": finite, "launch": launch},
        "autotune": {"shapes": {str(N): list(b) for N, (b, _) in tuned.items()},
                     "ms_per_pair": {str(N): {f"{c},{r}": round(ms, 4) for (c, r), ms in t.items()}
                                     for N, (_, t) in tuned.items()},
                     "default_shapes_value": pts * steps / sec_d,
                     "identical_to_default": same},
        "one_launch": {"value": pts * steps / sec_b, "ms_per_step": sec_b / steps * 1e3,
                       "max_rel_diff_vs_default_streams": rel_b,
                       "note": "one column and one row launch per RK4 stage for all sizes"},
        "launch_choice": "the faster of the autotuned per-size streams and the one-launch batch",
        "step_roofline": {"bound": "hbm", "achieved": B * steps / sec / 1e9,
                          "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                          "frac": B * steps / sec / PEAK_HBM, "bytes_per_step": B},
        "loop_baseline": {"value": loop_value, "unit": "member-grid-point RK4 steps/s",
                          "steps": loop_steps, "speedup": value / loop_value},
        "cpu_baseline": cpu}))


def bench_cnn(args):
    """The closure network alone on nc x nc images: the forward pass of one
    image (what a closed solver step evaluates) and training steps on a
    batch of 32 random fields."""
    import torch

    from pygsquig_amd.closure import ClosureNet

    dev = torch.device("cuda:0")
    H = args.nc
    B = 32
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    th, u, v, tg = (torch.randn((B + 1, H, H), dtype=torch.float64, device=dev, generator=g)
                    for _ in range(4))
    net = ClosureNet(H, H, max_batch=B, device=dev, seed=0)
    sel = torch.arange(B, dtype=torch.int32, device=dev)
    for _ in range(3):
        net.forward(th[:1], u[:1], v[:1], (1, 1, 1))
    n_fwd = 20
    _, sf = _timed(lambda: [net.forward(th[:1], u[:1], v[:1], (1, 1, 1)) for _ in range(n_fwd)],
                   n_fwd)
    net.train_step(th, u, v, tg, (1, 1, 1), 1.0, sel=sel)
    n_tr = 5
    _, st = _timed(lambda: [net.train_step(th, u, v, tg, (1, 1, 1), 1.0, sel=sel)
                            for _ in range(n_tr)], n_tr)
    flop = 2.0 * H * H * (27 * 64 + 4 * 576 * 64 + 576)
    fwd_tf = flop * n_fwd / sf / 1e12
    tr_tf = 3 * flop * B * n_tr / st / 1e12
    print(json.dumps({
        "metric": "closure network forward and training throughput", "value": fwd_tf,
        "unit": "TFLOP/s (forward, one image)", "n_gpus": 1, "steps": n_fwd, "warmup": 3,
        "ms_per_step": sf / n_fwd * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 (fp32 accumulation)", "data": "synthetic",
        "config": {"workload": f"6-layer 3x3 CNN, 64 channels, {H}x{H}", "batch_train": B},
        "roofline": {"bound": "mfma", "achieved": fwd_tf, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": fwd_tf / 2500.0, "traffic": None},
        "training": {"ms_per_step": st / n_tr * 1e3, "tflops": tr_tf, "frac": tr_tf / 2500.0},
        "cpu_baseline": None}))


def bench_closure(args):
    """DNS -> filtered snapshots -> training -> closed vs filtered DNS."""
    import numpy as np
    import torch

    import pygsquig_amd as pg
    from pygsquig_amd import closure as C
    from pygsquig_amd import ics

    dev = torch.device("cuda:0")
    L, alpha = 2 * np.pi, 1.0
    Nf, Nc = args.nf, args.nc
    dt_list = [float(x) for x in str(args.dt_coarse or "1e-4").split(",")]
    dt_f, dt_c = 1e-4, dt_list[0]
    nu = 1e-8 * (512 / Nf) ** 8 if args.nu is None else args.nu
    kf, dk, eps, mu = 40.0, 2.0, 0.1, 0.1
    t_start = time.perf_counter()

    def log(msg):
        print(f"[closure {time.perf_counter() - t_start:7.1f}s] {msg}", flush=True)

    def sync():
        torch.cuda.synchronize()

    gf, gc = pg.make_grid(Nf, L, dev), pg.make_grid(Nc, L, dev)
    forcing_f = pg.RingForcing(kf=kf, dk=dk, epsilon=eps, seed=42)
    damping = pg.CombinedDamping(mu=mu, kf=kf, nu_p=nu, p=8)
    dns = pg.gSQGSolver(gf, alpha, 0.0, 8, scheme="if")
    st = dns.initialize(torch.from_numpy(ics.kolmogorov_53(Nf, seed=2, alpha=alpha)).to(dev))
    dns_sec = 0.0

    def dns_run(state, n):
        nonlocal dns_sec
        sync()
        t0 = time.perf_counter()
        state = dns.multistep(state, n, dt_f, forcing=forcing_f, damping=damping)
        sync()
        dns_sec += time.perf_counter() - t0
        return state

    n_spin = int(round(args.t_spin / dt_f))
    chunk = 2000
    for i in range(0, n_spin, chunk):
        st = dns_run(st, min(chunk, n_spin - i))
        log(f"DNS N={Nf} spin-up step {min(i + chunk, n_spin)}/{n_spin}")
    # filtered snapshots
    sampler = C.SubgridSampler(gf, gc, alpha)
    store = C.SnapshotStore(args.snaps, Nc, dev)
    for k in range(args.snaps):
        st = dns_run(st, args.snap_every)
        _, th, u, v, tau = sampler.sample(st["theta_hat"])
        store.add(th, u, v, tau)
        if (k + 1) % 250 == 0 or k + 1 == args.snaps:
            log(f"snapshot {k + 1}/{args.snaps}")
    in_scale, tgt_scale = store.scales()
    rms_j = float(torch.sqrt(torch.mean(store.tau[: store.count] ** 2)))
    # training
    net = C.ClosureNet(Nc, Nc, max_batch=32, device=dev, seed=0)
    sync()
    t0 = time.perf_counter()
    hist = C.train(net, store, args.train_steps, batch=32, lr=1e-3, in_scale=in_scale,
                   tgt_scale=tgt_scale, seed=0, log_every=100,
                   log=lambda m: log(m))
    sync()
    train_sec = time.perf_counter() - t0
    flop_fwd = 2.0 * Nc * Nc * (27 * 64 + 4 * 576 * 64 + 576)
    train_tflops = 3 * flop_fwd * 32 * args.train_steps / train_sec / 1e12
    # inference throughput of one coarse field
    th1, u1, v1 = store.theta[:1], store.u[:1], store.v[:1]
    for _ in range(3):
        net.forward(th1, u1, v1, in_scale, 1.0)
    sync()
    t0 = time.perf_counter()
    n_inf = 50
    for _ in range(n_inf):
        net.forward(th1, u1, v1, in_scale, 1.0)
    sync()
    inf_ms = (time.perf_counter() - t0) / n_inf * 1e3
    log(f"trained: loss {hist[0][1]:.3e} -> {hist[-1][1]:.3e}; {train_tflops:.1f} TFLOP/s "
        f"training, forward {inf_ms:.3f} ms")
    # evaluation over t_eval: filtered DNS vs closed vs unclosed coarse runs
    t_eval = args.t_eval
    sample_dt = max(t_eval / 100, 0.01)
    n_samp = int(round(t_eval / sample_dt))
    nf_per = int(round(sample_dt / dt_f))
    thb0 = C.spectral_filter(st["theta_hat"], gc)
    seeds = [int(x) for x in str(args.coarse_seeds).split(",")]
    # The ring forcing's spectral amplitude is unnormalised (sqrt(eps) per mode), so
    # its physical field, the rate <theta f> and hence the reference's
    # |rate| > 1e-10 rescale threshold scale as 1/N^2.  The coarse forcing
    # takes amplitude sqrt(eps) (Nc/Nf)^2: the same physical forcing field and
    # threshold as the DNS (only the random phases differ).
    amp_c = np.sqrt(eps) * (Nc / Nf) ** 2 if args.match_forcing else None
    forcing_c = pg.RingForcing(kf=kf, dk=dk, epsilon=eps, seed=seeds[0], amplitude=amp_c)
    every_list = [int(x) for x in str(args.closure_every or "8").split(",")]
    closed_by = {m: C.ClosedSolver(gc, alpha, 0.0, 8, net=net, in_scale=in_scale,
                                   out_scale=1.0 / tgt_scale, scheme="if", every=m)
                 for m in every_list}
    plain = pg.gSQGSolver(gc, alpha, 0.0, 8, scheme="if")
    kmax = int(Nc // 3)

    def shell(th_c):
        return pg.compute_shell_spectrum(th_c, gc, alpha)[: kmax + 1]

    def run_coarse(solver, dtc, forcing=None):
        forcing = forcing_c if forcing is None else forcing
        nc_per = int(round(sample_dt / dtc))
        s = solver.initialize()
        s = {"theta_hat": thb0.clone(), "time": s["time"], "step": s["step"]}
        acc = np.zeros(kmax + 1)
        sec = 0.0
        for i in range(n_samp):
            sync()
            t0 = time.perf_counter()
            s = solver.multistep(s, nc_per, dtc, forcing=forcing, damping=damping)
            sync()
            sec += time.perf_counter() - t0
            acc += shell(s["theta_hat"])
        return acc / n_samp, sec, bool(torch.isfinite(torch.view_as_real(s["theta_hat"])).all())

    tuned = {}
    if not args.no_autotune:
        # Tile shapes of the coarse plans (N=4096 has one shape), tuned once
        # on the unclosed plan and applied to every coarse plan: the shapes
        # set the summation order of the forcing rate, so shared shapes keep
        # the runs comparable and the result independent of timing noise.
        best = plain.autotune(dt_c, forcing=forcing_c, damping=damping)[0]
        for m in every_list:
            closed_by[m]._plan(forcing_c, damping).set_shape(*best)
        tuned = {"all coarse plans": list(best)}
        log(f"autotuned coarse tile shapes {tuned}")
    runs = {}
    for dtc in dt_list:
        unclosed = run_coarse(plain, dtc)
        for m in every_list:
            runs[(dtc, m)] = (run_coarse(closed_by[m], dtc), unclosed)
            log(f"dt={dtc:g} every={m}: closed N={Nc} {runs[(dtc, m)][0][1]:.2f} s, unclosed "
                f"{unclosed[1]:.2f} s for {t_eval} time units")
    (E_cl, sec_cl, fin_cl), (E_pl, sec_pl, fin_pl) = runs[(dt_c, every_list[0])]
    spread = {}
    for sd in seeds[1:]:
        fo = pg.RingForcing(kf=kf, dk=dk, epsilon=eps, seed=sd, amplitude=amp_c)
        spread[sd] = run_coarse(plain, dt_c, fo)[0]
        log(f"unclosed N={Nc} with forcing seed {sd} done")
    dns_sec = 0.0
    E_dns = np.zeros(kmax + 1)
    for i in range(n_samp):
        st = dns_run(st, nf_per)
        E_dns += shell(C.spectral_filter(st["theta_hat"], gc))
        if (i + 1) % 10 == 0:
            log(f"DNS evaluation {i + 1}/{n_samp} samples")
    E_dns /= n_samp
    dns_eval_sec = dns_sec
    ks = np.arange(1, kmax + 1)
    speedup = dns_eval_sec / sec_cl
    def errs(E):
        r = np.abs(E[ks] - E_dns[ks]) / E_dns[ks]
        return {"mean_rel": float(r.mean()), "median_rel": float(np.median(r)),
                "max_rel": float(r.max()), "within_5pct_frac": float(np.mean(r <= 0.05)),
                "energy_weighted_rel": float(np.sum(np.abs(E[ks] - E_dns[ks])) / np.sum(E_dns[ks])),
                "mean_abs_log_ratio": float(np.mean(np.abs(np.log(E[ks] / E_dns[ks]))))}

    by_run = {}
    for (dtc, m), ((ec, sc, _), (ep, sp, _)) in runs.items():
        by_run[f"dt={dtc:g},every={m}"] = {"closed": errs(ec), "unclosed": errs(ep),
                                           "closed_s": sc, "unclosed_s": sp,
                                           "speedup": dns_eval_sec / sc,
                                           "unclosed_speedup": dns_eval_sec / sp,
                                           "spectrum_closed": ec[ks].tolist()}
    out = {
        "metric": "closure speedup at matched E(k) error", "value": speedup,
        "unit": "x (filtered-DNS wall-clock / closed-run wall-clock, same simulated time)",
        "n_gpus": 1, "steps": n_samp * int(round(sample_dt / dt_c)), "warmup": 0,
        "ms_per_step": sec_cl / (n_samp * int(round(sample_dt / dt_c))) * 1e3,
        "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 solver, bf16/fp32 network",
        "data": "synthetic (the solver's own forced DNS)",
        "config": {"workload": f"learned closure: DNS N={Nf} -> N={Nc}", "Nf": Nf, "Nc": Nc,
                   "alpha": alpha, "dt_dns": dt_f, "dt_coarse": dt_c, "nu_p": nu,
                   "snapshots": store.count, "snap_every": args.snap_every,
                   "train_steps": args.train_steps, "t_spin": args.t_spin, "t_eval": t_eval,
                   "coarse_forcing_amplitude": amp_c if amp_c is not None else float(np.sqrt(eps)),
                   "finite": fin_cl and fin_pl},
        "spectrum_error": {"k_max": kmax, "closed": errs(E_cl), "unclosed": errs(E_pl),
                           "closure_every": every_list[0]},
        "runs": by_run,
        "seed_spread": {str(sd): {"vs_dns": errs(E), "vs_unclosed_seed0": {
            "energy_weighted_rel": float(np.sum(np.abs(E[ks] - E_pl[ks])) / np.sum(E_pl[ks])),
            "median_rel": float(np.median(np.abs(E[ks] - E_pl[ks]) / E_pl[ks]))}}
            for sd, E in spread.items()},
        "wall_clock_s": {"dns_eval": dns_eval_sec, "closed": sec_cl, "unclosed": sec_pl,
                         "training": train_sec},
        "autotune": tuned,
        "training": {"loss_first": hist[0][1], "loss_last": hist[-1][1],
                     "tflops": train_tflops, "mfma_frac": train_tflops / 2500.0,
                     "forward_ms": inf_ms,
                     "forward_tflops": flop_fwd / (inf_ms * 1e-3) / 1e12,
                     "tau_rms": rms_j},
        "spectra": {"k": ks.tolist(), "dns": E_dns[ks].tolist(), "closed": E_cl[ks].tolist(),
                    "unclosed": E_pl[ks].tolist()},
        "cpu_baseline": None,
        "note": "cpu_baseline is reported for the default configuration only",
    }
    print(json.dumps(out))


def bench_closure_decay(args, emit=True):
    """Decaying DNS N=512 -> closed and unclosed N=256 runs on held-out
    realizations: mean log(E / E_DNS) per shell band and wall-clock per
    simulated time unit.  Prints the JSON line (emit) or returns it, its
    progress lines then going to stderr."""
    import numpy as np
    import torch

    import pygsquig_amd as pg
    from pygsquig_amd import closure as C
    from pygsquig_amd import ics

    dev = torch.device("cuda:0")
    L, alpha, nu, dt = 2 * np.pi, 1.0, 1e-15, 1e-3
    Nf, Nc = 512, 256
    n_real, n_test, n_snap, gap, spin, n_eval = 8, 2, 12, 50, 2000, 1000
    # the coarse grid's spacing is twice the DNS's: 2 dt is the DNS's CFL number
    dt_c = float(str(args.dt_coarse).split(",")[0]) if args.dt_coarse else 2 * dt
    every_list = [int(x) for x in str(args.closure_every or "4,0,8").split(",")]
    t_start = time.perf_counter()

    def log(msg):
        print(f"[closure-decay {time.perf_counter() - t_start:6.1f}s] {msg}", flush=True,
              file=sys.stdout if emit else sys.stderr)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    gf, gc = pg.make_grid(Nf, L, dev), pg.make_grid(Nc, L, dev)
    dns = pg.gSQGSolver(gf, alpha, nu, 8, scheme="if")
    sampler = C.SubgridSampler(gf, gc, alpha)
    store = C.SnapshotStore(n_real * n_snap, Nc, dev)
    ends = []
    for r in range(n_real):
        st = dns.initialize(torch.from_numpy(ics.kolmogorov_53(Nf, seed=50 + r)).to(dev))
        st = dns.multistep(st, spin, dt)
        for _ in range(n_snap):
            st = dns.multistep(st, gap, dt)
            _, th, u, v, tau = sampler.sample(st["theta_hat"])
            store.add(th, u, v, tau)
        ends.append(st)
    log(f"DNS N={Nf}: {n_real} realizations x {n_snap} snapshots")
    n_train = (n_real - n_test) * n_snap
    store.count = n_train
    in_scale, tgt_scale = store.scales()
    net = C.ClosureNet(Nc, Nc, max_batch=32, device=dev, seed=0)
    hist, train_sec = timed(lambda: C.train(net, store, 800, batch=32, lr=1e-3,
                                            in_scale=in_scale, tgt_scale=tgt_scale, seed=1))
    log(f"trained in {train_sec:.1f} s: loss {hist[-1][1]:.3e}")
    plain = pg.gSQGSolver(gc, alpha, nu, 8, scheme="if")
    closed = {m: C.ClosedSolver(gc, alpha, nu, 8, net=net, in_scale=in_scale,
                                out_scale=1.0 / tgt_scale, scheme="if", every=m)
              for m in every_list}
    kmax = Nc // 3
    bands = ((1, 20), (40, 69), (70, 85))
    n_c = int(round(n_eval * dt / dt_c))

    def shell(th_c):
        return pg.compute_shell_spectrum(th_c, gc, alpha)[: kmax + 1]

    def coarse_run(sol, thb0):
        s0 = sol.initialize()
        s0["theta_hat"] = thb0.clone()
        sol.multistep(s0, 4, dt_c)  # plan set-up and graph capture outside the clock
        return timed(lambda: sol.multistep(s0, n_c, dt_c))

    per_real, secs = [], {"dns": [], "unclosed": []}
    for m in every_list:
        secs[f"closed_every_{m}"] = []
    for r, st in enumerate(ends[n_real - n_test:]):
        thb0 = C.spectral_filter(st["theta_hat"], gc)
        dns.multistep(st, 4, dt)
        st1, t_dns = timed(lambda: dns.multistep(st, n_eval, dt))
        secs["dns"].append(t_dns)
        E_dns = shell(C.spectral_filter(st1["theta_hat"], gc))
        row = {}
        for name, sol in [("unclosed", plain)] + [(f"closed_every_{m}", c)
                                                   for m, c in closed.items()]:
            s1, sec = coarse_run(sol, thb0)
            secs[name].append(sec)
            E = shell(s1["theta_hat"])
            row[name] = {f"{lo}-{hi}": float(np.mean(np.log(E[lo:hi + 1] / E_dns[lo:hi + 1])))
                         for lo, hi in bands}
        per_real.append(row)
        log(f"held-out realization {r}: " + "; ".join(
            f"{k} " + ", ".join(f"{b} {v:+.3f}" for b, v in e.items()) for k, e in row.items()))
    mean_s = {k: float(np.mean(v)) for k, v in secs.items()}
    first = f"closed_every_{every_list[0]}"
    out = {
        "metric": "closure speedup at matched E(k) error",
        "value": mean_s["dns"] / mean_s[first],
        "unit": "x (filtered-DNS wall-clock / closed-run wall-clock, same simulated time)",
        "n_gpus": 1, "steps": n_c, "warmup": 4,
        "ms_per_step": mean_s[first] / n_c * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 solver, bf16/fp32 network",
        "data": "synthetic (the solver's own decaying DNS)",
        "config": {"workload": f"learned closure, decaying turbulence: DNS N={Nf} -> N={Nc}",
                   "Nf": Nf, "Nc": Nc, "alpha": alpha, "nu_p": nu, "dt_dns": dt,
                   "dt_coarse": dt_c, "realizations": n_real, "held_out": n_test,
                   "snapshots_each": n_snap, "train_steps": 800, "t_eval": n_eval * dt,
                   "closure_every": every_list},
        "mean_log_E_ratio_vs_dns": per_real,
        "wall_clock_s": mean_s,
        "speedup": {k: mean_s["dns"] / v for k, v in mean_s.items() if k != "dns"},
        "training": {"seconds": train_sec, "loss_last": hist[-1][1]},
        "cpu_baseline": None,
        "note": "cpu_baseline is reported for the default configuration only",
    }
    if not emit:
        return out
    print(json.dumps(out))


if __name__ == "__main__":
    main()
