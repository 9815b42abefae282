"""Benchmark: particle-steps/s of the headline workload on one MI355X.

Workload (the Kob-Andersen configuration): binary LJ 80:20, N = 65536 in a
cubic box at rho = 1.2 (simple-cubic start, types from a seeded permutation), SP cutoff 2.5 sigma_ij, cell-linked-list neighbour list with
skin 0.3, Nose-Hoover NVT at T = 0.45 (tau 0.2), dt = 0.004, float32, with a
1000-bin partial RDF every 100 steps.  The system is melted (T = 2.0) and
equilibrated at T = 0.45 before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A "step" is one MD timestep of all N particles; the timed region covers K
steps (the RDF updates that fall inside included) between two device
synchronisations.  Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
N_PARTICLES = 65536
RHO = 1.2
T_TARGET = 0.45
DT = 0.004
SKIN = 0.3
RDF_EVERY = 100
RDF_BINS = 1000
SIG = [[1.00, 0.80], [0.80, 0.88]]
EPS = [[1.00, 1.50], [1.50, 0.50]]
METRIC = "particle-steps/s KA-LJ N=65536; batched-replica & adjoint speedup vs plain"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_system(seed=0):
    import gamdpy_amd as gp
    conf = gp.Configuration(D=3)
    # cubic box L = (N/rho)^(1/3) = 37.94 (Configuration.make_positions, as examples/kablj.py)
    conf.make_positions(N=N_PARTICLES, rho=RHO)
    conf['m'] = 1.0
    N = conf.N
    rng = np.random.default_rng(seed)
    ptype = np.zeros(N, np.int32)
    ptype[rng.permutation(N)[:int(round(0.2 * N))]] = 1
    conf.ptype = ptype
    conf.randomize_velocities(temperature=2.0, seed=seed)
    func = gp.apply_shifted_potential_cutoff(gp.LJ_12_6_sigma_epsilon)
    pp = gp.PairPotential(func, params=[SIG, EPS, np.array(SIG) * 2.5], max_num_nbs=512)
    return conf, pp


def list_length_stats(conf, pp):
    nbl = pp.nblist.d_nblist.copy_to_host()
    return float(np.mean(nbl[:, -1])), int(np.max(nbl[:, -1]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=500)
    ap.add_argument("--equil", type=int, default=4000, help="melt + equilibration steps before warmup")
    ap.add_argument("--tp", type=int, default=None)
    ap.add_argument("--pb", type=int, default=None)
    ap.add_argument("--driver", choices=["gridsync", "graph", "plain"], default=None)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(independent replicas would shard one set per GPU)", file=sys.stderr)
        sys.exit(2)

    import gamdpy_amd as gp
    from gamdpy_amd import _native
    info = _native.device_info()
    log(f"[bench] device {info['name']} {info['arch']} CUs={info['compute_units']}")

    conf, pp = make_system()
    N = conf.N
    plan = gp.get_default_compute_plan(conf)
    plan['skin'] = SKIN
    plan['nblist'] = 'linked lists'
    if args.tp:
        plan['tp'] = args.tp
        plan['pb'] = max(1, 256 // args.tp)
    if args.pb:
        plan['pb'] = args.pb
    if args.driver:
        plan['gridsync'] = args.driver == "gridsync"
        plan['graph'] = args.driver == "graph"
    log(f"[bench] N={N} plan={plan}")

    # melt at T=2.0, then equilibrate at T=0.45 (untimed)
    half = max(RDF_EVERY, (args.equil // 2 // RDF_EVERY) * RDF_EVERY)
    melt = gp.NVT(temperature=2.0, tau=0.2, dt=DT)
    sim = gp.Simulation(conf, pp, melt, [gp.MomentumReset(100)], num_timeblocks=1, steps_per_timeblock=half,
                        compute_plan=dict(plan), storage='memory')
    for _ in sim.run_timeblocks():
        pass
    integ = gp.NVT(temperature=T_TARGET, tau=0.2, dt=DT)
    sim = gp.Simulation(conf, pp, integ, [gp.MomentumReset(100)], num_timeblocks=1, steps_per_timeblock=half,
                        compute_plan=dict(plan), storage='memory')
    for _ in sim.run_timeblocks():
        pass
    log(f"[bench] equilibrated {2 * half} steps; U/N={np.mean(conf['U']):.4f} "
        f"T={2 * np.sum(conf['K']) / (3 * N - 3):.4f}")

    # production: blocks of RDF_EVERY steps, RDF update after each block
    steps = args.steps
    nblocks = max(1, math.ceil(steps / RDF_EVERY))
    sim = gp.Simulation(conf, pp, integ, [gp.MomentumReset(100)], num_timeblocks=nblocks,
                        steps_per_timeblock=RDF_EVERY, compute_plan=dict(plan), storage='memory')
    calc = gp.CalculatorRadialDistribution(conf, bins=RDF_BINS)
    ctx = sim.ctx
    conf.copy_to_device()
    # warmup: W steps plus one RDF update (untimed)
    t_w = 0
    while t_w < args.warmup:
        k = min(RDF_EVERY, args.warmup - t_w)
        sim.integrate_self(0.0, k)
        t_w += k
    calc.update()
    rdf_h = calc._h
    nb0 = ctx.nbflag()[2]
    kernel_ms = []
    rdf_count = 0
    done = 0
    t0 = time.perf_counter()
    while done < steps:
        k = min(RDF_EVERY, steps - done)
        sim.integrate_self(np.float32(done * DT), k)
        kernel_ms.append((ctx.last_run_ms(), k))
        done += k
        if k == RDF_EVERY:
            rdf_h.compute_ctx(ctx)      # synchronous: returns the counts on the host
            rdf_count += 1
    elapsed = time.perf_counter() - t0
    rebuilds = int(ctx.nbflag()[2] - nb0)
    conf.copy_to_host()
    mean_nb, max_nb = list_length_stats(conf, pp)

    value = N * steps / elapsed
    tps = steps / elapsed
    # roofline of the dominant kernel: algorithmic bytes per particle-step
    # B_ps = 4 * mean list length + 92 (neighbour indices + per-particle state)
    n_list = mean_nb
    B_ps = 4.0 * n_list + 92.0
    step_ms = [ms / k for ms, k in kernel_ms]
    kernel_ms_per_step = float(np.mean(step_ms))
    achieved = B_ps * N / (kernel_ms_per_step * 1e-3) / 1e9
    log(f"[bench] {steps} steps in {elapsed:.3f}s: {tps:.1f} TPS, {value:.3e} particle-steps/s; "
        f"step kernel {kernel_ms_per_step * 1e3:.2f} us/step; rebuilds {rebuilds}; mean list {mean_nb:.1f}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_seconds)

    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "Kob-Andersen binary LJ 80:20, N=65536, cubic box L=37.94, rho=1.2, "
                               "SP cutoff 2.5 sigma_ij, linked-list nblist skin 0.3, NVT Nose-Hoover T=0.45 "
                               "tau=0.2, dt=0.004, 1000-bin partial RDF every 100 steps",
                   "N": N, "rho": RHO, "T": T_TARGET, "dt": DT, "skin": SKIN, "rdf_bins": RDF_BINS,
                   "rdf_every": RDF_EVERY, "equilibration_steps": 2 * half,
                   "plan": {k: plan[k] for k in ('pb', 'tp', 'gridsync', 'nblist')} | {"graph": plan.get('graph', True)}},
        "tps": tps,
        "kernel_us_per_step": kernel_ms_per_step * 1e3,
        "rdf_updates_timed": rdf_count,
        "nblist_rebuilds_timed": rebuilds,
        "mean_list_length": mean_nb,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_particle_step": B_ps,
                     "kernel": "k_persistent<LJ_sigma_epsilon,SP,NVT>" if plan['gridsync'] else "k_step<...>"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(budget_s):
    """The C/OpenMP restatement of the same algorithm (oracle/cpu_md.c) on the
    host cores, on a bounded sample of the same workload."""
    try:
        from oracle import c_oracle
    except Exception as e:  # pragma: no cover
        log(f"[bench] cpu baseline unavailable: {e}")
        return None
    conf, pp = make_system()
    table, max_cut = pp.param_table()
    threads = c_oracle.num_threads()
    state = dict(r=conf['r'], v=conf['v'])
    # warm-up (also melts the lattice a little)
    s = c_oracle.md_nvt_ll(state['r'], state['v'], conf['m'], conf.ptype, conf.simbox.lengths, table, max_cut, SKIN,
                           512, DT, T_TARGET, 0.2, 5)
    steps = 0
    t0 = time.perf_counter()
    r, v, xi = s['r'], s['v'], s['xi']
    while time.perf_counter() - t0 < budget_s:
        s = c_oracle.md_nvt_ll(r, v, conf['m'], conf.ptype, conf.simbox.lengths, table, max_cut, SKIN, 512, DT,
                               T_TARGET, 0.2, 50, xi=xi)
        r, v, xi = s['r'], s['v'], s['xi']
        steps += 50
    el = time.perf_counter() - t0
    val = conf.N * steps / el
    log(f"[bench] cpu baseline: {steps} steps in {el:.1f}s on {threads} threads: {val:.3e} particle-steps/s")
    return {"value": val, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"KA-LJ N={conf.N} NVT linked lists, {steps} steps ({el:.1f} s) after 5 warm-up steps, "
                      f"C/OpenMP float32 restatement (oracle/cpu_md.c), {threads} threads"}


if __name__ == "__main__":
    main()
