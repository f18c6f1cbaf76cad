"""Benchmark: s/turn of the strong-strong PIC collision at production shape on one MI355X.

Workload (production shape, FP64): e- 2.56M and p 1.024M macroparticles, flat Gaussian beams
(sigma_x 95 um, sigma_y 8.5 um; sigma_z 0.7 cm / 6 cm; sigma_dp 5.5e-4 / 6.6e-4; beta* 0.8 / 0.072 m),
15 :normal_quantile slices with centroid centres, 128x128 mesh, TSC deposition, 4th-order
derivative, and either
  --arm quadratic : interaction_grid = :slice_pair, slice_interpolation = :quadratic  (default)
  --arm node      : interaction_grid = :node,       slice_interpolation = :linear
(the reference rejects :node with :quadratic, src/tasks/strongstrong/pic_cpu.jl:560-574).
`--solver soft` runs the soft-Gaussian solver (GaussianPoissonSolver, gaussian.jl) on the same beams and
slicing instead (no mesh; the Hirata drift and the longitudinal kick, the reference defaults).
A turn is one collide! plus the linear 6D one-turn map of each beam (tunes (0.08, 0.06, 0.01) and
(0.228, 0.21, 0.01)).  W warm-up turns, then K timed turns; beams are synthetic (Philox counter RNG
on the device) and resident in HBM before the timed region.

Prints one JSON line.  After the warm-up, 10 untimed turns give the per-family kernel breakdown
(`kernel_ms_per_turn`, HIP events around every launch group); the timed turns then carry events around
the dominant modelled family only, and `roofline` is that family's algorithmic HBM bytes per launch
over its mean launch duration in the timed region.  `cpu_baseline` times the CPU oracle (a C++
restatement of the reference's CPU path) on the host cores.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "octopus.jl_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0

E_CFG = dict(n=2_560_000, sigma=(95e-6, 8.5e-6, 0.7e-2), beta=(0.8, 0.072, 0.7e-2 / 5.5e-4), alpha=(0.0, 0.0, 0.0),
             charge=-1.0, energy=10.0e9, npart=1.7203e11, rng_id=1, tunes=(0.08, 0.06, 0.01))
P_CFG = dict(n=1_024_000, sigma=(95e-6, 8.5e-6, 6.0e-2), beta=(0.8, 0.072, 6.0e-2 / 6.6e-4), alpha=(0.0, 0.0, 0.0),
             charge=1.0, energy=275.0e9, npart=0.6881e11, rng_id=2, tunes=(0.228, 0.21, 0.01))
NSLICES, GRID = 15, 128
BREAKDOWN_TURNS = 10
TRAFFIC_PROFILE = {"quadratic": "hbm_traffic_2p56M_1p024M_15sl_128_quadratic.md"}
COPY_PROBE_BYTES = 1 << 30


def make_solver(arm, precision="fp64", solver="pic"):
    from octopic import GaussianPoissonSolver, LongitudinalSlicing, PICPoissonSolver
    sl = LongitudinalSlicing(nslices=NSLICES, method="normal_quantile", center_position="centroid")
    if solver == "soft":
        return GaussianPoissonSolver(slicing=sl, precision=precision)
    if arm == "node":
        return PICPoissonSolver(grid=(GRID, GRID), deposit_method="TSC", field_derivative="fourth",
                                slice_interpolation="linear", interaction_grid="node", slicing=sl, precision=precision)
    return PICPoissonSolver(grid=(GRID, GRID), deposit_method="TSC", field_derivative="fourth",
                            slice_interpolation="quadratic", interaction_grid="slice_pair", slicing=sl,
                            precision=precision)


def params(cfg, mass):
    from octopic import beam_params
    return beam_params(cfg["charge"], mass, cfg["energy"], cfg["npart"])


def bytes_model(n1, n2, ns1, ns2, npl, solver="pic"):
    """Algorithmic HBM bytes per collide of each kernel family (FP64 coordinates)."""
    visits = n1 * ns2 + n2 * ns1          # particle-pair visits of each beam
    if solver == "soft":
        return {
            "geometry": 32.0 * visits,        # the source moments: x, px, y, py of both slices of every pair
            "kick": 96.0 * visits,            # read and write the six coordinates (the drifts move z)
            "slicing": 8.0 * 3 * (n1 + n2) + 2 * 48.0 * (n1 + n2),
            "store": 2 * 48.0 * (n1 + n2),    # states -> beams (6 coordinates)
        }
    return {
        "bounds": 40.0 * visits,             # x, px, y, py, z of both slices of every pair
        "deposit": 32.0 * visits,            # x, px, y, py once: every plane of a visit comes from one pre-kick read
        "kick": 88.0 * visits,               # read x, px, y, py, z, pz; write x, px, y, py, pz
        "slicing": 8.0 * 3 * (n1 + n2) + 2 * 48.0 * (n1 + n2),   # z passes + gather of the states
        "store": 2 * 40.0 * (n1 + n2),       # states -> beams (5 coordinates)
    }


def instance_configs(k, seed=7):
    """Batched form: k independent bunch pairs with e- N ~ U[0.3M, 2.56M], p N ~ U[0.12M, 1.024M],
    slices in {7, 11, 15}, meshes in {64, 96, 128}, per-instance seeds 1..k and tunes; the C1 beam
    shapes otherwise (TSC, 4th order, quadratic slice interpolation, FP32 form)."""
    import numpy as np
    rs = np.random.RandomState(seed)
    out = []
    for i in range(k):
        out.append(dict(n1=int(rs.uniform(0.3e6, 2.56e6)), n2=int(rs.uniform(0.12e6, 1.024e6)),
                        ns=int(rs.choice([7, 11, 15])), grid=int(rs.choice([64, 96, 128])), seed=i + 1,
                        dtune=0.002 * rs.uniform(-1, 1)))
    return out


def run_instances(args):
    """Aggregate turns/s of the batched form (all instances collided together, then their maps)."""
    import torch
    from octopic import (EMASS_EV, PMASS_EV, LongitudinalSlicing, PICCollider, PICPoissonSolver, collide_many,
                         generate_beam, linear6d_matrix)
    dev = torch.device("cuda:0")
    cfgs = instance_configs(args.instances)
    cols, b1, b2, maps = [], [], [], []
    for c in cfgs:
        sl = LongitudinalSlicing(nslices=c["ns"], method="normal_quantile", center_position="centroid")
        solver = PICPoissonSolver(grid=(c["grid"], c["grid"]), deposit_method="TSC", field_derivative="fourth",
                                  slice_interpolation="quadratic", slicing=sl, precision=args.precision)
        col = PICCollider(solver, device=0)
        cols.append(col)
        b1.append(generate_beam(col, c["n1"], params(E_CFG, EMASS_EV), E_CFG["beta"], E_CFG["alpha"], E_CFG["sigma"],
                                seed=c["seed"], rng_id=1, device=dev))
        b2.append(generate_beam(col, c["n2"], params(P_CFG, PMASS_EV), P_CFG["beta"], P_CFG["alpha"], P_CFG["sigma"],
                                seed=c["seed"], rng_id=2, device=dev))
        maps.append((linear6d_matrix(E_CFG["beta"], [2 * math.pi * (t + c["dtune"]) for t in E_CFG["tunes"]]),
                     linear6d_matrix(P_CFG["beta"], [2 * math.pi * (t + c["dtune"]) for t in P_CFG["tunes"]])))

    def turn(t):
        if args.one_by_one:
            lums = [col.collide(x, y, turn=t) for col, x, y in zip(cols, b1, b2)]
        else:
            lums = collide_many(cols, b1, b2, [t] * len(cols))
        for col, x, y, (mx, my) in zip(cols, b1, b2, maps):
            col.track_linear6d(x, mx)
            col.track_linear6d(y, my)
        return lums

    for t in range(args.warmup):
        turn(t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(args.warmup, args.warmup + args.steps):
        turn(t)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    particles = sum(c["n1"] + c["n2"] for c in cfgs)
    # algorithmic HBM bytes of one step: every instance's collide (the same per-visit model as the single
    # pair: bounds 40 B, deposit 32 B, kick 88 B per visit; slicing 28 B per particle) plus its two linear
    # maps (96 B per particle: read and write six coordinates)
    step_bytes = sum(160.0 * (c["n1"] * c["ns"] + c["n2"] * c["ns"]) + (28.0 + 96.0) * (c["n1"] + c["n2"]) for c in cfgs)
    achieved = step_bytes / dt / 1e9
    print(json.dumps({
        "metric": "aggregate turns/s of independent bunch-pair instances", "value": round(len(cfgs) / dt, 3),
        "unit": "turns/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32", "data": "synthetic",
        "config": {"workload": f"{len(cfgs)} instances, N e- in [0.3M, 2.56M], N p in [0.12M, 1.024M], slices "
                               "{7,11,15}, meshes {64,96,128}, TSC, 4th order, quadratic, one collide + maps each",
                   "instances": cfgs, "particles": particles,
                   "schedule": "one by one" if args.one_by_one else "collide_many (concurrent streams)"},
        "roofline": {"bound": "hbm", "kernel": "whole step (all instances' collides and maps)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_step": round(step_bytes)},
        "hbm_GBps_algorithmic": round(achieved, 1)}))


def measured_traffic(arm, family, precision):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (copy-kernel calibrated), when one exists for this workload, kernel and precision."""
    kernel = {("quadratic", "kick", "fp64"): "k_kick<1, 1, double>",
              ("quadratic", "deposit", "fp64"): "k_deposit<1, double>"}.get((arm, family, precision))
    path = os.path.join(ROOT, "profiles", TRAFFIC_PROFILE[arm]) if arm in TRAFFIC_PROFILE else None
    if kernel is None or path is None or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(arm):
    """The CPU oracle on the host cores, on the full workload: one cold collide (fills the Green
    cache), then timed warm turns (collide + the two linear maps), as the GPU turn: one on all the
    cores this process may use (at most 16, the box's CPU share), then one on a single thread."""
    import oracle as O
    import numpy as np
    from octopic import EMASS_EV, PMASS_EV, linear6d_matrix
    nproc = os.cpu_count() or 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    threads = max(1, min(16, usable))
    pe, pp = params(E_CFG, EMASS_EV), params(P_CFG, PMASS_EV)
    e = O.generate_beam(E_CFG["n"], 0, E_CFG["rng_id"], E_CFG["beta"], E_CFG["alpha"], E_CFG["sigma"])
    p = O.generate_beam(P_CFG["n"], 0, P_CFG["rng_id"], P_CFG["beta"], P_CFG["alpha"], P_CFG["sigma"])
    me = np.asarray(linear6d_matrix(E_CFG["beta"], [2 * math.pi * t for t in E_CFG["tunes"]]), dtype=np.float64)
    mp = np.asarray(linear6d_matrix(P_CFG["beta"], [2 * math.pi * t for t in P_CFG["tunes"]]), dtype=np.float64)
    col = O.Collider(make_solver(arm))
    t0 = time.perf_counter()
    col.collide(e, pe, p, pp, nthreads=threads)
    cold = time.perf_counter() - t0
    O.track_linear6d(me, e)
    O.track_linear6d(mp, p)

    def warm_turn(nthreads):
        t0 = time.perf_counter()
        col.collide(e, pe, p, pp, nthreads=nthreads)
        O.track_linear6d(me, e)
        O.track_linear6d(mp, p)
        return time.perf_counter() - t0

    arms = [{"cores": threads, "value": round(warm_turn(threads), 4)}]
    if threads > 1:
        arms.append({"cores": 1, "value": round(warm_turn(1), 4)})
    return {"value": arms[0]["value"], "unit": "s/turn", "cores": threads, "kind": "port", "arms": arms,
            "host_nproc": nproc, "host_usable_cpus": usable,
            "sample": f"CPU oracle (C++ restatement of the reference CPU path) on the full workload: one cold "
                      f"collide ({cold:.2f} s, fills the Green cache), then timed warm turns (collide + linear "
                      f"maps): {threads} threads over conflict-free slice-pair batches, then 1 thread; "
                      f"the host reports {nproc} CPUs, {usable} usable by this process"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--arm", choices=["quadratic", "node"], default="quadratic")
    ap.add_argument("--solver", choices=["pic", "soft"], default="pic",
                    help="soft: the soft-Gaussian solver on the same beams and slicing (no mesh)")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64",
                    help="fp32: the FP32 form (deposit weights, FFT, field, kick gather in FP32; FP64 beams)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="skip the per-kernel HIP event timing")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch shapes with the library's autotuner before the warm-up (results are unchanged)")
    ap.add_argument("--launch-shape", action="append", default=[], metavar="NAME=VALUE",
                    help="set a launch shape of the collider (repeatable; results are unchanged)")
    ap.add_argument("--instances", type=int, default=0,
                    help="batched form: this many independent bunch-pair instances per step (0: the single pair)")
    ap.add_argument("--one-by-one", action="store_true", help="batched form: collide the instances one at a time")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run plain device copies of 1 GiB (calibrates the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the collision path runs on one GPU: there is no multi-GPU implementation "
                                   "(particle decomposition with one integer all-reduce of the charge planes per "
                                   "batch is the design, not built)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    if args.copy_probe:
        src = torch.ones(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": COPY_PROBE_BYTES, "copies": 3}))
        return
    if args.instances > 0:
        run_instances(args)
        return
    from octopic import EMASS_EV, PMASS_EV, PICCollider, generate_beam, linear6d_matrix

    dev = torch.device("cuda:0")
    solver = make_solver(args.arm, args.precision, args.solver)
    col = PICCollider(solver, device=0)
    e = generate_beam(col, E_CFG["n"], params(E_CFG, EMASS_EV), E_CFG["beta"], E_CFG["alpha"], E_CFG["sigma"],
                      seed=0, rng_id=E_CFG["rng_id"], device=dev)
    p = generate_beam(col, P_CFG["n"], params(P_CFG, PMASS_EV), P_CFG["beta"], P_CFG["alpha"], P_CFG["sigma"],
                      seed=0, rng_id=P_CFG["rng_id"], device=dev)
    me = linear6d_matrix(E_CFG["beta"], [2 * math.pi * t for t in E_CFG["tunes"]])
    mp = linear6d_matrix(P_CFG["beta"], [2 * math.pi * t for t in P_CFG["tunes"]])
    for item in args.launch_shape:
        name, _, value = item.partition("=")
        col.set_launch_shape(name.strip(), int(value))
    tuned = col.autotune(e, p, trials=3) if args.autotune else None

    def turn(t):
        lum = col.collide(e, p, turn=t)
        col.track_linear6d(e, me)
        col.track_linear6d(p, mp)
        return lum

    for t in range(args.warmup):
        turn(t)
    torch.cuda.synchronize()

    npl = 3
    model = bytes_model(E_CFG["n"], P_CFG["n"], NSLICES, NSLICES, npl, args.solver)
    # Per-family breakdown on BREAKDOWN_TURNS untimed turns (events around every launch group); the
    # dominant modelled family is then the only one timed inside the timed region.
    fam = {}
    best = "kick"
    t_next = args.warmup
    if not args.no_timing:
        col.set_timing(True)
        for t in range(t_next, t_next + BREAKDOWN_TURNS):
            turn(t)
        torch.cuda.synchronize()
        fam = {k: (v[0] / BREAKDOWN_TURNS, v[1] / BREAKDOWN_TURNS) for k, v in col.kernel_times().items() if v[1]}
        t_next += BREAKDOWN_TURNS
        cands = [k for k in fam if k in model]
        if cands:
            best = max(cands, key=lambda k: fam[k][0])
        col.set_timing(True, best)
    lums = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(t_next, t_next + args.steps):
        lums.append(turn(t))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s_per_turn = elapsed / args.steps
    st = col.stats()
    roof = None
    if not args.no_timing:
        timed = col.kernel_times()
        col.set_timing(False)
        ms, launches = timed.get(best, (0.0, 0))
        if launches > 0 and ms > 0:
            per_launch_bytes = model[best] * args.steps / launches
            avg_s = ms / 1e3 / launches
            achieved = per_launch_bytes / avg_s / 1e9
            traffic = measured_traffic(args.arm, best, args.precision) if args.solver == "pic" else None
            roof = {"bound": "hbm", "kernel": best, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": round(traffic) if traffic else None,
                    "avg_launch_ms": round(avg_s * 1e3, 4), "launches": int(launches),
                    "bytes_per_launch": round(per_launch_bytes)}
    out = {
        "metric": "s/turn PIC (plain FP64 vs FP32+denoiser) at equal emittance growth; HBM %peak",
        "value": round(s_per_turn, 6),
        "unit": "s/turn",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_per_turn * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (Philox counter-RNG Gaussian beams generated on the device)",
        "config": {"workload": (f"plain PIC, {E_CFG['n']} e- / {P_CFG['n']} p, {NSLICES} slices, "
                                f"{GRID}x{GRID}, TSC, 4th-order, "
                                + ("slice_pair + quadratic" if args.arm == "quadratic" else "node + linear")
                                + (", FP32 form (FP64 beams, Green, moments, luminosity)" if args.precision == "fp32" else "")
                                if args.solver == "pic" else
                                f"soft-Gaussian solver, {E_CFG['n']} e- / {P_CFG['n']} p, {NSLICES} slices, Hirata "
                                f"drift, longitudinal kick"
                                + (", FP32 Bassetti-Erskine evaluation (FP64 beams, drifts, moments)"
                                   if args.precision == "fp32" else ""))
                               + ", one collide + linear 6D maps per turn",
                   "solver": args.solver,
                   "arm": args.arm, "precision": args.precision, "global_batch": E_CFG["n"] + P_CFG["n"], "seq_len": NSLICES,
                   "parallelism": "single GPU", "launch_shape": tuned if tuned is not None else col.launch_shape()},
        "roofline": roof,
        "luminosity_mean": sum(lums) / len(lums) if lums else None,
        "green_cache": {"hits": st["green_hits"], "misses": st["green_misses"], "rebuilds": st["green_rebuilds"]},
        "kernel_ms_per_turn": {k: round(v[0], 4) for k, v in fam.items()},
        # every modelled family's algorithmic HBM bytes per turn over its time per turn (breakdown turns)
        "family_hbm_frac": {k: round(model[k] / (fam[k][0] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                            for k in fam if k in model and fam[k][0] > 0},
    }
    if args.solver == "soft":
        out["metric"] = "s/turn soft-Gaussian solver; HBM %peak"
        out.pop("green_cache")
    if not args.no_cpu_baseline and args.precision == "fp64" and args.solver == "pic":
        try:
            out["cpu_baseline"] = cpu_baseline(args.arm)
        except Exception as exc:  # the CPU leg is reported, never the measurement
            out["cpu_baseline"] = {"error": str(exc)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
