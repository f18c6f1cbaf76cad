"""Benchmark: s/turn of the strong-strong PIC collision at production shape on one MI355X.

Workload (production shape, FP64): e- 2.56M and p 1.024M macroparticles, flat Gaussian beams
(sigma_x 95 um, sigma_y 8.5 um; sigma_z 0.7 cm / 6 cm; sigma_dp 5.5e-4 / 6.6e-4; beta* 0.8 / 0.072 m),
15 :normal_quantile slices with centroid centres, 128x128 mesh, TSC deposition, 4th-order
derivative, and either
  --arm quadratic : interaction_grid = :slice_pair, slice_interpolation = :quadratic  (default)
  --arm node      : interaction_grid = :node,       slice_interpolation = :linear
(the reference rejects :node with :quadratic, src/tasks/strongstrong/pic_cpu.jl:560-574).
A turn is one collide! plus the linear 6D one-turn map of each beam (tunes (0.08, 0.06, 0.01) and
(0.228, 0.21, 0.01)).  W warm-up turns, then K timed turns; beams are synthetic (Philox counter RNG
on the device) and resident in HBM before the timed region.

Prints one JSON line.  `roofline` is the dominant kernel family's algorithmic HBM bytes per launch
over its mean launch duration (HIP events on the library's stream); `cpu_baseline` times the CPU
oracle (a C++ restatement of the reference's CPU path) on the host cores on a bounded sample.
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


def make_solver(arm):
    from octopic import LongitudinalSlicing, PICPoissonSolver
    sl = LongitudinalSlicing(nslices=NSLICES, method="normal_quantile", center_position="centroid")
    if arm == "node":
        return PICPoissonSolver(grid=(GRID, GRID), deposit_method="TSC", field_derivative="fourth",
                                slice_interpolation="linear", interaction_grid="node", slicing=sl)
    return PICPoissonSolver(grid=(GRID, GRID), deposit_method="TSC", field_derivative="fourth",
                            slice_interpolation="quadratic", interaction_grid="slice_pair", slicing=sl)


def params(cfg, mass):
    from octopic import beam_params
    return beam_params(cfg["charge"], mass, cfg["energy"], cfg["npart"])


def bytes_model(n1, n2, ns1, ns2, npl):
    """Algorithmic HBM bytes per collide of each kernel family (FP64 coordinates)."""
    visits = n1 * ns2 + n2 * ns1          # particle-pair visits of each beam
    return {
        "bounds": 40.0 * visits,             # x, px, y, py, z of both slices of every pair
        "deposit": 32.0 * (npl + 1) * visits,   # x, px, y, py per field plane + the luminosity plane
        "kick": 88.0 * visits,               # read x, px, y, py, z, pz; write x, px, y, py, pz
        "slicing": 8.0 * 3 * (n1 + n2) + 2 * 48.0 * (n1 + n2),   # z passes + gather of the states
        "store": 2 * 40.0 * (n1 + n2),       # states -> beams (5 coordinates)
    }


def cpu_baseline(arm):
    """The CPU oracle on the host cores.  Bounded sample: one warm collide (after one cache-warming
    collide) at 1/40 and at 1/20 of the workload's macroparticles, extrapolated linearly in the
    particle count to the full workload (t = a + b N: the FFT and Green work is independent of N,
    the bounds/deposit/kick work is proportional to it)."""
    import oracle as O
    from octopic import EMASS_EV, PMASS_EV
    threads = min(16, os.cpu_count() or 1)
    pe, pp = params(E_CFG, EMASS_EV), params(P_CFG, PMASS_EV)
    times = {}
    for frac in (1.0 / 40, 1.0 / 20):
        n1, n2 = int(E_CFG["n"] * frac), int(P_CFG["n"] * frac)
        e = O.generate_beam(n1, 0, 1, E_CFG["beta"], E_CFG["alpha"], E_CFG["sigma"])
        p = O.generate_beam(n2, 0, 2, P_CFG["beta"], P_CFG["alpha"], P_CFG["sigma"])
        col = O.Collider(make_solver(arm))
        col.collide(e, pe, p, pp, nthreads=threads)          # warm the Green cache
        t0 = time.perf_counter()
        col.collide(e, pe, p, pp, nthreads=threads)
        times[frac] = time.perf_counter() - t0
    (f1, t1), (f2, t2) = sorted(times.items())
    slope = (t2 - t1) / (f2 - f1)
    full = t1 + slope * (1.0 - f1)
    return {"value": round(full, 3), "unit": "s/collide", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (C++ restatement of the reference CPU path), one warm collide at 1/40 "
                      f"({t1:.2f} s) and 1/20 ({t2:.2f} s) of the macroparticles, same slices/mesh/options, "
                      f"{threads} threads over conflict-free slice-pair batches, extrapolated linearly in N "
                      f"to the full workload"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--arm", choices=["quadratic", "node"], default="quadratic")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="skip the per-kernel HIP event timing")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; the collision path has no multi-GPU "
                                   "implementation (see DESIGN.md for how it would shard)", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    from octopic import EMASS_EV, PMASS_EV, PICCollider, generate_beam, linear6d_matrix

    dev = torch.device("cuda:0")
    solver = make_solver(args.arm)
    col = PICCollider(solver, device=0)
    e = generate_beam(col, E_CFG["n"], params(E_CFG, EMASS_EV), E_CFG["beta"], E_CFG["alpha"], E_CFG["sigma"],
                      seed=0, rng_id=E_CFG["rng_id"], device=dev)
    p = generate_beam(col, P_CFG["n"], params(P_CFG, PMASS_EV), P_CFG["beta"], P_CFG["alpha"], P_CFG["sigma"],
                      seed=0, rng_id=P_CFG["rng_id"], device=dev)
    me = linear6d_matrix(E_CFG["beta"], [2 * math.pi * t for t in E_CFG["tunes"]])
    mp = linear6d_matrix(P_CFG["beta"], [2 * math.pi * t for t in P_CFG["tunes"]])

    def turn(t):
        lum = col.collide(e, p, turn=t)
        col.track_linear6d(e, me)
        col.track_linear6d(p, mp)
        return lum

    for t in range(args.warmup):
        turn(t)
    torch.cuda.synchronize()
    if not args.no_timing:
        col.set_timing(True)
    lums = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(args.warmup, args.warmup + args.steps):
        lums.append(turn(t))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s_per_turn = elapsed / args.steps
    st = col.stats()
    fam = col.kernel_times() if not args.no_timing else {}

    npl = 3
    model = bytes_model(E_CFG["n"], P_CFG["n"], NSLICES, NSLICES, npl)
    roof = None
    if fam:
        best = None
        for name, (ms, launches) in fam.items():
            if name in model and launches > 0 and ms > 0:
                if best is None or ms > fam[best][0]:
                    best = name
        if best is not None:
            ms, launches = fam[best]
            per_launch_bytes = model[best] * args.steps / launches
            avg_s = ms / 1e3 / launches
            achieved = per_launch_bytes / avg_s / 1e9
            roof = {"bound": "hbm", "kernel": best, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
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
        "dtype": "f64",
        "data": "synthetic (Philox counter-RNG Gaussian beams generated on the device)",
        "config": {"workload": f"plain FP64 PIC, {E_CFG['n']} e- / {P_CFG['n']} p, {NSLICES} slices, "
                               f"{GRID}x{GRID}, TSC, 4th-order, "
                               + ("slice_pair + quadratic" if args.arm == "quadratic" else "node + linear")
                               + ", one collide + linear 6D maps per turn",
                   "arm": args.arm, "global_batch": E_CFG["n"] + P_CFG["n"], "seq_len": NSLICES,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "luminosity_mean": sum(lums) / len(lums) if lums else None,
        "green_cache": {"hits": st["green_hits"], "misses": st["green_misses"], "rebuilds": st["green_rebuilds"]},
        "kernel_ms_per_turn": {k: round(v[0] / args.steps, 4) for k, v in fam.items() if v[1]},
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(args.arm)
        except Exception as exc:  # the CPU leg is reported, never the measurement
            out["cpu_baseline"] = {"error": str(exc)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
