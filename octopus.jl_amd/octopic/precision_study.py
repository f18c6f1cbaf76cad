"""The FP32 form against FP64 across beam aspect ratios: kick and luminosity error, emittance growth.

Production shape (2.56M e- / 1.024M p, 15 :normal_quantile slices, 128 x 128, TSC, 4th order,
slice_pair + quadratic) with sigma_y = 8.5 um and sigma_x = aspect x sigma_y for both beams (beta_x
scaled to keep sigma_px), linear one-turn maps. For each aspect:

- one collide from the same start in both precisions: rms and max of (FP32 - FP64) kick over the
  rms FP64 kick, per plane, for the e- and p beams, and the luminosity's relative difference;
- ``turns`` turns in each precision: least-squares slope of eps(t)/eps(0) per turn in x and y of
  both beams;
- s/turn of each precision (wall time over ``timed`` turns).

    python -m octopic.precision_study --turns 1000 --out precision.json
"""
import argparse
import json
import math
import time

import numpy as np

from .denoise_study import E_CFG, NSLICES, P_CFG, _emittances


def _solver(precision, grid):
    from . import LongitudinalSlicing, PICPoissonSolver
    sl = LongitudinalSlicing(nslices=NSLICES, method="normal_quantile", center_position="centroid")
    return PICPoissonSolver(grid=(grid, grid), deposit_method="TSC", field_derivative="fourth",
                            slice_interpolation="quadratic", interaction_grid="slice_pair", slicing=sl,
                            precision=precision)


def _beams(col, aspect, n1, n2):
    import torch
    from . import EMASS_EV, PMASS_EV, beam_params, generate_beam
    dev = torch.device("cuda:0")
    out = []
    for cfg, mass, n in ((E_CFG, EMASS_EV, n1), (P_CFG, PMASS_EV, n2)):
        sy = cfg["sigma"][1]
        sx = aspect * sy
        # keep sigma_px = sigma_x / beta_x of the production optics
        bx = sx / (cfg["sigma"][0] / cfg["beta"][0])
        pars = beam_params(cfg["charge"], mass, cfg["energy"], cfg["npart"])
        out.append(generate_beam(col, n, pars, (bx, cfg["beta"][1], cfg["beta"][2]), cfg["alpha"],
                                 (sx, sy, cfg["sigma"][2]), seed=0, rng_id=cfg["rng_id"], device=dev))
    return out


def one_aspect(aspect, turns, timed, every, grid, n1, n2):
    import torch
    from . import PICCollider, linear6d_matrix
    res = {"aspect": aspect}
    cols = {p: PICCollider(_solver(p, grid)) for p in ("fp64", "fp32")}
    e0, p0 = _beams(cols["fp64"], aspect, n1, n2)
    kicks, lums = {}, {}
    for prec, col in cols.items():
        e = type(e0)(e0.coords.clone(), e0.params)
        p = type(p0)(p0.coords.clone(), p0.params)
        lums[prec] = col.collide(e, p)
        kicks[prec] = [(b.coords[k] - b0.coords[k]) for b, b0 in ((e, e0), (p, p0)) for k in (1, 3, 5)]
    names = ["e_px", "e_py", "e_pz", "p_px", "p_py", "p_pz"]
    err = {}
    for name, a, b in zip(names, kicks["fp32"], kicks["fp64"]):
        rms = float((b ** 2).mean().sqrt())
        d = a - b
        err[name] = {"rms": float((d ** 2).mean().sqrt()) / rms, "max": float(d.abs().max()) / rms}
    res["kick_error_fp32_vs_fp64"] = err
    res["luminosity_rel_diff"] = abs(lums["fp32"] - lums["fp64"]) / abs(lums["fp64"])
    # beta_x scaled with the aspect (sigma_px kept), other map parameters as produced
    bxe = aspect * E_CFG["sigma"][1] / (E_CFG["sigma"][0] / E_CFG["beta"][0])
    bxp = aspect * P_CFG["sigma"][1] / (P_CFG["sigma"][0] / P_CFG["beta"][0])
    me = linear6d_matrix((bxe, E_CFG["beta"][1], E_CFG["beta"][2]), [2 * math.pi * t for t in E_CFG["tunes"]])
    mp = linear6d_matrix((bxp, P_CFG["beta"][1], P_CFG["beta"][2]), [2 * math.pi * t for t in P_CFG["tunes"]])
    res["runs"] = {}
    for prec, col in cols.items():
        e = type(e0)(e0.coords.clone(), e0.params)
        p = type(p0)(p0.coords.clone(), p0.params)
        hist = [torch.stack(_emittances(e) + _emittances(p))]
        t_wall = 0.0
        for t in range(turns):
            if t == 5:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            col.collide(e, p, turn=t)
            col.track_linear6d(e, me)
            col.track_linear6d(p, mp)
            if t == 4 + timed:
                torch.cuda.synchronize()
                t_wall = (time.perf_counter() - t0) / timed
            if (t + 1) % every == 0:
                hist.append(torch.stack(_emittances(e) + _emittances(p)))
        h = torch.stack(hist).cpu().numpy()
        tt = np.arange(h.shape[0]) * every
        rel = h / h[0]
        slopes = [float(np.polyfit(tt, rel[:, k], 1)[0]) if len(tt) >= 2 else float("nan") for k in range(4)]
        res["runs"][prec] = {"s_per_turn": t_wall,
                             "growth_per_turn": dict(zip(("e_x", "e_y", "p_x", "p_y"), slopes)),
                             "final_rel_emittance": dict(zip(("e_x", "e_y", "p_x", "p_y"), map(float, rel[-1])))}
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--aspects", default="1,4,11,25")
    ap.add_argument("--turns", type=int, default=1000)
    ap.add_argument("--timed", type=int, default=30)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--n1", type=int, default=2_560_000)
    ap.add_argument("--n2", type=int, default=1_024_000)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    out = []
    for asp in (float(x) for x in a.aspects.split(",")):
        r = one_aspect(asp, a.turns, a.timed, a.every, a.grid, a.n1, a.n2)
        print(json.dumps(r), flush=True)
        out.append(r)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
