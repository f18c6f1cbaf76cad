"""The learned-denoiser form of the collision against plain PIC: speed, kick error, emittance growth.

Three runs on the production beams (flat Gaussians, 15 :normal_quantile slices, TSC, 4th order,
slice_pair + quadratic, on a ``grid`` x ``grid`` mesh, linear one-turn maps):

- ``reference``: plain PIC in FP64 at 2.56M e- / 1.024M p (four times the particles of the others);
- ``plain``:     plain PIC in FP32 at 640k / 256k;
- ``denoised``:  the FP32 form at 640k / 256k with the residual denoiser on every field plane.

For each: s/turn (collide + maps, HIP-synchronised wall time over ``timed`` turns), the kick error of
one collide (the e- kicks from the low-count p beam against the kicks from the 4x p beam, same e-
particles), and the emittance growth over ``turns`` turns (least-squares slope of eps(t)/eps(0) per
turn, eps = sqrt(<q^2><p^2> - <qp>^2) in x and y of both beams).

    python -m octopic.denoise_study --weights w.npz --turns 2000 --out study.json
"""
import argparse
import json
import math
import time

import numpy as np

E_CFG = dict(sigma=(95e-6, 8.5e-6, 0.7e-2), beta=(0.8, 0.072, 0.7e-2 / 5.5e-4), alpha=(0.0, 0.0, 0.0), charge=-1.0,
             energy=10.0e9, npart=1.7203e11, rng_id=1, tunes=(0.08, 0.06, 0.01))
P_CFG = dict(sigma=(95e-6, 8.5e-6, 6.0e-2), beta=(0.8, 0.072, 6.0e-2 / 6.6e-4), alpha=(0.0, 0.0, 0.0), charge=1.0,
             energy=275.0e9, npart=0.6881e11, rng_id=2, tunes=(0.228, 0.21, 0.01))
NSLICES = 15


def _solver(grid, precision):
    from . import LongitudinalSlicing, PICPoissonSolver
    sl = LongitudinalSlicing(nslices=NSLICES, method="normal_quantile", center_position="centroid")
    return PICPoissonSolver(grid=(grid, grid), deposit_method="TSC", field_derivative="fourth",
                            slice_interpolation="quadratic", interaction_grid="slice_pair", slicing=sl,
                            precision=precision)


def _beams(col, n1, n2, seed=0):
    import torch
    from . import EMASS_EV, PMASS_EV, beam_params, generate_beam
    dev = torch.device("cuda:0")
    pe = beam_params(E_CFG["charge"], EMASS_EV, E_CFG["energy"], E_CFG["npart"])
    pp = beam_params(P_CFG["charge"], PMASS_EV, P_CFG["energy"], P_CFG["npart"])
    e = generate_beam(col, n1, pe, E_CFG["beta"], E_CFG["alpha"], E_CFG["sigma"], seed=seed, rng_id=1, device=dev)
    p = generate_beam(col, n2, pp, P_CFG["beta"], P_CFG["alpha"], P_CFG["sigma"], seed=seed, rng_id=2, device=dev)
    return e, p


def _maps():
    from . import linear6d_matrix
    return (linear6d_matrix(E_CFG["beta"], [2 * math.pi * t for t in E_CFG["tunes"]]),
            linear6d_matrix(P_CFG["beta"], [2 * math.pi * t for t in P_CFG["tunes"]]))


def _emittances(b):
    c = b.coords
    out = []
    for q, p in ((0, 1), (2, 3)):
        x = c[q] - c[q].mean()
        px = c[p] - c[p].mean()
        out.append(((x * x).mean() * (px * px).mean() - (x * px).mean() ** 2).clamp(min=0).sqrt())
    return out


def _collider(grid, precision, denoiser, limit):
    from . import PICCollider
    col = PICCollider(_solver(grid, precision))
    if denoiser is not None:
        col.set_denoiser(denoiser, limit)
    return col


def run(kind, n1, n2, grid, precision, denoiser, limit, turns, timed, every):
    import torch
    col = _collider(grid, precision, denoiser, limit)
    e, p = _beams(col, n1, n2)
    me, mp = _maps()

    def turn(t):
        col.collide(e, p, turn=t)
        col.track_linear6d(e, me)
        col.track_linear6d(p, mp)

    hist = [torch.stack(_emittances(e) + _emittances(p))]
    t_wall = 0.0
    for t in range(turns):
        if t == 5:   # timed window after five warm turns
            torch.cuda.synchronize()
            t0 = time.perf_counter()
        turn(t)
        if t == 4 + timed:
            torch.cuda.synchronize()
            t_wall = (time.perf_counter() - t0) / timed
        if (t + 1) % every == 0:
            hist.append(torch.stack(_emittances(e) + _emittances(p)))
    torch.cuda.synchronize()
    # per-family device time of five more turns (events around every launch group)
    col.set_timing(True)
    for t in range(turns, turns + 5):
        turn(t)
    torch.cuda.synchronize()
    breakdown = {k: round(v[0] / 5, 4) for k, v in col.kernel_times().items() if v[1]}
    col.set_timing(False)
    h = torch.stack(hist).cpu().numpy()
    tt = np.arange(h.shape[0]) * every
    rel = h / h[0]
    slopes = [float(np.polyfit(tt, rel[:, k], 1)[0]) if len(tt) >= 2 else float("nan") for k in range(4)]
    st = col.stats()
    return dict(kind=kind, n1=n1, n2=n2, precision=precision, s_per_turn=t_wall, turns=turns,
                growth_per_turn={"e_x": slopes[0], "e_y": slopes[1], "p_x": slopes[2], "p_y": slopes[3]},
                final_rel_emittance={"e_x": float(rel[-1, 0]), "e_y": float(rel[-1, 1]), "p_x": float(rel[-1, 2]),
                                     "p_y": float(rel[-1, 3])},
                denoised_planes_last=int(st["denoised_planes"]), fallbacks_last=int(st["denoise_fallbacks"]),
                kernel_ms_per_turn=breakdown)


def kick_error(grid, denoiser, limit, n1, n2, factor):
    """rms over e- particles of (kick from the n2 p beam - kick from the factor x n2 p beam), relative to
    the rms reference kick, for plain FP32 and for the denoised FP32 form."""
    import torch
    ref_col = _collider(grid, "fp64", None, 0.0)
    e0, p_big = _beams(ref_col, n1, n2 * factor)
    _, p_small = _beams(ref_col, n1, n2)

    def kicks(col, p):
        e = type(e0)(e0.coords.clone(), e0.params)
        pp = type(p)(p.coords.clone(), p.params)
        col.collide(e, pp)
        return (e.coords[1] - e0.coords[1], e.coords[3] - e0.coords[3])

    rx, ry = kicks(ref_col, p_big)
    out = {}
    for name, dn in (("plain", None), ("denoised", denoiser)):
        kx, ky = kicks(_collider(grid, "fp32", dn, limit), p_small)
        out[name] = {"x": float(((kx - rx) ** 2).mean().sqrt() / (rx ** 2).mean().sqrt()),
                     "y": float(((ky - ry) ** 2).mean().sqrt() / (ry ** 2).mean().sqrt())}
    torch.cuda.synchronize()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--weights", default=None, help="trained parameters (.npz from octopic.train_denoiser)")
    ap.add_argument("--train-steps", type=int, default=3000, help="train first when no weights are given")
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--turns", type=int, default=2000)
    ap.add_argument("--timed", type=int, default=50)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--limit-factor", type=float, default=3.0)
    ap.add_argument("--skip-reference", action="store_true")
    ap.add_argument("--only", default=None, help="comma-separated subset of reference,plain,denoised")
    ap.add_argument("--skip-kick-error", action="store_true")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    from .denoise import Denoiser
    from . import train_denoiser as TD
    planes = 2 * NSLICES * 3
    d = Denoiser(a.grid, a.grid, max(planes, 16), seed=0)
    if a.weights:
        mon = TD.load(d, a.weights)
        train_info = {"weights": a.weights}
    else:
        t0 = time.time()
        losses, mon = TD.train(d, a.train_steps, 16, seed=1)
        train_info = {"steps": a.train_steps, "seconds": time.time() - t0, "final_loss": float(losses[-50:].mean())}
    limit = a.limit_factor * mon
    res = {"grid": a.grid, "monitor_range": mon, "monitor_limit": limit, "training": train_info}
    if not a.skip_kick_error:
        res["kick_error"] = kick_error(a.grid, d, limit, 640_000, 256_000, 4)
        print("kick error:", json.dumps(res["kick_error"]), flush=True)
    runs = []
    if not a.skip_reference:
        runs.append(("reference", 2_560_000, 1_024_000, "fp64", None))
    runs += [("plain", 640_000, 256_000, "fp32", None), ("denoised", 640_000, 256_000, "fp32", d)]
    if a.only:
        keep = set(a.only.split(","))
        runs = [r for r in runs if r[0] in keep]
    res["runs"] = []
    for kind, n1, n2, prec, dn in runs:
        r = run(kind, n1, n2, a.grid, prec, dn, limit, a.turns, a.timed, a.every)
        print(json.dumps(r), flush=True)
        res["runs"].append(r)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
