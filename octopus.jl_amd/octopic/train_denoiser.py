"""Train the residual denoiser on pairs the method generates itself, and measure what it gains.

Every step draws a fresh batch of training pairs on the device (octopic_denoiser_make_pairs: a noisy
deposit of 16k-256k macroparticles and a converged deposit of ``n_target`` macroparticles of the same
slice distribution, each Gaussian-subtracted) and takes one Adam step; the learning rate decays on a
cosine from ``lr`` to ``lr / 20``. The monitor range (the largest rms input residual seen in
training) is kept with the weights: in the collision the denoised plane is used only while the
plane's monitor stays below three times that range.

    python -m octopic.train_denoiser --steps 2000 --out weights.npz

writes the parameters and the monitor range to one ``.npz`` file (plain arrays, no pickles).
"""
import argparse
import math
import time

import numpy as np

from .denoise import Denoiser


def train(d: Denoiser, steps: int, batch: int, seed: int = 1, n_target: int = 16_000_000, lr: float = 2e-3,
          method: str = "TSC", log_every: int = 0, log=print):
    """Adam on fresh pairs; returns (loss per step as a numpy array, monitor range)."""
    t = d._torch
    losses = t.empty(steps, dtype=t.float32, device=f"cuda:{d.device}")
    mon_max = t.zeros((), dtype=t.float32, device=f"cuda:{d.device}")
    for s in range(steps):
        x, y, mon = d.make_pairs(seed, s * batch, batch, n_target=n_target, method=method)
        rate = lr * (0.05 + 0.95 * 0.5 * (1.0 + math.cos(math.pi * s / max(steps - 1, 1))))
        losses[s] = d.train_step(x, y, rate)[0]
        mon_max = t.maximum(mon_max, mon.max())
        if log_every and (s + 1) % log_every == 0:
            log(f"step {s + 1}/{steps}: loss {float(losses[s]):.4e} (lr {rate:.2e})")
    return losses.cpu().numpy(), float(mon_max)


def evaluate(d: Denoiser, seed: int, samples: int, batch: int, n_target: int = 16_000_000, method: str = "TSC"):
    """Held-out pairs (a seed not used in training): mean squared error against the converged residual
    of the raw noisy residual and of the denoised one, both over the subtracted Gaussian's peak."""
    raw = den = 0.0
    count = 0
    for first in range(0, samples, batch):
        b = min(batch, samples - first)
        x, y, _ = d.make_pairs(seed, first, b, n_target=n_target, method=method)
        out = d.forward(x)
        raw += float(((x[..., 0] - y) ** 2).sum())
        den += float(((out - y) ** 2).sum())
        count += y.numel()
    return raw / count, den / count


def save(d: Denoiser, path: str, monitor_range: float, **meta):
    np.savez(path, params=d.get_params(), monitor_range=np.float64(monitor_range),
             shape=np.array([d.height, d.width], np.int64), **{k: np.asarray(v) for k, v in meta.items()})


def load(d: Denoiser, path: str) -> float:
    """Loads parameters into d; returns the monitor range."""
    with np.load(path, allow_pickle=False) as f:
        d.set_params(f["params"])
        return float(f["monitor_range"])


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--n-target", type=int, default=16_000_000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--method", default="TSC", choices=["CIC", "TSC"])
    ap.add_argument("--eval-samples", type=int, default=256)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    d = Denoiser(a.grid, a.grid, a.batch, seed=a.init_seed)
    t0 = time.time()
    losses, mon = train(d, a.steps, a.batch, a.seed, a.n_target, a.lr, a.method, log_every=max(a.steps // 20, 1))
    d._torch.cuda.synchronize()
    t1 = time.time()
    raw, den = evaluate(d, a.seed + 1000, a.eval_samples, a.batch, a.n_target, a.method)
    print(f"trained {a.steps} steps x {a.batch} pairs in {t1 - t0:.1f} s; monitor range {mon:.4f}")
    print(f"held-out mse: raw residual {raw:.4e}, denoised {den:.4e}, gain {raw / den:.2f}x")
    if a.out:
        save(d, a.out, mon, steps=a.steps, batch=a.batch, n_target=a.n_target, seed=a.seed, method=a.method,
             heldout_raw=raw, heldout_denoised=den)
        print(f"wrote {a.out}")


if __name__ == "__main__":
    main()
