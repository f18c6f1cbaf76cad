"""End-to-end pin of the oracle against the reference's own production number.

The reference's CPU collide benchmark (profiling/benchmark_collide_cpu.jl:58-132, recorded in
docs/history/cpu_benchmark_history.tsv:2) collides the EIC production pair -- e- 2.56M
(sigma 106 um / 9.5 um / 0.7 cm, beta* 0.55 / 0.056 m), p 1.024M (95 um / 8.5 um / 6 cm, 0.8 / 0.072 m),
Philox seed 123456789, rng_id 1 / 2, cutoff 5 -- with 15 :normal_quantile centroid slices, a 128x128
CIC :integrated mesh and the slice-pair Green cache (0.5 / 0.25), through one persistent task: two
warm-up collides, then timed repeats from the restored beams.  It reports luminosity
1.0254361754701657e+30 per crossing, and its comments record 450 cache misses on the first collide and
450 hits on the second.  The oracle must reproduce both (the beams differ from the reference's only by
last-ulp libm differences, so the luminosity agrees to ~1e-15).
"""
import os

import pytest

from octopic import EMASS_EV, PMASS_EV, LongitudinalSlicing, PICPoissonSolver, beam_params

REFERENCE_LUMINOSITY = 1.0254361754701657e30
SEED = 123456789
ELE = dict(n=2_560_000, beta=(0.55, 0.056, 0.7e-2 / 5.5e-4), sigma=(106e-6, 9.5e-6, 0.7e-2))
PRO = dict(n=1_024_000, beta=(0.8, 0.072, 6.0e-2 / 6.6e-4), sigma=(95e-6, 8.5e-6, 6.0e-2))
PE = beam_params(-1.0, EMASS_EV, 10.0e9, 1.7203e11)
PP = beam_params(1.0, PMASS_EV, 275.0e9, 0.6881e11)


def production_solver():
    sl = LongitudinalSlicing(nslices=15, method="normal_quantile", center_position="centroid")
    return PICPoissonSolver(grid=(128, 128), deposit_method="CIC", green_type="integrated", green_cache="slice_pair",
                            slice_pair_green_min_ratio=0.5, slice_pair_green_growth=0.25, longitudinal_kick=True,
                            batch_mode="wavefront", slicing=sl)


def test_oracle_reproduces_reference_production_luminosity(oracle):
    e = oracle.generate_beam(ELE["n"], SEED, 1, ELE["beta"], (0, 0, 0), ELE["sigma"])
    p = oracle.generate_beam(PRO["n"], SEED, 2, PRO["beta"], (0, 0, 0), PRO["sigma"])
    col = oracle.Collider(production_solver())
    threads = min(16, os.cpu_count() or 1)
    lum_cold = col.collide(e.copy(), PE, p.copy(), PP, nthreads=threads)
    s1 = col.stats()
    lum_warm = col.collide(e.copy(), PE, p.copy(), PP, nthreads=threads)
    s2 = col.stats()
    assert (s1["misses"], s1["hits"], s1["rebuilds"]) == (450, 0, 0)
    assert (s2["misses"], s2["hits"], s2["rebuilds"]) == (450, 450, 0)
    assert lum_warm == pytest.approx(REFERENCE_LUMINOSITY, rel=1e-12)
    assert lum_cold == pytest.approx(REFERENCE_LUMINOSITY, rel=1e-12)
