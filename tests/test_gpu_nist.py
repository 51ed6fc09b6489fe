"""NIST Lennard-Jones validation over batched replicas (nist.py, the reference's bench/validate.py): the 16 state
points of N = 500 NVT runs (cutoff 3, dt 0.005, 10000 equilibration + 20000 production steps, FP64) advanced as
one md_batch, converted to NIST's conventions and compared with the NIST MC data and with the reference engine's
own published results.

The reference runs 8 replicas per state point and reports all 20 NIST MC comparisons within |z| < 2 and chi^2 =
17.2 over them (README.md:443-457). The batch runs 32 (seeds 1000 i + r + 1 as validate.py numbers them, so the
first 8 are the reference's seeds): 512 replicas in about 11 s, which halves the error bars. Bars: every |z| < 3
against the reference engine's table (engine against engine); against NIST MC, chi^2 over the 20 values below
45.3 (the 99.9% quantile of chi^2 with 20 degrees of freedom) and every |z| < 3.5 (a 1% chance for the largest of
20 normal deviates); each replica's mean temperature within 0.01 of its thermostat's target. The run is FP64 with
fixed seeds and fixed-order sums, so its result is the same on every run. The NIST MD comparison is printed for
context, as validate.py prints it (NIST's MD and MC data disagree at several points, validate.py:12-14).
"""
import math

import pytest

import nist

pytestmark = pytest.mark.gpu

REPLICAS = 32

# The reference engine's results (README.md:443-454): (T, rho, U, U err, P, P err)
REFERENCE_ENGINE = [
    (0.85, 0.776, -5.5120, 0.0009, -0.0002, 0.0048),
    (0.85, 0.780, -5.5382, 0.0013, 0.0413, 0.0076),
    (0.85, 0.820, -5.7921, 0.0013, 0.5588, 0.0074),
    (0.85, 0.860, -6.0276, 0.0011, 1.2768, 0.0060),
    (0.85, 0.900, -6.2355, 0.0015, 2.2427, 0.0083),
    (0.90, 0.776, -5.4671, 0.0015, 0.2412, 0.0076),
    (0.90, 0.780, -5.4950, 0.0013, 0.2756, 0.0078),
    (0.90, 0.820, -5.7453, 0.0009, 0.8186, 0.0042),
    (0.90, 0.860, -5.9742, 0.0011, 1.5762, 0.0062),
    (0.90, 0.900, -6.1733, 0.0015, 2.5995, 0.0087),
]


@pytest.fixture(scope="module")
def validation(ctx):
    print(f"nist: 16 state points x {REPLICAS} replicas, N = 500, 10000 + 20000 steps, FP64, graph replay",
          flush=True)
    results, info = nist.run_replicas(ctx, replicas=REPLICAS)
    print(f"nist: {info['replicas']} replicas in {info['seconds']:.2f} s "
          f"({info['atom_steps_per_s']:.3e} atom-steps/s, {info['rebuilds']} rebuilds)", flush=True)
    return results


def test_nist_mc_z_scores(validation):
    nist.compare(validation, nist.NIST_MD, "NIST MD")
    rows = nist.compare(validation, nist.NIST_MC, "NIST MC")
    z = [r[k] for r in rows for k in ("U_z", "P_z")]
    chi2 = sum(v * v for v in z)
    within = sum(abs(v) < 2.0 for v in z)
    print(f"NIST MC: {within}/{len(z)} |z| < 2, chi^2 = {chi2:.1f}", flush=True)
    assert len(z) == 20
    assert max(abs(v) for v in z) < 3.5
    assert chi2 < 45.3


def test_agrees_with_the_reference_engine(validation):
    rows = nist.compare(validation, REFERENCE_ENGINE, "reference engine (README.md:443-454)")
    z = [r[k] for r in rows for k in ("U_z", "P_z")]
    print(f"reference engine: {sum(abs(v) < 2.0 for v in z)}/{len(z)} |z| < 2, chi^2 = {sum(v * v for v in z):.1f}",
          flush=True)
    assert len(z) == 20 and max(abs(v) for v in z) < 3.0


def test_replica_temperatures_hold_the_target(validation):
    for (t, _), a in validation.items():
        assert len(a["t_means"]) == REPLICAS
        for tm in a["t_means"]:
            assert abs(tm - t) < 0.01, (t, tm)
        assert all(math.isfinite(v) for v in a["u_means"] + a["p_means"])
