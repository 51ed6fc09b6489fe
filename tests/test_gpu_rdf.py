"""The all-pairs g(r) histogram (cell pass): every pair within rmax, for rmax beyond the list radius (the g(r) to
5 sigma of the 256k-atom NVT workload, 250 bins, next to a 2.8 sigma list). The reference's
RadialDistribution::sample counts only list pairs (RadialDistribution.cpp:11-30), so beyond the list radius the
bar is the brute-force histogram of every pair (the oracle's, O(N^2)); counts are integers and must be bit-exact.
"""
import numpy as np
import pytest

import mdgpu
import oracle as O

pytestmark = pytest.mark.gpu

LJ = mdgpu.LJ(1.0, 1.0, 2.5)
DT = 0.005
RMAX, BINS = 5.0, 250


def oracle_of(g):
    st = g.state()
    return O.Sim.from_arrays(st["pos"][:, :3], g.box, vel=st["vel"].T)


@pytest.mark.parametrize("cells", [8, 20])
def test_all_pairs_histogram_matches_brute_force(ctx, cells):
    """N = 2048 (L = 13.4: two cells of 5 sigma per side, so every j > i is a candidate) and N = 32,000 (six cells
    per side, the 27-cell pass), after 60 steps of melting."""
    g = mdgpu.fcc_system(ctx, cells, 0.8442, 0.722, 1)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    nl.forces(g, LJ, mdgpu.MD_FP64)
    nl.run(g, LJ, DT, 60, precision=mdgpu.MD_FP64, record=False)
    got = np.zeros(BINS, dtype=np.uint64)
    g.rdf_sample_all(RMAX, BINS, got)
    want = oracle_of(g).pair_histogram(RMAX, BINS)
    assert np.array_equal(got, want) and got.sum() > 0
    # md_rdf_sample with rmax past the list radius counts every pair too; within it, the list pairs (all of them
    # right after a build)
    via_list = np.zeros(BINS, dtype=np.uint64)
    nl.rdf_sample(g, RMAX, BINS, via_list)
    assert np.array_equal(via_list, want)
    nl.build(g)
    near = np.zeros(140, dtype=np.uint64)
    nl.rdf_sample(g, 2.8, 140, near)
    assert np.array_equal(near, oracle_of(g).pair_histogram(2.8, 140))


@pytest.mark.parametrize("precision,thermostat", [(mdgpu.MD_FP64, False), (mdgpu.MD_MIXED, False),
                                                  (mdgpu.MD_FP64, True)])
def test_run_samples_every_k_steps(ctx, precision, thermostat):
    """md_run's histogram every 10 steps over 60 steps equals six 10-step runs each followed by md_rdf_sample_all
    (the runs give the same trajectory whether split or not), on the NVE (folded for mixed) and the Bussi path."""
    draws = mdgpu.BussiRng(5).draws(2048, 60) if thermostat else None
    out = []
    for split in (False, True):
        g = mdgpu.fcc_system(ctx, 8, 0.8442, 1.0, 3)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
        nl.build(g)
        nl.forces(g, LJ, precision)
        counts = np.zeros(BINS, dtype=np.uint64)
        if split:
            for k in range(6):
                d = None if draws is None else draws[10 * k:10 * (k + 1)]
                nl.run(g, LJ, DT, 10, precision=precision, draws=d, temperature=1.0, use_graph=False, record=False)
                g.rdf_sample_all(RMAX, BINS, counts)
            samples = 6
        else:
            _, samples = nl.run(g, LJ, DT, 60, precision=precision, draws=draws, temperature=1.0, use_graph=True,
                                record=False, rdf=(10, RMAX, counts))
        out.append((counts, samples, g.state_hash()))
    (a, sa, ha), (b, sb, hb) = out
    assert ha == hb
    assert sa == sb == 6
    assert np.array_equal(a, b) and a.sum() > 0


def test_config2_rdf_to_5_sigma(ctx):
    """256,000 atoms, Bussi tau 0.1, FP64, the histogram to 5 sigma in 250 bins taken inside md_run every 100
    steps: after 100 steps (the state the oracle reproduces, hash-checked) its one sample equals the brute-force
    count of every pair, and g(r) tends to 1 at 5 sigma."""
    steps = 100
    s = O.Sim.fcc(40, 0.8442, 1.0, 1)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    s.bussi(1.0, 0.1, 7)
    st = s.state()
    g = mdgpu.State.from_arrays(ctx, s.box, st["pos"], vel=st["vel"])
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    nl.forces(g, LJ, mdgpu.MD_FP64)
    counts = np.zeros(BINS, dtype=np.uint64)
    draws = mdgpu.BussiRng(7).draws(g.n, steps)
    _, samples = nl.run(g, LJ, DT, steps, precision=mdgpu.MD_FP64, draws=draws, temperature=1.0, tau=0.1,
                        rdf=(100, RMAX, counts))
    assert samples == 1
    print("256k: brute-force pair histogram", flush=True)
    s.list_build()
    s.forces_list()
    for _ in range(steps):
        s.step(DT)
        s.bussi_apply(DT)
    assert g.state_hash() == s.state_hash()
    assert np.array_equal(counts, s.pair_histogram(RMAX, BINS))
    width = RMAX / BINS
    lo = np.arange(BINS) * width
    shell = 4.0 / 3.0 * np.pi * ((lo + width) ** 3 - lo ** 3)
    gr = counts / (0.5 * g.n * (g.n / g.box ** 3) * shell)
    assert abs(gr[-20:].mean() - 1.0) < 0.05
