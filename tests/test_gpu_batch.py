"""Batched replicas (md_batch_*) against single-System runs of the same states and against the oracle.

Bars: in FP64 every replica's trajectory (positions, velocities, forces, ids, order after re-sorts) is bit-equal
to md_run on that System alone, NVE and Bussi, which the single-system tests pin to the reference; energies match
to 1e-12 relative (different fixed block order of the reductions); mixed precision max|dF|/max|F| <= 1e-6
against FP64 for every lanes-per-particle variant; graph replay bit-equal to eager launches.
"""
import numpy as np
import pytest

import mdgpu
import oracle as O

pytestmark = pytest.mark.gpu

LJ = mdgpu.LJ(1.0, 1.0, 2.5)
DT = 0.005
SKIN = 0.3

# (FCC cells, density, temperature, seed): cells 4 gives a 2-cell grid (every pair is a candidate), the rest
# 3- and 4-cell grids; uneven sizes 256..2048
SPECS = [(4, 0.84, 0.9, 3), (5, 0.75, 1.2, 4), (6, 0.84, 0.8, 5), (7, 0.9, 1.0, 6), (8, 0.8442, 0.722, 1)]


def gpu_state(ctx, spec):
    cells, rho, temp, seed = spec
    s = O.Sim.fcc(cells, rho, temp, seed)
    st = s.state()
    g = mdgpu.State.from_arrays(ctx, s.box, st["pos"], vel=st["vel"], mass=st["mass"], force=st["force"])
    g.upload(mdgpu.FIELD_ID, st["ids"])
    return g


def single_run(ctx, spec, steps, precision, draws=None):
    g = gpu_state(ctx, spec)
    nl = mdgpu.NeighbourList(ctx, 2.5, SKIN)
    nl.build(g)
    nl.forces(g, LJ, precision)
    en = nl.run(g, LJ, DT, steps, precision=precision, draws=draws, temperature=spec[2], tau=0.1, use_graph=False)
    return g, nl, en


def replica_state(ctx, batch, r):
    dst = mdgpu.State(ctx, batch.n[r], batch.boxes[r])
    rebuilds, pairs = batch.replica(r, dst)
    return dst, rebuilds, pairs


def assert_same_state(a, b):
    for f in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE, mdgpu.FIELD_MASS, mdgpu.FIELD_ID):
        assert np.array_equal(a.download(f), b.download(f)), f


@pytest.mark.parametrize("use_graph", [False, True])
def test_fp64_nve_replicas_match_single_runs(ctx, use_graph):
    steps = 150
    batch = mdgpu.Batch(ctx, [gpu_state(ctx, sp) for sp in SPECS], 2.5, SKIN)
    eb = batch.run(LJ, DT, steps, precision=mdgpu.MD_FP64, use_graph=use_graph)
    for r, sp in enumerate(SPECS):
        g, nl, en = single_run(ctx, sp, steps, mdgpu.MD_FP64)
        got, rebuilds, pairs = replica_state(ctx, batch, r)
        assert_same_state(got, g)
        info = nl.info()
        assert rebuilds == info.rebuilds and pairs == info.half_pairs
        assert np.allclose(eb[:, r, 0], en[:, 0], rtol=1e-12, atol=0)
        assert np.allclose(eb[:, r, 2], en[:, 2], rtol=1e-12, atol=1e-12)
        assert np.array_equal(eb[:, r, 3], en[:, 3])
        assert np.allclose(eb[:, r, 1], en[:, 1], rtol=1e-12, atol=0)


def test_fp64_bussi_replicas_match_single_runs(ctx):
    steps = 120
    draws = np.stack([mdgpu.BussiRng(100 + r).draws(4 * sp[0] ** 3, steps) for r, sp in enumerate(SPECS)])
    temps = np.array([sp[2] for sp in SPECS])
    batch = mdgpu.Batch(ctx, [gpu_state(ctx, sp) for sp in SPECS], 2.5, SKIN)
    eb = batch.run(LJ, DT, steps, precision=mdgpu.MD_FP64, temperatures=temps, draws=draws, use_graph=True)
    for r, sp in enumerate(SPECS):
        g, _, en = single_run(ctx, sp, steps, mdgpu.MD_FP64, draws=draws[r])
        got, _, _ = replica_state(ctx, batch, r)
        assert_same_state(got, g)
        assert np.array_equal(eb[:, r, 1], en[:, 1])  # the thermostat's K' is computed in the same order


def test_mixed_bussi_replicas_follow_single_runs(ctx):
    """Mixed Bussi replicas on the tiled kernels, whose force kernel also applies the closing kick and sums its kinetic
    energy per tile, against md_run of each System (separate kick, the kinetic energy in 1024-particle blocks): the
    same trajectory up to the summation order of the kinetic energy, so over a short run the records agree to
    rounding and the same pairs interact."""
    specs = SPECS[1:]
    steps = 20
    draws = np.stack([mdgpu.BussiRng(200 + r).draws(4 * sp[0] ** 3, steps) for r, sp in enumerate(specs)])
    temps = np.array([sp[2] for sp in specs])
    batch = mdgpu.Batch(ctx, [gpu_state(ctx, sp) for sp in specs], 2.5, SKIN)
    assert batch.info().tiled
    eb = batch.run(LJ, DT, steps, precision=mdgpu.MD_MIXED, temperatures=temps, draws=draws, use_graph=True)
    for r, sp in enumerate(specs):
        _, _, en = single_run(ctx, sp, steps, mdgpu.MD_MIXED, draws=draws[r])
        for k in (0, 1, 2):  # potential, the thermostat's K', virial
            assert np.allclose(eb[:, r, k], en[:, k], rtol=1e-9, atol=1e-9 * np.abs(en[:, k]).max()), k
        assert np.array_equal(eb[:, r, 3], en[:, 3])


def test_fp64_replica_matches_oracle_hash(ctx):
    """One replica against the oracle's deterministic mode directly (40 steps, re-sorts included)."""
    sp = SPECS[2]
    s = O.Sim.fcc(*sp)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, SKIN, reverse=True)
    s.list_build()
    s.forces_list()
    for _ in range(40):
        s.step(DT)
    batch = mdgpu.Batch(ctx, [gpu_state(ctx, SPECS[0]), gpu_state(ctx, sp)], 2.5, SKIN)
    batch.run(LJ, DT, 40, precision=mdgpu.MD_FP64)
    got, rebuilds, _ = replica_state(ctx, batch, 1)
    assert got.state_hash() == s.state_hash()


def liquid_state(ctx, spec):
    """The spec's FCC start melted by 60 oracle steps (forces on a perfect lattice are rounding noise)."""
    cells, rho, temp, seed = spec
    s = O.make_liquid(cells, 60, density=rho, temperature=temp, seed=seed, skin=SKIN)
    st = s.state()
    g = mdgpu.State.from_arrays(ctx, s.box, st["pos"], vel=st["vel"], mass=st["mass"], force=st["force"])
    g.upload(mdgpu.FIELD_ID, st["ids"])
    return g


@pytest.mark.parametrize("lanes", [1, 2, 4])
def test_mixed_forces_within_tolerance(ctx, lanes):
    batch = mdgpu.Batch(ctx, [liquid_state(ctx, sp) for sp in SPECS], 2.5, SKIN)
    batch.run(LJ, DT, 0, precision=mdgpu.MD_MIXED, lanes=lanes)
    for r, sp in enumerate(SPECS):
        g = liquid_state(ctx, sp)
        nl = mdgpu.NeighbourList(ctx, 2.5, SKIN)
        nl.build(g)
        nl.forces(g, LJ, mdgpu.MD_FP64)
        f64 = g.download(mdgpu.FIELD_FORCE)
        got, _, _ = replica_state(ctx, batch, r)
        err = float(np.abs(got.download(mdgpu.FIELD_FORCE) - f64).max() / np.abs(f64).max())
        assert err <= 1e-6, (r, err)


def test_mixed_graph_matches_eager(ctx):
    states_a = [liquid_state(ctx, sp) for sp in SPECS]
    states_b = [liquid_state(ctx, sp) for sp in SPECS]
    a = mdgpu.Batch(ctx, states_a, 2.5, SKIN)
    b = mdgpu.Batch(ctx, states_b, 2.5, SKIN)
    ea = a.run(LJ, DT, 80, precision=mdgpu.MD_MIXED, lanes=2, use_graph=False)
    eb = b.run(LJ, DT, 80, precision=mdgpu.MD_MIXED, lanes=2, use_graph=True)
    assert np.array_equal(ea, eb)
    for r in range(len(SPECS)):
        assert_same_state(replica_state(ctx, a, r)[0], replica_state(ctx, b, r)[0])
    # energy stays conserved in the mixed NVE replicas
    e = ea[:, :, 0] + ea[:, :, 1]
    assert np.all(np.abs(e[-1] - e[0]) <= 2e-3 * np.abs(e[0]))


def test_autotuner_picks_a_variant_and_leaves_the_batch(ctx):
    batch = mdgpu.Batch(ctx, [gpu_state(ctx, sp) for sp in SPECS], 2.5, SKIN)
    twin = mdgpu.Batch(ctx, [gpu_state(ctx, sp) for sp in SPECS], 2.5, SKIN)
    best, ms = batch.tune(LJ, DT, 20, 10, precision=mdgpu.MD_MIXED)
    assert best in (1, 2, 4) and np.all(ms > 0)
    assert batch.info().lanes == best
    # the trials ran on a copy: the batch still equals its untouched twin
    for r in range(len(SPECS)):
        assert_same_state(replica_state(ctx, batch, r)[0], replica_state(ctx, twin, r)[0])


def test_batch_rejects_small_boxes(ctx):
    small = gpu_state(ctx, (3, 0.84, 1.0, 1))  # L = 5.04 < 2 (rc + skin)
    with pytest.raises(mdgpu.MdError, match="Cutoff \\+ skin must be less than half the box length"):
        mdgpu.Batch(ctx, [gpu_state(ctx, SPECS[0]), small], 2.5, SKIN)
