"""Batch autotuner (md_batch_autotune) and list-radius changes (md_batch_set_skin).

Bars: after a skin change every replica, run in FP64, is bit-equal to md_run of that replica's state with a list
of the same skin (the batch/md_run equivalence the batch tests pin, under the new layout); runs of one state under
different skins agree in energy to 1e-10 relative over 60 FP64 steps (the same pair sums, summed in the orders
the re-sorts leave: rounding only); mixed runs after tuning stay within 1e-6 of FP64 in energy; the tuner keeps
every candidate's trial, picks the fastest and leaves the batch with that skin and padding.
"""
import numpy as np
import pytest

import mdgpu
import oracle as O

pytestmark = pytest.mark.gpu

LJ = mdgpu.LJ(1.0, 1.0, 2.5)
DT = 0.005
SPECS = [(4, 0.84, 0.9, 3), (5, 0.75, 1.2, 4), (6, 0.84, 0.8, 5), (7, 0.9, 1.0, 6), (8, 0.8442, 0.722, 1)]


def gpu_state(ctx, spec):
    cells, rho, temp, seed = spec
    s = O.Sim.fcc(cells, rho, temp, seed)
    st = s.state()
    g = mdgpu.State.from_arrays(ctx, s.box, st["pos"], vel=st["vel"], mass=st["mass"], force=st["force"])
    g.upload(mdgpu.FIELD_ID, st["ids"])
    return g


def fields(s):
    return [s.download(f) for f in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE, mdgpu.FIELD_MASS,
                                    mdgpu.FIELD_ID)]


def test_set_skin_replicas_match_single_runs_with_that_skin(ctx):
    batch = mdgpu.Batch(ctx, [gpu_state(ctx, sp) for sp in SPECS], 2.5, 0.3)
    batch.run(LJ, DT, 20, precision=mdgpu.MD_FP64, use_graph=False)
    batch.set_skin(0.55)
    # each replica's state as the relayout left it (re-sorted by the new cells): its own single-System run
    singles = []
    for r in range(len(SPECS)):
        g = mdgpu.State(ctx, batch.n[r], batch.boxes[r])
        batch.replica(r, g)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.55)
        nl.build(g)  # already sorted by these cells: the re-sort keeps the order
        nl.forces(g, LJ, mdgpu.MD_FP64)
        singles.append((g, nl))
    steps = 80
    eb = batch.run(LJ, DT, steps, precision=mdgpu.MD_FP64, use_graph=True)
    for r, (g, nl) in enumerate(singles):
        en = nl.run(g, LJ, DT, steps, precision=mdgpu.MD_FP64, use_graph=False)
        got = mdgpu.State(ctx, batch.n[r], batch.boxes[r])
        rebuilds, _ = batch.replica(r, got)
        for a, b in zip(fields(got), fields(g)):
            assert np.array_equal(a, b)
        assert np.allclose(eb[:, r, 0], en[:, 0], rtol=1e-12, atol=0)


def test_skins_agree_to_rounding(ctx):
    """The same replicas from the same state under skins 0.3 and 0.6: energies within 1e-10 relative."""
    steps = 60
    runs = []
    for skin in (0.3, 0.6):
        batch = mdgpu.Batch(ctx, [gpu_state(ctx, sp) for sp in SPECS], 2.5, skin)
        runs.append(batch.run(LJ, DT, steps, precision=mdgpu.MD_FP64, use_graph=False))
    a, b = runs
    e_a = a[:, :, 0] + a[:, :, 1]
    e_b = b[:, :, 0] + b[:, :, 1]
    assert np.max(np.abs(e_a - e_b) / np.abs(e_a)) < 1e-10
    assert np.array_equal(a[:, :, 3], b[:, :, 3])  # inside-pair counts: the same pairs interact


def test_autotune_picks_fastest_and_applies_it(ctx):
    specs = SPECS * 4
    temps = np.array([sp[2] for sp in specs])
    steps = 30
    draws = np.stack([mdgpu.BussiRng(200 + r).draws(4 * sp[0] ** 3, steps) for r, sp in enumerate(specs)])
    batch = mdgpu.Batch(ctx, [gpu_state(ctx, sp) for sp in specs], 2.5, 0.3)
    skins = [0.2, 0.3, 0.45]
    res = batch.autotune(LJ, DT, steps, 20, skins=skins, headroom=[1.15, 1.5], precision=mdgpu.MD_MIXED,
                         temperatures=temps, draws=draws)
    trials = res["trials"]
    assert len(trials) == len(skins) + 2
    assert [t["skin"] for t in trials[:3]] == skins
    assert all(t["ms_per_step"] > 0 for t in trials)
    best = min(trials, key=lambda t: t["ms_per_step"])
    assert res["ms_per_step"] == best["ms_per_step"]
    assert res["skin"] == best["skin"]
    assert batch.info().padding == res["padding"]
    # the tuned batch still runs the same physics: mixed within 1e-6 of FP64 in energy over 40 steps, from the
    # tuned state, against a twin batch run in FP64
    twin_states = []
    for r in range(len(specs)):
        g = mdgpu.State(ctx, batch.n[r], batch.boxes[r])
        batch.replica(r, g)
        twin_states.append(g)
    twin = mdgpu.Batch(ctx, twin_states, 2.5, 0.3)
    e_mixed = batch.run(LJ, DT, 40, precision=mdgpu.MD_MIXED, use_graph=True)
    e_fp64 = twin.run(LJ, DT, 40, precision=mdgpu.MD_FP64, use_graph=True)
    tot_m = e_mixed[:, :, 0] + e_mixed[:, :, 1]
    tot_d = e_fp64[:, :, 0] + e_fp64[:, :, 1]
    assert np.max(np.abs(tot_m - tot_d) / np.abs(tot_d)) < 1e-6


def test_set_skin_rejects_a_radius_beyond_half_the_box(ctx):
    batch = mdgpu.Batch(ctx, [gpu_state(ctx, sp) for sp in SPECS[:2]], 2.5, 0.3)
    with pytest.raises(mdgpu.MdError, match="Cutoff \\+ skin must be less than half the box length"):
        batch.set_skin(5.0)
