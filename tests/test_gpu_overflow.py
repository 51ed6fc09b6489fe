"""A neighbour-list row that outgrows its padding inside a device-driven run (NeighbourList.cpp:11-36 regrows the
reference's vectors on every build; here the rows are padded, so a device-side rebuild can overflow them).

Bar: a run started with rows exactly as long as the FCC start needs (78 entries at rc + skin = 2.8, rho 0.8442),
so that the first rebuild of the melting lattice overflows, finishes bit-equal to the same run started with the
automatic padding: state, per-step record, rebuild count. md_run halts its kernels at the overflowing rebuild,
grows the rows and resumes from it; md_batch_run does the same for every replica; md_adjoint_run repeats the
computation with longer rows.
"""
import numpy as np
import pytest

import mdgpu
import oracle as O

pytestmark = pytest.mark.gpu

LJ = mdgpu.LJ(1.0, 1.0, 2.5)
DT = 0.005
FCC_ROW = 78  # full-list neighbours of every FCC site within 2.8 sigma at rho 0.8442


def fcc_gpu(ctx, cells=8, rho=0.8442, temp=0.722, seed=1):
    s = O.Sim.fcc(cells, rho, temp, seed)
    st = s.state()
    g = mdgpu.State.from_arrays(ctx, s.box, st["pos"], vel=st["vel"], mass=st["mass"], force=st["force"])
    g.upload(mdgpu.FIELD_ID, st["ids"])
    return g


def run_pair(ctx, precision, use_graph, steps, draws=None, temp=0.722):
    out = []
    for padding in (0, FCC_ROW):
        g = fcc_gpu(ctx, temp=temp)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.3, max_neighbours=padding)
        nl.build(g)
        assert nl.info().max_neighbours == FCC_ROW
        nl.forces(g, LJ, precision)
        en = nl.run(g, LJ, DT, steps, precision=precision, draws=draws, temperature=temp, tau=0.1,
                    use_graph=use_graph)
        out.append((g, nl, en))
    return out


@pytest.mark.parametrize("precision,use_graph", [(mdgpu.MD_FP64, False), (mdgpu.MD_FP64, True),
                                                 (mdgpu.MD_MIXED, False), (mdgpu.MD_MIXED, True)])
def test_run_overflow_resumes_bit_equal(ctx, precision, use_graph):
    print(f"md_run overflow: precision {precision} graph {use_graph}", flush=True)
    (ga, la, ea), (gb, lb, eb) = run_pair(ctx, precision, use_graph, 200)
    assert lb.info().padding > FCC_ROW  # the short rows overflowed and grew
    assert la.info().rebuilds == lb.info().rebuilds > 3
    for f in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE, mdgpu.FIELD_ID):
        assert np.array_equal(ga.download(f), gb.download(f)), f
    assert np.array_equal(ea, eb)
    assert la.info().half_pairs == lb.info().half_pairs


def test_bussi_run_overflow_resumes_bit_equal(ctx):
    steps = 150
    draws = mdgpu.BussiRng(11).draws(2048, steps)
    (ga, la, ea), (gb, lb, eb) = run_pair(ctx, mdgpu.MD_FP64, True, steps, draws=draws, temp=1.0)
    assert lb.info().padding > FCC_ROW
    assert la.info().rebuilds == lb.info().rebuilds
    assert ga.state_hash() == gb.state_hash()
    assert np.array_equal(ea, eb)


def test_batch_run_overflow_resumes_bit_equal(ctx, monkeypatch):
    specs = [(6, 0.8442, 0.722, 2), (7, 0.8442, 1.0, 3), (8, 0.8442, 0.722, 1)]
    steps = 150
    temps = np.array([t for _, _, t, _ in specs])
    runs = []
    for short in (False, True):
        if short:
            monkeypatch.setenv("MDGPU_BATCH_PADDING", str(FCC_ROW))
        states = [fcc_gpu(ctx, c, rho, t, seed) for c, rho, t, seed in specs]
        batch = mdgpu.Batch(ctx, states, 2.5, 0.3)
        monkeypatch.delenv("MDGPU_BATCH_PADDING", raising=False)
        draws = np.stack([mdgpu.BussiRng(40 + r).draws(states[r].n, steps) for r in range(len(specs))])
        e1 = batch.run(LJ, DT, steps, precision=mdgpu.MD_FP64, use_graph=True)
        e2 = batch.run(LJ, DT, steps, precision=mdgpu.MD_MIXED, temperatures=temps, draws=draws, use_graph=False)
        runs.append((batch, e1, e2))
    (ba, a1, a2), (bb, b1, b2) = runs
    assert ba.info().padding > FCC_ROW and bb.info().padding > FCC_ROW  # the short rows grew
    assert np.array_equal(a1, b1) and np.array_equal(a2, b2)
    assert ba.info().rebuilds == bb.info().rebuilds
    for r, (c, *_rest) in enumerate(specs):
        x = mdgpu.State(ctx, 4 * c ** 3, ba.boxes[r])
        y = mdgpu.State(ctx, 4 * c ** 3, bb.boxes[r])
        ra = ba.replica(r, x)
        rb = bb.replica(r, y)
        assert ra == rb
        assert x.state_hash() == y.state_hash()


def test_adjoint_overflow_repeats_with_longer_rows(ctx):
    out = []
    for padding in (0, FCC_ROW):
        g = fcc_gpu(ctx, 6, temp=0.9)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.3, max_neighbours=padding)
        res, gp, gv = mdgpu.adjoint_run(g, nl, LJ, DT, 80, 20)
        out.append((res, gp, gv, nl.info().padding))
    (ra, gpa, gva, pa), (rb, gpb, gvb, pb) = out
    assert pb > FCC_ROW
    for k in ("loss", "mean_potential", "mean_pressure", "d_epsilon", "d_sigma"):
        assert getattr(ra, k) == getattr(rb, k), k
    assert np.array_equal(gpa, gpb) and np.array_equal(gva, gvb)


def test_list_copy_carries_device_rows(ctx):
    g = fcc_gpu(ctx)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    nl.run(g, LJ, DT, 40, precision=mdgpu.MD_FP64)
    r0 = nl.forces(g, LJ, mdgpu.MD_FP64)
    f0 = g.download(mdgpu.FIELD_FORCE)
    twin = mdgpu.NeighbourList(ctx, 1.0, 0.1)
    twin.copy_from(nl)
    a, b = nl.info(), twin.info()
    for k in ("half_pairs", "rebuilds", "padding", "particles", "max_neighbours", "cells_per_side", "tiled"):
        assert getattr(a, k) == getattr(b, k), k
    nl.close()  # the copy owns its own buffers
    r1 = twin.forces(g, LJ, mdgpu.MD_FP64)
    assert np.array_equal(g.download(mdgpu.FIELD_FORCE), f0)
    assert (r1.potential, r1.virial, r1.inside_pairs) == (r0.potential, r0.virial, r0.inside_pairs)
    r2 = twin.forces(g, LJ, mdgpu.MD_MIXED)
    assert abs(r2.potential - r0.potential) <= 1e-6 * abs(r0.potential)


@pytest.mark.parametrize("precision", [mdgpu.MD_FP64, mdgpu.MD_MIXED])
def test_step_overflow_recovers_bit_equal(ctx, precision):
    """md_step (VelocityVerletIntegrator::step): the rebuild runs gated on the device; rows that overflow halt the
    step's force and kick kernels and the host finishes the step with longer rows. Started short, the per-call
    steps end bit-equal to the same steps started with the automatic padding, and to md_run."""
    print(f"md_step overflow: precision {precision}", flush=True)
    steps = 150
    res = []
    for padding in (0, FCC_ROW):
        g = fcc_gpu(ctx)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.3, max_neighbours=padding)
        nl.build(g)
        nl.forces(g, LJ, precision)
        us = [nl.step(g, LJ, DT, precision).potential for _ in range(steps)]
        res.append((g, nl, np.array(us)))
    (ga, la, ua), (gb, lb, ub) = res
    assert lb.info().padding > FCC_ROW
    assert la.info().rebuilds == lb.info().rebuilds > 3
    for f in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE, mdgpu.FIELD_ID):
        assert np.array_equal(ga.download(f), gb.download(f)), f
    assert np.array_equal(ua, ub)
    assert la.info().half_pairs == lb.info().half_pairs
    # the same steps by md_run (eager launches, the separate-kernel path for FP64): the same state
    g = fcc_gpu(ctx)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    nl.forces(g, LJ, precision)
    en = nl.run(g, LJ, DT, steps, precision=precision, use_graph=False)
    assert nl.info().rebuilds == la.info().rebuilds
    for f in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_ID):
        assert np.array_equal(g.download(f), ga.download(f)), f
    assert np.array_equal(en[:, 0], ua)
