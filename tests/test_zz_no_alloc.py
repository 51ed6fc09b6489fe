"""md_run_static, the allocation-free md_run (its counters, records and Bussi draws in the library's static device
array), called through the C ABI, against md_run on the same inputs: bit-equal state and records, on two calls in
a row, the second with a smaller system and fewer steps (the shared array keeps the first call's contents). An NVE
md_run folds each step into the force kernel once the state has a spare position buffer, which md_run_static cannot
allocate: before that it takes the separate kernels, with the same trajectory and the kinetic energy summed in
another order."""
import ctypes as C

import numpy as np
import pytest

import mdgpu

pytestmark = pytest.mark.gpu

LJ = mdgpu.LJ(1.0, 1.0, 2.5)
DT = 0.005


def config(steps, precision, draws=None, temperature=1.0):
    cfg = mdgpu.RunConfig()
    cfg.dt = DT
    cfg.precision = precision
    cfg.steps = steps
    cfg.lj = LJ
    cfg.thermostat = 0 if draws is None else 1
    cfg.temperature = temperature
    cfg.tau = 0.1
    if draws is not None:
        cfg.draws = draws.ctypes.data_as(C.POINTER(C.c_double))
    cfg.use_graph = 0
    return cfg


def static_run(state, nl, steps, precision, draws=None, temperature=1.0):
    energies = np.empty((steps, 4))
    cfg = config(steps, precision, draws, temperature)
    mdgpu.check(mdgpu.lib().md_run_static(state.h, nl.h, C.byref(cfg), energies.ctypes.data_as(C.POINTER(C.c_double))))
    return energies


def twin(ctx, cells, temp, seed, precision):
    out = []
    for _ in range(2):
        g = mdgpu.fcc_system(ctx, cells, 0.8442, temp, seed)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
        nl.build(g)
        nl.forces(g, LJ, precision)
        out.append((g, nl))
    return out


def assert_same(a, la, ea, b, lb, eb):
    for f in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE, mdgpu.FIELD_ID):
        assert np.array_equal(a.download(f), b.download(f)), f
    assert np.array_equal(ea, eb)
    assert la.info().rebuilds == lb.info().rebuilds
    assert la.info().half_pairs == lb.info().half_pairs


def test_static_run_matches_md_run_twice(ctx):
    # first call: 2048 atoms, FP64 NVE, 300 steps
    print("md_run_static: 2048 atoms FP64 NVE, 300 steps", flush=True)
    (a, la), (b, lb) = twin(ctx, 8, 0.722, 1, mdgpu.MD_FP64)
    ea = la.run(a, LJ, DT, 300, precision=mdgpu.MD_FP64, use_graph=False)
    eb = static_run(b, lb, 300, mdgpu.MD_FP64)  # no spare position buffer on b: the separate kernels
    assert la.info().rebuilds > 3
    # md_run folds each NVE step into the force kernel (the kinetic energy summed per tile); the separate kernels
    # give the same trajectory and records, the kinetic energy summed in another order
    for f in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE, mdgpu.FIELD_ID):
        assert np.array_equal(a.download(f), b.download(f)), f
    assert np.array_equal(ea[:, [0, 2, 3]], eb[:, [0, 2, 3]])
    np.testing.assert_allclose(eb[:, 1], ea[:, 1], rtol=1e-12, atol=0)
    assert la.info().rebuilds == lb.info().rebuilds and la.info().half_pairs == lb.info().half_pairs
    # with a spare buffer on both (a short md_run each), md_run_static folds too: bit-equal records
    la.run(a, LJ, DT, 2, precision=mdgpu.MD_FP64, use_graph=False, record=False)
    lb.run(b, LJ, DT, 2, precision=mdgpu.MD_FP64, use_graph=False, record=False)
    ea = la.run(a, LJ, DT, 100, precision=mdgpu.MD_FP64, use_graph=False)
    eb = static_run(b, lb, 100, mdgpu.MD_FP64)
    assert_same(a, la, ea, b, lb, eb)
    # second call, smaller: 864 atoms, mixed with the Bussi thermostat, 120 steps
    print("md_run_static: 864 atoms mixed Bussi, 120 steps", flush=True)
    (a, la), (b, lb) = twin(ctx, 6, 1.0, 2, mdgpu.MD_MIXED)
    draws = mdgpu.BussiRng(9).draws(a.n, 120)
    ea = la.run(a, LJ, DT, 120, precision=mdgpu.MD_MIXED, draws=draws, temperature=1.0, use_graph=False)
    eb = static_run(b, lb, 120, mdgpu.MD_MIXED, draws=draws, temperature=1.0)
    assert_same(a, la, ea, b, lb, eb)


def test_static_mixed_nve_folded_after_md_run(ctx):
    """Mixed NVE: once an md_run has made the spare position buffer, md_run_static folds the step into the force
    kernel as md_run does; before that it takes the separate kernels (the same trajectory, the kinetic energy
    summed in another order)."""
    print("md_run_static: 2048 atoms mixed NVE", flush=True)
    (a, la), (b, lb) = twin(ctx, 8, 0.722, 4, mdgpu.MD_MIXED)
    ea = la.run(a, LJ, DT, 80, precision=mdgpu.MD_MIXED, use_graph=False)
    eb = static_run(b, lb, 80, mdgpu.MD_MIXED)  # no spare buffer yet: separate kernels
    for f in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE):
        assert np.array_equal(a.download(f), b.download(f)), f
    assert np.array_equal(ea[:, [0, 2, 3]], eb[:, [0, 2, 3]])
    np.testing.assert_allclose(eb[:, 1], ea[:, 1], rtol=1e-12, atol=0)
    lb.run(b, LJ, DT, 2, precision=mdgpu.MD_MIXED, use_graph=False, record=False)
    la.run(a, LJ, DT, 2, precision=mdgpu.MD_MIXED, use_graph=False, record=False)
    ea = la.run(a, LJ, DT, 60, precision=mdgpu.MD_MIXED, use_graph=False)
    eb = static_run(b, lb, 60, mdgpu.MD_MIXED)
    assert_same(a, la, ea, b, lb, eb)


def test_static_run_rejects_what_needs_an_allocation(ctx):
    g = mdgpu.fcc_system(ctx, 6, 0.8442, 0.722, 1)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    with pytest.raises(mdgpu.MdError, match="needs a list built"):
        static_run(g, nl, 10, mdgpu.MD_FP64)
    nl.build(g)
    with pytest.raises(mdgpu.MdError, match="at most 65536 steps"):
        static_run(g, nl, 70000, mdgpu.MD_FP64)
    cfg = config(10, mdgpu.MD_FP64)
    cfg.rdf_every = 5
    counts = np.zeros(10, dtype=np.uint64)
    cfg.rdf_bins, cfg.rdf_rmax, cfg.rdf_counts = 10, 3.0, counts.ctypes.data_as(C.POINTER(C.c_uint64))
    rc = mdgpu.lib().md_run_static(g.h, nl.h, C.byref(cfg), None)
    assert rc == 1 and b"pair histogram" in mdgpu.lib().md_last_error()
