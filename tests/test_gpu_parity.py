"""GPU parity of the mdgpu HIP path against the CPU oracle (oracle/) and the reference-produced golden vectors.

Bars: cell ids, sort permutation, re-sorted state and neighbour sets bit-exact; FP64 forces bit-exact against the
reference's deterministic mode and O(N^2) path; FP64 potential energy and virial within 1e-12 relative (a
different but fixed summation order); mixed precision max|dF|/max|F| <= 1e-6 against the oracle's mixed mode;
whole FP64 trajectories (integrator, re-sorts, rescaling, Bussi) bit-exact through the state hash.
"""
import math

import numpy as np
import pytest

import mdgpu
import oracle as O

pytestmark = pytest.mark.gpu

LJ = mdgpu.LJ(1.0, 1.0, 2.5)


def to_gpu(ctx, s):
    st = s.state()
    g = mdgpu.State.from_arrays(ctx, s.box, st["pos"], vel=st["vel"], mass=st["mass"], force=st["force"])
    g.upload(mdgpu.FIELD_ID, st["ids"])
    return g


def twin_det(cells, steps, skin=0.2, **kw):
    """An oracle liquid in deterministic mode with a reverse index, and the same state on the GPU."""
    s = O.make_liquid(cells, steps, skin=skin, **kw)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, skin, reverse=True)
    return s


def assert_state_equal(a, b, keys=("pos", "vel", "force", "mass", "ids")):
    for k in keys:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("min_cell", [2.7, 1.3, 0.45, 7.0])
def test_cells_bit_exact(ctx, min_cell):
    s = O.make_liquid(8, 50)
    g = to_gpu(ctx, s)
    a, b = s.cells(min_cell), g.cells(min_cell)
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("cells,skin", [(8, 0.2), (8, 0.3), (4, 0.2), (10, 0.3)])
def test_list_build_bit_exact(ctx, cells, skin):
    s = twin_det(cells, 60, skin=skin)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, skin)
    s.list_build()
    nl.build(g)
    assert_state_equal(s.state(), g.state())
    offs_o, nb_o = s.half_list()
    offs_g, nb_g = nl.half()
    assert np.array_equal(offs_o, offs_g) and np.array_equal(nb_o, nb_g)
    info = nl.info()
    assert info.half_pairs == s.pair_count and info.rebuilds == 1
    # the j < i part of every full row is the reference's reverse neighbours, in increasing index
    counts, half, rows = nl.full()
    roffs, rpairs = s.reverse_index()
    owner = np.repeat(np.arange(s.n), np.diff(offs_o).astype(np.int64))
    for i in range(0, s.n, max(1, s.n // 257)):
        lower = owner[rpairs[roffs[i]:roffs[i + 1]].astype(np.int64)]
        assert np.array_equal(rows[i, half[i]:counts[i]], lower)
    assert counts.sum() == 2 * info.half_pairs


@pytest.mark.parametrize("cells", [6, 8])
def test_fp64_forces_bit_exact(ctx, cells):
    s = twin_det(cells, 80)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.2)
    s.list_build()
    nl.build(g)
    u = s.forces_list()
    r = nl.forces(g, LJ, mdgpu.MD_FP64)
    assert np.array_equal(s.state()["force"], g.download(mdgpu.FIELD_FORCE))
    assert r.potential == pytest.approx(u, rel=1e-12, abs=0)
    assert r.virial == pytest.approx(s.virial, rel=1e-12, abs=0)
    assert r.inside_pairs == s.inside_pairs
    # and against the O(N^2) path, bit for bit
    ua = s.forces_allpairs()
    ra = g.forces_allpairs(LJ)
    assert np.array_equal(s.state()["force"], g.download(mdgpu.FIELD_FORCE))
    assert ra.potential == pytest.approx(ua, rel=1e-12, abs=0) and ra.inside_pairs == s.inside_pairs
    assert ra.virial == pytest.approx(s.virial, rel=1e-12, abs=0)


def test_mixed_forces_within_tolerance(ctx):
    s = twin_det(8, 80)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.2)
    s.list_build()
    nl.build(g)
    s.set_deterministic(False)
    s.set_mixed(True)
    um = s.forces_list()
    fm = s.state()["force"]
    wm, cm = s.virial, s.inside_pairs
    s.set_mixed(False)
    u64 = s.forces_list()
    w64 = s.virial
    f64 = s.state()["force"]
    r = nl.forces(g, LJ, mdgpu.MD_MIXED)
    fg = g.download(mdgpu.FIELD_FORCE)
    scale = np.abs(f64).max()
    assert np.abs(fg - fm).max() <= 1e-6 * scale  # against the oracle's mixed mode
    assert np.abs(fg - f64).max() <= 1e-6 * scale  # and against FP64
    assert r.potential == pytest.approx(um, rel=1e-6) and r.potential == pytest.approx(u64, rel=1e-6)
    # the virial cancels strongly; within 1e-6 of FP64, and no further from the oracle's mixed value than that
    # value is from FP64 (its float row sums) plus the same 1e-6
    assert r.virial == pytest.approx(w64, rel=1e-6)
    assert abs(r.virial - wm) <= abs(wm - w64) + 1e-6 * abs(w64)
    # inside counts may differ only for pairs within float rounding of rc^2
    assert abs(r.inside_pairs - cm) <= 2


def test_kinetic_energy_bit_exact(ctx):
    rng = np.random.default_rng(5)
    n = 100_003
    pos = rng.uniform(0, 50.0, size=(n, 3))
    vel = rng.normal(size=(n, 3))
    mass = rng.uniform(0.5, 2.0, size=n)
    s = O.Sim.from_arrays(pos, 50.0, vel=vel, mass=mass)
    g = mdgpu.State.from_arrays(ctx, 50.0, pos, vel=vel.T, mass=mass)
    assert g.kinetic_energy() == s.kinetic_energy()
    assert g.state_hash() == s.state_hash()


def test_reorder_matches_system_reorder(ctx):
    s = O.make_liquid(5, 20)
    g = to_gpu(ctx, s)
    order = np.random.default_rng(1).permutation(s.n).astype(np.uint64)
    s.reorder(order)
    g.reorder(order)
    assert_state_equal(s.state(), g.state())


def run_det_hash(step_fn, rescale_fn):
    for k in range(1, 401):
        step_fn()
        if k <= 200 and k % 10 == 0:
            rescale_fn()


def test_trajectory_matches_golden_deterministic_hash(ctx, golden):
    """The reference's runAndHash(threads, deterministic=true) (test_determinism.cpp:14-32) on the GPU."""
    s = O.Sim.fcc(6, 0.8442, 0.728, 42)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.2)
    nl.build(g)
    nl.forces(g, LJ, mdgpu.MD_FP64)
    run_det_hash(lambda: nl.step(g, LJ, 0.005, mdgpu.MD_FP64), lambda: g.rescale(0.728))
    assert f"{g.state_hash():016x}" == golden["det_hash_6"]["hash"]
    assert nl.info().rebuilds == golden["det_hash_6"]["rebuilds"]


def test_config1_trace_matches_golden(ctx, golden):
    """N = 2048, rho 0.8442, T 0.722, seed 1, skin 0.3, FP64 NVE, 1000 steps."""
    gd = golden["cfg1_det"]
    s = O.Sim.fcc(8, 0.8442, 0.722, 1)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    r = nl.forces(g, LJ, mdgpu.MD_FP64)
    rows = [[0, r.potential, g.kinetic_energy(), r.virial, r.inside_pairs]]
    for k in range(1, 1001):
        r = nl.step(g, LJ, 0.005, mdgpu.MD_FP64)
        if k % 50 == 0:
            rows.append([k, r.potential, g.kinetic_energy(), r.virial, r.inside_pairs])
    for got, want in zip(rows, gd["trace"]):
        assert got[0] == want[0]
        assert got[2] == want[2]  # kinetic energy: bit-exact trajectory, reference summation order
        assert got[1] == pytest.approx(want[1], rel=1e-12, abs=0)
        assert got[3] == pytest.approx(want[3], rel=1e-12, abs=0)
        assert got[4] == want[4]
    assert nl.info().rebuilds == gd["rebuilds"]
    assert f"{g.state_hash():016x}" == gd["hash"]


@pytest.mark.parametrize("use_graph", [False, True])
def test_run_matches_eager_steps(ctx, use_graph):
    s = O.Sim.fcc(8, 0.8442, 0.722, 1)
    ga, gb = to_gpu(ctx, s), to_gpu(ctx, s)
    la, lb = mdgpu.NeighbourList(ctx, 2.5, 0.3), mdgpu.NeighbourList(ctx, 2.5, 0.3)
    la.build(ga)
    la.forces(ga, LJ, mdgpu.MD_FP64)
    lb.build(gb)
    lb.forces(gb, LJ, mdgpu.MD_FP64)
    eager = [la.step(ga, LJ, 0.005, mdgpu.MD_FP64) for _ in range(300)]
    en = lb.run(gb, LJ, 0.005, 300, precision=mdgpu.MD_FP64, use_graph=use_graph)
    assert ga.state_hash() == gb.state_hash()
    assert la.info().rebuilds == lb.info().rebuilds
    assert np.array_equal(en[:, 0], [e.potential for e in eager])
    assert np.array_equal(en[:, 2], [e.virial for e in eager])
    assert np.array_equal(en[:, 3], [e.inside_pairs for e in eager])
    assert en[-1, 1] == pytest.approx(ga.kinetic_energy(), rel=1e-13)


def test_mixed_graph_run_matches_eager_run(ctx):
    """First md_run of fresh states on the tiled mixed path: the captured step (whose buffers must all exist before
    the capture) replays exactly the eager launches."""
    s = O.make_liquid(10, 40, skin=0.3)
    ga, gb = to_gpu(ctx, s), to_gpu(ctx, s)
    la, lb = mdgpu.NeighbourList(ctx, 2.5, 0.3), mdgpu.NeighbourList(ctx, 2.5, 0.3)
    for g, nl in ((ga, la), (gb, lb)):
        nl.build(g)
        nl.forces(g, LJ, mdgpu.MD_MIXED)
    ea = la.run(ga, LJ, 0.005, 200, precision=mdgpu.MD_MIXED, use_graph=False)
    eb = lb.run(gb, LJ, 0.005, 200, precision=mdgpu.MD_MIXED, use_graph=True)
    assert la.info().tiled == 1
    assert np.array_equal(ea, eb)
    assert ga.state_hash() == gb.state_hash()


@pytest.mark.parametrize("steps,use_graph", [(2, False), (151, False), (150, True), (151, True)])
def test_mixed_run_folded_step_matches_separate_kernels(ctx, monkeypatch, steps, use_graph):
    """Mixed NVE md_run carries each step inside the force kernel, with the positions in two alternating buffers.
    The run with the separate kick-kick-drift kernel (MDGPU_UNFOLDED=1) does the same arithmetic: positions,
    velocities and the final forces agree bit for bit, the list rebuilds at the same steps, and the recorded U, W
    and inside count are equal; the kinetic energy differs only by its summation order."""
    s = O.make_liquid(10, 40, skin=0.3)
    ga, gb = to_gpu(ctx, s), to_gpu(ctx, s)
    la, lb = mdgpu.NeighbourList(ctx, 2.5, 0.3), mdgpu.NeighbourList(ctx, 2.5, 0.3)
    for g, nl in ((ga, la), (gb, lb)):
        nl.build(g)
        nl.forces(g, LJ, mdgpu.MD_MIXED)
    monkeypatch.setenv("MDGPU_UNFOLDED", "1")
    ea = la.run(ga, LJ, 0.005, steps, precision=mdgpu.MD_MIXED, use_graph=use_graph)
    monkeypatch.delenv("MDGPU_UNFOLDED")
    eb = lb.run(gb, LJ, 0.005, steps, precision=mdgpu.MD_MIXED, use_graph=use_graph)
    assert lb.info().tiled == 1
    assert la.info().rebuilds == lb.info().rebuilds and (steps < 100 or lb.info().rebuilds > 2)
    for field in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE, mdgpu.FIELD_ID):
        assert np.array_equal(ga.download(field), gb.download(field))
    assert np.array_equal(ea[:, [0, 2, 3]], eb[:, [0, 2, 3]])
    np.testing.assert_allclose(eb[:, 1], ea[:, 1], rtol=1e-12, atol=0)
    assert ga.state_hash() == gb.state_hash()
    # a second run continues from the state the first left (positions in whichever buffer)
    ea2 = la.run(ga, LJ, 0.005, 31, precision=mdgpu.MD_MIXED, use_graph=False)
    eb2 = lb.run(gb, LJ, 0.005, 31, precision=mdgpu.MD_MIXED, use_graph=False)
    assert np.array_equal(ga.download(mdgpu.FIELD_POS), gb.download(mdgpu.FIELD_POS))
    assert np.array_equal(ea2[:, 0], eb2[:, 0])


def test_mixed_run_folded_step_with_unstaged_tiles(ctx, monkeypatch):
    """The folded step where some tiles' halos exceed the LDS capacity (108k atoms, skin 0.9: 14 cells of 3.6 sigma
    per side hold ~39 particles each, so a tile that crosses a cell column has a halo of up to ~2950 > 2176 slots;
    about two thirds of the tiles take the unstaged gather path, the rest stage their halo): mixed forces within
    1e-6 of FP64 on both paths, and the folded run bit-equal to the separate-kernel run."""
    runs = []
    for unfolded in (True, False):
        if unfolded:
            monkeypatch.setenv("MDGPU_UNFOLDED", "1")
        else:
            monkeypatch.delenv("MDGPU_UNFOLDED")
        g = mdgpu.fcc_system(ctx, 30, 0.8442, 0.722, 1)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.9)
        nl.build(g)
        nl.run(g, LJ, 0.005, 30, precision=mdgpu.MD_FP64, use_graph=False, record=False)
        info = nl.info()
        assert 0 < info.unstaged_tiles < info.tiles  # both paths are taken
        nl.forces(g, LJ, mdgpu.MD_FP64)
        f64 = g.download(mdgpu.FIELD_FORCE)
        nl.forces(g, LJ, mdgpu.MD_MIXED)
        assert np.abs(g.download(mdgpu.FIELD_FORCE) - f64).max() <= 1e-6 * np.abs(f64).max()
        en = nl.run(g, LJ, 0.005, 40, precision=mdgpu.MD_MIXED, use_graph=False)
        assert nl.info().tiled == 1 and nl.info().unstaged_tiles > 0
        runs.append((g, nl, en))
    (ga, la, ea), (gb, lb, eb) = runs
    assert la.info().rebuilds == lb.info().rebuilds
    for field in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE):
        assert np.array_equal(ga.download(field), gb.download(field))
    assert np.array_equal(ea[:, [0, 2, 3]], eb[:, [0, 2, 3]])
    np.testing.assert_allclose(eb[:, 1], ea[:, 1], rtol=1e-12, atol=0)


def test_fp64_tiled_forces_bit_exact_on_every_tile_kind(ctx):
    """FP64 forces through the tiled rows at 108k atoms, skin 0.9, after 30 steps: flat and wrapping tiles that
    stage their halo and tiles that gather through the int32 rows (about two thirds) are all bit-exact against the
    oracle's deterministic mode."""
    g = mdgpu.fcc_system(ctx, 30, 0.8442, 0.722, 1)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.9)
    nl.build(g)
    nl.forces(g, LJ, mdgpu.MD_FP64)
    nl.run(g, LJ, 0.005, 30, precision=mdgpu.MD_FP64, record=False)
    st = g.state()
    s = O.Sim.from_arrays(st["pos"][:, :3], g.box, vel=st["vel"].T)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.9, reverse=True)
    s.list_build()
    u = s.forces_list()
    h = mdgpu.State.from_arrays(ctx, g.box, st["pos"], vel=st["vel"])
    nh = mdgpu.NeighbourList(ctx, 2.5, 0.9)
    nh.build(h)
    info = nh.info()
    assert info.tiled == 1 and 0 < info.unstaged_tiles < info.tiles
    r = nh.forces(h, LJ, mdgpu.MD_FP64)
    assert np.array_equal(h.download(mdgpu.FIELD_FORCE), s.state()["force"])
    assert r.potential == pytest.approx(u, rel=1e-12, abs=0)
    assert r.virial == pytest.approx(s.virial, rel=1e-12, abs=0)
    assert r.inside_pairs == s.inside_pairs


def test_float_halo_build_matches_double_build(ctx, monkeypatch):
    """A 256k-atom perturbed FCC state has 24 cells per side, so its tiled build stages the halo as float frame
    coordinates and re-tests in double the candidates within rounding of the list radius. Its lists equal those of the
    double-halo build (MDGPU_BUILD_F64=1) entry for entry; so do the mixed forces read through the slot rows and a
    60-step mixed run with its device-side rebuilds."""
    out = []
    for f64 in ("1", "0"):
        monkeypatch.setenv("MDGPU_BUILD_F64", f64)
        g = mdgpu.fcc_system(ctx, 40, 0.8442, 0.722, 1)
        pos = g.download(mdgpu.FIELD_POS)
        pos[:, :3] = np.mod(pos[:, :3] + np.random.default_rng(7).normal(scale=0.06, size=(len(pos), 3)), g.box)
        g.upload(mdgpu.FIELD_POS, pos)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
        nl.build(g)
        assert nl.info().cells_per_side >= 20 and nl.info().tiled == 1
        counts, half, rows = nl.full()
        r = nl.forces(g, LJ, mdgpu.MD_MIXED)
        f = g.download(mdgpu.FIELD_FORCE)
        en = nl.run(g, LJ, 0.005, 60, precision=mdgpu.MD_MIXED, use_graph=False)
        out.append((counts, half, rows, f, r.potential, en, g.download(mdgpu.FIELD_POS), nl.info().rebuilds))
    a, b = out
    for k in range(3):
        assert np.array_equal(a[k], b[k])
    assert np.array_equal(a[3], b[3]) and a[4] == b[4]
    assert a[7] == b[7] and a[7] > 1
    assert np.array_equal(a[5], b[5]) and np.array_equal(a[6], b[6])


def test_int32_rows_rebuilt_after_mixed_run(ctx):
    """A mixed md_run rebuilds only the 16-bit slot rows on the device; the int32 rows are materialised from them
    when the FP64 kernel needs them, and hold every pair within the cutoff."""
    s = O.make_liquid(10, 40, skin=0.3)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    nl.forces(g, LJ, mdgpu.MD_MIXED)
    nl.run(g, LJ, 0.005, 150, precision=mdgpu.MD_MIXED, use_graph=True)
    assert nl.info().rebuilds > 2
    r = nl.forces(g, LJ, mdgpu.MD_FP64)
    f = g.download(mdgpu.FIELD_FORCE)
    ra = g.forces_allpairs(LJ)
    fa = g.download(mdgpu.FIELD_FORCE)
    assert r.inside_pairs == ra.inside_pairs
    assert np.abs(f - fa).max() <= 1e-12 * np.abs(fa).max()
    assert r.potential == pytest.approx(ra.potential, rel=1e-12, abs=0)


def test_bussi_run_matches_golden(ctx, golden):
    """FCC 6, T 1.0, Bussi tau 0.1 seed 7, FP64, 300 steps: the device-side thermostat reproduces the reference."""
    s = O.Sim.fcc(6, 0.8442, 1.0, 1)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    nl.forces(g, LJ, mdgpu.MD_FP64)
    draws = mdgpu.BussiRng(7).draws(g.n, 300)
    nl.run(g, LJ, 0.005, 300, precision=mdgpu.MD_FP64, draws=draws, temperature=1.0, tau=0.1)
    assert f"{g.state_hash():016x}" == golden["bussi_det"]["hash"]


def test_bussi_eager_matches_oracle(ctx):
    s = O.Sim.fcc(5, 0.8442, 1.0, 3)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    s.bussi(1.0, 0.1, 9)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    rng = mdgpu.BussiRng(9)
    s.list_build()
    s.forces_list()
    nl.build(g)
    nl.forces(g, LJ, mdgpu.MD_FP64)
    for _ in range(60):
        s.step(0.005)
        s.bussi_apply(0.005)
        nl.step(g, LJ, 0.005, mdgpu.MD_FP64)
        r1, sq = rng.draws(g.n, 1)[0]
        g.bussi_apply(1.0, 0.1, 0.005, r1, sq)
    assert g.state_hash() == s.state_hash()


def test_rdf_counts_match_oracle(ctx):
    s = O.make_liquid(8, 100)
    s.neighbour_list(2.5, 0.3)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    s.list_build()
    nl.build(g)
    bins, rmax = 250, 2.8
    s.rdf(rmax, bins)
    s.rdf_sample()
    _, g_ref = s.rdf_values()
    counts = np.zeros(bins, dtype=np.uint64)
    nl.rdf_sample(g, rmax, bins, counts)
    width = rmax / bins
    n = g.n
    density = n / g.box ** 3
    g_gpu = np.empty(bins)
    for b in range(bins):
        lo = b * width
        hi = lo + width
        shell = 4.0 / 3.0 * math.pi * (hi * hi * hi - lo * lo * lo)
        g_gpu[b] = float(counts[b]) / (1 * 0.5 * n * density * shell)
    assert np.array_equal(g_gpu, g_ref)
    assert counts.sum() > 0


def test_list_padding_regrows(ctx):
    s = twin_det(6, 30)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.2, max_neighbours=16)
    s.list_build()
    nl.build(g)
    info = nl.info()
    assert info.padding >= info.max_neighbours > 16
    s.forces_list()
    nl.forces(g, LJ, mdgpu.MD_FP64)
    assert np.array_equal(s.state()["force"], g.download(mdgpu.FIELD_FORCE))


def test_reference_error_messages(ctx):
    g = mdgpu.State.from_arrays(ctx, 4.0, np.array([[1.0, 2.0, 2.0], [2.0, 2.0, 2.0]]))
    with pytest.raises(mdgpu.MdError, match="Cutoff must be less than half the box length"):
        g.forces_allpairs(LJ)
    s = O.Sim.fcc(3, 0.8442, 0.728, 42)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.2)
    with pytest.raises(mdgpu.MdError, match="Cutoff \\+ skin must be less than half the box length"):
        nl.build(g)


def test_update_rebuilds_only_after_half_skin(ctx):
    s = O.make_liquid(6, 50)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.2)
    nl.build(g)
    p = g.download(mdgpu.FIELD_POS)
    p[0, 0] = s.wrap(p[0, 0] + 0.4 * 0.2)
    g.upload(mdgpu.FIELD_POS, p)
    assert not nl.update(g) and nl.info().rebuilds == 1
    p[0, 0] = s.wrap(p[0, 0] + 0.2 * 0.2)
    g.upload(mdgpu.FIELD_POS, p)
    assert nl.update(g) and nl.info().rebuilds == 2


def test_mixed_precision_conserves_energy(ctx):
    """test_forcefield.cpp:151-167 on the GPU: makeLiquid(5, 500), mixed, 2000 steps, max drift < 3e-4."""
    s = O.make_liquid(5, 500)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.2)
    nl.build(g)
    r = nl.forces(g, LJ, mdgpu.MD_MIXED)
    e0 = g.kinetic_energy() + r.potential
    en = nl.run(g, LJ, 0.005, 2000, precision=mdgpu.MD_MIXED)
    drift = np.abs((en[:, 0] + en[:, 1] - e0) / e0).max()
    assert drift < 3e-4
    assert nl.info().rebuilds > 1


def test_config2_size_invariants(ctx):
    """N = 256,000 (FCC 40^3 x 4): list symmetric, every sampled row equal to a brute-force search, forces
    equal to the oracle's deterministic list forces, total force zero."""
    s = O.Sim.fcc(40, 0.8442, 1.0, 1)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    s.list_build()
    nl.build(g)
    st = g.state()
    assert np.array_equal(s.state()["pos"], st["pos"])
    counts, half, rows = nl.full()
    assert counts.sum() == 2 * nl.info().half_pairs
    pos = st["pos"][:, :3]
    L = g.box
    for i in np.random.default_rng(2).choice(g.n, 200, replace=False):
        d = pos[i] - pos
        d -= L * np.round(d / L)
        r2 = (d * d).sum(1)
        want = set(np.nonzero(r2 < 2.8 * 2.8)[0].tolist()) - {int(i)}
        assert set(rows[i, :counts[i]].tolist()) == want
    s.forces_list()
    r = nl.forces(g, LJ, mdgpu.MD_FP64)
    f = g.download(mdgpu.FIELD_FORCE)
    assert np.array_equal(s.state()["force"], f)
    assert np.all(np.abs(f.sum(axis=1)) < 1e-9)
    assert r.inside_pairs == s.inside_pairs


def test_mixed_forces_large_box_within_tolerance(ctx):
    """N = 256,000 with every particle displaced (sigma 0.06) from FCC: the box is wide enough (24 cells) that
    most tiles stage without the periodic image term, and the rest (near the faces, or spanning a cell column
    end) with it. Mixed forces, energy and virial against the FP64 list path (bit-exact to the reference elsewhere),
    then a short mixed md_run against the same run in FP64."""
    s = O.Sim.fcc(40, 0.8442, 0.722, 1)
    st = s.state()
    pos = st["pos"].copy()
    pos[:, :3] = np.mod(pos[:, :3] + np.random.default_rng(5).normal(scale=0.06, size=(len(pos), 3)), s.box)
    def state():
        g = mdgpu.State.from_arrays(ctx, s.box, pos, vel=st["vel"], mass=st["mass"], force=st["force"])
        g.upload(mdgpu.FIELD_ID, st["ids"])
        return g
    g = state()
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    r64 = nl.forces(g, LJ, mdgpu.MD_FP64)
    f64 = g.download(mdgpu.FIELD_FORCE)
    rm = nl.forces(g, LJ, mdgpu.MD_MIXED)
    fm = g.download(mdgpu.FIELD_FORCE)
    assert np.abs(fm - f64).max() <= 1e-6 * np.abs(f64).max()
    assert rm.potential == pytest.approx(r64.potential, rel=1e-6)
    assert rm.virial == pytest.approx(r64.virial, rel=1e-6)
    assert abs(rm.inside_pairs - r64.inside_pairs) <= 2
    # 60 steps (a few rebuilds): mixed trajectory stays within 1e-6 of FP64 in energy
    ga, gb = state(), state()
    la, lb = mdgpu.NeighbourList(ctx, 2.5, 0.3), mdgpu.NeighbourList(ctx, 2.5, 0.3)
    la.build(ga)
    lb.build(gb)
    la.forces(ga, LJ, mdgpu.MD_FP64)
    lb.forces(gb, LJ, mdgpu.MD_MIXED)
    ea = la.run(ga, LJ, 0.005, 60, precision=mdgpu.MD_FP64)
    eb = lb.run(gb, LJ, 0.005, 60, precision=mdgpu.MD_MIXED)
    tot_a, tot_b = ea[:, 0] + ea[:, 1], eb[:, 0] + eb[:, 1]
    assert np.abs(tot_b - tot_a).max() <= 1e-6 * np.abs(tot_a).max()
