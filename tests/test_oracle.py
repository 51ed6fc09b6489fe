"""Pins the CPU oracle (oracle/liboracle.so): against the reference's own code built in place (oracle/_ref),
against golden vectors that code produced (tests/golden/reference_vectors.json), and against the closed forms
and invariants of the reference's test suite (tests/test_*.cpp). CPU only."""
import itertools
import math

import numpy as np
import pytest

import oracle as O

HAS_REF = O.available("ref")
needs_ref = pytest.mark.skipif(not HAS_REF, reason="oracle/_ref not built (no reference checkout)")


def lj_force(r):
    return 24.0 * (2.0 * r ** -13 - r ** -7)


def lj_energy(r, rc=2.5):
    v = lambda x: 4.0 * (x ** -12 - x ** -6)  # noqa: E731
    return v(r) - v(rc)


def pair_at(x0, x1, box):
    return O.Sim.from_arrays([[x0, 5.0, 5.0], [x1, 5.0, 5.0]], box)


# ---- closed forms (reference tests/test_forcefield.cpp, test_system.cpp, test_dynamics.cpp, test_observables.cpp)

@pytest.mark.parametrize("r", [0.95, 1.12, 1.5, 2.0, 2.45])
def test_two_particle_force_and_energy(r):
    s = pair_at(5.0, 5.0 + r, 20.0)
    s.forcefield(1.0, 1.0, 2.5)
    u = s.forces_allpairs()
    f = s.state()["force"]
    assert f[0, 0] == pytest.approx(-lj_force(r), abs=1e-10)
    assert f[0, 1] == pytest.approx(lj_force(r), abs=1e-10)
    assert abs(f[1, 0]) < 1e-12
    assert u == pytest.approx(lj_energy(r), abs=1e-12)


def test_force_vanishes_at_minimum_and_beyond_cutoff():
    s = pair_at(5.0, 5.0 + 2.0 ** (1 / 6), 20.0)
    s.forcefield(1.0, 1.0, 2.5)
    s.forces_allpairs()
    assert abs(s.state()["force"][0, 0]) < 1e-12
    s = pair_at(5.0, 5.0 + 2.5 + 1e-9, 20.0)
    s.forcefield(1.0, 1.0, 2.5)
    assert s.forces_allpairs() == 0.0
    assert s.state()["force"][0, 0] == 0.0


def test_interacts_across_boundary_and_rejects_large_cutoff():
    s = pair_at(0.5, 19.5, 20.0)
    s.forcefield(1.0, 1.0, 2.5)
    u = s.forces_allpairs()
    assert s.state()["force"][0, 0] == pytest.approx(lj_force(1.0), abs=1e-10)
    assert u == pytest.approx(lj_energy(1.0), abs=1e-12)
    s = pair_at(1.0, 2.0, 4.0)
    s.forcefield(1.0, 1.0, 2.5)
    with pytest.raises(O.OracleError, match="Cutoff must be less than half the box length"):
        s.forces_allpairs()


def test_two_particle_virial_and_inside_count():
    s = O.Sim.from_arrays([[5.0, 5.0, 5.0], [6.3, 5.0, 5.0]], 20.0)
    s.forcefield(1.0, 1.0, 2.5)
    s.forces_allpairs()
    assert s.virial == pytest.approx(1.3 * lj_force(1.3), abs=1e-12)
    assert s.inside_pairs == 1.0


def test_minimum_image_and_wrap():
    s = O.Sim.from_arrays(np.zeros((0, 3)), 10.0)
    assert np.allclose(s.min_image([0.5] * 3, [9.5] * 3), [1.0] * 3, atol=1e-12)
    assert np.allclose(s.min_image([3.0, 4.0, 5.0], [1.0, 6.0, 5.0]), [2.0, -2.0, 0.0], atol=1e-12)
    s7 = O.Sim.from_arrays(np.zeros((0, 3)), 7.0)
    for step in range(109):
        x = -20.0 + 0.37 * step
        assert np.all(np.abs(s7.min_image([x, 2 * x, -x], [0, 0, 0])) <= 3.5 + 1e-12)
    assert s.wrap(-0.5) == pytest.approx(9.5, abs=1e-12)
    assert s.wrap(10.0) == pytest.approx(0.0, abs=1e-12)
    assert s.wrap(23.0) == pytest.approx(3.0, abs=1e-12)
    for step in range(96):
        x = -35.0 + 0.73 * step
        for c in (x, -x, 0.5 * x):
            w = s.wrap(c)
            assert 0.0 <= w < 10.0


def test_kinetic_energy_temperature_and_reorder():
    s = O.Sim.from_arrays([[0, 0, 0], [1, 0, 0]], 10.0, vel=[[1, 0, 0], [-1, 0, 0]], mass=[1.0, 2.0])
    assert s.kinetic_energy() == 1.5
    assert s.temperature() == 2.0 * 1.5 / 3.0
    s = O.Sim.from_arrays([[float(i), 0, 0] for i in range(4)], 10.0)
    order = [2, 0, 3, 1]
    s.reorder(order)
    assert list(s.state()["pos"][:, 0]) == [float(o) for o in order]


def test_fcc_spacing_temperature_and_momentum():
    s = O.Sim.lattice(3, 1.5, 4.5)
    p = s.state()["pos"][:, :3]
    dmin = min(np.linalg.norm(s.min_image(p[i], p[j])) for i, j in itertools.combinations(range(len(p)), 2))
    assert len(p) == 108 and dmin == pytest.approx(1.5 / math.sqrt(2.0), abs=1e-12)
    s = O.Sim.fcc(5, 0.8442, 0.728, 42)
    assert s.temperature() == pytest.approx(0.728, abs=1e-12)
    st = s.state()
    assert np.all(np.abs((st["vel"] * st["mass"]).sum(axis=1)) < 1e-12)
    s.rescale(1.7)
    assert s.temperature() == pytest.approx(1.7, abs=1e-12)


# ---- brute force and invariants (test_neighbourlist.cpp, test_forcefield.cpp, test_dynamics.cpp)

def brute_pairs(s, radius):
    p = s.state()["pos"][:, :3]
    L = s.box
    d = p[:, None, :] - p[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d * d).sum(-1)
    i, j = np.nonzero(np.triu(r2 < radius * radius, 1))
    return set(zip(i.tolist(), j.tolist()))


@pytest.mark.parametrize("cells", [4, 8])
def test_half_list_equals_brute_force_set(cells):
    s = O.make_liquid(cells, 50)
    s.neighbour_list(2.5, 0.2)
    s.list_build()
    offs, neigh = s.half_list()
    got = set()
    for i in range(s.n):
        for j in neigh[offs[i]:offs[i + 1]]:
            assert j > i
            got.add((i, int(j)))
    assert got == brute_pairs(s, 2.7)
    assert s.pair_count == len(got)


def test_list_forces_match_allpairs_and_total_force_zero():
    s = O.make_liquid(6, 100)
    s.neighbour_list(2.5, 0.2, reverse=True)
    s.list_build()
    for det in (False, True):
        s.set_deterministic(det)
        u = s.forces_list()
        f_list = s.state()["force"].copy()
        w, c = s.virial, s.inside_pairs
        u_ref = s.forces_allpairs()
        f_ref = s.state()["force"]
        assert u == pytest.approx(u_ref, rel=1e-12)
        assert np.abs(f_list - f_ref).max() < 1e-11
        assert w == pytest.approx(s.virial, rel=1e-12) and c == s.inside_pairs
        assert np.all(np.abs(f_list.sum(axis=1)) < 1e-10)


def test_mixed_matches_double():
    s = O.make_liquid(6, 100)
    s.neighbour_list(2.5, 0.2)
    s.list_build()
    u = s.forces_list()
    f = s.state()["force"].copy()
    s.set_mixed(True)
    um = s.forces_list()
    fm = s.state()["force"]
    assert um == pytest.approx(u, rel=1e-5)
    assert np.abs(fm - f).max() < 1e-5 * np.abs(f).max()


def test_rebuild_only_after_half_skin():
    s = O.make_liquid(6, 50)
    s.neighbour_list(2.5, 0.2)
    s.list_build()
    b = s.rebuilds
    p = s.state()["pos"]
    p[0, 0] = s.wrap(p[0, 0] + 0.4 * 0.2)
    s.set_positions(p)
    assert not s.list_update() and s.rebuilds == b
    p[0, 0] = s.wrap(p[0, 0] + 0.2 * 0.2)
    s.set_positions(p)
    assert s.list_update() and s.rebuilds == b + 1


def test_nve_conservation_and_momentum():
    s = O.make_liquid(5, 500)
    s.neighbour_list(2.5, 0.2)
    s.list_build()
    u = s.forces_list()
    e0 = s.kinetic_energy() + u
    drift = 0.0
    for _ in range(2000):
        u = s.step(0.005)
        drift = max(drift, abs((s.kinetic_energy() + u - e0) / e0))
    assert drift < 3e-4 and s.rebuilds > 1
    st = s.state()
    assert np.all(np.abs((st["vel"] * st["mass"]).sum(axis=1)) < 1e-10)


def test_time_reversible_allpairs():
    s = O.make_liquid(4, 200)
    s.forces_allpairs()
    start = s.state()["pos"][:, :3].copy()
    for _ in range(100):
        s.step_allpairs(0.005)
    v = s.state()["vel"]
    s.set_velocities(-v)
    for _ in range(100):
        s.step_allpairs(0.005)
    end = s.state()["pos"][:, :3]
    err = max(np.abs(s.min_image(end[i], start[i])).max() for i in range(s.n))
    assert err < 1e-8


def test_bussi_canonical_moments():
    s = O.Sim.fcc(5, 0.8442, 0.728, 42)
    target, dof = 0.85, 3.0 * s.n - 3.0
    s.bussi(target, 0.05, 7)
    ks = []
    for k in range(100000):
        s.bussi_apply(0.005)
        if k >= 1000:
            ks.append(s.kinetic_energy())
    ks = np.array(ks)
    assert ks.mean() / (0.5 * dof) == pytest.approx(target, rel=0.005)
    assert ks.var() / (0.5 * dof * target * target) == pytest.approx(1.0, abs=0.05)


def test_rdf_ideal_gas_is_one():
    # 40 ideal-gas samples (test_observables.cpp:71-96); every sample has the same N and box, so the mean of the
    # single-sample g(r) equals the 40-sample g(r) of one accumulating RadialDistribution
    rng = np.random.default_rng(3)
    g_sum = np.zeros(25)
    for _ in range(40):
        s = O.Sim.from_arrays(rng.uniform(0.0, 12.0, size=(2000, 3)), 12.0)
        s.neighbour_list(2.5, 0.2)
        s.list_build()
        s.rdf(2.5, 25)
        s.rdf_sample()
        r, g = s.rdf_values()
        g_sum += g
    g = g_sum / 40
    assert np.all(np.abs(g[r > 1.0] - 1.0) < 0.06)


# ---- the oracle against the reference's own code

@needs_ref
def test_oracle_matches_reference_bitwise_on_liquid():
    O.set_threads("ref", 1)
    a, b = O.make_liquid(6, 300, "oracle"), O.make_liquid(6, 300, "ref")
    sa, sb = a.state(), b.state()
    for k in ("pos", "vel", "force", "ids"):
        assert np.array_equal(sa[k], sb[k]), k
    assert a.state_hash() == b.state_hash()


@needs_ref
def test_oracle_cells_and_lists_match_reference():
    O.set_threads("ref", 3)
    liq = O.make_liquid(8, 30, "oracle")
    st = liq.state()
    a, b = (O.Sim.from_arrays(st["pos"][:, :3], liq.box, vel=st["vel"].T, kind=k) for k in ("oracle", "ref"))
    for kind_a, kind_b in [(a, b)]:
        ca, cb = kind_a.cells(2.7), kind_b.cells(2.7)
        assert ca[0] == cb[0]
        for x, y in zip(ca[1:], cb[1:]):
            assert np.array_equal(x, y)
    for s in (a, b):
        s.neighbour_list(2.5, 0.3, reverse=True)
        s.list_build()
    assert np.array_equal(a.state()["pos"], b.state()["pos"])
    for x, y in zip(a.half_list(), b.half_list()):
        assert np.array_equal(x, y)
    for x, y in zip(a.reverse_index(), b.reverse_index()):
        assert np.array_equal(x, y)
    for side, cell in [(4, 0), (5, 124), (7, 57)]:
        assert np.array_equal(a.stencil(side, cell), b.stencil(side, cell))


@needs_ref
@pytest.mark.parametrize("mode", ["allpairs", "deterministic", "default", "mixed"])
def test_oracle_forces_match_reference(mode):
    O.set_threads("ref", 4)
    liq = O.make_liquid(6, 100, "oracle")
    st = liq.state()
    a, b = (O.Sim.from_arrays(st["pos"][:, :3], liq.box, vel=st["vel"].T, kind=k) for k in ("oracle", "ref"))
    res = []
    for s in (a, b):
        s.neighbour_list(2.5, 0.2, reverse=True)
        s.list_build()
        s.set_deterministic(mode == "deterministic")
        s.set_mixed(mode == "mixed")
        u = s.forces_allpairs() if mode == "allpairs" else s.forces_list()
        res.append((u, s.virial, s.inside_pairs, s.state()["force"]))
    (ua, wa, ca, fa), (ub, wb, cb, fb) = res
    assert ca == cb
    if mode in ("allpairs", "deterministic"):
        assert ua == ub and wa == wb and np.array_equal(fa, fb)
    else:  # per-thread arrays and SIMD reductions: equal to rounding
        tol = 1e-13 if mode == "default" else 1e-6
        assert ua == pytest.approx(ub, rel=tol) and wa == pytest.approx(wb, rel=tol)
        assert np.abs(fa - fb).max() <= tol * np.abs(fb).max()


@needs_ref
def test_oracle_bussi_and_rdf_match_reference():
    O.set_threads("ref", 2)
    out = []
    for kind in ("oracle", "ref"):
        s = O.Sim.fcc(5, 0.8442, 1.0, 1, kind)
        s.forcefield(1.0, 1.0, 2.5, deterministic=True)
        s.neighbour_list(2.5, 0.3, reverse=True)
        s.bussi(1.0, 0.1, 11)
        s.rdf(2.8, 100)
        s.list_build()
        s.forces_list()
        for k in range(1, 101):
            s.step(0.005)
            s.bussi_apply(0.005)
            if k % 10 == 0:
                s.rdf_sample()
        out.append((s.state_hash(), s.rdf_values()[1]))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


# ---- the oracle against the reference-produced golden vectors

def test_oracle_reproduces_golden_det_hash(golden):
    s = O.Sim.fcc(6, 0.8442, 0.728, 42)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.2, reverse=True)
    s.list_build()
    s.forces_list()
    for k in range(1, 401):
        s.step(0.005)
        if k <= 200 and k % 10 == 0:
            s.rescale(0.728)
    assert f"{s.state_hash():016x}" == golden["det_hash_6"]["hash"]
    assert s.rebuilds == golden["det_hash_6"]["rebuilds"]


def test_oracle_reproduces_golden_cfg1(golden):
    g = golden["cfg1_det"]
    s = O.Sim.fcc(8, 0.8442, 0.722, 1)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    s.list_build()
    u = s.forces_list()
    assert [u, s.kinetic_energy(), s.virial, s.inside_pairs] == g["trace"][0][1:]
    for k in range(1, 201):
        u = s.step(0.005)
        if k % 50 == 0:
            assert [k, u, s.kinetic_energy(), s.virial, s.inside_pairs] == g["trace"][k // 50]


def test_oracle_reproduces_golden_bussi(golden):
    s = O.Sim.fcc(6, 0.8442, 1.0, 1)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    s.bussi(1.0, 0.1, 7)
    s.list_build()
    s.forces_list()
    for _ in range(300):
        s.step(0.005)
        s.bussi_apply(0.005)
    assert f"{s.state_hash():016x}" == golden["bussi_det"]["hash"]


# ---- the oracle's two additions without a reference function ------------------------------------------------

def test_fixed_order_list_pairs_forces_and_rebuild_independence():
    """The fixed-order mode keeps the particle order, lists exactly the reference's pair set, and gives forces
    equal to the deterministic mode's to rounding. Its per-particle sums run over the pairs in increasing j
    (j > i part, then j < i part) and entries beyond rc add exact zeros, so a trajectory does not depend on when
    the list was rebuilt: forcing a rebuild after every step leaves the state hash unchanged."""
    base = O.make_liquid(6, 60, skin=0.3)
    ref = base.clone()
    ref.forcefield(1.0, 1.0, 2.5, deterministic=True)
    ref.neighbour_list(2.5, 0.3, reverse=True)
    ref.list_build()
    ref.forces_list()
    fx = base.clone()
    fx.forcefield(1.0, 1.0, 2.5)
    fx.neighbour_list(2.5, 0.3, fixed_order=True)
    ids_before = fx.state()["ids"].copy()
    fx.list_build()
    assert np.array_equal(fx.state()["ids"], ids_before)  # no re-sort
    offs, nb = fx.half_list()
    owner = np.repeat(np.arange(fx.n), np.diff(offs).astype(np.int64))
    ids = fx.state()["ids"]
    pairs_fixed = {(min(a, b), max(a, b)) for a, b in zip(ids[owner], ids[nb.astype(np.int64)])}
    ro, rn = ref.half_list()
    rown = np.repeat(np.arange(ref.n), np.diff(ro).astype(np.int64))
    rids = ref.state()["ids"]
    pairs_ref = {(min(a, b), max(a, b)) for a, b in zip(rids[rown], rids[rn.astype(np.int64)])}
    assert pairs_fixed == pairs_ref
    for i in range(fx.n):  # j > i part in increasing j
        row = nb[offs[i]:offs[i + 1]]
        assert np.all(row > i) and np.all(np.diff(row.astype(np.int64)) > 0)
    u = fx.forces_list()
    f = fx.state()["force"][:, np.argsort(ids)]
    fr = ref.state()["force"][:, np.argsort(rids)]
    assert np.abs(f - fr).max() <= 1e-13 * np.abs(fr).max()
    assert u == pytest.approx(ref.forces_list(), rel=1e-13)
    a, b = fx.clone(), fx.clone()
    for _ in range(80):
        a.step(0.005)
        b.step(0.005)
        b.list_build()
    assert a.rebuilds < b.rebuilds and a.state_hash() == b.state_hash()


def test_pair_histogram_matches_reference_rdf_counts_and_fcc_shells():
    """Brute-force pair histogram: below the list radius it equals the reference's list-pair counts right after
    a build (every pair within rmax is listed then); on a perfect FCC lattice it counts the coordination shells
    12, 6, 24, 12, 24, 8, 48 (N / 2 pairs per neighbour)."""
    kind = "ref" if HAS_REF else "oracle"
    s = O.make_liquid(6, 40, kind=kind, skin=0.3)
    s.neighbour_list(2.5, 0.3)
    s.list_build()
    bins, rmax = 140, 2.8
    s.rdf(rmax, bins)
    s.rdf_sample()
    _, g = s.rdf_values()
    st = s.state()
    o = O.Sim.from_arrays(st["pos"][:, :3], s.box, vel=st["vel"].T)
    counts = o.pair_histogram(rmax, bins)
    width = rmax / bins
    n = s.n
    density = n / s.box ** 3
    for b in range(bins):  # the reference's normalisation (RadialDistribution.cpp:45-60), one sample
        lo = b * width
        hi = lo + width
        shell = 4.0 / 3.0 * math.pi * (hi * hi * hi - lo * lo * lo)
        assert float(counts[b]) / (1 * 0.5 * n * density * shell) == g[b]
    assert counts.sum() > 0
    a = (4.0 / 0.8442) ** (1.0 / 3.0)
    lat = O.Sim.lattice(6, a, 6 * a)
    h = lat.pair_histogram(1.95 * a, 390)
    radii = (np.arange(390) + 0.5) * (1.95 * a / 390)
    got = {}
    for b in np.nonzero(h)[0]:
        got[round(radii[b] / a, 2)] = got.get(round(radii[b] / a, 2), 0) + int(h[b])
    shells = {0.71: 12, 1.0: 6, 1.22: 24, 1.41: 12, 1.58: 24, 1.73: 8, 1.87: 48}
    assert sorted(got) == sorted(shells)
    for r, c in shells.items():
        assert got[r] == c * lat.n // 2
