"""Parity at the sizes of the benchmark configurations (the CPU oracle runs them with its threads; its results do
not depend on the thread count).

- 4,000,000 atoms (FCC 100^3 x 4, the throughput workload), after 100 mixed steps on the GPU: the re-sort and the
  half list bit-exact against the oracle, FP64 forces bit-exact against the reference's deterministic mode,
  mixed forces within 1e-6 of FP64 (max|dF|/max|F|), and 200 mixed steps whose total energy stays within 1e-6 of
  the same 200 steps in FP64.
- 256,000 atoms, Bussi tau 0.1, FP64: 100 device-driven steps reproduce the oracle's state hash.
- The 512-replica batch (n ~ U{5..20}, rho ~ U[0.70, 0.90], T ~ U[0.7, 1.5]): replicas sampled at the extremes
  of size, density and temperature are bit-equal in FP64 to md_run on the same System (Bussi draws included),
  the smallest one to the oracle's state hash as well, and mixed forces are within 1e-6 of FP64.
- The adjoint at 32,000 atoms, 2000 steps, checkpoints every 50: the loss equals the oracle's forward loss over
  the same fixed-order trajectory to 1e-10; checkpoints every 100 give the same loss and gradients bit for bit;
  d/d(epsilon) and d/d(sigma) match oracle central finite differences to 1e-5 on a 10-step horizon.
"""
import numpy as np
import pytest

import mdgpu
import oracle as O

pytestmark = pytest.mark.gpu

LJ = mdgpu.LJ(1.0, 1.0, 2.5)
DT = 0.005


def say(msg):
    print(msg, flush=True)


def to_gpu(ctx, s):
    st = s.state()
    g = mdgpu.State.from_arrays(ctx, s.box, st["pos"], vel=st["vel"], mass=st["mass"], force=st["force"])
    g.upload(mdgpu.FIELD_ID, st["ids"])
    return g


def rel_force_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_config3_4m_lists_forces_and_mixed_run(ctx):
    say("4M: 100 mixed steps on the GPU")
    g0 = mdgpu.fcc_system(ctx, 100, 0.8442, 0.722, 1)
    l0 = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    l0.build(g0)
    l0.forces(g0, LJ, mdgpu.MD_MIXED)
    l0.run(g0, LJ, DT, 100, precision=mdgpu.MD_MIXED, record=False)
    pos, vel = g0.download(mdgpu.FIELD_POS), g0.download(mdgpu.FIELD_VEL)
    box = g0.box
    del l0, g0
    say("4M: oracle list build and deterministic forces")
    s = O.Sim.from_arrays(pos[:, :3], box, vel=vel.T)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    s.list_build()
    u = s.forces_list()
    so = s.state()
    g = mdgpu.State.from_arrays(ctx, box, pos, vel=vel)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    info = nl.info()
    say(f"4M: padding {info.padding}, longest row {info.max_neighbours}, cells/side {info.cells_per_side}, "
        f"tiles {info.tiles}, unstaged {info.unstaged_tiles}")
    assert info.tiled == 1 and info.cells_per_side == 59
    for k in ("pos", "vel", "ids"):
        assert np.array_equal(so[k], g.download({"pos": mdgpu.FIELD_POS, "vel": mdgpu.FIELD_VEL,
                                                  "ids": mdgpu.FIELD_ID}[k])), k
    offs_o, nb_o = s.half_list()
    offs_g, nb_g = nl.half()
    assert np.array_equal(offs_o, offs_g) and np.array_equal(nb_o, nb_g)
    del offs_o, nb_o, offs_g, nb_g
    r64 = nl.forces(g, LJ, mdgpu.MD_FP64)
    f64 = g.download(mdgpu.FIELD_FORCE)
    assert np.array_equal(f64, so["force"])
    assert r64.potential == pytest.approx(u, rel=1e-12, abs=0)
    assert r64.virial == pytest.approx(s.virial, rel=1e-12, abs=0)
    assert r64.inside_pairs == s.inside_pairs
    rm = nl.forces(g, LJ, mdgpu.MD_MIXED)
    assert rel_force_err(g.download(mdgpu.FIELD_FORCE), f64) <= 1e-6
    assert rm.potential == pytest.approx(r64.potential, rel=1e-6)
    assert rm.virial == pytest.approx(r64.virial, rel=1e-6)
    del s, so
    say("4M: 200 steps mixed and FP64 from the same state")
    runs = []
    for prec in (mdgpu.MD_FP64, mdgpu.MD_MIXED):
        gg = mdgpu.State.from_arrays(ctx, box, pos, vel=vel)
        ll = mdgpu.NeighbourList(ctx, 2.5, 0.3)
        ll.build(gg)
        r0 = ll.forces(gg, LJ, prec)
        e0 = r0.potential + gg.kinetic_energy()
        en = ll.run(gg, LJ, DT, 200, precision=prec)
        runs.append((e0, en[:, 0] + en[:, 1], ll.info().rebuilds))
        del gg, ll
    (e0a, ea, ra), (e0b, eb, rb) = runs
    assert e0b == pytest.approx(e0a, rel=1e-6)
    assert np.abs(eb - ea).max() <= 1e-6 * abs(e0a)
    say(f"4M: FP64 max|dE/E0| {np.abs((ea - e0a) / e0a).max():.3e}, mixed {np.abs((eb - e0b) / e0b).max():.3e}, "
        f"rebuilds {ra} / {rb}")


def test_config2_256k_bussi_fp64_matches_oracle_hash(ctx):
    steps = 100
    s = O.Sim.fcc(40, 0.8442, 1.0, 1)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    s.bussi(1.0, 0.1, 7)
    g = to_gpu(ctx, s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    nl.forces(g, LJ, mdgpu.MD_FP64)
    draws = mdgpu.BussiRng(7).draws(g.n, steps)
    en = nl.run(g, LJ, DT, steps, precision=mdgpu.MD_FP64, draws=draws, temperature=1.0, tau=0.1)
    say("256k: oracle Bussi run")
    s.list_build()
    s.forces_list()
    for _ in range(steps):
        s.step(DT)
        s.bussi_apply(DT)
    assert g.state_hash() == s.state_hash()
    assert nl.info().rebuilds == s.rebuilds
    assert en[-1, 1] == pytest.approx(s.kinetic_energy(), rel=1e-12)


def pick_replicas(specs):
    """Indices of the replicas at the extremes of size, density and temperature."""
    n = np.array([c for c, *_ in specs])
    rho = np.array([r for _, r, _, _ in specs])
    t = np.array([x for _, _, x, _ in specs])
    picks = [int(np.argmin(n)), int(np.argmax(n)), int(np.argmin(rho)), int(np.argmax(rho)), int(np.argmax(t)),
             int(np.argmin(t))]
    return sorted(set(picks))


def test_config4_512_replica_batch(ctx):
    specs = mdgpu.batched_replica_specs(512)
    picks = pick_replicas(specs)
    assert {specs[r][0] for r in picks} >= {5, 20}
    say(f"512 replicas: sampled {[(r, specs[r]) for r in picks]}")
    states = [mdgpu.fcc_system(ctx, c, rho, t, seed) for c, rho, t, seed in specs]
    n = [st.n for st in states]
    temps = np.array([t for _, _, t, _ in specs])
    steps = 40
    draws = np.stack([mdgpu.BussiRng(seed).draws(n[r], steps) for r, (*_, seed) in enumerate(specs)])
    batch = mdgpu.Batch(ctx, states, 2.5, 0.3)
    eb = batch.run(LJ, DT, steps, precision=mdgpu.MD_FP64, temperatures=temps, draws=draws, use_graph=True)
    del states
    for r in picks:
        c, rho, t, seed = specs[r]
        g = mdgpu.fcc_system(ctx, c, rho, t, seed)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
        nl.build(g)
        nl.forces(g, LJ, mdgpu.MD_FP64)
        en = nl.run(g, LJ, DT, steps, precision=mdgpu.MD_FP64, draws=draws[r], temperature=t, tau=0.1)
        got = mdgpu.State(ctx, n[r], batch.boxes[r])
        rebuilds, pairs = batch.replica(r, got)
        for f in (mdgpu.FIELD_POS, mdgpu.FIELD_VEL, mdgpu.FIELD_FORCE, mdgpu.FIELD_ID):
            assert np.array_equal(got.download(f), g.download(f)), (r, f)
        assert rebuilds == nl.info().rebuilds and pairs == nl.info().half_pairs
        assert np.array_equal(eb[:, r, 3], en[:, 3])
        np.testing.assert_allclose(eb[:, r, 0], en[:, 0], rtol=1e-12, atol=0)
    # the smallest replica against the oracle's own run
    r = min(picks, key=lambda k: n[k])
    c, rho, t, seed = specs[r]
    s = O.Sim.fcc(c, rho, t, seed)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    s.bussi(t, 0.1, seed)
    s.list_build()
    s.forces_list()
    for _ in range(steps):
        s.step(DT)
        s.bussi_apply(DT)
    got = mdgpu.State(ctx, n[r], batch.boxes[r])
    batch.replica(r, got)
    assert got.state_hash() == s.state_hash()
    # mixed forces within 1e-6 of FP64 at the same (evolved) positions; the first run of a batch evaluates the
    # forces of its state
    states = []
    for k in range(len(specs)):
        st = mdgpu.State(ctx, n[k], batch.boxes[k])
        batch.replica(k, st)
        states.append(st)
    b64, bmx = mdgpu.Batch(ctx, states, 2.5, 0.3), mdgpu.Batch(ctx, states, 2.5, 0.3)
    b64.run(LJ, DT, 0, precision=mdgpu.MD_FP64)
    bmx.run(LJ, DT, 0, precision=mdgpu.MD_MIXED, lanes=4)
    for r in picks:
        a, b = mdgpu.State(ctx, n[r], b64.boxes[r]), mdgpu.State(ctx, n[r], bmx.boxes[r])
        b64.replica(r, a)
        bmx.replica(r, b)
        assert rel_force_err(b.download(mdgpu.FIELD_FORCE), a.download(mdgpu.FIELD_FORCE)) <= 1e-6


def oracle_fixed_loss(cells, rho, temp, seed, steps, eps=1.0, sigma=1.0, target=1.0):
    """The adjoint's loss over the oracle's fixed-order trajectory: mean U + (mean P - target)^2, k = 1..steps."""
    s = O.Sim.fcc(cells, rho, temp, seed)
    s.forcefield(eps, sigma, 2.5)
    s.neighbour_list(2.5, 0.3, fixed_order=True)
    s.list_build()
    s.forces_list()
    vol = s.box ** 3
    us, ps = [], []
    for _ in range(steps):
        us.append(s.step(DT))
        ps.append((2.0 * s.kinetic_energy() + s.virial) / (3.0 * vol))
    return np.mean(us) + (np.mean(ps) - target) ** 2


def test_config5_adjoint_32k(ctx):
    spec = (20, 0.8442, 0.9, 1)
    out = {}
    for every in (50, 100):
        g = mdgpu.fcc_system(ctx, *spec)
        nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
        out[every] = mdgpu.adjoint_run(g, nl, LJ, DT, 2000, every)
    (r50, p50, v50), (r100, p100, v100) = out[50], out[100]
    say(f"32k adjoint: loss {r50.loss!r}, dL/deps {r50.d_epsilon!r}, dL/dsigma {r50.d_sigma!r}, "
        f"forward {r50.forward_ms:.1f} ms, reverse {r50.reverse_ms:.1f} ms")
    for k in ("loss", "d_epsilon", "d_sigma"):
        assert getattr(r50, k) == getattr(r100, k), k
    assert np.array_equal(p50, p100) and np.array_equal(v50, v100)
    say("32k adjoint: oracle forward over 2000 steps")
    want = oracle_fixed_loss(*spec, 2000)
    assert abs(r50.loss - want) <= 1e-10 * abs(want)
    # epsilon and sigma against central differences of the oracle's loss, 10 steps
    g = mdgpu.fcc_system(ctx, *spec)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    res, _, _ = mdgpu.adjoint_run(g, nl, LJ, DT, 10, 5)
    for name, grad in (("epsilon", res.d_epsilon), ("sigma", res.d_sigma)):
        def f(h):
            kw = {"eps": 1.0 + h} if name == "epsilon" else {"sigma": 1.0 + h}
            return oracle_fixed_loss(*spec, 10, **kw)
        h = 1e-5
        a = (f(h) - f(-h)) / (2 * h)
        b = (f(h / 2) - f(-h / 2)) / h
        assert abs(a - b) <= 1e-6 * abs(a) + 2e-6, f"finite differences disagree ({a} vs {b})"
        assert grad == pytest.approx(b, rel=1e-5), name
