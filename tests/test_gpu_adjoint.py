"""The GPU adjoint (md_adjoint_run) against FP64 central finite differences of the same loss on the CPU oracle,
through a torch.autograd.Function wrapper.

Loss: L = mean_k U_k + (mean_k P_k - 1)^2 over k = 1..K velocity-Verlet steps, P = (2 KE + W) / (3V).
The shifted LJ force (and so the virial) jumps at rc, which makes the discrete trajectory discontinuous where a
pair distance crosses rc between two perturbed runs; the checks use short horizons and small steps, and every
finite difference is confirmed by a second one at half the step, so a crossing shows up as a disagreement
instead of a false match.
"""
import numpy as np
import pytest

import mdgpu
import oracle as O

pytestmark = pytest.mark.gpu

DT, STEPS, EVERY = 0.005, 12, 5


def oracle_loss(base, eps=1.0, sigma=1.0, dpos=None, target=1.0):
    s = base.clone()
    if dpos is not None:
        p = s.state()["pos"]
        p[:, :3] += dpos
        s.set_positions(p)
    s.forcefield(eps, sigma, 2.5, deterministic=False)
    s.neighbour_list(2.5, 0.3)
    s.list_build()
    s.forces_list()
    vol = s.box ** 3
    us, ps = [], []
    for _ in range(STEPS):
        u = s.step(DT)
        us.append(u)
        ps.append((2.0 * s.kinetic_energy() + s.virial) / (3.0 * vol))
    return np.mean(us) + (np.mean(ps) - target) ** 2


def central(f, h):
    return (f(h) - f(-h)) / (2 * h)


def checked_fd(f, h):
    # the loss is O(1e3), so central differences carry ~1e-7 of rounding; a crossing of rc shows up far above it
    a, b = central(f, h), central(f, h / 2)
    assert abs(a - b) <= 1e-6 * abs(a) + 2e-6, f"finite differences disagree ({a} vs {b}): rc crossing"
    return b


@pytest.fixture(scope="module")
def system():
    s = O.Sim.fcc(5, 0.8442, 0.9, 3)  # N = 500, box 8.4: a 3x3x3 cell grid
    return s


def gpu_state(ctx, s):
    st = s.state()
    return mdgpu.State.from_arrays(ctx, s.box, st["pos"], vel=st["vel"], mass=st["mass"])


def test_adjoint_matches_finite_differences(ctx, system):
    g = gpu_state(ctx, system)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    lj = mdgpu.LJ(1.0, 1.0, 2.5)
    res, gpos, gvel = mdgpu.adjoint_run(g, nl, lj, DT, STEPS, EVERY, 1.0)
    assert res.loss == pytest.approx(oracle_loss(system), rel=1e-10)
    d_eps = checked_fd(lambda h: oracle_loss(system, eps=1.0 + h), 1e-6)
    d_sig = checked_fd(lambda h: oracle_loss(system, sigma=1.0 + h), 1e-6)
    assert res.d_epsilon == pytest.approx(d_eps, rel=1e-5)
    assert res.d_sigma == pytest.approx(d_sig, rel=1e-5)
    rng = np.random.default_rng(0)
    scale = np.abs(gpos).max()
    for p in rng.choice(system.n, 4, replace=False):
        for c in range(3):
            def f(h):
                d = np.zeros((system.n, 3))
                d[p, c] = h
                return oracle_loss(system, dpos=d)
            fd = checked_fd(f, 1e-6)
            assert abs(gpos[c, p] - fd) <= 1e-5 * max(abs(fd), 1e-2 * scale), (p, c, gpos[c, p], fd)


def test_adjoint_segments_do_not_change_the_gradient(ctx, system):
    g = gpu_state(ctx, system)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    lj = mdgpu.LJ(1.0, 1.0, 2.5)
    a = mdgpu.adjoint_run(g, nl, lj, DT, STEPS, 5)
    b = mdgpu.adjoint_run(g, nl, lj, DT, STEPS, STEPS)
    c = mdgpu.adjoint_run(g, nl, lj, DT, STEPS, 1)
    for r in (b, c):
        assert r[0].loss == a[0].loss
        assert r[0].d_epsilon == pytest.approx(a[0].d_epsilon, rel=1e-12)
        assert r[0].d_sigma == pytest.approx(a[0].d_sigma, rel=1e-12)
        assert np.allclose(r[1], a[1], rtol=1e-11, atol=1e-12 * np.abs(a[1]).max())


def test_torch_autograd_function(ctx, system):
    torch = pytest.importorskip("torch")

    class MDLoss(torch.autograd.Function):
        """L(eps, sigma, r0) through md_adjoint_run; backward returns the adjoint gradients."""

        @staticmethod
        def forward(fctx, eps, sigma, r0):
            g = mdgpu.State.from_arrays(ctx, system.box, r0.detach().double().numpy(), vel=vel0, mass=mass0)
            nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
            res, gpos, _ = mdgpu.adjoint_run(g, nl, mdgpu.LJ(float(eps), float(sigma), 2.5), DT, STEPS, EVERY)
            fctx.save_for_backward(torch.tensor(res.d_epsilon, dtype=torch.float64),
                                   torch.tensor(res.d_sigma, dtype=torch.float64), torch.from_numpy(gpos.T.copy()))
            return torch.tensor(res.loss, dtype=torch.float64)

        @staticmethod
        def backward(fctx, grad):
            de, ds, gp = fctx.saved_tensors
            return grad * de, grad * ds, grad * gp

    st = system.state()
    vel0, mass0 = st["vel"], st["mass"]
    eps = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    r0 = torch.tensor(st["pos"][:, :3], dtype=torch.float64, requires_grad=True)
    loss = MDLoss.apply(eps, sig, r0)
    loss.backward()
    d_eps = checked_fd(lambda h: oracle_loss(system, eps=1.0 + h), 1e-6)
    assert eps.grad.item() == pytest.approx(d_eps, rel=1e-5)
    assert r0.grad.shape == (system.n, 3)
    def shift_first_x(h):
        d = np.zeros((system.n, 3))
        d[0, 0] = h
        return oracle_loss(system, dpos=d)

    fd = checked_fd(shift_first_x, 1e-6)
    assert r0.grad[0, 0].item() == pytest.approx(fd, rel=1e-5, abs=1e-6)
