"""Child process of test_gpu_untiled.py: the per-particle (untiled) list and force path, which MDGPU_UNTILED=1
selects when the library first reads it (once per process). Checks the list, FP64 and mixed forces, a golden FP64
trajectory (eager steps with host-driven re-sorts, where k_cell_sort gathers the state) and a device-driven mixed
md_run (re-sorts inside the captured step) against the same run in FP64. Prints OK on success."""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "molecular-dynamics_amd", "python"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import mdgpu  # noqa: E402
import oracle as O  # noqa: E402

assert os.environ.get("MDGPU_UNTILED") == "1"
LJ = mdgpu.LJ(1.0, 1.0, 2.5)
ctx = mdgpu.Context(0)
golden = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_vectors.json")))


def to_gpu(s):
    st = s.state()
    g = mdgpu.State.from_arrays(ctx, s.box, st["pos"], vel=st["vel"], mass=st["mass"], force=st["force"])
    g.upload(mdgpu.FIELD_ID, st["ids"])
    return g


# list and FP64 forces, bit for bit (NeighbourList.cpp:38-134, Forcefield.cpp deterministic mode)
s = O.make_liquid(8, 60, skin=0.2)
s.forcefield(1.0, 1.0, 2.5, deterministic=True)
s.neighbour_list(2.5, 0.2, reverse=True)
g = to_gpu(s)
nl = mdgpu.NeighbourList(ctx, 2.5, 0.2)
s.list_build()
nl.build(g)
assert not nl.info().tiled, "the untiled path was not selected"
for k in ("pos", "vel", "force", "mass", "ids"):
    assert np.array_equal(s.state()[k], g.state()[k]), k
offs_o, nb_o = s.half_list()
offs_g, nb_g = nl.half()
assert np.array_equal(offs_o, offs_g) and np.array_equal(nb_o, nb_g)
u = s.forces_list()
r = nl.forces(g, LJ, mdgpu.MD_FP64)
assert np.array_equal(s.state()["force"], g.download(mdgpu.FIELD_FORCE))
assert abs(r.potential - u) <= 1e-12 * abs(u)
f64 = g.download(mdgpu.FIELD_FORCE)
nl.forces(g, LJ, mdgpu.MD_MIXED)
fm = g.download(mdgpu.FIELD_FORCE)
assert np.abs(fm - f64).max() <= 1e-6 * np.abs(f64).max()

# golden deterministic trajectory (test_determinism.cpp:14-32): 400 eager steps, rescaling, host-driven re-sorts
s = O.Sim.fcc(6, 0.8442, 0.728, 42)
g = to_gpu(s)
nl = mdgpu.NeighbourList(ctx, 2.5, 0.2)
nl.build(g)
nl.forces(g, LJ, mdgpu.MD_FP64)
for k in range(1, 401):
    nl.step(g, LJ, 0.005, mdgpu.MD_FP64)
    if k <= 200 and k % 10 == 0:
        g.rescale(0.728)
assert f"{g.state_hash():016x}" == golden["det_hash_6"]["hash"]
assert nl.info().rebuilds == golden["det_hash_6"]["rebuilds"]

# device-driven md_run with re-sorts inside the captured step: mixed within 1e-6 of FP64 in total energy
s = O.Sim.fcc(8, 0.8442, 0.722, 1)
runs = []
for prec in (mdgpu.MD_FP64, mdgpu.MD_MIXED):
    g = to_gpu(s)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(g)
    nl.forces(g, LJ, prec)
    e = nl.run(g, LJ, 0.005, 200, precision=prec)
    assert nl.info().rebuilds > 1
    runs.append(e[:, 0] + e[:, 1])
assert np.abs(runs[1] - runs[0]).max() <= 1e-6 * np.abs(runs[0]).max()
print("OK")
