"""Writes tests/golden/reference_vectors.json from oracle/_ref/libmdref.so — the reference engine's own sources
compiled in place (-ffp-contract=off) — so the tests can pin the oracle and the GPU path to numbers the
reference itself produced, also where oracle/_ref is not available.

Scenarios (the reference's own recipes, tests/TestHelpers.hpp:30-58 and tests/test_determinism.cpp:14-32):
  det_hash_6      FCC 6^3*4, seed 42, T 0.728, skin 0.2, deterministic forces, 400 steps with a rescale every
                  10 of the first 200 (runAndHash(threads, true)); final stateHash and rebuild count
  cfg1_det        FCC 8^3*4 (N = 2048), rho 0.8442, T 0.722, seed 1, rc 2.5, skin 0.3, dt 0.005, deterministic
                  FP64 NVE, 1000 steps: U and KE every 50 steps, rebuilds, final hash
  bussi_det       FCC 6^3*4, rho 0.8442, T 1.0, seed 1, Bussi tau 0.1 seed 7, deterministic FP64, 300 steps
  rdf_cfg1        g(r) of the cfg1 start, 250 bins to rMax = 2.8 over the half list of a skin 0.3 build
Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

KIND = "ref"


def det_hash_6():
    s = O.Sim.fcc(6, 0.8442, 0.728, 42, KIND)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.2, reverse=True)
    s.list_build()
    s.forces_list()
    for k in range(1, 401):
        s.step(0.005)
        if k <= 200 and k % 10 == 0:
            s.rescale(0.728)
    return {"hash": f"{s.state_hash():016x}", "rebuilds": s.rebuilds}


def cfg1_det():
    s = O.Sim.fcc(8, 0.8442, 0.722, 1, KIND)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    s.list_build()
    u = s.forces_list()
    trace = [[0, u, s.kinetic_energy(), s.virial, s.inside_pairs]]
    for k in range(1, 1001):
        u = s.step(0.005)
        if k % 50 == 0:
            trace.append([k, u, s.kinetic_energy(), s.virial, s.inside_pairs])
    return {"trace": trace, "rebuilds": s.rebuilds, "hash": f"{s.state_hash():016x}", "pairs": s.pair_count}


def bussi_det():
    s = O.Sim.fcc(6, 0.8442, 1.0, 1, KIND)
    s.forcefield(1.0, 1.0, 2.5, deterministic=True)
    s.neighbour_list(2.5, 0.3, reverse=True)
    s.bussi(1.0, 0.1, 7)
    s.list_build()
    s.forces_list()
    kes = []
    for k in range(1, 301):
        s.step(0.005)
        s.bussi_apply(0.005)
        if k % 50 == 0:
            kes.append([k, s.kinetic_energy()])
    return {"ke": kes, "hash": f"{s.state_hash():016x}", "rebuilds": s.rebuilds}


def rdf_cfg1():
    s = O.Sim.fcc(8, 0.8442, 0.722, 1, KIND)
    s.neighbour_list(2.5, 0.3)
    s.list_build()
    s.rdf(2.8, 250)
    s.rdf_sample()
    r, g = s.rdf_values()
    nz = [[int(b), float(g[b])] for b in range(len(g)) if g[b] != 0.0]
    return {"bins": 250, "rmax": 2.8, "nonzero": nz}


def main():
    if not O.available(KIND):
        raise SystemExit("oracle/_ref/libmdref.so is not built (make -C oracle with the reference checkout)")
    O.set_threads(KIND, 1)
    out = {"source": "oracle/_ref/libmdref.so (reference sources, -O2 -ffp-contract=off)",
           "det_hash_6": det_hash_6(), "cfg1_det": cfg1_det(), "bussi_det": bussi_det(), "rdf_cfg1": rdf_cfg1()}
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: (v if k != "rdf_cfg1" else "...") for k, v in out.items()}, indent=1)[:2000])


if __name__ == "__main__":
    main()
