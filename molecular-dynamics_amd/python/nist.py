"""NIST Lennard-Jones validation over batched replicas: the reference's validation driver (bench/validate.py) with
its 16 state points x 8 replicas advanced together as one md_batch instead of 128 md_sim processes.

What stays the reference's:
  - the state points, the NIST tables and the replica count (validate.py:33-61, 216);
  - the system of every replica: FCC with 5 cells per side (N = 500), cutoff 3, dt 0.005, skin 0.2, FP64, seed
    1000 * state_index + r + 1 with state points numbered from 1 in sorted (T, rho) order (validate.py:96-104,
    217-221), velocities drawn at the target temperature from that seed (main.cpp:168-170);
  - production: 20000 steps of NVT with the Bussi thermostat at tau 0.5 fed by that seed's own draw stream
    (main.cpp:242, 272-277);
  - the conversion to NIST's conventions (validate.py:114-120): the shift removed through the count of pairs
    inside the cutoff, the standard tail corrections added to U and P, P the virial pressure (main.cpp:209-211);
  - the statistics (validate.py:123-134, 137-152): replica means, the standard error of their spread, and
    z = (ours - NIST) / hypot(our error, NIST's error).

What differs, and why it cannot move the averages: the reference equilibrates with velocity rescaling every 10
steps (main.cpp:219-227), to T = 2 for the first third of its 10000 steps (--melt) and to the target for the rest.
The batch has no rescale step, so equilibration here runs the same two phases with the Bussi thermostat at a
strong coupling (tau = 0.05, ten steps, the rescaling cadence) drawing from a second stream per replica. Both
protocols only bring each replica to the target state before production starts; the production ensemble, and
with it every average compared to NIST, is the reference's.
"""
from __future__ import annotations

import argparse
import math
import statistics
import time

import numpy as np

import mdgpu

CUTOFF = 3.0
DT = 0.005
SKIN = 0.2
CELLS = 5                 # FCC cells per side: N = 500, as NIST
EQUIL_STEPS = 10000       # 50 time units, as NIST
PROD_STEPS = 20000        # 100 time units, as NIST
REPLICAS = 8
PROD_TAU = 0.5            # main.cpp:242
EQUIL_TAU = 0.05          # ten steps, the reference's rescaling cadence (main.cpp:148)
MELT_TEMPERATURE = 2.0    # main.cpp:225-226
EQUIL_STREAM = 1 << 40    # offset of the equilibration draw stream's seed from the replica's seed

# NIST NVE MD (T is their measured mean): (T, rho, U per particle, sigma U, P, sigma P) (validate.py:39-47)
NIST_MD = [
    (0.851, 0.776, -5.517, 0.001, 0.030, 0.006),
    (0.853, 0.780, -5.533, 0.001, 0.072, 0.005),
    (0.852, 0.820, -5.803, 0.001, 0.573, 0.005),
    (0.851, 0.840, -5.909, 0.001, 0.910, 0.006),
    (0.849, 0.860, -6.027, 0.001, 1.282, 0.005),
    (0.851, 0.900, -6.234, 0.001, 2.544, 0.006),
]

# NIST NVT Monte Carlo, liquid densities: (T, rho, U, sigma U, P, sigma P) (validate.py:49-61). The primary
# reference: it is at exactly the thermostat's temperatures.
NIST_MC = [
    (0.85, 0.776, -5.5121, 4.55e-4, 6.7714e-3, 1.77e-3),
    (0.85, 0.780, -5.5386, 7.26e-4, 4.7924e-2, 3.18e-3),
    (0.85, 0.820, -5.7947, 6.03e-4, 5.5355e-1, 4.13e-3),
    (0.85, 0.860, -6.0305, 2.38e-3, 1.2660, 1.36e-2),
    (0.85, 0.900, -6.2391, 5.27e-3, 2.2314, 2.72e-2),
    (0.90, 0.776, -5.4689, 4.20e-4, 2.4056e-1, 2.74e-3),
    (0.90, 0.780, -5.4956, 7.86e-4, 2.7851e-1, 2.97e-3),
    (0.90, 0.820, -5.7456, 7.51e-4, 8.2386e-1, 2.85e-3),
    (0.90, 0.860, -5.9753, 5.53e-4, 1.5781, 3.29e-3),
    (0.90, 0.900, -6.1773, 1.57e-3, 2.5848, 9.54e-3),
]


def shift_energy(rc: float) -> float:
    """4 (rc^-12 - rc^-6): the value the shifted potential subtracts from every pair inside the cutoff."""
    return 4.0 * (rc ** -12 - rc ** -6)


def tail_energy(rho: float, rc: float) -> float:
    """Long-range correction to U per particle with g(r) = 1 beyond the cutoff: 2 pi rho int_rc^inf r^2 u(r) dr."""
    return 8.0 / 3.0 * math.pi * rho * (rc ** -9 / 3.0 - rc ** -3)


def tail_pressure(rho: float, rc: float) -> float:
    """Long-range correction to P: -(2/3) pi rho^2 int_rc^inf r^3 u'(r) dr."""
    return 16.0 / 3.0 * math.pi * rho ** 2 * (2.0 / 3.0 * rc ** -9 - rc ** -3)


def state_points():
    """Every (T, rho) of both tables, sorted; numbered from 1 for the seeds (validate.py:216-221)."""
    return sorted({(t, rho) for t, rho, *_ in NIST_MD + NIST_MC})


def replica_seed(state_index: int, replica: int) -> int:
    """validate.py:103: one seed per replica of every state point."""
    return 1000 * state_index + replica + 1


def nist_series(records: np.ndarray, n: int, box: float, rho: float, rc: float = CUTOFF):
    """U per particle and P per step in NIST's conventions from a replica's per-step records (potential,
    kinetic, virial, inside pairs) (validate.py:114-120, main.cpp:209-216): the truncated, unshifted energy
    plus its tail correction, and the virial pressure (2 KE + W) / 3V plus its tail correction."""
    pe, ke, w, inside = records[:, 0], records[:, 1], records[:, 2], records[:, 3]
    u = pe / n + inside * shift_energy(rc) / n + tail_energy(rho, rc)
    p = (2.0 * ke + w) / (3.0 * box ** 3) + tail_pressure(rho, rc)
    return u, p


def summarise(u_means, p_means) -> dict:
    """Mean over replicas with the standard error of the replica means (validate.py:123-134)."""
    r = len(u_means)
    return {"u": statistics.mean(u_means), "u_err": statistics.stdev(u_means) / math.sqrt(r),
            "p": statistics.mean(p_means), "p_err": statistics.stdev(p_means) / math.sqrt(r)}


def z_scores(result: dict, u_ref: float, u_ref_err: float, p_ref: float, p_ref_err: float):
    """validate.py:143-144."""
    zu = (result["u"] - u_ref) / math.hypot(result["u_err"], u_ref_err)
    zp = (result["p"] - p_ref) / math.hypot(result["p_err"], p_ref_err)
    return zu, zp


def run_replicas(ctx, states=None, replicas: int = REPLICAS, equil: int = EQUIL_STEPS, prod: int = PROD_STEPS,
                 precision=mdgpu.MD_FP64, use_graph: bool = True, cells: int = CELLS):
    """Runs `replicas` replicas of every state point as one batch. Returns ({(T, rho): summary}, info) where
    info holds the batch's timing and counters."""
    states = state_points() if states is None else states
    index = {s: i for i, s in enumerate(state_points(), 1)}
    specs = [(t, rho, replica_seed(index[(t, rho)], r)) for t, rho in states for r in range(replicas)]
    systems = [mdgpu.fcc_system(ctx, cells, rho, t, seed) for t, rho, seed in specs]
    n = systems[0].n
    boxes = [s.box for s in systems]
    lj = mdgpu.LJ(1.0, 1.0, CUTOFF)
    batch = mdgpu.Batch(ctx, systems, CUTOFF, SKIN)
    del systems

    melt = equil // 3
    equil_rng = [mdgpu.BussiRng(seed + EQUIL_STREAM) for *_, seed in specs]
    phases = [(melt, np.full(len(specs), MELT_TEMPERATURE), EQUIL_TAU, equil_rng, False),
              (equil - melt, np.array([t for t, _, _ in specs]), EQUIL_TAU, equil_rng, False)]
    prod_rng = [mdgpu.BussiRng(seed) for *_, seed in specs]
    phases.append((prod, np.array([t for t, _, _ in specs]), PROD_TAU, prod_rng, True))

    phases = [(steps, temps, tau, np.stack([g.draws(n, steps) for g in rngs]), keep)
              for steps, temps, tau, rngs, keep in phases if steps > 0]

    ctx.synchronize()
    t0 = time.perf_counter()
    records = None
    for steps, temps, tau, draws, keep in phases:
        en = batch.run(lj, DT, steps, precision=precision, temperatures=temps, draws=draws, tau=tau,
                       use_graph=use_graph, record=keep)
        if keep:
            records = en
    ctx.synchronize()
    seconds = time.perf_counter() - t0

    out = {}
    for k, (t, rho) in enumerate(states):
        u_means, p_means, t_means = [], [], []
        for r in range(replicas):
            j = k * replicas + r
            u, p = nist_series(records[:, j, :], n, boxes[j], rho)
            u_means.append(float(np.mean(u)))
            p_means.append(float(np.mean(p)))
            t_means.append(float(np.mean(2.0 * records[:, j, 1] / (3 * n - 3))))  # System::temperature
        out[(t, rho)] = summarise(u_means, p_means) | {"u_means": u_means, "p_means": p_means, "t_means": t_means}
    bi = batch.info()
    info = {"seconds": seconds, "replicas": len(specs), "particles": bi.particles, "steps": equil + prod,
            "atom_steps_per_s": bi.particles * (equil + prod) / seconds, "rebuilds": bi.rebuilds}
    batch.close()
    return out, info


def compare(results: dict, reference, label: str, echo: bool = True):
    """Rows of validate.py's comparison table (validate.py:137-152) for the state points in `results`."""
    rows = []
    if echo:
        print(f"\n{label}")
        print(f"{'T':>6} {'rho':>6} | {'U (ours)':>17} {'U (NIST)':>17} {'z':>5} | "
              f"{'P (ours)':>15} {'P (NIST)':>15} {'z':>5}")
    for t, rho, u_ref, u_ref_err, p_ref, p_ref_err in reference:
        if (t, rho) not in results:
            continue
        a = results[(t, rho)]
        zu, zp = z_scores(a, u_ref, u_ref_err, p_ref, p_ref_err)
        if echo:
            print(f"{t:6.3f} {rho:6.3f} | {a['u']:9.4f} ±{a['u_err']:.4f} {u_ref:9.4f} ±{u_ref_err:.4f} {zu:5.1f} | "
                  f"{a['p']:7.4f} ±{a['p_err']:.4f} {p_ref:7.4f} ±{p_ref_err:.4f} {zp:5.1f}")
        rows.append({"reference": label, "T": t, "rho": rho, "U": a["u"], "U_err": a["u_err"], "U_nist": u_ref,
                     "U_nist_err": u_ref_err, "U_z": zu, "P": a["p"], "P_err": a["p_err"], "P_nist": p_ref,
                     "P_nist_err": p_ref_err, "P_z": zp})
    return rows


def write_rows(path: str, rows):
    keys = list(rows[0])
    with open(path, "w") as f:
        f.write(",".join(keys) + "\n")
        for r in rows:
            f.write(",".join(str(r[k]) for k in keys) + "\n")


def main(argv=None):
    parser = argparse.ArgumentParser(description="NIST LJ validation over batched replicas")
    parser.add_argument("--csv", help="write the comparison rows here")
    parser.add_argument("--mixed", action="store_true", help="mixed-precision forces instead of FP64")
    parser.add_argument("--eager", action="store_true", help="eager launches instead of hipGraph replay")
    parser.add_argument("--replicas", type=int, default=REPLICAS, help="replicas per state point")
    args = parser.parse_args(argv)
    ctx = mdgpu.Context(0)
    precision = mdgpu.MD_MIXED if args.mixed else mdgpu.MD_FP64
    results, info = run_replicas(ctx, replicas=args.replicas, precision=precision, use_graph=not args.eager)
    print(f"{info['replicas']} replicas x {info['steps']} steps in {info['seconds']:.2f} s "
          f"({info['atom_steps_per_s']:.3e} atom-steps/s, {info['rebuilds']} list rebuilds)")
    rows = compare(results, NIST_MD, "NIST MD") + compare(results, NIST_MC, "NIST MC")
    mc = [r for r in rows if r["reference"] == "NIST MC"]
    within = sum(abs(r[k]) < 2.0 for r in mc for k in ("U_z", "P_z"))
    chi2 = sum(r[k] ** 2 for r in mc for k in ("U_z", "P_z"))
    print(f"\nNIST MC: {within}/{2 * len(mc)} |z| < 2, chi^2 = {chi2:.1f} over {2 * len(mc)} values")
    if args.csv:
        write_rows(args.csv, rows)


if __name__ == "__main__":
    main()
