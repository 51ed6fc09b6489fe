"""Throughput of the MI355X Lennard-Jones step: atom-steps/s at N = 4,000,000 (FCC 100^3 x 4, rho* 0.8442,
T* 0.722, seed 1, rc 2.5, skin 0.3, dt 0.005, mixed precision, NVE, re-sort on every rebuild), one GPU.

A "step" is one velocity-Verlet step of the whole system (kick-drift, displacement test, re-sort + list rebuild
when stale, forces, kick), driven on the device by md_run. Inputs are generated on the host by the library's own
createFCCLattice / initializeVelocities, uploaded, and the first list built before the timed region.

Printed: one JSON line with the throughput, the force kernel's HBM roofline (event-timed inside the timed region;
with the kernel's VALU issue floor beside it, from the instruction count profiles/force_kernel_traffic.json holds)
and the reference's own OpenMP path (oracle/_ref/libmdref_fast.so) timed on this host on a bounded sample, over a
sweep of thread counts. After the timed region, untimed legs add the other figures of the metric: the energy
drift of 2000-step mixed and FP64 runs of the same 4M-atom state, the 512-replica batch against the same replicas
run one by one, the 32k-atom adjoint against its forward run, and the drop-in per-step path (md_step, the calls
VelocityVerletIntegrator::step makes) at 4M atoms.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "molecular-dynamics_amd", "python"))

import mdgpu  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    # name: cells, density, temperature, seed, skin, precision, description
    "lj4m": (100, 0.8442, 0.722, 1, 0.3, mdgpu.MD_MIXED,
             "LJ NVE N=4,000,000 (FCC 100^3x4) rho*=0.8442 T*=0.722 seed 1 rc=2.5 skin=0.3 dt=0.005 mixed"),
    "lj256k": (40, 0.8442, 1.0, 1, 0.3, mdgpu.MD_MIXED,
               "LJ NVE N=256,000 (FCC 40^3x4) rho*=0.8442 T*=1.0 seed 1 rc=2.5 skin=0.3 dt=0.005 mixed"),
    "lj4m_fp64": (100, 0.8442, 0.722, 1, 0.3, mdgpu.MD_FP64,
                  "LJ NVE N=4,000,000 (FCC 100^3x4) rho*=0.8442 T*=0.722 seed 1 rc=2.5 skin=0.3 dt=0.005 fp64"),
    "lj32m": (200, 0.8442, 0.722, 1, 0.3, mdgpu.MD_MIXED,
              "LJ NVE N=32,000,000 (FCC 200^3x4) rho*=0.8442 T*=0.722 seed 1 rc=2.5 skin=0.3 dt=0.005 mixed"),
    "lj108m": (300, 0.8442, 0.722, 1, 0.3, mdgpu.MD_MIXED,
               "LJ NVE N=108,000,000 (FCC 300^3x4) rho*=0.8442 T*=0.722 seed 1 rc=2.5 skin=0.3 dt=0.005 mixed"),
    "lj2048": (8, 0.8442, 0.722, 1, 0.3, mdgpu.MD_FP64,
               "LJ NVE N=2048 (FCC 8^3x4) rho*=0.8442 T*=0.722 seed 1 rc=2.5 skin=0.3 dt=0.005 fp64"),
}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "force_kernel_traffic.json")
# batched workload: the autotuner's candidates (skin sets the rebuild cadence; headroom the list padding as a
# multiple of the longest row the best skin built)
TUNE_SKINS = (0.2, 0.3, 0.4, 0.5, 0.6)
TUNE_HEADROOM = (1.1, 1.3)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def drift_stats(energies, e0, n):
    e = energies[:, 0] + energies[:, 1]
    rel = (e - e0) / abs(e0)
    steps = np.arange(1, len(e) + 1)
    slope = np.polyfit(steps, rel, 1)[0] if len(e) > 1 else 0.0
    return float(slope * 1000.0), float(np.abs(rel).max())


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(cells, density, temperature, seed, skin, mixed, budget_s):
    """The reference's own OpenMP list path, built from its sources (-O3, OpenMP), timed on this host over a
    bounded number of production steps of the same workload, for a sweep of thread counts: all the cores this
    process may run on, and fewer (the reference's per-thread force arrays cost O(threads x N) per step). Each
    count gets budget_s seconds of stepping after its own setup (FCC, first list, list forces)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    if not O.available("ref_fast"):
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    sweep = sorted({c for c in (cores, 64, 32, 16, 8) if c <= cores}, reverse=True)
    results = []
    for threads in sweep:
        O.set_threads("ref_fast", threads)
        s = O.Sim.fcc(cells, density, temperature, seed, "ref_fast")
        s.forcefield(1.0, 1.0, 2.5, mixed=mixed)
        s.neighbour_list(2.5, skin)
        s.list_build()
        s.forces_list()
        n = s.n
        steps, t0 = 0, time.perf_counter()
        while True:
            s.step(0.005)
            steps += 1
            el = time.perf_counter() - t0
            if el >= budget_s or steps >= 400:
                break
        results.append({"threads": threads, "value": n * steps / el, "steps": steps, "seconds": round(el, 2),
                        "builds": s.rebuilds})
        log(f"cpu baseline: {threads} threads, {steps} steps in {el:.1f} s: {n * steps / el:.4g} atom-steps/s")
        del s
    best = max(results, key=lambda r: r["value"])
    allc = next(r for r in results if r["threads"] == cores)
    return {"value": best["value"], "unit": "atom-steps/s", "cores": best["threads"], "kind": "reference",
            "sample": f"{best['steps']} production steps of the same workload (N={n}, "
                      f"{'mixed' if mixed else 'fp64'}, list path, initial forces via the list) in "
                      f"{best['seconds']} s on {best['threads']} host threads, the best of a sweep over "
                      f"{[r['threads'] for r in results]} threads; {best['builds']} list builds incl. the first",
            "all_cores": {"threads": cores, "value": allc["value"]},
            "sweep": results, "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "build": "-O3 -march=x86-64-v3 -fopenmp (oracle/Makefile)"}


def calibrate_copy(nbytes):
    """A plain device copy of known size, for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "launches": 5}))


def batched_replicas(count, base_seed=1):
    """The batched workload (mdgpu.batched_replica_specs)."""
    return mdgpu.batched_replica_specs(count, base_seed)


def measure_batched(ctx, replicas, steps, warm, plain_replicas, plain_steps, with_plain=True, use_graph=True):
    """Batched replicas (Bussi tau 0.1, mixed) in one md_batch, its skin and padding chosen by md_batch_autotune,
    against the same replicas advanced one after another by md_run with that skin (each captured as a hipGraph, the
    fastest plain path)."""
    lj = mdgpu.LJ(1.0, 1.0, 2.5)
    specs = batched_replicas(replicas)
    t_setup = time.perf_counter()
    states = [mdgpu.fcc_system(ctx, c, rho, t, seed) for c, rho, t, seed in specs]
    n_list = [s.n for s in states]
    n_total = sum(n_list)
    temps = np.array([t for _, _, t, _ in specs])
    rngs = [mdgpu.BussiRng(seed) for *_, seed in specs]
    draws_w = np.stack([rngs[r].draws(n_list[r], max(warm, 1)) for r in range(len(specs))])
    draws = np.stack([rngs[r].draws(n_list[r], steps) for r in range(len(specs))])
    batch = mdgpu.Batch(ctx, states, 2.5, 0.3)
    del states
    log(f"batched setup: {len(specs)} replicas, N_total={n_total} in {time.perf_counter() - t_setup:.1f} s")
    if warm:
        batch.run(lj, 0.005, warm, precision=mdgpu.MD_MIXED, temperatures=temps, draws=draws_w,
                  use_graph=use_graph, record=False)
    # the autotuner, from the warmed-up state: rebuild cadence (skin), then list padding for the best skin (trials
    # of 40 steps, each from that same state on a scratch copy); the batch then takes the fastest setting
    tune_steps = 40
    draws_t = np.stack([rngs[r].draws(n_list[r], tune_steps) for r in range(len(specs))])
    tuned = batch.autotune(lj, 0.005, tune_steps, tune_steps, skins=TUNE_SKINS, headroom=TUNE_HEADROOM,
                           tune_lanes=True, precision=mdgpu.MD_MIXED, temperatures=temps, draws=draws_t,
                           use_graph=use_graph)
    log("autotuner: " + ", ".join(f"skin {t['skin']:.2f} pad {t['padding']} lanes {t['lanes']}: "
                                  f"{t['ms_per_step']:.3f} ms" for t in tuned["trials"]) +
        f" -> skin {tuned['skin']:.2f}, padding {tuned['padding']}")
    ctx.synchronize()
    reb0 = batch.info().rebuilds
    t0 = time.perf_counter()
    en = batch.run(lj, 0.005, steps, precision=mdgpu.MD_MIXED, temperatures=temps, draws=draws,
                   use_graph=use_graph)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    info = batch.info()
    value = n_total * steps / elapsed
    log(f"batched: {steps} steps in {elapsed:.3f} s: {value:.4g} atom-steps/s, {info.rebuilds - reb0} rebuilds")
    # force-kernel roofline from an eager stretch of the same run (event-timed)
    k_steps = min(50, steps)
    draws_k = np.stack([rngs[r].draws(n_list[r], k_steps) for r in range(len(specs))])
    timing = []
    batch.run(lj, 0.005, k_steps, precision=mdgpu.MD_MIXED, temperatures=temps, draws=draws_k, timing=timing,
              record=False)
    k_full = 2.0 * info.half_pairs / n_total
    if info.tiled:  # own position, row length, slot row (2 B per entry), force
        bytes_per_atom, kernel = 32.0 + 4.0 + 2.0 * k_full + 24.0, "k_forces_tiled_mixed<false, true>"
    else:  # own position, row and half lengths, int32 row, force
        bytes_per_atom, kernel = 32.0 + 8.0 + 4.0 * k_full + 24.0, f"k_b_forces<true, {info.lanes}>"
    avg_ms = timing[0] / k_steps
    achieved = bytes_per_atom * n_total / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
                "avg_launch_ms": round(avg_ms, 5), "algorithmic_bytes_per_launch": round(bytes_per_atom * n_total)}
    mean_t = float(np.mean(2.0 * en[-1, :, 1] / (3.0 * np.array(n_list) - 3.0) / temps))
    del batch
    # plain: the first `plain_replicas` replicas one after another, each its own md_run
    sample = specs[:plain_replicas] if with_plain else []
    p_steps = min(steps, plain_steps)
    plain = None
    if sample:
        plain_states = [mdgpu.fcc_system(ctx, c, rho, t, seed) for c, rho, t, seed in sample]
        lists = []
        for st in plain_states:
            nl = mdgpu.NeighbourList(ctx, 2.5, tuned["skin"])  # the skin the batch runs with
            nl.build(st)
            nl.forces(st, lj, mdgpu.MD_MIXED)
            lists.append(nl)
        p_draws = [mdgpu.BussiRng(seed).draws(st.n, p_steps) for st, (*_, seed) in zip(plain_states, sample)]
        ctx.synchronize()
        t0 = time.perf_counter()
        for st, nl, d, (_, _, t, _) in zip(plain_states, lists, p_draws, sample):
            nl.run(st, lj, 0.005, p_steps, precision=mdgpu.MD_MIXED, draws=d, temperature=t, tau=0.1,
                   use_graph=True, record=False)
        ctx.synchronize()
        p_el = time.perf_counter() - t0
        p_atoms = sum(st.n for st in plain_states)
        plain = p_atoms * p_steps / p_el
        log(f"plain replica-by-replica: {len(sample)} replicas x {p_steps} steps in {p_el:.3f} s: {plain:.4g}")
    return {
        "value": value, "steps": steps, "warmup": warm, "ms_per_step": 1000.0 * elapsed / steps,
        "replicas": len(specs), "N_total": n_total, "skin": tuned["skin"], "padding": tuned["padding"],
        "particles_per_wavefront": 16 if info.tiled else 64 // info.lanes, "roofline": roofline,
        "plain_replica_by_replica": {"value": plain, "unit": "atom-steps/s", "replicas": len(sample),
                                     "steps": p_steps, "path": "md_run per replica, hipGraph"},
        "speedup_vs_plain": value / plain if plain else None,
        "autotuner": {"trial_steps": tune_steps, "trials": [{k: (round(v, 4) if isinstance(v, float) else v)
                                                            for k, v in t.items()} for t in tuned["trials"]]},
        "rebuilds": int(info.rebuilds - reb0), "k_full_mean": k_full,
        "mean_final_temperature_over_target": mean_t,
    }


def run_batched(args):
    ctx = mdgpu.Context(0)
    m = measure_batched(ctx, args.replicas, args.steps, args.warmup, args.plain_replicas, args.plain_steps,
                        not args.no_plain, not args.eager)
    out = {
        "metric": "batched atom-steps/s (512 Bussi replicas, 1 GPU)" if args.replicas == 512
        else "batched atom-steps/s (1 GPU)",
        "value": m["value"], "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair maths, f64 positions/accumulation",
        "data": "synthetic: per-replica FCC lattice + Maxwell-Boltzmann velocities generated by the library",
        "config": {"workload": f"{m['replicas']} LJ replicas, FCC n~U{{5..20}}, rho~U[0.70,0.90], T~U[0.7,1.5], "
                               f"Bussi tau=0.1, seed 1+r, rc=2.5 skin={m['skin']:g} (autotuned) dt=0.005 mixed",
                   "replicas": m["replicas"], "N_total": m["N_total"], "skin": m["skin"], "graph": not args.eager},
        "roofline": m["roofline"], "cpu_baseline": None,
    }
    out.update({k: m[k] for k in ("plain_replica_by_replica", "speedup_vs_plain", "autotuner", "padding",
                                  "particles_per_wavefront",
                                  "rebuilds", "k_full_mean", "mean_final_temperature_over_target")})
    print(json.dumps(out), flush=True)


def measure_adjoint(ctx, steps, every=50, with_plain=True):
    """Adjoint of the 32k-atom FP64 NVE run (FCC 20^3 x 4, rho 0.8442, T 0.9, seed 1, skin 0.3, dt 0.005),
    checkpoints every `every` steps: reverse-sweep time over forward time, and over a plain md_run of the same
    steps."""
    lj = mdgpu.LJ(1.0, 1.0, 2.5)
    state = mdgpu.fcc_system(ctx, 20, 0.8442, 0.9, 1)
    nl = mdgpu.NeighbourList(ctx, 2.5, 0.3)
    nl.build(state)
    nl.forces(state, lj, mdgpu.MD_FP64)
    # warm-up of both paths on a short horizon
    mdgpu.adjoint_run(state, nl, lj, 0.005, 50, 50)
    t0 = time.perf_counter()
    res, gp, gv = mdgpu.adjoint_run(state, nl, lj, 0.005, steps, every)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    plain_ms = None
    if with_plain:
        twin = mdgpu.fcc_system(ctx, 20, 0.8442, 0.9, 1)
        nl2 = mdgpu.NeighbourList(ctx, 2.5, 0.3)
        nl2.build(twin)
        nl2.forces(twin, lj, mdgpu.MD_FP64)
        ctx.synchronize()
        t0 = time.perf_counter()
        nl2.run(twin, lj, 0.005, steps, precision=mdgpu.MD_FP64, use_graph=True, record=False)
        ctx.synchronize()
        plain_ms = 1000.0 * (time.perf_counter() - t0)
    log(f"adjoint: forward {res.forward_ms:.1f} ms, reverse {res.reverse_ms:.1f} ms, plain md_run {plain_ms} ms")
    return {"adjoint_over_forward": res.reverse_ms / res.forward_ms, "forward_ms": res.forward_ms,
            "reverse_ms": res.reverse_ms, "wall_ms": 1000.0 * wall, "plain_md_run_ms": plain_ms,
            "reverse_over_plain_forward": res.reverse_ms / plain_ms if plain_ms else None,
            "N": state.n, "steps": steps, "checkpoint_every": every,
            "loss": res.loss, "d_epsilon": res.d_epsilon, "d_sigma": res.d_sigma}


def run_adjoint(args):
    ctx = mdgpu.Context(0)
    m = measure_adjoint(ctx, args.steps, 50, not args.no_plain)
    out = {
        "metric": "adjoint time / forward time (N=32k, 2000 steps, checkpoints every 50)",
        "value": m["adjoint_over_forward"], "unit": "ratio", "n_gpus": 1, "steps": args.steps, "warmup": 50,
        "ms_per_step": (m["forward_ms"] + m["reverse_ms"]) / args.steps, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: FCC lattice + Maxwell-Boltzmann velocities",
        "config": {"workload": "LJ NVE N=32,000 (FCC 20^3x4) rho*=0.8442 T*=0.9 seed 1 rc=2.5 skin=0.3 dt=0.005 "
                               "fp64, loss = mean U + (mean P - 1)^2, checkpoints every 50 steps", "N": m["N"]},
        "roofline": None, "cpu_baseline": None,
    }
    out.update({k: m[k] for k in ("forward_ms", "reverse_ms", "wall_ms", "plain_md_run_ms",
                                  "reverse_over_plain_forward", "loss", "d_epsilon", "d_sigma")})
    print(json.dumps(out), flush=True)


def measure_drift(ctx, cells, rho, temp, seed, skin, steps):
    """NVE energy drift of the same initial state run `steps` steps in mixed precision and in FP64 (the FP64 path
    is bit-equal to the reference's deterministic mode, so its drift is the reference's): per 1000 steps, the
    linear-fit slope of (E - E0) / |E0| (main.cpp:284-288 defines the relative deviation), and max |dE / E0|."""
    lj = mdgpu.LJ(1.0, 1.0, 2.5)
    out = {"steps": steps}
    for name, prec in (("fp64", mdgpu.MD_FP64), ("mixed", mdgpu.MD_MIXED)):
        st = mdgpu.fcc_system(ctx, cells, rho, temp, seed)
        nl = mdgpu.NeighbourList(ctx, 2.5, skin)
        nl.build(st)
        r0 = nl.forces(st, lj, prec)
        e0 = r0.potential + st.kinetic_energy()
        en = nl.run(st, lj, 0.005, steps, precision=prec, use_graph=True)
        slope, mx = drift_stats(en, e0, st.n)
        out[name] = {"drift_per_1000_steps": slope, "max_abs_dE_over_E0": mx, "E0": e0,
                     "rebuilds": nl.info().rebuilds}
        log(f"drift {name}: {slope:.3e} per 1000 steps, max |dE/E0| {mx:.3e} over {steps} steps")
        del st, nl
    f, m = out["fp64"], out["mixed"]
    out["mixed_over_fp64_drift"] = abs(m["drift_per_1000_steps"]) / abs(f["drift_per_1000_steps"]) \
        if f["drift_per_1000_steps"] else None
    out["mixed_over_fp64_max_dE"] = m["max_abs_dE_over_E0"] / f["max_abs_dE_over_E0"] \
        if f["max_abs_dE_over_E0"] else None
    return out


def measure_step_path(ctx, cells, rho, temp, seed, skin, prec, steps):
    """The drop-in per-step path: md_step (kick-drift, the list's half-skin check with its host read-back,
    rebuild when stale, forces with their host read-back, kick), the calls VelocityVerletIntegrator::step
    (Integrator.cpp:18-30) makes through the facade, one step per call."""
    lj = mdgpu.LJ(1.0, 1.0, 2.5)
    st = mdgpu.fcc_system(ctx, cells, rho, temp, seed)
    nl = mdgpu.NeighbourList(ctx, 2.5, skin)
    nl.build(st)
    nl.forces(st, lj, prec)
    for _ in range(3):
        nl.step(st, lj, 0.005, prec)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        nl.step(st, lj, 0.005, prec)
    ctx.synchronize()
    el = time.perf_counter() - t0
    value = st.n * steps / el
    log(f"md_step path: {steps} steps in {el:.3f} s: {value:.4g} atom-steps/s")
    return {"value": value, "unit": "atom-steps/s", "steps": steps, "ms_per_step": 1000.0 * el / steps,
            "path": "md_step per call (VelocityVerletIntegrator::step through the facade)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="lj4m", choices=sorted(WORKLOADS) + ["batched", "adjoint"])
    ap.add_argument("--replicas", type=int, default=512, help="batched workload: replica count")
    ap.add_argument("--plain-replicas", type=int, default=32, help="batched workload: replicas in the plain sample")
    ap.add_argument("--plain-steps", type=int, default=200, help="batched workload: steps of the plain sample")
    ap.add_argument("--eager", action="store_true",
                    help="batched workload: launch each step's kernels eagerly instead of replaying a hipGraph")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched / adjoint workloads: skip the plain md_run comparison")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="CPU baseline: seconds per thread count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="lj4m: skip the untimed drift, batched, adjoint and md_step legs")
    ap.add_argument("--drift-steps", type=int, default=2000)
    ap.add_argument("--graph", action="store_true", help="replay a captured step instead of eager launches")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (each box has one GPU).",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.workload == "batched":
        run_batched(args)
        return
    if args.workload == "adjoint":
        run_adjoint(args)
        return

    cells, rho, temp, seed, skin, prec, desc = WORKLOADS[args.workload]
    ctx = mdgpu.Context(0)
    t_setup = time.perf_counter()
    state = mdgpu.fcc_system(ctx, cells, rho, temp, seed)
    n = state.n
    nl = mdgpu.NeighbourList(ctx, 2.5, skin)
    lj = mdgpu.LJ(1.0, 1.0, 2.5)
    nl.build(state)
    nl.forces(state, lj, prec)
    log(f"setup N={n} in {time.perf_counter() - t_setup:.1f} s; padding {nl.info().padding}")

    if args.warmup:
        nl.run(state, lj, 0.005, args.warmup, precision=prec, use_graph=args.graph, record=False)
    ctx.synchronize()
    rebuilds0 = nl.info().rebuilds
    timing = None if args.graph else []
    t0 = time.perf_counter()
    en = nl.run(state, lj, 0.005, args.steps, precision=prec, use_graph=args.graph, timing=timing)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    info = nl.info()
    rebuilds = info.rebuilds - rebuilds0
    value = n * args.steps / elapsed
    log(f"timed {args.steps} steps in {elapsed:.3f} s: {value:.4g} atom-steps/s, {rebuilds} rebuilds")

    k_full = 2.0 * info.half_pairs / n
    # force kernel: per atom it reads its position (32 B), its row length(s) (4 B, 8 B with the half-row length
    # of the FP64 and untiled kernels) and its row (4 B per full-list neighbour as int32, 2 B as the tiled kernel's 16-bit
    # LDS slots), and writes its force (24 B); neighbour positions are re-reads of the same array. In an NVE run
    # the tiled kernels also carry the step for all but the last one: instead of writing the force they read the
    # velocity (24 B), mass (8 B) and list-build position (32 B) and write the velocity (24 B) and the next
    # position (32 B)
    tiled = bool(info.tiled) and (prec == mdgpu.MD_MIXED or os.environ.get("MDGPU_FP64_GATHER") != "1")
    lengths = 4.0 if tiled and prec == mdgpu.MD_MIXED else 8.0
    plain = 32.0 + lengths + (2.0 if tiled else 4.0) * k_full + 24.0
    folded = 32.0 + lengths + 2.0 * k_full + 24.0 + 8.0 + 32.0 + 24.0 + 32.0
    fold = tiled and os.environ.get("MDGPU_UNFOLDED") != "1" and args.steps >= 2
    bytes_per_atom = ((args.steps - 1) * folded + plain) / args.steps if fold else plain
    roofline = None
    if timing:
        avg_ms = timing[0] / args.steps
        achieved = bytes_per_atom * n / (avg_ms * 1e-3) / 1e9
        traffic = None
        valu = None
        if os.path.exists(TRAFFIC_FILE):
            with open(TRAFFIC_FILE) as f:
                tr = json.load(f).get("workloads", {}).get(args.workload, {})
            traffic = tr.get("hbm_bytes_per_launch")
            valu = tr.get("valu_instructions_per_launch")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": ("k_forces_tiled_mixed<true>" if fold else "k_forces_tiled_mixed" if tiled
                               else "k_forces_mixed") if prec == mdgpu.MD_MIXED
                    else ("k_forces_tiled_fp64_pair<STEP>" if fold else "k_forces_tiled_fp64_pair") if tiled
                    else "k_forces_fp64",
                    "avg_launch_ms": round(avg_ms, 5), "algorithmic_bytes_per_launch": round(bytes_per_atom * n)}
        if valu:
            # the kernel's issue floor: its wavefront VALU instructions (rocprofv3 SQ_INSTS_VALU of the same
            # workload) at 4 cycles each on 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock
            floor_ms = valu * 4.0 / (1024 * 2.4e9) * 1e3
            roofline["valu_issue"] = {"instructions": valu, "floor_ms": round(floor_ms, 4),
                                      "frac": round(floor_ms / avg_ms, 4)}
            # `frac` is measured against HBM bandwidth, but the kernel is held back by instruction issue
            # first: name the nearer ceiling
            if floor_ms / avg_ms > roofline["frac"]:
                roofline["limiter"] = "valu_issue"
    del state, nl
    extras = {}
    if not args.no_extras and args.workload == "lj4m":
        # untimed legs: the other figures of the metric
        extras["drift"] = measure_drift(ctx, cells, rho, temp, seed, skin, args.drift_steps)
        extras["batched"] = measure_batched(ctx, 512, 100, 20, 32, 100)
        extras["adjoint"] = measure_adjoint(ctx, 2000, 50)
        extras["md_step_path"] = measure_step_path(ctx, cells, rho, temp, seed, skin, prec, 20)
    cpu = None
    if not args.no_cpu:
        log("timing the reference's CPU path ...")
        cpu = cpu_baseline(cells, rho, temp, seed, skin, prec == mdgpu.MD_MIXED, args.cpu_seconds)
    out = {
        "metric": "atom-steps/s at N=4M LJ liquid (1 GPU)" if args.workload == "lj4m" else "atom-steps/s",
        "value": value, "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 pair maths, f64 positions/accumulation" if prec == mdgpu.MD_MIXED
        else "f64", "data": "synthetic: FCC lattice + Maxwell-Boltzmann velocities generated by the library",
        "config": {"workload": desc, "N": n, "steps_in_config": 2000, "graph": bool(args.graph)},
        "roofline": roofline, "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
        "interactions_per_s": value * k_full, "k_full_mean": k_full, "rebuilds": rebuilds,
        "steps_per_rebuild": (args.steps / rebuilds) if rebuilds else None,
    }
    if extras:
        d, b, a = extras["drift"], extras["batched"], extras["adjoint"]
        out["drift"] = d
        out["batched_vs_plain"] = b["speedup_vs_plain"]
        out["batched"] = {k: b[k] for k in ("value", "steps", "ms_per_step", "replicas", "N_total", "skin",
                                            "padding", "particles_per_wavefront", "plain_replica_by_replica",
                                            "autotuner", "roofline")}
        out["adjoint_over_forward"] = a["adjoint_over_forward"]
        out["adjoint"] = {k: a[k] for k in ("N", "steps", "checkpoint_every", "forward_ms", "reverse_ms",
                                            "plain_md_run_ms", "reverse_over_plain_forward")}
        out["md_step_path"] = extras["md_step_path"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
