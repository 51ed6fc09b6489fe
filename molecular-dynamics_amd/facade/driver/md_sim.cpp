// md_sim on the GPU: the reference's command-line driver (src/main.cpp) over the facade classes, with the
// list path's steps advanced on the device by md_run. Same options, console output, CSV rows, RDF file, XYZ
// frames and RESULT line (main.cpp:25-350), so the reference's benchmark scripts (bench/benchmark.py:42-60) read
// it unchanged.
//
// The arithmetic is the facade's: FP64 list forces bit-equal to the reference's deterministic mode, the initial
// forces over all pairs (main.cpp:190), kinetic energies in System::kineticEnergy's order, the Bussi draws of the
// same std::mt19937_64 stream. A --deterministic run therefore reproduces the reference's state hash, rebuild
// count and printed values; without --deterministic the reference sums forces in a thread-count dependent order
// while this driver keeps the deterministic one.
//
// md_run advances stretches of steps: equilibration 10 at a time (the velocity rescale between them, as the
// reference rescales every 10th step), production in one stretch, or up to the next XYZ frame. Each stretch
// returns every step's potential, kinetic energy (System::kineticEnergy of the step's final velocities), virial
// and inside-pair count, from which the rows and statistics follow; the production g(r) is sampled inside the
// run every 10 steps over every pair within its radius (the cut-off, so the list pairs the reference counts).
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string_view>
#include <vector>

#include "Forcefield.hpp"
#include "Initialisation.hpp"
#include "Integrator.hpp"
#include "MdgpuRuntime.hpp"
#include "NeighbourList.hpp"
#include "RadialDistribution.hpp"
#include "System.hpp"
#include "Thermostat.hpp"
#include "Trajectory.hpp"
#include "mdgpu.h"

namespace {

// Whole-string numbers with range checks (the reference's CommandLine.hpp semantics).
bool readDouble(const char* text, double& out) {
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(text, &end);
    if (end == text || *end != '\0' || errno != 0) return false;
    out = v;
    return true;
}

bool readLong(const char* text, long& out) {
    char* end = nullptr;
    errno = 0;
    const long v = std::strtol(text, &end, 10);
    if (end == text || *end != '\0' || errno != 0) return false;
    out = v;
    return true;
}

bool readInt(const char* text, int& out) {
    long v = 0;
    if (!readLong(text, v) || v < INT_MIN || v > INT_MAX) return false;
    out = static_cast<int>(v);
    return true;
}

bool readUnsigned(const char* text, unsigned& out) {
    long v = 0;
    if (!readLong(text, v) || v < 0 || v > static_cast<long>(UINT_MAX)) return false;
    out = static_cast<unsigned>(v);
    return true;
}

struct Options {
    int cellsPerSide = 5;
    bool useNeighbourList = true;
    double dt = 0.005;
    double skin = 0.2;
    double cutOff = 2.5;
    int equilibrationSteps = 2000;
    int productionSteps = 5000;
    const char* csvPath = nullptr;
    bool quiet = false;
    bool deterministic = false;
    bool mixed = false;
    double density = 0.8442;
    double temperature = 0.728;
    bool nvt = false;
    const char* rdfPath = nullptr;
    const char* xyzPath = nullptr;
    int xyzEvery = 100;
    unsigned seed = 42;
    bool melt = false;
    bool listInitialForces = false;  // driver extension: initial forces over the list instead of all pairs
};

void printUsage() {
    std::fprintf(
        stderr,
        "Usage: md_sim [cellsPerSide] [--brute] [--dt X] [--skin X] [--cutoff X] [--equil N] [--steps N] [--density X] "
        "[--temperature X] [--nvt] [--seed N] [--melt]\n"
        "              [--csv PATH] [--rdf PATH] [--xyz PATH] [--xyz-every N] [--deterministic] [--mixed] [--quiet]\n"
        "              [--list-initial-forces]\n");
}

bool parseOptions(int argc, char** argv, Options& o) {
    for (int a = 1; a < argc; ++a) {
        const std::string_view arg = argv[a];
        const bool hasValue = a + 1 < argc;
        bool ok = true;
        if (arg == "--brute") {
            o.useNeighbourList = false;
        } else if (arg == "--quiet") {
            o.quiet = true;
        } else if (arg == "--deterministic") {
            o.deterministic = true;
        } else if (arg == "--mixed") {
            o.mixed = true;
        } else if (arg == "--nvt") {
            o.nvt = true;
        } else if (arg == "--melt") {
            o.melt = true;
        } else if (arg == "--list-initial-forces") {
            o.listInitialForces = true;
        } else if (arg == "--seed" && hasValue) {
            ok = readUnsigned(argv[++a], o.seed);
        } else if (arg == "--density" && hasValue) {
            ok = readDouble(argv[++a], o.density);
        } else if (arg == "--temperature" && hasValue) {
            ok = readDouble(argv[++a], o.temperature);
        } else if (arg == "--rdf" && hasValue) {
            o.rdfPath = argv[++a];
        } else if (arg == "--xyz" && hasValue) {
            o.xyzPath = argv[++a];
        } else if (arg == "--xyz-every" && hasValue) {
            ok = readInt(argv[++a], o.xyzEvery);
        } else if (arg == "--dt" && hasValue) {
            ok = readDouble(argv[++a], o.dt);
        } else if (arg == "--skin" && hasValue) {
            ok = readDouble(argv[++a], o.skin);
        } else if (arg == "--cutoff" && hasValue) {
            ok = readDouble(argv[++a], o.cutOff);
        } else if (arg == "--equil" && hasValue) {
            ok = readInt(argv[++a], o.equilibrationSteps);
        } else if (arg == "--steps" && hasValue) {
            ok = readInt(argv[++a], o.productionSteps);
        } else if (arg == "--csv" && hasValue) {
            o.csvPath = argv[++a];
        } else if (!arg.empty() && arg[0] != '-') {
            ok = readInt(argv[a], o.cellsPerSide);
        } else {
            ok = false;
        }
        if (!ok) return false;
    }
    return o.cellsPerSide > 0 && o.dt > 0.0 && o.skin > 0.0 && o.cutOff > 0.0 && o.density > 0.0 &&
           o.temperature > 0.0 && o.xyzEvery > 0 && o.productionSteps > 0 && o.equilibrationSteps >= 0;
}

// One step's record: potential, kinetic energy, virial, inside pairs.
struct Record {
    double potential, kinetic, virial, inside;
};

int run(const Options& o) {
    const double density = o.density, temperature = o.temperature, cutOff = o.cutOff, skin = o.skin, dt = o.dt;
    const int equilibrationSteps = o.equilibrationSteps, productionSteps = o.productionSteps;
    const int rescaleEvery = 10, printEvery = 250, rdfEvery = 10;
    const bool quiet = o.quiet;

    std::FILE* csv = nullptr;
    if (o.csvPath) {
        csv = std::fopen(o.csvPath, "w");
        if (!csv) {
            std::fprintf(stderr, "Could not open %s: %s\n", o.csvPath, std::strerror(errno));
            return 1;
        }
        std::fprintf(
            csv, "step,time,phase,ke_per_particle,pe_per_particle,e_per_particle,temperature,pressure,inside_pairs\n");
    }

    const double latticeConstant = std::cbrt(4.0 / density);
    const double boxLength = o.cellsPerSide * latticeConstant;
    System system(createFCCLattice(o.cellsPerSide, latticeConstant), boxLength);
    initializeVelocities(system, temperature, o.seed);

    LennardJonesForceField forceField(1.0, 1.0, cutOff);
    VelocityVerletIntegrator integrator(dt);
    NeighbourList neighbourList(cutOff, skin, o.deterministic);
    forceField.setDeterministic(o.deterministic);
    forceField.setMixedPrecision(o.mixed);
    if (o.mixed && (o.deterministic || !o.useNeighbourList)) {
        std::fprintf(stderr, "--mixed applies to the default neighbour list path only\n");
        return 1;
    }

    double potential = 0.0;
    if (o.listInitialForces && o.useNeighbourList) {
        neighbourList.build(system);
        potential = forceField.computeForces(system, neighbourList);
    } else {
        potential = forceField.computeForces(system);
        neighbourList.build(system);
    }
    const std::size_t count = system.size();
    const double n = static_cast<double>(count);
    const double dof = 3.0 * n - 3.0;
    const int threads = 1;

    if (!quiet) {
        std::printf("N = %zu, L = %.4f, rho = %.4f, dt = %.4f, forces = %s, threads = %d\n", count, boxLength,
                    density, dt, o.useNeighbourList ? "neighbour list" : "all pairs", threads);
        std::printf("\nEquilibration (%d steps, rescale every %d)\n", equilibrationSteps, rescaleEvery);
        std::printf("%8s %12s %12s %12s %12s\n", "step", "KE/N", "PE/N", "E/N", "T");
    }

    const double volume = boxLength * boxLength * boxLength;
    auto pressure = [&](double kinetic, double virial) { return (2.0 * kinetic + virial) / (3.0 * volume); };
    auto writeRow = [&](int step, const char* phase, double kinetic, const Record& r, double t) {
        std::fprintf(csv, "%d,%.6f,%s,%.8f,%.8f,%.8f,%.8f,%.8f,%.0f\n", step, step * dt, phase, kinetic / n,
                     r.potential / n, (kinetic + r.potential) / n, t, pressure(kinetic, r.virial), r.inside);
    };

    const md_lj lj{1.0, 1.0, cutOff};
    BussiThermostat thermostat(temperature, 0.5, o.seed);
    const double tau = 0.5;
    RadialDistribution rdf(cutOff, 150);
    std::vector<std::uint64_t> rdfCounts(150);
    // `steps` steps from the current state; with `nvt` the thermostat after each, with `sampleRdf` the g(r)
    // histogram after every rdfEvery-th step of the stretch
    auto advance = [&](int steps, bool nvt, bool sampleRdf) {
        std::vector<Record> out(static_cast<std::size_t>(steps));
        if (!o.useNeighbourList) {
            for (int k = 0; k < steps; ++k) {
                const double u = integrator.step(system, forceField);
                if (nvt) thermostat.apply(system, dt);
                out[k] = {u, system.kineticEnergy(), forceField.getVirial(), forceField.getInsidePairs()};
            }
            return out;
        }
        std::vector<double> draws(nvt ? 2 * static_cast<std::size_t>(steps) : 0);
        if (nvt) thermostat.drawPairs(count, steps, draws.data());
        md_run_config cfg{};
        cfg.dt = dt;
        cfg.precision = forceField.precision();
        cfg.steps = steps;
        cfg.lj = lj;
        cfg.thermostat = nvt ? 1 : 0;
        cfg.temperature = temperature;
        cfg.tau = tau;
        cfg.draws = nvt ? draws.data() : nullptr;
        cfg.use_graph = 1;
        cfg.exact_kinetic = 1;
        int samples = 0;
        std::fill(rdfCounts.begin(), rdfCounts.end(), 0);
        if (sampleRdf) {
            cfg.rdf_every = rdfEvery;
            cfg.rdf_bins = static_cast<int>(rdfCounts.size());
            cfg.rdf_rmax = cutOff;
            cfg.rdf_counts = rdfCounts.data();
            cfg.rdf_samples = &samples;
        }
        mdgpu_facade::check(
            md_run(system.deviceState(), neighbourList.deviceList(), &cfg, reinterpret_cast<double*>(out.data())));
        system.deviceWrote(System::All);
        neighbourList.deviceUpdated();
        if (samples > 0) rdf.addCounts(rdfCounts.data(), samples, system);
        // md_run's records are potential, kinetic, virial, inside
        return out;
    };

    for (int done = 0; done < equilibrationSteps;) {
        const int stretch = std::min(rescaleEvery - done % rescaleEvery, equilibrationSteps - done);
        const std::vector<Record> rec = advance(stretch, false, false);
        for (int k = 0; k < stretch; ++k) {
            const int step = done + k + 1;
            potential = rec[k].potential;
            double kinetic = rec[k].kinetic;
            if (step % rescaleEvery == 0) {
                const bool melting = o.melt && step <= equilibrationSteps / 3;
                rescaleVelocities(system, melting ? 2.0 : temperature);
                kinetic = system.kineticEnergy();
            }
            if (csv) writeRow(step, "equilibration", kinetic, rec[k], 2.0 * kinetic / dof);
            if (!quiet && step % printEvery == 0)
                std::printf("%8d %12.6f %12.6f %12.6f %12.6f\n", step, kinetic / n, potential / n,
                            (kinetic + potential) / n, 2.0 * kinetic / dof);
        }
        done += stretch;
    }

    if (o.rdfPath && !o.useNeighbourList) {
        std::fprintf(stderr, "--rdf needs the neighbour list (drop --brute)\n");
        return 1;
    }
    const double initialEnergy = system.kineticEnergy() + potential;
    double sumT = 0.0, sumT2 = 0.0, maxDrift = 0.0;
    if (!quiet) {
        std::printf("\nProduction (%d steps, %s)\n", productionSteps, o.nvt ? "NVT, Bussi thermostat" : "NVE");
        std::printf("%8s %12s %12s %12s %12s %12s\n", "step", "KE/N", "PE/N", "E/N", "T", "dE/E0");
    }
    std::FILE* xyz = nullptr;
    if (o.xyzPath) {
        xyz = std::fopen(o.xyzPath, "w");
        if (!xyz) {
            std::fprintf(stderr, "Could not open %s: %s\n", o.xyzPath, std::strerror(errno));
            return 1;
        }
    }
    // stretches end at every XYZ frame: xyzEvery steps long when that is a multiple of rdfEvery, so that md_run
    // samples the g(r) every rdfEvery-th step inside them; otherwise gcd(xyzEvery, rdfEvery) long, and the g(r) is
    // sampled between stretches (over the list pairs, which hold every pair within the cut-off)
    const int stretchLength = xyz ? (o.xyzEvery % rdfEvery == 0 ? o.xyzEvery : std::gcd(o.xyzEvery, rdfEvery))
                                  : productionSteps;
    const bool rdfInRun = o.rdfPath && stretchLength % rdfEvery == 0;
    // frames are captured on the GPU and printed by a writer thread while the next stretch runs
    std::unique_ptr<XyzWriter> frames = xyz ? std::make_unique<XyzWriter>(xyz) : nullptr;

    const auto start = std::chrono::steady_clock::now();
    for (int done = 0; done < productionSteps;) {
        const int stretch = std::min(stretchLength, productionSteps - done);
        const std::vector<Record> rec = advance(stretch, o.nvt, rdfInRun);
        for (int k = 0; k < stretch; ++k) {
            const int step = done + k + 1;
            potential = rec[k].potential;
            const double kinetic = rec[k].kinetic;
            const double total = kinetic + potential;
            const double t = 2.0 * kinetic / dof;
            const double drift = (total - initialEnergy) / std::abs(initialEnergy);
            sumT += t;
            sumT2 += t * t;
            maxDrift = std::max(maxDrift, std::abs(drift));
            if (csv) writeRow(equilibrationSteps + step, "production", kinetic, rec[k], t);
            if (!quiet && step % printEvery == 0)
                std::printf("%8d %12.6f %12.6f %12.6f %12.6f %12.2e\n", step, kinetic / n, potential / n, total / n, t,
                            drift);
        }
        done += stretch;
        if (o.rdfPath && !rdfInRun && done % rdfEvery == 0) rdf.sample(system, neighbourList);
        if (frames && done % o.xyzEvery == 0) frames->write(system, (equilibrationSteps + done) * dt);
    }
    if (frames) frames->flush();
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();

    if (csv) std::fclose(csv);
    frames.reset();
    if (xyz) std::fclose(xyz);
    if (o.rdfPath && !rdf.write(o.rdfPath)) {
        std::fprintf(stderr, "Could not write %s: %s\n", o.rdfPath, std::strerror(errno));
        return 1;
    }
    const double meanT = sumT / productionSteps;
    const double stdT = std::sqrt(std::max(0.0, sumT2 / productionSteps - meanT * meanT));
    if (!quiet) {
        if (o.nvt)
            std::printf("\n<T> = %.4f +/- %.4f (target %.4f), NVT: total energy not conserved by design\n", meanT, stdT,
                        temperature);
        else
            std::printf("\n<T> = %.4f +/- %.4f (target %.4f), max |dE/E0| = %.2e\n", meanT, stdT, temperature,
                        maxDrift);
        if (o.useNeighbourList) {
            const int totalSteps = equilibrationSteps + productionSteps;
            std::printf("Neighbour list: %d rebuilds in %d steps (every %.1f steps), %zu pairs\n",
                        neighbourList.getRebuildCount(), totalSteps,
                        static_cast<double>(totalSteps) / neighbourList.getRebuildCount(),
                        neighbourList.getPairCount());
        }
        std::printf("Production took %.3f s (%.4f ms per step)\n", seconds, 1000.0 * seconds / productionSteps);
    }
    std::printf("RESULT n=%zu threads=%d forces=%s dt=%g steps=%d seconds=%.6f ms_per_step=%.6f mean_T=%.6f std_T=%.6f "
                "max_drift=%.6e skin=%g cutoff=%g rebuilds=%d deterministic=%d state_hash=%016llx device=gpu\n",
                count, threads, o.useNeighbourList ? "list" : "brute", dt, productionSteps, seconds,
                1000.0 * seconds / productionSteps, meanT, stdT, maxDrift, skin, cutOff,
                neighbourList.getRebuildCount(), o.deterministic ? 1 : 0,
                static_cast<unsigned long long>(system.stateHash()));
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Options options;
    if (!parseOptions(argc, argv, options)) {
        printUsage();
        return 1;
    }
    try {
        return run(options);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "md_sim: %s\n", e.what());
        return 1;
    }
}
