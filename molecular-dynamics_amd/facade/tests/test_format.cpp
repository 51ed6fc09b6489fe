// The XYZ line formatter (src/XyzFormat.hpp) against printf itself, the definition it must reproduce
// (Trajectory.cpp:21-22 prints with "%.8f" and "%zu"): exact ties (odd multiples of 2^-9, where |x| * 10^8 ends in
// exactly one half), their neighbours one ulp away, carries through every digit (9.999999995 and its powers of
// ten), zeros and negative values that round to zero, subnormals, the 1e10 edge of the direct range, integers,
// random bit patterns and the value ranges of positions and velocities. Runs on the CPU; returns the number of
// mismatches.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>
#include <string>

#include "../src/XyzFormat.hpp"

namespace {

long mismatches = 0, checked = 0;

void check(double x) {
    if (!xyz_format::fixed8_fits(x)) return;
    char want[64], got[64];
    const int n = std::snprintf(want, sizeof want, "%.8f", x);
    const char* end = xyz_format::fixed8(got, x);
    ++checked;
    if (end - got != n || std::memcmp(want, got, n) != 0) {
        if (++mismatches <= 20) {
            *const_cast<char*>(end) = '\0';
            std::printf("MISMATCH %a: printf %s, fixed8 %s\n", x, want, got);
        }
    }
}

void around(double x) {
    check(x);
    check(std::nextafter(x, 0.0));
    check(std::nextafter(x, 1e300));
    check(-x);
    check(-std::nextafter(x, 0.0));
    check(-std::nextafter(x, 1e300));
}

}  // namespace

int main() {
    // exact ties: q / 512 with q odd gives |x| * 10^8 = q * 195312.5
    for (std::int64_t q = 1; q < 2000000; q += 2) around(static_cast<double>(q) / 512.0);
    for (std::int64_t q = 1; q < 200000; q += 2) around(std::ldexp(static_cast<double>(q), -9) + 4096.0);
    // carries: 9.999999995, 99.999999995, ... and 0.000000005-style halves
    for (double p = 1.0; p < 1e10; p *= 10.0) {
        around(p * 9.999999995);
        around(p - 5e-9);
        around(p + 5e-9);
        around(p);
        around(5e-9 * p);
    }
    // zeros, tiny values, subnormals and the edge of the direct range
    for (double x : {0.0, -0.0, 5e-9, 4.9999999e-9, 5.0000001e-9, 1e-300, 4.9406564584124654e-324,
                     2.2250738585072014e-308, 9999999999.9999981, 9999999999.999998, 1e10, 0.5, 1.5, 2.5}) {
        around(x);
    }
    for (int e = -1074; e < 34; ++e) around(std::ldexp(1.0, e));
    // integers
    for (std::int64_t k = 0; k < 100000; ++k) around(static_cast<double>(k));
    std::mt19937_64 g(20241204);
    // random bit patterns (every exponent), then the ranges of positions and velocities
    for (int k = 0; k < 3000000; ++k) {
        const std::uint64_t bits = g();
        double x;
        std::memcpy(&x, &bits, sizeof x);
        check(x);
    }
    std::uniform_real_distribution<double> box(0.0, 600.0), vel(-6.0, 6.0), small(-1e-6, 1e-6);
    for (int k = 0; k < 3000000; ++k) {
        check(box(g));
        check(vel(g));
        check(small(g));
    }
    // whole lines, ids up to 2^64 - 1
    for (int k = 0; k < 300000; ++k) {
        const double r[3] = {box(g), box(g), box(g)}, v[3] = {vel(g), vel(g), vel(g)};
        const std::uint64_t id = k % 3 == 0 ? std::numeric_limits<std::uint64_t>::max() - k : g() >> (k % 64);
        char want[512], got[xyz_format::particle_line_max];
        const int n = std::snprintf(want, sizeof want, "Ar %.8f %.8f %.8f %.8f %.8f %.8f %" PRIu64 "\n", r[0], r[1],
                                    r[2], v[0], v[1], v[2], id);
        const char* end = xyz_format::particle_line(got, r, v, id);
        ++checked;
        if (end - got != n || std::memcmp(want, got, n) != 0) {
            if (++mismatches <= 20) std::printf("LINE MISMATCH: %s", want);
        }
    }
    std::printf("%ld values and lines checked, %ld mismatches\n", checked, mismatches);
    return mismatches == 0 ? 0 : 1;
}
