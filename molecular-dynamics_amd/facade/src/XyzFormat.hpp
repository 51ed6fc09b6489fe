#pragma once

// The text of one extended-XYZ particle line, "Ar %.8f %.8f %.8f %.8f %.8f %.8f %zu\n" (Trajectory.cpp:21-22),
// formatted without printf: the same characters, an order of magnitude faster, so that a writer thread keeps up
// with the GPU on large frames.
#include <cstdint>
#include <cstring>

namespace xyz_format {

// Appends what printf("%.8f", x) prints and returns the new end. A finite double is m * 2^e exactly (m < 2^53);
// |x| * 10^8 = m * 10^8 * 2^e, and m * 10^8 < 2^80 fits 128-bit integers, so N = |x| * 10^8 rounded to the
// nearest integer, ties to even (printf's rounding in the default rounding mode), is computed exactly from the
// quotient and remainder of a right shift by -e. The text is then N / 10^8, '.', and N mod 10^8 as eight digits,
// after a '-' for any negative x, -0 and values that round to zero included, as printf prints them. Needs
// |x| < 1e10 (finite): at most 20 characters.
// "00" "01" ... "99"
inline constexpr char digit_pairs[201] =
    "00010203040506070809101112131415161718192021222324252627282930313233343536373839404142434445464748495051525354"
    "555657585960616263646566676869707172737475767778798081828384858687888990919293949596979899";

inline bool fixed8_fits(double x) {
    return x < 1e10 && x > -1e10;  // false for NaN
}

inline char* fixed8(char* p, double x) {
    std::uint64_t bits;
    std::memcpy(&bits, &x, sizeof bits);
    const bool negative = bits >> 63;
    const int biased = static_cast<int>((bits >> 52) & 0x7ff);
    std::uint64_t m = bits & ((std::uint64_t(1) << 52) - 1);
    int e = -1074;
    if (biased != 0) {
        m |= std::uint64_t(1) << 52;
        e = biased - 1075;
    }
    using u128 = unsigned __int128;
    const u128 scaled = static_cast<u128>(m) * 100000000u;
    u128 q = 0;
    if (e >= 0) {
        q = scaled << e;  // not reached below 1e10 (there e <= -19); kept for completeness
    } else if (-e < 128) {
        const int shift = -e;
        q = scaled >> shift;
        const u128 rest = scaled - (q << shift);
        const u128 half = static_cast<u128>(1) << (shift - 1);
        if (rest > half || (rest == half && (q & 1))) ++q;
    }  // else: scaled < 2^80 <= 2^(shift - 48), below half a unit: N = 0
    const std::uint64_t rounded = static_cast<std::uint64_t>(q);  // below 1e18: |x| < 1e10
    std::uint64_t whole = rounded / 100000000u;
    std::uint32_t frac = static_cast<std::uint32_t>(rounded % 100000000u);
    if (negative) *p++ = '-';
    char digits[24];
    int k = 0;
    do {
        digits[k++] = static_cast<char>('0' + whole % 10);
        whole /= 10;
    } while (whole);
    while (k) *p++ = digits[--k];
    *p++ = '.';
    for (int d = 6; d >= 0; d -= 2) {
        std::memcpy(p + d, digit_pairs + 2 * (frac % 100), 2);
        frac /= 100;
    }
    return p + 8;
}

// Appends what printf("%zu", v) prints.
inline char* unsigned_decimal(char* p, std::uint64_t v) {
    char digits[24];
    int k = 0;
    do {
        digits[k++] = static_cast<char>('0' + v % 10);
        v /= 10;
    } while (v);
    while (k) *p++ = digits[--k];
    return p;
}

// One particle line of a frame whose six values all satisfy fixed8_fits; `out` needs room for 6 * 21 + 32
// characters. Returns the end.
constexpr int particle_line_max = 6 * 21 + 32;
inline char* particle_line(char* out, const double* r, const double* v, std::uint64_t id) {
    char* p = out;
    *p++ = 'A';
    *p++ = 'r';
    for (int k = 0; k < 3; ++k) {
        *p++ = ' ';
        p = fixed8(p, r[k]);
    }
    for (int k = 0; k < 3; ++k) {
        *p++ = ' ';
        p = fixed8(p, v[k]);
    }
    *p++ = ' ';
    p = unsigned_decimal(p, id);
    *p++ = '\n';
    return p;
}

}  // namespace xyz_format
