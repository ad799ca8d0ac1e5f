// Host check of div_by (titan_c-_amd/csrc/fdiv.h) against IEEE division: random operands over
// the exponent range the step sees, and operands built to sit on rounding boundaries.  Prints
// the number of mismatches and of cases tried; any mismatch is a failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../titan_c-_amd/csrc/fdiv.h"

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t next() {  // splitmix64
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double bits(uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; }
static uint64_t ubits(double d) { uint64_t b; std::memcpy(&b, &d, 8); return b; }

// a normal double with exponent in [lo, hi] and a random significand and sign
static double rnd(int lo, int hi) {
    const uint64_t m = next() & 0xFFFFFFFFFFFFFull;
    const int e = lo + (int)(next() % (uint64_t)(hi - lo + 1));
    return bits(((next() & 1) << 63) | ((uint64_t)(e + 1023) << 52) | m);
}

static long long bad = 0, tried = 0;
static void check(double a, double b) {
    const double y = 1.0 / b;
    const double q = titan::div_by(a, b, y);
    const double ref = a / b;
    tried++;
    if (ubits(q) != ubits(ref)) {
        if (bad < 10) std::printf("mismatch a=%a b=%a got=%a want=%a\n", a, b, q, ref);
        bad++;
    }
}

int main(int argc, char **argv) {
    const long long n = argc > 1 ? std::atoll(argv[1]) : 20000000;
    for (long long i = 0; i < n; i++) check(rnd(-300, 300), rnd(-300, 300));
    // denominators next to powers of two (all-ones and one-ulp significands), all numerators
    for (long long i = 0; i < n / 4; i++) {
        const int e = (int)(next() % 200) - 100;
        const double b1 = bits(((uint64_t)(e + 1023) << 52) | 0xFFFFFFFFFFFFFull);
        const double b2 = bits(((uint64_t)(e + 1023) << 52) | (next() & 7));
        check(rnd(-100, 100), b1);
        check(rnd(-100, 100), b2);
    }
    // numerators within a few ulps of an exact multiple of b: quotients on and next to
    // representable values and midpoints
    for (long long i = 0; i < n / 4; i++) {
        const double b = rnd(-50, 50);
        const double k = (double)(int64_t)(next() % 100000 + 1) * (next() & 1 ? 1.0 : 0.5);
        double a = k * b;
        const int d = (int)(next() % 7) - 3;
        uint64_t ab = ubits(a);
        ab += (int64_t)d;
        check(bits(ab), b);
    }
    // signed zeros and exact quotients
    check(0.0, 3.0); check(-0.0, 3.0); check(0.0, -3.0); check(-0.0, -3.0);
    check(6.0, 3.0); check(-1.0, 4.0); check(1.0, 3.0);
    std::printf("%lld %lld\n", bad, tried);
    return bad != 0;
}
