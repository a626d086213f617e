// Host check of fix_round128 (pic_device.h, read from the double's bits)
// against its definition rint(ldexp(a, q)) on the FPU, converted exactly to a
// 128-bit integer (frexp's mantissa and exponent): exact below 2^126,
// saturating at 2^126, 0 for non-finite values and overflow.  Random bit
// patterns (every class of double), random values near the grids the kernels
// use, and edge cases (ties, powers of two, 2^62 and 2^126 boundaries).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

#include "pic_device.h"

using pic::i128;


// t = rint(ldexp(a, q)) is an exact integer-valued double; beyond 2^62 it is
// converted exactly as its 53-bit mantissa shifted by its exponent (frexp).
static i128 by_definition(double a, int q) {
    const double t = std::rint(std::ldexp(a, q));
    if (std::fabs(t) < 0x1p62) return (i128)(long long)t;
    if (t != t || std::isinf(t)) return 0;
    if (!(std::fabs(t) < 0x1p126)) return t > 0 ? ((i128)1 << 126) : -((i128)1 << 126);
    int E;
    const double f = std::frexp(std::fabs(t), &E);          // |t| = f * 2^E, f in [0.5, 1)
    const i128 v = (i128)(long long)std::ldexp(f, 53) << (E - 53);
    return t < 0 ? -v : v;
}

int main() {
    const int qs[] = {-60, -1, 0, 1, 20, 22, 30, 40, 44, 50, 52, 53, 60, 62, 63, 64, 73, 80, 93,
                      100, 110, 126, 200, 500, 971, 1000, 1100};
    std::mt19937_64 rng(12345);
    long long n = 0, bad = 0;
    auto check = [&](double a, int q) {
        ++n;
        if (pic::fix_round128(a, q) != by_definition(a, q)) {
            if (bad < 10) std::printf("mismatch a=%a q=%d\n", a, q);
            ++bad;
        }
    };
    const double edge[] = {0.0, -0.0, 0.5, -0.5, 1.5, 2.5, -2.5, 0x1p62, -0x1p62, 0x1.fffffffffffffp61,
                           0x1p126, 0x1.fffffffffffffp125, -0x1.fffffffffffffp125, 0x1p-1074, -0x1p-1022,
                           INFINITY, -INFINITY, NAN, 0x1.fffffffffffffp1023, 1.0, -1.0, 3.0};
    for (double a : edge)
        for (int q : qs)
            for (int dq = -3; dq <= 3; ++dq) check(a, q + dq);
    for (int i = 0; i < 2000000; ++i) {
        const u64 bits = rng();
        double a;
        std::memcpy(&a, &bits, sizeof a);
        check(a, qs[i % (sizeof qs / sizeof qs[0])]);
    }
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (int i = 0; i < 3000000; ++i) {
        const int q = qs[i % (sizeof qs / sizeof qs[0])];
        const int k = (int)(rng() % 160) - 80;             // a*2^q from 2^(q-80) to 2^(q+80)
        check(std::ldexp(u(rng), k), q);
        // ties: (integer + 1/2) * 2^-q
        const double t = std::ldexp(std::floor(std::ldexp(u(rng), 50)) + 0.5, -q);
        check(t, q);
    }
    std::printf("%lld values, %lld mismatches\n", n, bad);
    return bad != 0;
}
