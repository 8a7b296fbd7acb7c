// Checks the jumped growth recurrence (exaca_amd/csrc/growth.hpp) against the step-by-step one on
// random and boundary cases; prints the number of mismatches and exits nonzero on any.
#include <cstdio>
#include <random>

#include "growth.hpp"

using exaca::advance_const;
using exaca::grow_step;
using exaca::reach_const;

static float naive_advance(float x, int n, double c) {
    for (int i = 0; i < n; i++) x = grow_step(x, c);
    return x;
}
static int naive_reach(float x, double c, float thr, int horizon) {
    for (int n = 1; n <= horizon; n++) {
        x = grow_step(x, c);
        if (x >= thr) return n;
    }
    return horizon;
}

int main() {
    std::mt19937_64 gen(12345);
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    std::vector<double> incs = {0.045, 0.045 * (1 + 1e-12), 1e-4, 0.01, 0.3, 1.5};
    // ties for x in [1, 2) (ulp 2^-23) and [2, 4): increments that are (q + 1/2) ulps
    incs.push_back(12.5 * std::ldexp(1.0, -23));
    incs.push_back(377.5 * std::ldexp(1.0, -22));
    incs.push_back(0.5 * std::ldexp(1.0, -23));
    long long cases = 0, bad = 0;
    for (int t = 0; t < 400000; t++) {
        double c = incs[t % incs.size()];
        if (t % 7 == 6) c = std::exp(std::log(1e-5) + u01(gen) * (std::log(2.0) - std::log(1e-5)));
        float x;
        if (t % 5 == 0) {  // next to a power of two
            const int e = static_cast<int>(u01(gen) * 12) - 6;
            x = std::nextafter(static_cast<float>(std::ldexp(1.0, e)), 0.0f);
            for (int k = static_cast<int>(u01(gen) * 40); k > 0; k--) x = std::nextafter(x, 0.0f);
        }
        else {
            x = static_cast<float>(std::exp(std::log(1e-3) + u01(gen) * (std::log(100.0) - std::log(1e-3))));
        }
        const int n = 1 + static_cast<int>(u01(gen) * 300);
        const float a = advance_const(x, n, c), b = naive_advance(x, n, c);
        cases++;
        if (__builtin_bit_cast(uint32_t, a) != __builtin_bit_cast(uint32_t, b)) {
            if (bad < 5) printf("advance x=%.9g n=%d c=%.17g: %.9g vs %.9g\n", x, n, c, a, b);
            bad++;
        }
        const float thr = static_cast<float>(x + u01(gen) * 4.0 * c * n);
        const int h = 1 + static_cast<int>(u01(gen) * 100);
        const int r1 = reach_const(x, c, thr, h), r2 = naive_reach(x, c, thr, h);
        cases++;
        if (r1 != r2) {
            if (bad < 5) printf("reach x=%.9g c=%.17g thr=%.9g h=%d: %d vs %d\n", x, c, thr, h, r1, r2);
            bad++;
        }
    }
    printf("cases %lld mismatches %lld\n", cases, bad);
    return bad ? 1 : 0;
}
