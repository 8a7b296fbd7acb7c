// Checks the frontier cell's threshold note (exaca_amd/csrc/threshold.hpp): for any least length
// thr and next-least length thr2 >= thr of a cell's liquid neighbours, the note's bound lies in
// [thr, thr2] (a cell below it reaches no neighbour), and within 1/256 of thr2 when the gap is
// below 1.99. A copy that rounds the gap up instead of down must fail the first property on the
// same cases, which shows that they reach it. Prints the counts; exits nonzero on a failure.
#include <cmath>
#include <cstdio>
#include <limits>
#include <random>
#include <vector>

#include "threshold.hpp"

using exaca::hint_bound;
using exaca::threshold_hint;

// the note with its gap rounded up (a planted mistake)
static float bound_rounded_up(float thr, float thr2) {
    const double gap = (static_cast<double>(thr2) - static_cast<double>(thr)) * 256.0;
    const int dq = gap >= 511.0 ? 511 : static_cast<int>(std::ceil(gap));
    return static_cast<float>(static_cast<double>(thr) + dq * (1.0 / 256.0));
}

int main() {
    std::mt19937_64 gen(2026);
    std::uniform_real_distribution<double> len(0.01, 4.0), gapd(0.0, 3.0);
    std::uniform_int_distribution<int> nb(0, 25), q(0, 1200), ulps(-3, 3);
    long long cases = 0, bad = 0, loose = 0, planted = 0;
    auto check = [&](float thr, float thr2, int at) {
        cases++;
        const uint16_t h = threshold_hint(at, thr, thr2);
        const float b = hint_bound(thr, h);
        if (((h >> 2) & 31) != static_cast<uint16_t>(at + 1)) bad++;
        if (!(b >= thr && b <= thr2)) bad++;
        if (thr2 - thr < 1.99f && thr2 - b > 1.0f / 256 + 1e-6f) loose++;
        if (bound_rounded_up(thr, thr2) > thr2) planted++;
    };
    for (int i = 0; i < 2000000; i++) {
        const float thr = static_cast<float>(len(gen));
        check(thr, static_cast<float>(thr + gapd(gen)), nb(gen));
        // next-least lengths a whole number of 1/256 above thr, give or take a few ulps: the
        // sums the note forms land next to them
        float t2 = static_cast<float>(static_cast<double>(thr) + q(gen) / 256.0);
        for (int u = ulps(gen); u > 0; u--) t2 = std::nextafter(t2, 10.0f);
        for (int u = ulps(gen); u < 0; u++) t2 = std::nextafter(t2, 0.0f);
        if (t2 >= thr) check(thr, t2, nb(gen));
    }
    // equal lengths, no second length, no neighbour
    check(1.25f, 1.25f, 3);
    check(1.25f, std::numeric_limits<float>::infinity(), 7);
    if (threshold_hint(-1, 1.0f, 2.0f) != 0 || hint_bound(1.0f, 0) != 1.0f) bad++;
    std::printf("cases %lld failures %lld loose %lld planted_failures %lld\n", cases, bad, loose, planted);
    return (bad == 0 && loose == 0 && planted > 0) ? 0 : 1;
}
