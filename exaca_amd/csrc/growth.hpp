// growth.hpp -- the half-diagonal growth recurrence of a CA cell, advanced many steps at once.
//
// Each executed step past a cell's liquidus time stores dl = float(double(dl) + g_k)
// (src/CAupdate.cpp:257-261). On the table's plateau (the 0.045 cap) g is a constant c. For dl a
// positive normal float in the binade [2^(e-1), 2^e), dl is a multiple of u = 2^(e-24), and the
// double sum dl + c rounds to float the same way for every such dl: the sum's bits below u come
// from c alone, at the same double precision across the binade. Unless that rounding is a tie,
// every step inside the binade therefore adds the same delta, and the interior of a binade is
// crossed in one jump. Steps next to a binade boundary, ties and non-normal values are taken one at
// a time. Host and device share these functions; the host copy is what the tests check against
// the step-by-step recurrence.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace exaca {

__host__ __device__ inline float grow_step(float x, double g) {
    return static_cast<float>(static_cast<double>(x) + g);
}

__host__ __device__ inline uint32_t float_bits(float x) { return __builtin_bit_cast(uint32_t, x); }

// e with 2^(e-1) <= x < 2^e for a positive normal float.
__host__ __device__ inline int binade_of(float x) { return static_cast<int>((float_bits(x) >> 23) & 0xFFu) - 126; }

// Whether the step x -> y = grow_step(x, c) gives the constant delta of x's binade; if so, delta
// and the number of further steps from y whose results stay two steps clear of the binade's top.
__host__ __device__ inline bool binade_run(float x, float y, double c, double &delta, int &room) {
    const uint32_t be = (float_bits(x) >> 23) & 0xFFu;
    if ((float_bits(x) >> 31) || be == 0 || be == 0xFFu) return false;
    const int e = binade_of(x);
    if (binade_of(y) != e) return false;
    const double u = std::ldexp(1.0, e - 24);
    if (std::fabs((static_cast<double>(x) + c) - static_cast<double>(y)) == 0.5 * u) return false;  // tie
    delta = static_cast<double>(y) - static_cast<double>(x);
    if (!(delta > 0)) return false;
    const double r = std::floor((std::ldexp(1.0, e) - 2.0 * delta - static_cast<double>(y)) / delta) - 1.0;
    room = r > 0 ? (r < 1.0e9 ? static_cast<int>(r) : 1000000000) : 0;
    return true;
}

// x after n steps of the constant increment c.
__host__ __device__ inline float advance_const(float x, int n, double c) {
    while (n > 0) {
        const float y = grow_step(x, c);
        n--;
        double delta = 0;
        int room = 0;
        if (n == 0 || !binade_run(x, y, c, delta, room)) {
            x = y;
            continue;
        }
        const int m = room < n ? room : n;
        x = static_cast<float>(static_cast<double>(y) + m * delta);
        n -= m;
    }
    return x;
}

// First n in 1 .. horizon at which x after n steps of the constant increment c reaches thr;
// horizon when it does not.
__host__ __device__ inline int reach_const(float x, double c, float thr, int horizon) {
    int n = 0;
    while (n < horizon) {
        const float y = grow_step(x, c);
        n++;
        if (y >= thr) return n;
        double delta = 0;
        int room = 0;
        if (n == horizon || !binade_run(x, y, c, delta, room)) {
            x = y;
            continue;
        }
        const int m = room < horizon - n ? room : horizon - n;
        // the values after further steps are y + i delta (i = 1 .. m), all exact in double
        const double yd = static_cast<double>(y), t = static_cast<double>(thr);
        double need = std::ceil((t - yd) / delta);
        if (need < 1) need = 1;
        if (need <= m) {
            int i = static_cast<int>(need);
            while (i > 1 && yd + (i - 1) * delta >= t) i--;
            while (yd + i * delta < t) i++;
            return n + i;
        }
        x = static_cast<float>(yd + m * delta);
        n += m;
    }
    return horizon;
}

}  // namespace exaca
