// threshold.hpp -- the note a frontier cell keeps about its capture threshold (host and device).
//
// After an evaluation a frontier cell keeps thr, the least critical length of its liquid
// neighbours not reached, and in bits 2-15 of its flags a note: bits 2-6 hold 1 + the neighbour
// whose length set thr (0: none), bits 7-15 a count d of 1/256 steps such that thr + d/256, in
// float, is at most the next-least length of those neighbours (0: no note). Liquid sets only
// shrink and the lengths are fixed, so when that neighbour has left the mask every liquid
// neighbour left is at least that far: a cell woken by thr whose half-diagonal is below the bound
// reaches none, and only its schedule changes (thr becomes the bound). Such a wake-up reads no row
// of lengths.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace exaca {

constexpr uint16_t kHintMask = 0xFFFC;

// The note for thr (set by neighbour at, -1: none) and the next-least length thr2 (+inf: none).
__host__ __device__ inline uint16_t threshold_hint(int at, float thr, float thr2) {
    if (at < 0) return 0;
    // the gap in whole 1/256 steps, rounded down: thr + dq/256 is exact in double and at most
    // thr2, and rounding it to float cannot pass thr2, which is a float (rounding is monotone)
    const double gap = (static_cast<double>(thr2) - static_cast<double>(thr)) * 256.0;  // +inf when none
    const int dq = gap >= 511.0 ? 511 : static_cast<int>(gap);
    return static_cast<uint16_t>(((at + 1) << 2) | (dq << 7));
}
// The bound the note gives (thr when there is none).
__host__ __device__ inline float hint_bound(float thr, uint16_t flags) {
    return static_cast<float>(static_cast<double>(thr) + (flags >> 7) * (1.0 / 256.0));
}

}  // namespace exaca
