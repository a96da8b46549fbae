// Extra device state of the single-large-world mode (big_world.hip): the
// linear BVH, candidate pairs, the existing-contact key set and the
// colouring.  The world itself lives in a one-world DeviceState.
#pragma once

#include <cstdint>

#include "layout.hpp"

namespace prx {

constexpr int kBigBlock = 1024;     // threads of the single-workgroup kernels
constexpr int kBigMaxColors = 64;   // colours (one bit each in a body's mask) before the serial bucket
constexpr int kBigColorRounds = 48; // colouring rounds before the rest goes to the serial bucket

// Indices into BigState::counters.
enum : int { kBigPairs = 0, kBigColors = 1, kBigTouching = 2, kBigMinSepBits = 3, kBigCounters = 8 };
// Indices into BigState::metrics.
enum : int { kBigMinSeparation = 0, kBigMetrics = 4 };

struct BigState {
    int P, B, cap, pairCap, hashMask;
    // Linear BVH: internal nodes 0..P-2 (0 is the root), leaf k at P-1+k.
    std::uint32_t* morton;
    std::uint32_t* morton2;
    int* order;   // sorted position -> proxy
    int* order2;
    int* left;
    int* right;
    int* parent;  // 2P-1
    int* escape;  // 2P-1
    int* visit;   // P-1
    float4* naabb;  // 2P-1
    float4* bounds;
    // Candidate pairs.
    int* pcount;   // P + 1
    int2* pairs;   // pairCap
    int* paccept;  // pairCap
    std::uint64_t* hash;  // hashMask + 1
    // Colouring and solve order.
    int* list;
    int* list2;
    int* color;
    int* claim;       // B * kBigMaxColors: lowest contact proposing (body, colour); kept at INT_MAX between rounds
    std::uint64_t* used;  // B: colours taken at the body
    int* tent;        // cap: proposed colour
    int* win;         // cap
    int* colorStart;  // kBigMaxColors + 2
    int* counters;    // kBigCounters
    float* metrics;   // kBigMetrics
};

}  // namespace prx
