// Extra device state of the single-large-world mode (big_world.hip): the
// linear BVH, candidate pairs, the existing-contact key set and the
// colouring.  The world itself lives in a one-world DeviceState.
#pragma once

#include <cstdint>

#include "layout.hpp"

namespace prx {

constexpr int kBigBlock = 1024;     // threads of the single-workgroup kernels
constexpr int kBigMaxColors = 16;   // colours (one bit each in a body's mask) before the serial bucket
// Early-exit slots of the sweeps (BigState::sweep): the largest impulse change
// of each velocity iteration (float bits) and the smallest separation of each
// position iteration (order-preserving keys).
constexpr int kBigVelSlot = 0, kBigPosSlot = 16, kBigSweepSlots = 32;
constexpr int kBigColorRounds = 40; // colouring rounds before the rest goes to the serial bucket

// Indices into BigState::counters.
// kBigEpoch counts colourings (claim keys are stamped with it); kBigLeft and
// kBigLeft + 1 hold the uncoloured count of the last two rounds; kBigRounds
// the rounds the last colouring took.
enum : int {
    kBigPairs = 0, kBigColors = 1, kBigTouching = 2, kBigMinSepBits = 3, kBigEpoch = 4, kBigLeft = 5,
    kBigRounds = 7, kBigCounters = 8
};
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
    // B * kBigMaxColors: the smallest key proposing (body, colour) in a round;
    // key = (~epoch << 32) | list position, so a later round's keys are
    // smaller than every stale one and no reset is needed.
    unsigned long long* claim;
    std::uint64_t* used;  // B: colours taken at the body
    int* tent;        // cap: proposed colour
    int* win;         // cap
    int* colorStart;  // kBigMaxColors + 2
    // B: the earliest time of impact against a static shape, (float bits << 32
    // | contact index), ~0 for none.
    unsigned long long* toiKey;
    float* toiLeft;  // B: the part of the step a body has left to move (1 before its first event)
    int* sweep;       // kBigSweepSlots
    int* counters;    // kBigCounters
    float* metrics;   // kBigMetrics
};

}  // namespace prx
