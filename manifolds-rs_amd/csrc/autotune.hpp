// Launch shapes chosen by measurement.
//
// A few kernels have more than one valid launch shape whose speed depends on
// the problem: the edge epoch's edges per wavefront (how many firing edges a
// wavefront compacts before it runs the negative-sampling body; the grid size
// follows) and the unpruned kNN sweep's database split count (how many
// workgroups share one query block's database).  Every shape of a family
// gives the same result: kNN lists are bit-identical, and edge-epoch spans
// change only the order of the FP32 atomic additions, which that strategy
// leaves unspecified anyway.
//
// The table maps (family, shape class) to a variant.  A shape class is a
// packed key of what decides the speed (dimension, edge count or rows as a
// power of two, precision).  The first call that meets an undecided class
// times every candidate on its own data with HIP events and records the
// fastest; later calls use the record.  The table can be saved to and loaded
// from a text file, so a later process starts with the decisions made.
// Tuning can be switched off (built-in choices) and a family's variant can be
// forced (tests run every variant that way).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mrs {

// family ids, as in include/mrs_amd.h
constexpr int kTuneEpochSpan = 1;
constexpr int kTuneKnnSplits = 2;

// ceil(log2(x)) for x >= 1
inline uint32_t ceil_log2(uint64_t x) {
    uint32_t r = 0;
    while ((1ull << r) < x) r++;
    return r;
}

// The forced variant of a family, or -1.
int tune_forced(int family);
// The recorded variant for (family, shape), or -1 if undecided.
int tune_lookup(int family, uint64_t shape);
// Whether undecided classes are tuned on first use.
bool tune_enabled();
// Record a decision: the candidates and their measured times (ms, same order).
// Returns the chosen (fastest) variant.
int tune_record(int family, uint64_t shape, const std::vector<int>& candidates,
                const std::vector<float>& ms);

}  // namespace mrs
