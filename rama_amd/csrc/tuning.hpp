// tuning.hpp — launch shapes chosen per stage and size bucket.
//
// A table maps (knob, size) to a value: each knob has rows (max_size, value)
// in increasing max_size, the first row whose max_size is at least the size
// applies, and max_size -1 covers every larger size.  The table is read once
// from the file named by RAMA_TUNING_TABLE, else from tuning/gfx950.table
// beside librama_hip.so; without a file (or with RAMA_TUNING_TABLE=none)
// every knob keeps its built-in default.  The tuner (`bench.py --mode tune`)
// writes the file.  Every knob changes only how work is laid out, never the
// result: sort paths are stable sorts of the same keys, scan chunking sums
// integers, grid bounds change which block does an element's work.
#pragma once

#include <cstdint>

namespace rama {
namespace tune {

enum Knob : int {
    // size = pairs; 0 = three launches per pass (histogram, scan, scatter),
    // 1 = one onesweep launch per pass, 2 = a top-digit onesweep pass then one
    // block per bucket (keys wider than 8 bits, at most kBucketPathMaxTiles
    // tiles; otherwise 1)
    SortPathU64 = 0,
    SortPathU32,
    // size = radix tiles; 1024 or 512 threads per 4096-pair tile in the
    // scatter and onesweep passes
    SortThreads,
    // size = items; blocks of the two-launch integer scan (64 .. 4096)
    ScanChunks,
    // size = symmetric CSR entries; grid bound of the matching's pick kernel
    PickBlocks,
    // size = vertices; grid bound of the matching's pairing kernel
    MatchBlocks,
    // size = message-passing edges; edge-kernel grids up to this many blocks
    // sum the lower bound themselves (larger grids finalise in a launch of
    // their own)
    MpFusedMaxBlocks,
    // size = expansions; chunked 4/5-cycle counts size the next chunk as this
    // percentage of the blocks the rate so far expects to reach the budget
    CycleChunkPct,
    KnobCount
};

constexpr int kBucketPathMaxTiles = 1024;

// Value of knob k for a given size (table row, else the built-in default).
int64_t get(Knob k, int64_t size);

// Name of the knob as the table file spells it.
const char* name(Knob k);

}  // namespace tune
}  // namespace rama
