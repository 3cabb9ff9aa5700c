// LSD radix sort of (u64 key, u32 value[, u64 aux]) pairs, 8-bit digits, stable.
//
// Three kernels per pass: per-tile digit histograms, one scan over (digit, tile), and a
// stable scatter that ranks equal digits inside each 64-lane wave with 8 ballots.
#pragma once

#include "td3_common.h"

namespace td3sort {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr int kRadix = 256;

inline i64 tiles(i64 n) { return (n + kTile - 1) / kTile; }

// Bytes of scratch one sort or scan of n items needs: the digit histograms (tiles x radix)
// and the per-tile sums of the scan over them.
inline size_t scratch_bytes(i64 n) {
    return (size_t)(tiles(n) * kRadix + tiles(tiles(n) * kRadix) + 64) * sizeof(u32);
}

// Sorts keys[0:n) (with values, and aux when non-null) by bits [bit_lo, bit_hi) of the key.
// The result lands in keys/vals/aux (the alt buffers are ping-pong space of the same
// size).  When vals_identity is set, the input values are taken as 0..n-1 and `vals`
// need not be initialised.
int sort_pairs(u64* keys, u32* vals, u64* aux, u64* keys_alt, u32* vals_alt, u64* aux_alt,
               i64 n, int bit_lo, int bit_hi, bool vals_identity, u32* scratch,
               hipStream_t stream);

// Exclusive scan of n u32 counts into `out` (may alias `in`); writes the total to *total
// (device) when non-null.  scratch: scratch_bytes(n) bytes.
int exclusive_scan_u32(const u32* in, u32* out, i64 n, u32* total, u32* scratch,
                       hipStream_t stream);

}  // namespace td3sort
