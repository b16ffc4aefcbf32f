// Device-wide exclusive scan and stable LSD radix sort (key/value u32 pairs),
// written for wave64: wave-level scans over 64 lanes, wave-local digit
// ranking with ballots, 4096-element tiles per 256-thread workgroup.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "workspace.hpp"

namespace splhip {

// out[i] = sum_{j<i} in[j] over n elements (in == out allowed).  If total is
// non-null the grand total is stored there (device).  Scratch: S_SCAN0+level.
// dn (device, optional): the element count is min(n, *dn), read on the device
// (launches sized for n elements; no host wait).
void scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *total, Workspace &ws,
              hipStream_t s, int level = 0, const uint32_t *dn = nullptr);

// Same with in[i] = ((flags[i] & bits) != 0).
void scan_u8bit(const uint8_t *flags, uint8_t bits, uint32_t *out, int64_t n, uint32_t *total,
                Workspace &ws, hipStream_t s);

// Stable sort of (keys, vals) by key bits [begin_bit, end_bit).  Ping-pongs
// between (k0,v0) and (k1,v1); returns 0 if the result is in (k0,v0), 1 if in
// (k1,v1).  Scratch: S_HIST and the scan slots.  dn (device, optional): the
// pair count is min(n, *dn), read on the device.
int radix_sort_pairs(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, int64_t n,
                     int begin_bit, int end_bit, Workspace &ws, hipStream_t s,
                     const uint32_t *dn = nullptr);

}  // namespace splhip
