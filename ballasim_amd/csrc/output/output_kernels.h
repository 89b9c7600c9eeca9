// Launch wrappers of the output stage's kernels (output.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "output.h"

namespace bsx {

// scratch: [0] kept count, [1..4] min/max of t_cross and t_reported (biased), [5] released count.
inline constexpr int kOutScratch = 8;

// Appends the kept raw events as OutEventDev at out (count in scratch[0], which the caller zeroes).
hipError_t launch_out_build(const RawEventDev* raw, size_t n, uint32_t width, const Readout& r, bool crop,
                            const CropBox& box, OutEventDev* out, unsigned long long* scratch, hipStream_t stream);
// min/max of t_cross and t_reported over ev[0 .. n_held + scratch[0]) (grid over n_max).
hipError_t launch_out_minmax(const OutEventDev* ev, size_t n_held, size_t n_max, unsigned long long* scratch,
                             hipStream_t stream);
// keys[i] of pass (0: t_cross - tc_min, 1: (x << y_bits) | y, 2: t_reported - tr_min) of ev[perm[i]] (perm
// null: identity, perm_init filled with it).
hipError_t launch_out_keys(const OutEventDev* ev, const uint32_t* perm, uint32_t* perm_init, size_t n, int pass,
                           int y_bits, int64_t tc_min, int64_t tr_min, uint64_t* keys, hipStream_t stream);
// Stable radix sort of (keys, perm) by the key's low `bits` bits (n < 2^32); the sorted copies land in keys_out /
// perm_out, and keys / perm are overwritten.
hipError_t launch_out_sort(uint64_t* keys, uint64_t* keys_out, uint32_t* perm, uint32_t* perm_out, size_t n, int bits,
                           DeviceBuffer& temp, hipStream_t stream);
// sorted[i] = ev[perm[i]], counting those reported before cut into scratch[5].
hipError_t launch_out_gather(const OutEventDev* ev, const uint32_t* perm, size_t n, int64_t cut, OutEventDev* sorted,
                             unsigned long long* scratch, hipStream_t stream);
// The C4 columns of ev[0 .. n).
hipError_t launch_out_columns(const OutEventDev* ev, size_t n, int64_t t_offset_us, uint16_t* x, uint16_t* y,
                              uint8_t* p, int64_t* t_us, uint16_t* t_ns_rem, int64_t* t_rep, int64_t* t_cross,
                              uint8_t* label, uint16_t* flags, hipStream_t stream);

// The compact stream records of ev[0 .. n) at dst.
hipError_t launch_out_pack(const OutEventDev* ev, size_t n, StreamEventDev* dst, hipStream_t stream);

}  // namespace bsx
