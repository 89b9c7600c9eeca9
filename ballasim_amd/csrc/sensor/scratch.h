// The library's static device scratch (scratch.hip) and its layout for the advance call that
// allocates nothing (SensorModel::advance_device_noalloc). Every call on one GPU shares that GPU's
// scratch and keeps the last call's contents: calls that overlap on one GPU (several host threads or
// streams at once) race.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace bsx {

// The largest camera the no-allocation call serves: 2^22 pixels (2048 x 2048; the 2048 x 1536 and
// 1920 x 1080 sensors fit).
inline constexpr size_t kScratchMaxPixels = size_t{1} << 22;
// Layout: [0, 16) the run checks' flags (4 x u32), [16, 32) the sample checks' counters (2 x u64),
// [64, 64 + 8 (1 + kStCount)) the call's counters, [256, 1024) the camera descriptor, [1024, 1280) the
// call header, [4096, 4096 + 4 P) row_of.
inline constexpr size_t kScratchFlags = 0, kScratchChecks = 16, kScratchCounters = 64, kScratchCamera = 256,
                        kScratchCall = 1024, kScratchRowOf = 4096;
inline constexpr size_t kScratchBytes = kScratchRowOf + 4 * kScratchMaxPixels;  // 16 MiB + 4 KiB

// This GPU's scratch (the current device's), its address taken once per GPU. Throws
// std::runtime_error when the address cannot be taken.
void* scratch_address();

}  // namespace bsx
