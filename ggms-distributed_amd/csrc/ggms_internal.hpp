// Library-internal limits shared by the host driver and the kernels.
#pragma once
#include <stdint.h>

namespace ggms {

// Largest bucket count of any pass (LDS histogram of 64 KB plus slack for the
// first pass's edge and one-key ranges).
constexpr uint32_t kMaxBuckets = 16384;
// Largest number of pieces of the learned / sketch splitter.
constexpr uint32_t kMaxPieces = 512;
// Largest grid of the streaming kernels (one histogram row per workgroup).
constexpr uint32_t kMaxWorkgroups = 4096;
// Largest vector (and batched problem) length: the reference's `int length`
// (DistributedSMOS/iterativeSMOS.cuh:218-220).  Positions are 32-bit in every
// kernel; below 2^31 no position plus one workgroup stride (<= 1024) wraps.
constexpr uint64_t kMaxN = 0x7FFFFFFFull;
// Smallest tile worth a workgroup of the streaming kernels.
constexpr uint32_t kMinTile = 16384;

}  // namespace ggms
