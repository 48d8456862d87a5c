// Launch epochs of the tagged words (shared by host and device code).
//
// A published word carries the epoch of the launch that wrote it; a consumer
// accepts a word only when its tag equals the current launch's epoch. Epoch 0
// tags memory never written (buffers start zeroed), so the sequence of
// launch epochs skips it when the 32-bit counter wraps: 1, 2, ..., 2^32 - 1,
// 1, 2, ... A never-written word can then never look current. (A written word
// looks current to a launch exactly 2^32 - 1 launches of the same buffer
// later only if none of those launches rewrote it, while the launch reads it
// before writing it; the launches of a buffer rewrite every word they read.)
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define DDRS_EPOCH_HD __host__ __device__
#else
#define DDRS_EPOCH_HD
#endif

namespace ddrs {

DDRS_EPOCH_HD inline uint32_t next_epoch(uint32_t e) {
    const uint32_t n = e + 1u;
    return n ? n : 1u;
}

}  // namespace ddrs
