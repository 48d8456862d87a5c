// Launch epochs of the tagged words (shared by host and device code).
//
// A published word carries the epoch of the launch that wrote it; a consumer
// accepts a word only when its tag equals the current launch's epoch. Tag 0
// marks memory never written (buffers start zeroed). The adjoint's (P, G)
// pairs carry row_tag(epoch, row), the epoch with the row's parity in its top
// bit, so that the pairs of two adjacent rows never carry the same tag; a
// copy of a pair left in LDS by the previous row's poll can then never pass
// for the current row's. Row tags are 0 only for epochs 0 and 2^31, so the
// sequence of launch epochs skips both: 1, ..., 2^31 - 1, 2^31 + 1, ...,
// 2^32 - 1, 1, ... A never-written word can then never look current. (A
// written word looks current to a launch about 2^32 launches of the same
// buffer later only if none of those launches rewrote it, while the launch
// reads it before writing it; the launches of a buffer rewrite every word they
// read.)
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
    return (n == 0u || n == 0x80000000u) ? n + 1u : n;
}

DDRS_EPOCH_HD inline uint32_t row_tag(uint32_t epoch, int row) {
    return epoch ^ (static_cast<uint32_t>(row & 1) << 31);
}

}  // namespace ddrs
