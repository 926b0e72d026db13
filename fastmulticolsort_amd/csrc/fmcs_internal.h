// fmcs_internal.h -- services of the context that the host composer layer
// (fmcs_composer.cpp) uses beyond the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fmcs.h"

namespace fmcs {
namespace internal {

constexpr int USER_SLOTS = 24;

hipStream_t stream_of(fmcs_ctx* ctx);
int set_error(fmcs_ctx* ctx, int code, const char* msg);
// Context-owned device buffer `slot` (0..USER_SLOTS-1), grown on demand.
int scratch(fmcs_ctx* ctx, int slot, size_t bytes, void** out);
// Synthetic bank: value i = top `bits` bits of SplitMix64(seed, i).
int fill_random(fmcs_ctx* ctx, void* p, uint32_t bank, uint64_t n, uint64_t seed, uint32_t bits);
// off[s] = min(s * size, n) for s in [0, nseg].
int uniform_offsets(fmcs_ctx* ctx, uint32_t* off, uint32_t nseg, uint64_t n, uint32_t size);

}  // namespace internal
}  // namespace fmcs
