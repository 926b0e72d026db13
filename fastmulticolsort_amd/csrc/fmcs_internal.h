// fmcs_internal.h -- services of the context that the host composer layer
// (fmcs_composer.cpp) uses beyond the public C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

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

// Address of the current device's copy of the static scratch array
// (fmcs_noalloc.hip), looked up once per device; the context's creation
// calls it with its device current.  On failure *err holds the reason.
int noalloc_setup(int device, void** ptr, std::string* err);
// The context's scratch address (FMCS_E_HIP if the lookup failed).
int scratch_of(fmcs_ctx* ctx, void** ptr);
// In-place LSD sort of (keys, oids) with every buffer carved out of
// [scratch, scratch + cap): no device allocation.
int sort_pairs_in_scratch(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, void* keys, uint32_t* oids,
                          void* scratch, size_t cap);

// Buffer arguments of an entry point.  Every non-null *_d pointer is
// classified (fmcs_pointer_kind) before anything is launched: device,
// pinned and managed memory pass; another GPU's memory fails with
// FMCS_E_ARG; pageable host memory fails with FMCS_E_ARG too, unless the
// context stages host buffers (fmcs_ctx_set_host_staging) and the entry
// point allows it, in which case the pointer is replaced by a device copy
// in a workspace slot: inputs are copied in at once, outputs are copied
// back (and the stream synchronized) by finish().
class Stager {
 public:
  enum Dir { IN = 1, OUT = 2, INOUT = 3 };
  explicit Stager(fmcs_ctx* ctx, bool allow_staging = true) : ctx_(ctx), allow_(allow_staging) {}
  template <class T>
  int arg(T*& p, uint64_t bytes, Dir dir, const char* name) {
    void* v = (void*)p;
    const int rc = arg_void(&v, bytes, dir, name);
    p = (T*)v;
    return rc;
  }
  int arg_void(void** p, uint64_t bytes, Dir dir, const char* name);
  // the copy-back of the output staged for the caller's pointer `host`
  // takes only `bytes` (an output whose length the call determines)
  void set_out_bytes(const void* host, uint64_t bytes);
  int finish();

 private:
  struct Out {
    void* host;
    void* dev;
    uint64_t bytes;
  };
  fmcs_ctx* ctx_;
  bool allow_;
  int next_ = 0;
  std::vector<Out> outs_;
};

}  // namespace internal
}  // namespace fmcs
