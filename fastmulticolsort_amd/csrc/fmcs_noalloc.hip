// fmcs_noalloc.hip -- the allocation-free form of fmcs_sort_pairs.
//
// Callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc) sort through
// fmcs_sort_pairs_noalloc.  Its buffers come from `scratch`, a file-scope
// device array of FMCS_NOALLOC_SCRATCH_BYTES that every GPU the library runs
// on holds until the process ends.  In host code the array's name is only a
// host-side stand-in, so its device address is looked up once per device
// (hipGetSymbolAddress, with that device current, from fmcs_ctx_create) and
// only that address is handed to kernels.
#include <mutex>
#include <string>

#include "../../include/fmcs.h"
#include "fmcs_internal.h"

static __device__ __align__(256) unsigned char scratch[FMCS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};
std::string g_err[MAX_DEVICES];
bool g_done[MAX_DEVICES] = {};

}  // namespace

namespace fmcs {
namespace internal {

int noalloc_setup(int device, void** ptr, std::string* err) {
  *ptr = nullptr;
  if (device < 0 || device >= MAX_DEVICES) {
    *err = "device index above the scratch table";
    return FMCS_E_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_done[device]) {
    void* p = nullptr;
    // ::scratch, the array: inside this namespace the bare name would find
    // the workspace function fmcs::internal::scratch instead
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      g_err[device] = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
      p = nullptr;
    }
    g_ptr[device] = p;
    g_done[device] = true;
  }
  *ptr = g_ptr[device];
  if (!*ptr) {
    *err = g_err[device];
    return FMCS_E_HIP;
  }
  return FMCS_OK;
}

}  // namespace internal
}  // namespace fmcs

extern "C" int fmcs_sort_pairs_noalloc(fmcs_ctx* ctx, uint32_t bank, uint32_t keybits, uint64_t n, void* keys_d,
                                       uint32_t* oids_d) {
  using namespace fmcs::internal;
  if (!ctx) return FMCS_E_ARG;
  if (!(bank == 2 || bank == 4 || bank == 8)) return set_error(ctx, FMCS_E_ARG, "bank must be 2, 4 or 8");
  if (keybits == 0 || keybits > 8 * bank) return set_error(ctx, FMCS_E_ARG, "bad keybits");
  if (n > FMCS_NOALLOC_MAX_ROWS)
    return set_error(ctx, FMCS_E_ARG, "fmcs_sort_pairs_noalloc: n exceeds FMCS_NOALLOC_MAX_ROWS (1048576)");
  if (n && (!keys_d || !oids_d)) return set_error(ctx, FMCS_E_ARG, "null keys or oids");
  if (n == 0) return FMCS_OK;
  // checked, never staged: staging would allocate
  Stager st(ctx, false);
  int rc = st.arg(keys_d, n * bank, Stager::INOUT, "keys_d");
  if (rc == FMCS_OK) rc = st.arg(oids_d, n * 4, Stager::INOUT, "oids_d");
  if (rc != FMCS_OK) return rc;
  void* p;
  rc = scratch_of(ctx, &p);
  if (rc != FMCS_OK) return rc;
  return sort_pairs_in_scratch(ctx, bank, keybits, n, keys_d, oids_d, p, FMCS_NOALLOC_SCRATCH_BYTES);
}
