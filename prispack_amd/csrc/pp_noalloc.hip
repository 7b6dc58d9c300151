// pp_noalloc.hip -- pp_solve_noalloc: the FP64 CG solve with every work
// vector in one fixed device array instead of library allocations, for
// callers that cannot allocate device memory (code under graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// The array is a file-scope device global of kScratchBytes; each GPU the
// library runs on holds its own copy until the process ends.  Its address on
// a device is looked up once, at context creation with that device current
// (noalloc_setup), and only that address is ever passed to kernels or copies:
// in host code the array's name is a host-side stand-in, not a device address.
//
// Every call on a GPU shares that GPU's array and it keeps the last call's
// contents: the solve writes every work vector before reading it (r = b - A x0
// and the vectors derived from it), so nothing has to start at zero.  Calls
// that overlap on one GPU (two contexts, two host threads) race on the array.
#include <mutex>

#include "pp_internal.h"

namespace {

constexpr size_t kScratchBytes = pp::kNoallocScratchBytes;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
constexpr int kMaxDevices = 64;
void *g_ptr[kMaxDevices] = {};

}  // namespace

namespace pp {

void noalloc_setup(int device) {
    PP_CHECK(device >= 0 && device < kMaxDevices, PP_ERR_INVALID, "device id beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return;
    int cur = -1;
    PP_HIP(hipGetDevice(&cur));
    PP_CHECK(cur == device, PP_ERR_INVALID, "noalloc_setup: the device is not current");
    void *p = nullptr;
    PP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PP_CHECK(p, PP_ERR_HIP, "no device address for the scratch array");
    g_ptr[device] = p;
}

void *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

}  // namespace pp
