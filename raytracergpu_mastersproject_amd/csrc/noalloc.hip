// noalloc.hip -- the static scratch array behind rt_render_noalloc.
//
// rt_render takes its scratch (per-sample storage, bounce-by-bounce path state and queues) from the
// context's cache, which allocates device memory when no cached block is large enough.  Callers that
// cannot allocate (code under HIP graph capture, loops that forbid hipMalloc) use rt_render_noalloc,
// whose scratch is this file-scope device array instead: kNoAllocScratchBytes per GPU, held from the
// loading of the library's code object until the process ends.  Every call on a GPU shares that
// GPU's array, so calls that overlap on one GPU race; the kernels clear the counters and queue
// lengths they use before each pass, and write every other word before reading it.
#include "rt_internal.hpp"

namespace rtgpu {
namespace {

constexpr uint64_t kNoAllocScratchBytes = 256ull << 20;  // one sample of a 2.2 Mpixel image bounce by bounce

__device__ __align__(256) char scratch[kNoAllocScratchBytes];

}  // namespace

// With the device whose copy is wanted current.  In host code `scratch` names a host-side stand-in,
// never the device memory: only the address returned here may be handed to kernels.
void *noalloc_scratch_address() {
    void *p = nullptr;
    RT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    RT_REQUIRE(p != nullptr, "hipGetSymbolAddress returned a null address for the static scratch array");
    return p;
}

uint64_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

}  // namespace rtgpu
