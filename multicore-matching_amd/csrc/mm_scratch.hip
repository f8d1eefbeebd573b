// The static device scratch of mm_graph_check_noalloc, in a source of its
// own: the runtime loads a source's device code (and sets up its device
// arrays) when one of its symbols is first used, so only callers of the
// no-allocation check pay for this array.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "mm_graph_impl.h"
#include "mm_internal.h"

namespace {

// A file-scope device array, one copy per GPU the library runs on, held until
// the process ends.  Host code must not pass `scratch` itself to a kernel (on
// the host it names a stand-in, not a device address): it uses the address
// hipGetSymbolAddress returns for the current device, fetched once per device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

}  // namespace

int32_t mm_noalloc_scratch(int device, char **out, size_t *bytes)
{
    *out = nullptr;
    *bytes = 0;
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return mm::set_error(MM_ERR_INVALID, "mm_graph_check_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return mm::set_errorf(MM_ERR_HIP, "mm_graph_check_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        if (!p) return mm::set_error(MM_ERR_HIP, "mm_graph_check_noalloc: no device address for the scratch array");
        g_scratch_ptr[device] = p;
    }
    *out = (char *)g_scratch_ptr[device];
    *bytes = NOALLOC_SCRATCH_BYTES;
    return MM_OK;
}
