// fta_noalloc.hip -- the device scratch of the allocation-free entry point
// (tracer_do_oce_adv_tra_noalloc): a file-scope device array, so callers that
// cannot allocate (code under graph capture, tight serving loops, harnesses
// that forbid hipMalloc) get a step whose scratch already exists.
//
// Each GPU the library runs on holds its own copy of the array until the
// process ends. Host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol): the device address of this
// GPU's copy is looked up once per device with hipGetSymbolAddress, with that
// device current, and that address is what the kernels receive. Every call on
// one GPU shares its copy, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "fta_engine.h"

namespace fta {

namespace {

constexpr size_t kScratchDoubles = (size_t)64 << 17;  // 64 MiB of doubles (8M)

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

double *noalloc_scratch(size_t *doubles)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range for the scratch table");
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    if (doubles) *doubles = kScratchDoubles;
    return g_scratch_ptr[dev];
}

}  // namespace fta
