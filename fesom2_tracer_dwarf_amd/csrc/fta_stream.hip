// fta_stream.hip -- a plain streaming copy on gfx950: the practical HBM
// ceiling that the step's bandwidth figures are set against.
//
// Each lane moves 16-byte words; a block keeps UNROLL words per lane in
// flight (all loads issued before the first store) and sweeps one contiguous
// range, and the block ranges are laid out so that each of the 8 XCDs (the
// hardware sends workgroup b to XCD b % 8) streams one contiguous eighth of
// the buffer. Loads and stores optionally carry the nontemporal hint (data
// touched once need not displace anything in the caches).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "fta_engine.h"

namespace fta {

namespace {

// a native 16-byte vector (the nontemporal builtins take clang vector types)
typedef unsigned int w16 __attribute__((ext_vector_type(4)));

template <int UNROLL, bool NT>
__global__ __launch_bounds__(256) void k_stream_copy(const w16 *__restrict__ src, w16 *__restrict__ dst,
                                                     long long n16, long long per_block)
{
    const unsigned nb = gridDim.x, b = blockIdx.x, per = nb >> 3;
    const unsigned lb = (b & 7u) * per + (b >> 3);
    const long long lo = (long long)lb * per_block;
    const long long hi = lo + per_block < n16 ? lo + per_block : n16;
    constexpr int kStep = 256 * UNROLL;
    for (long long i = lo + threadIdx.x; i < hi; i += kStep) {
        w16 v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const long long j = i + (long long)u * 256;
            if (j < hi) {
                if constexpr (NT) v[u] = __builtin_nontemporal_load(src + j);
                else v[u] = src[j];
            }
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const long long j = i + (long long)u * 256;
            if (j < hi) {
                if constexpr (NT) __builtin_nontemporal_store(v[u], dst + j);
                else dst[j] = v[u];
            }
        }
    }
}

#define HIP_CHECK(call)                                                                                       \
    do {                                                                                                      \
        hipError_t _e = (call);                                                                               \
        if (_e != hipSuccess) throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(_e));    \
    } while (0)

}  // namespace

// best read+write rate (bytes moved both ways / time, GB/s) of reps copies of
// `bytes` over the launch shapes and variants tried; best_variant receives
// {blocks, unroll, nontemporal} of the winner
double hbm_copy_rate(long long bytes, int reps, int best_variant[3])
{
    if (bytes < (1 << 20) || reps < 1) throw std::runtime_error("copy of at least 1 MiB and 1 repetition");
    const long long n16 = bytes / 16;
    w16 *a = nullptr, *b = nullptr;
    HIP_CHECK(hipMalloc(&a, (size_t)n16 * 16));
    if (hipMalloc(&b, (size_t)n16 * 16) != hipSuccess) {
        (void)hipFree(a);
        throw std::runtime_error("hipMalloc failed for the copy target");
    }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double best = 0.0;
    try {
        HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipMemsetAsync(a, 1, (size_t)n16 * 16, st));
        int cus = 256;
        {
            int dev = 0;
            HIP_CHECK(hipGetDevice(&dev));
            hipDeviceProp_t pr;
            HIP_CHECK(hipGetDeviceProperties(&pr, dev));
            cus = std::max(8, pr.multiProcessorCount);
        }
        for (int bpc : {2, 4, 8}) {
            const unsigned nb = (unsigned)((cus * bpc + 7) / 8 * 8);  // a multiple of 8 for the XCD layout
            const long long per_block = (n16 + nb - 1) / nb;
            for (int v = 0; v < 4; ++v) {
                const bool nt = v & 1;
                const int unroll = (v & 2) ? 8 : 4;
                auto launch = [&]() {
                    if (unroll == 4 && !nt)
                        hipLaunchKernelGGL((k_stream_copy<4, false>), dim3(nb), dim3(256), 0, st, a, b, n16, per_block);
                    else if (unroll == 4)
                        hipLaunchKernelGGL((k_stream_copy<4, true>), dim3(nb), dim3(256), 0, st, a, b, n16, per_block);
                    else if (!nt)
                        hipLaunchKernelGGL((k_stream_copy<8, false>), dim3(nb), dim3(256), 0, st, a, b, n16, per_block);
                    else
                        hipLaunchKernelGGL((k_stream_copy<8, true>), dim3(nb), dim3(256), 0, st, a, b, n16, per_block);
                };
                launch();  // warm-up
                HIP_CHECK(hipGetLastError());
                HIP_CHECK(hipEventRecord(e0, st));
                for (int r = 0; r < reps; ++r) launch();
                HIP_CHECK(hipEventRecord(e1, st));
                HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                const double gbs = 2.0 * (double)n16 * 16.0 * reps / ((double)ms * 1e-3) / 1e9;
                if (gbs > best) {
                    best = gbs;
                    best_variant[0] = (int)nb;
                    best_variant[1] = unroll;
                    best_variant[2] = nt ? 1 : 0;
                }
            }
        }
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
        (void)hipFree(a);
        (void)hipFree(b);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    HIP_CHECK(hipFree(a));
    HIP_CHECK(hipFree(b));
    return best;
}

}  // namespace fta
