// fta_hostcopy.h -- one large host memory copy split over a few threads.
// Tracer arrays run to hundreds of MB, and one core's memcpy reaches a
// fraction of the host's memory bandwidth.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <thread>
#include <vector>

namespace fta {

inline void copy_host(void *dst, const void *src, size_t bytes)
{
    const size_t min_part = (size_t)8 << 20;
    const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes / min_part));
    if (nt == 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes / nt + 63) & ~(size_t)63;
    std::vector<std::thread> th;
    size_t done = 0;  // bytes handed to threads so far (from the start)
    try {
        th.reserve(nt);
        for (int i = 0; i < nt; ++i) {
            const size_t o = (size_t)i * part;
            if (o >= bytes) break;
            const size_t n = std::min(part, bytes - o);
            th.emplace_back([=] { std::memcpy((char *)dst + o, (const char *)src + o, n); });
            done = o + n;
        }
    } catch (...) {
        // a thread could not be started: the caller's thread copies the rest
    }
    if (done < bytes) std::memcpy((char *)dst + done, (const char *)src + done, bytes - done);
    for (std::thread &t : th) t.join();
}

}  // namespace fta
