// Host-side copies of slabs between pageable memory and pinned staging
// (no HIP here: the planner's sanitizer build links this file too).
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "nf_internal.h"

namespace nf {

// memcpy split over nt threads in 64-byte aligned chunks of ceil(bytes / nt)
// rounded up to 64, so the nt chunks cover every byte (the last may be short
// or empty).
void par_memcpy_threads(void *dst, const void *src, size_t bytes, unsigned nt) {
    if (nt <= 1 || bytes == 0) {
        if (bytes) std::memcpy(dst, src, bytes);
        return;
    }
    const size_t chunk = ((bytes + nt - 1) / nt + 63) & ~size_t(63);
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; t++) {
        const size_t o = t * chunk;
        if (o >= bytes) break;
        th.emplace_back([=] { std::memcpy((char *)dst + o, (const char *)src + o, std::min(chunk, bytes - o)); });
    }
    std::memcpy(dst, src, std::min(chunk, bytes));
    for (auto &x : th) x.join();
}

// One thread does not reach the host's memory bandwidth, and a staged slab is
// tens of MB: up to 8 threads, at least 4 MiB each.
void par_memcpy(void *dst, const void *src, size_t bytes) {
    const size_t kMinChunk = 4u << 20;
    unsigned nt = std::min<unsigned>(8, std::max(1u, std::thread::hardware_concurrency() / 2));
    nt = (unsigned)std::min<size_t>(nt, std::max<size_t>(1, bytes / kMinChunk));
    par_memcpy_threads(dst, src, bytes, nt);
}

}  // namespace nf
