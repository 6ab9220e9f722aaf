// Checks nf::par_memcpy_threads (hostcopy.cpp): for every size and thread
// count tried, every byte of the source lands in the destination and no byte
// past the end is written.  Exit status 0 on success; the first failure is
// printed.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "nf_internal.h"

static int check(size_t bytes, unsigned nt) {
    const size_t guard = 256;
    std::vector<unsigned char> src(bytes + 1), dst(bytes + guard, 0xA5);
    for (size_t i = 0; i < bytes; i++) src[i] = (unsigned char)(i * 2654435761u >> 13) | 1u;
    nf::par_memcpy_threads(dst.data(), src.data(), bytes, nt);
    for (size_t i = 0; i < bytes; i++)
        if (dst[i] != src[i]) {
            std::printf("bytes %zu threads %u: byte %zu not copied\n", bytes, nt, i);
            return 1;
        }
    for (size_t i = bytes; i < bytes + guard; i++)
        if (dst[i] != 0xA5) {
            std::printf("bytes %zu threads %u: byte %zu past the end written\n", bytes, nt, i);
            return 1;
        }
    return 0;
}

int main() {
    int bad = 0;
    for (unsigned nt = 1; nt <= 9 && !bad; nt++)
        for (size_t b = 0; b <= 4096 && !bad; b++) bad |= check(b, nt);
    // 12-byte particles: slab sizes around 3.9 million particles, 8 threads
    // (12 * 3900459 bytes left 4 bytes uncopied when chunks rounded down).
    const size_t slabs[] = {12 * (size_t)3900459, 12 * (size_t)3900460, 12 * (size_t)1234567 + 4, 16 * (size_t)999999};
    for (size_t b : slabs)
        for (unsigned nt = 2; nt <= 8 && !bad; nt++) bad |= check(b, nt);
    if (!bad) std::printf("par_memcpy_threads: every size and thread count copied exactly\n");
    return bad;
}
