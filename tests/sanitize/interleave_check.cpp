// Checks nf::interleave_pos (nf_internal.h), the merge of a work list's run
// of M tiles with its run of T short tails: for every M, T up to 300 the
// positions are a permutation of [0, M + T), each run keeps its own order,
// the tails are spread evenly (the tiles between two consecutive tails, and
// before the first, differ in number by at most one; likewise the tails
// between consecutive tiles when T > M), and a nonempty tail run ends the
// list.  Exit status 0 on success; the first failure is printed.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "nf_internal.h"

static int check(int64_t M, int64_t T) {
    const int64_t n = M + T;
    std::vector<int> who(n, -1);  // 0: a tile, 1: a tail
    int64_t prev[2] = {-1, -1};
    for (int k = 0; k < 2; k++)
        for (int64_t i = 0; i < (k ? T : M); i++) {
            const int64_t p = nf::interleave_pos(k == 1, i, M, T);
            if (p < 0 || p >= n || who[p] != -1) {
                std::printf("M %lld T %lld: %s %lld at position %lld (out of range or taken)\n", (long long)M,
                            (long long)T, k ? "tail" : "tile", (long long)i, (long long)p);
                return 1;
            }
            if (p <= prev[k]) {
                std::printf("M %lld T %lld: %s run out of order at %lld\n", (long long)M, (long long)T,
                            k ? "tail" : "tile", (long long)i);
                return 1;
            }
            who[p] = k;
            prev[k] = p;
        }
    if (T > 0 && who[n - 1] != 1) {
        std::printf("M %lld T %lld: the list does not end with a tail\n", (long long)M, (long long)T);
        return 1;
    }
    // Lengths of the runs of one kind cut by the items of the other: at
    // most one apart, for the rarer kind spread through the more common one.
    // Tiles are counted in the run before each tail (none follow the last
    // tail); tails in the run after each tile (none precede the first tile).
    for (int k = 0; k < 2; k++) {
        const int sep = k, fill = 1 - k;  // runs of `fill` items cut by `sep` items
        const int64_t nsep = sep ? T : M, nfill = fill ? T : M;
        if (nsep == 0 || nfill < nsep) continue;
        std::vector<int64_t> runs(1, 0);  // nsep + 1 runs: before the first separator, ..., after the last
        for (int64_t p = 0; p < n; p++) {
            if (who[p] == fill) runs.back()++;
            else runs.push_back(0);
        }
        const size_t first = sep == 1 ? 0 : 1, last = sep == 1 ? runs.size() - 2 : runs.size() - 1;
        int64_t lo = INT64_MAX, hi = -1;
        for (size_t r = first; r <= last; r++) {
            lo = runs[r] < lo ? runs[r] : lo;
            hi = runs[r] > hi ? runs[r] : hi;
        }
        if (hi - lo > 1) {
            std::printf("M %lld T %lld: runs of %s between %ss range %lld..%lld\n", (long long)M, (long long)T,
                        fill ? "tails" : "tiles", sep ? "tail" : "tile", (long long)lo, (long long)hi);
            return 1;
        }
    }
    return 0;
}

int main() {
    int bad = 0;
    for (int64_t M = 0; M <= 300 && !bad; M++)
        for (int64_t T = 0; T <= 300 && !bad; T++) bad |= check(M, T);
    // Sizes of a launch group's lists: ~2 million tiles, ~100k tails.
    const int64_t big[][2] = {{2000003, 99991}, {1, 1000000}, {1000000, 1}, {1234567, 1234567}};
    for (const auto &mt : big)
        if (!bad) bad |= check(mt[0], mt[1]);
    if (!bad) std::printf("interleave_pos: every merge is an even permutation\n");
    return bad;
}
