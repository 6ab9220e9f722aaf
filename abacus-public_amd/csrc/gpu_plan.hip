// Device-side SIC planning: the SetInteractionCollection constructor
// (PencilOnPencil.cc:180-418, non-PARALLEL layout) computed on the GPU from
// the W slabs' cellinfo, plus the runtime's derived lists (the source blocks'
// parent pencils, the co-add chunks, the batched work list).  Every array has
// exactly the contents the host planner (planner.cpp) and the runtime's host
// code give it, except the work list, whose order within a cost bucket may
// differ (each sink belongs to exactly one item, so the results do not depend
// on the order).
//
// Four launches per SIC, none with a per-thread loop longer than the pencil
// width (items are counted in closed form, and a pencil's full tiles are
// placed as one contiguous run):
//   A0  zero the totals (and, for the first SIC of a group, the work-list
//       histogram and cursors);
//   A   per pencil: counts, blocks, chunks, items, pair totals, cell checks
//       -> totals (atomics); the host reads them to lay the plan out;
//   B1  sink pencils: SinkPlan, SinkSetCount, the work-list histogram;
//       source pencils: SourcePlan, SourceSetCount; and per 256-pencil
//       block the sums of padded blocks / chunks (for the prefix sums);
//   B2  each workgroup adds up the block sums before it, scans its own
//       pencils, and writes the starts, IdMax, block parents, interaction
//       list, y offsets, co-add chunks and its pencils' work items (bucket
//       offsets from an in-LDS scan of the histogram).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nf_internal.h"

namespace nf {
namespace {

constexpr int kB = NF_BLOCK_SIZE;
constexpr int kT = 256;  // threads per workgroup of the planner kernels
static_assert(kPlanBuckets == 2 * kT, "k_plan_b2 scans the histogram two buckets per thread");

__device__ __forceinline__ int wrapi(int i, int n) { return i >= n ? i - n : (i < 0 ? i + n : i); }
__device__ __forceinline__ int nblocks(int n) { return (n + kB - 1) / kB; }
__host__ __device__ inline int ngroups(int n) { return (n + kT - 1) / kT; }

// Sum over the wave, then one atomic per wave.
__device__ __forceinline__ void wave_add(uint64_t *p, uint64_t v) {
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, off), hi = __shfl_xor((uint32_t)(v >> 32), off);
        v += ((uint64_t)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicAdd((unsigned long long *)p, (unsigned long long)v);
}

// The same for a 32-bit counter.
__device__ __forceinline__ void wave_add32(int32_t *p, int32_t v) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(p, v);
}

// Whether a planned cell (start as the plan stores it) lies inside its slab.
__device__ __forceinline__ bool cell_ok(uint64_t start, int32_t n, uint64_t nslab) {
    return n >= 0 && (uint64_t)(uint32_t)start + (uint64_t)n <= nslab;
}

// Row of sink pencil (j, k) in the central slab: cells (j + j_low, k + c).
__device__ __forceinline__ const nf_cellinfo *sink_row(const GPlanIn &in, int j) {
    return in.ci[in.R] + (size_t)(j + in.j_low) * in.cpd;
}
// Cell of source pencil (jj, k) in each slab: y = jj + j_low - R, z =
// zcen(k) = k + R (both periodic).
__device__ __forceinline__ int source_cell(const GPlanIn &in, int jj, int k) {
    return wrapi(jj + in.j_low - in.R, in.cpd) * in.cpd + wrapi(k + in.R, in.cpd);
}
__device__ __forceinline__ int source_pencil_count(const GPlanIn &in, int jj, int k) {
    const int cell = source_cell(in, jj, k);
    int n = 0;
    for (int c = 0; c < in.W; c++) n += in.ci[c][cell].count;
    return n;
}
// Sources sink pencil (j, k) meets: its W source pencils (j + y, k).
__device__ __forceinline__ uint64_t sink_pencil_sources(const GPlanIn &in, int j, int k) {
    uint64_t n = 0;
    for (int y = 0; y < in.W; y++) n += (uint64_t)source_pencil_count(in, j + y, k);
    return n;
}

// The work items of a pencil of n sinks (meeting nsrc sources) for register
// tile S, as the runtime's host build_items makes them: `full` tiles of 64 S
// sinks (kind 0), then one-wave tiles (kind 2) while at least kTailLanes
// sinks remain -- q of 64 sinks and, if the rest s is at least kTailLanes,
// one of s -- then one short tail of s sinks (kind 1) if 0 < s < kTailLanes.
// A pencil whose register tile would cost more than split_cost takes no
// register tiles: one-wave tiles from its first sink.
struct ItemSplit {
    int full, tile, n2, q, s, tail;
    __device__ ItemSplit(int n, int S, uint64_t nsrc, uint64_t split_cost) {
        tile = kB * S;
        full = (S > 1 && (uint64_t)tile * nsrc > split_cost) ? 0 : n / tile;
        const int r = n - full * tile;
        q = r / kB;
        s = r - q * kB;
        n2 = q + (s >= kTailLanes ? 1 : 0);
        tail = (s > 0 && s < kTailLanes) ? 1 : 0;
    }
    __device__ int count() const { return full + n2 + tail; }
};

// 1/8-octave cost buckets, longest first (cost 0 last).
__device__ __forceinline__ int cost_bucket(uint64_t cost) {
    if (cost == 0) return kPlanBuckets - 1;
    const int b = (int)floor(8.0 * log2((double)cost));
    return kPlanBuckets - 1 - min(max(b, 0), kPlanBuckets - 1);
}

// ---- Workgroup scans (kT threads = 4 waves) --------------------------------

// Exclusive prefix of v over the workgroup; *total gets the sum.  Every
// thread must call it.
__device__ __forceinline__ int block_excl_scan(int v, int *total) {
    __shared__ int wsum[kT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(x, off);
        if (lane >= off) x += u;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < kT / 64; w++) {
        before += w < wave ? wsum[w] : 0;
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// Sum of part[0, n) over the workgroup (every thread gets it).
__device__ __forceinline__ int block_sum_prefix(const int32_t *part, int n) {
    int v = 0;
    for (int i = threadIdx.x; i < n; i += kT) v += part[i];
    int total;
    (void)block_excl_scan(v, &total);
    return total;
}

// ---- Phase A ---------------------------------------------------------------

__global__ void k_plan_zero(GPlanTotals *tot, int32_t *shared, int nshared) {
    uint64_t *p = reinterpret_cast<uint64_t *>(tot);
    constexpr int n = sizeof(GPlanTotals) / sizeof(uint64_t);
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
    for (int i = threadIdx.x; i < nshared; i += blockDim.x) shared[i] = 0;
}

// Threads [0, P): sink pencils; [P, P+Q): source pencils; bid is the
// workgroup's index within the SIC's grid.  group (or null): the launch
// group's shared area, which receives this SIC's work items per tile and its
// bad cells.
__device__ void plan_a(const GPlanIn &in, GPlanTotals *tot, int32_t *group, int bid) {
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const int t = bid * kT + threadIdx.x;
    uint64_t sink = 0, sblk = 0, ch = 0, it[3] = {0, 0, 0}, occ[4] = {0, 0, 0, 0}, src = 0, qblk = 0, bad = 0,
             d = 0, pd = 0;
    if (t < P) {
        const int j = t / in.cpd, k = t % in.cpd;
        const nf_cellinfo *row = sink_row(in, j);
        int n = 0;
        for (int c = 0; c < in.W; c++) {
            const nf_cellinfo &ci = row[wrapi(k + c, in.cpd)];
            n += ci.count;
            bad += cell_ok(ci.startindex, ci.count, in.nslab[in.R]) ? 0 : 1;
        }
        sink = (uint64_t)n;
        sblk = (uint64_t)nblocks(n);
        ch = (uint64_t)nblocks(row[k].count);  // pencil k starts with cell z = k: its chunks are that cell's
        const uint64_t nsrc = sink_pencil_sources(in, j, k);
        for (int ti = 0; ti < 3; ti++) it[ti] = (uint64_t)ItemSplit(n, 1 << ti, nsrc, in.split_cost[ti]).count();
        const int bin = n < 64 ? 0 : n < 256 ? 1 : n < 1024 ? 2 : 3;
        for (int i = 0; i < 4; i++) occ[i] = bin == i ? (uint64_t)n : 0;
        d = (uint64_t)n * nsrc;
        pd = sblk * kB * nsrc;
    } else if (t < P + Q) {
        const int u = t - P;
        const int cell = source_cell(in, u / in.cpd, u % in.cpd);
        int n = 0;
        for (int c = 0; c < in.W; c++) {
            const nf_cellinfo &ci = in.ci[c][cell];
            n += ci.count;
            bad += cell_ok(ci.startindex_with_ghost, ci.count, in.nslab[c]) ? 0 : 1;
        }
        src = (uint64_t)n;
        qblk = (uint64_t)nblocks(n);
    }
    wave_add(&tot->sink_total, sink);
    wave_add(&tot->sink_blocks, sblk);
    wave_add(&tot->chunks, ch);
    for (int i = 0; i < 3; i++) wave_add(&tot->items[i], it[i]);
    for (int i = 0; i < 4; i++) wave_add(&tot->occ[i], occ[i]);
    wave_add(&tot->source_total, src);
    wave_add(&tot->source_blocks, qblk);
    wave_add(&tot->bad, bad);
    wave_add(&tot->direct_total, d);
    wave_add(&tot->padded_direct_total, pd);
    if (group) {
        for (int i = 0; i < 3; i++) wave_add32(group + kPlanGroupItems + i, (int32_t)it[i]);
        wave_add32(group + kPlanGroupBad, bad ? 1 : 0);  // pencils with a bad cell
    }
}
__global__ __launch_bounds__(kT) void k_plan_a(GPlanIn in, GPlanTotals *tot, int32_t *group) {
    plan_a(in, tot, group, blockIdx.x);
}

// ---- Phase B ---------------------------------------------------------------

// Workgroups [0, GP): sink pencils kT at a time; [GP, GP+GQ): source pencils.
__device__ void plan_b1(const GPlanIn &in, const GPlanOut &out, int bid) {
    if (out.group && out.group[kPlanGroupBad]) return;  // a cell outside its slab: plan nothing
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const int GP = ngroups(P);
    if (bid < GP) {
        // The work-list histogram is built in LDS, then added to the group's
        // with one atomic per workgroup and bucket: nearly every pencil of a
        // uniform slab falls in the same bucket.
        __shared__ int32_t lh[kPlanLists * kPlanBuckets];
        for (int i = threadIdx.x; i < kPlanLists * kPlanBuckets; i += kT) lh[i] = 0;
        __syncthreads();
        const int s = bid * kT + threadIdx.x;
        int nb = 0, nc = 0;
        if (s < P) {
            const int j = s / in.cpd, k = s % in.cpd;
            const nf_cellinfo *row = sink_row(in, j);
            nf_sink_plan pl;
            int n = 0;
            for (int c = 0; c < NF_MAX_WIDTH; c++) {
                if (c < in.W) {
                    const nf_cellinfo &ci = row[wrapi(k + c, in.cpd)];
                    pl.cell[c].start = (uint32_t)ci.startindex;
                    pl.cell[c].N = ci.count;
                    n += ci.count;
                } else {
                    pl.cell[c].start = 0;
                    pl.cell[c].N = 0;
                }
            }
            // The cell that sits at z = 0 after the periodic wrap
            // (PencilPlan.cc:123-128): k + c = cpd, if within the pencil.
            const int wc = (in.cpd - k) % in.cpd;
            pl.wrapcell = wc < in.W ? wc : 0;
            out.sink_plan[s] = pl;
            out.sink_count[s] = n;
            nb = nblocks(n);
            nc = nblocks(pl.cell[0].N);
            if (n > 0) {  // the work-list histogram, per wanted tile
                const uint64_t nsrc = sink_pencil_sources(in, j, k);
                for (int ti = 0; ti < 3; ti++) {
                    if (!out.items[ti]) continue;
                    const ItemSplit sp(n, 1 << ti, nsrc, in.split_cost[ti]);
                    int32_t *h = lh + ti * kPlanBuckets;
                    if (sp.full) atomicAdd(h + cost_bucket((uint64_t)sp.tile * nsrc), sp.full);
                    if (sp.n2) atomicAdd(h + cost_bucket((uint64_t)kB * nsrc), sp.n2);
                    if (sp.tail) atomicAdd(h + 3 * kPlanBuckets + cost_bucket((uint64_t)sp.s * nsrc), 1);
                }
            }
        }
        int tb, tc;
        (void)block_excl_scan(nb, &tb);  // (its barriers also order the LDS histogram)
        (void)block_excl_scan(nc, &tc);
        if (threadIdx.x == 0) {
            out.part[bid] = tb;
            out.part[GP + bid] = tc;
        }
        for (int i = threadIdx.x; i < kPlanLists * kPlanBuckets; i += kT)
            if (lh[i]) atomicAdd(out.hist + i, lh[i]);
    } else {
        const int g = bid - GP;
        const int u = g * kT + threadIdx.x;
        int nb = 0;
        if (u < Q) {
            const int cell = source_cell(in, u / in.cpd, u % in.cpd);
            nf_source_plan pl;
            int n = 0;
            for (int c = 0; c < NF_MAX_WIDTH; c++) {
                if (c < in.W) {
                    const nf_cellinfo &ci = in.ci[c][cell];
                    pl.cell[c].start = (uint32_t)ci.startindex_with_ghost;
                    pl.cell[c].N = ci.count;
                    n += ci.count;
                } else {
                    pl.cell[c].start = 0;
                    pl.cell[c].N = 0;
                }
            }
            out.src_plan[u] = pl;
            out.src_count[u] = n;
            nb = nblocks(n);
        }
        int tb;
        (void)block_excl_scan(nb, &tb);
        if (threadIdx.x == 0) out.part[2 * GP + g] = tb;
    }
}
__global__ __launch_bounds__(kT) void k_plan_b1(GPlanIn in, GPlanOut out) { plan_b1(in, out, blockIdx.x); }

// The work items of the workgroup's pencils, taken in the host list's order
// (thread t: pencil (j, k) = (t % jw, t / jw), so neighbours in y -- which
// share 2R of their 2R+1 source pencils -- land next to each other in a
// bucket).  Each item run gets its place in two steps: an offset within the
// workgroup's share of its bucket (LDS atomics), then the share's base (one
// global atomic per workgroup and bucket).  boff: the buckets' offsets in
// each list's run of tiles (0-2) and of short tails (3-5); mt: the runs'
// lengths, M then T per list.  The two runs merge by interleave_pos.  Every
// thread of the workgroup must call it.
__device__ void place_items(const GPlanIn &in, const GPlanOut &out, const int32_t (*boff)[kPlanBuckets],
                            const int32_t *mt, int bid) {
    __shared__ int32_t lc[kPlanLists * kPlanBuckets];
    const int P = in.jw * in.cpd;
    for (int i = threadIdx.x; i < kPlanLists * kPlanBuckets; i += kT) lc[i] = 0;
    __syncthreads();
    const int t = bid * kT + threadIdx.x;
    const int s = t < P ? (t % in.jw) * in.cpd + t / in.jw : 0;
    const int n = t < P ? out.sink_count[s] : 0;
    const uint64_t nsrc = n > 0 ? sink_pencil_sources(in, s / in.cpd, s % in.cpd) : 0;
    int bk[3][3], lo[3][3];  // per tile: bucket and local offset of the full, one-wave and tail runs
    for (int ti = 0; ti < 3; ti++) {
        const ItemSplit sp(n, 1 << ti, nsrc, in.split_cost[ti]);
        const bool on = out.items[ti] != nullptr && n > 0;
        bk[ti][0] = cost_bucket((uint64_t)sp.tile * nsrc);
        bk[ti][1] = cost_bucket((uint64_t)kB * nsrc);
        bk[ti][2] = cost_bucket((uint64_t)sp.s * nsrc);
        int32_t *h = lc + ti * kPlanBuckets;
        lo[ti][0] = on && sp.full ? atomicAdd(h + bk[ti][0], sp.full) : 0;
        lo[ti][1] = on && sp.n2 ? atomicAdd(h + bk[ti][1], sp.n2) : 0;
        lo[ti][2] = on && sp.tail ? atomicAdd(h + 3 * kPlanBuckets + bk[ti][2], 1) : 0;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kPlanLists * kPlanBuckets; i += kT)
        if (lc[i]) lc[i] = atomicAdd(out.cursor + i, lc[i]);
    __syncthreads();
    if (n == 0) return;
    const int sic4 = 4 * out.sic;
    for (int ti = 0; ti < 3; ti++) {
        WorkItem *items = out.items[ti];
        if (!items) continue;
        const ItemSplit sp(n, 1 << ti, nsrc, in.split_cost[ti]);
        const int32_t *base = lc + ti * kPlanBuckets, *tbase = lc + (3 + ti) * kPlanBuckets;
        const int64_t M = mt[ti], T = mt[3 + ti];
        if (sp.full) {
            const int p = boff[ti][bk[ti][0]] + base[bk[ti][0]] + lo[ti][0];
            for (int i = 0; i < sp.full; i++)
                items[interleave_pos(false, p + i, M, T)] = WorkItem{s, i * sp.tile, sp.tile, 0 + sic4};
        }
        const int f = sp.full * sp.tile;
        if (sp.n2) {
            const int p = boff[ti][bk[ti][1]] + base[bk[ti][1]] + lo[ti][1];
            for (int i = 0; i < sp.n2; i++)
                items[interleave_pos(false, p + i, M, T)] = WorkItem{s, f + i * kB, i < sp.q ? kB : sp.s, 2 + sic4};
        }
        if (sp.tail) {
            const int p = boff[3 + ti][bk[ti][2]] + tbase[bk[ti][2]] + lo[ti][2];
            items[interleave_pos(true, p, M, T)] = WorkItem{s, f + sp.q * kB, sp.s, 1 + sic4};
        }
    }
}

// Same workgroup layout as k_plan_b1.
__device__ void plan_b2(const GPlanIn &in, const GPlanOut &out, int bid) {
    if (out.group && out.group[kPlanGroupBad]) return;
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const int GP = ngroups(P);
    if (bid < GP) {
        // Bucket offsets of the wanted work lists' runs of tiles and of
        // short tails: exclusive scans of the histograms in LDS (2 buckets
        // per thread), and the runs' lengths.
        __shared__ int32_t boff[kPlanLists][kPlanBuckets];
        __shared__ int32_t mt[kPlanLists];
        for (int li = 0; li < kPlanLists; li++) {
            if (!out.items[li % 3]) continue;
            const int32_t *h = out.hist + li * kPlanBuckets;
            const int a = h[2 * threadIdx.x], b = h[2 * threadIdx.x + 1];
            int total;
            const int e = block_excl_scan(a + b, &total);
            boff[li][2 * threadIdx.x] = e;
            boff[li][2 * threadIdx.x + 1] = e + a;
            if (threadIdx.x == 0) mt[li] = total;
        }
        const int s = bid * kT + threadIdx.x;
        const int n = s < P ? out.sink_count[s] : 0;
        const int n0 = s < P ? out.sink_plan[s].cell[0].N : 0;
        int tb, tc;
        const int b0 = block_sum_prefix(out.part, bid) + block_excl_scan(nblocks(n), &tb);
        const int c0 = block_sum_prefix(out.part + GP, bid) + block_excl_scan(nblocks(n0), &tc);
        place_items(in, out, boff, mt, bid);
        if (s >= P) return;
        const int j = s / in.cpd, k = s % in.cpd;
        out.sink_start[s] = b0 * kB;
        out.idmax[s] = (uint32_t)(b0 * kB + n);
        for (int q = 0, nb = nblocks(n); q < nb; q++) out.parent[b0 + q] = s;
        const float cellsize = (float)(1.0 / (double)in.cpd);
        for (int y = 0; y < in.W; y++) {
            out.ilist[in.W * s + y] = (j + y) * in.cpd + k;
            out.yoff[in.W * s + y] = (float)(y - in.R) * cellsize;
        }
        for (int f = 0, q = c0; f < n0; f += kB, q++) {
            out.chunks[2 * q] = s;
            out.chunks[2 * q + 1] = f;
        }
    } else {
        const int g = bid - GP;
        const int u = g * kT + threadIdx.x;
        const int n = u < Q ? out.src_count[u] : 0;
        int tb;
        const int b0 = block_sum_prefix(out.part + 2 * GP, g) + block_excl_scan(nblocks(n), &tb);
        if (u >= Q) return;
        out.src_start[u] = b0 * kB;
        for (int q = 0, nb = nblocks(n); q < nb; q++) out.src_parent[b0 + q] = u;
    }
}
__global__ __launch_bounds__(kT) void k_plan_b2(GPlanIn in, GPlanOut out) { plan_b2(in, out, blockIdx.x); }

// ---- A launch group in one launch per phase --------------------------------
// The SICs' inputs and outputs are arrays in device memory; first[s] is SIC
// s's first workgroup in the phase's grid (first[n]: the grid size).  Each
// workgroup finds its SIC by bisection and runs that SIC's workgroup
// bid = blockIdx.x - first[s], exactly as the one-SIC launch would.

__device__ __forceinline__ int group_sic(const int32_t *first, int n, int b) {
    int lo = 0, hi = n - 1;  // the last s with first[s] <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Zero the group's n totals (256 bytes apart) and its shared area.
__global__ void k_plan_zero_group(char *tot, int n, int32_t *shared, int nshared) {
    constexpr int m = sizeof(GPlanTotals) / sizeof(uint64_t);
    for (int i = threadIdx.x; i < n * m; i += blockDim.x)
        reinterpret_cast<uint64_t *>(tot + 256 * (size_t)(i / m))[i % m] = 0;
    for (int i = threadIdx.x; i < nshared; i += blockDim.x) shared[i] = 0;
}

__global__ __launch_bounds__(kT) void k_plan_a_group(const GPlanIn *__restrict__ ins, char *tot, int32_t *group,
                                                     const int32_t *__restrict__ first, int n) {
    const int s = group_sic(first, n, (int)blockIdx.x);
    plan_a(ins[s], reinterpret_cast<GPlanTotals *>(tot + 256 * (size_t)s), group, (int)blockIdx.x - first[s]);
}
__global__ __launch_bounds__(kT) void k_plan_b1_group(const GPlanIn *__restrict__ ins, const GPlanOut *__restrict__ outs,
                                                      const int32_t *__restrict__ first, int n) {
    const int s = group_sic(first, n, (int)blockIdx.x);
    plan_b1(ins[s], outs[s], (int)blockIdx.x - first[s]);
}
__global__ __launch_bounds__(kT) void k_plan_b2_group(const GPlanIn *__restrict__ ins, const GPlanOut *__restrict__ outs,
                                                      const int32_t *__restrict__ first, int n) {
    const int s = group_sic(first, n, (int)blockIdx.x);
    plan_b2(ins[s], outs[s], (int)blockIdx.x - first[s]);
}

}  // namespace

size_t gpu_plan_part_ints(int nsink_sets, int nsrc_sets) {
    return 2 * (size_t)ngroups(nsink_sets) + (size_t)ngroups(nsrc_sets);
}

int launch_gpu_plan_a(const GPlanIn &in, GPlanTotals *tot, int32_t *zero, int32_t *group, void *stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    hipLaunchKernelGGL(k_plan_zero, dim3(1), dim3(kT), 0, st, tot, zero, zero ? kPlanSharedInts : 0);
    hipLaunchKernelGGL(k_plan_a, dim3(ngroups(P + Q)), dim3(kT), 0, st, in, tot, group);
    return (int)hipGetLastError();
}

int launch_gpu_plan_b1(const GPlanIn &in, const GPlanOut &out, void *stream) {
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    hipLaunchKernelGGL(k_plan_b1, dim3(ngroups(P) + ngroups(Q)), dim3(kT), 0, static_cast<hipStream_t>(stream), in,
                       out);
    return (int)hipGetLastError();
}

int launch_gpu_plan_b2(const GPlanIn &in, const GPlanOut &out, void *stream) {
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    hipLaunchKernelGGL(k_plan_b2, dim3(ngroups(P) + ngroups(Q)), dim3(kT), 0, static_cast<hipStream_t>(stream), in,
                       out);
    return (int)hipGetLastError();
}

int gpu_plan_grid_a(const GPlanIn &in) { return ngroups((2 * in.jw + in.W - 1) * in.cpd); }
int gpu_plan_grid_b(const GPlanIn &in) { return ngroups(in.jw * in.cpd) + ngroups((in.jw + in.W - 1) * in.cpd); }

int launch_gpu_plan_group(int phase, const GPlanIn *ins, const GPlanOut *outs, GPlanTotals *tot, int32_t *shared,
                          const int32_t *first, int n, int grid, void *stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (n <= 0 || grid <= 0) return 0;
    char *t = reinterpret_cast<char *>(tot);
    switch (phase) {
        case 0:
            hipLaunchKernelGGL(k_plan_zero_group, dim3(1), dim3(kT), 0, st, t, n, shared, kPlanSharedInts);
            hipLaunchKernelGGL(k_plan_a_group, dim3(grid), dim3(kT), 0, st, ins, t, shared, first, n);
            break;
        case 1: hipLaunchKernelGGL(k_plan_b1_group, dim3(grid), dim3(kT), 0, st, ins, outs, first, n); break;
        case 2: hipLaunchKernelGGL(k_plan_b2_group, dim3(grid), dim3(kT), 0, st, ins, outs, first, n); break;
        default: return (int)hipErrorInvalidValue;
    }
    return (int)hipGetLastError();
}

}  // namespace nf
