// Device-side SIC planning: the SetInteractionCollection constructor
// (PencilOnPencil.cc:180-418, non-PARALLEL layout) computed on the GPU from
// the W slabs' cellinfo, plus the runtime's derived
// lists (the source blocks' parent pencils, the co-add chunks, the batched
// work list).  Every array has exactly the contents the host planner
// (planner.cpp) and the runtime's host code give it, except the work list,
// whose order within a cost bucket may differ (each sink belongs to exactly
// one item, so the results do not depend on the order).
//
// Phase A (launch_gpu_plan_a) accumulates the sizes the host needs to lay
// the plan out; phase B (launch_gpu_plan_b) writes the arrays by prefix sums
// over the pencils; launch_gpu_plan_items writes one work list.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "nf_internal.h"

namespace nf {
namespace {

constexpr int kB = NF_BLOCK_SIZE;
constexpr int kScanThreads = 1024;

__device__ __forceinline__ int wrapi(int i, int n) { return i >= n ? i - n : (i < 0 ? i + n : i); }
__device__ __forceinline__ int nblocks(int n) { return (n + kB - 1) / kB; }

// Scratch regions (int32 offsets) for P sink pencils and Q source pencils.
struct ScratchMap {
    int sblk, sscan;  // padded blocks per sink pencil, and their exclusive scan
    int qblk, qscan;  // the same per source pencil
    int cblk, cscan;  // co-add chunks per sink pencil, and their scan
    int hist;         // work-list bucket histogram, offsets, cursors
};
__host__ __device__ inline ScratchMap scratch_map(int P, int Q) {
    return ScratchMap{0, P, 2 * P, 2 * P + Q, 2 * P + 2 * Q, 3 * P + 2 * Q, 4 * P + 2 * Q};
}

// Sum over the wave, then one atomic per wave.
__device__ __forceinline__ void wave_add(uint64_t *p, uint64_t v) {
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)v, off), hi = __shfl_xor((uint32_t)(v >> 32), off);
        v += ((uint64_t)hi << 32) | lo;
    }
    if ((threadIdx.x & 63) == 0 && v) atomicAdd((unsigned long long *)p, (unsigned long long)v);
}

// Whether a planned cell (start as the plan stores it) lies inside its slab.
__device__ __forceinline__ bool cell_ok(uint64_t start, int32_t n, uint64_t nslab) {
    return n >= 0 && (uint64_t)(uint32_t)start + (uint64_t)n <= nslab;
}

// Sinks of sink pencil (j, k): cells (j + j_low, k + c) of the central slab;
// *bad counts its cells outside the slab.
__device__ __forceinline__ int sink_pencil_count(const GPlanIn &in, int j, int k, uint64_t *bad) {
    const nf_cellinfo *row = in.ci[in.R] + (size_t)(j + in.j_low) * in.cpd;
    int n = 0;
    for (int c = 0; c < in.W; c++) {
        const nf_cellinfo &ci = row[wrapi(k + c, in.cpd)];
        n += ci.count;
        *bad += cell_ok(ci.startindex, ci.count, in.nslab[in.R]) ? 0 : 1;
    }
    return n;
}
// Row and column of source pencil (jj, k): y = jj + j_low - R (periodic),
// z = zcen(k) = k + R (periodic), one cell in each of the W slabs.
__device__ __forceinline__ int source_cell(const GPlanIn &in, int jj, int k) {
    return wrapi(jj + in.j_low - in.R, in.cpd) * in.cpd + wrapi(k + in.R, in.cpd);
}
__device__ __forceinline__ int source_pencil_count(const GPlanIn &in, int jj, int k, uint64_t *bad) {
    const int cell = source_cell(in, jj, k);
    int n = 0;
    for (int c = 0; c < in.W; c++) {
        const nf_cellinfo &ci = in.ci[c][cell];
        n += ci.count;
        *bad += cell_ok(ci.startindex_with_ghost, ci.count, in.nslab[c]) ? 0 : 1;
    }
    return n;
}

// Work items of a pencil of n sinks for register tile S (the runtime's host
// build_items): full tiles of 64 S, one-wave tiles while they keep at least
// kTailLanes lanes busy, then one short tail.  f(first, nsink, kind).
template <typename F>
__device__ __forceinline__ void for_items(int n, int S, F f) {
    const int tile = kB * S;
    int first = 0;
    for (; first + tile <= n; first += tile) f(first, tile, 0);
    while (n - first >= kTailLanes) {
        const int m = min(kB, n - first);
        f(first, m, 2);
        first += m;
    }
    if (first < n) f(first, n - first, 1);
}

// 1/8-octave cost buckets, longest first (cost 0 last).
__device__ __forceinline__ int cost_bucket(uint64_t cost) {
    if (cost == 0) return kPlanBuckets - 1;
    const int b = (int)floor(8.0 * log2((double)cost));
    return kPlanBuckets - 1 - min(max(b, 0), kPlanBuckets - 1);
}

// ---- Phase A ---------------------------------------------------------------

__global__ void k_zero_totals(GPlanTotals *tot) {
    uint64_t *p = reinterpret_cast<uint64_t *>(tot);
    constexpr int n = sizeof(GPlanTotals) / sizeof(uint64_t);
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// Threads [0, P): sink pencils; [P, P+Q): source pencils.  Counts go to the
// sblk / qblk regions (phase B overwrites them), totals by atomics.
__global__ __launch_bounds__(256) void k_plan_a_counts(GPlanIn in, GPlanTotals *tot, int32_t *scratch) {
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const ScratchMap m = scratch_map(P, Q);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t sink = 0, sblk = 0, ch = 0, it[3] = {0, 0, 0}, occ[4] = {0, 0, 0, 0}, src = 0, qblk = 0, bad = 0;
    if (t < P) {
        const int j = t / in.cpd, k = t % in.cpd;
        const int n = sink_pencil_count(in, j, k, &bad);
        scratch[m.sblk + t] = n;
        sink = (uint64_t)n;
        sblk = (uint64_t)nblocks(n);
        // Pencil k starts with cell z = k: its chunks are that cell's.
        ch = (uint64_t)nblocks(in.ci[in.R][(size_t)(j + in.j_low) * in.cpd + k].count);
        for_items(n, 1, [&](int, int, int) { it[0]++; });
        for_items(n, 2, [&](int, int, int) { it[1]++; });
        for_items(n, 4, [&](int, int, int) { it[2]++; });
        const int bin = n < 64 ? 0 : n < 256 ? 1 : n < 1024 ? 2 : 3;
        for (int i = 0; i < 4; i++) occ[i] = bin == i ? (uint64_t)n : 0;
    } else if (t < P + Q) {
        const int u = t - P;
        const int n = source_pencil_count(in, u / in.cpd, u % in.cpd, &bad);
        scratch[m.qblk + u] = n;
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
}

// DirectTotal and PaddedDirectTotal over each sink pencil's W source pencils
// (j + y, k), y = 0..W-1.
__global__ __launch_bounds__(256) void k_plan_a_direct(GPlanIn in, GPlanTotals *tot, const int32_t *scratch) {
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const ScratchMap m = scratch_map(P, Q);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t d = 0, pd = 0;
    if (t < P) {
        const int j = t / in.cpd, k = t % in.cpd;
        const int ns = scratch[m.sblk + t];
        uint64_t nsrc = 0;
        for (int y = 0; y < in.W; y++) nsrc += (uint64_t)scratch[m.qblk + (j + y) * in.cpd + k];
        d = (uint64_t)ns * nsrc;
        pd = (uint64_t)nblocks(ns) * kB * nsrc;
    }
    wave_add(&tot->direct_total, d);
    wave_add(&tot->padded_direct_total, pd);
}

// ---- Phase B ---------------------------------------------------------------

// Exclusive prefix sum of n int32 (n of a few 10^5 at most) by one workgroup:
// each thread sums a contiguous run, a scan over the run sums, each thread
// writes its run.
__global__ __launch_bounds__(kScanThreads) void k_scan_excl(const int32_t *in, int32_t *out, int n) {
    __shared__ int64_t part[kScanThreads];
    const int t = threadIdx.x;
    const int per = (n + kScanThreads - 1) / kScanThreads;
    const int b = min(n, t * per), e = min(n, b + per);
    int64_t s = 0;
    for (int i = b; i < e; i++) s += in[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {  // inclusive scan of the run sums
        const int64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = part[t] - s;
    for (int i = b; i < e; i++) {
        const int32_t x = in[i];
        out[i] = (int32_t)run;
        run += x;
    }
}

// Sink pencils: SinkPlan (SinkPencilPlan::load, PencilPlan.cc:95-138),
// SinkSetCount, padded block and co-add chunk counts.
__global__ __launch_bounds__(256) void k_plan_b_sinks(GPlanIn in, GPlanOut out) {
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const ScratchMap m = scratch_map(P, Q);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P) return;
    const int j = s / in.cpd, k = s % in.cpd;
    const size_t row = (size_t)(j + in.j_low) * in.cpd;
    nf_sink_plan pl;
    int n = 0;
    for (int c = 0; c < NF_MAX_WIDTH; c++) {
        if (c < in.W) {
            const nf_cellinfo &ci = in.ci[in.R][row + wrapi(k + c, in.cpd)];
            pl.cell[c].start = (uint32_t)ci.startindex;
            pl.cell[c].N = ci.count;
            n += pl.cell[c].N;
        } else {
            pl.cell[c].start = 0;
            pl.cell[c].N = 0;
        }
    }
    // The cell that sits at z = 0 after the periodic wrap (PencilPlan.cc:123-128):
    // k + c = cpd, if that is within the pencil.
    const int wc = (in.cpd - k) % in.cpd;
    pl.wrapcell = wc < in.W ? wc : 0;
    out.sink_plan[s] = pl;
    out.sink_count[s] = n;
    out.scratch[m.sblk + s] = nblocks(n);
    out.scratch[m.cblk + s] = nblocks(pl.cell[0].N);
}

// Source pencils: SourcePlan (SourcePencilPlan::load, PencilPlan.cc:260-279),
// SourceSetCount, padded block counts.
__global__ __launch_bounds__(256) void k_plan_b_sources(GPlanIn in, GPlanOut out) {
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const ScratchMap m = scratch_map(P, Q);
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= Q) return;
    const int cell = source_cell(in, u / in.cpd, u % in.cpd);
    nf_source_plan pl;
    int n = 0;
    for (int c = 0; c < NF_MAX_WIDTH; c++) {
        if (c < in.W) {
            const nf_cellinfo &ci = in.ci[c][cell];
            pl.cell[c].start = (uint32_t)ci.startindex_with_ghost;
            pl.cell[c].N = ci.count;
            n += pl.cell[c].N;
        } else {
            pl.cell[c].start = 0;
            pl.cell[c].N = 0;
        }
    }
    out.src_plan[u] = pl;
    out.src_count[u] = n;
    out.scratch[m.qblk + u] = nblocks(n);
}

// After the scans.  Sink pencils: SinkSetStart, SinkSetIdMax, the block
// parents, the interaction list and y offsets (PencilOnPencil.cc:371-395),
// the co-add chunks.
__global__ __launch_bounds__(256) void k_plan_b_sinks2(GPlanIn in, GPlanOut out) {
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const ScratchMap m = scratch_map(P, Q);
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P) return;
    const int j = s / in.cpd, k = s % in.cpd;
    const int n = out.sink_count[s];
    const int b0 = out.scratch[m.sscan + s];
    out.sink_start[s] = b0 * kB;
    out.idmax[s] = (uint32_t)(b0 * kB + n);
    for (int q = 0, nb = nblocks(n); q < nb; q++) out.parent[b0 + q] = s;
    const float cellsize = (float)(1.0 / (double)in.cpd);
    for (int y = 0; y < in.W; y++) {
        out.ilist[in.W * s + y] = (j + y) * in.cpd + k;
        out.yoff[in.W * s + y] = (float)(y - in.R) * cellsize;
    }
    const int n0 = out.sink_plan[s].cell[0].N;
    for (int f = 0, q = out.scratch[m.cscan + s]; f < n0; f += kB, q++) {
        out.chunks[2 * q] = s;
        out.chunks[2 * q + 1] = f;
    }
}

// Source pencils: SourceSetStart and the blocks' parent pencils.
__global__ __launch_bounds__(256) void k_plan_b_sources2(GPlanIn in, GPlanOut out) {
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const ScratchMap m = scratch_map(P, Q);
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= Q) return;
    const int b0 = out.scratch[m.qscan + u];
    out.src_start[u] = b0 * kB;
    for (int q = 0, nb = nblocks(out.src_count[u]); q < nb; q++) out.src_parent[b0 + q] = u;
}

// ---- Work list -------------------------------------------------------------

// Thread t takes sink pencil (j, k) = (t % jw, t / jw): neighbours in y,
// which share 2R of their 2R+1 source pencils, sit on adjacent lanes and
// land close together in a bucket.
__device__ __forceinline__ int item_pencil(const GPlanIn &in, int t) { return (t % in.jw) * in.cpd + t / in.jw; }

__device__ __forceinline__ uint64_t pencil_sources(const GPlanIn &in, const GPlanOut &out, int s) {
    const int j = s / in.cpd, k = s % in.cpd;
    uint64_t nsrc = 0;
    for (int y = 0; y < in.W; y++) nsrc += (uint64_t)out.src_count[(j + y) * in.cpd + k];
    return nsrc;
}

__global__ __launch_bounds__(256) void k_items_hist(GPlanIn in, GPlanOut out, int S, int32_t *hist) {
    const int P = in.jw * in.cpd;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    const int s = item_pencil(in, t);
    const uint64_t nsrc = pencil_sources(in, out, s);
    for_items(out.sink_count[s], S, [&](int, int m, int kind) {
        atomicAdd(hist + cost_bucket((uint64_t)(kind == 2 ? kB : m) * nsrc), 1);
    });
}

__global__ __launch_bounds__(256) void k_items_place(GPlanIn in, GPlanOut out, int S, WorkItem *items,
                                                     const int32_t *offset, int32_t *cursor) {
    const int P = in.jw * in.cpd;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= P) return;
    const int s = item_pencil(in, t);
    const uint64_t nsrc = pencil_sources(in, out, s);
    for_items(out.sink_count[s], S, [&](int first, int m, int kind) {
        const int b = cost_bucket((uint64_t)(kind == 2 ? kB : m) * nsrc);
        items[offset[b] + atomicAdd(cursor + b, 1)] = WorkItem{s, first, m, kind + 4 * out.sic};
    });
}

__global__ __launch_bounds__(256) void k_zero_i32(int32_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

inline dim3 grid_for(int n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

size_t gpu_plan_scratch_ints(int nsink_sets, int nsrc_sets) {
    return (size_t)scratch_map(nsink_sets, nsrc_sets).hist + 3 * (size_t)kPlanBuckets;
}

int launch_gpu_plan_a(const GPlanIn &in, GPlanTotals *tot, int32_t *scratch, void *stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    hipLaunchKernelGGL(k_zero_totals, dim3(1), dim3(64), 0, st, tot);
    hipLaunchKernelGGL(k_plan_a_counts, grid_for(P + Q), dim3(256), 0, st, in, tot, scratch);
    hipLaunchKernelGGL(k_plan_a_direct, grid_for(P), dim3(256), 0, st, in, tot, scratch);
    return (int)hipGetLastError();
}

int launch_gpu_plan_b(const GPlanIn &in, const GPlanOut &out, void *stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    const ScratchMap m = scratch_map(P, Q);
    int32_t *sc = out.scratch;
    hipLaunchKernelGGL(k_plan_b_sinks, grid_for(P), dim3(256), 0, st, in, out);
    hipLaunchKernelGGL(k_plan_b_sources, grid_for(Q), dim3(256), 0, st, in, out);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(kScanThreads), 0, st, sc + m.sblk, sc + m.sscan, P);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(kScanThreads), 0, st, sc + m.qblk, sc + m.qscan, Q);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(kScanThreads), 0, st, sc + m.cblk, sc + m.cscan, P);
    hipLaunchKernelGGL(k_plan_b_sinks2, grid_for(P), dim3(256), 0, st, in, out);
    hipLaunchKernelGGL(k_plan_b_sources2, grid_for(Q), dim3(256), 0, st, in, out);
    return (int)hipGetLastError();
}

int launch_gpu_plan_items(const GPlanIn &in, const GPlanOut &out, int S, WorkItem *items, void *stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int P = in.jw * in.cpd, Q = (in.jw + in.W - 1) * in.cpd;
    int32_t *hist = out.scratch + scratch_map(P, Q).hist, *offset = hist + kPlanBuckets,
            *cursor = offset + kPlanBuckets;
    hipLaunchKernelGGL(k_zero_i32, grid_for(3 * kPlanBuckets), dim3(256), 0, st, hist, 3 * kPlanBuckets);
    hipLaunchKernelGGL(k_items_hist, grid_for(P), dim3(256), 0, st, in, out, S, hist);
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(kScanThreads), 0, st, hist, offset, kPlanBuckets);
    hipLaunchKernelGGL(k_items_place, grid_for(P), dim3(256), 0, st, in, out, S, items, offset, cursor);
    return (int)hipGetLastError();
}

}  // namespace nf
