// Device kernels of the near-field direct path, hand-written for gfx950.
//
//   k_gather_sinks / k_gather_sources   pencil gathers from device-resident
//       PosXYZ slabs into padded float4 pencil streams (the device-side form of
//       SinkPencilPlan/SourcePencilPlan::copy_into_pinned_memory,
//       PencilPlan.cc:22-90, 214-255), plus a bounding box per 64-source block.
//   k_gather_sources_binned  the batched path's source stream: each cell's
//       particles in x-bin order, FP64 cell offsets, block boxes.
//   k_plain      one 64-sink block per workgroup, one launch per SIC: the
//       structure of ComputeDirects (PencilKernel.cu:97-189) with the
//       reference's pair forms (DirectGPUKernels.cuh) and BLOCK_ACCUM tiles.
//   k_coadd_plain  FP32 co-add of the W partials per sink in the reference's
//       k order (PencilPlan.cc:140-207 as driven by PencilTask.cu:237-247).
//   k_batched    persistent kernel draining a work queue sorted by cost;
//       FP32 tile sums accumulated into FP64 per-pencil partials.
//   k_coadd_f64  FP64 reduction of the W partials per sink.
//
// Pair kernel, all forms: d = source - sink, a -= g(r) d, w += max(0, b2 - r^2).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "nf_internal.h"

namespace nf {
namespace {

constexpr int kWave = 64;
constexpr float kFar = 1.0e30f;  // x of a padding source: its pair terms are exact zeros

// Read-only data of a launch is addressed through the constant address space:
// the backend then loads wave-uniform addresses with scalar loads (s_load)
// and divergent ones with global loads, even when the pointer itself came
// from memory (the SIC view table) rather than from a kernel argument.
// (The qualifiers exist only in the device compilation pass of this file.)
#if defined(__HIP_DEVICE_COMPILE__)
#define NF_CONST __attribute__((address_space(4)))
#define NF_GLOBAL __attribute__((address_space(1)))
#else
#define NF_CONST
#define NF_GLOBAL
#endif
template <typename T>
__device__ __forceinline__ const NF_CONST T *cptr(const void *p) {
    return (const NF_CONST T *)(p);
}

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float clamp01(float x) { return __builtin_amdgcn_fmed3f(x, 0.0f, 1.0f); }

// A SIC's counts: from device memory when it was planned on the device
// without a host round trip (its host-side fields are then bounds), else
// the host's.  plan_bad: its launch group holds a cell outside its slab.
__device__ __forceinline__ bool plan_bad(const DevPlan &pl) { return pl.dbad && *cptr<int32_t>(pl.dbad) != 0; }
__device__ __forceinline__ int sink_blocks(const DevPlan &pl) {
    return pl.dtot ? (int)cptr<GPlanTotals>(pl.dtot)->sink_blocks : pl.nsink_blocks;
}
__device__ __forceinline__ int src_blocks(const DevPlan &pl) {
    return pl.dtot ? (int)cptr<GPlanTotals>(pl.dtot)->source_blocks : pl.nsrc_blocks;
}
__device__ __forceinline__ int coadd_chunks(const DevPlan &pl) {
    return pl.dtot ? (int)cptr<GPlanTotals>(pl.dtot)->chunks : pl.n_coadd_chunks;
}

// ---------------------------------------------------------------------------
// Gathers
// ---------------------------------------------------------------------------

// Position of padded index i within a pencil: cell c with pre[c] <= i < pre[c+1].
__device__ __forceinline__ int cell_of(int i, const int *pre, int W) {
    int c = 0;
    while (c < W - 1 && i >= pre[c + 1]) c++;
    return c;
}

// One wave per padded 64-sink block (the reference's SinkBlockParentPencil
// names its pencil), so a heavy pencil spreads over many waves.
// Position of particle idx of a slab of n particles: PosXYZ (X|Y|Z) or, with
// aos, posstruct (x,y,z interleaved).
__device__ __forceinline__ float3 slab_xyz(const float *P, uint64_t n, uint64_t idx, int aos) {
    if (aos) return make_float3(P[3 * idx], P[3 * idx + 1], P[3 * idx + 2]);
    return make_float3(P[idx], P[n + idx], P[2 * n + idx]);
}

// Padded blocks per gather wave.  A block's chain of dependent loads (its
// parent pencil, the pencil's plan, then the positions) is short on work;
// a wave that walks four blocks with their loads interleaved keeps four
// chains in flight.
constexpr int kGatherBlocks = 4;

__global__ __launch_bounds__(256) void k_gather_sinks(DevPlan pl, const float *__restrict__ slab, uint64_t n,
                                                      int aos, float4 *__restrict__ out, float cellsize) {
    const int b0 = (blockIdx.x * 4 + (int)(threadIdx.x / kWave)) * kGatherBlocks;
    const int nsb = sink_blocks(pl);
    if (b0 >= nsb || plan_bad(pl)) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int W = pl.W;
    int par[kGatherBlocks];
#pragma unroll
    for (int u = 0; u < kGatherBlocks; u++) par[u] = pl.parent[min(b0 + u, nsb - 1)];
    uint64_t idx[kGatherBlocks];
    int dcell[kGatherBlocks];
    bool live[kGatherBlocks];
#pragma unroll
    for (int u = 0; u < kGatherBlocks; u++) {
        const int s = par[u];
        const nf_sink_plan p = pl.sink_plan[s];
        int pre[NF_MAX_WIDTH + 1];
        pre[0] = 0;
        for (int c = 0; c < W; c++) pre[c + 1] = pre[c] + p.cell[c].N;
        const int i = (b0 + u) * kWave - pl.sink_set_start[s] + lane;  // index within the pencil
        live[u] = b0 + u < nsb && i < pre[W];
        const int c = live[u] ? cell_of(i, pre, W) : 0;
        idx[u] = live[u] ? (uint64_t)p.cell[c].start + (uint64_t)(i - pre[c]) : 0;
        dcell[u] = c - pl.R;
    }
    float3 q[kGatherBlocks];
#pragma unroll
    for (int u = 0; u < kGatherBlocks; u++)
        q[u] = live[u] ? slab_xyz(slab, n, idx[u], aos) : make_float3(0.f, 0.f, 0.f);
#pragma unroll
    for (int u = 0; u < kGatherBlocks; u++) {
        if (b0 + u >= nsb) break;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (live[u]) {
            float z = q[u].z;
            const float dz = (float)dcell[u] * cellsize;  // PencilPlan.cc:43-50: FLOAT, added if nonzero
            if (dz != 0.f) z = z + dz;
            v = make_float4(q[u].x, q[u].y, z, 0.f);
        }
        out[(int64_t)(b0 + u) * kWave + lane] = v;
    }
}

__device__ __forceinline__ const float *slab_ptr(const DevSlabs &sl, int c) {
    return c == 0 ? sl.pos[0] : c == 1 ? sl.pos[1] : c == 2 ? sl.pos[2] : c == 3 ? sl.pos[3] : sl.pos[4];
}
__device__ __forceinline__ uint64_t slab_n(const DevSlabs &sl, int c) {
    return c == 0 ? sl.n[0] : c == 1 ? sl.n[1] : c == 2 ? sl.n[2] : c == 3 ? sl.n[3] : sl.n[4];
}

// Wave-wide min / max (every lane of the wave active), wave-uniform result:
// four DPP steps leave each 16-lane row's extremum in all its lanes
// (quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, row_mirror), then
// the four rows' values are read out.  No LDS round trips, unlike a
// __shfl_xor butterfly (ds_bpermute per step).
template <bool MAX, int CTRL>
__device__ __forceinline__ float dpp_step(float v) {
    const int x = __builtin_bit_cast(int, v);
    const float u = __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(x, x, CTRL, 0xF, 0xF, false));
    return MAX ? fmaxf(v, u) : fminf(v, u);
}
template <bool MAX>
__device__ __forceinline__ float wave_reduce(float v) {
    v = dpp_step<MAX, 0xB1>(v);   // quad_perm [1,0,3,2]
    v = dpp_step<MAX, 0x4E>(v);   // quad_perm [2,3,0,1]
    v = dpp_step<MAX, 0x141>(v);  // row_half_mirror
    v = dpp_step<MAX, 0x140>(v);  // row_mirror
    const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0));
    const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 16));
    const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32));
    const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 48));
    return MAX ? fmaxf(fmaxf(r0, r1), fmaxf(r2, r3)) : fminf(fminf(r0, r1), fminf(r2, r3));
}
__device__ __forceinline__ float wave_min(float v) { return wave_reduce<false>(v); }
__device__ __forceinline__ float wave_max(float v) { return wave_reduce<true>(v); }

// Plain-path source stream, in the reference's order.  One wave per
// kGatherBlocks padded 64-source blocks; it writes each block and its
// bounding box (padding excluded): box[2b] = min, box[2b+1] = max.
__global__ __launch_bounds__(256) void k_gather_sources(DevPlan pl, DevSlabs sl, float4 *__restrict__ out,
                                                        float4 *__restrict__ box, float cellsize) {
    const int b0 = (blockIdx.x * 4 + (int)(threadIdx.x / kWave)) * kGatherBlocks;
    const int nqb = src_blocks(pl);
    if (b0 >= nqb || plan_bad(pl)) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int W = pl.W;
    int par[kGatherBlocks];
#pragma unroll
    for (int u = 0; u < kGatherBlocks; u++) par[u] = pl.src_parent[min(b0 + u, nqb - 1)];
    uint64_t idx[kGatherBlocks];
    int cc[kGatherBlocks];
    bool live[kGatherBlocks];
#pragma unroll
    for (int u = 0; u < kGatherBlocks; u++) {
        const int s = max(par[u], 0);  // a block no pencil covers (s < 0) is never read: left unwritten
        const nf_source_plan p = pl.src_plan[s];
        int pre[NF_MAX_WIDTH + 1];
        pre[0] = 0;
        for (int c = 0; c < W; c++) pre[c + 1] = pre[c] + p.cell[c].N;
        const int i = (b0 + u) * kWave - pl.src_set_start[s] + lane;
        live[u] = b0 + u < nqb && par[u] >= 0 && i < pre[W];
        const int c = live[u] ? cell_of(i, pre, W) : 0;
        idx[u] = live[u] ? (uint64_t)p.cell[c].start + (uint64_t)(i - pre[c]) : 0;
        cc[u] = c;
    }
    float3 q[kGatherBlocks];
#pragma unroll
    for (int u = 0; u < kGatherBlocks; u++)
        q[u] = live[u] ? slab_xyz(slab_ptr(sl, cc[u]), slab_n(sl, cc[u]), idx[u], sl.aos)
                       : make_float3(0.f, 0.f, 0.f);
    const float inf = __builtin_huge_valf();
#pragma unroll
    for (int u = 0; u < kGatherBlocks; u++) {
        const int b = b0 + u;
        if (b >= nqb) break;
        if (par[u] < 0) continue;
        float4 v = make_float4(kFar, 0.f, 0.f, 0.f);
        const bool real = live[u];
        if (real) {
            float x = q[u].x;
            const float dx = (float)(cc[u] - pl.R) * cellsize;  // PencilPlan.cc:226-238
            if (dx != 0.f) x = x + dx;
            v = make_float4(x, q[u].y, q[u].z, 0.f);
        }
        out[(int64_t)b * kWave + lane] = v;
        const float mnx = wave_min(real ? v.x : inf), mny = wave_min(real ? v.y : inf),
                    mnz = wave_min(real ? v.z : inf);
        const float mxx = wave_max(real ? v.x : -inf), mxy = wave_max(real ? v.y : -inf),
                    mxz = wave_max(real ? v.z : -inf);
        if (lane == 0) {
            box[2 * (int64_t)b] = make_float4(mnx, mny, mnz, 0.f);
            box[2 * (int64_t)b + 1] = make_float4(mxx, mxy, mxz, 0.f);
        }
    }
}

// Batched-path source stream.  Each cell's particles are ordered by
// kXBins equal x-bins of the cell (a stable counting sort: within a bin
// the reference's order), so a 64-source block that straddles two cells
// takes an x-slice of each rather than all of one: tighter block boxes, so
// more sink-block/source-block pairs clear the softening and density radii
// and run the fast pair form.  Cells above kBinnedCellMax = 128 particles
// keep the reference's order: a replay of the mode split on a clustered box
// gained little from sorting them.
// Cell offsets are added in FP64 (x + dc/cpd, rounded once).
//
// A wave fills a range of a pencil's 64-source blocks: it loads each binned
// cell that reaches into the range once (two 64-particle chunks each, all
// cells' loads in flight together), ranks every particle within its cell
// (ballots per bin), writes the ones whose slots fall in the range, and
// reduces the box of each block of the range over the particles that land
// in it (the pencil's slots run cell after cell, so its blocks are met in
// order).  The last block's padding slots take kFar.
//
// One wave per source pencil fills the pencil's first kSegBlocks blocks;
// one wave per kSegBlocks blocks of the stream fills the blocks that lie
// beyond the first kSegBlocks of their pencil.  So every block is written
// by one wave, and a pencil of large cells (a clustered box's halos put
// thousands of particles in a cell) spreads over many waves instead of
// holding one for the whole launch.
constexpr int kXBins = 8;
constexpr int kBinnedCellMax = 2 * kWave;
constexpr int kSegBlocks = kGatherSegBlocks;

__device__ __forceinline__ int xbin(float x, float cpdf) {  // x cell-centred; out-of-cell x clamps
    const float t = floorf((x * cpdf + 0.5f) * (float)kXBins);
    return (int)fminf(fmaxf(t, 0.f), (float)(kXBins - 1));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ int lanes_below(uint64_t m) {  // set bits of m below this lane
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// The box of the block being filled: each lane's running min/max over its
// sources in the block, reduced across the wave once, when the next block
// starts or the pencil ends.
struct BlockBox {
    int open = -1;
    float mn[3], mx[3];
    __device__ __forceinline__ void reset() {
        mn[0] = mn[1] = mn[2] = __builtin_huge_valf();
        mx[0] = mx[1] = mx[2] = -__builtin_huge_valf();
    }
    // Make b the open block (b >= open; the pencil's blocks are met in order).
    __device__ __forceinline__ void to(int b, float4 *box, int64_t gb0, int lane) {
        if (b == open) return;
        flush(box, gb0, lane);
        open = b;
        reset();
    }
    __device__ __forceinline__ void add(bool in, float x, float y, float z) {
        if (!in) return;
        mn[0] = fminf(mn[0], x), mn[1] = fminf(mn[1], y), mn[2] = fminf(mn[2], z);
        mx[0] = fmaxf(mx[0], x), mx[1] = fmaxf(mx[1], y), mx[2] = fmaxf(mx[2], z);
    }
    __device__ __forceinline__ void flush(float4 *box, int64_t gb0, int lane) {
        if (open < 0) return;
        const float4 lo = make_float4(wave_min(mn[0]), wave_min(mn[1]), wave_min(mn[2]), 0.f);
        const float4 hi = make_float4(wave_max(mx[0]), wave_max(mx[1]), wave_max(mx[2]), 0.f);
        if (lane == 0) {
            box[2 * (gb0 + open)] = lo;
            box[2 * (gb0 + open) + 1] = hi;
        }
    }
};

// Blocks [blo, bhi) of pencil s (pencil-relative; bhi may pass its end);
// RANGED false: the whole pencil, without the range tests.
template <bool RANGED>
__device__ __forceinline__ void gather_pencil_binned(const DevPlan &pl, const DevSlabs &sl, float4 *__restrict__ out,
                                                     float4 *__restrict__ box, float cpdf, double invcpd, int s,
                                                     int lane, int blo, int bhi) {
    const int W = pl.W;
    const nf_source_plan p = cptr<nf_source_plan>(pl.src_plan)[s];
    int pre[NF_MAX_WIDTH + 1];
    pre[0] = 0;
#pragma unroll
    for (int c = 0; c < NF_MAX_WIDTH; c++) pre[c + 1] = pre[c] + (c < W ? p.cell[c].N : 0);
    const int total = pre[NF_MAX_WIDTH];
    // the slots to fill
    const int lo = RANGED ? blo * kWave : 0, hi = !RANGED || bhi >= (total + kWave - 1) / kWave ? total : bhi * kWave;
    if (lo >= hi) return;
    const int start = cptr<int32_t>(pl.src_set_start)[s];  // the pencil's first slot, a multiple of 64
    float4 *dst = out + start;
    const int64_t gb0 = start / kWave;
    const int last = (total - 1) / kWave;
    if ((!RANGED || (last >= blo && last < bhi)) && last * kWave + lane >= total)
        dst[last * kWave + lane] = make_float4(kFar, 0.f, 0.f, 0.f);
    auto overlaps = [&](int c) { return !RANGED || (pre[c] < hi && pre[c + 1] > lo); };
    // the binned cells' particles (two chunks each) all load before any is ranked
    float3 q[NF_MAX_WIDTH][2];
#pragma unroll
    for (int c = 0; c < NF_MAX_WIDTH; c++) {
        const int n = pre[c + 1] - pre[c];
        q[c][0] = q[c][1] = make_float3(0.f, 0.f, 0.f);
        if (n > 0 && n <= kBinnedCellMax && overlaps(c)) {
            const float *P = slab_ptr(sl, c);
            const uint64_t ns = slab_n(sl, c), first = p.cell[c].start;
            if (lane < n) q[c][0] = slab_xyz(P, ns, first + (uint64_t)lane, sl.aos);
            if (lane + kWave < n) q[c][1] = slab_xyz(P, ns, first + (uint64_t)(lane + kWave), sl.aos);
        }
    }
    BlockBox bb;
    // The sources at slots s0 (and s1, if in1) of this lane: into the stream
    // and into the boxes of the blocks from slot a to slot z, as far as they
    // lie in the range.
    auto place = [&](int a, int z, bool in0, int s0, float3 v0, bool in1, int s1, float3 v1) {
        if (RANGED) {
            in0 = in0 && s0 >= lo && s0 < hi;
            in1 = in1 && s1 >= lo && s1 < hi;
            a = max(a, lo);
            z = min(z, hi - 1);
        }
        if (in0) dst[s0] = make_float4(v0.x, v0.y, v0.z, 0.f);
        if (in1) dst[s1] = make_float4(v1.x, v1.y, v1.z, 0.f);
        for (int b = a / kWave; b <= z / kWave; b++) {
            bb.to(b, box, gb0, lane);
            bb.add(in0 && s0 / kWave == b, v0.x, v0.y, v0.z);
            bb.add(in1 && s1 / kWave == b, v1.x, v1.y, v1.z);
        }
    };
    auto shifted = [&](float3 v, int c) {
        return make_float3((float)((double)v.x + (double)(c - pl.R) * invcpd), v.y, v.z);
    };
#pragma unroll
    for (int c = 0; c < NF_MAX_WIDTH; c++) {
        const int n = pre[c + 1] - pre[c];
        if (n == 0 || !overlaps(c)) continue;
        if (n <= kBinnedCellMax) {
            const int bin0 = lane < n ? xbin(q[c][0].x, cpdf) : -1;
            const int bin1 = lane + kWave < n ? xbin(q[c][1].x, cpdf) : -1;
            int rank0 = 0, rank1 = 0, acc = 0;
#pragma unroll
            for (int k = 0; k < kXBins; k++) {  // within a bin, chunk 0 ranks before chunk 1
                const uint64_t m0 = ballot(bin0 == k);
                if (bin0 == k) rank0 = acc + lanes_below(m0);
                acc += __builtin_popcountll(m0);
                if (n > kWave) {
                    const uint64_t m1 = ballot(bin1 == k);
                    if (bin1 == k) rank1 = acc + lanes_below(m1);
                    acc += __builtin_popcountll(m1);
                }
            }
            place(pre[c], pre[c + 1] - 1, bin0 >= 0, pre[c] + rank0, shifted(q[c][0], c), bin1 >= 0,
                  pre[c] + rank1, shifted(q[c][1], c));
            continue;
        }
        // a large cell: the reference's order, 64 at a time (the chunks
        // that reach into the range)
        const float *P = slab_ptr(sl, c);
        const uint64_t ns = slab_n(sl, c), first = p.cell[c].start;
        for (int i0 = RANGED ? max(lo - pre[c], 0) & ~(kWave - 1) : 0; i0 < n && (!RANGED || pre[c] + i0 < hi);
             i0 += kWave) {
            const bool in = i0 + lane < n;
            const float3 v = in ? slab_xyz(P, ns, first + (uint64_t)(i0 + lane), sl.aos) : make_float3(0.f, 0.f, 0.f);
            place(pre[c] + i0, pre[c] + min(n, i0 + kWave) - 1, in, pre[c] + i0 + lane, shifted(v, c), false, 0,
                  v);
        }
    }
    bb.flush(box, gb0, lane);
}

// Wave w of a SIC's source gather: w < nsrc_sets fills pencil w (its first
// kSegBlocks blocks if the gather is segmented); wave nsrc_sets + j fills,
// of the stream's blocks [j kSegBlocks, (j + 1) kSegBlocks), those beyond
// their pencil's first kSegBlocks.  Grid: gather_waves() waves.  SEG false:
// no SIC of the launch is segmented, and the kernel holds only the
// whole-pencil code (the segmented code's scalar registers would cost the
// whole-pencil waves an eighth of their occupancy).
template <bool SEG>
__device__ __forceinline__ void gather_wave(const DevPlan &pl, const DevSlabs &sl, float4 *__restrict__ out,
                                            float4 *__restrict__ box, float cpdf, double invcpd, int w, int lane) {
    if (!SEG) {
        gather_pencil_binned<false>(pl, sl, out, box, cpdf, invcpd, w, lane, 0, 0);
        return;
    }
    if (w < pl.nsrc_sets) {
        if (gather_segmented(pl.max_src_blocks))
            gather_pencil_binned<true>(pl, sl, out, box, cpdf, invcpd, w, lane, 0, kSegBlocks);
        else
            gather_pencil_binned<false>(pl, sl, out, box, cpdf, invcpd, w, lane, 0, 0);
        return;
    }
    // Only the pencil of the segment's first block can reach kSegBlocks past
    // its own start within the segment: a pencil that starts later in it
    // ends its first kSegBlocks blocks beyond the segment's end.
    const int b0 = (w - pl.nsrc_sets) * kSegBlocks;
    if (b0 >= src_blocks(pl)) return;
    const int s = cptr<int32_t>(pl.src_parent)[b0];
    if (s < 0) return;  // a block of no pencil
    const int f = cptr<int32_t>(pl.src_set_start)[s] / kWave;
    const int e = f + (cptr<int32_t>(pl.src_set_count)[s] + kWave - 1) / kWave;
    const int lo = max(b0, f + kSegBlocks), hi = min(b0 + kSegBlocks, e);
    if (lo < hi) gather_pencil_binned<true>(pl, sl, out, box, cpdf, invcpd, s, lane, lo - f, hi - f);
}

template <bool SEG>
__global__ __launch_bounds__(256) void k_gather_sources_binned(DevPlan pl, DevSlabs sl, float4 *__restrict__ out,
                                                               float4 *__restrict__ box, float cpdf,
                                                               double invcpd) {
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (int)(threadIdx.x / kWave));
    if (w >= gather_waves(pl.nsrc_sets, pl.nsrc_blocks, pl.max_src_blocks) || plan_bad(pl)) return;
    gather_wave<SEG>(pl, sl, out, box, cpdf, invcpd, w, threadIdx.x & (kWave - 1));
}

// The SIC of a group-wide launch that wave w belongs to: the last i with
// first[i] <= w (first[] ascending, first[n] the total), by bisection.
__device__ __forceinline__ int group_sic_of(const int32_t *first, int n, int w) {
    const NF_CONST int32_t *f = cptr<int32_t>(first);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (f[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Every SIC's source gather of a launch group in one launch: a group of
// small SICs (a clustered box's) fills the GPU where one SIC's pencils do not.
template <bool SEG>
__global__ __launch_bounds__(256) void k_gather_sources_binned_group(const GroupSic *__restrict__ g,
                                                                     const int32_t *__restrict__ first, int n,
                                                                     float cpdf, double invcpd) {
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (int)(threadIdx.x / kWave));
    if (w >= cptr<int32_t>(first)[n]) return;
    const int i = group_sic_of(first, n, w);
    const DevPlan pl = cptr<GroupSic>(g)[i].pl;
    const DevSlabs sl = cptr<GroupSic>(g)[i].sl;
    void *const sources4 = cptr<GroupSic>(g)[i].sources4;
    const int s = w - cptr<int32_t>(first)[i];
    if (s >= gather_waves(pl.nsrc_sets, pl.nsrc_blocks, pl.max_src_blocks) || plan_bad(pl)) return;
    float4 *src = static_cast<float4 *>(sources4);
    gather_wave<SEG>(pl, sl, src, src + (size_t)pl.nsrc_blocks * NF_BLOCK_SIZE, cpdf, invcpd, s,
                     threadIdx.x & (kWave - 1));  // the block boxes follow the stream
}

// ---------------------------------------------------------------------------
// Plain path: the reference's structure and pair forms
// ---------------------------------------------------------------------------

// DirectGPUKernels.cuh direct_density<0>: single spline :68-108, Plummer
// :144-174, cubic spline :20-66, cubic Plummer :110-143.  rsqrt is the
// hardware v_rsq_f32 (the reference uses its own platform's fast rsqrtf).
template <int SOFT>
__device__ __forceinline__ void ref_pair(float sx, float sy, float sz, float kx, float ky, float kz,
                                         const PairConst &pc, float &ax, float &ay, float &az, float &aw) {
    const float dx = sx - kx, dy = sy - ky, dz = sz - kz;
    const float r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < pc.b2) aw = aw + pc.b2 - r2;
    float f;
    if constexpr (SOFT == NF_SOFT_SINGLE_SPLINE) {
        const float q = r2 * pc.eps + 1e-32f;
        f = rsq(q);
        f = (q > 1.0f) ? f * f * f : (-15.0f * f + 6.0f) * q + 10.0f;
    } else if constexpr (SOFT == NF_SOFT_PLUMMER) {
        const float r = rsq(r2 + pc.eps);
        f = r * r * r;
    } else if constexpr (SOFT == NF_SOFT_CUBIC_SPLINE) {
        const float rinv = rsq(r2);
        const float u = pc.eps * __builtin_amdgcn_sqrtf(r2);  // eps_inv / rinv
        if (u >= 2.0f) {
            f = rinv * rinv * rinv;
        } else if (u >= 1.0f) {
            const float u3 = u * u * u;
            f = rinv * rinv * rinv *
                ((float)(-1.0 / 15.0) + (float)(8.0 / 3.0) * u3 - 3.0f * u3 * u + (float)(6.0 / 5.0) * u3 * u * u -
                 (float)(1.0 / 6.0) * u3 * u3);
        } else {
            f = pc.eps * pc.eps * pc.eps * ((float)(4.0 / 3.0) - (float)(6.0 / 5.0) * u * u + 0.5f * u * u * u);
        }
    } else {  // cubic Plummer
        float r = r2 + 1e-16f;
        r *= r * rsq(r);
        r += pc.eps;
        f = __builtin_amdgcn_rcpf(r);
    }
    ax -= f * dx;
    ay -= f * dy;
    az -= f * dz;
}

template <int SOFT>
__global__ __launch_bounds__(64) void k_plain(DevPlan pl, PairConst pc, const float4 *__restrict__ sinks,
                                              const float4 *__restrict__ src, float4 *__restrict__ out) {
    const int blk = blockIdx.x;
    if (blk >= sink_blocks(pl) || plan_bad(pl)) return;
    const int s = pl.parent[blk];
    const int id = blk * kWave + threadIdx.x;
    const float4 k = sinks[id];  // padded ids hold zeros (PencilKernel.cu:109-117)
    float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
    const int W = pl.W;
    for (int c = s * W; c < s * W + W; c++) {
        const int si = pl.ilist[c];
        const float yo = pl.yoff[c];
        const int st = pl.src_set_start[si], n = pl.src_set_count[si];
        const float4 *sp = src + st;
        const int nB = n / kWave;
        for (int b = 0; b < nB; b++) {  // FullDirectTile with BLOCK_ACCUM
            float tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll 16
            for (int i = 0; i < kWave; i++) {
                const float4 q = sp[b * kWave + i];
                ref_pair<SOFT>(q.x, q.y + yo, q.z, k.x, k.y, k.z, pc, tx, ty, tz, aw);
            }
            ax += tx;
            ay += ty;
            az += tz;
        }
        for (int i = nB * kWave; i < n; i++) {  // PartialDirectTile
            const float4 q = sp[i];
            ref_pair<SOFT>(q.x, q.y + yo, q.z, k.x, k.y, k.z, pc, ax, ay, az, aw);
        }
    }
    if ((uint32_t)id < pl.sink_idmax[s]) out[id] = make_float4(ax, ay, az, aw);
}

// Addresses of the W partials of the sinks of cell (j_low + j, z): the cell is
// at position c of pencil k = z - c (mod cpd).  Returns the cell's slab start
// and count; base[c] is the padded slot of its first particle in pencil k.
__device__ __forceinline__ void partial_slots(const DevPlan &pl, int j, int z, int64_t *base, uint32_t *cstart,
                                              int *cn) {
    const int Nk = pl.Nk, W = pl.W, cpd = pl.cpd;
    const nf_sink_plan &own = pl.sink_plan[j * Nk + z];  // pencil k = z starts with cell z
    *cstart = own.cell[0].start;
    *cn = own.cell[0].N;
    for (int c = 0; c < W; c++) {
        int k = z - c;
        if (k < 0) k += cpd;
        const int s = j * Nk + k;
        const nf_sink_plan &pk = pl.sink_plan[s];
        int off = 0;
        for (int c2 = 0; c2 < c; c2++) off += pk.cell[c2].N;
        base[c] = (int64_t)pl.sink_set_start[s] + off;
    }
}

// FP32 co-add in the reference's order: the pencils of a skewer are applied
// in increasing k, the first one setting the value and the rest adding.  For
// cell z that is c = min(z, W-1) .. 0, then W-1 .. z+1.  One wave per chunk
// of 64 particles of a sink cell.
// RescaleAndCoAddAcceleration (kick.cpp:96-105) on one sink's float4, with
// the FLOAT3p1 rules (float3p1.cpp:41-56): x, y, z = (a * inv_eps3 + far) *
// rescale in float, each operation rounded on its own (no contraction), and
// w -= DensityKernelRad2 in double.
__device__ __forceinline__ float4 apply_kick(float4 a, const KickArgs &k, int64_t row) {
#pragma clang fp contract(off)
    if (!k.on) return a;
    if (k.near_out)  // the NearAcc record: FLOAT3p1 *= FLOAT scales xyz only
        reinterpret_cast<float4 *>(k.near_out)[row] =
            make_float4(a.x * k.inv_eps3, a.y * k.inv_eps3, a.z * k.inv_eps3, a.w);
    const float *f = k.far + 3 * (row - k.far_base);
    a.x = (a.x * k.inv_eps3 + f[0]) * k.rescale;
    a.y = (a.y * k.inv_eps3 + f[1]) * k.rescale;
    a.z = (a.z * k.inv_eps3 + f[2]) * k.rescale;
    a.w = (float)((double)a.w - k.density_rad2);
    return a;
}

__global__ __launch_bounds__(256) void k_coadd_plain(DevPlan pl, const float4 *__restrict__ part,
                                                     float4 *__restrict__ acc, KickArgs kick) {
    const int q = blockIdx.x * 4 + (int)(threadIdx.x / kWave);
    if (q >= coadd_chunks(pl) || plan_bad(pl)) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int cell = pl.coadd_chunk[2 * q], first = pl.coadd_chunk[2 * q + 1];
    const int j = cell / pl.Nk, z = cell % pl.Nk;
    int64_t base[NF_MAX_WIDTH];
    uint32_t cstart;
    int cn;
    partial_slots(pl, j, z, base, &cstart, &cn);
    const int W = pl.W;
    int order[NF_MAX_WIDTH], no = 0;
    for (int c = (z < W - 1 ? z : W - 1); c >= 0; c--) order[no++] = c;
    for (int c = W - 1; c > z; c--) order[no++] = c;
    const int i = first + lane;
    if (i >= cn) return;
    float4 a = part[base[order[0]] + i];
    for (int t = 1; t < W; t++) {
        const float4 p = part[base[order[t]] + i];
        a.x += p.x;
        a.y += p.y;
        a.z += p.z;
        a.w += p.w;
    }
    acc[cstart + i] = apply_kick(a, kick, (int64_t)cstart + i);
}

// FP64 co-add: sum of the W double partials, scaled back to the reference's
// units (out_scale applies to a, not to w).  One wave per 64-particle chunk.
__device__ __forceinline__ void coadd_chunk_f64(const DevPlan &pl, const double *__restrict__ part, double out_scale,
                                                float4 *__restrict__ acc, const KickArgs &kick, int q, int lane) {
    const int cell = pl.coadd_chunk[2 * q], first = pl.coadd_chunk[2 * q + 1];
    const int j = cell / pl.Nk, z = cell % pl.Nk;
    int64_t base[NF_MAX_WIDTH];
    uint32_t cstart;
    int cn;
    partial_slots(pl, j, z, base, &cstart, &cn);
    const int W = pl.W;
    const int i = first + lane;
    if (i >= cn) return;
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int c = 0; c < W; c++) {
        const double *p = part + 4 * (base[c] + i);
        a0 += p[0];
        a1 += p[1];
        a2 += p[2];
        a3 += p[3];
    }
    const float4 a = make_float4((float)(a0 * out_scale), (float)(a1 * out_scale), (float)(a2 * out_scale), (float)a3);
    acc[cstart + i] = apply_kick(a, kick, (int64_t)cstart + i);
}

__global__ __launch_bounds__(256) void k_coadd_f64(DevPlan pl, const double *__restrict__ part, double out_scale,
                                                   float4 *__restrict__ acc, KickArgs kick) {
    const int q = blockIdx.x * 4 + (int)(threadIdx.x / kWave);
    if (q >= coadd_chunks(pl) || plan_bad(pl)) return;
    coadd_chunk_f64(pl, part, out_scale, acc, kick, q, threadIdx.x & (kWave - 1));
}

// Every SIC's FP64 co-add of a launch group in one launch.
__global__ __launch_bounds__(256) void k_coadd_f64_group(const GroupSic *__restrict__ g,
                                                         const int32_t *__restrict__ first, int n) {
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (int)(threadIdx.x / kWave));
    if (w >= cptr<int32_t>(first)[n]) return;
    const int i = group_sic_of(first, n, w);
    const DevPlan pl = cptr<GroupSic>(g)[i].pl;
    const int q = w - cptr<int32_t>(first)[i];
    if (q >= coadd_chunks(pl) || plan_bad(pl)) return;
    const KickArgs kick = cptr<GroupSic>(g)[i].kick;
    coadd_chunk_f64(pl, static_cast<const double *>(cptr<GroupSic>(g)[i].partial), cptr<GroupSic>(g)[i].out_scale,
                    reinterpret_cast<float4 *>(cptr<GroupSic>(g)[i].acc4), kick, q, threadIdx.x & (kWave - 1));
}

// ---------------------------------------------------------------------------
// Batched path
// ---------------------------------------------------------------------------

// Pair kernel in box units.  r2 carries pc.r2add; for the single spline the
// result is g/e^3 (the e^3 is restored in the co-add).  MODE selects what a
// source tile needs, decided from the tile's bounding box:
//   kNoDensity  unsoftened branch only, no FOF density (no pair within b)
//   kDensity    unsoftened branch only, with the FOF density term
//   kExact      the full softened form with the density term
// The unsoftened branch equals the full form for every pair with r > e.
constexpr int kNoDensity = 0, kDensity = 1, kExact = 2;

template <int SOFT, int MODE>
__device__ __forceinline__ void pair_box(float sx, float sy, float sz, float kx, float ky, float kz,
                                         const PairConst &pc, float &ax, float &ay, float &az, float &aw) {
    const float dx = sx - kx, dy = sy - ky, dz = sz - kz;
    float r2 = fmaf(dx, dx, pc.r2add);
    r2 = fmaf(dy, dy, r2);
    r2 = fmaf(dz, dz, r2);
    if constexpr (MODE != kNoDensity) aw += clamp01(pc.b2x - r2);
    float g;
    if constexpr (SOFT == NF_SOFT_SINGLE_SPLINE) {
        const float f = rsq(r2);
        const float f3 = f * f * f;
        if constexpr (MODE != kExact) g = f3;
        else g = fminf(f3, fmaf(fmaf(pc.c1, f, pc.c2), r2, pc.c3));
    } else if constexpr (SOFT == NF_SOFT_PLUMMER) {
        const float f = rsq(r2);
        g = f * f * f;
    } else if constexpr (SOFT == NF_SOFT_CUBIC_SPLINE) {
        const float rinv = rsq(r2);
        const float u = pc.eps * __builtin_amdgcn_sqrtf(r2);
        if (u >= 2.0f) {
            g = rinv * rinv * rinv;
        } else if (u >= 1.0f) {
            const float u3 = u * u * u;
            g = rinv * rinv * rinv *
                ((float)(-1.0 / 15.0) + (float)(8.0 / 3.0) * u3 - 3.0f * u3 * u + (float)(6.0 / 5.0) * u3 * u * u -
                 (float)(1.0 / 6.0) * u3 * u3);
        } else {
            g = pc.eps * pc.eps * pc.eps * ((float)(4.0 / 3.0) - (float)(6.0 / 5.0) * u * u + 0.5f * u * u * u);
        }
    } else {  // cubic Plummer; r2 already holds + TAU2.  r^3 as r2*sqrt(r2) so a
              // padding source (r2 = inf) gives 1/inf = 0, not inf*0.
        const float r = r2 * __builtin_amdgcn_sqrtf(r2) + pc.eps;
        g = __builtin_amdgcn_rcpf(r);
    }
    ax = fmaf(-g, dx, ax);
    ay = fmaf(-g, dy, ay);
    az = fmaf(-g, dz, az);
}

// TwoSum (Knuth): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    const float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void merge(float &h, float &l, float t) {
    float s, e;
    two_sum(h, t, s, e);
    h = s;
    l += e;
}

// Squared gap between two axis-aligned boxes (0 if they overlap).
__device__ __forceinline__ float box_gap2(float alx, float aly, float alz, float ahx, float ahy, float ahz, float4 bl,
                                          float4 bh) {
    const float gx = fmaxf(0.f, fmaxf(bl.x - ahx, alx - bh.x));
    const float gy = fmaxf(0.f, fmaxf(bl.y - ahy, aly - bh.y));
    const float gz = fmaxf(0.f, fmaxf(bl.z - ahz, alz - bh.z));
    return gx * gx + gy * gy + gz * gz;
}

struct BatchCtx {
    PairConst pc;
    double invcpd;
    float soft2;  // squared distance above which the unsoftened branch is exact (with margin)
    float dens2;  // squared distance above which the FOF density term is zero (with margin)
};

// A sink pencil's cells, for reading its sinks straight from the slab.
struct SinkCells {
    uint32_t start[NF_MAX_WIDTH];
    int pre[NF_MAX_WIDTH + 1];  // sinks before each cell
};
__device__ __forceinline__ SinkCells sink_cells(const SicView &v, int s) {
    const nf_sink_plan p = cptr<nf_sink_plan>(v.sink_plan)[s];
    SinkCells sc;
    sc.pre[0] = 0;
#pragma unroll
    for (int c = 0; c < NF_MAX_WIDTH; c++) {
        sc.start[c] = p.cell[c].start;
        sc.pre[c + 1] = sc.pre[c] + (c < v.W ? p.cell[c].N : 0);
    }
    return sc;
}
// Sink i of the pencil: its slab position with z moved into the pencil's
// central cell by the FP64 offset (c - R)/cpd, rounded once.
__device__ __forceinline__ float3 pencil_sink(const SicView &v, const SinkCells &sc, int i, double invcpd) {
    // The cell holding sink i, by selects over the (wave-uniform) cell table:
    // no indexed private array.
    int c = 0, pre = 0;
    uint32_t start = sc.start[0];
#pragma unroll
    for (int cc = 1; cc < NF_MAX_WIDTH; cc++) {
        const bool in = cc < v.W && i >= sc.pre[cc];
        c = in ? cc : c;
        pre = in ? sc.pre[cc] : pre;
        start = in ? sc.start[cc] : start;
    }
    const float3 q = slab_xyz(v.sink_slab, v.sink_n, (uint64_t)start + (uint64_t)(i - pre), v.aos);
    return make_float3(q.x, q.y, (float)((double)q.z + (double)(c - v.R) * invcpd));
}

// The mode a source tile needs, from the squared gap between its bounding box
// and the sinks' bounding box (wave-uniform).
template <int SOFT>
__device__ __forceinline__ int tile_mode(const BatchCtx &cx, float gap2) {
    if constexpr (SOFT == NF_SOFT_CUBIC_SPLINE || SOFT == NF_SOFT_CUBIC_PLUMMER) return kExact;
    if constexpr (SOFT == NF_SOFT_SINGLE_SPLINE)
        if (!cx.pc.fast_ok || gap2 <= cx.soft2) return kExact;
    return gap2 <= cx.dens2 ? kDensity : kNoDensity;
}

// Pair interactions a wave has computed in each pair form (kNoDensity,
// kDensity, kExact) and in short tails: wave-uniform, so the counts live in
// scalar registers and cost no VALU work.  Counted only by the census
// instantiations of k_batched: eight more 64-bit scalars push the single
// spline's S = 2 kernel into SGPR and VGPR spills that cost it 7%.
struct Census {
    uint64_t form[3];
    uint64_t tail;
};
__device__ __forceinline__ void census_add(Census &c, int mode, uint64_t n) {
    c.form[0] += mode == kNoDensity ? n : 0;
    c.form[1] += mode == kDensity ? n : 0;
    c.form[2] += mode == kExact ? n : 0;
}

// One source tile (cnt <= 64 sources) against the S sinks
// of every lane; sources are wave-uniform loads (scalar cache).  The force
// terms of each batch of kSrcBatch sources are summed on their own before
// they join the tile sum, for one add per component per batch.  On random
// terms a sum of 4 batch sums of 16 terms carries 1.7-2.3x less FP32
// rounding than one running sum of 64 terms, and 1.0-1.1x the rounding of 8
// batch sums of 8 terms, which cost 1% more kernel time.  Sink slots
// u < S/2 take pair form MLO, the others MHI (S = 1: MHI); every form gives
// the same value for the pairs its box test admits, so the split changes
// which code runs, not the result.  The last batch takes only the
// cnt % kSrcBatch real sources: the padding sources after them give exact
// zeros, so leaving them out changes no sum.
constexpr int kSrcBatch = 16;
template <int SOFT, int MLO, int MHI, int S>
__device__ __forceinline__ void full_tile(const NF_CONST float4 *sp, int cnt, const float *kx, const float *ky,
                                          const float *kz, const PairConst &pc, float *tx, float *ty, float *tz,
                                          float *tw) {
    const int fullb = cnt & ~(kSrcBatch - 1);
    for (int i = 0; i < fullb; i += kSrcBatch) {
        float bx[S], by[S], bz[S];
#pragma unroll
        for (int u = 0; u < S; u++) bx[u] = by[u] = bz[u] = 0.f;
#pragma unroll
        for (int q = 0; q < kSrcBatch; q++) {
            const float4 sv = sp[i + q];
#pragma unroll
            for (int u = 0; u < S / 2; u++)
                pair_box<SOFT, MLO>(sv.x, sv.y, sv.z, kx[u], ky[u], kz[u], pc, bx[u], by[u], bz[u], tw[u]);
#pragma unroll
            for (int u = S / 2; u < S; u++)
                pair_box<SOFT, MHI>(sv.x, sv.y, sv.z, kx[u], ky[u], kz[u], pc, bx[u], by[u], bz[u], tw[u]);
        }
#pragma unroll
        for (int u = 0; u < S; u++) {
            tx[u] += bx[u];
            ty[u] += by[u];
            tz[u] += bz[u];
        }
    }
    if (cnt > fullb) {
        float bx[S], by[S], bz[S];
#pragma unroll
        for (int u = 0; u < S; u++) bx[u] = by[u] = bz[u] = 0.f;
        for (int q = fullb; q < cnt; q++) {
            const float4 sv = sp[q];
#pragma unroll
            for (int u = 0; u < S / 2; u++)
                pair_box<SOFT, MLO>(sv.x, sv.y, sv.z, kx[u], ky[u], kz[u], pc, bx[u], by[u], bz[u], tw[u]);
#pragma unroll
            for (int u = S / 2; u < S; u++)
                pair_box<SOFT, MHI>(sv.x, sv.y, sv.z, kx[u], ky[u], kz[u], pc, bx[u], by[u], bz[u], tw[u]);
        }
#pragma unroll
        for (int u = 0; u < S; u++) {
            tx[u] += bx[u];
            ty[u] += by[u];
            tz[u] += bz[u];
        }
    }
}

// A tile with the modes of its two halves of sink slots.
template <int SOFT, int S>
__device__ __forceinline__ void mixed_tile(int mlo, int mhi, const NF_CONST float4 *sp, int cnt, const float *kx,
                                           const float *ky, const float *kz, const PairConst &pc, float *tx,
                                           float *ty, float *tz, float *tw) {
    if (S == 1 || mlo == mhi) {
        if (mhi == kNoDensity) full_tile<SOFT, kNoDensity, kNoDensity, S>(sp, cnt, kx, ky, kz, pc, tx, ty, tz, tw);
        else if (mhi == kDensity) full_tile<SOFT, kDensity, kDensity, S>(sp, cnt, kx, ky, kz, pc, tx, ty, tz, tw);
        else full_tile<SOFT, kExact, kExact, S>(sp, cnt, kx, ky, kz, pc, tx, ty, tz, tw);
    } else if (mlo == kNoDensity) {
        if (mhi == kDensity) full_tile<SOFT, kNoDensity, kDensity, S>(sp, cnt, kx, ky, kz, pc, tx, ty, tz, tw);
        else full_tile<SOFT, kNoDensity, kExact, S>(sp, cnt, kx, ky, kz, pc, tx, ty, tz, tw);
    } else if (mlo == kDensity) {
        if (mhi == kNoDensity) full_tile<SOFT, kDensity, kNoDensity, S>(sp, cnt, kx, ky, kz, pc, tx, ty, tz, tw);
        else full_tile<SOFT, kDensity, kExact, S>(sp, cnt, kx, ky, kz, pc, tx, ty, tz, tw);
    } else {
        if (mhi == kNoDensity) full_tile<SOFT, kExact, kNoDensity, S>(sp, cnt, kx, ky, kz, pc, tx, ty, tz, tw);
        else full_tile<SOFT, kExact, kDensity, S>(sp, cnt, kx, ky, kz, pc, tx, ty, tz, tw);
    }
}

// Register tile: 64*S sinks of one pencil on the lanes; sources are wave
// uniform and come through the scalar cache.
template <int SOFT, int S, bool CENSUS>
__device__ __forceinline__ void full_item(const BatchCtx &cx, const SicView &v, const WorkItem &w, int lane,
                                          Census &cen) {
    const NF_CONST float4 *src = cptr<float4>(v.sources4);
    const NF_CONST float4 *box = cptr<float4>(v.box4);
    const NF_CONST int32_t *sink_set_start = cptr<int32_t>(v.sink_set_start);
    const NF_CONST int32_t *src_set_start = cptr<int32_t>(v.src_set_start);
    const NF_CONST int32_t *src_set_count = cptr<int32_t>(v.src_set_count);
    const NF_CONST int32_t *ilist = cptr<int32_t>(v.ilist);
    NF_GLOBAL double *out = (NF_GLOBAL double *)v.partial;
    const int s = w.sinkset;
    const int base = sink_set_start[s] + w.first;
    const int nsink = w.nsink;  // 64*S, or fewer for a masked tail tile
    const uint64_t n_lo = (uint64_t)min(nsink, kWave * (S / 2)), n_hi = (uint64_t)nsink - n_lo;  // sinks per half
    const SinkCells sc = sink_cells(v, s);
    float kx[S], ky[S], kz[S];
    // The sinks' box: x and y over all slots, z per half of the slots (the
    // slots run along the pencil, z), so each half of a tile can take its
    // own pair form.
    float blx = __builtin_huge_valf(), bly = blx, bhx = -blx, bhy = -blx;
    float blz0 = blx, bhz0 = -blx, blz1 = blx, bhz1 = -blx;
#pragma unroll
    for (int u = 0; u < S; u++) {
        // Lanes past nsink repeat sink 0 (inside the box, results not stored).
        const int ki = lane + kWave * u < nsink ? lane + kWave * u : 0;
        const float3 k = pencil_sink(v, sc, w.first + ki, cx.invcpd);
        kx[u] = k.x;
        ky[u] = k.y;
        kz[u] = k.z;
        blx = fminf(blx, k.x);
        bly = fminf(bly, k.y);
        bhx = fmaxf(bhx, k.x);
        bhy = fmaxf(bhy, k.y);
        if (u < S / 2) {
            blz0 = fminf(blz0, k.z);
            bhz0 = fmaxf(bhz0, k.z);
        } else {
            blz1 = fminf(blz1, k.z);
            bhz1 = fmaxf(bhz1, k.z);
        }
    }
    blx = wave_min(blx);
    bly = wave_min(bly);
    bhx = wave_max(bhx);
    bhy = wave_max(bhy);
    blz1 = wave_min(blz1);
    bhz1 = wave_max(bhz1);
    if (S > 1) {
        blz0 = wave_min(blz0);
        bhz0 = wave_max(bhz0);
    }

    // Each block's FP32 tile sum joins an FP64 partial: a convert and an add
    // per component, which gfx950 issues at the FP32 rate (a TwoSum merge
    // into an FP32 (hi, lo) pair is 7 instructions and carries fewer bits).
    // The FOF density w too: in a dense halo a sink meets thousands of pairs
    // within b, and one FP32 running sum of them drifts by ~1e-5.
    double hx[S], hy[S], hz[S], hw[S];
#pragma unroll
    for (int u = 0; u < S; u++) hx[u] = hy[u] = hz[u] = hw[u] = 0.0;

    const int W = v.W;
    for (int y = 0; y < W; y++) {
        const int si = ilist[s * W + y];
        const int st = src_set_start[si];
        const int n = src_set_count[si];
        // Source y + (y-R)/cpd == sink y - (y-R)/cpd on the other side; the
        // cell-centre offset is formed in FP64.
        const double yo = (double)(y - v.R) * cx.invcpd;
        float kyy[S];
#pragma unroll
        for (int u = 0; u < S; u++) kyy[u] = (float)((double)ky[u] - yo);
        const float sly = (float)((double)bly - yo), shy = (float)((double)bhy - yo);
        const int nblk = (n + kWave - 1) / kWave;
        // Lane b decides the pair forms of block b of this source pencil (64
        // blocks per pass): one vector pass instead of a wave-uniform box test
        // and a dependent box load per block.
        int modes = 0;  // lane i: mlo | mhi << 2 of block b + i, for the pass starting at block b
        for (int b = 0; b < nblk; b++) {
            if ((b & (kWave - 1)) == 0) {
                const int bl = b + lane;
                modes = 0;
                if (bl < nblk) {
                    const int64_t gb = (int64_t)st / kWave + bl;
                    const float4 ql = box[2 * gb], qh = box[2 * gb + 1];
                    const float gx = fmaxf(0.f, fmaxf(ql.x - bhx, blx - qh.x));
                    const float gy = fmaxf(0.f, fmaxf(ql.y - shy, sly - qh.y));
                    const float gxy = gx * gx + gy * gy;
                    const float gz1 = fmaxf(0.f, fmaxf(ql.z - bhz1, blz1 - qh.z));
                    const int mh = tile_mode<SOFT>(cx, gxy + gz1 * gz1);
                    int ml = mh;
                    if (S > 1) {
                        const float gz0 = fmaxf(0.f, fmaxf(ql.z - bhz0, blz0 - qh.z));
                        ml = tile_mode<SOFT>(cx, gxy + gz0 * gz0);
                    }
                    modes = ml | (mh << 2);
                }
            }
            const int pm = __builtin_amdgcn_readlane(modes, b & (kWave - 1));
            const int mlo = pm & 3, mhi = pm >> 2;
            const NF_CONST float4 *sp = src + st + b * kWave;
            const int cnt = min(kWave, n - b * kWave);
            if constexpr (CENSUS) {
                census_add(cen, mlo, n_lo * (uint64_t)cnt);
                census_add(cen, mhi, n_hi * (uint64_t)cnt);
            }
            float tx[S], ty[S], tz[S], tw[S];
#pragma unroll
            for (int u = 0; u < S; u++) tx[u] = ty[u] = tz[u] = tw[u] = 0.f;
            mixed_tile<SOFT, S>(mlo, mhi, sp, cnt, kx, kyy, kz, cx.pc, tx, ty, tz, tw);
#pragma unroll
            for (int u = 0; u < S; u++) {
                hx[u] += (double)tx[u];
                hy[u] += (double)ty[u];
                hz[u] += (double)tz[u];
                hw[u] += (double)tw[u];
            }
        }
    }
#pragma unroll
    for (int u = 0; u < S; u++) {
        if (lane + kWave * u >= nsink) continue;
        NF_GLOBAL double *o = out + 4 * (int64_t)(base + lane + kWave * u);
        o[0] = hx[u];
        o[1] = hy[u];
        o[2] = hz[u];
        o[3] = hw[u];
    }
}

// Short tail of a pencil (fewer than kTailLanes sinks): the wave is four
// groups of 16 lanes; each group takes SU sinks of the pass (4*SU sinks per
// pass) and its lanes stride over the sources 16 apart.  Per-lane partial
// sums are TwoSum-reduced within each group by a 4-level butterfly, which
// reduces the four groups' sinks at once.
constexpr int kGroupLanes = 16;
constexpr int kGroups = kWave / kGroupLanes;

// TwoSum butterfly within aligned groups of kGroupLanes lanes.
__device__ __forceinline__ void group_two_sum(float &h, float &l) {
    for (int off = kGroupLanes / 2; off >= 1; off >>= 1) {
        const float h2 = __shfl_xor(h, off);
        const float l2 = __shfl_xor(l, off);
        float s, e;
        two_sum(h, h2, s, e);
        h = s;
        l = l + l2 + e;
    }
}
__device__ __forceinline__ double group_sum(double v) {
    for (int off = kGroupLanes / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

template <int SOFT, int SU, bool CENSUS>
__device__ __forceinline__ void rem_item(const BatchCtx &cx, const SicView &v, const WorkItem &w, int lane,
                                         Census &cen) {
    constexpr int TC = 16;  // 64-source blocks per TwoSum tile
    const NF_CONST float4 *box = cptr<float4>(v.box4);
    const NF_CONST int32_t *sink_set_start = cptr<int32_t>(v.sink_set_start);
    const NF_CONST int32_t *src_set_start = cptr<int32_t>(v.src_set_start);
    const NF_CONST int32_t *src_set_count = cptr<int32_t>(v.src_set_count);
    const NF_CONST int32_t *ilist = cptr<int32_t>(v.ilist);
    const float4 *src = static_cast<const float4 *>(v.sources4);
    NF_GLOBAL double *out = (NF_GLOBAL double *)v.partial;
    const int s = w.sinkset;
    const int base = sink_set_start[s] + w.first;
    const int r = w.nsink;
    const int W = v.W;
    const int grp = lane / kGroupLanes, gl = lane % kGroupLanes;
    const SinkCells sc = sink_cells(v, s);
    for (int g0 = 0; g0 < r; g0 += kGroups * SU) {
        // This group's sinks: g0 + grp*SU + u (repeats of the last sink past
        // the tail are computed and not stored).
        float kx[SU], ky[SU], kz[SU];
        float blx = __builtin_huge_valf(), bly = blx, blz = blx, bhx = -blx, bhy = -blx, bhz = -blx;
#pragma unroll
        for (int u = 0; u < SU; u++) {
            const int gi = min(g0 + grp * SU + u, r - 1);
            const float3 k = pencil_sink(v, sc, w.first + gi, cx.invcpd);
            kx[u] = k.x;
            ky[u] = k.y;
            kz[u] = k.z;
            blx = fminf(blx, k.x);
            bly = fminf(bly, k.y);
            blz = fminf(blz, k.z);
            bhx = fmaxf(bhx, k.x);
            bhy = fmaxf(bhy, k.y);
            bhz = fmaxf(bhz, k.z);
        }
        // The pass's box over all four groups: the tile mode is wave-uniform.
        blx = wave_min(blx);
        bly = wave_min(bly);
        blz = wave_min(blz);
        bhx = wave_max(bhx);
        bhy = wave_max(bhy);
        bhz = wave_max(bhz);
        float hx[SU], hy[SU], hz[SU], lx[SU], ly[SU], lz[SU], hw[SU];
#pragma unroll
        for (int u = 0; u < SU; u++) hx[u] = hy[u] = hz[u] = lx[u] = ly[u] = lz[u] = hw[u] = 0.f;
        for (int y = 0; y < W; y++) {
            const int si = ilist[s * W + y];
            const int st = src_set_start[si];
            const int n = src_set_count[si];
            const double yo = (double)(y - v.R) * cx.invcpd;
            float kyy[SU];
#pragma unroll
            for (int u = 0; u < SU; u++) kyy[u] = (float)((double)ky[u] - yo);
            const float sly = (float)((double)bly - yo), shy = (float)((double)bhy - yo);
            const int nblk = (n + kWave - 1) / kWave;
            for (int b0 = 0; b0 < nblk; b0 += TC) {
                const int b1 = min(nblk, b0 + TC);
                float tx[SU], ty[SU], tz[SU], tw[SU];
#pragma unroll
                for (int u = 0; u < SU; u++) tx[u] = ty[u] = tz[u] = tw[u] = 0.f;
                for (int b = b0; b < b1; b++) {
                    const int64_t gb = (int64_t)st / kWave + b;
                    const int mode =
                        tile_mode<SOFT>(cx, box_gap2(blx, sly, blz, bhx, shy, bhz, box[2 * gb], box[2 * gb + 1]));
                    if constexpr (CENSUS) {
                        const uint64_t pairs =
                            (uint64_t)min(kGroups * SU, r - g0) * (uint64_t)min(kWave, n - b * kWave);
                        census_add(cen, mode, pairs);
                        cen.tail += pairs;
                    }
                    // The block's 64 sources (padding included: exact zeros),
                    // 16 at a time on the group's lanes.
                    const float4 *bp = src + st + b * kWave + gl;
#pragma unroll
                    for (int c = 0; c < kGroups; c++) {
                        const float4 sv = bp[c * kGroupLanes];
                        if (mode == kNoDensity) {
#pragma unroll
                            for (int u = 0; u < SU; u++)
                                pair_box<SOFT, kNoDensity>(sv.x, sv.y, sv.z, kx[u], kyy[u], kz[u], cx.pc, tx[u],
                                                           ty[u], tz[u], tw[u]);
                        } else if (mode == kDensity) {
#pragma unroll
                            for (int u = 0; u < SU; u++)
                                pair_box<SOFT, kDensity>(sv.x, sv.y, sv.z, kx[u], kyy[u], kz[u], cx.pc, tx[u], ty[u],
                                                         tz[u], tw[u]);
                        } else {
#pragma unroll
                            for (int u = 0; u < SU; u++)
                                pair_box<SOFT, kExact>(sv.x, sv.y, sv.z, kx[u], kyy[u], kz[u], cx.pc, tx[u], ty[u],
                                                       tz[u], tw[u]);
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < SU; u++) {
                    merge(hx[u], lx[u], tx[u]);
                    merge(hy[u], ly[u], ty[u]);
                    merge(hz[u], lz[u], tz[u]);
                    hw[u] += tw[u];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < SU; u++) {
            group_two_sum(hx[u], lx[u]);
            group_two_sum(hy[u], ly[u]);
            group_two_sum(hz[u], lz[u]);
            const double ws = group_sum((double)hw[u]);  // (an FP64 lane sum spills at S = 2)
            const int gi = g0 + grp * SU + u;
            if (gl == 0 && gi < r) {
                NF_GLOBAL double *o = out + 4 * (int64_t)(base + gi);
                o[0] = (double)hx[u] + (double)lx[u];
                o[1] = (double)hy[u] + (double)ly[u];
                o[2] = (double)hz[u] + (double)lz[u];
                o[3] = ws;
            }
        }
    }
}

constexpr int kRemSinks = 2;  // sinks per 16-lane group per pass of a short tail

// Occupancy target WPE (waves per SIMD): more resident waves hide the v_rsq
// and scalar-load latencies, as long as the tile loops' registers fit.  The
// most waves at which each (pair form, sinks per lane) compiled with no VGPR
// spills for gfx950 (6 waves: 80 VGPRs; 5: 96; 4: 128; 2: 256); the single
// spline and Plummer forms keep 4 waves at S = 4, which measured faster than
// 5.  With FP64 tile accumulation the S = 1 forms spill a few words in their
// item set-up, outside the tile loops, at no measured cost.
constexpr int batched_waves_per_eu(int soft, int s) {
    switch (soft) {
        case NF_SOFT_SINGLE_SPLINE: return s == 4 ? 4 : 6;
        case NF_SOFT_PLUMMER: return s == 4 ? 4 : 5;
        case NF_SOFT_CUBIC_SPLINE: return s == 1 ? 5 : s == 2 ? 4 : 2;
        default: return s == 4 ? 5 : 6;  // cubic Plummer
    }
}

}  // namespace

int batched_kernel_waves(int softening, int sinks_per_lane) {
    return batched_waves_per_eu(softening, sinks_per_lane);
}

namespace {

// Persistent kernel: every wave pops items (of any SIC of the group) until
// the queue is empty.  CENSUS: count the pair interactions of each pair form
// into census (the diagnostic instantiations; the results are the same).
// counter[0] is the queue; nitems < 0 means a list of at most -nitems items
// whose length the prologue wrote into counter[1]; a census build adds its
// counts into the four uint64 at counter + kCensusOffset.  One queue for
// the whole grid: dealing the list to one queue per XCD, in interleaved
// chunks of 8 to 128 consecutive items (which share source pencils), was
// measured 11% slower.
constexpr int kCensusOffset = 16;  // int32 units: 64 bytes into the counter's slot
template <int SOFT, int S, int WPE, bool CENSUS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_batched(
    BatchCtx cx, const SicView *__restrict__ views, const WorkItem *__restrict__ items, int32_t nitems,
    int32_t *__restrict__ counter) {
    const int lane = threadIdx.x & (kWave - 1);
    if (nitems < 0) nitems = min(-nitems, *cptr<int32_t>(counter + 1));
    Census cen = {{0, 0, 0}, 0};
    for (;;) {
        int it = 0;
        if (lane == 0) it = atomicAdd(counter, 1);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= nitems) {
            if (CENSUS && lane == 0) {
                unsigned long long *census = reinterpret_cast<unsigned long long *>(counter + kCensusOffset);
                for (int f = 0; f < 3; f++)
                    if (cen.form[f]) atomicAdd(census + f, (unsigned long long)cen.form[f]);
                if (cen.tail) atomicAdd(census + 3, (unsigned long long)cen.tail);
            }
            return;
        }
        const WorkItem w = cptr<WorkItem>(items)[it];
        const int kind = w.kind_sic & 3;
        const SicView v = cptr<SicView>(views)[w.kind_sic >> 2];
        if (kind == 0) full_item<SOFT, S, CENSUS>(cx, v, w, lane, cen);
        else if (kind == 2) full_item<SOFT, 1, CENSUS>(cx, v, w, lane, cen);
        else rem_item<SOFT, kRemSinks, CENSUS>(cx, v, w, lane, cen);
    }
}

// The batched launch's prologue (a fill kernel, so that it can sit in a
// captured graph): clears the queue counter and, for a census build, the
// census; for a device-planned list, copies its length into counter[1], or
// 0 when the group holds a cell outside its slab.
__global__ void k_zero_counter(int32_t *c, const int32_t *dnitems, const int32_t *dbad, int census) {
    c[0] = 0;
    if (dnitems) c[1] = (dbad && *dbad != 0) ? 0 : *dnitems;
    if (census) {
        unsigned long long *cs = reinterpret_cast<unsigned long long *>(c + kCensusOffset);
        for (int f = 0; f < 4; f++) cs[f] = 0;
    }
}

// Pair-math probe: the register-tile loop of k_batched alone (16-source batch
// sums, 64-source FP32 tile sums accumulated in FP64), one pair form for
// every block, on sinks and sources that stay in the caches.  Each wave holds
// 64 S sinks and makes `passes` passes over `nsrc` sources (a multiple of
// 64); lane 0 of each wave stores a checksum so that nothing is optimised
// away.  Its rate is the ceiling of the tile loop at that tile and occupancy.
template <int SOFT, int S, int MODE, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void k_probe(
    PairConst pc, const float4 *__restrict__ src, int nsrc, int passes, double *__restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = blockIdx.x * (blockDim.x / kWave) + (int)(threadIdx.x / kWave);
    float kx[S], ky[S], kz[S];
#pragma unroll
    for (int u = 0; u < S; u++) {  // sinks spread over a cell-sized box around the sources
        const int i = lane + kWave * u + 7 * wave;
        kx[u] = 0.004f * (float)((i * 37) % 101) / 101.f - 0.002f;
        ky[u] = 0.004f * (float)((i * 53) % 103) / 103.f - 0.002f;
        kz[u] = 0.004f * (float)((i * 71) % 107) / 107.f - 0.002f;
    }
    double hx[S], hy[S], hz[S], hw[S];
#pragma unroll
    for (int u = 0; u < S; u++) hx[u] = hy[u] = hz[u] = hw[u] = 0.0;
    const NF_CONST float4 *sp = cptr<float4>(src);
    for (int p = 0; p < passes; p++)
        for (int b = 0; b < nsrc; b += kWave) {
            float tx[S], ty[S], tz[S], tw[S];
#pragma unroll
            for (int u = 0; u < S; u++) tx[u] = ty[u] = tz[u] = tw[u] = 0.f;
            full_tile<SOFT, MODE, MODE, S>(sp + b, kWave, kx, ky, kz, pc, tx, ty, tz, tw);
#pragma unroll
            for (int u = 0; u < S; u++) {
                hx[u] += (double)tx[u];
                hy[u] += (double)ty[u];
                hz[u] += (double)tz[u];
                hw[u] += (double)tw[u];
            }
        }
    double c = 0.0;
#pragma unroll
    for (int u = 0; u < S; u++) c += hx[u] + hy[u] + hz[u] + hw[u];
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off);
    if (lane == 0) out[wave] = c;
}

template <int S, int MODE, int WPE>
int run_probe_t(const PairConst &pc, double *rate) {
    constexpr int kSrc = 1024, kPasses = 2;  // ~0.3 s per launch at S = 2 on a full GPU
    int ncu = device_cu_count(), bpc = 0;
    if (ncu <= 0) return (int)hipErrorNoDevice;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_probe<NF_SOFT_SINGLE_SPLINE, S, MODE, WPE>,
                                                                256, 0);
    if (e != hipSuccess) return (int)e;
    const int grid = ncu * std::max(1, bpc);
    float4 h[kSrc];
    for (int i = 0; i < kSrc; i++)  // sources in the same box, none within the softening of a sink
        h[i] = make_float4(0.004f * (float)((i * 29) % 97) / 97.f - 0.002f + 2e-5f,
                           0.004f * (float)((i * 31) % 89) / 89.f - 0.002f + 2e-5f,
                           0.004f * (float)((i * 43) % 83) / 83.f - 0.002f + 2e-5f, 0.f);
    float4 *src = nullptr;
    double *out = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    float ms = 0.f;
    if ((e = hipMalloc(&src, sizeof h)) == hipSuccess && (e = hipMalloc(&out, sizeof(double) * grid * 4)) == hipSuccess &&
        (e = hipMemcpy(src, h, sizeof h, hipMemcpyHostToDevice)) == hipSuccess &&
        (e = hipEventCreate(&a)) == hipSuccess && (e = hipEventCreate(&b)) == hipSuccess) {
        // a warm-up launch, then the timed one
        hipLaunchKernelGGL((k_probe<NF_SOFT_SINGLE_SPLINE, S, MODE, WPE>), dim3(grid), dim3(256), 0, nullptr, pc, src,
                           kSrc, kPasses, out);
        (void)hipEventRecord(a, nullptr);
        hipLaunchKernelGGL((k_probe<NF_SOFT_SINGLE_SPLINE, S, MODE, WPE>), dim3(grid), dim3(256), 0, nullptr, pc, src,
                           kSrc, kPasses, out);
        (void)hipEventRecord(b, nullptr);
        e = hipEventSynchronize(b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (src) (void)hipFree(src);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) return (int)e;
    *rate = (double)grid * 4.0 * 64.0 * S * (double)kSrc * kPasses / (ms * 1e-3);
    return 0;
}

template <int SOFT>
int launch_plain_t(const DevPlan &pl, const PairConst &pc, const void *sinks4, const void *sources4, void *partial4,
                   hipStream_t st) {
    if (pl.nsink_blocks == 0) return 0;
    hipLaunchKernelGGL(k_plain<SOFT>, dim3(pl.nsink_blocks), dim3(kWave), 0, st, pl, pc,
                       static_cast<const float4 *>(sinks4), static_cast<const float4 *>(sources4),
                       static_cast<float4 *>(partial4));
    return (int)hipGetLastError();
}

}  // namespace

int device_cu_count() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return n;
}

int launch_gather(const DevPlan &pl, const DevSlabs &sl, bool batched_stream, void *sinks4, void *sources4,
                  void *stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    const float cellsize = (float)(1.0 / (double)pl.cpd);
    const double invcpd = 1.0 / (double)pl.cpd;
    float4 *box = static_cast<float4 *>(box_after(sources4, pl.nsrc_blocks));
    auto grid_of = [](int blocks, int per_wave = kGatherBlocks) {  // 4 waves per workgroup
        return dim3((unsigned)((blocks + 4 * per_wave - 1) / (4 * per_wave)));
    };
    if (sinks4 && pl.nsink_blocks > 0) {
        const dim3 grid = grid_of(pl.nsink_blocks);
        hipLaunchKernelGGL(k_gather_sinks, grid, dim3(256), 0, st, pl, sl.pos[pl.R], sl.n[pl.R], sl.aos,
                           static_cast<float4 *>(sinks4), cellsize);
    }
    if (pl.nsrc_blocks > 0) {
        const dim3 grid = grid_of(pl.nsrc_blocks);
        if (batched_stream) {
            const dim3 bgrid((unsigned)((gather_waves(pl.nsrc_sets, pl.nsrc_blocks, pl.max_src_blocks) + 3) / 4));
            if (gather_segmented(pl.max_src_blocks))
                hipLaunchKernelGGL(k_gather_sources_binned<true>, bgrid, dim3(256), 0, st, pl, sl,
                                   static_cast<float4 *>(sources4), box, (float)pl.cpd, invcpd);
            else
                hipLaunchKernelGGL(k_gather_sources_binned<false>, bgrid, dim3(256), 0, st, pl, sl,
                                   static_cast<float4 *>(sources4), box, (float)pl.cpd, invcpd);
        }
        else
            hipLaunchKernelGGL(k_gather_sources, grid, dim3(256), 0, st, pl, sl, static_cast<float4 *>(sources4),
                               box, cellsize);
    }
    return (int)hipGetLastError();
}

int launch_gather_group(const GroupSic *g, const int32_t *first, int n, int total, int cpd, bool segmented,
                        void *stream) {
    if (total <= 0) return 0;
    const dim3 grid((unsigned)((total + 3) / 4));
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (segmented)
        hipLaunchKernelGGL(k_gather_sources_binned_group<true>, grid, dim3(256), 0, st, g, first, n, (float)cpd,
                           1.0 / (double)cpd);
    else
        hipLaunchKernelGGL(k_gather_sources_binned_group<false>, grid, dim3(256), 0, st, g, first, n, (float)cpd,
                           1.0 / (double)cpd);
    return (int)hipGetLastError();
}

int launch_coadd_group(const GroupSic *g, const int32_t *first, int n, int total, void *stream) {
    if (total <= 0) return 0;
    hipLaunchKernelGGL(k_coadd_f64_group, dim3((unsigned)((total + 3) / 4)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), g, first, n);
    return (int)hipGetLastError();
}

int launch_plain(const DevPlan &pl, int softening, const PairConst &pc, const void *sinks4, const void *sources4,
                 void *partial4, void *stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (softening) {
        case NF_SOFT_PLUMMER: return launch_plain_t<NF_SOFT_PLUMMER>(pl, pc, sinks4, sources4, partial4, st);
        case NF_SOFT_SINGLE_SPLINE:
            return launch_plain_t<NF_SOFT_SINGLE_SPLINE>(pl, pc, sinks4, sources4, partial4, st);
        case NF_SOFT_CUBIC_SPLINE:
            return launch_plain_t<NF_SOFT_CUBIC_SPLINE>(pl, pc, sinks4, sources4, partial4, st);
        case NF_SOFT_CUBIC_PLUMMER:
            return launch_plain_t<NF_SOFT_CUBIC_PLUMMER>(pl, pc, sinks4, sources4, partial4, st);
        default: return (int)hipErrorInvalidValue;
    }
}

int launch_coadd_plain(const DevPlan &pl, const void *partial4, float *acc4, const KickArgs &kick, void *stream) {
    if (pl.n_coadd_chunks == 0) return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_coadd_plain, dim3((pl.n_coadd_chunks + 3) / 4), dim3(256), 0, st, pl,
                       static_cast<const float4 *>(partial4), reinterpret_cast<float4 *>(acc4), kick);
    return (int)hipGetLastError();
}

int launch_coadd_f64(const DevPlan &pl, const void *partial_d4, double out_scale, float *acc4,
                     const KickArgs &kick, void *stream) {
    if (pl.n_coadd_chunks == 0) return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_coadd_f64, dim3((pl.n_coadd_chunks + 3) / 4), dim3(256), 0, st, pl,
                       static_cast<const double *>(partial_d4), out_scale, reinterpret_cast<float4 *>(acc4), kick);
    return (int)hipGetLastError();
}

template <int SOFT, int S, bool CENSUS>
static int launch_batched_k(const BatchCtx &cx, const TileShape &ts, const SicView *views, const WorkItem *items,
                            int32_t nitems, int32_t *counter, hipStream_t st, const int32_t *dnitems,
                            const int32_t *dbad) {
    constexpr auto kern = k_batched<SOFT, S, batched_waves_per_eu(SOFT, S), CENSUS>;
    // Workgroups of 1, 2 or 4 waves; each wave pops its own items, so the
    // size only changes how the dispatcher places the persistent grid.
    const int wpw = ts.waves_per_wg == 1 || ts.waves_per_wg == 2 ? ts.waves_per_wg : 4;
    const int threads = kWave * wpw;
    int blocks_per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks_per_cu, kern, threads, 0) != hipSuccess ||
        blocks_per_cu <= 0)
        blocks_per_cu = 1;
    if (ts.waves_per_cu > 0) blocks_per_cu = std::max(1, std::min(blocks_per_cu, ts.waves_per_cu / wpw));
    const int ncu = device_cu_count();
    int grid = std::max(1, ncu) * blocks_per_cu;
    const int need = (nitems + wpw - 1) / wpw;
    if (grid > need) grid = std::max(1, need);
    static_assert(kCensusOffset * 4 == kCensusBytes && kCensusBytes + 32 <= 256, "census inside the counter's slot");
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, st, counter, dnitems, dbad, CENSUS ? 1 : 0);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, st, cx, views, items, dnitems ? -nitems : nitems, counter);
    return (int)hipGetLastError();
}

template <int SOFT, int S>
static int launch_batched_s(const BatchCtx &cx, const TileShape &ts, const SicView *views, const WorkItem *items,
                            int32_t nitems, int32_t *counter, hipStream_t st, const int32_t *dnitems,
                            const int32_t *dbad, bool census) {
    if (census) return launch_batched_k<SOFT, S, true>(cx, ts, views, items, nitems, counter, st, dnitems, dbad);
    return launch_batched_k<SOFT, S, false>(cx, ts, views, items, nitems, counter, st, dnitems, dbad);
}

template <int SOFT>
static int launch_batched_soft(const BatchCtx &cx, const TileShape &ts, const SicView *views, const WorkItem *items,
                               int32_t nitems, int32_t *counter, hipStream_t st, const int32_t *dnitems,
                               const int32_t *dbad, bool census) {
    switch (ts.sinks_per_lane) {
        case 1: return launch_batched_s<SOFT, 1>(cx, ts, views, items, nitems, counter, st, dnitems, dbad, census);
        case 2: return launch_batched_s<SOFT, 2>(cx, ts, views, items, nitems, counter, st, dnitems, dbad, census);
        case 4: return launch_batched_s<SOFT, 4>(cx, ts, views, items, nitems, counter, st, dnitems, dbad, census);
        default: return (int)hipErrorInvalidValue;
    }
}

// The single-spline pair-math probe at (sinks per lane, pair form, waves per
// SIMD): pair interactions per second over the whole GPU.
int run_pair_probe(const PairConst &pc, int sinks_per_lane, int mode, int waves_per_simd, double *rate) {
#define NF_PROBE(S_, M_, W_) \
    if (sinks_per_lane == S_ && mode == M_ && waves_per_simd == W_) return run_probe_t<S_, M_, W_>(pc, rate);
#define NF_PROBE_W(S_, M_) NF_PROBE(S_, M_, 4) NF_PROBE(S_, M_, 6) NF_PROBE(S_, M_, 8)
#define NF_PROBE_M(S_) NF_PROBE_W(S_, kNoDensity) NF_PROBE_W(S_, kDensity) NF_PROBE_W(S_, kExact)
    NF_PROBE_M(1)
    NF_PROBE_M(2)
    NF_PROBE_M(4)
#undef NF_PROBE_M
#undef NF_PROBE_W
#undef NF_PROBE
    return (int)hipErrorInvalidValue;
}

int launch_batched(const SicView *views, int cpd, int softening, const PairConst &pc, const TileShape &ts,
                   const WorkItem *items, int32_t nitems, int32_t *counter, void *stream, const int32_t *dnitems,
                   const int32_t *dbad, bool census) {
    if (nitems == 0) return 0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    BatchCtx cx;
    cx.pc = pc;
    cx.invcpd = 1.0 / (double)cpd;
    // The unsoftened branch is exact for r^2 > e^2 = 1/eps; the box test uses
    // a relative margin far above FP32 rounding of the box corners.
    cx.soft2 = (softening == NF_SOFT_SINGLE_SPLINE) ? (float)((1.0 / (double)pc.eps) * 1.01) : 0.f;
    // Density is nonzero only for r^2 + r2add < b2x; a 1% margin covers the
    // FP32 rounding of the box gap and of r^2.
    cx.dens2 = pc.b2x > 0.f ? (float)((double)pc.b2x * 1.01) : -1.f;
    switch (softening) {
        case NF_SOFT_PLUMMER: return launch_batched_soft<NF_SOFT_PLUMMER>(cx, ts, views, items, nitems, counter, st, dnitems, dbad, census);
        case NF_SOFT_SINGLE_SPLINE:
            return launch_batched_soft<NF_SOFT_SINGLE_SPLINE>(cx, ts, views, items, nitems, counter, st, dnitems, dbad, census);
        case NF_SOFT_CUBIC_SPLINE:
            return launch_batched_soft<NF_SOFT_CUBIC_SPLINE>(cx, ts, views, items, nitems, counter, st, dnitems, dbad, census);
        case NF_SOFT_CUBIC_PLUMMER:
            return launch_batched_soft<NF_SOFT_CUBIC_PLUMMER>(cx, ts, views, items, nitems, counter, st, dnitems, dbad, census);
        default: return (int)hipErrorInvalidValue;
    }
}

}  // namespace nf
