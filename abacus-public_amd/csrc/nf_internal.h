// Internal declarations shared by the host runtime (runtime.cpp), the
// planner (planner.cpp) and the device kernels (kernels.hip).
#ifndef NF_INTERNAL_H
#define NF_INTERNAL_H

#include <cstddef>
#include <cstdint>
#include <string>

#include "nf_direct.h"

namespace nf {

// Record an error message for nf_last_error() on this thread; returns code.
int fail(int code, const std::string &msg);

// nf_count_interactions' sums (planner.cpp); with strict false a negative
// count counts as zero.  max_pencil (or null): the largest source pencil.
int count_interactions(int32_t cpd, int32_t R, int32_t j_low, int32_t j_high, const nf_cellinfo *const *ci,
                       bool strict, uint64_t *out, int64_t *max_pencil);

// memcpy over nt threads, in 64-byte aligned chunks that cover every byte
// (hostcopy.cpp); par_memcpy picks nt from the host's cores and the size.
void par_memcpy_threads(void *dst, const void *src, size_t bytes, unsigned nt);
void par_memcpy(void *dst, const void *src, size_t bytes);

// Work item of the batched kernel: a run of `nsink` sinks starting at
// `first` within sink pencil `sinkset`.  kind 0: register tile of 64*S sinks,
// sinks on lanes; kind 2: one-wave tile of up to 64 sinks on the lanes (a
// pencil's tail; lanes past nsink idle); kind 1: a short tail (fewer than
// kTailLanes sinks), sources on the lanes.
// Measured on uniform and clustered boxes (thresholds 1-60): 32 is fastest
// on both, 3-4% ahead of 53 on a clustered box, 1.5% on a uniform one.
constexpr int kTailLanes = 32;  // below this many sinks a tail is cheaper with sources on the lanes

struct WorkItem {
    int32_t sinkset;
    int32_t first;
    int32_t nsink;
    int32_t kind_sic;  // kind + 4 * (index of the SIC in its launch group)
};

// A work list is two cost-sorted runs merged: the M register and one-wave
// tiles, and the T short tails (kind 1) spread evenly through them, tail t
// right after tile ceil((t+1) M / T) - 1.  A tail is mostly memory latency
// and a tile mostly FMA, so a persistent grid that pops them mixed keeps the
// SIMDs busier than one that runs every tail at the end (measured 4-6% on
// the whole launch).  Returns the list position of tile i (tail false) or of
// tail i.  (constexpr: host and device code both call it.)
constexpr int64_t interleave_pos(bool tail, int64_t i, int64_t M, int64_t T) {
    if (tail) return M == 0 ? i : ((i + 1) * M + T - 1) / T + i;
    return T == 0 ? i : i + (i * T) / M;
}

// What the batched kernel needs of one SIC of a launch group (device memory).
struct SicView {
    const int32_t *sink_set_start;
    const int32_t *src_set_start;
    const int32_t *src_set_count;
    const int32_t *ilist;
    const nf_sink_plan *sink_plan;  // sinks are read from the slab through the pencil plans
    const float *sink_slab;         // position slab of SlabId (layout: aos)
    uint64_t sink_n;                // its particle count
    const void *sources4;           // padded source stream, float4, followed by 2 float4 boxes per block
    const void *box4;
    double *partial;                // 4 doubles per padded sink slot
    int32_t W, R;
    int32_t aos;                    // 1: posstruct (x,y,z) per particle; 0: PosXYZ X|Y|Z
    int32_t pad;
};

struct GPlanTotals;

// Device views of the plan arrays of one SIC.  A SIC planned on the device
// without a host round trip has its arrays laid out for upper bounds of its
// sizes: then dtot (non-null) holds the actual block and chunk counts in
// device memory, which the kernels read instead of nsink_blocks,
// nsrc_blocks and n_coadd_chunks (those are the bounds: grid sizes and
// buffer offsets), and dbad the launch group's count of planned cells that
// lie outside their slabs (nonzero: every kernel of the group returns at
// once, and the tasks fail).
struct DevPlan {
    const int32_t *sink_set_start;
    const int32_t *sink_set_count;
    const uint32_t *sink_idmax;
    const nf_sink_plan *sink_plan;
    const int32_t *src_set_start;
    const int32_t *src_set_count;
    const nf_source_plan *src_plan;
    const int32_t *parent;
    const int32_t *ilist;
    const float *yoff;
    const int32_t *src_parent;   // source pencil of each padded 64-source block
    const int32_t *coadd_chunk;  // (j*Nk + z, first particle) per 64-particle chunk of a sink cell
    int32_t nsink_sets, nsrc_sets, nsink_blocks, nsrc_blocks;
    int32_t W, R, cpd, Nk, j_width, j_low;
    int32_t n_coadd_chunks;
    int32_t max_src_blocks;  // padded blocks of the largest source pencil (INT32_MAX: not known)
    const GPlanTotals *dtot;
    const int32_t *dbad;
};

// The slabs of one SIC on the device.
struct DevSlabs {
    const float *pos[NF_MAX_WIDTH];  // position slab of SlabId-R+c
    uint64_t n[NF_MAX_WIDTH];
    int32_t aos;                     // 1: posstruct (x,y,z) per particle; 0: PosXYZ X|Y|Z
};

// The consumer's RescaleAndCoAddAcceleration, applied by the co-add kernels
// to each sink's float4 (on == 0: off).  far[3 (row - far_base) + k] is the
// far-field acceleration of slab row `row`; near_out (or null) receives the
// near-field record (xyz * inv_eps3, w) at row `row`.
struct KickArgs {
    const float *far;
    float *near_out;
    int64_t far_base;
    float rescale, inv_eps3;
    double density_rad2;
    int32_t on;
};

// Constants of the pair kernel, derived on the host from (softening, eps, b2).
struct PairConst {
    float eps;        // the reference's eps parameter
    float b2;         // FOF density radius^2 as given
    float r2add;      // added to r^2 before rsqrt: tau (spline forms) or e^2 (Plummer)
    float b2x;        // density threshold against that r^2: b2 + r2add
    float c1, c2, c3; // single spline, box-unit polynomial (-15 e^-4, 6 e^-5, 10 e^-3)
    float theta;      // tile density sum below which no pair is softened (< 0: none)
    double out_scale; // e^3 for the box-unit spline, 1 otherwise
    int32_t fast_ok;  // 1 when the unsoftened fast path may be used
};

int derive_pair_const(int softening, float eps, float b2, PairConst *pc);

// Batched-path tuning knobs.  Every shape gives bitwise the same partials:
// a sink's sum does not depend on which lane, wave or workgroup computes it.
struct TileShape {
    int sinks_per_lane;   // S: 1, 2 or 4
    int tile;             // TwoSum tile depth in sources (sinks on lanes)
    int rem_sinks;        // uniform sinks per pass in the remainder path
    int waves_per_cu;     // persistent grid size (0: from occupancy)
    int waves_per_wg;     // workgroup size in waves: 1, 2 or 4 (0: 4)
};

// Kernel launchers (kernels.hip).  All enqueue on `stream` and return a HIP
// error code (0 on success).
// Source pencils always; sink pencils only if sinks4 is non-null (the plain
// path; the batched kernel reads its sinks from the slab itself).  With
// batched_stream the sources take FP64 cell offsets and each cell's particles
// are ordered by x-bin (tighter block boxes); otherwise the reference's
// order and FLOAT offsets.
int launch_gather(const DevPlan &pl, const DevSlabs &sl, bool batched_stream, void *sinks4,
                  void *sources4, void *stream);
int launch_plain(const DevPlan &pl, int softening, const PairConst &pc, const void *sinks4,
                 const void *sources4, void *partial4, void *stream);
int launch_coadd_plain(const DevPlan &pl, const void *partial4, float *acc4, const KickArgs &kick, void *stream);
// One persistent launch over the work items of a group of SICs.  With
// dnitems (device memory) non-null the list holds *dnitems items and nitems
// is only its bound; with dbad non-null and *dbad != 0 the kernel does
// nothing.  counter names a 256-byte slot: the queue counter, the device
// list's length and, with census, four uint64 at byte 64 (kCensusBytes)
// receiving the pair interactions computed in each pair form (unsoftened,
// with density, softened) and, in the fourth, those of short tails (also
// counted by form), by a diagnostic build of the kernel.
constexpr int kCensusBytes = 64;
int launch_batched(const SicView *views, int cpd, int softening, const PairConst &pc, const TileShape &ts,
                   const WorkItem *items, int32_t nitems, int32_t *counter, void *stream,
                   const int32_t *dnitems = nullptr, const int32_t *dbad = nullptr, bool census = false);
// Bounding boxes follow a padded source stream of nsrc_blocks blocks.
inline void *box_after(void *sources4, int32_t nsrc_blocks) {
    return static_cast<char *>(sources4) + (size_t)nsrc_blocks * NF_BLOCK_SIZE * 16;
}
int launch_coadd_f64(const DevPlan &pl, const void *partial_d4, double out_scale, float *acc4,
                     const KickArgs &kick, void *stream);
// One SIC of a launch group, for the group-wide batched-path gather and
// co-add launches (device memory, beside the group's view table).
struct GroupSic {
    DevPlan pl;
    DevSlabs sl;
    KickArgs kick;
    void *sources4;      // padded source stream; its block boxes follow it
    const void *partial;  // FP64 partials, 4 per padded sink slot
    float *acc4;
    double out_scale;
};
// Waves of each SIC in the group-wide launches: the gather takes one per
// source pencil, the co-add one per 64-particle chunk (its bound, for a
// device-planned SIC).
// Waves of a SIC's binned source gather: one per source pencil and, when a
// pencil may have more than kGatherSegBlocks blocks, one per
// kGatherSegBlocks blocks of the stream (the pencil waves then fill their
// first kGatherSegBlocks blocks, these the blocks past them).
constexpr int32_t kGatherSegBlocks = 8;
constexpr bool gather_segmented(int32_t max_src_blocks) { return max_src_blocks > kGatherSegBlocks; }
constexpr int32_t gather_waves(int32_t nsrc_sets, int32_t nsrc_blocks, int32_t max_src_blocks) {
    return nsrc_sets + (gather_segmented(max_src_blocks) ? (nsrc_blocks + kGatherSegBlocks - 1) / kGatherSegBlocks : 0);
}
inline int32_t group_gather_waves(const DevPlan &pl) {
    return gather_waves(pl.nsrc_sets, pl.nsrc_blocks, pl.max_src_blocks);
}
inline int32_t group_coadd_waves(const DevPlan &pl) { return pl.n_coadd_chunks; }
// The batched-path source gathers of n SICs as one launch, then (a second
// call) their FP64 co-adds as one launch.  first[i] is the first wave of SIC
// i and first[n] the total (device memory, like g).
// segmented: some SIC of the group has gather_segmented(max_src_blocks).
int launch_gather_group(const GroupSic *g, const int32_t *first, int n, int total, int cpd, bool segmented,
                        void *stream);
int launch_coadd_group(const GroupSic *g, const int32_t *first, int n, int total, void *stream);
// The single-spline pair-math probe (the batched tile loop alone, one pair
// form: 0 unsoftened without density, 1 with density, 2 the full spline) at
// a tile of S sinks per lane and an occupancy of waves_per_simd (4, 6 or 8):
// pair interactions per second over the whole GPU.  Synchronous.
int run_pair_probe(const PairConst &pc, int sinks_per_lane, int mode, int waves_per_simd, double *rate);
int device_cu_count();

// ---- Device-side SIC planning (gpu_plan.hip) -----------------------------
// Input: the cellinfo slabs of SlabId-R..SlabId+R in device memory (cpd*cpd
// cells each, cell (y, z) at y*cpd + z) and the SIC's sink rows.
struct GPlanIn {
    int32_t cpd, R, W, jw, j_low;
    const nf_cellinfo *ci[NF_MAX_WIDTH];
    uint64_t nslab[NF_MAX_WIDTH];  // particles per slab: every planned cell must lie inside its slab
    uint64_t split_cost[3];        // per S = 1, 2, 4: register tiles above this cost become one-wave tiles
};
// Totals of phase A (device; atomically accumulated, so zeroed first).
struct GPlanTotals {
    uint64_t sink_total, source_total, direct_total, padded_direct_total;
    uint64_t sink_blocks, source_blocks, chunks;
    uint64_t items[3];  // work items for S = 1, 2, 4
    uint64_t occ[4];    // sinks in pencils of <64, 64-255, 256-1023, >=1024 sinks
    uint64_t bad;       // planned cells with a negative count or outside their slab
};
// Where phase B writes (the plan buffer at the SIC's PlanLayout offsets).
// items[t]: the work list for sinks_per_lane 1 << t, or null if not wanted;
// hist and cursor (kPlanLists x kPlanBuckets int32 each: per tile, its tiles'
// buckets, then per tile its short tails' buckets; zeroed by phase A of the
// group's first SIC) are shared by the SICs of a launch group, whose lists
// are merged; part holds gpu_plan_part_ints() int32 of this SIC's block sums.
struct GPlanOut {
    int32_t *sink_start, *sink_count, *parent, *src_start, *src_count, *src_parent, *ilist, *chunks;
    uint32_t *idmax;
    nf_sink_plan *sink_plan;
    nf_source_plan *src_plan;
    float *yoff;
    WorkItem *items[3];
    int32_t *hist, *cursor, *part;
    const int32_t *group;  // the group's shared area (null: one SIC planned alone)
    int32_t sic;  // index of the SIC in its launch group (the work items' kind_sic)
};
constexpr int kPlanBuckets = 512;                   // 1/8-octave cost buckets of a device-built work list
constexpr int kPlanLists = 6;                       // tiles of S = 1, 2, 4, then short tails of S = 1, 2, 4
// The area a launch group's SICs share: hist, cursor, then the group's work
// items per tile (S = 1, 2, 4) and its count of bad cells.
constexpr int kPlanGroupItems = 2 * kPlanLists * kPlanBuckets;
constexpr int kPlanGroupBad = kPlanGroupItems + 3;
constexpr int kPlanSharedInts = kPlanGroupBad + 1;
size_t gpu_plan_part_ints(int nsink_sets, int nsrc_sets);
// Phase A: the SIC's totals (sizes), also added into the group's shared
// area `group` (its items per tile and bad cells); `zero` (the shared area,
// or null) is zeroed first, by the group's first SIC.  Phase B, with the
// plan laid out: B1 for every SIC of the group, then B2 for every SIC; both
// do nothing when the group has a bad cell.
int launch_gpu_plan_a(const GPlanIn &in, GPlanTotals *tot, int32_t *zero, int32_t *group, void *stream);
int launch_gpu_plan_b1(const GPlanIn &in, const GPlanOut &out, void *stream);
int launch_gpu_plan_b2(const GPlanIn &in, const GPlanOut &out, void *stream);
// The same phases for a whole launch group, one launch each: ins and outs
// are the n SICs' arrays in device memory, tot their totals 256 bytes apart,
// shared the group's shared area, first[s] SIC s's first workgroup in the
// phase's grid of `grid` workgroups (gpu_plan_grid_a/b per SIC).  phase 0
// zeroes the totals and the shared area, then counts (A); 1 and 2 are B1 and
// B2.
int gpu_plan_grid_a(const GPlanIn &in);
int gpu_plan_grid_b(const GPlanIn &in);
int launch_gpu_plan_group(int phase, const GPlanIn *ins, const GPlanOut *outs, GPlanTotals *tot, int32_t *shared,
                          const int32_t *first, int n, int grid, void *stream);
// The occupancy target (waves per SIMD) the batched kernel is built with for
// a pair form and tile; the third template argument of k_batched.
int batched_kernel_waves(int softening, int sinks_per_lane);

}  // namespace nf

#endif  // NF_INTERNAL_H
