// Host runtime of the near-field direct module: the C ABI of nf_direct.h.
//
// The MI355X-native counterpart of GPUControl.cu + PencilTask.cu: a pool of
// worker threads (DirectBPD per GPU), each owning one HIP stream, a growable
// device arena and pinned staging.  A task is one SetInteractionCollection:
// upload its plan arrays (and, in host mode, its slabs), gather the pencils on
// the device, run the direct kernel, co-add the W partials per sink on the
// device, and (host mode) copy the sink rows of the acceleration slab back.
// Completion is an atomic flag the caller polls (CheckCompletion).

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <sstream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "nf_direct.h"
#include "nf_internal.h"

namespace nf {

namespace {
thread_local std::string g_last_error;
}

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

#define NF_HIP(call)                                                                                      \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return fail(NF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));                    \
    } while (0)

int derive_pair_const(int softening, float eps, float b2, PairConst *pc) {
    std::memset(pc, 0, sizeof(*pc));
    pc->eps = eps;
    pc->b2 = b2;
    pc->out_scale = 1.0;
    pc->theta = -1.f;
    if (!(eps > 0.f) || !std::isfinite(eps)) return fail(NF_ERR_INVALID, "eps must be finite and positive");
    if (!(b2 >= 0.f) || !(b2 < 0.5f)) return fail(NF_ERR_INVALID, "b2 must be in [0, 0.5)");
    switch (softening) {
        case NF_SOFT_SINGLE_SPLINE: {
            // Box-unit form: with e = eps^-1/2, g(r)/e^3 = min(r^-3, (c1/r + c2) r^2 + c3).
            const double e = (double)eps;
            pc->r2add = 1e-24f;  // keeps r^-3 finite for the self pair; below any resolvable FP32 separation^2
            pc->c1 = (float)(-15.0 * e * e);
            pc->c2 = (float)(6.0 * e * e * std::sqrt(e));
            pc->c3 = (float)(10.0 * e * std::sqrt(e));
            pc->out_scale = 1.0 / (e * std::sqrt(e));
            if (!std::isfinite(pc->c1) || !std::isfinite(pc->c2) || !std::isfinite(pc->c3) || pc->c1 == 0.f)
                return fail(NF_ERR_INVALID, "single spline eps out of the FP32 box-unit range");
            pc->fast_ok = 1;
            break;
        }
        case NF_SOFT_PLUMMER: pc->r2add = eps; break;
        case NF_SOFT_CUBIC_SPLINE: pc->r2add = 0.f; break;
        case NF_SOFT_CUBIC_PLUMMER: pc->r2add = 1e-16f; break;  // TAU2, DirectGPUKernels.cuh:113
        default: return fail(NF_ERR_INVALID, "unknown softening");
    }
    pc->b2x = b2 > 0.f ? b2 + pc->r2add : 0.f;
    return NF_OK;
}

namespace {

using clk = std::chrono::steady_clock;

struct Task {
    nf_sic sic;
    // nf_submit_slab: the plan is computed on the device from the cellinfo
    // slabs; sic holds the scalars (and, after phase A, the totals) only.
    struct DevicePlanned {
        bool on = false;
        const nf_cellinfo *cellinfo[NF_MAX_WIDTH] = {};
        uint64_t r0 = 0, r1 = 0;  // the SIC's sink rows in its slab
        GPlanTotals tot{};        // phase A's sizes, read back with the group's results
        uint64_t chunks_bound = 0, items_bound = 0;  // bounds the plan is laid out for
        uint64_t occ[4] = {};     // sinks in pencils of <64, 64-255, 256-1023, >=1024 sinks
        uint64_t interactions = 0;  // DirectTotal from the cellinfo counts (nf_count_interactions)
        int32_t max_src_blocks = INT32_MAX;  // padded blocks of the largest source pencil, from the same sums
    } dp;
    std::atomic<int> state{0};  // 0 queued/running, 1 done, <0 failed
    std::string error;
    nf_task_stats stats{};
    std::mutex m;
    std::condition_variable cv;
};

// A device allocation that only grows.
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return NF_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 4 + (1 << 20);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(NF_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        cap = want;
        return NF_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return NF_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
        const size_t want = bytes + bytes / 4 + (1 << 20);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(NF_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
        }
        cap = want;
        return NF_OK;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct Worker {
    int id = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t plan_ready = nullptr;  // the plan image upload (copy stream) is done
    hipStream_t pstream = nullptr;     // device planning (nf_submit_slab)
    bool own_streams = true;           // stream and pstream are this worker's (else the runtime's)
    hipEvent_t gplan_ready = nullptr;  // the device-planned arrays and work lists are written
    DevBuf gscratch;                   // the device planner's totals and scratch
    PinBuf pin_tot;                    // phase A's totals, read back
    PinBuf pin_gplan;                  // the group planning launches' inputs, staged
    PinBuf pin_census;                 // the batched launch's interactions per pair form, read back
    hipGraphExec_t gexec = nullptr;  // the last group's executable graph (use_graphs)
    uint64_t graphs_instantiated = 0, graphs_updated = 0;
    DevBuf plan, work, slabs;
    PinBuf pin_plan, pin_acc, pin_far;
    std::thread th;
    bool ready = false;
};

struct Runtime {
    nf_setup_params params{};
    uint64_t max_buffer_bytes = 0;
    int ngpu = 0;
    std::vector<std::unique_ptr<Worker>> workers;
    std::deque<Task *> queue;
    std::mutex qm;
    std::condition_variable qcv;
    bool stopping = false;
    int inflight = 0;   // groups taken by workers and not finished (under qm)
    bool warm = false;  // work has been flowing since the queue was last empty (under qm)
    TileShape shape{2, 64, 4, 0, 4};
    std::vector<TileShape> candidates;  // the autotuner's search (empty: the shape is fixed)
    int max_group = 8;
    uint64_t group_interactions = 12000000000ull;
    bool autotune = false;
    std::mutex tune_m;
    struct TuneEntry {
        TileShape shape;
        std::vector<double> trial_ms;  // per candidate, in candidate order
        uint64_t interactions = 0;
        int uses = 0;
    };
    std::map<std::string, TuneEntry> tune;  // occupancy-histogram key -> tuned shape
    // The GPU throughput timer (GPUControl.cu:658-672): wall time during which
    // at least one launch group is on the GPU, and totals of completed tasks.
    struct Throughput {
        std::mutex m;
        int active = 0;
        std::chrono::steady_clock::time_point since;
        double active_s = 0.0;
        uint64_t direct = 0, padded = 0, to_dev = 0, from_dev = 0, tasks = 0, groups = 0;
    } tp;
    hipStream_t shared = nullptr;   // when gpu_streams == 1: the one stream all workers use
    hipStream_t pstream = nullptr;  // when gpu_streams == 1: the workers' one planning stream
    std::mutex enqueue;            // orders whole task sequences on the shared stream

    // Host-mode slabs resident in HBM.  A slot is reusable by address only
    // inside the stream session it was uploaded in (nf_stream_begin/end).
    struct Slot {
        const void *host = nullptr;
        uint64_t bytes = 0;
        uint64_t epoch = 0;
        bool valid = false;
        bool filling = false;  // an upload into it is being enqueued
        int refs = 0;
        uint64_t last = 0;
        DevBuf buf;
        hipEvent_t ready = nullptr;
        hipEvent_t t0 = nullptr, t1 = nullptr;  // brackets the last upload
    };
    static constexpr int kStages = 4;  // pinned staging buffers for pageable host slabs
    struct SlabCache {
        std::mutex m;
        std::condition_variable cv;  // a slot finished filling, or a staging buffer came free
        bool active = false;
        uint64_t epoch = 0, clock = 0;
        std::vector<std::unique_ptr<Slot>> slots;
        hipStream_t up = nullptr, down = nullptr;  // uploads; acceleration copy-backs
        std::mutex up_enqueue;                     // keeps each upload's event bracket contiguous
        PinBuf stage[kStages];
        hipEvent_t stage_done[kStages] = {};
        bool stage_busy[kStages] = {};
    } cache;
};

std::mutex g_rt_mutex;
Runtime *g_rt = nullptr;

// The runtime's streams (compute, upload, download, planning) live for the
// whole process: made by the first nf_setup that needs them and reused by
// every later one (under g_rt_mutex).  HIP maps a new stream onto one of its
// hardware queues when the stream is made; streams destroyed and made again
// by a second nf_setup can land two busy streams on one queue, where they
// serialize (a second driver in one process measured 3% slower).
struct StreamPool {
    hipStream_t compute = nullptr, up = nullptr, down = nullptr, plan = nullptr;
};
StreamPool g_streams;

hipError_t pooled_stream(hipStream_t *slot, hipStream_t *out) {
    if (!*slot) {
        const hipError_t e = hipStreamCreateWithFlags(slot, hipStreamNonBlocking);
        if (e != hipSuccess) {
            *slot = nullptr;
            return e;
        }
    }
    *out = *slot;
    return hipSuccess;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Where one SIC's plan arrays sit inside a group's plan upload.
struct PlanLayout {
    size_t o_sink_start, o_sink_count, o_idmax, o_sink_plan, o_src_start, o_src_count, o_src_plan, o_parent,
        o_ilist, o_yoff, o_src_parent, o_chunks, total;
    int32_t n_chunks;
};

// 64-particle chunks of the sink cells of rows [j_low, j_high): the co-add's
// work list.  Pencil k = z starts with cell z, so its first cell is cell z.
int32_t count_coadd_chunks(const nf_sic &s) {
    int64_t n = 0;
    for (int k = 0; k < s.NSinkSets; k++) n += (s.SinkPlan[k].cell[0].N + NF_BLOCK_SIZE - 1) / NF_BLOCK_SIZE;
    return (int32_t)n;
}

PlanLayout plan_layout(const nf_sic &s, size_t base, int32_t n_chunks) {
    PlanLayout L{};
    size_t o = base;
    auto take = [&](size_t bytes) {
        size_t r = o;
        o = align_up(o + bytes, 256);
        return r;
    };
    L.o_sink_start = take(4 * (size_t)s.NSinkSets);
    L.o_sink_count = take(4 * (size_t)s.NSinkSets);
    L.o_idmax = take(4 * (size_t)s.NSinkSets);
    L.o_sink_plan = take(sizeof(nf_sink_plan) * (size_t)s.NSinkSets);
    L.o_src_start = take(4 * (size_t)s.NSourceSets);
    L.o_src_count = take(4 * (size_t)s.NSourceSets);
    L.o_src_plan = take(sizeof(nf_source_plan) * (size_t)s.NSourceSets);
    L.o_parent = take(4 * (size_t)s.NSinkBlocks);
    L.o_ilist = take(4 * (size_t)s.InteractionCount);
    L.o_yoff = take(4 * (size_t)s.InteractionCount);
    L.o_src_parent = take(4 * (size_t)s.NSourceBlocks);
    L.n_chunks = n_chunks;
    L.o_chunks = take(8 * (size_t)L.n_chunks);
    L.total = o - base;
    return L;
}

// Where the device planner writes a SIC's arrays: the plan buffer `base` at
// the offsets of L; no work lists (the caller names them).
GPlanOut plan_out(char *base, const PlanLayout &L, int sic, int32_t *shared, int32_t *part) {
    GPlanOut o{};
    o.sink_start = reinterpret_cast<int32_t *>(base + L.o_sink_start);
    o.sink_count = reinterpret_cast<int32_t *>(base + L.o_sink_count);
    o.idmax = reinterpret_cast<uint32_t *>(base + L.o_idmax);
    o.sink_plan = reinterpret_cast<nf_sink_plan *>(base + L.o_sink_plan);
    o.src_start = reinterpret_cast<int32_t *>(base + L.o_src_start);
    o.src_count = reinterpret_cast<int32_t *>(base + L.o_src_count);
    o.src_plan = reinterpret_cast<nf_source_plan *>(base + L.o_src_plan);
    o.parent = reinterpret_cast<int32_t *>(base + L.o_parent);
    o.ilist = reinterpret_cast<int32_t *>(base + L.o_ilist);
    o.yoff = reinterpret_cast<float *>(base + L.o_yoff);
    o.src_parent = reinterpret_cast<int32_t *>(base + L.o_src_parent);
    o.chunks = reinterpret_cast<int32_t *>(base + L.o_chunks);
    o.hist = shared;
    o.cursor = shared ? shared + kPlanLists * kPlanBuckets : nullptr;
    o.part = part;
    o.sic = sic;
    return o;
}

// A host-planned SIC's arrays into the plan image `hp` at the offsets of L,
// with the two lists derived from them: the source pencil of each padded
// 64-source block (-1: a block of no pencil, skipped) and the co-add chunks.
void fill_plan_image(const nf_sic &s, const PlanLayout &L, char *hp) {
    std::memcpy(hp + L.o_sink_start, s.SinkSetStart, 4 * (size_t)s.NSinkSets);
    std::memcpy(hp + L.o_sink_count, s.SinkSetCount, 4 * (size_t)s.NSinkSets);
    std::memcpy(hp + L.o_idmax, s.SinkSetIdMax, 4 * (size_t)s.NSinkSets);
    std::memcpy(hp + L.o_sink_plan, s.SinkPlan, sizeof(nf_sink_plan) * (size_t)s.NSinkSets);
    std::memcpy(hp + L.o_src_start, s.SourceSetStart, 4 * (size_t)s.NSourceSets);
    std::memcpy(hp + L.o_src_count, s.SourceSetCount, 4 * (size_t)s.NSourceSets);
    std::memcpy(hp + L.o_src_plan, s.SourcePlan, sizeof(nf_source_plan) * (size_t)s.NSourceSets);
    if (s.NSinkBlocks) std::memcpy(hp + L.o_parent, s.SinkBlockParentPencil, 4 * (size_t)s.NSinkBlocks);
    std::memcpy(hp + L.o_ilist, s.SinkSourceInteractionList, 4 * (size_t)s.InteractionCount);
    std::memcpy(hp + L.o_yoff, s.SinkSourceYOffset, 4 * (size_t)s.InteractionCount);
    int32_t *sp = reinterpret_cast<int32_t *>(hp + L.o_src_parent);
    std::fill(sp, sp + s.NSourceBlocks, -1);
    for (int k = 0; k < s.NSourceSets; k++) {
        const int b0 = s.SourceSetStart[k] / NF_BLOCK_SIZE;
        const int nb = (s.SourceSetCount[k] + NF_BLOCK_SIZE - 1) / NF_BLOCK_SIZE;
        for (int b = 0; b < nb; b++) sp[b0 + b] = k;
    }
    int32_t *ch = reinterpret_cast<int32_t *>(hp + L.o_chunks);
    int64_t q = 0;
    for (int k = 0; k < s.NSinkSets; k++)
        for (int f = 0; f < s.SinkPlan[k].cell[0].N; f += NF_BLOCK_SIZE) {
            ch[2 * q] = k;
            ch[2 * q + 1] = f;
            q++;
        }
}

// The device views of a SIC's plan arrays at the offsets of L in `pd`.
DevPlan dev_plan(const nf_sic &s, const PlanLayout &L, char *pd) {
    DevPlan pl{};
    pl.sink_set_start = reinterpret_cast<const int32_t *>(pd + L.o_sink_start);
    pl.sink_set_count = reinterpret_cast<const int32_t *>(pd + L.o_sink_count);
    pl.sink_idmax = reinterpret_cast<const uint32_t *>(pd + L.o_idmax);
    pl.sink_plan = reinterpret_cast<const nf_sink_plan *>(pd + L.o_sink_plan);
    pl.src_set_start = reinterpret_cast<const int32_t *>(pd + L.o_src_start);
    pl.src_set_count = reinterpret_cast<const int32_t *>(pd + L.o_src_count);
    pl.src_plan = reinterpret_cast<const nf_source_plan *>(pd + L.o_src_plan);
    pl.parent = reinterpret_cast<const int32_t *>(pd + L.o_parent);
    pl.ilist = reinterpret_cast<const int32_t *>(pd + L.o_ilist);
    pl.yoff = reinterpret_cast<const float *>(pd + L.o_yoff);
    pl.src_parent = reinterpret_cast<const int32_t *>(pd + L.o_src_parent);
    pl.coadd_chunk = reinterpret_cast<const int32_t *>(pd + L.o_chunks);
    pl.n_coadd_chunks = L.n_chunks;
    pl.nsink_sets = s.NSinkSets;
    pl.nsrc_sets = s.NSourceSets;
    pl.nsink_blocks = s.NSinkBlocks;
    pl.nsrc_blocks = s.NSourceBlocks;
    // (a SIC planned on the device has no host counts: its kernels read the
    // largest pencil from the device totals)
    pl.max_src_blocks = INT32_MAX;
    if (s.SourceSetCount) {
        pl.max_src_blocks = 0;
        for (int k = 0; k < s.NSourceSets; k++)
            pl.max_src_blocks = std::max(pl.max_src_blocks, (s.SourceSetCount[k] + NF_BLOCK_SIZE - 1) / NF_BLOCK_SIZE);
    }
    pl.W = s.nfwidth;
    pl.R = s.nfradius;
    pl.cpd = s.cpd;
    pl.Nk = s.Nk;
    pl.j_width = s.j_width;
    pl.j_low = s.j_low;
    return pl;
}

// The scalar members of a SIC of slab `slab`, rows [j_low, j_high), as the
// SetInteractionCollection constructor sets them (PencilOnPencil.cc:180-215)
// for the non-PARALLEL layout.
void sic_header(nf_sic &s, int cpd, int R, int slab, int j_low, int j_high, float eps, float b2, int softening) {
    std::memset(&s, 0, sizeof(s));
    s.SlabId = slab;
    s.j_low = j_low;
    s.j_high = j_high;
    s.k_low = 0;
    s.k_high = cpd;
    s.Nk = cpd;
    s.j_width = j_high - j_low;
    s.cpd = cpd;
    s.nfradius = R;
    s.nfwidth = 2 * R + 1;
    s.eps = eps;
    s.b2 = b2;
    s.softening = softening;
    s.NSinkSets = s.j_width * s.Nk;
    s.NSourceSets = (s.j_width + s.nfwidth - 1) * s.Nk;
    s.InteractionCount = s.nfwidth * s.NSinkSets;
}

// The device planner's totals into the SIC's members.
int apply_totals(nf_sic &s, const GPlanTotals &t) {
    if (t.bad)
        return fail(NF_ERR_INVALID, std::to_string(t.bad) + " planned cells have a negative count or lie outside their slab");
    if (t.sink_blocks * NF_BLOCK_SIZE >= (uint64_t)INT32_MAX)
        return fail(NF_ERR_CAPACITY, "device plan: padded sink count overflows int32");
    if (t.source_blocks * NF_BLOCK_SIZE >= (uint64_t)INT32_MAX)
        return fail(NF_ERR_CAPACITY, "device plan: padded source count overflows int32");
    s.NSinkBlocks = (int32_t)t.sink_blocks;
    s.NSourceBlocks = (int32_t)t.source_blocks;
    s.DirectTotal = t.direct_total;
    s.PaddedDirectTotal = t.padded_direct_total;
    s.SinkTotal = t.sink_total;
    s.SourceTotal = t.source_total;
    s.PaddedSinkTotal = t.sink_blocks * NF_BLOCK_SIZE;
    s.PaddedSourceTotal = t.source_blocks * NF_BLOCK_SIZE;
    return NF_OK;
}

// Device working set of one SIC: sink stream, source stream + boxes, partials.
struct WorkLayout {
    size_t o_sinks, o_src, o_part, o_far, total;
};

WorkLayout work_layout(const nf_sic &s, int path, size_t base) {
    WorkLayout L{};
    const size_t psink = (size_t)s.NSinkBlocks * NF_BLOCK_SIZE;
    const size_t psrc = (size_t)s.NSourceBlocks * NF_BLOCK_SIZE;
    L.o_sinks = base;  // the plain path's sink stream (the batched kernel reads the slab)
    L.o_src = align_up(L.o_sinks + (path == NF_PATH_BATCHED ? 0 : 16 * psink), 256);
    L.o_part = align_up(L.o_src + 16 * psrc + 32 * (size_t)s.NSourceBlocks, 256);
    size_t end = align_up(L.o_part + psink * (path == NF_PATH_BATCHED ? 32 : 16), 256);
    L.o_far = end;  // host mode with a fused kick: the sink rows' far-field accelerations
    if (s.kick && s.mem == NF_MEM_HOST) end = align_up(end + 12 * (size_t)s.Nslab[s.nfradius], 256);
    L.total = end - base;
    return L;
}

// Rows [r0, r1) of the sink slab a SIC covers.  Cells are y-major, so rows
// j_low..j_high-1 are contiguous: from cell (j_low, 0) -- the first cell of
// pencil 0 -- to the end of cell (j_high-1, cpd-1) -- the first cell of the
// last pencil.
void sink_rows(const nf_sic &s, uint64_t *r0, uint64_t *r1) {
    *r0 = s.SinkPlan[0].cell[0].start;
    const nf_sink_plan &last = s.SinkPlan[(size_t)(s.j_width - 1) * s.Nk + (s.Nk - 1)];
    *r1 = (uint64_t)last.cell[0].start + (uint64_t)last.cell[0].N;
}
void task_rows(const Task &t, uint64_t *r0, uint64_t *r1) {
    if (t.dp.on) {
        *r0 = t.dp.r0;
        *r1 = t.dp.r1;
    } else {
        sink_rows(t.sic, r0, r1);
    }
}

// Device bytes a SIC adds to a group (plan + work + host-mode slabs).
size_t task_bytes(const nf_sic &s, int path, int32_t n_chunks) {
    size_t b = plan_layout(s, 0, n_chunks).total + work_layout(s, path, 0).total + sizeof(SicView) + 256;
    b += ((size_t)s.NSinkBlocks + 2 * (size_t)s.NSinkSets) * sizeof(WorkItem);  // at most one item per block + tails
    if (s.mem == NF_MEM_HOST) {
        for (int c = 0; c < s.nfwidth; c++) b += align_up(12 * (size_t)s.Nslab[c], 256);
        b += align_up(16 * (size_t)s.Nslab[s.nfradius], 256);
        if (s.kick && s.kick->NearAccOut) b += align_up(16 * (size_t)s.Nslab[s.nfradius], 256);
    }
    return b;
}

// A device-planned SIC before phase A: upper bounds of its sizes from its
// slab sizes (every pencil adds at most one partly filled block; a source
// row that wraps round the box appears twice).
// Sets s's block counts to those bounds; *chunks and *items get the bounds
// of its co-add chunks and of its work items (at most one per padded sink
// block plus one tail per pencil, for every tile).  The device planner's
// plan and working set are laid out for these.
void dp_bounds(nf_sic &s, uint64_t *chunks, uint64_t *items) {
    const uint64_t W = (uint64_t)s.nfwidth, cap = (uint64_t)INT32_MAX / NF_BLOCK_SIZE;
    uint64_t window = 0;
    for (int c = 0; c < s.nfwidth; c++) window += s.Nslab[c];
    const uint64_t wraps = ((uint64_t)(s.j_width + s.nfwidth - 1) + s.cpd - 1) / (uint64_t)s.cpd;
    s.NSinkBlocks = (int32_t)std::min<uint64_t>(cap, W * s.Nslab[s.nfradius] / NF_BLOCK_SIZE + s.NSinkSets);
    s.NSourceBlocks = (int32_t)std::min<uint64_t>(cap, wraps * window / NF_BLOCK_SIZE + s.NSourceSets);
    *chunks = std::min<uint64_t>(cap, s.Nslab[s.nfradius] / NF_BLOCK_SIZE + s.NSinkSets);
    *items = (uint64_t)s.NSinkBlocks + (uint64_t)s.NSinkSets;
}
size_t dp_task_bytes(const nf_sic &s0, int path) {
    nf_sic s = s0;
    uint64_t chunks, items;
    dp_bounds(s, &chunks, &items);
    return task_bytes(s, path, (int32_t)chunks);
}

// The sink-pencil occupancy histogram of a device-planned SIC, from the
// central slab's cellinfo on the host (a running sum over each row): the
// autotuner's key, before any plan exists.
void dp_occupancy(const Task &t, uint64_t occ[4]) {
    const nf_sic &s = t.sic;
    const nf_cellinfo *ci = t.dp.cellinfo[s.nfradius];
    for (int b = 0; b < 4; b++) occ[b] = 0;
    for (int j = s.j_low; j < s.j_high; j++) {
        const nf_cellinfo *row = ci + (size_t)j * s.cpd;
        int64_t n = 0;
        for (int c = 0; c < s.nfwidth; c++) n += std::max(row[c % s.cpd].count, 0);
        for (int z = 0; z < s.cpd; z++) {
            occ[n < 64 ? 0 : n < 256 ? 1 : n < 1024 ? 2 : 3] += (uint64_t)n;
            n += std::max(row[(z + s.nfwidth) % s.cpd].count, 0) - std::max(row[z].count, 0);
        }
    }
}
size_t queued_bytes(const Task &t, int path) {
    return t.dp.on ? dp_task_bytes(t.sic, path) : task_bytes(t.sic, path, count_coadd_chunks(t.sic));
}
uint64_t queued_interactions(const Task &t) { return t.dp.on ? t.dp.interactions : t.sic.DirectTotal; }

// The 1/8-octave bucket of an item's cost (sinks x sources), longest first:
// bucket 0 holds the most expensive items, kItemBuckets - 1 the cheapest.
constexpr int kItemBuckets = 1024;
uint16_t cost_bucket(uint64_t cost) {
    const uint64_t b = cost ? (uint64_t)std::floor(8.0 * std::log2((double)cost)) : 0;
    return (uint16_t)(kItemBuckets - 1 - std::min<uint64_t>(b, kItemBuckets - 1));
}
// Counting-sort key of an item: its bucket among the tiles, or kItemBuckets
// plus its bucket among the short tails.
int item_key(const WorkItem &w, uint16_t bucket) { return (int)bucket + ((w.kind_sic & 3) == 1 ? kItemBuckets : 0); }

// The batched work list of one SIC, appended in sink-pencil order z then y
// (pencils adjacent in y share 2R of their 2R+1 source pencils), with each
// item's cost bucket: per sink pencil, full register tiles of 64*S sinks,
// then the tail.  A pencil whose register tile would cost more than
// max_tile_cost (sinks x sources) takes one-wave tiles of 64 sinks instead:
// in a launch with little work per wave, one heavy tile would otherwise
// outlast every other wave.  Each sink's arithmetic is the same in both.
constexpr uint64_t kSplitMinCost = 1u << 18;  // a tile below this ends in well under a millisecond

// The cost above which a register tile of S sinks per lane is split, for a
// launch of `interactions` in all: a quarter of the work per wave of its
// persistent grid (CUs x 4 SIMDs x the kernel's waves per SIMD), at least
// kSplitMinCost.  (A quarter measured 0.5% faster than the whole share on a
// clustered box, and no different where no tile is that heavy.)
uint64_t tile_split_cost(int softening, int S, uint64_t interactions) {
    if (S <= 1) return UINT64_MAX;
    const uint64_t waves = (uint64_t)std::max(1, device_cu_count()) * 4 *
                           (uint64_t)std::max(1, batched_kernel_waves(softening, S));
    return std::max<uint64_t>(kSplitMinCost, interactions / waves / 4);
}
// Calls emit(item, bucket) for each work item of the SIC, in list order.
template <typename Emit>
void for_each_item(const nf_sic &s, int S, int sic, uint64_t max_tile_cost, Emit &&emit) {
    const int tile = NF_BLOCK_SIZE * S;
    const int W = s.nfwidth;
    for (int z = 0; z < s.Nk; z++)
        for (int j = 0; j < s.j_width; j++) {
            const int k = j * s.Nk + z;
            const int n = s.SinkSetCount[k];
            if (n == 0) continue;
            uint64_t nsrc = 0;
            for (int y = 0; y < W; y++) nsrc += (uint64_t)s.SourceSetCount[s.SinkSourceInteractionList[W * k + y]];
            int f = 0;
            if (S > 1 && (uint64_t)tile * nsrc > max_tile_cost) {
                const uint16_t b = cost_bucket((uint64_t)NF_BLOCK_SIZE * nsrc);
                for (; f + NF_BLOCK_SIZE <= n; f += NF_BLOCK_SIZE) emit(WorkItem{k, f, NF_BLOCK_SIZE, 2 + 4 * sic}, b);
            } else if (n >= tile) {
                const uint16_t b = cost_bucket((uint64_t)tile * nsrc);
                for (; f + tile <= n; f += tile) emit(WorkItem{k, f, tile, 0 + 4 * sic}, b);
            }
            // The tail: one-wave tiles while they keep at least kTailLanes
            // lanes busy, then the rest with sources on the lanes.
            if (n - f >= kTailLanes) {
                const uint16_t b = cost_bucket((uint64_t)NF_BLOCK_SIZE * nsrc);
                while (n - f >= kTailLanes) {
                    const int m = std::min(NF_BLOCK_SIZE, n - f);
                    emit(WorkItem{k, f, m, 2 + 4 * sic}, b);
                    f += m;
                }
            }
            if (f < n) emit(WorkItem{k, f, n - f, 1 + 4 * sic}, cost_bucket((uint64_t)(n - f) * nsrc));
        }
}

void build_items(const nf_sic &s, int S, int sic, uint64_t max_tile_cost, std::vector<WorkItem> &items,
                 std::vector<uint16_t> &bucket) {
    for_each_item(s, S, sic, max_tile_cost, [&](const WorkItem &w, uint16_t b) {
        items.push_back(w);
        bucket.push_back(b);
    });
}

int validate(const nf_sic &s) {
    if (s.nfradius < 0 || s.nfradius > NF_MAX_RADIUS || s.nfwidth != 2 * s.nfradius + 1)
        return fail(NF_ERR_INVALID, "nfradius/nfwidth out of range");
    if (s.cpd < s.nfwidth || (s.cpd % 2) == 0) return fail(NF_ERR_INVALID, "cpd must be odd and >= 2R+1");
    if (s.Nk != s.cpd || s.k_low != 0 || s.k_high != s.cpd)
        return fail(NF_ERR_INVALID, "only the non-PARALLEL layout (Nk = cpd) is supported");
    if (s.j_low < 0 || s.j_high > s.cpd || s.j_low >= s.j_high || s.j_width != s.j_high - s.j_low)
        return fail(NF_ERR_INVALID, "bad sink row range");
    if (s.NSinkSets != s.j_width * s.Nk || s.NSourceSets != (s.j_width + s.nfwidth - 1) * s.Nk ||
        s.InteractionCount != s.nfwidth * s.NSinkSets)
        return fail(NF_ERR_INVALID, "inconsistent set counts");
    if (!s.SinkSetStart || !s.SinkSetCount || !s.SinkSetIdMax || !s.SinkPlan || !s.SourceSetStart ||
        !s.SourceSetCount || !s.SourcePlan || !s.SinkSourceInteractionList || !s.SinkSourceYOffset ||
        (s.NSinkBlocks > 0 && !s.SinkBlockParentPencil))
        return fail(NF_ERR_INVALID, "missing plan array");
    if (!s.SinkPosSlab || !s.SinkAccSlab) return fail(NF_ERR_INVALID, "missing slab pointer");
    if (s.pos_layout != NF_POS_XYZ && s.pos_layout != NF_POS_AOS) return fail(NF_ERR_INVALID, "bad pos_layout");
    if (s.kick && (!s.kick->FarAccSlab || !std::isfinite(s.kick->rescale) || !std::isfinite(s.kick->inv_eps3) ||
                   !std::isfinite(s.kick->density_rad2)))
        return fail(NF_ERR_INVALID, "kick fusion needs a far-field slab and finite constants");
    for (int c = 0; c < s.nfwidth; c++)
        if (!s.SourcePosSlab[c]) return fail(NF_ERR_INVALID, "missing source slab pointer");
    if (s.mem != NF_MEM_HOST && s.mem != NF_MEM_DEVICE) return fail(NF_ERR_INVALID, "bad mem kind");
    // Every cell a plan names must lie inside its slab; the device gathers
    // index the slabs with these starts and counts.
    for (int k = 0; k < s.NSinkSets; k++) {
        int64_t tot = 0;
        for (int c = 0; c < s.nfwidth; c++) {
            const nf_cell_plan &cp = s.SinkPlan[k].cell[c];
            if (cp.N < 0 || (uint64_t)cp.start + (uint64_t)cp.N > s.Nslab[s.nfradius])
                return fail(NF_ERR_INVALID, "sink plan cell outside its slab");
            tot += cp.N;
        }
        if (tot != s.SinkSetCount[k]) return fail(NF_ERR_INVALID, "SinkSetCount disagrees with SinkPlan");
        if (s.SinkSetStart[k] % NF_BLOCK_SIZE != 0 || s.SinkSetStart[k] < 0 ||
            (int64_t)s.SinkSetStart[k] + (tot + NF_BLOCK_SIZE - 1) / NF_BLOCK_SIZE * NF_BLOCK_SIZE >
                (int64_t)s.NSinkBlocks * NF_BLOCK_SIZE ||
            (int64_t)s.SinkSetIdMax[k] != (int64_t)s.SinkSetStart[k] + tot)
            return fail(NF_ERR_INVALID, "bad SinkSetStart/IdMax");
    }
    for (int k = 0; k < s.NSourceSets; k++) {
        int64_t tot = 0;
        for (int c = 0; c < s.nfwidth; c++) {
            const nf_cell_plan &cp = s.SourcePlan[k].cell[c];
            if (cp.N < 0 || (uint64_t)cp.start + (uint64_t)cp.N > s.Nslab[c])
                return fail(NF_ERR_INVALID, "source plan cell outside its slab");
            tot += cp.N;
        }
        if (tot != s.SourceSetCount[k]) return fail(NF_ERR_INVALID, "SourceSetCount disagrees with SourcePlan");
        if (s.SourceSetStart[k] % NF_BLOCK_SIZE != 0 || s.SourceSetStart[k] < 0 ||
            (int64_t)s.SourceSetStart[k] + (tot + NF_BLOCK_SIZE - 1) / NF_BLOCK_SIZE * NF_BLOCK_SIZE >
                (int64_t)s.NSourceBlocks * NF_BLOCK_SIZE)
            return fail(NF_ERR_INVALID, "bad SourceSetStart");
    }
    for (int i = 0; i < s.InteractionCount; i++)
        if (s.SinkSourceInteractionList[i] < 0 || s.SinkSourceInteractionList[i] >= s.NSourceSets)
            return fail(NF_ERR_INVALID, "interaction list entry out of range");
    // Each padded block must lie inside the padded range of the pencil it
    // names: the sink gather indexes the pencil by the block's offset.
    for (int b = 0; b < s.NSinkBlocks; b++) {
        const int k = s.SinkBlockParentPencil[b];
        if (k < 0 || k >= s.NSinkSets) return fail(NF_ERR_INVALID, "SinkBlockParentPencil out of range");
        const int64_t lo = s.SinkSetStart[k];
        const int64_t hi = lo + (int64_t)(s.SinkSetCount[k] + NF_BLOCK_SIZE - 1) / NF_BLOCK_SIZE * NF_BLOCK_SIZE;
        if ((int64_t)b * NF_BLOCK_SIZE < lo || (int64_t)b * NF_BLOCK_SIZE >= hi)
            return fail(NF_ERR_INVALID, "SinkBlockParentPencil names a pencil that does not hold the block");
    }
    // The co-add reads pencil k = z's first cell as cell z of the skewer:
    // the plans must follow the reference's pencil geometry.
    for (int j = 0; j < s.j_width; j++)
        for (int k = 0; k < s.Nk; k++) {
            const nf_sink_plan &p = s.SinkPlan[j * s.Nk + k];
            for (int c = 0; c + 1 < s.nfwidth; c++) {
                const int z = (k + c) % s.cpd, z1 = (k + c + 1) % s.cpd;
                const nf_sink_plan &a = s.SinkPlan[j * s.Nk + z], &b = s.SinkPlan[j * s.Nk + z1];
                if (p.cell[c].start != a.cell[0].start || p.cell[c].N != a.cell[0].N ||
                    p.cell[c + 1].start != b.cell[0].start || p.cell[c + 1].N != b.cell[0].N)
                    return fail(NF_ERR_INVALID, "sink plans do not follow the pencil geometry");
            }
        }
    return NF_OK;
}

double event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

// ---- Autotuner -------------------------------------------------------

// Candidate tile shapes: sinks per lane S (sinks per wave 64 S) times
// waves per workgroup, each axis searched unless nf_setup fixes it (the
// persistent grid always fills the occupancy the kernel was built for:
// fewer resident waves, 8 per CU, never won on uniform or clustered slabs).
std::vector<TileShape> tile_candidates(int fixed_s, int fixed_wpw) {
    std::vector<TileShape> v;
    for (int S : {1, 2, 4})
        for (int wpw : {1, 2, 4})
            if ((!fixed_s || S == fixed_s) && (!fixed_wpw || wpw == fixed_wpw)) v.push_back(TileShape{S, 64, 4, 0, wpw});
    return v;
}

// The tuning key: a coarse histogram of sink-pencil occupancy (the fraction
// of the group's sinks in pencils of <64, 64-255, 256-1023 and >=1024 sinks,
// in quarters) and the group's size class (below 2^31, below 2^33, or more
// pair interactions: small launches are limited by their longest items,
// large ones by throughput).  Coarse on purpose: every new key costs one
// launch per candidate.
std::string occupancy_key(const std::vector<Task *> &g) {
    double h[4] = {0, 0, 0, 0};
    double tot = 0;
    uint64_t di = 0;
    for (Task *t : g) {
        di += queued_interactions(*t);  // device-planned: counted from the cellinfo, before any plan exists
        if (t->dp.on) {
            for (int b = 0; b < 4; b++) {
                h[b] += (double)t->dp.occ[b];
                tot += (double)t->dp.occ[b];
            }
            continue;
        }
        for (int k = 0; k < t->sic.NSinkSets; k++) {
            const int n = t->sic.SinkSetCount[k];
            h[n < 64 ? 0 : n < 256 ? 1 : n < 1024 ? 2 : 3] += n;
            tot += n;
        }
    }
    std::ostringstream os;
    for (int b = 0; b < 4; b++) os << (b ? "," : "") << (int)std::lround(4.0 * h[b] / std::max(tot, 1.0));
    os << ";" << (di < (1ull << 31) ? 0 : di < (1ull << 33) ? 1 : 2);
    return os.str();
}

// ---- Host-slab cache -------------------------------------------------

// Whether host memory was allocated pinned by HIP (then DMA reads it directly).
bool host_is_pinned(const void *p) {
    hipPointerAttribute_t a;
    const hipError_t e = hipPointerGetAttributes(&a, p);
    (void)hipGetLastError();  // pageable memory reports an error; do not leave it behind
    return e == hipSuccess && a.type == hipMemoryTypeHost;
}

// Find the HBM copy of a host slab (inside a stream session) or upload it
// into the least recently used free slot.  Returns the slot with a reference
// held; *uploaded says whether it was copied now (its t0/t1 events then
// bracket the copy).  Pageable host memory goes through a pinned staging
// buffer (copied by several threads); pinned memory is read by the DMA
// directly.  The copy into staging runs outside the cache lock; a second user
// of a slab being filled waits until its ready event is recorded, then waits
// on that event on the GPU.
int acquire_slab(Runtime &rt, const void *host, uint64_t bytes, Runtime::Slot **out, bool *uploaded) {
    auto &C = rt.cache;
    std::unique_lock<std::mutex> lk(C.m);
    *uploaded = false;
    if (C.active)
        for (auto &sp : C.slots) {
            Runtime::Slot &sl = *sp;
            if ((sl.valid || sl.filling) && sl.host == host && sl.bytes == bytes && sl.epoch == C.epoch) {
                sl.refs++;
                sl.last = ++C.clock;
                C.cv.wait(lk, [&] { return !sl.filling; });
                if (!sl.valid) {
                    sl.refs--;
                    return fail(NF_ERR_HIP, "slab upload by another task failed");
                }
                *out = &sl;
                return NF_OK;
            }
        }
    Runtime::Slot *v = nullptr;
    for (auto &sp : C.slots)
        if (sp->refs == 0 && (!v || sp->last < v->last)) v = sp.get();
    if (!v) {
        C.slots.push_back(std::make_unique<Runtime::Slot>());
        v = C.slots.back().get();
        NF_HIP(hipEventCreateWithFlags(&v->ready, hipEventDisableTiming));
        NF_HIP(hipEventCreate(&v->t0));
        NF_HIP(hipEventCreate(&v->t1));
    }
    v->valid = false;
    v->filling = true;
    v->refs = 1;
    v->last = ++C.clock;
    v->host = host;
    v->bytes = bytes;
    v->epoch = C.epoch;
    auto finish_fill = [&](int rc, int k) {  // under lk
        v->filling = false;
        v->valid = rc == NF_OK;
        if (rc != NF_OK) v->refs = 0;
        if (k >= 0) C.stage_busy[k] = false;
        C.cv.notify_all();
        return rc;
    };
    int rc = v->buf.ensure(bytes);
    if (rc) return finish_fill(rc, -1);
    const bool direct = bytes == 0 || host_is_pinned(host);
    int k = -1;
    if (!direct) {
        C.cv.wait(lk, [&] {
            for (int i = 0; i < Runtime::kStages; i++)
                if (!C.stage_busy[i]) return true;
            return false;
        });
        for (int i = 0; i < Runtime::kStages && k < 0; i++)
            if (!C.stage_busy[i]) k = i;
        C.stage_busy[k] = true;
    }
    lk.unlock();
    const void *src = host;
    hipError_t e = hipSuccess;
    if (!direct) {
        e = hipEventSynchronize(C.stage_done[k]);  // the last DMA out of this buffer is done
        if (e == hipSuccess) rc = C.stage[k].ensure(bytes);
        if (e == hipSuccess && rc == NF_OK) {
            par_memcpy(C.stage[k].p, host, bytes);
            src = C.stage[k].p;
        }
    }
    if (e == hipSuccess && rc == NF_OK) {
        std::lock_guard<std::mutex> ul(C.up_enqueue);
        e = hipEventRecord(v->t0, C.up);
        if (e == hipSuccess && bytes) e = hipMemcpyAsync(v->buf.p, src, bytes, hipMemcpyHostToDevice, C.up);
        if (e == hipSuccess) e = hipEventRecord(v->t1, C.up);
        if (e == hipSuccess && k >= 0) e = hipEventRecord(C.stage_done[k], C.up);
        if (e == hipSuccess) e = hipEventRecord(v->ready, C.up);
    }
    lk.lock();
    if (e != hipSuccess) rc = fail(NF_ERR_HIP, std::string("slab upload: ") + hipGetErrorString(e));
    if (rc == NF_OK) *uploaded = true;
    *out = v;
    return finish_fill(rc, k);
}

void release_slab(Runtime &rt, Runtime::Slot *sl) {
    auto &C = rt.cache;
    std::lock_guard<std::mutex> lk(C.m);
    if (--sl->refs == 0 && !(C.active && sl->epoch == C.epoch)) sl->valid = false;
}

// A roctx range around a launch group's work on its worker (visible with
// rocprofv3 --marker-trace; a no-op without a tool attached).
struct TraceRange {
    explicit TraceRange(const char *msg) { roctxRangePushA(msg); }
    ~TraceRange() { roctxRangePop(); }
};

// Counts a launch group as on the GPU from its first enqueue to its
// completion, for the throughput timer.
struct ActiveOnGpu {
    Runtime::Throughput &t;
    bool on = false;
    void start() {
        std::lock_guard<std::mutex> lk(t.m);
        if (t.active++ == 0) t.since = clk::now();
        on = true;
    }
    void stop() {
        if (!on) return;
        on = false;
        std::lock_guard<std::mutex> lk(t.m);
        if (--t.active == 0) t.active_s += std::chrono::duration<double>(clk::now() - t.since).count();
    }
    ~ActiveOnGpu() { stop(); }
};

// Drops the slab-cache references a group holds when it is done.
struct SlotRelease {
    Runtime &rt;
    std::vector<Runtime::Slot *> &v;
    ~SlotRelease() {
        for (auto *x : v)
            if (x) release_slab(rt, x);
    }
};

// The device planner's scratch for a group of SICs: each SIC's totals
// (256 bytes apiece, read back together), the work-list histogram and
// cursors the group shares, each SIC's block sums.
// Then the group-wide planning launches' inputs, uploaded from the worker's
// pinned staging (laid out the same from o_in): the SICs' GPlanIn and
// GPlanOut and the first workgroup of each SIC in phase A's and phase B's
// grids.
struct GPlanScratch {
    size_t o_shared = 0, total = 0;
    std::vector<size_t> o_part;
    size_t o_in = 0, o_out = 0, o_fa = 0, o_fb = 0;
};
GPlanScratch gplan_scratch(const std::vector<Task *> &g) {
    static_assert(sizeof(GPlanTotals) <= 256, "totals slot");
    GPlanScratch L;
    const size_t n = g.size();
    size_t o = 256 * n;
    L.o_shared = o;
    o = align_up(o + sizeof(int32_t) * kPlanSharedInts, 256);
    for (Task *t : g) {
        L.o_part.push_back(o);
        o = align_up(o + sizeof(int32_t) * gpu_plan_part_ints(t->sic.NSinkSets, t->sic.NSourceSets), 256);
    }
    L.o_in = o;
    L.o_out = align_up(L.o_in + sizeof(GPlanIn) * n, 256);
    L.o_fa = align_up(L.o_out + sizeof(GPlanOut) * n, 256);
    L.o_fb = align_up(L.o_fa + sizeof(int32_t) * (n + 1), 256);
    L.total = align_up(L.o_fb + sizeof(int32_t) * (n + 1), 256);
    return L;
}

// Device planning, phase A, of a group of tasks from nf_submit_slab: their
// cellinfo slabs go up through the slab cache (held in `cslot` until the
// group is done) and the counting kernels are enqueued on the planning
// stream.  Nothing waits for them: each SIC is laid out for the bounds of
// its sizes (dp_bounds, set here), the kernels that follow read phase A's
// counts from device memory, and the totals come back with the group's
// results.  Fills in[i] for phase B.
int plan_phase_a(Runtime &rt, Worker &w, std::vector<Task *> &g, std::vector<Runtime::Slot *> &cslot,
                 std::vector<GPlanIn> &in, const GPlanScratch &gs, uint64_t *to_dev) {
    const size_t n = g.size();
    int rc = w.gscratch.ensure(gs.total);
    if (rc || (rc = w.pin_tot.ensure(256 * n)) || (rc = w.pin_gplan.ensure(gs.total - gs.o_in))) return rc;
    char *sc = static_cast<char *>(w.gscratch.p);
    int32_t *shared = reinterpret_cast<int32_t *>(sc + gs.o_shared);
    in.assign(n, GPlanIn{});
    // Heavy register tiles are split against the group's work, counted from
    // the cellinfo at submission (the three planning phases all use this one
    // threshold, so the item counts of phase A are those the later phases
    // write).
    uint64_t group_est = 0;
    for (Task *t : g) group_est += t->dp.interactions;
    for (size_t i = 0; i < n; i++) {
        Task &t = *g[i];
        nf_sic &s = t.sic;
        dp_bounds(s, &t.dp.chunks_bound, &t.dp.items_bound);
        GPlanIn &gi = in[i];
        for (int ti = 0; ti < 3; ti++) gi.split_cost[ti] = tile_split_cost(s.softening, 1 << ti, group_est);
        gi.cpd = s.cpd;
        gi.R = s.nfradius;
        gi.W = s.nfwidth;
        gi.jw = s.j_width;
        gi.j_low = s.j_low;
        const uint64_t cbytes = sizeof(nf_cellinfo) * (uint64_t)s.cpd * (uint64_t)s.cpd;
        for (int c = 0; c < s.nfwidth; c++) {
            Runtime::Slot *sl = nullptr;
            bool up = false;
            if ((rc = acquire_slab(rt, t.dp.cellinfo[c], cbytes, &sl, &up))) return rc;
            cslot.push_back(sl);
            if (up) *to_dev += cbytes;
            gi.ci[c] = static_cast<const nf_cellinfo *>(sl->buf.p);
            gi.nslab[c] = s.Nslab[c];
            NF_HIP(hipStreamWaitEvent(w.pstream, sl->ready, 0));
        }
    }
    // One launch per phase for the whole group: its SICs' inputs and grid
    // offsets go up first.
    char *pg = static_cast<char *>(w.pin_gplan.p);
    std::memcpy(pg, in.data(), sizeof(GPlanIn) * n);
    int32_t *fa = reinterpret_cast<int32_t *>(pg + (gs.o_fa - gs.o_in));
    int grid = 0;
    for (size_t i = 0; i < n; i++) {
        fa[i] = grid;
        grid += gpu_plan_grid_a(in[i]);
    }
    fa[n] = grid;
    NF_HIP(hipMemcpyAsync(sc + gs.o_in, pg, sizeof(GPlanIn) * n, hipMemcpyHostToDevice, w.pstream));
    NF_HIP(hipMemcpyAsync(sc + gs.o_fa, fa, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice, w.pstream));
    const int e = launch_gpu_plan_group(0, reinterpret_cast<const GPlanIn *>(sc + gs.o_in), nullptr,
                                        reinterpret_cast<GPlanTotals *>(sc), shared,
                                        reinterpret_cast<const int32_t *>(sc + gs.o_fa), (int)n, grid, w.pstream);
    if (e) return fail(NF_ERR_HIP, std::string("device plan launch: ") + hipGetErrorString((hipError_t)e));
    return NF_OK;
}

// The device planner's totals, read back with the group's results, into its
// SICs' members; a cell outside its slab fails the group.
int apply_group_totals(Worker &w, std::vector<Task *> &g) {
    int rc = NF_OK;
    for (size_t i = 0; i < g.size(); i++) {
        std::memcpy(&g[i]->dp.tot, static_cast<char *>(w.pin_tot.p) + 256 * i, sizeof(GPlanTotals));
        const int r = apply_totals(g[i]->sic, g[i]->dp.tot);
        if (r && rc == NF_OK) rc = r;
    }
    return rc;
}

// Execute a group of compatible SICs on worker w: one plan upload, a gather
// per SIC, one persistent direct-kernel launch over all their work items
// (batched path; the plain path launches per SIC), a co-add per SIC.
int run_group(Runtime &rt, Worker &w, std::vector<Task *> &g) {
    const auto t0 = clk::now();
    NF_HIP(hipSetDevice(w.device));
    const int path = rt.params.path;
    TileShape ts = rt.shape;
    const int n = (int)g.size();
    char label[96];
    std::snprintf(label, sizeof label, "nf group: %d SIC%s from slab %d rows [%d,%d)%s", n, n > 1 ? "s" : "",
                  g[0]->sic.SlabId, g[0]->sic.j_low, g[0]->sic.j_high, g[0]->dp.on ? ", device-planned" : "");
    TraceRange trace_range(label);
    // Tasks from nf_submit_slab (grouped only with each other) are planned
    // on the device: phase A is enqueued first; everything below is laid out
    // for the bounds of the plans' sizes.
    const bool dplan = g[0]->dp.on;
    std::vector<GPlanIn> gin;
    GPlanScratch gs;
    std::vector<Runtime::Slot *> cslot;
    SlotRelease cell_release{rt, cslot};
    uint64_t cell_bytes = 0;
    if (dplan) {
        for (Task *t : g)
            if (!t->dp.on) return fail(NF_ERR_STATE, "device-planned and host-planned SICs in one group");
        gs = gplan_scratch(g);
        const int rc = plan_phase_a(rt, w, g, cslot, gin, gs, &cell_bytes);
        if (rc) return rc;
        if (path == NF_PATH_BATCHED && rt.autotune)
            for (Task *t : g) dp_occupancy(*t, t->dp.occ);
    }
    const auto t_plan = clk::now();
    // Autotuning: a known occupancy histogram gets its tuned shape; a new one
    // is tuned on this group's launch (every candidate computes the same
    // partials; the last launch's are co-added).
    std::string tkey;
    bool trial = false;
    if (path == NF_PATH_BATCHED && rt.autotune) {
        tkey = occupancy_key(g);
        std::lock_guard<std::mutex> lk(rt.tune_m);
        auto it = rt.tune.find(tkey);
        if (it != rt.tune.end()) {
            ts = it->second.shape;
            it->second.uses++;
        } else {
            trial = true;
        }
    }
    const nf_sic &s0 = g[0]->sic;
    PairConst pc;
    int rc = derive_pair_const(s0.softening, s0.eps, s0.b2, &pc);
    if (rc) return rc;

    // Work list over the whole group, longest items first: one list per
    // distinct sinks-per-lane value to be launched (several while tuning).
    // Items within 1/8 octave of cost count as equal and are kept in sink
    // pencil order (SIC, z, y): a stable counting sort by bucket, of the
    // tiles and of the short tails apart, merged by interleave_pos.
    uint64_t group_total = 0;
    for (int i = 0; i < n; i++) group_total += g[i]->sic.DirectTotal;
    auto make_items = [&](int S) {
        std::vector<WorkItem> raw, out;
        std::vector<uint16_t> bucket;
        const uint64_t max_tile_cost = tile_split_cost(s0.softening, S, group_total);
        for (int i = 0; i < n; i++) build_items(g[i]->sic, S, i, max_tile_cost, raw, bucket);
        std::vector<size_t> start(2 * kItemBuckets + 1, 0);
        for (size_t i = 0; i < raw.size(); i++) start[item_key(raw[i], bucket[i]) + 1]++;
        for (int b = 0; b < 2 * kItemBuckets; b++) start[b + 1] += start[b];
        out.resize(raw.size());
        const int64_t M = (int64_t)start[kItemBuckets], T = (int64_t)raw.size() - M;
        for (size_t i = 0; i < raw.size(); i++) {
            const int key = item_key(raw[i], bucket[i]);
            const int64_t r = (int64_t)start[key]++;
            out[key < kItemBuckets ? interleave_pos(false, r, M, T) : interleave_pos(true, r - M, M, T)] = raw[i];
        }
        return out;
    };
    std::vector<int> Svals;
    if (path == NF_PATH_BATCHED) {
        if (trial)
            for (const TileShape &c : rt.candidates)
                if (std::find(Svals.begin(), Svals.end(), c.sinks_per_lane) == Svals.end()) Svals.push_back(c.sinks_per_lane);
        if (std::find(Svals.begin(), Svals.end(), ts.sinks_per_lane) == Svals.end()) Svals.push_back(ts.sinks_per_lane);
    }
    // (A device-planned SIC's lists are written on the device; phase A
    // counted their items.)
    std::vector<std::vector<WorkItem>> itemsets;
    std::vector<size_t> nitems;
    for (int S : Svals) {
        itemsets.push_back(dplan ? std::vector<WorkItem>() : make_items(S));
        size_t m = itemsets.back().size();
        if (dplan)
            for (Task *t : g) m += (size_t)t->dp.items_bound;
        nitems.push_back(m);
    }
    auto items_of = [&](int S) -> int {
        for (size_t k = 0; k < Svals.size(); k++)
            if (Svals[k] == S) return (int)k;
        return 0;
    };
    size_t items_total = 0;
    for (size_t v : nitems) items_total += align_up(sizeof(WorkItem) * std::max<size_t>(v, 1), 256);

    // Plan upload: [views][group SICs][first waves][items][counter][plan
    // arrays of each SIC]; a device-planned group uploads the head (up to the
    // items) only.
    const size_t o_views = 0;
    const size_t o_gsic = align_up(o_views + sizeof(SicView) * (size_t)n, 64);
    const size_t o_gfirst = align_up(o_gsic + sizeof(GroupSic) * (size_t)n, 64);  // gather, then co-add
    const size_t head = align_up(o_gfirst + 2 * sizeof(int32_t) * (size_t)(n + 1), 256);
    const size_t o_items = head;
    std::vector<size_t> o_itemset(itemsets.size());
    {
        size_t o = o_items;
        for (size_t k = 0; k < itemsets.size(); k++) {
            o_itemset[k] = o;
            o += align_up(sizeof(WorkItem) * std::max<size_t>(nitems[k], 1), 256);
        }
    }
    const size_t o_counter = align_up(o_items + std::max<size_t>(items_total, 256), 256);
    size_t plan_total = align_up(o_counter + 256, 256);
    std::vector<PlanLayout> PL(n);
    for (int i = 0; i < n; i++) {
        PL[i] = plan_layout(g[i]->sic, plan_total,
                            g[i]->dp.on ? (int32_t)g[i]->dp.chunks_bound : count_coadd_chunks(g[i]->sic));
        plan_total += PL[i].total;
    }
    // Device working set of each SIC.
    std::vector<WorkLayout> WL(n);
    size_t work_total = 0;
    for (int i = 0; i < n; i++) {
        WL[i] = work_layout(g[i]->sic, path, work_total);
        work_total += WL[i].total;
    }
    // Host mode: every distinct host slab of the group once (through the
    // slab cache), plus a device acceleration slab per SIC.
    std::vector<const float *> hslab;
    std::vector<uint64_t> hslab_n;
    std::vector<size_t> acc_off(n, 0), near_off(n, 0);
    size_t slab_total = 0;
    if (s0.mem == NF_MEM_HOST) {
        for (int i = 0; i < n; i++) {
            const nf_sic &s = g[i]->sic;
            for (int c = 0; c < s.nfwidth; c++) {
                if (std::find(hslab.begin(), hslab.end(), s.SourcePosSlab[c]) != hslab.end()) continue;
                hslab.push_back(s.SourcePosSlab[c]);
                hslab_n.push_back(s.Nslab[c]);
            }
            acc_off[i] = slab_total;
            slab_total = align_up(slab_total + 16 * (size_t)s.Nslab[s.nfradius], 256);
            if (s.kick && s.kick->NearAccOut) {  // and one for the near-field record
                near_off[i] = slab_total;
                slab_total = align_up(slab_total + 16 * (size_t)s.Nslab[s.nfradius], 256);
            }
        }
    }
    if (plan_total + work_total + slab_total > rt.max_buffer_bytes)
        return fail(NF_ERR_CAPACITY, "SIC group needs " + std::to_string(plan_total + work_total + slab_total) +
                                         " device bytes, buffer holds " + std::to_string(rt.max_buffer_bytes));
    if ((rc = w.plan.ensure(plan_total)) || (rc = w.work.ensure(work_total)) || (rc = w.pin_census.ensure(32)) ||
        (rc = w.pin_plan.ensure(dplan ? head : plan_total)))
        return rc;
    if (s0.mem == NF_MEM_HOST && (rc = w.slabs.ensure(slab_total))) return rc;

    // Slabs: HBM copies through the cache (uploads overlap other groups' kernels).
    std::vector<Runtime::Slot *> hslot(hslab.size(), nullptr);
    SlotRelease slot_release{rt, hslot};
    int n_up = 0, n_reuse = 0;
    std::vector<char> fresh(hslab.size(), 0);
    for (size_t k = 0; k < hslab.size(); k++) {
        bool up = false;
        if ((rc = acquire_slab(rt, hslab[k], 12 * hslab_n[k], &hslot[k], &up))) return rc;
        fresh[k] = up;
        if (up) n_up++;
        else n_reuse++;
    }

    // ---- Fill the pinned plan image.
    char *hp = static_cast<char *>(w.pin_plan.p);
    char *pd = static_cast<char *>(w.plan.p);
    char *wb = static_cast<char *>(w.work.p);
    char *db = static_cast<char *>(w.slabs.p);
    std::vector<DevPlan> plans(n);
    std::vector<DevSlabs> slabs(n);
    std::vector<KickArgs> kicks(n);
    std::vector<float *> accd(n);
    SicView *views = reinterpret_cast<SicView *>(hp + o_views);
    for (int i = 0; i < n; i++) {
        const nf_sic &s = g[i]->sic;
        const PlanLayout &L = PL[i];
        if (!g[i]->dp.on) fill_plan_image(s, L, hp);  // host-planned: copy the arrays, derive the lists

        DevPlan &pl = plans[i];
        pl = dev_plan(s, L, pd);
        if (dplan) {  // the actual counts, in device memory; the group's bad cells
            pl.max_src_blocks = g[i]->dp.max_src_blocks;
            char *sc = static_cast<char *>(w.gscratch.p);
            pl.dtot = reinterpret_cast<const GPlanTotals *>(sc + 256 * (size_t)i);
            pl.dbad = reinterpret_cast<const int32_t *>(sc + gs.o_shared) + kPlanGroupBad;
        }

        DevSlabs &sl = slabs[i];
        for (int c = 0; c < s.nfwidth; c++) {
            sl.n[c] = s.Nslab[c];
            if (s.mem == NF_MEM_HOST) {
                const size_t k = std::find(hslab.begin(), hslab.end(), s.SourcePosSlab[c]) - hslab.begin();
                sl.pos[c] = static_cast<const float *>(hslot[k]->buf.p);
            } else {
                sl.pos[c] = s.SourcePosSlab[c];
            }
        }
        if (s.mem == NF_MEM_DEVICE) sl.pos[s.nfradius] = s.SinkPosSlab;
        sl.aos = s.pos_layout == NF_POS_AOS ? 1 : 0;
        accd[i] = s.mem == NF_MEM_HOST ? reinterpret_cast<float *>(db + acc_off[i]) : s.SinkAccSlab;
        KickArgs &ka = kicks[i];
        std::memset(&ka, 0, sizeof(ka));
        if (s.kick) {
            uint64_t r0, r1;
            task_rows(*g[i], &r0, &r1);
            ka.on = 1;
            ka.rescale = (float)s.kick->rescale;
            ka.inv_eps3 = (float)s.kick->inv_eps3;
            ka.density_rad2 = s.kick->density_rad2;
            ka.far = s.mem == NF_MEM_HOST ? reinterpret_cast<const float *>(wb + WL[i].o_far) : s.kick->FarAccSlab;
            if (s.kick->NearAccOut)
                ka.near_out = s.mem == NF_MEM_HOST ? reinterpret_cast<float *>(db + near_off[i]) : s.kick->NearAccOut;
            ka.far_base = s.mem == NF_MEM_HOST ? (int64_t)r0 : 0;
        }

        SicView &v = views[i];
        std::memset(&v, 0, sizeof(v));
        v.sink_set_start = pl.sink_set_start;
        v.src_set_start = pl.src_set_start;
        v.src_set_count = pl.src_set_count;
        v.ilist = pl.ilist;
        v.sink_plan = pl.sink_plan;
        v.sink_slab = sl.pos[s.nfradius];
        v.sink_n = s.Nslab[s.nfradius];
        v.aos = sl.aos;
        v.sources4 = wb + WL[i].o_src;
        v.box4 = box_after(wb + WL[i].o_src, s.NSourceBlocks);
        v.partial = reinterpret_cast<double *>(wb + WL[i].o_part);
        v.W = s.nfwidth;
        v.R = s.nfradius;
    }
    // The group-wide gather and co-add launches' view of each SIC.
    GroupSic *gsic = reinterpret_cast<GroupSic *>(hp + o_gsic);
    int32_t *gfirst = reinterpret_cast<int32_t *>(hp + o_gfirst), *cfirst = gfirst + (n + 1);
    gfirst[0] = cfirst[0] = 0;
    for (int i = 0; i < n; i++) {
        GroupSic &G = gsic[i];
        std::memset(&G, 0, sizeof(G));
        G.pl = plans[i];
        G.sl = slabs[i];
        G.kick = kicks[i];
        G.sources4 = wb + WL[i].o_src;
        G.partial = wb + WL[i].o_part;
        G.acc4 = accd[i];
        G.out_scale = pc.out_scale;
        gfirst[i + 1] = gfirst[i] + group_gather_waves(plans[i]);
        cfirst[i + 1] = cfirst[i] + group_coadd_waves(plans[i]);
    }
    for (size_t k = 0; k < itemsets.size(); k++)
        if (!itemsets[k].empty())
            std::memcpy(hp + o_itemset[k], itemsets[k].data(), sizeof(WorkItem) * itemsets[k].size());
    const auto t_fill = clk::now();
    // The plan image goes up on the copy stream, overlapping the kernels of
    // the group before (this worker's previous group has completed, so its
    // plan buffer is free).  A device-planned SIC's image is its view only;
    // phase B and the work lists fill the rest on the planning stream.
    if (dplan) {
        NF_HIP(hipMemcpyAsync(w.plan.p, hp, head, hipMemcpyHostToDevice, w.pstream));
        char *sc = static_cast<char *>(w.gscratch.p);
        int32_t *shared = reinterpret_cast<int32_t *>(sc + gs.o_shared);
        std::vector<GPlanOut> go(n);
        for (int i = 0; i < n; i++) {
            go[i] = plan_out(pd, PL[i], i, shared, reinterpret_cast<int32_t *>(sc + gs.o_part[i]));
            go[i].group = shared;
            for (size_t k = 0; k < Svals.size(); k++) {
                const int ti = Svals[k] == 1 ? 0 : Svals[k] == 2 ? 1 : 2;
                go[i].items[ti] = reinterpret_cast<WorkItem *>(pd + o_itemset[k]);
            }
        }
        char *pg = static_cast<char *>(w.pin_gplan.p);
        std::memcpy(pg + (gs.o_out - gs.o_in), go.data(), sizeof(GPlanOut) * (size_t)n);
        int32_t *fb = reinterpret_cast<int32_t *>(pg + (gs.o_fb - gs.o_in));
        int grid = 0;
        for (int i = 0; i < n; i++) {
            fb[i] = grid;
            grid += gpu_plan_grid_b(gin[i]);
        }
        fb[n] = grid;
        NF_HIP(hipMemcpyAsync(sc + gs.o_out, pg + (gs.o_out - gs.o_in), sizeof(GPlanOut) * (size_t)n,
                              hipMemcpyHostToDevice, w.pstream));
        NF_HIP(hipMemcpyAsync(sc + gs.o_fb, fb, sizeof(int32_t) * (size_t)(n + 1), hipMemcpyHostToDevice, w.pstream));
        const GPlanIn *ins_d = reinterpret_cast<const GPlanIn *>(sc + gs.o_in);
        const GPlanOut *outs_d = reinterpret_cast<const GPlanOut *>(sc + gs.o_out);
        const int32_t *fb_d = reinterpret_cast<const int32_t *>(sc + gs.o_fb);
        int e = launch_gpu_plan_group(1, ins_d, outs_d, nullptr, nullptr, fb_d, n, grid, w.pstream);
        if (!e) e = launch_gpu_plan_group(2, ins_d, outs_d, nullptr, nullptr, fb_d, n, grid, w.pstream);
        if (e) return fail(NF_ERR_HIP, std::string("device plan launch: ") + hipGetErrorString((hipError_t)e));
        NF_HIP(hipEventRecord(w.gplan_ready, w.pstream));
    } else {
        NF_HIP(hipMemcpyAsync(w.plan.p, hp, plan_total, hipMemcpyHostToDevice, rt.cache.up));
    }
    // Host mode with a fused kick: the far-field accelerations of the sink
    // rows go up beside the plan (pinned ones directly, else through staging).
    size_t far_bytes = 0;
    for (int i = 0; i < n; i++) {
        const nf_sic &s = g[i]->sic;
        if (!s.kick || s.mem != NF_MEM_HOST) continue;
        uint64_t r0, r1;
        task_rows(*g[i], &r0, &r1);
        if (r1 > r0) far_bytes += 12 * (size_t)(r1 - r0);
    }
    if (far_bytes) {
        if ((rc = w.pin_far.ensure(far_bytes))) return rc;
        size_t off = 0;
        for (int i = 0; i < n; i++) {
            const nf_sic &s = g[i]->sic;
            if (!s.kick || s.mem != NF_MEM_HOST) continue;
            uint64_t r0, r1;
            task_rows(*g[i], &r0, &r1);
            if (r1 <= r0) continue;
            const size_t nb = 12 * (size_t)(r1 - r0);
            const float *src = s.kick->FarAccSlab + 3 * r0;
            if (!host_is_pinned(src)) {
                par_memcpy(static_cast<char *>(w.pin_far.p) + off, src, nb);
                src = reinterpret_cast<const float *>(static_cast<char *>(w.pin_far.p) + off);
            }
            NF_HIP(hipMemcpyAsync(wb + WL[i].o_far, src, nb, hipMemcpyHostToDevice, rt.cache.up));
            off += nb;
        }
    }
    NF_HIP(hipEventRecord(w.plan_ready, rt.cache.up));

    // ---- Enqueue the group as one uninterrupted sequence.
    hipStream_t stream = rt.shared ? rt.shared : w.stream;
    std::unique_lock<std::mutex> enqueue_lock(rt.enqueue, std::defer_lock);
    if (rt.shared) enqueue_lock.lock();
    uint64_t to_dev = (dplan ? sizeof(SicView) * (size_t)n : plan_total) + far_bytes + cell_bytes, from_dev = 0;
    // With graphs, the sequence below is captured and launched as one graph.
    const bool use_graph = rt.params.use_graphs && !trial;
    struct CaptureGuard {  // ends an open capture on every early return
        hipStream_t s;
        bool on = false;
        ~CaptureGuard() {
            if (on) {
                hipGraph_t gr = nullptr;
                if (hipStreamEndCapture(s, &gr) == hipSuccess && gr) (void)hipGraphDestroy(gr);
            }
        }
    } capture{stream};
    NF_HIP(hipStreamWaitEvent(stream, w.plan_ready, 0));
    if (dplan) NF_HIP(hipStreamWaitEvent(stream, w.gplan_ready, 0));
    for (size_t k = 0; k < hslot.size(); k++) NF_HIP(hipStreamWaitEvent(stream, hslot[k]->ready, 0));
    for (size_t k = 0; k < hslab.size(); k++)
        if (fresh[k]) to_dev += 12 * hslab_n[k];
    ActiveOnGpu on_gpu{rt.tp};
    on_gpu.start();
    NF_HIP(hipEventRecord(w.ev[1], stream));
    // The captured region holds kernels only; copies and timing events stay
    // outside it, around the graph launch.
    if (use_graph) {
        (void)hipGetLastError();
        NF_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        capture.on = true;
    }
    int e = 0;
    const GroupSic *gsic_d = reinterpret_cast<const GroupSic *>(pd + o_gsic);
    const int32_t *gfirst_d = reinterpret_cast<const int32_t *>(pd + o_gfirst), *cfirst_d = gfirst_d + (n + 1);
    if (path == NF_PATH_BATCHED) {
        bool seg = false;
        for (int i = 0; i < n; i++) seg = seg || gather_segmented(plans[i].max_src_blocks);
        e = launch_gather_group(gsic_d, gfirst_d, n, gfirst[n], s0.cpd, seg, stream);
    }
    else
        for (int i = 0; i < n && !e; i++)
            e = launch_gather(plans[i], slabs[i], false, wb + WL[i].o_sinks, wb + WL[i].o_src, stream);
    if (e) return fail(NF_ERR_HIP, std::string("gather launch: ") + hipGetErrorString((hipError_t)e));
    if (!use_graph) NF_HIP(hipEventRecord(w.ev[2], stream));
    std::vector<hipEvent_t> tev;
    unsigned long long *census = nullptr;  // batched path: interactions per pair form (device)
    if (path == NF_PATH_BATCHED) {
        int32_t *counter = reinterpret_cast<int32_t *>(pd + o_counter);
        if (rt.params.census)  // in the counter's 256-byte slot
            census = reinterpret_cast<unsigned long long *>(pd + o_counter + kCensusBytes);
        // A device-planned group's list lengths and bad cells are in its
        // shared planning area.
        const int32_t *gshared =
            dplan ? reinterpret_cast<const int32_t *>(static_cast<char *>(w.gscratch.p) + gs.o_shared) : nullptr;
        auto launch = [&](const TileShape &shape) {
            const int k = items_of(shape.sinks_per_lane);
            const int ti = shape.sinks_per_lane == 1 ? 0 : shape.sinks_per_lane == 2 ? 1 : 2;
            return launch_batched(reinterpret_cast<const SicView *>(pd + o_views), s0.cpd, s0.softening, pc, shape,
                                  reinterpret_cast<const WorkItem *>(pd + o_itemset[k]),
                                  (int32_t)nitems[k], counter, stream,
                                  gshared ? gshared + kPlanGroupItems + ti : nullptr,
                                  gshared ? gshared + kPlanGroupBad : nullptr, census != nullptr);
        };
        if (trial) {
            // Time each candidate; the chosen shape runs last (every candidate
            // writes bitwise the same partials: a sink's sum does not depend on
            // the sinks per lane).
            const int ncand = (int)rt.candidates.size();
            tev.resize(ncand + 1);
            for (auto &ev : tev) NF_HIP(hipEventCreate(&ev));
            NF_HIP(hipEventRecord(tev[0], stream));
            for (int c = 0; c < ncand && !e; c++) {
                e = launch(rt.candidates[c]);
                NF_HIP(hipEventRecord(tev[c + 1], stream));
            }
        }
        if (!e) {
            if (trial) NF_HIP(hipEventRecord(w.ev[2], stream));
            e = launch(ts);
        }
    } else {
        for (int i = 0; i < n && !e; i++)
            e = launch_plain(plans[i], s0.softening, pc, wb + WL[i].o_sinks, wb + WL[i].o_src, wb + WL[i].o_part,
                             stream);
    }
    if (e) return fail(NF_ERR_HIP, std::string("direct kernel launch: ") + hipGetErrorString((hipError_t)e));
    if (!use_graph) NF_HIP(hipEventRecord(w.ev[3], stream));
    if (path == NF_PATH_BATCHED)
        e = launch_coadd_group(gsic_d, cfirst_d, n, cfirst[n], stream);
    else
        for (int i = 0; i < n && !e; i++)
            e = launch_coadd_plain(plans[i], wb + WL[i].o_part, accd[i], kicks[i], stream);
    if (e) return fail(NF_ERR_HIP, std::string("co-add launch: ") + hipGetErrorString((hipError_t)e));
    if (use_graph) {
        hipGraph_t graph = nullptr;
        capture.on = false;
        NF_HIP(hipStreamEndCapture(stream, &graph));
        bool updated = false;
        if (w.gexec) {
            hipGraphNode_t err_node = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            updated = hipGraphExecUpdate(w.gexec, graph, &err_node, &res) == hipSuccess &&
                      res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGraphExecDestroy(w.gexec);
                w.gexec = nullptr;
            }
        }
        hipError_t ge = hipSuccess;
        if (!updated) ge = hipGraphInstantiate(&w.gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ge != hipSuccess) {
            w.gexec = nullptr;
            return fail(NF_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ge));
        }
        if (updated) w.graphs_updated++;
        else w.graphs_instantiated++;
        NF_HIP(hipGraphLaunch(w.gexec, stream));
    }
    // A device-planned group's totals and the batched launch's census come
    // back behind its kernels.
    if (dplan) NF_HIP(hipMemcpyAsync(w.pin_tot.p, w.gscratch.p, 256 * (size_t)n, hipMemcpyDeviceToHost, stream));
    if (census) NF_HIP(hipMemcpyAsync(w.pin_census.p, census, 32, hipMemcpyDeviceToHost, stream));
    NF_HIP(hipEventRecord(w.ev[4], stream));
    if (enqueue_lock.owns_lock()) enqueue_lock.unlock();
    NF_HIP(hipEventSynchronize(w.ev[4]));
    on_gpu.stop();
    if (dplan && (rc = apply_group_totals(w, g))) return rc;
    const double kern_ms = event_ms(w.ev[1], w.ev[4]);
    // A graph is timed as a whole (gathers and co-adds included).
    const double direct_ms = use_graph ? kern_ms : event_ms(w.ev[2], w.ev[3]);
    double upload_ms = 0.0;  // the slabs this group copied (they finished before its kernels)
    for (size_t k = 0; k < hslot.size(); k++)
        if (fresh[k]) upload_ms += event_ms(hslot[k]->t0, hslot[k]->t1);
    if (trial) {
        Runtime::TuneEntry te;
        int best = 0;
        for (int c = 0; c < (int)rt.candidates.size(); c++) {
            te.trial_ms.push_back(event_ms(tev[c], tev[c + 1]));
            if (te.trial_ms[c] < te.trial_ms[best]) best = c;
        }
        for (auto &ev : tev) (void)hipEventDestroy(ev);
        te.shape = rt.candidates[best];
        for (Task *t : g) te.interactions += t->sic.DirectTotal;
        te.uses = 1;
        std::lock_guard<std::mutex> lk(rt.tune_m);
        rt.tune.emplace(tkey, te);
    }

    // ---- Host mode: the SICs' sink rows back into their AccSlabs.
    double copyback = 0.0, download_ms = 0.0;
    if (s0.mem == NF_MEM_HOST) {
        const auto tc = clk::now();
        for (int i = 0; i < n; i++) {
            const nf_sic &s = g[i]->sic;
            uint64_t r0, r1;
            task_rows(*g[i], &r0, &r1);
            if (r1 <= r0) continue;
            // The copy-back runs on the download stream, after the group's
            // kernels, while the next group's kernels use the compute stream:
            // straight into a pinned acceleration slab, else through staging.
            const size_t nb = 16 * (size_t)(r1 - r0);
            auto copy_rows = [&](float *slab, const float *dev) -> int {
                float *dst = slab + 4 * r0;
                const bool direct = host_is_pinned(dst);
                int rc2 = NF_OK;
                if (!direct && (rc2 = w.pin_acc.ensure(nb))) return rc2;
                NF_HIP(hipEventRecord(w.ev[0], rt.cache.down));
                NF_HIP(hipMemcpyAsync(direct ? (void *)dst : w.pin_acc.p, dev + 4 * r0, nb, hipMemcpyDeviceToHost,
                                      rt.cache.down));
                NF_HIP(hipEventRecord(w.ev[2], rt.cache.down));
                NF_HIP(hipEventSynchronize(w.ev[2]));
                download_ms += event_ms(w.ev[0], w.ev[2]);
                if (!direct) par_memcpy(dst, w.pin_acc.p, nb);
                from_dev += nb;
                return NF_OK;
            };
            if ((rc = copy_rows(s.SinkAccSlab, accd[i]))) return rc;
            if (s.kick && s.kick->NearAccOut &&
                (rc = copy_rows(s.kick->NearAccOut, reinterpret_cast<const float *>(db + near_off[i]))))
                return rc;
        }
        copyback = std::chrono::duration<double>(clk::now() - tc).count();
    }
    uint64_t gdi = 0;
    for (int i = 0; i < n; i++) gdi += g[i]->sic.DirectTotal;
    const double wall = std::chrono::duration<double>(clk::now() - t0).count();
    for (int i = 0; i < n; i++) {
        const nf_sic &s = g[i]->sic;
        const double share = gdi ? (double)s.DirectTotal / (double)gdi : 1.0 / n;
        nf_task_stats &st = g[i]->stats;
        st.device_thread_s = wall;
        st.fill_s = std::chrono::duration<double>(t_fill - t0).count();
        st.kernel_s = kern_ms * 1e-3 * share;
        st.direct_kernel_s = direct_ms * 1e-3 * share;
        st.copyback_s = copyback * share;
        st.bytes_to_device = (uint64_t)((double)to_dev * share);
        st.bytes_from_device = (uint64_t)((double)from_dev * share);
        st.direct_total = s.DirectTotal;
        st.padded_direct_total = s.PaddedDirectTotal;
        st.assigned_buffer = w.id;
        st.path = path;
        st.group_size = n;
        st.group_direct_kernel_s = direct_ms * 1e-3;
        st.group_direct_total = gdi;
        st.graph = use_graph ? 1 : 0;
        st.upload_s = upload_ms * 1e-3 * share;
        st.download_s = download_ms * 1e-3 * share;
        st.slabs_uploaded = n_up;
        st.slabs_reused = n_reuse;
        st.sinks_per_lane = path == NF_PATH_BATCHED ? ts.sinks_per_lane : 0;
        st.autotune_trial = trial ? 1 : 0;
        st.waves_per_eu = path == NF_PATH_BATCHED ? batched_kernel_waves(s0.softening, ts.sinks_per_lane) : 0;
        st.waves_per_wg = path == NF_PATH_BATCHED ? ts.waves_per_wg : 0;
        if (census) {
            const uint64_t *c = static_cast<const uint64_t *>(w.pin_census.p);
            for (int f = 0; f < 3; f++) st.group_form_interactions[f] = c[f];
            st.group_tail_interactions = c[3];
        }
        st.device_planned = g[i]->dp.on ? 1 : 0;
        st.plan_s = g[i]->dp.on ? std::chrono::duration<double>(t_plan - t0).count() : 0.0;
    }
    {
        std::lock_guard<std::mutex> lk(rt.tp.m);
        for (int i = 0; i < n; i++) {
            rt.tp.direct += g[i]->sic.DirectTotal;
            rt.tp.padded += g[i]->sic.PaddedDirectTotal;
        }
        rt.tp.to_dev += to_dev;
        rt.tp.from_dev += from_dev;
        rt.tp.tasks += (uint64_t)n;
        rt.tp.groups++;
    }
    return NF_OK;
}

void finish(Task &t, int rc) {
    {
        std::lock_guard<std::mutex> lk(t.m);
        if (rc != NF_OK) t.error = g_last_error;
        t.state.store(rc == NF_OK ? 1 : (rc < 0 ? rc : NF_ERR_HIP), std::memory_order_release);
    }
    t.cv.notify_all();
}

// SICs that may share one launch: same softening constants and memory kind.
bool compatible(const nf_sic &a, const nf_sic &b) {
    return a.softening == b.softening && a.eps == b.eps && a.b2 == b.b2 && a.mem == b.mem && a.cpd == b.cpd;
}

// How long a worker waits for a full group while earlier work is in flight
// (so the wait never leaves the GPU idle).  On a clustered 256^3 box 2 ms
// gave fuller groups than 1 ms (11.2 launches per step against 12, +0.6%)
// and 0.5 ms fewer; a uniform box's groups fill at once either way.
constexpr int kGroupFillWaitUs = 2000;

void worker_main(Runtime *rt, Worker *w) {
    for (;;) {
        std::vector<Task *> g;
        {
            std::unique_lock<std::mutex> lk(rt->qm);
            if (rt->queue.empty()) rt->warm = false;
            rt->qcv.wait(lk, [&] { return rt->stopping || !rt->queue.empty(); });
            if (rt->queue.empty()) return;  // stopping and drained
            // With work in flight, wait a little for a whole group's worth of
            // SICs: a launch of one or two small SICs runs at ~2/3 of the
            // kernel's rate (its longest items set its length), and the GPU
            // has a group to run meanwhile.
            if (rt->params.path == NF_PATH_BATCHED && rt->inflight > 0) {
                const auto deadline = std::chrono::steady_clock::now() + std::chrono::microseconds(kGroupFillWaitUs);
                auto full = [&] {
                    if (rt->stopping || rt->queue.empty() || (int)rt->queue.size() >= rt->max_group) return true;
                    uint64_t work = 0;
                    for (const Task *t : rt->queue) work += queued_interactions(*t);
                    return work >= rt->group_interactions;
                };
                while (!full() && rt->qcv.wait_until(lk, deadline) != std::cv_status::timeout) {
                }
                if (rt->queue.empty()) continue;  // another worker took them
            }
            g.push_back(rt->queue.front());
            rt->queue.pop_front();
            // Device-planned SICs are grouped with each other only, on
            // bounds of their sizes (the exact ones come with phase A).
            const int path = rt->params.path;
            size_t bytes = queued_bytes(*g[0], path);
            // After the queue ran dry the GPU is idle: start it on a single SIC
            // instead of waiting for a whole group to be prepared.
            const int limit = !rt->warm && rt->inflight == 0 ? 1 : rt->max_group;
            rt->warm = true;
            rt->inflight++;
            uint64_t work = queued_interactions(*g[0]);
            while (path == NF_PATH_BATCHED && (int)g.size() < limit && work < rt->group_interactions &&
                   !rt->queue.empty() && rt->queue.front()->dp.on == g[0]->dp.on &&
                   compatible(g[0]->sic, rt->queue.front()->sic)) {
                const size_t b = queued_bytes(*rt->queue.front(), path);
                if (bytes + b > rt->max_buffer_bytes) break;
                bytes += b;
                work += queued_interactions(*rt->queue.front());
                g.push_back(rt->queue.front());
                rt->queue.pop_front();
            }
        }
        // Invalid descriptors fail alone; the rest of the group runs.
        std::vector<Task *> ok;
        for (Task *t : g) {
            const int rc = t->dp.on ? NF_OK : validate(t->sic);  // nf_submit_slab checked its own
            if (rc) finish(*t, rc);
            else ok.push_back(t);
        }
        const int rc = ok.empty() ? NF_OK : run_group(*rt, *w, ok);
        {
            std::lock_guard<std::mutex> lk(rt->qm);
            rt->inflight--;
        }
        for (Task *t : ok) finish(*t, rc);
    }
}

// With a shared compute stream (gpu_streams = 1) the workers also share one
// planning stream, so the runtime holds four streams in all (compute, upload,
// download, planning): as many as HIP's hardware queues per process, so no
// stream queues behind another's work.  Otherwise each worker has its own.
int init_worker(Worker &w, hipStream_t shared_pstream) {
    NF_HIP(hipSetDevice(w.device));
    w.own_streams = shared_pstream == nullptr;
    if (w.own_streams) {
        NF_HIP(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
        NF_HIP(hipStreamCreateWithFlags(&w.pstream, hipStreamNonBlocking));
    } else {
        w.pstream = shared_pstream;
    }
    for (auto &e : w.ev) NF_HIP(hipEventCreate(&e));
    NF_HIP(hipEventCreateWithFlags(&w.plan_ready, hipEventDisableTiming));
    NF_HIP(hipEventCreateWithFlags(&w.gplan_ready, hipEventDisableTiming));
    w.ready = true;
    return NF_OK;
}

void destroy_worker(Worker &w) {
    if (!w.ready) return;
    (void)hipSetDevice(w.device);
    if (w.stream) (void)hipStreamSynchronize(w.stream);
    if (w.pstream) (void)hipStreamSynchronize(w.pstream);
    w.plan.release();
    w.gscratch.release();
    w.pin_tot.release();
    w.pin_gplan.release();
    w.pin_census.release();
    w.work.release();
    w.slabs.release();
    w.pin_plan.release();
    w.pin_acc.release();
    w.pin_far.release();
    if (w.gexec) (void)hipGraphExecDestroy(w.gexec);
    w.gexec = nullptr;
    for (auto &e : w.ev)
        if (e) (void)hipEventDestroy(e);
    if (w.plan_ready) (void)hipEventDestroy(w.plan_ready);
    if (w.gplan_ready) (void)hipEventDestroy(w.gplan_ready);
    if (w.own_streams) {
        if (w.pstream) (void)hipStreamDestroy(w.pstream);
        if (w.stream) (void)hipStreamDestroy(w.stream);
    }
    w.stream = w.pstream = nullptr;
    w.ready = false;
}

}  // namespace
}  // namespace nf

using namespace nf;

extern "C" const char *nf_last_error(void) { return g_last_error.c_str(); }

extern "C" int nf_get_ngpu(int32_t *n) {
    if (!n) return fail(NF_ERR_INVALID, "null argument");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *n = 0;
        return fail(NF_ERR_NODEV, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *n = c;
    return NF_OK;
}

extern "C" int nf_get_device_info(uint64_t *bytes, char *name, size_t name_len) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(NF_ERR_NODEV, "no HIP device");
    uint64_t mem = UINT64_MAX;
    hipDeviceProp_t p{};
    for (int g = 0; g < n; g++) {
        NF_HIP(hipGetDeviceProperties(&p, g));
        mem = std::min<uint64_t>(mem, (uint64_t)(0.8 * (double)p.totalGlobalMem));
    }
    if (bytes) *bytes = mem;
    if (name && name_len) {
        std::snprintf(name, name_len, "%s", p.gcnArchName);
    }
    return NF_OK;
}

extern "C" int nf_setup(const nf_setup_params *p, int32_t *max_sink_blocks, int32_t *max_source_blocks) {
    std::lock_guard<std::mutex> lk(g_rt_mutex);
    if (g_rt) return fail(NF_ERR_STATE, "nf_setup called twice without nf_reset");
    if (!p) return fail(NF_ERR_INVALID, "null params");
    if (p->n_gpu != 1) return fail(NF_ERR_INVALID, "this build drives exactly one GPU per process (n_gpu = 1)");
    if (p->buffers_per_device < 1 || p->buffers_per_device > 16)
        return fail(NF_ERR_INVALID, "buffers_per_device must be 1..16");
    if (p->path != NF_PATH_PLAIN && p->path != NF_PATH_BATCHED) return fail(NF_ERR_INVALID, "bad path");
    int ng = 0;
    if (hipGetDeviceCount(&ng) != hipSuccess || ng < 1) return fail(NF_ERR_NODEV, "no HIP device");
    auto rt = std::make_unique<Runtime>();
    rt->params = *p;
    rt->ngpu = 1;
    rt->max_buffer_bytes = p->max_buffer_bytes;
    if (p->max_group < 0 || p->max_group > 4096) return fail(NF_ERR_INVALID, "max_group must be 0..4096");
    rt->max_group = p->max_group ? p->max_group : 8;
    if (p->group_interactions) rt->group_interactions = p->group_interactions;
    if (p->sinks_per_lane != 0 && p->sinks_per_lane != 1 && p->sinks_per_lane != 2 && p->sinks_per_lane != 4)
        return fail(NF_ERR_INVALID, "sinks_per_lane must be 0, 1, 2 or 4");
    if (p->waves_per_wg != 0 && p->waves_per_wg != 1 && p->waves_per_wg != 2 && p->waves_per_wg != 4)
        return fail(NF_ERR_INVALID, "waves_per_wg must be 0, 1, 2 or 4");
    if (p->sinks_per_lane) rt->shape.sinks_per_lane = p->sinks_per_lane;
    if (p->waves_per_wg) rt->shape.waves_per_wg = p->waves_per_wg;
    // An axis left at 0 is searched by the autotuner per occupancy histogram.
    rt->candidates = tile_candidates(p->sinks_per_lane, p->waves_per_wg);
    rt->autotune = rt->candidates.size() > 1;
    if (p->source_tile != 0 && p->source_tile != 64) return fail(NF_ERR_INVALID, "source_tile must be 0 or 64");
    // Capacity per SIC, from the bytes a sink block brings with it (the
    // reference's estimate, GPUControl.cu:367-395, with this module's sizes):
    // 16 B sink + 32 B partial + W*(16 B source + 0.5 B box) per particle slot,
    // plus plan bytes per block.
    const double W = 2 * NF_MAX_RADIUS + 1;
    const double per_block = NF_BLOCK_SIZE * (16.0 + 32.0 + W * 16.5) + 4.0 + 8.0 * W + 64.0;
    const double usable = (double)p->max_buffer_bytes - 64.0 * (1 << 20);
    const double msb = usable > 0 ? usable / per_block : 0.0;
    const int32_t ms = (int32_t)std::min(msb, (double)(INT32_MAX / NF_BLOCK_SIZE / 8));
    if (ms < 2 * p->cpd) return fail(NF_ERR_CAPACITY, "max_buffer_bytes too small for one skewer per SIC");
    if (max_sink_blocks) *max_sink_blocks = ms;
    if (max_source_blocks) *max_source_blocks = (int32_t)std::min<double>((double)ms * W, INT32_MAX / NF_BLOCK_SIZE);
    if (p->gpu_streams != 0 && p->gpu_streams != 1) return fail(NF_ERR_INVALID, "gpu_streams must be 0 or 1");
    // Streams first (compute, upload, download, planning: the process's
    // pooled streams), then the workers.
    auto destroy_streams = [&] { rt->shared = rt->pstream = rt->cache.up = rt->cache.down = nullptr; };
    {
        hipError_t e = hipSuccess;
        if (p->gpu_streams == 1) e = pooled_stream(&g_streams.compute, &rt->shared);
        auto &C = rt->cache;
        if (e == hipSuccess) e = pooled_stream(&g_streams.up, &C.up);
        if (e == hipSuccess) e = pooled_stream(&g_streams.down, &C.down);
        if (e == hipSuccess && p->gpu_streams == 1) e = pooled_stream(&g_streams.plan, &rt->pstream);
        for (int k = 0; k < Runtime::kStages && e == hipSuccess; k++)
            e = hipEventCreateWithFlags(&C.stage_done[k], hipEventDisableTiming);
        if (e != hipSuccess) {
            destroy_streams();
            return fail(NF_ERR_HIP, std::string("stream setup: ") + hipGetErrorString(e));
        }
    }
    for (int i = 0; i < p->buffers_per_device; i++) {
        auto w = std::make_unique<Worker>();
        w->id = i;
        w->device = 0;
        int rc = init_worker(*w, rt->pstream);
        if (rc) {
            for (auto &x : rt->workers) destroy_worker(*x);
            destroy_worker(*w);
            destroy_streams();
            return rc;
        }
        rt->workers.push_back(std::move(w));
    }
    for (auto &w : rt->workers) w->th = std::thread(worker_main, rt.get(), w.get());
    g_rt = rt.release();
    return NF_OK;
}

extern "C" int nf_reset(void) {
    std::lock_guard<std::mutex> lk(g_rt_mutex);
    if (!g_rt) return fail(NF_ERR_STATE, "nf_reset without nf_setup");
    {
        std::lock_guard<std::mutex> q(g_rt->qm);
        g_rt->stopping = true;
    }
    g_rt->qcv.notify_all();
    for (auto &w : g_rt->workers)
        if (w->th.joinable()) w->th.join();
    for (auto &w : g_rt->workers) destroy_worker(*w);
    // The pooled streams stay for the next nf_setup, drained.
    for (hipStream_t st : {g_rt->shared, g_rt->pstream})
        if (st) (void)hipStreamSynchronize(st);
    {
        auto &C = g_rt->cache;
        for (hipStream_t st : {C.up, C.down})
            if (st) (void)hipStreamSynchronize(st);
        for (auto &sp : C.slots) {
            sp->buf.release();
            for (hipEvent_t ev : {sp->ready, sp->t0, sp->t1})
                if (ev) (void)hipEventDestroy(ev);
        }
        for (int k = 0; k < Runtime::kStages; k++) {
            C.stage[k].release();
            if (C.stage_done[k]) (void)hipEventDestroy(C.stage_done[k]);
        }
    }
    delete g_rt;
    g_rt = nullptr;
    return NF_OK;
}

extern "C" int nf_submit(const nf_sic *sic, int32_t blocking, nf_task **task) {
    if (!sic || !task) return fail(NF_ERR_INVALID, "null argument");
    Runtime *rt;
    {
        std::lock_guard<std::mutex> lk(g_rt_mutex);
        rt = g_rt;
    }
    if (!rt) return fail(NF_ERR_STATE, "nf_submit before nf_setup");
    Task *t = new Task();
    t->sic = *sic;
    *task = reinterpret_cast<nf_task *>(t);
    {
        std::lock_guard<std::mutex> q(rt->qm);
        rt->queue.push_back(t);
    }
    rt->qcv.notify_one();
    if (blocking) return nf_wait(*task);
    return NF_OK;
}

namespace {

// The checks nf_plan_sic makes of its arguments.
int check_plan_args(int32_t cpd, int32_t R, int32_t j_low, int32_t j_high, const nf_cellinfo *const *cellinfo) {
    if (!cellinfo) return fail(NF_ERR_INVALID, "null cellinfo");
    if (cpd <= 0 || (cpd % 2) == 0) return fail(NF_ERR_INVALID, "cpd must be odd and positive");
    if (R < 0 || R > NF_MAX_RADIUS || 2 * R + 1 > cpd) return fail(NF_ERR_INVALID, "nfradius out of range");
    if (j_low < 0 || j_high > cpd || j_low >= j_high) return fail(NF_ERR_INVALID, "bad sink row range");
    if ((int64_t)(j_high - j_low + 2 * R) * cpd * (2 * R + 1) >= INT32_MAX)
        return fail(NF_ERR_CAPACITY, "too many pencils for int32 indices");
    for (int c = 0; c < 2 * R + 1; c++)
        if (!cellinfo[c]) return fail(NF_ERR_INVALID, "missing cellinfo slab");
    return NF_OK;
}

}  // namespace

extern "C" int nf_submit_slab(const nf_slab_task *p, int32_t blocking, nf_task **task) {
    if (!p || !task) return fail(NF_ERR_INVALID, "null argument");
    Runtime *rt;
    {
        std::lock_guard<std::mutex> lk(g_rt_mutex);
        rt = g_rt;
    }
    if (!rt) return fail(NF_ERR_STATE, "nf_submit_slab before nf_setup");
    int rc = check_plan_args(p->cpd, p->nfradius, p->j_low, p->j_high, p->cellinfo);
    if (rc) return rc;
    if (p->cpd != rt->params.cpd) return fail(NF_ERR_INVALID, "cpd differs from nf_setup's");
    const int R = p->nfradius, W = 2 * R + 1;
    if (p->mem != NF_MEM_HOST && p->mem != NF_MEM_DEVICE) return fail(NF_ERR_INVALID, "bad mem kind");
    if (p->pos_layout != NF_POS_XYZ && p->pos_layout != NF_POS_AOS) return fail(NF_ERR_INVALID, "bad pos_layout");
    if (!p->SinkAccSlab) return fail(NF_ERR_INVALID, "missing acceleration slab");
    for (int c = 0; c < W; c++)
        if (!p->PosSlab[c]) return fail(NF_ERR_INVALID, "missing position slab");
    if (p->kick && (!p->kick->FarAccSlab || !std::isfinite(p->kick->rescale) || !std::isfinite(p->kick->inv_eps3) ||
                    !std::isfinite(p->kick->density_rad2)))
        return fail(NF_ERR_INVALID, "kick fusion needs a far-field slab and finite constants");
    PairConst pc;
    if ((rc = derive_pair_const(p->softening, p->eps, p->b2, &pc))) return rc;
    Task *t = new Task();
    nf_sic &s = t->sic;
    sic_header(s, p->cpd, R, p->SlabId, p->j_low, p->j_high, p->eps, p->b2, p->softening);
    s.mem = p->mem;
    s.pos_layout = p->pos_layout;
    s.kick = p->kick;
    s.SinkPosSlab = p->PosSlab[R];
    for (int c = 0; c < W; c++) {
        s.SourcePosSlab[c] = p->PosSlab[c];
        s.Nslab[c] = p->Nslab[c];
    }
    s.SinkAccSlab = p->SinkAccSlab;
    t->dp.on = true;
    for (int c = 0; c < W; c++) t->dp.cellinfo[c] = p->cellinfo[c];
    // The SIC's exact pair count sizes its launch group and the heavy-tile
    // split before the device plans it (a bad count fails the task in phase A).
    int64_t max_pencil = 0;
    if (count_interactions(p->cpd, R, p->j_low, p->j_high, p->cellinfo, false, &t->dp.interactions, &max_pencil) !=
        NF_OK)
        t->dp.interactions = 0;
    else
        t->dp.max_src_blocks = (int32_t)std::min<int64_t>(INT32_MAX, (max_pencil + NF_BLOCK_SIZE - 1) / NF_BLOCK_SIZE);
    // The sink rows run from cell (j_low, 0) to the end of cell (j_high-1,
    // cpd-1); phase A checks both cells against the slab.
    const nf_cellinfo &first = p->cellinfo[R][(size_t)p->j_low * p->cpd];
    const nf_cellinfo &last = p->cellinfo[R][(size_t)(p->j_high - 1) * p->cpd + (p->cpd - 1)];
    t->dp.r0 = (uint32_t)first.startindex;
    t->dp.r1 = (uint64_t)(uint32_t)last.startindex + (uint64_t)std::max(last.count, 0);
    *task = reinterpret_cast<nf_task *>(t);
    {
        std::lock_guard<std::mutex> q(rt->qm);
        rt->queue.push_back(t);
    }
    rt->qcv.notify_one();
    if (blocking) return nf_wait(*task);
    return NF_OK;
}

extern "C" int nf_plan_sic_device(int32_t cpd, int32_t R, int32_t slab, int32_t j_low, int32_t j_high,
                                  const nf_cellinfo *const *cellinfo, const float *const *pos_slab,
                                  const uint64_t *nslab, float eps, float b2, int32_t softening, void *arena_ptr,
                                  uint64_t arena_bytes, nf_sic *out) {
    if (!out || !arena_ptr || !nslab) return fail(NF_ERR_INVALID, "nf_plan_sic_device: null argument");
    int rc = check_plan_args(cpd, R, j_low, j_high, cellinfo);
    if (rc) return rc;
    const int W = 2 * R + 1;
    nf_sic s;
    sic_header(s, cpd, R, slab, j_low, j_high, eps, b2, softening);
    s.mem = NF_MEM_HOST;
    if (pos_slab) {
        s.SinkPosSlab = pos_slab[R];
        for (int c = 0; c < W; c++) s.SourcePosSlab[c] = pos_slab[c];
    }
    for (int c = 0; c < W; c++) s.Nslab[c] = nslab[c];

    // Device buffers: the cellinfo slabs, the totals and scratch, the arrays.
    struct Bufs {
        void *cells = nullptr, *scr = nullptr, *plan = nullptr;
        hipStream_t st = nullptr;
        ~Bufs() {
            if (st) (void)hipStreamSynchronize(st);
            if (cells) (void)hipFree(cells);
            if (scr) (void)hipFree(scr);
            if (plan) (void)hipFree(plan);
            if (st) (void)hipStreamDestroy(st);
        }
    } b;
    const size_t cbytes = sizeof(nf_cellinfo) * (size_t)cpd * (size_t)cpd;
    NF_HIP(hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking));
    NF_HIP(hipMalloc(&b.cells, cbytes * W));
    NF_HIP(hipMalloc(&b.scr, 256 + sizeof(int32_t) * gpu_plan_part_ints(s.NSinkSets, s.NSourceSets)));
    GPlanIn in{};
    for (uint64_t &c : in.split_cost) c = UINT64_MAX;  // (no work lists here)
    in.cpd = cpd;
    in.R = R;
    in.W = W;
    in.jw = s.j_width;
    in.j_low = j_low;
    for (int c = 0; c < W; c++) {
        in.ci[c] = reinterpret_cast<const nf_cellinfo *>(static_cast<char *>(b.cells) + cbytes * c);
        in.nslab[c] = nslab[c];
        NF_HIP(hipMemcpyAsync(const_cast<nf_cellinfo *>(in.ci[c]), cellinfo[c], cbytes, hipMemcpyHostToDevice, b.st));
    }
    GPlanTotals *tot_d = static_cast<GPlanTotals *>(b.scr);
    int32_t *part = reinterpret_cast<int32_t *>(static_cast<char *>(b.scr) + 256);
    int e = launch_gpu_plan_a(in, tot_d, nullptr, nullptr, b.st);
    if (e) return fail(NF_ERR_HIP, std::string("device plan launch: ") + hipGetErrorString((hipError_t)e));
    GPlanTotals tot;
    NF_HIP(hipMemcpyAsync(&tot, tot_d, sizeof(tot), hipMemcpyDeviceToHost, b.st));
    NF_HIP(hipStreamSynchronize(b.st));
    if ((rc = apply_totals(s, tot))) return rc;
    const PlanLayout L = plan_layout(s, 0, (int32_t)tot.chunks);
    NF_HIP(hipMalloc(&b.plan, L.total));
    const GPlanOut go = plan_out(static_cast<char *>(b.plan), L, 0, nullptr, part);
    e = launch_gpu_plan_b1(in, go, b.st);
    if (!e) e = launch_gpu_plan_b2(in, go, b.st);
    if (e) return fail(NF_ERR_HIP, std::string("device plan launch: ") + hipGetErrorString((hipError_t)e));

    // The arrays, carved from the arena as nf_plan_sic carves them.
    char *ap = static_cast<char *>(arena_ptr);
    uint64_t left = arena_bytes;
    auto take = [&](void **p, uint64_t bytes) {
        const uint64_t mis = reinterpret_cast<uintptr_t>(ap) % 4096;
        const uint64_t extra = mis ? 4096 - mis : 0;
        if (left < bytes + extra) return false;
        *p = ap + extra;
        ap += extra + bytes;
        left -= extra + bytes;
        return true;
    };
    struct Part {
        void **dst;
        size_t bytes, off;
    };
    const size_t nss = (size_t)s.NSinkSets, nqs = (size_t)s.NSourceSets, ic = (size_t)s.InteractionCount;
    const Part parts[] = {
        {(void **)&s.SinkSetStart, 4 * nss, L.o_sink_start},
        {(void **)&s.SinkSetCount, 4 * nss, L.o_sink_count},
        {(void **)&s.SinkPlan, sizeof(nf_sink_plan) * nss, L.o_sink_plan},
        {(void **)&s.SinkSetIdMax, 4 * nss, L.o_idmax},
        {(void **)&s.SourceSetStart, 4 * nqs, L.o_src_start},
        {(void **)&s.SourceSetCount, 4 * nqs, L.o_src_count},
        {(void **)&s.SourcePlan, sizeof(nf_source_plan) * nqs, L.o_src_plan},
        {(void **)&s.SinkBlockParentPencil, 4 * std::max<size_t>((size_t)s.NSinkBlocks, 1), L.o_parent},
        {(void **)&s.SinkSourceInteractionList, 4 * ic, L.o_ilist},
        {(void **)&s.SinkSourceYOffset, 4 * ic, L.o_yoff},
    };
    for (const Part &q : parts)
        if (!take(q.dst, q.bytes)) return fail(NF_ERR_CAPACITY, "nf_plan_sic_device: arena too small");
    for (const Part &q : parts) {
        const size_t nb = q.dst == (void **)&s.SinkBlockParentPencil ? 4 * (size_t)s.NSinkBlocks : q.bytes;
        if (nb) NF_HIP(hipMemcpyAsync(*q.dst, static_cast<char *>(b.plan) + q.off, nb, hipMemcpyDeviceToHost, b.st));
    }
    NF_HIP(hipStreamSynchronize(b.st));
    *out = s;
    return NF_OK;
}

extern "C" int nf_check_completion(nf_task *task) {
    if (!task) return fail(NF_ERR_INVALID, "null task");
    Task *t = reinterpret_cast<Task *>(task);
    const int s = t->state.load(std::memory_order_acquire);
    if (s < 0) fail(s, t->error);
    return s;
}

extern "C" int nf_wait(nf_task *task) {
    if (!task) return fail(NF_ERR_INVALID, "null task");
    Task *t = reinterpret_cast<Task *>(task);
    std::unique_lock<std::mutex> lk(t->m);
    t->cv.wait(lk, [&] { return t->state.load(std::memory_order_acquire) != 0; });
    const int s = t->state.load(std::memory_order_acquire);
    if (s < 0) return fail(s, t->error);
    return NF_OK;
}

extern "C" int nf_get_task_stats(nf_task *task, nf_task_stats *out) {
    if (!task || !out) return fail(NF_ERR_INVALID, "null argument");
    Task *t = reinterpret_cast<Task *>(task);
    if (t->state.load(std::memory_order_acquire) == 0) return fail(NF_ERR_STATE, "task not complete");
    *out = t->stats;
    return NF_OK;
}

extern "C" int nf_release(nf_task *task) {
    if (!task) return fail(NF_ERR_INVALID, "null task");
    Task *t = reinterpret_cast<Task *>(task);
    if (t->state.load(std::memory_order_acquire) == 0) return fail(NF_ERR_STATE, "task still running");
    delete t;
    return NF_OK;
}

extern "C" int nf_device_alloc(uint64_t bytes, void **ptr) {
    if (!ptr) return fail(NF_ERR_INVALID, "null argument");
    *ptr = nullptr;
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(NF_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return NF_OK;
}

extern "C" int nf_device_free(void *ptr) {
    if (ptr) NF_HIP(hipFree(ptr));
    return NF_OK;
}

extern "C" int nf_device_copy(void *dst, const void *src, uint64_t bytes, int32_t kind) {
    if (bytes == 0) return NF_OK;
    if (!dst || !src) return fail(NF_ERR_INVALID, "null argument");
    hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                      : hipMemcpyDeviceToDevice;
    if (kind < 0 || kind > 2) return fail(NF_ERR_INVALID, "bad copy kind");
    NF_HIP(hipMemcpy(dst, src, bytes, k));
    return NF_OK;
}

extern "C" int nf_device_synchronize(void) {
    NF_HIP(hipDeviceSynchronize());
    return NF_OK;
}

extern "C" int nf_host_alloc(uint64_t bytes, void **ptr) {
    if (!ptr) return fail(NF_ERR_INVALID, "null argument");
    *ptr = nullptr;
    hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(NF_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    }
    return NF_OK;
}

extern "C" int nf_host_free(void *ptr) {
    if (ptr) NF_HIP(hipHostFree(ptr));
    return NF_OK;
}

extern "C" int nf_get_throughput(nf_throughput *out) {
    if (!out) return fail(NF_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(g_rt_mutex);
    if (!g_rt) return fail(NF_ERR_STATE, "nf_get_throughput before nf_setup");
    auto &t = g_rt->tp;
    std::lock_guard<std::mutex> tl(t.m);
    out->gpu_active_s = t.active_s + (t.active ? std::chrono::duration<double>(clk::now() - t.since).count() : 0.0);
    out->direct_total = t.direct;
    out->padded_direct_total = t.padded;
    out->bytes_to_device = t.to_dev;
    out->bytes_from_device = t.from_dev;
    out->tasks = t.tasks;
    out->groups = t.groups;
    return NF_OK;
}

extern "C" int nf_autotune_table(char *buf, size_t len) {
    std::lock_guard<std::mutex> lk(g_rt_mutex);
    if (!g_rt) return fail(NF_ERR_STATE, "nf_autotune_table before nf_setup");
    std::ostringstream os;
    os << "[";
    {
        std::lock_guard<std::mutex> tk(g_rt->tune_m);
        bool first = true;
        for (const auto &kv : g_rt->tune) {
            const auto &e = kv.second;
            const size_t semi = kv.first.find(';');
            os << (first ? "" : ",") << "{\"occupancy_quarters\":[" << kv.first.substr(0, semi)
               << "],\"size_class\":" << kv.first.substr(semi + 1) << ",\"sinks_per_lane\":"
               << e.shape.sinks_per_lane << ",\"waves_per_wg\":" << e.shape.waves_per_wg
               << ",\"waves_per_cu\":" << e.shape.waves_per_cu << ",\"uses\":" << e.uses
               << ",\"interactions\":" << e.interactions << ",\"trial_ms\":[";
            for (size_t c = 0; c < e.trial_ms.size(); c++) os << (c ? "," : "") << e.trial_ms[c];
            os << "],\"candidates\":[";
            for (size_t c = 0; c < g_rt->candidates.size(); c++)
                os << (c ? "," : "") << "[" << g_rt->candidates[c].sinks_per_lane << ","
                   << g_rt->candidates[c].waves_per_wg << "]";
            os << "]}";
            first = false;
        }
    }
    os << "]";
    const std::string out = os.str();
    if (buf && len) std::snprintf(buf, len, "%s", out.c_str());
    return (int)out.size();
}

extern "C" int nf_stream_begin(void) {
    std::lock_guard<std::mutex> lk(g_rt_mutex);
    if (!g_rt) return fail(NF_ERR_STATE, "nf_stream_begin before nf_setup");
    std::lock_guard<std::mutex> ck(g_rt->cache.m);
    g_rt->cache.active = true;
    g_rt->cache.epoch++;
    return NF_OK;
}

extern "C" int nf_stream_end(void) {
    std::lock_guard<std::mutex> lk(g_rt_mutex);
    if (!g_rt) return fail(NF_ERR_STATE, "nf_stream_end before nf_setup");
    std::lock_guard<std::mutex> ck(g_rt->cache.m);
    g_rt->cache.active = false;
    for (auto &sp : g_rt->cache.slots)
        if (sp->refs == 0) sp->valid = false;  // slabs in use are dropped on release
    return NF_OK;
}

// ---- The entry point that allocates nothing ----------------------------

namespace nf {
namespace {

// Device scratch of nf_execute_sic_noalloc: a SIC's plan image (view, work
// list, queue counter, plan arrays), its source stream with the block boxes,
// and its FP64 partials.  Each GPU holds its own copy from the library's
// first use on it until the process ends.  In host code `scratch` names a
// host-side stand-in, never the device memory: kernels, copies and fills
// take the address hipGetSymbolAddress gives for the current device.
__device__ __align__(256) char scratch[NF_NOALLOC_SCRATCH_BYTES];

constexpr int kNoallocMaxDevices = 64;
char *g_noalloc_dev[kNoallocMaxDevices];  // per device: the address of `scratch` there

// Host image of the upload (view, work list, plan arrays) and the counting
// sort's cursors, static so that a call allocates no host memory either.
constexpr size_t kNoallocHostBytes = 32u << 20;
alignas(256) char g_noalloc_host[kNoallocHostBytes];
uint32_t g_noalloc_cursor[2 * kItemBuckets + 1];

// Calls share the scratch of their GPU and the host image: one at a time.
std::mutex g_noalloc_m;

// Where one SIC sits in the scratch: [view][items][counter][plan arrays]
// (the uploaded image) then [source stream + boxes][partials].
struct NoallocLayout {
    size_t o_items, o_counter, o_plan, image, o_work, total;
    size_t items_max;
    PlanLayout plan;
    WorkLayout work;
};

NoallocLayout noalloc_layout(const nf_sic &s) {
    NoallocLayout L{};
    // At most one item per padded sink block plus one tail per pencil,
    // whatever the tile split.
    L.items_max = (size_t)s.NSinkBlocks + (size_t)s.NSinkSets;
    L.o_items = align_up(sizeof(SicView), 256);
    L.o_counter = align_up(L.o_items + sizeof(WorkItem) * std::max<size_t>(L.items_max, 1), 256);
    L.o_plan = L.o_counter + 256;
    L.plan = plan_layout(s, L.o_plan, count_coadd_chunks(s));
    L.image = L.o_plan + L.plan.total;
    L.o_work = align_up(L.image, 256);
    L.work = work_layout(s, NF_PATH_BATCHED, L.o_work);
    L.total = L.o_work + L.work.total;
    return L;
}

// The address of `scratch` on device `dev` (current), looked up once.
int noalloc_scratch(int dev, char **out) {
    if (dev < 0 || dev >= kNoallocMaxDevices) return fail(NF_ERR_NODEV, "device index beyond the scratch table");
    if (!g_noalloc_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(NF_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_noalloc_dev[dev] = static_cast<char *>(p);
    }
    *out = g_noalloc_dev[dev];
    return NF_OK;
}

}  // namespace
}  // namespace nf

extern "C" int nf_execute_sic_noalloc(const nf_sic *sic, void *stream) {
    if (!sic) return fail(NF_ERR_INVALID, "null argument");
    const nf_sic &s = *sic;
    if (s.mem != NF_MEM_DEVICE)
        return fail(NF_ERR_INVALID, "nf_execute_sic_noalloc takes device-resident slabs (mem = NF_MEM_DEVICE)");
    int rc = validate(s);
    if (rc) return rc;
    PairConst pc;
    if ((rc = derive_pair_const(s.softening, s.eps, s.b2, &pc))) return rc;
    const NoallocLayout L = noalloc_layout(s);
    if (L.total > NF_NOALLOC_SCRATCH_BYTES)
        return fail(NF_ERR_CAPACITY, "SIC needs " + std::to_string(L.total) + " bytes of device scratch; "
                                     "nf_execute_sic_noalloc holds " + std::to_string((uint64_t)NF_NOALLOC_SCRATCH_BYTES) +
                                     " (split the slab's rows into smaller SICs)");
    if (L.image > kNoallocHostBytes)
        return fail(NF_ERR_CAPACITY, "SIC plan image of " + std::to_string(L.image) + " bytes exceeds the " +
                                         std::to_string((uint64_t)kNoallocHostBytes) + "-byte host image");

    std::lock_guard<std::mutex> lk(g_noalloc_m);
    int dev = 0;
    NF_HIP(hipGetDevice(&dev));
    char *d = nullptr;
    if ((rc = noalloc_scratch(dev, &d))) return rc;
    char *h = g_noalloc_host;

    // Plan arrays and the lists derived from them.
    fill_plan_image(s, L.plan, h);
    const DevPlan pl = dev_plan(s, L.plan, d);

    // The work list, longest items first (items within 1/8 octave of cost
    // keep sink-pencil order), short tails spread through the tiles: a
    // counting sort by cost bucket in two passes, as the batched path's.
    constexpr int S = 2;  // every tile shape gives bitwise the same partials
    const uint64_t split = tile_split_cost(s.softening, S, s.DirectTotal);
    std::fill(g_noalloc_cursor, g_noalloc_cursor + 2 * kItemBuckets + 1, 0u);
    size_t nitems = 0;
    for_each_item(s, S, 0, split, [&](const WorkItem &w, uint16_t b) {
        g_noalloc_cursor[item_key(w, b) + 1]++;
        nitems++;
    });
    if (nitems > L.items_max) return fail(NF_ERR_STATE, "work list longer than its bound");
    for (int b = 0; b < 2 * kItemBuckets; b++) g_noalloc_cursor[b + 1] += g_noalloc_cursor[b];
    const int64_t M = g_noalloc_cursor[kItemBuckets], T = (int64_t)nitems - M;
    WorkItem *items = reinterpret_cast<WorkItem *>(h + L.o_items);
    for_each_item(s, S, 0, split, [&](const WorkItem &w, uint16_t b) {
        const int key = item_key(w, b);
        const int64_t r = g_noalloc_cursor[key]++;
        items[key < kItemBuckets ? interleave_pos(false, r, M, T) : interleave_pos(true, r - M, M, T)] = w;
    });

    DevSlabs sl{};
    for (int c = 0; c < s.nfwidth; c++) {
        sl.n[c] = s.Nslab[c];
        sl.pos[c] = s.SourcePosSlab[c];
    }
    sl.pos[s.nfradius] = s.SinkPosSlab;
    sl.aos = s.pos_layout == NF_POS_AOS ? 1 : 0;
    KickArgs ka;
    std::memset(&ka, 0, sizeof(ka));
    if (s.kick) {
        ka.on = 1;
        ka.rescale = (float)s.kick->rescale;
        ka.inv_eps3 = (float)s.kick->inv_eps3;
        ka.density_rad2 = s.kick->density_rad2;
        ka.far = s.kick->FarAccSlab;
        ka.near_out = s.kick->NearAccOut;
        ka.far_base = 0;
    }
    char *src = d + L.work.o_src;
    SicView &v = *reinterpret_cast<SicView *>(h);
    std::memset(&v, 0, sizeof(v));
    v.sink_set_start = pl.sink_set_start;
    v.src_set_start = pl.src_set_start;
    v.src_set_count = pl.src_set_count;
    v.ilist = pl.ilist;
    v.sink_plan = pl.sink_plan;
    v.sink_slab = s.SinkPosSlab;
    v.sink_n = s.Nslab[s.nfradius];
    v.aos = sl.aos;
    v.sources4 = src;
    v.box4 = box_after(src, s.NSourceBlocks);
    v.partial = reinterpret_cast<double *>(d + L.work.o_part);
    v.W = s.nfwidth;
    v.R = s.nfradius;

    // Upload, gather, one persistent launch, co-add, on the caller's stream;
    // the queue counter is cleared by a kernel inside launch_batched.  The
    // host image stays untouched until the stream has drained.
    hipStream_t st = static_cast<hipStream_t>(stream);
    NF_HIP(hipMemcpyAsync(d, h, L.image, hipMemcpyHostToDevice, st));
    int e = launch_gather(pl, sl, true, nullptr, src, st);
    if (!e)
        e = launch_batched(reinterpret_cast<const SicView *>(d), s.cpd, s.softening, pc, TileShape{S, 64, 4, 0},
                           reinterpret_cast<const WorkItem *>(d + L.o_items), (int32_t)nitems,
                           reinterpret_cast<int32_t *>(d + L.o_counter), st);
    if (!e) e = launch_coadd_f64(pl, d + L.work.o_part, pc.out_scale, s.SinkAccSlab, ka, st);
    const hipError_t se = hipStreamSynchronize(st);
    if (e) return fail(NF_ERR_HIP, std::string("nf_execute_sic_noalloc launch: ") + hipGetErrorString((hipError_t)e));
    if (se != hipSuccess) return fail(NF_ERR_HIP, std::string("nf_execute_sic_noalloc: ") + hipGetErrorString(se));
    return NF_OK;
}

extern "C" int nf_probe_pair_rate(float eps, float b2, int32_t sinks_per_lane, int32_t mode, int32_t waves_per_simd,
                                  double *interactions_per_s) {
    if (!interactions_per_s) return fail(NF_ERR_INVALID, "null argument");
    if ((sinks_per_lane != 1 && sinks_per_lane != 2 && sinks_per_lane != 4) || mode < 0 || mode > 2 ||
        (waves_per_simd != 4 && waves_per_simd != 6 && waves_per_simd != 8))
        return fail(NF_ERR_INVALID, "probe shape: sinks_per_lane 1, 2 or 4; mode 0..2; waves_per_simd 4, 6 or 8");
    PairConst pc;
    int rc = derive_pair_const(NF_SOFT_SINGLE_SPLINE, eps, b2, &pc);
    if (rc) return rc;
    const int e = run_pair_probe(pc, sinks_per_lane, mode, waves_per_simd, interactions_per_s);
    if (e) return fail(NF_ERR_HIP, std::string("pair probe: ") + hipGetErrorString((hipError_t)e));
    return NF_OK;
}
