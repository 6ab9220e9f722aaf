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

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
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
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    DevBuf plan, work, slabs;
    PinBuf pin_plan, pin_acc;
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
    TileShape shape{2, 64, 4, 0};
};

std::mutex g_rt_mutex;
Runtime *g_rt = nullptr;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Bytes of plan arrays uploaded per task (plus the batched work list).
struct PlanLayout {
    size_t o_sink_start, o_sink_count, o_idmax, o_sink_plan, o_src_start, o_src_count, o_src_plan, o_parent,
        o_ilist, o_yoff, o_items, o_counter, total;
};

PlanLayout plan_layout(const nf_sic &s, int32_t nitems) {
    PlanLayout L{};
    size_t o = 0;
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
    L.o_items = take(sizeof(WorkItem) * (size_t)std::max(nitems, 1));
    L.o_counter = take(256);
    L.total = o;
    return L;
}

// The batched work list: per sink pencil, full register tiles of 64*S sinks
// and one remainder item, ordered by interaction count (longest first).
std::vector<WorkItem> build_items(const nf_sic &s, int S) {
    std::vector<WorkItem> items;
    std::vector<uint64_t> cost;
    items.reserve((size_t)s.NSinkSets * 2);
    const int tile = NF_BLOCK_SIZE * S;
    const int W = s.nfwidth;
    for (int k = 0; k < s.NSinkSets; k++) {
        const int n = s.SinkSetCount[k];
        if (n == 0) continue;
        uint64_t nsrc = 0;
        for (int y = 0; y < W; y++) nsrc += (uint64_t)s.SourceSetCount[s.SinkSourceInteractionList[W * k + y]];
        int f = 0;
        for (; f + tile <= n; f += tile) {
            items.push_back(WorkItem{k, f, tile, 0});
            cost.push_back((uint64_t)tile * nsrc);
        }
        if (f < n) {
            items.push_back(WorkItem{k, f, n - f, 1});
            cost.push_back((uint64_t)(n - f) * nsrc);
        }
    }
    std::vector<size_t> order(items.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
    std::vector<WorkItem> out(items.size());
    for (size_t i = 0; i < order.size(); i++) out[i] = items[order[i]];
    return out;
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
    for (int b = 0; b < s.NSinkBlocks; b++)
        if (s.SinkBlockParentPencil[b] < 0 || s.SinkBlockParentPencil[b] >= s.NSinkSets)
            return fail(NF_ERR_INVALID, "SinkBlockParentPencil out of range");
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

// Execute one SIC on worker w (called on the worker's thread, or on the
// caller's thread for blocking submissions).
int run_task(Runtime &rt, Worker &w, Task &t) {
    const auto t0 = clk::now();
    const nf_sic &s = t.sic;
    NF_HIP(hipSetDevice(w.device));
    int rc = validate(s);
    if (rc) return rc;
    PairConst pc;
    rc = derive_pair_const(s.softening, s.eps, s.b2, &pc);
    if (rc) return rc;
    const int path = rt.params.path;
    const TileShape ts = rt.shape;

    std::vector<WorkItem> items;
    if (path == NF_PATH_BATCHED) items = build_items(s, ts.sinks_per_lane);
    const int32_t nitems = (int32_t)items.size();
    const PlanLayout L = plan_layout(s, nitems);

    // Device working set: sink stream, source stream + boxes, partials.
    const size_t psink = (size_t)s.NSinkBlocks * NF_BLOCK_SIZE;
    const size_t psrc = (size_t)s.NSourceBlocks * NF_BLOCK_SIZE;
    const size_t part_bytes = psink * (path == NF_PATH_BATCHED ? 32 : 16);
    const size_t o_sinks = 0;
    const size_t o_src = align_up(o_sinks + 16 * psink, 256);
    const size_t o_part = align_up(o_src + 16 * psrc + 32 * (size_t)s.NSourceBlocks, 256);
    const size_t work_bytes = align_up(o_part + part_bytes, 256);

    // Host mode: the W slabs and the sink acceleration slab on the device.
    size_t slab_off[NF_MAX_WIDTH] = {0};
    size_t slab_bytes = 0, o_acc = 0;
    const uint64_t nsink_slab = s.Nslab[s.nfradius];
    if (s.mem == NF_MEM_HOST) {
        for (int c = 0; c < s.nfwidth; c++) {
            slab_off[c] = slab_bytes;
            slab_bytes = align_up(slab_bytes + 12 * (size_t)s.Nslab[c], 256);
        }
        o_acc = slab_bytes;
        slab_bytes = align_up(slab_bytes + 16 * (size_t)nsink_slab, 256);
    }
    if (L.total + work_bytes + slab_bytes > rt.max_buffer_bytes)
        return fail(NF_ERR_CAPACITY, "SIC needs " + std::to_string(L.total + work_bytes + slab_bytes) +
                                         " device bytes, buffer holds " + std::to_string(rt.max_buffer_bytes));
    if ((rc = w.plan.ensure(L.total)) || (rc = w.work.ensure(work_bytes)) || (rc = w.pin_plan.ensure(L.total)))
        return rc;
    if (s.mem == NF_MEM_HOST && ((rc = w.slabs.ensure(slab_bytes)) || (rc = w.pin_acc.ensure(16 * nsink_slab))))
        return rc;

    // ---- Fill: plan arrays through pinned staging (and slabs in host mode).
    char *hp = static_cast<char *>(w.pin_plan.p);
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
    if (nitems) std::memcpy(hp + L.o_items, items.data(), sizeof(WorkItem) * (size_t)nitems);
    NF_HIP(hipEventRecord(w.ev[0], w.stream));
    NF_HIP(hipMemcpyAsync(w.plan.p, hp, L.total, hipMemcpyHostToDevice, w.stream));
    uint64_t to_dev = L.total, from_dev = 0;

    DevSlabs sl{};
    float *acc_dev = nullptr;
    if (s.mem == NF_MEM_HOST) {
        char *db = static_cast<char *>(w.slabs.p);
        for (int c = 0; c < s.nfwidth; c++) {
            NF_HIP(hipMemcpyAsync(db + slab_off[c], s.SourcePosSlab[c], 12 * (size_t)s.Nslab[c],
                                  hipMemcpyHostToDevice, w.stream));
            to_dev += 12 * s.Nslab[c];
            sl.pos[c] = reinterpret_cast<const float *>(db + slab_off[c]);
            sl.n[c] = s.Nslab[c];
        }
        acc_dev = reinterpret_cast<float *>(db + o_acc);
    } else {
        for (int c = 0; c < s.nfwidth; c++) {
            sl.pos[c] = s.SourcePosSlab[c];
            sl.n[c] = s.Nslab[c];
        }
        // The sink slab is one of the source slabs (c = R); use the SIC's own pointer.
        sl.pos[s.nfradius] = s.SinkPosSlab;
        acc_dev = s.SinkAccSlab;
    }

    const char *pd = static_cast<const char *>(w.plan.p);
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
    pl.nsink_sets = s.NSinkSets;
    pl.nsrc_sets = s.NSourceSets;
    pl.nsink_blocks = s.NSinkBlocks;
    pl.nsrc_blocks = s.NSourceBlocks;
    pl.W = s.nfwidth;
    pl.R = s.nfradius;
    pl.cpd = s.cpd;
    pl.Nk = s.Nk;
    pl.j_width = s.j_width;
    pl.j_low = s.j_low;

    char *wb = static_cast<char *>(w.work.p);
    void *sinks4 = wb + o_sinks, *src4 = wb + o_src, *part = wb + o_part;
    // Accelerations of sink rows [j_low, j_high) are written through the
    // first cell of the first row: the co-add indexes the slab by cell start.
    const auto t_fill = clk::now();

    NF_HIP(hipEventRecord(w.ev[1], w.stream));
    int e = launch_gather(pl, sl, path == NF_PATH_BATCHED, sinks4, src4, w.stream);
    if (e) return fail(NF_ERR_HIP, std::string("gather launch: ") + hipGetErrorString((hipError_t)e));
    NF_HIP(hipEventRecord(w.ev[2], w.stream));
    if (path == NF_PATH_BATCHED) {
        int32_t *counter = reinterpret_cast<int32_t *>(static_cast<char *>(w.plan.p) + L.o_counter);
        const WorkItem *ditems = reinterpret_cast<const WorkItem *>(pd + L.o_items);
        e = launch_batched(pl, s.softening, pc, ts, sinks4, src4, ditems, nitems, counter, part, w.stream);
    } else {
        e = launch_plain(pl, s.softening, pc, sinks4, src4, part, w.stream);
    }
    if (e) return fail(NF_ERR_HIP, std::string("direct kernel launch: ") + hipGetErrorString((hipError_t)e));
    NF_HIP(hipEventRecord(w.ev[3], w.stream));
    if (path == NF_PATH_BATCHED) e = launch_coadd_f64(pl, part, pc.out_scale, acc_dev, w.stream);
    else e = launch_coadd_plain(pl, part, acc_dev, w.stream);
    if (e) return fail(NF_ERR_HIP, std::string("co-add launch: ") + hipGetErrorString((hipError_t)e));

    hipEvent_t evd = nullptr;
    NF_HIP(hipEventCreate(&evd));
    NF_HIP(hipEventRecord(evd, w.stream));
    NF_HIP(hipEventSynchronize(evd));
    const double kern_ms = event_ms(w.ev[1], evd);
    const double direct_ms = event_ms(w.ev[2], w.ev[3]);
    (void)hipEventDestroy(evd);

    double copyback = 0.0;
    if (s.mem == NF_MEM_HOST) {
        // Rows [j_low, j_high) are contiguous in the slab (cells are y-major).
        const uint64_t r0 = s.SinkPlan[0].cell[0].start;
        const nf_sink_plan &last = s.SinkPlan[(size_t)(s.j_width - 1) * s.Nk + (s.Nk - 1)];
        // pencil (j_high-1, cpd-1) holds cell cpd-1 of the last row at c = 0
        const uint64_t r1 = (uint64_t)last.cell[0].start + (uint64_t)last.cell[0].N;
        const auto tc = clk::now();
        if (r1 > r0) {
            NF_HIP(hipMemcpyAsync(static_cast<char *>(w.pin_acc.p), acc_dev + 4 * r0, 16 * (r1 - r0),
                                  hipMemcpyDeviceToHost, w.stream));
            NF_HIP(hipStreamSynchronize(w.stream));
            std::memcpy(s.SinkAccSlab + 4 * r0, w.pin_acc.p, 16 * (r1 - r0));
            from_dev += 16 * (r1 - r0);
        }
        copyback = std::chrono::duration<double>(clk::now() - tc).count();
    } else {
        NF_HIP(hipStreamSynchronize(w.stream));
    }
    t.stats.device_thread_s = std::chrono::duration<double>(clk::now() - t0).count();
    t.stats.fill_s = std::chrono::duration<double>(t_fill - t0).count();
    t.stats.kernel_s = kern_ms * 1e-3;
    t.stats.direct_kernel_s = direct_ms * 1e-3;
    t.stats.copyback_s = copyback;
    t.stats.bytes_to_device = to_dev;
    t.stats.bytes_from_device = from_dev;
    t.stats.direct_total = s.DirectTotal;
    t.stats.padded_direct_total = s.PaddedDirectTotal;
    t.stats.assigned_buffer = w.id;
    t.stats.path = path;
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

void worker_main(Runtime *rt, Worker *w) {
    for (;;) {
        Task *t = nullptr;
        {
            std::unique_lock<std::mutex> lk(rt->qm);
            rt->qcv.wait(lk, [&] { return rt->stopping || !rt->queue.empty(); });
            if (rt->queue.empty()) return;  // stopping and drained
            t = rt->queue.front();
            rt->queue.pop_front();
        }
        int rc = run_task(*rt, *w, *t);
        finish(*t, rc);
    }
}

int init_worker(Worker &w) {
    NF_HIP(hipSetDevice(w.device));
    NF_HIP(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
    for (auto &e : w.ev) NF_HIP(hipEventCreate(&e));
    w.ready = true;
    return NF_OK;
}

void destroy_worker(Worker &w) {
    if (!w.ready) return;
    (void)hipSetDevice(w.device);
    if (w.stream) (void)hipStreamSynchronize(w.stream);
    w.plan.release();
    w.work.release();
    w.slabs.release();
    w.pin_plan.release();
    w.pin_acc.release();
    for (auto &e : w.ev)
        if (e) (void)hipEventDestroy(e);
    if (w.stream) (void)hipStreamDestroy(w.stream);
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
    if (p->sinks_per_lane != 0) {
        if (p->sinks_per_lane != 1 && p->sinks_per_lane != 2 && p->sinks_per_lane != 4)
            return fail(NF_ERR_INVALID, "sinks_per_lane must be 0, 1, 2 or 4");
        rt->shape.sinks_per_lane = p->sinks_per_lane;
    }
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
    for (int i = 0; i < p->buffers_per_device; i++) {
        auto w = std::make_unique<Worker>();
        w->id = i;
        w->device = 0;
        int rc = init_worker(*w);
        if (rc) {
            for (auto &x : rt->workers) destroy_worker(*x);
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
