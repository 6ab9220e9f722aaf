/*
 * nf_direct.h -- C ABI of the MI355X near-field direct force module.
 *
 * This is the drop-in boundary that replaces Abacus's CUDA direct module
 * (singlestep/Direct/gpudirect.a, exported API in DeviceFunctions.h:8-26 and
 * SetInteractionCollection::GPUExecute).  A SetInteractionCollection (SIC) is
 * handed over as a plain-old-data descriptor (nf_sic) whose arrays are the
 * SIC's own members; the module computes the pencil-on-pencil directs on the
 * GPU and fills the sink slab's accelerations (float4: ax, ay, az, FOF
 * density), exactly as the reference's GPU thread does after its co-add.
 *
 * All entry points return NF_OK (0) or a negative nf_status; they never abort.
 * nf_last_error() describes the most recent failure of the calling thread.
 * No torch or C++ types cross this boundary.
 */
#ifndef NF_DIRECT_H
#define NF_DIRECT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compile-time maxima, as configure.ac:155-162 fixes NFRADIUS and NFBlockSize. */
#define NF_MAX_RADIUS 2
#define NF_MAX_WIDTH (2 * NF_MAX_RADIUS + 1)
#define NF_BLOCK_SIZE 64

typedef enum nf_status {
    NF_OK = 0,
    NF_ERR_INVALID = -1,   /* bad argument or inconsistent descriptor */
    NF_ERR_HIP = -2,       /* a HIP runtime call failed */
    NF_ERR_CAPACITY = -3,  /* task exceeds the buffer capacity given to nf_setup */
    NF_ERR_STATE = -4,     /* nf_setup not called, or called twice */
    NF_ERR_NODEV = -5,     /* no usable GPU */
    NF_ERR_NOMEM = -6      /* host or device allocation failed */
} nf_status;

/* Softening forms (configure.ac:315-328; kernels DirectGPUKernels.cuh). The
 * meaning of nf_sic.eps follows NearFieldDriver's ctor (directdriver.cpp:105-117):
 * PLUMMER eps = e^2, SINGLE_SPLINE eps = 1/e^2, CUBIC_SPLINE eps = 1/e,
 * CUBIC_PLUMMER eps = e^3, with e the internal softening length. */
typedef enum nf_softening {
    NF_SOFT_PLUMMER = 0,
    NF_SOFT_SINGLE_SPLINE = 1,
    NF_SOFT_CUBIC_SPLINE = 2,
    NF_SOFT_CUBIC_PLUMMER = 3
} nf_softening;

/* Execution path.  PLAIN: one launch per SIC over 64-sink blocks with FP32
 * tile accumulation and an FP32 k-ordered co-add (the reference's structure,
 * PencilKernel.cu:97-189 + PencilPlan.cc:140-207).  BATCHED: one persistent
 * launch per group of SICs draining a work queue sorted by interaction
 * count, FP32 tile sums accumulated in FP64, FP64 co-add of the per-pencil
 * partials. */
typedef enum nf_path { NF_PATH_PLAIN = 0, NF_PATH_BATCHED = 1 } nf_path;

/* Where the slab pointers of an nf_sic live. */
typedef enum nf_mem { NF_MEM_HOST = 0, NF_MEM_DEVICE = 1 } nf_mem;

/* Layout of the position slabs a SIC names.  NF_POS_XYZ is the PosXYZ slab
 * (X[n] | Y[n] | Z[n], List3, cellparticles.cpp:218-228).  NF_POS_AOS is the
 * PosSlab itself (posstruct x,y,z per particle, the same indexing): the
 * gathers read it directly, so the TransposePos pass (pipeline.cpp:107-124)
 * is not needed.  No counterpart in the reference's GPU module. */
typedef enum nf_pos_layout { NF_POS_XYZ = 0, NF_POS_AOS = 1 } nf_pos_layout;

/* Fusion of the consumer RescaleAndCoAddAcceleration (kick.cpp:74-109) into
 * the co-add; no counterpart in the reference's GPU module.  With it the SIC's
 * sink rows of SinkAccSlab receive the final acceleration (and, if asked, the
 * near-field record StoreForces == 2 writes to disk goes to NearAccOut),
 *   acc.xyz = (near.xyz * inv_eps3 + FarAcc.xyz) * rescale,
 *   acc.w   = near.w - density_rad2
 * (FLOAT3p1 arithmetic: w takes no part in the scaling, float3p1.cpp:41-56),
 * instead of the near-field sum: the same float operations, in the same
 * order, as running the module and then RescaleAndCoAddAcceleration. */
typedef struct nf_kick {
    const float *FarAccSlab; /* acc3struct (x, y, z) per particle of slab SlabId, in the SIC's nf_mem */
    double rescale;          /* -3 (Omega_M - Omega_Smooth) / (8 pi np), negated when making glass;
                                rounded to float, as the reference's FLOAT */
    double inv_eps3;         /* 1 / SofteningLengthInternal^3 for the single spline, else 1;
                                rounded to float */
    double density_rad2;     /* WriteState.DensityKernelRad2 (double, as the reference) */
    float *NearAccOut;       /* optional (NULL: none), float4 per particle of slab SlabId, in the
                                SIC's nf_mem: the StoreForces == 2 NearAcc record of the SIC's
                                rows, near.xyz * inv_eps3 and near.w (pipeline.cpp:374-399,
                                FLOAT3p1 *= FLOAT leaves w alone) */
} nf_kick;

/* CellPencilPlan (SetInteractionCollection.hh:76-87): start index of a cell
 * in its PosXYZ slab and its particle count.  Same layout as the reference. */
typedef struct nf_cell_plan {
    uint32_t start;
    int32_t N;
} nf_cell_plan;

/* SinkPencilPlan (SetInteractionCollection.hh:94-115), non-PARALLEL layout. */
typedef struct nf_sink_plan {
    nf_cell_plan cell[NF_MAX_WIDTH];
    int32_t wrapcell;
} nf_sink_plan;

/* SourcePencilPlan (SetInteractionCollection.hh:121-132). */
typedef struct nf_source_plan {
    nf_cell_plan cell[NF_MAX_WIDTH];
} nf_source_plan;

/* cellinfo (particlestruct.cpp:334-392), same layout as the reference. */
typedef struct nf_cellinfo {
    uint64_t startindex;
    uint64_t startindex_with_ghost;
    int32_t count;
    int32_t active;
    float mean_square_velocity;
    float max_component_velocity;
    float max_component_acceleration;
} nf_cellinfo;

/* The SIC descriptor: the members of SetInteractionCollection
 * (SetInteractionCollection.hh:171-283) that the GPU module reads. */
typedef struct nf_sic {
    int32_t SlabId, j_low, j_high, k_low, k_high, Nk, j_width;
    int32_t cpd, nfradius, nfwidth;
    float eps;           /* softening parameter, see nf_softening */
    float b2;            /* FOF density radius squared (0 disables w) */
    int32_t softening;   /* nf_softening */
    int32_t mem;         /* nf_mem of the three slab pointer fields below */

    const float *SinkPosSlab;                  /* PosXYZ of slab SlabId: X|Y|Z */
    const float *SourcePosSlab[NF_MAX_WIDTH];  /* PosXYZ of slabs SlabId-R..SlabId+R */
    uint64_t Nslab[NF_MAX_WIDTH];              /* particles per slab (with ghosts) */
    float *SinkAccSlab;                        /* float4 per particle of slab SlabId */

    int32_t NSinkSets;
    const int32_t *SinkSetStart;
    const int32_t *SinkSetCount;
    const uint32_t *SinkSetIdMax;
    const nf_sink_plan *SinkPlan;

    int32_t NSourceSets;
    const int32_t *SourceSetStart;
    const int32_t *SourceSetCount;
    const nf_source_plan *SourcePlan;

    int32_t NSinkBlocks, NSourceBlocks, InteractionCount;
    const int32_t *SinkBlockParentPencil;
    const int32_t *SinkSourceInteractionList;
    const float *SinkSourceYOffset;

    uint64_t DirectTotal, PaddedDirectTotal;
    uint64_t SinkTotal, SourceTotal, PaddedSinkTotal, PaddedSourceTotal;

    /* Extensions (zero for the reference's behaviour). */
    int32_t pos_layout;      /* nf_pos_layout of SinkPosSlab and SourcePosSlab */
    int32_t reserved;
    const nf_kick *kick;     /* non-NULL: write the final acceleration (see nf_kick) */
} nf_sic;

/* Per-task timers and counters, the analogue of the SIC's STimers and byte
 * counts that NearFieldDriver::Finalize aggregates (directdriver.cpp:453-474). */
typedef struct nf_task_stats {
    double device_thread_s;   /* wall time of the task on its worker */
    double fill_s;            /* plan upload + slab upload (host mode) */
    double kernel_s;          /* gather + direct kernel + co-add, GPU time */
    double direct_kernel_s;   /* the direct kernel alone, GPU time */
    double copyback_s;        /* acceleration copy back (host mode) */
    uint64_t bytes_to_device, bytes_from_device;
    uint64_t direct_total, padded_direct_total;
    int32_t assigned_buffer;
    int32_t path;
    int32_t group_size;            /* SICs that shared this task's persistent launch */
    double group_direct_kernel_s;  /* GPU time of that shared direct-kernel launch */
    uint64_t group_direct_total;   /* interactions of the whole group */
    int32_t graph;                 /* 1 if the group ran as a captured HIP graph */
    double upload_s;               /* GPU time of this task's share of slab uploads (host mode) */
    double download_s;             /* GPU time of its acceleration copy-back (host mode) */
    int32_t slabs_uploaded;        /* slabs this task's group uploaded (cache misses) */
    int32_t slabs_reused;          /* slabs it found resident (cache hits) */
    int32_t sinks_per_lane;        /* batched path: register tile of its launch (0: plain) */
    int32_t autotune_trial;        /* 1: its launch group timed every candidate tile */
    int32_t waves_per_eu;          /* batched path: occupancy target its kernel was built with */
    int32_t device_planned;        /* 1: submitted by nf_submit_slab (planned on the device) */
    double plan_s;                 /* nf_submit_slab: worker wall time from the start of its
                                      group to its planning kernels being enqueued */
    int32_t waves_per_wg;          /* batched path: workgroup size of its launch, in waves */
    uint64_t group_form_interactions[3]; /* batched path with nf_setup's census: pair
                                            interactions of its launch group computed in the
                                            unsoftened form without and with the FOF density,
                                            and in the full softened form */
    uint64_t group_tail_interactions;    /* of those, the ones in short pencil tails */
} nf_task_stats;

typedef struct nf_task nf_task;

/* Setup parameters, the arguments of GPUSetup (GPUControl.cu:357-363). */
typedef struct nf_setup_params {
    int32_t cpd;
    uint64_t max_buffer_bytes;    /* device bytes per buffer (MaxBufferSize) */
    int32_t n_gpu;                /* number of GPUs to use (1 on this build) */
    int32_t buffers_per_device;   /* DirectBPD: worker threads/streams per GPU */
    int32_t use_pinned;           /* UsePinnedGPUMemory: -1 auto, 0 no, 1 yes */
    int32_t path;                 /* nf_path used by nf_submit */
    int32_t sinks_per_lane;       /* batched path register tile: 1, 2 or 4 (0: auto) */
    int32_t source_tile;          /* batched path TwoSum tile depth (0: auto) */
    int32_t gpu_streams;          /* HIP streams the workers enqueue on: 0 one per
                                     worker; 1 one shared stream (tasks run back to
                                     back on the GPU while the workers still prepare
                                     them in parallel) and one shared planning
                                     stream: measured fastest */
    int32_t max_group;            /* batched path: SICs queued together may share one
                                     persistent launch, up to this many (0: 8) */
    int32_t use_graphs;           /* 1: each group's launch sequence (gathers, direct
                                     kernel, co-adds) is captured in a
                                     HIP graph and launched as one; a worker keeps its
                                     executable graph and updates it in place when the
                                     next group has the same shape */
    uint64_t group_interactions;  /* batched path: stop adding SICs to a launch group
                                     once it holds this many pair interactions
                                     (0: 1.2e10, about 4 ms of kernel time) */
    int32_t waves_per_wg;         /* batched path: workgroup size of the persistent
                                     launch, 1, 2 or 4 waves (0: auto) */
    int32_t census;               /* batched path: 1 counts the pair interactions of each
                                     pair form (nf_task_stats.group_form_interactions), with
                                     a diagnostic build of the kernel some percent slower;
                                     0: not counted (the fields read 0) */
} nf_setup_params;

/* GetNGPU (GPUControl.cu:267-271). */
int nf_get_ngpu(int32_t *n);

/* GetDeviceInfo (GPUControl.cu:274-285): min over GPUs of 0.8 * total
 * memory, and the device name. */
int nf_get_device_info(uint64_t *bytes, char *name, size_t name_len);

/* GPUSetup (GPUControl.cu:357-482): start buffers_per_device workers per GPU,
 * each owning one HIP stream and one device buffer, and report the largest
 * sink/source block counts a SIC may have (MaxSinkBlocks/MaxSourceBlocks). */
int nf_setup(const nf_setup_params *p, int32_t *max_sink_blocks, int32_t *max_source_blocks);

/* GPUReset (GPUControl.cu:492-510): stop the workers and free all buffers. */
int nf_reset(void);

/* SetInteractionCollection::GPUExecute (GPUControl.cu:235-261): queue the SIC
 * for the workers and return a task handle; with blocking != 0, also wait for
 * it.  On the batched path a worker takes every compatible SIC waiting in the
 * queue (up to max_group) into one persistent launch.  The slabs and the plan
 * arrays must stay alive until the task completes. */
int nf_submit(const nf_sic *sic, int32_t blocking, nf_task **task);

/* ExecuteSlabGPU (directdriver.cpp:242-387) for one split of a slab, with the
 * SetInteractionCollection constructor (PencilOnPencil.cc:180-418) run on the
 * device instead of the host: the task plans the SIC of slab SlabId, sink
 * rows [j_low, j_high), from the cellinfo slabs, then executes it as
 * nf_submit does.  The plan arrays exist on the device only; the task's
 * stats report the SIC's totals (direct_total ...).  The cellinfo slabs go up
 * through the host-slab cache like position slabs (inside a stream session
 * each is uploaded once).  A cell whose start + count falls outside its slab
 * fails the task with NF_ERR_INVALID.  Such a task is launched on its own,
 * never grouped with other SICs.  Every pointer must stay valid until the
 * task completes. */
typedef struct nf_slab_task {
    int32_t SlabId, j_low, j_high, cpd, nfradius;
    float eps, b2;                              /* as nf_sic */
    int32_t softening;                          /* nf_softening */
    int32_t mem;                                /* nf_mem of PosSlab, SinkAccSlab, kick->FarAccSlab */
    int32_t pos_layout;                         /* nf_pos_layout of PosSlab */
    const nf_cellinfo *cellinfo[NF_MAX_WIDTH];  /* host memory: slabs SlabId-R..SlabId+R, cpd*cpd cells, y-major */
    const float *PosSlab[NF_MAX_WIDTH];         /* position slabs SlabId-R..SlabId+R */
    uint64_t Nslab[NF_MAX_WIDTH];               /* particles per slab (with ghosts) */
    float *SinkAccSlab;                         /* float4 per particle of slab SlabId */
    const nf_kick *kick;                        /* optional, as nf_sic */
} nf_slab_task;
int nf_submit_slab(const nf_slab_task *t, int32_t blocking, nf_task **task);

/* SetInteractionCollection::CheckCompletion (PencilOnPencil.cc:434-436):
 * 1 when done, 0 while running, < 0 if the task failed. */
int nf_check_completion(nf_task *task);

/* Block until the task is done; returns its final status. */
int nf_wait(nf_task *task);

/* Copy out the task's timers and counters (valid once complete). */
int nf_get_task_stats(nf_task *task, nf_task_stats *out);

/* Release a completed task handle. */
int nf_release(nf_task *task);

/* Thread-local description of the last error. */
const char *nf_last_error(void);

/* The GPU throughput timer and totals (GPUControl.cu:658-672, read by
 * NearFieldDriver::AggregateStats, directdriver.cpp:495-520): wall seconds
 * during which at least one launch group was on the GPU, and totals of the
 * tasks completed since nf_setup.  direct_total / gpu_active_s is the
 * reference's "effective GDIPS" (reporting.cpp:233-236). */
typedef struct nf_throughput {
    double gpu_active_s;
    uint64_t direct_total, padded_direct_total;
    uint64_t bytes_to_device, bytes_from_device;
    uint64_t tasks, groups;
} nf_throughput;
int nf_get_throughput(nf_throughput *out);

/* The batched-path autotuner's table (no counterpart in the reference).  The
 * tile shape has two axes, sinks per lane (1, 2, 4: sinks per wave 64, 128,
 * 256) and waves per workgroup (1, 2, 4); nf_setup's sinks_per_lane and
 * waves_per_wg fix an axis when nonzero, and the autotuner searches the
 * others.  It keys every launch group on a coarse histogram of its
 * sink-pencil occupancies (the share of its sinks in pencils of <64, 64-255,
 * 256-1023 and >=1024 sinks, in quarters) and on its size class (<2^31,
 * <2^33 or more pair interactions); the first group with a new key times each
 * candidate shape on its own launch and the fastest is kept for every later
 * group with that key.  Every shape gives bitwise the same accelerations.
 * Writes the table as JSON into buf; returns the length it needs (excluding
 * the terminator), or < 0. */
int nf_autotune_table(char *buf, size_t len);

/* ---- host-slab streaming ---------------------------------------------
 * No counterpart in the reference.  Between nf_stream_begin and
 * nf_stream_end the caller promises that the host PosXYZ slabs named by
 * submitted SICs, and the cellinfo slabs named by nf_submit_slab tasks, do
 * not change (the reference's pipeline keeps a slab fixed while SICs that
 * read it are pending; one session per step is the usual bracket).  Inside
 * such a session the module keeps recently used host slabs in HBM keyed by
 * address, so consecutive slab SICs upload each slab once; uploads and acceleration
 * copy-backs run on a copy stream through double-buffered pinned staging
 * and overlap other groups' kernels.  Outside a session every group uploads
 * its slabs afresh.  nf_stream_end waits for no task; slabs still in use are
 * dropped when their tasks complete. */
int nf_stream_begin(void);
int nf_stream_end(void);

/* ---- device memory for HBM-resident slabs ------------------------------
 * No counterpart in the reference, whose slabs live in host memory.  These
 * are for callers that keep PosXYZ and acceleration slabs resident on the GPU
 * and submit SICs with mem = NF_MEM_DEVICE; they use the module's own HIP
 * runtime so a caller needs no second GPU library. */
int nf_device_alloc(uint64_t bytes, void **ptr);
int nf_device_free(void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device (synchronous). */
int nf_device_copy(void *dst, const void *src, uint64_t bytes, int32_t kind);
int nf_device_synchronize(void);
/* Pinned (page-locked) host memory.  Host-mode slabs and acceleration slabs
 * allocated here are read and written by the DMA engines directly; pageable
 * ones go through the module's pinned staging buffers. */
int nf_host_alloc(uint64_t bytes, void **ptr);
int nf_host_free(void *ptr);

/* ---- planning (host) ------------------------------------------------- */

/* ComputeSICSize (PencilOnPencil.cc:130-150): bytes of arena a SIC planned by
 * nf_plan_sic needs for sink rows [j_low, j_high), given the particle counts
 * of the slabs involved (an upper bound, like the reference's). */
uint64_t nf_sic_arena_bytes(int32_t cpd, int32_t nfradius, int32_t j_low, int32_t j_high,
                            uint64_t max_slab_particles);

/* SetInteractionCollection ctor (PencilOnPencil.cc:180-418) for the
 * non-PARALLEL build: plan the SIC of slab `slab`, sink rows [j_low, j_high).
 * cellinfo[c] is the cellinfo slab (cpd*cpd cells, y-major) of slab
 * slab-R+c, c = 0..2R; pos_slab[c] its PosXYZ slab and nslab[c] its size.
 * Arrays are carved from `arena` (arena_bytes long); *out is filled.  The
 * caller sets out->SinkAccSlab and out->mem afterwards if needed. */
int nf_plan_sic(int32_t cpd, int32_t nfradius, int32_t slab, int32_t j_low, int32_t j_high,
                const nf_cellinfo *const *cellinfo, const float *const *pos_slab,
                const uint64_t *nslab, float eps, float b2, int32_t softening, void *arena,
                uint64_t arena_bytes, nf_sic *out);

/* NearFieldDriver::ExecuteSlabGPU's split of a slab into SICs over y
 * (directdriver.cpp:248-303), with the skewer estimates ComputeSinkBlocks /
 * ComputeSourceBlocks (PencilOnPencil.cc:113-126): a new SIC starts where the
 * running worst-case sink or source blocks pass useMaxSink (the slab's total
 * over nbuffers, scaled by 1 + 1.1 nbuffers/cpd, at most max_sink_blocks) or
 * max_source_blocks.  cellinfo[c] and nslab[c] as for nf_plan_sic; nbuffers
 * is NGPU * DirectBPD; max_*_blocks are nf_setup's.  Writes each SIC's end
 * row into split (room for cpd entries) and their number into *nsplit;
 * NF_ERR_CAPACITY where the reference asserts (the last skewer overflows). */
int nf_split_slab(int32_t cpd, int32_t nfradius, const nf_cellinfo *const *cellinfo, const uint64_t *nslab,
                  int32_t nbuffers, int32_t max_sink_blocks, int32_t max_source_blocks, int32_t *split,
                  int32_t *nsplit);

/* The pair interactions of a SIC (its DirectTotal: self pairs included, no
 * padding) from the cellinfo counts alone, as CalcDI counts a slab's
 * (DirectBench.cpp:111-125), for the sink rows [j_low, j_high): each sink
 * cell's count times the particles in its (2 nfradius + 1)^3 cube of cells,
 * with the periodic wrap.  The cube sums are separable (over the slabs, then
 * z, then y), so a whole CPD 255 slab takes about 3 ms on one core.
 * nf_submit_slab uses it to size launch groups before any plan exists.
 * cellinfo[c] as for nf_plan_sic; NF_ERR_INVALID for a negative count. */
int nf_count_interactions(int32_t cpd, int32_t nfradius, int32_t j_low, int32_t j_high,
                          const nf_cellinfo *const *cellinfo, uint64_t *out);

/* nf_plan_sic computed by the device planner nf_submit_slab uses (same
 * arguments, same output; cellinfo in host memory; nslab must be given, as
 * the device planner checks every cell against its slab).  Synchronous, on
 * the current HIP device, with temporary device buffers.  No counterpart in
 * the reference: it exposes the device planner for callers that keep their
 * own plan arrays, and for checks against nf_plan_sic. */
int nf_plan_sic_device(int32_t cpd, int32_t nfradius, int32_t slab, int32_t j_low, int32_t j_high,
                       const nf_cellinfo *const *cellinfo, const float *const *pos_slab,
                       const uint64_t *nslab, float eps, float b2, int32_t softening, void *arena,
                       uint64_t arena_bytes, nf_sic *out);

/* ---- diagnostics --------------------------------------------------------
 * No counterpart in the reference.  The pair-math probe: the batched
 * kernel's register-tile loop alone (single spline; mode 0 the unsoftened
 * form without the FOF density, 1 with it, 2 the full softened form) with
 * sinks_per_lane 1, 2 or 4 and an occupancy of waves_per_simd 4, 6 or 8,
 * over sources that stay in the caches: the ceiling of the pair math at that
 * tile, in pair interactions per second over the whole GPU.  Synchronous, on
 * the current device; allocates its own small buffers. */
int nf_probe_pair_rate(float eps, float b2, int32_t sinks_per_lane, int32_t mode, int32_t waves_per_simd,
                       double *interactions_per_s);

/* ---- execution without allocation -------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): nf_submit(sic, blocking = 1) on the batched path, for a
 * SIC whose slabs are device-resident (mem = NF_MEM_DEVICE), with every
 * scratch buffer (plan image, work list, queue counter, source stream,
 * partials) in one fixed device array of NF_NOALLOC_SCRATCH_BYTES per GPU,
 * held from the first call on that GPU until the process ends.  It needs no
 * nf_setup, allocates no memory (device or host) when it succeeds, runs on
 * `stream` (a hipStream_t; NULL: the null stream) of the current device and
 * returns when the SIC's accelerations are in sic->SinkAccSlab: bitwise those
 * nf_submit gives on the batched path.  A SIC that needs more scratch fails with
 * NF_ERR_CAPACITY (split the slab's rows into smaller SICs, as
 * nf_split_slab does for nf_setup's buffers).  Calls share their GPU's array
 * and are serialized by a lock; the array keeps the last call's contents. */
#define NF_NOALLOC_SCRATCH_BYTES (256ull << 20)
int nf_execute_sic_noalloc(const nf_sic *sic, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* NF_DIRECT_H */
