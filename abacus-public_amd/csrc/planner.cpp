// Host-side SIC planner: builds the pencil plans and interaction lists of
// one SetInteractionCollection from the slabs' cellinfo.
//
// Restates, for the non-PARALLEL build (MPI_size_z == 1, node_z_start = 0,
// node_z_size = cpd, so Nk = cpd and no ghost cells):
//   SetInteractionCollection ctor          PencilOnPencil.cc:180-418
//   index_to_zcen                           PencilOnPencil.cc:37-43
//   NumPaddedBlocks                         PencilOnPencil.cc:20-22
//   ComputeSICSize                          PencilOnPencil.cc:130-150
//   SinkPencilPlan::load                    PencilPlan.cc:95-138
//   SourcePencilPlan::load                  PencilPlan.cc:260-279
//   ExecuteSlabGPU's y-split                directdriver.cpp:248-303
//   NumParticlesInSkewer, ComputeSinkBlocks,
//   ComputeSourceBlocks                     PencilOnPencil.cc:90-126
// The output arrays have exactly the reference's contents; they are carved
// page-aligned from a caller-owned arena as posix_memalign_wrap does
// (PencilOnPencil.cc:156-164).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "nf_direct.h"
#include "nf_internal.h"

namespace {

constexpr uint64_t kPage = 4096;

inline int wrap_index(int i, int cpd) {
    // grid::WrapSlab (grid.cpp:70-74)
    while (i >= cpd) i -= cpd;
    while (i < 0) i += cpd;
    return i;
}

inline int padded_blocks(int n) { return (n + NF_BLOCK_SIZE - 1) / NF_BLOCK_SIZE; }

struct Arena {
    char *p;
    uint64_t left;
    bool take(void **out, uint64_t bytes) {
        uint64_t mis = reinterpret_cast<uintptr_t>(p) % kPage;
        uint64_t extra = mis ? kPage - mis : 0;
        if (left < bytes + extra) return false;
        p += extra;
        left -= extra;
        *out = p;
        p += bytes;
        left -= bytes;
        return true;
    }
};

}  // namespace

extern "C" uint64_t nf_sic_arena_bytes(int32_t cpd, int32_t nfradius, int32_t j_low, int32_t j_high,
                                       uint64_t max_slab_particles) {
    const uint64_t W = 2 * (uint64_t)nfradius + 1;
    const uint64_t jw = (uint64_t)(j_high - j_low);
    const uint64_t nsink = jw * (uint64_t)cpd;
    const uint64_t nsrc = (jw + W - 1) * (uint64_t)cpd;
    // Each particle appears in W sink pencils; each pencil adds at most one
    // partially filled block (the reference's worst case, PencilOnPencil.cc:52-57).
    const uint64_t nblocks = W * max_slab_particles / NF_BLOCK_SIZE + nsink;
    uint64_t size = 0;
    size += (3 * sizeof(int32_t) + sizeof(nf_sink_plan)) * nsink;
    size += (2 * sizeof(int32_t) + sizeof(nf_source_plan)) * nsrc;
    size += (sizeof(int32_t) + sizeof(float)) * W * nsink;
    size += sizeof(int32_t) * nblocks;
    size += 16 * kPage;  // alignment of the 11 sub-arrays
    return size;
}

extern "C" int nf_plan_sic(int32_t cpd, int32_t R, int32_t slab, int32_t j_low, int32_t j_high,
                           const nf_cellinfo *const *cellinfo, const float *const *pos_slab,
                           const uint64_t *nslab, float eps, float b2, int32_t softening,
                           void *arena_ptr, uint64_t arena_bytes, nf_sic *out) {
    if (!out || !cellinfo || !arena_ptr) return nf::fail(NF_ERR_INVALID, "nf_plan_sic: null argument");
    if (cpd <= 0 || (cpd % 2) == 0) return nf::fail(NF_ERR_INVALID, "nf_plan_sic: cpd must be odd and positive");
    if (R < 0 || R > NF_MAX_RADIUS || 2 * R + 1 > cpd)
        return nf::fail(NF_ERR_INVALID, "nf_plan_sic: nfradius out of range");
    if (j_low < 0 || j_high > cpd || j_low >= j_high)
        return nf::fail(NF_ERR_INVALID, "nf_plan_sic: bad sink row range");
    const int W = 2 * R + 1;
    for (int c = 0; c < W; c++)
        if (!cellinfo[c]) return nf::fail(NF_ERR_INVALID, "nf_plan_sic: missing cellinfo slab");

    std::memset(out, 0, sizeof(*out));
    Arena ar{static_cast<char *>(arena_ptr), arena_bytes};
    out->SlabId = slab;
    out->j_low = j_low;
    out->j_high = j_high;
    out->k_low = 0;
    out->k_high = cpd;
    out->Nk = cpd;  // (MPI_size_z > 1) ? ... : k_high - k_low  (PencilOnPencil.cc:204)
    out->j_width = j_high - j_low;
    out->cpd = cpd;
    out->nfradius = R;
    out->nfwidth = W;
    out->eps = eps;
    out->b2 = b2;
    out->softening = softening;
    out->mem = NF_MEM_HOST;
    if (pos_slab) {
        out->SinkPosSlab = pos_slab[R];
        for (int c = 0; c < W; c++) out->SourcePosSlab[c] = pos_slab[c];
    }
    if (nslab)
        for (int c = 0; c < W; c++) out->Nslab[c] = nslab[c];

    const int Nk = out->Nk, jw = out->j_width;
    const int NSinkSets = jw * Nk;
    const int NSourceSets = Nk * (jw + W - 1);
    out->NSinkSets = NSinkSets;
    out->NSourceSets = NSourceSets;

    int32_t *SinkSetStart, *SinkSetCount, *SourceSetStart, *SourceSetCount;
    uint32_t *SinkSetIdMax;
    nf_sink_plan *SinkPlan;
    nf_source_plan *SourcePlan;
    bool ok = ar.take((void **)&SinkSetStart, sizeof(int32_t) * NSinkSets) &&
              ar.take((void **)&SinkSetCount, sizeof(int32_t) * NSinkSets) &&
              ar.take((void **)&SinkPlan, sizeof(nf_sink_plan) * NSinkSets) &&
              ar.take((void **)&SinkSetIdMax, sizeof(uint32_t) * NSinkSets) &&
              ar.take((void **)&SourceSetStart, sizeof(int32_t) * NSourceSets) &&
              ar.take((void **)&SourceSetCount, sizeof(int32_t) * NSourceSets) &&
              ar.take((void **)&SourcePlan, sizeof(nf_source_plan) * NSourceSets);
    if (!ok) return nf::fail(NF_ERR_CAPACITY, "nf_plan_sic: arena too small");

    // The central cell of sink pencil k: index_to_zcen (PencilOnPencil.cc:37-43).
    std::vector<int> zcen(Nk);
    for (int k = 0; k < Nk; k++) {
        int kk = k + R;
        if (kk >= cpd) kk -= cpd;
        zcen[k] = kk;
    }

    // ---- Sink pencils: W cells along z of slab `slab`, row y = j + j_low.
    const nf_cellinfo *ci_x = cellinfo[R];
    std::vector<uint64_t> skewer_blocks(jw + W, 0);
    uint64_t SinkTotal = 0;
    for (int j = 0; j < jw; j++) {
        const int y = j + j_low;
        for (int k = 0; k < Nk; k++) {
            nf_sink_plan &pl = SinkPlan[j * Nk + k];
            pl.wrapcell = 0;
            int total = 0;
            for (int c = 0; c < W; c++) {
                const int zc = zcen[k] + c - R;
                const nf_cellinfo &ci = ci_x[(uint64_t)y * cpd + wrap_index(zc, cpd)];
                pl.cell[c].start = (uint32_t)ci.startindex;
                pl.cell[c].N = ci.count;
                // First cell that sits at z = 0 after the periodic wrap
                // (PencilPlan.cc:123-128 with node_z_start = 0, node_z_size = cpd).
                if (wrap_index(zc - cpd, cpd) == 0 || wrap_index(zc, cpd) == 0) pl.wrapcell = c;
                total += ci.count;
            }
            for (int c = W; c < NF_MAX_WIDTH; c++) pl.cell[c] = nf_cell_plan{0, 0};
            SinkSetCount[j * Nk + k] = total;
            SinkTotal += (uint64_t)total;
            skewer_blocks[j] += (uint64_t)padded_blocks(total);
        }
    }
    std::vector<uint64_t> skewer_start(jw + W + 1, 0);
    for (int j = 0; j < jw; j++) skewer_start[j + 1] = skewer_start[j] + skewer_blocks[j];
    const uint64_t NSinkBlocks = skewer_start[jw];
    if (NSinkBlocks * NF_BLOCK_SIZE >= (uint64_t)INT32_MAX)
        return nf::fail(NF_ERR_CAPACITY, "nf_plan_sic: padded sink count overflows int32");
    out->NSinkBlocks = (int32_t)NSinkBlocks;

    int32_t *Parent;
    if (!ar.take((void **)&Parent, sizeof(int32_t) * (NSinkBlocks ? NSinkBlocks : 1)))
        return nf::fail(NF_ERR_CAPACITY, "nf_plan_sic: arena too small");
    for (int j = 0; j < jw; j++) {
        uint64_t b = skewer_start[j];
        for (int k = 0; k < Nk; k++) {
            const int s = j * Nk + k;
            const int nb = padded_blocks(SinkSetCount[s]);
            SinkSetStart[s] = (int32_t)(b * NF_BLOCK_SIZE);
            SinkSetIdMax[s] = (uint32_t)(SinkSetStart[s] + SinkSetCount[s]);
            for (int q = 0; q < nb; q++) Parent[b + q] = s;
            b += (uint64_t)nb;
        }
    }

    // ---- Source pencils: W cells along x at row y = j + j_low - R, z = zcen.
    std::fill(skewer_blocks.begin(), skewer_blocks.end(), 0);
    uint64_t SourceTotal = 0;
    for (int j = 0; j < jw + W - 1; j++) {
        const int y = wrap_index(j + j_low - R, cpd);
        for (int k = 0; k < Nk; k++) {
            nf_source_plan &pl = SourcePlan[j * Nk + k];
            int total = 0;
            for (int c = 0; c < W; c++) {
                const nf_cellinfo &ci = cellinfo[c][(uint64_t)y * cpd + zcen[k]];
                pl.cell[c].start = (uint32_t)ci.startindex_with_ghost;
                pl.cell[c].N = ci.count;
                total += ci.count;
            }
            for (int c = W; c < NF_MAX_WIDTH; c++) pl.cell[c] = nf_cell_plan{0, 0};
            SourceSetCount[j * Nk + k] = total;
            SourceTotal += (uint64_t)total;
            skewer_blocks[j] += (uint64_t)padded_blocks(total);
        }
    }
    skewer_start[0] = 0;
    for (int j = 0; j < jw + W - 1; j++) skewer_start[j + 1] = skewer_start[j] + skewer_blocks[j];
    const uint64_t NSourceBlocks = skewer_start[jw + W - 1];
    if (NSourceBlocks * NF_BLOCK_SIZE >= (uint64_t)INT32_MAX)
        return nf::fail(NF_ERR_CAPACITY, "nf_plan_sic: padded source count overflows int32");
    out->NSourceBlocks = (int32_t)NSourceBlocks;
    for (int j = 0; j < jw + W - 1; j++) {
        uint64_t b = skewer_start[j];
        for (int k = 0; k < Nk; k++) {
            const int s = j * Nk + k;
            SourceSetStart[s] = (int32_t)(b * NF_BLOCK_SIZE);
            b += (uint64_t)padded_blocks(SourceSetCount[s]);
        }
    }

    // ---- Interactions: each sink pencil meets W source pencils (j+y, k).
    const int IC = W * NSinkSets;
    out->InteractionCount = IC;
    int32_t *IL;
    float *YO;
    if (!ar.take((void **)&IL, sizeof(int32_t) * IC) || !ar.take((void **)&YO, sizeof(float) * IC))
        return nf::fail(NF_ERR_CAPACITY, "nf_plan_sic: arena too small");
    const float cellsize = (float)(1.0 / (double)cpd);  // FLOAT cellsize = CP->invcpd
    uint64_t DirectTotal = 0, PaddedDirectTotal = 0;
    for (int j = 0; j < jw; j++)
        for (int k = 0; k < Nk; k++) {
            const int s = j * Nk + k;
            const uint64_t ns = (uint64_t)SinkSetCount[s];
            const uint64_t nps = (uint64_t)padded_blocks(SinkSetCount[s]) * NF_BLOCK_SIZE;
            for (int y = 0; y < W; y++) {
                const int src = (j + y) * Nk + k;
                IL[W * s + y] = src;
                YO[W * s + y] = (float)(y - R) * cellsize;
                DirectTotal += ns * (uint64_t)SourceSetCount[src];
                PaddedDirectTotal += nps * (uint64_t)SourceSetCount[src];
            }
        }

    out->SinkSetStart = SinkSetStart;
    out->SinkSetCount = SinkSetCount;
    out->SinkSetIdMax = SinkSetIdMax;
    out->SinkPlan = SinkPlan;
    out->SourceSetStart = SourceSetStart;
    out->SourceSetCount = SourceSetCount;
    out->SourcePlan = SourcePlan;
    out->SinkBlockParentPencil = Parent;
    out->SinkSourceInteractionList = IL;
    out->SinkSourceYOffset = YO;
    out->DirectTotal = DirectTotal;
    out->PaddedDirectTotal = PaddedDirectTotal;
    out->SinkTotal = SinkTotal;
    out->SourceTotal = SourceTotal;
    out->PaddedSinkTotal = NSinkBlocks * NF_BLOCK_SIZE;
    out->PaddedSourceTotal = NSourceBlocks * NF_BLOCK_SIZE;
    return NF_OK;
}

// NearFieldDriver::ExecuteSlabGPU's y-split (directdriver.cpp:248-303) with
// the skewer estimates NumParticlesInSkewer / ComputeSinkBlocks /
// ComputeSourceBlocks (PencilOnPencil.cc:90-126).  Integer arithmetic as the
// reference's: int particle counts, int block counts, useMaxSink scaled by a
// double factor and truncated.
extern "C" int nf_split_slab(int32_t cpd, int32_t R, const nf_cellinfo *const *cellinfo, const uint64_t *nslab,
                             int32_t nbuffers, int32_t max_sink_blocks, int32_t max_source_blocks, int32_t *split,
                             int32_t *nsplit) {
    if (!cellinfo || !nslab || !split || !nsplit) return nf::fail(NF_ERR_INVALID, "nf_split_slab: null argument");
    if (cpd <= 0 || (cpd % 2) == 0) return nf::fail(NF_ERR_INVALID, "nf_split_slab: cpd must be odd and positive");
    if (R < 0 || R > NF_MAX_RADIUS || 2 * R + 1 > cpd)
        return nf::fail(NF_ERR_INVALID, "nf_split_slab: nfradius out of range");
    if (nbuffers <= 0) return nf::fail(NF_ERR_INVALID, "nf_split_slab: nbuffers must be positive");
    const int W = 2 * R + 1;
    for (int c = 0; c < W; c++)
        if (!cellinfo[c]) return nf::fail(NF_ERR_INVALID, "nf_split_slab: missing cellinfo slab");
    // NumParticlesInSkewer: next row's start (or the slab size) minus this row's.
    auto skewer = [&](int c, int j) -> uint64_t {
        const uint64_t start = cellinfo[c][(uint64_t)j * cpd].startindex_with_ghost;
        const uint64_t end = j < cpd - 1 ? cellinfo[c][(uint64_t)(j + 1) * cpd].startindex_with_ghost : nslab[c];
        return end - start;
    };
    std::vector<int> sink_blocks(cpd), source_blocks(cpd);
    int64_t tot_sink = 0;
    for (int j = 0; j < cpd; j++) {
        const uint64_t s0 = cellinfo[R][(uint64_t)j * cpd].startindex;
        const uint64_t s1 = j < cpd - 1 ? cellinfo[R][(uint64_t)(j + 1) * cpd].startindex : nslab[R];
        sink_blocks[j] = W * ((int)(s1 - s0) / NF_BLOCK_SIZE) + cpd;  // ComputeSinkBlocks (no ghosts)
        int npsrc = 0;
        for (int c = 0; c < W; c++) npsrc += (int)skewer(c, j);
        source_blocks[j] = npsrc / NF_BLOCK_SIZE + cpd;  // ComputeSourceBlocks
        tot_sink += sink_blocks[j];
    }
    int use_max_sink = (int)(tot_sink / nbuffers);
    use_max_sink = (int)(use_max_sink * (1 + 1.1 * nbuffers / cpd));
    use_max_sink = std::min(use_max_sink, max_sink_blocks);
    const int use_max_source = max_source_blocks;
    auto wrap = [&](int j) { return ((j % cpd) + cpd) % cpd; };
    int n = 0, this_sink = 0, this_source = 0;
    for (int j = 0; j < cpd; j++) {
        this_sink += sink_blocks[j];
        this_source += source_blocks[wrap(j + R)];
        if (j == 0 || this_sink > use_max_sink || this_source > use_max_source) {
            if (n > 0) split[n - 1] = j;
            this_sink = sink_blocks[j];
            this_source = 0;
            for (int c = -R; c <= R; c++) this_source += source_blocks[wrap(j - c)];
            n++;
            if (n >= 1024) return nf::fail(NF_ERR_CAPACITY, "nf_split_slab: too many splits");
        }
    }
    split[n - 1] = cpd;
    *nsplit = n;
    if (this_sink > max_sink_blocks || this_source > max_source_blocks)
        return nf::fail(NF_ERR_CAPACITY, "nf_split_slab: sinks or sources of the last skewer overflow the maxima");
    return NF_OK;
}

namespace nf {

// The separable cube sums behind nf_count_interactions; negative counts are
// an error (strict) or count as zero.  max_pencil (or null) receives the
// particles of the SIC's largest source pencil: the first sums, one per
// (y, z) over the W slabs, are exactly the source pencils' counts.
int count_interactions(int32_t cpd, int32_t R, int32_t j_low, int32_t j_high, const nf_cellinfo *const *ci,
                       bool strict, uint64_t *out, int64_t *max_pencil) {
    const int W = 2 * R + 1, rows = (j_high - j_low) + 2 * R;
    auto wrap = [&](int i) { return ((i % cpd) + cpd) % cpd; };
    // m: particles of the W slabs in each cell of the source rows j_low-R ..
    // j_high+R-1; then mz: m summed over z-R .. z+R.
    std::vector<int64_t> m((size_t)rows * cpd), mz((size_t)rows * cpd);
    for (int r = 0; r < rows; r++) {
        const size_t row = (size_t)wrap(j_low - R + r) * cpd;
        for (int z = 0; z < cpd; z++) {
            int64_t n = 0;
            for (int c = 0; c < W; c++) {
                const int32_t k = ci[c][row + z].count;
                if (k < 0) {
                    if (strict) return fail(NF_ERR_INVALID, "nf_count_interactions: negative cell count");
                    continue;
                }
                n += k;
            }
            m[(size_t)r * cpd + z] = n;
        }
        const int64_t *mr = &m[(size_t)r * cpd];
        int64_t run = 0;
        for (int dz = -R; dz <= R; dz++) run += mr[wrap(dz)];
        for (int z = 0; z < cpd; z++) {
            mz[(size_t)r * cpd + z] = run;
            run += mr[wrap(z + R + 1)] - mr[wrap(z - R)];
        }
    }
    if (max_pencil) *max_pencil = m.empty() ? 0 : *std::max_element(m.begin(), m.end());
    uint64_t total = 0;
    for (int j = j_low; j < j_high; j++) {
        const size_t row = (size_t)j * cpd;
        const int r0 = j - j_low;  // rows r0 .. r0 + 2R hold j-R .. j+R
        for (int z = 0; z < cpd; z++) {
            const int32_t k = ci[R][row + z].count;
            if (k <= 0) continue;
            int64_t cube = 0;
            for (int dy = 0; dy < W; dy++) cube += mz[(size_t)(r0 + dy) * cpd + z];
            total += (uint64_t)k * (uint64_t)cube;
        }
    }
    *out = total;
    return NF_OK;
}

}  // namespace nf

extern "C" int nf_count_interactions(int32_t cpd, int32_t nfradius, int32_t j_low, int32_t j_high,
                                     const nf_cellinfo *const *cellinfo, uint64_t *out) {
    if (!cellinfo || !out) return nf::fail(NF_ERR_INVALID, "nf_count_interactions: null argument");
    if (cpd <= 0 || (cpd % 2) == 0) return nf::fail(NF_ERR_INVALID, "nf_count_interactions: cpd must be odd and positive");
    if (nfradius < 0 || nfradius > NF_MAX_RADIUS || 2 * nfradius + 1 > cpd)
        return nf::fail(NF_ERR_INVALID, "nf_count_interactions: nfradius out of range");
    if (j_low < 0 || j_high > cpd || j_low >= j_high)
        return nf::fail(NF_ERR_INVALID, "nf_count_interactions: bad sink row range");
    for (int c = 0; c < 2 * nfradius + 1; c++)
        if (!cellinfo[c]) return nf::fail(NF_ERR_INVALID, "nf_count_interactions: missing cellinfo slab");
    return nf::count_interactions(cpd, nfradius, j_low, j_high, cellinfo, true, out, nullptr);
}
