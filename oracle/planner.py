"""numpy restatement of the SetInteractionCollection planner (non-PARALLEL build).

TEST INFRASTRUCTURE ONLY: this is the checker for the product's C++ planner;
it is imported by tests/ and never by the product path.

Restates:
  * SetInteractionCollection ctor, PencilOnPencil.cc:180-418
    (sink pencils :254-266, padded sink blocks :271-311, source pencils
    :329-341, padded source blocks :345-366, interaction list :390-416)
  * index_to_zcen, PencilOnPencil.cc:37-43
  * SinkPencilPlan::load, PencilPlan.cc:95-138 (wrapcell :123-128)
  * SourcePencilPlan::load, PencilPlan.cc:260-279
  * NumPaddedBlocks, PencilOnPencil.cc:20-22 (NFBlockSize = 64)

A box is described by per-slab cell counts ``count[x]`` (shape cpd*cpd, cell
(y,z) at y*cpd+z) and per-slab start offsets ``start[x]`` (prefix sums of
count[x]); a slab's particles are ordered by cell.
"""
import numpy as np

NF_BLOCK = 64


def num_padded_blocks(n):
    """PencilOnPencil.cc:20-22."""
    return (np.asarray(n, dtype=np.int64) + NF_BLOCK - 1) // NF_BLOCK


def plan_sic(cpd, R, x, j_low, j_high, count, start, block=NF_BLOCK):
    """Plan the SIC of slab ``x`` for sink rows [j_low, j_high).

    ``count`` / ``start`` are indexable by slab (wrapped here).  Returns a dict
    of numpy arrays named as the reference's members.
    """
    W = 2 * R + 1
    Nk = cpd                      # k_low=0, k_high=cpd; Nk = cpd when MPI_size_z == 1
    j_width = j_high - j_low
    wrap = lambda i: i % cpd
    ks = np.arange(Nk)
    zcen = ks + R                 # index_to_zcen: k + k_low + nfradius, wrapped once
    zcen = np.where(zcen >= cpd, zcen - cpd, zcen)
    cs = np.arange(W)

    # ---- sink pencils: cells (x, y, zcen+c-R), PencilPlan.cc:95-138
    nsink = j_width * Nk
    sink_start = np.zeros((nsink, W), np.uint32)
    sink_n = np.zeros((nsink, W), np.int32)
    wrapcell = np.zeros(nsink, np.int32)
    cnt_x = np.asarray(count[wrap(x)]).reshape(cpd, cpd)
    st_x = np.asarray(start[wrap(x)]).reshape(cpd, cpd)
    zc = zcen[:, None] + cs[None, :] - R                    # [Nk, W], unwrapped
    zw = zc % cpd
    wc = np.where(zw == 0, cs[None, :], 0).max(axis=1)     # first (only) c with zc == 0 mod cpd
    for j in range(j_width):
        y = j + j_low
        rows = slice(j * Nk, (j + 1) * Nk)
        sink_start[rows] = st_x[y][zw]
        sink_n[rows] = cnt_x[y][zw]
        wrapcell[rows] = wc
    sink_count = sink_n.sum(axis=1).astype(np.int64)
    # ---- padded sink blocks, PencilOnPencil.cc:271-311
    nb = num_padded_blocks(sink_count)
    sink_set_start = (np.concatenate([[0], np.cumsum(nb)[:-1]]) * block).astype(np.int64)
    nsink_blocks = int(nb.sum())
    parent = np.repeat(np.arange(nsink, dtype=np.int32), nb)
    idmax = (sink_set_start + sink_count).astype(np.uint32)

    # ---- source pencils: cells (x+c-R, j+j_low-R, zcen), PencilPlan.cc:260-279
    nsrc = (j_width + W - 1) * Nk
    src_start = np.zeros((nsrc, W), np.uint32)
    src_n = np.zeros((nsrc, W), np.int32)
    for j in range(j_width + W - 1):
        y = wrap(j + j_low - R)
        rows = slice(j * Nk, (j + 1) * Nk)
        for c in range(W):
            xs = wrap(x + c - R)
            cn = np.asarray(count[xs]).reshape(cpd, cpd)
            sn = np.asarray(start[xs]).reshape(cpd, cpd)
            src_start[rows, c] = sn[y][zcen]
            src_n[rows, c] = cn[y][zcen]
    src_count = src_n.sum(axis=1).astype(np.int64)
    nbs = num_padded_blocks(src_count)
    src_set_start = (np.concatenate([[0], np.cumsum(nbs)[:-1]]) * block).astype(np.int64)
    nsource_blocks = int(nbs.sum())

    # ---- interactions, PencilOnPencil.cc:390-416
    cellsize = np.float32(1.0 / cpd)            # FLOAT cellsize = CP->invcpd
    jj, kk = np.meshgrid(np.arange(j_width), ks, indexing="ij")
    sinkidx = (jj * Nk + kk).ravel()
    ilist = np.zeros(W * nsink, np.int32)
    yoff = np.zeros(W * nsink, np.float32)
    for y in range(W):
        ilist[W * sinkidx + y] = ((jj + y) * Nk + kk).ravel()
        yoff[W * sinkidx + y] = np.float32(y - R) * cellsize
    src_of = ilist.reshape(nsink, W)
    direct_total = int((sink_count[:, None] * src_count[src_of]).sum())
    padded_direct_total = int((nb[:, None] * block * src_count[src_of]).sum())
    return dict(
        cpd=cpd, R=R, W=W, slab=x, j_low=j_low, j_high=j_high, Nk=Nk, j_width=j_width,
        NSinkSets=nsink, NSourceSets=nsrc, NSinkBlocks=nsink_blocks, NSourceBlocks=nsource_blocks,
        InteractionCount=W * nsink,
        SinkPlanStart=sink_start, SinkPlanN=sink_n, WrapCell=wrapcell,
        SinkSetCount=sink_count.astype(np.int32), SinkSetStart=sink_set_start.astype(np.int32),
        SinkSetIdMax=idmax, SinkBlockParentPencil=parent,
        SourcePlanStart=src_start, SourcePlanN=src_n,
        SourceSetCount=src_count.astype(np.int32), SourceSetStart=src_set_start.astype(np.int32),
        SinkSourceInteractionList=ilist, SinkSourceYOffset=yoff,
        DirectTotal=direct_total, PaddedDirectTotal=padded_direct_total,
        SinkTotal=int(sink_count.sum()), SourceTotal=int(src_count.sum()),
        PaddedSinkTotal=nsink_blocks * block, PaddedSourceTotal=nsource_blocks * block,
    )


def skewer_particles(count_slab, cpd):
    """NumParticlesInSkewer (PencilOnPencil.cc:90-108): particles in each row
    (skewer) y of a slab, the difference of consecutive rows' start indices
    (no ghosts: startindex and startindex_with_ghost agree)."""
    return np.asarray(count_slab, np.int64).reshape(cpd, cpd).sum(axis=1)


def split_slab(cpd, R, counts, nbuffers, max_sink_blocks, max_source_blocks):
    """The y-split of NearFieldDriver::ExecuteSlabGPU (directdriver.cpp:248-303):
    per skewer the worst-case sink blocks W*(np/64) + cpd and source blocks
    (sum over the W slabs of np)/64 + cpd (ComputeSinkBlocks/SourceBlocks,
    PencilOnPencil.cc:113-126), then a greedy pass that starts a new SIC when
    the running sums exceed useMaxSink or MaxSourceBlocks.  ``counts`` are the
    cell counts of slabs x-R..x+R.  Returns the list of split points (each
    SIC's end row) and raises ValueError where the reference asserts."""
    W = 2 * R + 1
    sink = [W * (int(n) // NF_BLOCK) + cpd for n in skewer_particles(counts[R], cpd)]
    src_np = sum(skewer_particles(counts[c], cpd) for c in range(W))
    source = [int(n) // NF_BLOCK + cpd for n in src_np]
    use_max_sink = sum(sink) // nbuffers                 # int / int
    use_max_sink = int(use_max_sink * (1 + 1.1 * nbuffers / cpd))  # int *= double truncates
    use_max_sink = min(use_max_sink, max_sink_blocks)
    use_max_source = max_source_blocks
    split = []
    this_sink = this_source = 0
    for j in range(cpd):
        this_sink += sink[j]
        this_source += source[(j + R) % cpd]
        if j == 0 or this_sink > use_max_sink or this_source > use_max_source:
            if split:
                split[-1] = j
            this_sink = sink[j]
            this_source = sum(source[(j - c) % cpd] for c in range(-R, R + 1))
            split.append(None)
            if len(split) >= 1024:
                raise ValueError("too many splits")
    split[-1] = cpd
    if this_sink > max_sink_blocks or this_source > max_source_blocks:
        raise ValueError("sinks or sources of the last skewer overflow the maxima")
    return split
