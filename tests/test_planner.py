"""SIC planner: product (C++, nf_plan_sic) vs oracle (numpy restatement),
plus the structural invariants of the reference's pencil decomposition."""
import numpy as np
import pytest

import nearfield as nfm
import planner
import pyoracle as po
import synth
from nfhelpers import R, params, small_uniform

KEYS = ["SinkSetStart", "SinkSetCount", "SinkSetIdMax", "SinkPlanStart", "SinkPlanN", "WrapCell",
        "SourceSetStart", "SourceSetCount", "SourcePlanStart", "SourcePlanN", "SinkBlockParentPencil",
        "SinkSourceInteractionList", "SinkSourceYOffset"]
SCALARS = ["NSinkSets", "NSourceSets", "NSinkBlocks", "NSourceBlocks", "InteractionCount", "DirectTotal",
           "PaddedDirectTotal", "SinkTotal", "SourceTotal", "PaddedSinkTotal", "PaddedSourceTotal"]


def sparse_box(cpd=9, seed=7):
    """Ragged occupancy: empty cells and a few heavy ones."""
    rng = np.random.default_rng(seed)
    pts = rng.random((cpd ** 3 * 3, 3)) - 0.5
    pts = pts[(pts[:, 2] > 0.1) | (pts[:, 1] < -0.2)]
    heavy = rng.normal([0.1, 0.2, -0.3], 0.004, (700, 3))
    return synth.box_from_global(np.concatenate([pts, heavy]), cpd)


BOXES = {
    "uniform9": lambda: small_uniform(cpd=9, ppc=15, seed=1),
    "uniform5": lambda: synth.uniform_box(5 ** 3 * 30, 5, seed=2),
    "sparse9": sparse_box,
    "clustered": lambda: synth.clustered_box(24, 11, seed=3, n_halos=20),
}


def both(box, x, jl, jh):
    eps, b2, _ = params(box)
    prod = nfm.SetInteractionCollection(box, x, jl, jh, R, eps, b2, "single_spline").arrays()
    ref = planner.plan_sic(box.cpd, R, x, jl, jh, [s.count for s in box.slabs], [s.start for s in box.slabs])
    return prod, ref


@pytest.mark.parametrize("name", sorted(BOXES))
def test_product_planner_matches_oracle(name):
    box = BOXES[name]()
    cpd = box.cpd
    for x, jl, jh in [(0, 0, cpd), (cpd // 2, 0, cpd), (cpd - 1, 1, cpd - 2), (2, cpd - 1, cpd), (1, 0, 1)]:
        prod, ref = both(box, x, jl, jh)
        for k in KEYS:
            a, b = np.asarray(prod[k]), np.asarray(ref[k])
            assert a.shape == b.shape and np.array_equal(a, b.astype(a.dtype)), (name, x, jl, jh, k)
        for k in SCALARS:
            assert int(prod[k]) == int(ref[k]), (name, k)


@pytest.mark.parametrize("name", sorted(BOXES))
def test_structure_invariants(name):
    box = BOXES[name]()
    cpd, W = box.cpd, 2 * R + 1
    for x in (0, cpd // 2):
        jl, jh = 0, cpd
        p, _ = both(box, x, jl, jh)
        Nk, jw = p["Nk"], p["j_width"]
        assert p["NSinkSets"] == jw * Nk and p["NSourceSets"] == Nk * (jw + W - 1)
        assert p["InteractionCount"] == W * p["NSinkSets"]
        # padded starts: multiples of 64, per-skewer prefix of ceil(n/64)*64
        ss, sc = p["SinkSetStart"], p["SinkSetCount"]
        assert np.all(ss % 64 == 0)
        pad = (sc.astype(np.int64) + 63) // 64 * 64
        assert np.array_equal(ss, np.concatenate([[0], np.cumsum(pad)[:-1]]))
        assert np.array_equal(p["SinkSetIdMax"], (ss + sc).astype(np.uint32))
        # interaction list and y offsets, PencilOnPencil.cc:402-407
        j, k = np.divmod(np.arange(p["NSinkSets"]), Nk)
        for y in range(W):
            assert np.array_equal(p["SinkSourceInteractionList"][W * np.arange(p["NSinkSets"]) + y], (j + y) * Nk + k)
            assert np.all(p["SinkSourceYOffset"][y::W] == np.float32(y - R) * np.float32(1.0 / cpd))
        # every sink cell sits in exactly W sink pencils, at W distinct positions
        seen = {}
        for s in range(p["NSinkSets"]):
            for c in range(W):
                key = (int(p["SinkPlanStart"][s, c]), int(p["SinkPlanN"][s, c]), s // Nk)
                seen.setdefault(key, set()).add(c)
        for (st, n, jj), cs in seen.items():
            assert cs == set(range(W)), (st, n, jj)
        # each sink cell's neighbourhood = the (2R+1)^3 cube, each cell once
        counts = box.counts()
        for jj in (0, jw - 1):
            for z in (0, cpd - 1, cpd // 2):
                cells = []
                for c in range(W):
                    kk = (z - c) % cpd                       # pencil k holds cell z at position c
                    s = jj * Nk + kk
                    zc = (kk + R) % cpd
                    for y in range(W):
                        src = p["SinkSourceInteractionList"][W * s + y]
                        ys = (src // Nk + jl - R) % cpd
                        for cx in range(W):
                            cells.append(((x - R + cx) % cpd, ys, zc, c))
                # the W pencils of cell z see source z-planes zc = z - c + R; with
                # the sink's own z offset (c - R) this is the full z range.
                rel = sorted(((cx - x) % cpd, (ys - (jj + jl)) % cpd, (zc - z) % cpd) for cx, ys, zc, c in cells)
                cube = sorted((a % cpd, b % cpd, c % cpd) for a in range(-R, R + 1) for b in range(-R, R + 1)
                              for c in range(-R, R + 1))
                assert rel == cube
        # totals against the brute-force count (DirectBench.cpp CalcDI)
        assert p["DirectTotal"] == po.calc_di(box, R, x)
        assert p["PaddedDirectTotal"] == int((pad[:, None] * p["SourceSetCount"][
            p["SinkSourceInteractionList"].reshape(-1, W)]).sum())


def test_direct_total_sums_over_splits():
    box = BOXES["sparse9"]()
    cpd = box.cpd
    for x in (0, 4):
        whole, _ = both(box, x, 0, cpd)
        parts = [both(box, x, a, b)[0] for a, b in [(0, 3), (3, 4), (4, cpd)]]
        assert sum(q["DirectTotal"] for q in parts) == whole["DirectTotal"]


@pytest.mark.parametrize("name", sorted(BOXES))
def test_count_interactions_is_calc_di_and_direct_total(name):
    """nf_count_interactions (separable cube sums over the cellinfo counts)
    equals the oracle's brute-force CalcDI over whole slabs, and the
    planner's DirectTotal over row ranges, including ranges that wrap the
    source rows round the box and single rows."""
    box = BOXES[name]()
    cpd = box.cpd
    for x in (0, cpd // 2, cpd - 1):
        assert nfm.count_interactions(box, x) == po.calc_di(box, R, x), (name, x)
        for jl, jh in [(0, 1), (0, 3), (cpd - 2, cpd), (1, cpd - 1), (cpd // 2, cpd // 2 + 1)]:
            prod, _ = both(box, x, jl, jh)
            assert nfm.count_interactions(box, x, jl, jh) == int(prod["DirectTotal"]), (name, x, jl, jh)


def test_count_interactions_radius_one_and_errors():
    box = BOXES["uniform9"]()
    for x in (0, 5):
        assert nfm.count_interactions(box, x, nfradius=1) == po.calc_di(box, 1, x)
    bad = synth.Box(box.cpd, [synth.Slab(s.pos, s.count) for s in box.slabs], "negative count")
    bad.slabs[0].cellinfo["count"][7] = -1
    with pytest.raises(nfm.NFError, match="negative"):
        nfm.count_interactions(bad, 0)
    for jl, jh in [(0, 0), (-1, 2), (3, box.cpd + 1)]:
        with pytest.raises(nfm.NFError, match="row range"):
            nfm.count_interactions(box, 0, jl, jh)


def test_coadd_set_then_add_equals_sum():
    """The k-ordered set/add co-add (PencilPlan.cc:140-207) equals zero-then-add:
    with integer-valued partials every summation order is exact."""
    box = small_uniform(cpd=7, ppc=10, seed=3)
    x = 3
    p, _ = both(box, x, 0, box.cpd)
    rng = np.random.default_rng(0)
    part = rng.integers(-1000, 1000, (p["NSinkBlocks"] * 64, 4)).astype(np.float32)
    acc = po.coadd_f32(p, part, box.slab(x).n, R)
    expect = np.zeros((box.slab(x).n, 4))
    W = 2 * R + 1
    for s in range(p["NSinkSets"]):
        off = p["SinkSetStart"][s]
        for c in range(W):
            st, n = int(p["SinkPlanStart"][s, c]), int(p["SinkPlanN"][s, c])
            expect[st:st + n] += part[off:off + n]
            off += n
    assert np.array_equal(acc, expect.astype(np.float32))


def test_gather_restatement_layout():
    """Sink gather: each pencil is its W cells in order with z shifted by
    (c-R)*fl32(1/cpd); source gather: W slabs' cells with x shifted."""
    box = small_uniform(cpd=7, ppc=8, seed=5)
    x = 1
    p, _ = both(box, x, 0, box.cpd)
    k4 = po.gather_sinks(p, box, x, R)
    s4 = po.gather_sources(p, box, x, R)
    cs = np.float32(1.0 / box.cpd)
    sl = box.slab(x)
    for s in (0, 5, p["NSinkSets"] - 1):
        o = p["SinkSetStart"][s]
        for c in range(2 * R + 1):
            st, n = int(p["SinkPlanStart"][s, c]), int(p["SinkPlanN"][s, c])
            dz = np.float32(c - R) * cs
            xyz = sl.xyz[st:st + n]
            assert np.array_equal(k4[o:o + n, 0], xyz[:, 0])
            assert np.array_equal(k4[o:o + n, 2], (xyz[:, 2] + dz) if dz != 0 else xyz[:, 2])
            o += n
    for s in (0, 9, p["NSourceSets"] - 1):
        o = p["SourceSetStart"][s]
        for c in range(2 * R + 1):
            st, n = int(p["SourcePlanStart"][s, c]), int(p["SourcePlanN"][s, c])
            xyz = box.slab(x - R + c).xyz[st:st + n]
            dx = np.float32(c - R) * cs
            assert np.array_equal(s4[o:o + n, 0], (xyz[:, 0] + dx) if dx != 0 else xyz[:, 0])
            assert np.array_equal(s4[o:o + n, 1:3], xyz[:, 1:3])
            o += n


def c_split(box, x, nbuffers, max_sink, max_source, R_=R):
    import ctypes as C
    W = 2 * R_ + 1
    slabs = [box.slab(x - R_ + c) for c in range(W)]
    ci = (C.c_void_p * W)(*[s.cellinfo.ctypes.data for s in slabs])
    nsl = (C.c_uint64 * W)(*[s.n for s in slabs])
    out = (C.c_int32 * box.cpd)()
    n = C.c_int32(0)
    rc = nfm.lib().nf_split_slab(box.cpd, R_, ci, nsl, nbuffers, max_sink, max_source, out, C.byref(n))
    return rc, [int(out[i]) for i in range(n.value)]


def uniform_cells(cpd, per_cell):
    """A box with exactly per_cell particles in every cell."""
    slabs = []
    for _ in range(cpd):
        count = np.full(cpd * cpd, per_cell, np.int32)
        slabs.append(synth.Slab(np.zeros(3 * int(count.sum()), np.float32), count))
    return synth.Box(cpd, slabs)


def test_split_slab_known_answers():
    """ExecuteSlabGPU's split worked by hand: cpd 5, R 1, 64 particles per cell.
    Every skewer: 3*(320/64)+5 = 20 sink blocks and (3*320)/64+5 = 20 source
    blocks; the slab's 100 sink blocks over nbuffers, scaled by 1+1.1*nb/5."""
    box = uniform_cells(5, 64)
    ok, big = 0, 1000
    # nbuffers 2: useMaxSink = int(50 * 1.44) = 72 -> rows [0,3) then [3,5)
    assert planner.split_slab(5, 1, [s.count for s in box.slabs[:3]], 2, big, big) == [3, 5]
    assert c_split(box, 1, 2, big, big, R_=1) == (ok, [3, 5])
    # nbuffers 1: useMaxSink = int(100 * 1.22) = 122 -> one SIC
    assert planner.split_slab(5, 1, [s.count for s in box.slabs[:3]], 1, big, big) == [5]
    assert c_split(box, 1, 1, big, big, R_=1) == (ok, [5])
    # source cap 70: a skewer's 3-row source window is 60, two rows' 80 -> one row per SIC
    assert planner.split_slab(5, 1, [s.count for s in box.slabs[:3]], 1, big, 70) == [1, 2, 3, 4, 5]
    assert c_split(box, 1, 1, big, 70, R_=1) == (ok, [1, 2, 3, 4, 5])
    # sink cap 15 < one skewer's 20: the reference asserts; the product returns NF_ERR_CAPACITY
    with pytest.raises(ValueError):
        planner.split_slab(5, 1, [s.count for s in box.slabs[:3]], 1, 15, big)
    assert c_split(box, 1, 1, 15, big, R_=1)[0] == -3


@pytest.mark.parametrize("name", sorted(BOXES))
def test_split_slab_matches_oracle(name):
    """nf_split_slab == the oracle's restatement over worker counts and block
    maxima, and every split covers all rows in increasing order."""
    box = BOXES[name]()
    for x in (0, box.cpd // 2, box.cpd - 1):
        counts = [box.slab(x - R + c).count for c in range(2 * R + 1)]
        for nb in (1, 2, 3, 8):
            for max_sink, max_source in ((1 << 20, 1 << 20), (400, 1 << 20), (1 << 20, 250), (300, 200)):
                try:
                    ref = planner.split_slab(box.cpd, R, counts, nb, max_sink, max_source)
                except ValueError:
                    ref = None
                rc, got = c_split(box, x, nb, max_sink, max_source)
                if ref is None:
                    assert rc == -3, (name, x, nb, max_sink, max_source)
                    continue
                assert rc == 0 and got == ref, (name, x, nb, max_sink, max_source, got, ref)
                assert got[-1] == box.cpd and all(a < b for a, b in zip(got, got[1:]))
