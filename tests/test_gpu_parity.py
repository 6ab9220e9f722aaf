"""GPU parity of the HIP near-field path against the oracle (run with -m gpu).

Every test calls the product through the C ABI (libnfdirect.so via
nearfield.py) and checks it against oracle/ on the same seeded inputs.
"""
import os

import numpy as np
import pytest

import nearfield as nfm
import pyoracle as po
import synth
from nfhelpers import (R, not_soft_dominated, oracle_slab, params, rel_err, run_sics, small_uniform,
                       summarize)

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_box(box, path, softening="single_spline", slabs=None, nsplit=1, sinks_per_lane=0, bpd=2, waves_per_wg=0):
    """Run the listed slabs of a host-resident box; returns {slab: acc (n,4)}."""
    eps, b2, _ = params(box, softening)
    ell = box.mean_spacing()
    out = {}
    with nfm.NearFieldDriver(box.cpd, ell / 40.0, b2, nfradius=R, softening=softening, path=path,
                             buffers_per_device=bpd, max_buffer_bytes=2 << 30,
                             sinks_per_lane=sinks_per_lane, waves_per_wg=waves_per_wg) as drv:
        assert drv.eps == eps
        slabs = range(box.cpd) if slabs is None else slabs
        for x in slabs:
            acc = np.full((box.slab(x).n, 4), np.nan, np.float32)
            drv.execute_slab(box, x, acc, nsplit=nsplit)
            out[x] = acc
        for x in slabs:
            drv.finalize(x)
    return out


@pytest.fixture(scope="module")
def box9():
    return small_uniform(cpd=9, ppc=20, seed=11)


@pytest.fixture(scope="module")
def ref9(box9):
    eps, b2, _ = params(box9)
    return {x: oracle_slab(box9, x, "single_spline", eps, b2, with_soft=True) for x in (0, 4, 8)}


def test_plain_matches_pencil_restatement(box9):
    """Plain path == the pencil-structured FP32 restatement (same pencil
    decomposition, tile sums and co-add order), up to rsqrt rounding."""
    eps, b2, _ = params(box9)
    acc = run_box(box9, "plain", slabs=[0, 4, 8])
    for x in (0, 4, 8):
        sic = nfm.SetInteractionCollection(box9, x, 0, box9.cpd, R, eps, b2, "single_spline")
        plan = sic.arrays()
        k4 = po.gather_sinks(plan, box9, x, R)
        s4 = po.gather_sources(plan, box9, x, R)
        part = po.pencil_execute(plan, k4, s4, R, "single_spline", eps, b2)
        ref = po.coadd_f32(plan, part, box9.slab(x).n, R)
        assert np.isfinite(acc[x]).all()
        st = summarize(rel_err(acc[x], ref))
        assert st["rms"] < 1e-6 and st["max"] < 2e-5, st
        np.testing.assert_allclose(acc[x][:, 3], ref[:, 3], rtol=1e-5, atol=1e-6 * float(b2))


def test_plain_and_batched_vs_fp64(box9, ref9):
    eps, b2, e = params(box9)
    res = {p: run_box(box9, p, slabs=[0, 4, 8]) for p in ("plain", "batched")}
    for x in (0, 4, 8):
        ref, soft = ref9[x]
        keep = not_soft_dominated(ref, soft)
        for p in ("plain", "batched"):
            acc = res[p][x]
            assert np.isfinite(acc).all(), p
            st = summarize(rel_err(acc, ref), keep)
            lim_rms, lim_max = (3e-6, 1e-4) if p == "plain" else (1e-6, 1e-5)
            assert st["rms"] < lim_rms and st["max"] < lim_max, (p, x, st)
            # FOF density w (self term b2 included) against the FP64 sum
            np.testing.assert_allclose(acc[:, 3], ref[:, 3], rtol=2e-6, atol=1e-6 * float(b2))


def test_reference_ewald_positions():
    """The particle positions of the reference's own force-accuracy test
    (Tests/Ewald, 65,536 particles, tests/golden/ewald_pos_f32.npy) at CPD 11
    (~49 per cell): both paths against the FP64 oracle, at the same bars as
    the synthetic boxes."""
    pos = np.load(os.path.join(GOLDEN, "ewald_pos_f32.npy")).astype(np.float64)
    box = synth.box_from_global(pos, 11, name="ewald")
    eps, b2, _ = params(box)
    slabs = [0, 5, 10]
    res = {p: run_box(box, p, slabs=slabs) for p in ("plain", "batched")}
    for x in slabs:
        ref, soft = oracle_slab(box, x, "single_spline", eps, b2, with_soft=True)
        keep = not_soft_dominated(ref, soft)
        for p in ("plain", "batched"):
            assert np.isfinite(res[p][x]).all(), p
            st = summarize(rel_err(res[p][x], ref), keep)
            lim_rms, lim_max = (3e-6, 1e-4) if p == "plain" else (1e-6, 1e-5)
            assert st["rms"] < lim_rms and st["max"] < lim_max, (p, x, st)


@pytest.mark.parametrize("S", [1, 2, 4])
def test_batched_register_tiles_agree(box9, ref9, S):
    """Every register-tile shape gives the same physics (within FP32)."""
    acc = run_box(box9, "batched", slabs=[4], sinks_per_lane=S)[4]
    ref, soft = ref9[4]
    st = summarize(rel_err(acc, ref), not_soft_dominated(ref, soft))
    assert st["rms"] < 1e-6 and st["max"] < 1e-5, (S, st)


def test_register_tiles_bitwise_equal(box9):
    """A sink's sum does not depend on the tile shape: how many sinks a lane
    holds (S), how many waves share a workgroup, which half of a tile it
    sits in (each half picks its own pair form), or a heavy register tile
    being split into one-wave tiles.  All nine shapes (S = 1, 2, 4 times 1,
    2, 4 waves per workgroup) give bitwise the same accelerations.  The heavy
    cell (1,500 particles) makes register tiles costly enough to be split."""
    cpd = 7
    rng = np.random.default_rng(3)
    pts = rng.random((cpd ** 3 * 12, 3)) - 0.5
    heavy = synth.box_from_global(np.concatenate([pts, rng.normal(0.0, 0.01, (1500, 3))]), cpd)
    shapes = [(S, w) for S in (1, 2, 4) for w in (1, 2, 4)]
    for box, slabs in ((box9, [0, 4]), (heavy, [2, 3, 4])):
        res = {sh: run_box(box, "batched", slabs=slabs, sinks_per_lane=sh[0], waves_per_wg=sh[1]) for sh in shapes}
        for x in slabs:
            assert np.isfinite(res[(1, 4)][x]).all()
            for sh in shapes:
                assert np.array_equal(res[(1, 4)][x], res[sh][x]), (box.cpd, x, sh)


def test_split_sics_equal_whole_slab(box9):
    """A slab planned as 3 SICs over y, or split as ExecuteSlabGPU splits it
    (nf_split_slab, here into several SICs for 3 workers), gives bitwise the
    same accelerations as one SIC (each SIC covers whole skewers; co-add is
    per skewer)."""
    for path in ("plain", "batched"):
        a1 = run_box(box9, path, slabs=[2], nsplit=1)[2]
        a3 = run_box(box9, path, slabs=[2], nsplit=3)[2]
        ar = run_box(box9, path, slabs=[2], nsplit="reference", bpd=3)[2]
        assert np.array_equal(a1, a3), path
        assert np.array_equal(a1, ar), path
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, params(box9)[1], buffers_per_device=3,
                             max_buffer_bytes=2 << 30) as drv:
        assert len(drv.split_slab(box9, 2)) >= 2


def test_deterministic(box9):
    a = run_box(box9, "batched", slabs=[1])[1]
    b = run_box(box9, "batched", slabs=[1])[1]
    assert np.array_equal(a, b)


@pytest.mark.parametrize("softening", ["plummer", "cubic_spline", "cubic_plummer"])
def test_other_softenings_vs_fp64(softening):
    box = small_uniform(cpd=7, ppc=15, seed=5)
    eps, b2, _ = params(box, softening)
    for path in ("plain", "batched"):
        acc = run_box(box, path, softening=softening, slabs=[3])[3]
        ref, soft = oracle_slab(box, 3, softening, eps, b2, with_soft=True)
        st = summarize(rel_err(acc, ref), not_soft_dominated(ref, soft))
        assert st["rms"] < 3e-6 and st["max"] < 1e-4, (softening, path, st)


def test_empty_and_heavy_cells():
    """Ragged occupancy: empty cells, an empty pencil region and one heavy
    cell (1,500 particles) next to it."""
    cpd = 7
    rng = np.random.default_rng(3)
    pts = [rng.random((cpd ** 3 * 4, 3)) - 0.5]
    pts[0] = pts[0][pts[0][:, 0] > -0.2]                       # slabs 0-1 empty
    pts.append(rng.normal(0.0, 0.01, (1500, 3)))                # heavy cell at the centre
    box = synth.box_from_global(np.concatenate(pts), cpd)
    assert box.slab(0).n == 0 and box.counts().max() >= 1000
    eps, b2, _ = params(box)
    for path in ("plain", "batched"):
        res = run_box(box, path, slabs=[0, 1, 3])
        assert res[0].shape == (0, 4)
        for x in (1, 3):
            ref, soft = oracle_slab(box, x, "single_spline", eps, b2, with_soft=True)
            st = summarize(rel_err(res[x], ref), not_soft_dominated(ref, soft))
            assert st["rms"] < 3e-6, (path, x, st)


def test_heavy_pencils_split_gather():
    """Source pencils of more than 8 blocks (512 particles) have their
    batched-path gather split over several waves: the pencil's own wave fills
    its first 8 blocks, one wave per 8 blocks of the stream the rest.  Cells
    of 511, 512, 513, 1,100 and 3,000 particles (uniform within their cells)
    make pencils on both sides of that edge.  The group gather (nf_submit,
    host plans), the device-planned path (nf_submit_slab) and the per-SIC
    gather of nf_execute_sic_noalloc must give bitwise the same
    accelerations, and those must meet the FP64 oracle: a block written
    twice or not at all would miss or double 64 sources of a sink's sum."""
    cpd = 7
    rng = np.random.default_rng(29)
    w = 1.0 / cpd
    centre = lambda i: (np.asarray(i) + 0.5) * w - 0.5
    heavy = {(3, 1, 1): 511, (3, 1, 4): 512, (3, 4, 2): 513, (2, 5, 5): 1100, (3, 3, 6): 3000}
    bg = rng.random((cpd ** 3 * 20, 3)) - 0.5
    bg_cell = np.floor((bg + 0.5) * cpd).astype(int)
    pts = [bg[~np.isin(bg_cell @ [cpd * cpd, cpd, 1], [x * cpd * cpd + y * cpd + z for x, y, z in heavy])]]
    for cell, n in heavy.items():
        pts.append(centre(cell) + (rng.random((n, 3)) - 0.5) * w * 0.999)
    box = synth.box_from_global(np.concatenate(pts), cpd)
    for (x, y, z), n in heavy.items():
        assert box.slab(x).count[y * cpd + z] == n, (x, y, z)
    eps, b2, _ = params(box)
    slabs = [1, 3, 5]
    ell = box.mean_spacing()
    with nfm.NearFieldDriver(cpd, ell / 40.0, b2, nfradius=R, path="batched", buffers_per_device=2,
                             max_buffer_bytes=1 << 30) as drv:
        host = run_sics(drv, box, slabs, nfm.NF_MEM_DEVICE, planner="host")
        dev = run_sics(drv, box, slabs, nfm.NF_MEM_DEVICE, planner="device")
        # the per-SIC gather: slab 3 through nf_execute_sic_noalloc
        pos = [nfm.DeviceArray(host=box.slab(x).pos) for x in range(cpd)]
        acc = nfm.DeviceArray(host=np.full((box.slab(3).n, 4), np.nan, np.float32))
        try:
            sic = drv.plan(box, 3, pos_ptrs=[pos[(3 - R + c) % cpd].ptr for c in range(2 * R + 1)],
                           acc_ptr=acc.ptr, mem=nfm.NF_MEM_DEVICE)
            assert sic.arrays()["SourceSetCount"].max() > 8 * 64  # the gather is split
            nfm.execute_sic_noalloc(sic)
            noalloc = acc.to_host(np.float32, (box.slab(3).n, 4))
        finally:
            for d in pos + [acc]:
                d.free()
    assert np.array_equal(noalloc.view(np.uint32), host[3].view(np.uint32))
    for x in slabs:
        assert np.array_equal(host[x].view(np.uint32), dev[x].view(np.uint32)), x
        ref, soft = oracle_slab(box, x, "single_spline", eps, b2, with_soft=True)
        st = summarize(rel_err(host[x], ref), not_soft_dominated(ref, soft))
        assert st["rms"] < 1e-6 and st["max"] < 1e-4, (x, st)


def test_batched_source_order_regimes():
    """The batched path orders each source cell's particles by x-bin: cells of
    up to 64 particles (one chunk), 65-128 (two chunks), and larger cells kept
    in the reference's order.  Cells of 64, 65, 100, 128, 129 and 300
    particles, some particles exactly on bin edges and some just outside
    their cell: every source must land in the stream exactly once, so both
    paths meet the FP64 oracle at the bars of the uniform boxes."""
    cpd = 7
    rng = np.random.default_rng(21)
    w = 1.0 / cpd
    centre = lambda i: (np.asarray(i) + 0.5) * w - 0.5
    extra = {(2, 1, 1): 64, (2, 4, 5): 65, (3, 2, 2): 100, (3, 5, 0): 128, (4, 3, 3): 129, (3, 0, 6): 300}
    bg = rng.random((cpd ** 3 * 30, 3)) - 0.5
    bg_cell = np.floor((bg + 0.5) * cpd).astype(int)
    pts = [bg[~np.isin(bg_cell @ [cpd * cpd, cpd, 1], [x * cpd * cpd + y * cpd + z for x, y, z in extra])]]
    for cell, n in extra.items():
        p = centre(cell) + (rng.random((n, 3)) - 0.5) * w * 0.999
        k = np.arange(n // 8) % 7 + 1
        p[: n // 8, 0] = centre(cell)[0] + (k / 8.0 - 0.5) * w  # on the interior bin edges
        pts.append(p)
    box = synth.box_from_global(np.concatenate(pts), cpd)
    for (x, y, z), n in extra.items():
        assert box.slab(x).count[y * cpd + z] == n, (x, y, z)
    # cell-local x just past the cell edge (a particle that has drifted): the bins clamp
    s3 = box.slab(3)
    s3.pos[:5] = np.float32(0.5 * w * 1.01)
    eps, b2, _ = params(box)
    slabs = [2, 3, 4]
    res = {p: run_box(box, p, slabs=slabs) for p in ("plain", "batched")}
    for x in slabs:
        ref, soft = oracle_slab(box, x, "single_spline", eps, b2, with_soft=True)
        keep = not_soft_dominated(ref, soft)
        for p in ("plain", "batched"):
            assert np.isfinite(res[p][x]).all(), p
            st = summarize(rel_err(res[p][x], ref), keep)
            lim_rms, lim_max = (3e-6, 1e-4) if p == "plain" else (1e-6, 1e-5)
            assert st["rms"] < lim_rms and st["max"] < lim_max, (p, x, st)


def test_radius_one_vs_fp64():
    """NearFieldRadius 1 (pencils of W = 3 cells): both paths against the FP64
    oracle summed over the 3x3x3 neighbourhood, at the bars of radius 2."""
    box = small_uniform(cpd=7, ppc=20, seed=5)
    eps, b2, _ = params(box)
    slabs = [0, 3, 6]
    for path in ("plain", "batched"):
        with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, nfradius=1, path=path,
                                 buffers_per_device=2, max_buffer_bytes=1 << 30) as drv:
            accs = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in slabs}
            for x in slabs:
                drv.execute_slab(box, x, accs[x], nsplit=1)
            for x in slabs:
                drv.finalize(x)
        for x in slabs:
            sl, ix, ce = po.slab_sinks(box, x)
            ref, soft = po.direct_f64(box, 1, "single_spline", eps, b2, sl, ix, ce, with_soft=True)
            assert np.isfinite(accs[x]).all(), (path, x)
            st = summarize(rel_err(accs[x], ref), not_soft_dominated(ref, soft))
            lim_rms, lim_max = (3e-6, 1e-4) if path == "plain" else (1e-6, 1e-5)
            assert st["rms"] < lim_rms and st["max"] < lim_max, (path, x, st)


def test_clustered_box_vs_fp64():
    """A clustered box (Plummer halos in a GRF lattice; cells from empty to
    several hundred particles): softened pairs are common and pencils mix
    every source-order regime.  Against the FP64 oracle, with the bars of the
    precision study: rms, the max over sinks of condition number <= 100, and
    the max of error / condition number over all sinks (the condition number
    of a sink is the sum of its pair terms' magnitudes over |a|)."""
    box = synth.clustered_box(40, 11, seed=2, n_halos=60)
    assert box.counts().max() > 128
    eps, b2, _ = params(box)
    slabs = [0, 5, 10]
    res = {p: run_box(box, p, slabs=slabs) for p in ("plain", "batched")}
    for x in slabs:
        sl, ix, ce = po.slab_sinks(box, x)
        ref, soft, mag = po.direct_f64(box, R, "single_spline", eps, b2, sl, ix, ce, with_soft=True, with_abs=True)
        keep = not_soft_dominated(ref, soft)
        cond = mag / np.maximum(np.linalg.norm(ref[:, :3], axis=1), 1e-300)
        for p, (lim_rms, lim_well, lim_cond) in (("plain", (3e-6, 1e-4, 3e-6)), ("batched", (1e-6, 1e-5, 1e-6))):
            assert np.isfinite(res[p][x]).all(), p
            rel = rel_err(res[p][x], ref)
            st = summarize(rel, keep)
            well = summarize(rel, keep & (cond <= 100.0))
            assert st["rms"] < lim_rms and well["max"] < lim_well, (p, x, st, well)
            assert float((rel[keep] / cond[keep]).max()) < lim_cond, (p, x)


def test_density_disabled():
    """b2 = 0 turns the FOF density off: w is exactly 0 on both paths and the
    accelerations meet the FP64 oracle (computed with b2 = 0)."""
    box = small_uniform(cpd=7, ppc=20, seed=8)
    eps, _, _ = params(box)
    slabs = [1, 4]
    for path in ("plain", "batched"):
        with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, 0.0, nfradius=R, path=path,
                                 buffers_per_device=2, max_buffer_bytes=1 << 30) as drv:
            accs = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in slabs}
            for x in slabs:
                drv.execute_slab(box, x, accs[x], nsplit=1)
            for x in slabs:
                drv.finalize(x)
        for x in slabs:
            ref, soft = oracle_slab(box, x, "single_spline", eps, 0.0, with_soft=True)
            assert np.all(accs[x][:, 3] == 0.0), (path, x)
            st = summarize(rel_err(accs[x], ref), not_soft_dominated(ref, soft))
            lim_rms, lim_max = (3e-6, 1e-4) if path == "plain" else (1e-6, 1e-5)
            assert st["rms"] < lim_rms and st["max"] < lim_max, (path, x, st)


def test_minimum_cpd_wraps():
    """cpd = 2R+1: every pencil wraps, every source slab is distinct."""
    box = synth.uniform_box(5 ** 3 * 12, 5, seed=9)
    eps, b2, _ = params(box)
    for path in ("plain", "batched"):
        acc = run_box(box, path, slabs=[0])[0]
        ref, soft = oracle_slab(box, 0, "single_spline", eps, b2, with_soft=True)
        st = summarize(rel_err(acc, ref), not_soft_dominated(ref, soft))
        assert st["rms"] < 3e-6, (path, st)


def test_device_resident_matches_host():
    """NF_MEM_DEVICE (slabs and acceleration slab resident in HBM) gives
    bitwise the host-mode result."""
    box = small_uniform(cpd=7, ppc=20, seed=21)
    eps, b2, _ = params(box)
    host = run_box(box, "batched", slabs=[3])[3]
    with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                             max_buffer_bytes=1 << 30) as drv:
        dev = [nfm.DeviceArray(host=box.slab(x).pos) for x in range(box.cpd)]
        acc = nfm.DeviceArray(nbytes=box.slab(3).n * 16)
        acc.zero()
        ptrs = [dev[(3 - R + c) % box.cpd].ptr for c in range(2 * R + 1)]
        sic = drv.plan(box, 3, pos_ptrs=ptrs, acc_ptr=acc.ptr, mem=nfm.NF_MEM_DEVICE)
        t = drv.submit(sic)
        t.wait()
        t.release()
        got = acc.to_host(np.float32, (box.slab(3).n, 4))
        for d in dev + [acc]:
            d.free()
    assert np.array_equal(got, host)


def test_launch_groups_match_single_launches(box9):
    """Several SICs sharing one persistent launch give bitwise the results of
    one launch per SIC (work items are independent; order does not matter)."""
    eps, b2, _ = params(box9)
    out = {}
    for mg in (1, 8):
        with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                                 max_buffer_bytes=2 << 30, sinks_per_lane=2, max_group=mg) as drv:
            accs = {x: np.full((box9.slab(x).n, 4), np.nan, np.float32) for x in range(box9.cpd)}
            for x in range(box9.cpd):
                drv.execute_slab(box9, x, accs[x])
            sizes = []
            for x in range(box9.cpd):
                sizes += [s["group_size"] for s in drv.finalize(x)]
            out[mg] = (accs, sizes)
    assert max(out[1][1]) == 1 and max(out[8][1]) > 1
    for x in range(box9.cpd):
        assert np.array_equal(out[1][0][x], out[8][0][x]), x


def test_census_counts_every_pair():
    """The batched kernel's census of pair forms adds up to the launch
    group's DirectTotal: every pair is computed exactly once, in one form.
    With the FOF density off (b2 = 0) no pair takes the density form, and a
    softening far larger than the cells puts every pair in the softened form.
    The census build of the kernel gives bitwise the accelerations of the
    plain build."""
    box = small_uniform(cpd=7, ppc=20, seed=23)
    ell = box.mean_spacing()
    for soft_len, b2, check in ((ell / 40.0, synth.fof_b2(ell), None), (ell / 40.0, 0.0, "no_density"),
                                (3.0 / box.cpd, synth.fof_b2(ell), "all_exact")):
        accs = {}
        with nfm.NearFieldDriver(box.cpd, soft_len, b2, path="batched", buffers_per_device=1,
                                 max_buffer_bytes=1 << 30, max_group=1, census=1) as drv:
            for x in (0, 3):
                acc = accs[x] = np.full((box.slab(x).n, 4), np.nan, np.float32)
                drv.execute_slab(box, x, acc, blocking=True)
                for st in drv.finalize(x):
                    f = st["group_form_interactions"]
                    assert sum(f) == st["group_direct_total"] == st["direct_total"], (x, f, st["direct_total"])
                    assert 0 < st["group_tail_interactions"] < st["direct_total"]
                    if check == "no_density":
                        assert f[1] == 0 and f[0] > 0, f
                    if check == "all_exact":
                        assert f[2] == st["direct_total"], f
        if check is None:
            plain_build = run_box(box, "batched", slabs=[0, 3], bpd=1)
            for x in (0, 3):
                assert np.array_equal(accs[x], plain_build[x]), x


def test_autotuner_keys_on_occupancy(box9, ref9):
    eps, b2, _ = params(box9)
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                             max_buffer_bytes=2 << 30, sinks_per_lane=0, max_group=1) as drv:
        accs = {}
        for x in (0, 4, 8, 2):
            accs[x] = np.full((box9.slab(x).n, 4), np.nan, np.float32)
            drv.execute_slab(box9, x, accs[x], blocking=True)
            drv.finalize(x)
        table = drv.autotune_table()
    assert table and all(e["sinks_per_lane"] in (1, 2, 4) and e["waves_per_wg"] in (1, 2, 4) for e in table)
    assert sum(e["uses"] for e in table) == 4
    # two axes searched: sinks per lane x waves per workgroup
    assert all(len(e["candidates"]) == 9 and len(e["trial_ms"]) == 9 for e in table)
    assert {tuple(c) for c in table[0]["candidates"]} == {(S, w) for S in (1, 2, 4) for w in (1, 2, 4)}
    # Every pencil holds 5 cells x ~20 sinks: one key, tuned once, reused.
    assert len(table) == 1 and table[0]["uses"] == 4, table
    assert table[0]["occupancy_quarters"] == [0, 4, 0, 0] and table[0]["size_class"] == 0
    for x in (0, 4, 8):
        ref, soft = ref9[x]
        st = summarize(rel_err(accs[x], ref), not_soft_dominated(ref, soft))
        assert st["rms"] < 1e-6 and st["max"] < 1e-5, (x, st)


def test_graph_replays_match_eager(box9):
    """Groups captured as HIP graphs (re-captured, updated in place and
    replayed on later groups) give bitwise the eager results, on every
    replay, with eager device work and synchronisations between replays."""
    eps, b2, _ = params(box9)
    slabs = [2, 3, 4]
    eager = run_box(box9, "batched", slabs=slabs, sinks_per_lane=2)
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                             max_buffer_bytes=2 << 30, sinks_per_lane=2, waves_per_wg=4, max_group=1,
                             use_graphs=1) as drv:
        for rep in range(3):
            accs = {x: np.full((box9.slab(x).n, 4), np.nan, np.float32) for x in slabs}
            for x in slabs:
                drv.execute_slab(box9, x, accs[x], blocking=True)
                st = drv.finalize(x)
                assert all(s["graph"] == 1 for s in st)
            for x in slabs:
                assert np.array_equal(accs[x], eager[x]), (rep, x)
            # eager device work and a synchronisation between replays
            scratch = nfm.DeviceArray(host=np.arange(1 << 16, dtype=np.float32))
            assert np.array_equal(scratch.to_host(np.float32, (1 << 16,)), np.arange(1 << 16, dtype=np.float32))
            scratch.free()
            nfm.device_synchronize()


def test_inconsistent_descriptor_is_rejected(box9):
    """A descriptor whose arrays disagree is refused before any GPU work
    (error code, not a fault): block parents, counts and cell ranges."""
    eps, b2, _ = params(box9)
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                             max_buffer_bytes=1 << 30) as drv:
        acc = np.zeros((box9.slab(3).n, 4), np.float32)
        for field, mutate in (("SinkBlockParentPencil", lambda a: a.__setitem__(0, a[-1])),
                              ("SinkSetCount", lambda a: a.__setitem__(5, a[5] + 1)),
                              ("SinkSourceInteractionList", lambda a: a.__setitem__(3, 10 ** 6))):
            sic = drv.plan(box9, 3, acc_ptr=acc.ctypes.data)
            n = {"SinkBlockParentPencil": sic.desc.NSinkBlocks, "SinkSetCount": sic.desc.NSinkSets,
                 "SinkSourceInteractionList": sic.desc.InteractionCount}[field]
            arr = np.ctypeslib.as_array((nfm.C.c_int32 * n).from_address(getattr(sic.desc, field))).copy()
            mutate(arr)
            setattr(sic.desc, field, arr.ctypes.data)
            sic._keep = arr
            with pytest.raises(nfm.NFError):
                drv.submit(sic, blocking=True)
        # the driver still works afterwards
        t = drv.submit(drv.plan(box9, 3, acc_ptr=acc.ctypes.data), blocking=True)
        t.release()
        assert np.isfinite(acc).all()


def test_stream_session_reuses_slabs(box9):
    """Inside a stream session, host slabs already in HBM are reused by later
    SICs (cache hits), results are bitwise those of uploads every time, and a
    new session sees host data changed in between."""
    eps, b2, _ = params(box9)
    kw = dict(path="batched", buffers_per_device=2, max_buffer_bytes=2 << 30, sinks_per_lane=2, max_group=2)

    def run(drv, box, slabs, session):
        if session:
            drv.stream_begin()
        accs = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in slabs}
        for x in slabs:
            drv.execute_slab(box, x, accs[x], nsplit=2)
        stats = [s for x in slabs for s in drv.finalize(x)]
        if session:
            drv.stream_end()
        return accs, stats

    slabs = list(range(box9.cpd))
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, **kw) as drv:
        plain, st0 = run(drv, box9, slabs, False)
        cached, st1 = run(drv, box9, slabs, True)
    assert sum(s["slabs_reused"] for s in st0) == 0
    up1 = sum(s["slabs_uploaded"] for s in st1)
    assert sum(s["slabs_reused"] for s in st1) > 0 and up1 < sum(s["slabs_uploaded"] for s in st0)
    assert up1 >= box9.cpd
    assert all(s["upload_s"] >= 0 and s["download_s"] > 0 for s in st1)
    for x in slabs:
        assert np.array_equal(plain[x], cached[x]), x

    # A second session after slab 4 changed on the host: its new positions
    # are used (no stale copy), and the result equals a fresh driver's.
    mod = synth.Box(box9.cpd, [synth.Slab(s.pos.copy(), s.count) for s in box9.slabs])
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, **kw) as drv:
        first, _ = run(drv, mod, [3, 4], True)
        mod.slab(4).pos *= np.float32(0.5)          # stays inside each cell
        second, _ = run(drv, mod, [3, 4], True)
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, **kw) as drv:
        fresh, _ = run(drv, mod, [3, 4], False)
    for x in (3, 4):
        assert not np.array_equal(first[x], second[x]), x
        assert np.array_equal(second[x], fresh[x]), x


def test_pinned_host_slabs_match_pageable(box9):
    """Host-mode slabs and acceleration slabs in pinned memory (read and
    written by the DMA engines directly) give bitwise the results of pageable
    ones (copied through the pinned staging buffers)."""
    eps, b2, _ = params(box9)
    kw = dict(path="batched", buffers_per_device=2, max_buffer_bytes=2 << 30, sinks_per_lane=2)
    keep = []
    pinned_slabs = []
    for s in box9.slabs:
        pa = nfm.PinnedArray(s.pos.shape, np.float32)
        pa.array[:] = s.pos
        keep.append(pa)
        pinned_slabs.append(synth.Slab(pa.array, s.count))
    pbox = synth.Box(box9.cpd, pinned_slabs)
    assert all(ps.pos.ctypes.data == pa.ptr for ps, pa in zip(pbox.slabs, keep))
    slabs = [0, 4, 8]
    out = {}
    for name, box in (("pageable", box9), ("pinned", pbox)):
        with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, **kw) as drv:
            drv.stream_begin()
            accs = {}
            for x in slabs:
                if name == "pinned":
                    pa = nfm.PinnedArray((box.slab(x).n, 4), np.float32)
                    pa.array[:] = np.nan
                    keep.append(pa)
                    accs[x] = pa.array
                else:
                    accs[x] = np.full((box.slab(x).n, 4), np.nan, np.float32)
                drv.execute_slab(box, x, accs[x], nsplit=2)
            for x in slabs:
                drv.finalize(x)
            drv.stream_end()
            out[name] = {x: accs[x].copy() for x in slabs}
    for x in slabs:
        assert np.isfinite(out["pinned"][x]).all()
        assert np.array_equal(out["pageable"][x], out["pinned"][x]), x


@pytest.mark.parametrize("mem", [nfm.NF_MEM_HOST, nfm.NF_MEM_DEVICE])
@pytest.mark.parametrize("path", ["plain", "batched"])
def test_aos_positions_match_posxyz(box9, path, mem):
    """Position slabs given as posstruct (x,y,z per particle, no TransposePos)
    give bitwise the results of PosXYZ slabs."""
    eps, b2, _ = params(box9)
    slabs = [0, 4, 8]
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, path=path, buffers_per_device=2,
                             max_buffer_bytes=2 << 30, sinks_per_lane=2) as drv:
        soa = run_sics(drv, box9, slabs, mem)
        aos = run_sics(drv, box9, slabs, mem, pos_layout=nfm.NF_POS_AOS)
    for x in slabs:
        assert np.isfinite(soa[x]).all()
        assert np.array_equal(soa[x], aos[x]), x


@pytest.mark.parametrize("mem", [nfm.NF_MEM_HOST, nfm.NF_MEM_DEVICE])
@pytest.mark.parametrize("path", ["plain", "batched"])
def test_fused_kick_matches_consumer(box9, path, mem):
    """With the kick fused into the co-add, the acceleration slab holds
    bitwise what RescaleAndCoAddAcceleration (the oracle's restatement) makes
    of the unfused near field and the far field, and NearAccOut, if given,
    bitwise the NearAcc record StoreForces == 2 writes."""
    eps, b2, e = params(box9)
    slabs = [1, 4, 7]
    rng = np.random.default_rng(5)
    rescale = -3.0 * 0.3 / (8.0 * np.pi * box9.N)
    inv_eps3 = 1.0 / e ** 3
    dens = float(synth.fof_b2(box9.mean_spacing()))
    kick_args = {x: (rng.standard_normal((box9.slab(x).n, 3)).astype(np.float32) * 1e3, rescale, inv_eps3, dens)
                 for x in slabs}
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, path=path, buffers_per_device=2,
                             max_buffer_bytes=2 << 30, sinks_per_lane=2) as drv:
        near = run_sics(drv, box9, slabs, mem)
        fused = run_sics(drv, box9, slabs, mem, kick_args=kick_args)
        fused2, record = run_sics(drv, box9, slabs, mem, kick_args=kick_args, near_out=True)
    for x in slabs:
        expect = po.rescale_coadd(near[x], kick_args[x][0], rescale, inv_eps3, dens)
        assert np.array_equal(fused[x], expect), x
        # with NearAccOut the co-add also writes the StoreForces == 2 record
        assert np.array_equal(fused2[x], expect), x
        assert np.array_equal(record[x], po.near_acc_record(near[x], inv_eps3)), x
        assert np.abs(fused[x][:, 3]).max() < 10 * dens      # w minus its self term is small


def test_throughput_counters(box9):
    """nf_get_throughput: the GPU-active wall time and the totals of every
    completed task since setup (the reference's GPU throughput timer and
    AggregateStats totals)."""
    eps, b2, _ = params(box9)
    slabs = list(range(box9.cpd))
    with nfm.NearFieldDriver(box9.cpd, box9.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=2,
                             max_buffer_bytes=2 << 30, sinks_per_lane=2) as drv:
        assert drv.throughput()["tasks"] == 0
        import time
        t0 = time.perf_counter()
        accs = {x: np.full((box9.slab(x).n, 4), np.nan, np.float32) for x in slabs}
        for x in slabs:
            drv.execute_slab(box9, x, accs[x], nsplit=2)
        stats = [st for x in slabs for st in drv.finalize(x)]
        wall = time.perf_counter() - t0
        tp = drv.throughput()
    assert tp["tasks"] == len(stats) == 2 * box9.cpd
    assert 1 <= tp["groups"] <= tp["tasks"]
    assert tp["direct_total"] == sum(st["direct_total"] for st in stats)
    assert tp["padded_direct_total"] == sum(st["padded_direct_total"] for st in stats)
    assert 0 < tp["gpu_active_s"] <= wall and tp["gdips"] > 0
    assert abs(tp["bytes_to_device"] - sum(st["bytes_to_device"] for st in stats)) <= tp["tasks"]


def test_oversized_sic_fails_with_capacity_error():
    """A SIC whose working set exceeds the buffer fails with NF_ERR_CAPACITY
    (the reference asserts), the driver keeps working, and the same slab as
    one SIC per row succeeds with bitwise the accelerations of a whole-slab
    SIC on a driver with room."""
    box = synth.uniform_box(33 ** 3 * 300, 33, seed=31)  # ~107 MB working set per whole-slab SIC
    eps, b2, _ = params(box)
    x = 16
    small = 64 * (1 << 20) + (2 << 20)  # nf_setup keeps 64 MiB of headroom; ~28 MB per one-row SIC
    with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                             max_buffer_bytes=small, sinks_per_lane=2) as drv:
        acc = np.full((box.slab(x).n, 4), np.nan, np.float32)
        with pytest.raises(nfm.NFError) as ei:
            drv.execute_slab(box, x, acc, blocking=True)
        assert "device bytes" in str(ei.value) and "(-3)" in str(ei.value)
        drv.slab_tasks.pop(x % box.cpd, None)
        rows = np.full_like(acc, np.nan)
        drv.execute_slab(box, x, rows, nsplit=box.cpd)
        drv.finalize(x)
    with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                             max_buffer_bytes=2 << 30, sinks_per_lane=2) as drv:
        whole = np.full_like(acc, np.nan)
        drv.execute_slab(box, x, whole, blocking=True)
        drv.finalize(x)
    assert np.isfinite(whole).all() and np.array_equal(rows, whole)

