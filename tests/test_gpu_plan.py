"""The device SIC planner (nf_plan_sic_device, nf_submit_slab) against the
host planner nf_plan_sic (run with -m gpu).

The host planner's arrays are pinned by the oracle tests (test_planner.py);
the device planner must reproduce every array and total bit for bit, and a
SIC planned on the device must give bitwise the accelerations of the same
SIC planned on the host.
"""
import numpy as np
import pytest

import nearfield as nfm
import synth
from nfhelpers import (R, not_soft_dominated, oracle_slab, params, rel_err, run_sics, small_uniform,
                       summarize)

pytestmark = pytest.mark.gpu


def ragged_box():
    """Empty slabs 0-1, a heavy cell (1,500 particles) at the centre."""
    cpd = 7
    rng = np.random.default_rng(3)
    pts = rng.random((cpd ** 3 * 4, 3)) - 0.5
    pts = np.concatenate([pts[pts[:, 0] > -0.2], rng.normal(0.0, 0.01, (1500, 3))])
    return synth.box_from_global(pts, cpd)


def compare_plans(box, slab, j_low, j_high, nfradius=R):
    eps, b2, _ = params(box)
    host = nfm.SetInteractionCollection(box, slab, j_low, j_high, nfradius, eps, b2, "single_spline")
    dev = nfm.SetInteractionCollection(box, slab, j_low, j_high, nfradius, eps, b2, "single_spline",
                                       planner="device")
    a, b = host.arrays(), dev.arrays()
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), (slab, j_low, j_high, k)
        else:
            assert a[k] == b[k], (slab, j_low, j_high, k, a[k], b[k])
    for f in ("SlabId", "j_low", "j_high", "k_low", "k_high", "Nk", "j_width", "cpd", "nfradius", "nfwidth",
              "eps", "b2", "softening", "Nslab"):
        va, vb = getattr(host.desc, f), getattr(dev.desc, f)
        if f == "Nslab":
            va, vb = list(va), list(vb)
        assert va == vb, f
    return a


@pytest.mark.parametrize("rows", [(0, 9), (0, 3), (3, 7), (7, 9), (8, 9)])
def test_device_plan_equals_host_plan(rows):
    box = small_uniform(cpd=9, ppc=20, seed=11)
    for x in (0, 4, 8):
        compare_plans(box, x, *rows)


def test_device_plan_ragged_wrapping_and_radius():
    """Empty slabs (no sink blocks at all), a heavy cell, cpd = 2R+1 (every
    pencil wraps), radius 1, and a clustered box."""
    rb = ragged_box()
    for x in (0, 1, 3):
        a = compare_plans(rb, x, 0, rb.cpd)
        if x == 0:
            assert a["NSinkBlocks"] == 0
    small = synth.uniform_box(5 ** 3 * 12, 5, seed=9)
    for x in (0, 2):
        compare_plans(small, x, 0, 5)
        compare_plans(small, x, 1, 4)
    b7 = small_uniform(cpd=7, ppc=10, seed=4)
    for x in (0, 6):
        compare_plans(b7, x, 0, 7, nfradius=1)
        compare_plans(b7, x, 2, 5, nfradius=1)
    cl = synth.clustered_box(40, 11, seed=2, n_halos=60)
    for x in (0, 5, 10):
        compare_plans(cl, x, 0, 11)


@pytest.mark.parametrize("mem", [nfm.NF_MEM_HOST, nfm.NF_MEM_DEVICE])
@pytest.mark.parametrize("path", ["plain", "batched"])
def test_submit_slab_matches_submit(path, mem):
    """nf_submit_slab (planned on the device: plan arrays, source-block
    parents, co-add chunks and work list all built there) gives bitwise the
    accelerations of nf_plan_sic + nf_submit, with the fused kick and posstruct
    positions too."""
    box = small_uniform(cpd=9, ppc=20, seed=11)
    eps, b2, e = params(box)
    slabs = [0, 4, 8]
    rng = np.random.default_rng(5)
    rescale = -3.0 * 0.3 / (8.0 * np.pi * box.N)
    kick_args = {x: (rng.standard_normal((box.slab(x).n, 3)).astype(np.float32), rescale, 1.0 / e ** 3,
                     float(b2)) for x in slabs}
    with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path=path, buffers_per_device=2,
                             max_buffer_bytes=2 << 30, sinks_per_lane=2) as drv:
        host = run_sics(drv, box, slabs, mem)
        dev, st = run_sics(drv, box, slabs, mem, planner="device", with_stats=True)
        hk, hn = run_sics(drv, box, slabs, mem, pos_layout=nfm.NF_POS_AOS, kick_args=kick_args, near_out=True)
        dk, dn = run_sics(drv, box, slabs, mem, pos_layout=nfm.NF_POS_AOS, kick_args=kick_args, planner="device",
                          near_out=True)
    for i, x in enumerate(slabs):
        assert np.isfinite(dev[x]).all()
        assert np.array_equal(host[x], dev[x]), x
        assert np.array_equal(hk[x], dk[x]), x
        assert np.isfinite(dn[x]).all() and np.array_equal(hn[x], dn[x]), x
        sic = nfm.SetInteractionCollection(box, x, 0, box.cpd, R, eps, b2, "single_spline")
        assert st[i]["device_planned"] == 1 and st[i]["plan_s"] > 0
        assert st[i]["direct_total"] == sic.desc.DirectTotal
        assert st[i]["padded_direct_total"] == sic.desc.PaddedDirectTotal
        assert 1 <= st[i]["group_size"] <= 8


def test_device_planned_groups_match_single_launches():
    """Device-planned SICs sharing one persistent launch (their work lists
    merged on the device through one histogram) give bitwise the results of
    one launch per SIC and of host-planned groups."""
    box = small_uniform(cpd=9, ppc=20, seed=11)
    eps, b2, _ = params(box)
    out = {}
    for planner, mg in (("device", 1), ("device", 8), ("host", 8)):
        with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                                 max_buffer_bytes=2 << 30, sinks_per_lane=2, max_group=mg) as drv:
            accs = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in range(box.cpd)}
            for x in range(box.cpd):
                drv.execute_slab(box, x, accs[x], nsplit=2, planner=planner)
            sizes = []
            for x in range(box.cpd):
                sizes += [s["group_size"] for s in drv.finalize(x)]
            out[(planner, mg)] = (accs, sizes)
    assert max(out[("device", 1)][1]) == 1 and max(out[("device", 8)][1]) > 1
    for x in range(box.cpd):
        assert np.isfinite(out[("device", 8)][0][x]).all()
        assert np.array_equal(out[("device", 1)][0][x], out[("device", 8)][0][x]), x
        assert np.array_equal(out[("host", 8)][0][x], out[("device", 8)][0][x]), x


def test_submit_slab_splits_ragged_and_autotuned():
    """Split slabs, empty slabs and a heavy cell through nf_submit_slab:
    bitwise the host-planned result at a fixed tile; with the autotuner (which
    times every candidate tile on device-built work lists) within the
    single-precision bar of the FP64 oracle."""
    box = ragged_box()
    eps, b2, _ = params(box)
    slabs = [0, 1, 3, 5]
    out = {}
    for planner, spl in (("host", 2), ("device", 2), ("device", 0)):
        with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=2,
                                 max_buffer_bytes=2 << 30, sinks_per_lane=spl) as drv:
            accs = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in slabs}
            for x in slabs:
                drv.execute_slab(box, x, accs[x], nsplit=3, planner=planner)
            for x in slabs:
                drv.finalize(x)
            out[(planner, spl)] = accs
    for x in slabs:
        assert np.array_equal(out[("host", 2)][x], out[("device", 2)][x]), x
    assert out[("device", 0)][0].shape == (0, 4)
    for x in (1, 3, 5):
        ref, soft = oracle_slab(box, x, "single_spline", eps, b2, with_soft=True)
        st = summarize(rel_err(out[("device", 0)][x], ref), not_soft_dominated(ref, soft))
        assert st["rms"] < 3e-6, (x, st)


@pytest.mark.parametrize("mem", ["host", "device"])
def test_submit_slab_in_stream_session(mem):
    """Inside a stream session the cellinfo slabs, like host position slabs,
    go up once; results are bitwise those without a session.  With
    device-resident positions only the cellinfo crosses PCIe."""
    box = small_uniform(cpd=9, ppc=20, seed=11)
    eps, b2, _ = params(box)
    slabs = list(range(box.cpd))
    dev = mem == "device"
    pos = [nfm.DeviceArray(host=s.pos) for s in box.slabs] if dev else None
    res = {}
    for session in (False, True):  # one SIC per launch: a group would share its slabs without a session
        with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=2,
                                 max_buffer_bytes=2 << 30, sinks_per_lane=2, max_group=1) as drv:
            if session:
                drv.stream_begin()
            if dev:
                accd = {x: nfm.DeviceArray(nbytes=16 * box.slab(x).n) for x in slabs}
                tasks = [drv.submit_slab(box, x, acc_ptr=accd[x].ptr, mem=nfm.NF_MEM_DEVICE,
                                         pos_ptrs=[pos[(x - R + c) % box.cpd].ptr for c in range(2 * R + 1)])
                         for x in slabs]
            else:
                accs = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in slabs}
                tasks = [drv.submit_slab(box, x, acc_ptr=accs[x].ctypes.data) for x in slabs]
            st = []
            for t in tasks:
                t.wait()
                st.append(t.stats())
                t.release()
            if session:
                drv.stream_end()
            if dev:
                accs = {x: accd[x].to_host(np.float32, (box.slab(x).n, 4)) for x in slabs}
                for a in accd.values():
                    a.free()
            res[session] = (accs, st)
    for x in slabs:
        assert np.array_equal(res[False][0][x], res[True][0][x]), x
    cells = box.cpd * box.cpd * 40
    to_dev = [sum(s["bytes_to_device"] for s in res[k][1]) for k in (False, True)]
    # Without a session every SIC uploads its 2R+1 cellinfo (and, from the
    # host, position) slabs.  In one, a slab goes up again only when the
    # cache's least recently used free slot was taken by another slab in the
    # meantime (with one SIC per launch and two workers, 13 cellinfo uploads
    # for 9 slabs were measured).
    assert to_dev[0] - to_dev[1] >= (2 * R - 1) * box.cpd * cells
    if dev:
        for d in pos:
            d.free()


def test_submit_slab_rejects_cells_outside_slab():
    """A cellinfo whose cell runs past the end of its slab fails the task with
    NF_ERR_INVALID before any gather reads the slab."""
    box = small_uniform(cpd=7, ppc=10, seed=4)
    eps, b2, _ = params(box)
    bad = synth.Box(box.cpd, [synth.Slab(s.pos, s.count) for s in box.slabs])
    ci = box.slab(3).cellinfo.copy()
    ci["count"][ci.size // 2] += box.slab(3).n
    bad.slab(3).cellinfo = ci
    acc = np.full((box.slab(2).n, 4), np.nan, np.float32)
    with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                             max_buffer_bytes=1 << 30, sinks_per_lane=2) as drv:
        t = drv.submit_slab(bad, 2, acc_ptr=acc.ctypes.data)
        with pytest.raises(nfm.NFError, match="outside their slab"):
            t.wait()
        t.release()
        # the driver still works afterwards
        ok = np.full((box.slab(2).n, 4), np.nan, np.float32)
        drv.submit_slab(box, 2, acc_ptr=ok.ctypes.data, blocking=True).release()
    assert np.isnan(acc).all() and np.isfinite(ok).all()


def test_shared_streams_match_per_worker_streams():
    """gpu_streams = 1 (one compute stream and one planning stream shared by
    the workers: the layout the benchmark and the integration shim use) gives
    bitwise the accelerations of one stream pair per worker, host- and
    device-planned, with the workers' groups interleaved on the shared
    streams."""
    box = small_uniform(cpd=9, ppc=20, seed=11)
    eps, b2, _ = params(box)
    out = {}
    for gs in (0, 1):
        with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=2,
                                 max_buffer_bytes=2 << 30, sinks_per_lane=2, gpu_streams=gs) as drv:
            for planner in ("host", "device"):
                accs = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in range(box.cpd)}
                for x in range(box.cpd):
                    drv.execute_slab(box, x, accs[x], nsplit=2, planner=planner)
                for x in range(box.cpd):
                    drv.finalize(x)
                out[(gs, planner)] = accs
    ref = out[(0, "host")]
    for key, accs in out.items():
        for x in range(box.cpd):
            assert np.isfinite(accs[x]).all(), key
            assert np.array_equal(accs[x], ref[x]), (key, x)



def test_device_planned_graph_replays_match_eager():
    """Device-planned launch groups captured in HIP graphs (the gathers,
    persistent kernel and co-adds that read the planner's counts from device
    memory), re-captured, updated in place and replayed over three passes with
    eager work between them, give bitwise the eager host-planned results on
    every pass, including a slab split into several SICs."""
    box = small_uniform(cpd=9, ppc=20, seed=17)
    eps, b2, _ = params(box)
    slabs = [1, 4, 7]
    with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=2,
                             max_buffer_bytes=2 << 30, sinks_per_lane=2, waves_per_wg=4) as drv:
        eager = run_sics(drv, box, slabs, nfm.NF_MEM_HOST)
    with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                             max_buffer_bytes=2 << 30, sinks_per_lane=2, waves_per_wg=4, max_group=4,
                             use_graphs=1) as drv:
        for rep in range(3):
            got, st = run_sics(drv, box, slabs, nfm.NF_MEM_HOST, planner="device", with_stats=True)
            assert all(s["graph"] == 1 and s["device_planned"] == 1 for s in st), rep
            for x in slabs:
                assert np.array_equal(got[x], eager[x]), (rep, x)
            accs = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in slabs}
            for x in slabs:  # eager work between the passes: a slab in three device-planned SICs
                drv.execute_slab(box, x, accs[x], nsplit=3, planner="device")
            for x in slabs:
                drv.finalize(x)
                assert np.array_equal(accs[x], eager[x]), (rep, x)
