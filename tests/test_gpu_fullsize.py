"""Full-size parity on the benchmark's own realisations, against the FP64
direct-summation oracle on a seeded 1% subsample of sinks:

  * uniform N = 512^3, CPD 135, whole box resident in HBM (plain and batched);
  * clustered N = 256^3, CPD 69, cells from 0 to ~5,000 particles (plain and
    batched);
  * the clustered CPD 255 slab stream: host-resident slabs in a stream
    session, launch groups captured in HIP graphs, autotuned tiles; its first
    8 sink slabs (a slab's particles do not depend on which other slabs are
    realised, so they are the benchmark's).

For each: rms and max relative error over the sample, outside
softening-dominated sinks (softened pairs, r < e, carry more than half of
|a|) and over sinks of condition number <= 100 (the sum of the pair terms'
magnitudes over |a|), plus the batched path against the reference CPU directs
(restated, oracle/cpu_directs.c) on the sample's sinks of two slabs, and the
FP32 pair-term floor (oracle or_floor_f32: the batched kernel's pair terms,
positions rounded into the pencil frame as it rounds them, summed exactly)
on the same sinks.  The bars: rms <= 1e-6; max <= 1e-5 outside
softening-dominated sinks, or within 1.5x of the floor's max where FP32 pair
terms alone miss 1e-5 (they err by ~1e-7 x the condition number, more for a
sink dominated by one close pair across a cell boundary, whose separation
is formed from positions rounded into the pencil frame); on sinks of
condition number <= 100, max <= 1e-5.  NF_FULLSIZE_REPORT=<file> writes the
numbers as JSON.
"""
import json
import os

import numpy as np
import pytest

import nearfield as nfm
import pyoracle as po
import synth
from nfhelpers import R, not_soft_dominated, rel_err, summarize

pytestmark = pytest.mark.gpu
FRAC = 0.01
REPORT = {}


def _log(*a):
    print(*a, flush=True)


def _evaluate(name, box, accs, sample, cpu_slabs, eps, b2):
    """Errors of each path's accelerations {path: {slab: (n, 4)}} on the
    sampled sinks against the FP64 oracle (and against the CPU directs on the
    sample's sinks of cpu_slabs)."""
    sl, ix, ce = sample
    ref, soft, mag = po.direct_f64(box, R, "single_spline", eps, b2, sl, ix, ce, with_soft=True, with_abs=True)
    keep = not_soft_dominated(ref, soft)
    cond = mag / np.maximum(np.linalg.norm(ref[:, :3], axis=1), 1e-300)
    well = keep & (cond <= 100.0)
    rep = {"sinks": int(sl.size), "softening_dominated": int((~keep).sum()),
           "condition_number_above_100": int((cond > 100).sum())}
    floor = np.concatenate([po.floor_f32(box, R, eps, sl, ix, ce), ref[:, 3:]], axis=1)
    rel_fl = rel_err(floor, ref)
    rep["fp32_floor"] = {"outside_softening_dominated": summarize(rel_fl, keep, tail=True),
                         "condition_number_le_100": summarize(rel_fl, well, tail=True),
                         "max_error_over_condition_number": float((rel_fl[keep] / cond[keep]).max())}
    cpu = {x: po.execute_slab_cpu(box, R, "single_spline", eps, x)[0] for x in cpu_slabs}
    on_cpu = np.isin(sl, cpu_slabs)
    cpu_acc = np.zeros((int(on_cpu.sum()), 4), np.float32)
    for x in cpu_slabs:
        m = sl[on_cpu] == x
        cpu_acc[m] = cpu[x][ix[on_cpu][m]]
    rel_cpu = rel_err(cpu_acc, ref[on_cpu])
    rep["cpu_directs"] = {"vs_fp64": summarize(rel_cpu, keep[on_cpu], tail=True)}
    for path, acc in accs.items():
        got = np.zeros((sl.size, 4), np.float32)
        for x in np.unique(sl):
            m = sl == x
            got[m] = acc[int(x)][ix[m]]
        assert np.isfinite(got).all(), (name, path)
        rel = rel_err(got, ref)
        rep[path] = {"all": summarize(rel, tail=True), "outside_softening_dominated": summarize(rel, keep, tail=True),
                     "condition_number_le_100": summarize(rel, well, tail=True),
                     "max_error_over_condition_number": float((rel[keep] / cond[keep]).max()),
                     "vs_cpu_directs": summarize(rel_err(got[on_cpu], cpu_acc), keep[on_cpu], tail=True),
                     "vs_fp32_floor": summarize(rel_err(got, floor), keep, tail=True)}
        # the FOF density against FP64 (self term included): the batched path
        # sums it in FP32 per 64-source tile and in FP64 across tiles; the
        # plain path in FP32 throughout, as the reference does, which drifts
        # by ~1e-5 for sinks with thousands of pairs within b in dense halos
        rel_w = np.abs(got[:, 3].astype(np.float64) - ref[:, 3]) / np.maximum(np.abs(ref[:, 3]), float(b2))
        rep[path]["density_w"] = {"rms": float(np.sqrt(np.mean(rel_w ** 2))), "max": float(rel_w.max())}
        assert rel_w.max() <= (1e-5 if path == "batched" else 1e-4), (name, path, rep[path]["density_w"])
    REPORT[name] = rep
    _log(name, json.dumps(rep))
    out = os.environ.get("NF_FULLSIZE_REPORT")
    if out:
        with open(out, "w") as f:
            json.dump(REPORT, f, indent=1)
    return rep


def _check(rep):
    b, p, f = rep["batched"], rep.get("plain"), rep["fp32_floor"]
    bo, fo = b["outside_softening_dominated"], f["outside_softening_dominated"]
    assert bo["rms"] <= 1e-6, b
    assert bo["max"] <= max(1e-5, 1.5 * fo["max"]), (b, f)
    assert b["condition_number_le_100"]["max"] <= 1e-5, b
    assert b["max_error_over_condition_number"] <= max(1e-6, 1.5 * f["max_error_over_condition_number"]), (b, f)
    assert b["vs_cpu_directs"]["rms"] <= 1.5e-6, b
    if p is not None:
        assert p["outside_softening_dominated"]["rms"] <= 3e-6, p
        assert p["condition_number_le_100"]["max"] <= 1e-4, p
        assert b["outside_softening_dominated"]["rms"] < p["outside_softening_dominated"]["rms"]


def _resident_box(box, paths=("plain", "batched")):
    """Every slab's SIC of a box resident in HBM, per path -> {path: {slab: acc}}."""
    ell = box.mean_spacing()
    b2 = synth.fof_b2(ell)
    pos = [nfm.DeviceArray(host=s.pos) for s in box.slabs]
    acc = [nfm.DeviceArray(nbytes=16 * s.n) for s in box.slabs]
    out = {}
    for path in paths:
        _log(f"{box.name}: {path} path, {box.cpd} SICs")
        with nfm.NearFieldDriver(box.cpd, ell / 40.0, b2, nfradius=R, path=path, buffers_per_device=4,
                                 max_buffer_bytes=6 << 30, gpu_streams=1) as drv:
            tasks = []
            for x in range(box.cpd):
                ptrs = [pos[(x - R + c) % box.cpd].ptr for c in range(2 * R + 1)]
                tasks.append(drv.submit(drv.plan(box, x, pos_ptrs=ptrs, acc_ptr=acc[x].ptr, mem=nfm.NF_MEM_DEVICE)))
            for t in tasks:
                t.wait()
                t.release()
        out[path] = {x: acc[x].to_host(np.float32, (box.slab(x).n, 4)) for x in range(box.cpd)}
    for d in pos + acc:
        d.free()
    return out


def test_uniform512_fullsize_vs_fp64():
    _log("building uniform N=512^3 CPD 135 (seed 20260512)")
    box = synth.uniform_box(512 ** 3, 135, seed=20260512)
    eps, _ = synth.eps_param("single_spline", box.mean_spacing() / 40.0)
    b2 = synth.fof_b2(box.mean_spacing())
    accs = _resident_box(box)
    rep = _evaluate("uniform512", box, accs, po.sample_sinks(box, FRAC, seed=512), [0, 67], eps, b2)
    _check(rep)


def test_clustered256_fullsize_vs_fp64():
    _log("building clustered N=256^3 CPD 69 (seed 20260256)")
    box = synth.clustered_box(256, 69, seed=20260256)
    assert box.counts().max() > 2000
    eps, _ = synth.eps_param("single_spline", box.mean_spacing() / 40.0)
    b2 = synth.fof_b2(box.mean_spacing())
    accs = _resident_box(box)
    busiest = int(np.argmax([s.count.max() for s in box.slabs]))
    rep = _evaluate("clustered256", box, accs, po.sample_sinks(box, FRAC, seed=256), [busiest, 0], eps, b2)
    _check(rep)


def test_slabstream255_fullsize_vs_fp64():
    sinks = list(range(8))
    _log("building the clustered CPD 255 stream (N=998^3 recipe, seed 20260255), sink slabs 0-7")
    box = synth.clustered_slabs(998, 255, 20260255, range(-R, len(sinks) + R))
    ell = box.mean_spacing()
    eps, _ = synth.eps_param("single_spline", ell / 40.0)
    b2 = synth.fof_b2(ell)
    acc = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in sinks}
    with nfm.NearFieldDriver(box.cpd, ell / 40.0, b2, nfradius=R, path="batched", buffers_per_device=4,
                             max_buffer_bytes=6 << 30, gpu_streams=1, use_graphs=1) as drv:
        drv.stream_begin()
        tasks = [drv.submit(drv.plan(box, x, acc_ptr=acc[x].ctypes.data, mem=nfm.NF_MEM_HOST)) for x in sinks]
        stats = []
        for t in tasks:
            t.wait()
            stats.append(t.stats())
            t.release()
        drv.stream_end()
    assert all(s["graph"] == 1 or s["autotune_trial"] == 1 for s in stats)
    assert sum(s["slabs_reused"] for s in stats) > 0
    rep = _evaluate("slabstream255", box, {"batched": acc}, po.sample_sinks(box, FRAC, seed=255, slabs=sinks),
                    [0, 5], eps, b2)
    _check(rep)


def test_slabstream255_wrap_corners_vs_fp64():
    """Every sink of the cells within two of both periodic wraps (rows and
    columns 0-2 and 252-254) of sink slabs 206-210 of the CPD 255 stream,
    slabs in pageable host memory (copied into pinned staging by several
    threads).  Slab 208 holds 3,900,459 particles: 12 x that many bytes left
    the last particle's z uncopied when the threaded copy's chunks were
    rounded down, and the sinks around that particle, in these corner cells
    of slabs 206-210, erred by up to 4e-2."""
    sinks = list(range(206, 211))
    _log("building the clustered CPD 255 stream (N=998^3 recipe, seed 20260255), sink slabs 206-210")
    box = synth.clustered_slabs(998, 255, 20260255, range(sinks[0] - R, sinks[-1] + R + 1))
    assert box.slab(208).n == 3900459
    cpd = box.cpd
    ell = box.mean_spacing()
    eps, _ = synth.eps_param("single_spline", ell / 40.0)
    b2 = synth.fof_b2(ell)
    acc = {x: np.full((box.slab(x).n, 4), np.nan, np.float32) for x in sinks}
    with nfm.NearFieldDriver(cpd, ell / 40.0, b2, nfradius=R, path="batched", buffers_per_device=4,
                             max_buffer_bytes=6 << 30, gpu_streams=1, use_graphs=1) as drv:
        drv.stream_begin()
        tasks = [drv.submit(drv.plan(box, x, acc_ptr=acc[x].ctypes.data, mem=nfm.NF_MEM_HOST)) for x in sinks]
        for t in tasks:
            t.wait()
            t.release()
        drv.stream_end()
    rows = [0, 1, 2, cpd - 3, cpd - 2, cpd - 1]
    sl, ix, ce = [], [], []
    for x in sinks:
        s = box.slab(x)
        for j in rows:
            for k in rows:
                c = j * cpd + k
                n = int(s.count[c])
                sl.append(np.full(n, x, np.int32))
                ix.append(np.arange(s.start[c], s.start[c] + n, dtype=np.int64))
                ce.append(np.full(n, c, np.int32))
    sample = (np.concatenate(sl), np.concatenate(ix), np.concatenate(ce))
    assert sample[0].size > 5000
    rep = _evaluate("slabstream255_corners", box, {"batched": acc}, sample, [208], eps, b2)
    _check(rep)
