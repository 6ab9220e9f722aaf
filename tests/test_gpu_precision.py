"""Precision study on the 64^3, CPD 17 uniform workload (all 262,144 sinks).

Variants: the plain path (FP32 as the reference accumulates it: BLOCK_ACCUM
tile sums, FP32 offsets, FP32 co-add), the batched path (FP32 sums of
16-source batches and 64-source tiles, the tile sums accumulated in FP64, FP64
cell offsets, FP64 co-add; short tails keep TwoSum-compensated FP32 pairs),
the reference CPU directs (restated, oracle/cpu_directs.c), all against the
FP64 direct-summation oracle on the same FP32 cell-relative inputs.  Beside
them the FP32 pair-term floor (oracle or_floor_f32): the batched kernel's
pair terms formed in FP32 and summed exactly, the best any summation of
those terms can do.
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


def test_precision_study_uniform64():
    box = synth.uniform_box(64 ** 3, 17, seed=20260064)
    ell = box.mean_spacing()
    eps, _ = synth.eps_param("single_spline", ell / 40.0)
    b2 = synth.fof_b2(ell)
    res, rate = {}, {}
    for path in ("plain", "batched"):
        with nfm.NearFieldDriver(box.cpd, ell / 40.0, b2, path=path, buffers_per_device=2,
                                 max_buffer_bytes=4 << 30, sinks_per_lane=2, max_group=32) as drv:
            accs = {x: np.zeros((box.slab(x).n, 4), np.float32) for x in range(box.cpd)}
            for x in range(box.cpd):
                drv.execute_slab(box, x, accs[x])
            st = [s for x in range(box.cpd) for s in drv.finalize(x)]
        res[path] = np.concatenate([accs[x] for x in range(box.cpd)])
        rate[path] = sum(s["direct_total"] for s in st) / max(sum(s["direct_kernel_s"] for s in st), 1e-30)
    refs, softs, mags, cpu = [], [], [], []
    for x in range(box.cpd):
        sl, ix, ce = po.slab_sinks(box, x)
        r, s, m = po.direct_f64(box, R, "single_spline", eps, b2, sl, ix, ce, with_soft=True, with_abs=True)
        refs.append(r)
        softs.append(s)
        mags.append(m)
        cpu.append(po.execute_slab_cpu(box, R, "single_spline", eps, x)[0])
    ref, soft, mag, cpu = (np.concatenate(v) for v in (refs, softs, mags, cpu))
    parts = [po.slab_sinks(box, x) for x in range(box.cpd)]
    floor = po.floor_f32(box, R, eps, *(np.concatenate([p[i] for p in parts]) for i in range(3)))
    keep = not_soft_dominated(ref, soft)
    # Condition number of each sink's force sum: sum of term magnitudes / |a|.
    # FP32 terms carry relative errors of ~1e-7, so a sink's attainable
    # relative error grows with it.
    cond = mag / np.maximum(np.linalg.norm(ref[:, :3], axis=1), 1e-300)
    well = keep & (cond <= 100.0)
    report = {"workload": box.name, "sinks": int(ref.shape[0]),
              "softening_dominated_sinks": int((~keep).sum()),
              "condition_number": {"median": float(np.median(cond)), "p99": float(np.percentile(cond, 99)),
                                   "max": float(cond.max()), "sinks_above_100": int((cond > 100).sum())}}
    fl = np.concatenate([floor, ref[:, 3:]], axis=1)
    for name, acc in (("plain", res["plain"]), ("batched", res["batched"]), ("cpu_directs", cpu),
                      ("fp32_floor", fl)):
        rel = rel_err(acc, ref)
        report[name] = {"all": summarize(rel, tail=True),
                        "outside_softening_dominated": summarize(rel, keep, tail=True),
                        "condition_number_le_100": summarize(rel, well, tail=True),
                        "max_error_over_condition_number": float((rel[keep] / cond[keep]).max())}
        if name in res:
            report[name]["kernel_interactions_per_s"] = rate[name]
            report[name]["vs_cpu_directs"] = summarize(rel_err(acc, cpu), keep, tail=True)
            # distance from the exactly summed FP32 pair terms, per sink
            report[name]["vs_fp32_floor"] = summarize(rel_err(acc, fl), keep, tail=True)
    print(json.dumps(report))
    out = os.environ.get("NF_PRECISION_REPORT")
    if out:
        with open(out, "w") as f:
            json.dump(report, f, indent=1)
    b, p = report["batched"]["outside_softening_dominated"], report["plain"]["outside_softening_dominated"]
    c = report["cpu_directs"]["outside_softening_dominated"]
    # rms within 1e-6 of the FP64 oracle (measured 2.2e-7).  The max is set by
    # a few sinks whose net force nearly cancels (condition number 100-1000):
    # with FP32 pair terms their relative error is ~1e-7 x the condition
    # number, which the batch and tile sums do not change.  Bounded here, and
    # better than both the reference-structured FP32 path and the reference
    # CPU directs on the same inputs; on sinks with condition number <= 100
    # the max is within 1e-5.
    f = report["fp32_floor"]["outside_softening_dominated"]
    # The floor itself misses max <= 1e-5 on this box (tests/test_oracle.py);
    # the batched path stays within 1.5x of it.  Its distance from the exactly
    # summed FP32 terms grows with the condition number as well (the terms
    # differ: v_rsq_f32 is not the host's correctly rounded 1/sqrt): at most
    # 1e-6 per unit of condition number, and 1e-5 on sinks of condition
    # number <= 100.
    assert f["max"] > 1e-5 and b["max"] <= 1.5 * f["max"], (b, f)
    vf = rel_err(res["batched"], fl)
    report["batched"]["vs_fp32_floor"]["max_over_condition_number"] = float((vf[keep] / cond[keep]).max())
    report["batched"]["vs_fp32_floor"]["condition_number_le_100"] = summarize(vf, well)
    print(json.dumps({"batched_vs_fp32_floor": report["batched"]["vs_fp32_floor"]}))
    assert report["batched"]["vs_fp32_floor"]["max_over_condition_number"] <= 1e-6, report["batched"]["vs_fp32_floor"]
    assert report["batched"]["vs_fp32_floor"]["condition_number_le_100"]["max"] <= 1e-5, report["batched"]["vs_fp32_floor"]
    assert b["rms"] <= 4e-7 and b["max"] <= 3e-5, b
    assert b["rms"] < p["rms"] and b["rms"] < c["rms"]
    assert b["max"] < p["max"] and b["max"] < c["max"]
    assert b["above_1e5"] <= 2e-5 * b["n"]
    bw = report["batched"]["condition_number_le_100"]
    assert bw["max"] <= 1e-5 and bw["n"] >= 0.99 * b["n"], bw
    assert report["batched"]["max_error_over_condition_number"] <= 1e-6
    assert report["batched"]["vs_cpu_directs"]["rms"] <= 1e-6
    # the max against the CPU directs is theirs: they err by up to ~4e-5 against FP64 here
    assert report["batched"]["vs_cpu_directs"]["max"] <= 1.2 * report["cpu_directs"]["outside_softening_dominated"]["max"]
    # the FOF density w against FP64 (self term included)
    np.testing.assert_allclose(res["batched"][:, 3], ref[:, 3], rtol=2e-6, atol=1e-6 * float(b2))
