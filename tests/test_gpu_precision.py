"""Precision study on the 64^3, CPD 17 uniform workload (all 262,144 sinks).

Variants: the plain path (FP32 as the reference accumulates it: BLOCK_ACCUM
tile sums, FP32 offsets, FP32 co-add), the batched path (TwoSum-compensated
FP32 per pencil, FP64 cell offsets, FP64 co-add), the reference CPU directs
(restated, oracle/cpu_directs.c), all against the FP64 direct-summation
oracle on the same FP32 cell-relative inputs.
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
    refs, softs, cpu = [], [], []
    for x in range(box.cpd):
        sl, ix, ce = po.slab_sinks(box, x)
        r, s = po.direct_f64(box, R, "single_spline", eps, b2, sl, ix, ce, with_soft=True)
        refs.append(r)
        softs.append(s)
        cpu.append(po.execute_slab_cpu(box, R, "single_spline", eps, x)[0])
    ref, soft, cpu = np.concatenate(refs), np.concatenate(softs), np.concatenate(cpu)
    keep = not_soft_dominated(ref, soft)
    report = {"workload": box.name, "sinks": int(ref.shape[0]),
              "softening_dominated_sinks": int((~keep).sum())}
    for name, acc in (("plain", res["plain"]), ("batched", res["batched"]), ("cpu_directs", cpu)):
        rel = rel_err(acc, ref)
        report[name] = {"all": summarize(rel, tail=True),
                        "outside_softening_dominated": summarize(rel, keep, tail=True)}
        if name != "cpu_directs":
            report[name]["kernel_interactions_per_s"] = rate[name]
            report[name]["vs_cpu_directs"] = summarize(rel_err(acc, cpu))
    print(json.dumps(report))
    out = os.environ.get("NF_PRECISION_REPORT")
    if out:
        with open(out, "w") as f:
            json.dump(report, f, indent=1)
    b, p = report["batched"]["outside_softening_dominated"], report["plain"]["outside_softening_dominated"]
    c = report["cpu_directs"]["outside_softening_dominated"]
    # rms within 1e-6 of the FP64 oracle; the max is set by a few sinks whose
    # net force nearly cancels (FP32 tile sums of ~6,700 terms): bounded here,
    # and better than both the reference-structured FP32 path and the
    # reference CPU directs on the same inputs.
    assert b["rms"] <= 1e-6 and b["max"] <= 5e-5, b
    assert b["rms"] < p["rms"] and b["rms"] < c["rms"]
    assert b["max"] < p["max"] and b["max"] < c["max"]
    assert b["above_1e5"] <= 1e-4 * b["n"]
    assert report["batched"]["vs_cpu_directs"]["rms"] <= 1e-6
    # the FOF density w against FP64 (self term included)
    np.testing.assert_allclose(res["batched"][:, 3], ref[:, 3], rtol=2e-6, atol=1e-6 * float(b2))
