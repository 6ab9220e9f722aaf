"""Near-field direct path benchmark (one MI355X).

Workload (default): N = 512^3 equal-mass particles uniform in the periodic unit
box, CPD = 135, NearFieldRadius = 2, single-spline softening with Plummer-
equivalent length 1/40 of the mean spacing, FOF density radius 0.25 spacing.
The whole box (PosXYZ slabs and float4 acceleration slabs) is resident in HBM;
a step is one near-field evaluation of the whole box: every slab's
SetInteractionCollection submitted through the C ABI (device-resident mode),
gather + direct kernel + co-add on the GPU.  The SIC plans are built once from
the cellinfo before timing (host planning time is reported separately).

Prints one JSON line (rank 0).  Unit of work: unpadded pair interactions
including self pairs (DirectTotal); 20 FLOP per interaction for % of peak.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "abacus-public_amd")]

import nearfield as nfm  # noqa: E402
import synth  # noqa: E402

FLOP_PER_INTERACTION = 20
PEAK_FP32_TFLOPS = 157.3          # MI355X published FP32 vector peak (2 FLOP/FMA, packed)
METRIC = "near-field pair interactions/s, % FP32 VALU peak; batched vs plain at fixed error"

WORKLOADS = {
    # name: (builder, cpd, description)
    "uniform512": (lambda: synth.uniform_box(512 ** 3, 135, seed=20260512), 135,
                   "uniform N=512^3 CPD=135 R=2 single spline eps=l/40 b=0.25l, whole box resident"),
    "uniform64": (lambda: synth.uniform_box(64 ** 3, 17, seed=20260064), 17,
                  "uniform N=64^3 CPD=17 R=2 single spline eps=l/40 b=0.25l"),
    "clustered256": (lambda: synth.clustered_box(256, 69, seed=20260256), 69,
                     "clustered N=256^3 CPD=69 (70% GRF lattice, 30% in 2000 Plummer halos) R=2 single spline"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="uniform512", choices=sorted(WORKLOADS))
    ap.add_argument("--paths", default="plain,batched")
    ap.add_argument("--bpd", type=int, default=4, help="worker streams (DirectBPD)")
    ap.add_argument("--sinks-per-lane", type=int, default=0, help="0: the autotuner picks per occupancy histogram")
    ap.add_argument("--max-group", type=int, default=8, help="SICs per persistent launch (batched path)")
    ap.add_argument("--gpu-streams", type=int, default=1,
                    help="1: workers share one HIP stream (kernels back to back), 0: one per worker")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only copy this many MiB device-to-device twice (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by y-rows of each slab, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    build, cpd, desc = WORKLOADS[args.workload]
    t0 = time.time()
    box = build()
    log(f"[bench] built {box.name}: N={box.N} in {time.time() - t0:.1f}s")
    ell = box.mean_spacing()
    soft_len = ell / 40.0
    b2 = synth.fof_b2(ell)
    R = 2

    # Resident slabs in HBM.
    pos_d = [nfm.DeviceArray(host=s.pos) for s in box.slabs]
    acc_d = [nfm.DeviceArray(nbytes=16 * s.n) for s in box.slabs]
    for a in acc_d:
        a.zero()
    nfm.device_synchronize()
    log(f"[bench] uploaded {sum(s.n for s in box.slabs) * 12 / 1e9:.2f} GB of positions")

    results = {}
    paths = [p for p in args.paths.split(",") if p]
    for path in paths:
        with nfm.NearFieldDriver(cpd, soft_len, b2, nfradius=R, path=path, buffers_per_device=args.bpd,
                                 max_buffer_bytes=6 << 30, sinks_per_lane=args.sinks_per_lane,
                                 gpu_streams=args.gpu_streams, max_group=args.max_group) as drv:
            tp = time.time()
            sics = []
            for x in range(cpd):
                ptrs = [pos_d[(x - R + c) % cpd].ptr for c in range(2 * R + 1)]
                sics.append(drv.plan(box, x, pos_ptrs=ptrs, acc_ptr=acc_d[x].ptr, mem=nfm.NF_MEM_DEVICE))
            plan_s = time.time() - tp
            direct_total = sum(int(s.desc.DirectTotal) for s in sics)
            padded_total = sum(int(s.desc.PaddedDirectTotal) for s in sics)
            log(f"[bench] {path}: planned {cpd} SICs in {plan_s:.2f}s, DirectTotal={direct_total:.4e}")

            def step():
                tasks = [drv.submit(s) for s in sics]
                st = []
                for t in tasks:
                    t.wait()
                    st.append(t.stats())
                    t.release()
                return st

            for i in range(args.warmup):
                step()
                log(f"[bench] {path}: warmup {i + 1}/{args.warmup} done")
            nfm.device_synchronize()
            t1 = time.perf_counter()
            kstats = []
            for i in range(args.steps):
                kstats.extend(step())
                log(f"[bench] {path}: step {i + 1}/{args.steps} done")
            nfm.device_synchronize()
            wall = time.perf_counter() - t1
            kern = np.array([s["direct_kernel_s"] for s in kstats])      # each task's share of its launch
            tot = np.array([s["kernel_s"] for s in kstats])
            launches = float(sum(1.0 / s["group_size"] for s in kstats))
            per_slab_di = np.array([int(s.desc.DirectTotal) for s in sics] * args.steps, dtype=np.float64)
            results[path] = dict(
                wall_s=wall, ms_per_step=1e3 * wall / args.steps,
                interactions_per_s=direct_total * args.steps / wall,
                direct_kernel_s_per_step=float(kern.sum()) / args.steps,
                device_s_per_step=float(tot.sum()) / args.steps,
                kernel_interactions_per_s=float(per_slab_di.sum() / kern.sum()),
                kernel_avg_ms=float(kern.sum() / launches * 1e3),
                launches_per_step=launches / args.steps,
                interactions_per_launch=float(per_slab_di.sum() / launches),
                direct_total=direct_total, padded_direct_total=padded_total, plan_s=plan_s,
            )
            results[path]["acc_sample"] = [acc_d[x].to_host(np.float32, (box.slab(x).n, 4))[:2000]
                                           for x in (0, cpd // 2)]
            if path == "batched":
                results[path]["autotune"] = [{k: e[k] for k in ("sinks_per_lane", "waves_per_cu", "uses")}
                                             for e in drv.autotune_table()]
            log(f"[bench] {path}: {results[path]['interactions_per_s']:.4e} interactions/s wall, "
                f"kernel {results[path]['kernel_interactions_per_s']:.4e}/s")

    best = max(results, key=lambda p: results[p]["interactions_per_s"])
    r = results[best]
    achieved_tflops = FLOP_PER_INTERACTION * r["kernel_interactions_per_s"] / 1e12
    line = {
        "metric": METRIC,
        "value": r["interactions_per_s"],
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": desc, "N": box.N, "cpd": cpd, "nfradius": R, "path": best,
                   "softening": "single_spline", "interactions_per_step": r["direct_total"]},
        "pct_fp32_peak": 100.0 * FLOP_PER_INTERACTION * r["interactions_per_s"] / (PEAK_FP32_TFLOPS * 1e12),
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "k_batched" if best == "batched" else "k_plain",
                     "kernel_avg_ms": r["kernel_avg_ms"],
                     "flop_per_interaction": FLOP_PER_INTERACTION},
        "paths": {p: {k: v for k, v in d.items() if k != "acc_sample"} for p, d in results.items()},
    }
    if "plain" in results and "batched" in results:
        line["batched_over_plain"] = results["batched"]["interactions_per_s"] / results["plain"]["interactions_per_s"]

    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(box, R, soft_len, args.cpu_seconds, results)
    print(json.dumps(line), flush=True)


def calibrate_copy(mib):
    """A device-to-device copy of known size: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE give the counters' scale in bytes."""
    n = mib << 20
    a = nfm.DeviceArray(host=np.arange(n // 4, dtype=np.float32))
    b = nfm.DeviceArray(nbytes=n)
    for _ in range(2):
        nfm.check(nfm.lib().nf_device_copy(b.ptr, a.ptr, n, 2), "nf_device_copy")
    nfm.device_synchronize()
    print(json.dumps({"calibration_copy_bytes": n, "copies": 2}), flush=True)


def cpu_baseline(box, R, soft_len, budget_s, results):
    """The CPU-fallback restatement (oracle/cpu_directs.c, the reference's
    ExecuteSlabCPU loop + DirectCPUKernels.cc kernel, -Ofast AVX-512, OpenMP
    on all host cores), timed on whole slabs until the budget is used."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    eps, _ = synth.eps_param("single_spline", soft_len)
    cpd = box.cpd
    order = [0, cpd // 2] + [x for x in range(cpd) if x not in (0, cpd // 2)]
    done = 0
    di = 0
    t0 = time.perf_counter()
    agree = {}
    for x in order:
        acc, n = po.execute_slab_cpu(box, R, "single_spline", eps, x)
        di += n
        done += 1
        if x in (0, cpd // 2):
            k = 0 if x == 0 else 1
            for p, d in results.items():
                g = d["acc_sample"][k].astype(np.float64)
                c = acc[:2000].astype(np.float64)
                # the GPU carries the reference's output units (no e^-3), as does the CPU path
                rel = np.linalg.norm(g[:, :3] - c[:, :3], axis=1) / np.linalg.norm(c[:, :3], axis=1)
                agree.setdefault(p, []).append(rel)
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    out = {"value": di / el, "unit": "interactions/s", "cores": po.num_threads(), "kind": "port",
           "sample": f"{done} whole slabs of the same box ({di:.3e} interactions, {el:.1f}s)"}
    out["gpu_vs_cpu_rel_diff"] = {p: {"rms": float(np.sqrt(np.mean(np.concatenate(v) ** 2))),
                                      "max": float(np.concatenate(v).max())} for p, v in agree.items()}
    return out


if __name__ == "__main__":
    main()
