"""Near-field direct path benchmark (one MI355X).

Workload (default): N = 512^3 equal-mass particles uniform in the periodic unit
box, CPD = 135, NearFieldRadius = 2, single-spline softening with Plummer-
equivalent length 1/40 of the mean spacing, FOF density radius 0.25 spacing.
The whole box (PosXYZ slabs and float4 acceleration slabs) is resident in HBM;
a step is one near-field evaluation of the whole box: every slab's
SetInteractionCollection submitted through the C ABI (device-resident mode),
gather + direct kernel + co-add on the GPU.  Every step plans each slab's SIC
from the cellinfo inside the timed region, as ExecuteSlabGPU does: by default
it submits the slabs through nf_submit_slab, which plans them on the GPU with
no host round trip; with --planner host it plans each SIC with nf_plan_sic on
the host and submits it at once, so planning overlaps the GPU work already
queued.  5 timed steps by default; the line reports their total and their
median.

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


# Host-resident slab streams: (builder(slabs), cpd, n1d, description).  The
# clustered recipe at CPD 255 and ~1e9 particles, realised in the slabs a run
# of --stream-slabs sink slabs needs (plus the R-slab halo on each side).
STREAM_WORKLOADS = {
    "slabstream255": (lambda slabs: synth.clustered_slabs(998, 255, 20260255, slabs), 255, 998,
                      "clustered N=998^3 CPD=255 (70% GRF lattice, 30% in 100,937 Plummer halos), "
                      "host-resident slabs streamed through a 5-slab window"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="uniform512", choices=sorted(WORKLOADS) + sorted(STREAM_WORKLOADS))
    ap.add_argument("--stream-slabs", type=int, default=40, help="slab-stream workloads: sink slabs per pass")
    ap.add_argument("--pinned", action="store_true",
                    help="slab streams: host slabs in pinned memory (direct DMA) instead of pageable")
    ap.add_argument("--aos", action="store_true",
                    help="slab streams: hand the gathers posstruct (x,y,z per particle) slabs, skipping TransposePos")
    ap.add_argument("--fuse-kick", action="store_true",
                    help="slab streams: fuse RescaleAndCoAddAcceleration (far field in, final acceleration out)")
    ap.add_argument("--graphs", type=int, default=None,
                    help="capture each group's kernels in a HIP graph (default: on for slab streams)")
    ap.add_argument("--paths", default="plain,batched")
    ap.add_argument("--planner", default="device", choices=["device", "host", "host-once"],
                    help="device: each step submits the slabs with nf_submit_slab, which plans them on the GPU; "
                         "host: each step plans every slab's SIC with nf_plan_sic and submits it at once, so "
                         "planning the next SIC overlaps the GPU work of the ones before (ExecuteSlabGPU's "
                         "order); host-once: the SICs are planned once before timing and resubmitted "
                         "every step (a comparison only: real steps re-plan, particles move)")
    ap.add_argument("--bpd", type=int, default=4, help="worker streams (DirectBPD)")
    ap.add_argument("--sinks-per-lane", type=int, default=0, help="0: the autotuner picks per occupancy histogram")
    ap.add_argument("--waves-per-wg", type=int, default=0,
                    help="batched workgroup size in waves, 1, 2 or 4 (0: the autotuner picks)")
    ap.add_argument("--max-group", type=int, default=8, help="SICs per persistent launch (batched path)")
    ap.add_argument("--group-interactions", type=float, default=0,
                    help="pair interactions after which a launch group takes no more SICs (0: the library's default)")
    ap.add_argument("--gpu-streams", type=int, default=1,
                    help="1: workers share one HIP stream (kernels back to back), 0: one per worker")
    ap.add_argument("--no-configs", action="store_true",
                    help="whole-box workloads: skip the other workloads' lines (clustered 256^3 plain and "
                         "batched, the CPD 255 slab stream) that the default run adds under 'configs'")
    ap.add_argument("--config-steps", type=int, default=5, help="timed steps of each 'configs' workload")
    ap.add_argument("--config-stream-slabs", type=int, default=255,
                    help="sink slabs per pass of the slab stream under 'configs' (255: every slab of the "
                         "CPD 255 box, ~1e9 particles, ~28 GB of host memory)")
    ap.add_argument("--error-sample", type=float, default=0.01,
                    help="fraction of sinks checked against the FP64 oracle in the CPU leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample budget")
    ap.add_argument("--no-census", action="store_true",
                    help="whole-box workloads: skip the untimed census pass (share of pair interactions per "
                         "pair form) after the batched path's timed steps")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU leg (the CPU baseline and the FP64-oracle error checks)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="only copy this many MiB device-to-device twice (counter calibration) and exit")
    ap.add_argument("--probe", action="store_true",
                    help="only run the pair-math probe (nf_probe_pair_rate: the batched tile loop alone, per pair "
                         "form, sinks per lane and occupancy) and print its table")
    ap.add_argument("--pmc-summary", metavar="DIR",
                    help="only print the HBM traffic table from rocprofv3 --pmc databases under DIR and exit")
    return ap.parse_args(argv)


def main():
    args = parse_args()
    if args.pmc_summary:
        print_pmc_summary(args.pmc_summary, args.workload, args.calibrate_copy_mib or 1024)
        return
    if args.calibrate_copy_mib:
        calibrate_copy(args.calibrate_copy_mib)
        return
    if args.probe:
        pair_probe()
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard each slab's y-rows across GPUs, with a 2R-row source halo)",
              file=sys.stderr)
        sys.exit(2)
    R = 2
    if args.workload in STREAM_WORKLOADS:
        box = build_stream(args.workload, args.stream_slabs)
        res = run_stream(args, box, args.workload, args.stream_slabs, args.steps, args.warmup)
        line = stream_line(args, res)
        if not args.no_cpu and not args.fuse_kick:
            sample = {x: res["acc"][x][:2000] for x in (res["sinks"][0], res["sinks"][len(res["sinks"]) // 2])}
            line["cpu_baseline"] = cpu_baseline(box, R, res["soft_len"], args.cpu_seconds,
                                                {"batched": {"acc_sample": sample}}, slabs=res["sinks"])
            line["accuracy_vs_fp64"] = error_check(box, R, res["soft_len"], {"batched": res["acc"]},
                                                   args.error_sample, 255, slabs=res["sinks"])
        print(json.dumps(line), flush=True)
        return

    # The slab stream of the 'configs' lines is realised first: its generator
    # forks worker processes, which must happen before this process touches
    # the GPU.
    stream_box = None if args.no_configs else build_stream("slabstream255", args.config_stream_slabs)
    build, cpd, desc = WORKLOADS[args.workload]
    t0 = time.time()
    box = build()
    log(f"[bench] built {box.name}: N={box.N} in {time.time() - t0:.1f}s")
    paths = [p for p in args.paths.split(",") if p]
    res = run_resident(args, box, cpd, paths, args.steps, args.warmup, keep_acc=not args.no_cpu)
    line = resident_line(args, box, cpd, desc, res)
    extra = {}
    if not args.no_configs:
        cbuild, ccpd, cdesc = WORKLOADS["clustered256"]
        cbox = cbuild()
        log(f"[bench] built {cbox.name}: N={cbox.N}")
        cres = run_resident(args, cbox, ccpd, ["plain", "batched"], args.config_steps, 1,
                            keep_acc=not args.no_cpu)
        cl = {"workload": cdesc, "N": cbox.N, "cpd": ccpd, "steps": args.config_steps, "warmup": 1,
              "interactions_per_step": cres["plain"]["direct_total"],
              "max_cell_occupancy": int(cbox.counts().max()),
              "paths": {p: summary_of(d) for p, d in cres.items()},
              "batched_over_plain": cres["batched"]["interactions_per_s"] / cres["plain"]["interactions_per_s"],
              "batched_over_plain_kernel": cres["batched"]["kernel_interactions_per_s"]
              / cres["plain"]["kernel_interactions_per_s"]}
        sres = run_stream(args, stream_box, "slabstream255", args.config_stream_slabs, args.config_steps, 1)
        sl = stream_line(args, sres)
        st = {"workload": sl["config"]["workload"], "steps": args.config_steps, "warmup": 1, **sl["config"],
              **sl["end_to_end"], "autotune": sl["autotune"]}
        line["configs"] = {"clustered256": cl, "slabstream255": st}
        extra = {"clustered256": (cbox, cres), "slabstream255": (stream_box, sres)}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(box, R, box.mean_spacing() / 40.0, args.cpu_seconds, res)
        line["accuracy_vs_fp64"] = error_check(box, R, box.mean_spacing() / 40.0,
                                               {p: d["acc"] for p, d in res.items()}, args.error_sample, 512)
        if extra:
            cbox, cres = extra["clustered256"]
            line["configs"]["clustered256"]["accuracy_vs_fp64"] = error_check(
                cbox, R, cbox.mean_spacing() / 40.0, {p: d["acc"] for p, d in cres.items()}, args.error_sample, 256)
            sbox, sres = extra["slabstream255"]
            line["configs"]["slabstream255"]["accuracy_vs_fp64"] = error_check(
                sbox, R, sres["soft_len"], {"batched": sres["acc"]}, args.error_sample, 255, slabs=sres["sinks"])
    print(json.dumps(line), flush=True)


def run_resident(args, box, cpd, paths, steps, warmup, keep_acc=False):
    """Whole box resident in HBM; per path, `warmup` untimed and `steps` timed
    steps, each one near-field evaluation of every slab (planning included,
    per --planner).  Returns {path: results}, with each path's accelerations
    on the host under "acc" when keep_acc."""
    R = 2
    ell = box.mean_spacing()
    soft_len = ell / 40.0
    b2 = synth.fof_b2(ell)
    pos_d = [nfm.DeviceArray(host=s.pos) for s in box.slabs]
    acc_d = [nfm.DeviceArray(nbytes=16 * s.n) for s in box.slabs]
    for a in acc_d:
        a.zero()
    nfm.device_synchronize()
    log(f"[bench] uploaded {sum(s.n for s in box.slabs) * 12 / 1e9:.2f} GB of positions")
    results = {}
    for path in paths:
        with nfm.NearFieldDriver(cpd, soft_len, b2, nfradius=R, path=path, buffers_per_device=args.bpd,
                                 max_buffer_bytes=6 << 30, sinks_per_lane=args.sinks_per_lane, waves_per_wg=args.waves_per_wg,
                                 gpu_streams=args.gpu_streams, max_group=args.max_group,
                                 group_interactions=int(args.group_interactions),
                                 use_graphs=args.graphs or 0) as drv:
            ptrs_of = {x: [pos_d[(x - R + c) % cpd].ptr for c in range(2 * R + 1)] for x in range(cpd)}
            tp = time.time()
            sics = [drv.plan(box, x, pos_ptrs=ptrs_of[x], acc_ptr=acc_d[x].ptr, mem=nfm.NF_MEM_DEVICE)
                    for x in range(cpd)]
            plan_s = time.time() - tp
            direct_total = sum(int(s.desc.DirectTotal) for s in sics)
            padded_total = sum(int(s.desc.PaddedDirectTotal) for s in sics)
            log(f"[bench] {path}: planned {cpd} SICs in {plan_s:.2f}s, DirectTotal={direct_total:.4e}")
            sic_sizes = dict(padded_sinks=float(np.mean([s.desc.PaddedSinkTotal for s in sics])),
                             padded_sources=float(np.mean([s.desc.PaddedSourceTotal for s in sics])))
            per_slab_di = [int(s.desc.DirectTotal) for s in sics]
            if args.planner != "host-once":
                sics = None  # each step plans its own

            def step():
                # One stream session per step: the slabs and their cellinfo do
                # not change within a step, so each cellinfo slab goes up once
                # (device planner) instead of once per launch group.
                drv.stream_begin()
                if args.planner == "device":
                    tasks = [drv.submit_slab(box, x, pos_ptrs=ptrs_of[x], acc_ptr=acc_d[x].ptr,
                                             mem=nfm.NF_MEM_DEVICE) for x in range(cpd)]
                elif args.planner == "host":
                    tasks = [drv.submit(drv.plan(box, x, pos_ptrs=ptrs_of[x], acc_ptr=acc_d[x].ptr,
                                                 mem=nfm.NF_MEM_DEVICE)) for x in range(cpd)]
                else:
                    tasks = [drv.submit(s) for s in sics]
                st = []
                for t in tasks:
                    t.wait()
                    st.append(t.stats())
                    t.release()
                drv.stream_end()
                return st

            for i in range(warmup):
                step()
                log(f"[bench] {path}: warmup {i + 1}/{warmup} done")
            nfm.device_synchronize()
            t1 = time.perf_counter()
            kstats, step_walls = [], []
            for i in range(steps):
                ts = time.perf_counter()
                kstats.extend(step())  # returns once every task of the step has completed
                step_walls.append(time.perf_counter() - ts)
                log(f"[bench] {path}: step {i + 1}/{steps} done")
            nfm.device_synchronize()
            wall = time.perf_counter() - t1
            kern = np.array([s["direct_kernel_s"] for s in kstats])      # each task's share of its launch
            tot = np.array([s["kernel_s"] for s in kstats])
            launches = float(sum(1.0 / s["group_size"] for s in kstats))
            di_all = np.array(per_slab_di * steps, dtype=np.float64)
            # Per direct-kernel instantiation (tile shape): launches, time and work.
            shapes = {}
            for st_, di_ in zip(kstats, di_all):
                e = shapes.setdefault(kernel_name(path, st_["sinks_per_lane"], st_["waves_per_eu"]),
                                      dict(launches=0.0, kernel_s=0.0, interactions=0.0))
                e["launches"] += 1.0 / st_["group_size"]
                e["kernel_s"] += st_["direct_kernel_s"]
                e["interactions"] += di_
            for e in shapes.values():
                e["avg_ms"] = 1e3 * e["kernel_s"] / e["launches"]
                e["interactions_per_launch"] = e["interactions"] / e["launches"]
            results[path] = dict(
                wall_s=wall, ms_per_step=1e3 * wall / steps,
                median_step_ms=1e3 * float(np.median(step_walls)),
                interactions_per_s=direct_total * steps / wall,
                direct_kernel_s_per_step=float(kern.sum()) / steps,
                device_s_per_step=float(tot.sum()) / steps,
                kernel_interactions_per_s=float(di_all.sum() / kern.sum()),
                kernel_avg_ms=float(kern.sum() / launches * 1e3),
                launches_per_step=launches / steps,
                interactions_per_launch=float(di_all.sum() / launches),
                direct_total=direct_total, padded_direct_total=padded_total, plan_s_before_timing=plan_s,
                planner=args.planner,
                device_plan_s_per_step=float(sum(s["plan_s"] for s in kstats)) / steps,
                kernels=shapes, n_sics=cpd, sic_sizes=sic_sizes,
            )
            if keep_acc:
                results[path]["acc"] = {x: acc_d[x].to_host(np.float32, (box.slab(x).n, 4)) for x in range(cpd)}
                results[path]["acc_sample"] = {x: results[path]["acc"][x][:2000] for x in (0, cpd // 2)}
            if path == "batched":
                results[path]["autotune"] = [{k: e[k] for k in ("occupancy_quarters", "size_class", "sinks_per_lane",
                                                                "waves_per_wg", "uses")}
                                             for e in drv.autotune_table()]
            log(f"[bench] {path}: {results[path]['interactions_per_s']:.4e} interactions/s wall, "
                f"kernel {results[path]['kernel_interactions_per_s']:.4e}/s")
        if path == "batched" and not args.no_census:
            # One more, untimed pass with the census build of the kernel: the
            # share of pair interactions in each pair form.
            with nfm.NearFieldDriver(cpd, soft_len, b2, nfradius=R, path=path, buffers_per_device=args.bpd,
                                     max_buffer_bytes=6 << 30, sinks_per_lane=args.sinks_per_lane,
                                     waves_per_wg=args.waves_per_wg, gpu_streams=args.gpu_streams,
                                     max_group=args.max_group, group_interactions=int(args.group_interactions),
                                     census=1) as cdrv:
                tasks = [cdrv.submit(cdrv.plan(box, x, pos_ptrs=ptrs_of[x], acc_ptr=acc_d[x].ptr,
                                               mem=nfm.NF_MEM_DEVICE)) for x in range(cpd)]
                cst = []
                for t in tasks:
                    t.wait()
                    cst.append(t.stats())
                    t.release()
            results[path]["pair_forms"] = pair_forms(cst)
            log(f"[bench] {path}: pair forms {results[path]['pair_forms']}")
    for d in pos_d + acc_d:
        d.free()
    return results


def pair_forms(stats):
    """The batched launches' census over tasks (each task reports its launch
    group's counts): the fraction of pair interactions computed in each pair
    form and in short pencil tails, or None for the plain path."""
    forms, tail = np.zeros(3), 0.0
    for st in stats:
        n = max(1, st["group_size"])
        forms += np.asarray(st["group_form_interactions"], np.float64) / n
        tail += st["group_tail_interactions"] / n
    tot = forms.sum()
    if tot <= 0:
        return None
    return {"unsoftened": forms[0] / tot, "density": forms[1] / tot, "softened": forms[2] / tot,
            "short_tails": tail / tot}


def summary_of(d):
    """A path's results without the accelerations."""
    return {k: v for k, v in d.items() if k not in ("acc", "acc_sample")}


def resident_line(args, box, cpd, desc, results):
    R = 2
    best = max(results, key=lambda p: results[p]["interactions_per_s"])
    r = results[best]
    # The dominant kernel: the direct-kernel instantiation with the most time.
    kname, kd = max(r["kernels"].items(), key=lambda kv: kv[1]["kernel_s"])
    achieved_tflops = FLOP_PER_INTERACTION * kd["interactions_per_launch"] / (kd["avg_ms"] * 1e-3) / 1e12
    traffic, traffic_src = hbm_traffic(args.workload, kname)
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
                   "softening": "single_spline", "interactions_per_step": r["direct_total"],
                   "planner": args.planner, "graphs": args.graphs or 0},
        "pct_fp32_peak": 100.0 * FLOP_PER_INTERACTION * r["interactions_per_s"] / (PEAK_FP32_TFLOPS * 1e12),
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": kname, "kernel_avg_ms": kd["avg_ms"],
                     "interactions_per_launch": kd["interactions_per_launch"],
                     "flop_per_interaction": FLOP_PER_INTERACTION,
                     "algorithmic_bytes_per_launch": algorithmic_bytes(r, kd),
                     "traffic_source": traffic_src},
        "paths": {p: summary_of(d) for p, d in results.items()},
    }
    if "plain" in results and "batched" in results:
        line["batched_over_plain"] = results["batched"]["interactions_per_s"] / results["plain"]["interactions_per_s"]
    return line


def kernel_name(path, sinks_per_lane, waves_per_eu=0, softening=1):
    """The direct kernel's name as rocprofv3 prints it (template arguments:
    pair form, sinks per lane, occupancy target in waves per SIMD, census
    build: timed runs use the normal one)."""
    if path == "plain":
        return f"k_plain<{softening}>"
    return f"k_batched<{softening}, {sinks_per_lane}, {waves_per_eu}, false>"


def algorithmic_bytes(r, kd):
    """Bytes a batched launch must move at least: every padded sink slot's
    position read once from the slab (3 floats), every padded source (float4)
    and the two float4 bounding boxes per 64-source block read once, and a
    float64x4 partial written per padded sink; for a launch's mean number of
    SICs of mean size."""
    sics_per_launch = kd["interactions_per_launch"] / (r["direct_total"] / r["n_sics"])
    sz = r["sic_sizes"]
    per_sic = 12 * sz["padded_sinks"] + 16 * sz["padded_sources"] * (1 + 2 / 64) + 32 * sz["padded_sinks"]
    return per_sic * sics_per_launch


def hbm_traffic(workload, kname):
    """HBM bytes per launch of kernel kname, from the committed summary of the
    two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this workload; None
    when there is none."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 6 and cells[0] == kname:
                return float(cells[5]), os.path.relpath(path, ROOT)
    return None, None


def pmc_summary(prof_dir, workload):
    """Markdown table of HBM bytes per dispatch from the rocprofv3 --pmc
    databases under prof_dir: fetch_copy/write_copy (a device copy of known
    size, run with --calibrate-copy-mib) give each counter's scale;
    fetch_<workload>/write_<workload> give the kernels' counts."""
    import glob
    import sqlite3

    def per_kernel(run, counter):
        dbs = glob.glob(os.path.join(prof_dir, run, "*.db"))
        if not dbs:
            raise SystemExit(f"no database under {os.path.join(prof_dir, run)}")
        out = {}
        for db in dbs:
            con = sqlite3.connect(db)
            for name, n, avg in con.execute(
                    "select kernel_name, count(*), avg(value) from counters_collection "
                    "where counter_name = ? group by kernel_name", (counter,)):
                short = name.replace("(anonymous namespace)", "").replace("void ", "").split("(")[0]
                short = short.split("::")[-1].strip()
                out[short] = (n, avg)
            con.close()
        return out

    cal = {}
    for counter, run in (("FETCH_SIZE", "fetch_copy"), ("WRITE_SIZE", "write_copy")):
        k = per_kernel(run, counter)
        name = max(k, key=lambda s: k[s][0] * k[s][1])
        cal[counter] = (name, k[name][1])
    fetch = per_kernel(f"fetch_{workload}", "FETCH_SIZE")
    write = per_kernel(f"write_{workload}", "WRITE_SIZE")
    return cal, fetch, write


def sq_counters(prof_dir, run):
    """{kernel: {counter: (dispatches, sum over dispatches)}} of one --pmc run."""
    import glob
    import sqlite3
    out = {}
    for db in glob.glob(os.path.join(prof_dir, run, "*.db")):
        con = sqlite3.connect(db)
        for name, counter, n, tot in con.execute(
                "select kernel_name, counter_name, count(distinct dispatch_id), sum(value) "
                "from counters_collection group by kernel_name, counter_name"):
            short = name.replace("(anonymous namespace)", "").replace("void ", "").split("(")[0]
            out.setdefault(short.split("::")[-1].strip(), {})[counter] = (n, tot)
        con.close()
    return out


def print_valu_summary(prof_dir, workload):
    """VALU instruction mix and busy fraction of the direct kernels, from the
    valu_<workload> and busy_<workload> --pmc runs of a fixed-tile bench run
    (no autotuner trials), per pair interaction of the run."""
    runs = [r for r in (f"valu_{workload}", f"busy_{workload}") if os.path.isdir(os.path.join(prof_dir, r))]
    if not runs:
        return
    with open(os.path.join(prof_dir, f"valu_{workload}.log")) as f:
        line = json.loads([x for x in f if x.startswith("{")][-1])
    interactions = line["config"]["interactions_per_step"] * (line["steps"] + line["warmup"])
    print(f"\nVALU counters of the direct kernel over the whole run ({line['steps']} + {line['warmup']} steps, "
          f"{interactions:.4e} pair interactions, tile fixed so every launch is the same kernel):\n")
    print("| kernel | counter | dispatches | total | per interaction |")
    print("|---|---|---|---|---|")
    for run in runs:
        for k, cs in sorted(sq_counters(prof_dir, run).items()):
            if not k.startswith("k_batched") and not k.startswith("k_plain"):
                continue
            for c, (n, tot) in sorted(cs.items()):
                per = ""
                if c.startswith("SQ_INSTS_VALU") and "FLOPS" not in c:
                    per = f"{64 * tot / interactions:.3f} lane-ops"
                elif "FLOPS" in c or c in ("SQ_INSTS_SALU", "SQ_INSTS_SMEM"):
                    per = f"{tot / interactions:.4f}" if "FLOPS" in c else f"{64 * tot / interactions:.4f} (x64)"
                elif c in ("VALUBusy", "VALUUtilization"):
                    per = f"{tot / max(n, 1):.1f} % (mean over dispatches)"
                print(f"| {k} | {c} | {n} | {tot:.6g} | {per} |")


def print_pmc_summary(prof_dir, workload, copy_mib):
    print_valu_summary(prof_dir, workload)
    if not os.path.isdir(os.path.join(prof_dir, "fetch_copy")):
        return
    cal, fetch, write = pmc_summary(prof_dir, workload)
    copy_kib = copy_mib * 1024.0
    f_scale = copy_kib / cal["FETCH_SIZE"][1]   # true KiB per counted KiB
    w_scale = copy_kib / cal["WRITE_SIZE"][1]
    print(f"Counter calibration: a {copy_mib} MiB device-to-device copy ({cal['FETCH_SIZE'][0]}) reads "
          f"{cal['FETCH_SIZE'][1]:.0f} KiB of FETCH_SIZE and {cal['WRITE_SIZE'][1]:.0f} KiB of WRITE_SIZE per "
          f"dispatch: scale {f_scale:.3f} (fetch), {w_scale:.3f} (write).\n")
    print("| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | "
          "HBM read+write MB/dispatch (scaled) | HBM bytes/dispatch (scaled) |")
    print("|---|---|---|---|---|---|")
    for k in sorted(set(fetch) | set(write), key=lambda s: -fetch.get(s, (0, 0))[1]):
        nf, fa = fetch.get(k, (0, 0.0))
        nw, wa = write.get(k, (0, 0.0))
        b = 1024.0 * (fa * f_scale + wa * w_scale)
        print(f"| {k} | {max(nf, nw)} | {fa:.1f} | {wa:.1f} | {b / 1e6:.2f} | {b:.0f} |")


def build_stream(workload, nsinks):
    """The slabs a pass over sink slabs 0 .. nsinks-1 needs (R more on each
    side), realised by the forked-pool generator."""
    _, cpd, n1d, _ = STREAM_WORKLOADS[workload]
    R = 2
    t0 = time.time()
    step = max(1, (nsinks + 2 * R) // 40)  # about ten progress lines

    def progress(done, total):
        if done % step == 0 or done == total:
            log(f"[bench] stream slabs: {done}/{total} generator jobs done ({time.time() - t0:.0f}s)")

    box = synth.clustered_slabs_parallel(n1d, cpd, 20260255, range(-R, nsinks + R), progress=progress)
    log(f"[bench] built {box.name}: {sum(s.n for s in box.slabs)} particles in {time.time() - t0:.1f}s")
    return box


def run_stream(args, box, workload, nsinks, steps, warmup, options=True):
    """A slab stream: PosXYZ slabs and acceleration slabs in host memory, the
    sink slabs submitted in order inside one stream session (each source slab
    uploaded once while it is in the 2R+1 window), every launch group captured
    in a HIP graph, tiles autotuned.  A step is one pass over the sink slabs,
    each SIC planned in the step (per --planner); the timed region includes
    every host-device transfer.  options=False ignores --pinned, --aos and
    --fuse-kick (pageable PosXYZ slabs, plain near-field output)."""
    _, cpd, n1d, desc = STREAM_WORKLOADS[workload]
    R = 2
    pinned, use_aos, fuse = (args.pinned, args.aos, args.fuse_kick) if options else (False, False, False)
    sinks = list(range(nsinks))
    ell = box.mean_spacing()
    soft_len = ell / 40.0
    b2 = synth.fof_b2(ell)
    graphs = 1 if args.graphs is None else args.graphs
    keep = []

    def host_array(shape, like=None):
        if not pinned:
            return np.zeros(shape, np.float32) if like is None else like
        pa = nfm.PinnedArray(shape, np.float32)
        keep.append(pa)
        pa.array[:] = 0 if like is None else like
        return pa.array

    if pinned:
        box = synth.Box(cpd, [synth.Slab(host_array(s.pos.shape, s.pos), s.count) for s in box.slabs], box.name)
        box.N = n1d ** 3
    acc = {x: host_array((box.slab(x).n, 4)) for x in sinks}
    # posstruct copies of the slabs (what the reference's PosSlab holds before TransposePos)
    aos = {x: host_array((box.slab(x).n * 3,), box.slab(x).xyz.reshape(-1)) for x in range(-R, len(sinks) + R)} \
        if use_aos else None
    kicks = {}
    if fuse:
        rng = np.random.default_rng(7)
        _, e_int = synth.eps_param("single_spline", soft_len)
        for x in sinks:
            far = host_array((box.slab(x).n * 3,), rng.standard_normal(box.slab(x).n * 3).astype(np.float32))
            kicks[x] = nfm.nf_kick(FarAccSlab=far.ctypes.data, rescale=-3.0 * 0.3 / (8.0 * np.pi * box.N),
                                   inv_eps3=1.0 / e_int ** 3, density_rad2=float(b2))
            keep.append(far)

    def pos_ptrs(x):
        return None if aos is None else [aos[x - R + c].ctypes.data for c in range(2 * R + 1)]

    layout = nfm.NF_POS_AOS if aos is not None else nfm.NF_POS_XYZ
    with nfm.NearFieldDriver(cpd, soft_len, b2, nfradius=R, path="batched", buffers_per_device=args.bpd,
                             max_buffer_bytes=6 << 30, sinks_per_lane=args.sinks_per_lane, waves_per_wg=args.waves_per_wg,
                             gpu_streams=args.gpu_streams, max_group=args.max_group, use_graphs=graphs,
                             group_interactions=int(args.group_interactions)) as drv:
        tp = time.time()
        sics = [drv.plan(box, x, acc_ptr=acc[x].ctypes.data, mem=nfm.NF_MEM_HOST, pos_ptrs=pos_ptrs(x),
                         pos_layout=layout, kick=kicks.get(x)) for x in sinks]
        plan_s = time.time() - tp
        direct_total = sum(int(s.desc.DirectTotal) for s in sics)
        if args.planner != "host-once":
            sics = None  # each step plans its own

        def step():
            drv.stream_begin()
            if args.planner == "host":
                tasks = [drv.submit(drv.plan(box, x, acc_ptr=acc[x].ctypes.data, mem=nfm.NF_MEM_HOST,
                                             pos_ptrs=pos_ptrs(x), pos_layout=layout, kick=kicks.get(x)))
                         for x in sinks]
            elif args.planner == "device":
                tasks = [drv.submit_slab(box, x, acc_ptr=acc[x].ctypes.data, pos_ptrs=pos_ptrs(x),
                                         mem=nfm.NF_MEM_HOST, pos_layout=layout, kick=kicks.get(x))
                         for x in sinks]
            else:
                tasks = [drv.submit(s) for s in sics]
            st = []
            for t in tasks:
                t.wait()
                st.append(t.stats())
                t.release()
            drv.stream_end()
            return st

        for i in range(warmup):
            step()
            log(f"[bench] stream: warmup {i + 1}/{warmup} done")
        walls, stats = [], []
        for i in range(steps):
            t1 = time.perf_counter()
            stats.extend(step())
            walls.append(time.perf_counter() - t1)
            log(f"[bench] stream: step {i + 1}/{steps} done ({walls[-1] * 1e3:.1f} ms)")
        table = drv.autotune_table()
    return dict(desc=desc, cpd=cpd, n1d=n1d, sinks=sinks, box=box, acc=acc, soft_len=soft_len, walls=walls,
                stats=stats, steps=steps, warmup=warmup, direct_total=direct_total, plan_s=plan_s,
                table=table, graphs=graphs, pinned=pinned, aos=use_aos, fuse_kick=fuse, keep=keep)


def stream_line(args, res):
    R = 2
    box, stats, steps = res["box"], res["stats"], res["steps"]
    wall = float(sum(res["walls"]))
    up = sum(x["upload_s"] for x in stats)
    down = sum(x["download_s"] for x in stats)
    busy = sum(x["kernel_s"] for x in stats)
    kern = sum(x["direct_kernel_s"] for x in stats)
    xfer = up + down
    hidden = min(1.0, max(0.0, (busy + xfer - wall) / xfer)) if xfer > 0 else None
    di = res["direct_total"] * steps
    nsl = len(res["sinks"])
    return {
        "metric": METRIC,
        "value": di / wall,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": res["warmup"],
        "ms_per_step": 1e3 * wall / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": res["desc"], "cpd": res["cpd"], "N_nominal": res["n1d"] ** 3, "sink_slabs": nsl,
                   "particles_in_window": int(sum(box.slab(x).n for x in range(-R, nsl + R))),
                   "interactions_per_step": res["direct_total"], "path": "batched", "graphs": res["graphs"],
                   "softening": "single_spline", "mem": "host pinned" if res["pinned"] else "host pageable",
                   "positions": "posstruct (AoS)" if res["aos"] else "PosXYZ", "fused_kick": res["fuse_kick"],
                   "planner": args.planner},
        "end_to_end": {
            "interactions_per_s": di / wall,
            "median_step_ms": 1e3 * float(np.median(res["walls"])),
            "device_busy_s_per_step": busy / steps,
            "direct_kernel_interactions_per_s": di / kern if kern > 0 else None,
            "h2d_s_per_step": up / steps, "d2h_s_per_step": down / steps,
            "h2d_GB_per_step": sum(x["bytes_to_device"] for x in stats) / 1e9 / steps,
            "d2h_GB_per_step": sum(x["bytes_from_device"] for x in stats) / 1e9 / steps,
            "slabs_uploaded_per_step": sum(x["slabs_uploaded"] for x in stats) / steps,
            "slabs_reused_per_step": sum(x["slabs_reused"] for x in stats) / steps,
            "hidden_transfer_fraction": hidden,
            "plan_s_before_timing": res["plan_s"],
            "device_plan_s_per_step": sum(x["plan_s"] for x in stats) / steps,
        },
        "autotune": [{k: e[k] for k in ("occupancy_quarters", "size_class", "sinks_per_lane", "waves_per_wg", "uses")}
                     for e in res["table"]],
    }


def pair_probe():
    """The pair-math probe of the single-spline forms at the uniform 512^3
    workload's eps and b2: interactions/s of the batched tile loop alone for
    each pair form (0 unsoftened, 1 with the FOF density, 2 the full spline),
    sinks per lane and occupancy (waves per SIMD), and the fraction of the
    FP32 peak at 20 FLOP per interaction."""
    import ctypes as C
    ell = 1.0 / 512
    eps, _ = synth.eps_param("single_spline", ell / 40.0)
    b2 = synth.fof_b2(ell)
    lib = nfm.lib()
    lib.nf_probe_pair_rate.argtypes = [C.c_float, C.c_float, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    rows = []
    for S in (1, 2, 4):
        for wpe in (4, 6, 8):
            row = {"sinks_per_lane": S, "waves_per_simd": wpe}
            for mode, name in ((0, "fast"), (1, "density"), (2, "exact")):
                r = C.c_double(0.0)
                nfm.check(lib.nf_probe_pair_rate(float(eps), float(b2), S, mode, wpe, C.byref(r)), "nf_probe_pair_rate")
                row[name] = r.value
            log(f"[probe] S={S} waves/SIMD={wpe}: fast {row['fast']:.3e} density {row['density']:.3e} "
                f"exact {row['exact']:.3e} interactions/s")
            rows.append(row)
    best = max(rows, key=lambda r: r["fast"])
    print(json.dumps({"pair_probe": rows, "peak_tflops": PEAK_FP32_TFLOPS, "flop_per_interaction": FLOP_PER_INTERACTION,
                      "best_fast": best, "best_fast_frac_of_peak":
                          FLOP_PER_INTERACTION * best["fast"] / (PEAK_FP32_TFLOPS * 1e12)}), flush=True)


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


def cpu_baseline(box, R, soft_len, budget_s, results, slabs=None):
    """The reference CPU directs, restated (oracle/cpu_directs.c): the cell-pair
    loop of ExecuteSlabCPU with OpenMP over y-rows on all host cores, in two
    kernel forms -- the explicit AVX-512 kernel in the structure of
    AVX512Direct::compute (16 sources per vector, rsqrt14 + one Newton step),
    timed first with two thirds of the budget and reported as the value, and
    the production fallback's auto-vectorised DirectCPUKernels.cc kernel
    (-Ofast, 16-wide) with the rest.  Whole slabs, the GPU-sampled ones
    first; the GPU's accelerations are compared with the first form's on
    their first 2000 sinks."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    eps, _ = synth.eps_param("single_spline", soft_len)
    cpd = box.cpd
    slabs = list(range(cpd)) if slabs is None else list(slabs)
    first = list(next(iter(results.values()))["acc_sample"])
    order = first + [x for x in slabs if x not in first]

    def timed(fn, budget, compare):
        done = di = 0
        agree = {}
        t0 = time.perf_counter()
        for x in order:
            acc, n = fn(box, R, "single_spline", eps, x)
            di += n
            done += 1
            if compare and x in first:
                for p, d in results.items():
                    g = d["acc_sample"][x].astype(np.float64)
                    c = acc[:2000].astype(np.float64)
                    # the GPU carries the reference's output units (no e^-3), as does the CPU path
                    rel = np.linalg.norm(g[:, :3] - c[:, :3], axis=1) / np.linalg.norm(c[:, :3], axis=1)
                    agree.setdefault(p, []).append(rel)
            if time.perf_counter() - t0 > budget:
                break
        return di, done, time.perf_counter() - t0, agree

    di, done, el, agree = timed(po.execute_slab_avx512, budget_s * 2.0 / 3.0, True)
    di2, done2, el2, _ = timed(po.execute_slab_cpu, budget_s / 3.0, False)
    info = po.cpu_info()
    cg = po.cpu_directs_codegen() or {}
    out = {"value": di / el, "unit": "interactions/s", "cores": po.num_threads(), "kind": "port",
           "sample": f"{done} whole slabs of the same box ({di:.3e} interactions, {el:.1f}s), "
                     "explicit AVX-512 form",
           "cpu_model": info["model"], "logical_cpus": info["logical_cpus"], "avx512f": info["avx512f"],
           "isa": info["isa"],
           "kernel": {"form": "explicit AVX-512 (AVX512Direct::compute structure)",
                      "vector_lanes": cg.get("or_execute_slab_avx512", {}).get("vector_lanes"),
                      "rsqrt": "vrsqrt14ps + 1 Newton step"},
           "autovectorised_fallback": {"value": di2 / el2, "unit": "interactions/s",
                                       "sample": f"{done2} whole slabs ({di2:.3e} interactions, {el2:.1f}s)",
                                       "vector_lanes": cg.get("or_execute_slab_cpu", {}).get("vector_lanes"),
                                       "rsqrt": cg.get("or_execute_slab_cpu", {}).get("rsqrt")}}
    out["gpu_vs_cpu_rel_diff"] = {p: {"rms": float(np.sqrt(np.mean(np.concatenate(v) ** 2))),
                                      "max": float(np.concatenate(v).max())} for p, v in agree.items()}
    return out


def error_check(box, R, soft_len, accs, frac, seed, slabs=None):
    """Part of the CPU leg: the GPU's accelerations ({path: {slab: acc}})
    against the FP64 direct-summation oracle on a seeded sample of the sinks
    (a fraction ``frac`` of each slab's), as rms and max relative error over
    the sample, outside softening-dominated sinks, and over those of condition
    number <= 100 (oracle/pyoracle.py error_report)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    eps, _ = synth.eps_param("single_spline", soft_len)
    b2 = synth.fof_b2(box.mean_spacing())
    t0 = time.perf_counter()
    rep = po.error_report(box, R, "single_spline", eps, b2, accs, frac, seed, slabs)
    rep["oracle_s"] = time.perf_counter() - t0
    log(f"[bench] FP64 check of {box.name}: {rep['sample']['sinks']} sinks in {rep['oracle_s']:.1f}s")
    return rep


if __name__ == "__main__":
    main()
