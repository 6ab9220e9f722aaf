"""ctypes access to the oracle libraries (liboracle.so, libcpudirects.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg -- never by the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
KIND = {"plummer": 0, "single_spline": 1, "cubic_spline": 2, "cubic_plummer": 3}

_libs = {}


def _load(name):
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        _libs[name] = C.CDLL(path)
    return _libs[name]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _ptrs(arrs):
    return (C.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


class _BoxArgs:
    """Per-slab pointer tables of a box, kept alive with the arrays."""

    def __init__(self, box):
        s = box.slabs
        self.keep = s
        self.pos = _ptrs([x.pos for x in s])
        self.nslab = np.array([x.n for x in s], np.int64)
        self.start = _ptrs([x.start for x in s])
        self.count = _ptrs([x.count for x in s])


def kind_of(softening):
    return KIND[softening] if isinstance(softening, str) else int(softening)


# ---- pair kernels -----------------------------------------------------------

def pair_f64(kind, d, eps, b2):
    lib = _load("liboracle.so")
    out = np.zeros(4)
    lib.or_pair_f64(C.c_int(kind_of(kind)), C.c_double(d[0]), C.c_double(d[1]), C.c_double(d[2]),
                    C.c_double(eps), C.c_double(b2), _p(out))
    return out


def pair_gpu_f32(kind, src, sink, eps, b2):
    lib = _load("liboracle.so")
    out = np.zeros(4, np.float32)
    s = np.asarray(src, np.float32)
    k = np.asarray(sink, np.float32)
    lib.or_pair_gpu_f32(C.c_int(kind_of(kind)), _p(s), _p(k), C.c_float(eps), C.c_float(b2), _p(out))
    return out


def pair_cpu_f32(kind, src, sink, eps):
    lib = _load("libcpudirects.so")
    out = np.zeros(3, np.float32)
    s = np.asarray(src, np.float32)
    k = np.asarray(sink, np.float32)
    lib.or_pair_cpu_f32(C.c_int(kind_of(kind)), _p(s), _p(k), C.c_float(eps), _p(out))
    return out


# ---- box-level oracles --------------------------------------------------------

def direct_f64(box, R, kind, eps, b2, sink_slab, sink_index, sink_cell, with_soft=False, with_abs=False):
    """FP64 brute-force sum for the listed sinks -> (n, 4) float64; with
    ``with_soft`` also the norm of the softened-pair part of each sum, with
    ``with_abs`` also the sum of the pair terms' magnitudes (returned in that
    order after the sums)."""
    lib = _load("liboracle.so")
    a = _BoxArgs(box)
    ss = np.ascontiguousarray(sink_slab, np.int32)
    si = np.ascontiguousarray(sink_index, np.int64)
    sc = np.ascontiguousarray(sink_cell, np.int32)
    out = np.zeros((ss.size, 4))
    soft = np.zeros(ss.size) if with_soft else None
    mag = np.zeros(ss.size) if with_abs else None
    lib.or_direct_f64(C.c_int(box.cpd), a.pos, _p(a.nslab), a.start, a.count, C.c_int(R),
                      C.c_int(kind_of(kind)), C.c_double(float(eps)), C.c_double(float(b2)),
                      C.c_int64(ss.size), _p(ss), _p(si), _p(sc), _p(out),
                      _p(soft) if with_soft else None, _p(mag) if with_abs else None)
    extra = [v for v, on in ((soft, with_soft), (mag, with_abs)) if on]
    return tuple([out] + extra) if extra else out


def floor_f32(box, R, eps, sink_slab, sink_index, sink_cell):
    """The FP32 pair-term floor (or_floor_f32, single spline): each sink's
    force from pair terms formed in FP32 as the batched kernel forms them,
    summed in FP64 -> (n, 3) float64 in the reference's units."""
    lib = _load("liboracle.so")
    a = _BoxArgs(box)
    ss = np.ascontiguousarray(sink_slab, np.int32)
    si = np.ascontiguousarray(sink_index, np.int64)
    sc = np.ascontiguousarray(sink_cell, np.int32)
    out = np.zeros((ss.size, 3))
    lib.or_floor_f32(C.c_int(box.cpd), a.pos, _p(a.nslab), a.start, a.count, C.c_int(R),
                     C.c_double(float(eps)), C.c_int64(ss.size), _p(ss), _p(si), _p(sc), _p(out))
    return out


def slab_sinks(box, x):
    """(slab, index, cell) arrays for every particle of slab x."""
    s = box.slab(x)
    cell = np.repeat(np.arange(box.cpd * box.cpd, dtype=np.int32), s.count)
    return np.full(s.n, x % box.cpd, np.int32), np.arange(s.n, dtype=np.int64), cell


def sample_sinks(box, frac, seed, slabs=None):
    """A seeded random subsample of the sinks of the box (of the listed slabs
    only, if given): round(frac * n) of each slab's n, at least one."""
    rng = np.random.default_rng(seed)
    sl, ix, ce = [], [], []
    for x in (range(box.cpd) if slabs is None else [s % box.cpd for s in slabs]):
        s = box.slab(x)
        k = max(1, int(round(frac * s.n))) if s.n else 0
        if k == 0:
            continue
        pick = np.sort(rng.choice(s.n, size=k, replace=False))
        cell = np.repeat(np.arange(box.cpd * box.cpd, dtype=np.int32), s.count)
        sl.append(np.full(k, x, np.int32))
        ix.append(pick.astype(np.int64))
        ce.append(cell[pick])
    return np.concatenate(sl), np.concatenate(ix), np.concatenate(ce)


def error_report(box, R, kind, eps, b2, accs, frac, seed, slabs=None):
    """Relative acceleration errors |a - a_64| / |a_64| of each result in
    ``accs`` ({name: {slab: (n, 4) float32 acc}}) on a seeded sample of sinks
    (sample_sinks) against direct_f64: rms and max over the whole sample,
    outside softening-dominated sinks (pairs closer than the softening length
    carry more than half of |a_64|), and over the sinks outside those whose
    condition number (sum of the pair terms' magnitudes over |a_64|) is at
    most 100; the sinks above 1e-5 also against the FP32 pair-term floor
    (_above_floor)."""
    sl, ix, ce = sample_sinks(box, frac, seed, slabs)
    ref, soft, mag = direct_f64(box, R, kind, eps, b2, sl, ix, ce, with_soft=True, with_abs=True)
    an = np.linalg.norm(ref[:, :3], axis=1)
    keep = soft <= 0.5 * an
    cond = mag / np.maximum(an, 1e-300)
    out = {"sample": {"fraction": frac, "seed": seed, "sinks": int(sl.size),
                      "softening_dominated": int((~keep).sum()),
                      "condition_number_above_100": int((keep & (cond > 100)).sum())}}

    def stats(rel, mask):
        r = rel[mask]
        if r.size == 0:
            return None
        return {"n": int(r.size), "rms": float(np.sqrt(np.mean(r ** 2))), "max": float(r.max()),
                "above_1e-5": int((r > 1e-5).sum())}

    for name, acc in accs.items():
        got = np.empty((sl.size, 3))
        for x in np.unique(sl):
            m = sl == x
            got[m] = acc[int(x)][ix[m], :3]
        rel = np.linalg.norm(got - ref[:, :3], axis=1) / np.maximum(an, 1e-300)
        out[name] = {"all": stats(rel, np.ones(rel.size, bool)), "outside_softening_dominated": stats(rel, keep),
                     "condition_number_le_100": stats(rel, keep & (cond <= 100)),
                     "above_1e-5": _above_floor(box, R, eps, (sl, ix, ce), rel, ref, an, soft, cond, kind)}
    return out


def _above_floor(box, R, eps, sample, rel, ref, an, soft, cond, kind, bar=1e-5, worst=6):
    """The sinks of a result whose error exceeds ``bar``, against the FP32
    pair-term floor on those sinks (floor_f32, single spline only): how many
    err by no more than 1.5x the floor's own error, and the worst few with
    their floor error, condition number and softened fraction of |a|."""
    sl, ix, ce = sample
    bad = np.flatnonzero(rel > bar)
    rep = {"n": int(bad.size)}
    if bad.size == 0 or kind_of(kind) != KIND["single_spline"]:
        return rep
    fl = floor_f32(box, R, eps, sl[bad], ix[bad], ce[bad])
    fl_rel = np.linalg.norm(fl - ref[bad, :3], axis=1) / np.maximum(an[bad], 1e-300)
    rep["within_1.5x_fp32_floor"] = int((rel[bad] <= 1.5 * fl_rel).sum())
    order = np.argsort(-rel[bad])[:worst]
    rep["worst"] = [{"slab": int(sl[bad[i]]), "index": int(ix[bad[i]]), "rel": float(rel[bad[i]]),
                     "fp32_floor_rel": float(fl_rel[i]), "condition_number": float(cond[bad[i]]),
                     "softened_fraction": float(soft[bad[i]] / max(an[bad[i]], 1e-300))} for i in order]
    return rep


def calc_di(box, R, x):
    lib = _load("liboracle.so")
    lib.or_calc_di.restype = C.c_uint64
    a = _BoxArgs(box)
    return int(lib.or_calc_di(C.c_int(box.cpd), a.count, C.c_int(R), C.c_int(x)))


def execute_slab_cpu(box, R, kind, eps, x, y0=0, y1=None):
    """ExecuteSlabCPU restatement -> ((n, 4) float32 acc, interactions)."""
    lib = _load("libcpudirects.so")
    lib.or_execute_slab_cpu.restype = C.c_uint64
    a = _BoxArgs(box)
    if y1 is None:
        y1 = box.cpd
    acc = np.zeros((box.slab(x).n, 4), np.float32)
    di = lib.or_execute_slab_cpu(C.c_int(box.cpd), a.pos, _p(a.nslab), a.start, a.count, C.c_int(R),
                                 C.c_int(kind_of(kind)), C.c_float(eps), C.c_int(x % box.cpd), C.c_int(y0),
                                 C.c_int(y1), _p(acc))
    return acc, int(di)


def execute_slab_avx512(box, R, kind, eps, x, y0=0, y1=None):
    """The same cell-pair loop on the explicit AVX-512 kernel (16 sources per
    vector, rsqrt14 + one Newton step) -> ((n, 4) float32 acc, interactions)."""
    lib = _load("libcpudirects.so")
    lib.or_execute_slab_avx512.restype = C.c_uint64
    a = _BoxArgs(box)
    if y1 is None:
        y1 = box.cpd
    acc = np.zeros((box.slab(x).n, 4), np.float32)
    di = lib.or_execute_slab_avx512(C.c_int(box.cpd), a.pos, _p(a.nslab), a.start, a.count, C.c_int(R),
                                    C.c_int(kind_of(kind)), C.c_float(eps), C.c_int(x % box.cpd), C.c_int(y0),
                                    C.c_int(y1), _p(acc))
    return acc, int(di)


def num_threads():
    return int(_load("libcpudirects.so").or_num_threads())


def cpu_directs_codegen():
    """What the compiler made of the two CPU-directs loops: per OpenMP body,
    the widest vector register its FP32 arithmetic uses (16 lanes for zmm)
    and whether the reciprocal square root is the 14-bit estimate
    (vrsqrt14ps) refined in vector code.  None where objdump is missing."""
    import shutil
    import subprocess
    if shutil.which("objdump") is None:
        return None
    txt = subprocess.run(["objdump", "-d", "--no-show-raw-insn", os.path.join(_HERE, "libcpudirects.so")],
                         check=True, capture_output=True, text=True).stdout
    out, cur = {}, None
    for line in txt.splitlines():
        if line.endswith(">:"):
            name = line.split("<", 1)[1][:-2]
            cur = name.split("._omp_fn")[0] if "._omp_fn" in name else None
            if cur:
                out.setdefault(cur, {"ops": {}})
            continue
        if cur and "ps" in line:
            ins = line.split("\t")[-1].split()
            if ins:
                op, regs = ins[0], " ".join(ins[1:])
                width = 16 if "%zmm" in regs else 8 if "%ymm" in regs else 4 if "%xmm" in regs else 0
                if op.endswith("ps") and width:
                    out[cur]["ops"][op] = max(out[cur]["ops"].get(op, 0), width)
    for v in out.values():
        ops = v.pop("ops")
        arith = [w for o, w in ops.items() if o.startswith(("vfmadd", "vfnmadd", "vmul", "vsub", "vadd"))]
        v["vector_lanes"] = max(arith) if arith else 1
        v["rsqrt"] = "vrsqrt14ps" if "vrsqrt14ps" in ops else ("vsqrtps" if "vsqrtps" in ops else "scalar")
        v["rsqrt_lanes"] = ops.get("vrsqrt14ps", 0)
    return out


def cpu_info():
    """The host CPU: model name, logical CPUs, and whether it has avx512f."""
    model, flags, ncpu = "unknown", set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "processor":
                    ncpu += 1
                elif k == "model name" and model == "unknown":
                    model = v.strip()
                elif k == "flags" and not flags:
                    flags = set(v.split())
    except OSError:
        pass
    return {"model": model, "logical_cpus": ncpu, "avx512f": "avx512f" in flags,
            "isa": sorted(f for f in flags if f.startswith("avx512"))}


# ---- pencil-level restatements ---------------------------------------------

def gather_sinks(plan, box, slab, R):
    lib = _load("liboracle.so")
    s = box.slab(slab)
    W = 2 * R + 1
    out = np.zeros((plan["NSinkBlocks"] * 64, 4), np.float32)
    st = np.ascontiguousarray(plan["SinkPlanStart"], np.uint32)
    nn = np.ascontiguousarray(plan["SinkPlanN"], np.int32)
    wc = np.ascontiguousarray(plan["WrapCell"], np.int32)
    ss = np.ascontiguousarray(plan["SinkSetStart"], np.int32)
    lib.or_gather_sinks(C.c_int(plan["NSinkSets"]), C.c_int(W), C.c_int(R), _p(st), _p(nn), _p(wc), _p(ss),
                        _p(s.pos), C.c_int64(s.n), C.c_float(np.float32(1.0 / box.cpd)), _p(out))
    return out


def gather_sources(plan, box, slab, R):
    lib = _load("liboracle.so")
    W = 2 * R + 1
    slabs = [box.slab(slab - R + c) for c in range(W)]
    out = np.zeros((plan["NSourceBlocks"] * 64, 4), np.float32)
    st = np.ascontiguousarray(plan["SourcePlanStart"], np.uint32)
    nn = np.ascontiguousarray(plan["SourcePlanN"], np.int32)
    ss = np.ascontiguousarray(plan["SourceSetStart"], np.int32)
    nsl = np.array([x.n for x in slabs], np.int64)
    lib.or_gather_sources(C.c_int(plan["NSourceSets"]), C.c_int(W), C.c_int(R), _p(st), _p(nn), _p(ss),
                          _ptrs([x.pos for x in slabs]), _p(nsl), C.c_float(np.float32(1.0 / box.cpd)), _p(out))
    return out


def coadd_f32(plan, pinacc, nslab, R):
    lib = _load("liboracle.so")
    W = 2 * R + 1
    acc = np.zeros((nslab, 4), np.float32)
    st = np.ascontiguousarray(plan["SinkPlanStart"], np.uint32)
    nn = np.ascontiguousarray(plan["SinkPlanN"], np.int32)
    wc = np.ascontiguousarray(plan["WrapCell"], np.int32)
    ss = np.ascontiguousarray(plan["SinkSetStart"], np.int32)
    sc = np.ascontiguousarray(plan["SinkSetCount"], np.int32)
    pa = np.ascontiguousarray(pinacc, np.float32)
    lib.or_coadd_f32(C.c_int(plan["NSinkSets"]), C.c_int(plan["Nk"]), C.c_int(W), _p(st), _p(nn), _p(wc), _p(ss),
                     _p(sc), _p(pa), _p(acc))
    return acc


def pencil_execute(plan, sinks4, sources4, R, kind, eps, b2):
    lib = _load("liboracle.so")
    W = 2 * R + 1
    out = np.zeros((plan["NSinkBlocks"] * 64, 4), np.float32)
    par = np.ascontiguousarray(plan["SinkBlockParentPencil"], np.int32)
    idm = np.ascontiguousarray(plan["SinkSetIdMax"], np.uint32)
    ss = np.ascontiguousarray(plan["SourceSetStart"], np.int32)
    sc = np.ascontiguousarray(plan["SourceSetCount"], np.int32)
    il = np.ascontiguousarray(plan["SinkSourceInteractionList"], np.int32)
    yo = np.ascontiguousarray(plan["SinkSourceYOffset"], np.float32)
    k4 = np.ascontiguousarray(sinks4, np.float32)
    s4 = np.ascontiguousarray(sources4, np.float32)
    lib.or_pencil_execute(C.c_int(kind_of(kind)), C.c_int(W), C.c_int(plan["NSinkBlocks"]), C.c_int(64), _p(par),
                          _p(idm), _p(ss), _p(sc), _p(il), _p(yo), _p(k4), _p(s4), C.c_float(eps), C.c_float(b2),
                          _p(out))
    return out


# ---- consumer of the near field ---------------------------------------------

def rescale_coadd(nacc, facc, rescale, inv_eps3, density_rad2):
    """RescaleAndCoAddAcceleration (kick.cpp:74-109) for the single-spline,
    FOF-density build: nacc is the float4 near-field slab (FLOAT3p1), facc the
    float3 far-field slab (acc3struct).  FLOAT3p1 rules (float3p1.cpp:41-56):
    scaling by a float and adding a float3 leave w alone.  rescale and
    inv_eps3 are FLOAT (rounded to float32); each operation is rounded on its
    own; w -= DensityKernelRad2 is evaluated in double (a double member,
    statestructure.cpp:102) and rounded back."""
    n = np.asarray(nacc, np.float32).reshape(-1, 4)
    f = np.asarray(facc, np.float32).reshape(-1, 3)
    r, ie = np.float32(rescale), np.float32(inv_eps3)
    out = np.empty_like(n)
    out[:, :3] = ((n[:, :3] * ie) + f) * r
    out[:, 3] = (n[:, 3].astype(np.float64) - float(density_rad2)).astype(np.float32)
    return out


def near_acc_record(nacc, inv_eps3):
    """The NearAcc record StoreForces == 2 writes for a slab (pipeline.cpp:374-399):
    the AccSlab with the single-spline prefactor applied in place, nearacc[i] *=
    inv_eps3 with a FLOAT inv_eps3 -- FLOAT3p1 *= FLOAT (float3p1.cpp:50-51)
    scales x, y, z and leaves w.  Other softenings write the AccSlab as it is,
    which is inv_eps3 = 1."""
    n = np.asarray(nacc, np.float32).reshape(-1, 4)
    out = n.copy()
    out[:, :3] = n[:, :3] * np.float32(inv_eps3)
    return out
