"""Shared helpers for the near-field tests: workloads, error metrics."""
import numpy as np

import nearfield as nfm
import pyoracle as po
import synth

R = 2


def small_uniform(cpd=9, ppc=20, seed=11):
    return synth.uniform_box(cpd ** 3 * ppc, cpd, seed)


def params(box, softening="single_spline"):
    """(eps, b2, e_internal) with the workloads' conventions: Plummer-equivalent
    softening = 1/40 of the mean spacing, FOF b = 0.25 of it."""
    ell = box.mean_spacing()
    eps, e = synth.eps_param(softening, ell / 40.0)
    return eps, synth.fof_b2(ell), e


def rel_err(a, ref):
    """Per-sink relative acceleration error |a - ref| / |ref| (3 components)."""
    a = np.asarray(a, np.float64)[:, :3]
    r = np.asarray(ref, np.float64)[:, :3]
    den = np.linalg.norm(r, axis=1)
    num = np.linalg.norm(a - r, axis=1)
    return num / np.where(den > 0, den, 1.0)


def summarize(rel, mask=None, tail=False):
    if mask is not None:
        rel = rel[mask]
    out = dict(rms=float(np.sqrt(np.mean(rel ** 2))) if rel.size else 0.0,
               max=float(rel.max()) if rel.size else 0.0, n=int(rel.size))
    if tail and rel.size:
        out.update(p99=float(np.percentile(rel, 99)), p999=float(np.percentile(rel, 99.9)),
                   above_1e5=int((rel > 1e-5).sum()))
    return out


def not_soft_dominated(ref, soft):
    """Sinks whose softened pairs (r below the softening length) contribute at
    most half of |a_oracle|."""
    return soft <= 0.5 * np.linalg.norm(np.asarray(ref)[:, :3], axis=1)


def oracle_slab(box, x, softening, eps, b2, with_soft=False):
    sl, ix, ce = po.slab_sinks(box, x)
    return po.direct_f64(box, R, softening, eps, b2, sl, ix, ce, with_soft=with_soft)


def run_sics(drv, box, slabs, mem, pos_layout=nfm.NF_POS_XYZ, kick_args=None, planner="host",
             with_stats=False, near_out=False):
    """Whole-slab SICs for ``slabs`` in host or device mode; positions in the
    given layout, an optional fused kick per slab (far-field slab, constants).
    planner "host": nf_plan_sic + nf_submit; "device": nf_submit_slab.
    Returns {slab: (n, 4) float32}; with_stats=True also the tasks' stats;
    near_out=True (with a kick) also the NearAccOut records {slab: (n, 4)}."""
    cpd = box.cpd
    keep, out, tasks, stats, near = [], {}, [], [], {}
    pos = {}
    for x in range(cpd):
        s = box.slab(x)
        a = np.ascontiguousarray(s.xyz) if pos_layout == nfm.NF_POS_AOS else s.pos
        if mem == nfm.NF_MEM_DEVICE:
            d = nfm.DeviceArray(host=a)
            keep.append(d)
            pos[x] = d.ptr
        else:
            keep.append(a)
            pos[x] = a.ctypes.data
    for x in slabs:
        n = box.slab(x).n
        kick = None
        if kick_args is not None:
            far, rescale, inv_eps3, dens = kick_args[x]
            if mem == nfm.NF_MEM_DEVICE:
                fd = nfm.DeviceArray(host=far)
                keep.append(fd)
                fp = fd.ptr
            else:
                keep.append(far)
                fp = far.ctypes.data
            kick = nfm.nf_kick(FarAccSlab=fp, rescale=rescale, inv_eps3=inv_eps3, density_rad2=dens)
            if near_out:
                if mem == nfm.NF_MEM_DEVICE:
                    near[x] = nfm.DeviceArray(nbytes=16 * n)
                    keep.append(near[x])
                    kick.NearAccOut = near[x].ptr
                else:
                    near[x] = np.full((n, 4), np.nan, np.float32)
                    kick.NearAccOut = near[x].ctypes.data
        if mem == nfm.NF_MEM_DEVICE:
            acc = nfm.DeviceArray(nbytes=16 * n)
            keep.append(acc)
            aptr = acc.ptr
        else:
            acc = np.full((n, 4), np.nan, np.float32)
            aptr = acc.ctypes.data
        ptrs = [pos[(x - R + c) % cpd] for c in range(2 * R + 1)]
        if planner == "device":
            tasks.append(drv.submit_slab(box, x, acc_ptr=aptr, pos_ptrs=ptrs, mem=mem, pos_layout=pos_layout,
                                         kick=kick))
        else:
            sic = drv.plan(box, x, pos_ptrs=ptrs, acc_ptr=aptr, mem=mem, pos_layout=pos_layout, kick=kick)
            tasks.append(drv.submit(sic))
        out[x] = acc
    for t in tasks:
        t.wait()
        stats.append(t.stats())
        t.release()
    res = {x: (a.to_host(np.float32, (box.slab(x).n, 4)) if mem == nfm.NF_MEM_DEVICE else a)
           for x, a in out.items()}
    if near_out:
        return res, {x: (a.to_host(np.float32, (box.slab(x).n, 4)) if mem == nfm.NF_MEM_DEVICE else a)
                     for x, a in near.items()}
    return (res, stats) if with_stats else res
