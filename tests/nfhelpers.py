"""Shared helpers for the near-field tests: workloads, error metrics."""
import numpy as np

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
