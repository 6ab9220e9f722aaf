"""nf_execute_sic_noalloc, the entry point that allocates nothing: one
device-resident SIC on the batched path with the module's fixed device
scratch.  Its accelerations must equal, bitwise, those nf_submit gives for
the same SIC on the batched path, on two calls in a row, the second with a
smaller SIC (the scratch keeps the first call's contents)."""
import numpy as np
import pytest

import nearfield as nfm
import synth
from nfhelpers import R, params, small_uniform

pytestmark = pytest.mark.gpu


def _sic_rows(box, slab, j_low, j_high):
    """Particle rows [r0, r1) of the sink slab that rows j_low..j_high-1 cover."""
    s = box.slab(slab)
    cpd = box.cpd
    return int(s.start[j_low * cpd]), int(s.start[(j_high - 1) * cpd + cpd - 1] + s.count[(j_high - 1) * cpd + cpd - 1])


def _submit_and_noalloc(box, slab, j_low, j_high):
    """One device-resident SIC run through nf_submit (batched) and through
    nf_execute_sic_noalloc; both acceleration slabs start as the same NaN
    fill.  Returns (submit's, noalloc's) accelerations of the whole slab."""
    _, b2, _ = params(box)
    n = box.slab(slab).n
    init = np.full((n, 4), np.nan, np.float32)
    print(f"noalloc: {box.name} slab {slab} rows [{j_low},{j_high}) n={n}", flush=True)
    dev = [nfm.DeviceArray(host=box.slab(x).pos) for x in range(box.cpd)]
    acc_ref, acc_na = nfm.DeviceArray(host=init), nfm.DeviceArray(host=init)
    ptrs = [dev[(slab - R + c) % box.cpd].ptr for c in range(2 * R + 1)]
    try:
        with nfm.NearFieldDriver(box.cpd, box.mean_spacing() / 40.0, b2, path="batched", buffers_per_device=1,
                                 max_buffer_bytes=1 << 30) as drv:
            sic = drv.plan(box, slab, j_low, j_high, pos_ptrs=ptrs, acc_ptr=acc_ref.ptr, mem=nfm.NF_MEM_DEVICE)
            drv.submit(sic, blocking=True).release()
            sic_na = drv.plan(box, slab, j_low, j_high, pos_ptrs=ptrs, acc_ptr=acc_na.ptr, mem=nfm.NF_MEM_DEVICE)
        nfm.execute_sic_noalloc(sic_na)  # no nf_setup in force
        ref = acc_ref.to_host(np.float32, (n, 4))
        got = acc_na.to_host(np.float32, (n, 4))
    finally:
        for d in dev + [acc_ref, acc_na]:
            d.free()
    return ref, got


def test_noalloc_matches_submit_on_two_calls():
    big = synth.clustered_box(40, 11, seed=7)   # cells of 0 to several hundred particles
    small = small_uniform(cpd=7, ppc=20, seed=5)
    for box, slab, jl, jh in ((big, 5, 0, big.cpd), (small, 2, 1, 4)):
        ref, got = _submit_and_noalloc(box, slab, jl, jh)
        r0, r1 = _sic_rows(box, slab, jl, jh)
        assert r1 > r0
        assert np.isfinite(got[r0:r1]).all()
        assert np.isnan(got[:r0]).all() and np.isnan(got[r1:]).all()  # rows outside the SIC untouched
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (box.name, slab, jl, jh)
