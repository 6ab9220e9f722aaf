"""The C ABI: the library loads, exports every entry point include/*.h
declares, its structs match the ctypes mirror, and its host-side error
paths return codes instead of aborting (no GPU work)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import nearfield as nfm
from nfhelpers import R, params, small_uniform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nf_direct.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nf_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 12
    lib = C.CDLL(nfm.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layout_matches_c(tmp_path):
    """Compile a probe against the header with the host C compiler and compare
    sizes/offsets with the ctypes mirror used by nearfield.py."""
    fields = [f for f, _ in nfm.nf_sic._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "nf_direct.h"', 'int main(void){',
           'printf("%zu %zu %zu %zu %zu\\n", sizeof(nf_sic), sizeof(nf_sink_plan), sizeof(nf_source_plan),'
           ' sizeof(nf_cellinfo), sizeof(nf_task_stats));']
    for f in fields:
        src.append(f'printf("%zu\\n", offsetof(nf_sic, {f}));')
    slab_fields = [f for f, _ in nfm.nf_slab_task._fields_]
    for f in slab_fields:
        src.append(f'printf("%zu\\n", offsetof(nf_slab_task, {f}));')
    src.append('printf("%zu\\n", sizeof(nf_slab_task));')
    tp_fields = [f for f, _ in nfm.nf_throughput._fields_]
    for f in tp_fields:
        src.append(f'printf("%zu\\n", offsetof(nf_throughput, {f}));')
    src.append('printf("%zu\\n", sizeof(nf_throughput));')
    src.append('printf("%zu\\n", offsetof(nf_kick, NearAccOut));')
    src.append('printf("%zu %zu %zu\\n", offsetof(nf_kick, density_rad2), sizeof(nf_kick), sizeof(nf_setup_params));'
               ' return 0; }')
    c = tmp_path / "probe.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(c), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    sizes = list(map(int, out[:5]))
    assert sizes == [C.sizeof(nfm.nf_sic), C.sizeof(nfm.nf_sink_plan), C.sizeof(nfm.nf_source_plan), 40,
                     C.sizeof(nfm.nf_task_stats)]
    # the plan structs keep the reference's (non-PARALLEL) layout: 5*(4+4)+4, 5*(4+4)
    assert sizes[1] == 44 and sizes[2] == 40
    offs = list(map(int, out[5:5 + len(fields)]))
    assert offs == [getattr(nfm.nf_sic, f).offset for f in fields]
    k = 5 + len(fields)
    offs = list(map(int, out[k:k + len(slab_fields) + 1]))
    assert offs == [getattr(nfm.nf_slab_task, f).offset for f in slab_fields] + [C.sizeof(nfm.nf_slab_task)]
    k += len(slab_fields) + 1
    offs = list(map(int, out[k:k + len(tp_fields) + 1]))
    assert offs == [getattr(nfm.nf_throughput, f).offset for f in tp_fields] + [C.sizeof(nfm.nf_throughput)]
    assert list(map(int, out[-4:])) == [nfm.nf_kick.NearAccOut.offset, nfm.nf_kick.density_rad2.offset,
                                        C.sizeof(nfm.nf_kick), C.sizeof(nfm.nf_setup_params)]


def test_plan_errors_are_codes():
    lib = nfm.lib()
    box = small_uniform(cpd=7, ppc=4, seed=1)
    eps, b2, _ = params(box)
    W = 2 * R + 1
    ci = (C.c_void_p * W)(*[box.slab(c).cellinfo.ctypes.data for c in range(W)])
    arena = np.zeros(1 << 20, np.uint8)
    out = nfm.nf_sic()
    # even cpd, bad rows, null cellinfo, tiny arena
    assert lib.nf_plan_sic(8, R, 0, 0, 8, ci, None, None, eps, b2, 1, arena.ctypes.data, arena.size, C.byref(out)) == -1
    assert lib.nf_plan_sic(7, R, 0, 5, 3, ci, None, None, eps, b2, 1, arena.ctypes.data, arena.size, C.byref(out)) == -1
    assert lib.nf_plan_sic(7, 3, 0, 0, 7, ci, None, None, eps, b2, 1, arena.ctypes.data, arena.size, C.byref(out)) == -1
    assert lib.nf_plan_sic(7, R, 0, 0, 7, ci, None, None, eps, b2, 1, arena.ctypes.data, 64, C.byref(out)) == -3
    assert b"arena" in lib.nf_last_error()
    # a good plan
    assert lib.nf_plan_sic(7, R, 0, 0, 7, ci, None, None, eps, b2, 1, arena.ctypes.data, arena.size,
                           C.byref(out)) == 0
    assert out.NSinkSets == 49 and out.InteractionCount == 5 * 49
    # the device planner rejects the same arguments before touching a GPU
    nsl = (C.c_uint64 * W)(*[box.slab(c).n for c in range(W)])
    assert lib.nf_plan_sic_device(8, R, 0, 0, 8, ci, None, nsl, eps, b2, 1, arena.ctypes.data, arena.size,
                                  C.byref(out)) == -1
    assert lib.nf_plan_sic_device(7, R, 0, 5, 3, ci, None, nsl, eps, b2, 1, arena.ctypes.data, arena.size,
                                  C.byref(out)) == -1
    assert lib.nf_plan_sic_device(7, R, 0, 0, 7, ci, None, None, eps, b2, 1, arena.ctypes.data, arena.size,
                                  C.byref(out)) == -1


def test_runtime_errors_without_setup():
    lib = nfm.lib()
    h = C.c_void_p()
    assert lib.nf_submit(C.byref(nfm.nf_sic()), 0, C.byref(h)) == -4
    assert lib.nf_submit_slab(C.byref(nfm.nf_slab_task()), 0, C.byref(h)) == -4
    assert lib.nf_submit_slab(None, 0, C.byref(h)) == -1
    assert lib.nf_get_throughput(C.byref(nfm.nf_throughput())) == -4
    assert lib.nf_reset() == -4
    assert lib.nf_wait(None) == -1
    bad = nfm.nf_setup_params(cpd=7, max_buffer_bytes=1 << 30, n_gpu=2, buffers_per_device=1, path=1)
    a, b = C.c_int32(), C.c_int32()
    assert lib.nf_setup(C.byref(bad), C.byref(a), C.byref(b)) in (-1, -5)


def test_no_cpu_fallback_when_library_missing(tmp_path):
    with pytest.raises(nfm.NFError):
        nfm.load_library(str(tmp_path / "libnfdirect.so"))


def test_noalloc_rejects_before_touching_a_gpu():
    """nf_execute_sic_noalloc checks its SIC before any HIP call: a null or
    host-resident SIC is invalid, and one whose working set exceeds the fixed
    scratch fails with NF_ERR_CAPACITY naming both sizes."""
    import synth
    lib = nfm.lib()
    assert lib.nf_execute_sic_noalloc(None, None) == -1
    box = small_uniform(cpd=7, ppc=4, seed=1)
    eps, b2, _ = params(box)
    sic = nfm.SetInteractionCollection(box, 3, 0, 7, R, eps, b2, "single_spline", mem=nfm.NF_MEM_HOST)
    assert lib.nf_execute_sic_noalloc(C.byref(sic.desc), None) == -1
    assert b"NF_MEM_DEVICE" in lib.nf_last_error()
    # 30,000 particles in every cell of a CPD 7 box: about 2.5 times the
    # scratch.  The positions are never read (lazily zeroed pages).
    cpd, per_cell = 7, 30000
    slabs = [synth.Slab(np.zeros(3 * cpd * cpd * per_cell, np.float32), np.full(cpd * cpd, per_cell, np.int32))
             for _ in range(cpd)]
    heavy = synth.Box(cpd, slabs, "heavy")
    fake = [0x1000 * (c + 1) for c in range(2 * R + 1)]  # device addresses the check never follows
    sic = nfm.SetInteractionCollection(heavy, 3, 0, 7, R, eps, b2, "single_spline", pos_ptrs=fake,
                                       acc_ptr=0x100000, mem=nfm.NF_MEM_DEVICE)
    assert lib.nf_execute_sic_noalloc(C.byref(sic.desc), None) == -3
    msg = lib.nf_last_error().decode()
    assert "scratch" in msg and str(nfm.NOALLOC_SCRATCH_BYTES) in msg
