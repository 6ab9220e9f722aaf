"""Python host mirror of the near-field driver over the C ABI (include/nf_direct.h).

The classes follow the reference's host interface:
  * ``NearFieldDriver``  -- NearFieldDriver (directdriver.cpp:22-95): eps per
    softening form (:105-117), GPU setup (:137-202), ExecuteSlab (:395-409),
    SlabDone (:414-431), Finalize (:434-492).
  * ``SetInteractionCollection`` -- a planned SIC (PencilOnPencil.cc:180-418)
    held as an ``nf_sic`` descriptor plus the arena its arrays live in.

The compute path is libnfdirect.so (HIP for gfx950).  Nothing here falls back
to the CPU: if the library is missing, importing this module raises.
"""
import ctypes as C
import os

import numpy as np

from synth import CELLINFO_DTYPE, SOFTENINGS

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnfdirect.so")

NF_MAX_WIDTH = 5
NF_BLOCK_SIZE = 64
NF_MEM_HOST, NF_MEM_DEVICE = 0, 1
NF_PATH_PLAIN, NF_PATH_BATCHED = 0, 1
PATHS = {"plain": NF_PATH_PLAIN, "batched": NF_PATH_BATCHED}


class NFError(RuntimeError):
    pass


class nf_cell_plan(C.Structure):
    _fields_ = [("start", C.c_uint32), ("N", C.c_int32)]


class nf_sink_plan(C.Structure):
    _fields_ = [("cell", nf_cell_plan * NF_MAX_WIDTH), ("wrapcell", C.c_int32)]


class nf_source_plan(C.Structure):
    _fields_ = [("cell", nf_cell_plan * NF_MAX_WIDTH)]


_P = C.c_void_p


class nf_sic(C.Structure):
    _fields_ = [
        ("SlabId", C.c_int32), ("j_low", C.c_int32), ("j_high", C.c_int32), ("k_low", C.c_int32),
        ("k_high", C.c_int32), ("Nk", C.c_int32), ("j_width", C.c_int32),
        ("cpd", C.c_int32), ("nfradius", C.c_int32), ("nfwidth", C.c_int32),
        ("eps", C.c_float), ("b2", C.c_float), ("softening", C.c_int32), ("mem", C.c_int32),
        ("SinkPosSlab", _P), ("SourcePosSlab", _P * NF_MAX_WIDTH), ("Nslab", C.c_uint64 * NF_MAX_WIDTH),
        ("SinkAccSlab", _P),
        ("NSinkSets", C.c_int32), ("SinkSetStart", _P), ("SinkSetCount", _P), ("SinkSetIdMax", _P),
        ("SinkPlan", _P),
        ("NSourceSets", C.c_int32), ("SourceSetStart", _P), ("SourceSetCount", _P), ("SourcePlan", _P),
        ("NSinkBlocks", C.c_int32), ("NSourceBlocks", C.c_int32), ("InteractionCount", C.c_int32),
        ("SinkBlockParentPencil", _P), ("SinkSourceInteractionList", _P), ("SinkSourceYOffset", _P),
        ("DirectTotal", C.c_uint64), ("PaddedDirectTotal", C.c_uint64), ("SinkTotal", C.c_uint64),
        ("SourceTotal", C.c_uint64), ("PaddedSinkTotal", C.c_uint64), ("PaddedSourceTotal", C.c_uint64),
        ("pos_layout", C.c_int32), ("reserved", C.c_int32), ("kick", _P),
    ]


class nf_kick(C.Structure):
    _fields_ = [("FarAccSlab", _P), ("rescale", C.c_double), ("inv_eps3", C.c_double),
                ("density_rad2", C.c_double), ("NearAccOut", _P)]


NF_POS_XYZ, NF_POS_AOS = 0, 1


class nf_task_stats(C.Structure):
    _fields_ = [
        ("device_thread_s", C.c_double), ("fill_s", C.c_double), ("kernel_s", C.c_double),
        ("direct_kernel_s", C.c_double), ("copyback_s", C.c_double),
        ("bytes_to_device", C.c_uint64), ("bytes_from_device", C.c_uint64),
        ("direct_total", C.c_uint64), ("padded_direct_total", C.c_uint64),
        ("assigned_buffer", C.c_int32), ("path", C.c_int32), ("group_size", C.c_int32),
        ("group_direct_kernel_s", C.c_double), ("group_direct_total", C.c_uint64), ("graph", C.c_int32),
        ("upload_s", C.c_double), ("download_s", C.c_double), ("slabs_uploaded", C.c_int32),
        ("slabs_reused", C.c_int32), ("sinks_per_lane", C.c_int32), ("autotune_trial", C.c_int32),
        ("waves_per_eu", C.c_int32), ("device_planned", C.c_int32), ("plan_s", C.c_double),
        ("waves_per_wg", C.c_int32), ("group_form_interactions", C.c_uint64 * 3),
        ("group_tail_interactions", C.c_uint64),
    ]


class nf_throughput(C.Structure):
    _fields_ = [("gpu_active_s", C.c_double), ("direct_total", C.c_uint64), ("padded_direct_total", C.c_uint64),
                ("bytes_to_device", C.c_uint64), ("bytes_from_device", C.c_uint64), ("tasks", C.c_uint64),
                ("groups", C.c_uint64)]


class nf_slab_task(C.Structure):
    _fields_ = [
        ("SlabId", C.c_int32), ("j_low", C.c_int32), ("j_high", C.c_int32), ("cpd", C.c_int32),
        ("nfradius", C.c_int32), ("eps", C.c_float), ("b2", C.c_float), ("softening", C.c_int32),
        ("mem", C.c_int32), ("pos_layout", C.c_int32), ("cellinfo", _P * NF_MAX_WIDTH),
        ("PosSlab", _P * NF_MAX_WIDTH), ("Nslab", C.c_uint64 * NF_MAX_WIDTH), ("SinkAccSlab", _P),
        ("kick", _P),
    ]


class nf_setup_params(C.Structure):
    _fields_ = [
        ("cpd", C.c_int32), ("max_buffer_bytes", C.c_uint64), ("n_gpu", C.c_int32),
        ("buffers_per_device", C.c_int32), ("use_pinned", C.c_int32), ("path", C.c_int32),
        ("sinks_per_lane", C.c_int32), ("source_tile", C.c_int32), ("gpu_streams", C.c_int32),
        ("max_group", C.c_int32), ("use_graphs", C.c_int32), ("group_interactions", C.c_uint64),
        ("waves_per_wg", C.c_int32), ("census", C.c_int32),
    ]


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise NFError(f"{path} is missing: build it with `make -C abacus-public_amd` "
                      "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(path)
    I32, U64 = C.c_int32, C.c_uint64
    lib.nf_last_error.restype = C.c_char_p
    lib.nf_get_ngpu.argtypes = [C.POINTER(I32)]
    lib.nf_get_device_info.argtypes = [C.POINTER(U64), C.c_char_p, C.c_size_t]
    lib.nf_setup.argtypes = [C.POINTER(nf_setup_params), C.POINTER(I32), C.POINTER(I32)]
    lib.nf_reset.argtypes = []
    lib.nf_submit.argtypes = [C.POINTER(nf_sic), I32, C.POINTER(C.c_void_p)]
    lib.nf_check_completion.argtypes = [C.c_void_p]
    lib.nf_wait.argtypes = [C.c_void_p]
    lib.nf_get_task_stats.argtypes = [C.c_void_p, C.POINTER(nf_task_stats)]
    lib.nf_release.argtypes = [C.c_void_p]
    lib.nf_sic_arena_bytes.argtypes = [I32, I32, I32, I32, U64]
    lib.nf_sic_arena_bytes.restype = U64
    lib.nf_plan_sic.argtypes = [I32, I32, I32, I32, I32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                C.POINTER(U64), C.c_float, C.c_float, I32, C.c_void_p, U64,
                                C.POINTER(nf_sic)]
    lib.nf_plan_sic_device.argtypes = lib.nf_plan_sic.argtypes
    lib.nf_split_slab.argtypes = [I32, I32, C.POINTER(C.c_void_p), C.POINTER(U64), I32, I32, I32,
                                  C.POINTER(I32), C.POINTER(I32)]
    lib.nf_count_interactions.argtypes = [I32, I32, I32, I32, C.POINTER(C.c_void_p), C.POINTER(U64)]
    lib.nf_submit_slab.argtypes = [C.POINTER(nf_slab_task), I32, C.POINTER(C.c_void_p)]
    lib.nf_device_alloc.argtypes = [U64, C.POINTER(C.c_void_p)]
    lib.nf_device_free.argtypes = [C.c_void_p]
    lib.nf_device_copy.argtypes = [C.c_void_p, C.c_void_p, U64, I32]
    lib.nf_device_synchronize.argtypes = []
    lib.nf_host_alloc.argtypes = [U64, C.POINTER(C.c_void_p)]
    lib.nf_host_free.argtypes = [C.c_void_p]
    lib.nf_autotune_table.argtypes = [C.c_char_p, C.c_size_t]
    lib.nf_get_throughput.argtypes = [C.POINTER(nf_throughput)]
    lib.nf_stream_begin.argtypes = []
    lib.nf_stream_end.argtypes = []
    lib.nf_execute_sic_noalloc.argtypes = [C.POINTER(nf_sic), C.c_void_p]
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def check(rc, what):
    if rc != 0:
        msg = lib().nf_last_error()
        raise NFError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _view(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=dtype, count=n)


class DeviceArray:
    """A device allocation holding a copy of a numpy array's bytes (HBM-resident
    slab or acceleration slab), through the module's own HIP runtime."""

    def __init__(self, nbytes=0, host=None):
        if host is not None:
            host = np.ascontiguousarray(host)
            nbytes = host.nbytes
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(lib().nf_device_alloc(self.nbytes, C.byref(p)), "nf_device_alloc")
        self.ptr = p.value
        if host is not None and self.nbytes:
            check(lib().nf_device_copy(self.ptr, host.ctypes.data, self.nbytes, 0), "nf_device_copy")

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.nbytes
        if self.nbytes:
            check(lib().nf_device_copy(out.ctypes.data, self.ptr, self.nbytes, 1), "nf_device_copy")
        return out

    def zero(self):
        z = np.zeros(self.nbytes, np.uint8)
        if self.nbytes:
            check(lib().nf_device_copy(self.ptr, z.ctypes.data, self.nbytes, 0), "nf_device_copy")

    def free(self):
        if self.ptr:
            check(lib().nf_device_free(self.ptr), "nf_device_free")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """A numpy array in pinned host memory (nf_host_alloc): host-mode slabs the
    DMA engines read and write directly.  ``.array`` is the numpy view; keep
    this object alive while the view is in use."""

    def __init__(self, shape, dtype):
        dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(shape)) * dtype.itemsize
        p = C.c_void_p()
        check(lib().nf_host_alloc(self.nbytes, C.byref(p)), "nf_host_alloc")
        self.ptr = p.value
        buf = (C.c_char * max(self.nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            check(lib().nf_host_free(self.ptr), "nf_host_free")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_synchronize():
    check(lib().nf_device_synchronize(), "nf_device_synchronize")


# Device scratch of nf_execute_sic_noalloc (NF_NOALLOC_SCRATCH_BYTES).
NOALLOC_SCRATCH_BYTES = 256 << 20


def execute_sic_noalloc(sic, stream=None):
    """nf_execute_sic_noalloc: run one planned, device-resident SIC on the
    batched path with the module's fixed device scratch (no allocation, no
    nf_setup needed); returns when its accelerations are in place."""
    check(lib().nf_execute_sic_noalloc(C.byref(sic.desc), stream), "nf_execute_sic_noalloc")


class SetInteractionCollection:
    """A planned SIC: the ``nf_sic`` descriptor and the arena behind it."""

    def __init__(self, box, slab, j_low, j_high, nfradius, eps, b2, softening,
                 pos_ptrs=None, acc_ptr=None, mem=NF_MEM_HOST, pos_layout=NF_POS_XYZ, kick=None,
                 planner="host"):
        """planner "host": nf_plan_sic; "device": nf_plan_sic_device (the
        device planner nf_submit_slab uses, arrays copied back)."""
        cpd = box.cpd
        W = 2 * nfradius + 1
        slabs = [box.slab(slab - nfradius + c) for c in range(W)]
        self._slabs = slabs
        ci = (C.c_void_p * W)(*[s.cellinfo.ctypes.data for s in slabs])
        if pos_ptrs is None:
            pos_ptrs = [s.pos.ctypes.data for s in slabs]
        pos = (C.c_void_p * W)(*pos_ptrs)
        nsl = (C.c_uint64 * W)(*[s.n for s in slabs])
        maxn = max(s.n for s in slabs)
        nbytes = lib().nf_sic_arena_bytes(cpd, nfradius, j_low, j_high, maxn)
        self.arena = np.zeros(int(nbytes), np.uint8)
        self.desc = nf_sic()
        fn = {"host": "nf_plan_sic", "device": "nf_plan_sic_device"}[planner]
        check(getattr(lib(), fn)(cpd, nfradius, slab % cpd, j_low, j_high, ci, pos, nsl, float(eps),
                                 float(b2), SOFTENINGS[softening] if isinstance(softening, str) else softening,
                                 self.arena.ctypes.data, nbytes, C.byref(self.desc)), fn)
        self.desc.mem = mem
        if acc_ptr is not None:
            self.desc.SinkAccSlab = acc_ptr
        self.desc.pos_layout = pos_layout
        self.kick = None
        if kick is not None:
            # nf_kick(FarAccSlab=pointer, rescale=, inv_eps3=, density_rad2=); kept alive here
            self.kick = kick
            self.desc.kick = C.addressof(kick)

    # Views of the planned arrays (for tests and statistics).
    def arrays(self):
        d = self.desc
        ns, nr, nb, ic = d.NSinkSets, d.NSourceSets, d.NSinkBlocks, d.InteractionCount
        W = d.nfwidth
        sp = _view(d.SinkPlan, ns * 11, np.int32).reshape(ns, 11)
        rp = _view(d.SourcePlan, nr * 10, np.int32).reshape(nr, 10)
        return dict(
            SinkSetStart=_view(d.SinkSetStart, ns, np.int32).copy(),
            SinkSetCount=_view(d.SinkSetCount, ns, np.int32).copy(),
            SinkSetIdMax=_view(d.SinkSetIdMax, ns, np.uint32).copy(),
            SinkPlanStart=sp[:, 0:2 * W:2].astype(np.uint32), SinkPlanN=sp[:, 1:2 * W:2].copy(),
            WrapCell=sp[:, 10].copy(),
            SourceSetStart=_view(d.SourceSetStart, nr, np.int32).copy(),
            SourceSetCount=_view(d.SourceSetCount, nr, np.int32).copy(),
            SourcePlanStart=rp[:, 0:2 * W:2].astype(np.uint32), SourcePlanN=rp[:, 1:2 * W:2].copy(),
            SinkBlockParentPencil=_view(d.SinkBlockParentPencil, nb, np.int32).copy(),
            SinkSourceInteractionList=_view(d.SinkSourceInteractionList, ic, np.int32).copy(),
            SinkSourceYOffset=_view(d.SinkSourceYOffset, ic, np.float32).copy(),
            NSinkSets=ns, NSourceSets=nr, NSinkBlocks=nb, NSourceBlocks=d.NSourceBlocks,
            InteractionCount=ic, DirectTotal=d.DirectTotal, PaddedDirectTotal=d.PaddedDirectTotal,
            SinkTotal=d.SinkTotal, SourceTotal=d.SourceTotal, PaddedSinkTotal=d.PaddedSinkTotal,
            PaddedSourceTotal=d.PaddedSourceTotal, Nk=d.Nk, j_width=d.j_width,
        )


class Task:
    def __init__(self, handle, sic):
        self.h = handle
        self.sic = sic  # keeps the plan arena alive until completion

    def done(self):
        rc = lib().nf_check_completion(self.h)
        if rc < 0:
            check(rc, "task")
        return rc == 1

    def wait(self):
        check(lib().nf_wait(self.h), "nf_wait")

    def stats(self):
        s = nf_task_stats()
        check(lib().nf_get_task_stats(self.h, C.byref(s)), "nf_get_task_stats")
        d = {k: getattr(s, k) for k, _ in s._fields_}
        d["group_form_interactions"] = list(d["group_form_interactions"])
        return d

    def release(self):
        if self.h:
            check(lib().nf_release(self.h), "nf_release")
            self.h = None


class NearFieldDriver:
    """Near-field driver over the GPU module: plans SICs and submits them.

    ``softening_length`` is the Plummer-equivalent length in box units
    (SofteningLength/BoxSize); the internal length and eps follow
    singlestep.cpp:113-125 and directdriver.cpp:105-117.
    """

    _active = None

    def __init__(self, cpd, softening_length, b2, nfradius=2, softening="single_spline",
                 path="batched", buffers_per_device=3, max_buffer_bytes=None,
                 sinks_per_lane=0, gpu_streams=0, max_group=0, use_graphs=0, group_interactions=0,
                 waves_per_wg=0, census=0):
        from synth import eps_param
        if NearFieldDriver._active is not None:
            raise NFError("one NearFieldDriver per process (the GPU module is a singleton)")
        self.cpd, self.R, self.W = cpd, nfradius, 2 * nfradius + 1
        self.softening = softening
        self.eps, self.softening_internal = eps_param(softening, softening_length)
        self.b2 = np.float32(b2)
        self.path = path
        if max_buffer_bytes is None:
            free = C.c_uint64(0)
            name = C.create_string_buffer(64)
            check(lib().nf_get_device_info(C.byref(free), name, 64), "nf_get_device_info")
            max_buffer_bytes = int(free.value) // max(1, buffers_per_device)
        p = nf_setup_params(cpd=cpd, max_buffer_bytes=int(max_buffer_bytes), n_gpu=1,
                            buffers_per_device=buffers_per_device, use_pinned=-1, path=PATHS[path],
                            sinks_per_lane=sinks_per_lane, source_tile=0, gpu_streams=gpu_streams,
                            max_group=max_group, use_graphs=use_graphs,
                            group_interactions=int(group_interactions), waves_per_wg=waves_per_wg,
                            census=census)
        msb, mrb = C.c_int32(0), C.c_int32(0)
        check(lib().nf_setup(C.byref(p), C.byref(msb), C.byref(mrb)), "nf_setup")
        NearFieldDriver._active = self
        self.buffers_per_device = buffers_per_device
        self.max_sink_blocks, self.max_source_blocks = msb.value, mrb.value
        self.slab_tasks = {}

    def plan(self, box, slab, j_low=0, j_high=None, pos_ptrs=None, acc_ptr=None, mem=NF_MEM_HOST,
             pos_layout=NF_POS_XYZ, kick=None):
        if j_high is None:
            j_high = self.cpd
        return SetInteractionCollection(box, slab, j_low, j_high, self.R, self.eps, self.b2, self.softening,
                                        pos_ptrs=pos_ptrs, acc_ptr=acc_ptr, mem=mem, pos_layout=pos_layout,
                                        kick=kick)

    def submit(self, sic, blocking=False):
        h = C.c_void_p()
        check(lib().nf_submit(C.byref(sic.desc), 0, C.byref(h)), "nf_submit")
        t = Task(h, sic)
        if blocking:
            try:
                t.wait()
            except NFError:
                t.release()
                raise
        return t

    def submit_slab(self, box, slab, j_low=0, j_high=None, acc_ptr=None, pos_ptrs=None, mem=NF_MEM_HOST,
                    pos_layout=NF_POS_XYZ, kick=None, blocking=False):
        """nf_submit_slab: the SIC of ``slab``, rows [j_low, j_high), planned on
        the device from the box's cellinfo slabs and executed."""
        if j_high is None:
            j_high = self.cpd
        slabs = [box.slab(slab - self.R + c) for c in range(self.W)]
        t = nf_slab_task(SlabId=slab % self.cpd, j_low=j_low, j_high=j_high, cpd=self.cpd, nfradius=self.R,
                         eps=float(self.eps), b2=float(self.b2), softening=SOFTENINGS[self.softening] if isinstance(self.softening, str) else self.softening,
                         mem=mem, pos_layout=pos_layout, SinkAccSlab=acc_ptr)
        for c, sl in enumerate(slabs):
            t.cellinfo[c] = sl.cellinfo.ctypes.data
            t.PosSlab[c] = sl.pos.ctypes.data if pos_ptrs is None else pos_ptrs[c]
            t.Nslab[c] = sl.n
        if kick is not None:
            t.kick = C.addressof(kick)
        h = C.c_void_p()
        check(lib().nf_submit_slab(C.byref(t), 0, C.byref(h)), "nf_submit_slab")
        task = Task(h, (slabs, kick))  # the cellinfo, slabs and kick stay alive until completion
        if blocking:
            try:
                task.wait()
            except NFError:
                task.release()
                raise
        return task

    def split_slab(self, box, slab):
        """ExecuteSlabGPU's y-split of a slab (nf_split_slab): the end row of
        each SIC, for this driver's worker count and block maxima."""
        W = self.W
        slabs = [box.slab(slab - self.R + c) for c in range(W)]
        ci = (C.c_void_p * W)(*[s.cellinfo.ctypes.data for s in slabs])
        nsl = (C.c_uint64 * W)(*[s.n for s in slabs])
        out = (C.c_int32 * self.cpd)()
        n = C.c_int32(0)
        check(lib().nf_split_slab(self.cpd, self.R, ci, nsl, self.buffers_per_device, self.max_sink_blocks,
                                  self.max_source_blocks, out, C.byref(n)), "nf_split_slab")
        return [int(out[i]) for i in range(n.value)]

    def execute_slab(self, box, slab, acc, blocking=False, nsplit=1, planner="host"):
        """ExecuteSlab for a host-resident box: plan the slab as SICs over y
        (directdriver.cpp:242-391) and submit them.  ``nsplit``: a number of
        equal row ranges, or "reference" for ExecuteSlabGPU's own split
        (split_slab).  ``acc`` is the float32 (n, 4) acceleration array of the
        slab, filled on completion.  planner "device": each split is planned on
        the device (nf_submit_slab)."""
        cpd = self.cpd
        if nsplit == "reference":
            bounds = np.array([0] + self.split_slab(box, slab))
        else:
            bounds = np.linspace(0, cpd, nsplit + 1).round().astype(int)
        tasks = []
        for jl, jh in zip(bounds[:-1], bounds[1:]):
            if jh <= jl:
                continue
            if planner == "device":
                tasks.append(self.submit_slab(box, slab, int(jl), int(jh), acc_ptr=acc.ctypes.data,
                                              blocking=blocking))
                continue
            sic = self.plan(box, slab, int(jl), int(jh), acc_ptr=acc.ctypes.data)
            tasks.append(self.submit(sic, blocking))
        self.slab_tasks[slab % cpd] = tasks
        return tasks

    def slab_done(self, slab):
        return all(t.done() for t in self.slab_tasks.get(slab % self.cpd, []))

    def finalize(self, slab):
        tasks = self.slab_tasks.pop(slab % self.cpd, [])
        stats = []
        for t in tasks:
            t.wait()
            stats.append(t.stats())
            t.release()
        return stats

    def stream_begin(self):
        """Host slabs stay unchanged until stream_end: reuse their HBM copies."""
        check(lib().nf_stream_begin(), "nf_stream_begin")

    def stream_end(self):
        check(lib().nf_stream_end(), "nf_stream_end")

    def throughput(self):
        """The GPU throughput counters since setup, with the reference's
        "effective GDIPS" (giga direct interactions per GPU-active second)."""
        t = nf_throughput()
        check(lib().nf_get_throughput(C.byref(t)), "nf_get_throughput")
        d = {k: getattr(t, k) for k, _ in t._fields_}
        d["gdips"] = d["direct_total"] / d["gpu_active_s"] / 1e9 if d["gpu_active_s"] > 0 else 0.0
        return d

    def autotune_table(self):
        """The autotuner's table: occupancy histogram -> chosen tile shape."""
        import json
        n = lib().nf_autotune_table(None, 0)
        check(min(n, 0), "nf_autotune_table")
        buf = C.create_string_buffer(n + 1)
        lib().nf_autotune_table(buf, n + 1)
        return json.loads(buf.value.decode())

    def close(self):
        if NearFieldDriver._active is self:
            check(lib().nf_reset(), "nf_reset")
            NearFieldDriver._active = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def count_interactions(box, slab, j_low=0, j_high=None, nfradius=2):
    """nf_count_interactions: the pair interactions (DirectTotal) of the SIC of
    ``slab``, rows [j_low, j_high), from the box's cellinfo counts alone."""
    if j_high is None:
        j_high = box.cpd
    W = 2 * nfradius + 1
    ci = (C.c_void_p * W)(*[box.slab(slab - nfradius + c).cellinfo.ctypes.data for c in range(W)])
    out = C.c_uint64()
    check(lib().nf_count_interactions(box.cpd, nfradius, j_low, j_high, ci, C.byref(out)), "nf_count_interactions")
    return int(out.value)
