"""pytest plugin (loaded with -p san_plugin) that points the CPU tests at
sanitized builds: the planner entry points (nf_sic_arena_bytes, nf_plan_sic,
nf_split_slab, nf_count_interactions) from a host-only build of planner.cpp, and the oracle
libraries, all compiled with -fsanitize=address,undefined into the folder
NF_SANITIZED_LIBDIR names.  The process runs with libasan preloaded."""
import ctypes as C
import os
import sys

_DIR = os.environ["NF_SANITIZED_LIBDIR"]
_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for sub in ("abacus-public_amd", "oracle"):
    sys.path.insert(0, os.path.join(_ROOT, sub))

import nearfield as nfm  # noqa: E402
import pyoracle as po  # noqa: E402


def _planner_library(path):
    lib = C.CDLL(path)
    I32, U64 = C.c_int32, C.c_uint64
    lib.nf_last_error.restype = C.c_char_p
    lib.nf_sic_arena_bytes.argtypes = [I32, I32, I32, I32, U64]
    lib.nf_sic_arena_bytes.restype = U64
    lib.nf_plan_sic.argtypes = [I32, I32, I32, I32, I32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                C.POINTER(U64), C.c_float, C.c_float, I32, C.c_void_p, U64,
                                C.POINTER(nfm.nf_sic)]
    lib.nf_split_slab.argtypes = [I32, I32, C.POINTER(C.c_void_p), C.POINTER(U64), I32, I32, I32,
                                  C.POINTER(I32), C.POINTER(I32)]
    lib.nf_count_interactions.argtypes = [I32, I32, I32, I32, C.POINTER(C.c_void_p), C.POINTER(U64)]
    return lib


nfm._lib = _planner_library(os.path.join(_DIR, "libnfplan_san.so"))
po._HERE = _DIR


def pytest_terminal_summary(terminalreporter):
    """Names the sanitized libraries this process mapped, from /proc/self/maps."""
    with open("/proc/self/maps") as f:
        maps = f.read()
    names = sorted({os.path.basename(line.split()[-1]) for line in maps.splitlines()
                    if line.split()[-1].startswith(_DIR)})
    terminalreporter.write_line("sanitized libraries loaded: " + " ".join(names))
