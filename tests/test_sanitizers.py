"""The host-side C and C++ code under the address and undefined-behaviour
sanitizers: planner.cpp (nf_sic_arena_bytes, nf_plan_sic, nf_split_slab,
built without HIP) and the oracle's liboracle.so and libcpudirects.so, each
compiled with -fsanitize=address,undefined (undefined behaviour is fatal),
then the planner and oracle CPU tests run against those builds in a child
process with libasan preloaded; and hostcopy.cpp's threaded copy (the copy of
pageable slabs into pinned staging) checked byte for byte over sizes and
thread counts; and the work lists' merge of tiles and short tails
(interleave_pos) checked to be an even permutation.  A sanitizer report
fails the test.
CPU only: nothing here touches a GPU."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, (cmd, r.stdout[-3000:], r.stderr[-3000:])
    return r


@pytest.fixture(scope="module")
def sanitized_libs(tmp_path_factory):
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("gcc/g++ not found")
    out = tmp_path_factory.mktemp("sanitized")
    csrc = os.path.join(ROOT, "abacus-public_amd", "csrc")
    _run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", *SAN, "-I", os.path.join(ROOT, "include"), "-I", csrc,
          os.path.join(csrc, "planner.cpp"), os.path.join(ROOT, "tests", "sanitize", "fail_stub.cpp"),
          "-o", str(out / "libnfplan_san.so")])
    # the oracle's own flags (oracle/Makefile), plus the sanitizers
    _run(["gcc", "-O2", "-fPIC", "-shared", "-fno-fast-math", "-ffp-contract=off", "-fopenmp", "-std=c11", *SAN,
          os.path.join(ROOT, "oracle", "oracle.c"), "-o", str(out / "liboracle.so")])
    _run(["gcc", "-Ofast", "-march=x86-64-v4", "-mprefer-vector-width=512", "-fPIC", "-shared", "-fopenmp",
          "-std=c11", *SAN, os.path.join(ROOT, "oracle", "cpu_directs.c"), "-o", str(out / "libcpudirects.so")])
    return str(out)


def test_planner_and_oracle_under_sanitizers(sanitized_libs):
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("libasan.so not found")
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": asan,
        "NF_SANITIZED_LIBDIR": sanitized_libs,
        # Python's own allocations outlive the interpreter: leak reports
        # would be about CPython, not the code under test
        "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=0:halt_on_error=1",
        "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1",
        "PYTHONPATH": os.pathsep.join([os.path.join(ROOT, "tests", "sanitize"), env.get("PYTHONPATH", "")]),
        "OMP_NUM_THREADS": "4",
    })
    # the codegen check reads the vector width of an uninstrumented build
    r = subprocess.run([sys.executable, "-m", "pytest", "-p", "san_plugin", "-q", "-m", "not gpu",
                        "-p", "no:cacheprovider", os.path.join(ROOT, "tests", "test_planner.py"),
                        os.path.join(ROOT, "tests", "test_oracle.py"), "--deselect",
                        "tests/test_oracle.py::test_cpu_directs_are_16_wide_with_refined_rsqrt14"],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    report = r.stdout[-4000:] + r.stderr[-4000:]
    loaded = [ln for ln in r.stdout.splitlines() if ln.startswith("sanitized libraries loaded:")]
    assert loaded and set(loaded[0].split(":")[1].split()) == {"libnfplan_san.so", "liboracle.so",
                                                              "libcpudirects.so"}, report
    assert "AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, report
    assert r.returncode == 0, report
    assert " passed" in r.stdout, report


def test_threaded_host_copy_covers_every_byte(tmp_path):
    """par_memcpy_threads copies every byte and nothing past the end, for
    sizes 0-4096 over 1-9 threads and slab-sized copies over 2-8 threads.
    (Chunks of floor(bytes / threads) rounded up to 64 bytes once left the
    last bytes of a 12 x 3,900,459-byte slab uncopied at 8 threads.)"""
    if not shutil.which("g++"):
        pytest.skip("g++ not found")
    csrc = os.path.join(ROOT, "abacus-public_amd", "csrc")
    exe = str(tmp_path / "copy_check")
    _run(["g++", "-std=c++17", "-O1", "-pthread", *SAN, "-I", os.path.join(ROOT, "include"), "-I", csrc,
          os.path.join(csrc, "hostcopy.cpp"), os.path.join(ROOT, "tests", "sanitize", "copy_check.cpp"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="halt_on_error=1"))
    assert r.returncode == 0 and "copied exactly" in r.stdout, (r.stdout[-2000:], r.stderr[-2000:])


def test_work_list_interleave_is_an_even_permutation(tmp_path):
    """interleave_pos, which merges a work list's run of tiles with its run of
    short tails (host lists, the no-allocation entry point's list and the
    device planner's), places every item exactly once, keeps each run in its
    cost order and spreads the rarer kind evenly, for every M, T up to 300
    and launch-group sizes."""
    if not shutil.which("g++"):
        pytest.skip("g++ not found")
    csrc = os.path.join(ROOT, "abacus-public_amd", "csrc")
    exe = str(tmp_path / "interleave_check")
    _run(["g++", "-std=c++17", "-O1", *SAN, "-I", os.path.join(ROOT, "include"), "-I", csrc,
          os.path.join(ROOT, "tests", "sanitize", "interleave_check.cpp"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, UBSAN_OPTIONS="halt_on_error=1"))
    assert r.returncode == 0 and "even permutation" in r.stdout, (r.stdout[-2000:], r.stderr[-2000:])
