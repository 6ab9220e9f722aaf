"""bench.py's command-line contract, checked on the CPU (no GPU is touched)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True, timeout=300, cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    """The build runs on one GPU: --gpus N > 1 exits (status 2) with a message
    that says so, and prints no benchmark line."""
    for n in ("2", "8"):
        r = run("--gpus", n)
        assert r.returncode == 2, (n, r.returncode, r.stderr)
        assert "one GPU" in r.stderr and f"--gpus {n}" in r.stderr
        assert not any(line.startswith("{") for line in r.stdout.splitlines())


def test_help_lists_the_contract_options():
    """The options the benchmark contract and the measurements use."""
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--gpus", "--steps", "--warmup", "--workload", "--paths", "--planner", "--no-cpu",
                "--group-interactions", "--max-group", "--sinks-per-lane", "--waves-per-wg", "--gpu-streams",
                "--pmc-summary", "--probe", "--no-census", "--stream-slabs"):
        assert opt in r.stdout, opt


def test_default_planner_is_the_device():
    """The benchmark plans every SIC inside each timed step, on the device by
    default (nf_submit_slab); host-once, which reuses plans across steps, is
    only a comparison."""
    sys.path.insert(0, ROOT)
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_cli_probe", BENCH)
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    a = bench.parse_args([])
    assert a.planner == "device" and a.gpus == 1 and a.workload == "uniform512"
    assert bench.parse_args(["--planner", "host-once"]).planner == "host-once"
