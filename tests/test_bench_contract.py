"""bench.py's command-line contract and byte accounting, on the CPU: one GPU
only (a request for more exits with a message and status 2), the stream
option's choices, and the algorithmic bytes per amplitude of an order."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.parametrize("streams", ["0", "3"])
def test_stream_choice_is_one_or_two(streams):
    r = run_bench("--streams", streams)
    assert r.returncode == 2
    assert "--streams" in r.stderr


def test_algorithmic_bytes_per_order():
    # read T_{n-1} and T_{n-2} (16 + 16), the diagonal (8), read and write
    # psi (16 + 16), write T_n (16)
    assert bench.BYTES_PER_AMP_ORDER == 16 + 16 + 8 + 16 + 16 + 16
    n = 1 << 18
    assert bench.BYTES_PER_AMP_ORDER * n == 23_068_672
    # a step of K orders: order 1 (psi -> T_1), order 2 (fused psi update),
    # then K - 2 general orders
    k = 13
    assert bench.step_bytes(k, n) == (bench.BYTES_FIRST + bench.BYTES_SECOND
                                      + (k - 2) * bench.BYTES_PER_AMP_ORDER) * n


def test_every_priced_arrangement_names_a_committed_profile():
    """Each (order kernels, set size) the line can price names the committed
    profile that reproduces its time and traffic, and the default set size
    is one of them."""
    default_set = next(a.default for a in bench.build_parser()._actions if a.dest == "set")
    assert (bench.WHT, default_set) in bench.TRACE_PROFILE
    assert (bench.WHT, default_set, default_set) in bench.MEASURED_TRAFFIC
    for key, path in bench.TRACE_PROFILE.items():
        assert os.path.isfile(os.path.join(ROOT, path)), (key, path)
        assert (key[0], key[1], key[1]) in bench.MEASURED_TRAFFIC, key
