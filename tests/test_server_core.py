"""The serving loop's scheduling core (solana-perf-libs_amd/csrc/server_core.h)
on the CPU: tests/native/server_check.cpp drives it over a device double
(worker threads playing the slot streams, a host buffer playing the device
ring) through whole server lifetimes: two threads submitting verify and PoH
requests of random sizes, offsets past their packets, zero counts, empty and
oversized requests, requests whose packets the "device" reads in place, a
waiter with a time limit, stop() before or after the waits, and injected
device failures (a step that fails, a launch that fails
after enqueuing work).  Every result is checked exactly against the double's
stand-in kernels.

Built twice: with ThreadSanitizer (clang++; GCC 11's runtime does not
intercept pthread_cond_clockwait, which std::condition_variable::wait_for
uses, and reports false races), and optimised without it for more cycles."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "server_check.cpp")
HDRS = [os.path.join(ROOT, "solana-perf-libs_amd", "csrc", h) for h in ("server_core.h", "ring.h")]
BUILD = os.path.join(ROOT, "tests", "native", "_build")
CLANG = "/opt/rocm/llvm/bin/clang++"


def _build(name, cmd):
    out = os.path.join(BUILD, name)
    os.makedirs(BUILD, exist_ok=True)
    newest = max(os.path.getmtime(p) for p in [SRC] + HDRS)
    if not os.path.exists(out) or os.path.getmtime(out) < newest:
        r = subprocess.run(cmd + ["-o", out, SRC, "-lpthread"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
    return out


def _run(exe, cycles, seed, timeout):
    r = subprocess.run([exe, str(cycles), str(seed)], capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, (r.returncode, r.stderr[-4000:])
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    f = r.stdout.split()
    return dict(zip(f[0::2], map(int, f[1::2])))


def test_server_core_threadsanitizer():
    if not os.path.exists(CLANG):
        pytest.skip("clang++ with ThreadSanitizer not found")
    exe = _build("server_check_tsan", [CLANG, "-O1", "-g", "-std=c++17", "-fsanitize=thread"])
    total = {}
    for seed in (11, 12):
        c = _run(exe, 300, seed, 600)
        for k, v in c.items():
            total[k] = total.get(k, 0) + v
    # the cycles reach every path: refusals, device failures of both kinds,
    # many batches of both kinds
    assert total["requests"] > 20000 and total["refused"] > 500
    assert total["injected_step_failures"] > 50 and total["injected_launch_failures"] > 10
    assert total["failed"] > 0 and total["chains"] > 50000 and total["signatures"] > 100000
    assert total["read_in_place"] > 5000


def test_server_core_optimised():
    cxx = shutil.which("g++")
    exe = _build("server_check", [cxx, "-O2", "-std=c++17", "-Wall", "-Wextra"])
    c = _run(exe, 800, 21, 600)
    assert c["requests"] > 30000 and c["read_in_place"] > 5000
