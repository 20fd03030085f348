"""The serving loop's ring allocator (solana-perf-libs_amd/csrc/ring.h), built
for the host with g++ and driven through random allocation/release sequences
(tests/native/ring_check.cpp): blocks stay inside the ring and never overlap a
live block, an empty ring admits any size up to its capacity (the property
the loop's admission relies on), and the peak matches an interval model."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "ring_check.cpp")
HDR = os.path.join(ROOT, "solana-perf-libs_amd", "csrc", "ring.h")
OUT = os.path.join(ROOT, "tests", "native", "_build", "ring_check")


def _build():
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(SRC),
                                                              os.path.getmtime(HDR)):
        r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-o", OUT, SRC],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
    return OUT


def test_ring_allocator_random_sequences():
    exe = _build()
    for seed in (1, 2, 3):
        r = subprocess.run([exe, "3000", str(seed)], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        counts = dict(zip(r.stdout.split()[0::2], map(int, r.stdout.split()[1::2])))
        # the sequences exercise wrap-around and refusals, not only the easy path
        assert counts["grants"] > 100000 and counts["refusals"] > 1000 and counts["wraps"] > 1000
