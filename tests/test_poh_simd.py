"""CPU tests of libpoh-simd.so (the host multi-buffer PoH library that stands
in for the reference's ISPC libpoh-simd.so).

Every kernel this CPU supports is run through every entry point and checked
against the reference's own PoH fixture (src/poh-verify-test, 129 chains),
against hashlib's SHA-256 iterated by hand, and against the oracle's
poh_verify_many on random chains with uneven counts.  A count of 0 must leave
its 32 bytes untouched (src/poh-simd/poh-verify.ispc:23).
"""
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, load_json
import poh_simd


@pytest.fixture(scope="module")
def simd():
    s = poh_simd.load()
    chosen = s.kernel()              # the one the library picked by itself
    yield s
    assert s.use_kernel(chosen)


def _kernels(simd):
    ks = [k for k in poh_simd.KERNELS if simd.use_kernel(k)]
    assert "portable" in ks
    return ks


def _iterated(h, n):
    x = h.tobytes()
    for _ in range(n):
        x = hashlib.sha256(x).digest()
    return np.frombuffer(x, np.uint8)


def test_exports_match_header(simd):
    hdr = open(os.path.join(ROOT, "include", "poh_simd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9 ]+\*?\s*\b([a-z_0-9]+)\(", hdr, re.M))
    assert declared == set(poh_simd.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", poh_simd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert not declared - exported, declared - exported


def test_selection(simd):
    ks = _kernels(simd)
    assert simd.kernel() in ks
    assert not simd.use_kernel("no-such-kernel")
    for k in ks:
        assert simd.use_kernel(k) and simd.kernel() == k


@pytest.mark.parametrize("target", sorted(poh_simd.GROUPS))
def test_reference_fixture(simd, target):
    """The reference's 129-chain PoH vectors, grouped and zero-padded as the
    reference's caller does, through each entry point and kernel."""
    fx = load_json("poh_reference.json")
    inputs = np.frombuffer(bytes.fromhex("".join(fx["inputs"])), np.uint8)
    counts = np.array(fx["num_hashes"], np.uint64)
    want = "".join(fx["outputs"])
    for k in _kernels(simd):
        if k == "portable" and target != "avx512skx":
            continue              # ~7 s a pass; one entry point covers it
        simd.use_kernel(k)
        h = inputs.copy()
        simd.verify_chains(h, counts, target)
        assert h.tobytes().hex() == want, (k, target)


def test_uneven_counts_and_zero(simd):
    rng = np.random.default_rng(11)
    h0 = rng.integers(0, 256, 32 * 16, dtype=np.uint8)
    # zeros, ones, repeats, a lone long chain: every segment shape of the
    # lane kernels and every pairing of the interleaved one
    cases = [np.array([0, 1, 2, 3, 5, 8, 100, 0, 7, 7, 7, 64, 1, 33, 200, 2], np.uint64),
             np.zeros(16, np.uint64),
             np.array([0] * 15 + [77], np.uint64),
             np.array([9] * 16, np.uint64),
             np.array([1] + [0] * 15, np.uint64)]
    for counts in cases:
        want = np.concatenate([_iterated(h0[32 * i:32 * i + 32], int(c))
                               for i, c in enumerate(counts)])
        for k in _kernels(simd):
            simd.use_kernel(k)
            for target, n in poh_simd.GROUPS.items():
                h = h0.copy()
                for g in range(0, 16, n):
                    simd.group(target, h[32 * g:32 * (g + n)], counts[g:g + n])
                assert np.array_equal(h, want), (k, target, counts)


def test_unaligned_buffer(simd):
    """Chains at an odd address (the ABI takes any uint8_t pointer)."""
    rng = np.random.default_rng(12)
    raw = rng.integers(0, 256, 32 * 16 + 1, dtype=np.uint8)
    counts = rng.integers(0, 40, 16).astype(np.uint64)
    want = np.concatenate([_iterated(raw[1 + 32 * i:1 + 32 * i + 32], int(c))
                           for i, c in enumerate(counts)])
    for k in _kernels(simd):
        simd.use_kernel(k)
        buf = raw.copy()
        view = buf[1:]
        assert view.ctypes.data % 2 == 1
        poh_simd.load().lib.poh_verify_many_simd_avx512skx(view.ctypes.data,
                                                          counts.ctypes.data)
        assert buf[0] == raw[0] and np.array_equal(view, want), k


def test_many_threads_vs_oracle(simd, oracle):
    """poh_verify_many_simd over a ragged batch on several thread counts
    equals the oracle's poh_verify_many."""
    rng = np.random.default_rng(13)
    for m in (1, 15, 16, 17, 1000):
        h0 = rng.integers(0, 256, 32 * m, dtype=np.uint8)
        counts = rng.integers(0, 300, m).astype(np.uint64)
        counts[rng.random(m) < 0.1] = 0
        want = h0.copy()
        oracle.poh_verify_many(want, counts)
        for k in _kernels(simd):
            simd.use_kernel(k)
            for threads in (1, 3, 0):
                h = h0.copy()
                simd.many(h, counts, threads)
                assert np.array_equal(h, want), (k, m, threads)
    simd.many(np.zeros(0, np.uint8), np.zeros(0, np.uint64))       # empty batch
