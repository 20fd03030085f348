"""GPU parity of poh_verify_many (and its device-resident form) against the
oracle and the reference's own PoH fixture; bit-exact end hashes."""
import numpy as np
import pytest

from conftest import load_json
import workloads as W

pytestmark = pytest.mark.gpu


def test_reference_fixture(gpu_lib):
    assert gpu_lib.poh_init() is True
    fx = load_json("poh_reference.json")
    h = np.frombuffer(bytes.fromhex("".join(fx["inputs"])), np.uint8).copy()
    counts = np.array(fx["num_hashes"], np.uint64)
    for nds in (0, 1):
        x = h.copy()
        assert gpu_lib.poh_verify_many(x, counts, nds) == 0
        assert x.tobytes().hex() == "".join(fx["outputs"])


def test_empty_and_zero_counts(gpu_lib, oracle):
    x = np.zeros(0, np.uint8)
    assert gpu_lib.poh_verify_many(x, np.zeros(0, np.uint64)) == 0
    h, _ = W.poh_chains(7, 0, seed=9)
    counts = np.array([0, 1, 0, 2, 5, 0, 33], np.uint64)
    want = h.copy()
    oracle.poh_verify_many(want, counts)
    got = h.copy()
    assert gpu_lib.poh_verify_many(got, counts) == 0
    assert np.array_equal(got, want)
    assert np.array_equal(got[:32], h[:32])          # count 0: unchanged


def test_ragged_counts_vs_oracle(gpu_lib, oracle):
    rng = np.random.default_rng(4)
    n = 1000
    h, _ = W.poh_chains(n, 0, seed=10)
    counts = rng.integers(0, 700, n).astype(np.uint64)
    want = h.copy()
    oracle.poh_verify_many(want, counts, threads=8)
    got = h.copy()
    assert gpu_lib.poh_verify_many(got, counts) == 0
    assert np.array_equal(got, want)


def test_c2_shape_vs_oracle(gpu_lib, oracle):
    """The plain poh_verify_many workload at full size: 4096 chains of 12,500
    hashes from random 32-byte states.  The end hashes equal the oracle's, and
    with 5 % of the expected hashes given one flipped bit, exactly those chains
    are reported as mismatches."""
    h, counts = W.poh_chains(4096, 12500, seed=2)
    assert (counts == 12500).all()
    expected = h.copy()
    oracle.poh_verify_many(expected, counts, threads=16)
    got = h.copy()
    assert gpu_lib.poh_verify_many(got, counts) == 0
    assert np.array_equal(got, expected)

    rng = np.random.default_rng(12)
    flip = rng.random(4096) < 0.05
    flip[0] = True
    for i in np.nonzero(flip)[0]:
        expected[32 * i + int(rng.integers(0, 32))] ^= np.uint8(1 << int(rng.integers(0, 8)))
    mask = (got.reshape(-1, 32) == expected.reshape(-1, 32)).all(axis=1)
    assert np.array_equal(mask, ~flip)


def test_device_resident_form(gpu_lib, oracle):
    import hip_rt
    h, counts = W.poh_chains(3000, 37, seed=8)
    want = h.copy()
    oracle.poh_verify_many(want, counts, threads=8)
    dh = hip_rt.DeviceBuffer.from_array(h)
    dc = hip_rt.DeviceBuffer.from_array(counts)
    s = hip_rt.Stream()
    assert gpu_lib.poh_verify_many_device(dh.ptr, dc.ptr, 3000, s.handle) == 0
    s.synchronize()
    assert np.array_equal(dh.download(), want)


def test_both_kernels_vs_oracle(gpu_lib, oracle):
    """Up to 32 chains per SIMD (32,768 on 256 CUs) run on two waves per 64
    chains (rounds and message schedule split); more run one chain per lane.
    Both sides of the switch, with ragged counts including zeros, and a group
    whose one long chain makes the other 63 idle."""
    rng = np.random.default_rng(14)
    for n in (32768, 32769, 40000):
        h, _ = W.poh_chains(n, 0, seed=n)
        counts = rng.integers(0, 24, n).astype(np.uint64)
        counts[rng.random(n) < 0.05] = 0
        counts[64:128] = 0
        counts[100] = 300
        want = h.copy()
        oracle.poh_verify_many(want, counts, threads=8)
        got = h.copy()
        assert gpu_lib.poh_verify_many(got, counts) == 0
        assert np.array_equal(got, want), n
        assert np.array_equal(got[32 * 64:32 * 100], h[32 * 64:32 * 100])
