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


def _simds():
    import hip_rt
    return 4 * hip_rt.device_attribute(hip_rt.HIP_ATTR_CUS)


def _ragged(n, rng, seed):
    h, _ = W.poh_chains(n, 0, seed=seed)
    counts = rng.integers(0, 24, n).astype(np.uint64)
    counts[rng.random(n) < 0.05] = 0
    return h, counts


def test_kernel_switches_vs_oracle(gpu_lib, oracle):
    """poh_verify_many picks the pair kernel up to 8 chains per SIMD, the split
    kernel up to 32 and one chain per lane beyond: both sides of each switch,
    with ragged counts including zeros, and a group whose one long chain makes
    the others idle (they keep their end states)."""
    rng = np.random.default_rng(14)
    simds = _simds()
    for n, kernel in ((8 * simds, "pair"), (8 * simds + 1, "split"), (32 * simds, "split"),
                      (32 * simds + 1, "chain"), (40 * simds, "chain")):
        assert gpu_lib.poh_kernel_choice(n) == kernel, n
        h, counts = _ragged(n, rng, seed=n)
        counts[64:128] = 0
        counts[100] = 300
        want = h.copy()
        oracle.poh_verify_many(want, counts, threads=8)
        got = h.copy()
        assert gpu_lib.poh_verify_many(got, counts) == 0
        assert np.array_equal(got, want), n
        assert np.array_equal(got[32 * 64:32 * 100], h[32 * 64:32 * 100])


@pytest.mark.parametrize("kernel", ["chain", "split", "pair"])
def test_every_kernel_vs_oracle(gpu_lib, oracle, kernel):
    """Each PoH kernel, named through spl_poh_device_kernel, on chain counts
    around its group sizes (32 chains for the pair kernel, 64 for the others),
    ragged counts with zeros, one chain far longer than its group, and the
    full 4096-chain shape at 2,000 hashes: every end state equals the
    oracle's."""
    import hip_rt
    rng = np.random.default_rng(15)
    s = hip_rt.Stream()
    for n in (1, 2, 31, 32, 33, 63, 64, 65, 100, 2049):
        h, counts = _ragged(n, rng, seed=1000 + n)
        counts[n // 2] = 61
        want = h.copy()
        oracle.poh_verify_many(want, counts, threads=8)
        dh = hip_rt.DeviceBuffer.from_array(h)
        dc = hip_rt.DeviceBuffer.from_array(counts)
        assert gpu_lib.poh_device_kernel(dh.ptr, dc.ptr, n, kernel, s.handle) == 0
        s.synchronize()
        assert np.array_equal(dh.download(), want), (kernel, n)
    h, counts = W.poh_chains(4096, 2000, seed=16)
    want = h.copy()
    oracle.poh_verify_many(want, counts, threads=16)
    dh = hip_rt.DeviceBuffer.from_array(h)
    dc = hip_rt.DeviceBuffer.from_array(counts)
    assert gpu_lib.poh_device_kernel(dh.ptr, dc.ptr, 4096, kernel, s.handle) == 0
    s.synchronize()
    assert np.array_equal(dh.download(), want), kernel
