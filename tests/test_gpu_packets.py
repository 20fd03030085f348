"""GPU tests of the packet path: wire-format packets parsed on the device into
the verify tables, verified, and folded into one verdict per packet.

Expected tables, signature counts and verdicts come from the oracle
(oracle/packets.py for parsing, the C oracle for verification); the builder's
own record of where it put each field is checked against the oracle on the
CPU (test_packets_oracle.py)."""
import numpy as np
import pytest

import workloads as W
from oracle import packets as P

pytestmark = pytest.mark.gpu


def _wire(oracle, **kw):
    args = dict(seed=31, num_batches=16, max_packets=96, malformed_frac=0.15, minimal_frac=0.05,
                num_keys=128)
    args.update(kw)
    return W.wire_packets(oracle.sign, oracle.create_keypair, **args)


def _expected(oracle, packets, stride, sizes):
    sig_lens, lens, pk, sig, msg = P.generate_offsets(packets, stride, sizes)
    mask = oracle.verify_many(packets, lens, pk, sig, msg, threads=8)
    return sig_lens, (lens, pk, sig, msg), mask, P.packet_verdicts(sig_lens, mask)


def _device_offsets(lib, packets, stride, sizes, cap):
    import hip_rt
    num = sizes.size
    d_pk = hip_rt.DeviceBuffer.from_array(packets)
    d_sz = hip_rt.DeviceBuffer.from_array(np.ascontiguousarray(sizes, np.uint32))
    bufs = [hip_rt.DeviceBuffer(4 * max(1, num)) for _ in range(2)]
    tabs = [hip_rt.DeviceBuffer(4 * max(1, cap)) for _ in range(4)]
    d_counts = hip_rt.DeviceBuffer(16)
    rc = lib.packet_offsets_device(d_pk.ptr, num, stride, d_sz.ptr, cap, bufs[0].ptr, bufs[1].ptr,
                                   *[t.ptr for t in tabs], d_counts.ptr)
    assert rc == 0
    hip_rt.synchronize()
    counts = d_counts.download(np.uint32, 2)
    rows = int(counts[1])
    return (bufs[0].download(np.uint32, num), bufs[1].download(np.uint32, num),
            [t.download(np.uint32, rows) for t in tabs], counts)


def test_offsets_vs_oracle(gpu_lib, oracle):
    b = _wire(oracle)
    sizes = b.meta["sizes"]
    want_lens, want_tabs, _, _ = _expected(oracle, b.packets, b.message_size, sizes)
    cap = int(sizes.size * 12)
    sig_lens, sig_base, tabs, counts = _device_offsets(gpu_lib, b.packets, b.message_size, sizes,
                                                       cap)
    assert np.array_equal(sig_lens, want_lens)
    assert np.array_equal(sig_base, np.concatenate([[0], np.cumsum(want_lens)[:-1]]))
    assert int(counts[0]) == int(counts[1]) == want_tabs[0].size
    for got, want in zip(tabs, want_tabs):
        assert np.array_equal(got, want)
    # a capacity below the batch's count: only the first rows are written
    small = int(want_tabs[0].size // 3)
    _, _, tabs2, counts2 = _device_offsets(gpu_lib, b.packets, b.message_size, sizes, small)
    assert int(counts2[0]) == want_tabs[0].size and int(counts2[1]) == small
    for got, want in zip(tabs2, want_tabs):
        assert np.array_equal(got, want[:small])


def test_offsets_fuzzed_headers(gpu_lib, oracle):
    """Valid packets with random bytes written over the fields the parser
    reads, and random sizes: the device and the oracle agree packet by
    packet."""
    b = _wire(oracle, seed=32, num_batches=8, max_packets=128, malformed_frac=0.0)
    stride = b.message_size
    rng = np.random.default_rng(33)
    pk = b.packets.copy().reshape(-1, stride)
    sizes = b.meta["sizes"].copy()
    num = sizes.size
    for i in range(num):
        n = int(b.meta["sig_lens"][i])
        m = 1 + 64 * n
        spots = [0, 1, 2, m, m + 1, m + 2, m + 3, m + 4, m + 5, m + 6]
        for _ in range(int(rng.integers(0, 3))):
            pos = spots[int(rng.integers(0, len(spots)))]
            pk[i, pos] = rng.choice([0, 1, 2, 3, 4, 5, 0x7F, 0x80, 0x81, 0xFF, rng.integers(0, 256)])
        u = rng.random()
        if u < 0.1:
            sizes[i] = int(rng.integers(0, stride + 40))
        elif u < 0.2:
            sizes[i] = max(0, int(sizes[i]) + int(rng.integers(-40, 4)))
    flat = pk.reshape(-1)
    want_lens, want_tabs, _, _ = _expected(oracle, flat, stride, sizes)
    assert 0 < (want_lens > 0).sum() < num
    sig_lens, _, tabs, counts = _device_offsets(gpu_lib, flat, stride, sizes, num * 12)
    assert np.array_equal(sig_lens, want_lens), np.nonzero(sig_lens != want_lens)[0][:10]
    for got, want in zip(tabs, want_tabs):
        assert np.array_equal(got, want)


def test_host_abi_staging_past_one_chunk(gpu_lib, oracle):
    """The host entry points stage packets through pinned memory in 4 MiB
    chunks, each sent to the device as soon as it is full.  Here the packets
    span several chunks and 60 elems, so chunk boundaries fall inside elems:
    verdicts and the per-signature mask still equal the oracle's."""
    b = _wire(oracle, seed=35, num_batches=60, max_packets=256, malformed_frac=0.05)
    print(f"staging: {b.packets.size} bytes in {len(b.batches)} elems", flush=True)
    assert b.packets.size > 2 * (4 << 20)
    sizes = b.meta["sizes"]
    want_lens, want_tabs, mask, want_ok = _expected(oracle, b.packets, b.message_size, sizes)
    rc, ok, lens = gpu_lib.verify_packets(b.batches, b.message_size, sizes)
    assert rc == 0
    assert np.array_equal(lens, want_lens)
    assert np.array_equal(ok, want_ok)
    assert np.array_equal(gpu_lib.verify_many(b.batches, b.message_size, *want_tabs), mask)


def test_verify_packets_host_abi(gpu_lib, oracle):
    b = _wire(oracle, seed=34, num_batches=24, max_packets=128)
    sizes = b.meta["sizes"]
    want_lens, want_tabs, mask, want_ok = _expected(oracle, b.packets, b.message_size, sizes)
    assert 0 < want_ok.sum() < want_ok.size
    rc, ok, lens = gpu_lib.verify_packets(b.batches, b.message_size, sizes)
    assert rc == 0
    assert np.array_equal(lens, want_lens)
    assert np.array_equal(ok, want_ok), np.nonzero(ok != want_ok)[0][:10]
    # the per-signature mask of the same rows through the reference ABI
    got_mask = gpu_lib.verify_many(b.batches, b.message_size, *want_tabs)
    assert np.array_equal(got_mask, mask)


def test_verify_packets_device_twice(gpu_lib, oracle):
    """Two calls on one stream, the second smaller (scratch reuse), each equal
    to the oracle; the signature counts are optional."""
    import hip_rt
    for seed, nb, with_lens in ((35, 20, True), (36, 3, False)):
        b = _wire(oracle, seed=seed, num_batches=nb, max_packets=128)
        sizes = b.meta["sizes"]
        want_lens, _, _, want_ok = _expected(oracle, b.packets, b.message_size, sizes)
        num = sizes.size
        d_pk = hip_rt.DeviceBuffer.from_array(b.packets)
        d_sz = hip_rt.DeviceBuffer.from_array(sizes)
        d_ok = hip_rt.DeviceBuffer(num)
        d_ln = hip_rt.DeviceBuffer(4 * num) if with_lens else None
        s = hip_rt.Stream()
        rc = gpu_lib.verify_packets_device(d_pk.ptr, num, b.message_size, d_sz.ptr, d_ok.ptr,
                                           d_ln.ptr if d_ln else None, s.handle)
        assert rc == 0
        s.synchronize()
        assert np.array_equal(d_ok.download(np.uint8, num), want_ok), seed
        if d_ln:
            assert np.array_equal(d_ln.download(np.uint32, num), want_lens)


def test_packet_edges(gpu_lib, oracle):
    import hip_rt
    # nothing to do
    rc, ok, lens = gpu_lib.verify_packets([np.zeros(0, np.uint8)], 1232, np.zeros(0, np.uint32))
    assert rc == 0 and ok.size == 0
    d_counts = hip_rt.DeviceBuffer(16)
    d_counts.upload(np.full(4, 7, np.uint32))
    assert gpu_lib.packet_offsets_device(None, 0, 1232, None, 0, None, None, None, None, None,
                                         None, d_counts.ptr) == 0
    hip_rt.synchronize()
    assert list(d_counts.download(np.uint32, 2)) == [0, 0]
    # bad arguments: a batch of 4 GiB or more, a zero stride
    assert gpu_lib.verify_packets_device(None, 1 << 22, 1 << 10, None, None) == 2
    assert gpu_lib.verify_packets_device(None, 4, 0, None, None) == 2
    # slots too small to hold any signature: every packet rejected
    b = _wire(oracle, seed=37, num_batches=2, max_packets=16, malformed_frac=0.0)
    small = 100
    pk = b.packets.reshape(-1, b.message_size)[:, :small].copy().reshape(-1)
    sizes = np.minimum(b.meta["sizes"], small).astype(np.uint32)
    rc, ok, lens = gpu_lib.verify_packets([pk], small, sizes)
    assert rc == 0 and not ok.any() and not lens.any()
    # a single packet, split over several gpu_Elems with an empty one between
    b = _wire(oracle, seed=38, num_batches=3, max_packets=5, malformed_frac=0.0)
    sizes = b.meta["sizes"]
    _, _, _, want_ok = _expected(oracle, b.packets, b.message_size, sizes)
    parts = b.batches[:1] + [np.zeros(0, np.uint8)] + b.batches[1:]
    rc, ok, _ = gpu_lib.verify_packets(parts, b.message_size, sizes)
    assert rc == 0 and np.array_equal(ok, want_ok)
