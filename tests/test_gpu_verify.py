"""GPU parity of ed25519_verify_many (ABI and device-resident forms) against
the oracle: bit-identical accept masks on valid, bit-flipped, non-canonical-S,
small-order-key and small-order-R inputs, ragged multi-signature packets and
malformed offset tables."""
import numpy as np
import pytest

from conftest import load_json
import workloads as W

pytestmark = pytest.mark.gpu


def _oracle_mask(oracle, b):
    return oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off, threads=8)


def _gpu_mask(lib, b, nds=1):
    return lib.verify_many(b.batches, b.message_size, b.lens, b.pk_off, b.sig_off, b.msg_off, nds)


def test_plain_1024(gpu_lib, oracle):
    """One signature per packet, 1232-byte stride, 128-byte messages, 10% bit flips."""
    b = W.verify_plain(1024, oracle.sign, oracle.create_keypair, seed=1)
    want = _oracle_mask(oracle, b)
    assert want.sum() == 1024 - b.meta["corrupted"].sum()
    for nds in (0, 1):
        assert np.array_equal(_gpu_mask(gpu_lib, b, nds), want)


def test_multiple_elems_split(gpu_lib, oracle):
    b = W.verify_plain(300, oracle.sign, oracle.create_keypair, seed=5,
                       batch_sizes=[1, 100, 57, 142])
    assert np.array_equal(_gpu_mask(gpu_lib, b), _oracle_mask(oracle, b))


def test_validator_shape_small(gpu_lib, oracle):
    b = W.verify_validator(oracle.sign, oracle.create_keypair, seed=7, num_batches=12,
                           num_keys=256, flip_frac=0.1, noncanon_frac=0.05,
                           small_order_frac=0.05)
    want = _oracle_mask(oracle, b)
    got = _gpu_mask(gpu_lib, b)
    assert np.array_equal(got, want)
    kinds = np.array(b.meta["kinds"])
    assert got[kinds == "valid"].all()
    assert not got[kinds == "noncanonical_s"].any()


def test_validator_shape_full(gpu_lib, oracle):
    """128 batches of U[1,256] packets, 1-4 signatures each (~41k)."""
    b = W.verify_validator(oracle.sign, oracle.create_keypair, seed=3)
    assert b.n > 30000
    assert np.array_equal(_gpu_mask(gpu_lib, b), _oracle_mask(oracle, b))


def test_reference_fixture_cases(gpu_lib):
    """The seeded cases whose verdicts came from the reference's own code."""
    cases = load_json("ref_quirks.json")["cases"]
    stride = 1600
    pk = np.zeros((len(cases), stride), np.uint8)
    lens = []
    for i, c in enumerate(cases):
        sig, key, msg = (bytes.fromhex(c[k]) for k in ("sig", "pk", "msg"))
        pk[i, :64] = np.frombuffer(sig, np.uint8)
        pk[i, 64:96] = np.frombuffer(key, np.uint8)
        pk[i, 96:96 + len(msg)] = np.frombuffer(msg, np.uint8)
        lens.append(len(msg))
    base = np.arange(len(cases), dtype=np.uint32) * stride
    got = gpu_lib.verify_many([pk.reshape(-1)], stride, lens, base + 64, base, base + 96)
    assert got.tolist() == [c["verify"] for c in cases]


def test_small_order_and_torsion(gpu_lib, oracle):
    """Every order <= 8 encoding as key and as R, with unaligned offsets."""
    pk0, sk0 = oracle.create_keypair(bytes(range(32)))
    msg = b"torsion" * 20
    sig = oracle.sign(msg, pk0, sk0)
    rows = [(sig, pk0)]
    for enc in W.SMALL_ORDER_KEYS:
        rows += [(sig, enc), (enc + sig[32:], pk0), (enc + sig[32:], enc)]
    stride = 301                                   # odd stride: every field unaligned
    buf = np.zeros((len(rows), stride), np.uint8)
    for i, (s, k) in enumerate(rows):
        buf[i, 1:65] = np.frombuffer(s, np.uint8)
        buf[i, 67:99] = np.frombuffer(k, np.uint8)
        buf[i, 101:101 + len(msg)] = np.frombuffer(msg, np.uint8)
    base = np.arange(len(rows), dtype=np.uint32) * stride
    lens = np.full(len(rows), len(msg), np.uint32)
    got = gpu_lib.verify_many([buf.reshape(-1)], stride, lens, base + 67, base + 1, base + 101)
    want = oracle.verify_many(buf.reshape(-1), lens, base + 67, base + 1, base + 101)
    assert np.array_equal(got, want)
    assert got[0] == 1 and got[1:].sum() == 0


def test_message_length_edges(gpu_lib, oracle):
    """Lengths around the SHA-512 block boundaries, incl. 0."""
    lens = [0, 1, 3, 4, 5, 47, 48, 49, 63, 64, 65, 111, 112, 113, 127, 128, 129, 175, 176,
            177, 239, 240, 241, 1100, 1167, 1168]
    stride = 1300
    buf = np.zeros((len(lens), stride), np.uint8)
    rng = np.random.default_rng(6)
    for i, n in enumerate(lens):
        pk, sk = oracle.create_keypair(rng.integers(0, 256, 32, dtype=np.uint8).tobytes())
        m = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        off = 130 + (i % 4)
        buf[i, off:off + n] = np.frombuffer(m, np.uint8)
        buf[i, 0:64] = np.frombuffer(oracle.sign(m, pk, sk), np.uint8)
        buf[i, 64:96] = np.frombuffer(pk, np.uint8)
    base = np.arange(len(lens), dtype=np.uint32) * stride
    msg_off = base + 130 + (np.arange(len(lens)) % 4).astype(np.uint32)
    got = gpu_lib.verify_many([buf.reshape(-1)], stride, lens, base + 64, base, msg_off)
    assert got.tolist() == [1] * len(lens)


def test_empty_and_bad_offsets(gpu_lib, oracle):
    out = gpu_lib.verify_many([np.zeros(0, np.uint8)], 1232, [], [], [], [])
    assert out.size == 0
    b = W.verify_plain(8, oracle.sign, oracle.create_keypair, seed=2, corrupt_frac=0)
    sig_off = b.sig_off.copy()
    sig_off[3] = b.packets.size - 10               # runs past the packet buffer
    lens = b.lens.copy()
    lens[5] = 10**6                                # message past the buffer
    got = gpu_lib.verify_many(b.batches, b.message_size, lens, b.pk_off, sig_off, b.msg_off)
    assert got.tolist() == [1, 1, 1, 0, 1, 0, 1, 1]


def test_allocation_failure_rejects_and_recovers(gpu_lib, oracle):
    """Calls whose buffers cannot be allocated on the device (a 1 TiB packet
    area or hash array, more than the GPU holds) fail safely instead of
    aborting as the reference does on a runtime error (gpu_common.h:15-24):
    verify rejects every signature, sign writes zero signatures, PoH returns 1
    with the hashes untouched; nothing of the caller's buffers is read.  The
    same thread's next calls, on every queue of the pool, give exact results."""
    import ctypes
    import perf_libs
    L = gpu_lib.lib
    u32p = ctypes.POINTER(ctypes.c_uint32)
    u8p = ctypes.POINTER(ctypes.c_uint8)
    empty = np.zeros(64, np.uint8)
    elems = (perf_libs.GpuElems * 1)()
    elems[0].elems = empty.ctypes.data_as(u8p)
    elems[0].num = 0                                  # no packet bytes to read
    n = 4
    tabs = [np.zeros(n, np.uint32) for _ in range(4)]
    tp = [t.ctypes.data_as(u32p) for t in tabs]
    for _ in range(8):                                # every queue of the pool
        out = np.full(n, 7, np.uint8)
        L.ed25519_verify_many(elems, 1, 1 << 20, 1 << 20, n, *tp, out.ctypes.data_as(u8p), 1)
        assert out.tolist() == [0] * n
        sigs = np.full(64 * n, 7, np.uint8)
        L.ed25519_sign_many(elems, 1, 1 << 20, 1 << 20, n, *tp, sigs.ctypes.data_as(u8p), 1)
        assert not sigs.any()
        h = np.arange(64, dtype=np.uint8)
        counts = np.ones(2, np.uint64)
        assert L.poh_verify_many(h.ctypes.data_as(u8p),
                                 counts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                 1 << 35, 1) == 1
        assert np.array_equal(h, np.arange(64, dtype=np.uint8))
    b = W.verify_plain(96, oracle.sign, oracle.create_keypair, seed=5, corrupt_frac=0.2)
    want = _oracle_mask(oracle, b)
    h0, counts = W.poh_chains(40, 17, seed=5)
    want_h = h0.copy()
    oracle.poh_verify_many(want_h, counts)
    for _ in range(8):
        got = gpu_lib.verify_many(b.batches, b.message_size, b.lens, b.pk_off, b.sig_off,
                                  b.msg_off)
        assert np.array_equal(got, want)
        h = h0.copy()
        assert gpu_lib.poh_verify_many(h, counts) == 0
        assert np.array_equal(h, want_h)


def test_pinned_packet_batches(gpu_lib, oracle):
    """Packet batches the caller holds in pinned memory are read by the copy
    kernel straight from the caller's pages; others go through the staging
    buffer.  A call mixing pinned batches, pageable ones and a pinned batch at
    an address that is not 16-byte aligned (staged) gives the oracle's mask,
    for verify, sign and the wire-format packet entry point alike."""
    import hip_rt
    b = W.verify_validator(oracle.sign, oracle.create_keypair, seed=31, num_batches=12,
                           num_keys=64, flip_frac=0.1)
    want = _oracle_mask(oracle, b)
    mixed = []
    for k, x in enumerate(b.batches):
        if k % 3 == 0:
            mixed.append(hip_rt.pinned_copy(x))               # pinned, aligned
        elif k % 3 == 1:
            mixed.append(x.copy())                            # pageable
        else:
            pad = hip_rt.pinned_copy(np.concatenate([np.zeros(8, np.uint8), x]))
            mixed.append(pad[8:])                             # pinned, misaligned: staged
    assert any(m.ctypes.data % 16 for m in mixed)
    # more pinned batches than one gathered copy launch takes (128): 2-packet
    # pieces of the same concatenation
    cat = np.concatenate(b.batches)
    npk = cat.size // b.message_size
    pieces = [hip_rt.pinned_copy(cat[i * b.message_size:min(i + 2, npk) * b.message_size])
              for i in range(0, npk, 2)]
    assert len(pieces) > 2 * 128
    for batches in (mixed, [hip_rt.pinned_copy(x) for x in b.batches], pieces):
        for _ in range(3):
            got = gpu_lib.verify_many(batches, b.message_size, b.lens, b.pk_off, b.sig_off,
                                      b.msg_off)
            assert np.array_equal(got, want)
    s = W.sign_batch(300, oracle.create_keypair, seed=33)
    want_s = oracle.sign_many(s.packets, s.lens, s.pk_off, s.sig_off, s.msg_off)
    got_s = gpu_lib.sign_many([hip_rt.pinned_copy(x) for x in s.batches], s.message_size,
                              s.lens, s.pk_off, s.sig_off, s.msg_off)
    assert np.array_equal(got_s, want_s)
    wp = W.wire_packets(oracle.sign, oracle.create_keypair, seed=34, num_batches=6,
                        malformed_frac=0.05)
    rc, ok_ref, _ = gpu_lib.verify_packets(wp.batches, wp.message_size, wp.meta["sizes"])
    assert rc == 0
    rc, ok_pinned, _ = gpu_lib.verify_packets([hip_rt.pinned_copy(x) for x in wp.batches],
                                              wp.message_size, wp.meta["sizes"])
    assert rc == 0 and np.array_equal(ok_pinned, ok_ref) and 0 < ok_ref.sum() < ok_ref.size


def test_device_resident_form(gpu_lib, oracle):
    import hip_rt
    b = W.verify_validator(oracle.sign, oracle.create_keypair, seed=11, num_batches=6,
                           num_keys=64)
    want = _oracle_mask(oracle, b)
    dp = hip_rt.DeviceBuffer.from_array(b.packets)
    tabs = [hip_rt.DeviceBuffer.from_array(x) for x in (b.lens, b.pk_off, b.sig_off, b.msg_off)]
    dout = hip_rt.DeviceBuffer(b.n)
    s = hip_rt.Stream()
    rc = gpu_lib.verify_many_device(dp.ptr, b.packets.size, *[x.ptr for x in tabs], b.n,
                                    dout.ptr, s.handle)
    assert rc == 0
    s.synchronize()
    assert np.array_equal(dout.download(), want)


def test_concurrent_callers(gpu_lib, oracle):
    """Any thread may call the ABI concurrently (the reference serialises per
    context, gpu_ctx.cu:46-73): 8 threads, each with its own batch shape, call
    ed25519_verify_many, ed25519_sign_many and poh_verify_many at once, three
    times each, and every result equals the oracle's."""
    import threading
    jobs = []
    for t in range(8):
        kind = ("verify", "verify", "verify", "sign", "poh")[t % 5]
        if kind == "verify":
            b = W.verify_plain(200 + 300 * t, oracle.sign, oracle.create_keypair, seed=40 + t,
                               batch_sizes=[50, 150 + 300 * t])
            want = _oracle_mask(oracle, b)
        elif kind == "sign":
            b = W.sign_batch(500 + 100 * t, oracle.create_keypair, seed=50 + t)
            want = oracle.sign_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off)
        else:
            h, counts = W.poh_chains(300 + 50 * t, 200 + t, seed=60 + t)
            want = h.copy()
            oracle.poh_verify_many(want, counts, threads=4)
            b = (h, counts)
        jobs.append((kind, b, want))
    results = [[] for _ in jobs]
    errors = []

    def run(i):
        kind, b, _ = jobs[i]
        try:
            for _ in range(3):
                if kind == "verify":
                    got = gpu_lib.verify_many(b.batches, b.message_size, b.lens, b.pk_off,
                                              b.sig_off, b.msg_off)
                elif kind == "sign":
                    got = gpu_lib.sign_many(b.batches, b.message_size, b.lens, b.pk_off,
                                            b.sig_off, b.msg_off)
                else:
                    got = b[0].copy()
                    assert gpu_lib.poh_verify_many(got, b[1]) == 0
                results[i].append(got)
        except Exception as exc:          # reported below, on the main thread
            errors.append((i, repr(exc)))

    threads = [threading.Thread(target=run, args=(i,)) for i in range(len(jobs))]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
        assert not th.is_alive(), "a caller did not return"
    assert not errors, errors
    for (kind, _, want), got in zip(jobs, results):
        assert len(got) == 3
        for g in got:
            assert np.array_equal(g, want), kind
