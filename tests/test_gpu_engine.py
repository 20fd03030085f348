"""GPU tests of the verify engine: the graph-captured launch replayed for
different batch sizes, the autotuner, and bit-identical masks against the
oracle for every configuration it can pick."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _load(engine, b):
    import hip_rt
    hip_rt.upload_to(engine.d_packets, b.packets)
    for dptr, tab in ((engine.d_lens, b.lens), (engine.d_pk, b.pk_off),
                      (engine.d_sig, b.sig_off), (engine.d_msg, b.msg_off)):
        hip_rt.upload_to(dptr, tab)


def _mask(engine, n):
    import hip_rt
    return hip_rt.download_from(engine.d_out, n)


def test_engine_replay_autotune(gpu_lib, oracle, tmp_path):
    import perf_libs
    b = W.verify_validator(oracle.sign, oracle.create_keypair, seed=21, num_batches=24,
                           num_keys=128, flip_frac=0.1, noncanon_frac=0.03,
                           small_order_frac=0.03)
    want = oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off, threads=8)
    eng = perf_libs.VerifyEngine(gpu_lib, b.n + 100, b.packets.size)
    try:
        _load(eng, b)
        assert eng.run(b.n) == 0 and eng.sync() == 0
        assert np.array_equal(_mask(eng, b.n), want), eng.config()
        report = eng.autotune(b.n)
        lines = [l for l in report.splitlines() if l.strip()]
        assert len(lines) >= 8, report
        # every tuned configuration is exercised by the next runs
        for n in (b.n, b.n // 3, 1, 65, b.n):
            assert eng.run(n) == 0 and eng.sync() == 0
            assert np.array_equal(_mask(eng, n), want[:n]), (n, eng.config())
    finally:
        eng.close()
    # the ABI call now uses the tuned one-shot configuration of this size class
    got = gpu_lib.verify_many(b.batches, b.message_size, b.lens, b.pk_off, b.sig_off, b.msg_off)
    assert np.array_equal(got, want)
    # the record survives a save / load round trip, and the device's own
    # architecture and CU count are written with it
    tuned, plain = gpu_lib.tune_query(0, b.n)
    assert tuned and "persistent" not in plain
    path = tmp_path / "tune.txt"
    assert gpu_lib.tune_save(str(path)) >= 1
    line = [l for l in path.read_text().splitlines() if " class %d " % b.n.bit_length() in l]
    assert len(line) == 1 and line[0].startswith("device 0 gfx950 ") and line[0].endswith(plain)
    assert gpu_lib.tune_load(str(path)) >= 1
    assert gpu_lib.tune_query(0, b.n) == (True, plain)


def _load_set(engine, k, b):
    import hip_rt
    bufs = engine.set_buffers(k)
    for dptr, tab in zip(bufs[:5], (b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off)):
        hip_rt.upload_to(dptr, tab)


def test_engine_batches_in_flight(gpu_lib, oracle):
    """An engine of 3 buffer sets with a different batch in each: rounds of
    next() / run_set() keep up to 3 launches in flight, sets are refilled
    with other batches between rounds, and every set's mask equals the
    oracle's for the batch it holds; the same after autotuning with every set
    in flight, in a persistent configuration (per-set work counters) and in
    the tuned one."""
    import hip_rt
    import perf_libs
    batches = [W.verify_validator(oracle.sign, oracle.create_keypair, seed=70 + k,
                                  num_batches=6 + 3 * k, num_keys=64, flip_frac=0.1,
                                  noncanon_frac=0.03, small_order_frac=0.03)
               for k in range(4)]
    wants = [oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off, threads=8)
             for b in batches]
    cap = max(b.n for b in batches)
    eng = perf_libs.VerifyEngine(gpu_lib, cap, max(b.packets.size for b in batches), depth=3)
    try:
        held = [None] * 3

        def rounds(count, refill):
            for r in range(count):
                k = eng.next()
                assert 0 <= k < 3
                # the set's previous batch is finished: check it, then maybe refill
                if held[k] is not None:
                    got = hip_rt.download_from(eng.set_buffers(k)[5], batches[held[k]].n)
                    assert np.array_equal(got, wants[held[k]]), (k, held[k], eng.config())
                if held[k] is None or refill:
                    held[k] = (held[k] + 1 if held[k] is not None else k) % len(batches)
                    _load_set(eng, k, batches[held[k]])
                assert eng.run_set(k, batches[held[k]].n) == 0
            assert eng.sync() == 0
            for k in range(3):
                got = hip_rt.download_from(eng.set_buffers(k)[5], batches[held[k]].n)
                assert np.array_equal(got, wants[held[k]]), (k, held[k], eng.config())

        rounds(9, refill=False)
        rounds(12, refill=True)
        names = eng.candidates()
        eng.use(names.index("persistent W=4 scalars=half sort=1 waves_per_cu=4"))
        rounds(9, refill=True)
        # autotune with every set in flight (each set holds a batch of its own)
        report = eng.autotune(min(batches[h].n for h in held))
        assert len([l for l in report.splitlines() if l.strip()]) == len(names)
        rounds(9, refill=True)
    finally:
        eng.close()


def test_engine_bounds(gpu_lib, oracle):
    import perf_libs
    eng = perf_libs.VerifyEngine(gpu_lib, 64, 64 * 1232)
    try:
        assert eng.run(65) == 2          # above capacity: rejected, nothing launched
        assert eng.run(0) == 0
    finally:
        eng.close()


def _torsion_batch(oracle, n=384, seed=5):
    """Signatures under keys A + T and with R + T for each of the 8 points T of
    order dividing 8: the reference accepts A + T exactly when [h]T = O, so
    these rows separate an exact check from a cofactored one."""
    import ed25519_math as M
    rng = np.random.default_rng(seed)
    tors = [M.decode_lenient(t) for t in W.TORSION_CANONICAL]
    stride = 256
    buf = np.zeros((n, stride), np.uint8)
    for i in range(n):
        pk, sk = oracle.create_keypair(rng.integers(0, 256, 32, dtype=np.uint8).tobytes())
        T = tors[i % 8]
        kind = (i // 8) % 3
        key = M.encode(M.add(M.decode_lenient(pk), T)) if kind != 1 else pk
        msg = rng.integers(0, 256, int(rng.integers(0, 120)), dtype=np.uint8).tobytes()
        sig = oracle.sign(msg, key, sk)
        if kind >= 1:
            sig = M.encode(M.add(M.decode_lenient(sig[:32]), T)) + sig[32:]
        buf[i, :64] = np.frombuffer(sig, np.uint8)
        buf[i, 64:96] = np.frombuffer(key, np.uint8)
        buf[i, 96:96 + len(msg)] = np.frombuffer(msg, np.uint8)
        buf[i, 255] = len(msg)
    base = np.arange(n, dtype=np.uint32) * stride
    lens = buf[:, 255].astype(np.uint32)
    flat = buf.reshape(-1)
    return W.Batch(flat, stride, [flat], lens, base + 64, base.copy(), base + 96)


P25519 = 2 ** 255 - 19


def _undecodable_batch(oracle, n=320, seed=7):
    """Signatures whose key, R or both are replaced by random 32-byte strings
    (about half of them are not on the curve), or by encodings of y = p + k
    (non-canonical y, accepted by the reference's decoder for keys), mixed in
    every wave with untouched signatures.  The half-scalar kernels decode A
    and R together, so this exercises each decode failing alone and both at
    once."""
    rng = np.random.default_rng(seed)
    stride = 256
    buf = np.zeros((n, stride), np.uint8)
    for i in range(n):
        pk, sk = oracle.create_keypair(rng.integers(0, 256, 32, dtype=np.uint8).tobytes())
        msg = rng.integers(0, 256, int(rng.integers(0, 150)), dtype=np.uint8).tobytes()
        sig = oracle.sign(msg, pk, sk)
        kind = i % 6
        rand = lambda: rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        if kind in (1, 3):
            pk = rand()
        if kind in (2, 3):
            sig = rand() + sig[32:]
        if kind in (4, 5):
            y = P25519 + int(rng.integers(2, 19)) + (int(rng.integers(0, 2)) << 255)
            enc = y.to_bytes(32, "little")
            if kind == 4:
                pk = enc
            else:
                sig = enc + sig[32:]
        buf[i, :64] = np.frombuffer(sig, np.uint8)
        buf[i, 64:96] = np.frombuffer(pk, np.uint8)
        buf[i, 96:96 + len(msg)] = np.frombuffer(msg, np.uint8)
        buf[i, 255] = len(msg)
    base = np.arange(n, dtype=np.uint32) * stride
    lens = buf[:, 255].astype(np.uint32)
    flat = buf.reshape(-1)
    return W.Batch(flat, stride, [flat], lens, base + 64, base.copy(), base + 96)


def test_every_configuration(gpu_lib, oracle):
    """Each candidate configuration (one-shot / persistent, W = 4 / 5, full or
    half-size scalars, register targets, bucketing) gives the oracle's mask on
    a validator-shaped batch, on torsion-shifted keys and R, and on keys and R
    that do not decode."""
    import perf_libs
    batches = [
        W.verify_validator(oracle.sign, oracle.create_keypair, seed=23, num_batches=10,
                           num_keys=64, flip_frac=0.1, noncanon_frac=0.03,
                           small_order_frac=0.03),
        _torsion_batch(oracle),
        _undecodable_batch(oracle),
    ]
    wants = [oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off, threads=8)
             for b in batches]
    assert 0 < wants[1].sum() < wants[1].size
    # only the untouched rows of the undecodable batch verify
    assert np.array_equal(np.nonzero(wants[2])[0], np.arange(0, wants[2].size, 6))
    cap = max(b.n for b in batches)
    eng = perf_libs.VerifyEngine(gpu_lib, cap, max(b.packets.size for b in batches))
    try:
        names = eng.candidates()
        assert any("scalars=half" in nm for nm in names)
        for k, name in enumerate(names):
            eng.use(k)
            for b, want in zip(batches, wants):
                _load(eng, b)
                assert eng.run(b.n) == 0 and eng.sync() == 0
                got = _mask(eng, b.n)
                assert np.array_equal(got, want), (name, b.n, np.nonzero(got != want)[0][:10])
            print(name, "ok", flush=True)
    finally:
        eng.close()


def test_staged_form_sizes(gpu_lib, oracle):
    """The staged form (hash wave + point wave per 32 signatures, handing over
    through LDS at two barriers) and the shared-hash form (three point waves +
    a hash wave per 192 signatures, three barriers) on batch sizes around
    their groups: a partly filled last group, and groups wholly past the batch
    in a graph captured for the engine's capacity (they leave before any
    barrier).  The messages (100-1200 bytes) end on both sides of the 6 (W = 4)
    and 8 (W = 5) SHA-512 blocks the shared-hash form's hash wave runs before
    handing over."""
    import perf_libs
    b = W.verify_validator(oracle.sign, oracle.create_keypair, seed=27, num_batches=12,
                           num_keys=64, flip_frac=0.1, noncanon_frac=0.03,
                           small_order_frac=0.03)
    want = oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off, threads=8)
    eng = perf_libs.VerifyEngine(gpu_lib, b.n, b.packets.size)
    try:
        _load(eng, b)
        names = eng.candidates()
        staged = [k for k, nm in enumerate(names) if "staged" in nm]
        shared = [k for k, nm in enumerate(names) if "shared-hash" in nm]
        assert len(staged) == 2 and len(shared) == 2, names
        blocks = (81 + b.lens.astype(np.int64) + 127) // 128
        assert (blocks <= 6).any() and (blocks > 8).any()
        sizes = {k: (1, 2, 31, 32, 33, 64, 65, 500, b.n - 1, b.n) for k in staged}
        for k in shared:
            sizes[k] = (1, 63, 64, 65, 191, 192, 193, 384, 385, b.n - 1, b.n)
        for k, ns in sizes.items():
            eng.use(k)
            for n in ns:
                assert eng.run(n) == 0 and eng.sync() == 0
                got = _mask(eng, n)
                assert np.array_equal(got, want[:n]), (names[k], n,
                                                       np.nonzero(got != want[:n])[0][:10])
    finally:
        eng.close()


def _fixture_batch(rows, stride=512):
    """The rows of a verify fixture as one batch, one signature per packet:
    [0,64) signature, [64,96) key, [96,96+len) message."""
    n = len(rows)
    buf = np.zeros((n, stride), np.uint8)
    lens = np.zeros(n, np.uint32)
    for i, r in enumerate(rows):
        sig, msg, pk = (np.frombuffer(bytes.fromhex(r[k]), np.uint8) for k in ("sig", "msg", "pk"))
        assert 96 + msg.size <= stride
        buf[i, :64], buf[i, 64:96], buf[i, 96:96 + msg.size] = sig, pk, msg
        lens[i] = msg.size
    base = np.arange(n, dtype=np.uint32) * stride
    flat = buf.reshape(-1)
    return W.Batch(flat, stride, [flat], lens, base + 64, base.copy(), base + 96)


def test_every_configuration_torsion_fixture(gpu_lib):
    """Every engine configuration, the drop-in ed25519_verify_many and the
    device entry point give the compiled reference's own verdicts (not the
    oracle's) on keys A + T, R + T and both, non-canonical keys and R, and
    S in [2^252, L): the rows where a cofactored or torsion-blind check would
    differ from the reference's exact one."""
    import hip_rt
    import perf_libs
    from conftest import load_json
    rows = load_json("ref_torsion.json")["rows"]
    want = np.array([r["verify"] for r in rows], np.uint8)
    assert 0 < want.sum() < want.size
    b = _fixture_batch(rows)
    # tiled, so each configuration also sees the rows at every lane position of a wave
    reps = 5
    packets = np.tile(b.packets, reps)
    shift = np.repeat(np.arange(reps, dtype=np.uint32) * np.uint32(b.packets.size), b.n)
    big = W.Batch(packets, b.message_size, [packets], np.tile(b.lens, reps),
                  np.tile(b.pk_off, reps) + shift, np.tile(b.sig_off, reps) + shift,
                  np.tile(b.msg_off, reps) + shift)
    big_want = np.tile(want, reps)
    eng = perf_libs.VerifyEngine(gpu_lib, big.n, big.packets.size)
    try:
        for k, name in enumerate(eng.candidates()):
            eng.use(k)
            for bb, ww in ((b, want), (big, big_want)):
                _load(eng, bb)
                assert eng.run(bb.n) == 0 and eng.sync() == 0
                got = _mask(eng, bb.n)
                bad = np.nonzero(got != ww)[0]
                assert bad.size == 0, (name, [(rows[i % b.n]["class"], int(ww[i])) for i in bad[:8]])
    finally:
        eng.close()
    got = gpu_lib.verify_many(big.batches, big.message_size, big.lens, big.pk_off, big.sig_off,
                              big.msg_off)
    assert np.array_equal(got, big_want)
    d_p = hip_rt.DeviceBuffer.from_array(big.packets)
    d_t = [hip_rt.DeviceBuffer.from_array(x) for x in (big.lens, big.pk_off, big.sig_off,
                                                       big.msg_off)]
    d_o = hip_rt.DeviceBuffer(big.n)
    s = hip_rt.Stream()
    assert gpu_lib.verify_many_device(d_p.ptr, big.packets.size, *[x.ptr for x in d_t], big.n,
                                      d_o.ptr, s.handle) == 0
    s.synchronize()
    assert np.array_equal(d_o.download(), big_want)


def test_every_configuration_message_lengths(gpu_lib, oracle):
    """Every engine configuration against the oracle on messages whose
    lengths sit on every SHA-512 block boundary of R || A || M (64 + len + 17
    = 128 k, k = 1..10, and one byte either side), including the 6- and
    8-block handovers of the shared-hash forms, and on short messages from 0
    bytes; each message starts 0-3 bytes past a 4-byte boundary, a fifth of
    the signatures have one bit flipped, and the last message ends at the last
    byte of the packet buffer."""
    import perf_libs
    rng = np.random.default_rng(41)
    lengths = [m for k in range(1, 11) for m in (128 * k - 82, 128 * k - 81, 128 * k - 80)]
    lengths += [0, 1, 2, 3, 4, 5, 7, 8, 16, 31, 32, 33, 63, 64, 65, 127, 128, 129]
    rows = [(m, sh) for m in lengths for sh in range(4)]
    stride = 1304
    n = len(rows)
    buf = np.zeros((n, stride), np.uint8)
    lens = np.zeros(n, np.uint32)
    msg_off = np.zeros(n, np.uint32)
    base = np.arange(n, dtype=np.uint32) * stride
    for i, (m, sh) in enumerate(rows):
        pk, sk = oracle.create_keypair(rng.bytes(32))
        msg = rng.bytes(m)
        sig = bytearray(oracle.sign(msg, pk, sk))
        start = stride - m if i == n - 1 else 96 + sh
        mb = bytearray(msg)
        if rng.random() < 0.2:
            bit = int(rng.integers(0, (64 + m) * 8))
            if bit < 512:
                sig[bit // 8] ^= 1 << (bit % 8)
            else:
                mb[bit // 8 - 64] ^= 1 << (bit % 8)
        buf[i, :64] = np.frombuffer(bytes(sig), np.uint8)
        buf[i, 64:96] = np.frombuffer(pk, np.uint8)
        buf[i, start:start + m] = np.frombuffer(bytes(mb), np.uint8)
        lens[i] = m
        msg_off[i] = base[i] + start
    flat = buf.reshape(-1)
    b = W.Batch(flat, stride, [flat], lens, base + 64, base.copy(), msg_off)
    assert int(b.msg_off[-1]) + int(b.lens[-1]) == b.packets.size
    want = oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off)
    assert 0 < want.sum() < want.size
    eng = perf_libs.VerifyEngine(gpu_lib, n, b.packets.size)
    try:
        _load(eng, b)
        for k, name in enumerate(eng.candidates()):
            eng.use(k)
            assert eng.run(n) == 0 and eng.sync() == 0
            got = _mask(eng, n)
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, (name, [rows[i] for i in bad[:8]])
    finally:
        eng.close()


def test_every_configuration_full_batches(gpu_lib, oracle):
    """Every candidate configuration against the oracle on full-size
    validator-shaped batches (about 41k signatures each: 5 % flipped bits,
    1 % S + L, 1 % small-order keys).  SPL_CONFIG_SWEEP_BATCHES sets how many
    seeded batches (default 3; a 150-batch run, 6.2M signatures x 28
    configurations, matched the oracle everywhere, and so did a 100-batch run,
    4.1M signatures x 32 configurations, with the staged and shared-hash
    forms)."""
    import os
    import perf_libs
    count = int(os.environ.get("SPL_CONFIG_SWEEP_BATCHES", "3"))
    eng = perf_libs.VerifyEngine(gpu_lib, 70000, 40 << 20)
    try:
        names = eng.candidates()
        for seed in range(100, 100 + count):
            b = W.verify_validator(oracle.sign, oracle.create_keypair, seed=seed)
            want = oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off,
                                      threads=8)
            assert b.n <= 70000 and b.packets.size <= 40 << 20
            print(f"seed {seed}: {b.n} signatures, {int(want.sum())} accepted", flush=True)
            _load(eng, b)
            for k, name in enumerate(names):
                eng.use(k)
                assert eng.run(b.n) == 0 and eng.sync() == 0
                got = _mask(eng, b.n)
                assert np.array_equal(got, want), (seed, name, np.nonzero(got != want)[0][:10])
    finally:
        eng.close()
