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


def test_engine_replay_autotune(gpu_lib, oracle):
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


def test_engine_bounds(gpu_lib, oracle):
    import perf_libs
    eng = perf_libs.VerifyEngine(gpu_lib, 64, 64 * 1232)
    try:
        assert eng.run(65) == 2          # above capacity: rejected, nothing launched
        assert eng.run(0) == 0
    finally:
        eng.close()
