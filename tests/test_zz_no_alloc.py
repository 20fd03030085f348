"""ed25519_verify_many_device_noalloc, the verify entry point that allocates
nothing (fixed device scratch), against ed25519_verify_many_device on the same
inputs: two calls in a row, the second with a smaller batch, each mask
bit-identical to the allocating entry point's and to the oracle's; a batch
above the fixed maximum is refused before anything is enqueued."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

MAX_SIGS = 16384


def _device_batch(b):
    import hip_rt
    d_p = hip_rt.DeviceBuffer.from_array(b.packets)
    d_t = [hip_rt.DeviceBuffer.from_array(x) for x in (b.lens, b.pk_off, b.sig_off, b.msg_off)]
    return d_p, d_t


def test_noalloc_matches_allocating_entry(gpu_lib, oracle):
    import hip_rt
    stream = hip_rt.Stream()
    # first call: a validator-shaped batch near the maximum; second: a smaller one
    big = W.verify_validator(oracle.sign, oracle.create_keypair, seed=61, num_batches=48,
                             num_keys=256)
    small = W.verify_validator(oracle.sign, oracle.create_keypair, seed=62, num_batches=9,
                               num_keys=64)
    assert small.n < big.n <= MAX_SIGS, (small.n, big.n)
    for b in (big, small):
        print(f"noalloc: {b.n} signatures", flush=True)
        want = oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off, threads=8)
        d_p, d_t = _device_batch(b)
        d_ref, d_got = hip_rt.DeviceBuffer(b.n), hip_rt.DeviceBuffer(b.n)
        d_got.upload(np.full(b.n, 0xEE, np.uint8))
        args = (d_p.ptr, b.packets.size, *[x.ptr for x in d_t], b.n)
        assert gpu_lib.verify_many_device(*args, d_ref.ptr, stream.handle) == 0
        assert gpu_lib.verify_many_device_noalloc(*args, d_got.ptr, stream.handle) == 0
        stream.synchronize()
        ref, got = d_ref.download(), d_got.download()
        assert np.array_equal(ref, want)
        assert np.array_equal(got, ref), np.nonzero(got != ref)[0][:10]


def test_noalloc_refuses_above_maximum(gpu_lib):
    import hip_rt
    d_out = hip_rt.DeviceBuffer(16)
    d_out.upload(np.full(16, 0x5A, np.uint8))
    # refused on the host before any launch: the pointers are never read
    assert gpu_lib.verify_many_device_noalloc(d_out.ptr, 16, d_out.ptr, d_out.ptr, d_out.ptr,
                                              d_out.ptr, MAX_SIGS + 1, d_out.ptr, None) == 2
    assert gpu_lib.verify_many_device_noalloc(0, 0, 0, 0, 0, 0, 0, 0, None) == 0
    assert np.array_equal(d_out.download(), np.full(16, 0x5A, np.uint8))
