"""GPU parity of ed25519_sign_many against the oracle signer: bit-identical
64-byte signatures for random keys and message lengths."""
import numpy as np
import pytest

from conftest import load_json
import workloads as W

pytestmark = pytest.mark.gpu


def test_sign_many_random(gpu_lib, oracle):
    b = W.sign_batch(2000, oracle.create_keypair, seed=4)
    got = gpu_lib.sign_many(b.batches, b.message_size, b.lens, b.pk_off, b.sig_off, b.msg_off)
    want = oracle.sign_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off)
    assert np.array_equal(got, want)


def test_sign_many_edges(gpu_lib, oracle):
    """RFC 8032 vectors, empty messages and unaligned offsets."""
    vecs = load_json("rfc8032.json")
    stride = 211
    rows = []
    for v in vecs:
        pk, sk = oracle.create_keypair(bytes.fromhex(v["secret"]))
        rows.append((pk, sk, bytes.fromhex(v["message"]), bytes.fromhex(v["signature"])))
    buf = np.zeros((len(rows), stride), np.uint8)
    for i, (pk, sk, m, _) in enumerate(rows):
        buf[i, 3:35] = np.frombuffer(pk, np.uint8)
        buf[i, 37:101] = np.frombuffer(sk, np.uint8)
        buf[i, 105:105 + len(m)] = np.frombuffer(m, np.uint8)
    base = np.arange(len(rows), dtype=np.uint32) * stride
    lens = [len(r[2]) for r in rows]
    got = gpu_lib.sign_many([buf.reshape(-1)], stride, lens, base + 3, base + 37, base + 105)
    assert [got[64 * i:64 * i + 64].tobytes() for i in range(len(rows))] == [r[3] for r in rows]


def test_sign_many_full_batch(gpu_lib, oracle):
    """The full batch-sign workload: 65,536 seeded keypairs, messages of
    32-512 B, every signature bit-identical to the oracle signer's."""
    b = W.sign_batch(65536, oracle.create_keypair, seed=4)
    print(f"sign_many: {b.n} signatures, messages {b.lens.min()}-{b.lens.max()} B", flush=True)
    got = gpu_lib.sign_many(b.batches, b.message_size, b.lens, b.pk_off, b.sig_off, b.msg_off)
    want = oracle.sign_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off)
    assert np.array_equal(got, want)
