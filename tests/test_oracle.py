"""Pins the CPU oracle (oracle/liboracle.so) before anything is checked
against it: hashlib, the reference's own PoH fixture, RFC 8032 vectors, a
textbook big-int Ed25519 (tests/ed25519_math.py), and outputs of the
reference's own primitives compiled from its sources (committed fixture, plus
the live library when it is built)."""
import hashlib
import random

import numpy as np
import pytest

import ed25519_math as E
from conftest import load_json
import workloads as W


def test_sha256_sha512_vs_hashlib(oracle):
    rng = random.Random(11)
    for n in list(range(0, 260)) + [1000, 4095, 4096, 4097]:
        m = bytes(rng.randrange(256) for _ in range(n))
        assert oracle.sha256(m) == hashlib.sha256(m).digest()
        assert oracle.sha512(m) == hashlib.sha512(m).digest()


def test_poh_reference_fixture(oracle):
    fx = load_json("poh_reference.json")
    hashes = np.frombuffer(bytes.fromhex("".join(fx["inputs"])), np.uint8).copy()
    counts = np.array(fx["num_hashes"], np.uint64)
    assert fx["num_chains"] == 129 and sorted(set(fx["num_hashes"])) == [2451, 16901, 19353]
    oracle.poh_verify_many(hashes, counts)
    assert hashes.tobytes().hex() == "".join(fx["outputs"])


def test_poh_closed_form_and_zero_count(oracle):
    rng = np.random.default_rng(5)
    h = rng.integers(0, 256, 32 * 6, dtype=np.uint8)
    counts = np.array([0, 1, 2, 3, 17, 64], np.uint64)
    want = []
    for i, c in enumerate(counts):
        x = h[32 * i:32 * i + 32].tobytes()
        for _ in range(int(c)):
            x = hashlib.sha256(x).digest()
        want.append(x)
    oracle.poh_verify_many(h, counts)
    assert h.tobytes() == b"".join(want)


def test_sc_reduce_muladd_vs_bigint(oracle):
    rng = random.Random(3)
    specials = [0, 1, E.L - 1, E.L, E.L + 1, 2 * E.L, 2**512 - 1, 2**256, 2**253 - 1]
    for k in range(300):
        x = specials[k] if k < len(specials) else rng.getrandbits(rng.choice([64, 253, 256, 512]))
        assert oracle.sc_reduce(x.to_bytes(64, "little")) == (x % E.L).to_bytes(32, "little")
    for _ in range(200):
        a, b, c = (rng.getrandbits(256) for _ in range(3))
        got = oracle.sc_muladd(*(v.to_bytes(32, "little") for v in (a, b, c)))
        assert got == ((a * b + c) % E.L).to_bytes(32, "little")


def test_rfc8032_vectors(oracle):
    for v in load_json("rfc8032.json"):
        seed, msg = bytes.fromhex(v["secret"]), bytes.fromhex(v["message"])
        pk, sk = oracle.create_keypair(seed)
        assert pk.hex() == v["public"]
        sig = oracle.sign(msg, pk, sk)
        assert sig.hex() == v["signature"]
        assert oracle.verify(sig, msg, pk) == 1
        # and the textbook implementation agrees on the same vector
        assert E.keypair(seed)[0].hex() == v["public"]
        assert E.sign(msg, pk, sk).hex() == v["signature"]


def test_keypair_sign_vs_bigint(oracle):
    rng = random.Random(7)
    for _ in range(12):
        seed = bytes(rng.randrange(256) for _ in range(32))
        msg = bytes(rng.randrange(256) for _ in range(rng.randrange(0, 200)))
        pk, sk = oracle.create_keypair(seed)
        assert (pk, sk) == E.keypair(seed)
        assert oracle.sign(msg, pk, sk) == E.sign(msg, pk, sk)


def test_verify_closed_form_adversarial(oracle):
    """Verify decision vs the closed-form predicate on valid, corrupted,
    non-canonical and small-order inputs."""
    rng = random.Random(9)
    special = W.SMALL_ORDER_KEYS
    for k in range(40):
        seed = bytes(rng.randrange(256) for _ in range(32))
        msg = bytes(rng.randrange(256) for _ in range(rng.randrange(0, 150)))
        pk, sk = E.keypair(seed)
        sig = bytearray(E.sign(msg, pk, sk))
        kind = k % 6
        if kind == 1:
            sig[rng.randrange(64)] ^= 1 << rng.randrange(8)
        elif kind == 2:
            sig[32:] = (int.from_bytes(sig[32:], "little") + E.L).to_bytes(32, "little")
        elif kind == 3:
            pk = special[k % len(special)]
        elif kind == 4:
            sig[:32] = special[k % len(special)]
        elif kind == 5:
            # R replaced by a non-canonical encoding of the true R (y + p when it fits)
            y = int.from_bytes(sig[:32], "little") & ((1 << 255) - 1)
            if y < 19:
                y += E.P
            sig[:32] = (y | (sig[31] >> 7 << 255)).to_bytes(32, "little")
        got = oracle.verify(bytes(sig), msg, pk)
        assert got == E.verify_rule(bytes(sig), msg, pk), (k, kind)
        if kind == 0:
            assert got == 1


def test_small_order_rule_on_torsion(oracle):
    for enc in W.SMALL_ORDER_KEYS:
        assert oracle.check_packed_ge_small_order(enc) == E.small_order_rule(enc)
    # the four order-8 points are not flagged; order <= 4 always is
    for enc in W.TORSION_CANONICAL[4:]:
        assert oracle.check_packed_ge_small_order(enc) == 0
    for enc in W.TORSION_CANONICAL[:4] + W.TORSION_NONCANONICAL:
        assert oracle.check_packed_ge_small_order(enc) == 1


def test_oracle_vs_reference_fixture(oracle):
    fx = load_json("ref_quirks.json")
    for c in fx["get_checked_scalar"]:
        rc, out = oracle.get_checked_scalar(bytes.fromhex(c["in"]), bytes.fromhex(c["prior"]))
        assert (rc, out.hex()) == (c["rc"], c["out"]), c
    for c in fx["small_order"]:
        assert oracle.check_packed_ge_small_order(bytes.fromhex(c["enc"])) == c["check"]
        # the radix-free rule reproduces the raw-limb test wherever it decodes
        if c["raw"] >= 0:
            assert E.low_order_y(bytes.fromhex(c["enc"])) == bool(c["raw"])
    for c in fx["cases"]:
        pk, sk = oracle.create_keypair(bytes.fromhex(c["seed"]))
        assert sk.hex() == c["sk"]
        sig, msg, cpk = bytes.fromhex(c["sig"]), bytes.fromhex(c["msg"]), bytes.fromhex(c["pk"])
        assert oracle.verify(sig, msg, cpk) == c["verify"], c["kind"]
        if c["kind"] in (0, 6, 7):
            assert oracle.sign(msg, pk, sk) == sig


def test_oracle_vs_torsion_fixture(oracle):
    """The oracle and the independent big-int closed form against the compiled
    reference's verdicts on keys A + T, R + T and both, non-canonical keys and
    R, and S in [2^252, L) (ref_torsion.json, 268 rows, both verdicts present
    in every torsion class)."""
    fx = load_json("ref_torsion.json")
    rows = fx["rows"]
    assert len(rows) >= 200
    seen = set()
    for r in rows:
        sig, msg, pk = bytes.fromhex(r["sig"]), bytes.fromhex(r["msg"]), bytes.fromhex(r["pk"])
        assert oracle.verify(sig, msg, pk) == r["verify"], r
        assert E.verify_rule(sig, msg, pk) == r["verify"], r
        seen.add((r["class"], r["verify"]))
    for cls in ("a_plus_t", "mixed"):
        assert (cls, 0) in seen and (cls, 1) in seen


def test_oracle_vs_live_reference(oracle, reflib):
    rng = np.random.default_rng(13)
    for k in range(200):
        seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        pk, sk = reflib.create_keypair(seed)
        assert oracle.create_keypair(seed) == (pk, sk)
        msg = rng.integers(0, 256, int(rng.integers(0, 1300)), dtype=np.uint8).tobytes()
        sig = reflib.sign(msg, pk, sk)
        assert oracle.sign(msg, pk, sk) == sig
        bad = bytearray(sig)
        bad[int(rng.integers(0, 64))] ^= 1 << int(rng.integers(0, 8))
        for s in (sig, bytes(bad)):
            assert oracle.verify(s, msg, pk) == reflib.verify(s, msg, pk)
        prior = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        inp = bytearray(rng.integers(0, 256, 32, dtype=np.uint8).tobytes())
        inp[31] = 0x10 + k % 0x70
        assert oracle.get_checked_scalar(bytes(inp), prior) == \
            reflib.get_checked_scalar(bytes(inp), prior, 0)
        x = rng.integers(0, 256, 64, dtype=np.uint8).tobytes()
        assert oracle.sc_reduce(x) == reflib.sc_reduce(x)
    for enc in W.SMALL_ORDER_KEYS:
        assert oracle.check_packed_ge_small_order(enc) == reflib.check_small_order(enc)


def test_oracle_poh_vs_live_reference(oracle, reflib):
    h1 = np.random.default_rng(1).integers(0, 256, 32 * 40, dtype=np.uint8)
    h2 = h1.copy()
    counts = np.arange(40, dtype=np.uint64) * 7
    oracle.poh_verify_many(h1, counts)
    reflib.poh_verify_many(h2, counts)
    assert np.array_equal(h1, h2)


@pytest.mark.parametrize("threads", [1, 3])
def test_oracle_batched_forms(oracle, threads):
    b = W.verify_plain(40, oracle.sign, oracle.create_keypair, seed=21, corrupt_frac=0.3)
    mask = oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off, threads=threads)
    for i in range(b.n):
        s, p, m = int(b.sig_off[i]), int(b.pk_off[i]), int(b.msg_off[i])
        assert mask[i] == oracle.verify(b.packets[s:s + 64].tobytes(),
                                        b.packets[m:m + int(b.lens[i])].tobytes(),
                                        b.packets[p:p + 32].tobytes())
    assert mask.sum() == b.n - b.meta["corrupted"].sum()
