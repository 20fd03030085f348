"""CPU tests of the packet-parsing oracle (oracle/packets.py) and of the
wire-format builder the GPU tests use.

The parsing rules come from the reference's caller (agave perf/src/sigverify.rs,
absent here), so they are pinned by construction: the builder records where it
put every signature, key and message, malformed packets break exactly one
rule each, and the short-vec decoder is checked against the closed form "a
byte string decodes iff it starts with the shortest encoding of a u16".
"""
import numpy as np

import workloads as W
from oracle import packets as P


def _closed_form(b):
    """decode via the shortest-encoding rule: the first byte below 0x80 ends
    the value; it decodes iff the value fits 16 bits and re-encodes to the
    same bytes."""
    for ln in range(1, 4):
        if len(b) < ln:
            return None
        if b[ln - 1] < 0x80:
            v = sum((b[i] & 0x7F) << (7 * i) for i in range(ln))
            if v > 0xFFFF or W.short_u16(v) != bytes(b[:ln]):
                return None
            return v, ln
    return None


def test_short_u16_closed_form():
    for v in list(range(0, 300)) + [0x3FFF, 0x4000, 0xFFFF]:
        e = W.short_u16(v)
        assert P.decode_shortu16_len(e, 0, len(e)) == (v, len(e))
        assert P.encode_shortu16(v) == e
        assert len(e) == (1 if v < 0x80 else 2 if v < 0x4000 else 3)
    # every string of one and two bytes, then three-byte strings with the
    # third byte at the values where the rules change, plus a random sample
    cases = [bytes([a]) for a in range(256)]
    cases += [bytes([a, b]) for a in range(256) for b in range(256)]
    for c in (0x00, 0x01, 0x02, 0x03, 0x04, 0x7F, 0x80, 0x81, 0x83, 0xFF):
        cases += [bytes([a, b, c]) for a in range(0x80, 0x100) for b in range(0x80, 0x100, 7)]
    rng = np.random.default_rng(1)
    cases += [bytes(x) for x in rng.integers(0, 256, (50000, 3), dtype=np.uint8)]
    for c in cases:
        assert P.decode_shortu16_len(c, 0, len(c)) == _closed_form(c), c.hex()
    # running out of bytes: the same string cut short fails
    assert P.decode_shortu16_len(bytes([0x80, 0x01]), 0, 1) is None
    assert P.decode_shortu16_len(bytes([0xFF, 0xFF, 0x03]), 0, 3) == (0xFFFF, 3)
    assert P.decode_shortu16_len(bytes([0xFF, 0xFF, 0x04]), 0, 3) is None     # > 0xFFFF
    assert P.decode_shortu16_len(bytes([0x80, 0x80, 0x00]), 0, 3) is None     # alias


def _batch(oracle, **kw):
    args = dict(seed=3, num_batches=6, max_packets=60, malformed_frac=0.25, minimal_frac=0.1,
                num_keys=64)
    args.update(kw)
    return W.wire_packets(oracle.sign, oracle.create_keypair, **args)


def test_builder_vs_oracle_offsets(oracle):
    b = _batch(oracle)
    sizes = b.meta["sizes"]
    sig_lens, lens, pk, sig, msg = P.generate_offsets(b.packets, b.message_size, sizes)
    assert np.array_equal(sig_lens, b.meta["sig_lens"])
    for got, want in ((lens, b.lens), (pk, b.pk_off), (sig, b.sig_off), (msg, b.msg_off)):
        assert np.array_equal(got, want)
    bad = [k for k in b.meta["malformed"] if k]
    assert len(set(bad)) == len(W.MALFORMED_KINDS), sorted(set(bad))
    for i, kind in enumerate(b.meta["malformed"]):
        assert (sig_lens[i] == 0) == (kind is not None), (i, kind)


def test_builder_signatures_verify(oracle):
    """Rows of valid packets verify exactly when the builder left the
    signature intact, and the packet verdict is the AND over its rows."""
    b = _batch(oracle, seed=4, malformed_frac=0.1)
    mask = oracle.verify_many(b.packets, b.lens, b.pk_off, b.sig_off, b.msg_off)
    kinds = np.array(b.meta["kinds"])
    assert np.array_equal(mask == 1, kinds == "valid")
    ok = P.packet_verdicts(b.meta["sig_lens"], mask)
    at = 0
    for i, n in enumerate(b.meta["sig_lens"]):
        want = n > 0 and all(kinds[at:at + n] == "valid")
        assert ok[i] == want
        at += n


def test_rule_boundaries(oracle):
    """Each size limit passes at the boundary and fails one byte short."""
    b = _batch(oracle, seed=5, malformed_frac=0.0, minimal_frac=1.0, num_batches=2)
    stride = b.message_size
    for i, size in enumerate(b.meta["sizes"]):
        row = b.packets[i * stride:(i + 1) * stride]
        n = int(b.meta["sig_lens"][i])
        r = P.packet_offsets(row, int(size), stride)
        assert r is not None and r[0] == n
        _, a, m, p = r
        k = P.decode_shortu16_len(row, p - 1, p)[0] if row[p - 1] < 0x80 else None
        assert k is not None and p + 32 * k == size          # minimal: keys end the packet
        assert P.packet_offsets(row, int(size) - 1, stride) is None
        assert P.packet_offsets(row, int(size), int(size)) is not None   # size == buffer
        assert P.packet_offsets(row, int(size), int(size) - 1) is None   # size > buffer
    # the 65-byte floor: a packet that would parse at 65 bytes but not at 64
    row = np.zeros(128, np.uint8)
    row[0] = 0                    # no signatures: m = 1
    row[1:4] = [0, 0, 0]          # required 0 <= readonly 0: fails anyway
    assert P.packet_offsets(row, 65, 128) is None


def test_verdicts_edge_cases():
    assert list(P.packet_verdicts([], np.zeros(0, np.uint8))) == []
    assert list(P.packet_verdicts([0, 2, 1], np.array([1, 1, 0], np.uint8))) == [0, 1, 0]
    assert list(P.packet_verdicts([1, 0, 1], np.array([1, 1], np.uint8))) == [1, 0, 1]
