"""Seeded synthetic inputs in the libcuda-crypt.so packet layout.

Every generator returns plain numpy arrays: the packet buffer (the
concatenation of gpu_Elems batches, stride `message_size`), the four offset
tables and, for verify, the list of per-batch packet arrays.  Signing is done
by a caller-supplied function so tests can use an independent signer and the
benchmark can use the library under test for input preparation.

Workload shapes (signatures sign the message bytes they are checked against):
  verify_plain(n)         one signature per packet: [0,64) signature,
                          [64,96) public key, [96,96+msg_len) message;
                          a fraction of packets gets one bit flipped in the
                          signature or the message.
  verify_validator(...)   batches of U[1,256] packets; packet = [0] count n,
                          n x 64-byte signatures, then the message, whose
                          header holds the n public keys at 3 + 32 k; message
                          lengths U[max(100, 3+32n), min(1200, 1231-64n)];
                          corruption mix: bit-flipped signatures, S := S + L
                          (non-canonical), public key := an order-<=8 point.
  wire_packets(...)       Solana wire-format transactions, one per packet
                          (legacy or version-0 messages), for the packet
                          parser: short-vec signature count, signatures,
                          message header, short-vec key count, keys, recent
                          blockhash, one instruction; valid packets plus
                          malformed ones that break one parsing rule each.
  poh_chains(n, hashes)   n random 32-byte states, one count per chain.
"""
import numpy as np

L_ORDER = 2**252 + 27742317777372353535851937790883648493
P_FIELD = 2**255 - 19

# The 8-torsion points of Edwards25519 (orders 1, 2, 4, 4, 8, 8, 8, 8) as
# canonical encodings, plus non-canonical encodings of the order <= 4 points.
TORSION_CANONICAL = [
    bytes.fromhex("0100000000000000000000000000000000000000000000000000000000000000"),
    bytes.fromhex("ecffffffffffffffffffffffffffffffffffffffffffffffffffffffffffff7f"),
    bytes.fromhex("0000000000000000000000000000000000000000000000000000000000000000"),
    bytes.fromhex("0000000000000000000000000000000000000000000000000000000000000080"),
    bytes.fromhex("26e8958fc2b227b045c3f489f2ef98f0d5dfac05d3c63339b13802886d53fc05"),
    bytes.fromhex("26e8958fc2b227b045c3f489f2ef98f0d5dfac05d3c63339b13802886d53fc85"),
    bytes.fromhex("c7176a703d4dd84fba3c0b760d10670f2a2053fa2c39ccc64ec7fd7792ac037a"),
    bytes.fromhex("c7176a703d4dd84fba3c0b760d10670f2a2053fa2c39ccc64ec7fd7792ac03fa"),
]
TORSION_NONCANONICAL = [
    (P_FIELD).to_bytes(32, "little"),                       # y = p   (= 0)
    (P_FIELD | (1 << 255)).to_bytes(32, "little"),
    (P_FIELD + 1).to_bytes(32, "little"),                   # y = p+1 (= 1)
    (P_FIELD + 1 | (1 << 255)).to_bytes(32, "little"),
    (1 | (1 << 255)).to_bytes(32, "little"),                # x = 0 with sign bit
    (P_FIELD - 1 | (1 << 255)).to_bytes(32, "little"),
]
SMALL_ORDER_KEYS = TORSION_CANONICAL + TORSION_NONCANONICAL


class Batch:
    """A verify/sign batch in the ABI layout."""

    def __init__(self, packets, message_size, batches, lens, pk_off, sig_off, msg_off, meta=None):
        self.packets = packets            # uint8 [total_packets * message_size]
        self.message_size = message_size
        self.batches = batches            # list of uint8 arrays (views of packets)
        self.lens = lens                  # uint32 [n]
        self.pk_off = pk_off              # uint32 [n]
        self.sig_off = sig_off            # uint32 [n]
        self.msg_off = msg_off            # uint32 [n]
        self.meta = meta or {}

    @property
    def n(self):
        return int(self.lens.size)

    @property
    def total_packets(self):
        return self.packets.size // self.message_size


def _split(packets, message_size, counts):
    views, start = [], 0
    for c in counts:
        views.append(packets[start * message_size:(start + c) * message_size])
        start += c
    return views


def _seeds(rng, n):
    return rng.integers(0, 256, size=(n, 32), dtype=np.uint8)


def verify_plain(n, sign_fn, keypair_fn, seed=1, msg_len=128, stride=1232,
                 corrupt_frac=0.10, batch_sizes=None):
    """One signature per packet (the reference driver's layout, config-1 shape).

    sign_fn(msg, pk, sk) -> 64 bytes; keypair_fn(seed32) -> (pk, sk)."""
    rng = np.random.default_rng(seed)
    pkts = np.zeros((n, stride), np.uint8)
    seeds = _seeds(rng, n)
    msgs = rng.integers(0, 256, size=(n, msg_len), dtype=np.uint8)
    for i in range(n):
        pk, sk = keypair_fn(seeds[i].tobytes())
        m = msgs[i].tobytes()
        pkts[i, 0:64] = np.frombuffer(sign_fn(m, pk, sk), np.uint8)
        pkts[i, 64:96] = np.frombuffer(pk, np.uint8)
        pkts[i, 96:96 + msg_len] = msgs[i]
    bad = rng.random(n) < corrupt_frac
    for i in np.nonzero(bad)[0]:
        bit = int(rng.integers(0, (64 + msg_len) * 8))
        byte = bit // 8
        pos = byte if byte < 64 else 96 + (byte - 64)
        pkts[i, pos] ^= np.uint8(1 << (bit % 8))
    flat = pkts.reshape(-1)
    base = np.arange(n, dtype=np.uint64) * stride
    counts = batch_sizes or [n]
    return Batch(flat, stride, _split(flat, stride, counts),
                 np.full(n, msg_len, np.uint32), (base + 64).astype(np.uint32),
                 base.astype(np.uint32), (base + 96).astype(np.uint32),
                 {"corrupted": bad})


def verify_validator(sign_fn, keypair_fn, seed=3, num_batches=128, max_packets=256,
                     max_sigs=4, stride=1232, flip_frac=0.05, noncanon_frac=0.01,
                     small_order_frac=0.01, num_keys=4096):
    """Validator-shaped batch (config-3 shape).  A pool of `num_keys` keypairs
    is drawn once and reused, as a validator sees repeat signers."""
    rng = np.random.default_rng(seed)
    key_seeds = _seeds(rng, num_keys)
    keys = [keypair_fn(key_seeds[k].tobytes()) for k in range(num_keys)]
    counts = [int(c) for c in rng.integers(1, max_packets + 1, size=num_batches)]
    total = sum(counts)
    pkts = np.zeros((total, stride), np.uint8)
    lens, pko, sgo, mso, kinds = [], [], [], [], []
    for p in range(total):
        ns = int(rng.integers(1, max_sigs + 1))
        lo, hi = max(100, 3 + 32 * ns), min(1200, stride - 1 - 64 * ns)
        mlen = int(rng.integers(lo, hi + 1))
        msg = bytearray(rng.integers(0, 256, size=mlen, dtype=np.uint8).tobytes())
        msg[0], msg[1], msg[2] = ns, 0, 0
        who = rng.integers(0, num_keys, size=ns)
        sig_kind = []
        for k in range(ns):
            pk, _ = keys[who[k]]
            u = rng.random()
            if u < small_order_frac:
                pk = SMALL_ORDER_KEYS[int(rng.integers(0, len(SMALL_ORDER_KEYS)))]
                sig_kind.append("small_order_key")
            elif u < small_order_frac + noncanon_frac:
                sig_kind.append("noncanonical_s")
            elif u < small_order_frac + noncanon_frac + flip_frac:
                sig_kind.append("bit_flip")
            else:
                sig_kind.append("valid")
            msg[3 + 32 * k:3 + 32 * (k + 1)] = pk
        msg = bytes(msg)
        base = p * stride
        mstart = 1 + 64 * ns
        pkts[p, 0] = ns
        pkts[p, mstart:mstart + mlen] = np.frombuffer(msg, np.uint8)
        for k in range(ns):
            _, sk = keys[who[k]]
            pk = msg[3 + 32 * k:3 + 32 * (k + 1)]
            sig = bytearray(sign_fn(msg, keys[who[k]][0], sk))
            kind = sig_kind[k]
            if kind == "noncanonical_s":
                s = int.from_bytes(sig[32:], "little") + L_ORDER
                sig[32:] = s.to_bytes(32, "little")
            elif kind == "bit_flip":
                bit = int(rng.integers(0, 512))
                sig[bit // 8] ^= 1 << (bit % 8)
            pkts[p, 1 + 64 * k:1 + 64 * (k + 1)] = np.frombuffer(bytes(sig), np.uint8)
            lens.append(mlen)
            sgo.append(base + 1 + 64 * k)
            pko.append(base + mstart + 3 + 32 * k)
            mso.append(base + mstart)
            kinds.append(kind)
    flat = pkts.reshape(-1)
    u32 = lambda a: np.asarray(a, dtype=np.uint32)
    return Batch(flat, stride, _split(flat, stride, counts), u32(lens), u32(pko), u32(sgo),
                 u32(mso), {"kinds": kinds, "batch_counts": counts})


def sign_batch(n, keypair_fn, seed=4, min_len=32, max_len=512, stride=None):
    """Packets holding [pk(32) | sk(64) | message] for ed25519_sign_many."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(min_len, max_len + 1, size=n).astype(np.uint32)
    stride = stride or (96 + max_len)
    pkts = np.zeros((n, stride), np.uint8)
    seeds = _seeds(rng, n)
    for i in range(n):
        pk, sk = keypair_fn(seeds[i].tobytes())
        pkts[i, 0:32] = np.frombuffer(pk, np.uint8)
        pkts[i, 32:96] = np.frombuffer(sk, np.uint8)
        pkts[i, 96:96 + lens[i]] = rng.integers(0, 256, size=int(lens[i]), dtype=np.uint8)
    flat = pkts.reshape(-1)
    base = np.arange(n, dtype=np.uint64) * stride
    return Batch(flat, stride, [flat], lens, base.astype(np.uint32),
                 (base + 32).astype(np.uint32), (base + 96).astype(np.uint32))


def poh_chains(n, num_hashes, seed=2):
    rng = np.random.default_rng(seed)
    hashes = rng.integers(0, 256, size=n * 32, dtype=np.uint8)
    counts = np.full(n, num_hashes, np.uint64) if np.isscalar(num_hashes) else \
        np.asarray(num_hashes, np.uint64)
    return hashes, counts


def short_u16(v):
    """Short-vec encoding of a u16: 7 bits per byte, low group first."""
    out = bytearray()
    while True:
        b, v = v & 0x7F, v >> 7
        out.append(b | (0x80 if v else 0))
        if not v:
            return bytes(out)


# Ways a packet can break exactly one parsing rule (wire_packets' malformed mix).
MALFORMED_KINDS = ("short", "oversize", "count_alias", "count_too_big", "version",
                   "payer_readonly", "count_mismatch", "keys_fewer", "keys_overrun",
                   "keys_alias", "header_cut")


def wire_packets(sign_fn, keypair_fn, seed=7, num_batches=128, max_packets=256, max_sigs=4,
                 stride=1232, v0_frac=0.5, flip_frac=0.05, noncanon_frac=0.01,
                 small_order_frac=0.01, malformed_frac=0.02, minimal_frac=0.02,
                 num_keys=4096):
    """Solana wire-format packets in batches of U[1, max_packets].

    Packet = short_u16(n) || n signatures || message, message = [0x80 for
    version 0] || header (n, readonly signed < n, readonly unsigned) ||
    short_u16(k) || k keys (the n signers first) || 32-byte blockhash || one
    instruction (|| an empty lookup-table list for version 0).  Messages are
    sized U[100, min(1200, stride - 1 - 64 n)] like the validator-shaped
    batch; `minimal_frac` of packets carry only header and keys (the keys end
    exactly at the packet end).  Signatures cover the message bytes and get
    the same corruption mix as verify_validator.  `malformed_frac` of packets
    then break one parsing rule (MALFORMED_KINDS) and must get no signatures.

    Returns a Batch whose tables hold the rows the parser must produce (known
    from where the builder put each field), with meta: sizes (uint32 per
    packet), sig_lens (expected signatures per packet), kinds (per row),
    malformed (per packet: kind or None), batch_counts."""
    rng = np.random.default_rng(seed)
    key_seeds = _seeds(rng, num_keys)
    keys = [keypair_fn(key_seeds[k].tobytes()) for k in range(num_keys)]
    counts = [int(c) for c in rng.integers(1, max_packets + 1, size=num_batches)]
    total = sum(counts)
    pkts = np.zeros((total, stride), np.uint8)
    sizes = np.zeros(total, np.uint32)
    sig_lens = np.zeros(total, np.uint32)
    malformed = [None] * total
    lens, pko, sgo, mso, kinds = [], [], [], [], []
    for p in range(total):
        ns = int(rng.integers(1, max_sigs + 1))
        v0 = rng.random() < v0_frac
        minimal = rng.random() < minimal_frac
        extra = 0 if minimal else int(rng.integers(1, 4))
        k = ns + extra
        who = rng.integers(0, num_keys, size=ns)
        sig_kind, signer_pks = [], []
        for j in range(ns):
            pk = keys[who[j]][0]
            u = rng.random()
            if u < small_order_frac:
                pk = SMALL_ORDER_KEYS[int(rng.integers(0, len(SMALL_ORDER_KEYS)))]
                sig_kind.append("small_order_key")
            elif u < small_order_frac + noncanon_frac:
                sig_kind.append("noncanonical_s")
            elif u < small_order_frac + noncanon_frac + flip_frac:
                sig_kind.append("bit_flip")
            else:
                sig_kind.append("valid")
            signer_pks.append(pk)
        prefix = b"\x80" if v0 else b""
        header = bytes([ns, int(rng.integers(0, ns)), int(rng.integers(0, extra + 1))])
        key_bytes = b"".join(signer_pks) + rng.integers(0, 256, 32 * extra, np.uint8).tobytes()
        head = prefix + header + short_u16(k) + key_bytes
        if minimal:
            msg = head
        else:
            lo = max(100, len(head) + 32 + 8)
            hi = min(1200, stride - 1 - 64 * ns)
            target = int(rng.integers(lo, max(lo, hi) + 1))
            fixed = len(head) + 32 + 1 + 1 + 1 + 2 + (1 if v0 else 0)
            dlen = max(0, target - fixed - 1)
            if len(short_u16(dlen)) == 2:
                dlen = max(0, dlen - 1)
            ix = (bytes([k - 1]) + short_u16(2) + bytes([0, 1 % k]) + short_u16(dlen)
                  + rng.integers(0, 256, dlen, np.uint8).tobytes())
            msg = (head + rng.integers(0, 256, 32, np.uint8).tobytes() + short_u16(1) + ix
                   + (short_u16(0) if v0 else b""))
        a = len(short_u16(ns))
        m = a + 64 * ns
        size = m + len(msg)
        assert size <= stride, (size, stride)
        row = pkts[p]
        row[:a] = np.frombuffer(short_u16(ns), np.uint8)
        row[m:size] = np.frombuffer(msg, np.uint8)
        for j in range(ns):
            _, sk = keys[who[j]]
            sig = bytearray(sign_fn(msg, keys[who[j]][0], sk))
            if sig_kind[j] == "noncanonical_s":
                sig[32:] = (int.from_bytes(sig[32:], "little") + L_ORDER).to_bytes(32, "little")
            elif sig_kind[j] == "bit_flip":
                bit = int(rng.integers(0, 512))
                sig[bit // 8] ^= 1 << (bit % 8)
            row[a + 64 * j:a + 64 * (j + 1)] = np.frombuffer(bytes(sig), np.uint8)
        sizes[p] = size
        h = m + len(prefix)
        keys_at = h + 3 + len(short_u16(k))
        if rng.random() < malformed_frac:
            bad = MALFORMED_KINDS[int(rng.integers(0, len(MALFORMED_KINDS)))]
            _break_rule(row, sizes, p, bad, ns, a, m, h, k, keys_at, size, stride, rng)
            malformed[p] = bad
            continue
        base = p * stride
        sig_lens[p] = ns
        for j in range(ns):
            lens.append(size - m)
            sgo.append(base + a + 64 * j)
            pko.append(base + keys_at + 32 * j)
            mso.append(base + m)
            kinds.append(sig_kind[j])
    flat = pkts.reshape(-1)
    u32 = lambda x: np.asarray(x, dtype=np.uint32)
    return Batch(flat, stride, _split(flat, stride, counts), u32(lens), u32(pko), u32(sgo),
                 u32(mso), {"sizes": sizes, "sig_lens": sig_lens, "kinds": kinds,
                            "malformed": malformed, "batch_counts": counts})


def _break_rule(row, sizes, p, bad, ns, a, m, h, k, keys_at, size, stride, rng):
    """Breaks exactly one parsing rule of the packet in row[:size]."""
    if bad == "short":                       # fewer than 1 + 64 bytes
        sizes[p] = int(rng.integers(0, 65))
    elif bad == "oversize":                  # more bytes than the slot holds
        sizes[p] = stride + int(rng.integers(1, 64))
    elif bad == "count_alias":               # a non-shortest count: n | 0x80, 0x00
        row[0] = ns | 0x80
        row[1] = 0
    elif bad == "count_too_big":             # signatures would reach past the end
        row[0] = 0x7F
    elif bad == "version":                   # message version 1..127
        row[m] = 0x80 | int(rng.integers(1, 128))
    elif bad == "payer_readonly":            # readonly signed >= required
        row[h + 1] = row[h] + int(rng.integers(0, 3))
    elif bad == "count_mismatch":            # header's required != n
        row[h] = ns + 1
        row[h + 1] = 0
    elif bad == "keys_fewer":                # fewer keys than signatures
        row[h + 3] = ns - 1
    elif bad == "keys_overrun":              # the keys run past the end
        row[h + 3] = 0x7F
    elif bad == "keys_alias":                # key count 0x80 | k, then 0x00
        row[h + 3] = 0x80 | (k & 0x7F)
        row[h + 4] = 0
    elif bad == "header_cut":                # the packet ends inside the header
        sizes[p] = h + 3
    else:
        raise ValueError(bad)
