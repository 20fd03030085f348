"""CPU oracle for the packet-parsing step before the verify path.

TEST INFRASTRUCTURE ONLY.  Loaded by tests/ as the checker; the product
library never uses it.

The reference consumes per-signature tables (message length, public-key,
signature and message offsets; src/cuda-ecc-ed25519/gpu_ctx.cu:133-136,
verify.cu:240-340) but does not build them.  Its caller does, from Solana
wire-format packets: agave v2.0 `perf/src/sigverify.rs`
(`do_get_packet_offsets`, `generate_offsets`, `copy_return_values`) with
`decode_shortu16_len` from the `solana-short-vec` crate.  That code is not in
the reference tree and not in this image; this module restates its published
algorithm:

decode_shortu16_len(bytes): a u16 in up to three bytes, 7 bits per byte, low
  group first, bit 7 = "more".  Fails when the bytes run out, when a byte
  after the first is 0x00 (an alias: the encoding is not the shortest), when
  the third byte has bit 7 set, or when the value exceeds 0xFFFF.

packet_offsets(packet bytes [0, size), base = packet's offset in the batch):
  1. size >= 1 + 64, and size <= the packet buffer;
  2. (n, a) = decode_shortu16_len(bytes)          signature count, its length
  3. m = a + 64 n; require m < size                message start
  4. h = m, or m + 1 when bytes[m] has bit 7 set (a versioned message); a
     version other than 0 (bytes[m] & 0x7F) fails
  5. require h + 1 + 3 <= size                     header and one count byte
  6. require bytes[h] > bytes[h+1]  (num_required_signatures above
     num_readonly_signed_accounts: the fee payer must be writable)
  7. require bytes[h] == n
  8. (k, b) = decode_shortu16_len(bytes[h+3 :])    account-key count
  9. p = h + 3 + b; require p + 32 k <= size and k >= n
  10. signatures at base + a + 64 i, keys at base + p + 32 i, message at
      base + m with length size - m, for i < n
  Any failure gives the packet no signatures.

generate_offsets: the packets of all batches in order, base advancing by the
  packet stride; one table row per signature.

copy_return_values: a packet is kept when it has at least one signature and
  every one of its signatures verified.

Parity with the live agave code is not checkable here (it is absent).  The
restatement is pinned by construction instead: the tests build packets whose
signature, key and message positions are known from the builder, mutate one
field at a time across each rule's boundary, and check the short-vec decoder
against the closed form (a byte string decodes iff it starts with the
shortest encoding of some u16).
"""
import numpy as np


def decode_shortu16_len(buf, pos, end):
    """(value, length) of the short-vec u16 at buf[pos:end], or None."""
    val = 0
    for k in range(3):
        if pos + k >= end:
            return None
        b = int(buf[pos + k])
        if b == 0 and k != 0:
            return None                       # alias
        if k == 2 and b & 0x80:
            return None                       # third byte continues
        val |= (b & 0x7F) << (7 * k)
        if val > 0xFFFF:
            return None                       # overflow
        if not b & 0x80:
            return val, k + 1
    return None


def encode_shortu16(v):
    """The shortest encoding of v (what a serializer writes)."""
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def packet_offsets(buf, size, buf_len):
    """Packet-relative (n, sig_start, msg_start, key_start) or None, for the
    packet whose bytes are buf[0:buf_len] with `size` valid bytes."""
    if size < 65 or size > buf_len:
        return None
    d = decode_shortu16_len(buf, 0, size)
    if d is None:
        return None
    n, a = d
    m = a + 64 * n
    if m >= size:
        return None
    h = m
    if buf[m] & 0x80:
        if buf[m] & 0x7F:
            return None                       # only version 0 is known
        h = m + 1
    if h + 1 + 3 > size:
        return None
    if buf[h] <= buf[h + 1]:
        return None                           # payer not writable
    if buf[h] != n:
        return None
    d = decode_shortu16_len(buf, h + 3, size)
    if d is None:
        return None
    k, b = d
    p = h + 3 + b
    if p + 32 * k > size or k < n:
        return None
    return n, a, m, p


def generate_offsets(packets, stride, sizes):
    """Tables for a concatenation of packets of `stride` bytes.  Returns
    (sig_lens per packet, message_lens, public_key_offsets, signature_offsets,
    message_start_offsets) as uint32 arrays."""
    packets = np.asarray(packets, np.uint8)
    num = len(sizes)
    sig_lens = np.zeros(num, np.uint32)
    lens, pks, sigs, msgs = [], [], [], []
    for i in range(num):
        base = i * stride
        r = packet_offsets(packets[base:base + stride], int(sizes[i]), stride)
        if r is None:
            continue
        n, a, m, p = r
        if base + p + 32 * n > 0xFFFFFFFF:
            continue                          # offsets are 32-bit
        sig_lens[i] = n
        for k in range(n):
            sigs.append(base + a + 64 * k)
            pks.append(base + p + 32 * k)
            msgs.append(base + m)
            lens.append(int(sizes[i]) - m)
    u32 = lambda x: np.asarray(x, np.uint32)
    return sig_lens, u32(lens), u32(pks), u32(sigs), u32(msgs)


def packet_verdicts(sig_lens, sig_mask):
    """1 for a packet with signatures that all verified, else 0."""
    out = np.zeros(len(sig_lens), np.uint8)
    at = 0
    for i, n in enumerate(np.asarray(sig_lens, np.int64)):
        out[i] = 1 if n and all(sig_mask[at:at + n] == 1) else 0
        at += n
    return out
