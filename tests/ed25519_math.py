"""Textbook Edwards25519 / Ed25519 in Python integers (RFC 8032, section 5.1).

Used by the CPU tests as an independent closed form: it shares no code with
the C oracle or the HIP library.  Slow (milliseconds per scalar
multiplication), so only for small cases.
"""
import hashlib

P = 2**255 - 19
L = 2**252 + 27742317777372353535851937790883648493
D = (-121665 * pow(121666, P - 2, P)) % P
SQRTM1 = pow(2, (P - 1) // 4, P)
IDENT = (0, 1, 1, 0)


def inv(x):
    return pow(x, P - 2, P)


def add(p, q):
    x1, y1, z1, t1 = p
    x2, y2, z2, t2 = q
    a = (y1 - x1) * (y2 - x2) % P
    b = (y1 + x1) * (y2 + x2) % P
    c = 2 * D * t1 * t2 % P
    d = 2 * z1 * z2 % P
    e, f, g, h = b - a, d - c, d + c, b + a
    return (e * f % P, g * h % P, f * g % P, e * h % P)


def mul(k, p):
    r = IDENT
    while k:
        if k & 1:
            r = add(r, p)
        p = add(p, p)
        k >>= 1
    return r


def neg(p):
    x, y, z, t = p
    return ((-x) % P, y, z, (-t) % P)


def eq(p, q):
    return (p[0] * q[2] - q[0] * p[2]) % P == 0 and (p[1] * q[2] - q[1] * p[2]) % P == 0


def affine(p):
    zi = inv(p[2])
    return p[0] * zi % P, p[1] * zi % P


def encode(p):
    x, y = affine(p)
    return (y | ((x & 1) << 255)).to_bytes(32, "little")


def recover_x(y, sign):
    """x with the given parity for y, or None (ref10 leniency: y is taken mod p
    by the caller; x = 0 with sign = 1 is allowed)."""
    x2 = (y * y - 1) * inv(D * y * y + 1) % P
    if x2 == 0:
        return 0
    x = pow(x2, (P + 3) // 8, P)
    if (x * x - x2) % P != 0:
        x = x * SQRTM1 % P
    if (x * x - x2) % P != 0:
        return None
    if (x & 1) != sign:
        x = P - x
    return x


def decode_lenient(s):
    """The reference's decoder: y = low 255 bits mod p (so y >= p is
    accepted), x = 0 with the sign bit set is accepted.  Returns the point
    (not negated) or None."""
    v = int.from_bytes(s, "little")
    sign = v >> 255
    y = (v & ((1 << 255) - 1)) % P
    x = recover_x(y, sign)
    if x is None:
        return None
    return (x, y, 1, x * y % P)


def decode_strict(s):
    """RFC 8032 decoding: rejects y >= p and x = 0 with the sign bit set."""
    v = int.from_bytes(s, "little")
    sign, y = v >> 255, v & ((1 << 255) - 1)
    if y >= P:
        return None
    x = recover_x(y, sign)
    if x is None or (x == 0 and sign == 1):
        return None
    return (x, y, 1, x * y % P)


BY = 4 * inv(5) % P
B = (recover_x(BY, 0), BY, 1, recover_x(BY, 0) * BY % P)


def sha512_int(*parts):
    return int.from_bytes(hashlib.sha512(b"".join(parts)).digest(), "little")


def expand(seed):
    h = hashlib.sha512(seed).digest()
    a = int.from_bytes(h[:32], "little")
    a &= (1 << 254) - 8
    a |= 1 << 254
    return a, h[32:]


def keypair(seed):
    """(public key, 64-byte expanded private key) per the reference layout."""
    a, prefix = expand(seed)
    return encode(mul(a, B)), a.to_bytes(32, "little") + prefix


def sign(msg, pk, sk):
    a = int.from_bytes(sk[:32], "little")
    r = sha512_int(sk[32:], msg) % L
    R = encode(mul(r, B))
    h = sha512_int(R, pk, msg) % L
    s = (r + h * a) % L
    return R + s.to_bytes(32, "little")


def low_order_y(enc):
    """y mod p in {0, 1, p-1}."""
    y = (int.from_bytes(enc, "little") & ((1 << 255) - 1)) % P
    return y in (0, 1, P - 1)


def small_order_rule(enc):
    """The reference's check_packed_ge_small_order as a closed form:
    1 if the encoding does not decode or has y mod p in {0, 1, p-1}."""
    if decode_lenient(enc) is None:
        return 1
    return 1 if low_order_y(enc) else 0


def verify_rule(sig, msg, pk):
    """The reference's verify decision as a closed form (1 = accept)."""
    R, S = sig[:32], sig[32:]
    if S[31] & 0xF0:
        return 0
    if small_order_rule(R):
        return 0
    if small_order_rule(pk):
        return 0
    A = decode_lenient(pk)
    h = sha512_int(R, pk, msg) % L
    Rp = add(mul(int.from_bytes(S, "little"), B), mul(h, neg(A)))
    return 1 if encode(Rp) == R else 0


def order(p):
    """Order of a point of small order (1..8), or 0 if larger."""
    q = p
    for k in range(1, 9):
        if eq(q, IDENT):
            return k
        q = add(q, p)
    return 0
