"""CPU oracle for the Ed25519 verify/sign and SHA-256 PoH paths.

TEST INFRASTRUCTURE ONLY.  Loaded by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the product library never uses it.

`Oracle` wraps liboracle.so (our own C restatement of the reference's
algorithms, built by oracle/Makefile).  `RefLib` wraps _ref/libref.so, built
from the reference's own primitive sources when that tree is available; it is
used to pin the restatement and as the "reference" CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBORACLE = os.path.join(HERE, "liboracle.so")
LIBREF = os.path.join(HERE, "_ref", "libref.so")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


def build(ref=True):
    """Compile liboracle.so (and _ref/libref.so when the reference is mounted)."""
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-B", "-C", HERE, target], check=True)


def _ptr(a, t=_u8p):
    return a.ctypes.data_as(t)


def _buf(b):
    """bytes-like -> contiguous uint8 array (kept alive by the caller)."""
    return np.ascontiguousarray(np.frombuffer(bytes(b), dtype=np.uint8))


class Oracle:
    def __init__(self, path=LIBORACLE):
        if not os.path.exists(path):
            build(ref=False)
        self.lib = lib = ctypes.CDLL(path)
        lib.oracle_sha256.argtypes = [_u8p, _u8p, ctypes.c_size_t]
        lib.oracle_sha512.argtypes = [_u8p, _u8p, ctypes.c_size_t]
        lib.oracle_poh_verify_many.argtypes = [_u8p, _u64p, ctypes.c_size_t]
        lib.oracle_poh_verify_many_mt.argtypes = [_u8p, _u64p, ctypes.c_size_t, ctypes.c_int]
        lib.oracle_sc_reduce.argtypes = [_u8p]
        lib.oracle_sc_muladd.argtypes = [_u8p, _u8p, _u8p, _u8p]
        lib.oracle_create_keypair.argtypes = [_u8p, _u8p, _u8p]
        lib.oracle_sign.argtypes = [_u8p, _u8p, ctypes.c_size_t, _u8p, _u8p]
        lib.oracle_verify.argtypes = [_u8p, _u8p, ctypes.c_uint32, _u8p]
        lib.oracle_verify.restype = ctypes.c_int
        lib.oracle_check_packed_ge_small_order.argtypes = [_u8p]
        lib.oracle_check_packed_ge_small_order.restype = ctypes.c_int
        lib.oracle_get_checked_scalar.argtypes = [_u8p, _u8p]
        lib.oracle_get_checked_scalar.restype = ctypes.c_int
        lib.oracle_verify_many.argtypes = [_u8p, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p, _u8p]
        lib.oracle_verify_many_mt.argtypes = [_u8p, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p,
                                              _u8p, ctypes.c_int]
        lib.oracle_sign_many.argtypes = [_u8p, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p, _u8p]

    def sha256(self, msg):
        m = _buf(msg)
        out = np.zeros(32, np.uint8)
        self.lib.oracle_sha256(_ptr(out), _ptr(m), len(m))
        return out.tobytes()

    def sha512(self, msg):
        m = _buf(msg)
        out = np.zeros(64, np.uint8)
        self.lib.oracle_sha512(_ptr(out), _ptr(m), len(m))
        return out.tobytes()

    def poh_verify_many(self, hashes, num_hashes, threads=1):
        """hashes: uint8 array [n*32] updated in place; num_hashes: uint64 [n]."""
        nh = np.ascontiguousarray(num_hashes, dtype=np.uint64)
        assert hashes.dtype == np.uint8 and hashes.flags.c_contiguous
        assert hashes.size == 32 * nh.size
        if threads > 1:
            self.lib.oracle_poh_verify_many_mt(_ptr(hashes), _ptr(nh, _u64p), nh.size, threads)
        else:
            self.lib.oracle_poh_verify_many(_ptr(hashes), _ptr(nh, _u64p), nh.size)

    def sc_reduce(self, s64):
        b = np.array(bytearray(s64), np.uint8)
        self.lib.oracle_sc_reduce(_ptr(b))
        return b[:32].tobytes()

    def sc_muladd(self, a, b, c):
        out = np.zeros(32, np.uint8)
        aa, bb, cc = _buf(a), _buf(b), _buf(c)
        self.lib.oracle_sc_muladd(_ptr(out), _ptr(aa), _ptr(bb), _ptr(cc))
        return out.tobytes()

    def create_keypair(self, seed):
        pk = np.zeros(32, np.uint8)
        sk = np.zeros(64, np.uint8)
        s = _buf(seed)
        self.lib.oracle_create_keypair(_ptr(pk), _ptr(sk), _ptr(s))
        return pk.tobytes(), sk.tobytes()

    def sign(self, msg, pk, sk):
        sig = np.zeros(64, np.uint8)
        m, p, s = _buf(msg), _buf(pk), _buf(sk)
        self.lib.oracle_sign(_ptr(sig), _ptr(m), len(m), _ptr(p), _ptr(s))
        return sig.tobytes()

    def verify(self, sig, msg, pk):
        s, m, p = _buf(sig), _buf(msg), _buf(pk)
        return self.lib.oracle_verify(_ptr(s), _ptr(m), len(m), _ptr(p))

    def check_packed_ge_small_order(self, enc):
        e = _buf(enc)
        return self.lib.oracle_check_packed_ge_small_order(_ptr(e))

    def get_checked_scalar(self, inp, out_prior=bytes(32)):
        out = np.array(bytearray(out_prior), np.uint8)
        i = _buf(inp)
        rc = self.lib.oracle_get_checked_scalar(_ptr(out), _ptr(i))
        return rc, out.tobytes()

    def verify_many(self, packets, message_lens, pk_off, sig_off, msg_off, threads=1):
        n = len(sig_off)
        out = np.zeros(n, np.uint8)
        args = [np.ascontiguousarray(a, dtype=np.uint32) for a in
                (message_lens, pk_off, sig_off, msg_off)]
        pk = np.ascontiguousarray(packets, dtype=np.uint8)
        if threads > 1:
            self.lib.oracle_verify_many_mt(_ptr(pk), n, *[_ptr(a, _u32p) for a in args],
                                           _ptr(out), threads)
        else:
            self.lib.oracle_verify_many(_ptr(pk), n, *[_ptr(a, _u32p) for a in args], _ptr(out))
        return out

    def sign_many(self, packets, message_lens, pk_off, sk_off, msg_off):
        n = len(sk_off)
        out = np.zeros(64 * n, np.uint8)
        args = [np.ascontiguousarray(a, dtype=np.uint32) for a in
                (message_lens, pk_off, sk_off, msg_off)]
        pk = np.ascontiguousarray(packets, dtype=np.uint8)
        self.lib.oracle_sign_many(_ptr(pk), n, *[_ptr(a, _u32p) for a in args], _ptr(out))
        return out


class RefLib:
    """The reference's own primitives compiled from its sources (when present)."""

    def __init__(self, path=LIBREF):
        self.lib = lib = ctypes.CDLL(path)
        lib.ref_get_checked_scalar.argtypes = [_u8p, _u8p, ctypes.c_uint8]
        lib.ref_get_checked_scalar.restype = ctypes.c_int
        lib.ref_check_small_order.argtypes = [_u8p]
        lib.ref_check_small_order.restype = ctypes.c_int
        lib.ref_small_order_raw.argtypes = [_u8p]
        lib.ref_small_order_raw.restype = ctypes.c_int
        lib.ref_verify.argtypes = [_u8p, _u8p, ctypes.c_uint32, _u8p]
        lib.ref_verify.restype = ctypes.c_int
        lib.ref_create_keypair.argtypes = [_u8p, _u8p, _u8p]
        lib.ref_sign.argtypes = [_u8p, _u8p, ctypes.c_size_t, _u8p, _u8p]
        lib.ref_sign_repeat.argtypes = [_u8p, _u8p, ctypes.c_size_t, _u8p, _u8p, ctypes.c_uint32]
        lib.ref_sc_reduce.argtypes = [_u8p]
        lib.ref_sc_muladd.argtypes = [_u8p, _u8p, _u8p, _u8p]
        lib.ref_poh.argtypes = [_u8p, _u64p, ctypes.c_size_t]
        lib.ref_verify_many.argtypes = [_u8p, ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p, _u8p]

    @staticmethod
    def available(path=LIBREF):
        return os.path.exists(path)

    def get_checked_scalar(self, inp, out_prior=bytes(32), extra=0):
        out = np.array(bytearray(out_prior), np.uint8)
        i = _buf(inp)
        rc = self.lib.ref_get_checked_scalar(_ptr(out), _ptr(i), extra)
        return rc, out.tobytes()

    def check_small_order(self, enc):
        e = _buf(enc)
        return self.lib.ref_check_small_order(_ptr(e))

    def small_order_raw(self, enc):
        e = _buf(enc)
        return self.lib.ref_small_order_raw(_ptr(e))

    def verify(self, sig, msg, pk):
        s, m, p = _buf(sig), _buf(msg), _buf(pk)
        return self.lib.ref_verify(_ptr(s), _ptr(m), len(m), _ptr(p))

    def create_keypair(self, seed):
        pk = np.zeros(32, np.uint8)
        sk = np.zeros(64, np.uint8)
        s = _buf(seed)
        self.lib.ref_create_keypair(_ptr(pk), _ptr(sk), _ptr(s))
        return pk.tobytes(), sk.tobytes()

    def sign(self, msg, pk, sk):
        sig = np.zeros(64, np.uint8)
        m, p, s = _buf(msg), _buf(pk), _buf(sk)
        self.lib.ref_sign(_ptr(sig), _ptr(m), len(m), _ptr(p), _ptr(s))
        return sig.tobytes()

    def sign_repeat(self, msg, pk, sk, count):
        """`count` signatures of one message (a timing loop in C)."""
        sig = np.zeros(64, np.uint8)
        m, p, s = _buf(msg), _buf(pk), _buf(sk)
        self.lib.ref_sign_repeat(_ptr(sig), _ptr(m), len(m), _ptr(p), _ptr(s), count)
        return sig.tobytes()

    def sc_reduce(self, s64):
        b = np.array(bytearray(s64), np.uint8)
        self.lib.ref_sc_reduce(_ptr(b))
        return b[:32].tobytes()

    def sc_muladd(self, a, b, c):
        out = np.zeros(32, np.uint8)
        aa, bb, cc = _buf(a), _buf(b), _buf(c)
        self.lib.ref_sc_muladd(_ptr(out), _ptr(aa), _ptr(bb), _ptr(cc))
        return out.tobytes()

    def poh_verify_many(self, hashes, num_hashes):
        nh = np.ascontiguousarray(num_hashes, dtype=np.uint64)
        self.lib.ref_poh(_ptr(hashes), _ptr(nh, _u64p), nh.size)

    def verify_many(self, packets, message_lens, pk_off, sig_off, msg_off):
        n = len(sig_off)
        out = np.zeros(n, np.uint8)
        args = [np.ascontiguousarray(a, dtype=np.uint32) for a in
                (message_lens, pk_off, sig_off, msg_off)]
        pk = np.ascontiguousarray(packets, dtype=np.uint8)
        self.lib.ref_verify_many(_ptr(pk), n, *[_ptr(a, _u32p) for a in args], _ptr(out))
        return out
