"""ctypes binding of libpoh-simd.so, the CPU multi-buffer PoH library.

The reference's consumer (solana_entry) binds poh_verify_many_simd_avx512skx
and poh_verify_many_simd_avx2 from libpoh-simd.so and calls them on groups of
16 or 8 chains, padding the last group with zero counts
(src/poh-simd/poh-verify.ispc:3-8 is the interface).  This module binds the
same symbols and the extensions; it raises if the library is not built.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpoh-simd.so")

# every symbol include/poh_simd.h declares, with its chains per call
GROUPS = {"sse2": 4, "sse4": 4, "avx1": 8, "avx2": 8, "avx512skx": 16}
EXPORTS = tuple(f"poh_verify_many_simd_{k}" for k in GROUPS) + (
    "poh_simd_kernel", "poh_simd_use_kernel", "poh_verify_many_simd")
KERNELS = ("sha-ni", "avx512", "avx2", "portable")


class PohSimd:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} is not built; run __graft_entry__.build() (make -C solana-perf-libs_amd)")
        self.lib = lib = ctypes.CDLL(path)
        for k in GROUPS:
            f = getattr(lib, f"poh_verify_many_simd_{k}")
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
            f.restype = None
        lib.poh_simd_kernel.restype = ctypes.c_char_p
        lib.poh_simd_use_kernel.argtypes = [ctypes.c_char_p]
        lib.poh_simd_use_kernel.restype = ctypes.c_int
        lib.poh_verify_many_simd.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                             ctypes.c_uint32]
        lib.poh_verify_many_simd.restype = None

    def kernel(self):
        return self.lib.poh_simd_kernel().decode()

    def use_kernel(self, name):
        return self.lib.poh_simd_use_kernel(name.encode()) == 0

    def group(self, target, hashes, counts):
        """One call of poh_verify_many_simd_<target> on exactly GROUPS[target]
        chains: hashes (uint8, 32 per chain) are advanced in place."""
        n = GROUPS[target]
        counts = np.ascontiguousarray(counts, np.uint64)
        assert hashes.dtype == np.uint8 and hashes.flags.c_contiguous and hashes.size == 32 * n
        assert counts.size == n
        getattr(self.lib, f"poh_verify_many_simd_{target}")(hashes.ctypes.data, counts.ctypes.data)

    def verify_chains(self, hashes, counts, target="avx512skx"):
        """Any number of chains through one entry point, grouped and padded with
        zero counts as the reference's caller does; in place."""
        n = GROUPS[target]
        counts = np.ascontiguousarray(counts, np.uint64)
        m = counts.size
        assert hashes.dtype == np.uint8 and hashes.size == 32 * m
        pad = (-m) % n
        h = np.concatenate([hashes, np.zeros(32 * pad, np.uint8)]) if pad else hashes
        c = np.concatenate([counts, np.zeros(pad, np.uint64)]) if pad else counts
        f = getattr(self.lib, f"poh_verify_many_simd_{target}")
        for g in range(0, m + pad, n):
            f(h[32 * g:].ctypes.data, c[g:].ctypes.data)
        if pad:
            hashes[:] = h[:32 * m]

    def many(self, hashes, counts, threads=0):
        """poh_verify_many_simd: all chains, on `threads` threads (0 = all
        CPUs this process may use); in place."""
        counts = np.ascontiguousarray(counts, np.uint64)
        assert hashes.dtype == np.uint8 and hashes.flags.c_contiguous
        assert hashes.size == 32 * counts.size
        self.lib.poh_verify_many_simd(hashes.ctypes.data, counts.ctypes.data, counts.size,
                                      threads)


_LIB = None


def load():
    """The process-wide PohSimd instance (raises if the library is not built)."""
    global _LIB
    if _LIB is None:
        _LIB = PohSimd()
    return _LIB
