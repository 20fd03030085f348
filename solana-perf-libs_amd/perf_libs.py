"""ctypes binding of libcuda-crypt.so (the gfx950 build), shaped like the
reference's consumer `solana_perf::perf_libs`: it binds the same exported C
symbols with the same argument meaning, plus the device-resident extensions.

Importing this module loads `libcuda-crypt.so` from this directory and raises
if it is missing: there is no fallback implementation.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcuda-crypt.so")

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p


class GpuElems(ctypes.Structure):
    """gpu_Elems (src/cuda-ecc-ed25519/ed25519.h:35-38)."""
    _fields_ = [("elems", _u8p), ("num", ctypes.c_uint32)]


# every symbol include/cuda_crypt.h declares
EXPORTS = (
    "ed25519_create_seed", "ed25519_create_keypair", "ed25519_sign", "ed25519_sign_many",
    "ed25519_verify", "ed25519_verify_many", "ed25519_set_verbose", "ed25519_license",
    "ed25519_init", "cuda_host_register", "cuda_host_unregister",
    "ed25519_get_checked_scalar", "ed25519_check_packed_ge_small_order",
    "ed25519_free_gpu_mem", "poh_verify_many_set_verbose", "poh_verify_many",
    "ed25519_verify_many_device", "ed25519_sign_many_device", "poh_verify_many_device",
    "spl_poh_device_kernel", "spl_poh_kernel_choice",
    "ed25519_verify_many_device_noalloc",
    "spl_verify_engine_create", "spl_verify_engine_destroy", "spl_verify_engine_buffers",
    "spl_verify_engine_create_pipelined", "spl_verify_engine_depth",
    "spl_verify_engine_set_buffers", "spl_verify_engine_next", "spl_verify_engine_run_set",
    "spl_verify_engine_sync_set", "spl_verify_tune_save", "spl_verify_tune_load",
    "spl_verify_tune_query",
    "spl_verify_engine_run", "spl_verify_engine_sync", "spl_verify_engine_autotune",
    "spl_verify_engine_config", "spl_verify_engine_candidates", "spl_verify_engine_use",
    "spl_server_create", "spl_server_destroy",
    "spl_server_submit_verify", "spl_server_submit_poh", "spl_server_wait", "spl_server_stats",
    "spl_packet_offsets_device", "spl_packet_verdicts_device", "spl_verify_packets_device",
    "spl_verify_packets",
)

# Declared with C++ linkage, as the reference exports them: name -> symbol.
CXX_EXPORTS = {"poh_init": "_Z8poh_initv"}


def _p(a, t=_u8p):
    return a.ctypes.data_as(t)


class PerfLibs:
    def __init__(self, path=LIB_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(
                f"{path} is not built; run __graft_entry__.build() (make -C solana-perf-libs_amd)")
        self.lib = lib = ctypes.CDLL(path)
        lib.ed25519_create_seed.argtypes = [_u8p]
        lib.ed25519_create_seed.restype = ctypes.c_int
        lib.ed25519_create_keypair.argtypes = [_u8p, _u8p, _u8p]
        lib.ed25519_sign.argtypes = [_u8p, _u8p, ctypes.c_size_t, _u8p, _u8p]
        lib.ed25519_verify.argtypes = [_u8p, _u8p, ctypes.c_uint32, _u8p]
        lib.ed25519_verify.restype = ctypes.c_int
        many = [ctypes.POINTER(GpuElems), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                ctypes.c_uint32, _u32p, _u32p, _u32p, _u32p, _u8p, ctypes.c_uint8]
        lib.ed25519_verify_many.argtypes = many
        lib.ed25519_verify_many.restype = None
        lib.ed25519_sign_many.argtypes = many
        lib.ed25519_sign_many.restype = None
        lib.ed25519_set_verbose.argtypes = [ctypes.c_bool]
        lib.poh_verify_many_set_verbose.argtypes = [ctypes.c_bool]
        lib.ed25519_license.restype = ctypes.c_char_p
        lib.ed25519_init.restype = ctypes.c_bool
        self._poh_init = getattr(lib, CXX_EXPORTS["poh_init"])
        self._poh_init.argtypes = []
        self._poh_init.restype = ctypes.c_bool
        lib.cuda_host_register.argtypes = [_vp, ctypes.c_size_t, ctypes.c_uint]
        lib.cuda_host_register.restype = ctypes.c_int
        lib.cuda_host_unregister.argtypes = [_vp]
        lib.cuda_host_unregister.restype = ctypes.c_int
        lib.ed25519_get_checked_scalar.argtypes = [_u8p, _u8p]
        lib.ed25519_get_checked_scalar.restype = ctypes.c_int
        lib.ed25519_check_packed_ge_small_order.argtypes = [_u8p]
        lib.ed25519_check_packed_ge_small_order.restype = ctypes.c_int
        lib.ed25519_free_gpu_mem.restype = None
        lib.poh_verify_many.argtypes = [_u8p, _u64p, ctypes.c_size_t, ctypes.c_uint8]
        lib.poh_verify_many.restype = ctypes.c_int
        dev = [_vp, ctypes.c_uint64, _vp, _vp, _vp, _vp, ctypes.c_uint32, _vp, _vp]
        lib.ed25519_verify_many_device.argtypes = dev
        lib.ed25519_verify_many_device.restype = ctypes.c_int
        lib.ed25519_sign_many_device.argtypes = dev
        lib.ed25519_sign_many_device.restype = ctypes.c_int
        lib.ed25519_verify_many_device_noalloc.argtypes = dev
        lib.ed25519_verify_many_device_noalloc.restype = ctypes.c_int
        lib.poh_verify_many_device.argtypes = [_vp, _vp, ctypes.c_size_t, _vp]
        lib.poh_verify_many_device.restype = ctypes.c_int
        lib.spl_poh_device_kernel.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp]
        lib.spl_poh_device_kernel.restype = ctypes.c_int
        lib.spl_poh_kernel_choice.argtypes = [ctypes.c_size_t]
        lib.spl_poh_kernel_choice.restype = ctypes.c_int
        lib.spl_verify_engine_create.argtypes = [ctypes.c_uint32, ctypes.c_uint64]
        lib.spl_verify_engine_create.restype = _vp
        lib.spl_verify_engine_create_pipelined.argtypes = [ctypes.c_uint32, ctypes.c_uint64,
                                                           ctypes.c_uint32]
        lib.spl_verify_engine_create_pipelined.restype = _vp
        lib.spl_verify_engine_destroy.argtypes = [_vp]
        lib.spl_verify_engine_destroy.restype = None
        lib.spl_verify_engine_depth.argtypes = [_vp]
        lib.spl_verify_engine_depth.restype = ctypes.c_int
        lib.spl_verify_engine_set_buffers.argtypes = [_vp, ctypes.c_uint32] + [ctypes.POINTER(_vp)] * 7
        lib.spl_verify_engine_set_buffers.restype = ctypes.c_int
        lib.spl_verify_engine_next.argtypes = [_vp]
        lib.spl_verify_engine_next.restype = ctypes.c_int
        lib.spl_verify_engine_run_set.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32]
        lib.spl_verify_engine_run_set.restype = ctypes.c_int
        lib.spl_verify_engine_sync_set.argtypes = [_vp, ctypes.c_uint32]
        lib.spl_verify_engine_sync_set.restype = ctypes.c_int
        lib.spl_verify_tune_save.argtypes = [ctypes.c_char_p]
        lib.spl_verify_tune_save.restype = ctypes.c_int
        lib.spl_verify_tune_load.argtypes = [ctypes.c_char_p]
        lib.spl_verify_tune_load.restype = ctypes.c_int
        lib.spl_verify_tune_query.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_char_p,
                                              ctypes.c_size_t]
        lib.spl_verify_tune_query.restype = ctypes.c_int
        lib.spl_verify_engine_buffers.argtypes = [_vp] + [ctypes.POINTER(_vp)] * 7
        lib.spl_verify_engine_buffers.restype = ctypes.c_int
        lib.spl_verify_engine_run.argtypes = [_vp, ctypes.c_uint32]
        lib.spl_verify_engine_run.restype = ctypes.c_int
        lib.spl_verify_engine_sync.argtypes = [_vp]
        lib.spl_verify_engine_sync.restype = ctypes.c_int
        lib.spl_verify_engine_autotune.argtypes = [_vp, ctypes.c_uint32, ctypes.c_char_p,
                                                   ctypes.c_size_t]
        lib.spl_verify_engine_autotune.restype = ctypes.c_int
        lib.spl_verify_engine_config.argtypes = [_vp, ctypes.c_char_p, ctypes.c_size_t]
        lib.spl_verify_engine_config.restype = ctypes.c_int
        lib.spl_verify_engine_candidates.argtypes = [_vp, ctypes.c_char_p, ctypes.c_size_t]
        lib.spl_verify_engine_candidates.restype = ctypes.c_int
        lib.spl_verify_engine_use.argtypes = [_vp, ctypes.c_int]
        lib.spl_verify_engine_use.restype = ctypes.c_int
        lib.spl_server_create.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint32]
        lib.spl_server_create.restype = _vp
        lib.spl_server_destroy.argtypes = [_vp]
        lib.spl_server_destroy.restype = None
        lib.spl_server_submit_verify.argtypes = [_vp, _vp, ctypes.c_uint64, _vp, _vp, _vp, _vp,
                                                 ctypes.c_uint32, _vp]
        lib.spl_server_submit_verify.restype = ctypes.c_int64
        lib.spl_server_submit_poh.argtypes = [_vp, _vp, _vp, ctypes.c_size_t]
        lib.spl_server_submit_poh.restype = ctypes.c_int64
        lib.spl_server_wait.argtypes = [_vp, ctypes.c_int64, ctypes.c_int, _u64p]
        lib.spl_server_wait.restype = ctypes.c_int
        lib.spl_server_stats.argtypes = [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_int]
        lib.spl_server_stats.restype = ctypes.c_int
        u32 = ctypes.c_uint32
        lib.spl_packet_offsets_device.argtypes = [_vp, u32, u32, _vp, u32] + [_vp] * 7 + [_vp]
        lib.spl_packet_offsets_device.restype = ctypes.c_int
        lib.spl_packet_verdicts_device.argtypes = [_vp, _vp, u32, _vp, u32, _vp, _vp]
        lib.spl_packet_verdicts_device.restype = ctypes.c_int
        lib.spl_verify_packets_device.argtypes = [_vp, u32, u32, _vp, _vp, _vp, _vp]
        lib.spl_verify_packets_device.restype = ctypes.c_int
        lib.spl_verify_packets.argtypes = [ctypes.POINTER(GpuElems), u32, u32, u32, _u32p, _u8p,
                                           _u32p]
        lib.spl_verify_packets.restype = ctypes.c_int

    # ------------------------------------------------------------ host API
    def init(self):
        return bool(self.lib.ed25519_init())

    def poh_init(self):
        """The reference's C++-linkage poh_init() (poh_verify.cu:66-71)."""
        return bool(self._poh_init())

    def license(self):
        return self.lib.ed25519_license().decode()

    def set_verbose(self, v):
        self.lib.ed25519_set_verbose(bool(v))
        self.lib.poh_verify_many_set_verbose(bool(v))

    def create_keypair(self, seed):
        pk = np.zeros(32, np.uint8)
        sk = np.zeros(64, np.uint8)
        s = np.frombuffer(bytes(seed), np.uint8).copy()
        self.lib.ed25519_create_keypair(_p(pk), _p(sk), _p(s))
        return pk.tobytes(), sk.tobytes()

    def sign(self, msg, pk, sk):
        sig = np.zeros(64, np.uint8)
        m = np.frombuffer(bytes(msg) + b"\0", np.uint8).copy()
        p = np.frombuffer(bytes(pk), np.uint8).copy()
        s = np.frombuffer(bytes(sk), np.uint8).copy()
        self.lib.ed25519_sign(_p(sig), _p(m), len(msg), _p(p), _p(s))
        return sig.tobytes()

    def verify(self, sig, msg, pk):
        s = np.frombuffer(bytes(sig), np.uint8).copy()
        m = np.frombuffer(bytes(msg) + b"\0", np.uint8).copy()
        p = np.frombuffer(bytes(pk), np.uint8).copy()
        return self.lib.ed25519_verify(_p(s), _p(m), len(msg), _p(p))

    def get_checked_scalar(self, inp, out_prior=bytes(32)):
        out = np.frombuffer(bytes(out_prior), np.uint8).copy()
        i = np.frombuffer(bytes(inp), np.uint8).copy()
        rc = self.lib.ed25519_get_checked_scalar(_p(out), _p(i))
        return rc, out.tobytes()

    def check_packed_ge_small_order(self, enc):
        e = np.frombuffer(bytes(enc), np.uint8).copy()
        return self.lib.ed25519_check_packed_ge_small_order(_p(e))

    # ------------------------------------------------------------ batch API
    def verify_many(self, packet_batches, message_size, message_lens, pk_off, sig_off, msg_off,
                    use_non_default_stream=1):
        """packet_batches: list of uint8 arrays, each a whole number of packets of
        message_size bytes (the gpu_Elems list)."""
        return self._many(self.lib.ed25519_verify_many, packet_batches, message_size,
                          message_lens, pk_off, sig_off, msg_off, 1, use_non_default_stream)

    def sign_many(self, packet_batches, message_size, message_lens, pk_off, sk_off, msg_off,
                  use_non_default_stream=1):
        return self._many(self.lib.ed25519_sign_many, packet_batches, message_size,
                          message_lens, pk_off, sk_off, msg_off, 64, use_non_default_stream)

    @staticmethod
    def _prepare_many(fn, batches, message_size, lens, t1, t2, t3, out_width, nds):
        batches = [np.ascontiguousarray(b, dtype=np.uint8) for b in batches]
        elems = (GpuElems * max(1, len(batches)))()
        total_packets = 0
        for k, b in enumerate(batches):
            assert b.size % message_size == 0
            elems[k].elems = _p(b)
            elems[k].num = b.size // message_size
            total_packets += b.size // message_size
        tabs = [np.ascontiguousarray(t, dtype=np.uint32) for t in (lens, t1, t2, t3)]
        n = tabs[0].size
        out = np.zeros(max(1, n * out_width), np.uint8)
        args = (elems, len(batches), message_size, total_packets, n,
                *[_p(t, _u32p) for t in tabs], _p(out), nds)
        return fn, args, (batches, tabs), out[: n * out_width]

    @classmethod
    def _many(cls, fn, batches, message_size, lens, t1, t2, t3, out_width, nds):
        fn, args, _keep, out = cls._prepare_many(fn, batches, message_size, lens, t1, t2, t3,
                                                 out_width, nds)
        fn(*args)
        return out

    def prepare_verify_many(self, packet_batches, message_size, message_lens, pk_off, sig_off,
                            msg_off, use_non_default_stream=1):
        """The arguments of one ed25519_verify_many call, converted once (the
        gpu_Elems list, table pointers and an output buffer).  Each
        call_prepared() of the result is then the C call alone, as a compiled
        caller makes it; the returned mask is overwritten by the next call."""
        return self._prepare_many(self.lib.ed25519_verify_many, packet_batches, message_size,
                                  message_lens, pk_off, sig_off, msg_off, 1,
                                  use_non_default_stream)

    @staticmethod
    def call_prepared(prepared):
        fn, args, _keep, out = prepared
        fn(*args)
        return out

    def verify_packets(self, packet_batches, packet_stride, packet_sizes):
        """spl_verify_packets: wire-format packets (gpu_Elems list of whole
        slots) and their lengths in; (rc, verdict per packet, signatures per
        packet) out."""
        batches = [np.ascontiguousarray(b, dtype=np.uint8) for b in packet_batches]
        elems = (GpuElems * max(1, len(batches)))()
        total = 0
        for k, b in enumerate(batches):
            assert b.size % packet_stride == 0
            elems[k].elems = _p(b)
            elems[k].num = b.size // packet_stride
            total += b.size // packet_stride
        sizes = np.ascontiguousarray(packet_sizes, dtype=np.uint32)
        assert sizes.size == total
        ok = np.zeros(max(1, total), np.uint8)
        lens = np.zeros(max(1, total), np.uint32)
        rc = self.lib.spl_verify_packets(elems, len(batches), packet_stride, total,
                                         _p(sizes, _u32p), _p(ok), _p(lens, _u32p))
        return rc, ok[:total], lens[:total]

    def poh_verify_many(self, hashes, num_hashes, use_non_default_stream=1):
        """hashes: uint8 [n*32], updated in place. Returns the ABI's int."""
        nh = np.ascontiguousarray(num_hashes, dtype=np.uint64)
        assert hashes.dtype == np.uint8 and hashes.flags.c_contiguous and hashes.size == 32 * nh.size
        return self.lib.poh_verify_many(_p(hashes), _p(nh, _u64p), nh.size, use_non_default_stream)

    # ------------------------------------------------------------ device API
    def verify_many_device(self, d_packets, packets_size, d_lens, d_pk, d_sig, d_msg, n, d_out,
                           stream=0):
        return self.lib.ed25519_verify_many_device(d_packets, packets_size, d_lens, d_pk, d_sig,
                                                   d_msg, n, d_out, stream)

    def tune_save(self, path):
        """Writes the autotuner's records to `path`; returns how many."""
        return self.lib.spl_verify_tune_save(os.fsencode(path))

    def tune_load(self, path):
        """Takes the records in `path`; returns how many (< 0: none, bad file)."""
        return self.lib.spl_verify_tune_load(os.fsencode(path))

    def tune_query(self, device, n):
        """(from_record, configuration name) used for n signatures on `device`."""
        buf = ctypes.create_string_buffer(256)
        rc = self.lib.spl_verify_tune_query(device, n, buf, len(buf))
        if rc < 0:
            raise ValueError(f"spl_verify_tune_query({device}, {n}) = {rc}")
        return bool(rc), buf.value.decode()

    def verify_many_device_noalloc(self, d_packets, packets_size, d_lens, d_pk, d_sig, d_msg, n,
                                   d_out, stream=0):
        """ed25519_verify_many_device with the library's fixed scratch (no
        allocation; at most 16,384 signatures; needs init() first)."""
        return self.lib.ed25519_verify_many_device_noalloc(d_packets, packets_size, d_lens, d_pk,
                                                           d_sig, d_msg, n, d_out, stream)

    def sign_many_device(self, d_packets, packets_size, d_lens, d_pk, d_sk, d_msg, n, d_out,
                         stream=0):
        return self.lib.ed25519_sign_many_device(d_packets, packets_size, d_lens, d_pk, d_sk,
                                                 d_msg, n, d_out, stream)

    def poh_verify_many_device(self, d_hashes, d_counts, n, stream=0):
        return self.lib.poh_verify_many_device(d_hashes, d_counts, n, stream)

    # kernel names for poh_device_kernel (spl_poh_device_kernel's numbering)
    POH_KERNELS = {"auto": 0, "chain": 1, "split": 2, "pair": 3}

    def poh_device_kernel(self, d_hashes, d_counts, n, kernel, stream=0):
        """poh_verify_many_device with the kernel named ("chain", "split",
        "pair" or "auto")."""
        return self.lib.spl_poh_device_kernel(d_hashes, d_counts, n, self.POH_KERNELS[kernel],
                                              stream)

    def poh_kernel_choice(self, n):
        """Name of the PoH kernel the library uses for n chains ("chain",
        "split", "pair"), or None without a GPU."""
        k = self.lib.spl_poh_kernel_choice(n)
        return next((name for name, v in self.POH_KERNELS.items() if v == k and k), None)

    def packet_offsets_device(self, d_packets, num, stride, d_sizes, max_sigs, d_sig_lens,
                              d_sig_base, d_lens, d_pk, d_sig, d_msg, d_counts, stream=0):
        return self.lib.spl_packet_offsets_device(d_packets, num, stride, d_sizes, max_sigs,
                                                  d_sig_lens, d_sig_base, d_lens, d_pk, d_sig,
                                                  d_msg, d_counts, stream)

    def packet_verdicts_device(self, d_sig_lens, d_sig_base, num, d_sig_out, num_sigs, d_ok,
                               stream=0):
        return self.lib.spl_packet_verdicts_device(d_sig_lens, d_sig_base, num, d_sig_out,
                                                   num_sigs, d_ok, stream)

    def verify_packets_device(self, d_packets, num, stride, d_sizes, d_ok, d_sig_lens=None,
                              stream=0):
        return self.lib.spl_verify_packets_device(d_packets, num, stride, d_sizes, d_ok,
                                                  d_sig_lens, stream)


class VerifyEngine:
    """spl_verify_engine: device buffers + graph-captured tuned verify launch,
    in `depth` buffer sets (batches in flight).  The d_* attributes are set
    0's buffers; set_buffers(k) gives set k's."""

    def __init__(self, lib, max_signatures, max_packet_bytes, depth=1):
        self.lib = lib.lib
        self.h = self.lib.spl_verify_engine_create_pipelined(max_signatures, max_packet_bytes,
                                                             depth)
        if not self.h:
            raise RuntimeError("spl_verify_engine_create failed")
        (self.d_packets, self.d_lens, self.d_pk, self.d_sig, self.d_msg, self.d_out,
         self.stream) = self.set_buffers(0)
        self.cap = max_signatures
        self.depth = depth

    def set_buffers(self, k):
        """(packets, lens, pk, sig, msg, out, stream) device addresses of set k."""
        ptrs = [_vp() for _ in range(7)]
        rc = self.lib.spl_verify_engine_set_buffers(self.h, k, *[ctypes.byref(p) for p in ptrs])
        assert rc == 0
        return tuple(p.value for p in ptrs)

    def run(self, n):
        return self.lib.spl_verify_engine_run(self.h, n)

    def next(self):
        """Index of the next buffer set, once its previous batch has finished."""
        return self.lib.spl_verify_engine_next(self.h)

    def run_set(self, k, n):
        return self.lib.spl_verify_engine_run_set(self.h, k, n)

    def sync_set(self, k):
        return self.lib.spl_verify_engine_sync_set(self.h, k)

    def sync(self):
        return self.lib.spl_verify_engine_sync(self.h)

    def autotune(self, n):
        buf = ctypes.create_string_buffer(8192)
        rc = self.lib.spl_verify_engine_autotune(self.h, n, buf, len(buf))
        if rc != 0:
            raise RuntimeError(f"autotune failed: {rc}")
        return buf.value.decode()

    def candidates(self):
        buf = ctypes.create_string_buffer(8192)
        n = self.lib.spl_verify_engine_candidates(self.h, buf, len(buf))
        names = [l for l in buf.value.decode().splitlines() if l]
        assert len(names) == n
        return names

    def use(self, index):
        rc = self.lib.spl_verify_engine_use(self.h, index)
        if rc != 0:
            raise RuntimeError(f"spl_verify_engine_use({index}) failed: {rc}")

    def config(self):
        buf = ctypes.create_string_buffer(256)
        self.lib.spl_verify_engine_config(self.h, buf, len(buf))
        return buf.value.decode()

    def close(self):
        if self.h:
            self.lib.spl_verify_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """spl_server: continuous-batching request loop for verify and PoH requests.

    The numpy arrays of a request are referenced here until its wait()
    returns, as the C ABI requires."""

    def __init__(self, lib, hbm_budget_bytes=0, max_batch_signatures=0, verify_streams=0,
                 poh_streams=0):
        self.lib = lib.lib
        self.h = self.lib.spl_server_create(hbm_budget_bytes, max_batch_signatures,
                                            verify_streams, poh_streams)
        if not self.h:
            raise RuntimeError("spl_server_create failed (no GPU, or bad arguments)")
        self._live = {}
        self._times = (ctypes.c_uint64 * 3)()

    def prepare_verify(self, packets, lens, pk_off, sig_off, msg_off, out=None):
        """The arguments of one verify request (ed25519_verify_many layout,
        offsets relative to `packets`), converted for the C call ahead of
        time.  Returns (prepared, out); submit with submit_prepared()."""
        tabs = [np.ascontiguousarray(t, dtype=np.uint32) for t in (lens, pk_off, sig_off, msg_off)]
        packets = np.ascontiguousarray(packets, dtype=np.uint8)
        n = tabs[0].size
        if out is None:
            out = np.zeros(max(n, 1), np.uint8)
        args = (packets.ctypes.data, packets.size, *[x.ctypes.data for x in tabs], n,
                out.ctypes.data)
        return (self.lib.spl_server_submit_verify, args, (packets, tabs, out)), out

    def prepare_poh(self, hashes, counts):
        """hashes: uint8 [32*n], hashed in place on completion."""
        counts = np.ascontiguousarray(counts, dtype=np.uint64)
        assert hashes.dtype == np.uint8 and hashes.flags.c_contiguous
        assert hashes.size == 32 * counts.size
        args = (hashes.ctypes.data, counts.ctypes.data, counts.size)
        return (self.lib.spl_server_submit_poh, args, (hashes, counts))

    def submit_prepared(self, prepared):
        """Submits a request from prepare_verify / prepare_poh; returns the
        ticket (> 0) or a refusal code."""
        fn, args, keep = prepared
        t = fn(self.h, *args)
        if t > 0:
            self._live[t] = keep
        return t

    def submit_verify(self, packets, lens, pk_off, sig_off, msg_off, out=None):
        """One request in the ed25519_verify_many layout (offsets relative to
        `packets`).  Returns (ticket, out); out is filled on completion."""
        prepared, out = self.prepare_verify(packets, lens, pk_off, sig_off, msg_off, out)
        return self.submit_prepared(prepared), out

    def submit_poh(self, hashes, counts):
        """hashes: uint8 [32*n], hashed in place on completion."""
        return self.submit_prepared(self.prepare_poh(hashes, counts))

    def wait(self, ticket, timeout_ms=-1):
        """Returns (rc, (t_submit, t_admit, t_done) in CLOCK_MONOTONIC ns)."""
        rc = self.lib.spl_server_wait(self.h, ticket, timeout_ms, self._times)
        times = tuple(int(x) for x in self._times)
        if rc in (0, 3):
            self._live.pop(ticket, None)
        return rc, times

    def stats(self, reset=False):
        import json
        buf = ctypes.create_string_buffer(4096)
        rc = self.lib.spl_server_stats(self.h, buf, len(buf), int(reset))
        assert rc == 0
        return json.loads(buf.value.decode())

    def close(self):
        if self.h:
            self.lib.spl_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def abandon(self):
        """Drop the handle without spl_server_destroy (which waits for every
        submitted request): for a server whose work did not complete.  The
        request buffers are kept alive with it."""
        if self.h:
            _ABANDONED.append((self.h, self._live))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_LIB = None
_ABANDONED = []


def load():
    """The process-wide PerfLibs instance (raises if the library is not built)."""
    global _LIB
    if _LIB is None:
        _LIB = PerfLibs()
    return _LIB
