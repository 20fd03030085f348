// Packet parsing on the device: the step before the verify path, and the
// per-packet verdict after it.
//
// The reference consumes per-signature tables (message length, public-key,
// signature and message offsets: gpu_ctx.cu:133-136, verify.cu:240-340) and
// leaves building them to its caller, which walks every Solana wire-format
// packet on the host (agave perf/src/sigverify.rs `generate_offsets`, one row
// per signature) and afterwards folds the per-signature mask into one verdict
// per packet (`copy_return_values`).  Here both steps run on the GPU, so raw
// packets plus their lengths go in and one byte per packet comes out:
//
//   packet_parse_kernel   one lane per packet: the short-vec signature count,
//                         the optional version-0 prefix, the message header
//                         and the account-key count; the rules are those of
//                         `do_get_packet_offsets` (a packet that breaks any of
//                         them gets no signatures); per-block signature sums
//   packet_scan_kernel    one block: exclusive prefix of the block sums and
//                         the batch total
//   packet_emit_kernel    block-local exclusive scan + the block prefix gives
//                         each packet its first row; it writes one row per
//                         signature
//   (verify)              the persistent verify kernel, whose signature count
//                         is read from device memory, so nothing waits on
//                         the host between parsing and verifying
//   packet_verdict_kernel one lane per packet: at least one signature, and
//                         every one of them verified
//
// All of it is HBM/latency-bound byte work on a few cache lines per packet;
// the verify kernel between them is ~1000x the cost.
#include <hip/hip_runtime.h>

#include "verify_kernels.h"
#include "../../include/cuda_crypt.h"

namespace spl {

constexpr uint32_t kPacketThreads = 256;
constexpr uint32_t kScanThreads = 1024;

// The largest signature count a packet of `stride` bytes can pass with: the
// count prefix (>= 1 byte), 64 n signature bytes, the 3-byte header, the key
// count (>= 1 byte) and n keys of 32 bytes must fit, so 96 n + 5 <= size.
__host__ __device__ inline uint32_t max_sigs_per_packet(uint32_t stride) {
    return stride >= 101 ? (stride - 5) / 96 : 0;
}

// decode_shortu16_len: up to three bytes of 7 bits, low group first; fails on
// running out of bytes, a zero byte after the first, a third byte with the
// continuation bit, or a value above 0xFFFF.
__device__ inline bool short_u16(const uint8_t *p, uint32_t pos, uint32_t end, uint32_t &val,
                                 uint32_t &len) {
    uint32_t v = 0;
    for (uint32_t k = 0; k < 3; ++k) {
        if (pos + k >= end) return false;
        const uint32_t b = p[pos + k];
        if (b == 0 && k != 0) return false;
        if (k == 2 && (b & 0x80u)) return false;
        v |= (b & 0x7Fu) << (7 * k);
        if (v > 0xFFFFu) return false;
        if (!(b & 0x80u)) {
            val = v;
            len = k + 1;
            return true;
        }
    }
    return false;
}

// (n, signature start, message start, key start), packet-relative; n = 0 when
// the packet breaks a rule.
__device__ inline uint4 parse_packet(const uint8_t *p, uint32_t size, uint32_t stride) {
    const uint4 none = make_uint4(0, 0, 0, 0);
    if (size < 65 || size > stride) return none;
    uint32_t n, a;
    if (!short_u16(p, 0, size, n, a)) return none;
    const uint32_t m = a + 64u * n;
    if (m >= size) return none;
    uint32_t h = m;
    const uint32_t prefix = p[m];
    if (prefix & 0x80u) {
        if (prefix & 0x7Fu) return none;       // only message version 0
        h = m + 1;
    }
    if (h + 4 > size) return none;
    const uint32_t required = p[h], readonly_signed = p[h + 1];
    if (required <= readonly_signed || required != n) return none;
    uint32_t k, b;
    if (!short_u16(p, h + 3, size, k, b)) return none;
    const uint32_t keys = h + 3 + b;
    if (keys + 32u * k > size || k < n) return none;
    return make_uint4(n, a, m, keys);
}

// Exclusive prefix of v over the block (kPacketThreads lanes); *total gets the
// block sum.
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t buf[2][kPacketThreads];
    const uint32_t t = threadIdx.x;
    int cur = 0;
    buf[0][t] = v;
    __syncthreads();
    for (uint32_t d = 1; d < kPacketThreads; d <<= 1) {
        const uint32_t x = buf[cur][t] + (t >= d ? buf[cur][t - d] : 0u);
        buf[cur ^ 1][t] = x;
        cur ^= 1;
        __syncthreads();
    }
    const uint32_t incl = buf[cur][t];
    *total = buf[cur][kPacketThreads - 1];
    return incl - v;
}

__global__ void __launch_bounds__(kPacketThreads)
packet_parse_kernel(const uint8_t *packets, uint32_t num, uint32_t stride, const uint32_t *sizes,
                    uint4 *info, uint32_t *block_sums) {
    const uint32_t i = blockIdx.x * kPacketThreads + threadIdx.x;
    uint4 r = make_uint4(0, 0, 0, 0);
    if (i < num) {
        r = parse_packet(packets + (size_t)i * stride, sizes[i], stride);
        info[i] = r;
    }
    uint32_t total;
    (void)block_exclusive_scan(r.x, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// block_sums[0..nb) -> exclusive prefixes; counts[0] = batch total,
// counts[1] = min(total, cap) (the rows that exist).
__global__ void __launch_bounds__(kScanThreads)
packet_scan_kernel(uint32_t *block_sums, uint32_t nb, uint32_t cap, uint32_t *counts) {
    __shared__ uint32_t buf[2][kScanThreads];
    __shared__ uint32_t carry;
    const uint32_t t = threadIdx.x;
    if (t == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += kScanThreads) {
        const uint32_t v = base + t < nb ? block_sums[base + t] : 0u;
        int cur = 0;
        buf[0][t] = v;
        __syncthreads();
        for (uint32_t d = 1; d < kScanThreads; d <<= 1) {
            const uint32_t x = buf[cur][t] + (t >= d ? buf[cur][t - d] : 0u);
            buf[cur ^ 1][t] = x;
            cur ^= 1;
            __syncthreads();
        }
        const uint32_t c = carry;
        if (base + t < nb) block_sums[base + t] = c + buf[cur][t] - v;
        __syncthreads();
        if (t == 0) carry = c + buf[cur][kScanThreads - 1];
        __syncthreads();
    }
    if (t == 0) {
        counts[0] = carry;
        counts[1] = carry < cap ? carry : cap;
    }
}

struct PacketTables {
    uint32_t *sig_lens, *sig_base;             // per packet
    uint32_t *lens, *pk, *sig, *msg;           // per signature row
    uint32_t cap;                              // rows available
};

__global__ void __launch_bounds__(kPacketThreads)
packet_emit_kernel(const uint4 *info, const uint32_t *block_prefix, const uint32_t *sizes,
                   uint32_t num, uint32_t stride, PacketTables out) {
    const uint32_t i = blockIdx.x * kPacketThreads + threadIdx.x;
    const uint4 r = i < num ? info[i] : make_uint4(0, 0, 0, 0);
    uint32_t total;
    const uint32_t first = block_exclusive_scan(r.x, &total) + block_prefix[blockIdx.x];
    if (i >= num) return;
    out.sig_lens[i] = r.x;
    out.sig_base[i] = first;
    const uint32_t base = i * stride;            // the batch is < 4 GiB
    const uint32_t len = sizes[i] - r.z;
    for (uint32_t k = 0; k < r.x; ++k) {
        const uint32_t j = first + k;
        if (j >= out.cap) break;
        out.lens[j] = len;
        out.pk[j] = base + r.w + 32u * k;
        out.sig[j] = base + r.y + 64u * k;
        out.msg[j] = base + r.z;
    }
}

__global__ void __launch_bounds__(kPacketThreads)
packet_verdict_kernel(const uint32_t *sig_lens, const uint32_t *sig_base, uint32_t num,
                      const uint8_t *sig_out, uint32_t cap, uint8_t *ok) {
    const uint32_t i = blockIdx.x * kPacketThreads + threadIdx.x;
    if (i >= num) return;
    const uint32_t n = sig_lens[i], b = sig_base[i];
    uint32_t v = n != 0;
    for (uint32_t k = 0; k < n && v; ++k) {
        const uint32_t j = b + k;
        v = j < cap && sig_out[j] == 1;
    }
    ok[i] = (uint8_t)v;
}

// ------------------------------------------------------------------ host side

static bool packet_args_ok(uint32_t num, uint32_t stride) {
    return stride != 0 && (uint64_t)num * stride <= 0xFFFFFFFFull;
}

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Scratch for the parse kernels: info[num] (16 B) and the block sums.
struct ParseScratch {
    size_t info, sums, bytes;
};

static ParseScratch parse_scratch(uint32_t num) {
    ParseScratch s;
    const uint32_t nb = (num + kPacketThreads - 1) / kPacketThreads;
    s.info = 0;
    s.sums = align256(16ull * num);
    s.bytes = s.sums + align256(4ull * (nb + 1));
    return s;
}

// Parse, scan and emit; counts[0] = signatures in the batch, counts[1] = rows
// written (at most out.cap).
static hipError_t enqueue_offsets(const uint8_t *d_packets, uint32_t num, uint32_t stride,
                                  const uint32_t *d_sizes, const PacketTables &out,
                                  uint32_t *counts, uint8_t *scratch, const ParseScratch &ps,
                                  hipStream_t s) {
    uint4 *info = reinterpret_cast<uint4 *>(scratch + ps.info);
    uint32_t *sums = reinterpret_cast<uint32_t *>(scratch + ps.sums);
    const uint32_t nb = (num + kPacketThreads - 1) / kPacketThreads;
    if (nb)
        hipLaunchKernelGGL(packet_parse_kernel, dim3(nb), dim3(kPacketThreads), 0, s, d_packets,
                           num, stride, d_sizes, info, sums);
    hipLaunchKernelGGL(packet_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, sums, nb, out.cap,
                       counts);
    if (nb)
        hipLaunchKernelGGL(packet_emit_kernel, dim3(nb), dim3(kPacketThreads), 0, s, info, sums,
                           d_sizes, num, stride, out);
    return hipGetLastError();
}

// Everything spl_verify_packets* needs beyond its inputs and outputs.
struct VerifyPacketsScratch {
    ParseScratch ps;
    size_t parse, counts, tabs, sig_lens, sig_base, sig_out, work, verify, bytes;
    uint32_t cap;
};

// Parsing is followed by the persistent verify launch (signature count read
// from device memory): W = 4, half-size scalars, no sorting, 4 waves per CU.
static const VerifyConfig kPacketVerifyConfig = {1, 4, 1, 0, 4, 1, 1};

static VerifyPacketsScratch verify_packets_scratch(uint32_t num, uint32_t stride, int cus) {
    VerifyPacketsScratch v;
    v.ps = parse_scratch(num);
    v.cap = num * max_sigs_per_packet(stride);
    v.parse = 0;
    v.counts = align256(v.ps.bytes);
    v.tabs = v.counts + 256;
    v.sig_lens = v.tabs + align256(16ull * v.cap);
    v.sig_base = v.sig_lens + align256(4ull * num);
    v.sig_out = v.sig_base + align256(4ull * num);
    v.work = v.sig_out + align256((size_t)v.cap + 16);
    v.verify = v.work + 256;
    v.bytes = v.verify + verify_scratch_bytes(kPacketVerifyConfig, 0, cus);
    return v;
}

static hipError_t enqueue_verify_packets(const ed25519_tables *T, int cus,
                                         const uint8_t *d_packets, uint32_t num, uint32_t stride,
                                         const uint32_t *d_sizes, uint8_t *d_ok,
                                         uint32_t *d_sig_lens_out, uint8_t *scratch,
                                         const VerifyPacketsScratch &v, hipStream_t s) {
    uint32_t *tab = reinterpret_cast<uint32_t *>(scratch + v.tabs);
    PacketTables out;
    out.cap = v.cap;
    out.lens = tab;
    out.pk = tab + v.cap;
    out.sig = tab + 2ull * v.cap;
    out.msg = tab + 3ull * v.cap;
    out.sig_lens = d_sig_lens_out ? d_sig_lens_out
                                  : reinterpret_cast<uint32_t *>(scratch + v.sig_lens);
    out.sig_base = reinterpret_cast<uint32_t *>(scratch + v.sig_base);
    uint32_t *counts = reinterpret_cast<uint32_t *>(scratch + v.counts);
    hipError_t e = enqueue_offsets(d_packets, num, stride, d_sizes, out, counts,
                                   scratch + v.parse, v.ps, s);
    if (e != hipSuccess) return e;
    uint8_t *sig_out = scratch + v.sig_out;
    if (v.cap) {
        VerifyJob job = make_verify_job(d_packets, (uint64_t)num * stride, out.lens, out.pk,
                                        out.sig, out.msg, sig_out, scratch + v.verify, T);
        job.n_dev = counts + 1;
        job.work = reinterpret_cast<uint32_t *>(scratch + v.work);
        job.order = nullptr;
        if ((e = launch_verify_cfg(kPacketVerifyConfig, job, v.cap, cus, s)) != hipSuccess)
            return e;
    }
    const uint32_t nb = (num + kPacketThreads - 1) / kPacketThreads;
    hipLaunchKernelGGL(packet_verdict_kernel, dim3(nb), dim3(kPacketThreads), 0, s, out.sig_lens,
                       out.sig_base, num, sig_out, v.cap, d_ok);
    return hipGetLastError();
}

}  // namespace spl

using namespace spl;

extern "C" {

int spl_packet_offsets_device(const uint8_t *d_packets, uint32_t num_packets,
                              uint32_t packet_stride, const uint32_t *d_packet_sizes,
                              uint32_t max_signatures, uint32_t *d_sig_lens,
                              uint32_t *d_sig_base, uint32_t *d_message_lens,
                              uint32_t *d_public_key_offsets, uint32_t *d_signature_offsets,
                              uint32_t *d_message_start_offsets, uint32_t *d_counts,
                              void *stream) {
    if (!packet_args_ok(num_packets, packet_stride) || !d_counts) return 2;
    if (num_packets && (!d_packets || !d_packet_sizes || !d_sig_lens || !d_sig_base)) return 2;
    if (max_signatures && (!d_message_lens || !d_public_key_offsets || !d_signature_offsets ||
                           !d_message_start_offsets))
        return 2;
    if (pool_init() <= 0) return 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    const ParseScratch ps = parse_scratch(num_packets);
    uint8_t *scr = static_cast<uint8_t *>(stream_scratch(dev, (hipStream_t)stream, ps.bytes));
    if (!scr) {
        fprintf(stderr, "spl_packet_offsets_device: scratch allocation failed\n");
        return 3;
    }
    PacketTables out{d_sig_lens, d_sig_base, d_message_lens, d_public_key_offsets,
                     d_signature_offsets, d_message_start_offsets, max_signatures};
    hipError_t e = enqueue_offsets(d_packets, num_packets, packet_stride, d_packet_sizes, out,
                                   d_counts, scr, ps, (hipStream_t)stream);
    if (e != hipSuccess) {
        fprintf(stderr, "spl_packet_offsets_device: launch failed: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

int spl_packet_verdicts_device(const uint32_t *d_sig_lens, const uint32_t *d_sig_base,
                               uint32_t num_packets, const uint8_t *d_sig_out,
                               uint32_t num_signatures, uint8_t *d_packet_ok, void *stream) {
    if (num_packets == 0) return 0;
    if (!d_sig_lens || !d_sig_base || !d_packet_ok || (num_signatures && !d_sig_out)) return 2;
    if (pool_init() <= 0) return 1;
    const uint32_t nb = (num_packets + kPacketThreads - 1) / kPacketThreads;
    hipLaunchKernelGGL(packet_verdict_kernel, dim3(nb), dim3(kPacketThreads), 0,
                       (hipStream_t)stream, d_sig_lens, d_sig_base, num_packets, d_sig_out,
                       num_signatures, d_packet_ok);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "spl_packet_verdicts_device: launch failed: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

int spl_verify_packets_device(const uint8_t *d_packets, uint32_t num_packets,
                              uint32_t packet_stride, const uint32_t *d_packet_sizes,
                              uint8_t *d_packet_ok, uint32_t *d_sig_lens, void *stream) {
    if (num_packets == 0) return 0;
    if (!packet_args_ok(num_packets, packet_stride) || !d_packets || !d_packet_sizes ||
        !d_packet_ok)
        return 2;
    if (pool_init() <= 0) return 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    const ed25519_tables *T = device_tables(dev);
    if (!T) return 1;
    const int cus = device_cus(dev);
    const VerifyPacketsScratch v = verify_packets_scratch(num_packets, packet_stride, cus);
    uint8_t *scr = static_cast<uint8_t *>(stream_scratch(dev, (hipStream_t)stream, v.bytes));
    if (!scr) {
        fprintf(stderr, "spl_verify_packets_device: scratch allocation failed\n");
        return 3;
    }
    hipError_t e = enqueue_verify_packets(T, cus, d_packets, num_packets, packet_stride,
                                          d_packet_sizes, d_packet_ok, d_sig_lens, scr, v,
                                          (hipStream_t)stream);
    if (e != hipSuccess) {
        fprintf(stderr, "spl_verify_packets_device: launch failed: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

int spl_verify_packets(const gpu_Elems *elems, uint32_t num_elems, uint32_t packet_stride,
                       uint32_t total_packets, const uint32_t *packet_sizes, uint8_t *packet_ok,
                       uint32_t *sig_lens) {
    if (total_packets == 0) return 0;
    if (!packet_args_ok(total_packets, packet_stride) || !elems || !packet_sizes || !packet_ok)
        return 2;
    Lease l = acquire_queue();
    if (!l.q) {
        fprintf(stderr, "spl_verify_packets: no GPU available; all %u packets rejected\n",
                total_packets);
        memset(packet_ok, 0, total_packets);
        if (sig_lens) memset(sig_lens, 0, 4ull * total_packets);
        return 1;
    }
    QueueCtx &q = *l.q;
    hipStream_t s = q.stream;
    const int cus = device_cus(l.dev->id);
    const uint64_t pbytes = (uint64_t)total_packets * packet_stride;
    const uint64_t sizes_at = (pbytes + 15) & ~15ull;
    const uint64_t in_bytes = sizes_at + 4ull * total_packets;
    const uint64_t ok_at = (in_bytes + 15) & ~15ull;
    const uint64_t lens_at = (ok_at + total_packets + 15) & ~15ull;
    const VerifyPacketsScratch v = verify_packets_scratch(total_packets, packet_stride, cus);
    uint8_t *h = nullptr, *d_in = nullptr, *d_res = nullptr;
    uint64_t sent = 0;
    SPL_HIP_TRY(q.h_stage.reserve(lens_at + 4ull * total_packets + 16));
    SPL_HIP_TRY(q.d_packets.reserve(in_bytes + 16));
    SPL_HIP_TRY(q.d_out.reserve(lens_at - ok_at + 4ull * total_packets + 16));
    SPL_HIP_TRY(q.d_scratch.reserve(v.bytes));
    h = q.h_stage.as<uint8_t>();
    d_in = q.d_packets.as<uint8_t>();
    d_res = q.d_out.as<uint8_t>();
    SPL_HIP_TRY(stage_elems(h, d_in, elems, num_elems, packet_stride, pbytes, &sent, s));
    memset(h + pbytes, 0, sizes_at - pbytes);
    memcpy(h + sizes_at, packet_sizes, 4ull * total_packets);
    SPL_HIP_TRY(launch_copy16(d_in + sent, h + sent, in_bytes - sent, s));
    SPL_HIP_TRY(enqueue_verify_packets(l.dev->d_tables, cus, d_in, total_packets, packet_stride,
                                       reinterpret_cast<const uint32_t *>(d_in + sizes_at), d_res,
                                       reinterpret_cast<uint32_t *>(d_res + (lens_at - ok_at)),
                                       q.d_scratch.as<uint8_t>(), v, s));
    SPL_HIP_TRY(launch_copy16(h + ok_at, d_res, lens_at - ok_at + 4ull * total_packets, s));
    SPL_HIP_TRY(hipStreamSynchronize(s));
    memcpy(packet_ok, h + ok_at, total_packets);
    if (sig_lens) memcpy(sig_lens, h + lens_at, 4ull * total_packets);
    release_queue(l);
    return 0;
fail:
    memset(packet_ok, 0, total_packets);
    if (sig_lens) memset(sig_lens, 0, 4ull * total_packets);
    release_queue(l);
    return 3;
}

}  // extern "C"
