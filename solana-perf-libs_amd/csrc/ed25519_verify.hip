// Batched Ed25519 verification behind the reference's ed25519_verify_many ABI.
//
// Reference: src/cuda-ecc-ed25519/verify.cu:176-340 — three kernels
// (verify_kernel: S check, R check, SHA-512, sc_reduce; gen_lookup_kernel:
// decode A, build the 8-entry table in global memory; scalar_double_kernel:
// double-scalar multiplication and compare) with a per-signature mask threaded
// through global memory.  Here one fused kernel does the whole check per
// signature; the odd-multiple table of -A lives in the lane's private segment.
#include "runtime.h"
#include "../../include/cuda_crypt.h"

namespace spl {

__global__ void __launch_bounds__(64)
verify_kernel(const uint8_t *__restrict__ packets, uint64_t packets_size,
              const uint32_t *__restrict__ lens, const uint32_t *__restrict__ pk_off,
              const uint32_t *__restrict__ sig_off, const uint32_t *__restrict__ msg_off,
              uint32_t n, uint8_t *__restrict__ out, const ed25519_tables *__restrict__ T) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t so = sig_off[i], po = pk_off[i], mo = msg_off[i], ml = lens[i];
    if (so + 64 > packets_size || po + 32 > packets_size || mo + ml > packets_size) {
        out[i] = 0;    // offsets outside the packet buffer: reject
        return;
    }
    uint32_t sig[16], pk[8];
    load_words(sig, packets + so, 16);
    load_words(pk, packets + po, 8);
    MsgReader msg{packets + mo, (uint32_t)ml};
    ge_cached Ai[9];
    out[i] = (uint8_t)verify_one(sig, pk, msg, *T, Ai);
}

hipError_t launch_verify(const uint8_t *d_packets, uint64_t packets_size, const uint32_t *d_lens,
                         const uint32_t *d_pk, const uint32_t *d_sig, const uint32_t *d_msg,
                         uint32_t n, uint8_t *d_out, const ed25519_tables *d_tables,
                         hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t block = 64;
    hipLaunchKernelGGL(verify_kernel, dim3((n + block - 1) / block), dim3(block), 0, s,
                       d_packets, packets_size, d_lens, d_pk, d_sig, d_msg, n, d_out, d_tables);
    return hipGetLastError();
}

// Packs gpu_Elems and the four offset tables into the queue's pinned staging
// buffer and copies them to the device (the reference's setup_gpu_ctx,
// gpu_ctx.cu:75-144, with one copy per table instead of one per elem).
hipError_t stage_packets(QueueCtx &q, const gpu_Elems *elems, uint32_t num_elems,
                         uint32_t message_size, uint64_t packets_bytes, uint32_t n,
                         const uint32_t *t0, const uint32_t *t1, const uint32_t *t2,
                         const uint32_t *t3, hipStream_t s) {
    hipError_t e;
    uint64_t tab_bytes = 16ull * n;
    if ((e = q.d_packets.reserve(packets_bytes + 16)) != hipSuccess) return e;
    if ((e = q.d_offsets.reserve(tab_bytes)) != hipSuccess) return e;
    if ((e = q.h_stage.reserve(packets_bytes + tab_bytes + 64 * (uint64_t)n)) != hipSuccess)
        return e;
    uint8_t *h = q.h_stage.as<uint8_t>();
    uint64_t cur = 0;
    for (uint32_t k = 0; k < num_elems; k++) {
        uint64_t b = (uint64_t)elems[k].num * message_size;
        if (cur + b > packets_bytes) b = packets_bytes - cur;   // never overrun
        if (b) memcpy(h + cur, elems[k].elems, b);
        cur += b;
    }
    if (cur < packets_bytes) memset(h + cur, 0, packets_bytes - cur);
    uint32_t *tab = reinterpret_cast<uint32_t *>(h + packets_bytes);
    memcpy(tab, t0, 4ull * n);
    memcpy(tab + n, t1, 4ull * n);
    memcpy(tab + 2ull * n, t2, 4ull * n);
    memcpy(tab + 3ull * n, t3, 4ull * n);
    if ((e = hipMemcpyAsync(q.d_packets.ptr, h, packets_bytes, hipMemcpyHostToDevice, s)) !=
        hipSuccess)
        return e;
    return hipMemcpyAsync(q.d_offsets.ptr, tab, tab_bytes, hipMemcpyHostToDevice, s);
}

}  // namespace spl

using namespace spl;

extern "C" {

void ed25519_set_verbose(bool val) { g_verbose = val; }

void ed25519_verify_many(const gpu_Elems *elems, uint32_t num_elems, uint32_t message_size,
                         uint32_t total_packets, uint32_t total_signatures,
                         const uint32_t *message_lens, const uint32_t *public_key_offsets,
                         const uint32_t *signature_offsets,
                         const uint32_t *message_start_offsets, uint8_t *out,
                         uint8_t use_non_default_stream) {
    SPL_LOG("ed25519_verify_many: elems %u sigs %u packets %u message_size %u\n", num_elems,
            total_signatures, total_packets, message_size);
    if (total_packets == 0) return;
    uint32_t n = total_signatures;
    Lease l = acquire_queue();
    if (!l.q) {
        fprintf(stderr, "ed25519_verify_many: no GPU available; all %u signatures rejected\n",
                n);
        if (n) memset(out, 0, n);
        return;
    }
    QueueCtx &q = *l.q;
    hipStream_t s = use_non_default_stream ? q.stream : nullptr;
    uint64_t pbytes = (uint64_t)total_packets * message_size;
    const uint32_t *tab = nullptr;
    uint8_t *hout = nullptr;
    SPL_HIP_TRY(stage_packets(q, elems, num_elems, message_size, pbytes, n, message_lens,
                              public_key_offsets, signature_offsets, message_start_offsets, s));
    SPL_HIP_TRY(q.d_out.reserve(n));
    tab = q.d_offsets.as<uint32_t>();
    SPL_HIP_TRY(launch_verify(q.d_packets.as<uint8_t>(), pbytes, tab, tab + n, tab + 2ull * n,
                              tab + 3ull * n, n, q.d_out.as<uint8_t>(), l.dev->d_tables, s));
    hout = q.h_stage.as<uint8_t>();
    SPL_HIP_TRY(hipMemcpyAsync(hout, q.d_out.ptr, n, hipMemcpyDeviceToHost, s));
    SPL_HIP_TRY(hipStreamSynchronize(s));
    memcpy(out, hout, n);
    release_queue(l);
    return;
fail:
    if (n) memset(out, 0, n);
    release_queue(l);
}

int ed25519_verify_many_device(const uint8_t *d_packets, uint64_t packets_size,
                               const uint32_t *d_message_lens, const uint32_t *d_public_key_offsets,
                               const uint32_t *d_signature_offsets,
                               const uint32_t *d_message_start_offsets, uint32_t total_signatures,
                               uint8_t *d_out, void *stream) {
    if (total_signatures == 0) return 0;
    if (pool_init() <= 0) return 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    const ed25519_tables *tabs = device_tables(dev);
    if (!tabs) return 1;
    hipError_t e = launch_verify(d_packets, packets_size, d_message_lens, d_public_key_offsets,
                                 d_signature_offsets, d_message_start_offsets, total_signatures,
                                 d_out, tabs, (hipStream_t)stream);
    if (e != hipSuccess) {
        fprintf(stderr, "ed25519_verify_many_device: launch failed: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

}  // extern "C"
