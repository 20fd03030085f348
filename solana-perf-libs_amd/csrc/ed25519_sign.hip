// Batched Ed25519 signing behind the reference's ed25519_sign_many ABI.
//
// Reference: src/cuda-ecc-ed25519/sign.cu:50-145.  One lane per signature;
// the fixed-base multiplication scans each 8-entry table row with
// constant-time selects, because the nonce r is secret.
#include "runtime.h"
#include "../../include/cuda_crypt.h"

namespace spl {

hipError_t stage_packets(QueueCtx &q, const gpu_Elems *elems, uint32_t num_elems,
                         uint32_t message_size, uint64_t packets_bytes, uint32_t n,
                         const uint32_t *t0, const uint32_t *t1, const uint32_t *t2,
                         const uint32_t *t3, hipStream_t s);

__global__ void __launch_bounds__(64)
sign_kernel(const uint8_t *__restrict__ packets, uint64_t packets_size,
            const uint32_t *__restrict__ lens, const uint32_t *__restrict__ pk_off,
            const uint32_t *__restrict__ sk_off, const uint32_t *__restrict__ msg_off, uint32_t n,
            uint8_t *__restrict__ sigs, const ed25519_tables *__restrict__ T) {
    raise_wave_priority();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t *o = reinterpret_cast<uint32_t *>(sigs + 64ull * i);
    uint64_t ko = sk_off[i], po = pk_off[i], mo = msg_off[i], ml = lens[i];
    if (ko + 64 > packets_size || po + 32 > packets_size || mo + ml > packets_size) {
#pragma unroll
        for (int k = 0; k < 16; k++) o[k] = 0;   // offsets outside the buffer
        return;
    }
    uint32_t sk[16], pk[8], sig[16];
    load_words(sk, packets + ko, 16);
    load_words(pk, packets + po, 8);
    MsgReader msg{packets + mo, (uint32_t)ml};
    sign_one(sig, pk, sk, msg, *T);
#pragma unroll
    for (int k = 0; k < 16; k++) o[k] = sig[k];
}

hipError_t launch_sign(const uint8_t *d_packets, uint64_t packets_size, const uint32_t *d_lens,
                       const uint32_t *d_pk, const uint32_t *d_sk, const uint32_t *d_msg,
                       uint32_t n, uint8_t *d_sigs, const ed25519_tables *d_tables,
                       hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint32_t block = 64;
    hipLaunchKernelGGL(sign_kernel, dim3((n + block - 1) / block), dim3(block), 0, s, d_packets,
                       packets_size, d_lens, d_pk, d_sk, d_msg, n, d_sigs, d_tables);
    return hipGetLastError();
}

}  // namespace spl

using namespace spl;

extern "C" {

void ed25519_sign_many(const gpu_Elems *elems, uint32_t num_elems, uint32_t message_size,
                       uint32_t total_packets, uint32_t total_signatures,
                       const uint32_t *message_lens, const uint32_t *public_key_offsets,
                       const uint32_t *private_key_offsets, const uint32_t *message_start_offsets,
                       uint8_t *signatures_out, uint8_t use_non_default_stream) {
    SPL_LOG("ed25519_sign_many: elems %u sigs %u packets %u message_size %u\n", num_elems,
            total_signatures, total_packets, message_size);
    if (total_packets == 0) return;
    uint32_t n = total_signatures;
    Lease l = acquire_queue();
    if (!l.q) {
        fprintf(stderr, "ed25519_sign_many: no GPU available; no signatures produced\n");
        if (n) memset(signatures_out, 0, 64ull * n);
        return;
    }
    QueueCtx &q = *l.q;
    hipStream_t s = use_non_default_stream ? q.stream : nullptr;
    uint64_t pbytes = (uint64_t)total_packets * message_size;
    const uint32_t *tab = nullptr;
    uint8_t *hout = nullptr;
    // table order on the device: lens, pk, sk, msg
    SPL_HIP_TRY(stage_packets(q, elems, num_elems, message_size, pbytes, n, message_lens,
                              public_key_offsets, private_key_offsets, message_start_offsets, s));
    SPL_HIP_TRY(q.d_out.reserve(64ull * n));
    tab = q.d_offsets.as<uint32_t>();
    SPL_HIP_TRY(launch_sign(q.d_packets.as<uint8_t>(), pbytes, tab, tab + n, tab + 2ull * n,
                            tab + 3ull * n, n, q.d_out.as<uint8_t>(), l.dev->d_tables, s));
    hout = q.h_stage.as<uint8_t>();
    SPL_HIP_TRY(launch_copy16(hout, q.d_out.ptr, 64ull * n, s));
    SPL_HIP_TRY(hipStreamSynchronize(s));
    memcpy(signatures_out, hout, 64ull * n);
    release_queue(l);
    return;
fail:
    if (n) memset(signatures_out, 0, 64ull * n);
    release_queue(l);
}

int ed25519_sign_many_device(const uint8_t *d_packets, uint64_t packets_size,
                             const uint32_t *d_message_lens, const uint32_t *d_public_key_offsets,
                             const uint32_t *d_private_key_offsets,
                             const uint32_t *d_message_start_offsets, uint32_t total_signatures,
                             uint8_t *d_signatures_out, void *stream) {
    if (total_signatures == 0) return 0;
    if (pool_init() <= 0) return 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    const ed25519_tables *tabs = device_tables(dev);
    if (!tabs) return 1;
    hipError_t e = launch_sign(d_packets, packets_size, d_message_lens, d_public_key_offsets,
                               d_private_key_offsets, d_message_start_offsets, total_signatures,
                               d_signatures_out, tabs, (hipStream_t)stream);
    if (e != hipSuccess) {
        fprintf(stderr, "ed25519_sign_many_device: launch failed: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

}  // extern "C"
