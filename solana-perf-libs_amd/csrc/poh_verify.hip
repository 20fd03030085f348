// Proof-of-History chain verification: hash_i <- SHA-256(hash_i), repeated
// num_hashes[i] times, one chain per lane.
//
// Reference: src/cuda-poh-verify/poh_verify.cu:12-27 (kernel), :80-144 (ABI).
// The 32-byte state stays in eight VGPRs as big-endian words for the whole
// chain: SHA-256 output words are exactly the next message words, so no byte
// shuffling happens inside the loop, and the padding half of the block is a
// compile-time constant.
#include "runtime.h"
#include "../../include/cuda_crypt.h"

namespace spl {

__global__ void __launch_bounds__(256) poh_chain_kernel(uint32_t *__restrict__ hashes,
                                                        const uint64_t *__restrict__ counts,
                                                        uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t *h = hashes + 8 * i;
    uint32_t st[8];
#pragma unroll
    for (int k = 0; k < 8; k++) st[k] = bswap32(h[k]);
    uint64_t reps = counts[i];
    for (uint64_t r = 0; r < reps; r++) sha256_32byte(st);
#pragma unroll
    for (int k = 0; k < 8; k++) h[k] = bswap32(st[k]);
}

hipError_t launch_poh(uint32_t *d_hashes, const uint64_t *d_counts, uint64_t n,
                      hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t block = 64;
    uint64_t grid = (n + block - 1) / block;
    hipLaunchKernelGGL(poh_chain_kernel, dim3((uint32_t)grid), dim3(block), 0, stream,
                       d_hashes, d_counts, n);
    return hipGetLastError();
}

}  // namespace spl

using namespace spl;

extern "C" {

void poh_verify_many_set_verbose(bool val) { g_verbose = val; }

int poh_verify_many(uint8_t *hashes, const uint64_t *num_hashes_arr, size_t num_elems,
                    uint8_t use_non_default_stream) {
    SPL_LOG("poh_verify_many: num_elems %zu\n", num_elems);
    if (num_elems == 0) return 0;
    Lease l = acquire_queue();
    if (!l.q) {
        SPL_LOG("poh_verify_many: no GPU\n");
        return 1;
    }
    int rc = 1;
    QueueCtx &q = *l.q;
    hipStream_t s = use_non_default_stream ? q.stream : nullptr;
    size_t hb = num_elems * 32, cb = num_elems * 8;
    SPL_HIP_TRY(q.d_hashes.reserve(hb));
    SPL_HIP_TRY(q.d_counts.reserve(cb + 16));
    SPL_HIP_TRY(q.h_stage.reserve(hb + cb + 16));
    memcpy(q.h_stage.as<uint8_t>(), hashes, hb);
    memcpy(q.h_stage.as<uint8_t>() + hb, num_hashes_arr, cb);
    // copy kernels, not DMA copies (see launch_copy16)
    SPL_HIP_TRY(launch_copy16(q.d_hashes.ptr, q.h_stage.ptr, hb, s));
    SPL_HIP_TRY(launch_copy16(q.d_counts.ptr, q.h_stage.as<uint8_t>() + hb, cb, s));
    SPL_HIP_TRY(launch_poh(q.d_hashes.as<uint32_t>(), q.d_counts.as<uint64_t>(), num_elems, s));
    SPL_HIP_TRY(launch_copy16(q.h_stage.ptr, q.d_hashes.ptr, hb, s));
    SPL_HIP_TRY(hipStreamSynchronize(s));
    memcpy(hashes, q.h_stage.ptr, hb);
    rc = 0;
fail:
    release_queue(l);
    return rc;
}

int poh_verify_many_device(uint8_t *d_hashes, const uint64_t *d_num_hashes, size_t num_elems,
                           void *stream) {
    if (num_elems == 0) return 0;
    if (((uintptr_t)d_hashes & 3) != 0 || ((uintptr_t)d_num_hashes & 7) != 0) {
        fprintf(stderr, "poh_verify_many_device: misaligned buffers\n");
        return 2;
    }
    if (pool_init() <= 0) return 1;
    hipError_t e = launch_poh((uint32_t *)d_hashes, d_num_hashes, num_elems, (hipStream_t)stream);
    if (e != hipSuccess) {
        fprintf(stderr, "poh_verify_many_device: launch failed: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

}  // extern "C"
