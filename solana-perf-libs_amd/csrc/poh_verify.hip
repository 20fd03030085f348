// Proof-of-History chain verification: hash_i <- SHA-256(hash_i), repeated
// num_hashes[i] times.
//
// Reference: src/cuda-poh-verify/poh_verify.cu:12-27 (kernel), :80-144 (ABI).
// The 32-byte state stays in eight VGPRs as big-endian words for the whole
// chain: SHA-256 output words are exactly the next message words, so no byte
// shuffling happens inside the loop, and the padding half of the block is a
// compile-time constant.
//
// Two kernels, by how many chains there are:
//   poh_chain_kernel  one chain per lane.  Once the chains fill every SIMD
//                     (64 per wave, one wave per SIMD) this is the fastest
//                     form: a wave issues its 1,322 instructions per hash
//                     back to back.
//   poh_split_kernel  two waves per 64 chains, for fewer chains than that.
//                     A wave64 VALU instruction occupies its SIMD for 4
//                     cycles, so one wave per chain group leaves a chain's
//                     time at 4 cycles x the instructions of one hash, and
//                     the other SIMDs idle.  Here the second wave of the
//                     workgroup (on another SIMD) computes the message
//                     schedule W16..W63 (plus K) of the same hash and hands
//                     it over through LDS, so the rounds wave issues 881
//                     instructions per hash instead of 1,322.  Four barriers
//                     per hash keep the two waves in step; every one of them
//                     sits in control flow that both waves run identically
//                     (the hash count is the workgroup's maximum, and lanes
//                     past their own count keep their state).
#include "runtime.h"
#include "../../include/cuda_crypt.h"

namespace spl {

// Hashes of a chain with `count` hashes in all that fall in the window
// [skip, skip + limit) of one launch: a chain is hashed in slices by
// successive launches with skip = 0, limit, 2 limit, ... (the state stays in
// place between them); skip = 0, limit = all hashes is the whole chain.
__device__ __forceinline__ uint64_t slice_count(uint64_t count, uint64_t skip, uint64_t limit) {
    const uint64_t left = count > skip ? count - skip : 0;
    return left < limit ? left : limit;
}

__global__ void __launch_bounds__(256) poh_chain_kernel(uint32_t *__restrict__ hashes,
                                                        const uint64_t *__restrict__ counts,
                                                        uint64_t n, uint64_t skip,
                                                        uint64_t limit) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t *h = hashes + 8 * i;
    uint32_t st[8];
#pragma unroll
    for (int k = 0; k < 8; k++) st[k] = bswap32(h[k]);
    uint64_t reps = slice_count(counts[i], skip, limit);
    for (uint64_t r = 0; r < reps; r++) sha256_32byte(st);
#pragma unroll
    for (int k = 0; k < 8; k++) h[k] = bswap32(st[k]);
}

struct Sha256Vars {
    uint32_t a, b, c, d, e, f, g, h;
};

// One SHA-256 round; kw = K[t] + W[t].
//
// h + kw is known rounds ahead (h is e of three rounds back), so it is summed
// first and kept as one value: the empty asm stops the compiler from
// reassociating it into (h + S1) + ch + kw, which put one more dependent add
// on the e chain (now e -> rotates -> xor3 -> add3 -> add).  Measured on the
// split kernel: 0.4-0.7 % faster at 512-32,768 chains; the rounds wave is
// bound by a lone wave's issue rate, not by this chain.
__device__ __forceinline__ void sha256_round(Sha256Vars &s, uint32_t kw) {
    const uint32_t S1 = xor3_32(rotr32(s.e, 6), rotr32(s.e, 11), rotr32(s.e, 25));
    const uint32_t ch = ch32(s.e, s.f, s.g);
    uint32_t u = s.h + kw;
    asm("" : "+v"(u));
    const uint32_t t1 = u + S1 + ch;
    const uint32_t S0 = xor3_32(rotr32(s.a, 2), rotr32(s.a, 13), rotr32(s.a, 22));
    const uint32_t t2 = S0 + maj32(s.a, s.b, s.c);
    s.h = s.g; s.g = s.f; s.f = s.e; s.e = s.d + t1;
    s.d = s.c; s.c = s.b; s.b = s.a; s.a = t1 + t2;
}

// W[t] for t in [t0, t0 + 16) into the ring w (w[t & 15]), with K[t] + W[t]
// stored to kw[t - 16][lane].
template <int T0>
__device__ __forceinline__ void sha256_schedule16(uint32_t (&w)[16], uint32_t (*kw)[64],
                                                  uint32_t lane) {
    const uint32_t K[64] = SPL_SHA256_K;
#pragma unroll
    for (int t = T0; t < T0 + 16; t++) {
        const uint32_t w15 = w[(t - 15) & 15], w2 = w[(t - 2) & 15];
        const uint32_t s0 = xor3_32(rotr32(w15, 7), rotr32(w15, 18), w15 >> 3);
        const uint32_t s1 = xor3_32(rotr32(w2, 17), rotr32(w2, 19), w2 >> 10);
        const uint32_t wt = w[t & 15] + s0 + w[(t - 7) & 15] + s1;
        w[t & 15] = wt;
        kw[t - 16][lane] = wt + K[t];
    }
}

constexpr uint32_t kPohSplitThreads = 128;   // wave 0: rounds, wave 1: schedule

// Rounds in the rounds wave's first phase, the one that overlaps the schedule
// wave reading the digest from LDS and computing W16..31.  Timed on 512 to
// 32,768 chains: 14 is 0.7-1 % faster than 12 or 13 (the schedule wave's
// first phase was the longer one) and than 15 or 16 (fewer rounds are left
// to cover the LDS reads after each barrier).
constexpr int kPohPhase1Rounds = 14;
static_assert(kPohPhase1Rounds >= 9 && kPohPhase1Rounds <= 16,
              "rounds 0..8 (state words and the padding bit) always run in phase 1");

__global__ void __launch_bounds__(kPohSplitThreads)
poh_split_kernel(uint32_t *__restrict__ hashes, const uint64_t *__restrict__ counts, uint64_t n,
                 uint64_t skip, uint64_t limit) {
    __shared__ uint32_t kw[48][64];              // K[t] + W[t], t = 16..63
    __shared__ uint32_t msg[8][64];              // message words of the coming hash
    __shared__ unsigned long long reps_max;
    const uint32_t K[64] = SPL_SHA256_K;
    const uint32_t H0[8] = SPL_SHA256_H0;
    const uint32_t lane = threadIdx.x & 63;
    const bool rounds = threadIdx.x < 64;        // wave-uniform role
    const uint64_t i = (uint64_t)blockIdx.x * 64 + lane;
    const bool valid = i < n;
    const uint64_t cnt = valid ? slice_count(counts[i], skip, limit) : 0;
    if (threadIdx.x == 0) reps_max = 0;
    __syncthreads();
    if (rounds) atomicMax(&reps_max, (unsigned long long)cnt);
    uint32_t st[8];
#pragma unroll
    for (int k = 0; k < 8; k++) st[k] = rounds && valid ? bswap32(hashes[8 * i + k]) : 0u;
    if (rounds) {
#pragma unroll
        for (int k = 0; k < 8; k++) msg[k][lane] = st[k];
    }
    __syncthreads();
    const uint64_t reps = reps_max;
    Sha256Vars v{};
    uint32_t w[16];
    // rounds wave: K + W of t = 16..31, 32..47 and 48..63, one array per
    // phase so the four rounds carried into the next phase need no copies
    uint32_t qa[16], qb[16], qc[16];
    // The rounds wave's phases end at rounds P1, P1 + 16, P1 + 32 and 64;
    // right after each barrier it loads K + W for the rest of that phase and
    // the next phase's first 16 - P1 rounds, so every LDS read has rounds of
    // work in front of its first use.  The schedule wave writes W16..31,
    // W32..47 and W48..63 in the first three phases and waits out the fourth.
    constexpr int P1 = kPohPhase1Rounds;
    for (uint64_t r = 0; r < reps; r++) {
        // phase 1: rounds 0..P1-1 from the state | W16..31
        if (rounds) {
            v = Sha256Vars{H0[0], H0[1], H0[2], H0[3], H0[4], H0[5], H0[6], H0[7]};
#pragma unroll
            for (int t = 0; t < 8; t++) sha256_round(v, st[t] + K[t]);
            sha256_round(v, 0x80000000u + K[8]);
#pragma unroll
            for (int t = 9; t < P1; t++) sha256_round(v, t == 15 ? 256u + K[15] : K[t]);
        } else {
#pragma unroll
            for (int k = 0; k < 8; k++) w[k] = msg[k][lane];
            w[8] = 0x80000000u;
#pragma unroll
            for (int k = 9; k < 15; k++) w[k] = 0;
            w[15] = 256;
            sha256_schedule16<16>(w, kw, lane);
        }
        __syncthreads();
        // phase 2: rounds P1..P1+15 | W32..47
        if (rounds) {
#pragma unroll
            for (int t = 16; t < 32; t++) qa[t - 16] = kw[t - 16][lane];
            __builtin_amdgcn_sched_barrier(0);   // keep the loads ahead of the rounds
#pragma unroll
            for (int t = P1; t < 16; t++) sha256_round(v, t == 15 ? 256u + K[15] : K[t]);
#pragma unroll
            for (int t = 16; t < P1 + 16; t++) sha256_round(v, qa[t - 16]);
        } else {
            sha256_schedule16<32>(w, kw, lane);
        }
        __syncthreads();
        // phase 3: rounds P1+16..P1+31 | W48..63
        if (rounds) {
#pragma unroll
            for (int t = 32; t < 48; t++) qb[t - 32] = kw[t - 16][lane];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = P1 + 16; t < 32; t++) sha256_round(v, qa[t - 16]);
#pragma unroll
            for (int t = 32; t < P1 + 32; t++) sha256_round(v, qb[t - 32]);
        } else {
            sha256_schedule16<48>(w, kw, lane);
        }
        __syncthreads();
        // phase 4: rounds P1+32..63, feed-forward, next message
        if (rounds) {
#pragma unroll
            for (int t = 48; t < 64; t++) qc[t - 48] = kw[t - 16][lane];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = P1 + 32; t < 48; t++) sha256_round(v, qb[t - 32]);
#pragma unroll
            for (int t = 48; t < 64; t++) sha256_round(v, qc[t - 48]);
            if (r < cnt) {
                st[0] = H0[0] + v.a; st[1] = H0[1] + v.b; st[2] = H0[2] + v.c;
                st[3] = H0[3] + v.d; st[4] = H0[4] + v.e; st[5] = H0[5] + v.f;
                st[6] = H0[6] + v.g; st[7] = H0[7] + v.h;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) msg[k][lane] = st[k];
        }
        __syncthreads();
    }
    if (rounds && valid) {
#pragma unroll
        for (int k = 0; k < 8; k++) hashes[8 * i + k] = bswap32(st[k]);
    }
}

// Up to 32 chains per SIMD the split kernel's two waves per 64 chains still
// give every wave a SIMD of its own.
static bool use_split_kernel(uint64_t n) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    return n <= 32ull * 4ull * (uint64_t)device_cus(dev);
}

hipError_t launch_poh(uint32_t *d_hashes, const uint64_t *d_counts, uint64_t n,
                      hipStream_t stream, uint64_t skip, uint64_t limit) {
    if (n == 0 || limit == 0) return hipSuccess;
    const uint64_t groups = (n + 63) / 64;
    if (use_split_kernel(n))
        hipLaunchKernelGGL(poh_split_kernel, dim3((uint32_t)groups), dim3(kPohSplitThreads), 0,
                           stream, d_hashes, d_counts, n, skip, limit);
    else
        hipLaunchKernelGGL(poh_chain_kernel, dim3((uint32_t)groups), dim3(64), 0, stream,
                           d_hashes, d_counts, n, skip, limit);
    return hipGetLastError();
}

}  // namespace spl

using namespace spl;

// C++ linkage on purpose: the reference exports it as _Z8poh_initv.
bool poh_init() { return pool_init() > 0; }

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
