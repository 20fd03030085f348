// Proof-of-History chain verification: hash_i <- SHA-256(hash_i), repeated
// num_hashes[i] times.
//
// Reference: src/cuda-poh-verify/poh_verify.cu:12-27 (kernel), :80-144 (ABI).
// The 32-byte state stays in eight VGPRs as big-endian words for the whole
// chain: SHA-256 output words are exactly the next message words, so no byte
// shuffling happens inside the loop, and the padding half of the block is a
// compile-time constant.
//
// Three kernels, by how many chains there are:
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
//   poh_pair_kernel   the split kernel's two waves per group with two lanes
//                     per chain in each (32 chains per group): the rounds
//                     wave issues 11 instructions per round instead of 14,
//                     about 730 per hash.  Fastest while a quarter of the
//                     SIMDs or fewer carry a wave.
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

// ---------------------------------------------------------------------------
// poh_pair_kernel: two lanes per chain in the rounds wave.
//
// A SHA-256 round is two half-rounds of the same shape: the e side computes
// Sigma1(e) and Ch(e, f, g), the a side Sigma0(a) and Maj(a, b, c).  Here one
// lane holds e, f, g, h and its partner lane a, b, c, d of the same chain, and
// one instruction stream computes both halves: the rotate counts of Sigma are
// per-lane registers, and Ch and Maj are both Q ^ (x & P) with P = y ^ z,
// Q = z (Ch) or y & z (Maj).  What one side needs of the other crosses over
// by DPP within a 16-lane row.  Per round:
//
//   S = Sigma(r1)                         3 x V_ALIGNBIT + V_BITOP3
//   F = Q ^ (r1 & P)                      V_XOR, V_BITOP3 (Q), V_BITOP3
//   Y = e side: h + d + K + W, a side: -d V_ADD_DPP (d from the a lane), V_XAD
//   u = Y + S + F                         V_ADD3: e' on the e side, T2 - d on the a side
//   a side: a' = u + e' (partner's u)     V_ADD_DPP
//
// 11 instructions per round against 14 when one lane holds all eight words.
// Lanes 0-3 and 8-11 of each row are e lanes, 4-7 and 12-15 their a-lane
// partners (lane ^ 4), so each crossing is one row shift by 4 with a bank
// mask: the lanes that must keep their value are simply not written.
//
// The second wave computes the message schedule of the same 32 chains, also
// two lanes per chain: W[t] = sigma0(W[t-15]) + W[t-7] + W[t-16] on the e
// lane plus sigma1(W[t-2]) on the a lane, each lane one sigma with per-lane
// counts.  The a lane keeps its 16-word ring rotated by 13 slots, so the same
// ring slot holds W[t-15] on the e lane and W[t-2] on the a lane.  It hands
// K[t] + W[t] to the rounds wave through LDS as in the split kernel.
constexpr uint32_t kPohPairThreads = 128;    // wave 0: rounds, wave 1: schedule
constexpr uint32_t kPohPairChains = 32;      // chains per workgroup
constexpr int kDppRowShl4 = 0x104;           // lane i reads lane i + 4 of its row
constexpr int kDppRowShr4 = 0x114;           // lane i reads lane i - 4 of its row
constexpr int kBanksE = 0x5, kBanksA = 0xA;  // lanes 0-3, 8-11 / 4-7, 12-15 of a row

// e lanes take `src` from their a-lane partner; a lanes return `old`.
__device__ __forceinline__ uint32_t from_a_lane(uint32_t old, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, kDppRowShl4, 0xf, kBanksE,
                                                 false);
}

// a lanes take `src` from their e-lane partner; e lanes return `old`.
__device__ __forceinline__ uint32_t from_e_lane(uint32_t old, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, kDppRowShr4, 0xf, kBanksA,
                                                 false);
}

struct PairLane {
    uint32_t lane;    // lane in the wave
    uint32_t chain;   // chain within the workgroup, 0..31
    bool a;           // holds a, b, c, d (else e, f, g, h)
};

__device__ __forceinline__ PairLane pair_lane(uint32_t lane) {
    return {lane, (lane >> 4) * 8 + ((lane >> 3) & 1) * 4 + (lane & 3), ((lane >> 2) & 1) != 0};
}

struct PairRound {
    uint32_t s1, s2, s3;   // Sigma rotate counts: e 6, 11, 25; a 2, 13, 22
    uint32_t zmask;        // e: all ones (Q = z), a: 0 (Q = y & z)
    uint32_t nmask;        // e: 0 (Y adds h), a: all ones (Y = ~d + 1 = -d)
};

// One round on r1..r4 = (e, f, g, h) or (a, b, c, d).  kw is K[t] + W[t] on
// an e lane and 1 on an a lane.
__device__ __forceinline__ void pair_round(uint32_t &r1, uint32_t &r2, uint32_t &r3, uint32_t &r4,
                                           uint32_t kw, const PairRound &c) {
    const uint32_t S = xor3_32(__builtin_amdgcn_alignbit(r1, r1, c.s1),
                               __builtin_amdgcn_alignbit(r1, r1, c.s2),
                               __builtin_amdgcn_alignbit(r1, r1, c.s3));
    const uint32_t P = r2 ^ r3;
    const uint32_t Q = __builtin_amdgcn_bitop3_b32(r3, r2, c.zmask, 0xe0);   // z & (y | zmask)
    const uint32_t F = __builtin_amdgcn_bitop3_b32(r1, P, Q, 0x6a);          // Q ^ (x & P)
    // e: h + (K + W + d), a: ~d + 1; summed apart from the S/F chain.  The
    // empty asm keeps kw + d a two-operand add, which takes the DPP operand in
    // one V_ADD_U32_DPP (a three-operand add cannot).
    uint32_t C = kw + from_a_lane(0, r4);
    asm("" : "+v"(C));
    uint32_t Y = (r4 ^ c.nmask) + C;   // one V_XAD_U32
    asm("" : "+v"(Y));
    const uint32_t u = Y + S + F;
    const uint32_t n = u + from_e_lane(0, u);   // a: (T2 - d) + e' = T1 + T2
    r4 = r3; r3 = r2; r2 = r1; r1 = n;
}

// Rounds 0..3 start from the constant initial state, so in each of them every
// term but the message word folds into one per-lane constant, fixed at
// compile time: round t adds y[t] = h + K[t] + d (e lane) or -d (a lane), and
// round 0 also its Sigma and Ch / Maj; round 1 takes its P and Q from p1, q1.
struct PairStart {
    uint32_t y[4];
    uint32_t p1, q1;
};

constexpr uint32_t crotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

constexpr PairStart pair_start(bool a) {
    constexpr uint32_t H[8] = SPL_SHA256_H0;
    constexpr uint32_t Kc[4] = {0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u};
    const uint32_t x0 = a ? H[0] : H[4], x1 = a ? H[1] : H[5], x2 = a ? H[2] : H[6];
    PairStart st{};
    for (int t = 0; t < 4; t++) st.y[t] = a ? 0u - H[3 - t] : H[7 - t] + Kc[t] + H[3 - t];
    const uint32_t S = a ? crotr(x0, 2) ^ crotr(x0, 13) ^ crotr(x0, 22)
                         : crotr(x0, 6) ^ crotr(x0, 11) ^ crotr(x0, 25);
    const uint32_t F = a ? (x0 & x1) ^ (x0 & x2) ^ (x1 & x2) : (x0 & x1) ^ (~x0 & x2);
    st.y[0] += S + F;
    st.p1 = x0 ^ x1;
    st.q1 = a ? x0 & x1 : x1;   // Q of round 1: r3 = x1, r2 = x0
    return st;
}

// The rest of a round once P, Q and Y are known.
__device__ __forceinline__ uint32_t pair_round_tail(uint32_t r1, uint32_t P, uint32_t Q, uint32_t Y,
                                                    const PairRound &c) {
    const uint32_t S = xor3_32(__builtin_amdgcn_alignbit(r1, r1, c.s1),
                               __builtin_amdgcn_alignbit(r1, r1, c.s2),
                               __builtin_amdgcn_alignbit(r1, r1, c.s3));
    const uint32_t F = __builtin_amdgcn_bitop3_b32(r1, P, Q, 0x6a);
    const uint32_t u = Y + S + F;
    return u + from_e_lane(0, u);
}

// The e lane's message word w (taken from `src` of the a-lane partner when
// partner is set, else its own) plus y; an a lane returns y.  One
// V_ADD_U32_DPP.
template <bool Partner>
__device__ __forceinline__ uint32_t add_e_word(uint32_t y, uint32_t src) {
    uint32_t v = y + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)src,
                                                           Partner ? kDppRowShl4 : 0xe4, 0xf,
                                                           kBanksE, false);
    asm("" : "+v"(v));
    return v;
}

struct PairSchedule {
    uint32_t q1, q2, q3;   // sigma counts: e (sigma0) 7, 18, >>3; a (sigma1) 17, 19, >>10
};

// W[t] for t in [T0, T1).  Ring w: on an e lane w[s] = W[j] with
// j = s (mod 16), on an a lane j = s + 13 (mod 16).  kwv[t - 16] receives
// K[t] + W[t] (meaningful on the e lanes).
template <int T0, int T1>
__device__ __forceinline__ void pair_schedule(uint32_t (&w)[16], uint32_t (&kwv)[48],
                                              const PairSchedule &c) {
    const uint32_t K[64] = SPL_SHA256_K;
#pragma unroll
    for (int t = T0; t < T1; t++) {
        const uint32_t x = w[(t - 15) & 15];   // e: W[t-15], a: W[t-2]
        const uint32_t sg = xor3_32(__builtin_amdgcn_alignbit(x, x, c.q1),
                                    __builtin_amdgcn_alignbit(x, x, c.q2), x >> c.q3);
        uint32_t u = sg + w[(t - 7) & 15] + w[(t - 16) & 15];
        asm("" : "+v"(u));                            // keeps the next add one V_ADD_U32_DPP
        const uint32_t wt = u + from_a_lane(0, sg);   // e: W[t]
        // e slot t held W[t-16] (used above); a slot t holds W[t-3], no longer needed
        w[t & 15] = wt;
        // a slot t+3 takes W[t]; e slot t+3 keeps W[t-13]
        w[(t + 3) & 15] = from_e_lane(w[(t + 3) & 15], wt);
        kwv[t - 16] = wt + K[t];
    }
}

// The schedule wave writes K + W in three blocks, W[16..B1), W[B1..B2) and
// W[B2..64), one per phase; the rounds wave's phases end after rounds P1, P2
// and P3 (round 64 ends the fourth) and it loads each block right after the
// barrier that publishes it.  Round t of phase k + 1 may only use a block the
// schedule wrote in phases 1..k, hence P2 <= B1 and P3 <= B2; rounds 16 - P1,
// B1 - P2 and B2 - P3 run from registers while the loads are in flight.
// Timed on 4096 x 12,500 (20.66 ms here): blocks ending at 31/47, 30/46,
// 31/46 or 30/47 with phase ends two rounds before them are 1.3-4 % slower,
// and so are phase ends 15/30/46, 14/29/45, 14/30/46 and others with these
// blocks (0.6-2.4 %), and P/P+16/P+32 for P = 11-14 or 16 (0.5-6 %): a second
// round of cover for the loads does not pay for the rounds it moves.
constexpr int kPohPairBlockEnd[2] = {32, 48};
constexpr int kPohPairPhaseEnd[3] = {15, 31, 47};
static_assert(kPohPairPhaseEnd[0] >= 8 && kPohPairPhaseEnd[0] <= 16 &&
                  kPohPairPhaseEnd[1] >= 16 && kPohPairPhaseEnd[1] <= kPohPairBlockEnd[0] &&
                  kPohPairPhaseEnd[2] >= kPohPairBlockEnd[0] &&
                  kPohPairPhaseEnd[2] <= kPohPairBlockEnd[1] && kPohPairBlockEnd[0] > 16 &&
                  kPohPairBlockEnd[1] > kPohPairBlockEnd[0] && kPohPairBlockEnd[1] < 64,
              "phase 1 holds the message rounds; no round runs before its block is loaded");

__global__ void __launch_bounds__(kPohPairThreads)
poh_pair_kernel(uint32_t *__restrict__ hashes, const uint64_t *__restrict__ counts, uint64_t n,
                uint64_t skip, uint64_t limit) {
    __shared__ uint32_t kw[48][64];      // K[t] + W[t], t = 16..63, by rounds-wave lane (a lanes: 1)
    __shared__ uint32_t msg[32][32];     // W[0..15] of the coming hash, twice, by chain
    __shared__ unsigned long long reps_max;
    const uint32_t K[64] = SPL_SHA256_K;
    const uint32_t H0[8] = SPL_SHA256_H0;
    const PairLane pl = pair_lane(threadIdx.x & 63);
    const bool rounds = threadIdx.x < 64;   // wave-uniform role
    const uint64_t i = (uint64_t)blockIdx.x * kPohPairChains + pl.chain;
    const bool valid = i < n;
    const uint64_t cnt = valid ? slice_count(counts[i], skip, limit) : 0;
    const uint32_t half = pl.a ? 0 : 4;     // this lane's digest words: half .. half + 3
    if (threadIdx.x == 0) reps_max = 0;
    __syncthreads();
    if (rounds) atomicMax(&reps_max, (unsigned long long)cnt);
    uint32_t dg[4];
#pragma unroll
    for (int k = 0; k < 4; k++) dg[k] = rounds && valid ? bswap32(hashes[8 * i + half + k]) : 0u;
    if (rounds) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            msg[half + k][pl.chain] = dg[k];
            msg[half + k + 16][pl.chain] = dg[k];
        }
        if (pl.a) {
#pragma unroll
            for (int t = 0; t < 48; t++) kw[t][pl.lane] = 1u;
        }
    } else {
        // the constant second half of the block: the padding bit and the length
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t row = 8 + (k & 3) + 16 * (k >> 2) + 4 * (pl.lane >> 5);
            const uint32_t wv = row == 8 || row == 24 ? 0x80000000u
                                : row == 15 || row == 31 ? 256u : 0u;
            msg[row][pl.lane & 31] = wv;
        }
    }
    __syncthreads();
    const uint64_t reps = reps_max;
    const PairRound rc = pl.a ? PairRound{2, 13, 22, 0u, ~0u} : PairRound{6, 11, 25, ~0u, 0u};
    const PairSchedule sc = pl.a ? PairSchedule{17, 19, 10} : PairSchedule{7, 18, 3};
    uint32_t hl[4];   // this lane's initial words: H0[half .. half + 3]
#pragma unroll
    for (int k = 0; k < 4; k++) hl[k] = pl.a ? H0[k] : H0[4 + k];
    // K[t] + W[t] of rounds 0..15 on e lanes (1 on a lanes): rounds 8..15 are
    // constants, rounds 0..7 add the message words
    uint32_t ke[16];
#pragma unroll
    for (int t = 0; t < 16; t++)
        ke[t] = pl.a ? 1u : K[t] + (t == 8 ? 0x80000000u : t == 15 ? 256u : 0u);
    constexpr PairStart kStartA = pair_start(true), kStartE = pair_start(false);
    uint32_t ys[4];
#pragma unroll
    for (int t = 0; t < 4; t++) ys[t] = pl.a ? kStartA.y[t] : kStartE.y[t];
    const uint32_t p1c = pl.a ? kStartA.p1 : kStartE.p1, q1c = pl.a ? kStartA.q1 : kStartE.q1;
    uint32_t r1 = 0, r2 = 0, r3 = 0, r4 = 0;
    uint32_t w[16], kwv[48];
    uint32_t k0[8], q[48];
    // K[t] + W[t] of round t: message words (k0), constants (ke), then the
    // schedule wave's three blocks, each loaded right after the barrier that
    // publishes it.  t is a constant in every unrolled use.
    auto kwt = [&](int t) -> uint32_t { return t < 8 ? k0[t] : t < 16 ? ke[t] : q[t - 16]; };
    constexpr int P1 = kPohPairPhaseEnd[0], P2 = kPohPairPhaseEnd[1], P3 = kPohPairPhaseEnd[2];
    constexpr int B1 = kPohPairBlockEnd[0], B2 = kPohPairBlockEnd[1];
    const uint32_t srow = pl.a ? 13 : 0;   // schedule ring rotation of this lane
    for (uint64_t r = 0; r < reps; r++) {
        // phase 1: rounds 0..P1-1 | W[16..B1)
        if (rounds) {
#pragma unroll
            for (int t = 4; t < 8; t++) k0[t] = add_e_word<false>(ke[t], dg[t - 4]);
            // rounds 0..3 from the constant state (the e lane's words 0..3 are
            // its partner's digest words)
            const uint32_t n0 = add_e_word<true>(ys[0], dg[0]);
            const uint32_t s0 = n0 + from_e_lane(0, n0);
            const uint32_t s1 = pair_round_tail(s0, p1c, q1c, add_e_word<true>(ys[1], dg[1]), rc);
            const uint32_t s2 = pair_round_tail(
                s1, s0 ^ hl[0], __builtin_amdgcn_bitop3_b32(hl[0], s0, rc.zmask, 0xe0),
                add_e_word<true>(ys[2], dg[2]), rc);
            const uint32_t s3 = pair_round_tail(
                s2, s1 ^ s0, __builtin_amdgcn_bitop3_b32(s0, s1, rc.zmask, 0xe0),
                add_e_word<true>(ys[3], dg[3]), rc);
            r1 = s3; r2 = s2; r3 = s1; r4 = s0;
#pragma unroll
            for (int t = 4; t < P1; t++) pair_round(r1, r2, r3, r4, kwt(t), rc);
        } else {
#pragma unroll
            for (int s = 0; s < 16; s++) w[s] = msg[s + srow][pl.chain];
            pair_schedule<16, B1>(w, kwv, sc);
            if (!pl.a) {
#pragma unroll
                for (int t = 16; t < B1; t++) kw[t - 16][pl.lane] = kwv[t - 16];
            }
        }
        __syncthreads();
        // phase 2: rounds P1..P2-1 | W[B1..B2)
        if (rounds) {
#pragma unroll
            for (int t = 16; t < B1; t++) q[t - 16] = kw[t - 16][pl.lane];
            __builtin_amdgcn_sched_barrier(0);   // keep the loads ahead of the rounds
#pragma unroll
            for (int t = P1; t < P2; t++) pair_round(r1, r2, r3, r4, kwt(t), rc);
        } else {
            pair_schedule<B1, B2>(w, kwv, sc);
            if (!pl.a) {
#pragma unroll
                for (int t = B1; t < B2; t++) kw[t - 16][pl.lane] = kwv[t - 16];
            }
        }
        __syncthreads();
        // phase 3: rounds P2..P3-1 | W[B2..64)
        if (rounds) {
#pragma unroll
            for (int t = B1; t < B2; t++) q[t - 16] = kw[t - 16][pl.lane];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = P2; t < P3; t++) pair_round(r1, r2, r3, r4, kwt(t), rc);
        } else {
            pair_schedule<B2, 64>(w, kwv, sc);
            if (!pl.a) {
#pragma unroll
                for (int t = B2; t < 64; t++) kw[t - 16][pl.lane] = kwv[t - 16];
            }
        }
        __syncthreads();
        // phase 4: rounds P3..63, feed-forward, next message
        if (rounds) {
#pragma unroll
            for (int t = B2; t < 64; t++) q[t - 16] = kw[t - 16][pl.lane];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = P3; t < 64; t++) pair_round(r1, r2, r3, r4, kwt(t), rc);
            if (r < cnt) {
                dg[0] = hl[0] + r1; dg[1] = hl[1] + r2; dg[2] = hl[2] + r3; dg[3] = hl[3] + r4;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                msg[half + k][pl.chain] = dg[k];
                msg[half + k + 16][pl.chain] = dg[k];
            }
        }
        __syncthreads();
    }
    if (rounds && valid) {
#pragma unroll
        for (int k = 0; k < 4; k++) hashes[8 * i + half + k] = bswap32(dg[k]);
    }
}

// Which kernel runs n chains.  Up to 32 chains per SIMD the split kernel's
// two waves per 64 chains still give every wave a SIMD of its own; up to 8
// per SIMD the pair kernel's four waves per 64 chains do too, and it is the
// faster of the two (timed on 512 to 65,536 chains: 9-10 % faster up to
// 8,192 on 256 CUs, 17 % slower at 16,384).
enum PohKernel : int { kPohAuto = 0, kPohChain = 1, kPohSplit = 2, kPohPair = 3 };

static PohKernel pick_poh_kernel(uint64_t n) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return kPohChain;
    const uint64_t simds = 4ull * (uint64_t)device_cus(dev);
    return n <= 8ull * simds ? kPohPair : n <= 32ull * simds ? kPohSplit : kPohChain;
}

static hipError_t launch_poh_kernel(PohKernel kind, uint32_t *d_hashes, const uint64_t *d_counts,
                                    uint64_t n, hipStream_t stream, uint64_t skip,
                                    uint64_t limit) {
    if (n == 0 || limit == 0) return hipSuccess;
    if (kind == kPohAuto) kind = pick_poh_kernel(n);
    const uint64_t groups = (n + 63) / 64;
    switch (kind) {
    case kPohSplit:
        hipLaunchKernelGGL(poh_split_kernel, dim3((uint32_t)groups), dim3(kPohSplitThreads), 0,
                           stream, d_hashes, d_counts, n, skip, limit);
        break;
    case kPohPair:
        hipLaunchKernelGGL(poh_pair_kernel,
                           dim3((uint32_t)((n + kPohPairChains - 1) / kPohPairChains)),
                           dim3(kPohPairThreads), 0, stream, d_hashes, d_counts, n, skip, limit);
        break;
    default:
        hipLaunchKernelGGL(poh_chain_kernel, dim3((uint32_t)groups), dim3(64), 0, stream,
                           d_hashes, d_counts, n, skip, limit);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_poh(uint32_t *d_hashes, const uint64_t *d_counts, uint64_t n,
                      hipStream_t stream, uint64_t skip, uint64_t limit) {
    return launch_poh_kernel(kPohAuto, d_hashes, d_counts, n, stream, skip, limit);
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

int spl_poh_kernel_choice(size_t num_elems) {
    if (pool_init() <= 0) return 0;
    return (int)pick_poh_kernel(num_elems);
}

int spl_poh_device_kernel(uint8_t *d_hashes, const uint64_t *d_num_hashes, size_t num_elems,
                          int kernel, void *stream) {
    if (kernel < kPohAuto || kernel > kPohPair) return 2;
    if (num_elems == 0) return 0;
    if (((uintptr_t)d_hashes & 3) != 0 || ((uintptr_t)d_num_hashes & 7) != 0) return 2;
    if (pool_init() <= 0) return 1;
    hipError_t e = launch_poh_kernel((PohKernel)kernel, (uint32_t *)d_hashes, d_num_hashes,
                                     num_elems, (hipStream_t)stream, 0, ~0ull);
    if (e != hipSuccess) {
        fprintf(stderr, "spl_poh_device_kernel: launch failed: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

}  // extern "C"
