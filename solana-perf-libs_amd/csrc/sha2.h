// FIPS 180-4 SHA-256 / SHA-512 cores for gfx950 (and the host).
//
// Replaces the reference's LibTomCrypt-derived code
// (src/cuda-sha256/sha256.cu:63-277, src/cuda-ecc-ed25519/sha512.cu:21-280).
// The cores work on message words that the caller has already assembled, so a
// kernel can build its blocks straight from packet memory without a byte
// buffer; SHA-512's 64-bit words are held as 32-bit halves by the compiler
// (V_ALIGNBIT_B32 pairs for the rotates, add/addc pairs for the sums).
#pragma once
#include <utility>

#include "fe25519.h"

namespace spl {

SPL_HD uint32_t rotr32(uint32_t x, int n) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(x, x, n);
#else
    return (x >> n) | (x << (32 - n));
#endif
}

// 64-bit rotate and shift on the two 32-bit halves: a rotate is two
// V_ALIGNBIT_B32 (the compiler otherwise builds it from two 64-bit shifts and
// two ORs), a right shift by n < 32 one V_ALIGNBIT_B32 and one 32-bit shift.
// n is a compile-time constant at every use, so the n < 32 test folds.
SPL_HD uint64_t rotr64(uint64_t x, int n) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    uint32_t rl, rh;
    if (n < 32) {
        rl = __builtin_amdgcn_alignbit(hi, lo, n);
        rh = __builtin_amdgcn_alignbit(lo, hi, n);
    } else {
        rl = __builtin_amdgcn_alignbit(lo, hi, n - 32);
        rh = __builtin_amdgcn_alignbit(hi, lo, n - 32);
    }
    return ((uint64_t)rh << 32) | rl;
#else
    return (x >> n) | (x << (64 - n));
#endif
}

SPL_HD uint64_t shr64(uint64_t x, int n) {   // 0 < n < 32
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    return ((uint64_t)(hi >> n) << 32) | __builtin_amdgcn_alignbit(hi, lo, n);
#else
    return x >> n;
#endif
}

// a ^ b ^ c in one V_BITOP3_B32 (truth table 0x96) on gfx950; the compiler
// otherwise emits two V_XOR_B32 for each of the Sigma/sigma functions.
// Measured on the PoH loop: 1552 -> 1402 instructions per SHA-256, and
// 1344 with the majority below (the compiler then also forms Ch with it).
SPL_HD uint32_t xor3_32(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

// Majority of three (truth table 0xe8, symmetric in its inputs).
SPL_HD uint32_t maj32(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0xe8);
#else
    return (a & b) | (c & (a | b));
#endif
}

// Choose: e ? f : g, bit by bit (truth table 0xca; V_BITOP3_B32 indexes its
// table by (S0, S1, S2) with S0 as the high bit).  Written as
// g ^ (e & (f ^ g)) the compiler forms one instruction in the one-lane PoH
// and SHA-512 rounds but two (V_XOR_B32 + V_BITOP3_B32) in the split PoH
// kernel's rounds.
SPL_HD uint32_t ch32(uint32_t e, uint32_t f, uint32_t g) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(e, f, g, 0xca);
#else
    return g ^ (e & (f ^ g));
#endif
}

SPL_HD uint64_t maj64(uint64_t a, uint64_t b, uint64_t c) {
    return ((uint64_t)maj32((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32))
            << 32) |
           maj32((uint32_t)a, (uint32_t)b, (uint32_t)c);
}

SPL_HD uint64_t xor3_64(uint64_t a, uint64_t b, uint64_t c) {
    return ((uint64_t)xor3_32((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32))
            << 32) |
           xor3_32((uint32_t)a, (uint32_t)b, (uint32_t)c);
}

SPL_HD uint32_t bswap32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(0, x, 0x00010203u);
#else
    return __builtin_bswap32(x);
#endif
}

// ---------------------------------------------------------------- SHA-256

#define SPL_SHA256_K                                                              \
    {0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, \
     0x923f82a4u, 0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, \
     0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, \
     0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, \
     0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, \
     0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, \
     0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu, \
     0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, \
     0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, \
     0x5b9cca4fu, 0x682e6ff3u, 0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, \
     0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u}

#define SPL_SHA256_H0                                                       \
    {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au, 0x510e527fu, 0x9b05688cu, \
     0x1f83d9abu, 0x5be0cd19u}

// One SHA-256 of a 32-byte message whose big-endian words are st[0..7],
// replacing st with the digest words.  With the message fixed at 32 bytes
// the second half of the block is the constant padding
// (0x80000000, 0 x 6, 256), which the compiler folds into the schedule.
SPL_HD void sha256_32byte(uint32_t st[8]) {
    const uint32_t K[64] = SPL_SHA256_K;
    const uint32_t H0[8] = SPL_SHA256_H0;
    uint32_t w[16];
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = st[i];
    w[8] = 0x80000000u;
#pragma unroll
    for (int i = 9; i < 15; i++) w[i] = 0;
    w[15] = 256;
    uint32_t a = H0[0], b = H0[1], c = H0[2], d = H0[3];
    uint32_t e = H0[4], f = H0[5], g = H0[6], h = H0[7];
#pragma unroll
    for (int t = 0; t < 64; t++) {
        uint32_t wt;
        if (t < 16) {
            wt = w[t];
        } else {
            uint32_t w15 = w[(t - 15) & 15], w2 = w[(t - 2) & 15];
            uint32_t s0 = xor3_32(rotr32(w15, 7), rotr32(w15, 18), w15 >> 3);
            uint32_t s1 = xor3_32(rotr32(w2, 17), rotr32(w2, 19), w2 >> 10);
            wt = w[t & 15] + s0 + w[(t - 7) & 15] + s1;
            w[t & 15] = wt;
        }
        uint32_t S1 = xor3_32(rotr32(e, 6), rotr32(e, 11), rotr32(e, 25));
        uint32_t ch = g ^ (e & (f ^ g));
        uint32_t t1 = h + S1 + ch + K[t] + wt;
        uint32_t S0 = xor3_32(rotr32(a, 2), rotr32(a, 13), rotr32(a, 22));
        uint32_t maj = maj32(a, b, c);
        uint32_t t2 = S0 + maj;
        h = g; g = f; f = e; e = d + t1;
        d = c; c = b; b = a; a = t1 + t2;
    }
    st[0] = H0[0] + a; st[1] = H0[1] + b; st[2] = H0[2] + c; st[3] = H0[3] + d;
    st[4] = H0[4] + e; st[5] = H0[5] + f; st[6] = H0[6] + g; st[7] = H0[7] + h;
}

// ---------------------------------------------------------------- SHA-512

#define SPL_SHA512_K                                                  \
    {0x428a2f98d728ae22ull, 0x7137449123ef65cdull, 0xb5c0fbcfec4d3b2full, \
     0xe9b5dba58189dbbcull, 0x3956c25bf348b538ull, 0x59f111f1b605d019ull, \
     0x923f82a4af194f9bull, 0xab1c5ed5da6d8118ull, 0xd807aa98a3030242ull, \
     0x12835b0145706fbeull, 0x243185be4ee4b28cull, 0x550c7dc3d5ffb4e2ull, \
     0x72be5d74f27b896full, 0x80deb1fe3b1696b1ull, 0x9bdc06a725c71235ull, \
     0xc19bf174cf692694ull, 0xe49b69c19ef14ad2ull, 0xefbe4786384f25e3ull, \
     0x0fc19dc68b8cd5b5ull, 0x240ca1cc77ac9c65ull, 0x2de92c6f592b0275ull, \
     0x4a7484aa6ea6e483ull, 0x5cb0a9dcbd41fbd4ull, 0x76f988da831153b5ull, \
     0x983e5152ee66dfabull, 0xa831c66d2db43210ull, 0xb00327c898fb213full, \
     0xbf597fc7beef0ee4ull, 0xc6e00bf33da88fc2ull, 0xd5a79147930aa725ull, \
     0x06ca6351e003826full, 0x142929670a0e6e70ull, 0x27b70a8546d22ffcull, \
     0x2e1b21385c26c926ull, 0x4d2c6dfc5ac42aedull, 0x53380d139d95b3dfull, \
     0x650a73548baf63deull, 0x766a0abb3c77b2a8ull, 0x81c2c92e47edaee6ull, \
     0x92722c851482353bull, 0xa2bfe8a14cf10364ull, 0xa81a664bbc423001ull, \
     0xc24b8b70d0f89791ull, 0xc76c51a30654be30ull, 0xd192e819d6ef5218ull, \
     0xd69906245565a910ull, 0xf40e35855771202aull, 0x106aa07032bbd1b8ull, \
     0x19a4c116b8d2d0c8ull, 0x1e376c085141ab53ull, 0x2748774cdf8eeb99ull, \
     0x34b0bcb5e19b48a8ull, 0x391c0cb3c5c95a63ull, 0x4ed8aa4ae3418acbull, \
     0x5b9cca4f7763e373ull, 0x682e6ff3d6b2b8a3ull, 0x748f82ee5defb2fcull, \
     0x78a5636f43172f60ull, 0x84c87814a1f0ab72ull, 0x8cc702081a6439ecull, \
     0x90befffa23631e28ull, 0xa4506cebde82bde9ull, 0xbef9a3f7b2c67915ull, \
     0xc67178f2e372532bull, 0xca273eceea26619cull, 0xd186b8c721c0c207ull, \
     0xeada7dd6cde0eb1eull, 0xf57d4f7fee6ed178ull, 0x06f067aa72176fbaull, \
     0x0a637dc5a2c898a6ull, 0x113f9804bef90daeull, 0x1b710b35131c471bull, \
     0x28db77f523047d84ull, 0x32caab7b40c72493ull, 0x3c9ebe0a15c9bebcull, \
     0x431d67c49c100d4cull, 0x4cc5d4becb3e42b6ull, 0x597f299cfc657e2aull, \
     0x5fcb6fab3ad6faecull, 0x6c44198c4a475817ull}

SPL_HD void sha512_init(uint64_t st[8]) {
    st[0] = 0x6a09e667f3bcc908ull; st[1] = 0xbb67ae8584caa73bull;
    st[2] = 0x3c6ef372fe94f82bull; st[3] = 0xa54ff53a5f1d36f1ull;
    st[4] = 0x510e527fade682d1ull; st[5] = 0x9b05688c2b3e6c1full;
    st[6] = 0x1f83d9abfb41bd6bull; st[7] = 0x5be0cd19137e2179ull;
}

// Round T of SHA-512 on the working variables v[8], which rotate roles
// instead of values: a is v[(8 - T mod 8) mod 8], b the next one, and so on.
// The round writes the new a into h's register and the new e into d's.
template <int T>
SPL_HD void sha512_round(uint64_t v[8], uint64_t w[16]) {
    constexpr uint64_t K[80] = SPL_SHA512_K;
    constexpr int A = (8 - T % 8) % 8;
    uint64_t wt;
    if (T < 16) {
        wt = w[T];
    } else {
        const uint64_t w15 = w[(T - 15) & 15], w2 = w[(T - 2) & 15];
        const uint64_t s0 = xor3_64(rotr64(w15, 1), rotr64(w15, 8), shr64(w15, 7));
        const uint64_t s1 = xor3_64(rotr64(w2, 19), rotr64(w2, 61), shr64(w2, 6));
        wt = w[T & 15] + s0 + w[(T - 7) & 15] + s1;
        w[T & 15] = wt;
    }
    const uint64_t a = v[A], b = v[(A + 1) % 8], c = v[(A + 2) % 8];
    const uint64_t e = v[(A + 4) % 8], f = v[(A + 5) % 8], g = v[(A + 6) % 8];
    const uint64_t S1 = xor3_64(rotr64(e, 14), rotr64(e, 18), rotr64(e, 41));
    const uint64_t ch = g ^ (e & (f ^ g));
    const uint64_t t1 = v[(A + 7) % 8] + S1 + ch + K[T] + wt;
    const uint64_t S0 = xor3_64(rotr64(a, 28), rotr64(a, 34), rotr64(a, 39));
    v[(A + 3) % 8] += t1;                        // e of the next round
    v[(A + 7) % 8] = t1 + S0 + maj64(a, b, c);   // a of the next round
}

template <int... Ts>
SPL_HD void sha512_rounds(std::integer_sequence<int, Ts...>, uint64_t v[8], uint64_t w[16]) {
    using expand = int[];
    (void)expand{0, (sha512_round<Ts>(v, w), 0)...};
}

// One compression; w[0..15] are the block's big-endian 64-bit words and are
// overwritten (rolling schedule).  The 80 rounds are straight-line code by
// construction (no loop for the unroller to weigh).
SPL_HD void sha512_block(uint64_t st[8], uint64_t w[16]) {
    uint64_t v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = st[i];
    sha512_rounds(std::make_integer_sequence<int, 80>{}, v, w);
#pragma unroll
    for (int i = 0; i < 8; i++) st[i] += v[i];
}

// Digest words (big-endian 64-bit) -> 16 little-endian 32-bit words of the
// 64-byte digest, the layout the scalar code reads.
SPL_HD void sha512_digest_le_words(uint32_t out[16], const uint64_t st[8]) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
        out[2 * i] = bswap32((uint32_t)(st[i] >> 32));
        out[2 * i + 1] = bswap32((uint32_t)st[i]);
    }
}

}  // namespace spl
