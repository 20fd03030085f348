// Scalars modulo L = 2^252 + 27742317777372353535851937790883648493, radix 2^32.
//
// The reference reduces with ref10's 21-bit-limb signed arithmetic
// (src/cuda-ecc-ed25519/sc.cu:238-550 sc_reduce, :565-1012 sc_muladd).  Here
// both are Barrett reductions (HAC 14.42, b = 2^32, k = 8) on 32-bit words so
// they run on V_MAD_U64_U32; the results are the same integers mod L.
#pragma once
#include "fe25519.h"

namespace spl {

#define SPL_L_WORDS {0x5cf5d3edu, 0x5812631au, 0xa2f79cd6u, 0x14def9deu, 0u, 0u, 0u, 0x10000000u}
// mu = floor(2^512 / L), 9 words
#define SPL_MU_WORDS {0x0a2c131bu, 0xed9ce5a3u, 0x086329a7u, 0x2106215du, 0xffffffebu, \
                      0xffffffffu, 0xffffffffu, 0xffffffffu, 0x0000000fu}

// r (8 words) = x (16 words) mod L
SPL_HD void sc_reduce512(uint32_t r[8], const uint32_t x[16]) {
    const uint32_t L[8] = SPL_L_WORDS;
    const uint32_t MU[9] = SPL_MU_WORDS;
    // q2 = q1 * mu with q1 = x >> 224 (words 7..15); only words 9..17 of q2
    // are needed (q3 = q2 >> 288), but the lower words feed the carries.
    uint32_t q2[18];
#pragma unroll
    for (int i = 0; i < 18; i++) q2[i] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 9; j++) {
            c = (uint64_t)x[7 + i] * MU[j] + q2[i + j] + hi32(c);
            q2[i + j] = lo32(c);
        }
        q2[i + 9] = hi32(c);
    }
    // r2 = (q3 * L) mod 2^288, q3 = q2[9..17]
    uint32_t r2[9];
#pragma unroll
    for (int i = 0; i < 9; i++) r2[i] = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = 0; j < 8 && i + j < 9; j++) {
            c = (uint64_t)q2[9 + i] * L[j] + r2[i + j] + hi32(c);
            r2[i + j] = lo32(c);
        }
        if (i == 0) r2[8] = hi32(c);   // row 0's carry is still below 2^288
    }
    // r = (x mod 2^288) - r2 (mod 2^288); 0 <= r < 3L
    uint32_t t[9];
    int64_t b = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        b = (int64_t)x[i] - r2[i] + (b >> 32);
        t[i] = (uint32_t)b;
    }
    // subtract L while r >= L (at most twice)
#pragma unroll
    for (int rep = 0; rep < 2; rep++) {
        uint32_t s[9];
        int64_t bb = 0;
#pragma unroll
        for (int i = 0; i < 9; i++) {
            bb = (int64_t)t[i] - (i < 8 ? L[i] : 0u) + (bb >> 32);
            s[i] = (uint32_t)bb;
        }
        uint32_t keep = (uint32_t)(bb >> 32) & 1u;   // borrow: t < L
        uint32_t m = keep - 1u;                       // all ones when t >= L
#pragma unroll
        for (int i = 0; i < 9; i++) t[i] = (s[i] & m) | (t[i] & ~m);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) r[i] = t[i];
}

// r = (a*b + c) mod L, all 8-word little-endian (a, b, c < 2^256).
SPL_HD void sc_muladd(uint32_t r[8], const uint32_t a[8], const uint32_t b[8],
                      const uint32_t c[8]) {
    uint32_t t[16];
#pragma unroll
    for (int i = 0; i < 16; i++) t[i] = i < 8 ? c[i] : 0u;
    // operand scanning: row i adds a_i*b into t[i..i+7] and leaves its carry
    // in t[i+8], which no earlier row has written.
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t cc = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            cc = (uint64_t)a[i] * b[j] + t[i + j] + hi32(cc);
            t[i + j] = lo32(cc);
        }
        t[i + 8] = hi32(cc);
    }
    sc_reduce512(r, t);
}

}  // namespace spl
