// GF(2^255 - 19) arithmetic for gfx950, radix 2^32.
//
// The reference keeps field elements as 10 signed limbs in radix 2^25.5
// (src/cuda-ecc-ed25519/fe.h:7-16, fe.cu) so that CUDA's 32x32->64 IMAD can
// accumulate 100 limb products per multiply.  On CDNA4 the natural multiplier
// is V_MAD_U64_U32 (32x32 + 64 -> 64, one instruction), so the radix here is
// 2^32: 8 unsigned words, 64 limb products per multiply and 36 per square.
//
// Representation: w[0..7] little-endian words of an integer v < 2^256 with
// v == x (mod p).  Values are kept only "weakly" reduced (mod 2^256 - 38);
// fe_canon() produces the unique representative in [0, p).
//
// Every routine is __host__ __device__ so the exported host helpers
// (ed25519_verify, ed25519_check_packed_ge_small_order, ...) run the same code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SPL_HD __host__ __device__ __forceinline__

namespace spl {

struct fe {
    uint32_t w[8];
};

SPL_HD uint32_t lo32(uint64_t x) { return (uint32_t)x; }
SPL_HD uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }

SPL_HD void fe_set(fe &r, uint32_t v) {
    r.w[0] = v;
#pragma unroll
    for (int i = 1; i < 8; i++) r.w[i] = 0;
}

// r = a + b (mod 2^256 - 38)
SPL_HD void fe_add(fe &r, const fe &a, const fe &b) {
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        c = (uint64_t)a.w[i] + b.w[i] + hi32(c);
        r.w[i] = lo32(c);
    }
    // 2^256 == 38: fold the carry, then fold the (rare) second carry, which
    // can only occur when the result is below 38 and so cannot propagate.
    c = (uint64_t)r.w[0] + 38u * hi32(c);
    r.w[0] = lo32(c);
#pragma unroll
    for (int i = 1; i < 8; i++) {
        c = (uint64_t)r.w[i] + hi32(c);
        r.w[i] = lo32(c);
    }
    r.w[0] += 38u * hi32(c);
}

// r = a - b (mod 2^256 - 38)
SPL_HD void fe_sub(fe &r, const fe &a, const fe &b) {
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        c = (int64_t)a.w[i] - b.w[i] + (c >> 32);
        r.w[i] = (uint32_t)c;
    }
    // a borrow means 2^256 was added: take 38 back off (twice at most; the
    // second time only touches w[0], which is then >= 2^32 - 38).
    uint32_t bw = (uint32_t)(-(c >> 32));
    c = (int64_t)r.w[0] - 38 * (int64_t)bw;
    r.w[0] = (uint32_t)c;
#pragma unroll
    for (int i = 1; i < 8; i++) {
        c = (int64_t)r.w[i] + (c >> 32);
        r.w[i] = (uint32_t)c;
    }
    bw = (uint32_t)(-(c >> 32));
    r.w[0] -= 38u * bw;
}

SPL_HD void fe_neg(fe &r, const fe &a) {
    fe z;
    fe_set(z, 0);
    fe_sub(r, z, a);
}

// 512-bit t[0..15] -> r (mod 2^256 - 38): r = t_lo + 38 t_hi.
SPL_HD void fe_fold512(fe &r, const uint32_t t[16]) {
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        c = (uint64_t)t[8 + i] * 38u + t[i] + hi32(c);
        r.w[i] = lo32(c);
    }
    // hi32(c) < 39: fold once more with a full carry chain, then the last
    // carry (only when the value is tiny) into w[0].
    c = (uint64_t)hi32(c) * 38u + r.w[0];
    r.w[0] = lo32(c);
#pragma unroll
    for (int i = 1; i < 8; i++) {
        c = (uint64_t)r.w[i] + hi32(c);
        r.w[i] = lo32(c);
    }
    r.w[0] += 38u * hi32(c);
}

// r = a * b
SPL_HD void fe_mul(fe &r, const fe &a, const fe &b) {
    uint32_t t[16];
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        c = (uint64_t)a.w[0] * b.w[j] + hi32(c);
        t[j] = lo32(c);
    }
    t[8] = hi32(c);
#pragma unroll
    for (int i = 1; i < 8; i++) {
        c = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            c = (uint64_t)a.w[i] * b.w[j] + t[i + j] + hi32(c);
            t[i + j] = lo32(c);
        }
        t[i + 8] = hi32(c);
    }
    fe_fold512(r, t);
}

// r = a^2: 28 cross products once, doubled by a shift, plus 8 squares.
SPL_HD void fe_sq(fe &r, const fe &a) {
    uint32_t t[16];
#pragma unroll
    for (int i = 0; i < 16; i++) t[i] = 0;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = i + 1; j < 8; j++) {
            c = (uint64_t)a.w[i] * a.w[j] + t[i + j] + hi32(c);
            t[i + j] = lo32(c);
        }
        t[i + 8] = hi32(c);
    }
    // double: t <<= 1 (t < 2^511 before doubling)
    t[15] = (t[15] << 1) | (t[14] >> 31);
#pragma unroll
    for (int i = 14; i > 0; i--) t[i] = (t[i] << 1) | (t[i - 1] >> 31);
    t[0] <<= 1;
    // add the diagonal a_i^2 at word 2i
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint64_t sq = (uint64_t)a.w[i] * a.w[i];
        c = (uint64_t)t[2 * i] + lo32(sq) + hi32(c);
        t[2 * i] = lo32(c);
        c = (uint64_t)t[2 * i + 1] + hi32(sq) + hi32(c);
        t[2 * i + 1] = lo32(c);
    }
    fe_fold512(r, t);
}

SPL_HD void fe_sqn(fe &r, const fe &a, int n) {
    fe_sq(r, a);
    for (int i = 1; i < n; i++) fe_sq(r, r);
}

// r = a * k for a small constant k < 2^26
SPL_HD void fe_mul_small(fe &r, const fe &a, uint32_t k) {
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        c = (uint64_t)a.w[i] * k + hi32(c);
        r.w[i] = lo32(c);
    }
    c = (uint64_t)hi32(c) * 38u + r.w[0];
    r.w[0] = lo32(c);
#pragma unroll
    for (int i = 1; i < 8; i++) {
        c = (uint64_t)r.w[i] + hi32(c);
        r.w[i] = lo32(c);
    }
    r.w[0] += 38u * hi32(c);
}

// Unique representative in [0, p).
SPL_HD void fe_canon(fe &r, const fe &a) {
    // fold bit 255: v = (v mod 2^255) + 19 * bit255  (< 2^255 + 19)
    uint32_t top = a.w[7] >> 31;
    uint64_t c = (uint64_t)a.w[0] + 19u * top;
    r.w[0] = lo32(c);
#pragma unroll
    for (int i = 1; i < 7; i++) {
        c = (uint64_t)a.w[i] + hi32(c);
        r.w[i] = lo32(c);
    }
    c = (uint64_t)(a.w[7] & 0x7fffffffu) + hi32(c);
    r.w[7] = lo32(c);
    // subtract p once if v >= p, i.e. if v + 19 >= 2^255
    c = (uint64_t)r.w[0] + 19u;
    uint32_t s[8];
    s[0] = lo32(c);
#pragma unroll
    for (int i = 1; i < 8; i++) {
        c = (uint64_t)r.w[i] + hi32(c);
        s[i] = lo32(c);
    }
    uint32_t ge = s[7] >> 31;
    uint32_t m = 0u - ge;
#pragma unroll
    for (int i = 0; i < 7; i++) r.w[i] = (s[i] & m) | (r.w[i] & ~m);
    r.w[7] = ((s[7] & 0x7fffffffu) & m) | (r.w[7] & ~m);
}

SPL_HD uint32_t fe_iszero(const fe &a) {
    fe c;
    fe_canon(c, a);
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) o |= c.w[i];
    return o == 0;
}

SPL_HD uint32_t fe_isneg(const fe &a) {
    fe c;
    fe_canon(c, a);
    return c.w[0] & 1;
}

// bytes -> field element, ignoring bit 255 (non-canonical y >= p is kept).
SPL_HD void fe_from_words(fe &r, const uint32_t w[8]) {
#pragma unroll
    for (int i = 0; i < 7; i++) r.w[i] = w[i];
    r.w[7] = w[7] & 0x7fffffffu;
}

SPL_HD void fe_to_words(uint32_t w[8], const fe &a) {
    fe c;
    fe_canon(c, a);
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = c.w[i];
}

SPL_HD void fe_cmov(fe &r, const fe &a, uint32_t flag) {
    uint32_t m = 0u - flag;
#pragma unroll
    for (int i = 0; i < 8; i++) r.w[i] = (r.w[i] & ~m) | (a.w[i] & m);
}

// z^(2^250 - 1) and z^11: the common prefix of the inversion and square-root
// exponent chains.
SPL_HD void fe_chain250(fe &z250, fe &z11, const fe &z) {
    fe z2, z9, t, z5, z10, z20, z50, z100;
    fe_sq(z2, z);
    fe_sqn(t, z2, 2);
    fe_mul(z9, t, z);
    fe_mul(z11, z9, z2);
    fe_sq(t, z11);
    fe_mul(z5, t, z9);
    fe_sqn(t, z5, 5);
    fe_mul(z10, t, z5);
    fe_sqn(t, z10, 10);
    fe_mul(z20, t, z10);
    fe_sqn(t, z20, 20);
    fe_mul(t, t, z20);        // 2^40 - 1
    fe_sqn(t, t, 10);
    fe_mul(z50, t, z10);
    fe_sqn(t, z50, 50);
    fe_mul(z100, t, z50);
    fe_sqn(t, z100, 100);
    fe_mul(t, t, z100);       // 2^200 - 1
    fe_sqn(t, t, 50);
    fe_mul(z250, t, z50);
}

// z^(p-2)
SPL_HD void fe_invert(fe &r, const fe &z) {
    fe z250, z11;
    fe_chain250(z250, z11, z);
    fe_sqn(z250, z250, 5);
    fe_mul(r, z250, z11);
}

// z^((p-5)/8) = z^(2^252 - 3)
SPL_HD void fe_pow22523(fe &r, const fe &z) {
    fe z250, z11;
    fe_chain250(z250, z11, z);
    fe_sqn(z250, z250, 2);
    fe_mul(r, z250, z);
}

}  // namespace spl
