// GF(2^255 - 19) arithmetic for gfx950.
//
// Radix choice (measured on MI355X, gfx950):
//   V_MAD_U64_U32 / V_MAD_I64_I32 (32x32 + 64 -> 64) issue at ~0.8-0.9 of the
//   int32 VALU peak, i.e. a full-rate multiply-accumulate.  What decides the
//   cost of a field multiply is therefore the bookkeeping around it:
//     - radix 2^32 (8 words): 64 products, but every product needs a carry
//       out of the 64-bit accumulator (V_MAD_U64_U32's SGPR carry + an ADDC,
//       with SGPR hazard NOPs between them) and every add/sub needs a carry
//       chain plus a fold of 2^256 = 38;
//     - radix 2^25.5 (10 signed limbs of alternately 26 and 25 bits): 100
//       products, none of which can overflow a signed 64-bit column sum, the
//       wrap-around factor 19 is folded into a 32-bit multiplier operand
//       (19 * limb < 2^31), squares need only 55 products, and add/sub are 10
//       independent 32-bit adds.
//   Both multiplies measured ~175 lane-cycles; the 2^25.5 form wins on
//   squares (55 vs 36 products + carries + doubling) and on add/sub, which
//   dominate the point formulas, so it is used here.
//
// Element: v[i] has weight 2^ceil(25.5 i).  "Tight" elements (outputs of mul,
// sq, carry) have |v_even| <= ~2^25 and |v_odd| <= ~2^24; a sum or difference
// of at most three tight elements is a valid input to mul/sq (the column sums
// then stay below 2^62).  fe_to_words / fe_canon give the unique
// representative in [0, p).
//
// Every routine is __host__ __device__ so the exported host helpers run the
// same arithmetic as the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define SPL_HD __host__ __device__ __forceinline__

namespace spl {

// Calls f(std::integral_constant<int, j>{}) for j = 0 .. N-1, expanded at
// compile time.  Used for the lockstep forms below: a "#pragma unroll" loop
// whose body holds several field multiplications can be left rolled by the
// unroller, which puts the per-element arrays in scratch memory.
template <int N, class F>
SPL_HD void for_each_n(F &&f) {
    if constexpr (N > 0) {
        for_each_n<N - 1>(f);
        f(std::integral_constant<int, N - 1>{});
    }
}

struct fe {
    int32_t v[10];
};

SPL_HD uint32_t lo32(uint64_t x) { return (uint32_t)x; }
SPL_HD uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }

// limb i holds 26 bits when i is even, 25 bits when odd
#define SPL_LIMB_BITS(i) (((i) & 1) ? 25 : 26)

SPL_HD void fe_set(fe &r, int32_t small) {
    r.v[0] = small;
#pragma unroll
    for (int i = 1; i < 10; i++) r.v[i] = 0;
}

SPL_HD void fe_add(fe &r, const fe &a, const fe &b) {
#pragma unroll
    for (int i = 0; i < 10; i++) r.v[i] = a.v[i] + b.v[i];
}

SPL_HD void fe_sub(fe &r, const fe &a, const fe &b) {
#pragma unroll
    for (int i = 0; i < 10; i++) r.v[i] = a.v[i] - b.v[i];
}

SPL_HD void fe_neg(fe &r, const fe &a) {
#pragma unroll
    for (int i = 0; i < 10; i++) r.v[i] = -a.v[i];
}

SPL_HD void fe_cmov(fe &r, const fe &a, uint32_t flag) {
    int32_t m = -(int32_t)flag;
#pragma unroll
    for (int i = 0; i < 10; i++) r.v[i] ^= (r.v[i] ^ a.v[i]) & m;
}

// Rounding carry of limb i into limb i+1 (limb 9 wraps into limb 0 times 19).
template <typename T>
SPL_HD void carry_round(T c[10], int i) {
    const int b = SPL_LIMB_BITS(i);
    T cy = (c[i] + ((T)1 << (b - 1))) >> b;
    c[i] -= cy * ((T)1 << b);
    if (i == 9) c[0] += 19 * cy;
    else c[i + 1] += cy;
}

// Rounding offsets for the column sums.  Column i's carry is rounded when the
// column holds value + 2^(b_i - 1) as that carry is taken.  An even column
// starts at its own offset, 2^25, plus the next (odd) column's, 2^24, times
// 2^26: that multiple of 2^26 leaves the even limb's bits alone and arrives
// in the odd column as +2^24 inside the carry, so odd columns start at 0.
// (The compiler adds a column's starting value with a 64-bit add of its own,
// so five offsets cost half the adds of ten.)
#define SPL_COL_BIAS(i) \
    (((i) & 1) ? (int64_t)0 : ((int64_t)1 << 25) + ((int64_t)1 << 50))

// First carry of a biased column (c[i] = value + 2^(b-1) + a multiple of
// 2^b that is carried on): the rounded carry is c[i] >> b and the limb is
// (c[i] mod 2^b) - 2^(b-1), in [-2^(b-1), 2^(b-1)).
SPL_HD void carry_biased(int64_t c[10], int i) {
    const int b = SPL_LIMB_BITS(i);
    const int64_t cy = c[i] >> b;
    c[i] = (c[i] & (((int64_t)1 << b) - 1)) - ((int64_t)1 << (b - 1));
    if (i == 9) c[0] += 19 * cy;
    else c[i + 1] += cy;
}

// Biased 64-bit column sums (SPL_COL_BIAS added) -> tight limbs.  The carries
// run 0 -> 9, so each odd column has its offset (from the even column below)
// before its own carry is taken.  One carry
// chain 0 -> 9, each column taking its carry-in before its own carry, then
// the wrap 9 -> 0 (times 19) and one more carry out of limb 0, which leaves
// limb 1 at most ~2^15 above its tight range.  Eleven carries: two chains
// started at limbs 0 and 4 need twelve (limb 4 is carried again after the
// first chain reaches it), and with one wave per SIMD the shorter dependency
// depth buys nothing (measured on gfx950: 586 against 612 cycles per
// multiply or square).
SPL_HD void fe_reduce_cols(fe &h, int64_t c[10]) {
    carry_biased(c, 0); carry_biased(c, 1); carry_biased(c, 2); carry_biased(c, 3);
    carry_biased(c, 4); carry_biased(c, 5); carry_biased(c, 6); carry_biased(c, 7);
    carry_biased(c, 8); carry_biased(c, 9);
    carry_round(c, 0);
#pragma unroll
    for (int i = 0; i < 10; i++) h.v[i] = (int32_t)c[i];
}

// Loose 32-bit limbs -> tight limbs.
SPL_HD void fe_carry(fe &h) {
    int32_t *c = h.v;
    carry_round(c, 0); carry_round(c, 4);
    carry_round(c, 1); carry_round(c, 5);
    carry_round(c, 2); carry_round(c, 6);
    carry_round(c, 3); carry_round(c, 7);
    carry_round(c, 4); carry_round(c, 8);
    carry_round(c, 9);
    carry_round(c, 0);
}

// h = f * g.  Column k collects f_i g_j for i + j = k (weight 1, or 2 when i
// and j are both odd: their weights add up half a bit above limb k's) and, with
// the factor 19 for 2^255 = 19, the products with i + j = k + 10.
SPL_HD void fe_mul(fe &h, const fe &f, const fe &g) {
    int32_t g19[10], f2[10];
#pragma unroll
    for (int i = 0; i < 10; i++) {
        g19[i] = 19 * g.v[i];
        f2[i] = (i & 1) ? 2 * f.v[i] : f.v[i];
    }
    int64_t c[10];
#pragma unroll
    for (int k = 0; k < 10; k++) {
        int64_t s = SPL_COL_BIAS(k);
#pragma unroll
        for (int i = 0; i < 10; i++) {
            int j = k - i;
            const bool wrap = j < 0;
            if (wrap) j += 10;
            const int32_t a = ((i & 1) && (j & 1)) ? f2[i] : f.v[i];
            const int32_t b = wrap ? g19[j] : g.v[j];
            s += (int64_t)a * b;
        }
        c[k] = s;
    }
    fe_reduce_cols(h, c);
}

// Column sums of f^2 plus SPL_COL_BIAS >> bias_shift: each unordered pair
// once, the factor 2 of the cross terms on the left operand and the odd-odd
// factor 2 / wrap factor 19 on the right one (38 f_j only for odd, 25-bit,
// limbs so it fits in 31 bits).
SPL_HD void fe_sq_cols(int64_t c[10], const fe &f, int bias_shift = 0) {
    int32_t f2[10], f19[10], f38[10];
#pragma unroll
    for (int i = 0; i < 10; i++) {
        f2[i] = 2 * f.v[i];
        f19[i] = 19 * f.v[i];
        f38[i] = 38 * f.v[i];
    }
#pragma unroll
    for (int k = 0; k < 10; k++) c[k] = SPL_COL_BIAS(k) >> bias_shift;
#pragma unroll
    for (int i = 0; i < 10; i++) {
#pragma unroll
        for (int j = i; j < 10; j++) {
            const bool wrap = i + j >= 10;
            const bool oo = (i & 1) && (j & 1);
            const int32_t a = (i == j) ? f.v[i] : f2[i];
            int32_t b;
            if (wrap) b = oo ? f38[j] : f19[j];
            else b = oo ? f2[j] : f.v[j];
            c[(i + j) % 10] += (int64_t)a * b;
        }
    }
}

SPL_HD void fe_sq(fe &h, const fe &f) {
    int64_t c[10];
    fe_sq_cols(c, f);
    fe_reduce_cols(h, c);
}

// h = 2 f^2 (half the bias before doubling: the doubled sums carry it whole)
SPL_HD void fe_sq2(fe &h, const fe &f) {
    int64_t c[10];
    fe_sq_cols(c, f, 1);
#pragma unroll
    for (int k = 0; k < 10; k++) c[k] += c[k];
    fe_reduce_cols(h, c);
}

// Unique representative in [0, p), as limbs in [0, 2^26) / [0, 2^25).
SPL_HD void fe_canon(fe &r, const fe &a) {
    int32_t h[10];
#pragma unroll
    for (int i = 0; i < 10; i++) h[i] = a.v[i];
    // tight, then + 2p so every limb is positive
    fe t;
#pragma unroll
    for (int i = 0; i < 10; i++) t.v[i] = h[i];
    fe_carry(t);
#pragma unroll
    for (int i = 0; i < 10; i++)
        h[i] = t.v[i] + ((i & 1) ? 2 * 0x1ffffff : (i == 0 ? 2 * 0x3ffffed : 2 * 0x3ffffff));
    // two floor-carry passes: limbs in range, value < 2^255
#pragma unroll
    for (int pass = 0; pass < 2; pass++) {
#pragma unroll
        for (int i = 0; i < 9; i++) {
            const int b = SPL_LIMB_BITS(i);
            int32_t cy = h[i] >> b;
            h[i] &= (1 << b) - 1;
            h[i + 1] += cy;
        }
        int32_t cy = h[9] >> 25;
        h[9] &= (1 << 25) - 1;
        h[0] += 19 * cy;
    }
    // subtract p once if value >= p, i.e. if value + 19 >= 2^255
    int32_t s[10];
    int32_t cy = 19;
#pragma unroll
    for (int i = 0; i < 10; i++) {
        const int b = SPL_LIMB_BITS(i);
        int32_t x = h[i] + cy;
        cy = x >> b;
        s[i] = x & ((1 << b) - 1);
    }
    const int32_t m = -cy;   // cy is 1 exactly when value >= p
#pragma unroll
    for (int i = 0; i < 10; i++) r.v[i] = (h[i] & ~m) | (s[i] & m);
}

// 8 little-endian words of the canonical value.
SPL_HD void fe_to_words(uint32_t w[8], const fe &a) {
    fe c;
    fe_canon(c, a);
    const uint32_t *h = reinterpret_cast<const uint32_t *>(c.v);
    w[0] = h[0] | (h[1] << 26);
    w[1] = (h[1] >> 6) | (h[2] << 19);
    w[2] = (h[2] >> 13) | (h[3] << 13);
    w[3] = (h[3] >> 19) | (h[4] << 6);
    w[4] = h[5] | (h[6] << 25);
    w[5] = (h[6] >> 7) | (h[7] << 19);
    w[6] = (h[7] >> 13) | (h[8] << 12);
    w[7] = (h[8] >> 20) | (h[9] << 6);
}

// Low 255 bits of 8 little-endian words (bit 255 ignored; a value >= p is
// kept as-is and reduced by the arithmetic).
SPL_HD void fe_from_words(fe &r, const uint32_t w[8]) {
    r.v[0] = (int32_t)(w[0] & 0x3ffffff);
    r.v[1] = (int32_t)(((w[0] >> 26) | (w[1] << 6)) & 0x1ffffff);
    r.v[2] = (int32_t)(((w[1] >> 19) | (w[2] << 13)) & 0x3ffffff);
    r.v[3] = (int32_t)(((w[2] >> 13) | (w[3] << 19)) & 0x1ffffff);
    r.v[4] = (int32_t)((w[3] >> 6) & 0x3ffffff);
    r.v[5] = (int32_t)(w[4] & 0x1ffffff);
    r.v[6] = (int32_t)(((w[4] >> 25) | (w[5] << 7)) & 0x3ffffff);
    r.v[7] = (int32_t)(((w[5] >> 19) | (w[6] << 13)) & 0x1ffffff);
    r.v[8] = (int32_t)(((w[6] >> 12) | (w[7] << 20)) & 0x3ffffff);
    r.v[9] = (int32_t)((w[7] >> 6) & 0x1ffffff);
}

SPL_HD uint32_t fe_iszero(const fe &a) {
    fe c;
    fe_canon(c, a);
    int32_t o = 0;
#pragma unroll
    for (int i = 0; i < 10; i++) o |= c.v[i];
    return o == 0;
}

SPL_HD uint32_t fe_isneg(const fe &a) {
    fe c;
    fe_canon(c, a);
    return (uint32_t)c.v[0] & 1u;
}

// The exponent chains below take N independent elements through the same
// steps in lockstep: every step is applied to all N before the next one, so a
// loop of squarings holds N independent squarings per iteration and the
// scheduler can overlap their carry chains (a single chain is one long
// dependency: each squaring's carries wait for the previous one).

// r[j] <- r[j]^(2^n) for every j.
template <int N>
SPL_HD void fe_sqn_n(fe (&r)[N], int n) {
    for (int i = 0; i < n; i++)
        for_each_n<N>([&](auto j) { fe_sq(r[j], r[j]); });
}

// z^(2^250 - 1) and z^11: the common prefix of the inversion and square-root
// exponent chains.
template <int N>
SPL_HD void fe_chain250_n(fe (&z250)[N], fe (&z11)[N], const fe (&z)[N]) {
    fe z2[N], z9[N], t[N], z5[N], z10[N], z20[N], z50[N], z100[N];
    for_each_n<N>([&](auto j) { fe_sq(z2[j], z[j]); t[j] = z2[j]; });
    fe_sqn_n(t, 2);
    for_each_n<N>([&](auto j) {
        fe_mul(z9[j], t[j], z[j]);
        fe_mul(z11[j], z9[j], z2[j]);
        fe_sq(t[j], z11[j]);
        fe_mul(z5[j], t[j], z9[j]);
        t[j] = z5[j];
    });
    fe_sqn_n(t, 5);
    for_each_n<N>([&](auto j) { fe_mul(z10[j], t[j], z5[j]); t[j] = z10[j]; });
    fe_sqn_n(t, 10);
    for_each_n<N>([&](auto j) { fe_mul(z20[j], t[j], z10[j]); t[j] = z20[j]; });
    fe_sqn_n(t, 20);
    for_each_n<N>([&](auto j) { fe_mul(t[j], t[j], z20[j]); });     // 2^40 - 1
    fe_sqn_n(t, 10);
    for_each_n<N>([&](auto j) { fe_mul(z50[j], t[j], z10[j]); t[j] = z50[j]; });
    fe_sqn_n(t, 50);
    for_each_n<N>([&](auto j) { fe_mul(z100[j], t[j], z50[j]); t[j] = z100[j]; });
    fe_sqn_n(t, 100);
    for_each_n<N>([&](auto j) { fe_mul(t[j], t[j], z100[j]); });    // 2^200 - 1
    fe_sqn_n(t, 50);
    for_each_n<N>([&](auto j) { fe_mul(z250[j], t[j], z50[j]); });
}

// z^((p-5)/8) = z^(2^252 - 3)
template <int N>
SPL_HD void fe_pow22523_n(fe (&r)[N], const fe (&z)[N]) {
    fe z250[N], z11[N];
    fe_chain250_n(z250, z11, z);
    fe_sqn_n(z250, 2);
    for_each_n<N>([&](auto j) { fe_mul(r[j], z250[j], z[j]); });
}

// z^(p-2)
SPL_HD void fe_invert(fe &r, const fe &z) {
    fe z250[1], z11[1];
    const fe in[1] = {z};
    fe_chain250_n(z250, z11, in);
    fe_sqn_n(z250, 5);
    fe_mul(r, z250[0], z11[0]);
}

SPL_HD void fe_pow22523(fe &r, const fe &z) {
    fe out[1];
    const fe in[1] = {z};
    fe_pow22523_n(out, in);
    r = out[0];
}

}  // namespace spl
