// Half-size scalars for verification: dimension-2 lattice basis reduction.
//
// Verification asks whether R' = [S]B + [h](-A) equals R.  For any y that is
// invertible modulo the group order 8L (y odd and not a multiple of L), the
// map P -> [y]P is a bijection of the curve group, so R' = R exactly when
//     [y S mod L]B + [x](-A) + [y](-R) = O,   x = y h (mod 8L)
// (B has order L; A and R may carry an 8-torsion component, hence 8L).  The
// pairs (x, y) with x = y h (mod 8L) form a lattice of determinant 8L ~ 2^255;
// its reduced basis has vectors of about 2^127.5 per coordinate, so the two
// variable-base multiplications need about 128 doublings instead of 253.
//
// half_scalars() finds such a pair with y odd and positive.  Exactness never
// depends on the reduction being optimal: every operation below is an exact
// unimodular integer transform of the basis {(8L, 0), (h, 1)}, so x = y h
// (mod 8L) holds by construction whatever the approximations decide; the
// approximations only affect how small x and y come out.  A result that is
// not small enough (only for inputs with an unusually short lattice vector) is
// replaced by (h, 1), which is the plain [h](-A) + [S]B check.
//
// Reduction: Euclid on (8L, h) with Lehmer-style rounds.  Each round takes the
// top 62 bits of both remainders at a common exponent, runs binary
// shift-and-subtract steps on those 64-bit approximations while recording the
// 2x2 cofactor matrix (entries kept below 2^29 in magnitude, and stopping
// while the approximated remainder is still >= 2^34 so the truncation error,
// at most the cofactor size, cannot flip a decision by more than one step),
// then applies the matrix exactly to the full remainders and cofactors.  It
// stops when the smaller remainder drops below 2^128; that remainder and its
// cofactor give the answer when the cofactor is odd; otherwise the balanced
// combination of the last two basis vectors (one of whose cofactors is odd)
// is used.
#pragma once
#include "sc25519.h"

namespace spl {

// 8L, little-endian words
#define SPL_8L_WORDS {0xe7ae9f68u, 0xc09318d2u, 0x17bce6b2u, 0xa6f7cef5u, 0u, 0u, 0u, 0x80000000u}

// Signed integers of 9 words (two's complement, 288 bits).
SPL_HD void s9_set_u8(uint32_t r[9], const uint32_t a[8]) {
#pragma unroll
    for (int i = 0; i < 8; i++) r[i] = a[i];
    r[8] = 0;
}

SPL_HD uint32_t s9_is_neg(const uint32_t a[9]) { return a[8] >> 31; }

// r = neg ? -a : a
SPL_HD void s9_cneg(uint32_t r[9], const uint32_t a[9], uint32_t neg) {
    const uint32_t m = 0u - neg;
    uint64_t c = neg;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        c += (uint64_t)(a[i] ^ m);
        r[i] = lo32(c);
        c >>= 32;
    }
}

// r = m0 x + m1 y for |m0|, |m1| < 2^30 (exact as long as the result fits).
SPL_HD void s9_lin2(uint32_t r[9], const uint32_t x[9], const uint32_t y[9], int32_t m0,
                    int32_t m1) {
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int64_t xi = i < 8 ? (int64_t)x[i] : (int64_t)(int32_t)x[i];
        const int64_t yi = i < 8 ? (int64_t)y[i] : (int64_t)(int32_t)y[i];
        c += xi * m0 + yi * m1;   // |xi m0| + |yi m1| < 2^63 - 2^33
        r[i] = (uint32_t)c;
        c >>= 32;                 // arithmetic: signed carry
    }
}

// a < b, both non-negative
SPL_HD uint32_t s9_lt(const uint32_t a[9], const uint32_t b[9]) {
    int64_t c = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        c = (int64_t)a[i] - (int64_t)b[i] + (c >> 32);
    }
    return (uint32_t)(c >> 63) & 1u;
}

SPL_HD int bitlen32(uint32_t x) { return x ? 32 - __builtin_clz(x) : 0; }
SPL_HD int bitlen64(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// Bit length of a non-negative 9-word integer.
SPL_HD int s9_bitlen(const uint32_t a[9]) {
    int r = 0;
#pragma unroll
    for (int i = 0; i < 9; i++)
        if (a[i]) r = 32 * i + bitlen32(a[i]);
    return r;
}

// Bits [e, e + 64) of a non-negative 9-word integer, 0 <= e < 288.
SPL_HD uint64_t s9_bits64(const uint32_t a[9], int e) {
    const int k = e >> 5, sh = e & 31;
    uint32_t w0 = 0, w1 = 0, w2 = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        w0 = i == k ? a[i] : w0;
        w1 = i == k + 1 ? a[i] : w1;
        w2 = i == k + 2 ? a[i] : w2;
    }
    const uint64_t lo = ((uint64_t)w1 << 32) | w0;
    const uint64_t hi = w2;
    return sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
}

// Shift-and-subtract Euclid steps on 64-bit approximations; returns the
// cofactor matrix [[m00, m01], [m10, m11]] applied to (x, y).
SPL_HD void lehmer_steps(uint64_t x, uint64_t y, uint64_t stop_below, int32_t m[4]) {
    int32_t m00 = 1, m01 = 0, m10 = 0, m11 = 1;
    const uint64_t floor_y = 1ull << 34;
    for (int it = 0; it < 96; it++) {
        if (y < stop_below || y < floor_y) break;
        int s = bitlen64(x) - bitlen64(y);
        if ((y << s) > x) s--;
        const uint32_t big1 = (uint32_t)(m10 < 0 ? -m10 : m10) | (uint32_t)(m11 < 0 ? -m11 : m11);
        const uint32_t big0 = (uint32_t)(m00 < 0 ? -m00 : m00) | (uint32_t)(m01 < 0 ? -m01 : m01);
        const int g = bitlen32(big1);    // = max of the two bit lengths
        const int g0 = bitlen32(big0);
        if (g + s >= 28 || g0 >= 29) break;   // keeps every entry below 2^29
        x -= y << s;
        m00 -= m10 * (1 << s);
        m01 -= m11 * (1 << s);
        if (x < y) {
            uint64_t t = x; x = y; y = t;
            int32_t u = m00; m00 = m10; m10 = u;
            u = m01; m01 = m11; m11 = u;
        }
    }
    m[0] = m00; m[1] = m01; m[2] = m10; m[3] = m11;
}

// Replaces (a, ta), (b, tb) by M applied to them, keeping a, b >= 0 (a sign
// flip of a row keeps the congruence) and a >= b.
SPL_HD void lattice_apply(uint32_t a[9], uint32_t ta[9], uint32_t b[9], uint32_t tb[9],
                          const int32_t m[4]) {
    uint32_t na[9], nb[9], nta[9], ntb[9];
    s9_lin2(na, a, b, m[0], m[1]);
    s9_lin2(nb, a, b, m[2], m[3]);
    s9_lin2(nta, ta, tb, m[0], m[1]);
    s9_lin2(ntb, ta, tb, m[2], m[3]);
    const uint32_t fa = s9_is_neg(na), fb = s9_is_neg(nb);
    s9_cneg(na, na, fa);
    s9_cneg(nta, nta, fa);
    s9_cneg(nb, nb, fb);
    s9_cneg(ntb, ntb, fb);
    const uint32_t sw = s9_lt(na, nb);
#pragma unroll
    for (int i = 0; i < 9; i++) {
        a[i] = sw ? nb[i] : na[i];
        b[i] = sw ? na[i] : nb[i];
        ta[i] = sw ? ntb[i] : nta[i];
        tb[i] = sw ? nta[i] : ntb[i];
    }
}

// One exact binary step a <- a - (b << s), ta <- ta - (tb << s) with
// s = bitlen(a) - bitlen(b) - 1 (>= 0), for when the quotient is too large for
// the 64-bit rounds; then restores a >= b.
SPL_HD void lattice_big_step(uint32_t a[9], uint32_t ta[9], uint32_t b[9], uint32_t tb[9]) {
    int s = s9_bitlen(a) - s9_bitlen(b) - 1;
    if (s < 0) s = 0;
    // (b << s) and (tb << s) as 9-word values: multiply by 2^(s mod 29) in
    // pieces of at most 2^28 so s9_lin2's operand bound holds
    uint32_t bs[9], ts[9], zero[9];
#pragma unroll
    for (int i = 0; i < 9; i++) { bs[i] = b[i]; ts[i] = tb[i]; zero[i] = 0; }
    while (s > 0) {
        const int k = s < 28 ? s : 28;
        s9_lin2(bs, bs, zero, 1 << k, 0);
        s9_lin2(ts, ts, zero, 1 << k, 0);
        s -= k;
    }
    uint32_t na[9], nta[9];
    s9_lin2(na, a, bs, 1, -1);
    s9_lin2(nta, ta, ts, 1, -1);
    const int32_t id[4] = {1, 0, 0, 1};
#pragma unroll
    for (int i = 0; i < 9; i++) { a[i] = na[i]; ta[i] = nta[i]; }
    lattice_apply(a, ta, b, tb, id);
}

// x = y h (mod 8L) with y odd and positive, |x| and y < 2^254; x is returned
// as magnitude + sign.  Falls back to (h, 1).
SPL_HD void half_scalars(const uint32_t h[8], uint32_t x[8], uint32_t &x_neg, uint32_t y[8]) {
    const uint32_t N[8] = SPL_8L_WORDS;
    uint32_t a[9], b[9], ta[9], tb[9];
    s9_set_u8(a, N);
    s9_set_u8(b, h);
#pragma unroll
    for (int i = 0; i < 9; i++) { ta[i] = 0; tb[i] = i == 0; }
    for (int round = 0; round < 12; round++) {
        const int lb = s9_bitlen(b);
        if (lb <= 128) break;
        const int la = s9_bitlen(a);
        if (la - lb >= 28) {
            lattice_big_step(a, ta, b, tb);
            continue;
        }
        const int e = la - 62;
        const uint64_t xa = s9_bits64(a, e), xb = s9_bits64(b, e);
        const uint64_t stop = e >= 128 ? 1ull : (1ull << (128 - e));
        int32_t m[4];
        lehmer_steps(xa, xb, stop, m);
        if (m[0] == 1 && m[1] == 0 && m[2] == 0 && m[3] == 1) {
            lattice_big_step(a, ta, b, tb);   // no 64-bit progress possible
            continue;
        }
        lattice_apply(a, ta, b, tb, m);
    }
    // candidate vectors: (b, tb) when tb is odd, else a balanced a - k b
    uint32_t vx[9], vy[9];
    if (tb[0] & 1u) {
#pragma unroll
        for (int i = 0; i < 9; i++) { vx[i] = b[i]; vy[i] = tb[i]; }
    } else {
        // k = round(<(a, ta), (b, tb)> / <(b, tb), (b, tb)>) (the projection),
        // from 30-bit leading parts: num at scale 2^(ea + eb), den at 2^(2 eb)
        uint32_t mta[9], mtb[9];
        const uint32_t sa = s9_is_neg(ta), sb = s9_is_neg(tb);
        s9_cneg(mta, ta, sa);
        s9_cneg(mtb, tb, sb);
        const int E = s9_bitlen(a) > s9_bitlen(mta) ? s9_bitlen(a) : s9_bitlen(mta);
        const int F = s9_bitlen(b) > s9_bitlen(mtb) ? s9_bitlen(b) : s9_bitlen(mtb);
        const int ea = E > 30 ? E - 30 : 0, eb = F > 30 ? F - 30 : 0;
        const int64_t A0 = (int64_t)(s9_bits64(a, ea) & 0x3fffffffu);
        const int64_t A1 = (int64_t)(s9_bits64(mta, ea) & 0x3fffffffu) * (sa ? -1 : 1);
        const int64_t B0 = (int64_t)(s9_bits64(b, eb) & 0x3fffffffu);
        const int64_t B1 = (int64_t)(s9_bits64(mtb, eb) & 0x3fffffffu) * (sb ? -1 : 1);
        const int64_t num = A0 * B0 + A1 * B1;    // |num| < 2^61
        const int64_t den = B0 * B0 + B1 * B1;    // < 2^61
        const int d = ea - eb;                    // k = num / den * 2^d
        const uint64_t an = (uint64_t)(num < 0 ? -num : num);
        uint64_t q = 0;
        if (den > 0) {
            if (d >= 0) {
                const uint64_t dd = d >= 62 ? 0 : (uint64_t)den >> d;
                q = dd == 0 ? (1ull << 29) : (an + dd / 2) / dd;
            } else {
                q = -d >= 62 ? 0 : ((an >> (-d)) + (uint64_t)den / 2) / (uint64_t)den;
            }
        }
        if (q > (1ull << 29)) q = 1ull << 29;
        const int64_t k = num < 0 ? -(int64_t)q : (int64_t)q;
        s9_lin2(vx, a, b, 1, (int32_t)-k);
        s9_lin2(vy, ta, tb, 1, (int32_t)-k);
    }
    // y > 0
    const uint32_t fy = s9_is_neg(vy);
    s9_cneg(vx, vx, fy);
    s9_cneg(vy, vy, fy);
    x_neg = s9_is_neg(vx);
    s9_cneg(vx, vx, x_neg);
    const int bits = s9_bitlen(vx) > s9_bitlen(vy) ? s9_bitlen(vx) : s9_bitlen(vy);
    if (bits > 253 || (vy[0] & 1u) == 0) {
#pragma unroll
        for (int i = 0; i < 8; i++) { x[i] = h[i]; y[i] = i == 0; }
        x_neg = 0;
        return;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) { x[i] = vx[i]; y[i] = vy[i]; }
}

}  // namespace spl
