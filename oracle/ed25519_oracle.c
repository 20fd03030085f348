/*
 * CPU ORACLE — test infrastructure only (see oracle.h).
 *
 * Restates the reference's Ed25519 path (src/cuda-ecc-ed25519/) in plain C:
 * same decision rules, same scalar-multiplication structure (signed sliding
 * windows with 8-entry odd-multiple tables for verify, signed radix-16 with a
 * 32x8 table for the base point), same keypair/sign composition.  The field
 * representation differs on purpose (radix 2^51, 5 x uint64 with 128-bit
 * products, instead of ref10's radix 2^25.5); the group law is exact, so every
 * result is the same integer/point.  The one reference decision that depends
 * on the raw radix-2^25.5 limbs — the "small order" test of ge.cu:60-82 — is
 * restated as its radix-free rule (decodes and y mod p in {0, 1, p-1}), which
 * the tests pin against the compiled reference.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef uint64_t fe[5];
#define M51 ((1ULL << 51) - 1)

/* ------------------------------------------------------------ field mod p */

static uint64_t ld64(const uint8_t *s) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; i--) v = (v << 8) | s[i];
    return v;
}

static void fe_copy(fe h, const fe f) { memcpy(h, f, sizeof(fe)); }
static void fe_0(fe h) { memset(h, 0, sizeof(fe)); }
static void fe_1(fe h) { fe_0(h); h[0] = 1; }

static void fe_carry(fe h) {
    uint64_t c;
    c = h[0] >> 51; h[0] &= M51; h[1] += c;
    c = h[1] >> 51; h[1] &= M51; h[2] += c;
    c = h[2] >> 51; h[2] &= M51; h[3] += c;
    c = h[3] >> 51; h[3] &= M51; h[4] += c;
    c = h[4] >> 51; h[4] &= M51; h[0] += 19 * c;
}

/* low 255 bits of s, as limbs (the value may be >= p: non-canonical input is
 * kept as-is and only reduced by the arithmetic; fe.cu:297-362 ignores the top
 * bit the same way). */
static void fe_frombytes(fe h, const uint8_t s[32]) {
    h[0] = ld64(s) & M51;
    h[1] = (ld64(s + 6) >> 3) & M51;
    h[2] = (ld64(s + 12) >> 6) & M51;
    h[3] = (ld64(s + 19) >> 1) & M51;
    h[4] = (ld64(s + 24) >> 12) & M51;
}

/* canonical little-endian encoding of h mod p (fe.cu:1386-1490 semantics). */
static void fe_tobytes(uint8_t s[32], const fe f) {
    fe h;
    fe_copy(h, f);
    fe_carry(h);
    fe_carry(h);
    uint64_t q = (h[0] + 19) >> 51;
    q = (h[1] + q) >> 51;
    q = (h[2] + q) >> 51;
    q = (h[3] + q) >> 51;
    q = (h[4] + q) >> 51;
    h[0] += 19 * q;
    uint64_t c;
    c = h[0] >> 51; h[0] &= M51; h[1] += c;
    c = h[1] >> 51; h[1] &= M51; h[2] += c;
    c = h[2] >> 51; h[2] &= M51; h[3] += c;
    c = h[3] >> 51; h[3] &= M51; h[4] += c;
    h[4] &= M51;
    uint64_t w[4];
    w[0] = h[0] | (h[1] << 51);
    w[1] = (h[1] >> 13) | (h[2] << 38);
    w[2] = (h[2] >> 26) | (h[3] << 25);
    w[3] = (h[3] >> 39) | (h[4] << 12);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 8; k++) s[8 * i + k] = (uint8_t)(w[i] >> (8 * k));
}

static void fe_add(fe h, const fe f, const fe g) {
    for (int i = 0; i < 5; i++) h[i] = f[i] + g[i];
    fe_carry(h);
}

/* h = f - g, computed as f + 4p - g so no limb underflows. */
static void fe_sub(fe h, const fe f, const fe g) {
    h[0] = f[0] + 0x1fffffffffffb4ULL - g[0];
    h[1] = f[1] + 0x1ffffffffffffcULL - g[1];
    h[2] = f[2] + 0x1ffffffffffffcULL - g[2];
    h[3] = f[3] + 0x1ffffffffffffcULL - g[3];
    h[4] = f[4] + 0x1ffffffffffffcULL - g[4];
    fe_carry(h);
}

static void fe_neg(fe h, const fe f) {
    fe z;
    fe_0(z);
    fe_sub(h, z, f);
}

static void fe_mul(fe h, const fe f, const fe g) {
    uint64_t g1 = 19 * g[1], g2 = 19 * g[2], g3 = 19 * g[3], g4 = 19 * g[4];
    u128 r0 = (u128)f[0] * g[0] + (u128)f[1] * g4 + (u128)f[2] * g3 +
              (u128)f[3] * g2 + (u128)f[4] * g1;
    u128 r1 = (u128)f[0] * g[1] + (u128)f[1] * g[0] + (u128)f[2] * g4 +
              (u128)f[3] * g3 + (u128)f[4] * g2;
    u128 r2 = (u128)f[0] * g[2] + (u128)f[1] * g[1] + (u128)f[2] * g[0] +
              (u128)f[3] * g4 + (u128)f[4] * g3;
    u128 r3 = (u128)f[0] * g[3] + (u128)f[1] * g[2] + (u128)f[2] * g[1] +
              (u128)f[3] * g[0] + (u128)f[4] * g4;
    u128 r4 = (u128)f[0] * g[4] + (u128)f[1] * g[3] + (u128)f[2] * g[2] +
              (u128)f[3] * g[1] + (u128)f[4] * g[0];
    r1 += (uint64_t)(r0 >> 51);
    r2 += (uint64_t)(r1 >> 51);
    r3 += (uint64_t)(r2 >> 51);
    r4 += (uint64_t)(r3 >> 51);
    uint64_t c = (uint64_t)(r4 >> 51);
    h[0] = ((uint64_t)r0 & M51) + 19 * c;
    h[1] = (uint64_t)r1 & M51;
    h[2] = (uint64_t)r2 & M51;
    h[3] = (uint64_t)r3 & M51;
    h[4] = (uint64_t)r4 & M51;
    fe_carry(h);
}

static void fe_sq(fe h, const fe f) { fe_mul(h, f, f); }

static void fe_sqn(fe h, const fe f, int n) {
    fe_sq(h, f);
    for (int i = 1; i < n; i++) fe_sq(h, h);
}

/* z^(2^250 - 1) and z^11, the shared prefix of both exponent chains
 * (fe.cu:364-455 and :881-975 use the same addition chain shape). */
static void fe_chain_250(fe z250, fe z11, const fe z) {
    fe z2, z9, t, z5, z10, z20, z40, z50, z100, z200;
    fe_sq(z2, z);
    fe_sqn(t, z2, 2);
    fe_mul(z9, t, z);
    fe_mul(z11, z9, z2);
    fe_sq(t, z11);
    fe_mul(z5, t, z9);        /* 2^5 - 1 */
    fe_sqn(t, z5, 5);
    fe_mul(z10, t, z5);       /* 2^10 - 1 */
    fe_sqn(t, z10, 10);
    fe_mul(z20, t, z10);      /* 2^20 - 1 */
    fe_sqn(t, z20, 20);
    fe_mul(z40, t, z20);      /* 2^40 - 1 */
    fe_sqn(t, z40, 10);
    fe_mul(z50, t, z10);      /* 2^50 - 1 */
    fe_sqn(t, z50, 50);
    fe_mul(z100, t, z50);     /* 2^100 - 1 */
    fe_sqn(t, z100, 100);
    fe_mul(z200, t, z100);    /* 2^200 - 1 */
    fe_sqn(t, z200, 50);
    fe_mul(z250, t, z50);     /* 2^250 - 1 */
}

/* z^(p-2) = z^(2^255 - 21) */
static void fe_invert(fe out, const fe z) {
    fe z250, z11, t;
    fe_chain_250(z250, z11, z);
    fe_sqn(t, z250, 5);
    fe_mul(out, t, z11);
}

/* z^((p-5)/8) = z^(2^252 - 3) */
static void fe_pow22523(fe out, const fe z) {
    fe z250, z11, t;
    fe_chain_250(z250, z11, z);
    fe_sqn(t, z250, 2);
    fe_mul(out, t, z);
}

static int fe_isnegative(const fe f) {
    uint8_t s[32];
    fe_tobytes(s, f);
    return s[0] & 1;
}

static int fe_isnonzero(const fe f) {
    uint8_t s[32];
    fe_tobytes(s, f);
    uint8_t r = 0;
    for (int i = 0; i < 32; i++) r |= s[i];
    return r != 0;
}

static void fe_small(fe h, uint64_t v) { fe_0(h); h[0] = v; }

/* ------------------------------------------------------------ group */

typedef struct { fe X, Y, Z; } ge_p2;
typedef struct { fe X, Y, Z, T; } ge_p3;
typedef struct { fe X, Y, Z, T; } ge_p1p1;       /* x = X/Z, y = Y/T */
typedef struct { fe yplusx, yminusx, xy2d; } ge_precomp;
typedef struct { fe YplusX, YminusX, Z, T2d; } ge_cached;

static fe fe_d, fe_d2, fe_sqrtm1;
static ge_precomp Bi[8];            /* (2k+1)B, the table of ge.cu:134-171 */
static ge_precomp base_tab[32][8];  /* (j+1) 256^i B, precomp_data.h:45-46 shape */
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

static void ge_p3_0(ge_p3 *h) { fe_0(h->X); fe_1(h->Y); fe_1(h->Z); fe_0(h->T); }

static void p1p1_to_p2(ge_p2 *r, const ge_p1p1 *p) {
    fe_mul(r->X, p->X, p->T);
    fe_mul(r->Y, p->Y, p->Z);
    fe_mul(r->Z, p->Z, p->T);
}

static void p1p1_to_p3(ge_p3 *r, const ge_p1p1 *p) {
    fe_mul(r->X, p->X, p->T);
    fe_mul(r->Y, p->Y, p->Z);
    fe_mul(r->Z, p->Z, p->T);
    fe_mul(r->T, p->X, p->Y);
}

static void p3_to_p2(ge_p2 *r, const ge_p3 *p) {
    fe_copy(r->X, p->X); fe_copy(r->Y, p->Y); fe_copy(r->Z, p->Z);
}

static void p3_to_cached(ge_cached *r, const ge_p3 *p) {
    fe_add(r->YplusX, p->Y, p->X);
    fe_sub(r->YminusX, p->Y, p->X);
    fe_copy(r->Z, p->Z);
    fe_mul(r->T2d, p->T, fe_d2);
}

/* doubling of a projective point (a = -1 twisted Edwards, HWCD 2008). */
static void p2_dbl(ge_p1p1 *r, const ge_p2 *p) {
    fe xx, yy, b, a, aa;
    fe_sq(xx, p->X);
    fe_sq(yy, p->Y);
    fe_sq(b, p->Z);
    fe_add(b, b, b);
    fe_add(a, p->X, p->Y);
    fe_sq(aa, a);
    fe_add(r->Y, yy, xx);
    fe_sub(r->Z, yy, xx);
    fe_sub(r->X, aa, r->Y);
    fe_sub(r->T, b, r->Z);
}

static void p3_dbl(ge_p1p1 *r, const ge_p3 *p) {
    ge_p2 q;
    p3_to_p2(&q, p);
    p2_dbl(r, &q);
}

/* r = p + q (sign = +1) or p - q (sign = -1) with q in cached form. */
static void ge_addsub(ge_p1p1 *r, const ge_p3 *p, const ge_cached *q, int sign) {
    fe a, b, c, dd;
    fe_add(a, p->Y, p->X);
    fe_sub(b, p->Y, p->X);
    fe_mul(a, a, sign > 0 ? q->YplusX : q->YminusX);
    fe_mul(b, b, sign > 0 ? q->YminusX : q->YplusX);
    fe_mul(c, q->T2d, p->T);
    fe_mul(dd, p->Z, q->Z);
    fe_add(dd, dd, dd);
    fe_sub(r->X, a, b);
    fe_add(r->Y, a, b);
    if (sign > 0) { fe_add(r->Z, dd, c); fe_sub(r->T, dd, c); }
    else          { fe_sub(r->Z, dd, c); fe_add(r->T, dd, c); }
}

/* r = p + q (sign = +1) or p - q (sign = -1) with q in affine precomp form. */
static void ge_maddsub(ge_p1p1 *r, const ge_p3 *p, const ge_precomp *q, int sign) {
    fe a, b, c, dd;
    fe_add(a, p->Y, p->X);
    fe_sub(b, p->Y, p->X);
    fe_mul(a, a, sign > 0 ? q->yplusx : q->yminusx);
    fe_mul(b, b, sign > 0 ? q->yminusx : q->yplusx);
    fe_mul(c, q->xy2d, p->T);
    fe_add(dd, p->Z, p->Z);
    fe_sub(r->X, a, b);
    fe_add(r->Y, a, b);
    if (sign > 0) { fe_add(r->Z, dd, c); fe_sub(r->T, dd, c); }
    else          { fe_sub(r->Z, dd, c); fe_add(r->T, dd, c); }
}

static void ge_p3_add(ge_p3 *r, const ge_p3 *p, const ge_p3 *q) {
    ge_cached c;
    ge_p1p1 t;
    p3_to_cached(&c, q);
    ge_addsub(&t, p, &c, 1);
    p1p1_to_p3(r, &t);
}

static void p3_to_precomp(ge_precomp *r, const ge_p3 *p) {
    fe zi, x, y;
    fe_invert(zi, p->Z);
    fe_mul(x, p->X, zi);
    fe_mul(y, p->Y, zi);
    fe_add(r->yplusx, y, x);
    fe_sub(r->yminusx, y, x);
    fe_mul(r->xy2d, x, y);
    fe_mul(r->xy2d, r->xy2d, fe_d2);
}

static void ge_p2_tobytes(uint8_t s[32], const fe X, const fe Y, const fe Z) {
    fe zi, x, y;
    fe_invert(zi, Z);
    fe_mul(x, X, zi);
    fe_mul(y, Y, zi);
    fe_tobytes(s, y);
    s[31] ^= (uint8_t)(fe_isnegative(x) << 7);
}

/* ref10 decompression that returns the NEGATED point (ge.cu:182-222):
 * y = low 255 bits of s (kept even if >= p); x from
 * x = u v^3 (u v^7)^((p-5)/8), u = y^2 - 1, v = d y^2 + 1; if v x^2 != u try
 * x * sqrt(-1) when v x^2 == -u, else fail; then x := -x when
 * isnegative(x) == bit 255 (so x = 0 with bit 255 set is accepted). */
static int ge_frombytes_negate(ge_p3 *h, const uint8_t s[32]) {
    fe u, v, v3, vxx, chk;
    fe_frombytes(h->Y, s);
    fe_1(h->Z);
    fe_sq(u, h->Y);
    fe_mul(v, u, fe_d);
    fe_sub(u, u, h->Z);
    fe_add(v, v, h->Z);
    fe_sq(v3, v);
    fe_mul(v3, v3, v);
    fe_sq(h->X, v3);
    fe_mul(h->X, h->X, v);
    fe_mul(h->X, h->X, u);
    fe_pow22523(h->X, h->X);
    fe_mul(h->X, h->X, v3);
    fe_mul(h->X, h->X, u);
    fe_sq(vxx, h->X);
    fe_mul(vxx, vxx, v);
    fe_sub(chk, vxx, u);
    if (fe_isnonzero(chk)) {
        fe_add(chk, vxx, u);
        if (fe_isnonzero(chk)) return -1;
        fe_mul(h->X, h->X, fe_sqrtm1);
    }
    if (fe_isnegative(h->X) == (s[31] >> 7)) fe_neg(h->X, h->X);
    fe_mul(h->T, h->X, h->Y);
    return 0;
}

static void oracle_init(void) {
    fe a, b;
    /* d = -121665 / 121666 */
    fe_small(a, 121666);
    fe_invert(b, a);
    fe_small(a, 121665);
    fe_mul(b, b, a);
    fe_neg(fe_d, b);
    fe_add(fe_d2, fe_d, fe_d);
    /* sqrt(-1) = 2^((p-1)/4) = 2^(2^253 - 5): 2^(2^252-3) squared, times 2 */
    fe_small(a, 2);
    fe_pow22523(b, a);     /* 2^(2^252 - 3) */
    fe_sq(b, b);           /* 2^(2^253 - 6) */
    fe_mul(fe_sqrtm1, b, a);
    /* base point B: y = 4/5, x even; ge_frombytes_negate returns -B */
    uint8_t enc[32];
    memset(enc, 0x66, 32);
    enc[0] = 0x58;
    ge_p3 B, nB;
    ge_frombytes_negate(&nB, enc);
    fe_neg(B.X, nB.X); fe_copy(B.Y, nB.Y); fe_copy(B.Z, nB.Z); fe_neg(B.T, nB.T);
    /* Bi[k] = (2k+1) B */
    ge_p3 B2, acc;
    ge_p1p1 t;
    p3_dbl(&t, &B);
    p1p1_to_p3(&B2, &t);
    acc = B;
    for (int k = 0; k < 8; k++) {
        p3_to_precomp(&Bi[k], &acc);
        ge_p3_add(&acc, &acc, &B2);
    }
    /* base_tab[i][j] = (j+1) 256^i B */
    ge_p3 row = B;
    for (int i = 0; i < 32; i++) {
        ge_p3 m = row;
        for (int j = 0; j < 8; j++) {
            p3_to_precomp(&base_tab[i][j], &m);
            ge_p3_add(&m, &m, &row);
        }
        for (int k = 0; k < 8; k++) { p3_dbl(&t, &row); p1p1_to_p3(&row, &t); }
    }
}

static void ensure_init(void) { pthread_once(&g_once, oracle_init); }

/* signed sliding window recoding of a 256-bit scalar (ge.cu:25-58): digits are
 * odd, |d| <= 15, and each nonzero digit is followed by at least ... the same
 * greedy merge over a 6-bit look-ahead as the reference. */
static void slide(signed char r[256], const uint8_t a[32]) {
    for (int i = 0; i < 256; i++) r[i] = 1 & (a[i >> 3] >> (i & 7));
    for (int i = 0; i < 256; i++) {
        if (!r[i]) continue;
        for (int b = 1; b <= 6 && i + b < 256; b++) {
            if (!r[i + b]) continue;
            int hi = r[i + b] << b;
            if (r[i] + hi <= 15) {
                r[i] = (signed char)(r[i] + hi);
                r[i + b] = 0;
            } else if (r[i] - hi >= -15) {
                r[i] = (signed char)(r[i] - hi);
                for (int k = i + b; k < 256; k++) {
                    if (!r[k]) { r[k] = 1; break; }
                    r[k] = 0;
                }
            } else {
                break;
            }
        }
    }
}

/* r = a*A' + b*B where Ai[k] = (2k+1)A' (ge.cu:134-171). */
static void double_scalarmult(ge_p2 *r, const uint8_t a[32], const ge_cached Ai[8],
                              const uint8_t b[32]) {
    signed char as[256], bs[256];
    slide(as, a);
    slide(bs, b);
    fe_0(r->X); fe_1(r->Y); fe_1(r->Z);
    int i = 255;
    while (i >= 0 && !as[i] && !bs[i]) i--;
    ge_p1p1 t;
    ge_p3 u;
    for (; i >= 0; i--) {
        p2_dbl(&t, r);
        if (as[i]) {
            p1p1_to_p3(&u, &t);
            ge_addsub(&t, &u, &Ai[(as[i] > 0 ? as[i] : -as[i]) / 2], as[i] > 0 ? 1 : -1);
        }
        if (bs[i]) {
            p1p1_to_p3(&u, &t);
            ge_maddsub(&t, &u, &Bi[(bs[i] > 0 ? bs[i] : -bs[i]) / 2], bs[i] > 0 ? 1 : -1);
        }
        p1p1_to_p2(r, &t);
    }
}

/* h = a*B for a < 2^255, signed radix-16 digits (ge.cu:452-500). */
static void scalarmult_base(ge_p3 *h, const uint8_t a[32]) {
    signed char e[64];
    for (int i = 0; i < 32; i++) {
        e[2 * i] = a[i] & 15;
        e[2 * i + 1] = (a[i] >> 4) & 15;
    }
    signed char carry = 0;
    for (int i = 0; i < 63; i++) {
        e[i] = (signed char)(e[i] + carry);
        carry = (signed char)((e[i] + 8) >> 4);
        e[i] = (signed char)(e[i] - (carry << 4));
    }
    e[63] = (signed char)(e[63] + carry);
    ge_p3_0(h);
    ge_p1p1 t;
    for (int pass = 1; pass >= 0; pass--) {
        for (int i = pass; i < 64; i += 2) {
            int d = e[i];
            if (!d) continue;          /* adding the identity changes nothing */
            ge_maddsub(&t, h, &base_tab[i / 2][(d > 0 ? d : -d) - 1], d > 0 ? 1 : -1);
            p1p1_to_p3(h, &t);
        }
        if (pass == 1)
            for (int k = 0; k < 4; k++) { p3_dbl(&t, h); p1p1_to_p3(h, &t); }
    }
}

static void ge_p3_tobytes(uint8_t s[32], const ge_p3 *h) {
    ge_p2_tobytes(s, h->X, h->Y, h->Z);
}

/* ------------------------------------------------------------ scalars mod L */

static const uint64_t Lw[4] = {0x5812631a5cf5d3edULL, 0x14def9dea2f79cd6ULL, 0,
                               0x1000000000000000ULL};

/* x (nbytes little-endian) mod L, one bit at a time. */
static void mod_l(uint8_t out[32], const uint8_t *x, int nbytes) {
    uint64_t r[4] = {0, 0, 0, 0};
    for (int bit = nbytes * 8 - 1; bit >= 0; bit--) {
        uint64_t in = (x[bit >> 3] >> (bit & 7)) & 1;
        r[3] = (r[3] << 1) | (r[2] >> 63);
        r[2] = (r[2] << 1) | (r[1] >> 63);
        r[1] = (r[1] << 1) | (r[0] >> 63);
        r[0] = (r[0] << 1) | in;
        int ge = 1;
        for (int k = 3; k >= 0; k--) {
            if (r[k] != Lw[k]) { ge = r[k] > Lw[k]; break; }
        }
        if (ge) {
            u128 bor = 0;
            for (int k = 0; k < 4; k++) {
                u128 d = (u128)r[k] - Lw[k] - bor;
                r[k] = (uint64_t)d;
                bor = (d >> 64) ? 1 : 0;
            }
        }
    }
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 8; k++) out[8 * i + k] = (uint8_t)(r[i] >> (8 * k));
}

void oracle_sc_reduce(uint8_t s[64]) {
    uint8_t r[32];
    mod_l(r, s, 64);
    memcpy(s, r, 32);
}

/* s = (a*b + c) mod L (sc.cu:552-1012). */
void oracle_sc_muladd(uint8_t s[32], const uint8_t a[32], const uint8_t b[32],
                      const uint8_t c[32]) {
    uint64_t aw[4], bw[4], prod[9] = {0};
    for (int i = 0; i < 4; i++) { aw[i] = ld64(a + 8 * i); bw[i] = ld64(b + 8 * i); }
    for (int i = 0; i < 4; i++) {
        u128 carry = 0;
        for (int j = 0; j < 4; j++) {
            u128 t = (u128)aw[i] * bw[j] + prod[i + j] + carry;
            prod[i + j] = (uint64_t)t;
            carry = t >> 64;
        }
        prod[i + 4] += (uint64_t)carry;
    }
    u128 carry = 0;
    for (int i = 0; i < 9; i++) {
        u128 t = (u128)prod[i] + (i < 4 ? ld64(c + 8 * i) : 0) + carry;
        prod[i] = (uint64_t)t;
        carry = t >> 64;
    }
    uint8_t buf[72];
    for (int i = 0; i < 9; i++)
        for (int k = 0; k < 8; k++) buf[8 * i + k] = (uint8_t)(prod[i] >> (8 * k));
    mod_l(s, buf, 72);
}

/* -------------------------------------------------- the reference's quirks */

/* Restatement of scalar32_reduce (sc.cu:12-224): a Montgomery multiply of
 * the 5 x 52-bit unpacked input by R = 2^260 mod L followed by Montgomery
 * reduction and one conditional subtraction of L.  Two reference behaviours
 * are kept on purpose:
 *   - column 6 of the schoolbook product is out[5] + a4*b2 instead of
 *     a2*b4 + a3*b3 + a4*b2 (sc.cu:58-59);
 *   - limb 4 is read as 52 bits starting at bit 208, i.e. it takes the low
 *     nibble of the byte after the 32-byte buffer (sc.cu:79); that byte is
 *     taken to be 0 here.
 * Result packed little-endian modulo 2^256 (sc.cu:85-130). */
static void scalar32_reduce_quirk(uint8_t s[32]) {
    static const uint64_t R52[5] = {0x000f48bd6721e6edULL, 0x0003bab5ac67e45aULL,
                                    0x000fffffeb35e51bULL, 0x000fffffffffffffULL,
                                    0x00000fffffffffffULL};
    static const uint64_t L52[5] = {0x0002631a5cf5d3edULL, 0x000dea2f79cd6581ULL,
                                    0x000000000014def9ULL, 0, 0x0000100000000000ULL};
    const uint64_t LF = 0x51da312547e1bULL;
    const uint64_t M52 = (1ULL << 52) - 1;
    uint8_t ext[40];
    memset(ext, 0, sizeof ext);
    memcpy(ext, s, 32);
    uint64_t a[5];
    a[0] = ld64(ext) & M52;
    a[1] = (ld64(ext + 6) >> 4) & M52;
    a[2] = ld64(ext + 13) & M52;
    a[3] = (ld64(ext + 19) >> 4) & M52;
    a[4] = ld64(ext + 26) & M52;
    u128 z[9];
    for (int k = 0; k < 9; k++) {
        z[k] = 0;
        for (int i = 0; i < 5; i++) {
            int j = k - i;
            if (j >= 0 && j < 5) z[k] += (u128)a[i] * R52[j];
        }
    }
    z[6] = z[5] + (u128)a[4] * R52[2];
    uint64_t n[5];
    u128 c;
    u128 sum = z[0];
    n[0] = ((uint64_t)sum * LF) & M52;
    c = (sum + (u128)n[0] * L52[0]) >> 52;
    sum = (u128)n[0] * L52[1] + z[1] + c;
    n[1] = ((uint64_t)sum * LF) & M52;
    c = (sum + (u128)n[1] * L52[0]) >> 52;
    sum = (u128)n[0] * L52[2] + (u128)n[1] * L52[1] + z[2] + c;
    n[2] = ((uint64_t)sum * LF) & M52;
    c = (sum + (u128)n[2] * L52[0]) >> 52;
    sum = (u128)n[1] * L52[2] + (u128)n[2] * L52[1] + z[3] + c;
    n[3] = ((uint64_t)sum * LF) & M52;
    c = (sum + (u128)n[3] * L52[0]) >> 52;
    sum = (u128)n[0] * L52[4] + (u128)n[3] * L52[1] + (u128)n[2] * L52[2] + z[4] + c;
    n[4] = ((uint64_t)sum * LF) & M52;
    c = (sum + (u128)n[4] * L52[0]) >> 52;
    uint64_t r[5];
    sum = c + z[5] + (u128)n[1] * L52[4] + (u128)n[3] * L52[2] + (u128)n[4] * L52[1];
    r[0] = (uint64_t)sum & M52; c = sum >> 52;
    sum = c + z[6] + (u128)n[2] * L52[4] + (u128)n[4] * L52[2];
    r[1] = (uint64_t)sum & M52; c = sum >> 52;
    sum = c + z[7] + (u128)n[3] * L52[4];
    r[2] = (uint64_t)sum & M52; c = sum >> 52;
    sum = c + z[8] + (u128)n[4] * L52[4];
    r[3] = (uint64_t)sum & M52; c = sum >> 52;
    r[4] = (uint64_t)c;
    /* r - L, adding L back when the difference is negative */
    uint64_t d[5], borrow = 0;
    for (int i = 0; i < 5; i++) {
        borrow = r[i] - (L52[i] + (borrow >> 63));
        d[i] = borrow & M52;
    }
    uint64_t mask = ((borrow >> 63) ^ 1) - 1;
    uint64_t cc = 0;
    for (int i = 0; i < 5; i++) {
        cc = (cc >> 52) + d[i] + (L52[i] & mask);
        d[i] = cc & M52;
    }
    uint64_t w[4];
    w[0] = d[0] | (d[1] << 52);
    w[1] = (d[1] >> 12) | (d[2] << 40);
    w[2] = (d[2] >> 24) | (d[3] << 28);
    w[3] = (d[3] >> 36) | (d[4] << 16);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 8; k++) s[8 * i + k] = (uint8_t)(w[i] >> (8 * k));
}

/* ed25519_get_checked_scalar (verify.cu:60-85): 0 = ok. */
int oracle_get_checked_scalar(uint8_t out[32], const uint8_t in[32]) {
    if ((in[31] & 0xf0) == 0) {
        memcpy(out, in, 32);
        return 0;
    }
    if (in[31] >> 7) return 1;
    scalar32_reduce_quirk(out);
    return memcmp(out, in, 32) == 0 ? 0 : 1;
}

/* y mod p in {0, 1, p-1}: the decision made by ge_is_small_order's raw-limb
 * identity test after three doublings (ge.cu:60-82, fe.cu:46-67). */
static int y_is_low_order(const uint8_t enc[32]) {
    fe y;
    uint8_t c[32];
    fe_frombytes(y, enc);
    fe_tobytes(c, y);
    int zero = 1, one = c[0] == 1, pm1 = c[0] == 0xec && c[31] == 0x7f;
    for (int i = 0; i < 32; i++) zero &= c[i] == 0;
    for (int i = 1; i < 32; i++) one &= c[i] == 0;
    for (int i = 1; i < 31; i++) pm1 &= c[i] == 0xff;
    return zero || one || pm1;
}

/* check_packed_ge_small_order (verify.cu:88-107): 0 = decodes and not small. */
int oracle_check_packed_ge_small_order(const uint8_t enc[32]) {
    ensure_init();
    ge_p3 p;
    if (ge_frombytes_negate(&p, enc) != 0) return 1;
    return y_is_low_order(enc) ? 1 : 0;
}

/* ------------------------------------------------------------ keys, sign */

void oracle_create_keypair(uint8_t pk[32], uint8_t sk[64], const uint8_t seed[32]) {
    ensure_init();
    oracle_sha512(sk, seed, 32);
    sk[0] &= 248;
    sk[31] &= 63;
    sk[31] |= 64;
    ge_p3 A;
    scalarmult_base(&A, sk);
    ge_p3_tobytes(pk, &A);
}

static void sha512_3(uint8_t out[64], const uint8_t *a, size_t la, const uint8_t *b,
                     size_t lb, const uint8_t *c, size_t lc) {
    uint8_t stackbuf[2048];
    size_t n = la + lb + lc;
    uint8_t *buf = n <= sizeof stackbuf ? stackbuf : (uint8_t *)malloc(n);
    if (la) memcpy(buf, a, la);
    if (lb) memcpy(buf + la, b, lb);
    if (lc) memcpy(buf + la + lb, c, lc);
    oracle_sha512(out, buf, n);
    if (buf != stackbuf) free(buf);
}

/* RFC 8032 sign with the 64-byte expanded private key (sign.cu:9-38). */
void oracle_sign(uint8_t sig[64], const uint8_t *msg, size_t len,
                 const uint8_t pk[32], const uint8_t sk[64]) {
    ensure_init();
    uint8_t r[64], hram[64];
    sha512_3(r, sk + 32, 32, msg, len, NULL, 0);
    oracle_sc_reduce(r);
    ge_p3 R;
    scalarmult_base(&R, r);
    ge_p3_tobytes(sig, &R);
    sha512_3(hram, sig, 32, pk, 32, msg, len);
    oracle_sc_reduce(hram);
    oracle_sc_muladd(sig + 32, hram, sk, r);
}

/* ------------------------------------------------------------ verify */

/* ed25519_verify (verify.cu:109-174): S top nibble clear, R decodes and is not
 * small, A decodes and is not small, encode(h(-A) + S B) == R bytes. */
int oracle_verify(const uint8_t sig[64], const uint8_t *msg, uint32_t len,
                  const uint8_t pk[32]) {
    ensure_init();
    if (sig[63] & 0xf0) return 0;
    if (oracle_check_packed_ge_small_order(sig) != 0) return 0;
    uint8_t h[64];
    sha512_3(h, sig, 32, pk, 32, msg, len);
    oracle_sc_reduce(h);
    ge_p3 A, A2, u;
    if (ge_frombytes_negate(&A, pk) != 0) return 0;
    if (y_is_low_order(pk)) return 0;
    ge_cached Ai[8];
    ge_p1p1 t;
    p3_to_cached(&Ai[0], &A);
    p3_dbl(&t, &A);
    p1p1_to_p3(&A2, &t);
    for (int k = 1; k < 8; k++) {
        ge_addsub(&t, &A2, &Ai[k - 1], 1);
        p1p1_to_p3(&u, &t);
        p3_to_cached(&Ai[k], &u);
    }
    ge_p2 R;
    double_scalarmult(&R, h, Ai, sig + 32);
    uint8_t chk[32];
    ge_p2_tobytes(chk, R.X, R.Y, R.Z);
    return memcmp(chk, sig, 32) == 0 ? 1 : 0;
}

void oracle_verify_many(const uint8_t *packets, uint32_t total_signatures,
                        const uint32_t *message_lens,
                        const uint32_t *public_key_offsets,
                        const uint32_t *signature_offsets,
                        const uint32_t *message_start_offsets, uint8_t *out) {
    for (uint32_t i = 0; i < total_signatures; i++)
        out[i] = (uint8_t)oracle_verify(packets + signature_offsets[i],
                                        packets + message_start_offsets[i],
                                        message_lens[i],
                                        packets + public_key_offsets[i]);
}

void oracle_sign_many(const uint8_t *packets, uint32_t total_signatures,
                      const uint32_t *message_lens,
                      const uint32_t *public_key_offsets,
                      const uint32_t *private_key_offsets,
                      const uint32_t *message_start_offsets, uint8_t *sigs_out) {
    for (uint32_t i = 0; i < total_signatures; i++)
        oracle_sign(sigs_out + 64 * (size_t)i, packets + message_start_offsets[i],
                    message_lens[i], packets + public_key_offsets[i],
                    packets + private_key_offsets[i]);
}

typedef struct {
    const uint8_t *packets;
    const uint32_t *ml, *pk, *sg, *ms;
    uint8_t *out;
    uint32_t begin, end;
} verify_job;

static void *verify_worker(void *arg) {
    verify_job *j = (verify_job *)arg;
    uint32_t b = j->begin;
    oracle_verify_many(j->packets, j->end - b, j->ml + b, j->pk + b, j->sg + b,
                       j->ms + b, j->out + b);
    return NULL;
}

void oracle_verify_many_mt(const uint8_t *packets, uint32_t total_signatures,
                           const uint32_t *message_lens,
                           const uint32_t *public_key_offsets,
                           const uint32_t *signature_offsets,
                           const uint32_t *message_start_offsets, uint8_t *out,
                           int threads) {
    ensure_init();
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    verify_job jobs[256];
    for (int t = 0; t < threads; t++) {
        verify_job *j = &jobs[t];
        j->packets = packets;
        j->ml = message_lens; j->pk = public_key_offsets;
        j->sg = signature_offsets; j->ms = message_start_offsets;
        j->out = out;
        j->begin = (uint32_t)((uint64_t)total_signatures * t / threads);
        j->end = (uint32_t)((uint64_t)total_signatures * (t + 1) / threads);
        pthread_create(&tid[t], NULL, verify_worker, j);
    }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
}
