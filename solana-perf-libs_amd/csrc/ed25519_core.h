// Ed25519 verify / sign / keypair cores shared by the gfx950 kernels and the
// exported host helpers.
//
// Acceptance rule of the reference's verify path
// (src/cuda-ecc-ed25519/verify.cu:109-174, ge.cu:60-222):
//   1. S = sig[32..64): reject when its top nibble is set (S >= 2^252);
//   2. R = sig[0..32): must decode and not be "small order";
//   3. A = public key: must decode and not be "small order";
//   4. h = SHA-512(R || A || M) mod L;
//   5. accept iff encode([h](-A) + [S]B) == R (byte comparison).
// "Small order" is the reference's raw-limb identity test after three
// doublings; it fires exactly when the encoding decodes and y mod p is 0, 1 or
// p-1 (y_low_order).  Because encode() only produces canonical encodings of
// curve points, step 5 implies that R decodes, and for a canonical R the
// small-order test reduces to y_low_order(R).  So step 2 is evaluated as
// "not y_low_order(R)" without decompressing R: same mask, one square root
// (about 260 field operations) less per signature.
//
// Scalar multiplication uses windows at the same bit positions in every lane
// of a wave: [h](-A) in signed radix-2^W digits (W = 4 or 5; an add every W
// bits from a (2^(W-1)+1)-entry table 0..2^(W-1) x (-A)), and [S]B as a comb
// with no doublings: S's 32 signed radix-256 digits each select from a
// precomputed table d 256^m B (d = 0..128), accumulated in a second point that
// is independent of the -A chain (instruction-level parallelism).  A zero digit
// adds the identity.  Sliding windows, as in the reference (ge.cu:25-58,
// 134-171), put the adds at data-dependent positions; across 64 lanes nearly
// every bit position then needs an add in some lane and the whole wave pays
// for it.  The group law is exact, so the result is the same point either way.
#pragma once
#include "ge25519.h"
#include "lattice.h"
#include "sc25519.h"
#include "sha2.h"

namespace spl {

struct ed25519_tables {
    curve_consts k;
    ge_precomp base[32][8];     // (j+1) 256^i B: signing (constant-time scans)
    ge_precomp comb[32][129];   // d 256^m B, d = 0..128 (entry 0 = identity): verify
};

// Signed digits.  Digit j of a scalar in radix 2^W is e_j = n_j - 2^W t_j +
// t_{j-1}, where n_j is the unsigned digit and t_j its top bit:
// sum e_j 2^(W j) = sum n_j 2^(W j) (the t terms telescope), |e_j| <= 2^(W-1),
// exact as long as the top digit's t is 0.  Each digit needs only its own bits
// and the bit just below them, so digits are produced on the fly.

// Consumes digits from the top: the scalar is held left-aligned in 8 words;
// each call returns the top digit and shifts the scalar left by W bits.
template <int W>
SPL_HD int take_top_digit(uint32_t a[8]) {
    const uint32_t n = a[7] >> (32 - W);
    const uint32_t below = (a[7] >> (31 - W)) & 1u;
#pragma unroll
    for (int i = 7; i > 0; i--) a[i] = (a[i] << W) | (a[i - 1] >> (32 - W));
    a[0] <<= W;
    return (int)n - (int)((n >> (W - 1)) << W) + (int)below;
}

// Consumes radix-256 digits from the bottom; `below` carries the previous
// byte's top bit between calls (start with 0).
SPL_HD int take_low_byte_digit(uint32_t a[8], uint32_t &below) {
    const uint32_t n = a[0] & 0xffu;
    const int d = (int)n - (int)((n >> 7) << 8) + (int)below;
    below = n >> 7;
#pragma unroll
    for (int i = 0; i < 7; i++) a[i] = (a[i] >> 8) | (a[i + 1] << 24);
    a[7] >>= 8;
    return d;
}

// Variable-base window geometry for [h](-A) with h < 2^253.
template <int W>
struct VarWindow {
    static constexpr int kWindows = (253 + W - 1) / W;   // digits 0..kWindows-1
    static constexpr int kTable = (1 << (W - 1)) + 1;    // entries 0..2^(W-1)
    static constexpr int kAlign = 32 * 8 - W * kWindows; // left shift to top-align
    static_assert(W * kWindows - 1 >= 253, "top digit must carry nothing out");
};

// Keeps the device's instruction scheduler from moving anything across this
// point (a no-op on the host).
SPL_HD void sched_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_sched_barrier(0);
#endif
}

// ------------------------------------------------------------ message reader

// Little-endian 32-bit word m of "M || 0x80 || 0 ...": the message bytes
// 4m..4m+3 followed by the SHA-2 padding byte at position len.
struct MsgReader {
    const uint8_t *p;   // first message byte
    uint32_t len;

    SPL_HD uint32_t word(uint32_t m) const {
        uint32_t j = 4u * m;
        if (j > len) return 0u;
        if (j == len) return 0x80u;
        uint32_t w;
#if defined(__HIP_DEVICE_COMPILE__)
        uintptr_t addr = (uintptr_t)(p + j);
        const uint32_t *a = (const uint32_t *)(addr & ~(uintptr_t)3);
        uint32_t sh = (uint32_t)(addr & 3);
        uint32_t lo = a[0];
        uint32_t hi = (sh != 0 && j + 4u - sh < len) ? a[1] : 0u;
        w = __builtin_amdgcn_alignbyte(hi, lo, sh);
#else
        w = 0;
        for (uint32_t b = 0; b < 4 && j + b < len; b++) w |= (uint32_t)p[j + b] << (8 * b);
#endif
        uint32_t rem = len - j;
        if (rem < 4u) {
            uint32_t keep = (1u << (8u * rem)) - 1u;
            w = (w & keep) | (0x80u << (8u * rem));
        }
        return w;
    }
};

// n little-endian words from an arbitrary byte address (inside one buffer).
SPL_HD void load_words(uint32_t *out, const uint8_t *src, int n) {
#if defined(__HIP_DEVICE_COMPILE__)
    uintptr_t addr = (uintptr_t)src;
    const uint32_t *a = (const uint32_t *)(addr & ~(uintptr_t)3);
    uint32_t sh = (uint32_t)(addr & 3);
    uint32_t prev = a[0];
    for (int i = 0; i < n; i++) {
        // the dword after the field is read only when the field straddles it
        uint32_t next = (i + 1 < n || sh != 0) ? a[i + 1] : 0u;
        out[i] = __builtin_amdgcn_alignbyte(next, prev, sh);
        prev = next;
    }
#else
    for (int i = 0; i < n; i++)
        out[i] = (uint32_t)src[4 * i] | ((uint32_t)src[4 * i + 1] << 8) |
                 ((uint32_t)src[4 * i + 2] << 16) | ((uint32_t)src[4 * i + 3] << 24);
#endif
}

#if defined(__HIP_DEVICE_COMPILE__)
// Dwords q0 .. q0+32 of the message's dword array (the message read from its
// address rounded down to 4), each index clamped into [0, last dword holding
// a message byte] so every load stays inside the message.  The 33 loads are
// independent and unconditional (a lane with an empty message loads
// nothing), so they are all in flight together.
__device__ __forceinline__ void msg_window(uint32_t d[33], const MsgReader &msg, int q0) {
    if (msg.len == 0) {
#pragma unroll
        for (int k = 0; k < 33; k++) d[k] = 0u;
        return;
    }
    const uintptr_t addr = (uintptr_t)msg.p;
    const uint32_t *a = (const uint32_t *)(addr & ~(uintptr_t)3);
    const int qmax = (int)(((uint64_t)(addr & 3) + msg.len - 1u) >> 2);
#pragma unroll
    for (int k = 0; k < 33; k++) {
        int q = q0 + k;
        q = q < 0 ? 0 : (q > qmax ? qmax : q);
        d[k] = a[q];
    }
}

// Word m of "M || 0x80 || 0 ..." from the window starting at dword q0
// (0 <= m - q0 < 32): the bytes of the message, the padding byte at position
// len, zeros after; the same word MsgReader::word returns.
__device__ __forceinline__ uint32_t msg_window_word(const uint32_t d[33], int i, int m,
                                                    uint32_t sh, uint32_t len) {
    const uint32_t w = __builtin_amdgcn_alignbyte(d[i + 1], d[i], sh);
    const int64_t rem = (int64_t)len - 4 * (int64_t)m;
    const uint32_t keep = rem >= 4 ? ~0u : (rem > 0 ? (1u << (8 * (uint32_t)rem)) - 1u : 0u);
    const uint32_t term = (rem >= 0 && rem < 4) ? 0x80u << (8 * (uint32_t)rem) : 0u;
    return (w & keep) | term;
}
#endif

// SHA-512 over prefix (npw little-endian words, npw even) || M: blocks
// [b0, b1) of the padded stream (b1 clamped to its block count), continuing
// from st (which holds the initial state when b0 = 0).  Splitting a hash at a
// block boundary between two calls gives the same state as one call.
SPL_HD void sha512_prefixed_blocks(uint64_t st[8], const uint32_t *prefix, uint32_t npw,
                                   const MsgReader &msg, uint32_t b0, uint32_t b1) {
    uint64_t total = 4ull * npw + msg.len;
    uint32_t nblocks = (uint32_t)((total + 17 + 127) / 128);
    if (b1 > nblocks) b1 = nblocks;
#if defined(__HIP_DEVICE_COMPILE__)
    // Block b holds message words 32 b - npw .. + 31.  Each block's message
    // dwords are loaded one block ahead, while the previous block compresses,
    // instead of one dependent load (and wait) per word.
    const uint32_t sh = (uint32_t)((uintptr_t)msg.p & 3);
    uint32_t win[33];
    msg_window(win, msg, 32 * (int)b0 - (int)npw);
    for (uint32_t b = b0; b < b1; b++) {
        const int q0 = 32 * (int)b - (int)npw;
        uint32_t nxt[33];
        msg_window(nxt, msg, q0 + 32);
        sched_fence();
        uint64_t w[16];
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const uint32_t k = 32u * b + 2u * t;   // stream word index
            const uint32_t s0 = k < npw ? prefix[k]
                                        : msg_window_word(win, 2 * t, q0 + 2 * t, sh, msg.len);
            const uint32_t s1 = k + 1 < npw
                                    ? prefix[k + 1]
                                    : msg_window_word(win, 2 * t + 1, q0 + 2 * t + 1, sh, msg.len);
            w[t] = ((uint64_t)bswap32(s0) << 32) | bswap32(s1);
        }
        if (b == nblocks - 1) {
            w[14] = 0;
            w[15] = total * 8;
        }
        sha512_block(st, w);
#pragma unroll
        for (int k = 0; k < 33; k++) win[k] = nxt[k];
    }
    return;
#endif
    for (uint32_t b = b0; b < b1; b++) {
        uint64_t w[16];
        for (int t = 0; t < 16; t++) {
            uint32_t k = 32u * b + 2u * t;   // stream word index
            uint32_t s0 = k < npw ? prefix[k] : msg.word(k - npw);
            uint32_t s1 = k + 1 < npw ? prefix[k + 1] : msg.word(k + 1 - npw);
            w[t] = ((uint64_t)bswap32(s0) << 32) | bswap32(s1);
        }
        if (b == nblocks - 1) {
            w[14] = 0;
            w[15] = total * 8;
        }
        sha512_block(st, w);
    }
}

// SHA-512 over prefix (npw little-endian words, npw even) || M.
SPL_HD void sha512_prefixed(uint64_t st[8], const uint32_t *prefix, uint32_t npw,
                            const MsgReader &msg) {
    sha512_init(st);
    sha512_prefixed_blocks(st, prefix, npw, msg, 0u, ~0u);
}

// ------------------------------------------------------------ verify

// Stage 1 of verify: every check that does not need the scalar
// multiplication, plus h = SHA-512(R || A || M) mod L and the decoded -A.
// Returns 0 when the signature is already rejected.
SPL_HD int verify_prepare(const uint32_t sig[16], const uint32_t pk[8], const MsgReader &msg,
                          const ed25519_tables &T, uint32_t h[8], ge_p3 &A) {
    if (sig[15] & 0xf0000000u) return 0;           // S >= 2^252
    if (y_low_order(sig)) return 0;                // R of order <= 4
    uint32_t pre[16];
#pragma unroll
    for (int i = 0; i < 8; i++) { pre[i] = sig[i]; pre[8 + i] = pk[i]; }
    uint64_t st[8];
    sha512_prefixed(st, pre, 16, msg);
    uint32_t hw[16];
    sha512_digest_le_words(hw, st);
    sc_reduce512(h, hw);
    if (ge_decode_neg(A, pk, T.k) != 0) return 0;
    if (y_low_order(pk)) return 0;
    return 1;
}

// Ai[k] = k A for k = 0..n-1 in cached form (Ai[0] is the identity).  Each
// entry is A2 + the entry two back, which is kept in registers: reading it
// back from Ai (global scratch on the device) would put a store-to-load round
// trip through memory on every step.
SPL_HD void build_multiples(ge_cached *Ai, int n, const ge_p3 &A, const fe &d2) {
    ge_p1p1 t;
    ge_p3 u, A2;
    ge_cached back2, back1, cur;
    fe_set(Ai[0].YplusX, 1); fe_set(Ai[0].YminusX, 1);
    fe_set(Ai[0].Z, 1); fe_set(Ai[0].T2d, 0);
    ge_p3_to_cached(back2, A, d2);
    Ai[1] = back2;
    ge_p3_dbl(t, A);
    ge_p1p1_to_p3(A2, t);
    ge_p3_to_cached(back1, A2, d2);
    Ai[2] = back1;
    for (int k = 3; k < n; k++) {
        ge_add_cached(t, A2, back2, 0);
        ge_p1p1_to_p3(u, t);
        ge_p3_to_cached(cur, u, d2);
        Ai[k] = cur;
        back2 = back1;
        back1 = cur;
    }
}

// Stage 2: R' = [h](-A) + [S]B with the table Ai = 0..2^(W-1) x (-A); returns
// 1 when encode(R') equals the R bytes of the signature.
template <int W>
SPL_HD int verify_finish(const uint32_t sig[16], const uint32_t h[8], const ge_cached *Ai,
                         const ed25519_tables &T) {
    typedef VarWindow<W> G;
    uint32_t a[8], s[8];
#pragma unroll
    for (int i = 0; i < 8; i++) { a[i] = h[i]; s[i] = sig[8 + i]; }
    if (G::kAlign) {
#pragma unroll
        for (int i = 7; i > 0; i--) a[i] = (a[i] << G::kAlign) | (a[i - 1] >> (32 - G::kAlign));
        a[0] <<= G::kAlign;
    }
    uint32_t sbelow = 0;
    ge_p3 pb, u;
    ge_p3_0(pb);
    ge_p1p1 t, tb;
    ge_p2 r;
    ge_p2_0(r);
    for (int j = G::kWindows - 1; j >= 0; j--) {
        const int dv = take_top_digit<W>(a);
        const ge_cached e = Ai[dv < 0 ? -dv : dv];
        if (j < 32) {
            // comb digit m = 31 - j of S (S < 2^252: the last digit carries 0)
            const int df = take_low_byte_digit(s, sbelow);
            ge_add_precomp(tb, pb, T.comb[31 - j][df < 0 ? -df : df], df < 0);
            ge_p1p1_to_p3(pb, tb);
        }
        if (j != G::kWindows - 1) {
#pragma unroll
            for (int k = 0; k < W - 1; k++) {
                ge_p2_dbl(t, r);
                ge_p1p1_to_p2(r, t);
            }
            ge_p2_dbl(t, r);
            ge_p1p1_to_p3(u, t);
        } else {
            ge_p3_0(u);
        }
        ge_add_cached(t, u, e, dv < 0);
        if (j != 0) ge_p1p1_to_p2(r, t);
    }
    ge_cached cb;
    ge_p1p1_to_p3(u, t);
    ge_p3_to_cached(cb, pb, T.k.d2);
    ge_add_cached(t, u, cb, 0);
    ge_p1p1_to_p2(r, t);
    uint32_t enc[8];
    ge_encode(enc, r.X, r.Y, r.Z);
    uint32_t diff = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) diff |= enc[j] ^ sig[j];
    return diff == 0;
}

// Returns 1 when the signature verifies under the rule above.  Ai is caller
// storage for VarWindow<W>::kTable cached points.
template <int W = 4>
SPL_HD int verify_one(const uint32_t sig[16], const uint32_t pk[8], const MsgReader &msg,
                      const ed25519_tables &T, ge_cached *Ai) {
    uint32_t h[8];
    ge_p3 A;
    if (!verify_prepare(sig, pk, msg, T, h, A)) return 0;
    build_multiples(Ai, VarWindow<W>::kTable, A, T.k.d2);
    return verify_finish<W>(sig, h, Ai, T);
}

// ------------------------------------------------------------ verify, half-size scalars

// y (the low 255 bits of an encoding) < p = 2^255 - 19
SPL_HD uint32_t enc_y_canonical(const uint32_t enc[8]) {
    uint32_t mid = 0xffffffffu;
#pragma unroll
    for (int i = 1; i < 7; i++) mid &= enc[i];
    const uint32_t ge_p = ((enc[7] & 0x7fffffffu) == 0x7fffffffu) & (mid == 0xffffffffu) &
                          (enc[0] >= 0xffffffedu);
    return ge_p ^ 1u;
}

// a <<= n for 0 <= n < 256 (bits shifted out are dropped)
SPL_HD void shl256(uint32_t a[8], int n) {
    const int q = n >> 5, r = n & 31;
    uint32_t w[8];
#pragma unroll
    for (int i = 7; i >= 0; i--) {
        const uint32_t lo = i > 0 ? a[i - 1] : 0u;
        w[i] = r ? (a[i] << r) | (lo >> (32 - r)) : a[i];
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint32_t v = 0;
#pragma unroll
        for (int j = 0; j < 8; j++) v = (j + q == i) ? w[j] : v;
        a[i] = v;
    }
}

// Stage 1 of the half-size verify (see lattice.h): the same acceptance checks
// as verify_prepare, plus -R (decoded: the equation below needs the point), the
// half scalars x, y of h and sp = y S mod L.  P1 receives -A, or A when x < 0,
// so the multiplication below always uses |x|.  Returns 0 when rejected.
//   accept <=> [sp]B + [|x|]P1 + [y](-R) = O
// R's bytes must be the canonical encoding of a decodable point that is not of
// small order: encode() only produces canonical encodings, so the reference's
// byte comparison rejects every other R.
SPL_HD int verify_prepare_half(const uint32_t sig[16], const uint32_t pk[8], const MsgReader &msg,
                               const ed25519_tables &T, uint32_t x[8], uint32_t y[8],
                               uint32_t sp[8], ge_p3 &P1, ge_p3 &R) {
    if (sig[15] & 0xf0000000u) return 0;           // S >= 2^252
    if (!enc_y_canonical(sig)) return 0;           // R not canonical
    if (y_low_order(sig)) return 0;                // R of order <= 4
    uint32_t pre[16];
#pragma unroll
    for (int i = 0; i < 8; i++) { pre[i] = sig[i]; pre[8 + i] = pk[i]; }
    uint64_t st[8];
    sha512_prefixed(st, pre, 16, msg);
    uint32_t hw[16], h[8];
    sha512_digest_le_words(hw, st);
    sc_reduce512(h, hw);
    {
        // -A and -R in one lockstep decode (two independent exponentiations)
        ge_p3 pts[2];
        const uint32_t *const enc[2] = {pk, sig};
        uint32_t bad[2];
        ge_decode_neg_n(pts, enc, T.k, bad);
        if (bad[0] | bad[1]) return 0;
        P1 = pts[0];
        R = pts[1];
    }
    if (y_low_order(pk)) return 0;
    uint32_t x_neg;
    half_scalars(h, x, x_neg, y);
    const uint32_t zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    sc_muladd(sp, y, sig + 8, zero);
    if (x_neg) {
        fe_neg(P1.X, P1.X);
        fe_neg(P1.T, P1.T);
    }
    return 1;
}

// Bits needed by the windowed multiplication of stage 2 for this lane.
SPL_HD int half_bits(const uint32_t x[8], const uint32_t y[8]) {
    uint32_t o[8];
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = x[i] | y[i];
    int r = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
        if (o[i]) r = 32 * i + bitlen32(o[i]);
    return r;
}

// Windows of W bits covering `bits` (one extra bit for the signed digits).
template <int W>
SPL_HD int half_windows(int bits) { return (bits + W) / W; }

// Stage 2: K windows of W bits (K W <= 256, the same K in every lane of a
// wave); T1, Tr hold 0..2^(W-1) times P1 and -R.  Returns 1 when
// [sp]B + [x]P1 + [y](-R) is the identity.
//
// PF: load each window's three table entries (two per-signature tables in
// global scratch, far larger than the L2 for a full batch, and one comb
// entry) before its W doublings, with a scheduling barrier so they stay
// there and the doublings cover their latency.  Measured on gfx950 the
// compiler's own schedule already hid most of it (the window loop runs at the
// single-wave issue rate either way); the barrier makes the order explicit.
// It holds ~110 more registers live across the doublings, so it suits
// kernels with a whole SIMD's register file (one wave per SIMD).
template <int W, bool PF = false>
SPL_HD int verify_finish_half(const uint32_t x[8], const uint32_t y[8], const uint32_t sp[8],
                              int K, const ge_cached *T1, const ge_cached *Tr,
                              const ed25519_tables &T) {
    uint32_t a[8], b[8], s[8];
#pragma unroll
    for (int i = 0; i < 8; i++) { a[i] = x[i]; b[i] = y[i]; s[i] = sp[i]; }
    shl256(a, 256 - K * W);
    shl256(b, 256 - K * W);
    uint32_t sbelow = 0;
    int m = 0;                        // comb digits of sp consumed
    ge_p3 pb, u;
    ge_p3_0(pb);
    ge_p1p1 t, tb;
    ge_p2 r;
    ge_p2_0(r);
    for (int j = K - 1; j >= 0; j--) {
        const int d1 = take_top_digit<W>(a);
        const int d2 = take_top_digit<W>(b);
        if (PF) {
            const ge_cached c1 = T1[d1 < 0 ? -d1 : d1];
            const ge_cached c2 = Tr[d2 < 0 ? -d2 : d2];
            const bool comb = m < 32;              // the same in every lane
            int df = 0;
            ge_precomp cp;
            if (comb) {
                df = take_low_byte_digit(s, sbelow);
                cp = T.comb[m][df < 0 ? -df : df];
            }
            sched_fence();
            if (j != K - 1) {
#pragma unroll
                for (int k = 0; k < W - 1; k++) {
                    ge_p2_dbl(t, r);
                    ge_p1p1_to_p2(r, t);
                }
                ge_p2_dbl(t, r);
                ge_p1p1_to_p3(u, t);
            } else {
                ge_p3_0(u);
            }
            if (comb) {
                ge_add_precomp(tb, pb, cp, df < 0);
                ge_p1p1_to_p3(pb, tb);
                m++;
            }
            ge_add_cached(t, u, c1, d1 < 0);
            ge_p1p1_to_p3(u, t);
            ge_add_cached(t, u, c2, d2 < 0);
            if (j != 0) ge_p1p1_to_p2(r, t);
            continue;
        }
        if (m < 32) {
            const int df = take_low_byte_digit(s, sbelow);
            ge_add_precomp(tb, pb, T.comb[m][df < 0 ? -df : df], df < 0);
            ge_p1p1_to_p3(pb, tb);
            m++;
        }
        if (j != K - 1) {
#pragma unroll
            for (int k = 0; k < W - 1; k++) {
                ge_p2_dbl(t, r);
                ge_p1p1_to_p2(r, t);
            }
            ge_p2_dbl(t, r);
            ge_p1p1_to_p3(u, t);
        } else {
            ge_p3_0(u);
        }
        ge_add_cached(t, u, T1[d1 < 0 ? -d1 : d1], d1 < 0);
        ge_p1p1_to_p3(u, t);
        ge_add_cached(t, u, Tr[d2 < 0 ? -d2 : d2], d2 < 0);
        if (j != 0) ge_p1p1_to_p2(r, t);
    }
    for (; m < 32; m++) {             // comb digits left when K < 32
        const int df = take_low_byte_digit(s, sbelow);
        ge_add_precomp(tb, pb, T.comb[m][df < 0 ? -df : df], df < 0);
        ge_p1p1_to_p3(pb, tb);
    }
    ge_cached cb;
    ge_p1p1_to_p3(u, t);
    ge_p3_to_cached(cb, pb, T.k.d2);
    ge_add_cached(t, u, cb, 0);
    ge_p1p1_to_p2(r, t);
    fe dz;
    fe_sub(dz, r.Y, r.Z);
    return (int)(fe_iszero(r.X) & fe_iszero(dz));
}

// The half-size verify for one signature with its own window count (host
// and per-lane use).  Tables: caller storage for 2 * VarWindow<W>::kTable
// cached points.
template <int W = 4>
SPL_HD int verify_half_one(const uint32_t sig[16], const uint32_t pk[8], const MsgReader &msg,
                           const ed25519_tables &T, ge_cached *tables) {
    uint32_t x[8], y[8], sp[8];
    ge_p3 P1, R;
    if (!verify_prepare_half(sig, pk, msg, T, x, y, sp, P1, R)) return 0;
    constexpr int kT = VarWindow<W>::kTable;
    build_multiples(tables, kT, P1, T.k.d2);
    build_multiples(tables + kT, kT, R, T.k.d2);
    return verify_finish_half<W>(x, y, sp, half_windows<W>(half_bits(x, y)), tables,
                                 tables + kT, T);
}

// ------------------------------------------------------------ fixed-base mult

// h = a B for a 256-bit scalar with a < 2^255 (signed radix-16, constant-time
// table scans: the scalar is secret when signing).  Same digit recoding and
// table shape as the reference's ge_scalarmult_base (ge.cu:452-500).
SPL_HD void ge_scalarmult_base(ge_p3 &h, const uint32_t a[8], const ed25519_tables &T) {
    int8_t e[64];
    for (int i = 0; i < 64; i++) e[i] = (int8_t)((a[i >> 3] >> (4 * (i & 7))) & 15);
    int carry = 0;
    for (int i = 0; i < 63; i++) {
        int v = e[i] + carry;
        carry = (v + 8) >> 4;
        e[i] = (int8_t)(v - (carry << 4));
    }
    e[63] = (int8_t)(e[63] + carry);
    ge_p3_0(h);
    ge_p1p1 r;
    for (int pass = 1; pass >= 0; pass--) {
        for (int i = pass; i < 64; i += 2) {
            int d = e[i];
            uint32_t neg = (uint32_t)d >> 31;             // 1 when d < 0
            uint32_t ad = (uint32_t)(d < 0 ? -d : d);      // 0..8
            ge_precomp sel;
            fe_set(sel.yplusx, 1); fe_set(sel.yminusx, 1); fe_set(sel.xy2d, 0);
            for (uint32_t j = 0; j < 8; j++) {
                uint32_t hit = (ad == j + 1);
                const ge_precomp &c = T.base[i >> 1][j];
                fe_cmov(sel.yplusx, c.yplusx, hit);
                fe_cmov(sel.yminusx, c.yminusx, hit);
                fe_cmov(sel.xy2d, c.xy2d, hit);
            }
            ge_precomp ng;
            ng.yplusx = sel.yminusx;
            ng.yminusx = sel.yplusx;
            fe_neg(ng.xy2d, sel.xy2d);
            fe_cmov(sel.yplusx, ng.yplusx, neg);
            fe_cmov(sel.yminusx, ng.yminusx, neg);
            fe_cmov(sel.xy2d, ng.xy2d, neg);
            ge_madd_ct(r, h, sel);
            ge_p1p1_to_p3(h, r);
        }
        if (pass == 1) {
            ge_p2 s;
            ge_p3_dbl(r, h);
            ge_p1p1_to_p2(s, r);
            ge_p2_dbl(r, s);
            ge_p1p1_to_p2(s, r);
            ge_p2_dbl(r, s);
            ge_p1p1_to_p2(s, r);
            ge_p2_dbl(r, s);
            ge_p1p1_to_p3(h, r);
        }
    }
}

// ------------------------------------------------------------ sign, keypair

// RFC 8032 signature with the 64-byte expanded private key
// (clamped scalar || prefix), the reference's composition (sign.cu:9-38).
SPL_HD void sign_one(uint32_t sig[16], const uint32_t pk[8], const uint32_t sk[16],
                     const MsgReader &msg, const ed25519_tables &T) {
    uint64_t st[8];
    uint32_t dw[16], r[8], hram[8];
    sha512_prefixed(st, sk + 8, 8, msg);
    sha512_digest_le_words(dw, st);
    sc_reduce512(r, dw);
    ge_p3 R;
    ge_scalarmult_base(R, r, T);
    ge_encode(sig, R.X, R.Y, R.Z);
    uint32_t pre[16];
#pragma unroll
    for (int i = 0; i < 8; i++) { pre[i] = sig[i]; pre[8 + i] = pk[i]; }
    sha512_prefixed(st, pre, 16, msg);
    sha512_digest_le_words(dw, st);
    sc_reduce512(hram, dw);
    sc_muladd(sig + 8, hram, sk, r);
}

// ed25519_create_keypair (keypair.cu:6-16): sk = clamp(SHA-512(seed)[0:32]) ||
// SHA-512(seed)[32:64], pk = encode(sk_scalar B).
SPL_HD void keypair_from_seed(uint32_t pk[8], uint32_t sk[16], const uint32_t seed[8],
                              const ed25519_tables &T) {
    MsgReader none{nullptr, 0};
    uint64_t st[8];
    sha512_prefixed(st, seed, 8, none);
    sha512_digest_le_words(sk, st);
    sk[0] &= ~7u;
    sk[7] &= 0x3fffffffu;
    sk[7] |= 0x40000000u;
    ge_p3 A;
    ge_scalarmult_base(A, sk, T);
    ge_encode(pk, A.X, A.Y, A.Z);
}

}  // namespace spl
