// Edwards25519 group arithmetic (-x^2 + y^2 = 1 + d x^2 y^2) over fe25519.h.
//
// Coordinate systems follow the extended twisted Edwards formulas of
// Hisil, Wong, Carter and Dawson (2008) with a = -1, the same ones the
// reference uses (src/cuda-ecc-ed25519/ge.h:21-54, ge.cu:225-549):
//   p2     (X:Y:Z)            x = X/Z, y = Y/Z
//   p3     (X:Y:Z:T)          additionally XY = ZT
//   p1p1   ((X:Z),(Y:T))      "completed" output of add/double
//   cached (Y+X, Y-X, Z, 2dT) right operand of a projective add
//   precomp(y+x, y-x, 2dxy)   affine right operand (Z = 1)
// These formulas are complete on the whole curve, so results do not depend on
// the order of operations: any correct evaluation gives the same point.
#pragma once
#include "fe25519.h"

namespace spl {

struct ge_p2 { fe X, Y, Z; };
struct ge_p3 { fe X, Y, Z, T; };
struct ge_p1p1 { fe X, Y, Z, T; };
struct ge_precomp { fe yplusx, yminusx, xy2d; };
struct ge_cached { fe YplusX, YminusX, Z, T2d; };

// Curve constants (canonical field elements).
struct curve_consts {
    fe d;        // -121665/121666
    fe d2;       // 2d
    fe sqrtm1;   // 2^((p-1)/4)
};

SPL_HD void ge_p3_0(ge_p3 &h) {
    fe_set(h.X, 0); fe_set(h.Y, 1); fe_set(h.Z, 1); fe_set(h.T, 0);
}

SPL_HD void ge_p2_0(ge_p2 &h) { fe_set(h.X, 0); fe_set(h.Y, 1); fe_set(h.Z, 1); }

SPL_HD void ge_p1p1_to_p2(ge_p2 &r, const ge_p1p1 &p) {
    fe_mul(r.X, p.X, p.T);
    fe_mul(r.Y, p.Y, p.Z);
    fe_mul(r.Z, p.Z, p.T);
}

SPL_HD void ge_p1p1_to_p3(ge_p3 &r, const ge_p1p1 &p) {
    fe_mul(r.X, p.X, p.T);
    fe_mul(r.Y, p.Y, p.Z);
    fe_mul(r.Z, p.Z, p.T);
    fe_mul(r.T, p.X, p.Y);
}

SPL_HD void ge_p3_to_cached(ge_cached &r, const ge_p3 &p, const fe &d2) {
    fe_add(r.YplusX, p.Y, p.X);
    fe_sub(r.YminusX, p.Y, p.X);
    r.Z = p.Z;
    fe_mul(r.T2d, p.T, d2);
}

// r = 2p, p projective: 3 squares (one doubled in the column sums so that
// T stays a valid multiplier input) + 1 square of (X+Y)
SPL_HD void ge_p2_dbl(ge_p1p1 &r, const ge_p2 &p) {
    fe xx, yy, b, aa;
    fe_sq(xx, p.X);
    fe_sq(yy, p.Y);
    fe_sq2(b, p.Z);
    fe_add(aa, p.X, p.Y);
    fe_sq(aa, aa);
    fe_add(r.Y, yy, xx);
    fe_sub(r.Z, yy, xx);
    fe_sub(r.X, aa, r.Y);
    fe_sub(r.T, b, r.Z);
}

SPL_HD void ge_p3_dbl(ge_p1p1 &r, const ge_p3 &p) {
    ge_p2 q;
    q.X = p.X; q.Y = p.Y; q.Z = p.Z;
    ge_p2_dbl(r, q);
}

// r = p + q (neg = 0) or p - q (neg = 1), q cached.
SPL_HD void ge_add_cached(ge_p1p1 &r, const ge_p3 &p, const ge_cached &q, uint32_t neg) {
    fe a, b, c, dd;
    fe_add(a, p.Y, p.X);
    fe_sub(b, p.Y, p.X);
    fe_mul(a, a, neg ? q.YminusX : q.YplusX);
    fe_mul(b, b, neg ? q.YplusX : q.YminusX);
    fe_mul(c, q.T2d, p.T);
    fe_mul(dd, p.Z, q.Z);
    fe_add(dd, dd, dd);
    fe_sub(r.X, a, b);
    fe_add(r.Y, a, b);
    if (neg) { fe_sub(r.Z, dd, c); fe_add(r.T, dd, c); }
    else     { fe_add(r.Z, dd, c); fe_sub(r.T, dd, c); }
}

// r = p + q (neg = 0) or p - q (neg = 1), q affine precomp.
SPL_HD void ge_add_precomp(ge_p1p1 &r, const ge_p3 &p, const ge_precomp &q, uint32_t neg) {
    fe a, b, c, dd;
    fe_add(a, p.Y, p.X);
    fe_sub(b, p.Y, p.X);
    fe_mul(a, a, neg ? q.yminusx : q.yplusx);
    fe_mul(b, b, neg ? q.yplusx : q.yminusx);
    fe_mul(c, q.xy2d, p.T);
    fe_add(dd, p.Z, p.Z);
    fe_sub(r.X, a, b);
    fe_add(r.Y, a, b);
    if (neg) { fe_sub(r.Z, dd, c); fe_add(r.T, dd, c); }
    else     { fe_add(r.Z, dd, c); fe_sub(r.T, dd, c); }
}

// Constant-time (branch-free) form of ge_add_precomp for secret operands.
SPL_HD void ge_madd_ct(ge_p1p1 &r, const ge_p3 &p, const ge_precomp &q) {
    fe a, b, c, dd;
    fe_add(a, p.Y, p.X);
    fe_sub(b, p.Y, p.X);
    fe_mul(a, a, q.yplusx);
    fe_mul(b, b, q.yminusx);
    fe_mul(c, q.xy2d, p.T);
    fe_add(dd, p.Z, p.Z);
    fe_sub(r.X, a, b);
    fe_add(r.Y, a, b);
    fe_add(r.Z, dd, c);
    fe_sub(r.T, dd, c);
}

// Encoding: y with the sign of x in bit 255 (canonical).
SPL_HD void ge_encode(uint32_t out[8], const fe &X, const fe &Y, const fe &Z) {
    fe zi, x, y;
    fe_invert(zi, Z);
    fe_mul(x, X, zi);
    fe_mul(y, Y, zi);
    fe_to_words(out, y);
    out[7] ^= fe_isneg(x) << 31;
}

// Decompression returning the NEGATED point, with the reference's acceptance
// rule (ge.cu:182-222): y is the low 255 bits (a value >= p is accepted),
// x = u v^3 (u v^7)^((p-5)/8) with u = y^2 - 1, v = d y^2 + 1; if v x^2 != u
// then x *= sqrt(-1) when v x^2 == -u, otherwise the encoding is rejected;
// finally x is negated when its parity equals bit 255, so x = 0 with bit 255
// set is accepted.
//
// N encodings are decoded in lockstep (the exponentiation dominates; see
// fe_chain250_n).  The root selection and the sign fix are branch-free, so
// the N chains stay in one instruction stream.  bad[j] is 1 when enc[j] is
// rejected; h[j] is then unspecified.
template <int N>
SPL_HD void ge_decode_neg_n(ge_p3 (&h)[N], const uint32_t *const (&enc)[N],
                            const curve_consts &k, uint32_t (&bad)[N]) {
    fe u[N], v[N], v3[N], x[N];
    for_each_n<N>([&](auto j) {
        fe_from_words(h[j].Y, enc[j]);
        fe_set(h[j].Z, 1);
        fe_sq(u[j], h[j].Y);
        fe_mul(v[j], u[j], k.d);
        fe_sub(u[j], u[j], h[j].Z);
        fe_add(v[j], v[j], h[j].Z);
        fe_sq(v3[j], v[j]);
        fe_mul(v3[j], v3[j], v[j]);
        fe_sq(x[j], v3[j]);
        fe_mul(x[j], x[j], v[j]);
        fe_mul(x[j], x[j], u[j]);
    });
    fe_pow22523_n(x, x);
    for_each_n<N>([&](auto j) {
        fe vxx, chk, xi, nx;
        fe_mul(x[j], x[j], v3[j]);
        fe_mul(x[j], x[j], u[j]);
        fe_sq(vxx, x[j]);
        fe_mul(vxx, vxx, v[j]);
        fe_sub(chk, vxx, u[j]);
        const uint32_t root = fe_iszero(chk);          // v x^2 == u
        fe_add(chk, vxx, u[j]);
        const uint32_t flip = fe_iszero(chk);          // v x^2 == -u
        bad[j] = (root | flip) ^ 1u;
        fe_mul(xi, x[j], k.sqrtm1);
        fe_cmov(x[j], xi, root ^ 1u);
        fe_neg(nx, x[j]);
        fe_cmov(x[j], nx, fe_isneg(x[j]) == (enc[j][7] >> 31));
        h[j].X = x[j];
        fe_mul(h[j].T, h[j].X, h[j].Y);
    });
}

// One encoding.  Returns 0 on success.
SPL_HD int ge_decode_neg(ge_p3 &h, const uint32_t enc[8], const curve_consts &k) {
    ge_p3 out[1];
    const uint32_t *const in[1] = {enc};
    uint32_t bad[1];
    ge_decode_neg_n(out, in, k, bad);
    h = out[0];
    return bad[0] ? -1 : 0;
}

// The reference's "small order" decision (ge.cu:60-82 with the raw-limb
// identity test of fe.cu:46-67) fires exactly for encodings that decode and
// whose y (taken mod p) is 0, 1 or p-1, i.e. the points of order 1, 2 and 4.
// Order-8 points are not flagged.  Input: the 32-byte encoding as words.
SPL_HD uint32_t y_low_order(const uint32_t enc[8]) {
    fe y;
    uint32_t w[8];
    fe_from_words(y, enc);
    fe_to_words(w, y);
    uint32_t mid0 = 0, mid1 = 0xffffffffu;   // OR and AND of words 1..6
#pragma unroll
    for (int i = 1; i < 7; i++) { mid0 |= w[i]; mid1 &= w[i]; }
    uint32_t is0 = (w[0] == 0) & (mid0 == 0) & (w[7] == 0);
    uint32_t is1 = (w[0] == 1) & (mid0 == 0) & (w[7] == 0);
    uint32_t ispm1 = (w[0] == 0xffffffecu) & (mid1 == 0xffffffffu) & (w[7] == 0x7fffffffu);
    return is0 | is1 | ispm1;
}

}  // namespace spl
