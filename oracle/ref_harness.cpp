// CPU ORACLE support — test infrastructure only.
//
// Builds the reference's own host-compilable primitives (fe.cu, ge.cu, sc.cu,
// sha512.cu, keypair.cu from src/cuda-ecc-ed25519 and sha256.cu from
// src/cuda-sha256) straight from /root/reference with g++ (the CUDA
// qualifiers are defined away on the command line by oracle/Makefile), and
// exposes them to the tests with C linkage.  verify.cu and sign.cu cannot be
// compiled without the CUDA runtime (they hold kernel launches), so the short
// host compositions of those files are written out here as calls into the
// reference primitives:
//   ref_verify                  <- verify.cu:60-174
//   ref_get_checked_scalar      <- verify.cu:60-85 (with a 33-byte buffer,
//                                  see below)
//   ref_check_small_order       <- verify.cu:88-107
//   ref_sign                    <- sign.cu:9-38
//   ref_poh                     <- poh_verify.cu:12-27 loop over sha256.cu
// Output goes only to oracle/_ref/ (git-ignored).
#include <assert.h>
#include <stdint.h>
#include <string.h>

#include "sc.cu"
#include "fe.cu"
#include "ge.cu"
#include "sha512.cu"
#include "sha256.cu"

extern "C" void ed25519_create_keypair(unsigned char *public_key,
                                       unsigned char *private_key,
                                       const unsigned char *seed);

static int ref_equal32(const unsigned char *x, const unsigned char *y) {
    unsigned char r = 0;
    for (int i = 0; i < 32; i++) r |= x[i] ^ y[i];
    return r == 0;
}

extern "C" {

// scalar32_reduce reads 33 bytes (limb 4 is a 7-byte load at offset 26), so
// the caller's 32-byte buffer is widened by one byte holding `extra`.
int ref_get_checked_scalar(unsigned char *out, const unsigned char *in,
                           unsigned char extra) {
    if ((in[31] & 0xf0) == 0) {
        memcpy(out, in, 32);
        return 0;
    }
    if ((in[31] >> 7) != 0) return 1;
    unsigned char buf[40];
    memset(buf, 0, sizeof buf);
    memcpy(buf, out, 32);
    buf[32] = extra;
    scalar32_reduce(buf);
    memcpy(out, buf, 32);
    return ref_equal32(out, in) ? 0 : 1;
}

int ref_check_small_order(const unsigned char *enc) {
    ge_p3 p;
    if (ge_frombytes_negate_vartime(&p, enc) != 0) return 1;
    if (ge_is_small_order(&p) != 0) return 1;
    return 0;
}

// 1 when ge_is_small_order fires on a point that decodes, 0 when it does not,
// -1 when the encoding does not decode.
int ref_small_order_raw(const unsigned char *enc) {
    ge_p3 p;
    if (ge_frombytes_negate_vartime(&p, enc) != 0) return -1;
    return ge_is_small_order(&p) ? 1 : 0;
}

int ref_verify(const unsigned char *sig, const unsigned char *msg,
               uint32_t len, const unsigned char *pk) {
    unsigned char checker[32];
    memset(checker, 0, sizeof checker);
    unsigned char sbuf[40];
    memset(sbuf, 0, sizeof sbuf);
    // S gate: top nibble clear copies S; otherwise the reference reduces an
    // uninitialised buffer and compares — modelled as a zero buffer.
    if ((sig[63] & 0xf0) != 0) {
        if ((sig[63] >> 7) != 0) return 0;
        scalar32_reduce(sbuf);
        if (!ref_equal32(sbuf, sig + 32)) return 0;
    }
    if (ref_check_small_order(sig) != 0) return 0;
    sha512_context hash;
    unsigned char h[64];
    sha512_init(&hash);
    sha512_update(&hash, sig, 32);
    sha512_update(&hash, pk, 32);
    sha512_update(&hash, msg, len);
    sha512_final(&hash, h);
    sc_reduce(h);
    ge_cached Ai[GE_LOOKUP_SIZE];
    if (ge_gen_lookup(pk, Ai) == 0) return 0;
    ge_p2 R;
    ge_double_scalarmult_vartime(&R, h, Ai, sig + 32);
    ge_tobytes(checker, &R);
    return ref_equal32(checker, sig) ? 1 : 0;
}

void ref_create_keypair(unsigned char *pk, unsigned char *sk,
                        const unsigned char *seed) {
    ed25519_create_keypair(pk, sk, seed);
}

void ref_sign(unsigned char *sig, const unsigned char *msg, size_t len,
              const unsigned char *pk, const unsigned char *sk) {
    sha512_context hash;
    unsigned char hram[64];
    unsigned char r[64];
    ge_p3 R;
    sha512_init(&hash);
    sha512_update(&hash, sk + 32, 32);
    sha512_update(&hash, msg, len);
    sha512_final(&hash, r);
    sc_reduce(r);
    ge_scalarmult_base(&R, r);
    ge_p3_tobytes(sig, &R);
    sha512_init(&hash);
    sha512_update(&hash, sig, 32);
    sha512_update(&hash, pk, 32);
    sha512_update(&hash, msg, len);
    sha512_final(&hash, hram);
    sc_reduce(hram);
    sc_muladd(sig + 32, hram, sk, r);
}

void ref_sc_reduce(unsigned char *s) { sc_reduce(s); }

void ref_sc_muladd(unsigned char *s, const unsigned char *a,
                   const unsigned char *b, const unsigned char *c) {
    sc_muladd(s, a, b, c);
}

void ref_poh(unsigned char *hashes, const uint64_t *num_hashes, size_t num) {
    for (size_t i = 0; i < num; i++) {
        unsigned char *hash = hashes + 32 * i;
        for (uint64_t k = 0; k < num_hashes[i]; k++) {
            hash_state st;
            sha256_init(&st);
            sha256_process(&st, hash, 32);
            sha256_done(&st, hash);
        }
    }
}

// `count` signatures of one message (timing loop for the CPU baseline).
void ref_sign_repeat(unsigned char *sig, const unsigned char *msg, size_t len,
                     const unsigned char *pk, const unsigned char *sk, uint32_t count) {
    for (uint32_t i = 0; i < count; i++) ref_sign(sig, msg, len, pk, sk);
}

void ref_verify_many(const unsigned char *packets, uint32_t n,
                     const uint32_t *message_lens, const uint32_t *pk_off,
                     const uint32_t *sig_off, const uint32_t *msg_off,
                     unsigned char *out) {
    for (uint32_t i = 0; i < n; i++)
        out[i] = (unsigned char)ref_verify(packets + sig_off[i],
                                           packets + msg_off[i],
                                           message_lens[i], packets + pk_off[i]);
}

}  // extern "C"
