// Host-side exports of libcuda-crypt.so: single-shot Ed25519 operations, the
// two check helpers, init/verbosity/licence and host-memory registration.
// They run the same __host__ __device__ cores as the kernels.
#include <stdio.h>
#include <string.h>

#include "runtime.h"
#include "../../include/cuda_crypt.h"

namespace spl {

static void bytes_to_words(uint32_t *w, const unsigned char *b, int n) {
    for (int i = 0; i < n; i++)
        w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) |
               ((uint32_t)b[4 * i + 2] << 16) | ((uint32_t)b[4 * i + 3] << 24);
}

static void words_to_bytes(unsigned char *b, const uint32_t *w, int n) {
    for (int i = 0; i < n; i++) {
        b[4 * i] = (unsigned char)w[i];
        b[4 * i + 1] = (unsigned char)(w[i] >> 8);
        b[4 * i + 2] = (unsigned char)(w[i] >> 16);
        b[4 * i + 3] = (unsigned char)(w[i] >> 24);
    }
}

static uint64_t le_load(const unsigned char *s, int n) {
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; i--) v = (v << 8) | s[i];
    return v;
}

// The reference's scalar32_reduce (src/cuda-ecc-ed25519/sc.cu:12-224): the
// input, unpacked into five 52-bit limbs, is multiplied by R = 2^260 mod L and
// Montgomery-reduced (factor 2^-260), then L is subtracted once if that does
// not go negative.  Two of its behaviours decide the exported
// ed25519_get_checked_scalar result and are kept: the schoolbook product's
// column 6 is (column 5) + a4*R2 rather than a2*R4 + a3*R3 + a4*R2
// (sc.cu:58-59), and limb 4 spans bits 208..259, whose top nibble would come
// from the byte after the 32-byte buffer (sc.cu:79); that byte is taken as 0.
static void scalar32_reduce_compat(unsigned char s[32]) {
    typedef unsigned __int128 u128;
    const uint64_t R52[5] = {0x000f48bd6721e6edull, 0x0003bab5ac67e45aull,
                             0x000fffffeb35e51bull, 0x000fffffffffffffull,
                             0x00000fffffffffffull};
    const uint64_t L52[5] = {0x0002631a5cf5d3edull, 0x000dea2f79cd6581ull,
                             0x000000000014def9ull, 0, 0x0000100000000000ull};
    const uint64_t LF = 0x51da312547e1bull;   // -L^-1 mod 2^52
    const uint64_t M = (1ull << 52) - 1;
    unsigned char e[40] = {0};
    memcpy(e, s, 32);
    uint64_t a[5] = {le_load(e, 7) & M, (le_load(e + 6, 7) >> 4) & M, le_load(e + 13, 7) & M,
                     (le_load(e + 19, 7) >> 4) & M, le_load(e + 26, 7) & M};
    u128 z[9];
    for (int k = 0; k < 9; k++) {
        z[k] = 0;
        for (int i = 0; i < 5; i++)
            if (k - i >= 0 && k - i < 5) z[k] += (u128)a[i] * R52[k - i];
    }
    z[6] = z[5] + (u128)a[4] * R52[2];
    // Montgomery reduction, one 52-bit limb at a time
    uint64_t nq[5];
    u128 c = 0;
    const int rows[5][3][2] = {
        {{-1, -1}, {-1, -1}, {-1, -1}},
        {{0, 1}, {-1, -1}, {-1, -1}},
        {{0, 2}, {1, 1}, {-1, -1}},
        {{1, 2}, {2, 1}, {-1, -1}},
        {{0, 4}, {3, 1}, {2, 2}},
    };
    for (int k = 0; k < 5; k++) {
        u128 sum = z[k] + c;
        for (int t = 0; t < 3; t++)
            if (rows[k][t][0] >= 0) sum += (u128)nq[rows[k][t][0]] * L52[rows[k][t][1]];
        nq[k] = ((uint64_t)sum * LF) & M;
        c = (sum + (u128)nq[k] * L52[0]) >> 52;
    }
    uint64_t r[5];
    u128 sum;
    sum = c + z[5] + (u128)nq[1] * L52[4] + (u128)nq[3] * L52[2] + (u128)nq[4] * L52[1];
    r[0] = (uint64_t)sum & M; c = sum >> 52;
    sum = c + z[6] + (u128)nq[2] * L52[4] + (u128)nq[4] * L52[2];
    r[1] = (uint64_t)sum & M; c = sum >> 52;
    sum = c + z[7] + (u128)nq[3] * L52[4];
    r[2] = (uint64_t)sum & M; c = sum >> 52;
    sum = c + z[8] + (u128)nq[4] * L52[4];
    r[3] = (uint64_t)sum & M; c = sum >> 52;
    r[4] = (uint64_t)c;
    uint64_t d[5], borrow = 0;
    for (int i = 0; i < 5; i++) {
        borrow = r[i] - (L52[i] + (borrow >> 63));
        d[i] = borrow & M;
    }
    uint64_t mask = ((borrow >> 63) ^ 1) - 1;
    uint64_t cc = 0;
    for (int i = 0; i < 5; i++) {
        cc = (cc >> 52) + d[i] + (L52[i] & mask);
        d[i] = cc & M;
    }
    uint64_t w[4] = {d[0] | (d[1] << 52), (d[1] >> 12) | (d[2] << 40), (d[2] >> 24) | (d[3] << 28),
                     (d[3] >> 36) | (d[4] << 16)};
    for (int i = 0; i < 32; i++) s[i] = (unsigned char)(w[i >> 3] >> (8 * (i & 7)));
}

}  // namespace spl

using namespace spl;

extern "C" {

int ed25519_create_seed(unsigned char *seed) {
    FILE *f = fopen("/dev/urandom", "rb");
    if (!f) return 1;
    size_t got = fread(seed, 1, 32, f);
    fclose(f);
    return got == 32 ? 0 : 1;
}

void ed25519_create_keypair(unsigned char *public_key, unsigned char *private_key,
                            const unsigned char *seed) {
    uint32_t s[8], pk[8], sk[16];
    bytes_to_words(s, seed, 8);
    keypair_from_seed(pk, sk, s, host_tables());
    words_to_bytes(public_key, pk, 8);
    words_to_bytes(private_key, sk, 16);
}

void ed25519_sign(unsigned char *signature, const unsigned char *message, size_t message_len,
                  const unsigned char *public_key, const unsigned char *private_key) {
    uint32_t pk[8], sk[16], sig[16];
    bytes_to_words(pk, public_key, 8);
    bytes_to_words(sk, private_key, 16);
    MsgReader msg{message, (uint32_t)message_len};
    sign_one(sig, pk, sk, msg, host_tables());
    words_to_bytes(signature, sig, 16);
}

int ed25519_verify(const unsigned char *signature, const unsigned char *message,
                   uint32_t message_len, const unsigned char *public_key) {
    uint32_t sig[16], pk[8];
    bytes_to_words(sig, signature, 16);
    bytes_to_words(pk, public_key, 8);
    MsgReader msg{message, message_len};
    ge_cached Ai[9];
    return verify_one(sig, pk, msg, host_tables(), Ai);
}

int ed25519_get_checked_scalar(unsigned char *out_scalar, const unsigned char *in_scalar) {
    if ((in_scalar[31] & 0xf0) == 0) {
        memcpy(out_scalar, in_scalar, 32);
        return 0;
    }
    if (in_scalar[31] >> 7) return 1;
    scalar32_reduce_compat(out_scalar);
    return memcmp(out_scalar, in_scalar, 32) == 0 ? 0 : 1;
}

int ed25519_check_packed_ge_small_order(const unsigned char *packed_group_element) {
    uint32_t enc[8];
    bytes_to_words(enc, packed_group_element, 8);
    ge_p3 p;
    if (ge_decode_neg(p, enc, host_tables().k) != 0) return 1;
    return y_low_order(enc) ? 1 : 0;
}

const char *ed25519_license(void) {
    return "Copyright (c) 2018 Solana Labs, Inc. "
           "Licensed under the Apache License, Version 2.0 "
           "<http://www.apache.org/licenses/LICENSE-2.0>";
}

bool ed25519_init(void) { return pool_init() > 0; }

void ed25519_free_gpu_mem(void) { pool_free_buffers(); }

int cuda_host_register(void *ptr, size_t size, unsigned int flags) {
    hipError_t e = hipHostRegister(ptr, size, flags);
    if (e != hipSuccess) (void)hipGetLastError();   // do not leave the error pending
    return (int)e;
}

int cuda_host_unregister(void *ptr) {
    hipError_t e = hipHostUnregister(ptr);
    if (e != hipSuccess) (void)hipGetLastError();
    return (int)e;
}

}  // extern "C"
