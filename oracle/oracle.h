/*
 * CPU ORACLE — test infrastructure only.
 *
 * A plain-C restatement of the reference's algorithms for the batched
 * Ed25519 verify/sign and SHA-256 Proof-of-History paths of
 * solana-perf-libs.  Only the test-suite, the smoke check and the CPU
 * baseline leg of the benchmark load this library; the product library
 * (libcuda-crypt.so) never links or calls it.
 *
 * Parity pins (what each function is checked against in tests/):
 *   oracle_sha256 / oracle_poh_chain      hashlib.sha256 + the reference's
 *                                         129-chain PoH fixture
 *   oracle_sha512                         hashlib.sha512
 *   oracle_sc_reduce / oracle_sc_muladd   Python big-int mod L
 *   oracle_create_keypair / oracle_sign   RFC 8032 section 7.1 vectors,
 *                                         a Python big-int RFC 8032 signer,
 *                                         and the compiled reference
 *   oracle_verify                         closed-form big-int predicate,
 *                                         the compiled reference
 *   oracle_check_packed_ge_small_order    the compiled reference on the
 *                                         8-torsion + non-canonical forms
 *   oracle_get_checked_scalar             the compiled reference (only pin
 *                                         for its scalar32_reduce quirk)
 */
#ifndef SPL_ORACLE_H
#define SPL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void oracle_sha256(uint8_t out[32], const uint8_t *msg, size_t len);
void oracle_sha512(uint8_t out[64], const uint8_t *msg, size_t len);

/* hashes[i*32..] <- SHA256^n[i](hashes[i*32..]) for i in [0, num). */
void oracle_poh_verify_many(uint8_t *hashes, const uint64_t *num_hashes,
                            size_t num);

void oracle_sc_reduce(uint8_t s[64]);  /* in place, result in s[0..32) */
void oracle_sc_muladd(uint8_t s[32], const uint8_t a[32], const uint8_t b[32],
                      const uint8_t c[32]);

void oracle_create_keypair(uint8_t pk[32], uint8_t sk[64], const uint8_t seed[32]);
void oracle_sign(uint8_t sig[64], const uint8_t *msg, size_t len,
                 const uint8_t pk[32], const uint8_t sk[64]);
int oracle_verify(const uint8_t sig[64], const uint8_t *msg, uint32_t len,
                  const uint8_t pk[32]);

int oracle_check_packed_ge_small_order(const uint8_t enc[32]);
int oracle_get_checked_scalar(uint8_t out[32], const uint8_t in[32]);

/* Batched forms over the libcuda-crypt.so packet layout (offsets are byte
 * offsets into the concatenated packet buffer). */
void oracle_verify_many(const uint8_t *packets, uint32_t total_signatures,
                        const uint32_t *message_lens,
                        const uint32_t *public_key_offsets,
                        const uint32_t *signature_offsets,
                        const uint32_t *message_start_offsets, uint8_t *out);
void oracle_sign_many(const uint8_t *packets, uint32_t total_signatures,
                      const uint32_t *message_lens,
                      const uint32_t *public_key_offsets,
                      const uint32_t *private_key_offsets,
                      const uint32_t *message_start_offsets, uint8_t *sigs_out);

/* Multi-threaded driver used only for the CPU baseline timing: runs
 * oracle_verify_many / oracle_poh_verify_many split over `threads`
 * POSIX threads (one contiguous range each). */
void oracle_verify_many_mt(const uint8_t *packets, uint32_t total_signatures,
                           const uint32_t *message_lens,
                           const uint32_t *public_key_offsets,
                           const uint32_t *signature_offsets,
                           const uint32_t *message_start_offsets, uint8_t *out,
                           int threads);
void oracle_poh_verify_many_mt(uint8_t *hashes, const uint64_t *num_hashes,
                               size_t num, int threads);

#ifdef __cplusplus
}
#endif
#endif
