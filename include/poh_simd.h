/*
 * libpoh-simd.so — C ABI of the CPU multi-buffer PoH library (x86-64 host,
 * no GPU).
 *
 * Drop-in for the reference's ISPC-built libpoh-simd.so
 * (src/poh-simd/poh-verify.ispc:3-8, src/poh-simd/Makefile:15-40), which
 * solana_entry binds for GPU-less verification of PoH entries.  Each entry point
 * advances exactly the number of chains named in its comment (the ISPC gang
 * width of that target in the reference's Makefile): chain i's 32-byte state
 * is hashes[32*i .. 32*i+32), and it is replaced in place by SHA-256 applied
 * num_hashes_arr[i] times to it (poh-verify.ispc:23-177).  A count of 0
 * leaves those 32 bytes untouched (poh-verify.ispc:23).
 *
 * The suffix fixes only the number of chains.  Every entry point runs the
 * fastest kernel this CPU supports (SHA extensions, AVX-512, AVX2 or portable
 * C++), chosen on the first call; the results are the same bytes whichever
 * runs.  These five calls are thread-safe and allocate nothing.
 */
#ifndef SPL_POH_SIMD_H
#define SPL_POH_SIMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* poh-verify.ispc:3-8 with --target=sse2-i32x4 (Makefile:18-20): 4 chains. */
void poh_verify_many_simd_sse2(uint8_t *hashes, const uint64_t *num_hashes_arr);

/* poh-verify.ispc:3-8 with --target=sse4-i32x4 (Makefile:22-24): 4 chains. */
void poh_verify_many_simd_sse4(uint8_t *hashes, const uint64_t *num_hashes_arr);

/* poh-verify.ispc:3-8 with --target=avx1-i32x8 (Makefile:26-28): 8 chains. */
void poh_verify_many_simd_avx1(uint8_t *hashes, const uint64_t *num_hashes_arr);

/* poh-verify.ispc:3-8 with --target=avx2-i32x8 (Makefile:30-32): 8 chains. */
void poh_verify_many_simd_avx2(uint8_t *hashes, const uint64_t *num_hashes_arr);

/* poh-verify.ispc:3-8 with --target=avx512skx-i32x16 (Makefile:34-36):
 * 16 chains. */
void poh_verify_many_simd_avx512skx(uint8_t *hashes, const uint64_t *num_hashes_arr);

/* ------------------------------------------------------------------------
 * Extensions (no counterpart in the reference).
 */

/* Name of the kernel the entry points above run: "sha-ni", "avx512", "avx2"
 * or "portable". */
const char *poh_simd_kernel(void);

/* Selects the kernel by name for every later call in this process.  Returns
 * 0, or -1 when the name is unknown or this CPU lacks the instructions (the
 * selection is then unchanged). */
int poh_simd_use_kernel(const char *name);

/* Advances num_chains chains (any count, same layout and count rule as
 * above) on up to `threads` threads (0 = one per CPU this process may run
 * on), in groups of 16 chains.  For callers that hold a whole batch. */
void poh_verify_many_simd(uint8_t *hashes, const uint64_t *num_hashes_arr, uint64_t num_chains,
                          uint32_t threads);

#ifdef __cplusplus
}
#endif

#endif
