/*
 * libcuda-crypt.so — C ABI of the MI355X (gfx950) build.
 *
 * Drop-in for the exports of solana-perf-libs' libcuda-crypt.so that
 * solana_perf::perf_libs binds.  Every declaration below cites the reference
 * interface it replaces (paths relative to the reference tree).  The packet
 * layout is unchanged: packets are the concatenation of
 * elems[i].elems[0 .. elems[i].num * message_size), and every offset is a
 * byte offset into that concatenation.
 *
 * Entry points after "Extensions" have no counterpart in the reference.
 */
#ifndef SPL_CUDA_CRYPT_H
#define SPL_CUDA_CRYPT_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#define SHA512_SIZE 64
#define PUB_KEY_SIZE 32
#define PRIV_KEY_SIZE 64
#define SEED_SIZE 32
#define SCALAR_SIZE 32
#define SIG_SIZE 64

/* src/cuda-ecc-ed25519/ed25519.h:35-38 */
typedef struct {
    uint8_t *elems;
    uint32_t num;
} gpu_Elems;

/* ed25519.h:25, seed.cu:12-41 — 32 bytes from /dev/urandom; 0 = ok. */
int ed25519_create_seed(unsigned char *seed);

/* ed25519.h:40, keypair.cu:6-16 — private key is the 64-byte expanded key
 * clamp(SHA-512(seed)[0:32]) || SHA-512(seed)[32:64]. */
void ed25519_create_keypair(unsigned char *public_key, unsigned char *private_key,
                            const unsigned char *seed);

/* ed25519.h:41, sign.cu:40-46 — host, single signature. */
void ed25519_sign(unsigned char *signature, const unsigned char *message, size_t message_len,
                  const unsigned char *public_key, const unsigned char *private_key);

/* ed25519.h:43-53, sign.cu:76-145 — GPU batch sign; signatures_out[64*i].
 * private_key_offsets point at 64-byte expanded private keys in the packets. */
void ed25519_sign_many(const gpu_Elems *elems, uint32_t num_elems, uint32_t message_size,
                       uint32_t total_packets, uint32_t total_signatures,
                       const uint32_t *message_lens, const uint32_t *public_key_offsets,
                       const uint32_t *private_key_offsets, const uint32_t *message_start_offsets,
                       uint8_t *signatures_out, uint8_t use_non_default_stream);

/* ed25519.h:55, verify.cu:154-174 — host, single signature; 1 = valid. */
int ed25519_verify(const unsigned char *signature, const unsigned char *message,
                   uint32_t message_len, const unsigned char *public_key);

/* ed25519.h:57-67, verify.cu:240-340 — GPU batch verify; out[i] = 1 valid,
 * 0 invalid.  Synchronous.  total_packets == 0 returns without touching out.
 * (The reference header names the 8th parameter private_key_offsets; its
 * definition, like this one, uses it as the signature offsets.)  On a device
 * error every out[i] is written 0 and the error is printed; the call never
 * aborts the process. */
void ed25519_verify_many(const gpu_Elems *elems, uint32_t num_elems, uint32_t message_size,
                         uint32_t total_packets, uint32_t total_signatures,
                         const uint32_t *message_lens, const uint32_t *public_key_offsets,
                         const uint32_t *signature_offsets, const uint32_t *message_start_offsets,
                         uint8_t *out, uint8_t use_non_default_stream);

/* ed25519.h:71, verify.cu:234-238 */
void ed25519_set_verbose(bool val);

/* ed25519.h:73, verify.cu:342-347 */
const char *ed25519_license(void);

/* ed25519.h:74, gpu_ctx.cu:38-44 — true iff at least one GPU is usable. */
bool ed25519_init(void);

/* ed25519.h:76-77, verify.cu:349-355 — hipHostRegister/hipHostUnregister;
 * returns the runtime error code, 0 = success. */
int cuda_host_register(void *ptr, size_t size, unsigned int flags);
int cuda_host_unregister(void *ptr);

/* ed25519.h:79, verify.cu:60-85 — 0 when the scalar passes the reference's
 * canonicity gate (top nibble clear: copied to out_scalar). */
int ed25519_get_checked_scalar(unsigned char *out_scalar, const unsigned char *in_scalar);

/* ed25519.h:81, verify.cu:88-107 — 0 when the encoding decodes and is not of
 * order <= 4; 1 otherwise. */
int ed25519_check_packed_ge_small_order(const unsigned char *packed_group_element);

/* gpu_ctx.h:39, gpu_ctx.cu:151-167 — releases the cached device buffers. */
void ed25519_free_gpu_mem(void);

/* src/cuda-poh-verify/poh_verify.cu:76-78 */
void poh_verify_many_set_verbose(bool val);

/* poh_verify.cu:80-144 — hashes[32*i..] <- SHA-256^num_hashes_arr[i], in place.
 * Returns 0 on success, 1 when no GPU is available (hashes untouched). */
int poh_verify_many(uint8_t *hashes, const uint64_t *num_hashes_arr, size_t num_elems,
                    uint8_t use_non_default_stream);

/* ------------------------------------------------------------------------
 * Extensions: device-resident forms (no counterpart in the reference).
 * All pointers are device pointers on the current HIP device; `stream` is a
 * hipStream_t (NULL = null stream).  The calls only enqueue work.
 * Return 0 on success, 1 when no GPU is usable, 2 on bad arguments, 3 on a
 * launch failure. */
int ed25519_verify_many_device(const uint8_t *d_packets, uint64_t packets_size,
                               const uint32_t *d_message_lens, const uint32_t *d_public_key_offsets,
                               const uint32_t *d_signature_offsets,
                               const uint32_t *d_message_start_offsets, uint32_t total_signatures,
                               uint8_t *d_out, void *stream);

int ed25519_sign_many_device(const uint8_t *d_packets, uint64_t packets_size,
                             const uint32_t *d_message_lens, const uint32_t *d_public_key_offsets,
                             const uint32_t *d_private_key_offsets,
                             const uint32_t *d_message_start_offsets, uint32_t total_signatures,
                             uint8_t *d_signatures_out, void *stream);

int poh_verify_many_device(uint8_t *d_hashes, const uint64_t *d_num_hashes, size_t num_elems,
                           void *stream);

/* No counterpart in the reference.  poh_verify_many_device with the kernel
 * named instead of chosen by chain count, for tests and tuning: 0 = the
 * automatic choice, 1 = one chain per lane, 2 = rounds and message schedule
 * in two waves, 3 = as 2 with two lanes per chain in both waves.  Every kernel
 * gives the same end states.  Returns as poh_verify_many_device; 2 also for an
 * unknown kernel. */
int spl_poh_device_kernel(uint8_t *d_hashes, const uint64_t *d_num_hashes, size_t num_elems,
                          int kernel, void *stream);

/* No counterpart in the reference.  The kernel (numbered as above, 1-3) that
 * poh_verify_many and poh_verify_many_device use for num_elems chains on the
 * current device; 0 when no GPU is usable. */
int spl_poh_kernel_choice(size_t num_elems);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ed25519_verify_many_device with the same arguments,
 * kernels and results, whose per-signature table scratch is one fixed device
 * array of 45 MiB per GPU (held from library load to process end) instead of
 * a buffer the library allocates.  At most 16,384 signatures per call
 * (more: returns 2, nothing enqueued).  Makes no allocating call; requires
 * ed25519_init() first (else returns 1).  Every call on one GPU shares the
 * array, so calls whose work overlaps on the device race: order them, e.g.
 * on one stream. */
int ed25519_verify_many_device_noalloc(const uint8_t *d_packets, uint64_t packets_size,
                                       const uint32_t *d_message_lens,
                                       const uint32_t *d_public_key_offsets,
                                       const uint32_t *d_signature_offsets,
                                       const uint32_t *d_message_start_offsets,
                                       uint32_t total_signatures, uint8_t *d_out, void *stream);

/* ------------------------------------------------------------------------
 * Extension: verify engine (no counterpart in the reference).
 * A verify engine owns device buffers for up to max_signatures signatures and
 * max_packet_bytes of packets, a HIP stream, and a hipGraph of the tuned
 * launch sequence (a persistent kernel fed by a device-side work queue, or a
 * one-shot kernel; both read the signature count from device memory, so one
 * captured graph serves every batch size up to the capacity).  The caller
 * writes packets and the four offset tables into the buffers returned by
 * spl_verify_engine_buffers (same layout as ed25519_verify_many_device),
 * calls spl_verify_engine_run, then spl_verify_engine_sync and reads the mask.
 * spl_verify_engine_autotune times every launch configuration (one-shot or
 * persistent, window width 4/5, full 253-bit or half-size ~128-bit scalars,
 * register target, length bucketing, waves per CU) on the batch currently in
 * the buffers, keeps the fastest and records it
 * for that batch-size class on the device, where the ABI calls pick it up.
 * An engine is used by one thread at a time.  Functions returning int: 0 ok,
 * 2 bad arguments, 3 device error.
 *
 * Batches in flight: spl_verify_engine_create_pipelined makes an engine of
 * `depth` (1..8) buffer sets, each with its own buffers, stream and graph, so
 * consecutive batches overlap on the device (one validator-sized launch
 * leaves a third of the SIMDs idle; the next set's launch fills them).
 * spl_verify_engine_next waits until the next set in round-robin order has
 * finished its previous batch and returns its index; the caller fills that
 * set's buffers (spl_verify_engine_set_buffers), starts it with
 * spl_verify_engine_run_set, and reads its mask after
 * spl_verify_engine_sync_set (or the next time next() hands the set out).
 * The single-set calls (buffers, run) address set 0, and sync waits for
 * every set.  Autotuning an engine of several sets times each configuration
 * with every set in flight (load the batch into each set first) and records
 * nothing for the ABI calls. */
typedef struct spl_verify_engine spl_verify_engine;
spl_verify_engine *spl_verify_engine_create(uint32_t max_signatures, uint64_t max_packet_bytes);
spl_verify_engine *spl_verify_engine_create_pipelined(uint32_t max_signatures,
                                                      uint64_t max_packet_bytes, uint32_t depth);
void spl_verify_engine_destroy(spl_verify_engine *engine);
int spl_verify_engine_depth(spl_verify_engine *engine);
int spl_verify_engine_set_buffers(spl_verify_engine *engine, uint32_t set, uint8_t **d_packets,
                                  uint32_t **d_lens, uint32_t **d_pk_off, uint32_t **d_sig_off,
                                  uint32_t **d_msg_off, uint8_t **d_out, void **stream);
int spl_verify_engine_next(spl_verify_engine *engine);
int spl_verify_engine_run_set(spl_verify_engine *engine, uint32_t set, uint32_t n);
int spl_verify_engine_sync_set(spl_verify_engine *engine, uint32_t set);
int spl_verify_engine_buffers(spl_verify_engine *engine, uint8_t **d_packets, uint32_t **d_lens,
                              uint32_t **d_public_key_offsets, uint32_t **d_signature_offsets,
                              uint32_t **d_message_start_offsets, uint8_t **d_out, void **stream);
int spl_verify_engine_run(spl_verify_engine *engine, uint32_t num_signatures);
int spl_verify_engine_sync(spl_verify_engine *engine);
int spl_verify_engine_autotune(spl_verify_engine *engine, uint32_t num_signatures, char *report,
                               size_t report_len);
int spl_verify_engine_config(spl_verify_engine *engine, char *buf, size_t len);
/* The autotuner's candidate configurations: returns their number and writes
 * their names, one per line, to buf (if not NULL).  spl_verify_engine_use
 * makes candidate `index` the engine's configuration (recapturing its graph)
 * without timing anything. */
int spl_verify_engine_candidates(spl_verify_engine *engine, char *buf, size_t len);
int spl_verify_engine_use(spl_verify_engine *engine, int index);

/* Tuning records (no counterpart in the reference).  The autotuner of a
 * one-set engine records, per device and batch-size class (bit length of
 * n), the fastest configuration and the fastest one-shot configuration; the
 * ABI and device entry points use the latter.  spl_verify_tune_save writes
 * every record to a text file (one line per record: device id, architecture,
 * CU count, class, both configurations) and returns how many it wrote, < 0
 * on a file error.  spl_verify_tune_load takes the records of such a file,
 * so a restarted process need not tune again; records made on a device of
 * another architecture or CU count than the same device id here are
 * skipped.  It returns the number taken, or < 0 (nothing taken) when the
 * file cannot be read or holds a malformed line or an unknown
 * configuration.  spl_verify_tune_query writes the name of the one-shot
 * configuration used for n signatures on `device` and returns 1 when it
 * comes from a record, 0 for the untuned rule. */
int spl_verify_tune_save(const char *path);
int spl_verify_tune_load(const char *path);
int spl_verify_tune_query(int device, uint32_t n, char *buf, size_t len);

/* ------------------------------------------------------------------------
 * Extension: serving loop (no counterpart in the reference, whose callers
 * make one synchronous ed25519_verify_many / poh_verify_many call per batch).
 * A server owns a device ring of hbm_budget_bytes (0: 1 GiB, at most 4 GiB)
 * with a pinned host mirror, verify_streams + poh_streams streams (0: 3 and 1;
 * at most 16 together), and one loop thread.  Requests are submitted from any
 * thread and return a ticket (> 0) at once; the loop admits pending requests
 * in arrival order into ring space as it frees up and, whenever a stream is
 * idle, launches every admitted request of that kind as one batch (up to
 * max_batch_signatures signatures or chains; 0: 65,536), so batches on
 * different streams overlap on the device.  The streams are plain ones and
 * share the process's hardware queues; PoH batches are hashed 256 hashes per
 * chain per launch, one launch in flight per stream, so a verify batch
 * queued behind PoH work waits for at most one such launch.
 *
 * spl_server_submit_verify takes one request in the ed25519_verify_many
 * layout, with offsets relative to `packets` (packets_size bytes); a signature
 * whose offsets fall outside its own packets is rejected (0).  `out` receives
 * num_signatures mask bytes.  Packets in pinned host memory (allocated
 * pinned or registered, with a 16-byte-aligned address and size) are read by
 * the device in place; others are copied by the loop.  spl_server_submit_poh
 * hashes `hashes` in place like poh_verify_many.  Every buffer passed must
 * stay valid and unchanged until spl_server_wait reports the ticket
 * complete.  Tickets < 0: -2 bad arguments or a request larger than the ring
 * or max_batch_signatures.
 *
 * spl_server_wait blocks up to timeout_ms (< 0: no limit) and returns 0 when
 * the results are written, 1 on time-out (the ticket stays valid), 2 for an
 * unknown ticket, 3 when a device error failed the batch (verify masks are
 * written 0, PoH hashes are left unchanged).  On 0 and 3 the ticket is
 * retired and times_ns (if not NULL) receives submit, admit and completion
 * times from CLOCK_MONOTONIC.  spl_server_stats writes a JSON object with
 * per-kind request counts (for verify, also how many were read in place) and
 * p50/p99/max latency (submit to completion), and the device operation the
 * loop thread is inside at that moment with how long it has been there
 * ("loop_in", "loop_in_ms"), and optionally resets the counts.
 * spl_server_destroy finishes every submitted request first. */
typedef struct spl_server spl_server;
spl_server *spl_server_create(uint64_t hbm_budget_bytes, uint32_t max_batch_signatures,
                              uint32_t verify_streams, uint32_t poh_streams);
void spl_server_destroy(spl_server *server);
int64_t spl_server_submit_verify(spl_server *server, const uint8_t *packets, uint64_t packets_size,
                                 const uint32_t *message_lens, const uint32_t *public_key_offsets,
                                 const uint32_t *signature_offsets,
                                 const uint32_t *message_start_offsets, uint32_t num_signatures,
                                 uint8_t *out);
int64_t spl_server_submit_poh(spl_server *server, uint8_t *hashes, const uint64_t *num_hashes_arr,
                              size_t num_elems);
int spl_server_wait(spl_server *server, int64_t ticket, int timeout_ms, uint64_t *times_ns);
int spl_server_stats(spl_server *server, char *buf, size_t len, int reset);

/* ------------------------------------------------------------------------
 * Extension: packets in, verdicts out (no counterpart in the reference, which
 * consumes the four offset tables, gpu_ctx.cu:133-136, and leaves building
 * them to its caller).  The caller's host pass over Solana wire-format packets
 * (agave perf/src/sigverify.rs: generate_offsets / do_get_packet_offsets,
 * and copy_return_values afterwards) runs on the device instead.
 *
 * Packets are num_packets slots of packet_stride bytes (< 4 GiB in all);
 * packet i holds packet_sizes[i] valid bytes.  A packet gets its signatures
 * when: size >= 65 and size <= packet_stride; a short-vec signature count n
 * (u16 in 1-3 bytes, shortest form) is followed by n x 64 signature bytes and
 * at least one more byte; an optional 0x80 prefix (message version 0; other
 * versions fail) precedes a 3-byte header whose first byte equals n and is
 * above its second; a short-vec account-key count k >= n follows, and the k x
 * 32 key bytes fit in the packet.  Signature j of such a packet is checked
 * against key j over the message from its first byte (the prefix, if any) to
 * the end of the packet.  Otherwise the packet has no signatures.
 *
 * spl_packet_offsets_device writes, per packet, its signature count and the
 * row of its first signature (d_sig_lens, d_sig_base), and one row per
 * signature of the four tables ed25519_verify_many_device takes (at most
 * max_signatures rows; further rows are not written); d_counts[0] receives
 * the batch's signature count and d_counts[1] the rows written.
 * spl_packet_verdicts_device writes d_packet_ok[i] = 1 when packet i has at
 * least one signature and all of them verified in d_sig_out (num_signatures
 * entries), else 0.  spl_verify_packets_device does both with the verify in
 * between, on one stream with no host synchronisation (d_sig_lens may be
 * NULL), and spl_verify_packets is its synchronous host form, taking the
 * packets as gpu_Elems like ed25519_verify_many.  All return 0 ok, 1 no GPU
 * (spl_verify_packets then writes 0 verdicts), 2 bad arguments, 3 device
 * error. */
int spl_packet_offsets_device(const uint8_t *d_packets, uint32_t num_packets,
                              uint32_t packet_stride, const uint32_t *d_packet_sizes,
                              uint32_t max_signatures, uint32_t *d_sig_lens,
                              uint32_t *d_sig_base, uint32_t *d_message_lens,
                              uint32_t *d_public_key_offsets, uint32_t *d_signature_offsets,
                              uint32_t *d_message_start_offsets, uint32_t *d_counts,
                              void *stream);
int spl_packet_verdicts_device(const uint32_t *d_sig_lens, const uint32_t *d_sig_base,
                               uint32_t num_packets, const uint8_t *d_sig_out,
                               uint32_t num_signatures, uint8_t *d_packet_ok, void *stream);
int spl_verify_packets_device(const uint8_t *d_packets, uint32_t num_packets,
                              uint32_t packet_stride, const uint32_t *d_packet_sizes,
                              uint8_t *d_packet_ok, uint32_t *d_sig_lens, void *stream);
int spl_verify_packets(const gpu_Elems *elems, uint32_t num_elems, uint32_t packet_stride,
                       uint32_t total_packets, const uint32_t *packet_sizes, uint8_t *packet_ok,
                       uint32_t *sig_lens);

#ifdef __cplusplus
}

/* poh_verify.cu:66-71 — sets up the device pool; true when at least one GPU
 * is usable.  C++ linkage, as in the reference (the symbol is _Z8poh_initv).
 * Calling it is optional: every entry point sets the pool up on first use.
 *
 * Not exported: get_gpu_ctx, setup_gpu_ctx and release_gpu_ctx
 * (gpu_ctx.h:12-54).  They hand out the reference's internal pool records,
 * whose layout holds a cudaStream_t, ge_cached tables and a pthread mutex,
 * and no caller outside that library uses them.  The pool here is internal;
 * ed25519_init and ed25519_free_gpu_mem are its public controls. */
bool poh_init();
#endif

#endif /* SPL_CUDA_CRYPT_H */
