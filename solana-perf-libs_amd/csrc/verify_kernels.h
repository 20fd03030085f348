// Verify kernel family and its launch configurations.
#pragma once
#include "runtime.h"

namespace spl {

// One launch configuration of the verify path; the autotuner picks among them.
struct VerifyConfig {
    uint8_t persistent;   // 0: one lane per signature, grid covers the batch;
                          // 1: fixed grid of waves pulling 64-signature chunks
                          //    from a device-side work queue
    uint8_t window;       // variable-base window width W (4 or 5)
    uint8_t waves_per_eu; // register-allocation target (1: up to 512 VGPRs per
                          // lane, 2: up to 256)
    uint8_t sort;         // persistent only: process signatures grouped by
                          // SHA-512 block count, longest first
    uint16_t waves_per_cu;// persistent only: resident waves per CU in the grid
    uint8_t half;         // 0: [h](-A) + [S]B with a 253-bit h, compared by
                          //    encoding; 1: half-size scalars (lattice.h),
                          //    [y S]B + [x](+-A) + [y](-R) = O, about 128-bit x, y
    uint8_t lanes;        // lanes per signature: 1, or 2 (half only, one-shot only):
                          //    one lane runs the x(+-A) chain, its neighbour the
                          //    y(-R) chain, each with half of the comb digits;
                          //    3 (half only, one-shot only): the staged form, the
                          //    two chain lanes in one wave and the hashing, the
                          //    lattice step and the comb in a second wave;
                          //    4 (half only, one-shot only): the
                          //    shared-hash form, three one-lane point waves and
                          //    a wave that starts their hashes and sums their
                          //    combs, 192 signatures per workgroup
};

constexpr VerifyConfig kDefaultVerifyConfig = {0, 4, 1, 0, 0, 1, 1};
constexpr VerifyConfig kPairVerifyConfig = {0, 4, 1, 0, 0, 1, 2};
constexpr VerifyConfig kStagedVerifyConfig = {0, 5, 1, 0, 0, 1, 3};
constexpr VerifyConfig kStagedVerifyConfigW4 = {0, 4, 1, 0, 0, 1, 3};
constexpr VerifyConfig kSharedHashVerifyConfig = {0, 5, 1, 0, 0, 1, 4};
constexpr VerifyConfig kSharedHashVerifyConfigW4 = {0, 4, 1, 0, 0, 1, 4};

// Whether the shared-hash form (48 signatures per SIMD per round, 0.75 ms a
// round) finishes n signatures sooner than one lane per signature (64 per
// SIMD, 0.88 ms a round).
inline bool shared_hash_pays(uint32_t n, uint32_t simds) {
    const uint64_t shared_rounds = (n + 48ull * simds - 1) / (48ull * simds);
    const uint64_t plain_rounds = (n + 64ull * simds - 1) / (64ull * simds);
    return 75 * shared_rounds <= 88 * plain_rounds;
}

// The untuned one-shot choice for n signatures on a device with `simds`
// SIMDs: the staged form while its two waves per 32 signatures get a SIMD
// each, two lanes per signature while those waves do; above that, the
// shared-hash form (four waves per 192 signatures) or one lane per signature,
// whichever the launch times below give fewer milliseconds for in whole
// rounds of the device's SIMDs.  Measured per launch on validator-shaped
// signatures (256 CUs): staged 0.48-0.51 ms, two lanes 0.64-0.66 ms,
// shared-hash 0.75 ms, one lane 0.87-0.88 ms.
inline VerifyConfig untuned_plain_config(uint32_t n, uint32_t simds) {
    if (16ull * simds >= n) return kStagedVerifyConfig;
    if (32ull * simds >= n) return kPairVerifyConfig;
    return shared_hash_pays(n, simds) ? kSharedHashVerifyConfig : kDefaultVerifyConfig;
}

// Device-side description of one verify batch.  All pointers are device
// pointers; n lives in device memory so a captured graph can be replayed for
// any batch size up to the capacity it was built for.
struct VerifyJob {
    const uint8_t *packets;
    uint64_t packets_size;
    const uint32_t *lens, *pk_off, *sig_off, *msg_off;
    uint8_t *out;
    ge_cached *scratch;           // table storage (see verify_scratch_bytes)
    const ed25519_tables *T;
    const uint32_t *n_dev;        // number of signatures
    uint32_t *work;               // [0]: queue head, [1..16]: bucket counts,
                                  // [17..32]: bucket cursors
    uint32_t *order;              // sort: n entries
};

// Bytes of scratch a launch of `cfg` needs for n signatures on this device.
size_t verify_scratch_bytes(const VerifyConfig &cfg, uint32_t n, int num_cus);

// Enqueues one verify of job on stream s.  The signature count is *job.n_dev
// when job.n_dev is set (required for persistent configs: write it with
// launch_set_n first), otherwise n_grid; plain configs launch a grid covering
// n_grid signatures.  Persistent configs need job.work (33 words) and, with
// sorting, job.order (n words).  Persistent configs run with waves_per_eu 1.
hipError_t launch_verify_cfg(const VerifyConfig &cfg, const VerifyJob &job, uint32_t n_grid,
                             int num_cus, hipStream_t s);

// Writes n to *n_dev in stream order.
void launch_set_n(uint32_t *n_dev, uint32_t n, hipStream_t s);

// Configurations the autotuner tries.
int verify_config_candidates(VerifyConfig *out, int max);

// Human-readable name of a configuration.
void config_name(const VerifyConfig &c, char *buf, size_t len);

// The fastest one-shot configuration the autotuner recorded for n's size
// class on HIP device dev; when none was recorded, untuned_plain_config.
VerifyConfig tuned_plain_config(int dev, uint32_t n);

// Compute units of HIP device dev (256 if the query fails).
int device_cus(int dev);

// Builds a VerifyJob for a one-shot launch (n_dev, work and order unset).
VerifyJob make_verify_job(const uint8_t *packets, uint64_t size, const uint32_t *lens,
                          const uint32_t *pk, const uint32_t *sig, const uint32_t *msg,
                          uint8_t *out, void *scratch, const ed25519_tables *T);

}  // namespace spl
