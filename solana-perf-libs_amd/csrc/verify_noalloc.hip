// ed25519_verify_many_device_noalloc: the device-resident verify entry point
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// No counterpart in the reference.  It enqueues the same kernels as
// ed25519_verify_many_device, which takes its per-signature tables of point
// multiples from a per-stream buffer the library grows with hipMalloc; here
// they live in `scratch`, one file-scope device array of fixed size, which
// each GPU holds from the moment the library's code is loaded on it until the
// process ends.
//
// The array's name in host code is only the host-side handle of the symbol,
// not its device address.  The address is taken once per device at library
// set-up (noalloc_bind_device, called by the pool's initialisation with that
// device current) and kept in the device's pool record; launches use that
// address and nothing else.
//
// Every call on one GPU shares the array: calls whose launches overlap on the
// device (different streams, or host threads racing) race on it, so callers
// must order them, for example on one stream.  The kernels write every table
// entry they read, so nothing needs clearing between calls.
#include "verify_kernels.h"

namespace spl {

// Two tables of 9 cached points (W = 4, half-size scalars) per signature,
// the most any configuration this entry point runs needs per signature.
constexpr uint32_t kNoAllocMaxSignatures = 16384;
constexpr size_t kNoAllocBytesPerSig = 2 * 9 * sizeof(ge_cached);
constexpr size_t kNoAllocScratchBytes = kNoAllocMaxSignatures * kNoAllocBytesPerSig;
static_assert(sizeof(ge_cached) == 160, "table layout");

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

hipError_t noalloc_bind_device(void **ptr) {
    *ptr = nullptr;
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

}  // namespace spl

using namespace spl;

extern "C" int ed25519_verify_many_device_noalloc(
    const uint8_t *d_packets, uint64_t packets_size, const uint32_t *d_message_lens,
    const uint32_t *d_public_key_offsets, const uint32_t *d_signature_offsets,
    const uint32_t *d_message_start_offsets, uint32_t total_signatures, uint8_t *d_out,
    void *stream) {
    if (total_signatures == 0) return 0;
    if (total_signatures > kNoAllocMaxSignatures) {
        fprintf(stderr,
                "ed25519_verify_many_device_noalloc: %u signatures exceed the fixed scratch "
                "(at most %u per call)\n",
                total_signatures, kNoAllocMaxSignatures);
        return 2;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    const ed25519_tables *tabs = nullptr;
    void *scr = nullptr;
    if (!pool_device_ready(dev, &tabs, &scr)) {
        fprintf(stderr, "ed25519_verify_many_device_noalloc: device %d not set up "
                        "(call ed25519_init() first)\n", dev);
        return 1;
    }
    if (!scr) {
        fprintf(stderr, "ed25519_verify_many_device_noalloc: no scratch address on device %d\n",
                dev);
        return 1;
    }
    VerifyConfig cfg = tuned_plain_config(dev, total_signatures);
    if (verify_scratch_bytes(cfg, total_signatures, 0) > kNoAllocScratchBytes) {
        // the tuned configuration needs wider tables than the array holds:
        // the W = 4 forms fit, chosen by the untuned rule's sizes
        const uint32_t simds = 4u * (uint32_t)device_cus(dev);
        cfg = 16ull * simds >= total_signatures          ? kStagedVerifyConfigW4
              : 32ull * simds >= total_signatures        ? kPairVerifyConfig
              : shared_hash_pays(total_signatures, simds) ? kSharedHashVerifyConfigW4
                                                         : kDefaultVerifyConfig;
    }
    VerifyJob job = make_verify_job(d_packets, packets_size, d_message_lens, d_public_key_offsets,
                                    d_signature_offsets, d_message_start_offsets, d_out, scr,
                                    tabs);
    const hipError_t e = launch_verify_cfg(cfg, job, total_signatures, 0, (hipStream_t)stream);
    if (e != hipSuccess) {
        fprintf(stderr, "ed25519_verify_many_device_noalloc: launch failed: %s\n",
                hipGetErrorString(e));
        return 3;
    }
    return 0;
}
