// Device context pool and host-side plumbing for libcuda-crypt.so on gfx950.
//
// Mirrors the reference's pool semantics (src/cuda-ecc-ed25519/gpu_ctx.cu:6-167,
// src/cuda-poh-verify/poh_verify.cu:29-72): up to 8 devices x 8 queues, picked
// round-robin per call, one mutex per queue held for the whole call, grow-only
// device buffers.  Differences, by design:
//   - errors are reported (stderr + a return code / zeroed mask), never abort;
//   - packets and offset tables are packed into one pinned staging buffer and
//     sent with a single H2D copy instead of one copy per gpu_Elems;
//   - each device holds one copy of the precomputed base-point tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "ed25519_core.h"
#include "../../include/cuda_crypt.h"

namespace spl {

extern bool g_verbose;

#define SPL_LOG(...)                      \
    do {                                  \
        if (::spl::g_verbose) {           \
            fprintf(stderr, __VA_ARGS__); \
        }                                 \
    } while (0)

// Evaluate a HIP call; on failure print it and jump to the caller's `fail`.
#define SPL_HIP_TRY(expr)                                                       \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            fprintf(stderr, "libcuda-crypt: %s failed: %s (%s:%d)\n", #expr,   \
                    hipGetErrorString(e_), __FILE__, __LINE__);                 \
            goto fail;                                                          \
        }                                                                       \
    } while (0)

// Short-lived kernels (verify, sign, copies) raise their waves' priority above
// the default that PoH chain waves keep.  Among the waves sharing a SIMD the
// instruction arbiter serves higher priority first and then the older wave; a
// PoH wave always has a VALU instruction ready, so without this a verify wave
// placed on the same SIMD as a running chain waits until that chain finishes
// (measured: 1.1 ms verify launches stretched to 22-32 ms beside 33 ms chains).
__device__ __forceinline__ void raise_wave_priority() { __builtin_amdgcn_s_setprio(1); }

constexpr int kMaxDevices = 8;
constexpr int kQueuesPerDevice = 8;

// A grow-only buffer; device or pinned host memory.
struct Buf {
    void *ptr = nullptr;
    size_t cap = 0;
    bool host = false;
    hipError_t reserve(size_t bytes);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

struct QueueCtx {
    std::mutex mu;
    hipStream_t stream = nullptr;
    Buf d_packets, d_offsets, d_out, d_scratch, d_hashes, d_counts;
    Buf h_stage;     // pinned host staging
};

struct DeviceCtx {
    int id = -1;
    ed25519_tables *d_tables = nullptr;
    void *noalloc_scratch = nullptr;   // device address of verify_noalloc.hip's array
    QueueCtx q[kQueuesPerDevice];
};

// Host copy of the tables (computed once, on first use).
const ed25519_tables &host_tables();

// Lazily initialises the pool; returns the number of usable devices (0 when
// there is no GPU or HIP fails to initialise).
int pool_init();

// Picks (device, queue) round-robin, locks the queue and makes its device
// current.  Returns nullptr when there is no device.  Release with
// release_queue().
struct Lease {
    DeviceCtx *dev = nullptr;
    QueueCtx *q = nullptr;
};
Lease acquire_queue();
void release_queue(const Lease &l);

// Device copy of the tables for HIP device `hip_device` (nullptr if that
// device is not in the pool).
const ed25519_tables *device_tables(int hip_device);

// Without initialising anything: true when the pool is set up and holds HIP
// device `hip_device`, with its tables and the address of the fixed verify
// scratch (nullptr if that address could not be taken).
bool pool_device_ready(int hip_device, const ed25519_tables **tables, void **noalloc_scratch);

// Address of verify_noalloc.hip's fixed scratch array on the current device
// (called once per device by the pool's initialisation).
hipError_t noalloc_bind_device(void **ptr);

// Device scratch for launches enqueued by the device-resident entry points:
// one grow-only buffer per (device, stream), so launches on one stream reuse
// it in stream order and launches on different streams never share it.
// Returns nullptr on allocation failure.
void *stream_scratch(int hip_device, hipStream_t stream, size_t bytes);

// Copies `bytes` rounded up to 16 from src to dst on stream s with a kernel
// (device or pinned host memory; both 16-byte aligned, both valid for the
// rounded length).  Used instead of DMA copies for host<->device transfers.
hipError_t launch_copy16(void *dst, const void *src, uint64_t bytes, hipStream_t s);

// The device address of the host bytes [p, p + n) when all of them lie in
// one allocation the GPU can read directly: pinned host memory, allocated as
// such or registered (cuda_host_register).  nullptr for anything else.
const uint8_t *pinned_device_range(const uint8_t *p, uint64_t n);

// Up to kGatherMax copies of 16-byte-aligned ranges (device or pinned host
// memory) into one device buffer, done by one kernel launch whose blocks each
// copy one kGatherChunk16-word chunk of one range; the list travels as the
// kernel's argument.  first[k] is range k's first block, first[count] the
// grid size.
constexpr uint32_t kGatherMax = 128;
constexpr uint32_t kGatherChunk16 = 1024;    // 16 KB per block
struct GatherList {
    uint32_t count;
    uint32_t first[kGatherMax + 1];
    uint32_t n16[kGatherMax];
    uint64_t dst_off[kGatherMax];
    const uint4 *src[kGatherMax];
    GatherList() : count(0) { first[0] = 0; }
    bool full() const { return count == kGatherMax; }
    // Appends [src, src + 16 n16) -> d + dst_off (0 < n16 < 2^32, list not
    // full; dst_off and src 16-byte aligned).
    void push(uint64_t off, const void *src_, uint64_t words) {
        dst_off[count] = off;
        src[count] = static_cast<const uint4 *>(src_);
        n16[count] = (uint32_t)words;
        first[count + 1] = first[count] + (uint32_t)((words + kGatherChunk16 - 1) / kGatherChunk16);
        count++;
    }
};
static_assert(sizeof(GatherList) <= 4096, "kernel argument size");
// Enqueues the list's copies into d on s (nothing when it is empty) and
// empties it.
hipError_t launch_gather(uint8_t *d, GatherList &list, hipStream_t s);

// Copies the packets of elems[0..num_elems) (elems[k].num packets of
// `stride` bytes each, concatenated, cut at `total` bytes) into pinned host
// memory h, and sends them on to device memory d with copy kernels on s as
// they land: each time another kStageChunk bytes are in h, that chunk's copy
// kernel is enqueued, so the PCIe transfer of one chunk overlaps the host copy
// of the next.  On return h[0, total) is written (zeros after the elems'
// bytes if they hold fewer), and the first *sent bytes (a multiple of 16) are
// enqueued for d; the caller fills what follows in h and copies [*sent, end)
// itself.  An elem whose bytes are already in pinned host memory (allocated
// pinned, or registered with cuda_host_register) is copied to d straight from
// the caller's pages instead, when its offset in the concatenation, its size
// and its address are multiples of 16.
constexpr uint64_t kStageChunk = 4ull << 20;
hipError_t stage_elems(uint8_t *h, uint8_t *d, const gpu_Elems *elems, uint32_t num_elems,
                       uint32_t stride, uint64_t total, uint64_t *sent, hipStream_t s);

// Compute units of a HIP device (256 when the query fails).
int device_cus(int dev);

// Enqueues SHA-256 chains: hashes[8*i..8*i+8) <- SHA-256^m of itself, in
// place, for i < n, where m is the part of counts[i] that falls in
// [skip, skip + limit) (poh_verify.hip).  The defaults hash whole chains;
// successive launches with skip = 0, limit, 2 limit, ... hash them in slices.
// Few chains take the two-waves-per-64 kernel, many the one-lane-per-chain
// kernel.
hipError_t launch_poh(uint32_t *d_hashes, const uint64_t *d_counts, uint64_t n,
                      hipStream_t stream, uint64_t skip = 0, uint64_t limit = ~0ull);

// Frees every cached device/pinned buffer (ed25519_free_gpu_mem).
void pool_free_buffers();

}  // namespace spl
