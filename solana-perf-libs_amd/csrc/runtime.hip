// Device context pool, precomputed tables, pinned staging.
#include "runtime.h"

#include <string.h>

#include <vector>

namespace spl {

bool g_verbose = false;

// Copies n16 16-byte words between device memory and pinned host memory (or
// within either).  Host<->device transfers use this kernel rather than DMA
// copies: on the MI355X, hipMemcpyAsync between pinned host memory and the
// device was measured to stall for about 7 ms now and then (the same 1 MiB
// copy otherwise takes 0.1 ms), while this kernel, reading or writing the
// pinned memory over PCIe, took 0.27 ms every time.
__global__ void __launch_bounds__(256) copy16_kernel(uint4 *__restrict__ dst,
                                                     const uint4 *__restrict__ src,
                                                     uint64_t n16) {
    raise_wave_priority();
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += step)
        dst[i] = src[i];
}

hipError_t launch_copy16(void *dst, const void *src, uint64_t bytes, hipStream_t s) {
    const uint64_t n16 = (bytes + 15) / 16;
    if (n16 == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)(n16 < 256ull * 1024 ? (n16 + 255) / 256 : 1024);
    hipLaunchKernelGGL(copy16_kernel, dim3(grid), dim3(256), 0, s, static_cast<uint4 *>(dst),
                       static_cast<const uint4 *>(src), n16);
    return hipGetLastError();
}

// Measured on MI355X: for pinned memory the runtime reports the queried
// address itself as the device address (also inside the allocation) and the
// allocation's start and size; ordinary memory reports type "unregistered"
// and fails the range queries.
const uint8_t *pinned_device_range(const uint8_t *p, uint64_t n) {
    if (n == 0) return nullptr;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (a.type != hipMemoryTypeHost || !a.devicePointer) return nullptr;
    uint64_t start = 0, size = 0;
    if (hipPointerGetAttribute(&start, HIP_POINTER_ATTRIBUTE_RANGE_START_ADDR,
                               (hipDeviceptr_t)p) != hipSuccess ||
        hipPointerGetAttribute(&size, HIP_POINTER_ATTRIBUTE_RANGE_SIZE, (hipDeviceptr_t)p) !=
            hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    const uint64_t lo = (uint64_t)(uintptr_t)p;
    if (lo < start || lo + n > start + size) return nullptr;
    return static_cast<const uint8_t *>(a.devicePointer);
}

// One block per 16 KB chunk of a range, so a large range gets as many blocks
// reading over PCIe at once as it has chunks, and a list of small ranges (a
// validator batch arrives as ~128 packet batches) goes in one launch: one
// launch per range left the copies waiting on each other's launch and PCIe
// round trips.
__global__ void __launch_bounds__(256) gather_copy16_kernel(uint4 *__restrict__ dst,
                                                            GatherList list) {
    raise_wave_priority();
    // the range holding this block's chunk: the last k with first[k] <= block
    // (every range has at least one block, so first[] rises strictly)
    const uint32_t blk = blockIdx.x;
    uint32_t lo = 0, hi = list.count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (list.first[mid] <= blk) lo = mid;
        else hi = mid;
    }
    const uint64_t c0 = (uint64_t)(blk - list.first[lo]) * kGatherChunk16;
    const uint64_t n16 = list.n16[lo];
    const uint64_t end = c0 + kGatherChunk16 < n16 ? c0 + kGatherChunk16 : n16;
    uint4 *out = dst + list.dst_off[lo] / 16;
    const uint4 *in = list.src[lo];
    for (uint64_t i = c0 + threadIdx.x; i < end; i += blockDim.x) out[i] = in[i];
}

hipError_t launch_gather(uint8_t *d, GatherList &list, hipStream_t s) {
    if (list.count == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_copy16_kernel, dim3(list.first[list.count]), dim3(256), 0, s,
                       reinterpret_cast<uint4 *>(d), list);
    list.count = 0;
    return hipGetLastError();
}

hipError_t stage_elems(uint8_t *h, uint8_t *d, const gpu_Elems *elems, uint32_t num_elems,
                       uint32_t stride, uint64_t total, uint64_t *sent, hipStream_t s) {
    uint64_t cur = 0, done = 0;
    GatherList pinned;
    for (uint32_t k = 0; k < num_elems && cur < total; k++) {
        uint64_t b = (uint64_t)elems[k].num * stride;
        if (cur + b > total) b = total - cur;   // never overrun
        const uint8_t *src = elems[k].elems;
        // A batch already in pinned memory is copied to the device by the
        // copy kernel straight from the caller's pages, without passing
        // through the staging buffer (when its place in the concatenation,
        // its size and its address allow 16-byte copies).
        const uint8_t *dsrc = nullptr;
        if (b && cur % 16 == 0 && b % 16 == 0 && (uintptr_t)src % 16 == 0 &&
            b / 16 < (1ull << 32))
            dsrc = pinned_device_range(src, b);
        if ((uintptr_t)dsrc % 16) dsrc = nullptr;
        if (dsrc) {
            // send what is staged so far (its region ends here), then leave
            // this elem's region to the gathered copy
            hipError_t e = hipSuccess;
            if (cur > done) e = launch_copy16(d + done, h + done, cur - done, s);
            if (e == hipSuccess && pinned.full()) e = launch_gather(d, pinned, s);
            if (e != hipSuccess) return e;
            pinned.push(cur, dsrc, b / 16);
            cur += b;
            done = cur;
            continue;
        }
        while (b) {
            // fill up to the next chunk boundary, then send every whole chunk
            const uint64_t room = kStageChunk - (cur - done) % kStageChunk;
            const uint64_t m = b < room ? b : room;
            memcpy(h + cur, src, m);
            cur += m;
            src += m;
            b -= m;
            if (cur - done >= kStageChunk) {
                const uint64_t len = (cur - done) & ~15ull;
                const hipError_t e = launch_copy16(d + done, h + done, len, s);
                if (e != hipSuccess) return e;
                done += len;
            }
        }
    }
    // fewer packets in elems than total_packets says: the rest reads as zeros
    if (cur < total) memset(h + cur, 0, total - cur);
    *sent = done;
    return launch_gather(d, pinned, s);
}

hipError_t Buf::reserve(size_t bytes) {
    if (bytes <= cap && ptr) return hipSuccess;
    release();
    size_t want = bytes < 256 ? 256 : bytes;
    // pinned host memory is coherent: copy kernels write results into it
    hipError_t e = host ? hipHostMalloc(&ptr, want, hipHostMallocCoherent)
                        : hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        // The failure is returned here; it must not stay behind as the
        // thread's last error, which the launch helpers read to detect a
        // failed launch (a later call on this thread would fail with it).
        (void)hipGetLastError();
        ptr = nullptr;
        cap = 0;
        return e;
    }
    cap = want;
    return hipSuccess;
}

void Buf::release() {
    if (ptr) {
        if (host) (void)hipHostFree(ptr);
        else (void)hipFree(ptr);
    }
    ptr = nullptr;
    cap = 0;
}

// ------------------------------------------------------------ tables

static ed25519_tables g_tables;
static std::once_flag g_tables_once;

static void p3_add(ge_p3 &r, const ge_p3 &p, const ge_p3 &q, const fe &d2) {
    ge_cached c;
    ge_p1p1 t;
    ge_p3_to_cached(c, q, d2);
    ge_add_cached(t, p, c, 0);
    ge_p1p1_to_p3(r, t);
}

// Affine (y+x, y-x, 2dxy) forms of n points with one field inversion
// (Montgomery's simultaneous-inversion trick).
static void batch_to_precomp(ge_precomp *out, const ge_p3 *in, int n, const fe &d2) {
    std::vector<fe> prefix(n);
    prefix[0] = in[0].Z;
    for (int i = 1; i < n; i++) fe_mul(prefix[i], prefix[i - 1], in[i].Z);
    fe inv;
    fe_invert(inv, prefix[n - 1]);
    for (int i = n - 1; i >= 0; i--) {
        fe zi;
        if (i > 0) {
            fe_mul(zi, inv, prefix[i - 1]);
            fe_mul(inv, inv, in[i].Z);
        } else {
            zi = inv;
        }
        fe x, y;
        fe_mul(x, in[i].X, zi);
        fe_mul(y, in[i].Y, zi);
        fe_canon(x, x);
        fe_canon(y, y);
        ge_precomp &r = out[i];
        fe_add(r.yplusx, y, x);
        fe_sub(r.yminusx, y, x);
        fe_mul(r.xy2d, x, y);
        fe_mul(r.xy2d, r.xy2d, d2);
        fe_canon(r.yplusx, r.yplusx);
        fe_canon(r.yminusx, r.yminusx);
        fe_canon(r.xy2d, r.xy2d);
    }
}

static void compute_tables(ed25519_tables &T) {
    fe a, b;
    // d = -121665 / 121666
    fe_set(a, 121666);
    fe_invert(b, a);
    fe_set(a, 121665);
    fe_mul(b, b, a);
    fe_neg(T.k.d, b);
    fe_canon(T.k.d, T.k.d);
    fe_add(T.k.d2, T.k.d, T.k.d);
    fe_canon(T.k.d2, T.k.d2);
    // sqrt(-1) = 2^((p-1)/4) = (2^((p-5)/8))^2 * 2
    fe_set(a, 2);
    fe_pow22523(b, a);
    fe_sq(b, b);
    fe_mul(T.k.sqrtm1, b, a);
    fe_canon(T.k.sqrtm1, T.k.sqrtm1);
    // B: y = 4/5 with x even (encoding 0x58 0x66 ... 0x66)
    uint32_t enc[8];
    for (int i = 0; i < 8; i++) enc[i] = 0x66666666u;
    enc[0] = 0x66666658u;
    ge_p3 B;
    ge_decode_neg(B, enc, T.k);
    fe_neg(B.X, B.X);
    fe_neg(B.T, B.T);
    // comb[m][d] = d 256^m B for d = 1..128; comb[m][0] = identity
    std::vector<ge_p3> pts(32 * 128);
    ge_p3 row = B;
    ge_p1p1 t;
    for (int m = 0; m < 32; m++) {
        ge_p3 acc = row;
        for (int d = 1; d <= 128; d++) {
            pts[m * 128 + d - 1] = acc;
            p3_add(acc, acc, row, T.k.d2);
        }
        for (int k = 0; k < 8; k++) {
            ge_p3_dbl(t, row);
            ge_p1p1_to_p3(row, t);
        }
    }
    std::vector<ge_precomp> pre(32 * 128);
    batch_to_precomp(pre.data(), pts.data(), 32 * 128, T.k.d2);
    for (int m = 0; m < 32; m++) {
        fe_set(T.comb[m][0].yplusx, 1);
        fe_set(T.comb[m][0].yminusx, 1);
        fe_set(T.comb[m][0].xy2d, 0);
        for (int d = 1; d <= 128; d++) T.comb[m][d] = pre[m * 128 + d - 1];
        // signing table: (j+1) 256^m B, j = 0..7
        for (int j = 0; j < 8; j++) T.base[m][j] = T.comb[m][j + 1];
    }
}

const ed25519_tables &host_tables() {
    std::call_once(g_tables_once, [] { compute_tables(g_tables); });
    return g_tables;
}

// ------------------------------------------------------------ pool

static std::mutex g_pool_mu;
static int g_num_devices = -1;
static DeviceCtx g_dev[kMaxDevices];
static uint32_t g_next_dev = 0;
static uint32_t g_next_q[kMaxDevices] = {0};

static int pool_init_locked() {
    if (g_num_devices >= 0) return g_num_devices;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        g_num_devices = 0;
        SPL_LOG("libcuda-crypt: no HIP device\n");
        return 0;
    }
    if (n > kMaxDevices) n = kMaxDevices;
    const ed25519_tables &T = host_tables();
    int ok = 0;
    for (int d = 0; d < n; d++) {
        DeviceCtx &dc = g_dev[ok];
        dc.id = d;
        SPL_HIP_TRY(hipSetDevice(d));
        SPL_HIP_TRY(hipMalloc(&dc.d_tables, sizeof(ed25519_tables)));
        SPL_HIP_TRY(hipMemcpy(dc.d_tables, &T, sizeof(ed25519_tables), hipMemcpyHostToDevice));
        {
            // the fixed scratch of the allocation-free verify entry point;
            // the device stays usable for every other entry point without it
            const hipError_t e = noalloc_bind_device(&dc.noalloc_scratch);
            if (e != hipSuccess) {
                fprintf(stderr, "libcuda-crypt: device %d: no address for the fixed verify "
                                "scratch: %s\n", d, hipGetErrorString(e));
                dc.noalloc_scratch = nullptr;
                (void)hipGetLastError();
            }
        }
        for (int q = 0; q < kQueuesPerDevice; q++) {
            SPL_HIP_TRY(hipStreamCreate(&dc.q[q].stream));
            dc.q[q].h_stage.host = true;
        }
        ok++;
        continue;
    fail:
        fprintf(stderr, "libcuda-crypt: device %d unusable, skipped\n", d);
    }
    g_num_devices = ok;
    SPL_LOG("libcuda-crypt: %d device(s)\n", ok);
    return ok;
}

int pool_init() {
    std::lock_guard<std::mutex> g(g_pool_mu);
    return pool_init_locked();
}

Lease acquire_queue() {
    Lease l;
    {
        std::lock_guard<std::mutex> g(g_pool_mu);
        int n = pool_init_locked();
        if (n <= 0) return l;
        uint32_t d = g_next_dev;
        g_next_dev = (g_next_dev + 1) % (uint32_t)n;
        uint32_t q = g_next_q[d];
        g_next_q[d] = (g_next_q[d] + 1) % kQueuesPerDevice;
        l.dev = &g_dev[d];
        l.q = &g_dev[d].q[q];
    }
    l.q->mu.lock();
    if (hipSetDevice(l.dev->id) != hipSuccess) {
        fprintf(stderr, "libcuda-crypt: hipSetDevice(%d) failed\n", l.dev->id);
        l.q->mu.unlock();
        return Lease{};
    }
    return l;
}

void release_queue(const Lease &l) {
    if (l.q) l.q->mu.unlock();
}

const ed25519_tables *device_tables(int hip_device) {
    std::lock_guard<std::mutex> g(g_pool_mu);
    for (int d = 0; d < g_num_devices; d++)
        if (g_dev[d].id == hip_device) return g_dev[d].d_tables;
    return nullptr;
}

bool pool_device_ready(int hip_device, const ed25519_tables **tables, void **noalloc_scratch) {
    std::lock_guard<std::mutex> g(g_pool_mu);
    for (int d = 0; d < g_num_devices; d++)
        if (g_dev[d].id == hip_device) {
            *tables = g_dev[d].d_tables;
            *noalloc_scratch = g_dev[d].noalloc_scratch;
            return true;
        }
    return false;
}

struct StreamScratch {
    int dev;
    hipStream_t stream;
    Buf buf;
};
static std::mutex g_scratch_mu;
static std::vector<StreamScratch *> g_scratch;

void *stream_scratch(int hip_device, hipStream_t stream, size_t bytes) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    StreamScratch *e = nullptr;
    for (StreamScratch *x : g_scratch)
        if (x->dev == hip_device && x->stream == stream) e = x;
    if (!e) {
        e = new StreamScratch{hip_device, stream, Buf{}};
        g_scratch.push_back(e);
    }
    if (e->buf.reserve(bytes) != hipSuccess) return nullptr;
    return e->buf.ptr;
}

void pool_free_buffers() {
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        for (StreamScratch *x : g_scratch) {
            if (hipSetDevice(x->dev) == hipSuccess) {
                (void)hipDeviceSynchronize();
                x->buf.release();
            }
        }
    }
    std::lock_guard<std::mutex> g(g_pool_mu);
    if (g_num_devices <= 0) return;
    for (int d = 0; d < g_num_devices; d++) {
        if (hipSetDevice(g_dev[d].id) != hipSuccess) continue;
        for (int q = 0; q < kQueuesPerDevice; q++) {
            QueueCtx &c = g_dev[d].q[q];
            std::lock_guard<std::mutex> qg(c.mu);
            (void)hipStreamSynchronize(c.stream);
            c.d_packets.release();
            c.d_offsets.release();
            c.d_out.release();
            c.d_scratch.release();
            c.d_hashes.release();
            c.d_counts.release();
            c.h_stage.release();
        }
    }
}

}  // namespace spl
