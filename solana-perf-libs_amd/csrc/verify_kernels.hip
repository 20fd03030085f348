// Verify kernels: one-shot and persistent (work-queue) forms, window widths 4
// and 5, two register-allocation targets, and the SHA-length bucketing
// prepass.
//
// Reference: src/cuda-ecc-ed25519/verify.cu:176-232 launches three kernels of
// 64 threads, one thread per signature.  Here a single fused kernel checks a
// signature per lane.  The persistent form keeps a fixed grid of waves that
// pull 64-signature chunks from a device-side queue head (an atomic counter),
// so one launch (or one captured graph) serves any batch size, and it can walk
// the signatures grouped by SHA-512 block count so the lanes of a wave hash
// messages of similar length.
#include "verify_kernels.h"

namespace spl {

constexpr int kBuckets = 16;

SPL_HD uint32_t sha_bucket(uint32_t len) {
    // SHA-512 blocks for R || A || M: ceil((64 + len + 17) / 128)
    uint32_t b = (81u + len + 127u) / 128u;
    return b < kBuckets - 1 ? b : kBuckets - 1;
}

// Maximum of v over the active lanes of the wave, 0 <= v < 128.
__device__ __forceinline__ int wave_max7(int v) {
    int r = 0;
#pragma unroll
    for (int b = 6; b >= 0; b--) {
        const int c = r | (1 << b);
        if (__ballot(v >= c) != 0ull) r = c;
    }
    return r;
}

// Cached points of scratch per lane.
template <int W, int HALF>
SPL_HD constexpr int lane_tables() { return (HALF ? 2 : 1) * VarWindow<W>::kTable; }

template <int W, int HALF>
__device__ __forceinline__ void verify_sig(const VerifyJob &job, uint32_t i, ge_cached *tabs) {
    const uint64_t so = job.sig_off[i], po = job.pk_off[i], mo = job.msg_off[i],
                   ml = job.lens[i];
    if (so + 64 > job.packets_size || po + 32 > job.packets_size ||
        mo + ml > job.packets_size) {
        job.out[i] = 0;   // offsets outside the packet buffer: reject
        return;
    }
    uint32_t sig[16], pk[8];
    load_words(sig, job.packets + so, 16);
    load_words(pk, job.packets + po, 8);
    MsgReader msg{job.packets + mo, (uint32_t)ml};
    if (!HALF) {
        job.out[i] = (uint8_t)verify_one<W>(sig, pk, msg, *job.T, tabs);
        return;
    }
    uint32_t x[8], y[8], sp[8];
    ge_p3 P1, R;
    if (!verify_prepare_half(sig, pk, msg, *job.T, x, y, sp, P1, R)) {
        job.out[i] = 0;
        return;
    }
    constexpr int kT = VarWindow<W>::kTable;
    build_multiples(tabs, kT, P1, job.T->k.d2);
    build_multiples(tabs + kT, kT, R, job.T->k.d2);
    // the same window count in every lane still running keeps the adds at
    // the same bit positions across the wave
    const int K = wave_max7(half_windows<W>(half_bits(x, y)));
    job.out[i] = (uint8_t)verify_finish_half<W>(x, y, sp, K, tabs, tabs + kT, *job.T);
}

template <int W, int WPE, int HALF>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
verify_plain_kernel(VerifyJob job, uint32_t n_arg) {
    raise_wave_priority();
    const uint32_t n = job.n_dev ? *job.n_dev : n_arg;
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    verify_sig<W, HALF>(job, i, job.scratch + (size_t)i * lane_tables<W, HALF>());
}

template <int W, int WPE, int HALF>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
verify_persistent_kernel(VerifyJob job) {
    raise_wave_priority();
    const uint32_t n = *job.n_dev;
    const uint32_t lane = threadIdx.x;
    ge_cached *tabs = job.scratch + ((size_t)blockIdx.x * 64u + lane) * lane_tables<W, HALF>();
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(job.work, 64u);
        base = __builtin_amdgcn_readfirstlane(base);
        if (base >= n) break;   // wave-uniform: every wave exits once the queue is drained
        const uint32_t k = base + lane;
        if (k < n) verify_sig<W, HALF>(job, job.order ? job.order[k] : k, tabs);
    }
}

__global__ void verify_set_n_kernel(uint32_t *n_dev, uint32_t n) {
    if (threadIdx.x == 0) *n_dev = n;
}

// Clears the queue head and the bucket counters/cursors (a kernel, not a
// memset, so it can sit inside a captured graph).
__global__ void verify_reset_kernel(uint32_t *work) {
    if (threadIdx.x < 1 + 2 * kBuckets) work[threadIdx.x] = 0;
}

constexpr int kPrepassBlocks = 128, kPrepassThreads = 256;

__global__ void __launch_bounds__(kPrepassThreads)
verify_bucket_count_kernel(const uint32_t *lens, const uint32_t *n_dev, uint32_t *work) {
    __shared__ uint32_t hist[kBuckets];
    if (threadIdx.x < kBuckets) hist[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t n = *n_dev;
    for (uint32_t i = blockIdx.x * kPrepassThreads + threadIdx.x; i < n;
         i += kPrepassBlocks * kPrepassThreads)
        atomicAdd(&hist[sha_bucket(lens[i])], 1u);
    __syncthreads();
    if (threadIdx.x < kBuckets && hist[threadIdx.x]) atomicAdd(&work[1 + threadIdx.x], hist[threadIdx.x]);
}

// Bucket cursors: longest messages first.
__global__ void verify_bucket_scan_kernel(uint32_t *work) {
    if (threadIdx.x == 0) {
        uint32_t pos = 0;
        for (int b = kBuckets - 1; b >= 0; b--) {
            work[1 + kBuckets + b] = pos;
            pos += work[1 + b];
        }
    }
}

__global__ void __launch_bounds__(kPrepassThreads)
verify_bucket_scatter_kernel(const uint32_t *lens, const uint32_t *n_dev, uint32_t *work,
                             uint32_t *order) {
    __shared__ uint32_t hist[kBuckets], base[kBuckets];
    const uint32_t n = *n_dev;
    const uint32_t stride = kPrepassBlocks * kPrepassThreads;
    for (uint32_t start = blockIdx.x * kPrepassThreads; start < n; start += stride) {
        if (threadIdx.x < kBuckets) hist[threadIdx.x] = 0;
        __syncthreads();
        const uint32_t i = start + threadIdx.x;
        uint32_t b = 0, local = 0;
        if (i < n) {
            b = sha_bucket(lens[i]);
            local = atomicAdd(&hist[b], 1u);
        }
        __syncthreads();
        if (threadIdx.x < kBuckets && hist[threadIdx.x])
            base[threadIdx.x] = atomicAdd(&work[1 + kBuckets + threadIdx.x], hist[threadIdx.x]);
        __syncthreads();
        if (i < n) order[base[b] + local] = i;
        __syncthreads();
    }
}

template <int W>
static size_t table_bytes() {
    return (size_t)VarWindow<W>::kTable * sizeof(ge_cached);
}

size_t verify_scratch_bytes(const VerifyConfig &cfg, uint32_t n, int num_cus) {
    const size_t per = (cfg.window == 5 ? table_bytes<5>() : table_bytes<4>()) * (cfg.half ? 2 : 1);
    const size_t lanes = cfg.persistent ? (size_t)num_cus * cfg.waves_per_cu * 64u : n;
    return per * (lanes ? lanes : 1);
}

template <int W, int WPE, int HALF>
static void launch_plain(const VerifyJob &job, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL((verify_plain_kernel<W, WPE, HALF>), dim3((n + 63) / 64), dim3(64), 0, s,
                       job, n);
}

template <int W, int HALF>
static void launch_persistent(const VerifyJob &job, uint32_t waves, hipStream_t s) {
    hipLaunchKernelGGL((verify_persistent_kernel<W, 1, HALF>), dim3(waves), dim3(64), 0, s, job);
}

template <int HALF>
static void launch_plain_cfg(const VerifyConfig &cfg, const VerifyJob &job, uint32_t n,
                             hipStream_t s) {
    const bool w5 = cfg.window == 5, wpe2 = cfg.waves_per_eu == 2;
    if (w5) wpe2 ? launch_plain<5, 2, HALF>(job, n, s) : launch_plain<5, 1, HALF>(job, n, s);
    else    wpe2 ? launch_plain<4, 2, HALF>(job, n, s) : launch_plain<4, 1, HALF>(job, n, s);
}

void launch_set_n(uint32_t *n_dev, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(verify_set_n_kernel, dim3(1), dim3(64), 0, s, n_dev, n);
}

hipError_t launch_verify_cfg(const VerifyConfig &cfg, const VerifyJob &job, uint32_t n_grid,
                             int num_cus, hipStream_t s) {
    if (n_grid == 0) return hipSuccess;
    const bool w5 = cfg.window == 5;
    if (!cfg.persistent) {
        if (cfg.half) launch_plain_cfg<1>(cfg, job, n_grid, s);
        else          launch_plain_cfg<0>(cfg, job, n_grid, s);
        return hipGetLastError();
    }
    if (!job.work || !job.n_dev || (cfg.sort && !job.order)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(verify_reset_kernel, dim3(1), dim3(64), 0, s, job.work);
    VerifyJob j = job;
    if (cfg.sort) {
        hipLaunchKernelGGL(verify_bucket_count_kernel, dim3(kPrepassBlocks), dim3(kPrepassThreads),
                           0, s, job.lens, job.n_dev, job.work);
        hipLaunchKernelGGL(verify_bucket_scan_kernel, dim3(1), dim3(64), 0, s, job.work);
        hipLaunchKernelGGL(verify_bucket_scatter_kernel, dim3(kPrepassBlocks),
                           dim3(kPrepassThreads), 0, s, job.lens, job.n_dev, job.work, job.order);
    } else {
        j.order = nullptr;
    }
    const uint32_t waves = (uint32_t)num_cus * cfg.waves_per_cu;
    if (cfg.half) w5 ? launch_persistent<5, 1>(j, waves, s) : launch_persistent<4, 1>(j, waves, s);
    else          w5 ? launch_persistent<5, 0>(j, waves, s) : launch_persistent<4, 0>(j, waves, s);
    return hipGetLastError();
}

void config_name(const VerifyConfig &c, char *buf, size_t len) {
    if (c.persistent)
        snprintf(buf, len, "persistent W=%u scalars=%s sort=%u waves_per_cu=%u", c.window,
                 c.half ? "half" : "full", c.sort, c.waves_per_cu);
    else
        snprintf(buf, len, "one-shot W=%u scalars=%s waves_per_eu=%u", c.window,
                 c.half ? "half" : "full", c.waves_per_eu);
}

int verify_config_candidates(VerifyConfig *out, int max) {
    int k = 0;
    for (int half = 0; half <= 1; half++) {
        for (int w = 4; w <= 5; w++)
            for (int wpe = 1; wpe <= 2; wpe++)
                if (k < max)
                    out[k++] = VerifyConfig{0, (uint8_t)w, (uint8_t)wpe, 0, 0, (uint8_t)half};
        for (int w = 4; w <= 5; w++)
            for (int wpc = 4; wpc <= 8; wpc *= 2)
                for (int sort = 0; sort <= 1; sort++)
                    if (k < max)
                        out[k++] = VerifyConfig{1, (uint8_t)w, 1, (uint8_t)sort, (uint16_t)wpc,
                                                (uint8_t)half};
    }
    return k;
}

}  // namespace spl
