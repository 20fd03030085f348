// libpoh-simd.so: PoH chains on the host CPU, several chains per core.
//
// Replaces the reference's ISPC library (src/poh-simd/poh-verify.ispc:3-179,
// built by src/poh-simd/Makefile:6-40 for five x86 targets).  The recurrence
// is the one poh_verify_many runs on the GPU: a chain's 32-byte state x is
// replaced by SHA-256(x), count times.  A 32-byte message is one padded block
// whose words 8..15 are fixed (0x80000000, six zeros, 0x100 = 256 bits), and
// the digest words of one step are the message words of the next, so every
// kernel here keeps a chain in native-endian words from the first load to the
// last store.
//
// Kernels (one is picked per process; all give the same bytes):
//   sha-ni    SHA extensions (SHA256RNDS2/MSG1/MSG2), kShaLanes chains
//             interleaved so one chain's round latency hides behind the
//             others'
//   avx512    16 chains, one per 32-bit lane of a zmm register (the layout
//             the reference's avx512skx-i32x16 target uses)
//   avx2      8 chains in ymm lanes
//   portable  scalar C++, one chain at a time
// The lane kernels run a group of chains for the smallest remaining count at
// a time and blend only the lanes still running, so a count of 0 leaves its
// chain's bytes untouched (poh-verify.ispc:23,172-177).
#include <immintrin.h>
#include <sched.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "../../include/poh_simd.h"

namespace {

// FIPS 180-4 section 4.2.2 and 5.3.3.
alignas(64) constexpr uint32_t kK[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u,
    0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu,
    0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu,
    0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u,
    0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu,
    0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
    0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u,
    0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u,
    0xc67178f2u};
constexpr uint32_t kIV[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                             0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
// Words 8..15 of every PoH block: the 0x80 pad byte and the 256-bit length.
constexpr uint32_t kPad[8] = {0x80000000u, 0, 0, 0, 0, 0, 0, 256};

inline uint32_t load_be(const uint8_t *p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}
inline void store_be(uint8_t *p, uint32_t v) {
    p[0] = (uint8_t)(v >> 24); p[1] = (uint8_t)(v >> 16); p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
}

// ---------------------------------------------------------------------------
// portable
inline uint32_t rotr(uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }

void step_portable(uint32_t s[8]) {
    uint32_t w[64];
    for (int i = 0; i < 8; ++i) { w[i] = s[i]; w[8 + i] = kPad[i]; }
    for (int t = 16; t < 64; ++t) {
        uint32_t a = w[t - 15], b = w[t - 2];
        w[t] = w[t - 16] + (rotr(a, 7) ^ rotr(a, 18) ^ (a >> 3)) + w[t - 7] +
               (rotr(b, 17) ^ rotr(b, 19) ^ (b >> 10));
    }
    uint32_t a = kIV[0], b = kIV[1], c = kIV[2], d = kIV[3];
    uint32_t e = kIV[4], f = kIV[5], g = kIV[6], h = kIV[7];
    for (int t = 0; t < 64; ++t) {
        uint32_t t1 = h + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) +
                      kK[t] + w[t];
        uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    s[0] = a + kIV[0]; s[1] = b + kIV[1]; s[2] = c + kIV[2]; s[3] = d + kIV[3];
    s[4] = e + kIV[4]; s[5] = f + kIV[5]; s[6] = g + kIV[6]; s[7] = h + kIV[7];
}

void chains_portable(uint8_t *hashes, const uint64_t *counts, int lanes) {
    for (int i = 0; i < lanes; ++i) {
        if (counts[i] == 0) continue;
        uint32_t s[8];
        for (int k = 0; k < 8; ++k) s[k] = load_be(hashes + 32 * i + 4 * k);
        for (uint64_t j = 0; j < counts[i]; ++j) step_portable(s);
        for (int k = 0; k < 8; ++k) store_be(hashes + 32 * i + 4 * k, s[k]);
    }
}

// ---------------------------------------------------------------------------
// sha-ni.  State registers hold (F,E,B,A) and (H,G,D,C) in lanes 0..3, the
// order SHA256RNDS2 takes; message registers hold W[4i..4i+3] in lanes 0..3.
#define SPL_SHA __attribute__((target("sha,sse4.1")))

constexpr int kShaLanes = 2;

// Two rounds per SHA256RNDS2: the second takes the upper half of wk.  The
// old (A,B,E,F) becomes the new (C,D,G,H), so the two registers swap roles.
#define SHA_4ROUNDS(s0, s1, wk)                                   \
    do {                                                          \
        s1 = _mm_sha256rnds2_epu32(s1, s0, wk);                   \
        s0 = _mm_sha256rnds2_epu32(s0, s1, _mm_shuffle_epi32(wk, 0x0E)); \
    } while (0)

template <int N>
SPL_SHA void sha_steps(__m128i (&lo)[N], __m128i (&hi)[N], uint64_t n) {
    const __m128i iv0 = _mm_set_epi32((int)kIV[0], (int)kIV[1], (int)kIV[4], (int)kIV[5]);
    const __m128i iv1 = _mm_set_epi32((int)kIV[2], (int)kIV[3], (int)kIV[6], (int)kIV[7]);
    const __m128i *kv = reinterpret_cast<const __m128i *>(kK);
    // rounds 8..15 see only the fixed pad words
    const __m128i wk2 = _mm_add_epi32(_mm_set_epi32(0, 0, 0, (int)kPad[0]), kv[2]);
    const __m128i wk3 = _mm_add_epi32(_mm_set_epi32((int)kPad[7], 0, 0, 0), kv[3]);
    const __m128i pad2 = _mm_set_epi32(0, 0, 0, (int)kPad[0]);
    const __m128i pad3 = _mm_set_epi32((int)kPad[7], 0, 0, 0);
    for (uint64_t j = 0; j < n; ++j) {
        __m128i s0[N], s1[N], m0[N], m1[N], m2[N], m3[N];
#pragma GCC unroll 4
        for (int c = 0; c < N; ++c) {
            s0[c] = iv0; s1[c] = iv1;
            m0[c] = lo[c]; m1[c] = hi[c]; m2[c] = pad2; m3[c] = pad3;
            SHA_4ROUNDS(s0[c], s1[c], _mm_add_epi32(m0[c], kv[0]));
        }
#pragma GCC unroll 4
        for (int c = 0; c < N; ++c) SHA_4ROUNDS(s0[c], s1[c], _mm_add_epi32(m1[c], kv[1]));
#pragma GCC unroll 4
        for (int c = 0; c < N; ++c) SHA_4ROUNDS(s0[c], s1[c], wk2);
#pragma GCC unroll 4
        for (int c = 0; c < N; ++c) SHA_4ROUNDS(s0[c], s1[c], wk3);
        // rounds 16..63: W[t..t+3] from W[t-16..t-1], four words at a time;
        // the registers rotate m0 <- m1 <- m2 <- m3 <- new.
#pragma GCC unroll 12
        for (int q = 4; q < 16; ++q) {
#pragma GCC unroll 4
            for (int c = 0; c < N; ++c) {
                __m128i x = _mm_sha256msg1_epu32(m0[c], m1[c]);
                x = _mm_add_epi32(x, _mm_alignr_epi8(m3[c], m2[c], 4));
                x = _mm_sha256msg2_epu32(x, m3[c]);
                m0[c] = m1[c]; m1[c] = m2[c]; m2[c] = m3[c]; m3[c] = x;
                SHA_4ROUNDS(s0[c], s1[c], _mm_add_epi32(x, kv[q]));
            }
        }
#pragma GCC unroll 4
        for (int c = 0; c < N; ++c) {
            __m128i a = _mm_add_epi32(s0[c], iv0);   // F E B A
            __m128i b = _mm_add_epi32(s1[c], iv1);   // H G D C
            lo[c] = _mm_shuffle_epi32(_mm_unpackhi_epi64(b, a), 0x1B);   // A B C D
            hi[c] = _mm_shuffle_epi32(_mm_unpacklo_epi64(b, a), 0x1B);   // E F G H
        }
    }
}

SPL_SHA void chains_sha(uint8_t *hashes, const uint64_t *counts, int lanes) {
    const __m128i bswap = _mm_set_epi8(12, 13, 14, 15, 8, 9, 10, 11, 4, 5, 6, 7, 0, 1, 2, 3);
    // Chains are taken kShaLanes at a time in order of count, so each group
    // runs its shortest chain's count interleaved and the rest alone.
    int order[16];
    for (int i = 0; i < lanes; ++i) order[i] = i;
    std::sort(order, order + lanes, [&](int a, int b) { return counts[a] < counts[b]; });
    int i0 = 0;
    while (i0 < lanes && counts[order[i0]] == 0) ++i0;
    for (; i0 < lanes; i0 += kShaLanes) {
        const int g = std::min(kShaLanes, lanes - i0);
        __m128i lo[kShaLanes], hi[kShaLanes];
        for (int c = 0; c < kShaLanes; ++c) {
            const uint8_t *p = hashes + 32 * order[i0 + std::min(c, g - 1)];
            lo[c] = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i *)p), bswap);
            hi[c] = _mm_shuffle_epi8(_mm_loadu_si128((const __m128i *)(p + 16)), bswap);
        }
        const uint64_t first = counts[order[i0]];
        if (g == kShaLanes) sha_steps<kShaLanes>(lo, hi, first);
        for (int c = 0; c < g; ++c) {
            __m128i l1[1] = {lo[c]}, h1[1] = {hi[c]};
            sha_steps<1>(l1, h1, counts[order[i0 + c]] - (g == kShaLanes ? first : 0));
            uint8_t *p = hashes + 32 * order[i0 + c];
            _mm_storeu_si128((__m128i *)p, _mm_shuffle_epi8(l1[0], bswap));
            _mm_storeu_si128((__m128i *)(p + 16), _mm_shuffle_epi8(h1[0], bswap));
        }
    }
}

// ---------------------------------------------------------------------------
// Lane kernels: one chain per 32-bit lane.  poh_lanes.inc holds the kernel
// over the lane operations of the namespace it is included into, compiled
// for that namespace's instruction set.
}  // namespace

#pragma GCC push_options
#pragma GCC target("avx512f,avx512bw")
namespace spl_poh_avx512 {
typedef __m512i T;
constexpr int L = 16;
static inline T set1(uint32_t x) { return _mm512_set1_epi32((int)x); }
static inline T add(T a, T b) { return _mm512_add_epi32(a, b); }
static inline T ror(T a, int r) { return _mm512_ror_epi32(a, r); }
static inline T shr(T a, int r) { return _mm512_srli_epi32(a, r); }
static inline T xor3(T a, T b, T c) { return _mm512_ternarylogic_epi32(a, b, c, 0x96); }
static inline T ch(T e, T f, T g) { return _mm512_ternarylogic_epi32(e, f, g, 0xCA); }
static inline T maj(T a, T b, T c) { return _mm512_ternarylogic_epi32(a, b, c, 0xE8); }
struct Live {
    __mmask16 m;
    explicit Live(uint32_t live) : m((__mmask16)live) {}
    T blend(T old, T x) const { return _mm512_mask_mov_epi32(old, m, x); }
};
#include "poh_lanes.inc"
}  // namespace spl_poh_avx512
#pragma GCC pop_options

#pragma GCC push_options
#pragma GCC target("avx2")
namespace spl_poh_avx2 {
typedef __m256i T;
constexpr int L = 8;
static inline T set1(uint32_t x) { return _mm256_set1_epi32((int)x); }
static inline T add(T a, T b) { return _mm256_add_epi32(a, b); }
static inline T ror(T a, int r) {
    return _mm256_or_si256(_mm256_srli_epi32(a, r), _mm256_slli_epi32(a, 32 - r));
}
static inline T shr(T a, int r) { return _mm256_srli_epi32(a, r); }
static inline T xor3(T a, T b, T c) { return _mm256_xor_si256(_mm256_xor_si256(a, b), c); }
static inline T ch(T e, T f, T g) {
    return _mm256_xor_si256(_mm256_and_si256(e, _mm256_xor_si256(f, g)), g);
}
static inline T maj(T a, T b, T c) {
    return _mm256_or_si256(_mm256_and_si256(a, b), _mm256_and_si256(c, _mm256_or_si256(a, b)));
}
struct Live {
    T m;
    explicit Live(uint32_t live)
        : m(_mm256_cmpgt_epi32(_mm256_and_si256(_mm256_set1_epi32((int)live),
                                                _mm256_setr_epi32(1, 2, 4, 8, 16, 32, 64, 128)),
                               _mm256_setzero_si256())) {}
    T blend(T old, T x) const { return _mm256_blendv_epi8(old, x, m); }
};
#include "poh_lanes.inc"
}  // namespace spl_poh_avx2
#pragma GCC pop_options

namespace {

// ---------------------------------------------------------------------------
// Selection
typedef void (*ChainsFn)(uint8_t *, const uint64_t *, int);

struct Kernel {
    const char *name;
    ChainsFn fn;
    bool (*supported)();
};

bool has_sha() {
    __builtin_cpu_init();
    unsigned a, b, c, d;
    if (!__builtin_cpu_supports("sse4.1") || !__builtin_cpu_supports("ssse3")) return false;
    __asm__("cpuid" : "=a"(a), "=b"(b), "=c"(c), "=d"(d) : "a"(7), "c"(0));
    return (b >> 29) & 1;   // CPUID.(EAX=7,ECX=0):EBX.SHA[bit 29]
}
bool has_avx512() {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
}
bool has_avx2() {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx2");
}
bool always() { return true; }

const Kernel kKernels[] = {
    {"sha-ni", chains_sha, has_sha},
    {"avx512", spl_poh_avx512::chains, has_avx512},
    {"avx2", spl_poh_avx2::chains, has_avx2},
    {"portable", chains_portable, always},
};
constexpr int kNumKernels = sizeof(kKernels) / sizeof(kKernels[0]);

// Fastest supported kernel by a short timing: 16 chains of 64 steps each.
int calibrate() {
    int best = kNumKernels - 1;
    double best_t = 1e30;
    for (int k = 0; k < kNumKernels; ++k) {
        if (!kKernels[k].supported()) continue;
        uint8_t h[16 * 32];
        uint64_t n[16];
        for (int i = 0; i < 16 * 32; ++i) h[i] = (uint8_t)i;
        double t = 1e30;
        for (int rep = 0; rep < 3; ++rep) {
            std::fill(n, n + 16, 64);
            auto t0 = std::chrono::steady_clock::now();
            kKernels[k].fn(h, n, 16);
            t = std::min(t, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                                .count());
        }
        if (t < best_t) { best_t = t; best = k; }
    }
    return best;
}

std::atomic<int> g_kernel{-1};

const Kernel &kernel() {
    int k = g_kernel.load(std::memory_order_acquire);
    if (k < 0) {
        int c = calibrate();
        int expected = -1;
        g_kernel.compare_exchange_strong(expected, c, std::memory_order_acq_rel);
        k = g_kernel.load(std::memory_order_acquire);
    }
    return kKernels[k];
}

int allowed_cpus() {
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) return std::max(1, CPU_COUNT(&set));
    return (int)std::max(1u, std::thread::hardware_concurrency());
}

}  // namespace

extern "C" {

void poh_verify_many_simd_sse2(uint8_t *hashes, const uint64_t *num_hashes_arr) {
    kernel().fn(hashes, num_hashes_arr, 4);
}
void poh_verify_many_simd_sse4(uint8_t *hashes, const uint64_t *num_hashes_arr) {
    kernel().fn(hashes, num_hashes_arr, 4);
}
void poh_verify_many_simd_avx1(uint8_t *hashes, const uint64_t *num_hashes_arr) {
    kernel().fn(hashes, num_hashes_arr, 8);
}
void poh_verify_many_simd_avx2(uint8_t *hashes, const uint64_t *num_hashes_arr) {
    kernel().fn(hashes, num_hashes_arr, 8);
}
void poh_verify_many_simd_avx512skx(uint8_t *hashes, const uint64_t *num_hashes_arr) {
    kernel().fn(hashes, num_hashes_arr, 16);
}

const char *poh_simd_kernel(void) { return kernel().name; }

int poh_simd_use_kernel(const char *name) {
    if (!name) return -1;
    for (int k = 0; k < kNumKernels; ++k)
        if (strcmp(name, kKernels[k].name) == 0) {
            if (!kKernels[k].supported()) return -1;
            g_kernel.store(k, std::memory_order_release);
            return 0;
        }
    return -1;
}

void poh_verify_many_simd(uint8_t *hashes, const uint64_t *num_hashes_arr, uint64_t num_chains,
                          uint32_t threads) {
    if (num_chains == 0) return;
    const ChainsFn fn = kernel().fn;
    const uint64_t groups = (num_chains + 15) / 16;
    uint64_t nt = threads ? threads : (uint64_t)allowed_cpus();
    nt = std::min(nt, groups);
    std::atomic<uint64_t> next{0};
    auto work = [&] {
        for (uint64_t g; (g = next.fetch_add(1, std::memory_order_relaxed)) < groups;) {
            const uint64_t c0 = 16 * g;
            fn(hashes + 32 * c0, num_hashes_arr + c0, (int)std::min<uint64_t>(16, num_chains - c0));
        }
    };
    std::vector<std::thread> pool;
    for (uint64_t t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
}

}  // extern "C"
