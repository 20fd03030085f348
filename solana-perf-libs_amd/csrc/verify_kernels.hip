// Verify kernels: one-shot and persistent (work-queue) forms, window widths 4
// and 5, two register-allocation targets, and the SHA-length bucketing
// prepass; two lanes per signature; and two forms that give a signature's
// stages to different waves of a workgroup (staged: a hash wave and a point
// wave per 32 signatures; shared-hash: three point waves and one hash/comb
// wave per 192 signatures).
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

template <int W, int HALF, bool PF>
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
    job.out[i] = (uint8_t)verify_finish_half<W, PF>(x, y, sp, K, tabs, tabs + kT, *job.T);
}

template <int W, int WPE, int HALF>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
verify_plain_kernel(VerifyJob job, uint32_t n_arg) {
    raise_wave_priority();
    const uint32_t n = job.n_dev ? *job.n_dev : n_arg;
    const uint32_t i = blockIdx.x * 64u + threadIdx.x;
    if (i >= n) return;
    // one wave per SIMD leaves the registers to prefetch table entries
    verify_sig<W, HALF, WPE == 1>(job, i, job.scratch + (size_t)i * lane_tables<W, HALF>());
}

// The value held by lane ^ 1 (ds_swizzle, bit mode: and 0x1f, or 0, xor 1).
// (HIP's generic __shfl_xor here sent the compiler into a >20-minute build.)
__device__ __forceinline__ int32_t pair_swap(int32_t v) {
    return __builtin_amdgcn_ds_swizzle(v, 0x041f);
}

// Two lanes per signature (half-size scalars).  Lane role 0 ("A lane")
// multiplies |x| by +-A, role 1 ("R lane") y by -R; the comb digits of
// sp = y S mod L are split between them (bytes 0..15 and 16..31: the signed
// radix-256 recoding carries only the top bit of byte 15 across the split).
// Each lane ends with its partial sum; the two are exchanged within the pair
// and both lanes test their total for the identity.  Every exit is taken by
// both lanes of a pair together, so the exchange never reads an inactive lane.
template <int W>
__device__ __forceinline__ void verify_sig_pair(const VerifyJob &job, uint32_t i, uint32_t role,
                                                ge_cached *tab) {
    const ed25519_tables &T = *job.T;
    const uint64_t so = job.sig_off[i], po = job.pk_off[i], mo = job.msg_off[i],
                   ml = job.lens[i];
    if (so + 64 > job.packets_size || po + 32 > job.packets_size ||
        mo + ml > job.packets_size) {
        if (!role) job.out[i] = 0;
        return;
    }
    uint32_t sig[16], pk[8];
    load_words(sig, job.packets + so, 16);
    load_words(pk, job.packets + po, 8);
    MsgReader msg{job.packets + mo, (uint32_t)ml};
    // the same checks as verify_prepare_half, evaluated by both lanes
    if ((sig[15] & 0xf0000000u) || !enc_y_canonical(sig) || y_low_order(sig) ||
        y_low_order(pk)) {
        if (!role) job.out[i] = 0;
        return;
    }
    uint32_t pre[16];
#pragma unroll
    for (int k = 0; k < 8; k++) { pre[k] = sig[k]; pre[8 + k] = pk[k]; }
    uint64_t st[8];
    sha512_prefixed(st, pre, 16, msg);
    uint32_t hw[16], h[8];
    sha512_digest_le_words(hw, st);
    sc_reduce512(h, hw);
    ge_p3 P;
    uint32_t enc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) enc[k] = role ? sig[k] : pk[k];
    int dec = ge_decode_neg(P, enc, T.k) == 0;      // -A or -R
    dec &= pair_swap(dec);
    if (!dec) {
        if (!role) job.out[i] = 0;
        return;
    }
    uint32_t x[8], y[8], sp[8], x_neg;
    half_scalars(h, x, x_neg, y);
    const uint32_t zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    sc_muladd(sp, y, sig + 8, zero);
    uint32_t a[8], s[8], below = role ? sp[3] >> 31 : 0u;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        a[k] = role ? y[k] : x[k];
        s[k] = role ? (k < 4 ? sp[k + 4] : 0u) : sp[k];
    }
    if (!role && x_neg) {
        fe_neg(P.X, P.X);
        fe_neg(P.T, P.T);
    }
    constexpr int kT = VarWindow<W>::kTable;
    build_multiples(tab, kT, P, T.k.d2);
    int bits = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (a[k]) bits = 32 * k + bitlen32(a[k]);
    const int K = wave_max7(half_windows<W>(bits));
    shl256(a, 256 - K * W);
    const int mbase = role ? 16 : 0;
    int m = 0;
    ge_p3 pb, u;
    ge_p3_0(pb);
    ge_p1p1 t, tb;
    ge_p2 r;
    ge_p2_0(r);
    for (int j = K - 1; j >= 0; j--) {
        const int d = take_top_digit<W>(a);
        if (m < 16) {
            const int df = take_low_byte_digit(s, below);
            ge_add_precomp(tb, pb, T.comb[mbase + m][df < 0 ? -df : df], df < 0);
            ge_p1p1_to_p3(pb, tb);
            m++;
        }
        if (j != K - 1) {
#pragma unroll
            for (int k = 0; k < W - 1; k++) {
                ge_p2_dbl(t, r);
                ge_p1p1_to_p2(r, t);
            }
            ge_p2_dbl(t, r);
            ge_p1p1_to_p3(u, t);
        } else {
            ge_p3_0(u);
        }
        ge_add_cached(t, u, tab[d < 0 ? -d : d], d < 0);
        if (j != 0) ge_p1p1_to_p2(r, t);
    }
    for (; m < 16; m++) {
        const int df = take_low_byte_digit(s, below);
        ge_add_precomp(tb, pb, T.comb[mbase + m][df < 0 ? -df : df], df < 0);
        ge_p1p1_to_p3(pb, tb);
    }
    // this lane's part: chain + comb half
    ge_cached c;
    ge_p1p1_to_p3(u, t);
    ge_p3_to_cached(c, pb, T.k.d2);
    ge_add_cached(t, u, c, 0);
    ge_p1p1_to_p3(u, t);
    // add the partner's part
    ge_p3 v;
#pragma unroll
    for (int k = 0; k < 10; k++) {
        v.X.v[k] = pair_swap(u.X.v[k]);
        v.Y.v[k] = pair_swap(u.Y.v[k]);
        v.Z.v[k] = pair_swap(u.Z.v[k]);
        v.T.v[k] = pair_swap(u.T.v[k]);
    }
    ge_p3_to_cached(c, v, T.k.d2);
    ge_add_cached(t, u, c, 0);
    ge_p1p1_to_p2(r, t);
    fe dz;
    fe_sub(dz, r.Y, r.Z);
    const uint32_t ok = fe_iszero(r.X) & fe_iszero(dz);
    if (!role) job.out[i] = (uint8_t)ok;
}

template <int W, int WPE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
verify_pair_kernel(VerifyJob job, uint32_t n_arg) {
    raise_wave_priority();
    const uint32_t n = job.n_dev ? *job.n_dev : n_arg;
    const uint32_t g = blockIdx.x * 64u + threadIdx.x;
    const uint32_t i = g >> 1;
    if (i >= n) return;                    // both lanes of the pair
    verify_sig_pair<W>(job, i, g & 1u, job.scratch + (size_t)g * VarWindow<W>::kTable);
}

// Staged form: the two-lane form with the stages that do not need the points
// moved to a second wave.  A workgroup covers 32 signatures with two waves:
//   hash wave  (lanes 0..31, one per signature): the byte-level checks, the
//              SHA-512 of R || A || M, h mod L, the half scalars x, y and
//              sp = y S mod L, then the whole fixed-base sum [sp]B (32 comb
//              additions);
//   point wave (two lanes per signature, as in the two-lane form): decode
//              -A or -R and build its table of multiples, then its doubling
//              chain [|x|](+-A) or [y](-R) with no comb additions.
// The two waves hand over through LDS at two barriers: x, y and the checks
// once the point wave has its tables (it decodes while the hash wave
// hashes), and [sp]B once the chains are done.  A signature's critical path
// becomes max(hash + lattice, decode + table) + chain, against their sum in
// the two-lane form, and no stage runs twice.  Every lane of both waves
// reaches both barriers (lanes past the batch compute on and write nothing,
// rejected signatures compute on and write 0), so the barriers sit in control
// flow that both waves run.
struct StagedShared {
    uint32_t x[8][32], y[8][32];   // |x| and y, by word
    uint32_t flags[32];            // bit 0: passed the byte checks, bit 1: x < 0
    int32_t comb[40][32];          // [sp]B as a ge_cached, by limb
};

template <int W>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(1, 1)))
verify_staged_kernel(VerifyJob job, uint32_t n_arg) {
    raise_wave_priority();
    __shared__ StagedShared sh;
    const ed25519_tables &T = *job.T;
    const uint32_t n = job.n_dev ? *job.n_dev : n_arg;
    // a graph captured for a larger capacity launches workgroups past the
    // batch: both waves of such a group leave here, before either barrier
    if (blockIdx.x * 32u >= n) return;
    const uint32_t lane = threadIdx.x & 63;
    constexpr int kT = VarWindow<W>::kTable;
    if (threadIdx.x < 64) {
        // ---- hash wave
        const uint32_t j = lane, i = blockIdx.x * 32u + lane;
        const bool mine = j < 32 && i < n;
        uint32_t x[8] = {0, 0, 0, 0, 0, 0, 0, 0}, y[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t sp[8] = {0, 0, 0, 0, 0, 0, 0, 0}, x_neg = 0, ok = 0;
        if (mine) {
            const uint64_t so = job.sig_off[i], po = job.pk_off[i], mo = job.msg_off[i],
                           ml = job.lens[i];
            if (so + 64 <= job.packets_size && po + 32 <= job.packets_size &&
                mo + ml <= job.packets_size) {
                uint32_t sig[16], pk[8];
                load_words(sig, job.packets + so, 16);
                load_words(pk, job.packets + po, 8);
                MsgReader msg{job.packets + mo, (uint32_t)ml};
                ok = !(sig[15] & 0xf0000000u) && enc_y_canonical(sig) && !y_low_order(sig) &&
                     !y_low_order(pk);
                uint32_t pre[16];
#pragma unroll
                for (int k = 0; k < 8; k++) { pre[k] = sig[k]; pre[8 + k] = pk[k]; }
                uint64_t st[8];
                sha512_prefixed(st, pre, 16, msg);
                uint32_t hw[16], h[8];
                sha512_digest_le_words(hw, st);
                sc_reduce512(h, hw);
                half_scalars(h, x, x_neg, y);
                const uint32_t zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                sc_muladd(sp, y, sig + 8, zero);
            }
        }
        if (j < 32) {
#pragma unroll
            for (int k = 0; k < 8; k++) { sh.x[k][j] = x[k]; sh.y[k][j] = y[k]; }
            sh.flags[j] = ok | (x_neg << 1);
        }
        __syncthreads();
        // [sp]B from the comb while the point wave runs its chains
        ge_p3 pb;
        ge_p3_0(pb);
        if (j < 32) {
            ge_p1p1 tb;
            uint32_t below = 0;
            for (int m = 0; m < 32; m++) {
                const int df = take_low_byte_digit(sp, below);
                ge_add_precomp(tb, pb, T.comb[m][df < 0 ? -df : df], df < 0);
                ge_p1p1_to_p3(pb, tb);
            }
            ge_cached cb;
            ge_p3_to_cached(cb, pb, T.k.d2);
            const int32_t *cw = reinterpret_cast<const int32_t *>(&cb);
#pragma unroll
            for (int k = 0; k < 40; k++) sh.comb[k][j] = cw[k];
        }
        __syncthreads();
        return;
    }
    // ---- point wave: lanes 2j (-A) and 2j + 1 (-R) of signature j
    const uint32_t j = lane >> 1, role = lane & 1u, i = blockIdx.x * 32u + j;
    const bool valid = i < n;
    // out-of-range lanes read (never write) the first lane's table of the group
    ge_cached *tab = job.scratch + (size_t)(blockIdx.x * 64u + (valid ? lane : 0u)) * kT;
    uint32_t enc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (valid) {
        const uint64_t off = role ? job.sig_off[i] : job.pk_off[i];
        if (off + 32 <= job.packets_size) load_words(enc, job.packets + off, 8);
    }
    ge_p3 P;
    int dec = ge_decode_neg(P, enc, T.k) == 0;      // -A or -R
    dec &= pair_swap(dec);
    if (valid) build_multiples(tab, kT, P, T.k.d2);
    __syncthreads();
    uint32_t a[8];
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = role ? sh.y[k][j] : sh.x[k][j];
    const uint32_t flags = sh.flags[j];
    // the A lane's table holds multiples of -A; for x < 0 it needs +A
    const uint32_t flip = role ? 0u : (flags >> 1) & 1u;
    int bits = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (a[k]) bits = 32 * k + bitlen32(a[k]);
    const int K = wave_max7(half_windows<W>(bits));
    shl256(a, 256 - K * W);
    ge_p3 u;
    ge_p1p1 t;
    ge_p2 r;
    ge_p2_0(r);
    for (int w = K - 1; w >= 0; w--) {
        const int d = take_top_digit<W>(a);
        if (w != K - 1) {
#pragma unroll
            for (int k = 0; k < W - 1; k++) {
                ge_p2_dbl(t, r);
                ge_p1p1_to_p2(r, t);
            }
            ge_p2_dbl(t, r);
            ge_p1p1_to_p3(u, t);
        } else {
            ge_p3_0(u);
        }
        ge_add_cached(t, u, tab[d < 0 ? -d : d], (uint32_t)(d < 0) ^ flip);
        if (w != 0) ge_p1p1_to_p2(r, t);
    }
    ge_p1p1_to_p3(u, t);
    __syncthreads();
    if (!role) {
        // the A lane adds the fixed-base part
        ge_cached cb;
        int32_t *cw = reinterpret_cast<int32_t *>(&cb);
#pragma unroll
        for (int k = 0; k < 40; k++) cw[k] = sh.comb[k][j];
        ge_add_cached(t, u, cb, 0);
        ge_p1p1_to_p3(u, t);
    }
    // add the partner's part
    ge_p3 v;
#pragma unroll
    for (int k = 0; k < 10; k++) {
        v.X.v[k] = pair_swap(u.X.v[k]);
        v.Y.v[k] = pair_swap(u.Y.v[k]);
        v.Z.v[k] = pair_swap(u.Z.v[k]);
        v.T.v[k] = pair_swap(u.T.v[k]);
    }
    ge_cached c;
    ge_p3_to_cached(c, v, T.k.d2);
    ge_add_cached(t, u, c, 0);
    ge_p1p1_to_p2(r, t);
    fe dz;
    fe_sub(dz, r.Y, r.Z);
    const uint32_t id = fe_iszero(r.X) & fe_iszero(dz);
    if (valid && !role) job.out[i] = (uint8_t)(id & (uint32_t)dec & (flags & 1u));
}

// Shared-hash form, for batches too large for the staged form: a workgroup
// covers 192 signatures with four waves, each on a SIMD of its own (one
// workgroup per CU):
//   point waves 0-2 (one lane per signature, 64 each): the byte checks, the
//              lockstep decode of -A and -R and both tables of multiples;
//              then the rest of the SHA-512 after the hash wave's first KH
//              blocks, h mod L, the half scalars and sp = y S mod L; then the
//              doubling chain [|x|](+-A) + [y](-R) with no comb additions;
//   hash wave  (lane j: signatures j, 64 + j and 128 + j of the group): the
//              first KH SHA-512 blocks of its three messages while the point
//              waves decode; then, once they have sp, the three comb sums
//              [sp]B while they run their chains.
// Three barriers order the handovers through LDS (48 KB per group): the
// hash states, sp, the comb sums.  Every lane of every wave reaches all three
// (lanes past the batch compute on and write nothing, rejected signatures
// compute on and write 0), and a group wholly past the batch leaves before
// the first.  A negative x flips the
// sign of each -A table entry as it is added, so the tables need not wait for
// the hash.  KH balances the two: on 41,626 validator-shaped signatures
// (messages of 100-1200 bytes, 2-11 blocks) one launch took 0.771, 0.756,
// 0.753, 0.774 and 0.789 ms with W = 5 and KH = 6, 7, 8, 9 and 11, and
// 0.757 and 0.765 ms with W = 4 (half the table building) and KH = 6 and 7,
// against 0.882 ms for one lane per signature.
template <int W>
SPL_HD constexpr int shared_hash_blocks() { return W == 5 ? 8 : 6; }

struct SharedHashShared {
    uint32_t st[16][192];   // SHA-512 state after KH blocks, as 32-bit halves
    uint32_t sp[8][192];    // y S mod L
    int32_t comb[40][192];  // [sp]B as a ge_cached, by limb
};

// [|x|](P1 or -P1) + [y](Pr) from tables of P1 and Pr multiples, K windows of
// W bits, the same K in every lane of the wave.
template <int W>
__device__ __forceinline__ void verify_chain_half(ge_p3 &u, const uint32_t x[8],
                                                  const uint32_t y[8], int K, const ge_cached *T1,
                                                  const ge_cached *Tr, uint32_t flip) {
    uint32_t a[8], b[8];
#pragma unroll
    for (int k = 0; k < 8; k++) { a[k] = x[k]; b[k] = y[k]; }
    shl256(a, 256 - K * W);
    shl256(b, 256 - K * W);
    ge_p1p1 t;
    ge_p2 r;
    ge_p2_0(r);
    for (int j = K - 1; j >= 0; j--) {
        const int d1 = take_top_digit<W>(a);
        const int d2 = take_top_digit<W>(b);
        if (j != K - 1) {
#pragma unroll
            for (int k = 0; k < W - 1; k++) {
                ge_p2_dbl(t, r);
                ge_p1p1_to_p2(r, t);
            }
            ge_p2_dbl(t, r);
            ge_p1p1_to_p3(u, t);
        } else {
            ge_p3_0(u);
        }
        ge_add_cached(t, u, T1[d1 < 0 ? -d1 : d1], (uint32_t)(d1 < 0) ^ flip);
        ge_p1p1_to_p3(u, t);
        ge_add_cached(t, u, Tr[d2 < 0 ? -d2 : d2], d2 < 0);
        if (j != 0) ge_p1p1_to_p2(r, t);
    }
    ge_p1p1_to_p3(u, t);
}

template <int W, int KH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
verify_shared_hash_kernel(VerifyJob job, uint32_t n_arg) {
    raise_wave_priority();
    __shared__ SharedHashShared sh;
    const ed25519_tables &T = *job.T;
    const uint32_t n = job.n_dev ? *job.n_dev : n_arg;
    const uint32_t base = blockIdx.x * 192u;
    // a graph captured for a larger capacity launches groups past the batch:
    // all four waves of such a group leave here, before any barrier
    if (base >= n) return;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wave == 3) {
        // ---- hash wave: the first KH blocks of three messages
        for (int q = 0; q < 3; q++) {
            const uint32_t j = 64u * q + lane, i = base + j;
            uint64_t st[8];
            sha512_init(st);
            if (i < n) {
                const uint64_t so = job.sig_off[i], po = job.pk_off[i], mo = job.msg_off[i],
                               ml = job.lens[i];
                if (so + 64 <= job.packets_size && po + 32 <= job.packets_size &&
                    mo + ml <= job.packets_size) {
                    uint32_t pre[16];
                    load_words(pre, job.packets + so, 8);
                    load_words(pre + 8, job.packets + po, 8);
                    MsgReader msg{job.packets + mo, (uint32_t)ml};
                    sha512_prefixed_blocks(st, pre, 16, msg, 0u, (uint32_t)KH);
                }
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                sh.st[2 * k][j] = (uint32_t)st[k];
                sh.st[2 * k + 1][j] = (uint32_t)(st[k] >> 32);
            }
        }
        __syncthreads();   // the hash states are out
        __syncthreads();   // the point waves' sp are in
        for (int q = 0; q < 3; q++) {
            const uint32_t j = 64u * q + lane;
            uint32_t sp[8];
#pragma unroll
            for (int k = 0; k < 8; k++) sp[k] = sh.sp[k][j];
            ge_p3 pb;
            ge_p3_0(pb);
            ge_p1p1 tb;
            uint32_t below = 0;
            for (int m = 0; m < 32; m++) {
                const int df = take_low_byte_digit(sp, below);
                ge_add_precomp(tb, pb, T.comb[m][df < 0 ? -df : df], df < 0);
                ge_p1p1_to_p3(pb, tb);
            }
            ge_cached cb;
            ge_p3_to_cached(cb, pb, T.k.d2);
            const int32_t *cw = reinterpret_cast<const int32_t *>(&cb);
#pragma unroll
            for (int k = 0; k < 40; k++) sh.comb[k][j] = cw[k];
        }
        __syncthreads();   // the comb sums are out
        return;
    }
    // ---- point wave `wave`: signature j of the group, one lane each
    constexpr int kT = VarWindow<W>::kTable;
    const uint32_t j = 64u * wave + lane, i = base + j;
    const bool valid = i < n;
    // out-of-range lanes read (never write) the group's first tables
    ge_cached *tabs = job.scratch + (size_t)(valid ? i : base) * lane_tables<W, 1>();
    uint32_t sig[16], pk[8];
#pragma unroll
    for (int k = 0; k < 16; k++) sig[k] = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) pk[k] = 0;
    MsgReader msg{job.packets, 0};
    uint32_t ok = 0;
    bool in_range = false;   // the offsets lie in the packets (the hash wave hashed them)
    if (valid) {
        const uint64_t so = job.sig_off[i], po = job.pk_off[i], mo = job.msg_off[i],
                       ml = job.lens[i];
        if (so + 64 <= job.packets_size && po + 32 <= job.packets_size &&
            mo + ml <= job.packets_size) {
            in_range = true;
            load_words(sig, job.packets + so, 16);
            load_words(pk, job.packets + po, 8);
            msg = MsgReader{job.packets + mo, (uint32_t)ml};
            ok = !(sig[15] & 0xf0000000u) && enc_y_canonical(sig) && !y_low_order(sig) &&
                 !y_low_order(pk);
        }
    }
    {
        ge_p3 pts[2];
        const uint32_t *const enc[2] = {pk, sig};
        uint32_t bad[2];
        ge_decode_neg_n(pts, enc, T.k, bad);
        ok &= !(bad[0] | bad[1]);
        if (valid) {
            build_multiples(tabs, kT, pts[0], T.k.d2);
            build_multiples(tabs + kT, kT, pts[1], T.k.d2);
        }
    }
    __syncthreads();   // the hash states are in
    uint64_t st[8];
#pragma unroll
    for (int k = 0; k < 8; k++)
        st[k] = (uint64_t)sh.st[2 * k][j] | ((uint64_t)sh.st[2 * k + 1][j] << 32);
    if (in_range) {
        uint32_t pre[16];
#pragma unroll
        for (int k = 0; k < 8; k++) { pre[k] = sig[k]; pre[8 + k] = pk[k]; }
        sha512_prefixed_blocks(st, pre, 16, msg, (uint32_t)KH, ~0u);
    }
    uint32_t hw[16], h[8], x[8], y[8], sp[8], x_neg;
    sha512_digest_le_words(hw, st);
    sc_reduce512(h, hw);
    half_scalars(h, x, x_neg, y);
    const uint32_t zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    sc_muladd(sp, y, sig + 8, zero);
#pragma unroll
    for (int k = 0; k < 8; k++) sh.sp[k][j] = sp[k];
    __syncthreads();   // sp is out
    const int K = wave_max7(half_windows<W>(half_bits(x, y)));
    ge_p3 u;
    verify_chain_half<W>(u, x, y, K, tabs, tabs + kT, x_neg);
    __syncthreads();   // the comb sums are in
    ge_cached cb;
    int32_t *cw = reinterpret_cast<int32_t *>(&cb);
#pragma unroll
    for (int k = 0; k < 40; k++) cw[k] = sh.comb[k][j];
    ge_p1p1 t;
    ge_add_cached(t, u, cb, 0);
    ge_p2 r;
    ge_p1p1_to_p2(r, t);
    fe dz;
    fe_sub(dz, r.Y, r.Z);
    const uint32_t id = fe_iszero(r.X) & fe_iszero(dz);
    if (valid) job.out[i] = (uint8_t)(ok & id);
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
        if (k < n) verify_sig<W, HALF, true>(job, job.order ? job.order[k] : k, tabs);
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
    // a half-scalar signature needs two tables, in one lane or spread over two
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

template <int W, int WPE>
static void launch_pair(const VerifyJob &job, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL((verify_pair_kernel<W, WPE>), dim3((n + 31) / 32), dim3(64), 0, s, job, n);
}

template <int W>
static void launch_staged(const VerifyJob &job, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL((verify_staged_kernel<W>), dim3((n + 31) / 32), dim3(128), 0, s, job, n);
}

template <int W>
static void launch_shared_hash(const VerifyJob &job, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL((verify_shared_hash_kernel<W, shared_hash_blocks<W>()>),
                       dim3((n + 191) / 192), dim3(256), 0, s, job, n);
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
    if (!cfg.persistent && cfg.lanes == 4) {
        w5 ? launch_shared_hash<5>(job, n_grid, s) : launch_shared_hash<4>(job, n_grid, s);
        return hipGetLastError();
    }
    if (!cfg.persistent && cfg.lanes == 3) {
        w5 ? launch_staged<5>(job, n_grid, s) : launch_staged<4>(job, n_grid, s);
        return hipGetLastError();
    }
    if (!cfg.persistent && cfg.lanes == 2) {
        const bool wpe2 = cfg.waves_per_eu == 2;
        if (w5) wpe2 ? launch_pair<5, 2>(job, n_grid, s) : launch_pair<5, 1>(job, n_grid, s);
        else    wpe2 ? launch_pair<4, 2>(job, n_grid, s) : launch_pair<4, 1>(job, n_grid, s);
        return hipGetLastError();
    }
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
    if (c.lanes == 4)
        snprintf(buf, len, "one-shot W=%u scalars=half shared-hash", c.window);
    else if (c.lanes == 3)
        snprintf(buf, len, "one-shot W=%u scalars=half staged", c.window);
    else if (c.lanes == 2)
        snprintf(buf, len, "one-shot W=%u scalars=half lanes=2 waves_per_eu=%u", c.window,
                 c.waves_per_eu);
    else if (c.persistent)
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
                    out[k++] = VerifyConfig{0, (uint8_t)w, (uint8_t)wpe, 0, 0, (uint8_t)half, 1};
        for (int w = 4; w <= 5; w++)
            for (int wpc = 4; wpc <= 8; wpc *= 2)
                for (int sort = 0; sort <= 1; sort++)
                    if (k < max)
                        out[k++] = VerifyConfig{1, (uint8_t)w, 1, (uint8_t)sort, (uint16_t)wpc,
                                                (uint8_t)half, 1};
    }
    for (int w = 4; w <= 5; w++)
        for (int wpe = 1; wpe <= 2; wpe++)
            if (k < max) out[k++] = VerifyConfig{0, (uint8_t)w, (uint8_t)wpe, 0, 0, 1, 2};
    for (int w = 4; w <= 5; w++)
        if (k < max) out[k++] = VerifyConfig{0, (uint8_t)w, 1, 0, 0, 1, 3};
    for (int w = 4; w <= 5; w++)
        if (k < max) out[k++] = VerifyConfig{0, (uint8_t)w, 1, 0, 0, 1, 4};
    return k;
}

}  // namespace spl
