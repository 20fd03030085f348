// CPU check of the half-size-scalar verify (lattice.h, verify_half_one in
// ed25519_core.h), built by tests/test_half_scalars.py with the library's own
// headers and linked against libcuda-crypt.so for the point tables.
//
//   half_verify_check lattice N SEED  -> prints "h x x_neg y" (hex, big-endian)
//                                        for N scalars h (random and edge values)
//   half_verify_check field N SEED    -> limbs of f, g, f g, f^2, 2 f^2 for N
//                                        pairs of loose field elements
//   half_verify_check verify N SEED   -> for N signatures (valid, corrupted, with
//                                        small-order and mixed-torsion keys and R),
//                                        compares verify_half_one<4>/<5> with
//                                        verify_one<4>; prints counts, exits 1 on
//                                        any difference
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../solana-perf-libs_amd/csrc/ed25519_core.h"

namespace spl {
const ed25519_tables &host_tables();
}
using namespace spl;

static uint64_t g_state;
static uint64_t next_u64() {   // splitmix64
    uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static void rand_words(uint32_t *w, int n) {
    for (int i = 0; i < n; i++) w[i] = (uint32_t)next_u64();
}
static void print_hex(const uint32_t *w, int n) {
    for (int i = n - 1; i >= 0; i--) printf("%08x", w[i]);
}

static int run_lattice(int n) {
    const uint32_t L[8] = SPL_L_WORDS;
    for (int k = 0; k < n; k++) {
        uint32_t h[8], x[8], y[8], xn;
        if (k < 40) {
            // edge values: 0..15, L-1-j, 2^j, 2^128 +- 1, 2^127.5-ish
            memset(h, 0, sizeof h);
            if (k < 16) h[0] = (uint32_t)k;
            else if (k < 24) { memcpy(h, L, sizeof h); h[0] -= (uint32_t)(k - 15); }
            else if (k < 32) { int b = (k - 24) * 31 + 7; h[b >> 5] = 1u << (b & 31); }
            else if (k < 36) { h[4] = 1; if (k & 1) h[0] = 1; else { h[4] = 0; for (int i = 0; i < 4; i++) h[i] = 0xffffffffu; } }
            else { h[3] = 0xb504f333u; h[2] = (uint32_t)k; }
        } else {
            uint32_t w[16];
            rand_words(w, 16);
            sc_reduce512(h, w);
        }
        half_scalars(h, x, xn, y);
        print_hex(h, 8); printf(" "); print_hex(x, 8); printf(" %u ", xn); print_hex(y, 8); printf("\n");
    }
    return 0;
}

// encoding of the point decoded from `enc` plus the point decoded from `t`
// (decode gives the negation; the sum of two negations is negated back)
static void add_encodings(uint32_t out[8], const uint32_t enc[8], const uint32_t t[8],
                          const ed25519_tables &T) {
    ge_p3 a, b, s;
    ge_decode_neg(a, enc, T.k);
    ge_decode_neg(b, t, T.k);
    ge_cached bc;
    ge_p3_to_cached(bc, b, T.k.d2);
    ge_p1p1 r;
    ge_add_cached(r, a, bc, 0);
    ge_p1p1_to_p3(s, r);
    fe_neg(s.X, s.X);
    ge_encode(out, s.X, s.Y, s.Z);
}

static const char *kTorsion[8] = {
    // the 8 points of order dividing 8 (little-endian hex of the encodings)
    "0100000000000000000000000000000000000000000000000000000000000000",
    "ecffffffffffffffffffffffffffffffffffffffffffffffffffffffffffff7f",
    "0000000000000000000000000000000000000000000000000000000000000080",
    "0000000000000000000000000000000000000000000000000000000000000000",
    "c7176a703d4dd84fba3c0b760d10670f2a2053fa2c39ccc64ec7fd7792ac037a",
    "c7176a703d4dd84fba3c0b760d10670f2a2053fa2c39ccc64ec7fd7792ac03fa",
    "26e8958fc2b227b045c3f489f2ef98f0d5dfac05d3c63339b13802886d53fc05",
    "26e8958fc2b227b045c3f489f2ef98f0d5dfac05d3c63339b13802886d53fc85",
};
static void hex_words(uint32_t w[8], const char *hex) {
    uint8_t b[32];
    for (int i = 0; i < 32; i++) {
        unsigned v;
        sscanf(hex + 2 * i, "%2x", &v);
        b[i] = (uint8_t)v;
    }
    for (int i = 0; i < 8; i++)
        w[i] = b[4 * i] | (b[4 * i + 1] << 8) | (b[4 * i + 2] << 16) | ((uint32_t)b[4 * i + 3] << 24);
}

static int run_verify(int n) {
    const ed25519_tables &T = host_tables();
    int kinds[12] = {0}, acc_old[12] = {0}, diff = 0;
    uint8_t msg[600];
    for (int k = 0; k < n; k++) {
        uint32_t seed[8], pk[8], sk[16], sig[16], tor[8];
        rand_words(seed, 8);
        keypair_from_seed(pk, sk, seed, T);
        const uint32_t mlen = (uint32_t)(next_u64() % 500);
        for (uint32_t i = 0; i < mlen; i++) msg[i] = (uint8_t)next_u64();
        const int kind = k % 12;
        hex_words(tor, kTorsion[next_u64() % 8]);
        uint32_t pk_used[8];
        memcpy(pk_used, pk, sizeof pk);
        if (kind == 6) memcpy(pk_used, tor, sizeof tor);              // small-order key
        if (kind == 7 || kind == 8) add_encodings(pk_used, pk, tor, T); // A + T (mixed)
        MsgReader m{msg, mlen};
        sign_one(sig, pk_used, sk, m, T);   // hash uses the key as given
        switch (kind) {
            case 1: sig[next_u64() % 16] ^= 1u << (next_u64() % 32); break;   // flipped sig bit
            case 2: if (mlen) msg[next_u64() % mlen] ^= 1u << (next_u64() % 8); break;
            case 3: {   // S + L (non-canonical S)
                const uint32_t L[8] = SPL_L_WORDS;
                uint64_t c = 0;
                for (int i = 0; i < 8; i++) { c += (uint64_t)sig[8 + i] + L[i]; sig[8 + i] = (uint32_t)c; c >>= 32; }
                break;
            }
            case 4: add_encodings(sig, sig, tor, T); break;               // R + T
            case 5: memcpy(sig, tor, 32); break;                          // R = torsion point
            case 9: rand_words(sig, 8); break;                            // random R bytes
            case 10: rand_words(pk_used, 8); break;                       // random key bytes
            case 11: sig[7] ^= 0x80000000u; break;                        // R sign bit
            default: break;
        }
        ge_cached tabs[2 * VarWindow<5>::kTable];
        ge_cached Ai[VarWindow<4>::kTable];
        const int o = verify_one<4>(sig, pk_used, m, T, Ai);
        const int n4 = verify_half_one<4>(sig, pk_used, m, T, tabs);
        const int n5 = verify_half_one<5>(sig, pk_used, m, T, tabs);
        kinds[kind]++;
        acc_old[kind] += o;
        if (o != n4 || o != n5) {
            diff++;
            if (diff < 10) printf("DIFF case %d kind %d: old %d half4 %d half5 %d\n", k, kind, o, n4, n5);
        }
    }
    for (int i = 0; i < 12; i++) printf("kind %d: %d cases, %d accepted\n", i, kinds[i], acc_old[i]);
    printf("differences: %d\n", diff);
    return diff ? 1 : 0;
}

// Field products on loose inputs: limbs up to k times their tight range
// (k = 1, 2, 3; 3 is the most the multiply's 32-bit operand 19 g allows),
// random and at the extremes.  Prints, per case, the limbs of f, g,
// f g, f^2 and 2 f^2 (ten signed integers each).
static void rand_fe(fe &f, int kind) {
    const int k = 1 + kind % 3;
    for (int i = 0; i < 10; i++) {
        const int64_t m = (int64_t)k << ((i & 1) ? 24 : 25);   // limb in (-m, m)
        const uint64_t r = next_u64();
        int64_t v;
        switch (kind / 3) {
        case 0: v = (int64_t)(r % (uint64_t)(2 * m - 1)) - (m - 1); break;   // random
        case 1: v = m - 1; break;                                             // all +max
        case 2: v = -(m - 1); break;                                          // all -max
        case 3: v = (i & 1) ? -(m - 1) : m - 1; break;                        // alternating
        default: v = (r & 1) ? m - 1 : -(m - 1); break;                       // random signs
        }
        f.v[i] = (int32_t)v;
    }
}

static void print_fe(const fe &f) {
    for (int i = 0; i < 10; i++) printf(" %d", f.v[i]);
}

static int run_field(int n) {
    for (int c = 0; c < n; c++) {
        fe f, g, m, s, s2;
        rand_fe(f, c % 15);
        rand_fe(g, (c / 15) % 15);
        fe_mul(m, f, g);
        fe_sq(s, f);
        fe_sq2(s2, f);
        print_fe(f); print_fe(g); print_fe(m); print_fe(s); print_fe(s2);
        printf("\n");
    }
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    const int n = atoi(argv[2]);
    g_state = strtoull(argv[3], nullptr, 10);
    if (!strcmp(argv[1], "lattice")) return run_lattice(n);
    if (!strcmp(argv[1], "verify")) return run_verify(n);
    if (!strcmp(argv[1], "field")) return run_field(n);
    return 2;
}
