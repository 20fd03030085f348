// CPU check of the serving loop's ring allocator (solana-perf-libs_amd/csrc/
// ring.h), built with g++ by tests/test_ring.py.
//
//   ring_check N SEED   runs N random sequences of allocations and releases
//                       (releases in random order, as batches complete), and
//                       checks every step against an interval model; prints
//                       counts, exits 1 on the first violation
//
// Checked: a granted block lies in [0, cap) and overlaps no live block; an
// empty ring grants every size from 1 to cap; peak() is the largest sum of
// live sizes; a refused request is granted once every block is released.
#include <stdint.h>
#include <stdio.h>

#include <algorithm>
#include <vector>

#include "../../solana-perf-libs_amd/csrc/ring.h"

static uint64_t g_state;
static uint64_t next_u64() {   // splitmix64
    uint64_t z = (g_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static uint64_t below(uint64_t n) { return next_u64() % n; }

struct Live {
    uint64_t off, size;
};

#define FAIL(...)                                            \
    do {                                                     \
        fprintf(stderr, "seq %d step %d: ", seq, step);      \
        fprintf(stderr, __VA_ARGS__);                        \
        fprintf(stderr, "\n");                               \
        return 1;                                            \
    } while (0)

int main(int argc, char **argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 200;
    g_state = argc > 2 ? strtoull(argv[2], nullptr, 10) : 1;
    uint64_t grants = 0, refusals = 0, wraps = 0;
    for (int seq = 0; seq < n; seq++) {
        int step = 0;
        // capacities from tiny (every request large against the ring) to roomy
        const uint64_t cap = 1 + below(seq % 3 == 0 ? 64 : 4096);
        spl::Ring ring(cap);
        std::vector<Live> live;
        uint64_t used = 0, last_off = 0;
        // an empty ring grants every size up to cap (the largest of them is
        // the peak so far)
        uint64_t peak = std::min<uint64_t>(cap, 300);
        for (uint64_t s = 1; s <= peak; s++) {
            uint64_t off = 77;
            if (!ring.alloc(s, &off) || off != 0) FAIL("empty ring refused size %llu", (unsigned long long)s);
            ring.release(off);
        }
        uint64_t off = 0;
        if (ring.alloc(0, &off) || ring.alloc(cap + 1, &off)) FAIL("granted size 0 or cap + 1");
        for (step = 0; step < 400; step++) {
            if (live.empty() || below(100) < 55) {
                // sizes skewed small, with the occasional large one
                const uint64_t size = 1 + (below(4) ? below(std::max<uint64_t>(1, cap / 6)) : below(cap));
                if (ring.alloc(size, &off)) {
                    if (off + size > cap) FAIL("block [%llu, +%llu) past cap %llu", (unsigned long long)off, (unsigned long long)size, (unsigned long long)cap);
                    for (const Live &l : live)
                        if (off < l.off + l.size && l.off < off + size)
                            FAIL("block [%llu, +%llu) overlaps live [%llu, +%llu)", (unsigned long long)off, (unsigned long long)size, (unsigned long long)l.off, (unsigned long long)l.size);
                    if (!live.empty() && off < last_off) wraps++;
                    last_off = off;
                    live.push_back({off, size});
                    used += size;
                    peak = std::max(peak, used);
                    grants++;
                } else {
                    refusals++;
                    if (live.empty()) FAIL("empty ring refused size %llu of cap %llu", (unsigned long long)size, (unsigned long long)cap);
                    if (below(2) == 0) {
                        // the server's situation: wait for every batch in flight, then
                        // the same request must be admitted
                        for (const Live &l : live) ring.release(l.off);
                        live.clear();
                        used = 0;
                        if (!ring.alloc(size, &off) || off != 0) FAIL("refused size %llu after all released", (unsigned long long)size);
                        live.push_back({off, size});
                        used = size;
                        peak = std::max(peak, used);
                        last_off = off;
                        grants++;
                    }
                }
            } else {
                const size_t k = below(live.size());   // completion order is arbitrary
                ring.release(live[k].off);
                used -= live[k].size;
                live.erase(live.begin() + k);
            }
            if (ring.peak() != peak) FAIL("peak %llu, model %llu", (unsigned long long)ring.peak(), (unsigned long long)peak);
        }
    }
    printf("sequences %d grants %llu refusals %llu wraps %llu\n", n, (unsigned long long)grants,
           (unsigned long long)refusals, (unsigned long long)wraps);
    return 0;
}
