// CPU check of the serving loop's scheduling core
// (solana-perf-libs_amd/csrc/server_core.h), built with g++ (with and without
// -fsanitize=thread) by tests/test_server_core.py.
//
//   server_check CYCLES SEED   runs CYCLES server lifetimes; prints counts,
//                              exits 1 on the first wrong result
//
// The core runs over a device double: one worker thread per slot plays the
// slot's stream, and its own buffer plays the device ring.  Every step it
// copies the batch's input bytes in (as the ingest copy kernel does: from the
// ring's host copy, and, for requests marked as read in place, from the
// submitter's own packet buffer), computes in the device buffer, and copies
// the results back to the host copy (the egress copy kernel), after a random
// delay.  The
// "kernels" are stand-ins with exact expected values:
//   verify: out[i] = 1 + (FNV-1a of the signature, key and message bytes) % 255
//           for a signature whose fields lie in its request's packets, else 0
//   PoH:    a 32-byte state mixed by a fixed permutation once per "hash",
//           sliced exactly as the HIP kernels slice a chain
// so a batch that reads another request's bytes, loses a slice, or has its
// ring block reused while the worker still copies it gives a wrong result
// (and, under ThreadSanitizer, a data race report).  Half the verify requests
// are read in place; once such a request's wait returns, the main thread
// overwrites its packets, which races with any read the "device" still makes
// of them.
//
// Each cycle picks a ring size, batch cap, slot counts and PoH slice length,
// then two threads submit verify and PoH requests (random sizes, offsets past
// their packets, zero counts, empty requests, requests too large for the
// ring) while the main thread waits on them with a time limit; stop() is
// called either after every wait or before them.  Failures are injected: a
// step that reports failure, and a launch that fails after enqueuing work
// (the core must drain the slot before its block is reused).
#include <assert.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../solana-perf-libs_amd/csrc/server_core.h"

using namespace spl::serve;

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {   // splitmix64
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

uint8_t verify_value(const uint8_t *sig, const uint8_t *pk, const uint8_t *msg, uint32_t len) {
    uint32_t h = 2166136261u;
    auto mix = [&h](const uint8_t *p, uint32_t n) {
        for (uint32_t i = 0; i < n; i++) h = (h ^ p[i]) * 16777619u;
    };
    mix(sig, 64);
    mix(pk, 32);
    mix(msg, len);
    return (uint8_t)(1 + h % 255);
}

void poh_step(uint8_t *st) {
    uint64_t w[4];
    memcpy(w, st, 32);
    for (int k = 0; k < 4; k++) {
        uint64_t z = w[k] + 0x9e3779b97f4a7c15ull * (k + 1) + w[(k + 1) & 3];
        z = (z ^ (z >> 31)) * 0xbf58476d1ce4e5b9ull;
        w[k] = z ^ (z >> 29);
    }
    memcpy(st, w, 32);
}

// One step of a batch, as the worker executes it.
struct Step {
    Batch *b;
    bool ingest, kernel, egress, fail;
    uint64_t skip, limit;       // PoH slice: hashes [skip, skip + limit) of each chain
    uint32_t delay_us;
};

struct FakeDev {
    struct Queue {
        std::mutex mu;
        std::condition_variable cv;
        std::deque<Step> q;
        uint64_t enq = 0;                       // loop thread only
        std::atomic<uint64_t> done{0};          // release: the step's writes are visible
        std::atomic<bool> failed{false};
        bool stop = false;
        std::thread th;
    };

    uint8_t *h_ring, *d_ring;
    uint64_t slice;
    std::vector<Queue *> qs;
    Rng rng;                                    // loop thread only
    int fail_step_pct, fail_launch_pct;
    std::atomic<uint64_t> injected_step{0}, injected_launch{0}, steps{0};

    FakeDev(uint8_t *h, uint8_t *d, int nslots, uint64_t slice_, uint64_t seed, int fs, int fl)
        : h_ring(h), d_ring(d), slice(slice_), rng{seed}, fail_step_pct(fs), fail_launch_pct(fl) {
        for (int k = 0; k < nslots; k++) {
            Queue *q = new Queue();
            q->th = std::thread([this, q] { work(q); });
            qs.push_back(q);
        }
    }

    ~FakeDev() {
        for (Queue *q : qs) {
            {
                std::lock_guard<std::mutex> g(q->mu);
                q->stop = true;
            }
            q->cv.notify_one();
            q->th.join();
            delete q;
        }
    }

    void run_step(const Step &s) {
        Batch &b = *s.b;
        uint8_t *d = d_ring + b.off, *h = h_ring + b.off;
        if (s.ingest && b.kind == kVerify)
            ingest_ranges(h, b, [d](const uint8_t *src, uint64_t off, uint64_t n) {
                assert(off % 16 == 0 && n % 16 == 0);
                memcpy(d + off, src, n);
            });
        else if (s.ingest)
            memcpy(d, h, b.in_bytes);
        if (s.kernel && !s.fail) {
            if (b.kind == kVerify) {
                const uint64_t tab = align_up(b.packets_bytes, 256);
                const uint32_t *t = reinterpret_cast<const uint32_t *>(d + tab);
                const uint32_t n = b.n;
                uint8_t *out = d + b.out_off;
                for (uint32_t i = 0; i < n; i++) {
                    const uint64_t ml = t[i], po = t[n + i], so = t[2 * n + i], mo = t[3 * n + i];
                    // the HIP kernel's own bounds rule over the batch's packet area
                    const bool ok = so + 64 <= b.packets_bytes && po + 32 <= b.packets_bytes &&
                                    mo + ml <= b.packets_bytes;
                    out[i] = ok ? verify_value(d + so, d + po, d + mo, (uint32_t)ml) : 0;
                }
            } else {
                const uint64_t *cnt = reinterpret_cast<const uint64_t *>(d + 32ull * b.n);
                for (uint32_t i = 0; i < b.n; i++) {
                    const uint64_t c = cnt[i];
                    uint64_t m = c > s.skip ? c - s.skip : 0;
                    if (m > s.limit) m = s.limit;
                    for (uint64_t r = 0; r < m; r++) poh_step(d + 32ull * i);
                }
            }
        }
        if (s.egress && !s.fail) memcpy(h + b.out_off, d + b.out_off, b.out_bytes);
    }

    void work(Queue *q) {
        for (;;) {
            Step s;
            {
                std::unique_lock<std::mutex> lk(q->mu);
                q->cv.wait(lk, [q] { return q->stop || !q->q.empty(); });
                if (q->q.empty()) return;
                s = q->q.front();
            }
            if (s.delay_us) std::this_thread::sleep_for(std::chrono::microseconds(s.delay_us));
            run_step(s);
            if (s.fail) q->failed.store(true, std::memory_order_relaxed);
            {
                std::lock_guard<std::mutex> g(q->mu);
                q->q.pop_front();
            }
            q->done.fetch_add(1, std::memory_order_release);
            steps.fetch_add(1, std::memory_order_relaxed);
        }
    }

    void enqueue(int slot, const Step &s) {
        Queue *q = qs[slot];
        q->enq++;
        {
            std::lock_guard<std::mutex> g(q->mu);
            q->q.push_back(s);
        }
        q->cv.notify_one();
    }

    Step step_for(Batch &b, bool ingest, bool kernel, bool egress, uint64_t skip) {
        Step s{&b, ingest, kernel, egress, false, skip, slice, (uint32_t)rng.below(120)};
        if ((int)rng.below(100) < fail_step_pct) {
            s.fail = true;
            injected_step++;
        }
        return s;
    }

    // --- the policy the core calls
    void thread_init() {}

    bool launch(int slot, Batch &b) {
        qs[slot]->failed.store(false, std::memory_order_relaxed);
        if (b.kind == kVerify) {
            stage_verify(h_ring, b);
            enqueue(slot, step_for(b, true, true, true, 0));
        } else {
            stage_poh(h_ring, b);
            const bool last = b.poh_max <= slice;
            enqueue(slot, step_for(b, true, b.poh_max > 0, last, 0));
            b.poh_done = b.poh_max > 0 ? std::min(slice, b.poh_max) : 0;
        }
        // a launch that fails after part of the batch is enqueued
        if ((int)rng.below(100) < fail_launch_pct) {
            injected_launch++;
            return false;
        }
        return true;
    }

    Poll poll(int slot) {
        Queue *q = qs[slot];
        if (q->done.load(std::memory_order_acquire) != q->enq) return Poll::kBusy;
        return q->failed.load(std::memory_order_relaxed) ? Poll::kFailed : Poll::kDone;
    }

    int next(int slot, Batch &b) {
        if (b.kind == kVerify || b.poh_done >= b.poh_max) return 0;
        const uint64_t skip = b.poh_done;
        b.poh_done = std::min(b.poh_max, skip + slice);
        enqueue(slot, step_for(b, false, true, b.poh_done >= b.poh_max, skip));
        return 1;
    }

    void drain(int slot) {
        Queue *q = qs[slot];
        while (q->done.load(std::memory_order_acquire) != q->enq) std::this_thread::yield();
    }
};

// A request as its submitter built it, with the result it must get.
struct Sub {
    Kind kind;
    int64_t ticket = 0;
    bool refused_expected = false;
    bool direct = false;                // packets read in place by the "device"
    std::vector<uint8_t> packets, out, want;
    std::vector<uint32_t> lens, pk, sig, msg;
    std::vector<uint8_t> hashes, start;
    std::vector<uint64_t> counts;
};

void make_verify(Rng &r, Sub &s, uint64_t ring_cap, uint32_t max_batch) {
    s.kind = kVerify;
    const uint32_t n = r.below(10) == 0 ? 0 : 1 + (uint32_t)r.below(r.below(4) == 0 ? 300 : 24);
    const uint32_t stride = 96 + (uint32_t)r.below(200);
    const uint32_t npk = 1 + (uint32_t)r.below(n ? n : 1);
    // read in place: the size must be a multiple of 16 (the address is, as
    // operator new aligns to 16)
    s.direct = r.below(2) == 0;
    const size_t bytes = (size_t)npk * stride + r.below(15);
    s.packets.resize(s.direct ? align_up(bytes, 16) : bytes);
    for (auto &x : s.packets) x = (uint8_t)r.next();
    s.lens.resize(n);
    s.pk.resize(n);
    s.sig.resize(n);
    s.msg.resize(n);
    s.want.assign(n, 0);
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t base = (uint64_t)r.below(npk) * stride;
        s.sig[i] = (uint32_t)base;
        s.pk[i] = (uint32_t)(base + 64);
        s.msg[i] = (uint32_t)(base + 96);
        s.lens[i] = (uint32_t)r.below(stride - 95);
        switch (r.below(12)) {              // fields past the request's own packets
            case 0: s.lens[i] = (uint32_t)(s.packets.size() - s.msg[i] + 1); break;
            case 1: s.sig[i] = (uint32_t)s.packets.size() - 63; break;
            case 2: s.pk[i] = 0xFFFFFFF0u; break;
            default: break;
        }
        const uint64_t sz = s.packets.size();
        const bool ok = (uint64_t)s.sig[i] + 64 <= sz && (uint64_t)s.pk[i] + 32 <= sz &&
                        (uint64_t)s.msg[i] + s.lens[i] <= sz;
        if (ok)
            s.want[i] = verify_value(&s.packets[s.sig[i]], &s.packets[s.pk[i]],
                                     &s.packets[s.msg[i]], s.lens[i]);
    }
    s.out.assign(n ? n : 1, 0xEE);
    s.refused_expected = verify_block_bytes(align_up(s.packets.size(), 16), n) > ring_cap ||
                         n > max_batch;
}

void make_poh(Rng &r, Sub &s, uint64_t ring_cap, uint32_t max_batch) {
    s.kind = kPoh;
    const uint32_t n = r.below(10) == 0 ? 0 : 1 + (uint32_t)r.below(r.below(4) == 0 ? 200 : 12);
    s.hashes.resize(32ull * n);
    for (auto &x : s.hashes) x = (uint8_t)r.next();
    s.start = s.hashes;
    s.counts.resize(n);
    s.want = s.hashes;
    for (uint32_t i = 0; i < n; i++) {
        s.counts[i] = r.below(5) == 0 ? 0 : r.below(40);
        for (uint64_t k = 0; k < s.counts[i]; k++) poh_step(&s.want[32ull * i]);
    }
    s.refused_expected = poh_block_bytes(n) > ring_cap || n > max_batch;
}

int64_t submit(Core<FakeDev> &core, Sub &s) {
    Request *q = new Request();
    q->kind = s.kind;
    if (s.kind == kVerify) {
        q->n = (uint32_t)s.lens.size();
        q->packets = s.packets.data();
        q->packets_size = s.packets.size();
        q->lens = s.lens.data();
        q->pk = s.pk.data();
        q->sig = s.sig.data();
        q->msg = s.msg.data();
        q->out = s.out.data();
        if (s.direct && !s.packets.empty()) {
            assert((uintptr_t)s.packets.data() % 16 == 0);
            q->direct = true;
            q->packets_dev = s.packets.data();
        }
    } else {
        q->n = (uint32_t)s.counts.size();
        q->hashes = s.hashes.data();
        q->counts = s.counts.data();
    }
    return core.submit(q);
}

struct Counts {
    uint64_t requests = 0, refused = 0, failed = 0, verify_sigs = 0, chains = 0, steps = 0,
             inj_step = 0, inj_launch = 0, direct = 0;
};

// A failed check ends the process at once (the loop and worker threads are
// still running, so nothing is unwound).
#define CHECK(cond, ...)                                          \
    do {                                                          \
        if (!(cond)) {                                            \
            fprintf(stderr, "cycle %d: ", cycle);                 \
            fprintf(stderr, __VA_ARGS__);                         \
            fprintf(stderr, "\n");                                \
            fflush(stderr);                                       \
            _exit(1);                                             \
        }                                                         \
    } while (0)

bool check_result(int cycle, const Sub &s, int rc) {
    CHECK(rc == 0 || rc == 3, "wait returned %d", rc);
    if (s.kind == kVerify) {
        for (size_t i = 0; i < s.want.size(); i++)
            CHECK(s.out[i] == (rc == 0 ? s.want[i] : 0),
                  "verify result %zu: got %u want %u (rc %d)", i, s.out[i],
                  rc == 0 ? s.want[i] : 0, rc);
    } else {
        CHECK(s.hashes == (rc == 0 ? s.want : s.start), "PoH states differ (rc %d)", rc);
    }
    return true;
}

bool cycle_once(int cycle, Rng &r, Counts &c) {
    const uint64_t cap = align_up(4096 + r.below(1 << 20), 256);
    const uint32_t max_batch = 16 + (uint32_t)r.below(600);
    const int nv = 1 + (int)r.below(4), np = 1 + (int)r.below(2);
    const uint64_t slice = 1 + r.below(16);
    const bool faults = r.below(3) == 0;
    std::vector<Kind> kinds;
    for (int k = 0; k < nv; k++) kinds.push_back(kVerify);
    for (int k = 0; k < np; k++) kinds.push_back(kPoh);
    std::vector<uint8_t> h(cap), d(cap);
    FakeDev dev(h.data(), d.data(), nv + np, slice, r.next(), faults ? 4 : 0, faults ? 4 : 0);
    std::vector<Sub> subs[2];
    const int per = 8 + (int)r.below(40);
    for (int t = 0; t < 2; t++) {
        subs[t].resize(per);
        for (int i = 0; i < per; i++) {
            if ((r.below(2) == 0) == (t == 0)) make_verify(r, subs[t][i], cap, max_batch);
            else make_poh(r, subs[t][i], cap, max_batch);
        }
        // now and then one far larger than the ring
        if (r.below(4) == 0) {
            Sub &s = subs[t][r.below(per)];
            s.kind = kPoh;
            s.counts.assign(cap / 40 + 8, 1);
            s.hashes.assign(32 * s.counts.size(), 0);
            s.start = s.want = s.hashes;
            s.refused_expected = true;
        }
    }
    Core<FakeDev> core(dev, h.data(), cap, max_batch, kinds);
    core.start();
    std::thread feeders[2];
    for (int t = 0; t < 2; t++)
        feeders[t] = std::thread([&core, &subs, t] {
            for (Sub &s : subs[t]) {
                s.ticket = submit(core, s);
                if ((s.ticket & 7) == 3) std::this_thread::yield();
            }
        });
    const bool stop_first = r.below(4) == 0;
    for (auto &f : feeders) f.join();
    if (stop_first) core.stop();
    CHECK(core.wait(987654321, 0, nullptr) == 2, "unknown ticket not reported");
    for (int t = 0; t < 2; t++)
        for (Sub &s : subs[t]) {
            if (s.refused_expected) {
                CHECK(s.ticket == -2, "oversized request not refused (ticket %lld)",
                      (long long)s.ticket);
                c.refused++;
                continue;
            }
            CHECK(s.ticket > 0, "request refused (ticket %lld)", (long long)s.ticket);
            uint64_t times[3];
            const int rc = core.wait(s.ticket, 20000, times);
            CHECK(rc != 1, "request %lld not finished within 20 s", (long long)s.ticket);
            CHECK(times[0] <= times[1] && times[1] <= times[2], "times out of order");
            if (!check_result(cycle, s, rc)) return false;
            if (s.direct) {
                memset(s.packets.data(), 0x5A, s.packets.size());
                c.direct++;
            }
            c.requests++;
            if (rc == 3) c.failed++;
            if (s.kind == kVerify) c.verify_sigs += s.want.size();
            else c.chains += s.counts.size();
            CHECK(core.wait(s.ticket, 0, nullptr) == 2, "ticket still known after its wait");
        }
    if (!stop_first) core.stop();
    char buf[4096];
    CHECK(core.stats(buf, sizeof buf, true) == 0, "stats did not fit");
    c.steps += dev.steps.load();
    c.inj_step += dev.injected_step.load();
    c.inj_launch += dev.injected_launch.load();
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    const int cycles = argc > 1 ? atoi(argv[1]) : 200;
    Rng r{argc > 2 ? strtoull(argv[2], nullptr, 10) : 1};
    Counts c;
    for (int k = 0; k < cycles; k++)
        if (!cycle_once(k, r, c)) return 1;
    printf("cycles %d requests %llu refused %llu failed %llu signatures %llu chains %llu "
           "steps %llu injected_step_failures %llu injected_launch_failures %llu "
           "read_in_place %llu\n",
           cycles, (unsigned long long)c.requests, (unsigned long long)c.refused,
           (unsigned long long)c.failed, (unsigned long long)c.verify_sigs,
           (unsigned long long)c.chains, (unsigned long long)c.steps,
           (unsigned long long)c.inj_step, (unsigned long long)c.inj_launch,
           (unsigned long long)c.direct);
    return 0;
}
