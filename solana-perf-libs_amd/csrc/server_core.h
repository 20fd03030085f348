// Scheduling core of the serving loop, free of HIP: requests, pending queues,
// batch forming, the ring, completion, wait and stop.
//
// No counterpart in the reference: its callers invoke ed25519_verify_many /
// poh_verify_many synchronously, one call per batch
// (src/cuda-ecc-ed25519/verify.cu:240-340, src/cuda-poh-verify/poh_verify.cu:80-144).
//
// The core is a template over a device policy, so the same code runs over
// HIP streams (server.hip) and over a CPU double that executes the "device
// steps" on threads of its own (tests/native/server_check.cpp, run under
// ThreadSanitizer).  A policy `Dev` provides:
//
//   void thread_init();              called first on the loop thread
//   bool launch(int slot, Batch &b); stage b into its ring block and enqueue
//                                    its first device step on `slot`
//   Poll poll(int slot);             state of the step last enqueued on slot
//   int  next(int slot, Batch &b);   after that step finished: 1 = another
//                                    step of b enqueued, 0 = b is finished,
//                                    -1 = error
//   void drain(int slot);            wait until everything enqueued on slot
//                                    has finished
//
// A verify batch is one step (ingest copy, verify kernel, egress copy).  The
// loop stages each request's packets into the ring's host copy, except those
// of a request whose packets already sit in pinned host memory: the ingest
// copy reads those straight from the caller's pages (ingest_ranges), and the
// loop writes only their rebased offset tables.  A PoH batch is one step per
// slice of at most `poh_slice` hashes per chain (the first also copies the
// chains in, the last copies them out); the chain state stays in the ring
// block between slices.  The loop enqueues the next slice only when the
// previous one has finished, so a slot's queue never holds more than one
// slice: any other work behind it waits at most one slice, however long the
// chains are.
//
// Scheduling.  One loop thread owns every device operation.  A slot has at
// most one batch in flight.  Whenever a slot is free and requests of its
// kind are pending, the loop takes them in arrival order (up to the batch cap
// and the ring space free right now), packs them into one block of the ring
// and launches it.  Requests that arrive meanwhile wait only for the next
// free slot.  Each request's offsets are checked against its own packets and
// rebased into the batch on the host, so a request can never read another
// request's packets and its results do not depend on what it is batched with.
//
// Locking.  `mu` guards the pending queues, the request records, the ring and
// the statistics; the loop takes it only to move work between them, never
// across a device call, so submitters and waiters are never held up by a
// poll.  A batch's ring block is released only after its device work has
// finished or been drained, so a block is never handed to a new batch while
// the device may still read or write it.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ring.h"

namespace spl {
namespace serve {

inline uint64_t now_ns() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (uint64_t)ts.tv_sec * 1000000000ull + (uint64_t)ts.tv_nsec;
}

constexpr uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

enum Kind { kVerify = 0, kPoh = 1 };
enum State { kPending = 0, kAdmitted = 1, kDone = 2, kFailed = 3 };

struct Request {
    int64_t id = 0;
    Kind kind = kVerify;
    uint32_t n = 0;                       // signatures or chains
    // verify
    const uint8_t *packets = nullptr;
    uint64_t packets_size = 0;
    // The device reads the packets in place, from packets_dev (the device's
    // address of `packets`: pinned host memory), instead of the loop staging
    // them; set by the submitter only when packets_dev and packets_size are
    // multiples of 16.
    bool direct = false;
    const uint8_t *packets_dev = nullptr;
    const uint32_t *lens = nullptr, *pk = nullptr, *sig = nullptr, *msg = nullptr;
    uint8_t *out = nullptr;
    // PoH
    uint8_t *hashes = nullptr;
    const uint64_t *counts = nullptr;
    uint64_t t_submit = 0, t_admit = 0, t_done = 0;
    State state = kPending;
};

struct Batch {
    Kind kind = kVerify;
    std::vector<Request *> reqs;
    std::vector<uint64_t> pbase;          // verify: packet base of each request
    std::vector<uint32_t> ubase;          // first signature / chain of each request
    uint32_t n = 0;                       // signatures or chains
    uint64_t packets_bytes = 0;           // verify: bytes of the packet area
    uint32_t direct = 0;                  // verify: requests read in place
    uint64_t off = 0, size = 0;           // ring block
    uint64_t in_bytes = 0;                // bytes copied in from the block start
    uint64_t out_off = 0, out_bytes = 0;  // result area inside the block
    uint64_t poh_max = 0;                 // PoH: largest hash count of the batch
    uint64_t poh_done = 0;                // PoH: hashes per chain enqueued so far
    uint32_t steps = 0;                   // device steps enqueued
    uint64_t t_formed = 0, t_launched = 0, t_device_done = 0;
};

// Bytes of ring a verify batch needs: packets (256-aligned), the four u32
// tables and the u8 mask.
inline uint64_t verify_block_bytes(uint64_t packets_bytes, uint64_t n) {
    return align_up(align_up(packets_bytes, 256) + 16 * n + n, 256);
}
// PoH: 32-byte states then u64 counts.
inline uint64_t poh_block_bytes(uint64_t n) { return align_up(40 * n, 256); }

// Packs a verify batch into its block of the ring's host copy: each request's
// packets (16-byte aligned; a request read in place leaves its area of the
// host copy unwritten), then lens, pk, sig, msg tables with offsets rebased
// into the batch; a signature whose fields fall outside its own request's
// packets gets a signature offset past the batch (the kernel rejects it).
inline void stage_verify(uint8_t *h_ring, Batch &b) {
    uint8_t *h = h_ring + b.off;
    const uint64_t tab_off = align_up(b.packets_bytes, 256);
    const uint32_t N = b.n;
    uint32_t *lens = reinterpret_cast<uint32_t *>(h + tab_off);
    uint32_t *pk = lens + N, *sig = lens + 2ull * N, *msg = lens + 3ull * N;
    b.direct = 0;
    for (size_t k = 0; k < b.reqs.size(); k++) {
        const Request &r = *b.reqs[k];
        const uint64_t base = b.pbase[k];
        if (r.direct) {
            b.direct++;
        } else {
            memcpy(h + base, r.packets, r.packets_size);
            const uint64_t pad = align_up(r.packets_size, 16) - r.packets_size;
            if (pad) memset(h + base + r.packets_size, 0, pad);
        }
        const uint32_t u0 = b.ubase[k];
        for (uint32_t i = 0; i < r.n; i++) {
            const uint64_t so = r.sig[i], po = r.pk[i], mo = r.msg[i], ml = r.lens[i];
            const bool ok = so + 64 <= r.packets_size && po + 32 <= r.packets_size &&
                            mo + ml <= r.packets_size;
            lens[u0 + i] = ok ? (uint32_t)ml : 0;
            pk[u0 + i] = ok ? (uint32_t)(po + base) : 0;
            msg[u0 + i] = ok ? (uint32_t)(mo + base) : 0;
            sig[u0 + i] = ok ? (uint32_t)(so + base) : 0xFFFFFFFFu;
        }
    }
    b.in_bytes = tab_off + 16ull * N;
    b.out_off = b.in_bytes;
    b.out_bytes = N;
}

// The ingest copies of a staged verify batch into its ring block, as
// emit(source, offset in the block, bytes) calls, every one a multiple of 16
// bytes at a 16-byte-aligned offset: each direct request's packets from its
// own pages, and everything between them (staged packets, padding, the
// tables) from the block's host copy h_block.  Without direct requests that
// is one range, [0, in_bytes).
template <class F>
inline void ingest_ranges(const uint8_t *h_block, const Batch &b, F &&emit) {
    uint64_t from = 0;                    // start of the host-copy range not yet emitted
    for (size_t k = 0; b.direct && k < b.reqs.size(); k++) {
        const Request &r = *b.reqs[k];
        if (!r.direct || r.packets_size == 0) continue;
        if (b.pbase[k] > from) emit(h_block + from, from, b.pbase[k] - from);
        emit(r.packets_dev, b.pbase[k], r.packets_size);
        from = b.pbase[k] + r.packets_size;
    }
    if (b.in_bytes > from) emit(h_block + from, from, b.in_bytes - from);
}

// Packs a PoH batch: the chains' 32-byte states, then their u64 counts.
inline void stage_poh(uint8_t *h_ring, Batch &b) {
    uint8_t *h = h_ring + b.off;
    const uint32_t N = b.n;
    uint64_t *counts = reinterpret_cast<uint64_t *>(h + 32ull * N);
    uint64_t mx = 0;
    for (size_t k = 0; k < b.reqs.size(); k++) {
        const Request &r = *b.reqs[k];
        memcpy(h + 32ull * b.ubase[k], r.hashes, 32ull * r.n);
        memcpy(counts + b.ubase[k], r.counts, 8ull * r.n);
        for (uint32_t i = 0; i < r.n; i++) mx = std::max(mx, r.counts[i]);
    }
    b.in_bytes = 40ull * N;
    b.out_off = 0;
    b.out_bytes = 32ull * N;
    b.poh_max = mx;
    b.poh_done = 0;
}

// Copies a finished batch's results from the ring's host copy to the callers'
// buffers (a failed verify batch rejects every signature; a failed PoH batch
// leaves the callers' states as they were).
inline void deliver(const uint8_t *h_ring, const Batch &b, bool ok) {
    const uint8_t *res = h_ring + b.off + b.out_off;
    for (size_t k = 0; k < b.reqs.size(); k++) {
        Request &r = *b.reqs[k];
        if (b.kind == kVerify) {
            if (ok) memcpy(r.out, res + b.ubase[k], r.n);
            else memset(r.out, 0, r.n);
        } else if (ok) {
            memcpy(r.hashes, res + 32ull * b.ubase[k], 32ull * r.n);
        }
    }
}

struct LatencyStats {
    std::vector<uint64_t> ns;             // per request: submit -> results delivered
    // per batch: formed -> enqueued (host staging), enqueued -> completion
    // seen (every device step), completion -> results delivered
    std::vector<uint64_t> stage_ns, device_ns, deliver_ns;
    uint64_t requests = 0, units = 0, batches = 0, failed = 0, steps = 0;
    uint64_t in_place = 0;                // verify requests read in place
    void reset() { *this = LatencyStats(); }
};

// p50, p99 and max of v in milliseconds.
inline void percentiles(std::vector<uint64_t> v, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    if (v.empty()) return;
    const double ps[2] = {0.50, 0.99};
    for (int i = 0; i < 2; i++) {
        size_t k = (size_t)(ps[i] * (double)(v.size() - 1) + 0.5);
        std::nth_element(v.begin(), v.begin() + k, v.end());
        out[i] = v[k] / 1e6;
    }
    out[2] = *std::max_element(v.begin(), v.end()) / 1e6;
}

inline void append_class(std::string &s, const char *name, const LatencyStats &st) {
    double lat[3], stg[3], dev[3], fet[3];
    percentiles(st.ns, lat);
    percentiles(st.stage_ns, stg);
    percentiles(st.device_ns, dev);
    percentiles(st.deliver_ns, fet);
    char buf[1024];
    snprintf(buf, sizeof buf,
             "\"%s\": {\"requests\": %llu, \"units\": %llu, \"batches\": %llu, "
             "\"device_steps\": %llu, \"failed\": %llu, \"read_in_place\": %llu, "
             "\"p50_ms\": %.4f, \"p99_ms\": %.4f, "
             "\"max_ms\": %.4f, \"batch_stage_ms\": [%.4f, %.4f, %.4f], "
             "\"batch_device_ms\": [%.4f, %.4f, %.4f], \"batch_deliver_ms\": [%.4f, %.4f, %.4f]}",
             name, (unsigned long long)st.requests, (unsigned long long)st.units,
             (unsigned long long)st.batches, (unsigned long long)st.steps,
             (unsigned long long)st.failed, (unsigned long long)st.in_place, lat[0], lat[1],
             lat[2], stg[0], stg[1], stg[2],
             dev[0], dev[1], dev[2], fet[0], fet[1], fet[2]);
    s += buf;
}

enum class Poll { kBusy, kDone, kFailed };

template <class Dev>
class Core {
  public:
    // h_ring: the ring's host copy (ring_cap bytes), which the policy's
    // launch() stages into and whose result areas are read on completion.
    Core(Dev &dev, uint8_t *h_ring, uint64_t ring_cap, uint32_t max_batch,
         const std::vector<Kind> &slot_kinds)
        : dev_(dev), h_ring_(h_ring), ring_(ring_cap), max_batch_(max_batch) {
        for (Kind k : slot_kinds) slots_.push_back(Slot{k, nullptr});
    }

    // stop() must have returned (every admitted batch has then completed).
    ~Core() {
        for (auto &kv : reqs_) delete kv.second;
    }

    void start() { loop_ = std::thread([this] { run(); }); }

    // Refuses new requests, lets the loop finish every submitted one and
    // joins it.
    void stop() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
            wake_.notify_one();
        }
        if (loop_.joinable()) loop_.join();
    }

    uint64_t ring_cap() const { return ring_.cap(); }

    // Takes ownership of r.  Returns its ticket (> 0), or -2 when it can
    // never fit (larger than the ring or than the batch cap) or the core is
    // stopping.
    int64_t submit(Request *r) {
        const uint64_t bytes = r->kind == kVerify
                                   ? verify_block_bytes(align_up(r->packets_size, 16), r->n)
                                   : poh_block_bytes(r->n);
        if (bytes > ring_.cap() || r->n > max_batch_) {
            delete r;
            return -2;
        }
        std::lock_guard<std::mutex> g(mu_);
        if (stop_) {
            delete r;
            return -2;
        }
        r->id = next_id_++;
        r->t_submit = now_ns();
        reqs_[r->id] = r;
        if (r->n == 0) {
            r->t_admit = r->t_done = r->t_submit;
            r->state = kDone;
            return r->id;
        }
        pending_[r->kind].push_back(r);
        submitted_.fetch_add(1, std::memory_order_relaxed);
        wake_.notify_one();
        return r->id;
    }

    // 0: done, 3: failed (both forget the ticket); 1: not finished within
    // timeout_ms (< 0: no limit); 2: unknown ticket.
    int wait(int64_t ticket, int timeout_ms, uint64_t *times_ns) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(ticket);
        if (it == reqs_.end()) return 2;
        Request *r = it->second;
        auto finished = [r] { return r->state == kDone || r->state == kFailed; };
        if (timeout_ms < 0) {
            done_cv_.wait(lk, finished);
        } else if (!done_cv_.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished)) {
            return 1;
        }
        if (times_ns) {
            times_ns[0] = r->t_submit;
            times_ns[1] = r->t_admit;
            times_ns[2] = r->t_done;
        }
        const int rc = r->state == kDone ? 0 : 3;
        reqs_.erase(it);
        delete r;
        return rc;
    }

    // Statistics as one JSON object; 0 when it fit in buf.
    int stats(char *buf, size_t len, bool reset) {
        std::lock_guard<std::mutex> g(mu_);
        std::string s = "{";
        append_class(s, "verify", stats_[kVerify]);
        s += ", ";
        append_class(s, "poh", stats_[kPoh]);
        char tail[256];
        snprintf(tail, sizeof tail,
                 ", \"ring_bytes\": %llu, \"ring_peak_bytes\": %llu, "
                 "\"max_inflight_batches\": %u}",
                 (unsigned long long)ring_.cap(), (unsigned long long)ring_.peak(),
                 max_inflight_);
        s += tail;
        snprintf(buf, len, "%s", s.c_str());
        if (reset) {
            stats_[kVerify].reset();
            stats_[kPoh].reset();
            max_inflight_ = 0;
        }
        return s.size() < len ? 0 : 2;
    }

  private:
    struct Slot {
        Kind kind;
        Batch *batch;
    };

    // Takes pending requests of `kind` in arrival order into one batch that
    // fits the batch cap and the ring's free space.  Called with mu_ held.
    Batch *form_batch(Kind kind) {
        std::deque<Request *> &q = pending_[kind];
        if (q.empty()) return nullptr;
        size_t take = 0;
        uint64_t units = 0, pbytes = 0;
        while (take < q.size()) {
            const Request &r = *q[take];
            if (take > 0 && units + r.n > max_batch_) break;
            units += r.n;
            if (kind == kVerify) pbytes += align_up(r.packets_size, 16);
            take++;
        }
        // shrink until the block fits the ring (the first request alone
        // always fits an empty ring: submit checked it against the capacity)
        uint64_t off = 0, size = 0;
        for (;;) {
            size = kind == kVerify ? verify_block_bytes(pbytes, units) : poh_block_bytes(units);
            if (ring_.alloc(size, &off)) break;
            if (take == 1) return nullptr;   // wait for space to free up
            take--;
            const Request &r = *q[take];
            units -= r.n;
            if (kind == kVerify) pbytes -= align_up(r.packets_size, 16);
        }
        Batch *b = new Batch();
        b->t_formed = now_ns();
        b->kind = kind;
        b->off = off;
        b->size = size;
        b->n = (uint32_t)units;
        b->packets_bytes = pbytes;
        uint64_t pb = 0;
        uint32_t ub = 0;
        for (size_t k = 0; k < take; k++) {
            Request *r = q.front();
            q.pop_front();
            r->state = kAdmitted;
            r->t_admit = b->t_formed;
            b->reqs.push_back(r);
            b->pbase.push_back(pb);
            b->ubase.push_back(ub);
            if (kind == kVerify) pb += align_up(r->packets_size, 16);
            ub += r->n;
        }
        return b;
    }

    // Marks a batch's requests finished and frees its ring block; the device
    // no longer touches the block.  Called with mu_ held.
    void complete(Batch *b, bool ok) {
        const uint64_t t = now_ns();
        LatencyStats &st = stats_[b->kind];
        st.batches++;
        st.steps += b->steps;
        if (ok && b->t_device_done) {
            st.stage_ns.push_back(b->t_launched - b->t_formed);
            st.device_ns.push_back(b->t_device_done - b->t_launched);
            st.deliver_ns.push_back(t - b->t_device_done);
        }
        for (Request *r : b->reqs) {
            r->t_done = t;
            r->state = ok ? kDone : kFailed;
            st.requests++;
            st.units += r->n;
            if (r->direct) st.in_place++;
            if (!ok) st.failed++;
            st.ns.push_back(t - r->t_submit);
        }
        ring_.release(b->off);
        delete b;
        done_cv_.notify_all();
    }

    // Delivers and completes a batch whose device work has ended (ok) or
    // been drained (failed).  Called without mu_.
    void finish(Batch *b, bool ok) {
        b->t_device_done = now_ns();
        deliver(h_ring_, *b, ok);
        std::lock_guard<std::mutex> g(mu_);
        complete(b, ok);
    }

    void run() {
        dev_.thread_init();
        for (;;) {
            bool progress = false;
            for (size_t k = 0; k < slots_.size(); k++) {
                Slot &s = slots_[k];
                if (!s.batch) continue;
                const Poll p = dev_.poll((int)k);
                if (p == Poll::kBusy) continue;
                Batch *b = s.batch;
                progress = true;
                int nx = -1;
                if (p == Poll::kDone) {
                    nx = dev_.next((int)k, *b);
                    if (nx == 1) {                // the batch's next step is enqueued
                        b->steps++;
                        continue;
                    }
                }
                if (nx != 0) {
                    fprintf(stderr, "spl_server: batch of %u failed on the device\n", b->n);
                    dev_.drain((int)k);           // nothing may still touch the block
                }
                s.batch = nullptr;
                {
                    std::lock_guard<std::mutex> g(mu_);
                    inflight_--;
                }
                finish(b, nx == 0);
            }
            for (size_t k = 0; k < slots_.size(); k++) {
                Slot &s = slots_[k];
                if (s.batch) continue;
                Batch *b = nullptr;
                {
                    std::lock_guard<std::mutex> g(mu_);
                    if (!pending_[s.kind].empty()) b = form_batch(s.kind);
                    if (b) {
                        inflight_++;
                        max_inflight_ = std::max(max_inflight_, inflight_);
                    }
                }
                if (!b) continue;
                progress = true;
                if (dev_.launch((int)k, *b)) {
                    b->steps = 1;
                    b->t_launched = now_ns();
                    s.batch = b;
                    continue;
                }
                // Part of the batch may already be enqueued: wait for it
                // before the block can go to another batch.
                fprintf(stderr, "spl_server: launch of a batch of %u failed\n", b->n);
                dev_.drain((int)k);
                {
                    std::lock_guard<std::mutex> g(mu_);
                    inflight_--;
                }
                b->t_launched = now_ns();
                deliver(h_ring_, *b, false);
                std::lock_guard<std::mutex> g(mu_);
                complete(b, false);
            }
            if (progress) continue;
            // Busy-poll rather than sleep while work is in flight: a sleeping
            // thread's core drops into an idle state, and waking it (and the
            // memcpy that follows) was measured to add up to a millisecond
            // at moderate request rates.
            bool idle;
            {
                std::lock_guard<std::mutex> g(mu_);
                idle = inflight_ == 0;
                if (idle && stop_ && pending_[0].empty() && pending_[1].empty()) break;
            }
            if (!idle) {
                std::this_thread::yield();
                continue;
            }
            // idle: keep polling for a while before blocking
            const uint64_t seen = submitted_.load(std::memory_order_relaxed);
            const uint64_t until = now_ns() + kIdleSpinNs;
            while (submitted_.load(std::memory_order_relaxed) == seen && now_ns() < until)
                std::this_thread::yield();
            std::unique_lock<std::mutex> lk(mu_);
            wake_.wait(lk, [this] {
                return stop_ || !pending_[0].empty() || !pending_[1].empty();
            });
        }
    }

    static constexpr uint64_t kIdleSpinNs = 1000000;

    Dev &dev_;
    uint8_t *h_ring_;
    std::vector<Slot> slots_;             // the loop thread's only
    std::mutex mu_;
    std::condition_variable wake_, done_cv_;
    Ring ring_;
    const uint32_t max_batch_;
    std::deque<Request *> pending_[2];
    std::unordered_map<int64_t, Request *> reqs_;
    int64_t next_id_ = 1;
    bool stop_ = false;
    std::atomic<uint64_t> submitted_{0};  // requests ever queued (polled without mu_)
    LatencyStats stats_[2];
    uint32_t inflight_ = 0, max_inflight_ = 0;
    std::thread loop_;
};

}  // namespace serve
}  // namespace spl
