// Serving loop: verify and PoH requests of any size arrive from any thread,
// are admitted into a device-resident ring by HBM budget, and are batched
// continuously onto a small set of streams ("slots").
//
// No counterpart in the reference: its callers invoke ed25519_verify_many /
// poh_verify_many synchronously, one call per batch
// (src/cuda-ecc-ed25519/verify.cu:240-340, src/cuda-poh-verify/poh_verify.cu:80-144).
//
// Scheduling.  One loop thread owns every device operation.  A slot is a
// stream with at most one batch in flight.  Whenever a slot is free and
// requests of its kind are pending, the loop takes the pending requests in
// arrival order (up to the slot's signature cap and the ring space free right
// now), packs them into one block of the ring's pinned host mirror, and
// enqueues ingest copy kernel -> verify or PoH kernel -> egress copy kernel ->
// event on the slot's stream.  Each slot's stream has a hardware queue of its
// own, so a long PoH launch never sits in front of a verify batch.  Requests that arrive meanwhile wait only for the next free
// slot, not for a fixed batch boundary, and batches on different slots run
// concurrently on the device (a one-shot verify launch covers ceil(n/64)
// waves, a PoH launch ceil(chains/64)).
//
// Each request's offsets are checked against its own packet buffer and
// rebased into the batch on the host, so a request can never read another
// request's packets, and its verdicts do not depend on what it is batched with.
#include <hip/hip_runtime.h>
#include <time.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <deque>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ring.h"
#include "verify_kernels.h"
#include "../../include/cuda_crypt.h"

namespace spl {
namespace {

uint64_t now_ns() {
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
    uint64_t off = 0, size = 0;           // ring block
    uint64_t in_bytes = 0;                // bytes copied H2D from the block start
    uint64_t out_off = 0, out_bytes = 0;  // result area inside the block
    uint64_t t_formed = 0, t_launched = 0, t_device_done = 0;
};

// Bytes of ring a verify batch needs.
uint64_t verify_block_bytes(uint64_t packets_bytes, uint64_t n) {
    return align_up(align_up(packets_bytes, 256) + 16 * n + n, 256);
}
uint64_t poh_block_bytes(uint64_t n) { return align_up(40 * n, 256); }

struct LatencyStats {
    std::vector<uint64_t> ns;             // per request: submit -> results delivered
    // per batch: formed -> enqueued (host staging), enqueued -> completion
    // seen (copy-in, kernel, copy-out), completion -> results delivered
    std::vector<uint64_t> stage_ns, device_ns, deliver_ns;
    uint64_t requests = 0, units = 0, batches = 0, failed = 0;
    void reset() { *this = LatencyStats(); }
};

// Hardware resources of servers, kept for the life of the process and handed
// to the next server created on the same device: a slot's stream (CU-masked,
// so a hardware queue of its own), its completion event and its verify
// scratch, and a ring's device buffer and pinned mirror.  Creating a server
// then makes these only when no free ones exist, and destroying it makes no
// HIP call at all: it only hands them back.
struct SlotRes {
    int dev = 0;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    bool in_use = false;
};

struct RingRes {
    int dev = 0;
    uint64_t cap = 0;
    uint8_t *d = nullptr, *h = nullptr;
    bool in_use = false;
};

std::mutex g_res_mu;
std::vector<SlotRes *> g_slot_res;
std::vector<RingRes *> g_ring_res;

SlotRes *take_slot_res(int dev, const std::vector<uint32_t> &cu_mask) {
    std::lock_guard<std::mutex> g(g_res_mu);
    for (SlotRes *r : g_slot_res)
        if (r->dev == dev && !r->in_use) {
            r->in_use = true;
            return r;
        }
    auto *r = new SlotRes();
    r->dev = dev;
    // A stream created with a CU mask gets a hardware queue of its own rather
    // than one shared round-robin with the process's other streams, so a long
    // PoH kernel never sits in front of a verify batch.
    if (hipExtStreamCreateWithCUMask(&r->stream, (uint32_t)cu_mask.size(), cu_mask.data()) !=
            hipSuccess ||
        hipEventCreateWithFlags(&r->done, hipEventDisableTiming) != hipSuccess) {
        if (r->stream) (void)hipStreamDestroy(r->stream);
        delete r;
        return nullptr;
    }
    r->in_use = true;
    g_slot_res.push_back(r);
    return r;
}

// A free ring of at least `cap` bytes on `dev`, else a new one.
RingRes *take_ring_res(int dev, uint64_t cap) {
    std::lock_guard<std::mutex> g(g_res_mu);
    for (RingRes *r : g_ring_res)
        if (r->dev == dev && !r->in_use && r->cap >= cap) {
            r->in_use = true;
            return r;
        }
    auto *r = new RingRes();
    r->dev = dev;
    r->cap = cap;
    // coherent: the egress kernel's writes are visible to the host once the
    // batch's event has fired
    if (hipMalloc(&r->d, cap) != hipSuccess ||
        hipHostMalloc(&r->h, cap, hipHostMallocCoherent) != hipSuccess) {
        if (r->d) (void)hipFree(r->d);
        delete r;
        return nullptr;
    }
    // Touch every page of the mirror now: the first CPU write to a page of
    // pinned memory was measured at 2.7 GB/s, later ones at 33-52 GB/s, and
    // staging must not pay that on the request path.
    memset(r->h, 0, cap);
    r->in_use = true;
    g_ring_res.push_back(r);
    return r;
}

void give_back(SlotRes *r) {
    std::lock_guard<std::mutex> g(g_res_mu);
    r->in_use = false;
}

void give_back(RingRes *r) {
    std::lock_guard<std::mutex> g(g_res_mu);
    r->in_use = false;
}

}  // namespace
}  // namespace spl

using namespace spl;

struct spl_server {
    struct Slot {
        Kind kind = kVerify;
        SlotRes *r = nullptr;                 // stream, event, scratch (pooled)
        Batch *batch = nullptr;
    };

    int dev = 0, cus = 256;
    uint32_t max_batch = 0;
    Ring ring;
    RingRes *ring_res = nullptr;              // the ring's memory (pooled)
    uint8_t *d_ring = nullptr, *h_ring = nullptr;
    const ed25519_tables *T = nullptr;
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable wake, done_cv;
    std::deque<Request *> pending[2];
    std::unordered_map<int64_t, Request *> reqs;
    int64_t next_id = 1;
    bool stop = false;
    std::atomic<uint64_t> submitted{0};     // requests ever queued (polled without mu)
    static constexpr uint64_t kIdleSpinNs = 1000000;
    std::thread loop;

    LatencyStats stats[2];
    uint32_t inflight = 0, max_inflight = 0;

    // Every batch has completed by now (the loop only exits idle), so the
    // pooled resources go back with no work on them.
    ~spl_server() {
        for (auto &s : slots)
            if (s.r) give_back(s.r);
        if (ring_res) give_back(ring_res);
        for (auto &kv : reqs) delete kv.second;
    }

    uint64_t request_bytes(const Request &r) const {
        return r.kind == kVerify ? verify_block_bytes(align_up(r.packets_size, 16), r.n)
                                 : poh_block_bytes(r.n);
    }

    // Takes pending requests of `kind` in arrival order into one batch that
    // fits the slot's cap and the ring's free space.  Called with mu held.
    Batch *form_batch(Kind kind) {
        std::deque<Request *> &q = pending[kind];
        if (q.empty()) return nullptr;
        size_t take = 0;
        uint64_t units = 0, pbytes = 0;
        while (take < q.size()) {
            const Request &r = *q[take];
            if (take > 0 && units + r.n > max_batch) break;
            units += r.n;
            if (kind == kVerify) pbytes += align_up(r.packets_size, 16);
            take++;
        }
        // shrink until the block fits the ring (the first request alone
        // always fits an empty ring: submit checked it against the budget)
        uint64_t off = 0, size = 0;
        for (;;) {
            size = kind == kVerify ? verify_block_bytes(pbytes, units) : poh_block_bytes(units);
            if (ring.alloc(size, &off)) break;
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
        const uint64_t t = now_ns();
        uint64_t pb = 0;
        uint32_t ub = 0;
        for (size_t k = 0; k < take; k++) {
            Request *r = q.front();
            q.pop_front();
            r->state = kAdmitted;
            r->t_admit = t;
            b->reqs.push_back(r);
            b->pbase.push_back(pb);
            b->ubase.push_back(ub);
            if (kind == kVerify) pb += align_up(r->packets_size, 16);
            ub += r->n;
        }
        return b;
    }

    // Packs a verify batch into the pinned mirror of its ring block.
    void stage_verify(Batch &b) {
        uint8_t *h = h_ring + b.off;
        const uint64_t tab_off = align_up(b.packets_bytes, 256);
        const uint32_t N = b.n;
        uint32_t *lens = reinterpret_cast<uint32_t *>(h + tab_off);
        uint32_t *pk = lens + N, *sig = lens + 2ull * N, *msg = lens + 3ull * N;
        for (size_t k = 0; k < b.reqs.size(); k++) {
            const Request &r = *b.reqs[k];
            const uint64_t base = b.pbase[k];
            memcpy(h + base, r.packets, r.packets_size);
            const uint64_t pad = align_up(r.packets_size, 16) - r.packets_size;
            if (pad) memset(h + base + r.packets_size, 0, pad);
            const uint32_t u0 = b.ubase[k];
            for (uint32_t i = 0; i < r.n; i++) {
                const uint64_t so = r.sig[i], po = r.pk[i], mo = r.msg[i], ml = r.lens[i];
                const bool ok = so + 64 <= r.packets_size && po + 32 <= r.packets_size &&
                                mo + ml <= r.packets_size;
                lens[u0 + i] = ok ? (uint32_t)ml : 0;
                pk[u0 + i] = ok ? (uint32_t)(po + base) : 0;
                msg[u0 + i] = ok ? (uint32_t)(mo + base) : 0;
                // an offset past the batch's packet area makes the kernel reject
                sig[u0 + i] = ok ? (uint32_t)(so + base) : 0xFFFFFFFFu;
            }
        }
        b.in_bytes = tab_off + 16ull * N;
        b.out_off = b.in_bytes;
        b.out_bytes = N;
    }

    void stage_poh(Batch &b) {
        uint8_t *h = h_ring + b.off;
        const uint32_t N = b.n;
        uint64_t *counts = reinterpret_cast<uint64_t *>(h + 32ull * N);
        for (size_t k = 0; k < b.reqs.size(); k++) {
            const Request &r = *b.reqs[k];
            memcpy(h + 32ull * b.ubase[k], r.hashes, 32ull * r.n);
            memcpy(counts + b.ubase[k], r.counts, 8ull * r.n);
        }
        b.in_bytes = 40ull * N;
        b.out_off = 0;
        b.out_bytes = 32ull * N;
    }

    hipError_t ensure_scratch(Slot &s, size_t bytes) {
        SlotRes &r = *s.r;
        if (bytes <= r.scratch_bytes) return hipSuccess;
        if (r.scratch) {
            hipError_t e = hipStreamSynchronize(r.stream);
            if (e != hipSuccess) return e;
            (void)hipFree(r.scratch);
            r.scratch = nullptr;
            r.scratch_bytes = 0;
        }
        hipError_t e = hipMalloc(&r.scratch, bytes);
        if (e == hipSuccess) r.scratch_bytes = bytes;
        return e;
    }

    // Stages and enqueues a batch on its slot: ingest copy kernel (pinned
    // mirror -> ring block), the verify or PoH kernel, egress copy kernel
    // (results -> pinned mirror), event.  Called without mu.
    hipError_t launch(Slot &s, Batch &b) {
        hipError_t e;
        uint8_t *d = d_ring + b.off;
        uint8_t *h = h_ring + b.off;
        if (b.kind == kVerify) {
            stage_verify(b);
            const VerifyConfig cfg = tuned_plain_config(dev, b.n);
            if ((e = ensure_scratch(s, verify_scratch_bytes(cfg, b.n, cus))) != hipSuccess)
                return e;
            if ((e = launch_copy16(d, h, b.in_bytes, s.r->stream)) != hipSuccess) return e;
            const uint64_t tab_off = align_up(b.packets_bytes, 256);
            const uint32_t *tab = reinterpret_cast<const uint32_t *>(d + tab_off);
            VerifyJob job = make_verify_job(d, b.packets_bytes, tab, tab + b.n, tab + 2ull * b.n,
                                            tab + 3ull * b.n, d + b.out_off, s.r->scratch, T);
            if ((e = launch_verify_cfg(cfg, job, b.n, cus, s.r->stream)) != hipSuccess) return e;
        } else {
            stage_poh(b);
            if ((e = launch_copy16(d, h, b.in_bytes, s.r->stream)) != hipSuccess) return e;
            if ((e = launch_poh(reinterpret_cast<uint32_t *>(d),
                                reinterpret_cast<const uint64_t *>(d + 32ull * b.n), b.n,
                                s.r->stream)) != hipSuccess)
                return e;
        }
        if ((e = launch_copy16(h + b.out_off, d + b.out_off, b.out_bytes, s.r->stream)) !=
            hipSuccess)
            return e;
        b.t_launched = now_ns();
        return hipEventRecord(s.r->done, s.r->stream);
    }

    // Delivers a finished batch's results.  Called without mu.
    void deliver(Batch &b, bool ok) {
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

    // Marks a batch's requests complete and frees its ring block.  Called
    // with mu held.
    void complete(Batch *b, bool ok) {
        const uint64_t t = now_ns();
        LatencyStats &st = stats[b->kind];
        st.batches++;
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
            if (!ok) st.failed++;
            st.ns.push_back(t - r->t_submit);
        }
        ring.release(b->off);
        delete b;
        done_cv.notify_all();
    }

    // The loop thread owns the slots; mu guards the pending queues, the
    // request records and the ring, and is taken only to move work between
    // them, never while polling events, so submitters and waiters are not
    // starved by the poll.
    void run() {
        (void)hipSetDevice(dev);
        for (;;) {
            bool progress = false;
            for (Slot &s : slots) {
                if (!s.batch) continue;
                const hipError_t q = hipEventQuery(s.r->done);
                if (q == hipErrorNotReady) continue;
                Batch *b = s.batch;
                b->t_device_done = now_ns();
                s.batch = nullptr;
                if (q != hipSuccess)
                    fprintf(stderr, "spl_server: batch of %u failed: %s\n", b->n,
                            hipGetErrorString(q));
                deliver(*b, q == hipSuccess);
                std::lock_guard<std::mutex> g(mu);
                inflight--;
                complete(b, q == hipSuccess);
                progress = true;
            }
            for (Slot &s : slots) {
                if (s.batch) continue;
                Batch *b = nullptr;
                {
                    std::lock_guard<std::mutex> g(mu);
                    if (!pending[s.kind].empty()) b = form_batch(s.kind);
                }
                if (!b) continue;
                const hipError_t e = launch(s, *b);
                if (e != hipSuccess) {
                    fprintf(stderr, "spl_server: launch of a batch of %u failed: %s\n", b->n,
                            hipGetErrorString(e));
                    deliver(*b, false);
                }
                std::lock_guard<std::mutex> g(mu);
                if (e == hipSuccess) {
                    s.batch = b;
                    inflight++;
                    max_inflight = std::max(max_inflight, inflight);
                } else {
                    complete(b, false);
                }
                progress = true;
            }
            if (progress) continue;
            // Busy-poll rather than sleep: a sleeping thread's core drops into
            // an idle state, and waking it (and the memcpy that follows) was
            // measured to add up to a millisecond at moderate request rates.
            bool idle;
            {
                std::lock_guard<std::mutex> g(mu);
                idle = inflight == 0;
                if (idle && stop && pending[0].empty() && pending[1].empty()) break;
            }
            if (!idle) {
                std::this_thread::yield();
                continue;
            }
            // idle: keep polling for a while before blocking
            const uint64_t seen = submitted.load(std::memory_order_relaxed);
            const uint64_t until = now_ns() + kIdleSpinNs;
            while (submitted.load(std::memory_order_relaxed) == seen && now_ns() < until)
                std::this_thread::yield();
            std::unique_lock<std::mutex> lk(mu);
            wake.wait(lk, [this] {
                return stop || !pending[0].empty() || !pending[1].empty();
            });
        }
    }

    int64_t submit(Request *r) {
        if (request_bytes(*r) > ring.cap() || r->n > max_batch) {
            delete r;
            return -2;
        }
        std::lock_guard<std::mutex> g(mu);
        if (stop) {
            delete r;
            return -2;
        }
        r->id = next_id++;
        r->t_submit = now_ns();
        reqs[r->id] = r;
        if (r->n == 0) {
            r->t_admit = r->t_done = r->t_submit;
            r->state = kDone;
            return r->id;
        }
        pending[r->kind].push_back(r);
        submitted.fetch_add(1, std::memory_order_relaxed);
        wake.notify_one();
        return r->id;
    }
};

namespace {

// p50, p99 and max of v in milliseconds.
void percentiles(std::vector<uint64_t> v, double out[3]) {
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

void append_class(std::string &s, const char *name, LatencyStats &st) {
    double lat[3], stg[3], dev[3], fet[3];
    percentiles(st.ns, lat);
    percentiles(st.stage_ns, stg);
    percentiles(st.device_ns, dev);
    percentiles(st.deliver_ns, fet);
    char buf[1024];
    snprintf(buf, sizeof buf,
             "\"%s\": {\"requests\": %llu, \"units\": %llu, \"batches\": %llu, "
             "\"failed\": %llu, \"p50_ms\": %.4f, \"p99_ms\": %.4f, \"max_ms\": %.4f, "
             "\"batch_stage_ms\": [%.4f, %.4f, %.4f], \"batch_device_ms\": [%.4f, %.4f, %.4f], "
             "\"batch_deliver_ms\": [%.4f, %.4f, %.4f]}",
             name, (unsigned long long)st.requests, (unsigned long long)st.units,
             (unsigned long long)st.batches, (unsigned long long)st.failed, lat[0], lat[1], lat[2],
             stg[0], stg[1], stg[2], dev[0], dev[1], dev[2], fet[0], fet[1], fet[2]);
    s += buf;
}

}  // namespace

extern "C" {

spl_server *spl_server_create(uint64_t hbm_budget_bytes, uint32_t max_batch_signatures,
                              uint32_t verify_streams, uint32_t poh_streams) {
    if (hbm_budget_bytes == 0) hbm_budget_bytes = 1ull << 30;
    if (max_batch_signatures == 0) max_batch_signatures = 65536;
    if (verify_streams == 0) verify_streams = 6;
    if (poh_streams == 0) poh_streams = 2;
    if (verify_streams + poh_streams > 16 || hbm_budget_bytes > (1ull << 32)) return nullptr;
    if (pool_init() <= 0) return nullptr;
    auto *sv = new spl_server();
    bool ok = hipGetDevice(&sv->dev) == hipSuccess;
    if (ok) sv->T = device_tables(sv->dev);
    ok = ok && sv->T;
    if (ok) {
        sv->cus = device_cus(sv->dev);
        sv->max_batch = max_batch_signatures;
        sv->ring = Ring(align_up(hbm_budget_bytes, 256));
        sv->ring_res = take_ring_res(sv->dev, sv->ring.cap());
        ok = sv->ring_res != nullptr;
        if (ok) {
            sv->d_ring = sv->ring_res->d;
            sv->h_ring = sv->ring_res->h;
        }
    }
    std::vector<uint32_t> all_cus((sv->cus + 31) / 32, 0xFFFFFFFFu);
    for (uint32_t k = 0; ok && k < verify_streams + poh_streams; k++) {
        spl_server::Slot s;
        s.kind = k < verify_streams ? kVerify : kPoh;
        s.r = take_slot_res(sv->dev, all_cus);
        ok = s.r != nullptr;
        if (ok && s.kind == kVerify) {
            // sized for a full batch under the widest window, so the loop
            // never allocates in steady state
            const VerifyConfig widest = {0, 5, 1, 0, 0, 1, 1};
            ok = sv->ensure_scratch(s, verify_scratch_bytes(widest, max_batch_signatures,
                                                            sv->cus)) == hipSuccess;
        }
        if (s.r) sv->slots.push_back(s);
    }
    if (!ok) {
        fprintf(stderr, "spl_server_create: failed\n");
        delete sv;
        return nullptr;
    }
    sv->loop = std::thread([sv] { sv->run(); });
    return sv;
}

void spl_server_destroy(spl_server *sv) {
    if (!sv) return;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
        sv->wake.notify_one();
    }
    if (sv->loop.joinable()) sv->loop.join();
    delete sv;
}

int64_t spl_server_submit_verify(spl_server *sv, const uint8_t *packets, uint64_t packets_size,
                                 const uint32_t *message_lens, const uint32_t *public_key_offsets,
                                 const uint32_t *signature_offsets,
                                 const uint32_t *message_start_offsets, uint32_t num_signatures,
                                 uint8_t *out) {
    if (!sv) return -2;
    if (num_signatures && (!packets || !message_lens || !public_key_offsets ||
                           !signature_offsets || !message_start_offsets || !out))
        return -2;
    if (packets_size > 0xFFFFFFFFull) return -2;
    auto *r = new Request();
    r->kind = kVerify;
    r->n = num_signatures;
    r->packets = packets;
    r->packets_size = packets_size;
    r->lens = message_lens;
    r->pk = public_key_offsets;
    r->sig = signature_offsets;
    r->msg = message_start_offsets;
    r->out = out;
    return sv->submit(r);
}

int64_t spl_server_submit_poh(spl_server *sv, uint8_t *hashes, const uint64_t *num_hashes_arr,
                              size_t num_elems) {
    if (!sv) return -2;
    if (num_elems && (!hashes || !num_hashes_arr)) return -2;
    if (num_elems > 0xFFFFFFFFull) return -2;
    auto *r = new Request();
    r->kind = kPoh;
    r->n = (uint32_t)num_elems;
    r->hashes = hashes;
    r->counts = num_hashes_arr;
    return sv->submit(r);
}

int spl_server_wait(spl_server *sv, int64_t ticket, int timeout_ms, uint64_t *times_ns) {
    if (!sv) return 2;
    std::unique_lock<std::mutex> lk(sv->mu);
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return 2;
    Request *r = it->second;
    auto finished = [r] { return r->state == kDone || r->state == kFailed; };
    if (timeout_ms < 0) {
        sv->done_cv.wait(lk, finished);
    } else if (!sv->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished)) {
        return 1;
    }
    if (times_ns) {
        times_ns[0] = r->t_submit;
        times_ns[1] = r->t_admit;
        times_ns[2] = r->t_done;
    }
    const int rc = r->state == kDone ? 0 : 3;
    sv->reqs.erase(it);
    delete r;
    return rc;
}

int spl_server_stats(spl_server *sv, char *buf, size_t len, int reset) {
    if (!sv || !buf || !len) return 2;
    std::lock_guard<std::mutex> g(sv->mu);
    std::string s = "{";
    append_class(s, "verify", sv->stats[kVerify]);
    s += ", ";
    append_class(s, "poh", sv->stats[kPoh]);
    char tail[256];
    snprintf(tail, sizeof tail,
             ", \"ring_bytes\": %llu, \"ring_peak_bytes\": %llu, \"max_inflight_batches\": %u}",
             (unsigned long long)sv->ring.cap(), (unsigned long long)sv->ring.peak(),
             sv->max_inflight);
    s += tail;
    snprintf(buf, len, "%s", s.c_str());
    if (reset) {
        sv->stats[kVerify].reset();
        sv->stats[kPoh].reset();
        sv->max_inflight = 0;
    }
    return s.size() < len ? 0 : 2;
}

}  // extern "C"
