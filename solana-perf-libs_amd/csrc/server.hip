// Serving loop: verify and PoH requests of any size arrive from any thread,
// are admitted into a device-resident ring by HBM budget, and are batched
// continuously onto a small set of streams ("slots").
//
// No counterpart in the reference: its callers invoke ed25519_verify_many /
// poh_verify_many synchronously, one call per batch
// (src/cuda-ecc-ed25519/verify.cu:240-340, src/cuda-poh-verify/poh_verify.cu:80-144).
//
// The scheduling (admission, batch forming, completion, wait, stop) is the
// HIP-free core in server_core.h; this file is its device policy over HIP
// streams and the C ABI around it.
//
// Per batch, on the slot's stream: ingest copy kernel (pinned mirror -> ring
// block; a verify request whose packets are already in pinned memory is
// gathered from the caller's pages instead), the verify kernel or one PoH
// slice, egress copy kernel (results -> pinned mirror), event.  The loop
// thread polls the events.
//
// Hardware queues.  Every slot stream is a plain, non-blocking, default-
// priority stream, so all of them share the process's GPU_MAX_HW_QUEUES
// hardware queues with its other streams: creating a server adds no queue.
// Kernels of streams that share a hardware queue run in enqueue order, so a
// verify batch may find a PoH launch in front of it.  That wait is bounded
// by slicing PoH: a PoH batch is hashed kPohSliceHashes per chain per launch,
// the state kept in the ring block between launches, and the loop enqueues a
// slot's next slice only once the previous one has finished, so no queue
// ever holds more than one slice.
//
// HIP calls.  spl_server_create makes hipGetDevice / hipDeviceGetAttribute
// and, only when the process holds no free pooled resource of the kind, the
// allocations of the pooled resources below (hipMalloc of the ring and the
// verify scratch, hipHostMalloc of the pinned mirror, hipStreamCreateWithFlags
// and hipEventCreateWithFlags per slot).  The loop thread makes hipSetDevice,
// kernel launches, hipEventRecord and hipEventQuery, and hipStreamSynchronize
// on a failed batch.  spl_server_submit_verify asks hipPointerGetAttributes
// (on the submitting thread) whether the packets lie in pinned memory.  spl_server_destroy makes none: it joins the loop, and
// the resources go back to the pool for the next server on the device.
#include <hip/hip_runtime.h>

#include <atomic>
#include <memory>
#include <string>
#include <vector>

#include "server_core.h"
#include "verify_kernels.h"
#include "../../include/cuda_crypt.h"

namespace spl {
namespace {

using serve::align_up;
using serve::Batch;
using serve::Kind;
using serve::kPoh;
using serve::kVerify;
using serve::Poll;
using serve::Request;

// Hashes per chain per PoH launch.  The two-wave split kernel (every PoH
// batch the server forms has at most 65,536 chains, most far fewer) takes
// about 1.9 us per hash on an idle GPU and up to twice that beside verify
// waves, which run at higher priority: 256 hashes is 0.5-1 ms per launch,
// the most a verify batch sharing the hardware queue waits for.
constexpr uint64_t kPohSliceHashes = 256;

// Hardware resources of servers, kept for the life of the process and handed
// to the next server created on the same device: a slot's stream, its
// completion event and its verify scratch, and a ring's device buffer and
// pinned mirror.
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

SlotRes *take_slot_res(int dev) {
    std::lock_guard<std::mutex> g(g_res_mu);
    for (SlotRes *r : g_slot_res)
        if (r->dev == dev && !r->in_use) {
            r->in_use = true;
            return r;
        }
    auto *r = new SlotRes();
    r->dev = dev;
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess ||
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

hipError_t ensure_scratch(SlotRes &r, size_t bytes) {
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

// The core's device policy over HIP streams.
struct HipDev {
    int dev = 0, cus = 256;
    const ed25519_tables *T = nullptr;
    uint8_t *d_ring = nullptr, *h_ring = nullptr;
    std::vector<SlotRes *> res;
    // The device operation the loop thread is inside, if any, and since when
    // (read by spl_server_stats from any thread): a request that never
    // completes then names the call the loop is waiting in.
    std::atomic<const char *> in_call{nullptr};
    std::atomic<uint64_t> in_call_since{0};

    struct InCall {
        HipDev &d;
        InCall(HipDev &dv, const char *what) : d(dv) {
            d.in_call_since.store(serve::now_ns(), std::memory_order_relaxed);
            d.in_call.store(what, std::memory_order_release);
        }
        ~InCall() { d.in_call.store(nullptr, std::memory_order_release); }
    };

    static bool ok(hipError_t e, const char *what) {
        if (e == hipSuccess) return true;
        fprintf(stderr, "spl_server: %s: %s\n", what, hipGetErrorString(e));
        return false;
    }

    void thread_init() {
        InCall c(*this, "hipSetDevice");
        (void)hipSetDevice(dev);
    }

    // One PoH slice of batch b: hashes [b.poh_done, + kPohSliceHashes) of
    // every chain, then the egress copy if that was the last slice.
    bool poh_slice(SlotRes &r, Batch &b) {
        uint8_t *d = d_ring + b.off, *h = h_ring + b.off;
        const uint64_t skip = b.poh_done;
        if (!ok(launch_poh(reinterpret_cast<uint32_t *>(d),
                           reinterpret_cast<const uint64_t *>(d + 32ull * b.n), b.n, r.stream,
                           skip, kPohSliceHashes),
                "PoH launch"))
            return false;
        b.poh_done = std::min(b.poh_max, skip + kPohSliceHashes);
        if (b.poh_done >= b.poh_max &&
            !ok(launch_copy16(h + b.out_off, d + b.out_off, b.out_bytes, r.stream), "egress copy"))
            return false;
        return ok(hipEventRecord(r.done, r.stream), "event record");
    }

    // A staged verify batch's input into its ring block: one copy from the
    // host copy, or, with requests read in place, gathered copies from the
    // host copy and the callers' pinned pages (one launch per kGatherMax
    // ranges).
    hipError_t ingest_verify(SlotRes &r, const Batch &b) {
        uint8_t *d = d_ring + b.off;
        const uint8_t *h = h_ring + b.off;
        if (!b.direct) return launch_copy16(d, h, b.in_bytes, r.stream);
        GatherList g;
        hipError_t e = hipSuccess;
        serve::ingest_ranges(h, b, [&](const uint8_t *src, uint64_t off, uint64_t n) {
            if (e == hipSuccess && g.full()) e = launch_gather(d, g, r.stream);
            if (e == hipSuccess) g.push(off, src, n / 16);
        });
        return e == hipSuccess ? launch_gather(d, g, r.stream) : e;
    }

    bool launch(int s, Batch &b) {
        InCall c(*this, b.kind == kVerify ? "verify launch" : "PoH launch");
        SlotRes &r = *res[s];
        uint8_t *d = d_ring + b.off, *h = h_ring + b.off;
        if (b.kind == kVerify) {
            serve::stage_verify(h_ring, b);
            const VerifyConfig cfg = tuned_plain_config(dev, b.n);
            if (!ok(ensure_scratch(r, verify_scratch_bytes(cfg, b.n, cus)), "scratch") ||
                !ok(ingest_verify(r, b), "ingest copy"))
                return false;
            const uint64_t tab_off = align_up(b.packets_bytes, 256);
            const uint32_t *tab = reinterpret_cast<const uint32_t *>(d + tab_off);
            VerifyJob job = make_verify_job(d, b.packets_bytes, tab, tab + b.n, tab + 2ull * b.n,
                                            tab + 3ull * b.n, d + b.out_off, r.scratch, T);
            return ok(launch_verify_cfg(cfg, job, b.n, cus, r.stream), "verify launch") &&
                   ok(launch_copy16(h + b.out_off, d + b.out_off, b.out_bytes, r.stream),
                      "egress copy") &&
                   ok(hipEventRecord(r.done, r.stream), "event record");
        }
        serve::stage_poh(h_ring, b);
        if (!ok(launch_copy16(d, h, b.in_bytes, r.stream), "ingest copy")) return false;
        if (b.poh_max == 0)                      // every count is 0: states go back as they are
            return ok(launch_copy16(h + b.out_off, d + b.out_off, b.out_bytes, r.stream),
                      "egress copy") &&
                   ok(hipEventRecord(r.done, r.stream), "event record");
        return poh_slice(r, b);
    }

    Poll poll(int s) {
        hipError_t q;
        {
            InCall c(*this, "hipEventQuery");
            q = hipEventQuery(res[s]->done);
        }
        if (q == hipErrorNotReady) return Poll::kBusy;
        if (q == hipSuccess) return Poll::kDone;
        fprintf(stderr, "spl_server: batch failed: %s\n", hipGetErrorString(q));
        return Poll::kFailed;
    }

    int next(int s, Batch &b) {
        if (b.kind == kVerify || b.poh_done >= b.poh_max) return 0;
        InCall c(*this, "PoH slice launch");
        return poh_slice(*res[s], b) ? 1 : -1;
    }

    void drain(int s) {
        InCall c(*this, "hipStreamSynchronize");
        (void)hipStreamSynchronize(res[s]->stream);
    }

    // `"loop_in": "<operation>" | null, "loop_in_ms": <ms>`
    std::string state_json() const {
        const char *what = in_call.load(std::memory_order_acquire);
        char buf[160];
        if (!what) return "\"loop_in\": null, \"loop_in_ms\": 0";
        const double ms =
            (double)(serve::now_ns() - in_call_since.load(std::memory_order_relaxed)) / 1e6;
        snprintf(buf, sizeof buf, "\"loop_in\": \"%s\", \"loop_in_ms\": %.3f", what, ms);
        return buf;
    }
};

}  // namespace
}  // namespace spl

using namespace spl;

struct spl_server {
    HipDev dev;
    RingRes *ring_res = nullptr;
    std::unique_ptr<serve::Core<HipDev>> core;

    // The loop has stopped (or never started): the pooled resources go back
    // with no work on them.
    ~spl_server() {
        core.reset();
        for (SlotRes *r : dev.res)
            if (r) give_back(r);
        if (ring_res) give_back(ring_res);
    }
};

extern "C" {

spl_server *spl_server_create(uint64_t hbm_budget_bytes, uint32_t max_batch_signatures,
                              uint32_t verify_streams, uint32_t poh_streams) {
    if (hbm_budget_bytes == 0) hbm_budget_bytes = 1ull << 30;
    if (max_batch_signatures == 0) max_batch_signatures = 65536;
    if (verify_streams == 0) verify_streams = 3;
    if (poh_streams == 0) poh_streams = 1;
    if (verify_streams + poh_streams > 16 || hbm_budget_bytes > (1ull << 32)) return nullptr;
    if (pool_init() <= 0) return nullptr;
    auto *sv = new spl_server();
    HipDev &hd = sv->dev;
    bool ok = hipGetDevice(&hd.dev) == hipSuccess;
    if (ok) hd.T = device_tables(hd.dev);
    ok = ok && hd.T;
    const uint64_t cap = align_up(hbm_budget_bytes, 256);
    if (ok) {
        hd.cus = device_cus(hd.dev);
        sv->ring_res = take_ring_res(hd.dev, cap);
        ok = sv->ring_res != nullptr;
        if (ok) {
            hd.d_ring = sv->ring_res->d;
            hd.h_ring = sv->ring_res->h;
        }
    }
    std::vector<Kind> kinds;
    for (uint32_t k = 0; ok && k < verify_streams + poh_streams; k++) {
        const Kind kind = k < verify_streams ? kVerify : kPoh;
        SlotRes *r = take_slot_res(hd.dev);
        ok = r != nullptr;
        if (!ok) break;
        hd.res.push_back(r);
        kinds.push_back(kind);
        if (kind == kVerify) {
            // sized for a full batch under the widest one-shot configuration,
            // so the loop never allocates in steady state
            const VerifyConfig widest = {0, 5, 1, 0, 0, 1, 1};
            ok = ensure_scratch(*r, verify_scratch_bytes(widest, max_batch_signatures, hd.cus)) ==
                 hipSuccess;
        }
    }
    if (!ok) {
        fprintf(stderr, "spl_server_create: failed\n");
        delete sv;
        return nullptr;
    }
    sv->core.reset(new serve::Core<HipDev>(hd, hd.h_ring, cap, max_batch_signatures, kinds));
    sv->core->start();
    return sv;
}

void spl_server_destroy(spl_server *sv) {
    if (!sv) return;
    sv->core->stop();
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
    // packets already in pinned memory (how solana_perf holds its packet
    // batches) are read by the device in place, not staged by the loop
    if (num_signatures && packets_size % 16 == 0 && (uintptr_t)packets % 16 == 0) {
        const uint8_t *dp = pinned_device_range(packets, packets_size);
        if (dp && (uintptr_t)dp % 16 == 0) {
            r->direct = true;
            r->packets_dev = dp;
        }
    }
    return sv->core->submit(r);
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
    return sv->core->submit(r);
}

int spl_server_wait(spl_server *sv, int64_t ticket, int timeout_ms, uint64_t *times_ns) {
    if (!sv) return 2;
    return sv->core->wait(ticket, timeout_ms, times_ns);
}

int spl_server_stats(spl_server *sv, char *buf, size_t len, int reset) {
    if (!sv || !buf || !len) return 2;
    std::string s(65536, '\0');
    if (sv->core->stats(&s[0], s.size(), reset != 0) != 0) return 2;
    s.resize(strlen(s.c_str()));
    // the core's object, with the loop thread's device operation added
    s.pop_back();
    s += ", " + sv->dev.state_json() + "}";
    snprintf(buf, len, "%s", s.c_str());
    return s.size() < len ? 0 : 2;
}

}  // extern "C"
