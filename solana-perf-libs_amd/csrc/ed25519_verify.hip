// Batched Ed25519 verification: the reference's ed25519_verify_many ABI, the
// device-resident entry point, and the verify engine (graph-captured
// persistent launch + autotuner).
//
// Reference: src/cuda-ecc-ed25519/verify.cu:240-340 (ABI, three launches),
// gpu_ctx.cu:75-144 (per-elem H2D ingest).
#include <hip/hip_runtime.h>

#include <stdio.h>

#include <string>
#include <vector>

#include "verify_kernels.h"
#include "../../include/cuda_crypt.h"

namespace spl {

// ------------------------------------------------------------ tuning cache

// Best configuration per (device, size class = bit length of n), filled by
// the autotuner; the ABI and device entry points use the best one-shot
// configuration of the class, the engine uses the overall best.
struct TuneEntry {
    bool valid = false;
    VerifyConfig best, best_plain;
};
static std::mutex g_tune_mu;
static TuneEntry g_tune[kMaxDevices][33];

static int size_class(uint32_t n) { return n ? 32 - __builtin_clz(n) : 0; }

VerifyConfig tuned_plain_config(int dev, uint32_t n) {
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        if (dev >= 0 && dev < kMaxDevices) {
            const TuneEntry &e = g_tune[dev][size_class(n)];
            if (e.valid) return e.best_plain;
        }
    }
    // Untuned: a launch that fits one wave per SIMD is bound by one
    // signature's latency, which the staged and two-lane forms shorten;
    // beyond that their extra waves cost more than they save.
    return untuned_plain_config(n, 4u * (uint32_t)device_cus(dev));
}

int device_cus(int dev) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
        cus = 256;
    return cus;
}

// Packs gpu_Elems and the four offset tables into the queue's pinned staging
// buffer and copies them to the device with two copy kernels (the reference's
// setup_gpu_ctx, gpu_ctx.cu:75-144, makes one DMA copy per elem).
hipError_t stage_packets(QueueCtx &q, const gpu_Elems *elems, uint32_t num_elems,
                         uint32_t message_size, uint64_t packets_bytes, uint32_t n,
                         const uint32_t *t0, const uint32_t *t1, const uint32_t *t2,
                         const uint32_t *t3, hipStream_t s) {
    hipError_t e;
    const uint64_t tab_bytes = 16ull * n;
    const uint64_t tab_at = (packets_bytes + 15) & ~15ull;   // 16-byte aligned for the copy
    if ((e = q.d_packets.reserve(packets_bytes + 16)) != hipSuccess) return e;
    if ((e = q.d_offsets.reserve(tab_bytes)) != hipSuccess) return e;
    if ((e = q.h_stage.reserve(tab_at + tab_bytes + 64 * (uint64_t)n + 16)) != hipSuccess)
        return e;
    uint8_t *h = q.h_stage.as<uint8_t>();
    uint8_t *d = q.d_packets.as<uint8_t>();
    uint64_t sent = 0;
    if ((e = stage_elems(h, d, elems, num_elems, message_size, packets_bytes, &sent, s)) !=
        hipSuccess)
        return e;
    memset(h + packets_bytes, 0, tab_at - packets_bytes);   // pad to the tables
    uint32_t *tab = reinterpret_cast<uint32_t *>(h + tab_at);
    memcpy(tab, t0, 4ull * n);
    memcpy(tab + n, t1, 4ull * n);
    memcpy(tab + 2ull * n, t2, 4ull * n);
    memcpy(tab + 3ull * n, t3, 4ull * n);
    if ((e = launch_copy16(d + sent, h + sent, packets_bytes - sent, s)) != hipSuccess) return e;
    return launch_copy16(q.d_offsets.ptr, tab, tab_bytes, s);
}

VerifyJob make_verify_job(const uint8_t *packets, uint64_t size, const uint32_t *lens,
                          const uint32_t *pk, const uint32_t *sig, const uint32_t *msg,
                          uint8_t *out, void *scratch, const ed25519_tables *T) {
    VerifyJob j{};
    j.packets = packets;
    j.packets_size = size;
    j.lens = lens;
    j.pk_off = pk;
    j.sig_off = sig;
    j.msg_off = msg;
    j.out = out;
    j.scratch = static_cast<ge_cached *>(scratch);
    j.T = T;
    return j;
}

}  // namespace spl

// ------------------------------------------------------------ verify engine

// Device buffers sized for a capacity and one instantiated graph of the tuned
// launch sequence, in `depth` buffer sets.  Each set has its own stream,
// buffers, scratch, work counters and graph, so batches in different sets run
// concurrently: one launch of a validator batch leaves a third of the SIMDs
// idle (651 waves of 64 signatures on 1,024 SIMDs), and the next set's launch
// fills them.  Inputs are written by the caller straight into a set's
// buffers; run() sets the signature count in device memory (outside the
// graph) and replays the set's graph.
struct spl_verify_engine {
    struct Set {
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;           // recorded after each run of the set
        bool ran = false;
        spl::Buf packets, tabs, out, scratch, work, order, n_dev;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };

    int dev = 0;
    int cus = 256;
    uint32_t cap_sigs = 0;
    uint64_t cap_bytes = 0;
    std::vector<Set> sets;
    uint32_t next_set = 0;
    spl::VerifyConfig cfg = spl::kDefaultVerifyConfig;
    const spl::ed25519_tables *T = nullptr;

    spl::VerifyJob job(const Set &s) const {
        const uint32_t *t = s.tabs.as<uint32_t>();
        spl::VerifyJob j = spl::make_verify_job(s.packets.as<uint8_t>(), cap_bytes, t,
                                                t + cap_sigs, t + 2ull * cap_sigs,
                                                t + 3ull * cap_sigs, s.out.as<uint8_t>(),
                                                s.scratch.ptr, T);
        j.n_dev = s.n_dev.as<uint32_t>();
        j.work = s.work.as<uint32_t>();
        j.order = s.order.as<uint32_t>();
        return j;
    }

    hipError_t alloc_set(Set &s) {
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking)) != hipSuccess ||
            (e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming)) != hipSuccess ||
            (e = s.packets.reserve(cap_bytes + 16)) != hipSuccess ||
            (e = s.tabs.reserve(16ull * cap_sigs)) != hipSuccess ||
            (e = s.out.reserve(cap_sigs)) != hipSuccess ||
            (e = s.work.reserve(64 * 4)) != hipSuccess ||
            (e = s.order.reserve(4ull * cap_sigs)) != hipSuccess ||
            (e = s.n_dev.reserve(4)) != hipSuccess)
            return e;
        return hipSuccess;
    }

    void free_set(Set &s) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        drop_graph(s);
        s.packets.release();
        s.tabs.release();
        s.out.release();
        s.scratch.release();
        s.work.release();
        s.order.release();
        s.n_dev.release();
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
        s = Set{};
    }

    static void drop_graph(Set &s) {
        if (s.exec) (void)hipGraphExecDestroy(s.exec);
        if (s.graph) (void)hipGraphDestroy(s.graph);
        s.exec = nullptr;
        s.graph = nullptr;
    }

    // Captures the launch sequence of `cfg` for one set (grid sized for the
    // capacity; the kernels read the count from the set's n_dev).
    hipError_t capture(Set &s) {
        drop_graph(s);
        hipError_t e = s.scratch.reserve(spl::verify_scratch_bytes(cfg, cap_sigs, cus));
        if (e != hipSuccess) return e;
        e = hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t le = spl::launch_verify_cfg(cfg, job(s), cap_sigs, cus, s.stream);
        e = hipStreamEndCapture(s.stream, &s.graph);
        if (le != hipSuccess) return le;
        if (e != hipSuccess) return e;
        return hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0);
    }

    hipError_t capture_all() {
        for (Set &s : sets) {
            if (s.stream) {
                const hipError_t e = hipStreamSynchronize(s.stream);
                if (e != hipSuccess) return e;
            }
            const hipError_t e = capture(s);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }

    hipError_t run(Set &s, uint32_t n) {
        spl::launch_set_n(s.n_dev.as<uint32_t>(), n, s.stream);
        hipError_t e = hipGraphLaunch(s.exec, s.stream);
        if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
        if (e == hipSuccess) s.ran = true;
        return e;
    }
};

using namespace spl;

extern "C" {

void ed25519_set_verbose(bool val) { g_verbose = val; }

void ed25519_verify_many(const gpu_Elems *elems, uint32_t num_elems, uint32_t message_size,
                         uint32_t total_packets, uint32_t total_signatures,
                         const uint32_t *message_lens, const uint32_t *public_key_offsets,
                         const uint32_t *signature_offsets,
                         const uint32_t *message_start_offsets, uint8_t *out,
                         uint8_t use_non_default_stream) {
    SPL_LOG("ed25519_verify_many: elems %u sigs %u packets %u message_size %u\n", num_elems,
            total_signatures, total_packets, message_size);
    if (total_packets == 0) return;
    uint32_t n = total_signatures;
    Lease l = acquire_queue();
    if (!l.q) {
        fprintf(stderr, "ed25519_verify_many: no GPU available; all %u signatures rejected\n",
                n);
        if (n) memset(out, 0, n);
        return;
    }
    QueueCtx &q = *l.q;
    hipStream_t s = use_non_default_stream ? q.stream : nullptr;
    uint64_t pbytes = (uint64_t)total_packets * message_size;
    const uint32_t *tab = nullptr;
    uint8_t *hout = nullptr;
    const VerifyConfig cfg = tuned_plain_config(l.dev->id, n);
    VerifyJob job;
    SPL_HIP_TRY(stage_packets(q, elems, num_elems, message_size, pbytes, n, message_lens,
                              public_key_offsets, signature_offsets, message_start_offsets, s));
    SPL_HIP_TRY(q.d_out.reserve(n + 16));
    SPL_HIP_TRY(q.d_scratch.reserve(verify_scratch_bytes(cfg, n, 0)));
    tab = q.d_offsets.as<uint32_t>();
    job = make_verify_job(q.d_packets.as<uint8_t>(), pbytes, tab, tab + n, tab + 2ull * n,
                          tab + 3ull * n, q.d_out.as<uint8_t>(), q.d_scratch.ptr,
                          l.dev->d_tables);
    SPL_HIP_TRY(launch_verify_cfg(cfg, job, n, 0, s));
    hout = q.h_stage.as<uint8_t>();
    SPL_HIP_TRY(launch_copy16(hout, q.d_out.ptr, n, s));
    SPL_HIP_TRY(hipStreamSynchronize(s));
    memcpy(out, hout, n);
    release_queue(l);
    return;
fail:
    if (n) memset(out, 0, n);
    release_queue(l);
}

int ed25519_verify_many_device(const uint8_t *d_packets, uint64_t packets_size,
                               const uint32_t *d_message_lens, const uint32_t *d_public_key_offsets,
                               const uint32_t *d_signature_offsets,
                               const uint32_t *d_message_start_offsets, uint32_t total_signatures,
                               uint8_t *d_out, void *stream) {
    if (total_signatures == 0) return 0;
    if (pool_init() <= 0) return 1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1;
    const ed25519_tables *tabs = device_tables(dev);
    if (!tabs) return 1;
    const VerifyConfig cfg = tuned_plain_config(dev, total_signatures);
    void *scr = stream_scratch(dev, (hipStream_t)stream,
                               verify_scratch_bytes(cfg, total_signatures, 0));
    if (!scr) {
        fprintf(stderr, "ed25519_verify_many_device: scratch allocation failed\n");
        return 3;
    }
    VerifyJob job = make_verify_job(d_packets, packets_size, d_message_lens,
                                    d_public_key_offsets, d_signature_offsets,
                                    d_message_start_offsets, d_out, scr, tabs);
    hipError_t e = launch_verify_cfg(cfg, job, total_signatures, 0, (hipStream_t)stream);
    if (e != hipSuccess) {
        fprintf(stderr, "ed25519_verify_many_device: launch failed: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

spl_verify_engine *spl_verify_engine_create_pipelined(uint32_t max_signatures,
                                                      uint64_t max_packet_bytes, uint32_t depth) {
    if (max_signatures == 0 || max_packet_bytes == 0 || depth == 0 || depth > 8) return nullptr;
    if (pool_init() <= 0) return nullptr;
    auto *eng = new spl_verify_engine();
    bool ok = hipGetDevice(&eng->dev) == hipSuccess;
    if (ok) eng->T = device_tables(eng->dev);
    ok = ok && eng->T;
    if (ok) {
        eng->cus = device_cus(eng->dev);
        eng->cap_sigs = max_signatures;
        eng->cap_bytes = max_packet_bytes;
        eng->sets.resize(depth);
        for (auto &st : eng->sets) ok = ok && eng->alloc_set(st) == hipSuccess;
    }
    if (ok && depth == 1) {
        std::lock_guard<std::mutex> g(g_tune_mu);
        const TuneEntry &t = g_tune[eng->dev][size_class(max_signatures)];
        if (t.valid) eng->cfg = t.best;
    }
    ok = ok && eng->capture_all() == hipSuccess;
    if (!ok) {
        fprintf(stderr, "spl_verify_engine_create: failed\n");
        spl_verify_engine_destroy(eng);
        return nullptr;
    }
    return eng;
}

spl_verify_engine *spl_verify_engine_create(uint32_t max_signatures, uint64_t max_packet_bytes) {
    return spl_verify_engine_create_pipelined(max_signatures, max_packet_bytes, 1);
}

void spl_verify_engine_destroy(spl_verify_engine *eng) {
    if (!eng) return;
    for (auto &st : eng->sets) eng->free_set(st);
    delete eng;
}

int spl_verify_engine_depth(spl_verify_engine *eng) { return eng ? (int)eng->sets.size() : -2; }

int spl_verify_engine_set_buffers(spl_verify_engine *eng, uint32_t set, uint8_t **d_packets,
                                  uint32_t **d_lens, uint32_t **d_pk_off, uint32_t **d_sig_off,
                                  uint32_t **d_msg_off, uint8_t **d_out, void **stream) {
    if (!eng || set >= eng->sets.size()) return 2;
    const spl_verify_engine::Set &st = eng->sets[set];
    uint32_t *t = st.tabs.as<uint32_t>();
    if (d_packets) *d_packets = st.packets.as<uint8_t>();
    if (d_lens) *d_lens = t;
    if (d_pk_off) *d_pk_off = t + eng->cap_sigs;
    if (d_sig_off) *d_sig_off = t + 2ull * eng->cap_sigs;
    if (d_msg_off) *d_msg_off = t + 3ull * eng->cap_sigs;
    if (d_out) *d_out = st.out.as<uint8_t>();
    if (stream) *stream = st.stream;
    return 0;
}

int spl_verify_engine_buffers(spl_verify_engine *eng, uint8_t **d_packets, uint32_t **d_lens,
                              uint32_t **d_pk_off, uint32_t **d_sig_off, uint32_t **d_msg_off,
                              uint8_t **d_out, void **stream) {
    return spl_verify_engine_set_buffers(eng, 0, d_packets, d_lens, d_pk_off, d_sig_off,
                                         d_msg_off, d_out, stream);
}

int spl_verify_engine_next(spl_verify_engine *eng) {
    if (!eng) return -2;
    const uint32_t k = eng->next_set;
    spl_verify_engine::Set &st = eng->sets[k];
    if (st.ran && hipEventSynchronize(st.done) != hipSuccess) return -3;
    eng->next_set = (k + 1) % (uint32_t)eng->sets.size();
    return (int)k;
}

int spl_verify_engine_run_set(spl_verify_engine *eng, uint32_t set, uint32_t n) {
    if (!eng || set >= eng->sets.size()) return 2;
    if (n > eng->cap_sigs) return 2;
    if (n == 0) return 0;
    const hipError_t e = eng->run(eng->sets[set], n);
    if (e != hipSuccess) {
        fprintf(stderr, "spl_verify_engine_run: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

int spl_verify_engine_run(spl_verify_engine *eng, uint32_t n) {
    return spl_verify_engine_run_set(eng, 0, n);
}

int spl_verify_engine_sync_set(spl_verify_engine *eng, uint32_t set) {
    if (!eng || set >= eng->sets.size()) return 2;
    return hipStreamSynchronize(eng->sets[set].stream) == hipSuccess ? 0 : 3;
}

int spl_verify_engine_sync(spl_verify_engine *eng) {
    if (!eng) return 2;
    int rc = 0;
    for (uint32_t k = 0; k < eng->sets.size(); k++)
        if (spl_verify_engine_sync_set(eng, k) != 0) rc = 3;
    return rc;
}

int spl_verify_engine_config(spl_verify_engine *eng, char *buf, size_t len) {
    if (!eng || !buf || !len) return 2;
    config_name(eng->cfg, buf, len);
    return 0;
}

int spl_verify_engine_candidates(spl_verify_engine *eng, char *buf, size_t len) {
    if (!eng) return -2;
    VerifyConfig cand[64];
    const int nc = verify_config_candidates(cand, 64);
    if (buf && len) {
        std::string all;
        for (int i = 0; i < nc; i++) {
            char line[128];
            config_name(cand[i], line, sizeof line);
            all += line;
            all += "\n";
        }
        snprintf(buf, len, "%s", all.c_str());
    }
    return nc;
}

int spl_verify_engine_use(spl_verify_engine *eng, int index) {
    if (!eng) return 2;
    VerifyConfig cand[64];
    const int nc = verify_config_candidates(cand, 64);
    if (index < 0 || index >= nc) return 2;
    eng->cfg = cand[index];
    return eng->capture_all() == hipSuccess ? 0 : 3;
}

// Times every candidate configuration on the batch currently in the engine's
// buffer sets (n signatures in each), keeps the fastest for this engine
// (recapturing its graphs) and, for an engine of one set, records it for n's
// size class on this device.
int spl_verify_engine_autotune(spl_verify_engine *eng, uint32_t n, char *report,
                               size_t report_len) {
    if (!eng || n == 0 || n > eng->cap_sigs) return 2;
    VerifyConfig cand[64];
    const int nc = verify_config_candidates(cand, 64);
    const uint32_t depth = (uint32_t)eng->sets.size();
    hipStream_t s0 = eng->sets[0].stream;
    hipEvent_t e0, e1, mark[8];
    int made = 0;
    bool ev_ok = hipEventCreate(&e0) == hipSuccess;
    if (ev_ok && hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        ev_ok = false;
    }
    for (; ev_ok && made < (int)depth; made++)
        if (hipEventCreateWithFlags(&mark[made], hipEventDisableTiming) != hipSuccess) break;
    if (!ev_ok || made < (int)depth) {
        for (int k = 0; k < made; k++) (void)hipEventDestroy(mark[k]);
        if (ev_ok) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        return 3;
    }
    float best_ms = 1e30f, best_plain_ms = 1e30f;
    const VerifyConfig before = eng->cfg;
    VerifyConfig best = eng->cfg, best_plain = kDefaultVerifyConfig;
    std::string rep;
    int rc = 0;
    // Each candidate is timed as the engine runs it: its launch sequence
    // captured and instantiated as every set's graph, and steps (write n,
    // replay the graph) issued back to back, round-robin over the sets: one
    // per set to warm the caches, then kTimedSteps per set between the events
    // (every set's stream starts after e0, and e1 follows the last step of
    // every set); the time per step is the candidate's time.  Timing the
    // kernels launched one at a time, or the best of single replays,
    // favoured the persistent forms (their prepass kernels are separate graph
    // nodes, and their steps vary more): one tied the one-shot form there and
    // then ran back to back 1 % slower.
    constexpr int kTimedSteps = 4;
    for (int i = 0; i < nc && rc == 0; i++) {
        const VerifyConfig &c = cand[i];
        eng->cfg = c;
        if (eng->capture_all() != hipSuccess) continue;
        float ms = 0;
        for (auto &st : eng->sets)
            if (eng->run(st, n) != hipSuccess) rc = 3;
        for (auto &st : eng->sets)
            if (rc == 0 && hipStreamSynchronize(st.stream) != hipSuccess) rc = 3;
        if (rc == 0 && hipEventRecord(e0, s0) != hipSuccess) rc = 3;
        for (uint32_t k = 1; k < depth && rc == 0; k++)
            if (hipStreamWaitEvent(eng->sets[k].stream, e0, 0) != hipSuccess) rc = 3;
        for (int r = 0; r < kTimedSteps * (int)depth && rc == 0; r++)
            if (eng->run(eng->sets[r % depth], n) != hipSuccess) rc = 3;
        for (uint32_t k = 1; k < depth && rc == 0; k++)
            if (hipEventRecord(mark[k], eng->sets[k].stream) != hipSuccess ||
                hipStreamWaitEvent(s0, mark[k], 0) != hipSuccess)
                rc = 3;
        if (rc == 0 && (hipEventRecord(e1, s0) != hipSuccess ||
                        hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = 3;
        ms /= (float)(kTimedSteps * depth);
        if (rc) break;
        char name[128], line[160];
        config_name(c, name, sizeof name);
        snprintf(line, sizeof line, "%s: %.3f ms\n", name, ms);
        rep += line;
        if (ms < best_ms) { best_ms = ms; best = c; }
        if (!c.persistent && ms < best_plain_ms) { best_plain_ms = ms; best_plain = c; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < made; k++) (void)hipEventDestroy(mark[k]);
    if (rc) {
        eng->cfg = before;                     // leave the engine as it was
        (void)eng->capture_all();
        return rc;
    }
    eng->cfg = best;
    if (eng->capture_all() != hipSuccess) return 3;
    if (depth == 1) {
        std::lock_guard<std::mutex> g(g_tune_mu);
        TuneEntry &t = g_tune[eng->dev][size_class(n)];
        t.valid = true;
        t.best = best;
        t.best_plain = best_plain;
    }
    if (report && report_len) snprintf(report, report_len, "%s", rep.c_str());
    return 0;
}

// ------------------------------------------------------------ tuning records

}  // extern "C"

namespace {

// "<arch> <cus>" of HIP device dev, or "" when it cannot be queried.
std::string device_identity(int dev) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        return "";
    }
    std::string arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch.resize(colon);   // drop feature flags
    return arch + " " + std::to_string(prop.multiProcessorCount);
}

// Index of the candidate configuration called `name`, or -1.
int candidate_index(const std::string &name) {
    VerifyConfig cand[64];
    const int nc = verify_config_candidates(cand, 64);
    for (int i = 0; i < nc; i++) {
        char buf[128];
        config_name(cand[i], buf, sizeof buf);
        if (name == buf) return i;
    }
    return -1;
}

int hip_device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

}  // namespace

extern "C" {

// One line per recorded (device, size class):
//   device <id> <arch> <cus> class <c> best <configuration> | plain <configuration>
// where a configuration is spelled as spl_verify_engine_config prints it.
int spl_verify_tune_save(const char *path) {
    if (!path) return -2;
    FILE *f = fopen(path, "w");
    if (!f) return -3;
    fprintf(f, "# libcuda-crypt verify tuning records, one per device and batch-size class\n");
    int count = 0;
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (int d = 0; d < kMaxDevices; d++) {
        std::string id;
        for (int c = 0; c < 33; c++) {
            const TuneEntry &t = g_tune[d][c];
            if (!t.valid) continue;
            if (id.empty()) id = device_identity(d);
            char best[128], plain[128];
            config_name(t.best, best, sizeof best);
            config_name(t.best_plain, plain, sizeof plain);
            fprintf(f, "device %d %s class %d best %s | plain %s\n", d,
                    id.empty() ? "unknown 0" : id.c_str(), c, best, plain);
            count++;
        }
    }
    return fclose(f) == 0 ? count : -3;
}

// Returns the number of records taken, or < 0 when the file cannot be read
// or holds a malformed line (then nothing is taken).  A record for a device
// this process sees with another architecture or CU count is skipped.
int spl_verify_tune_load(const char *path) {
    if (!path) return -2;
    FILE *f = fopen(path, "r");
    if (!f) return -3;
    struct Rec {
        int dev, cls;
        VerifyConfig best, plain;
    };
    std::vector<Rec> recs;
    VerifyConfig cand[64];
    verify_config_candidates(cand, 64);
    const int ndev = hip_device_count();
    char line[512];
    bool bad = false;
    while (fgets(line, sizeof line, f)) {
        std::string ln(line);
        while (!ln.empty() && (ln.back() == '\n' || ln.back() == '\r')) ln.pop_back();
        if (ln.empty() || ln[0] == '#') continue;
        int dev = -1, cus = 0, cls = -1, used = 0;
        char arch[64];
        if (sscanf(ln.c_str(), "device %d %63s %d class %d best %n", &dev, arch, &cus, &cls,
                   &used) != 4 ||
            used == 0 || dev < 0 || dev >= kMaxDevices || cls < 0 || cls > 32) {
            bad = true;
            break;
        }
        const std::string rest = ln.substr((size_t)used);
        const size_t bar = rest.find(" | plain ");
        if (bar == std::string::npos) {
            bad = true;
            break;
        }
        const int ib = candidate_index(rest.substr(0, bar));
        const int ip = candidate_index(rest.substr(bar + 9));
        if (ib < 0 || ip < 0 || cand[ip].persistent) {
            bad = true;
            break;
        }
        if (dev < ndev && device_identity(dev) != std::string(arch) + " " + std::to_string(cus))
            continue;                                 // recorded on another kind of GPU
        recs.push_back(Rec{dev, cls, cand[ib], cand[ip]});
    }
    fclose(f);
    if (bad) return -4;
    std::lock_guard<std::mutex> g(g_tune_mu);
    for (const Rec &r : recs) {
        TuneEntry &t = g_tune[r.dev][r.cls];
        t.valid = true;
        t.best = r.best;
        t.best_plain = r.plain;
    }
    return (int)recs.size();
}

// The one-shot configuration the ABI and device entry points use for n
// signatures on HIP device `device`; returns 1 when it comes from a tuning
// record, 0 when it is the untuned rule, < 0 on bad arguments.
int spl_verify_tune_query(int device, uint32_t n, char *buf, size_t len) {
    if (device < 0 || device >= kMaxDevices || !buf || !len) return -2;
    bool tuned;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        tuned = g_tune[device][size_class(n)].valid;
    }
    config_name(tuned_plain_config(device, n), buf, len);
    return tuned ? 1 : 0;
}

}  // extern "C"
