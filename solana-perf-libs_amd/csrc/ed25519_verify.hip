// Batched Ed25519 verification: the reference's ed25519_verify_many ABI, the
// device-resident entry point, and the verify engine (graph-captured
// persistent launch + autotuner).
//
// Reference: src/cuda-ecc-ed25519/verify.cu:240-340 (ABI, three launches),
// gpu_ctx.cu:75-144 (per-elem H2D ingest).
#include <hip/hip_runtime.h>

#include <string>

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
    // Untuned: a launch that fits one wave per SIMD is bound by one lane's
    // latency, which two lanes per signature shorten (measured 0.69-0.75 ms
    // against 0.93-0.98 ms for 512-16384 signatures); beyond that the pair
    // kernel's extra total work costs more than it saves.
    const uint32_t simds = 4u * (uint32_t)device_cus(dev);
    if (2ull * n <= 64ull * simds) return kPairVerifyConfig;
    return kDefaultVerifyConfig;
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

// Device buffers sized for a capacity, one stream, and one instantiated graph
// of the tuned launch sequence.  Inputs are written by the caller straight
// into the engine's buffers; run() sets the signature count in device memory
// (outside the graph) and replays the graph.
struct spl_verify_engine {
    int dev = 0;
    int cus = 256;
    hipStream_t stream = nullptr;
    uint32_t cap_sigs = 0;
    uint64_t cap_bytes = 0;
    spl::Buf packets, tabs, out, scratch, work, order, n_dev;
    spl::VerifyConfig cfg = spl::kDefaultVerifyConfig;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const spl::ed25519_tables *T = nullptr;

    spl::VerifyJob job() const {
        const uint32_t *t = tabs.as<uint32_t>();
        spl::VerifyJob j = spl::make_verify_job(packets.as<uint8_t>(), cap_bytes, t,
                                                t + cap_sigs, t + 2ull * cap_sigs,
                                                t + 3ull * cap_sigs, out.as<uint8_t>(),
                                                scratch.ptr, T);
        j.n_dev = n_dev.as<uint32_t>();
        j.work = work.as<uint32_t>();
        j.order = order.as<uint32_t>();
        return j;
    }

    hipError_t ensure_scratch(const spl::VerifyConfig &c) {
        return scratch.reserve(spl::verify_scratch_bytes(c, cap_sigs, cus));
    }

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }

    // Captures the launch sequence of `cfg` (grid sized for the capacity; the
    // kernels read the count from n_dev).
    hipError_t capture() {
        drop_graph();
        hipError_t e = ensure_scratch(cfg);
        if (e != hipSuccess) return e;
        e = hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        hipError_t le = spl::launch_verify_cfg(cfg, job(), cap_sigs, cus, stream);
        e = hipStreamEndCapture(stream, &graph);
        if (le != hipSuccess) return le;
        if (e != hipSuccess) return e;
        return hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
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

spl_verify_engine *spl_verify_engine_create(uint32_t max_signatures, uint64_t max_packet_bytes) {
    if (max_signatures == 0 || max_packet_bytes == 0) return nullptr;
    if (pool_init() <= 0) return nullptr;
    auto *eng = new spl_verify_engine();
    bool ok = hipGetDevice(&eng->dev) == hipSuccess;
    if (ok) eng->T = device_tables(eng->dev);
    ok = ok && eng->T;
    if (ok) {
        eng->cus = device_cus(eng->dev);
        eng->cap_sigs = max_signatures;
        eng->cap_bytes = max_packet_bytes;
        ok = hipStreamCreate(&eng->stream) == hipSuccess &&
             eng->packets.reserve(max_packet_bytes + 16) == hipSuccess &&
             eng->tabs.reserve(16ull * max_signatures) == hipSuccess &&
             eng->out.reserve(max_signatures) == hipSuccess &&
             eng->work.reserve(64 * 4) == hipSuccess &&
             eng->order.reserve(4ull * max_signatures) == hipSuccess &&
             eng->n_dev.reserve(4) == hipSuccess;
    }
    if (ok) {
        std::lock_guard<std::mutex> g(g_tune_mu);
        const TuneEntry &t = g_tune[eng->dev][size_class(max_signatures)];
        if (t.valid) eng->cfg = t.best;
    }
    ok = ok && eng->capture() == hipSuccess;
    if (!ok) {
        fprintf(stderr, "spl_verify_engine_create: failed\n");
        spl_verify_engine_destroy(eng);
        return nullptr;
    }
    return eng;
}

void spl_verify_engine_destroy(spl_verify_engine *eng) {
    if (!eng) return;
    if (eng->stream) (void)hipStreamSynchronize(eng->stream);
    eng->drop_graph();
    eng->packets.release();
    eng->tabs.release();
    eng->out.release();
    eng->scratch.release();
    eng->work.release();
    eng->order.release();
    eng->n_dev.release();
    if (eng->stream) (void)hipStreamDestroy(eng->stream);
    delete eng;
}

int spl_verify_engine_buffers(spl_verify_engine *eng, uint8_t **d_packets, uint32_t **d_lens,
                              uint32_t **d_pk_off, uint32_t **d_sig_off, uint32_t **d_msg_off,
                              uint8_t **d_out, void **stream) {
    if (!eng) return 2;
    uint32_t *t = eng->tabs.as<uint32_t>();
    if (d_packets) *d_packets = eng->packets.as<uint8_t>();
    if (d_lens) *d_lens = t;
    if (d_pk_off) *d_pk_off = t + eng->cap_sigs;
    if (d_sig_off) *d_sig_off = t + 2ull * eng->cap_sigs;
    if (d_msg_off) *d_msg_off = t + 3ull * eng->cap_sigs;
    if (d_out) *d_out = eng->out.as<uint8_t>();
    if (stream) *stream = eng->stream;
    return 0;
}

int spl_verify_engine_run(spl_verify_engine *eng, uint32_t n) {
    if (!eng) return 2;
    if (n > eng->cap_sigs) return 2;
    if (n == 0) return 0;
    launch_set_n(eng->n_dev.as<uint32_t>(), n, eng->stream);
    hipError_t e = hipGraphLaunch(eng->exec, eng->stream);
    if (e != hipSuccess) {
        fprintf(stderr, "spl_verify_engine_run: %s\n", hipGetErrorString(e));
        return 3;
    }
    return 0;
}

int spl_verify_engine_sync(spl_verify_engine *eng) {
    if (!eng) return 2;
    return hipStreamSynchronize(eng->stream) == hipSuccess ? 0 : 3;
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
    return eng->capture() == hipSuccess ? 0 : 3;
}

// Times every candidate configuration on the batch currently in the engine's
// buffers (n signatures), keeps the fastest for this engine (recapturing its
// graph) and records it for n's size class on this device.
int spl_verify_engine_autotune(spl_verify_engine *eng, uint32_t n, char *report,
                               size_t report_len) {
    if (!eng || n == 0 || n > eng->cap_sigs) return 2;
    VerifyConfig cand[64];
    const int nc = verify_config_candidates(cand, 64);
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return 3;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return 3;
    }
    float best_ms = 1e30f, best_plain_ms = 1e30f;
    const VerifyConfig before = eng->cfg;
    VerifyConfig best = eng->cfg, best_plain = kDefaultVerifyConfig;
    std::string rep;
    int rc = 0;
    // Each candidate is timed as the engine runs it: its launch sequence
    // captured and instantiated as the graph, and steps of
    // spl_verify_engine_run (write n, replay the graph) issued back to back:
    // one to warm the caches, then four between the events, whose mean is the
    // candidate's time.  Timing the kernels launched one at a time, or the
    // best of single replays, favoured the persistent forms (their prepass
    // kernels are separate graph nodes, and their steps vary more): one tied
    // the one-shot form there and then ran back to back 1 % slower.
    constexpr int kTimedSteps = 4;
    for (int i = 0; i < nc && rc == 0; i++) {
        const VerifyConfig &c = cand[i];
        eng->cfg = c;
        if (eng->capture() != hipSuccess) continue;
        float ms = 0;
        launch_set_n(eng->n_dev.as<uint32_t>(), n, eng->stream);
        if (hipGraphLaunch(eng->exec, eng->stream) != hipSuccess ||
            hipStreamSynchronize(eng->stream) != hipSuccess ||
            hipEventRecord(e0, eng->stream) != hipSuccess)
            rc = 3;
        for (int r = 0; r < kTimedSteps && rc == 0; r++) {
            launch_set_n(eng->n_dev.as<uint32_t>(), n, eng->stream);
            if (hipGraphLaunch(eng->exec, eng->stream) != hipSuccess) rc = 3;
        }
        if (rc == 0 && (hipEventRecord(e1, eng->stream) != hipSuccess ||
                        hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = 3;
        ms /= kTimedSteps;
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
    if (rc) {
        eng->cfg = before;                     // leave the engine as it was
        (void)eng->capture();
        return rc;
    }
    eng->cfg = best;
    if (eng->capture() != hipSuccess) return 3;
    {
        std::lock_guard<std::mutex> g(g_tune_mu);
        TuneEntry &t = g_tune[eng->dev][size_class(n)];
        t.valid = true;
        t.best = best;
        t.best_plain = best_plain;
    }
    if (report && report_len) snprintf(report, report_len, "%s", rep.c_str());
    return 0;
}

}  // extern "C"
