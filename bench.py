"""Benchmark: batched Ed25519 verify and SHA-256 PoH on one MI355X.

A step is one ed25519 verify pass over one validator-shaped batch (128 batches
of U[1,256] packets, 1-4 signatures per packet, 100-1200 byte messages in a
1232-byte stride, 5% bit-flipped signatures, 1% S+L, 1% small-order keys),
packets and offset tables already resident in HBM, launched through the
device-resident entry point on one stream.  `value` = signatures verified per
second over K timed steps.  Also reported (same JSON line):
  poh:           poh_verify_many_device on 4096 chains x 12,500 hashes, and a
                 saturating 262,144 x 500 run
  plain_verify:  the plain ed25519_verify_many ABI (host buffers, PCIe in and
                 out) on 1024 one-signature packets: p50/p99 call latency
  roofline:      verify kernel, int32 VALU ops per second vs 3.93e13
  cpu_baseline:  the reference's own CPU code (oracle/_ref, built from its
                 sources) on the same batch, all host cores
Inputs are synthetic and seeded; signing for input preparation uses the
library under test.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "solana-perf-libs_amd")
sys.path.insert(0, PKG)

VALU_PEAK = 256 * 64 * 2.4e9          # int32 lane-ops/s, MI355X
W_SHA256 = 2296                        # int32 ops per SHA-256 compression
W_SHA512 = 5840                        # int32 ops per SHA-512 block
W_VERIFY_128 = 830810                  # int32 ops per verify, 128-byte message


def verify_ops(lens):
    """Algorithmic int32 ops for verifying signatures over messages of `lens`."""
    blocks = (81 + np.asarray(lens, np.int64) + 127) // 128
    return float(np.sum(W_VERIFY_128 + (blocks - 2) * W_SHA512))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def run_cpu_baseline(b, seconds_cap=20.0):
    """The reference CPU path on the same batch (or a prefix of it bounded by
    seconds_cap), all host cores, one contiguous slice per thread."""
    sys.path.insert(0, ROOT)
    from concurrent.futures import ThreadPoolExecutor
    from oracle import Oracle, RefLib
    kind = "reference" if RefLib.available() else "port"
    impl = RefLib() if kind == "reference" else Oracle()
    threads = cpu_threads()
    # calibrate on 256 signatures, then size the sample to the time cap
    t = time.perf_counter()
    impl.verify_many(b.packets, b.lens[:256], b.pk_off[:256], b.sig_off[:256], b.msg_off[:256])
    per_sig = (time.perf_counter() - t) / 256
    n = int(min(b.n, max(1024, seconds_cap * threads / max(per_sig, 1e-9) * 0.5)))
    chunks = np.array_split(np.arange(n), threads)
    out = np.zeros(n, np.uint8)

    def work(ix):
        if ix.size:
            sl = slice(int(ix[0]), int(ix[-1]) + 1)
            out[sl] = impl.verify_many(b.packets, b.lens[sl], b.pk_off[sl], b.sig_off[sl],
                                       b.msg_off[sl])
    t = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, chunks))
    dt = time.perf_counter() - t
    return {"value": n / dt, "unit": "sig-verifies/s", "cores": threads, "kind": kind,
            "sample": f"first {n} signatures of the validator-shaped batch "
                      f"({'oracle/_ref: reference sources built with g++ -O3' if kind == 'reference' else 'oracle restatement'}), "
                      f"{threads} threads, {dt:.2f} s"}, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the PoH and plain-ABI legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(signatures and PoH chains shard trivially across GPUs; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import hip_rt
    import perf_libs
    import workloads as W
    lib = perf_libs.load()
    assert lib.init(), "no usable GPU"
    stream = hip_rt.Stream()

    log("bench: generating validator-shaped batch")
    b = W.verify_validator(lib.sign, lib.create_keypair, seed=3)
    d_packets = hip_rt.DeviceBuffer.from_array(b.packets)
    d_tabs = [hip_rt.DeviceBuffer.from_array(x) for x in (b.lens, b.pk_off, b.sig_off, b.msg_off)]
    d_out = hip_rt.DeviceBuffer(b.n)

    def step():
        rc = lib.verify_many_device(d_packets.ptr, b.packets.size, *[x.ptr for x in d_tabs], b.n,
                                    d_out.ptr, stream.handle)
        assert rc == 0

    for _ in range(args.warmup):
        step()
    stream.synchronize()
    log(f"bench: timing {args.steps} steps of {b.n} signatures")
    ev0, ev1 = hip_rt.Event(), hip_rt.Event()
    hip_rt.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    ev1.synchronize()
    hip_rt.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_ms(ev1) / args.steps
    mask = d_out.download()
    ms_per_step = wall * 1e3 / args.steps
    value = b.n / (ms_per_step / 1e3)
    ops = verify_ops(b.lens)
    achieved = ops / (kern_ms / 1e3)
    result = {
        "metric": "sig verifies/s + PoH SHA-256 hashes/s on 1 MI355X; p99 request latency",
        "value": value, "unit": "sig-verifies/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32/uint32 (exact integer)",
        "data": "synthetic, seeded (numpy PCG64); keys/signatures made by the library's host "
                "ed25519_create_keypair/ed25519_sign",
        "config": {"workload": "validator-shaped batched verify: 128 batches x U[1,256] packets, "
                               "1-4 sigs/packet, msg 100-1200 B, stride 1232 B, 5% flipped, "
                               "1% S+L, 1% small-order keys; device-resident",
                   "signatures": b.n, "packets": b.total_packets,
                   "accepted": int(mask.sum()), "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": VALU_PEAK / 1e12,
                     "unit": "TOP/s (int32 lane-ops)", "frac": achieved / VALU_PEAK,
                     "traffic": None, "kernel": "verify_kernel",
                     "ops_per_launch": ops, "launch_ms": kern_ms},
    }

    if not args.no_extra:
        # PoH: the plain 4096 x 12,500 workload and a saturating one
        poh = {}
        for name, n, reps in (("c2_4096x12500", 4096, 12500), ("sat_262144x500", 262144, 500)):
            h, counts = W.poh_chains(n, reps, seed=2)
            dh = hip_rt.DeviceBuffer.from_array(h)
            dc = hip_rt.DeviceBuffer.from_array(counts)
            lib.poh_verify_many_device(dh.ptr, dc.ptr, n, stream.handle)
            stream.synchronize()
            e0, e1 = hip_rt.Event(), hip_rt.Event()
            e0.record(stream)
            lib.poh_verify_many_device(dh.ptr, dc.ptr, n, stream.handle)
            e1.record(stream)
            e1.synchronize()
            ms = e0.elapsed_ms(e1)
            hps = n * reps / (ms / 1e3)
            poh[name] = {"hashes_per_s": hps, "ms": ms,
                         "roofline_frac": hps * W_SHA256 / VALU_PEAK}
            log(f"bench: poh {name}: {hps:.3e} hashes/s")
        result["poh"] = poh
        # plain ABI call latency (host buffers, PCIe both ways)
        p = W.verify_plain(1024, lib.sign, lib.create_keypair, seed=1)
        lat = []
        for k in range(40):
            t = time.perf_counter()
            out = lib.verify_many(p.batches, p.message_size, p.lens, p.pk_off, p.sig_off,
                                  p.msg_off)
            lat.append((time.perf_counter() - t) * 1e3)
        lat = np.array(lat[5:])
        result["plain_verify_1024"] = {
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "sig_verifies_per_s": 1024 / (np.median(lat) / 1e3),
            "accepted": int(out.sum())}
        log(f"bench: plain 1024 p50 {result['plain_verify_1024']['p50_ms']:.2f} ms")

    if not args.no_cpu:
        log("bench: CPU baseline")
        cb, cpu_mask = run_cpu_baseline(b)
        result["cpu_baseline"] = cb
        result["parity_vs_cpu"] = bool(np.array_equal(cpu_mask, mask[:cpu_mask.size]))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
