"""Open-loop load generator for the serving loop (spl_server).

Served mixed load: a Poisson stream of verify requests (log-uniform 1-2048
packets in the validator-shaped layout: 1-4 signatures per packet, messages of
100-1200 B in a 1232 B stride) and PoH requests (log-uniform 1-512 chains of
12,500 hashes), at a stepped offered load in signatures per second.

Requests are cut from two pools whose expected results are known: verify
requests are runs of consecutive packets of one validator-shaped batch, PoH
requests are runs of chains of one chain pool.  A verify request carries its
own packet bytes (a view into the pool) and offset tables rebased to them, as
a caller of ed25519_verify_many would build them.  With pinned pools the
packet pool lives in pinned host memory, as solana_perf keeps its packet
batches, so the server and the plain calls read requests in place.
"""
import gc
import time

import numpy as np

import workloads as W


class RequestPools:
    def __init__(self, lib, verify_batch, verify_mask, poh_in, poh_counts, poh_out,
                 pinned=False):
        self.lib = lib
        self.b = verify_batch
        self.pinned = pinned
        if pinned:
            import hip_rt
            self.packets = hip_rt.pinned_copy(self.b.packets)
        else:
            self.packets = self.b.packets
        self.mask = np.asarray(verify_mask, np.uint8)
        stride = self.b.message_size
        pkt = (self.b.sig_off // stride).astype(np.int64)      # packet of each signature
        npk = self.b.total_packets
        # signatures of packet p are [first[p], first[p+1])
        self.first = np.searchsorted(pkt, np.arange(npk + 1), side="left")
        self.poh_in = poh_in
        self.poh_counts = poh_counts
        self.poh_out = poh_out

    @property
    def num_packets(self):
        return self.b.total_packets

    @property
    def num_chains(self):
        return self.poh_counts.size

    def verify_request(self, p0, npk):
        """Packets [p0, p0+npk) as one request: (packets, lens, pk, sig, msg,
        expected mask)."""
        stride = self.b.message_size
        s0, s1 = int(self.first[p0]), int(self.first[p0 + npk])
        base = np.uint32(p0 * stride)
        packets = self.packets[p0 * stride:(p0 + npk) * stride]
        return (packets, self.b.lens[s0:s1].copy(), self.b.pk_off[s0:s1] - base,
                self.b.sig_off[s0:s1] - base, self.b.msg_off[s0:s1] - base,
                self.mask[s0:s1])

    def poh_request(self, c0, nch):
        """Chains [c0, c0+nch): (hashes to hash in place, counts, expected)."""
        return (self.poh_in[32 * c0:32 * (c0 + nch)].copy(), self.poh_counts[c0:c0 + nch],
                self.poh_out[32 * c0:32 * (c0 + nch)])


def build_pools(lib, seed=5, num_batches=128, poh_chains=512, poh_hashes=12500,
                expected=None, pinned=False):
    """Pools for served load.  `expected(batch, poh_in, counts)` returns the
    verify mask and the PoH end states; by default they come from the
    library's plain entry points (ed25519_verify_many / poh_verify_many).
    pinned: keep the packet pool in pinned host memory (hipHostMalloc)."""
    b = W.verify_validator(lib.sign, lib.create_keypair, seed=seed, num_batches=num_batches)
    poh_in, counts = W.poh_chains(poh_chains, poh_hashes, seed=seed + 1)
    if expected is None:
        mask = lib.verify_many(b.batches, b.message_size, b.lens, b.pk_off, b.sig_off,
                               b.msg_off)
        out = poh_in.copy()
        assert lib.poh_verify_many(out, counts) == 0
    else:
        mask, out = expected(b, poh_in, counts)
    return RequestPools(lib, b, mask, poh_in, counts, out, pinned=pinned)


def log_uniform_int(rng, lo, hi):
    return int(min(hi, max(lo, np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi + 1)))))))


def make_trace(pools, rng, sig_rate, duration_s, poh_rate=50.0, max_packets=2048,
               max_chains=512):
    """Poisson arrivals over duration_s: verify requests at sig_rate / (mean
    signatures per request) per second, PoH requests at poh_rate per second.
    Returns a list of (arrival_s, kind, payload) sorted by arrival."""
    sigs_per_packet = pools.b.n / pools.num_packets
    mean_packets = (max_packets - 1) / np.log(max_packets)     # log-uniform on [1, max]
    v_rate = sig_rate / (mean_packets * sigs_per_packet)
    trace = []
    for rate, kind in ((v_rate, "verify"), (poh_rate, "poh")):
        if rate <= 0:
            continue
        t = rng.exponential(1.0 / rate)
        while t < duration_s:
            if kind == "verify":
                npk = log_uniform_int(rng, 1, min(max_packets, pools.num_packets))
                p0 = int(rng.integers(0, pools.num_packets - npk + 1))
                trace.append((t, kind, pools.verify_request(p0, npk)))
            else:
                nch = log_uniform_int(rng, 1, min(max_chains, pools.num_chains))
                c0 = int(rng.integers(0, pools.num_chains - nch + 1))
                trace.append((t, kind, pools.poh_request(c0, nch)))
            t += rng.exponential(1.0 / rate)
    trace.sort(key=lambda r: r[0])
    return trace


class ServedTimeout(RuntimeError):
    """A request did not complete within the wait limit."""


def run_served(server, trace, lead_s=0.05, wait_ms=60000, keep_results=False):
    """Submits every request at its arrival time (open loop) and waits for all
    of them.  Returns per-request records: kind, units, latency from the
    scheduled arrival to completion (ms), submit lateness (ms), and whether the
    result equals the expected one (with keep_results, also the result and
    the request itself)."""
    # convert every request's arguments before the clock starts, so that a
    # submission at its arrival time is one C call
    prepared = []
    for arr, kind, req in trace:
        if kind == "verify":
            packets, lens, pk, sig, msg, _ = req
            prep, res = server.prepare_verify(packets, lens, pk, sig, msg)
        else:
            hashes, counts, _ = req
            prep, res = server.prepare_poh(hashes, counts), hashes
        prepared.append((arr, kind, req, prep, res))
    t0 = time.monotonic_ns() + int(lead_s * 1e9)
    tickets = []
    # No cyclic garbage collection while submitting: a full collection over
    # tens of thousands of live request objects pauses this thread for
    # milliseconds, and every request due meanwhile would be submitted late.
    gc_was_enabled = gc.isenabled()
    gc.disable()
    try:
        for arr, kind, req, prep, res in prepared:
            target = t0 + int(arr * 1e9)
            while True:
                now = time.monotonic_ns()
                if now >= target:
                    break
                if target - now > 300_000:
                    time.sleep((target - now - 200_000) / 1e9)
            tickets.append((server.submit_prepared(prep), target, kind, req, res))
    finally:
        if gc_was_enabled:
            gc.enable()
    t_last = time.monotonic_ns()
    recs = []
    for t, target, kind, req, res in tickets:
        assert t > 0, "submit refused"
        rc, (ts, ta, td) = server.wait(t, wait_ms)
        if rc == 1:
            raise ServedTimeout(f"request {t} ({kind}) not complete after {wait_ms} ms; "
                                f"server stats {server.stats()}")
        if kind == "verify":
            exp = req[5]
            ok = rc == 0 and np.array_equal(res[:exp.size], exp)
            units = exp.size
        else:
            exp = req[2]
            ok = rc == 0 and np.array_equal(res, exp)
            units = exp.size // 32
        rec = {"kind": kind, "units": units, "latency_ms": (td - target) / 1e6,
               "late_ms": (ts - target) / 1e6, "done_ns": td, "ok": bool(ok)}
        if keep_results:
            rec["result"] = res[:exp.size].copy() if kind == "verify" else res.copy()
            rec["request"] = req
        recs.append(rec)
    span_s = (max(r["done_ns"] for r in recs) - t0) / 1e9 if recs else 0.0
    return recs, span_s, (t_last - t0) / 1e9


def summarize(recs, span_s):
    out = {}
    for kind in ("verify", "poh"):
        r = [x for x in recs if x["kind"] == kind]
        if not r:
            continue
        lat = np.array([x["latency_ms"] for x in r])
        late = np.array([x["late_ms"] for x in r])
        units = sum(x["units"] for x in r)
        out[kind] = {"requests": len(r), "units": units,
                     "units_per_s": units / span_s if span_s > 0 else 0.0,
                     "p50_ms": float(np.percentile(lat, 50)),
                     "p99_ms": float(np.percentile(lat, 99)),
                     "max_ms": float(lat.max()),
                     "submit_late_p99_ms": float(np.percentile(late, 99)),
                     "all_exact": all(x["ok"] for x in r)}
    return out


def run_back_to_back(lib, trace):
    """The same verify requests through the plain ed25519_verify_many, one call
    after another (closed loop), each call's ctypes arguments converted before
    the clock starts.  Returns (signatures/s, p50 ms, p99 ms, all exact)."""
    calls = [(lib.prepare_verify_many([req[0]], 1232, *req[1:5]), req[5])
             for _, kind, req in trace if kind == "verify"]
    lat, units, ok = [], 0, True
    t_start = time.perf_counter()
    for prep, exp in calls:
        t = time.perf_counter()
        mask = lib.call_prepared(prep)
        lat.append((time.perf_counter() - t) * 1e3)
        units += exp.size
        ok = ok and np.array_equal(mask, exp)
    wall = time.perf_counter() - t_start
    lat = np.array(lat)
    return {"sig_verifies_per_s": units / wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "calls": int(lat.size),
            "all_exact": bool(ok)}
