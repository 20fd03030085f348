"""CPU tests of libcuda-crypt.so: it loads, exports every symbol the header
declares, and its host-side exports (which run the same field/group/scalar
code as the kernels, compiled for the host) agree with the oracle and the
committed fixtures.  No GPU compute is called here."""
import os
import re
import subprocess
import sys

import numpy as np

from conftest import ROOT, load_json
import ed25519_math as E
import perf_libs
import workloads as W


def test_exports_match_header(lib):
    hdr = open(os.path.join(ROOT, "include", "cuda_crypt.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    declared = set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9 ]+\*?\s*\b([a-z_0-9]+)\(", hdr, re.M))
    declared -= {"if", "return"}
    assert declared == set(perf_libs.EXPORTS) | set(perf_libs.CXX_EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", perf_libs.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    symbols = {perf_libs.CXX_EXPORTS.get(name, name) for name in declared}
    missing = symbols - exported
    assert not missing, missing


def test_license_string(lib):
    assert "Apache License, Version 2.0" in lib.license()


def test_rfc8032(lib):
    for v in load_json("rfc8032.json"):
        pk, sk = lib.create_keypair(bytes.fromhex(v["secret"]))
        assert pk.hex() == v["public"]
        msg = bytes.fromhex(v["message"])
        assert lib.sign(msg, pk, sk).hex() == v["signature"]
        assert lib.verify(bytes.fromhex(v["signature"]), msg, pk) == 1


def test_host_exports_vs_reference_fixture(lib):
    fx = load_json("ref_quirks.json")
    for c in fx["get_checked_scalar"]:
        rc, out = lib.get_checked_scalar(bytes.fromhex(c["in"]), bytes.fromhex(c["prior"]))
        assert (rc, out.hex()) == (c["rc"], c["out"]), c
    for c in fx["small_order"]:
        assert lib.check_packed_ge_small_order(bytes.fromhex(c["enc"])) == c["check"]
    for c in fx["cases"]:
        pk, sk = lib.create_keypair(bytes.fromhex(c["seed"]))
        assert sk.hex() == c["sk"]
        sig, msg, cpk = bytes.fromhex(c["sig"]), bytes.fromhex(c["msg"]), bytes.fromhex(c["pk"])
        assert lib.verify(sig, msg, cpk) == c["verify"], c["kind"]
        if c["kind"] in (0, 6, 7):
            assert lib.sign(msg, pk, sk) == sig


def test_host_exports_vs_oracle_random(lib, oracle):
    rng = np.random.default_rng(17)
    for k in range(60):
        seed = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        pk, sk = lib.create_keypair(seed)
        assert (pk, sk) == oracle.create_keypair(seed)
        msg = rng.integers(0, 256, int(rng.integers(0, 1300)), dtype=np.uint8).tobytes()
        sig = lib.sign(msg, pk, sk)
        assert sig == oracle.sign(msg, pk, sk)
        assert lib.verify(sig, msg, pk) == 1
        bad = bytearray(sig)
        bad[int(rng.integers(0, 64))] ^= 1 << int(rng.integers(0, 8))
        assert lib.verify(bytes(bad), msg, pk) == oracle.verify(bytes(bad), msg, pk)
        enc = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        assert lib.check_packed_ge_small_order(enc) == oracle.check_packed_ge_small_order(enc)
        inp = bytearray(rng.integers(0, 256, 32, dtype=np.uint8).tobytes())
        inp[31] = [0x05, 0x17, 0x42, 0x7f, 0x80, 0xff][k % 6]
        prior = rng.integers(0, 256, 32, dtype=np.uint8).tobytes()
        assert lib.get_checked_scalar(bytes(inp), prior) == \
            oracle.get_checked_scalar(bytes(inp), prior)


def test_verify_edge_encodings(lib):
    """Non-canonical R never verifies; small-order keys and R rejected;
    S >= 2^252 rejected (host build of the verify core)."""
    seed = bytes(range(32))
    pk, sk = lib.create_keypair(seed)
    msg = b"edge cases"
    sig = lib.sign(msg, pk, sk)
    assert lib.verify(sig, msg, pk) == 1
    s_plus_l = (int.from_bytes(sig[32:], "little") + E.L).to_bytes(32, "little")
    assert lib.verify(sig[:32] + s_plus_l, msg, pk) == 0
    for enc in W.SMALL_ORDER_KEYS:
        assert lib.verify(sig, msg, enc) == E.verify_rule(sig, msg, enc) == 0
        assert lib.verify(enc + sig[32:], msg, pk) == E.verify_rule(enc + sig[32:], msg, pk)


def _run_without_gpu(code):
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1",
               CUDA_VISIBLE_DEVICES="-1")
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          env=env, cwd=ROOT, timeout=120)


def test_no_gpu_behaviour():
    """Without a device: init and poh_init are false, poh_verify_many returns 1 and leaves the
    hashes alone, verify_many rejects everything instead of crashing
    (the reference dereferences NULL there, verify.cu:265-267), and the engine
    and server constructors refuse."""
    code = r"""
import sys, numpy as np
sys.path.insert(0, 'solana-perf-libs_amd')
import perf_libs
L = perf_libs.load()
assert L.init() is False
assert L.poh_init() is False
h = np.arange(64, dtype=np.uint8); h0 = h.copy()
assert L.poh_verify_many(h, np.array([3, 4], np.uint64)) == 1
assert np.array_equal(h, h0)
pk, sk = L.create_keypair(bytes(32))
pkt = np.zeros(256, np.uint8)
pkt[:64] = np.frombuffer(L.sign(b'x' * 100, pk, sk), np.uint8)
pkt[64:96] = np.frombuffer(pk, np.uint8); pkt[96:196] = ord('x')
out = L.verify_many([pkt], 256, [100], [64], [0], [96])
assert out.tolist() == [0]
assert L.lib.spl_server_create(0, 0, 0, 0) is None
assert L.lib.spl_verify_engine_create(16, 4096) is None
# PoH kernel hook: an unknown kernel is refused before any device work; a
# known one reports no GPU
assert L.lib.spl_poh_device_kernel(None, None, 4, 4, None) == 2
assert L.lib.spl_poh_device_kernel(None, None, 4, -1, None) == 2
assert L.lib.spl_poh_device_kernel(None, None, 0, 3, None) == 0
assert L.lib.spl_poh_device_kernel(16, 16, 4, 3, None) == 1
assert L.poh_kernel_choice(4096) is None
try:
    perf_libs.Server(L)
    raise SystemExit('Server() must raise without a GPU')
except RuntimeError:
    pass
print('ok')
"""
    r = _run_without_gpu(code)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr


def test_tune_records_file(tmp_path):
    """The autotuner's records in a file: load takes well-formed records
    (the ABI then uses the recorded one-shot configuration for that
    batch-size class), save writes back the same records, and a malformed
    line or an unknown configuration makes load take nothing.  Without a
    device, so no record can be skipped as made on another GPU."""
    good = tmp_path / "tune.txt"
    good.write_text(
        "# comment\n"
        "device 0 gfx950 256 class 16 best persistent W=4 scalars=half sort=1 waves_per_cu=4 "
        "| plain one-shot W=4 scalars=half waves_per_eu=1\n"
        "device 0 gfx950 256 class 10 best one-shot W=5 scalars=half lanes=2 waves_per_eu=2 "
        "| plain one-shot W=5 scalars=half lanes=2 waves_per_eu=2\n")
    bad_line = tmp_path / "bad.txt"
    bad_line.write_text("device 0 gfx950 256 class 12 best\n")
    bad_name = tmp_path / "badname.txt"
    bad_name.write_text("device 0 gfx950 256 class 12 best one-shot W=7 scalars=half "
                        "waves_per_eu=1 | plain one-shot W=4 scalars=half waves_per_eu=1\n")
    persistent_plain = tmp_path / "pp.txt"
    persistent_plain.write_text(
        "device 0 gfx950 256 class 12 best one-shot W=4 scalars=half waves_per_eu=1 "
        "| plain persistent W=4 scalars=half sort=1 waves_per_cu=4\n")
    saved = tmp_path / "saved.txt"
    code = rf"""
import sys
sys.path.insert(0, 'solana-perf-libs_amd')
import perf_libs
L = perf_libs.load()
before = L.tune_query(0, 3000)
assert before[0] is False, before
for p in ({str(bad_line)!r}, {str(bad_name)!r}, {str(persistent_plain)!r}):
    assert L.tune_load(p) == -4, p
assert L.tune_query(0, 3000) == before
assert L.tune_load({str(tmp_path / 'missing.txt')!r}) == -3
assert L.tune_load({str(good)!r}) == 2
assert L.tune_query(0, 40000) == (True, 'one-shot W=4 scalars=half waves_per_eu=1')
assert L.tune_query(0, 33000) == (True, 'one-shot W=4 scalars=half waves_per_eu=1')
assert L.tune_query(0, 600) == (True, 'one-shot W=5 scalars=half lanes=2 waves_per_eu=2')
assert L.tune_query(0, 3000) == before
assert L.tune_query(1, 40000)[0] is False
assert L.tune_save({str(saved)!r}) == 2
print('ok')
"""
    r = _run_without_gpu(code)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
    lines = [l for l in saved.read_text().splitlines() if not l.startswith("#")]
    assert lines == [
        "device 0 unknown 0 class 10 best one-shot W=5 scalars=half lanes=2 waves_per_eu=2 "
        "| plain one-shot W=5 scalars=half lanes=2 waves_per_eu=2",
        "device 0 unknown 0 class 16 best persistent W=4 scalars=half sort=1 waves_per_cu=4 "
        "| plain one-shot W=4 scalars=half waves_per_eu=1"], lines


def test_check_helpers_full_mix(lib, oracle):
    """Both host check helpers over 100k random 32-byte encodings, 1 % of them
    replaced by small-order encodings (canonical and non-canonical) and 1 %
    by non-reduced scalars (top byte 0x10-0x7F), with zeroed out buffers:
    every return code and output equals the oracle's."""
    rng = np.random.default_rng(6)
    n = 100_000
    encs = rng.integers(0, 256, size=(n, 32), dtype=np.uint8)
    pick = rng.random(n)
    small = np.nonzero(pick < 0.01)[0]
    for k, i in enumerate(small):
        encs[i] = np.frombuffer(W.SMALL_ORDER_KEYS[k % len(W.SMALL_ORDER_KEYS)], np.uint8)
    unreduced = np.nonzero((pick >= 0.01) & (pick < 0.02))[0]
    encs[unreduced, 31] = 0x10 | (encs[unreduced, 31] & 0x6f)
    so_rejects = gc_rejects = 0
    for i in range(n):
        e = encs[i].tobytes()
        so = lib.check_packed_ge_small_order(e)
        assert so == oracle.check_packed_ge_small_order(e), i
        gc = lib.get_checked_scalar(e)
        assert gc == oracle.get_checked_scalar(e), i
        so_rejects += so
        gc_rejects += gc[0]
    # every planted case is rejected: small-order ones by the closed-form
    # rule (decodes, y mod p in {0, 1, p-1}), non-reduced scalars outright
    planted = sum(E.small_order_rule(W.SMALL_ORDER_KEYS[k % len(W.SMALL_ORDER_KEYS)])
                  for k in range(small.size))
    assert planted > 0 and so_rejects >= planted
    assert gc_rejects >= unreduced.size
