"""Half-size scalars (csrc/lattice.h) and the verify built on them, checked on
the CPU with the library's own headers compiled for the host
(tests/native/half_verify_check.cpp):

- every (x, y) that half_scalars returns satisfies x = y h (mod 8L) with y odd
  and positive, checked with Python integers, for random h and edge values;
  their sizes match the lattice's (about 128 bits);
- verify_half_one<4> and <5> give the same verdict as verify_one<4> (the path
  the GPU parity tests pin to the oracle) on valid, corrupted, small-order,
  mixed-torsion-key and torsion-R signatures;
- the field products under them are exact mod p on loose inputs up to the
  limb sizes the 32-bit operands allow.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "half_verify_check.cpp")
OUT = os.path.join(ROOT, "tests", "native", "_build", "half_verify_check")
LIBDIR = os.path.join(ROOT, "solana-perf-libs_amd")
L = 2**252 + 27742317777372353535851937790883648493


@pytest.fixture(scope="module")
def checker():
    if not os.path.exists(os.path.join(LIBDIR, "libcuda-crypt.so")):
        pytest.fail("libcuda-crypt.so is not built")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    deps = [SRC] + [os.path.join(LIBDIR, "csrc", f) for f in os.listdir(os.path.join(LIBDIR, "csrc"))]
    if not os.path.exists(OUT) or os.path.getmtime(OUT) < max(os.path.getmtime(d) for d in deps):
        cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O1", "-std=c++17",
               "-Wno-unused-result", "-o", OUT, SRC, "-L" + LIBDIR, "-lcuda-crypt",
               "-Wl,-rpath," + LIBDIR]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
    return OUT


def test_half_scalars_congruence_and_size(checker):
    out = subprocess.run([checker, "lattice", "6000", "7"], capture_output=True, text=True,
                         check=True).stdout
    rows = [l.split() for l in out.splitlines()]
    assert len(rows) == 6000
    bits = []
    for hh, xx, xn, yy in rows:
        h, x, y = int(hh, 16), int(xx, 16), int(yy, 16)
        if xn == "1":
            x = -x
        assert (x - y * h) % (8 * L) == 0, hh
        assert y > 0 and y % 2 == 1, hh
        bits.append(max(abs(x).bit_length(), y.bit_length()))
    rnd = bits[40:]                     # after the edge values
    assert max(rnd) <= 136 and sum(rnd) / len(rnd) < 129
    # h = 0..15 are their own half scalars (x = h, y = 1)
    assert bits[:16] == [max(1, h.bit_length()) for h in range(16)]


P = 2**255 - 19
LIMB_SHIFT = [0, 26, 51, 77, 102, 128, 153, 179, 204, 230]   # radix 2^25.5


def test_field_products_on_loose_limbs(checker):
    """fe_mul, fe_sq and fe_sq2 (fe25519.h, compiled for the host from the
    same source as the kernels) against Python integers mod p, on inputs
    whose limbs reach 1, 2 and 3 times their tight range, at random and at
    every extreme (all +max, all -max, alternating, random signs): the column
    sums with their rounding offsets must not overflow, and every output limb
    must come back to its tight range (limb 1 at most 2^15 over)."""
    out = subprocess.run([checker, "field", "2250", "5"], capture_output=True, text=True,
                         check=True).stdout
    rows = [list(map(int, l.split())) for l in out.splitlines()]
    assert len(rows) == 2250

    def val(l):
        return sum(v << s for v, s in zip(l, LIMB_SHIFT))

    for r in rows:
        f, g, m, s, s2 = (r[10 * k:10 * k + 10] for k in range(5))
        vf, vg = val(f), val(g)
        assert (val(m) - vf * vg) % P == 0, (f, g)
        assert (val(s) - vf * vf) % P == 0, f
        assert (val(s2) - 2 * vf * vf) % P == 0, f
        for out_limbs in (m, s, s2):
            for i, v in enumerate(out_limbs):
                bound = (1 << (25 if i % 2 == 0 else 24)) + ((1 << 15) if i == 1 else 0)
                assert -bound <= v <= bound, (i, v, f, g)


def test_half_verify_matches_full_verify(checker):
    r = subprocess.run([checker, "verify", "3000", "11"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "differences: 0" in r.stdout
    # the mixed-torsion-key rows include accepts and rejects
    acc = {int(l.split()[1][:-1]): int(l.split()[4]) for l in r.stdout.splitlines()
           if l.startswith("kind")}
    assert 0 < acc[7] < 250 and acc[0] == 250
