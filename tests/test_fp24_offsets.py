"""CPU check of the FP16 M2L stage's pair-result addressing (ibim_fp24.h):
offsets into the FP24 buffer are formed in 64 bits, so pair rows beyond
2^32 / 384 (about 11.2M at p = 5, where a row is 384 dwords) address the
right place.  The helper is compiled for the host only and compared with
exact integer arithmetic."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def offsets_exe(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("fp24") / "fp24_offsets"
    subprocess.run([HIPCC, "--offload-host-only", "-x", "hip", "-std=c++17", "-O2",
                    os.path.join(HERE, "native", "fp24_offsets.cpp"), "-o", str(out)], check=True)
    return str(out)


def expected(pair, ys, v):
    return pair * (ys // 4) * 3 + (v // 4) * 3, (pair + 1) * (ys // 4) * 3


# (pair row, row stride YS, value index): p = 5 (4 p^3 = 500 -> YS 512),
# p = 6 (864 -> 864), p = 3 (108 -> 128); rows around the 32-bit limits
CASES = [(12_000_000, 512, 448), (11_184_811, 512, 508), (5_592_406, 512, 4), (2**31 - 2, 512, 508),
         (30_000_000, 864, 860), (3, 128, 124), (0, 512, 0)]


def test_offsets_are_64_bit(offsets_exe):
    args = [str(a) for c in CASES for a in c]
    out = subprocess.run([offsets_exe] + args, check=True, capture_output=True, text=True).stdout.split("\n")
    for (pair, ys, v), line in zip(CASES, out):
        got = tuple(int(t) for t in line.split())
        assert got == expected(pair, ys, v), (pair, ys, v, got)
    # the case that wrapped with 32-bit offsets: 12M rows of 384 dwords
    assert expected(12_000_000, 512, 448)[0] > 2**32
