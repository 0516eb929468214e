"""The shared-reciprocal division of the device code equals IEEE division.

The sweeps and the CFL divide several quotients by one divisor d with
r = RN(1/d), q0 = RN(a r), rem = a - q0 d (FMA), q = RN(q0 + rem r) (FMA),
q0 kept when rem == 0 (fargo3d_amd/csrc/hip/f3d_device_body.h, rdiv/rdivd).
The FP64 path's bit-for-bit parity with the CPU restatement rests on
q == RN(a/d).  Checked here in exact integer arithmetic for every pair of
significands of small binary formats, where the uncorrected q0 is wrong in
about a quarter of the pairs (so the check has teeth), and in binary64 on
random and near-midpoint operands.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("recip") / "recip_division_check")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", os.path.join(HERE, "recip_division_check.c"),
                    "-o", exe, "-lm"], check=True)
    return exe


def test_every_significand_pair_of_small_formats(checker):
    out = subprocess.run([checker, "exhaustive", "4", "12"], check=True, capture_output=True, text=True).stdout
    rows = [tuple(int(x) for x in line.split()) for line in out.split("\n") if line.strip()]
    assert [r[0] for r in rows] == list(range(4, 13))
    for p, pairs, bad, q0bad in rows:
        assert pairs == 4 ** (p - 1)
        assert bad == 0, (p, bad)
        assert q0bad > pairs // 5, (p, q0bad)  # the correction step is needed


def test_binary64_random_and_near_midpoint(checker):
    out = subprocess.run([checker, "binary64", "2000000", "20241203"], check=True, capture_output=True,
                         text=True).stdout
    samples, bad = (int(x) for x in out.split())
    assert samples == 12_000_000
    assert bad == 0
