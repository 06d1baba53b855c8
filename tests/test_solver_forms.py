"""The library's HLLD has two forms in dragon_amd/csrc/dg_math.hpp: the
reference's branching control flow (HLLD.cpp:38-205), run by the FP64
kernels, and a wavefront-uniform instruction stream, run by the FP32 kernels
of the mixed form.  tests/cpp/hlld_forms compiles both for the host and
compares them bit for bit on random faces covering every wave region, in FP64
and FP32 (the FP64 form itself is pinned to the oracle bit for bit by the GPU
parity tests, and the oracle to the reference's golden vectors)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "hlld_forms")


def test_hlld_forms_bit_identical():
    if not os.path.exists(EXE):
        pytest.fail("tests/cpp/hlld_forms is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE, "200000"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if "faces" in l]
    assert len(lines) == 4
    for l in lines:
        assert " 0 mismatches" in l
        census = dict((k, int(v)) for k, v in re.findall(r"(supL|supR|zeroB|starL|starR|dstar) (\d+)", l))
        assert len(census) == 6 and min(census.values()) > 1000, l  # every region exercised
