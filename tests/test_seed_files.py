"""get_initial_value (Initialization.cpp:242-338) of the compat layer on the
reference's own seed files (tests/golden/initData: DIS, BCC, FCC, HPC, LAM, DG,
copied unchanged from the reference's initData/), against a reading of the same
files written here from that function's description:

  * the first integer is the mode count n, then 4 integers per mode: f0 f1 f2
    and an integer amplitude a (dimCpt + 1 = 4 per record);
  * mode i goes to index k_d = round(f_d / l * L) (C rounding: halves away from
    zero), wrapped to N + k_d when negative, where l = 1, sqrt 2, sqrt 3,
    sqrt 2, 1, sqrt 6 for DIS, BCC, FCC, HPC, LAM, DG;
  * its value is a * 0.1 in the real part; every other entry is zero, and a
    later mode at the same index overwrites an earlier one.

LAM has 3 integers per line, so its second record runs past the end of the
file: the reference's fscanf then leaves those integers unset (whatever the
allocation held); the compat layer reads them as 0, and so does this reading.
The host-only mode of the compat test program does the work (no GPU call).
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "tests", "compat", "compat_parity")
NAMES = ["DIS", "BCC", "FCC", "HPC", "LAM", "DG"]
SCALE = [1.0, math.sqrt(2.0), math.sqrt(3.0), math.sqrt(2.0), 1.0, math.sqrt(6.0)]


def _c_round(x):
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def _expected(choice, N, L):
    toks = [int(t) for t in open(os.path.join(GOLDEN, "initData", NAMES[choice])).read().split()]
    n, vals = toks[0], toks[1:]
    vals = vals + [0] * max(0, 4 * n - len(vals))
    out = {}
    for i in range(n):
        f = vals[4 * i: 4 * i + 4]
        k = [_c_round(f[d] / SCALE[choice] * L) for d in range(3)]
        k = [N + x if x < 0 else x for x in k]
        out[(k[0] * N + k[1]) * N + k[2]] = f[3] * 0.1
    return {idx: v for idx, v in out.items() if v != 0.0}


def _run(N, L):
    if not os.path.exists(EXE):
        pytest.skip("compat test program not built (__graft_entry__.build())")
    r = subprocess.run([EXE, "seeds", str(N), repr(L)], cwd=GOLDEN, capture_output=True, text=True, timeout=120)
    return r


@pytest.mark.parametrize("N", [64, 128])
def test_seed_files_match_the_reference_reading(N):
    L = 26.0
    r = _run(N, L)
    assert r.returncode == 0, r.stderr[-2000:]
    got = json.loads(r.stdout.strip().splitlines()[-1])
    for choice in range(6):
        entries = {int(idx): (re, im) for idx, re, im in got[str(choice)]}
        want = _expected(choice, N, L)
        assert set(entries) == set(want), NAMES[choice]
        for idx, v in want.items():
            assert entries[idx] == (v, 0.0), (NAMES[choice], idx)
    # every mode of the crystal seeds sits on its star: |k| = |f| / l for each listed mode
    for choice in (1, 2, 3, 5):
        for idx in got[str(choice)]:
            k = np.array(np.unravel_index(idx[0], (N, N, N)))
            k = np.where(k >= N // 2, k - N, k)
            assert abs(np.linalg.norm(k) / L - 1.0) < 0.05, NAMES[choice]


def test_seed_mode_outside_the_grid_is_an_error():
    """At N = 16, L = 26 the BCC modes land at index 18 > N: the reference writes
    outside its array there; the compat layer stops with a message."""
    r = _run(16, 26.0)
    assert r.returncode != 0
    assert "outside the 16^3 grid" in r.stderr
