"""The reference-signature C++ layer (include/compat/) on the GPU: each
function against the CPU oracle through a C++ test program, and the stage-1
driver (main_SPM.cpp's call sequence) end to end."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(cmd, cwd=None, timeout=240):
    print("[case]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd, timeout=timeout)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


@pytest.mark.parametrize("N", [16, 24])
def test_compat_functions_match_oracle(N):
    exe = os.path.join(ROOT, "tests", "compat", "compat_parity")
    assert os.path.exists(exe), "build the test program first (__graft_entry__.build())"
    out = _run([exe, str(N)])
    res = json.loads([l for l in out.splitlines() if l.startswith("{")][-1])
    assert res["fft"] < 1e-15 and res["ifft"] < 1e-11
    assert res["energy_rel"] < 1e-12
    for key in ("ngrad", "hv", "precond", "tv", "natural", "minmax"):
        assert res[key] < 1e-12, (key, res[key])
    assert res["inpfp"] < 1e-14
    assert res["gradflow"] < 1e-10
    assert res["spring"] < 1e-11
    assert res["file"] == 0.0
    # at N = 24 the lowest shell |n|^2 ~ 258 (|k| ~ 0.618) is a cluster of
    # near-degenerate eigenvalues and 12 x 21 iterations of block 2 do not
    # resolve it to 1e-9; N = 16 has a separated lowest eigenvalue
    assert res["lobpcg_lowest"] < (1e-9 if N == 16 else 1e-6)


def test_stage1_driver(seeds, tmp_path):
    """relax -> LOBPCG -> SPM -> saddle file, through the reference's call sequence."""
    N = 16
    phi = seeds.gaussian(N, 0.05, 0)
    src = tmp_path / "phi0.txt"
    np.savetxt(src, phi.ravel(), fmt="%.16e ")
    out_file = tmp_path / "S1.txt"
    exe = os.path.join(ROOT, "spm-iqc_amd", "lib", "spm_stage1")
    out = _run([exe, "--N", str(N), "--phi", str(src), "--k", "2", "--lobpcg", "2", "--maxiter", "20",
                "--tol", "1e-9", "--out", str(out_file)], cwd=str(tmp_path))
    res = json.loads([l for l in out.splitlines() if l.startswith("{")][-1])
    assert np.isfinite(res["E_min"]) and np.isfinite(res["E_saddle"])
    saddle = np.loadtxt(out_file)
    assert saddle.shape == (N ** 3,) and np.isfinite(saddle).all()
