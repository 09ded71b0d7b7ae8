"""The reference-signature C++ layer (include/compat/) on the GPU: each
function against the CPU oracle through a C++ test program, and the stage-1
and stage-2 drivers (the call sequences of main_SPM.cpp and grad.cpp) end to end."""
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
    # hiosd_lobpcg_sieh: 5 HiSD steps (x-step + one LOBPCG sweep each)
    assert res["hisd_x"] < 1e-10
    assert res["hisd_v"] < 1e-8
    assert res["hisd_res_rel"] < 1e-8
    # hiosd_lobpcg_sieh_V0: 3 GOSD steps with a 2-field null space
    assert res["gosd_x"] < 1e-10
    assert res["gosd_v"] < 1e-8
    assert res["gosd_perp"] < 1e-10


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


def test_stage2_driver(seeds, tmp_path):
    """saddle file -> its modes -> the minimum it leads to -> the modes there (grad.cpp's
    sequence). The first run computes the saddle's modes and saves them; the second finds the
    file and reads them back (grad.cpp:92-113), and must reach the same minimum."""
    N = 16
    phi = seeds.gaussian(N, 0.05, 0)
    src = tmp_path / "phi0.txt"
    np.savetxt(src, phi.ravel(), fmt="%.16e ")
    s1 = tmp_path / "S1.txt"
    _run([os.path.join(ROOT, "spm-iqc_amd", "lib", "spm_stage1"), "--N", str(N), "--phi", str(src), "--k", "2",
          "--lobpcg", "2", "--maxiter", "20", "--tol", "1e-9", "--out", str(s1)], cwd=str(tmp_path))
    exe = os.path.join(ROOT, "spm-iqc_amd", "lib", "spm_stage2")
    vsad = tmp_path / "V1_saddle.txt"
    res = []
    for tag in ("a", "b"):
        out = _run([exe, "--N", str(N), "--in", str(s1), "--vsaddle", str(vsad), "--k", "3", "--lobpcg1", "2",
                    "--lobpcg2", "2", "--maxiter", "200", "--min", str(tmp_path / ("Min_%s.txt" % tag)),
                    "--vmin", str(tmp_path / ("V0_Min_%s.txt" % tag))], cwd=str(tmp_path))
        res.append(json.loads([l for l in out.splitlines() if l.startswith("{")][-1]))
    assert res[0]["modes_loaded"] == 0 and res[1]["modes_loaded"] == 1
    assert np.loadtxt(vsad).shape == (3 * N ** 3,)
    mins = [np.loadtxt(tmp_path / ("Min_%s.txt" % t)) for t in ("a", "b")]
    assert mins[0].shape == (N ** 3,) and np.isfinite(mins[0]).all()
    # "%.16e" round-trips a double, so the read-back modes are the computed ones
    assert np.array_equal(mins[0], mins[1]) and res[0]["E_min"] == res[1]["E_min"]
    assert res[0]["E_min"] < res[0]["E_saddle"]  # gradient flow from the saddle goes downhill
    vmin = np.loadtxt(tmp_path / "V0_Min_a.txt")
    assert vmin.shape == (3 * N ** 3,) and np.isfinite(vmin).all()
