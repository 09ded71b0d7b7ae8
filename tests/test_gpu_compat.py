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
    assert res["energy_uf_rel"] < 1e-12  # get_Ene reads u_F_local (LPsrc.cpp:27-35)
    assert res["basic_ops"] < 1e-13  # BasicOperators.h:6-50 against host loops
    assert res["copy_option_V"] == 0.0  # LOBPCG.cpp:489-520, all three overloads
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


@pytest.mark.parametrize("driver", ["critical_point_HISD", "critical_point_HISD_v", "critical_point",
                                    "critical_point_v"])
def test_saddle_drivers_match_oracle(driver):
    """critical_point_HISD(_v) (HISD.cpp:13-83) and critical_point(_v) (GOSD.cpp:13-120)
    through the compat layer, in the {110} cell at 16^3: u_R_local is the index-1
    saddle Spring_Pairing reaches, option1 its lowest 4 modes.  The oracle runs
    hiosd_lobpcg_sieh(_V0) with the driver's perturbation, directions, null space and
    parameters from the same inputs.  Both must end at the same energy, the oracle
    must certify the GPU's end point as critical to the driver's tolerance, and the
    two end points may differ only as far as their stopping tests allow; the drivers
    that read seed files
    (FCC, HPC) find them in tests/golden/initData.

    The HiSD drivers climb back to the saddle along its one unstable mode, which V
    must match.  The GOSD drivers' null space V0 is option1's first three modes,
    the saddle's unstable mode among them, so they relax to the ordered minimum
    below it, where V ends in the (degenerate) translation modes: only the field
    (up to a translation) and the energy are unique there."""
    exe = os.path.join(ROOT, "tests", "compat", "compat_parity")
    out = _run([exe, "drivers", driver, "16"], cwd=os.path.join(ROOT, "tests", "golden"))
    res = json.loads([l for l in out.splitlines() if l.startswith("{")][-1])
    print(res, flush=True)
    assert res["res_oracle"] < res["epsf"] and res["iters_oracle"] < res["maxiter"]
    assert res["res"] < res["epsf"] * 1.01
    # the oracle certifies the GPU's end point as a critical point to the driver's tolerance
    assert res["res_oracle_at_gpu"] < res["epsf"] * 1.01
    # Both runs stop anywhere inside the ball ||ngrad|| < epsf, so their end points
    # may differ by dx along directions the Hessian barely resists (the minimum's
    # translations and soft modes): to first order H dx is the difference of the two
    # end gradients, so ||H dx|| <= 2 epsf, and dx itself must stay small (one
    # critical point, not two).
    assert res["hdx"] <= 2.02 * res["epsf"]
    assert res["x"] < 1e-4
    if driver.startswith("critical_point_HISD"):
        # climbing back to the saddle the two runs also agree pointwise to epsf
        assert res["x"] < max(1e-8, res["epsf"])
    # at a point where ||grad F|| < epsf, E moves by at most about epsf |dx| with the field
    assert abs(res["E"] - res["E_oracle"]) <= 1e-10 * abs(res["E_oracle"]) + res["epsf"] * res["x"]
    if driver.startswith("critical_point_HISD"):
        assert res["v"] < 1e-8
        assert abs(res["E"] - res["E_S"]) < 1e-5 * res["E_S"]  # the saddle, to epsf 1e-7
    else:
        assert res["E"] < -2.4e-6  # the ordered minimum (-2.40342e-6)
