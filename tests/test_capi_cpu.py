"""CPU checks of the product library (no GPU calls): it loads, exports every
entry point declared in include/spm_iqc.h, the Python binding declares the
same set, and the host-only helpers agree with the reference's formulas."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "spm_iqc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(spm_\w+)\s*\(", src)))


def test_library_exports_header(spm):
    names = _declared()
    assert len(names) >= 20
    L = ctypes.CDLL(spm.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header(spm):
    assert sorted(spm.SIGNATURES) == _declared()
    spm.lib()  # every declared signature binds


def test_get_GK_matches_oracle(spm, po):
    for k2 in np.linspace(0.0, 5.0, 101):
        assert spm.get_GK(k2) == po.get_GK(k2)


def test_seeds(seeds):
    q = seeds.icosahedral_star()
    assert q.shape == (30, 3)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0)
    # closed under inversion; five-fold-compatible angle set
    for v in q:
        assert np.min(np.linalg.norm(q + v, axis=1)) < 1e-12
    dots = np.round(q @ q.T, 9)
    allowed = {0.0, 0.5, -0.5, 1.0, -1.0}
    g = (1 + 5 ** 0.5) / 2
    allowed |= {round(s * g / 2, 9) for s in (1, -1)} | {round(s / (2 * g), 9) for s in (1, -1)}
    assert set(np.unique(dots)) <= allowed
    m = seeds.icosahedral_nucleus(32, seed=3)
    assert m.shape == (32, 32, 32) and np.isfinite(m).all()
    # |n| = 26 for the reference metric: the minimum of G at |k| = 1
    n = seeds.grid_modes(q, 26.0, 1.0)
    assert np.all(np.abs(np.linalg.norm(n, axis=1) - 26) < 1.0)
    x = seeds.gaussian(16, 0.05, 0)
    assert abs(x.mean()) < 1e-17 and abs(x.std() - 0.05) < 0.005


def test_step_passes_validated_before_any_device_work(spm):
    """spm_create rejects a bad step_passes value (and the two-pass step at an N
    other than 64, 128 and 256) during parameter checks, before it touches a device."""
    import ctypes

    lib = spm.lib()
    h = ctypes.c_void_p()
    for N, passes, want in ((64, 3, -1), (96, 2, -4), (192, 2, -4), (256, -1, -1)):
        p = spm._params(N, spm.FP64, 26.0, 1.0, -0.0061, 0.30, 0.3, 0, passes)
        assert lib.spm_create(ctypes.byref(p), ctypes.byref(h)) == want
        assert not h.value


def test_compat_library_defines_reference_api():
    """every function include/compat/spm_reference_api.h declares (the
    reference's signatures) is defined in the compat library"""
    import shutil
    import subprocess

    import pytest

    so = os.path.join(ROOT, "spm-iqc_amd", "lib", "libspm_iqc_compat.so")
    if not os.path.exists(so) or shutil.which("nm") is None:
        pytest.skip("compat library not built or nm missing")
    src = open(os.path.join(ROOT, "include", "compat", "spm_reference_api.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    src = re.sub(r"#[^\n]*", "", src)
    src = re.sub(r"struct\s+\w+\s*\{.*?\};", "", src, flags=re.S)
    names = set()
    for stmt in src.split(";"):
        stmt = " ".join(stmt.split())
        if stmt.startswith(("extern", "typedef")):
            continue
        m = re.match(r"[\w\s\*]+?\b(\w+)\s*\(", stmt)
        if m:
            names.add(m.group(1))
    assert len(names) >= 40, sorted(names)
    out = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    defined = set(re.findall(r" T (\w+)\(", out))
    missing = sorted(names - defined)
    assert not missing, missing
