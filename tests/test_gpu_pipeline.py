"""The stage-1 pipeline (main_SPM.cpp:15-163: relax, lowest modes, SPM) in FP64
and in mixed precision (FP32 fields with FP64 sums and an FP64 polish).

Pins: the relaxed field and energy against the oracle's get_min_u_F run for
the same number of iterations; the eigenvalues at the relaxed (disordered,
phi = 0) state against the closed form G(k) - Tau of the lowest grid shells
(hv_cam at phi = 0 is diagonal in Fourier space); mixed against FP64 for every
reported quantity.  Tolerances: energy 1e-12 relative (FP64), eigenvalues
1e-9 absolute, saddle energy 1e-10."""
import importlib
import itertools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 24
KS = 2.0 * np.pi  # k = 2 pi n / L: the |q| = 1 star is resolved on a 24^3 grid
PARAMS = dict(L=26.0, kscale=KS, Tau=-0.0061, Gamma=0.30, dt=0.3)
# relaxed far enough that the first-order shift 2 Gamma phi of the Hessian at
# the minimum (|phi| ~ ||grad F|| / 0.006) stays far below 1e-9
GN_TOL = 1e-12
MAX_ITER = 30000


@pytest.fixture(scope="module")
def pipe():
    return importlib.import_module("spm-iqc_amd.pipeline")


@pytest.fixture(scope="module")
def seed_field(seeds):
    return seeds.icosahedral_nucleus(N, 26.0, KS, 0.1, 0)


def _closed_form_lowest(po, n):
    """Lowest n eigenvalues of hv_cam at phi = 0: G(kscale^2 |m|^2 / L^2) - Tau
    over the nonzero grid modes m (each Fourier mode is one real eigenvector)."""
    h = N // 2
    vals = []
    for a, b, c in itertools.product(range(-h + 1, h + 1), repeat=3):
        m2 = a * a + b * b + c * c
        if m2:
            vals.append(po.get_GK(KS * KS * m2 / (PARAMS["L"] ** 2)) - PARAMS["Tau"])
    return np.sort(vals)[:n]


@pytest.fixture(scope="module")
def runs(pipe, seed_field):
    """relax + lowest modes in both precisions; the SPM stage of both starts
    from the FP64 minimum and mode (the lowest shell is degenerate, so the two
    runs may return different vectors of it)."""
    out = {}
    for p in ("fp64", "mixed"):
        phi, ri = pipe.relax(N, seed_field, p, gn_tol=GN_TOL, max_iter=MAX_ITER, polish=1000, params=PARAMS)
        ritz, V, mi = pipe.lowest_modes(N, phi, p, k=8, nwant=4, calls=6, modes_polish=1, params=PARAMS)
        out[p] = {"relax": ri, "eigenvalues": ritz, "modes": mi, "phi": phi, "v1": V[0].clone()}
    phi, v1 = out["fp64"]["phi"], out["fp64"]["v1"]
    for p in ("fp64", "mixed"):
        p0, p1, si = pipe.saddle(N, phi, v1, p, offset=0.01, maxiter=300, tol=1e-9, spm_polish=100, params=PARAMS)
        out[p].update(saddle=si, phi0=p0, phi1=p1)
    for r in out.values():
        for k in ("phi", "v1", "phi0", "phi1"):
            r[k] = r[k].cpu().numpy()
    return out


def test_relax_fp64_matches_oracle(po, runs, seed_field):
    r = runs["fp64"]
    it = r["relax"]["iters"]
    assert r["relax"]["gn"] <= GN_TOL and 0 < it < MAX_ITER
    o = po.Oracle(N, **PARAMS)
    uR = seed_field.astype(np.complex128)
    uF = o.my_fft(uR)
    n, _, _, _ = o.get_min_u_F(uR, uF, it, -1.0)
    assert n == it
    E_ref = o.energy_of(uR.real)
    assert abs(r["relax"]["E"] - E_ref) <= 1e-12 * abs(E_ref) + 1e-20
    assert np.abs(r["phi"] - uR.real).max() <= 1e-10


def test_relax_mixed_vs_fp64(runs):
    r64, rmx = runs["fp64"], runs["mixed"]
    assert rmx["relax"]["iters_fp32"] > 0
    assert rmx["relax"]["gn"] <= GN_TOL
    # both reach the same minimum, to what the tolerance leaves open
    assert np.abs(rmx["phi"] - r64["phi"]).max() <= 1e-7
    assert abs(rmx["relax"]["E"] - r64["relax"]["E"]) <= 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_eigenvalues_closed_form(po, runs, prec):
    r = runs[prec]
    assert np.abs(r["phi"]).max() < 1e-8  # the seed dissolves: the minimum is phi = 0
    ref = _closed_form_lowest(po, 4)
    got = np.asarray(r["eigenvalues"])
    assert np.abs(got - ref).max() <= 1e-9, (got, ref)
    assert np.abs(np.asarray(r["modes"]["rayleigh"]) - ref).max() <= 1e-9
    if prec == "mixed":
        assert r["modes"]["calls_fp32"] == 5 and r["modes"]["calls_fp64"] == 1


def test_saddle_fp64_matches_oracle(po, runs):
    """Spring_Pairing from the same replica pair on the oracle."""
    r = runs["fp64"]
    o = po.Oracle(N, **PARAMS)
    P0 = (r["phi"] - 0.01 * r["v1"]).astype(np.complex128)
    P1 = (r["phi"] + 0.01 * r["v1"]).astype(np.complex128)
    it, res = o.Spring_Pairing(P0, P1, 300, 1e-9)
    s = r["saddle"]
    assert it == s["iters"]
    E_ref = o.energy_of(P1.real)
    assert abs(s["E"] - E_ref) <= 1e-10 * abs(E_ref) + 1e-18
    assert np.abs(r["phi1"] - P1.real).max() <= 1e-9
    assert abs(s["res"] - res) <= 1e-9 * abs(res) + 1e-15


def test_saddle_mixed_vs_fp64(runs):
    """Mixed: 200 FP32 iterations and 100 FP64 ones against 300 FP64 ones
    from the same pair.  From phi = 0 the min-max force climbs along the mode
    (the pair moves away from the minimum: growth 1 + lambda / (1 + G) per
    iteration), so the FP32 rounding of every iteration accumulates instead
    of being damped; measured 5e-5 relative in E.  The bound states it."""
    r64, rmx = runs["fp64"], runs["mixed"]
    assert rmx["saddle"]["iters_fp32"] == 200 and rmx["saddle"]["iters_fp64"] == 100
    E64 = r64["saddle"]["E"]
    assert abs(rmx["saddle"]["E"] - E64) <= 1e-3 * abs(E64)
    assert np.abs(rmx["phi1"] - r64["phi1"]).max() <= 1e-3 * np.abs(r64["phi1"]).max()


@pytest.mark.parametrize("Nc", [64, 128])
def test_cell_workflow_mixed_matches_fp64_at_converged_states(pipe, Nc):
    """Both stages of the workflow in the {110} cell (pipeline.cell_workflow):
    the index-1 saddle (Spring_Pairing to res 1e-9) and the ordered minimum below
    it (gradient flow to ||grad F|| 1e-8), each reached by its tolerance in both
    precisions.  Mixed (FP32 stages handing over to FP64) must give the FP64
    energies to 1e-9 relative and the same Hessian eigenvalues to 1e-9, with
    index 1 at the saddle and 0 at the minimum."""
    print("[case] cell workflow fp64 vs mixed N=%d" % Nc, flush=True)
    rep = {p: pipe.cell_workflow(Nc, p)[0] for p in ("fp64", "mixed")}
    for p in ("fp64", "mixed"):
        s, m = rep[p]["saddle"], rep[p]["minimum"]
        assert s["res"] <= 1e-9 and s["iters"] < 30000 and s["index"] == 1, (p, s)
        assert m["gn"] <= 1e-8 and m["index"] == 0 and m["E"] < 0, (p, m)
    f, x = rep["fp64"], rep["mixed"]
    assert x["saddle"]["iters_fp32"] > 0 and x["minimum"]["iters_fp32"] > 0  # the FP32 stages did the bulk
    for stage in ("saddle", "minimum"):
        assert abs(x[stage]["E"] - f[stage]["E"]) <= 1e-9 * abs(f[stage]["E"]), stage
        assert np.abs(np.array(x[stage]["ritz"]) - np.array(f[stage]["ritz"])).max() < 1e-9, stage
