"""GPU parity of the HIP path (through the C ABI) against the CPU FP64 oracle.

Tolerances: FP64 energy 1e-12 relative, gradient fields 1e-10 in
max-norm.  Transforms are held to 1e-13 relative (FP64) and 2e-6 (FP32).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FFT_NS = [4, 8, 12, 16, 24, 32, 48, 64, 96, 128]


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available(), "GPU tests need a GPU"
    return _t


@pytest.mark.parametrize("N", FFT_NS)
@pytest.mark.parametrize("prec", [0, 1])
def test_my_fft_c2c(spm, po, torch, N, prec):
    """My_FFT.cpp:4-21 c2c forward (1/N^3) and inverse against the oracle."""
    print(f"[case] c2c N={N} prec={prec}", flush=True)
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    o = po.Oracle(N)
    ref_f = o.my_fft(x)
    ref_i = o.my_ifft(x)
    with spm.Grid(N, prec) as g:
        xt = torch.as_tensor(x.astype(np.complex128 if prec == 0 else np.complex64)).cuda()
        f = _np(g.my_fft(xt))
        i = _np(g.my_ifft(xt))
    tol = 1e-13 if prec == 0 else 3e-6
    assert _rel(f, ref_f) < tol
    assert _rel(i, ref_i) < tol


@pytest.mark.parametrize("N", FFT_NS + [192])
@pytest.mark.parametrize("prec", [0, 1])
def test_r2c_c2r(spm, po, torch, N, prec):
    """Half-spectrum transforms equal the oracle's c2c of a real field on kz <= N/2."""
    print(f"[case] r2c/c2r N={N} prec={prec}", flush=True)
    rng = np.random.default_rng(100 + N)
    x = rng.standard_normal((N, N, N))
    o = po.Oracle(N)
    F = o.my_fft(x + 0j)
    with spm.Grid(N, prec) as g:
        xt = torch.as_tensor(x).to(g.dtype).cuda()
        S = _np(g.r2c(xt))
        tol = 1e-13 if prec == 0 else 3e-6
        assert _rel(S[:, :, : N // 2 + 1], F[:, :, : N // 2 + 1]) < tol
        # c2r of the oracle's spectrum equals the oracle's inverse
        St = torch.zeros((N, N, g.nhp), dtype=g.cdtype, device="cuda")
        St[:, :, : N // 2 + 1] = torch.as_tensor(F[:, :, : N // 2 + 1]).to(g.cdtype)
        y = _np(g.c2r(St))
        assert _rel(y, o.my_ifft(F).real) < tol
        # c2r must not modify its input
        assert _rel(_np(St)[:, :, : N // 2 + 1], F[:, :, : N // 2 + 1]) < tol


@pytest.mark.parametrize("prec", [0, 1])
def test_roundtrip_256(spm, torch, prec):
    """Size-independent property at the full size: c2r(r2c(x)) == x."""
    print(f"[case] roundtrip 256 prec={prec}", flush=True)
    N = 256
    with spm.Grid(N, prec) as g:
        gen = torch.Generator(device="cuda").manual_seed(3)
        x = torch.randn((N,) * 3, dtype=g.dtype, device="cuda", generator=gen)
        y = g.c2r(g.r2c(x))
        err = (y - x).abs().max().item() / x.abs().max().item()
        assert err < (2e-15 * 64 if prec == 0 else 3e-6)


def _field(seeds, N, kind, seed=0):
    if kind == "gauss":
        return seeds.gaussian(N, 0.05, seed)
    if kind == "ico":
        return seeds.icosahedral_nucleus(N, seed=seed)
    return seeds.crystal(N, kind, seed=seed)


@pytest.mark.parametrize("N,kind", [(16, "gauss"), (32, "gauss"), (64, "ico"), (48, "bcc"), (96, "fcc")])
def test_energy_ngrad(spm, po, seeds, torch, N, kind):
    """get_Ene (LPsrc.cpp:8-38) and ngrad_cam (LPsrc.cpp:119-173), FP64."""
    print(f"[case] energy/ngrad N={N} {kind}", flush=True)
    phi = _field(seeds, N, kind)
    o = po.Oracle(N)
    E_ref = o.energy_of(phi)
    g_ref = o.ngrad_cam(phi + 0j).real
    n_ref = o.nrmp(g_ref + 0j)
    with spm.Grid(N, 0) as g:
        pt = torch.as_tensor(phi).cuda()
        E = g.energy(pt)
        gr, nrm = g.ngrad(pt)
        gr = _np(gr)
    assert abs(E - E_ref) <= 1e-12 * abs(E_ref)
    assert np.abs(gr - g_ref).max() <= 1e-10
    assert abs(nrm - n_ref) <= 1e-10 * max(n_ref, 1e-30) + 1e-15


@pytest.mark.parametrize("N,kind", [(16, "gauss"), (32, "gauss"), (64, "ico")])
def test_hv_precond(spm, po, seeds, torch, N, kind):
    """hv_cam and precond_camnew (LOBPCG.cpp:113-199), FP64."""
    print(f"[case] hv/precond N={N} {kind}", flush=True)
    phi = _field(seeds, N, kind, 1)
    v = np.random.default_rng(7).uniform(-1, 1, (N, N, N))
    o = po.Oracle(N)
    hv_ref = o.hv_cam(phi + 0j, v + 0j).real
    pc_ref = o.precond_camnew(v + 0j).real
    with spm.Grid(N, 0) as g:
        pt, vt = torch.as_tensor(phi).cuda(), torch.as_tensor(v).cuda()
        hv = _np(g.hv(pt, vt))
        w = vt.clone()
        pc = _np(g.precond(w))
    assert _rel(hv, hv_ref) < 1e-12
    assert _rel(pc, pc_ref) < 1e-12


@pytest.mark.parametrize("N,kind,steps", [(32, "gauss", 60), (64, "gauss", 40), (64, "ico", 30)])
def test_gradflow_fp64(spm, po, seeds, torch, N, kind, steps):
    """get_min_u_F (LPsrc.cpp:40-117): per-step E, ||grad F||, residual and the
    final field against the oracle run from the same field."""
    print(f"[case] gradflow N={N} {kind} steps={steps}", flush=True)
    phi = _field(seeds, N, kind)
    o = po.Oracle(N)
    uR = phi.astype(np.complex128)
    uF = o.my_fft(uR)
    it, E, res, gn = o.get_min_u_F(uR, uF, steps, 1e-16, gnorm=True)
    with spm.Grid(N, 0) as g:
        pt = torch.as_tensor(phi).cuda()
        itg, hist = g.gradflow(pt, steps, 1e-16, 0.0, check_every=7)
        out = _np(pt)
    assert itg == it == steps
    assert np.all(np.abs(hist[:, 0] - E) <= 1e-12 * np.abs(E))
    assert np.all(np.abs(hist[:, 1] - gn) <= 1e-9 * gn + 1e-14)
    assert np.all(np.abs(hist[:, 2] - res) <= 1e-8 * res + 1e-15)
    assert np.abs(out - uR.real).max() <= 1e-10


@pytest.mark.parametrize("kscale", [1.0, 2.0 * np.pi])
def test_config1_64_500_steps(spm, po, seeds, torch, kscale):
    """The 64^3 workload of the benchmark configurations: an i.i.d. N(0, 0.05^2)
    seed (seed 0, mean removed), 500 get_min_u_F steps (LPsrc.cpp:40-117) with
    E and ||grad F|| logged at every step, against the oracle (default step:
    two passes at 64^3), on the reference's k = n / L grid and on the
    k = 2 pi n / L grid the configuration names."""
    print("[case] config 1: 64^3, 500 steps, kscale %g" % kscale, flush=True)
    N, steps = 64, 500
    phi = seeds.gaussian(N, 0.05, 0)
    o = po.Oracle(N, kscale=kscale)
    uR = phi.astype(np.complex128)
    uF = o.my_fft(uR)
    it, E, res, gn = o.get_min_u_F(uR, uF, steps, -1.0, gnorm=True)
    with spm.Grid(N, 0, kscale=kscale) as g:
        pt = torch.as_tensor(phi).cuda()
        itg, hist = g.gradflow(pt, steps, -1.0, 0.0, check_every=64)
        out = _np(pt)
    assert itg == it == steps
    assert np.all(np.abs(hist[:, 0] - E) <= 1e-12 * np.abs(E))
    assert np.all(np.abs(hist[:, 1] - gn) <= 1e-9 * gn + 1e-14)
    assert np.abs(out - uR.real).max() <= 1e-10


def test_gradflow_stops_like_reference(spm, po, seeds, torch):
    """The residual stop (LPsrc.cpp:52): same iteration count as the oracle."""
    N = 16
    phi = seeds.gaussian(N, 0.05, 4)
    o = po.Oracle(N)
    uR = phi.astype(np.complex128)
    uF = o.my_fft(uR)
    tol = 1e-9
    it, E, res, _ = o.get_min_u_F(uR, uF, 5000, tol)
    with spm.Grid(N, 0) as g:
        pt = torch.as_tensor(phi).cuda()
        itg, hist = g.gradflow(pt, 5000, tol, 0.0, check_every=50)
    assert it < 5000 and itg == it
    assert np.abs(_np(pt) - uR.real).max() <= 1e-10


@pytest.mark.parametrize("N", [32, 64])
def test_gradflow_fp32(spm, po, seeds, torch, N):
    """Mixed precision: FP32 fields and FFTs, FP64 sums; energy within 1e-6."""
    print(f"[case] gradflow fp32 N={N}", flush=True)
    phi = seeds.gaussian(N, 0.05, 0)
    o = po.Oracle(N)
    uR = phi.astype(np.complex128)
    uF = o.my_fft(uR)
    steps = 30
    _, E, _, _ = o.get_min_u_F(uR, uF, steps, 1e-16)
    with spm.Grid(N, 1) as g:
        pt = torch.as_tensor(phi).float().cuda()
        itg, hist = g.gradflow(pt, steps, 1e-16, 0.0, check_every=10)
    assert itg == steps
    assert np.all(np.abs(hist[:, 0] - E) <= 1e-5 * np.abs(E))


@pytest.mark.parametrize("prec", [0, 1])
def test_gradflow_graph_replay_equals_single_launches(spm, seeds, prec):
    """Iterations replayed from the captured graph (16 at a time, host checks
    and copies between replays) give exactly the single-launch iterations."""
    import torch

    N, n = 32, 53
    phi = seeds.crystal(N, "bcc", 26.0, 2 * np.pi, 0.1, 4)
    out = []
    for check in (1, 32):  # 1: every iteration launched on its own; 32: two graph replays per check
        with spm.Grid(N, prec, kscale=2 * np.pi, Tau=0.005) as g:
            t = torch.as_tensor(phi).to(g.dtype).cuda()
            it, hist = g.gradflow(t, n, -1.0, 0.0, check_every=check)
            out.append((it, hist.copy(), t.cpu().numpy()))
    (i1, h1, f1), (i2, h2, f2) = out
    assert i1 == i2 == n
    assert np.array_equal(h1, h2) and np.array_equal(f1, f2)
