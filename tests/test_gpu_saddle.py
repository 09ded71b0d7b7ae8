"""A certified index-1 saddle on the GPU, by the reference's two saddle searches.

Box: one cell of the {110} shell in the reference's metric (k = n / L, L = sqrt 2),
so the 12 wave vectors of |n|^2 = 2 sit at |k| = 1, the minimum of G; Tau, Gamma
and dt are the reference's (Initialization.cpp:7-42).  The disordered field phi = 0
is a minimum whose lowest Hessian eigenvalue, G(1) - Tau, is 12-fold degenerate.

  * main_SPM.cpp:66-122: relax, 10 x 21 LOBPCG iterations with k = 2, then
    Spring_Pairing(30000, 1e-9) from (phi, phi + 0.001 V_1);
  * HISD.cpp:13-30 (critical_point_HISD): HiSD with k = 1 from phi + 0.01 V_1
    along V_1, here to epsf 1e-10 (the driver's 1e-7 stops 2.6e-6 short in E).

Each result is certified on the oracle: its get_Ene and nrmp(ngrad_cam) at the
GPU field (up to 256^3), and (at 32^3) its own LOBPCG at the GPU field, which must show
exactly one negative eigenvalue beside the three translation modes (zero to
1e-7).  Both searches must reach the same saddle energy, and the SPM replicas
must have closed to under 1e-9.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 2 ** 0.5
TAU = -0.0061
# the saddle energy, as the oracle's main_SPM sequence finds it at 32^3
E_SADDLE = 4.65847807e-07


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available()
    return _t


def _np(t):
    return t.detach().cpu().numpy()


def _modes_at_minimum(g, torch, N, seed):
    """relax phi = 0 (it is the disordered minimum), then the main_SPM LOBPCG calls."""
    phi = torch.zeros((N,) * 3, dtype=torch.float64, device="cuda")
    it, hist = g.gradflow(phi, 100, 1e-16)
    assert float(phi.abs().max()) == 0.0
    start = torch.as_tensor(np.random.default_rng(seed).uniform(-1, 1, (2, N, N, N))).cuda()
    V = g.lobpcg_init(start)
    for _ in range(10):
        ritz = g.lobpcg(phi, V, 21)
    return phi, V, ritz


def _certify(po, g, torch, N, x, res_gpu, tol, oracle_index):
    o = po.Oracle(N, L=L)
    xr = _np(x).astype(np.float64)
    gn = o.nrmp(o.ngrad_cam(xr + 0j))
    E_o = o.energy_of(xr)
    assert gn <= tol, gn
    assert abs(gn - res_gpu) <= 1e-6 * res_gpu + 1e-14
    assert abs(g.energy(x) - E_o) <= 1e-12 * abs(E_o)
    assert abs(E_o - E_SADDLE) <= 1e-8 * E_SADDLE, E_o
    # index: the lowest four Hessian eigenvalues at the saddle
    start = np.random.default_rng(7).uniform(-1, 1, (4, N, N, N))
    W = g.lobpcg_init(torch.as_tensor(start).cuda())
    for _ in range(10):
        ritz = g.lobpcg(x, W, 21)
    assert ritz[0] < -4e-3 and ritz[0] > -4.1e-3, ritz
    assert np.all(np.abs(ritz[1:]) < 1e-7), ritz  # translations, no second negative mode
    if oracle_index:
        V = o.lobpcg_initialization(start)
        for _ in range(10):
            ritz_o = o.lobpcg_inip(xr + 0j, V, 21)
        assert np.abs(ritz_o - ritz).max() < 1e-9, (ritz_o, ritz)
    return E_o


@pytest.mark.parametrize("N", [32, 64, 256])
def test_main_spm_workflow_reaches_index1_saddle(spm, po, torch, N):
    """main_SPM.cpp:66-122 on the GPU, certified on the oracle."""
    print("[case] main_SPM workflow N=%d" % N, flush=True)
    with spm.Grid(N, spm.FP64, L=L) as g:
        phi, V, ritz = _modes_at_minimum(g, torch, N, 0)
        # the 12-fold lowest eigenvalue at phi = 0 is G(|k| = 1) - Tau
        assert np.abs(ritz - (spm.get_GK(1.0) - TAU)).max() < 1e-10
        p0 = phi.clone()
        p1 = (phi + 0.001 * V[0]).contiguous()
        it, res, hist = g.spring_pairing(p0, p1, 30000, 1e-9)
        print("spring pairing: %d iterations, res %.3e" % (it, res), flush=True)
        assert it < 30000 and res <= 1e-9
        assert np.sqrt(float(((p1 - p0) ** 2).mean())) < 1e-9  # the replicas have closed
        _certify(po, g, torch, N, p1, res, 1e-9 * (1 + 1e-6), oracle_index=(N == 32))


@pytest.mark.parametrize("N", [32, 64])
def test_critical_point_hisd_reaches_the_same_saddle(spm, po, torch, N):
    """critical_point_HISD (HISD.cpp:13-30) on the GPU: HiSD (k = 1) from the
    minimum perturbed along its lowest mode, certified on the oracle.

    The lowest mode at phi = 0 is any unit field of a 12-fold eigenspace, fixed
    by the LOBPCG start.  From some of them the sweep turns V away from x and x
    relaxes back to phi = 0, in the oracle as on the GPU.  So for four starts the
    outcome (saddle or minimum) and its energy must be the oracle's from the same
    start (at 32^3), and at least one start must climb to the saddle, which is
    then certified."""
    print("[case] critical_point_HISD N=%d" % N, flush=True)
    climbed = []
    with spm.Grid(N, spm.FP64, L=L) as g:
        for seed in range(4):
            phi, V, ritz = _modes_at_minimum(g, torch, N, seed)
            x = (phi + 0.01 * V[0]).contiguous()
            V1 = V[:1].clone()
            x0, v0 = _np(x).copy(), _np(V1).copy()
            it, res = g.hisd(x, V1, dt=0.8, betat=8.0, betau=0.02, epsf=1e-10, maxiter=10000)
            E = g.energy(x)
            print("seed %d: hisd %d steps, res %.3e, E %.12e" % (seed, it, res, E), flush=True)
            assert it < 10000 and res < 1e-10
            saddle = E > 0.5 * E_SADDLE
            assert saddle or abs(E) < 1e-15  # the saddle or the minimum, nothing else
            if N == 32:  # the oracle from the same start
                o = po.Oracle(N, L=L)
                xo, vo = (x0 + 0j).copy(), (v0 + 0j).copy()
                it_o, res_o = o.hiosd_lobpcg_sieh(xo, vo, dt=0.8, betat=8.0, betau=0.02, epsf=1e-10, maxiter=10000)
                E_o = o.energy_of(xo.real)
                assert res_o < 1e-10
                assert (E_o > 0.5 * E_SADDLE) == saddle
                assert abs(E_o - E) <= 1e-9 * E_SADDLE
            if saddle:
                climbed.append(seed)
                if len(climbed) == 1:
                    _certify(po, g, torch, N, x, res, 1e-10, oracle_index=(N == 32))
    assert climbed, "no start climbed to the saddle"


@pytest.mark.parametrize("prec", [0, 1])
def test_lobpcg_block8_256_closed_form(spm, torch, prec):
    """Block-8 LOBPCG (hiosd_lobpcg_inip, LOBPCG.cpp:207-487) at 256^3, the
    configuration's block size: at phi = 0 in the {110} cell the Hessian is
    G(k) - Tau on each Fourier mode, and its lowest eigenvalue G(1) - Tau is
    12-fold (|n|^2 = 2), with the next shell (|n|^2 = 1) 0.07 above.  All eight
    Ritz values must equal it, in FP64 fields and in FP32 fields (FP64 Gram,
    Cholesky-QR and Rayleigh-Ritz)."""
    print("[case] block-8 LOBPCG 256^3 prec=%d" % prec, flush=True)
    N = 256
    with spm.Grid(N, prec, L=L) as g:
        phi = torch.zeros((N,) * 3, dtype=g.dtype, device="cuda")
        start = np.random.default_rng(3).uniform(-1, 1, (8, N, N, N))
        V = g.lobpcg_init(torch.as_tensor(start).to(g.dtype).cuda())
        del start
        for _ in range(3):
            ritz = g.lobpcg(phi, V, 21)
    lam = spm.get_GK(1.0) - TAU
    assert np.abs(ritz - lam).max() < (1e-12 if prec == 0 else 1e-9), ritz - lam
