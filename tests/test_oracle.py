"""CPU tests that pin the FP64 oracle (oracle/lp_oracle.c) independently of
its own formulas: the DFT definition (numpy's pocketfft and brute force),
closed forms of the model on single Fourier modes, finite differences, the
exactly diagonal Hessian at phi = 0 and a dense-Hessian brute force.

The reference ships no tests, golden vectors or fixtures for this path (its
result/ directory holds plotting scripts only), so these checks are the pins.
"""
import numpy as np
import pytest

L, TAU, GAMMA, DT = 26.0, -0.0061, 0.30, 0.3


def _mode(N, n, A, axis=0, phase=0.0):
    i = np.arange(N)
    c = A * np.cos(2 * np.pi * n * i / N + phase)
    shape = [1, 1, 1]
    shape[axis] = N
    return np.broadcast_to(c.reshape(shape), (N, N, N)).copy()


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("N", [2, 3, 4, 5, 6, 8, 10, 12, 16, 24, 32, 48])
def test_fft_matches_numpy(po, N):
    """My_FFT.cpp:4-21: forward / N^3 (FFTW_FORWARD sign), inverse unnormalised."""
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N,) * 3) + 1j * rng.standard_normal((N,) * 3)
    o = po.Oracle(N)
    assert np.abs(o.my_fft(x) - np.fft.fftn(x) / N ** 3).max() < 1e-15 * N ** 1.5
    assert np.abs(o.my_ifft(x) - np.fft.ifftn(x) * N ** 3).max() < 1e-13 * N ** 1.5


@pytest.mark.parametrize("N", [3, 4, 6])
def test_fft_brute_force_dft(po, N):
    """The DFT definition itself, independent of any FFT library."""
    rng = np.random.default_rng(10 + N)
    x = rng.standard_normal((N,) * 3) + 1j * rng.standard_normal((N,) * 3)
    idx = np.array(np.unravel_index(np.arange(N ** 3), (N,) * 3)).T
    ph = np.exp(-2j * np.pi * (idx @ idx.T) / N)
    ref = (ph @ x.ravel()).reshape((N,) * 3) / N ** 3
    assert np.abs(po.Oracle(N).my_fft(x) - ref).max() < 1e-14


def test_parseval(po):
    N = 16
    x = np.random.default_rng(0).standard_normal((N,) * 3) + 0j
    X = po.Oracle(N).my_fft(x)
    assert abs(np.sum(np.abs(x) ** 2) / N ** 3 - np.sum(np.abs(X) ** 2)) < 1e-13


# ---------------------------------------------------------------- pair kernel
def test_get_GK_values(po):
    """Initialization.cpp:150-166: G(0) = d0 + 0.08; near-zero minima at |k| = 1
    and |k| = 0.618 (the two IQC length scales)."""
    assert po.get_GK(0.0) == pytest.approx(2.04918291908034, abs=1e-14)
    G = np.vectorize(po.get_GK)
    for k0 in (1.0, 0.618):
        k = np.linspace(k0 - 0.05, k0 + 0.05, 2001)
        g = G(k * k)
        assert g.min() < 1e-7
        assert abs(k[np.argmin(g)] - k0) < 2e-3
    assert po.get_GK(100.0) == pytest.approx(0.08, abs=1e-12)  # the constant tail


def test_tables_follow_reference_metric(po):
    """k^2 = sum (n_i / L)^2 with n in FFT order (Initialization.cpp:97-148),
    ikt2 = 1 / (G + |Tau| + 0.02)."""
    N = 8
    o = po.Oracle(N)
    ksq, G, ikt2 = o.tables()
    n = np.fft.fftfreq(N, 1.0 / N)
    n[N // 2] = -N // 2
    ref = (n[:, None, None] ** 2 + n[None, :, None] ** 2 + n[None, None, :] ** 2) / L ** 2
    assert np.abs(ksq - ref).max() < 1e-16
    assert np.abs(G - np.vectorize(po.get_GK)(ref)).max() < 1e-15
    assert np.abs(ikt2 - 1.0 / (G + abs(TAU) + 0.02)).max() < 1e-15


# ---------------------------------------------------------------- energy and gradient
@pytest.mark.parametrize("N,n,A", [(16, 3, 0.1), (32, 5, 0.3), (24, 2, 0.05)])
def test_energy_single_mode_closed_form(po, N, n, A):
    """E[A cos(k.x)] = A^2 G(k)/4 - Tau A^2/4 + 3 A^4/32 (no aliasing: 4n < N/2)."""
    phi = _mode(N, n, A, axis=1)
    E = po.Oracle(N).energy_of(phi)
    ref = A * A * po.get_GK((n / L) ** 2) / 4 - TAU * A * A / 4 + 3 * A ** 4 / 32
    assert abs(E - ref) < 1e-14


def test_ngrad_single_mode_closed_form(po):
    """-grad F for phi = A cos(t): -[(G - Tau + 3A^2/4) A cos t
    - (Gamma A^2/2) cos 2t + (A^3/4) cos 3t]."""
    N, n, A = 32, 3, 0.2
    phi = _mode(N, n, A, axis=2)
    t = 2 * np.pi * n * np.arange(N) / N
    G = po.get_GK((n / L) ** 2)
    g1 = (G - TAU + 0.75 * A * A) * A * np.cos(t) - 0.5 * GAMMA * A * A * np.cos(2 * t) + 0.25 * A ** 3 * np.cos(3 * t)
    ref = -np.broadcast_to(g1[None, None, :], (N, N, N))
    out = po.Oracle(N).ngrad_cam(phi + 0j)
    assert np.abs(out.real - ref).max() < 1e-15
    assert np.abs(out.imag).max() == 0.0


def test_ngrad_finite_difference(po, seeds):
    """dE/de (phi + e v) = -<ngrad, v> for mean-free v (inpfp metric)."""
    N = 16
    o = po.Oracle(N)
    phi = seeds.gaussian(N, 0.1, 1)
    v = seeds.gaussian(N, 1.0, 2)
    g = o.ngrad_cam(phi + 0j).real
    eps = 1e-4
    d = (o.energy_of(phi + eps * v) - o.energy_of(phi - eps * v)) / (2 * eps)
    assert abs(d + np.mean(g * v)) < 1e-9 * max(1.0, abs(d))


def test_hv_finite_difference(po, seeds):
    """hv_cam(phi, v) = -d/de ngrad(phi + e v) for mean-free v.  ngrad is cubic
    in phi, so Richardson's combination of two central differences is exact."""
    N = 16
    o = po.Oracle(N)
    phi = seeds.gaussian(N, 0.1, 3)
    v = seeds.gaussian(N, 1.0, 4)

    def D(eps):
        return -(o.ngrad_cam(phi + eps * v + 0j).real - o.ngrad_cam(phi - eps * v + 0j).real) / (2 * eps)

    fd = (4 * D(1e-3) - D(2e-3)) / 3
    hv = o.hv_cam(phi + 0j, v + 0j).real
    assert np.abs(hv - fd).max() < 1e-9 * np.abs(hv).max()


def test_precond_single_mode(po):
    """precond_camnew on a single mode: divides by G + |Tau| + 0.02."""
    N, n = 16, 4
    v = _mode(N, n, 1.0, axis=0, phase=0.3)
    out = po.Oracle(N).precond_camnew(v + 0j).real
    ref = v / (po.get_GK((n / L) ** 2) + abs(TAU) + 0.02)
    assert np.abs(out - ref).max() < 1e-14


def test_gradflow_one_step_per_coefficient(po):
    """One get_min_u_F step on phi = A cos(t) (LPsrc.cpp:56-75), coefficient by
    coefficient: u'(m) = (u(m) + dt f(m)) / (1 + dt G(m)), DC = 0."""
    N, n, A = 32, 3, 0.2
    o = po.Oracle(N)
    phi = _mode(N, n, A, axis=0)
    uR = phi + 0j
    uF = o.my_fft(uR)
    it, E, res, _ = o.get_min_u_F(uR, uF, 1, -1.0)
    Gk = lambda m: po.get_GK((m / L) ** 2)
    ref = {
        n: (A / 2 + DT * (TAU * A - 0.75 * A ** 3) / 2) / (1 + DT * Gk(n)),
        2 * n: DT * (GAMMA * A * A / 4) / (1 + DT * Gk(2 * n)),
        3 * n: DT * (-(A ** 3) / 8) / (1 + DT * Gk(3 * n)),
    }
    for m, val in ref.items():
        assert abs(uF[m, 0, 0] - val) < 1e-15
        assert abs(uF[N - m, 0, 0] - val) < 1e-15
    assert abs(uF[0, 0, 0]) == 0.0
    rest = uF.copy()
    for m in ref:
        rest[m, 0, 0] = rest[N - m, 0, 0] = 0
    assert np.abs(rest).max() < 1e-16
    assert E[0] == pytest.approx(A * A * Gk(n) / 4 - TAU * A * A / 4 + 3 * A ** 4 / 32, abs=1e-15)


def test_gradflow_energy_decreases(po, seeds):
    """Semi-implicit gradient flow with dt = 0.3 lowers the energy each step."""
    N = 16
    o = po.Oracle(N)
    uR = seeds.gaussian(N, 0.05, 0) + 0j
    uF = o.my_fft(uR)
    _, E, res, gn = o.get_min_u_F(uR, uF, 40, 1e-16, gnorm=True)
    assert np.all(np.diff(E) <= 1e-15)
    assert gn[-1] < gn[0]


# ---------------------------------------------------------------- Hessian / LOBPCG
def test_sym_eig_matches_numpy(po):
    rng = np.random.default_rng(5)
    for n in (1, 2, 5, 13, 24):
        A = rng.standard_normal((n, n))
        A = A + A.T
        w, Z = po.sym_eig(A)
        assert np.abs(w - np.linalg.eigvalsh(A)).max() < 1e-12
        assert np.abs(A @ Z - Z * w).max() < 1e-11


def test_lobpcg_phi0_closed_form(po):
    """At phi = 0 the Hessian is diag(G(k) - Tau) on mean-free fields."""
    N, k = 12, 4
    i = np.arange(N)
    K = np.where(i < N // 2, i, i - N)
    n2 = K[:, None, None] ** 2 + K[None, :, None] ** 2 + K[None, None, :] ** 2
    ev = np.sort(np.vectorize(po.get_GK)(n2[n2 > 0] / L ** 2) - TAU)
    o = po.Oracle(N)
    V = o.lobpcg_initialization(np.random.default_rng(1).uniform(-1, 1, (k, N, N, N)))
    ritz = o.lobpcg_inip(np.zeros((N,) * 3, complex), V, 40)
    assert np.abs(ritz - ev[:k]).max() < 1e-12


def test_lobpcg_dense_bruteforce(po, seeds):
    """Lowest eigenvalues of the dense Hessian (N = 6, mean-free subspace)."""
    N, k = 6, 3
    R = N ** 3
    phi = seeds.gaussian(N, 0.3, 5)
    o = po.Oracle(N)
    _, G, _ = o.tables()
    E = np.eye(R).reshape(R, N, N, N)
    conv = np.fft.ifftn(G[None] * np.fft.fftn(E, axes=(1, 2, 3)), axes=(1, 2, 3)).real.reshape(R, R)
    A = conv + np.diag((-TAU - 2 * GAMMA * phi + 3 * phi * phi).ravel())
    Q, _ = np.linalg.qr(np.eye(R)[:, 1:] - 1.0 / R)
    ev = np.linalg.eigvalsh(Q.T @ (0.5 * (A + A.T)) @ Q)
    V = o.lobpcg_initialization(np.random.default_rng(2).uniform(-1, 1, (k, N, N, N)))
    for _ in range(10):
        ritz = o.lobpcg_inip(phi + 0j, V, 21)
    assert np.abs(ritz - ev[:k]).max() < 1e-10
    # rayleighq of a converged vector
    assert abs(o.rayleighq(phi + 0j, V[0]) - ev[0]) < 1e-10


def test_lobpcg_init_orthonormal(po):
    N, k = 8, 5
    V = po.Oracle(N).lobpcg_initialization(np.random.default_rng(3).uniform(-1, 1, (k, N, N, N))).real
    Gm = np.einsum("aijk,bijk->ab", V, V) / N ** 3
    assert np.abs(Gm - np.eye(k)).max() < 1e-13
    assert np.abs(V.mean(axis=(1, 2, 3))).max() < 1e-15


# ---------------------------------------------------------------- SPM
def test_spm_forces_fix_a_critical_point(po, seeds):
    """At a critical point (ngrad = 0) both force steps leave the field
    unchanged: (x_hat + nln_hat)/(1 + G) = x_hat when G x_hat = nln_hat."""
    N = 16
    o = po.Oracle(N)
    uR = seeds.gaussian(N, 0.05, 6) + 0j
    uF = o.my_fft(uR)
    o.get_min_u_F(uR, uF, 5000, 1e-17)
    assert o.nrmp(o.ngrad_cam(uR)) < 1e-12
    v = o.get_Tv(uR, uR + seeds.gaussian(N, 1.0, 7))
    for f in (o.Natural_Forces, o.MinMax_Forces):
        x = uR.copy()
        f(x, v)
        assert np.abs(x.real - uR.real).max() < 1e-12


def test_spm_tangent_unit(po, seeds):
    N = 8
    o = po.Oracle(N)
    a, b = seeds.gaussian(N, 0.1, 1) + 0j, seeds.gaussian(N, 0.1, 2) + 0j
    v = o.get_Tv(a, b)
    assert abs(o.nrmp(v) - 1.0) < 1e-15
    assert np.abs(v.real * o.nrmp(b - a) - (b - a).real).max() < 1e-15


def test_spring_pairing_exit_semantics(po, seeds):
    """Spring_Pairing's loop (SPM.cpp:191-230) keeps the largest residual with
    `if (res < resi)`, so a replica that turns non-finite leaves res at 0 and
    ends the loop: from a random direction at the disordered minimum the
    min-max ascent has no saddle to find and runs away."""
    N = 8
    o = po.Oracle(N)
    uR = seeds.gaussian(N, 0.05, 1) + 0j
    uF = o.my_fft(uR)
    o.get_min_u_F(uR, uF, 50, 1e-16)
    p0, p1 = uR.copy(), uR + 0.001 * seeds.gaussian(N, 1.0, 2)
    it1, r1 = o.Spring_Pairing(p0.copy(), p1.copy(), 1, 1e-30)
    it5, r5 = o.Spring_Pairing(p0.copy(), p1.copy(), 5, 1e-30)
    assert it1 == 1 and it5 == 5 and 0 < r1 < r5
    a, b = p0.copy(), p1.copy()
    it, r = o.Spring_Pairing(a, b, 400, 1e-30)
    assert it < 400 and r == 0.0 and np.abs(b.real).max() > 1e100


# ---------------------------------------------------------------- HiSD
def _flip(a):
    """a(-r) on the periodic grid."""
    return np.roll(a[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))


def test_hisd_step_is_gradient_flow_for_even_field_and_odd_mode(po, seeds):
    """hiosd_lobpcg_sieh (HISD.cpp:188-268) on an even field with an odd mode v:
    <f, v> = 0, so the reflected force is f itself, and the first Barzilai-Borwein
    step is |<f + dt, f>| / <f, f> = 1 (mean(f) = 0). The first x-step is then one
    get_min_u_F step with dt = 1 (LPsrc.cpp:56-84)."""
    N = 16
    y = seeds.gaussian(N, 0.05, 3)
    x0 = 0.5 * (y + _flip(y))
    z = seeds.gaussian(N, 1.0, 4)
    v = 0.5 * (z - _flip(z))
    v /= np.sqrt(np.mean(v * v))
    o = po.Oracle(N, dt=1.0)
    x = (x0 + 0j).copy()
    V = (v + 0j).reshape((1, N, N, N)).copy()
    it, res = o.hiosd_lobpcg_sieh(x, V, dt=0.8, maxiter=1, epsf=-1.0)
    assert it == 1
    uR = (x0 + 0j).copy()
    uF = o.my_fft(uR)
    o.get_min_u_F(uR, uF, 1, -1.0)
    assert np.abs(x.real - uR.real).max() < 1e-14
    assert np.abs(x.imag).max() < 1e-18  # rounding of the complex inverse transform


def test_hisd_step_keeps_the_even_part(po, seeds):
    """The x-step keeps only the real part of the new spectrum (HISD.cpp:254-258),
    so from any start the field after one step is even: x(-r) = x(r)."""
    N = 12
    o = po.Oracle(N)
    x = (seeds.gaussian(N, 0.05, 5) + 0j).copy()
    V = (seeds.gaussian(N, 1.0, 6) + 0j).reshape((1, N, N, N)).copy()
    o.hiosd_lobpcg_sieh(x, V, maxiter=1, epsf=-1.0)
    assert np.abs(x.real - _flip(x.real)).max() < 1e-15


def test_hisd_eigen_sweep_is_lobpcg(po, seeds):
    """The eigen-direction part of a HiSD step (HISD.cpp:282-441) is one
    hiosd_lobpcg_inip sweep (LOBPCG.cpp:262-449) at the new x, started from the
    same block with Vp = 0."""
    N, k = 12, 2
    o = po.Oracle(N)
    x = (seeds.gaussian(N, 0.05, 7) + 0j).copy()
    V0 = np.stack([seeds.gaussian(N, 1.0, 8 + i) for i in range(k)]) + 0j
    V = V0.copy()
    o.hiosd_lobpcg_sieh(x, V, maxiter=1, epsf=-1.0)
    W = V0.copy()
    o.lobpcg_inip(x, W, maxiter=1)
    assert np.abs(V - W).max() < 1e-12


def test_gosd_without_null_space_is_hisd(po, seeds):
    """hiosd_lobpcg_sieh_V0 (GOSD.cpp:123-531) with k0 = 0 is hiosd_lobpcg_sieh."""
    N = 12
    o = po.Oracle(N)
    x = (seeds.gaussian(N, 0.05, 11) + 0j).copy()
    V = (seeds.gaussian(N, 1.0, 12) + 0j).reshape((1, N, N, N)).copy()
    x2, V2 = x.copy(), V.copy()
    r1 = o.hiosd_lobpcg_sieh(x, V, maxiter=3, epsf=-1.0)
    r2 = o.hiosd_lobpcg_sieh_V0(x2, V2, np.zeros((0, N, N, N), complex), dt=0.8, betat=8.0, betau=0.02,
                                 maxiter=3, epsf=-1.0)
    assert r1 == r2
    assert np.array_equal(x, x2) and np.array_equal(V, V2)


@pytest.mark.parametrize("k0", [1, 2])
def test_gosd_deflates_the_null_space_dense(po, seeds, k0):
    """With a tiny dt the x-step keeps an even mean-free x (to 1e-12), so the GOSD sweeps
    run at a fixed Hessian: with V0 the k0 lowest eigenvectors of the dense
    mean-free Hessian (N = 6), V converges to eigenvector k0 + 1, and stays
    orthogonal to V0 (GOSD.cpp:206, 227, 347)."""
    N = 6
    R = N ** 3
    y = seeds.gaussian(N, 0.5, 18)  # eigenvalues 1 - 4 about 0.02 apart
    phi = 0.5 * (y + _flip(y))
    phi -= phi.mean()  # the x-step sets the mean to zero (HISD.cpp:260-261)
    o = po.Oracle(N)
    _, G, _ = o.tables()
    E = np.eye(R).reshape(R, N, N, N)
    conv = np.fft.ifftn(G[None] * np.fft.fftn(E, axes=(1, 2, 3)), axes=(1, 2, 3)).real.reshape(R, R)
    A = conv + np.diag((-TAU - 2 * GAMMA * phi + 3 * phi * phi).ravel())
    Q, _ = np.linalg.qr(np.eye(R)[:, 1:] - 1.0 / R)
    ev, Z = np.linalg.eigh(Q.T @ (0.5 * (A + A.T)) @ Q)
    assert ev[k0] - ev[k0 - 1] > 1e-3 and ev[k0 + 1] - ev[k0] > 1e-3  # separated
    vecs = (Q @ Z[:, : k0 + 1]).T * np.sqrt(R)  # unit in the inpfp norm (mean of squares)
    V0 = (vecs[:k0].reshape(k0, N, N, N) + 0j).copy()
    x = (phi + 0j).copy()
    V = (seeds.gaussian(N, 1.0, 14) + 0j).reshape((1, N, N, N)).copy()
    o.hiosd_lobpcg_sieh_V0(x, V, V0, dt=1e-16, betat=8.0, betau=0.02, maxiter=200, epsf=-1.0)
    assert np.abs(x.real - phi).max() < 1e-12
    v = V[0].real
    assert np.abs(np.einsum("aijk,ijk->a", V0.real, v) / R).max() < 1e-12
    assert abs(o.rayleighq(x, V[0]) - ev[k0]) < 1e-10
    assert abs(abs(np.mean(v * vecs[k0].reshape(N, N, N))) - 1.0) < 1e-9


def test_hisd_step_clamp_keeps_a_nan_step(po):
    """HISD.cpp:203-208 clamps bta with std::min(bta, betat dt) then std::max(bta,
    betau dt), which return their first argument when it is NaN. At a critical
    point (x = 0: f = 0, so Dg = 0 and bta = 0/0) the step is NaN, and so is x."""
    N = 8
    o = po.Oracle(N)
    x = np.zeros((N, N, N), complex)
    V = (np.random.default_rng(3).standard_normal((1, N, N, N)) + 0j).copy()
    with np.errstate(all="ignore"):
        o.hiosd_lobpcg_sieh(x, V, maxiter=1, epsf=-1.0)
    assert np.isnan(x.real).all()
