"""Closed-form pins of the oracle's SPM force steps, HiSD x-step and LOBPCG
start, and proof that each pin catches the mistake it is there for.

Each pin compares one oracle routine with what the mathematics fixes on an
input where that is known exactly:
  * SPM.cpp:34-168 on phi = A cos(t), t = 2 pi n i / N, with the unit tangent
    v = cos(t) + cos(s), s = 2 pi m j / N: the force step changes the Fourier
    coefficients at +-n, +-2n, +-3n (axis 0) and +-m (axis 1) only, to values
    written below from the single-mode expansions of ngrad and of the
    nonlinearity (the coefficient of v's s-part is the tangent term alone);
  * HISD.cpp:188-268, one step from the same input with V = v: the first
    Barzilai-Borwein step is 1 and the force is reflected by 2 <f, v> v;
  * HISD.cpp:199-212, the second step from an even field with an odd V (no
    reflection): the textbook BB step |<Dx, Dg>| / <Dg, Dg> of the first step's
    displacement and force change (Barzilai & Borwein 1988), clamped to
    [betau dt, betat dt], then a gradient-flow step of that size;
  * LOBPCG.cpp:10-109: the start block is the Q factor of the mean-removed
    random block with a positive R diagonal (in-order Gram-Schmidt).

The tests below then rebuild the oracle with each planted mistake
of lp_oracle.c (LPO_PLANT) and check that the pin for that term fails.
"""
import numpy as np
import pytest

L, TAU, GAMMA = 26.0, -0.0061, 0.30


def _cos(N, n, axis):
    c = np.cos(2 * np.pi * n * np.arange(N) / N)
    shape = [1, 1, 1]
    shape[axis] = N
    return np.broadcast_to(c.reshape(shape), (N, N, N)).copy()


def _spec(x):
    return np.fft.fftn(x) / x.size


def _check_coefficients(out, want, tol):
    """out: real field; want: {(i, j, k): value} for the nonzero coefficients
    (each listed index and its negative)."""
    N = out.shape[0]
    S = _spec(out)
    rest = S.copy()
    for idx, val in want.items():
        for sgn in (1, -1):
            m = tuple((sgn * a) % N for a in idx)
            assert abs(S[m] - val) < tol, (idx, S[m], val)
            rest[m] = 0
    assert np.abs(rest).max() < tol, np.abs(rest).max()


def _single_mode_setup(po, make, N=16, n=2, m=3, A=0.2):
    o = make(N)
    G = lambda q: po.get_GK((q / L) ** 2)
    phi = A * _cos(N, n, 0)
    v = _cos(N, n, 0) + _cos(N, m, 1)  # unit in the inpfp norm: mean(v^2) = 1
    a1 = -(G(n) - TAU + 0.75 * A * A) * A  # cos(t) coefficient of ngrad(phi)
    p = a1 / 2  # <ngrad(phi), v> = mean(f v)
    return o, G, phi, v, a1, p


def pin_spm_force(po, make, minmax):
    """x_hat' = (x_hat + beta nln_hat - beta c <f,v> v_hat) / (1 + beta G), DC = 0,
    beta = 1, c = 1 (Natural_Forces) or 2 (MinMax_Forces)."""
    N, n, m, A = 16, 2, 3, 0.2
    o, G, phi, v, a1, p = _single_mode_setup(po, make, N, n, m, A)
    c = 2.0 if minmax else 1.0
    x = (phi + 0j).copy()
    (o.MinMax_Forces if minmax else o.Natural_Forces)(x, v + 0j)
    want = {
        (n, 0, 0): (A / 2 + (TAU * A - 0.75 * A ** 3) / 2 - c * p / 2) / (1 + G(n)),
        (0, m, 0): (-c * p / 2) / (1 + G(m)),
        (2 * n, 0, 0): (GAMMA * A * A / 4) / (1 + G(2 * n)),
        (3 * n, 0, 0): (-(A ** 3) / 8) / (1 + G(3 * n)),
    }
    _check_coefficients(x.real, want, 1e-15)


def pin_hisd_reflection(po, make):
    """First HiSD step with V = v: g = f - 2 <f,v> v, step 1 (Dx = g + dt, mean g = 0),
    x_hat' = (x_hat + nln_hat - 2 <f,v> v_hat) / (1 + G), DC = 0."""
    N, n, m, A = 16, 2, 3, 0.2
    o, G, phi, v, a1, p = _single_mode_setup(po, make, N, n, m, A)
    x = (phi + 0j).copy()
    V = (v + 0j).reshape((1, N, N, N)).copy()
    o.hiosd_lobpcg_sieh(x, V, dt=0.8, betat=8.0, betau=0.02, maxiter=1, epsf=-1.0)
    want = {
        (n, 0, 0): (A / 2 + (TAU * A - 0.75 * A ** 3) / 2 - 2 * p / 2) / (1 + G(n)),
        (0, m, 0): (-2 * p / 2) / (1 + G(m)),
        (2 * n, 0, 0): (GAMMA * A * A / 4) / (1 + G(2 * n)),
        (3 * n, 0, 0): (-(A ** 3) / 8) / (1 + G(3 * n)),
    }
    _check_coefficients(x.real, want, 1e-15)


def _flip(a):
    return np.roll(a[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))


def _even_odd(seeds, N):
    y = seeds.gaussian(N, 0.05, 21)
    x0 = 0.5 * (y + _flip(y))
    z = seeds.gaussian(N, 1.0, 22)
    v = 0.5 * (z - _flip(z))
    return x0, v / np.sqrt(np.mean(v * v))


def _gf_step(po, N, x, dt):
    """One get_min_u_F step of size dt (pinned coefficient by coefficient in
    test_oracle.py)."""
    o = po.Oracle(N, dt=dt)
    uR = (x + 0j).copy()
    uF = o.my_fft(uR)
    o.get_min_u_F(uR, uF, 1, -1.0)
    return uR.real.copy()


def _bb_reference(po, seeds, N, dt, betat, betau):
    x0, _ = _even_odd(seeds, N)
    clamp = lambda b: max(min(b, betat * dt), betau * dt)
    o = po.Oracle(N)
    x1 = _gf_step(po, N, x0, clamp(1.0))  # step 1: Dx = f + dt, Dg = f, so BB = 1
    f0 = o.ngrad_cam(x0 + 0j).real
    f1 = o.ngrad_cam(x1 + 0j).real
    dx, dg = x1 - x0, f1 - f0
    bb = abs(np.mean(dx * dg) / np.mean(dg * dg))
    return x1, bb, clamp(bb), _gf_step(po, N, x1, clamp(bb))


def pin_hisd_bb_step(po, make, seeds, clamped):
    """Two HiSD steps from an even field with an odd V (the reflection vanishes,
    V stays odd through its LOBPCG sweeps): gradient-flow steps of size 1, then
    the clamped BB step of the first step's displacement and force change."""
    N, dt, betau = 12, 0.8, 0.02
    _, bb, _, _ = _bb_reference(po, seeds, N, dt, 1e3, betau)
    betat = 0.5 * bb / dt if clamped else 1e3  # the upper clamp active or not
    x1, bb, bta2, x2 = _bb_reference(po, seeds, N, dt, betat, betau)
    assert (bta2 < bb) == clamped
    x0, v = _even_odd(seeds, N)
    o = make(N)
    x = (x0 + 0j).copy()
    V = (v + 0j).reshape((1, N, N, N)).copy()
    o.hiosd_lobpcg_sieh(x, V, dt=dt, betat=betat, betau=betau, maxiter=2, epsf=-1.0)
    assert np.abs(x.real - x2).max() < 1e-13 * np.abs(x2).max()


def pin_lobpcg_init(po, make):
    """V = Q sqrt(R): Q R_ = the mean-removed start block (fields as columns),
    diag(R_) > 0, so V_i is field i orthonormalised against fields 0..i-1."""
    N, k = 8, 5
    R = N ** 3
    start = np.random.default_rng(3).uniform(-1, 1, (k, N, N, N))
    V = make(N).lobpcg_initialization(start).real.reshape(k, R)
    S = start.reshape(k, R)
    S = S - S.mean(axis=1, keepdims=True)
    Q, Rm = np.linalg.qr(S.T)
    Q = Q * np.sign(np.diag(Rm))
    assert np.abs(V - Q.T * np.sqrt(R)).max() < 1e-12


PINS = {
    "natural_force": lambda po, make, seeds: pin_spm_force(po, make, minmax=False),
    "minmax_force": lambda po, make, seeds: pin_spm_force(po, make, minmax=True),
    "hisd_reflection": lambda po, make, seeds: pin_hisd_reflection(po, make),
    "hisd_bb_step": lambda po, make, seeds: pin_hisd_bb_step(po, make, seeds, clamped=False),
    "hisd_bb_step_clamped": lambda po, make, seeds: pin_hisd_bb_step(po, make, seeds, clamped=True),
    "lobpcg_init": lambda po, make, seeds: pin_lobpcg_init(po, make),
}

# planted mistake (LPO_PLANT in lp_oracle.c) -> the pins that must catch it
PLANTS = {
    1: ("natural_force", "minmax_force"),  # SPM beta 0.5
    2: ("minmax_force",),  # MinMax coefficient 1
    3: ("natural_force",),  # Natural coefficient 2
    4: ("hisd_reflection",),  # reflection factor 1
    5: ("hisd_reflection", "hisd_bb_step"),  # BB denominator <Dx, Dx>
    6: ("hisd_bb_step_clamped",),  # no upper clamp
    7: ("lobpcg_init",),  # Gram-Schmidt last to first
    8: ("lobpcg_init",),  # means kept
    9: ("natural_force", "minmax_force"),  # tangent term added
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_oracle_pin(po, seeds, name):
    PINS[name](po, lambda N: po.Oracle(N), seeds)


@pytest.fixture(scope="module")
def planted(po, tmp_path_factory):
    d = tmp_path_factory.mktemp("planted")
    return {n: po.build_planted(n, str(d)) for n in PLANTS}


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_pin_catches_planted_mistake(po, seeds, planted, plant):
    for name in PLANTS[plant]:
        make = lambda N, p=planted[plant]: po.Oracle(N, libpath=p)
        with pytest.raises(AssertionError):
            PINS[name](po, make, seeds)
