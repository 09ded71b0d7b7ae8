"""Seeded initial fields of the benchmark workloads (host, numpy).

The reference seeds from Fourier-mode files (initData/*, read by
get_initial_value, Initialization.cpp:242-338); the workloads here are seeded
in real space instead, so every field is real and Hermitian by construction.
A Fourier mode n = (n0, n1, n2) on the N^3 grid is the plane wave
cos(2 pi (n . idx) / N + theta); with the reciprocal metric k = kscale * n / L a
wave vector q of length |q| maps to n = round(q * L / kscale).
"""
from __future__ import annotations

import itertools

import numpy as np

GOLDEN = (1.0 + 5.0 ** 0.5) / 2.0


def gaussian(N: int, sigma: float = 0.05, seed: int = 0) -> np.ndarray:
    """Zero-mean i.i.d. Gaussian field (N^3 float64)."""
    x = np.random.default_rng(seed).standard_normal((N, N, N)) * sigma
    return x - x.mean()


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def icosahedral_star() -> np.ndarray:
    """The 30 two-fold directions of the icosahedron (edge midpoints /
    icosidodecahedron vertices), unit length: (+-1,0,0) and permutations, and
    the cyclic permutations of (+-1/2, +-phi/2, +-1/(2 phi))."""
    vs = []
    for a in range(3):
        for s in (1.0, -1.0):
            v = [0.0, 0.0, 0.0]
            v[a] = s
            vs.append(v)
    base = (0.5, GOLDEN / 2.0, 1.0 / (2.0 * GOLDEN))
    for s0, s1, s2 in itertools.product((1, -1), repeat=3):
        b = (s0 * base[0], s1 * base[1], s2 * base[2])
        for r in range(3):
            vs.append([b[(0 - r) % 3], b[(1 - r) % 3], b[(2 - r) % 3]])
    v = _unit(vs)
    assert v.shape == (30, 3)
    return v


def bcc_110() -> np.ndarray:
    vs = {tuple(p) for p in itertools.permutations((1, 1, 0))} | {tuple(p) for p in itertools.permutations((1, -1, 0))}
    vs |= {tuple(-x for x in p) for p in vs}
    return _unit(sorted(vs))


def fcc_111_200():
    q111 = _unit([p for p in itertools.product((1, -1), repeat=3)])
    q200 = _unit([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]) * (2.0 / np.sqrt(3.0))
    return np.vstack([q111, q200])


def lamellar_z() -> np.ndarray:
    return np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])


STARS = {"bcc": bcc_110, "fcc": fcc_111_200, "lam": lamellar_z, "ico": icosahedral_star}


def grid_modes(q: np.ndarray, L: float, kscale: float) -> np.ndarray:
    """Round wave vectors |q| (physical units) to integer grid modes."""
    return np.rint(np.asarray(q) * L / kscale).astype(np.int64)


def plane_waves(N: int, modes: np.ndarray, amp: float, seed: int) -> np.ndarray:
    """sum_q amp * cos(2 pi n_q . idx / N + theta_q), theta_q ~ U[0, 2 pi)."""
    theta = np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, size=len(modes))
    i = np.arange(N)
    out = np.zeros((N, N, N))
    for (a, b, c), th in zip(modes, theta):
        ph = 2.0 * np.pi * ((a * i)[:, None, None] + (b * i)[None, :, None] + (c * i)[None, None, :]) / N
        out += amp * np.cos(ph + th)
    return out


def sphere_mask(N: int, L: float, radius: float = 6.0, width: float = 1.0) -> np.ndarray:
    """0.5 * (1 - tanh((|r - c| - radius) / width)), r = idx * L / N, c = L/2."""
    r = np.arange(N) * L / N - L / 2.0
    d = np.sqrt(r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2)
    return 0.5 * (1.0 - np.tanh((d - radius) / width))


def icosahedral_nucleus(N: int, L: float = 26.0, kscale: float = 1.0, amp: float = 0.1, seed: int = 0,
                        radius: float = 6.0, width: float = 1.0) -> np.ndarray:
    """Icosahedral 30-vector star (|q| = 1) with seeded phases inside a sphere
    of radius `radius` with a tanh edge, in a phi = 0 background."""
    modes = grid_modes(icosahedral_star(), L, kscale)
    return sphere_mask(N, L, radius, width) * plane_waves(N, modes, amp, seed)


def crystal(N: int, kind: str, L: float = 26.0, kscale: float = 1.0, amp: float = 0.1, seed: int = 0) -> np.ndarray:
    """Periodic seed from one of the stars in STARS (amplitude amp, seeded phases)."""
    modes = grid_modes(STARS[kind](), L, kscale)
    return plane_waves(N, modes, amp, seed)


def device_field(torch, N: int, kind: str, L: float = 26.0, kscale: float = 1.0, amp: float = 0.1, seed: int = 0,
                 dtype=None, radius: float = 6.0, width: float = 1.0):
    """The same seeds evaluated on the GPU in FP64 (phases drawn on the host
    exactly as plane_waves draws them): kind is a key of STARS (periodic
    crystal) or "nucleus" (the icosahedral star inside the tanh sphere).
    Used where seeding must not hold the host (serving, batched sweeps)."""
    star = icosahedral_star() if kind == "nucleus" else STARS[kind]()
    modes = grid_modes(star, L, kscale)
    theta = np.random.default_rng(seed).uniform(0.0, 2.0 * np.pi, size=len(modes))
    i = torch.arange(N, dtype=torch.float64, device="cuda") * (2.0 * np.pi / N)
    out = torch.zeros((N, N, N), dtype=torch.float64, device="cuda")
    for (a, b, c), th in zip(modes, theta):
        out += amp * torch.cos((a * i)[:, None, None] + (b * i)[None, :, None] + (c * i)[None, None, :] + float(th))
    if kind == "nucleus":
        r = torch.arange(N, dtype=torch.float64, device="cuda") * (L / N) - L / 2.0
        d = torch.sqrt(r[:, None, None] ** 2 + r[None, :, None] ** 2 + r[None, None, :] ** 2)
        out *= 0.5 * (1.0 - torch.tanh((d - radius) / width))
    return out if dtype is None else out.to(dtype)
