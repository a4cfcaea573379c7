"""Seeded synthetic inputs for the benchmark configurations (numpy, host).

All randomness comes from one SplitMix64 stream per seed, so the GPU path and
the CPU oracle receive identical inputs.

  membranes(...)   ellipses with tether springs to a circle of radius sqrt(ab)
  fourier_velocity(...)  u = curl(psi), psi = sum_{0<|k|<=kmax} c_k |k|^-3
                   cos(2 pi k.x + phi_k): exactly divergence-free on the MAC
                   grid (discrete curl), velocity ~ |k|^-2, normalised to RMS 1
  fourier_rhs(...) Poisson right-hand side: sum of modes with |k|^-2 decay
"""
import math

import numpy as np


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))

    def normal(self):
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def ellipse(cx, cy, a, b, npts, k=1.0):
    """One membrane: markers X on the ellipse, rest positions R on the circle
    of radius sqrt(ab) (equal area), stiffness k, arclength weight ds =
    2 pi r0 / npts (the convention of test/ib_test.cc:55)."""
    th = 2.0 * np.pi * np.arange(npts) / npts
    r0 = math.sqrt(a * b)
    X = np.stack([cx + a * np.cos(th), cy + b * np.sin(th)], axis=1)
    R = np.stack([cx + r0 * np.cos(th), cy + r0 * np.sin(th)], axis=1)
    ds = np.full(npts, 2.0 * np.pi * r0 / npts)
    return X, R, np.full(npts, k), ds


def membranes(count, npts, seed, center=(0.1, 0.9), axes=(0.02, 0.05), k=1.0):
    rng = SplitMix64(seed)
    Xs, Rs, ks, dss = [], [], [], []
    for _ in range(count):
        cx, cy = rng.uniform(*center), rng.uniform(*center)
        a, b = rng.uniform(*axes), rng.uniform(*axes)
        X, R, kk, ds = ellipse(cx, cy, a, b, npts, k)
        Xs.append(X); Rs.append(R); ks.append(kk); dss.append(ds)
    return np.concatenate(Xs), np.concatenate(Rs), np.concatenate(ks), np.concatenate(dss)


def single_ellipse(npts=512):
    """The CPU reference case: centre (0.5, 0.5), semi-axes 0.20 / 0.15,
    tethered to the circle of radius 0.1732."""
    X, R, k, ds = ellipse(0.5, 0.5, 0.20, 0.15, npts)
    r0 = 0.1732
    th = 2.0 * np.pi * np.arange(npts) / npts
    R = np.stack([0.5 + r0 * np.cos(th), 0.5 + r0 * np.sin(th)], axis=1)
    ds = np.full(npts, 2.0 * np.pi * r0 / npts)
    return X, R, k, ds


def fourier_velocity(n, kmax, seed, rms=1.0):
    """Divergence-free MAC velocity: u = d psi/dy, v = -d psi/dx as discrete
    differences of psi sampled at cell corners, so div u = 0 to round-off."""
    rng = SplitMix64(seed)
    x = np.arange(n) / n
    X, Y = np.meshgrid(x, x, indexing="xy")     # corners (i/n, j/n), shape (ny, nx)
    psi = np.zeros((n, n))
    for kx in range(-kmax, kmax + 1):
        for ky in range(-kmax, kmax + 1):
            kk = math.hypot(kx, ky)
            if kk == 0 or kk > kmax:
                continue
            c = rng.normal() * kk ** -3
            ph = rng.uniform(0.0, 2.0 * np.pi)
            psi += c * np.cos(2.0 * np.pi * (kx * X + ky * Y) + ph)
    # u at x-face (i, j+1/2): (psi[i, j+1] - psi[i, j]) n ; v at y-face (i+1/2, j): -(psi[i+1, j] - psi[i, j]) n
    u = (np.roll(psi, -1, axis=0) - psi) * n
    v = -(np.roll(psi, -1, axis=1) - psi) * n
    s = math.sqrt((np.mean(u * u) + np.mean(v * v)) / 2.0)
    return u * (rms / s), v * (rms / s)


def fourier_rhs(n, modes, kmax, seed):
    """Sum of `modes` random Fourier modes with N(0,1) |k|^-2 amplitudes,
    cell-centred, mean removed."""
    rng = SplitMix64(seed)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="xy")
    f = np.zeros((n, n))
    for _ in range(modes):
        while True:
            kx = int(rng.next() % (2 * kmax + 1)) - kmax
            ky = int(rng.next() % (2 * kmax + 1)) - kmax
            if kx or ky:
                break
        a = rng.normal() * math.hypot(kx, ky) ** -2
        ph = rng.uniform(0.0, 2.0 * np.pi)
        f += a * np.cos(2.0 * np.pi * (kx * X + ky * Y) + ph)
    return f - f.mean()


def fourier_modes(modes, kmax, seed):
    """The (kx, ky, amplitude, phase) lists of fourier_rhs, in its order."""
    rng = SplitMix64(seed)
    kxs, kys, amps, phs = [], [], [], []
    for _ in range(modes):
        while True:
            kx = int(rng.next() % (2 * kmax + 1)) - kmax
            ky = int(rng.next() % (2 * kmax + 1)) - kmax
            if kx or ky:
                break
        amps.append(rng.normal() * math.hypot(kx, ky) ** -2)
        phs.append(rng.uniform(0.0, 2.0 * np.pi))
        kxs.append(kx)
        kys.append(ky)
    return kxs, kys, amps, phs


def mode_sum_device(n, kxs, kys, amps, phs, device, offset=0.5):
    """sum_m a_m cos(2 pi (kx_m x + ky_m y) + ph_m) at x = (i + offset)/n,
    y = (j + offset)/n on the GPU (torch float64), as [y, x]: each cos(a + b)
    is expanded into two outer products, so the field is one matrix product."""
    import torch
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=device)
    x = (torch.arange(n, dtype=torch.float64, device=device) + offset) / n
    ax = 2.0 * np.pi * x[:, None] * t(kxs)[None, :] + t(phs)[None, :]      # n x modes
    by = 2.0 * np.pi * x[:, None] * t(kys)[None, :]
    a = t(amps)
    return (torch.cos(by) * a) @ torch.cos(ax).T - (torch.sin(by) * a) @ torch.sin(ax).T


def fourier_rhs_device(n, modes, kmax, seed, device):
    """fourier_rhs built on the GPU: the same modes from the same SplitMix64
    stream (to rounding of the sums)."""
    f = mode_sum_device(n, *fourier_modes(modes, kmax, seed), device)
    return f - f.mean()


def fourier_velocity_device(n, kmax, seed, device, rms=1.0):
    """fourier_velocity built on the GPU: the same psi (same modes, same
    stream order) sampled at the cell corners by mode_sum_device, then the
    same discrete curl and normalisation."""
    import torch
    rng = SplitMix64(seed)
    kxs, kys, amps, phs = [], [], [], []
    for kx in range(-kmax, kmax + 1):
        for ky in range(-kmax, kmax + 1):
            kk = math.hypot(kx, ky)
            if kk == 0 or kk > kmax:
                continue
            amps.append(rng.normal() * kk ** -3)
            phs.append(rng.uniform(0.0, 2.0 * np.pi))
            kxs.append(kx)
            kys.append(ky)
    psi = mode_sum_device(n, kxs, kys, amps, phs, device, offset=0.0)
    u = (torch.roll(psi, -1, 0) - psi) * n
    v = -(torch.roll(psi, -1, 1) - psi) * n
    s = math.sqrt((float(torch.mean(u * u)) + float(torch.mean(v * v))) / 2.0)
    return u * (rms / s), v * (rms / s)


# The benchmark and parity configurations (sizes and physics) used by bench.py
# and the tests.
CONFIGS = {
    "c1": dict(n=128, membranes=None, mu=0.01, rho=1.0, dt=1e-4, steps=200, tol=1e-10),
    "c2": dict(n=2048, membranes=(64, 256), seed=2, kmax=8, mu=1e-3, rho=1.0, dt=2e-5, steps=100, tol=1e-10),
}


def clustered_structures(n_membranes=512, n_filaments=512, npts=256, blobs=8, sigma=0.03, seed=5):
    """The combined stress case: closed membranes and open filaments whose
    centres are drawn from `blobs` Gaussian clusters (sigma), clipped to
    [0.15, 0.85]^2 so every marker stays > 3h from the periodic seams.
    Filaments are straight segments of length U[0.05, 0.1] at a random angle,
    tethered to their initial positions; ds = length / npts."""
    rng = SplitMix64(seed)
    cb = [(rng.uniform(0.3, 0.7), rng.uniform(0.3, 0.7)) for _ in range(blobs)]

    def centre():
        b = cb[int(rng.next() % blobs)]
        return (min(0.85, max(0.15, b[0] + sigma * rng.normal())), min(0.85, max(0.15, b[1] + sigma * rng.normal())))

    Xs, Rs, ks, dss = [], [], [], []
    for _ in range(n_membranes):
        cx, cy = centre()
        a, b = rng.uniform(0.02, 0.05), rng.uniform(0.02, 0.05)
        X, R, k, ds = ellipse(cx, cy, a, b, npts)
        Xs.append(X); Rs.append(R); ks.append(k); dss.append(ds)
    for _ in range(n_filaments):
        cx, cy = centre()
        L, th = rng.uniform(0.05, 0.1), rng.uniform(0.0, np.pi)
        s = (np.arange(npts) / (npts - 1) - 0.5) * L
        X = np.stack([cx + s * np.cos(th), cy + s * np.sin(th)], axis=1)
        Xs.append(X); Rs.append(X.copy()); ks.append(np.ones(npts)); dss.append(np.full(npts, L / npts))
    return np.concatenate(Xs), np.concatenate(Rs), np.concatenate(ks), np.concatenate(dss)


def config2_inputs(n=2048, count=64, npts=256, seed=2, kmax=8):
    X, R, k, ds = membranes(count, npts, seed)
    u, v = fourier_velocity(n, kmax, seed + 1)
    return u, v, X, R, k, ds
