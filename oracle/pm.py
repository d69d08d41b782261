"""FP64 particle-mesh long-range force (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module, as the checker of the PM kernels in
blitzar_amd/csrc/k_pm.hip. The reference's periodic PM (engine/physics/
treepm/cuda/fft/TpmFft.inl, TpmField.inl) pairs with no short-range walk
(its periodic mode has r_s = 0, TpmCosmology.inl:52), so this is the
definition the GPU mesh is held to, restated from the standard TreePM split:

  * cloud-in-cell deposit of mass onto an N^3 mesh of cell h = L / N, cell
    centres at (i + 1/2) h, periodic;
  * density FFT; long-range potential
        phi_k = -4 pi G rho_k exp(-k^2 r_s^2) / k^2 / W(k)^2,
    W(k) = prod_axes sinc^2(k_a h / 2) (one CIC window for the deposit and
    one for the interpolation), phi_0 = 0;
  * acceleration a_k = -i k phi_k (spectral gradient), inverse FFT;
  * cloud-in-cell interpolation back to the particles.

The exp(-k^2 r_s^2) filter is the complement of the short-range weight
erfc(r / 2 r_s) + r / (r_s sqrt(pi)) exp(-r^2 / 4 r_s^2) used by the walk.

Pinned in tests/test_oracle_pm.py: the long-range force of a point mass
matches m (1 - w(r)) / r^2 between a few r_s and a tenth of the box; total
force on the mesh is zero (momentum); a uniform lattice feels no force;
translation by whole cells shifts the field exactly.
"""
from __future__ import annotations

import numpy as np


def _cic(pos, L, N):
    u = np.asarray(pos, np.float64) / L * N - 0.5
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    return i0, f


def deposit(pos, mass, L, N):
    rho = np.zeros(N * N * N)
    i0, f = _cic(pos, L, N)
    mass = np.asarray(mass, np.float64)
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1 - f[:, 0]
        ix = (i0[:, 0] + dx) % N
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1 - f[:, 1]
            iy = (i0[:, 1] + dy) % N
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1 - f[:, 2]
                iz = (i0[:, 2] + dz) % N
                # summed per cell in FP64 (bincount adds the weights of equal indices)
                rho += np.bincount((ix * N + iy) * N + iz, weights=mass * wx * wy * wz,
                                   minlength=N * N * N)
    return rho.reshape(N, N, N) / (L / N) ** 3


def interpolate(field, pos, L, N):
    i0, f = _cic(pos, L, N)
    out = np.zeros(len(pos))
    for dx in (0, 1):
        wx = f[:, 0] if dx else 1 - f[:, 0]
        for dy in (0, 1):
            wy = f[:, 1] if dy else 1 - f[:, 1]
            for dz in (0, 1):
                wz = f[:, 2] if dz else 1 - f[:, 2]
                out += field[(i0[:, 0] + dx) % N, (i0[:, 1] + dy) % N,
                             (i0[:, 2] + dz) % N] * wx * wy * wz
    return out


def wavenumbers(L, N):
    k = 2 * np.pi / L * np.fft.fftfreq(N, d=1.0 / N)
    kz = 2 * np.pi / L * np.arange(N // 2 + 1)
    return k, kz


def long_range_accel(pos, mass, L, N, rs, G=1.0, targets=None):
    """Long-range acceleration at every particle, or at pos[targets] only
    (the deposit always takes every particle)."""
    rho_k = np.fft.rfftn(deposit(pos, mass, L, N))
    k, kz = wavenumbers(L, N)
    KX, KY, KZ = np.meshgrid(k, k, kz, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    h = L / N
    W = (np.sinc(KX * h / (2 * np.pi)) * np.sinc(KY * h / (2 * np.pi)) *
         np.sinc(KZ * h / (2 * np.pi))) ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        phi_k = -4 * np.pi * G * rho_k * np.exp(-k2 * rs * rs) / k2 / (W * W)
    phi_k[0, 0, 0] = 0.0
    del rho_k, W, k2
    at = pos if targets is None else np.asarray(pos)[targets]
    acc = np.zeros((len(at), 3))
    for a, K in enumerate((KX, KY, KZ)):
        field = np.fft.irfftn(-1j * K * phi_k, s=(N, N, N), axes=(0, 1, 2))
        acc[:, a] = interpolate(field, at, L, N)
    return acc


def short_range_weight(r, rs):
    from math import erfc, exp, pi, sqrt
    a = 0.5 * r / rs
    return erfc(a) + r / (rs * sqrt(pi)) * exp(-a * a)
