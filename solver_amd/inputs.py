"""Seeded synthetic inputs of the adjoint workloads (host side, numpy).

The reference has no perturbed or forced Taylor-Green case, so these generators define them:

* `perturbation(n, seed, rms)`: a random MAC velocity field on the periodic n^3 grid of
  [0, 2pi)^3, discretely divergence-free, with RMS amplitude `rms`. Complex Gaussian Fourier
  modes of amplitude |k|^-2 exp(-|k|^2 / 64) (integer wavenumbers, k = 0 removed) are drawn
  from a splitmix64 stream, transformed to physical space in FP64, projected onto the kernel of
  the MAC divergence exactly in Fourier space (the periodic MAC operators are diagonal there),
  and scaled. Used for the 128^3 initial-state perturbation (amplitude 1e-2) and for the
  finite-difference directions (seed + 1 .. seed + 8, RMS 1).
* `forcing(n, seed, amplitude)`: f = (A sin(kx x) sin(ky y), 0, 0) on the u faces with (kx, ky)
  drawn from {1..4}^2 by splitmix64.

Arrays are in the reference storage layout: shape (sz, sy, sx), one ghost layer, active
entries 1..n (+1 along a face field's normal axis; the periodic duplicate face holds the
same value as face 1), ghosts zero.
"""
from __future__ import annotations

import hashlib
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.fft

GOLDEN_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Outputs offset .. offset+count-1 of the splitmix64 stream started at `seed` (Steele,
    Lea and Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014): the state
    advances by the golden gamma and each state is passed through the variant-13 mixer."""
    k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * GOLDEN_GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Uniform doubles in (0, 1): the top 53 bits of each output, centred in their interval."""
    return ((splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def _normals_into(out: np.ndarray, seed: int, offset: int) -> None:
    u = uniform01(seed, out.size, offset)
    r = np.sqrt(-2.0 * np.log(u[0::2]))
    t = 2.0 * np.pi * u[1::2]
    out[0::2] = r * np.cos(t)
    out[1::2] = r * np.sin(t)


def normals(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Standard normal deviates offset .. offset+count-1 (offset and count even): the Box-Muller
    transform of uniform pair i gives deviates 2i (cosine) and 2i+1 (sine), so any aligned
    sub-range of the stream can be generated on its own; large requests are split into aligned
    chunks generated on several threads (the values do not depend on the split)."""
    if offset % 2 or count % 2:
        raise ValueError("normals are generated in pairs: offset and count must be even")
    out = np.empty(count)
    chunk = 1 << 22
    if count <= chunk:
        _normals_into(out, seed, offset)
        return out
    with ThreadPoolExecutor(_threads()) as ex:
        list(ex.map(lambda b: _normals_into(out[b:b + chunk], seed, offset + b), range(0, count, chunk)))
    return out


def _storage(n, loc):
    s = [n + 2, n + 2, n + 2]  # (z, y, x)
    if loc:
        s[3 - loc] += 1
    return np.zeros(s)


def perturbation(n: int, seed: int, rms: float):
    """Discretely divergence-free random MAC velocity (u, v, w) on the periodic n^3 grid."""
    h = 2.0 * np.pi / n
    kk = np.fft.fftfreq(n, 1.0 / n)  # integer wavenumbers
    # broadcastable wavenumbers along (z, y, x)
    k3 = [kk[None, None, :], kk[None, :, None], kk[:, None, None]]  # x, y, z
    kmag2 = k3[0] ** 2 + k3[1] ** 2 + k3[2] ** 2
    nz = kmag2 > 0
    amp = np.zeros(kmag2.shape)
    amp[nz] = np.exp(-kmag2[nz] / 64.0) / kmag2[nz]
    # component c uses deviates [2 c n^3, 2 (c+1) n^3): real parts, then imaginary parts
    n3 = n * n * n
    coef = []
    for c in range(3):
        g = normals(seed, 2 * n3, offset=2 * n3 * c).reshape(2, n, n, n)
        coef.append((g[0] + 1j * g[1]) * (np.sqrt(0.5) * amp))
        del g
    # MAC operators in Fourier space: D_a = (e^{i th} - 1)/h (faces -> cells), G_a =
    # (1 - e^{-i th})/h (cells -> faces), D.G = sum (2 cos th - 2)/h^2
    th = [2.0 * np.pi * k / n for k in k3]
    D = [(np.exp(1j * t) - 1.0) / h for t in th]
    G = [(1.0 - np.exp(-1j * t)) / h for t in th]
    lap = (2.0 * np.cos(th[0]) - 2.0) / (h * h) + (2.0 * np.cos(th[1]) - 2.0) / (h * h) + \
        (2.0 * np.cos(th[2]) - 2.0) / (h * h)
    phi = D[0] * coef[0] + D[1] * coef[1] + D[2] * coef[2]
    phi[nz] = phi[nz] / lap[nz]
    phi[~nz] = 0.0
    comps = []
    for a in range(3):
        c = coef[a]
        c -= G[a] * phi
        coef[a] = None
        comps.append(np.real(scipy.fft.ifftn(c, workers=_threads())))
        del c
    del phi
    scale = rms / np.sqrt(sum(float(np.sum(c * c)) for c in comps) / (3.0 * n ** 3))
    out = []
    for a in range(3):
        f = _storage(n, a + 1)
        c = comps[a] * scale
        f[1:n + 1, 1:n + 1, 1:n + 1] = c
        if a == 0:
            f[1:n + 1, 1:n + 1, n + 1] = c[:, :, 0]
        elif a == 1:
            f[1:n + 1, n + 1, 1:n + 1] = c[:, 0, :]
        else:
            f[n + 1, 1:n + 1, 1:n + 1] = c[0, :, :]
        out.append(f)
    return out


def forcing_wavenumbers(seed: int):
    z = int(splitmix64(seed, 1)[0])
    return 1 + (z & 3), 1 + ((z >> 32) & 3)


def forcing(n: int, seed: int, amplitude: float):
    """(A sin(kx x) sin(ky y), 0, 0) on the u faces (x = (i-1) h, y = (j - 1/2) h)."""
    kx, ky = forcing_wavenumbers(seed)
    h = 2.0 * np.pi / n
    fu = _storage(n, 1)
    x = (np.arange(1, n + 2) - 1) * h
    y = (np.arange(1, n + 1) - 0.5) * h
    fu[1:n + 1, 1:n + 1, 1:n + 2] = amplitude * np.sin(ky * y)[None, :, None] * np.sin(kx * x)[None, None, :]
    return [fu, _storage(n, 2), _storage(n, 3)]


def sha256_fields(fields) -> str:
    h = hashlib.sha256()
    for f in fields:
        h.update(np.ascontiguousarray(f, dtype="<f8").tobytes())
    return h.hexdigest()


def mac_divergence(fields, n):
    """Discrete MAC divergence on the active cells (host check of the generator)."""
    h = 2.0 * np.pi / n
    u, v, w = fields
    s = slice(1, n + 1)
    return ((u[s, s, 2:n + 2] - u[s, s, 1:n + 1]) + (v[s, 2:n + 2, s] - v[s, 1:n + 1, s]) +
            (w[2:n + 2, s, s] - w[1:n + 1, s, s])) / h
