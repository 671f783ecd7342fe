"""Host-side CKKS encoding for the device plaintexts the library builds
itself (bootstrap linear transforms, constants).

Slots follow the canonical embedding with OpenFHE's slot order: slot j of a
polynomial m(X) in Z[X]/(X^N + 1) is m(zeta^(5^j mod 2N)), zeta = exp(i pi / N),
j < N/2.  Encoding inverts this with one length-N FFT: the N values
m(zeta^(2t+1)) (slot values at t = (5^j - 1)/2, conjugates at the mirrored
points) determine m through m_i = (1/N) sum_t v_t zeta^(-(2t+1) i).

Integer coefficients are returned as int64 when they fit, otherwise as Python
integers (object arrays); `to_rns` reduces either form modulo each prime.
"""
from __future__ import annotations

import numpy as np


def slot_points(N: int) -> np.ndarray:
    """t_j with zeta^(2 t_j + 1) = zeta^(5^j mod 2N), j < N/2."""
    n = N // 2
    out = np.empty(n, np.int64)
    g = 1
    for j in range(n):
        out[j] = (g - 1) // 2
        g = g * 5 % (2 * N)
    return out


_POINTS = {}


def _points(N):
    if N not in _POINTS:
        _POINTS[N] = slot_points(N)
    return _POINTS[N]


def encode_real(z, scale: float, N: int) -> np.ndarray:
    """Slots (up to N/2 complex values; fewer are replicated) -> unrounded
    real coefficients of m with m(zeta^(5^j)) = scale * z_j."""
    n = N // 2
    z = np.asarray(z, np.complex128)
    if z.size < n:
        if n % z.size:
            raise ValueError("slot count must divide N/2")
        z = np.tile(z, n // z.size)
    t = _points(N)
    v = np.zeros(N, np.complex128)
    v[t] = z
    v[N - 1 - t] = np.conj(z)       # zeta^(2(N-1-t)+1) = conj(zeta^(2t+1))
    # m_i = (1/N) sum_t v_t zeta^{-(2t+1) i} = zeta^{-i} (1/N) sum_t v_t w^{-t i}, w = zeta^2
    i = np.arange(N)
    m = (np.fft.fft(v) / N) * np.exp(-1j * np.pi * i / N)
    return m.real * scale


def encode(z, scale: float, N: int):
    """Rounded integer coefficients (int64 array, or object array of Python
    ints when some |coefficient| >= 2^62)."""
    m = np.rint(encode_real(z, scale, N))
    if np.max(np.abs(m)) < 2.0 ** 62:
        return m.astype(np.int64)
    return np.array([int(x) for x in m], dtype=object)


def decode(coeffs, scale: float, N: int) -> np.ndarray:
    """Integer (or float) coefficients -> N/2 slot values / scale."""
    m = np.array([float(x) for x in coeffs], np.float64) / scale
    i = np.arange(N)
    v = np.fft.ifft(m * np.exp(1j * np.pi * i / N)) * N
    return v[_points(N)]


def to_rns(coeffs, primes) -> np.ndarray:
    """Coefficients (int64 or Python ints, any sign) -> uint64 residues [len(primes)][N]."""
    c = np.asarray(coeffs)
    out = np.empty((len(primes), c.size), np.uint64)
    if c.dtype == object:
        for j, q in enumerate(primes):
            out[j] = np.array([int(x) % int(q) for x in c], dtype=np.uint64)
    else:
        c = c.astype(np.int64)
        for j, q in enumerate(primes):
            out[j] = np.mod(c, np.int64(q)).astype(np.uint64)
    return out
