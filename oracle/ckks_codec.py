"""CKKS encode/decode, key generation, encryption and decryption on the CPU.

TEST INFRASTRUCTURE ONLY (imported by tests/ and __graft_entry__.smoke()).
It plays the role OpenFHE plays for the reference: it makes keys and
ciphertexts that both the HIP library and the oracle consume, and it
decrypts results to report precision against an FP64 plaintext computation.

Sampling uses numpy's PCG64 generator from an explicit seed:
  - secret: uniform ternary {-1, 0, 1} (OpenFHE UNIFORM_TERNARY), or sparse
    ternary of Hamming weight h;
  - errors: rounded normal with sigma = 3.19 (OpenFHE default DGG sigma);
  - uniform a: integers uniform in [0, q_i) per limb.
OpenFHE's own PRNG streams cannot be reproduced without OpenFHE, so keys and
encryptions are this file's own (documented) draws.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O

SIGMA = 3.19


# ----------------------------------------------------------------- sampling
def sample_ternary(rng: np.random.Generator, N: int, h: int | None = None) -> np.ndarray:
    if h is None:
        return rng.integers(-1, 2, size=N, dtype=np.int64)
    s = np.zeros(N, np.int64)
    idx = rng.choice(N, size=h, replace=False)
    s[idx] = rng.choice(np.array([-1, 1], np.int64), size=h)
    return s


def sample_error(rng: np.random.Generator, N: int) -> np.ndarray:
    return np.rint(rng.normal(0.0, SIGMA, size=N)).astype(np.int64)


def sample_uniform(rng: np.random.Generator, primes, N: int) -> np.ndarray:
    out = np.empty((len(primes), N), np.uint64)
    for i, q in enumerate(primes):
        out[i] = rng.integers(0, int(q), size=N, dtype=np.uint64)
    return out


def to_rns(coeffs: np.ndarray, primes) -> np.ndarray:
    """Signed integer coefficients (int64) -> residues [len(primes)][N]."""
    c = np.asarray(coeffs, dtype=np.int64)
    out = np.empty((len(primes), c.size), np.uint64)
    for i, q in enumerate(primes):
        assert int(q) < (1 << 63)
        out[i] = np.mod(c, np.int64(q)).astype(np.uint64)
    return out


def to_rns_big(coeffs, primes) -> np.ndarray:
    """Python-int coefficients -> residues (for values beyond int64)."""
    out = np.empty((len(primes), len(coeffs)), np.uint64)
    for i, q in enumerate(primes):
        out[i] = np.array([int(x) % q for x in coeffs], dtype=np.uint64)
    return out


def crt_centered(limbs: np.ndarray, primes) -> list:
    """Residues [l][N] (coefficient form) -> centred Python ints."""
    Q = 1
    for q in primes:
        Q *= int(q)
    terms = []
    for i, q in enumerate(primes):
        qh = Q // int(q)
        terms.append((qh * pow(qh % int(q), -1, int(q))) % Q)
    N = limbs.shape[1]
    res = []
    cols = [limbs[i].tolist() for i in range(len(primes))]
    for k in range(N):
        x = 0
        for i in range(len(primes)):
            x += cols[i][k] * terms[i]
        x %= Q
        if x > Q // 2:
            x -= Q
        res.append(x)
    return res


# -------------------------------------------------------------- encoding
def _slot_index(N: int) -> np.ndarray:
    """t_j with 2 t_j + 1 = 5^j mod 2N, for j < N/2 (canonical embedding)."""
    M = 2 * N
    idx = np.empty(N // 2, np.int64)
    g = 1
    for j in range(N // 2):
        idx[j] = (g - 1) // 2
        g = (g * 5) % M
    return idx


def encode(z: np.ndarray, scale: float, N: int) -> np.ndarray:
    """Slots (N/2 complex) -> rounded integer coefficients (int64) of
    m(X) with m(zeta^(5^j)) ~= scale * z_j, zeta = exp(i pi / N).  Past
    2^62 the coefficients come back as Python ints (object array)."""
    z = np.asarray(z, np.complex128)
    n = N // 2
    full = np.zeros(n, np.complex128)
    full[: z.size] = z
    if z.size < n:  # sparse packing: replicate (as OpenFHE does for slots | N/2)
        rep = n // z.size
        full = np.tile(z, rep)
    v = np.zeros(N, np.complex128)
    t = _slot_index(N)
    v[t] = full
    v[N - 1 - t] = np.conj(full)  # zeta^(2N - (2t+1)) = conj
    # m_i zeta^i = (1/N) sum_t v_t omega^{-it} -> numpy fft(v)/N
    w = np.fft.fft(v) / N
    i = np.arange(N)
    zeta_inv = np.exp(-1j * np.pi * i / N)
    m = np.rint((w * zeta_inv).real * scale)
    if np.max(np.abs(m)) >= 2.0 ** 62:  # big scaling factors (FLEXIBLEAUTOEXT top level)
        return np.array([int(x) for x in m], dtype=object)
    return m.astype(np.int64)


def decode(coeffs, scale: float, N: int, slots: int | None = None) -> np.ndarray:
    m = np.array([float(x) for x in coeffs], np.float64) / scale
    i = np.arange(N)
    zeta = np.exp(1j * np.pi * i / N)
    v = np.fft.ifft(m * zeta) * N  # m(zeta^(2t+1)) for all t
    t = _slot_index(N)
    out = v[t]
    if slots is not None:
        out = out[:slots]
    return out


# ------------------------------------------------------------------ keys
class Keys:
    """Secret key plus evaluation keys for an oracle Context."""

    def __init__(self, ctx: O.Context, seed: int, h: int | None = None):
        self.ctx = ctx
        self.rng = np.random.default_rng(seed)
        N = ctx.N
        self.s_coef = sample_ternary(self.rng, N, h)
        allp = ctx.primes
        pids = np.arange(len(allp), dtype=np.uint32)
        self.s_eval = ctx.ntt_limbs(to_rns(self.s_coef, allp), pids)  # [L+K][N]

    def _switch_key(self, s_new: np.ndarray, s_old: np.ndarray) -> np.ndarray:
        ctx = self.ctx
        LK = ctx.L + ctx.K
        a = np.stack([sample_uniform(self.rng, ctx.primes, ctx.N) for _ in range(ctx.dnum)])
        pids = np.arange(LK, dtype=np.uint32)
        e = np.stack([ctx.ntt_limbs(to_rns(sample_error(self.rng, ctx.N), ctx.primes), pids)
                      for _ in range(ctx.dnum)])
        return ctx.keygen_switch(s_new, s_old, a, e)

    def relin_key(self) -> np.ndarray:
        ctx = self.ctx
        pids = np.arange(ctx.L + ctx.K, dtype=np.uint32)
        s2 = ctx.mul_limbs(self.s_eval, self.s_eval, pids)
        return self._switch_key(self.s_eval, s2)

    def rotation_key(self, k: int) -> np.ndarray:
        ctx = self.ctx
        kinv = pow(k, -1, 2 * ctx.N)
        s_perm = ctx.automorph(self.s_eval, kinv)
        return self._switch_key(s_perm, self.s_eval)

    def encapsulation_keys(self, h: int) -> tuple:
        """Sparse-secret encapsulation (reference src/CKKS/Bootstrap.cu:254-260,
        592-600, BootstrapPrecomputation.cuh:37-38): a second ternary secret
        s' of Hamming weight h and the two switching keys, s -> s' (used at
        the bottom level before ModRaise) and s' -> s (after it)."""
        ctx = self.ctx
        sp = sample_ternary(self.rng, ctx.N, h)
        pids = np.arange(len(ctx.primes), dtype=np.uint32)
        sp_eval = ctx.ntt_limbs(to_rns(sp, ctx.primes), pids)
        return self._switch_key(sp_eval, self.s_eval), self._switch_key(self.s_eval, sp_eval)

    # ---------------------------------------------------------- enc / dec
    def encrypt(self, m_coef: np.ndarray, l: int) -> np.ndarray:
        """Secret-key encryption of integer coefficients at l Q limbs -> [2][l][N]."""
        ctx = self.ctx
        qs = ctx.q[:l]
        pids = np.arange(l, dtype=np.uint32)
        a = sample_uniform(self.rng, qs, ctx.N)
        e = sample_error(self.rng, ctx.N)
        if m_coef.dtype == object:
            coef = to_rns_big([int(x) + int(y) for x, y in zip(m_coef, e)], qs)
        else:
            coef = to_rns(m_coef.astype(np.int64) + e, qs)
        me = ctx.ntt_limbs(coef, pids)
        as_ = ctx.mul_limbs(a, self.s_eval[:l], pids)
        c0 = ctx.sub_limbs(me, as_, pids)
        return np.stack([c0, a])

    def decrypt_coef(self, ct: np.ndarray, l: int) -> list:
        ctx = self.ctx
        pids = np.arange(l, dtype=np.uint32)
        m = ctx.add_limbs(ct[0], ctx.mul_limbs(ct[1], self.s_eval[:l], pids), pids)
        mc = ctx.intt_limbs(m, pids)
        return crt_centered(mc, ctx.q[:l])


def precision_bits(got: np.ndarray, want: np.ndarray) -> float:
    err = float(np.max(np.abs(np.asarray(got) - np.asarray(want))))
    return float("inf") if err == 0 else -np.log2(err)
