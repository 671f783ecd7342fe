"""Parameter selection on the host: NTT-friendly primes and parameter sets.

The reference takes primes, roots and the digit partition from OpenFHE
(src/CKKS/openfhe-interface/RawCiphertext.cu:213-643, GetRawParams) and
falls back to fixed tables (api/CryptoContext.cpp:53-113).  OpenFHE is not
available here, so this module picks primes itself with a documented rule:

  primes q = 1 (mod 2N) of exactly `bits` bits, scanned downward from
  2^bits + 1 in steps of 2N, first `count` that pass a deterministic
  Miller-Rabin test, skipping any listed in `exclude`.

The special-prime count follows OpenFHE's HYBRID rule (enough 60-bit
special primes that P exceeds the largest digit product); any count can be
passed explicitly.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field

_MR_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n: int) -> bool:
    """Deterministic Miller-Rabin for n < 3.3e24 (covers 64-bit)."""
    if n < 2:
        return False
    for p in _MR_BASES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _MR_BASES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def ntt_primes(bits: int, count: int, logN: int, exclude=()) -> list[int]:
    step = 2 << logN
    x = ((1 << bits) // step) * step + 1
    if x >= (1 << bits):
        x -= step
    out = []
    ex = set(int(e) for e in exclude)
    while len(out) < count:
        if x.bit_length() != bits:
            raise ValueError("ran out of %d-bit primes = 1 mod 2N" % bits)
        if x not in ex and is_prime(x):
            out.append(x)
        x -= step
    return out


def special_prime_count(q: list[int], dnum: int, pbits: int = 60) -> int:
    """Smallest K with K * pbits > max digit bit size (OpenFHE HYBRID rule)."""
    alpha = (len(q) + dnum - 1) // dnum
    maxbits = 0
    for j in range(dnum):
        prod = 1
        for x in q[j * alpha:(j + 1) * alpha]:
            prod *= x
        maxbits = max(maxbits, prod.bit_length())
    return math.ceil((maxbits + 1) / pbits)


@dataclass
class ParamSet:
    """One CKKS parameter set: ring degree, Q and P primes, digits."""
    name: str
    logN: int
    q: list
    p: list
    dnum: int
    scale_bits: int
    roots: list | None = None
    notes: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return 1 << self.logN

    @property
    def L(self) -> int:
        return len(self.q)

    @property
    def K(self) -> int:
        return len(self.p)

    @property
    def alpha(self) -> int:
        return (self.L + self.dnum - 1) // self.dnum

    def digits(self, l: int) -> int:
        return (l + self.alpha - 1) // self.alpha


def ckks_params(logN: int, first_bits: int, scale_bits: int, n_scale: int, dnum: int,
                special_bits: int = 60, K: int | None = None, name: str = "") -> ParamSet:
    """Q = one first_bits prime + n_scale scale_bits primes; P = K special primes."""
    q0 = ntt_primes(first_bits, 1, logN)
    qs = ntt_primes(scale_bits, n_scale, logN, exclude=q0)
    q = q0 + qs
    if K is None:
        K = special_prime_count(q, dnum, special_bits)
    p = ntt_primes(special_bits, K, logN, exclude=q)
    return ParamSet(name or "logN%d_L%d_dnum%d" % (logN, len(q), dnum), logN, q, p, dnum, scale_bits)


def reference_primes() -> dict:
    """FIDESlib's own test prime tables (p64, sp64), shipped as a fixture."""
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "reference_primes.json")) as f:
        return json.load(f)


# --- the workloads named by the headline benchmark --------------------------

def config1() -> ParamSet:
    """N=2^13, one 60-bit + three 40-bit Q primes, scale 2^40, dnum 2."""
    return ckks_params(13, 60, 40, 3, dnum=2, name="cfg1_N2^13_L4_dnum2")


def config2() -> ParamSet:
    """N=2^16, 32 Q limbs + 8 special limbs from FIDESlib's 51-bit test primes."""
    pr = reference_primes()["p64"]
    return ParamSet("cfg2_N2^16_32+8", 16, pr[1:33], pr[33:41], dnum=4, scale_bits=51)


def config3(K: int | None = None) -> ParamSet:
    """N=2^16, 30 Q limbs (60 + 29x59 bits), dnum 3, scale 2^59."""
    return ckks_params(16, 60, 59, 29, dnum=3, K=K, name="cfg3_N2^16_L29_dnum3")


def serve_small() -> ParamSet:
    """N=2^12, one 60-bit + eleven 50-bit Q primes, dnum 3: deep enough for
    the degree-15 Chebyshev request at small N (request-loop tests)."""
    return ckks_params(12, 60, 50, 11, dnum=3, name="serve_N2^12_L12_dnum3")


def boot_small() -> ParamSet:
    """N=2^12 with the headline prime sizes (60 + 21x59 bits, dnum 3): deep
    enough for a full bootstrap at small N (bootstrap tests)."""
    return ckks_params(12, 60, 59, 21, dnum=3, name="boot_N2^12_L22_dnum3")


def flexible_scaling_factors(q) -> list:
    """OpenFHE's FLEXIBLEAUTO scaling factor per level (index = limbs - 1), as
    the reference copies them into RawParams (RawCiphertext.cu:260-272):
    sf[L-1] = q[L-1], sf[l-1] = sf[l]^2 / q[l]."""
    L = len(q)
    sf = [0.0] * L
    sf[L - 1] = float(q[L - 1])
    for l in range(L - 1, 0, -1):
        sf[l - 1] = sf[l] * sf[l] / float(q[l])
    return sf
