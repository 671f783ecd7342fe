"""The product's CKKS encoder (fideslib_amd/codec.py, used for the device
plaintexts the library builds itself) against the oracle's encoder
(oracle/ckks_codec.py) and against the canonical embedding evaluated
directly (CPU only).

Both encoders are restatements of the same definition, so each is also
pinned independently: slot j of m(X) is m(zeta^(5^j mod 2N)),
zeta = exp(i pi / N), checked by evaluating the polynomial at those points
(O(N^2), small N).  Slot order follows OpenFHE (reference
src/CKKS/Context.cu rotation index 5^r; the encoder itself lives in
OpenFHE's CKKSPackedEncoding, absent from the reference tree).
"""
import numpy as np
import pytest

from fideslib_amd import codec as PC
from fideslib_amd import params as P
from oracle import ckks_codec as OC


def _slots(rng, n):
    return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)


def _embed(coeffs, N, scale):
    """m(zeta^(5^j)) / scale for j < N/2 by direct evaluation."""
    m = np.array([float(c) for c in coeffs])
    out = np.empty(N // 2, np.complex128)
    g = 1
    for j in range(N // 2):
        zeta = np.exp(1j * np.pi * g / N)
        out[j] = np.polyval(m[::-1], zeta)
        g = g * 5 % (2 * N)
    return out / scale


@pytest.mark.parametrize("logN", [4, 6, 10, 13])
def test_encode_equals_oracle_full_slots(logN):
    N = 1 << logN
    rng = np.random.default_rng(logN)
    z = _slots(rng, N // 2)
    for scale in (2.0 ** 20, 2.0 ** 40, 2.0 ** 59):
        a = PC.encode(z, scale, N)
        b = OC.encode(z, scale, N)
        assert a.dtype == b.dtype
        assert np.array_equal(a, b)
        assert np.array_equal(PC.slot_points(N), OC._slot_index(N))


@pytest.mark.parametrize("slots", [1, 2, 8, 64])
def test_encode_equals_oracle_sparse_slots(slots):
    N = 1 << 10
    z = _slots(np.random.default_rng(slots), slots)
    a = PC.encode(z, 2.0 ** 40, N)
    b = OC.encode(z, 2.0 ** 40, N)
    assert np.array_equal(a, b)
    # sparse packing replicates the slots over N/2
    assert np.array_equal(a, PC.encode(np.tile(z, (N // 2) // slots), 2.0 ** 40, N))


def test_encode_big_coefficients_equal_oracle():
    # past 2^62 both return Python integers (object arrays)
    N = 1 << 8
    z = _slots(np.random.default_rng(3), N // 2) * 4
    a = PC.encode(z, 2.0 ** 70, N)
    b = OC.encode(z, 2.0 ** 70, N)
    assert a.dtype == object and b.dtype == object
    assert [int(x) for x in a] == [int(x) for x in b]


@pytest.mark.parametrize("logN", [3, 5, 7])
def test_encode_is_the_canonical_embedding(logN):
    # evaluate the encoded polynomial at zeta^(5^j): the slots come back up to
    # the rounding of the coefficients (|error| <= N/2 / scale per slot)
    N = 1 << logN
    scale = 2.0 ** 30
    z = _slots(np.random.default_rng(logN + 100), N // 2)
    got = _embed(PC.encode(z, scale, N), N, scale)
    assert np.max(np.abs(got - z)) < N / scale
    # a wrong slot order (5^j replaced by 3^j) or a missing conjugate half fails
    bad = np.empty(N // 2, np.complex128)
    m = PC.encode(z, scale, N).astype(float)
    g = 1
    for j in range(N // 2):
        bad[j] = np.polyval(m[::-1], np.exp(1j * np.pi * g / N)) / scale
        g = g * 3 % (2 * N)
    if N > 8:
        assert np.max(np.abs(bad - z)) > 1e-3


@pytest.mark.parametrize("logN", [6, 12])
def test_decode_equals_oracle_and_inverts_encode(logN):
    N = 1 << logN
    scale = 2.0 ** 40
    z = _slots(np.random.default_rng(logN + 7), N // 2)
    c = PC.encode(z, scale, N)
    a = PC.decode(c, scale, N)
    b = OC.decode(c, scale, N)
    assert np.array_equal(a, b)
    assert np.max(np.abs(a - z)) < 1e-9


def test_to_rns_equals_oracle():
    prm = P.config1()
    N = 1 << 9
    rng = np.random.default_rng(5)
    c = rng.integers(-(1 << 62), 1 << 62, N, dtype=np.int64)
    assert np.array_equal(PC.to_rns(c, prm.q), OC.to_rns(c, prm.q))
    big = [int(x) * (1 << 40) + 12345 for x in c[:64]]
    obj = np.array(big, dtype=object)
    assert np.array_equal(PC.to_rns(obj, prm.q), OC.to_rns_big(big, prm.q))
