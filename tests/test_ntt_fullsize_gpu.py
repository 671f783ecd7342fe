"""The forward/inverse NTT at the roofline launch's full size, on the GPU:
one fides_ntt call over 960 limbs of N = 2^16 (the 30 Q primes of the
key-switching workload, 32 limbs each, grouped by prime as the batched
executor issues them).

The oracle checks a sample of the limbs; the rest is held by properties that
do not depend on the size: INTT(NTT(x)) = x, linearity NTT(x + y) = NTT(x) +
NTT(y) mod q, and the closed form of the impulse (NTT of delta_0 is all ones
in every limb, for any root and bit-reversed order).
"""
import numpy as np
import pytest

from fideslib_amd import capi, params
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LIMBS = 960


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def full():
    prm = params.config3()
    nq = 30
    pids = [i for i in range(nq) for _ in range(LIMBS // nq)]
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    q = np.array([prm.q[p] for p in pids], np.uint64)[:, None]
    return prm, pids, q, octx, gctx


def test_fullsize_roundtrip_linearity_and_oracle_sample(full):
    prm, pids, q, octx, gctx = full
    N = prm.N
    rng = np.random.default_rng(960)
    x = (rng.integers(0, 1 << 62, (LIMBS, N), dtype=np.uint64) % q).astype(np.uint64)
    y = (rng.integers(0, 1 << 62, (LIMBS, N), dtype=np.uint64) % q).astype(np.uint64)
    say("full-size NTT: %d limbs x 2^%d" % (LIMBS, prm.logN))
    bx = capi.DeviceBuffer.from_array(x)
    gctx.ntt(bx, pids)
    X = bx.download(x.shape)
    # a sample of limbs (every prime appears) against the oracle
    sample = list(range(0, LIMBS, 31))
    want = octx.ntt_limbs(x[sample], [pids[t] for t in sample])
    assert np.array_equal(X[sample], want)
    # round trip
    gctx.intt(bx, pids)
    assert np.array_equal(bx.download(x.shape), x)
    # linearity: NTT(x + y) = NTT(x) + NTT(y) mod q
    s = x + y
    s = np.where(s >= q, s - q, s).astype(np.uint64)
    del x
    by = capi.DeviceBuffer.from_array(y)
    gctx.ntt(by, pids)
    Y = by.download(y.shape)
    del by, y
    bs = capi.DeviceBuffer.from_array(s)
    gctx.ntt(bs, pids)
    S = bs.download(s.shape)
    XY = X + Y
    XY = np.where(XY >= q, XY - q, XY).astype(np.uint64)
    assert np.array_equal(S, XY)
    # a planted error in one residue breaks the check (the comparison is live)
    XY[LIMBS // 2, 12345] ^= 1
    assert not np.array_equal(S, XY)


def test_fullsize_impulse_closed_form(full):
    prm, pids, q, octx, gctx = full
    N = prm.N
    d = np.zeros((LIMBS, N), np.uint64)
    d[:, 0] = 1
    bd = capi.DeviceBuffer.from_array(d)
    gctx.ntt(bd, pids)
    assert np.array_equal(bd.download(d.shape), np.ones((LIMBS, N), np.uint64))
    gctx.intt(bd, pids)
    assert np.array_equal(bd.download(d.shape), d)
