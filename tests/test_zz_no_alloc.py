"""The allocation-free key switch (fides_keyswitch_noalloc) against
fides_keyswitch, through the C ABI: bit-identical outputs on two calls in a
row, the second one smaller (its scratch array keeps the first call's
contents), no growth of device memory use across calls, and a clear refusal
of a key switch larger than the static scratch.  N = 2^16, 30 Q limbs
(60 + 29 x 59 bits), K = 10, dnum = 3; the smaller call is also checked
against the oracle."""
import ctypes

import numpy as np
import pytest

from fideslib_amd import capi, params
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rand_limbs(rng, primes, N):
    return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])


def free_device_bytes() -> int:
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_keyswitch_noalloc_equals_keyswitch():
    prm = params.config3()
    N = prm.N
    print("no-alloc key switch: N=2^16 L=%d K=%d dnum=%d" % (prm.L, prm.K, prm.dnum), flush=True)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(4242)
    key = np.stack([np.stack([rand_limbs(rng, prm.q + prm.p, N) for _ in range(prm.dnum)]) for _ in range(2)])
    gkey = gctx.upload_key(key)
    outs = {}
    for call, l in enumerate((30, 12)):
        print("call %d: l=%d" % (call, l), flush=True)
        d = rand_limbs(rng, prm.q[:l], N)
        din = capi.DeviceBuffer.from_array(d)
        a0, a1 = capi.DeviceBuffer(l * N), capi.DeviceBuffer(l * N)
        b0, b1 = capi.DeviceBuffer(l * N), capi.DeviceBuffer(l * N)
        gctx.keyswitch(din, l, gkey, a0, a1)
        gctx.keyswitch_noalloc(din, l, gkey, b0, b1)
        gctx.sync()
        want = (a0.download((l, N)), a1.download((l, N)))
        got = (b0.download((l, N)), b1.download((l, N)))
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), l
        outs[l] = (d, got)
    # the smaller call against the oracle
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    d, got = outs[12]
    w0, w1 = octx.keyswitch(d, 12, key)
    assert np.array_equal(got[0], w0) and np.array_equal(got[1], w1)


def test_keyswitch_noalloc_allocates_nothing():
    prm = params.config3()
    N, l = prm.N, prm.L
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(4243)
    key = np.stack([np.stack([rand_limbs(rng, prm.q + prm.p, N) for _ in range(prm.dnum)]) for _ in range(2)])
    gkey = gctx.upload_key(key)
    din = capi.DeviceBuffer.from_array(rand_limbs(rng, prm.q[:l], N))
    o0, o1 = capi.DeviceBuffer(l * N), capi.DeviceBuffer(l * N)
    gctx.keyswitch_noalloc(din, l, gkey, o0, o1)  # first launch loads the kernels
    gctx.sync()
    before = free_device_bytes()
    for _ in range(3):
        gctx.keyswitch_noalloc(din, l, gkey, o0, o1)
    gctx.sync()
    after = free_device_bytes()
    # the allocating entry point would take a 180 MiB arena here
    assert before - after < (16 << 20), (before, after)


def test_keyswitch_noalloc_refuses_oversize():
    """N = 2^17, 20 Q limbs, dnum 2: a key switch needs 166 N words (170 MiB),
    more than the 160 MiB scratch; refused with FIDES_ERR_ARG, nothing run."""
    prm = params.ckks_params(17, 60, 50, 19, dnum=2, K=6)
    N, l = prm.N, prm.L
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    key = np.zeros((2, prm.dnum, prm.L + prm.K, N), np.uint64)
    gkey = gctx.upload_key(key)
    din = capi.DeviceBuffer(l * N)
    o0, o1 = capi.DeviceBuffer(l * N), capi.DeviceBuffer(l * N)
    with pytest.raises(capi.FidesError, match="scratch"):
        gctx.keyswitch_noalloc(din, l, gkey, o0, o1)
