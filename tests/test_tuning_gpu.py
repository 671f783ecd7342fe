"""Launch-shape tuning on the GPU.

Every setting of fides_tuning (conversion chunk per size class, NTT task
grouping by prime, fused or separate row pass of the inner product) must give
the same bits: a batch of HMult+relin+Rescale and Rotate at mixed levels
(N = 2^16, 30 Q limbs, K = 10, dnum = 3, random residues and a random key) is
run under each setting and compared with the default.  The autotuner keeps a
valid setting in the context, and results after it are unchanged.  Invalid
settings, and the autotuner during a graph capture, are refused; settings
belong to their context.
"""
import hashlib

import numpy as np
import pytest

from fideslib_amd import capi, params

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def setup():
    prm = params.config3()
    N = prm.N
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(99)
    primes = list(prm.q) + list(prm.p)
    key = np.empty((2, prm.dnum, len(primes), N), dtype=np.uint64)
    for m, q in enumerate(primes):
        key[:, :, m, :] = rng.integers(0, q, (2, prm.dnum, N), dtype=np.uint64)
    gkey = gctx.upload_key(key)
    # 24 ciphertexts: enough conversion targets to reach every size class
    # when run whole, and a 3-ciphertext prefix for the small class
    levels = [30, 6] + [int(x) for x in rng.integers(5, prm.L + 1, 22)]
    cts = []
    for l in levels:
        cts.append(capi.DeviceBuffer.from_array(
            np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in prm.q[:l]]) for _ in range(2)])))
    return prm, gctx, gkey, levels, cts


def run_batch(setup, n):
    prm, gctx, gkey, levels, cts = setup
    N = prm.N
    lv = levels[:n]
    A = cts[:n]
    B = [cts[(i + 1) % n] if levels[(i + 1) % n] == levels[i] else cts[i] for i in range(n)]
    OM = [capi.DeviceBuffer(2 * (l - 1) * N) for l in lv]
    OR = [capi.DeviceBuffer(2 * l * N) for l in lv]
    gctx.hmult_batch(A, B, lv, gkey, OM, rescale=True)
    gctx.rotate_batch(A, lv, [gkey] * n, [pow(5, 3, 2 * N)] * n, OR)
    gctx.sync()
    h = hashlib.sha256()
    for o, r, l in zip(OM, OR, lv):
        h.update(o.download((2, l - 1, N)).tobytes())
        h.update(r.download((2, l, N)).tobytes())
    return h.hexdigest()


DEFAULT = {"conv_chunk": [8, 16, 32], "group_by_prime": [1, 1, 1], "fuse_row_dot": [1, 1, 1]}
SETTINGS = [
    {"conv_chunk": [8, 16, 32], "group_by_prime": [1, 1, 1], "fuse_row_dot": [0, 0, 0]},
    {"conv_chunk": [8, 16, 32], "group_by_prime": [0, 0, 0], "fuse_row_dot": [1, 1, 1]},
    {"conv_chunk": [1, 1, 1], "group_by_prime": [1, 0, 1], "fuse_row_dot": [1, 0, 0]},
    {"conv_chunk": [32, 32, 32], "group_by_prime": [0, 0, 0], "fuse_row_dot": [0, 0, 0]},
    {"conv_chunk": [5, 7, 3], "group_by_prime": [0, 1, 0], "fuse_row_dot": [0, 1, 1]},
    {"conv_chunk": [96, 64, 48], "group_by_prime": [1, 1, 1], "fuse_row_dot": [1, 1, 1]},
]


@pytest.mark.parametrize("n", [3, 24])
def test_every_setting_gives_the_same_bits(setup, n):
    gctx = setup[1]
    default = gctx.tuning
    assert default == DEFAULT
    want = run_batch(setup, n)
    say("n=%d default %s" % (n, want[:16]))
    try:
        for s in SETTINGS:
            gctx.tuning = s
            assert gctx.tuning == s
            got = run_batch(setup, n)
            say("n=%d %s %s" % (n, s, got[:16]))
            assert got == want, s
    finally:
        gctx.tuning = default


def test_invalid_settings_refused(setup):
    gctx = setup[1]
    before = gctx.tuning
    for bad in (dict(DEFAULT, conv_chunk=[0, 16, 32]), dict(DEFAULT, conv_chunk=[8, 97, 32]),
                dict(DEFAULT, group_by_prime=[1, 2, 1]), dict(DEFAULT, fuse_row_dot=[1, 1, 7])):
        with pytest.raises(capi.FidesError):
            gctx.tuning = bad
    assert gctx.tuning == before
    with pytest.raises(capi.FidesError):
        gctx.autotune([1])  # rescale needs two limbs


def test_autotune_keeps_a_valid_setting(setup):
    prm, gctx = setup[0], setup[1]
    want = run_batch(setup, 24)
    chosen = gctx.autotune(setup[3], reps=2)
    say("autotune chose %s" % chosen)
    assert gctx.tuning == chosen
    assert all(1 <= v <= 96 for v in chosen["conv_chunk"])
    assert all(v in (0, 1) for v in chosen["group_by_prime"] + chosen["fuse_row_dot"])
    # the batch reaches classes 1 (ModDown) and 2 (ModUp): class 0 is untouched
    assert chosen["conv_chunk"][0] == 8 and chosen["group_by_prime"][0] == 1 and chosen["fuse_row_dot"][0] == 1
    assert run_batch(setup, 24) == want
    gctx.tuning = DEFAULT


def test_autotune_refused_during_capture_and_settings_per_context(setup):
    prm, gctx = setup[0], setup[1]
    levels, cts = setup[3], setup[4]
    st = capi.Stream()
    out = capi.DeviceBuffer(2 * levels[1] * prm.N)
    gctx.graph_begin(st, 1 << 26, 1 << 20)
    with pytest.raises(capi.FidesError):
        gctx.autotune([prm.L])
    gctx.add_batch([cts[1]], [cts[1]], [levels[1]], [out], stream=st.h)
    g = gctx.graph_end()
    g.launch(st)
    st.sync()
    g.close()
    st.close()
    other = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx.tuning = dict(DEFAULT, conv_chunk=[4, 4, 4])
    assert other.tuning == DEFAULT  # settings belong to their context
    gctx.tuning = DEFAULT
