"""The benchmark's own headline workload, bit for bit: bench.py's 64
ciphertexts (N = 2^16, 30 Q limbs, K = 10, dnum = 3, limb counts drawn
U{6..30} from the bench seed; HMult+relin+Rescale then Rotate by 5^r, r ~
U[1, 2^15), each with its own resident key) run through the batched path,
the plain per-ciphertext path and the per-ciphertext fused path, under the
default launch shapes and under the shapes the autotuner picked on the
driver's bench run.  Every output of every path must be identical, and the
l = 30 and smallest-level ciphertexts must equal the oracle (the reference's
ASSERT_EQ_CIPHERTEXT limb check, test/ParametrizedTest.cuh:543-558)."""
import numpy as np
import pytest

import bench
from fideslib_amd import capi, params
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TUNINGS = [None, {"conv_chunk": [8, 16, 32], "group_by_prime": [0, 1, 0], "fuse_row_dot": [1, 1, 1]}]


@pytest.fixture(scope="module")
def workload():
    prm = params.config3()
    ctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    print("bench workload: building", flush=True)
    w = bench.Workload(prm, ctx, np.random.default_rng(bench.SEED))
    return prm, ctx, w


@pytest.mark.parametrize("tuning", TUNINGS, ids=["default", "autotuned"])
def test_bench_workload_batched_equals_plain(workload, tuning):
    prm, ctx, w = workload
    if tuning is not None:
        ctx.tuning = tuning
    print("tuning %s" % ctx.tuning, flush=True)
    w.step_batched(ctx)
    batched = w.outputs("batched")
    w.step_plain(ctx)
    plain = w.outputs("plain")
    w.step_plain_fused(ctx)
    fused = w.outputs("plain")
    for i, l in enumerate(w.levels):
        for k in range(2):
            assert np.array_equal(batched[i][k], plain[i][k]), (i, l, k)
            assert np.array_equal(batched[i][k], fused[i][k]), (i, l, k)
    if tuning is None:
        octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
        for i in (w.levels.index(max(w.levels)), w.levels.index(min(w.levels))):
            l = w.levels[i]
            print("oracle: ciphertext %d, l=%d" % (i, l), flush=True)
            a = np.ascontiguousarray(w.pool_a[:, :l])
            b = np.ascontiguousarray(w.pool_b[:, :l])
            r = octx.rescale_ct(octx.hmult_relin(a, b, l, w.key_host), l)
            assert np.array_equal(batched[i][0], r), (i, l)
            assert np.array_equal(batched[i][1], octx.rotate(r, l - 1, w.key_host, w.kidx[i])), (i, l)
