"""Request loop on the GPU.

* N = 2^12 with real keys: every request served in a mix equals the oracle
  running the same program op by op (bit-exact), equals the same request
  served alone, and decrypts to the FP64 evaluation of its program
  (Chebyshev series / rotate-and-sum) to a stated precision.
* N = 2^16, 30 limbs, synthetic keys: Chebyshev, rotate-and-sum and
  bootstrap requests (FLEXIBLEAUTO, sparse-secret encapsulation) give
  identical results alone and in a mix (exact arithmetic: any difference is
  a race or a storage-reuse bug).
"""
import numpy as np
import pytest

from fideslib_amd import capi, params
from fideslib_amd import serve as S
from oracle import ckks_codec as C
from oracle import oracle as O
from serve_oracle import run_oracle

pytestmark = pytest.mark.gpu

SIGMOID = lambda x: 1.0 / (1.0 + np.exp(-x))


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def small():
    prm = params.serve_small()
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 123)
    kmax = 6
    rlk = keys.relin_key()
    rot = {pow(5, 1 << i, 2 * prm.N): None for i in range(kmax)}
    rot = {k: keys.rotation_key(k) for k in rot}
    be = S.GpuBackend(gctx, gctx.upload_key(rlk), {k: gctx.upload_key(v) for k, v in rot.items()}, 24 << 20)
    return prm, octx, gctx, keys, rlk, rot, be, kmax


def _requests(prm, keys, rng, n, kmax):
    N = prm.N
    coef = S.chebyshev_coefficients(SIGMOID, 15, -8, 8)
    out = []
    for i in range(n):
        if i % 3 != 2:
            l = int(rng.integers(S.chebyshev_depth(15) + 1, prm.L + 1))
            P = S.chebyshev_program(prm, l, coef, -8, 8)
            z = rng.uniform(-8, 8, N // 2)
            want = np.polynomial.chebyshev.chebval(z / 8, coef)
        else:
            l = int(rng.integers(1, prm.L + 1))
            k = int(rng.integers(1, kmax + 1))
            P = S.rotsum_program(prm, l, k)
            z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
            want = sum(np.roll(z, -t) for t in range(1 << k))
        ct = keys.encrypt(C.encode(z, 2.0 ** prm.scale_bits, N), l)
        out.append((P, ct, want))
    return out


def test_serve_small_mix_oracle_alone_precision(small):
    prm, octx, gctx, keys, rlk, rot, be, kmax = small
    say("serve small mix")
    rng = np.random.default_rng(9)
    work = _requests(prm, keys, rng, 9, kmax)
    arr = np.cumsum(rng.exponential(0.002, len(work)))
    reqs = [S.Request(i, P, float(arr[i]), inputs=[ct], keep_result=True) for i, (P, ct, _) in enumerate(work)]
    srv = S.Server(be, prm.N)
    srv.run(reqs, realtime=True)
    assert srv.stats["launches"] < srv.stats["ops"]
    for r, (P, ct, want) in zip(reqs, work):
        assert r.status == "done"
        # bit-exact against the oracle
        ref = run_oracle(octx, P, [ct], rlk, rot)
        assert np.array_equal(r.result, ref), r.rid
        # decrypted against FP64
        dec = C.decode(keys.decrypt_coef(r.result, P.output.level), P.output.scale, prm.N)
        bits = C.precision_bits(dec.real if P.name.startswith("cheb") else dec, want)
        assert bits > 15, (r.rid, P.name, bits)
    # each request alone gives the same bits
    for r, (P, ct, _) in list(zip(reqs, work))[:4]:
        alone = S.Request(100 + r.rid, P, 0.0, inputs=[ct], keep_result=True)
        S.Server(be, prm.N).run([alone], realtime=False)
        assert np.array_equal(alone.result, r.result), r.rid


def test_serve_small_memory_pressure_and_rejection(small):
    """An arena that holds only a few requests at once still serves all of
    them (queued until storage frees), and one larger than the arena is
    rejected rather than failing the loop."""
    prm, octx, gctx, keys, rlk, rot, be, kmax = small
    rng = np.random.default_rng(10)
    work = _requests(prm, keys, rng, 6, kmax)
    need = max(S._compile(P).block for P, _, _ in work) * prm.N * 8
    tight = S.GpuBackend(gctx, be.rlk, be.keys, int(need * 1.5))
    reqs = [S.Request(i, P, 0.0, inputs=[ct], keep_result=True) for i, (P, ct, _) in enumerate(work)]
    big = S.Program(prm.q, prm.logN, prm.scale_bits, "big")
    x = [big.input(prm.L) for _ in range(40)]
    big.set_output(big.sum_tree(x))
    reqs.append(S.Request(99, big, 0.0, inputs=[np.zeros((2, prm.L, prm.N), np.uint64)] * 40))
    S.Server(tight, prm.N).run(reqs, realtime=False)
    assert reqs[-1].status == "rejected"
    for r, (P, ct, _) in zip(reqs, work):
        assert r.status == "done"
        assert np.array_equal(r.result, run_oracle(octx, P, [ct], rlk, rot)), r.rid


def test_serve_cfg3_alone_equals_mix():
    """N=2^16, 30 limbs, synthetic keys and ciphertexts; two of the ten
    requests are bootstraps."""
    from fideslib_amd import bootstrap as B
    prm = params.config3()
    say("serve cfg3 determinism")
    N = prm.N
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(31)
    LK = prm.L + prm.K

    def rkey():
        return np.stack([np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in prm.q + prm.p])
                                   for _ in range(prm.dnum)]) for _ in range(2)])

    kmax = 3
    rlk = gctx.upload_key(rkey())
    keys = {pow(5, 1 << i, 2 * N): gctx.upload_key(rkey()) for i in range(kmax)}
    bp = B.BootstrapPlan(prm, gctx, B.BootConfig.encapsulated(flexible=True))
    bp.build_plaintexts()
    shared = rkey()  # one synthetic key, a distinct device copy per bootstrap key
    for r in bp.rotations() + [None]:
        k = pow(5, r, 2 * N) if r is not None else 2 * N - 1
        if k not in keys:
            keys[k] = gctx.upload_key(shared)
    for k in bp.switch_keys():
        keys[k] = gctx.upload_key(shared)
    sf = params.flexible_scaling_factors(prm.q)
    be = S.GpuBackend(gctx, rlk, keys, 24 << 30)
    coef = S.chebyshev_coefficients(SIGMOID, 15, -8, 8)
    work = []
    for i in range(8):
        if i % 2 == 0:
            P = S.chebyshev_program(prm, int(rng.integers(S.chebyshev_depth(15) + 1, prm.L + 1)), coef, -8, 8)
        else:
            P = S.rotsum_program(prm, int(rng.integers(1, prm.L + 1)), int(rng.integers(1, kmax + 1)))
        l = P.inputs[0].level
        ct = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in prm.q[:l]]) for _ in range(2)])
        work.append((P, ct))
    for lb in (3, 17):
        P = bp.program(lb, sf[lb - 1])
        ct = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in prm.q[:lb]]) for _ in range(2)])
        work.insert(3 if lb == 3 else 7, (P, ct))
    reqs = [S.Request(i, P, 0.001 * i, inputs=[ct], keep_result=True) for i, (P, ct) in enumerate(work)]
    srv = S.Server(be, N)
    srv.run(reqs, realtime=True)
    say("cfg3 mix: %d steps, %d launches, %d ops" % (srv.stats["steps"], srv.stats["launches"], srv.stats["ops"]))
    for r, (P, ct) in zip(reqs, work):
        alone = S.Request(100 + r.rid, P, 0.0, inputs=[ct], keep_result=True)
        S.Server(be, N).run([alone], realtime=False)
        assert np.array_equal(alone.result, r.result), (r.rid, P.name)
