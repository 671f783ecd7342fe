"""Per-shape graph replay in the request loop (serve.GraphCache) on the GPU.

* N = 2^12 with real keys: a Poisson mix where part of the shapes are
  captured; every request, whether served by a graph replay (on the cache's
  own stream, concurrent with the batched steps) or by the steps, equals the
  oracle running its program op by op, bit for bit.  Two requests of one
  captured shape in a row check that a replay reads its new inputs.
* N = 2^16, 30 limbs, synthetic keys: captured Chebyshev and bootstrap
  (FLEXIBLEAUTO, sparse-secret encapsulation) shapes, on two lanes (streams
  with their own scratch), give the same bits as the batched loop without
  graphs.
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


def make_cache(gctx, rlk, keys, programs, scratch_bytes, copies=1, lanes=1):
    """Programs spread round-robin over `lanes` lanes (stream, register
    memory and scratch each)."""
    N = programs[0].N
    split = [programs[i::lanes] for i in range(lanes)]
    cache, keep = None, []
    for i, progs in enumerate(split):
        words = S.GraphCache.words_needed(progs, copies)
        be = S.GpuBackend(gctx, rlk, keys, words * 8, stream=capi.Stream())
        scratch = capi.DeviceBuffer(scratch_bytes // 8)
        keep.append(scratch)
        if cache is None:
            cache = S.GraphCache(be, N, scratch_bytes, scratch_ptr=scratch.ptr)
        else:
            assert cache.add_lane(be, scratch_ptr=scratch.ptr) == i
        for P in progs:
            cache.add(P, copies, lane=i)
    cache._scratch = keep   # keep alive with the cache
    return cache


def test_graph_replays_equal_oracle_small():
    prm = params.serve_small()
    say("graph cache small: N=2^%d L=%d" % (prm.logN, prm.L))
    N = prm.N
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 321)
    kmax = 4
    rlk = keys.relin_key()
    rot = {pow(5, 1 << i, 2 * N): None for i in range(kmax)}
    rot = {k: keys.rotation_key(k) for k in rot}
    g_rlk = gctx.upload_key(rlk)
    g_keys = {k: gctx.upload_key(v) for k, v in rot.items()}
    be = S.GpuBackend(gctx, g_rlk, g_keys, 24 << 20)
    coef = S.chebyshev_coefficients(SIGMOID, 15, -8, 8)
    dmin = S.chebyshev_depth(15) + 1
    rng = np.random.default_rng(77)
    progs = []
    for i in range(12):
        if i % 3 != 2:
            progs.append(S.chebyshev_program(prm, int(rng.integers(dmin, prm.L + 1)), coef, -8, 8))
        else:
            progs.append(S.rotsum_program(prm, int(rng.integers(1, prm.L + 1)), int(rng.integers(1, kmax + 1))))
    # two requests of one shape back to back (the second replays with new inputs)
    progs.insert(1, S.chebyshev_program(prm, progs[0].inputs[0].level, coef, -8, 8)
                 if progs[0].name.startswith("cheb") else progs[0])
    cached = [progs[0]] + progs[2::2]
    cache = make_cache(gctx, g_rlk, g_keys, cached, 64 << 20)
    say("captured %s" % cache.stats)
    cts = [keys.encrypt(C.encode(rng.uniform(-1, 1, N // 2), 2.0 ** prm.scale_bits, N), P.inputs[0].level)
           for P in progs]
    arr = np.cumsum(rng.exponential(0.001, len(progs)))
    reqs = [S.Request(i, P, float(arr[i]), inputs=[ct], keep_result=True) for i, (P, ct) in enumerate(zip(progs, cts))]
    srv = S.Server(be, N)
    srv.run(reqs, realtime=True, graphs=cache)
    say("graph replays %d of %d requests" % (cache.stats["replays"], len(reqs)))
    assert cache.stats["replays"] >= 2
    assert reqs[0].graph
    for r, P, ct in zip(reqs, progs, cts):
        assert r.status == "done"
        assert np.array_equal(r.result, run_oracle(octx, P, [ct], rlk, rot)), (r.rid, r.graph, P.name)
    # the same mix again, every request through the cache when its shape is free
    reqs2 = [S.Request(i, P, 0.0, inputs=[ct], keep_result=True) for i, (P, ct) in enumerate(zip(progs, cts))]
    S.Server(be, N).run(reqs2, realtime=False, graphs=cache)
    for r, r2 in zip(reqs, reqs2):
        assert np.array_equal(r.result, r2.result), r.rid
    cache.close()


def test_graph_replays_equal_batched_cfg3():
    from fideslib_amd import bootstrap as B
    prm = params.config3()
    N = prm.N
    say("graph cache cfg3")
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(41)

    def rkey():
        return np.stack([np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in prm.q + prm.p])
                                   for _ in range(prm.dnum)]) for _ in range(2)])

    rlk = gctx.upload_key(rkey())
    shared = rkey()
    bp = B.BootstrapPlan(prm, gctx, B.BootConfig.encapsulated(flexible=True))
    bp.build_plaintexts()
    keys = {}
    for r in bp.rotations() + [None]:
        k = pow(5, r, 2 * N) if r is not None else 2 * N - 1
        keys[k] = gctx.upload_key(shared)
    for k in bp.switch_keys():
        keys[k] = gctx.upload_key(shared)
    sf = params.flexible_scaling_factors(prm.q)
    coef = S.chebyshev_coefficients(SIGMOID, 15, -8, 8)
    cheb = S.chebyshev_program(prm, 24, coef, -8, 8)
    boot = bp.program(9, sf[8])
    cache = make_cache(gctx, rlk, keys, [cheb, boot], 2 << 30, lanes=2)   # the bootstrap on a lane of its own
    say("captured %s" % cache.stats)
    be = S.GpuBackend(gctx, rlk, keys, 8 << 30)
    progs = [cheb, boot, S.chebyshev_program(prm, 24, coef, -8, 8), S.chebyshev_program(prm, 20, coef, -8, 8),
             bp.program(9, sf[8])]
    cts = [np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in prm.q[:P.inputs[0].level]])
                     for _ in range(2)]) for P in progs]
    mk = lambda: [S.Request(i, P, 0.0005 * i, inputs=[ct], keep_result=True) for i, (P, ct) in enumerate(zip(progs, cts))]
    with_g = mk()
    S.Server(be, N).run(with_g, realtime=True, graphs=cache)
    without = mk()
    S.Server(be, N).run(without, realtime=False)
    say("replays %d" % cache.stats["replays"])
    assert with_g[0].graph and with_g[1].graph
    for a, b in zip(with_g, without):
        assert np.array_equal(a.result, b.result), (a.rid, a.graph, a.program.name)
    cache.close()
