"""Graph capture on the GPU.

A closed batch of requests (two bootstraps with sparse-secret encapsulation
and a Chebyshev series, N = 2^12) is recorded into one HIP graph by
Server.capture and replayed three times, with other eager work and
synchronizations between the replays; every replay must equal the eager
request loop bit for bit.  Operations on another stream during the capture
are refused with an error (nothing is recorded or run for them).
"""
import numpy as np
import pytest

from fideslib_amd import bootstrap as B
from fideslib_amd import capi, params
from fideslib_amd import serve as S
from oracle import ckks_codec as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def test_graph_replays_equal_eager_loop():
    prm = params.boot_small()
    N = prm.N
    say("graph capture N=2^12")
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 21, h=None)
    bp = B.BootstrapPlan(prm, gctx, B.BootConfig.encapsulated())
    dev_keys = {}
    for r in bp.rotations() + [None]:
        k = pow(5, r, 2 * N) if r is not None else 2 * N - 1
        dev_keys[k] = gctx.upload_key(keys.rotation_key(k))
    for i in range(3):
        dev_keys.setdefault(pow(5, 1 << i, 2 * N), gctx.upload_key(keys.rotation_key(pow(5, 1 << i, 2 * N))))
    down, up = keys.encapsulation_keys(bp.cfg.encaps_h)
    dev_keys[bp.ENCAPS_DOWN], dev_keys[bp.ENCAPS_UP] = gctx.upload_key(down), gctx.upload_key(up)
    rlk = gctx.upload_key(keys.relin_key())
    rng = np.random.default_rng(17)
    work = []
    for lvl in (3, 11):
        z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
        work.append((bp.program(lvl), keys.encrypt(C.encode(z, 2.0 ** prm.scale_bits, N), lvl)))
    coef = S.chebyshev_coefficients(lambda x: np.tanh(x / 4), 15, -8, 8)
    xz = rng.uniform(-8, 8, N // 2)
    work.append((S.chebyshev_program(prm, 12, coef, -8, 8), keys.encrypt(C.encode(xz, 2.0 ** prm.scale_bits, N), 12)))

    def requests():
        return [S.Request(i, P, 0.0, inputs=[ct], keep_result=True) for i, (P, ct) in enumerate(work)]

    eager_be = S.GpuBackend(gctx, rlk, dev_keys, 1 << 30)
    eager = requests()
    S.Server(eager_be, N).run(eager, realtime=False)
    want = [r.result for r in eager]

    stream = capi.Stream()
    be = S.GpuBackend(gctx, rlk, dev_keys, 1 << 30, stream=stream)
    srv = S.Server(be, N)
    cap = srv.capture(requests(), scratch_bytes=1 << 30, task_bytes=64 << 20)
    say("captured %d steps, %d launches" % (srv.stats["steps"], srv.stats["launches"]))
    other = S.rotsum_program(prm, 5, 3)
    ct_other = keys.encrypt(C.encode(rng.uniform(-1, 1, N // 2), 2.0 ** prm.scale_bits, N), 5)
    for rep in range(3):
        cap.replay()
        got = cap.results()
        for i, (g, w) in enumerate(zip(got, want)):
            assert np.array_equal(g, w), (rep, i)
        say("replay %d equals the eager loop" % rep)
        # eager work in between (default stream), then the next replay
        extra = S.Request(50 + rep, other, 0.0, inputs=[ct_other], keep_result=True)
        S.Server(eager_be, N).run([extra], realtime=False)
        gctx.sync()
    cap.close()
    stream.close()


def test_graph_refuses_other_stream():
    prm = params.serve_small()
    N = prm.N
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(3)
    l = 4
    a = capi.DeviceBuffer.from_array(np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64)
                                                         for q in prm.q[:l]]) for _ in range(2)]))
    o = capi.DeviceBuffer(2 * l * N)
    s1, s2 = capi.Stream(), capi.Stream()
    gctx.graph_begin(s1, 1 << 26, 1 << 20)
    with pytest.raises(capi.FidesError):
        gctx.add_batch([a], [a], [l], [o], stream=s2.h)
    gctx.add_batch([a], [a], [l], [o], stream=s1.h)
    g = gctx.graph_end()
    g.launch(s1)
    s1.sync()
    x = a.download((2, l, N))
    want = np.stack([np.stack([(x[c, j] * 2) % np.uint64(prm.q[j]) for j in range(l)]) for c in range(2)])
    assert np.array_equal(o.download((2, l, N)), want)
    g.close()
    s1.close()
    s2.close()
