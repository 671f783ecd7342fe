"""Sparse-slot bootstrapping on the GPU (slots < N/2: partial trace,
2 slots-periodic transforms, one EvalMod).

* N = 2^12 through the request loop: 64 slots (weight-64 secret) and 256 slots
  (dense secret, encapsulation, FLEXIBLEAUTO) equal the oracle running the
  same program, bit for bit, and decrypt to > 20 bits in every slot.
* The C++ API's EvalBootstrap with sparse slots: driven by the Python plan's
  exported tables it equals the oracle running that plan bit for bit; its own
  plan (BootstrapConfig.slots) asks for the same rotations and decrypts.
* N = 2^16, 4096 slots (h = 192, encapsulation, FLEXIBLEAUTO): bit-exact
  against the oracle and >= 20 bits in every slot.
"""
import time

import numpy as np
import pytest

import flex_oracle as F
from cpp_script import Script
from fideslib_amd import bootstrap as B
from fideslib_amd import capi, params
from fideslib_amd import serve as S
from oracle import ckks_codec as C
from oracle import oracle as O
from serve_oracle import run_oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def host_keys(bp, keys, N):
    rot = {}
    for r in bp.rotations() + [None]:
        k = pow(5, r, 2 * N) if r is not None else 2 * N - 1
        rot[k] = keys.rotation_key(k)
    if bp.cfg.encaps_h:
        rot[bp.ENCAPS_DOWN], rot[bp.ENCAPS_UP] = keys.encapsulation_keys(bp.cfg.encaps_h)
    return rot


def served(prm, gctx, bp, keys, work, arena):
    """Serve [(program, ciphertext)] together; returns the requests and the
    host keys the GPU used (every call of keys.relin_key() or
    keys.rotation_key() samples a new key, so the oracle must get these)."""
    rot = host_keys(bp, keys, prm.N)
    rlk = keys.relin_key()
    be = S.GpuBackend(gctx, gctx.upload_key(rlk), {k: gctx.upload_key(v) for k, v in rot.items()}, arena)
    reqs = [S.Request(i, P, 0.0, inputs=[ct], keep_result=True) for i, (P, ct) in enumerate(work)]
    S.Server(be, prm.N).run(reqs, realtime=False)
    return reqs, rot, rlk


@pytest.mark.parametrize("ns,encaps", [(64, False), (256, True)])
def test_sparse_small_equals_oracle(ns, encaps):
    prm = params.boot_small()
    N = prm.N
    say("sparse bootstrap N=2^12, %d slots, encapsulation %s" % (ns, encaps))
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 30 + ns, h=None if encaps else 64)
    cfg = B.BootConfig.encapsulated(flexible=True, slots=ns) if encaps else B.BootConfig(slots=ns)
    bp = B.BootstrapPlan(prm, gctx, cfg, keep_host=True)
    sf = params.flexible_scaling_factors(prm.q) if encaps else [2.0 ** prm.scale_bits] * prm.L
    rng = np.random.default_rng(ns)
    work, msgs = [], []
    for lvl in (3, 9):
        z = rng.uniform(-1, 1, ns) + 1j * rng.uniform(-1, 1, ns)
        work.append((bp.program(lvl, sf[lvl - 1] if encaps else None),
                     keys.encrypt(C.encode(z, sf[lvl - 1], N), lvl)))
        msgs.append(z)
    reqs, rot, rlk = served(prm, gctx, bp, keys, work, 1 << 30)
    for r, (P, ct), z in zip(reqs, work, msgs):
        assert r.status == "done"
        assert np.array_equal(r.result, run_oracle(octx, P, [ct], rlk, rot)), r.rid
        dec = C.decode(keys.decrypt_coef(r.result, P.output.level), P.output.scale, N)
        bits = C.precision_bits(dec, np.tile(z, N // 2 // ns))
        say("  from level %d: %.1f bits" % (P.inputs[0].level, bits))
        assert bits > 20, bits


@pytest.mark.parametrize("source", ["precomputed", "own_plan"])
def test_cpp_sparse_bootstrap(tmp_path, source):
    prm = params.boot_small()
    N, ns = prm.N, 128
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 77, h=None)
    cfg = B.BootConfig.encapsulated(flexible=True, slots=ns)
    bp = B.BootstrapPlan(prm, None, cfg, keep_host=True)
    rot = host_keys(bp, keys, N)
    rlk = keys.relin_key()
    sf = params.flexible_scaling_factors(prm.q)
    rng = np.random.default_rng(78)
    z = rng.uniform(-1, 1, ns) + 1j * rng.uniform(-1, 1, ns)
    lvl = 4
    ct = keys.encrypt(C.encode(z, sf[lvl - 1], N), lvl)
    s = Script(tmp_path, prm, F.FLEXIBLEAUTO)
    s.rlk(rlk)
    for k, v in rot.items():
        if isinstance(k, int):
            s.autokey(k, v)
    s.encapkeys(rot[bp.ENCAPS_DOWN], rot[bp.ENCAPS_UP])
    if source == "precomputed":
        s.bootsetup_pre(bp.export_precomputation(), cfg)
    else:
        s.op("bootsetup", 3, 3, cfg.degree, cfg.double_angle, cfg.K, cfg.correction, 1, ns)
    s.ct("x", ct, 1, sf[lvl - 1], ns)
    s.op("boot", "y", "x")
    # a second input at another level: EvalBootstrapMany gives each one's bits
    z2 = rng.uniform(-1, 1, ns) + 1j * rng.uniform(-1, 1, ns)
    ct2 = keys.encrypt(C.encode(z2, sf[5], N), 6)
    s.ct("x2", ct2, 1, sf[5], ns)
    s.op("boot", "y2", "x2")
    s.op("bootmany", "m", "x", "x2")
    s.store("y", "y2", "m0", "m1")
    res = s.run()
    assert np.array_equal(res["m0"]["data"], res["y"]["data"])
    assert np.array_equal(res["m1"]["data"], res["y2"]["data"])
    assert res["rots"] == bp.rotations()
    y = res["y"]
    lo = prm.L - bp.depth
    assert y["limbs"] == lo and y["sf"] == sf[lo - 1]
    if source == "precomputed":
        want = run_oracle(octx, bp.program(lvl, sf[lvl - 1]), [ct], rlk, rot)
        assert np.array_equal(y["data"], want)
    dec = C.decode(keys.decrypt_coef(y["data"], lo), y["sf"], N)
    bits = C.precision_bits(dec, np.tile(z, N // 2 // ns))
    say("C++ sparse EvalBootstrap (%s, %d slots): %.1f bits" % (source, ns, bits))
    assert bits > 20


def test_sparse_cfg4_shape_equals_oracle():
    prm = params.config3()
    N, ns = prm.N, 1 << 12
    t0 = time.time()
    say("sparse bootstrap N=2^16, %d slots: keys" % ns)
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 4193, h=192)
    bp = B.BootstrapPlan(prm, gctx, B.BootConfig.encapsulated(flexible=True, slots=ns), keep_host=True)
    sf = params.flexible_scaling_factors(prm.q)
    rng = np.random.default_rng(45)
    z = rng.uniform(-1, 1, ns) + 1j * rng.uniform(-1, 1, ns)
    lvl = 5
    work = [(bp.program(lvl, sf[lvl - 1]), keys.encrypt(C.encode(z, sf[lvl - 1], N), lvl))]
    reqs, rot, rlk = served(prm, gctx, bp, keys, work, 8 << 30)
    say("keys, plan and one bootstrap: %.1f s" % (time.time() - t0))
    P, ct = work[0]
    r = reqs[0]
    assert r.status == "done"
    t0 = time.time()
    assert np.array_equal(r.result, run_oracle(octx, P, [ct], rlk, rot))
    say("oracle: %.1f s" % (time.time() - t0))
    assert P.output.scale == sf[P.output.level - 1]
    dec = C.decode(keys.decrypt_coef(r.result, P.output.level), P.output.scale, N)
    bits = C.precision_bits(dec, np.tile(z, N // 2 // ns))
    say("N=2^16 sparse bootstrap (%d slots, h=192, encapsulation, FLEXIBLEAUTO): %.2f bits" % (ns, bits))
    assert bits >= 20.0, bits
