"""Bootstrapping as the headline configuration names it, on the GPU.

N = 2^16, 2^15 slots, 30 limbs (60 + 29 x 59 bits), K = 10, dnum = 3,
ternary secret of Hamming weight 192, sparse-secret encapsulation (weight 32),
FLEXIBLEAUTO (input and output at their levels' FLEXIBLEAUTO scaling
factors), level budget {3, 3}:
* one bootstrap through the request loop equals the oracle running the same
  program, bit for bit, and decrypts to >= 19 bits of precision;
* eight bootstraps captured in one HIP graph: two replays equal the eager
  request loop bit for bit.
And the host C++ API's EvalBootstrap (N = 2^12): driven by the Python plan's
exported tables it equals the oracle running that plan, bit for bit
(FIXEDMANUAL; FLEXIBLEAUTO with encapsulation); its own plan with
encapsulation under FLEXIBLEAUTO lands on the FLEXIBLEAUTO factor and
decrypts to > 20 bits; two iterations (numIterations = 2) gain at least 10
bits over one.
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


# ------------------------------------------------------------------ N = 2^16
@pytest.fixture(scope="module")
def cfg4():
    prm = params.config3()
    N = prm.N
    t0 = time.time()
    say("cfg4: keys")
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 4192, h=192)
    bp = B.BootstrapPlan(prm, gctx, B.BootConfig.encapsulated(flexible=True), keep_host=True)
    rot = host_keys(bp, keys, N)
    rlk = keys.relin_key()
    dev = {k: gctx.upload_key(v) for k, v in rot.items()}
    drlk = gctx.upload_key(rlk)
    say("cfg4: %d keys in %.1f s; plaintexts" % (len(rot), time.time() - t0))
    bp.build_plaintexts()
    sf = params.flexible_scaling_factors(prm.q)
    return prm, octx, gctx, keys, bp, rot, rlk, dev, drlk, sf


def test_cfg4_bootstrap_bit_exact_and_precise(cfg4):
    prm, octx, gctx, keys, bp, rot, rlk, dev, drlk, sf = cfg4
    N = prm.N
    assert bp.out_scale == sf[bp.output_level() - 1]
    rng = np.random.default_rng(41)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    lvl = 4
    ct = keys.encrypt(C.encode(z, sf[lvl - 1], N), lvl)
    P = bp.program(lvl, sf[lvl - 1])
    be = S.GpuBackend(gctx, drlk, dev, 8 << 30)
    req = S.Request(0, P, 0.0, inputs=[ct], keep_result=True)
    S.Server(be, N).run([req], realtime=False)
    assert req.status == "done"
    t0 = time.time()
    want = run_oracle(octx, P, [ct], rlk, rot)
    say("oracle bootstrap: %.1f s" % (time.time() - t0))
    assert np.array_equal(req.result, want)
    assert P.output.scale == sf[P.output.level - 1]
    dec = C.decode(keys.decrypt_coef(req.result, P.output.level), P.output.scale, N)
    bits = C.precision_bits(dec, z)
    say("N=2^16 bootstrap (h=192, encapsulation, FLEXIBLEAUTO): %.2f bits" % bits)
    assert bits >= 19.0, bits


def test_cfg4_eight_bootstraps_in_one_graph(cfg4):
    prm, octx, gctx, keys, bp, rot, rlk, dev, drlk, sf = cfg4
    N = prm.N
    rng = np.random.default_rng(42)
    work = []
    for lvl in (2, 3, 5, 8, 12, 17, 23, 30):
        z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
        work.append((bp.program(lvl, sf[lvl - 1]), keys.encrypt(C.encode(z, sf[lvl - 1], N), lvl), z))

    def requests():
        return [S.Request(i, P, 0.0, inputs=[ct], keep_result=True) for i, (P, ct, _) in enumerate(work)]
    say("eager loop: 8 bootstraps")
    eager_be = S.GpuBackend(gctx, drlk, dev, 48 << 30)
    eager = requests()
    S.Server(eager_be, N).run(eager, realtime=False)
    want = [r.result for r in eager]
    stream = capi.Stream()
    foot = S.Server(eager_be, N).footprint_bytes(work[-1][0])
    be = S.GpuBackend(gctx, drlk, dev, 8 * foot + (256 << 20), stream=stream)
    srv = S.Server(be, N)
    cap = srv.capture(requests(), scratch_bytes=40 << 30, task_bytes=1 << 30)
    say("captured 8 bootstraps: %d steps, %d launches" % (srv.stats["steps"], srv.stats["launches"]))
    for rep in range(2):
        cap.replay()
        got = cap.results()
        for i, (g, w) in enumerate(zip(got, want)):
            assert np.array_equal(g, w), (rep, i)
        say("replay %d equals the eager loop" % rep)
    cap.close()
    stream.close()
    P, _, z = work[5]
    dec = C.decode(keys.decrypt_coef(want[5], P.output.level), P.output.scale, N)
    assert C.precision_bits(dec, z) >= 19.0


# ------------------------------------------------------------------ C++ EvalBootstrap (N = 2^12)
@pytest.mark.parametrize("variant", ["fixed", "flexible_encapsulated"])
def test_cpp_bootstrap_equals_the_plan_on_the_oracle(tmp_path, variant):
    prm = params.boot_small()
    N = prm.N
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    flex = variant != "fixed"
    keys = C.Keys(octx, 7, h=None if flex else 64)
    cfg = B.BootConfig.encapsulated(flexible=True) if flex else B.BootConfig()
    bp = B.BootstrapPlan(prm, None, cfg, keep_host=True)
    rot = host_keys(bp, keys, N)
    rlk = keys.relin_key()
    sf = params.flexible_scaling_factors(prm.q) if flex else [2.0 ** prm.scale_bits] * prm.L
    rng = np.random.default_rng(43)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    lvl = 6
    ct = keys.encrypt(C.encode(z, sf[lvl - 1], N), lvl)
    s = Script(tmp_path, prm, F.FLEXIBLEAUTO if flex else F.FIXEDMANUAL)
    s.rlk(rlk)
    for k, v in rot.items():
        if isinstance(k, int):
            s.autokey(k, v)
    if flex:
        s.encapkeys(rot[bp.ENCAPS_DOWN], rot[bp.ENCAPS_UP])
    s.bootsetup_pre(bp.export_precomputation(), cfg)
    s.ct("x", ct, 1, sf[lvl - 1], N // 2)
    s.op("boot", "y", "x")
    s.store("y")
    res = s.run()
    assert res["rots"] == bp.rotations()
    P = bp.program(lvl, sf[lvl - 1])
    want = run_oracle(octx, P, [ct], rlk, rot)
    y = res["y"]
    assert y["limbs"] == P.output.level and y["deg"] == 1
    assert abs(y["sf"] - P.output.scale) <= 1e-12 * P.output.scale
    assert np.array_equal(y["data"], want)
    dec = C.decode(keys.decrypt_coef(want, P.output.level), P.output.scale, N)
    bits = C.precision_bits(dec, z)
    say("C++ EvalBootstrap (%s) = oracle; %.1f bits" % (variant, bits))
    assert bits > 20


def test_cpp_bootstrap_own_plan_flexible_encapsulated(tmp_path):
    prm = params.boot_small()
    N = prm.N
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 9, h=None)
    cfg = B.BootConfig.encapsulated(flexible=True)
    bp = B.BootstrapPlan(prm, None, cfg)  # the same rotation set as the C++ plan
    rot = host_keys(bp, keys, N)
    sf = params.flexible_scaling_factors(prm.q)
    rng = np.random.default_rng(44)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    lvl = 3
    ct = keys.encrypt(C.encode(z, sf[lvl - 1], N), lvl)
    s = Script(tmp_path, prm, F.FLEXIBLEAUTO)
    s.rlk(keys.relin_key())
    for k, v in rot.items():
        if isinstance(k, int):
            s.autokey(k, v)
    s.encapkeys(rot[bp.ENCAPS_DOWN], rot[bp.ENCAPS_UP])
    s.op("bootsetup", 3, 3, cfg.degree, cfg.double_angle, cfg.K, cfg.correction, 1)
    s.ct("x", ct, 1, sf[lvl - 1], N // 2)
    s.op("boot", "y", "x")
    s.store("y")
    res = s.run()
    assert res["rots"] == bp.rotations()
    y = res["y"]
    lo = prm.L - bp.depth
    assert y["limbs"] == lo and y["sf"] == sf[lo - 1]
    dec = C.decode(keys.decrypt_coef(y["data"], lo), y["sf"], N)
    bits = C.precision_bits(dec, z)
    say("C++ EvalBootstrap, own plan, FLEXIBLEAUTO + encapsulation: %.1f bits" % bits)
    assert bits > 20


@pytest.mark.parametrize("variant", ["fixed", "flexible_encapsulated"])
def test_cpp_bootstrap_many_equals_single(tmp_path, variant):
    """EvalBootstrapMany over inputs at different levels (and, under
    FLEXIBLEAUTO, different scaling factors) gives each input's
    EvalBootstrap result bit for bit; one of them equals the oracle running
    the plan.  The same script times n single bootstraps against one batch."""
    prm = params.boot_small()
    N = prm.N
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    flex = variant != "fixed"
    keys = C.Keys(octx, 13, h=None if flex else 64)
    cfg = B.BootConfig.encapsulated(flexible=True) if flex else B.BootConfig()
    bp = B.BootstrapPlan(prm, None, cfg, keep_host=True)
    rot = host_keys(bp, keys, N)
    sf = params.flexible_scaling_factors(prm.q) if flex else [2.0 ** prm.scale_bits] * prm.L
    rng = np.random.default_rng(47)
    lvls = [6, 3, 5]
    zs = [rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2) for _ in lvls]
    cts = [keys.encrypt(C.encode(z, sf[l - 1], N), l) for z, l in zip(zs, lvls)]
    rlk = keys.relin_key()
    s = Script(tmp_path, prm, F.FLEXIBLEAUTO if flex else F.FIXEDMANUAL)
    s.rlk(rlk)
    for k, v in rot.items():
        if isinstance(k, int):
            s.autokey(k, v)
    if flex:
        s.encapkeys(rot[bp.ENCAPS_DOWN], rot[bp.ENCAPS_UP])
    s.bootsetup_pre(bp.export_precomputation(), cfg)
    for i, (ct, l) in enumerate(zip(cts, lvls)):
        s.ct("x%d" % i, ct, 1, sf[l - 1], N // 2)
        s.op("boot", "y%d" % i, "x%d" % i)
    s.op("releasemem")  # idle cached blocks back to the device; the next calls allocate again
    s.op("bootmany", "m", *["x%d" % i for i in range(len(lvls))])
    s.op("boottime", 2, 8, "x0")
    s.store(*["y%d" % i for i in range(len(lvls))], *["m%d" % i for i in range(len(lvls))])
    res = s.run()
    for i in range(len(lvls)):
        y, m = res["y%d" % i], res["m%d" % i]
        assert m["limbs"] == y["limbs"] and m["sf"] == y["sf"] and m["deg"] == y["deg"]
        assert np.array_equal(m["data"], y["data"]), i
    P = bp.program(lvls[0], sf[lvls[0] - 1])
    want = run_oracle(octx, P, [cts[0]], rlk, rot)
    assert np.array_equal(res["m0"]["data"], want)
    n, one, many = res["boottime"][0]
    say("C++ bootstrap (%s), N=2^%d: %d one by one %.1f ms, EvalBootstrapMany %.1f ms (%.2fx)"
        % (variant, prm.logN, n, one, many, one / many))


def test_cpp_bootstrap_many_rate_cfg4(tmp_path):
    """The C++ API at the cfg4 shape (N = 2^16, 30 limbs, dnum 3, FLEXIBLEAUTO,
    encapsulation, the library's own plan): 8 bootstraps of a 3-limb input
    one by one and in one EvalBootstrapMany, with synthetic keys and
    residues (the rate does not depend on their values).  Both must finish;
    the rates are reported."""
    prm = params.config3()
    N = prm.N
    cfg = B.BootConfig.encapsulated(flexible=True)
    sf = params.flexible_scaling_factors(prm.q)
    s = Script(tmp_path, prm, F.FLEXIBLEAUTO)
    s.op("bootsetup", 3, 3, cfg.degree, cfg.double_angle, cfg.K, cfg.correction, 1)
    s.op("synthkeys", 1)
    s.op("synthct", "x", 3, 5, sf[2], N // 2)
    s.op("boottime", 1, 8, "x")
    res = s.run(timeout=900)
    n, one, many = res["boottime"][0]
    say("C++ bootstrap cfg4 (N=2^16, FLEXIBLEAUTO + encapsulation): %d one by one %.1f ms (%.1f/s), "
        "EvalBootstrapMany %.1f ms (%.1f/s)" % (n, one, 1e3 * n / one, many, 1e3 * n / many))
    assert many > 0 and one > 0


@pytest.mark.parametrize("variant", ["fixed", "flexible_encapsulated"])
def test_cpp_bootstrap_two_iterations(tmp_path, variant):
    """EvalBootstrap(ct, 2, p): the error 2^p (x - BTS(x)) is bootstrapped
    and added back; the result carries the output factor times 2^p."""
    prm = params.boot_small()
    N = prm.N
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    flex = variant != "fixed"
    keys = C.Keys(octx, 11, h=None if flex else 64)
    cfg = B.BootConfig.encapsulated(flexible=True) if flex else B.BootConfig()
    bp = B.BootstrapPlan(prm, None, cfg)  # the rotation set of the C++ plan
    rot = host_keys(bp, keys, N)
    sf = params.flexible_scaling_factors(prm.q) if flex else [2.0 ** prm.scale_bits] * prm.L
    rng = np.random.default_rng(46)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    lvl, p = 3, 18
    ct = keys.encrypt(C.encode(z, sf[lvl - 1], N), lvl)
    s = Script(tmp_path, prm, F.FLEXIBLEAUTO if flex else F.FIXEDMANUAL)
    s.rlk(keys.relin_key())
    for k, v in rot.items():
        if isinstance(k, int):
            s.autokey(k, v)
    if flex:
        s.encapkeys(rot[bp.ENCAPS_DOWN], rot[bp.ENCAPS_UP])
    s.op("bootsetup", 3, 3, cfg.degree, cfg.double_angle, cfg.K, cfg.correction, 1 if flex else 0)
    s.ct("x", ct, 1, sf[lvl - 1], N // 2)
    s.op("boot", "y1", "x")
    s.op("boot", "y2", "x", 2, p)
    s.store("y1", "y2")
    res = s.run()
    bits = []
    for name in ("y1", "y2"):
        y = res[name]
        bits.append(C.precision_bits(C.decode(keys.decrypt_coef(y["data"], y["limbs"]), y["sf"], N), z))
    lo = prm.L - bp.depth
    assert res["y2"]["limbs"] == lo
    assert res["y2"]["sf"] == res["y1"]["sf"] * 2.0 ** p
    say("C++ EvalBootstrap (%s): one iteration %.1f bits, two iterations %.1f bits" % (variant, bits[0], bits[1]))
    assert bits[1] >= bits[0] + 10, bits
