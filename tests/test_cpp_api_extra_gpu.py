"""The host C++ API's level/scale management and the rest of its hot-path
surface (api/CryptoContext.hpp), driven through tests/cpp/api_script and
checked limb for limb against the oracle:

- FIXEDAUTO / FLEXIBLEAUTO alignment of EvalAdd/EvalSub/EvalMult on operands
  at different levels and scale degrees, scalar ops and SetLevel, against
  tests/flex_oracle.py (the reference's adjustScaleAndLevel family restated
  over the oracle's exact operations);
- EvalAdd/EvalSub with plaintexts, EvalAddMany, the Mutable aliases;
- EvalFastRotationExt and AccumulateSum against tests/ext_oracle.py;
- EvalChebyshevSeries against the oracle running the same program
  (fideslib_amd/serve.py chebyshev_program), GetChebyshevCoefficients
  against the reference's interpolation formula.
Small parameters (N = 2^12, 12 limbs; N = 2^13, 4 limbs) and N = 2^16 with
30 limbs."""
import math

import numpy as np
import pytest

import ext_oracle as X
import flex_oracle as F
import serve_oracle as SO
from cpp_script import Script
from fideslib_amd import params
from fideslib_amd import serve as S
from oracle import ckks_codec as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def same(res, want: F.Ct, what):
    assert res["limbs"] == want.limbs, (what, res["limbs"], want.limbs)
    assert res["deg"] == want.deg, (what, res["deg"], want.deg)
    assert abs(res["sf"] - want.sf) <= 1e-12 * abs(want.sf), (what, res["sf"], want.sf)
    assert np.array_equal(res["data"], want.data), what


@pytest.fixture(scope="module")
def small():
    prm = params.serve_small()  # N = 2^12, 60 + 11 x 50 bits, dnum 3
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 91)
    return prm, octx, keys, keys.relin_key()


def _inputs(prm, octx, keys, rlk, fl, rng, spec):
    """Ciphertexts named by spec {name: (limbs, deg)}: fresh encryptions at the
    level's scaling factor (degree 1) or products of two (degree 2)."""
    N = prm.N
    out = {}
    for name, (l, deg) in spec.items():
        # FLEXIBLEAUTOEXT: a fresh ciphertext at the top level carries the big
        # factor (the level ElemForEvalAddOrSub treats apart, Context.cu:477-481)
        ext_top = fl.tech == F.FLEXIBLEAUTOEXT and l == fl.L
        sf = fl.sfbig[l - 1] if ext_top else fl.sf[l - 1]
        z = [rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2) for _ in range(deg)]
        if ext_top and deg == 2:
            # sfBig = sf^2 is already the degree-2 factor of that level: the
            # message encrypted there directly (a product of two would carry sf^4)
            zz = z[0] * z[1]
            out[name] = (F.Ct(keys.encrypt(C.encode(zz, sf, N), l), 2, sf), zz)
            continue
        cts = [keys.encrypt(C.encode(v, sf, N), l) for v in z]
        if deg == 1:
            out[name] = (F.Ct(cts[0], 1, sf), z[0])
        else:
            out[name] = (F.Ct(octx.hmult_relin(cts[0], cts[1], l, rlk), 2, sf * sf), z[0] * z[1])
    return out


@pytest.mark.parametrize("tech", [F.FLEXIBLEAUTO, F.FIXEDAUTO, F.FLEXIBLEAUTOEXT])
def test_alignment_against_the_restated_rules(tmp_path, small, tech):
    prm, octx, keys, rlk = small
    say("alignment technique %d" % tech)
    fl = F.Flex(octx, tech, rlk)
    rng = np.random.default_rng(100 + tech)
    spec = {"a12": (12, 1), "b9": (9, 1), "c11": (11, 1), "d12": (12, 2), "e9": (9, 2), "f11": (11, 1),
            "g8": (8, 1), "h10": (10, 1), "i10": (10, 2), "j7": (7, 2)}
    inp = _inputs(prm, octx, keys, rlk, fl, rng, spec)
    s = Script(tmp_path, prm, tech)
    s.rlk(rlk)
    for n, (c, _) in inp.items():
        s.ct(n, c.data, c.deg, c.sf, prm.N // 2)
    pairs = [("a12", "b9"), ("a12", "c11"), ("d12", "e9"), ("d12", "f11"), ("d12", "g8"), ("a12", "e9"),
             ("h10", "i10"), ("b9", "a12"), ("e9", "d12"), ("i10", "h10"), ("j7", "a12"), ("g8", "d12")]
    want = {}
    for i, (x, y) in enumerate(pairs):
        A, B = inp[x][0], inp[y][0]
        s.op("add", "add%d" % i, x, y)
        want["add%d" % i] = (fl.add(A, B), inp[x][1] + inp[y][1])
        s.op("sub", "sub%d" % i, x, y)
        want["sub%d" % i] = (fl.add(A, B, sub=True), inp[x][1] - inp[y][1])
        s.op("mul", "mul%d" % i, x, y)
        want["mul%d" % i] = (fl.mult(A, B), inp[x][1] * inp[y][1])
    for n, c in (("a12", 0.75), ("d12", -1.25), ("j7", 3.0)):
        A = inp[n][0]
        s.op("adds", "adds_" + n, n, c)
        want["adds_" + n] = (fl.add_scalar(A, c), inp[n][1] + c)
        s.op("muls", "muls_" + n, n, c)
        want["muls_" + n] = (fl.mult_scalar(A, c), inp[n][1] * c)
    for n, lev in (("a12", 3), ("d12", 5), ("b9", 4)):
        s.op("setlevel", "lev_%s_%d" % (n, lev), n, lev)
        want["lev_%s_%d" % (n, lev)] = (fl.set_level(inp[n][0], lev), inp[n][1])
    s.store(*want)
    res = s.run()
    assert np.allclose(res["sf"], fl.sf, rtol=1e-15)
    for name, (w, z) in want.items():
        same(res[name], w, name)
        dec = C.decode(keys.decrypt_coef(res[name]["data"], w.limbs), w.sf, prm.N)
        assert C.precision_bits(dec, z) > 12, name


def test_plaintext_add_sub(tmp_path, small):
    """EvalAdd/EvalSub(ct, pt) and pt - ct: FIXEDMANUAL (plaintext with more
    limbs, its first limbs used), FIXEDAUTO and FLEXIBLEAUTO (plaintext two
    levels up: dropped, resp. brought down by adjustScaleAndLevel)."""
    prm, octx, keys, rlk = small
    N = prm.N
    rng = np.random.default_rng(110)
    for tech in (F.FIXEDMANUAL, F.FIXEDAUTO, F.FLEXIBLEAUTO):
        say("plaintext add technique %d" % tech)
        fl = F.Flex(octx, tech, rlk)
        l, lp = 8, 10
        z, zp = [rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2) for _ in range(2)]
        sfc = fl.sf[l - 1]
        sfp = fl.sf[lp - 1] if tech != F.FIXEDMANUAL else sfc
        ct = F.Ct(keys.encrypt(C.encode(z, sfc, N), l), 1, sfc)
        coef = C.to_rns(C.encode(zp, sfp, N), prm.q[:lp])
        pe = octx.ntt_limbs(np.ascontiguousarray(coef), list(range(lp)))
        P = F.Ct(np.stack([pe, np.zeros_like(pe)]), 1, sfp)
        s = Script(tmp_path, prm, tech)
        s.ct("c", ct.data, 1, sfc, N // 2)
        s.pt("p", coef, 0, 1, sfp)
        s.op("addpt", "r_add", "c", "p")
        s.op("subpt", "r_sub", "c", "p")
        s.op("ptsub", "r_psub", "p", "c")
        s.store("r_add", "r_sub", "r_psub")
        res = s.run()
        same(res["r_add"], fl.add_pt(ct, P), "add %d" % tech)
        same(res["r_sub"], fl.add_pt(ct, P, sub=True), "sub %d" % tech)
        neg = F.Ct(fl._mul_res(ct, [q - 1 for q in prm.q[:l]]), 1, sfc)
        same(res["r_psub"], fl.add_pt(neg, P), "psub %d" % tech)
        for name, want in (("r_add", z + zp), ("r_sub", z - zp), ("r_psub", zp - z)):
            dec = C.decode(keys.decrypt_coef(res[name]["data"], res[name]["limbs"]), res[name]["sf"], N)
            assert C.precision_bits(dec, want) > 12, (tech, name)


def test_add_many_mutable_and_ext_rotation(tmp_path, small):
    """EvalAddMany at one level (fused sums) and at mixed levels under
    FLEXIBLEAUTO (the reference's pairwise tree of EvalAdd); EvalMultMutable /
    EvalSquareMutable equal EvalMult / EvalSquare; EvalFastRotationExt then
    its ModDown equal the oracle composition."""
    prm, octx, keys, rlk = small
    N = prm.N
    rng = np.random.default_rng(120)
    fl = F.Flex(octx, F.FLEXIBLEAUTO, rlk)
    spec = {"u%d" % i: (9, 1) for i in range(19)}
    spec.update({"m0": (12, 1), "m1": (9, 2), "m2": (11, 1), "m3": (7, 1)})
    inp = _inputs(prm, octx, keys, rlk, fl, rng, spec)
    s = Script(tmp_path, prm, F.FLEXIBLEAUTO)
    s.rlk(rlk)
    ks = [pow(5, 3, 2 * N), pow(5, 17, 2 * N)]
    hk = {k: keys.rotation_key(k) for k in ks}
    for k, v in hk.items():
        s.autokey(k, v)
    for n, (c, _) in inp.items():
        s.ct(n, c.data, c.deg, c.sf, N // 2)
    uni = ["u%d" % i for i in range(19)]
    s.op("addmany", "sum_u", *uni)
    s.op("addmany", "sum_m", "m0", "m1", "m2", "m3")
    s.op("mul", "mul", "m0", "m2")
    s.op("mulmut", "mulmut", "m0", "m2")
    s.op("sq", "sq", "m2")
    s.op("sqmut", "sqmut", "m2")
    s.op("fastrotext", "rx3", "u0", 3, 1)
    s.op("fastrotext", "rx17", "u0", 17, 1)
    s.op("add", "rxsum", "rx3", "rx17")
    s.store("sum_u", "sum_m", "mul", "mulmut", "sq", "sqmut", "rx3", "rxsum")
    res = s.run()
    acc = inp[uni[0]][0]
    for n in uni[1:]:
        acc = fl.add(acc, inp[n][0])
    same(res["sum_u"], acc, "sum_u")
    c = [inp["m%d" % i][0] for i in range(4)]
    t = [fl.add(c[0], c[1]), fl.add(c[2], c[3])]
    same(res["sum_m"], fl.add(t[0], t[1]), "sum_m")
    assert np.array_equal(res["mul"]["data"], res["mulmut"]["data"])
    assert np.array_equal(res["sq"]["data"], res["sqmut"]["data"])
    u0 = inp["u0"][0]
    l = u0.limbs
    r3 = X.rotate_ext(octx, u0.data, l, hk[ks[0]], ks[0])
    r17 = X.rotate_ext(octx, u0.data, l, hk[ks[1]], ks[1])
    assert res["rx3"]["ext"] == 1  # the register stays in Q_l u P; StoreCiphertext takes a copy back
    assert np.array_equal(res["rx3"]["data"], X.moddown_ext(octx, r3, l))
    assert np.array_equal(res["rxsum"]["data"], X.moddown_ext(octx, X.add_ext(octx, r3, r17, l), l))
    dec = C.decode(keys.decrypt_coef(res["rxsum"]["data"], l), u0.sf, N)
    z = inp["u0"][1]
    assert C.precision_bits(dec, np.roll(z, -3) + np.roll(z, -17)) > 12


@pytest.mark.parametrize("slots,stride,start", [(16, 1, 1), (32, 2, 1), (64, 1, 4)])
def test_accumulate_sum(tmp_path, small, slots, stride, start):
    prm, octx, keys, rlk = small
    N, l = prm.N, prm.L
    rng = np.random.default_rng(130 + slots)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    sc = 2.0 ** prm.scale_bits
    ct = keys.encrypt(C.encode(z, sc, N), l)
    rots = set()
    s_ = start
    while s_ < slots:
        rots |= {j * s_ * stride for j in (1, 2, 3) if j * s_ * stride < stride * slots}
        s_ *= 4
    hk = {pow(5, r, 2 * N): keys.rotation_key(pow(5, r, 2 * N)) for r in sorted(rots)}
    s = Script(tmp_path, prm, F.FIXEDMANUAL)
    s.rlk(rlk)
    for k, v in hk.items():
        s.autokey(k, v)
    s.ct("x", ct, 1, sc, N // 2)
    if start == 1:
        s.op("accum", "acc", "x", slots, stride)
    else:
        s.op("accum", "acc", "x", slots, stride, start)
    s.store("acc")
    res = s.run()
    want = X.accumulate(octx, ct, l, hk, slots, stride, N, start=start)
    assert np.array_equal(res["acc"]["data"], want)
    dec = C.decode(keys.decrypt_coef(want, l), sc, N)
    ref = sum(np.roll(z, -j * stride) for j in range(0, slots, start))
    assert C.precision_bits(dec, ref) > 10


def _cheb_check(tmp_path, prm, octx, keys, rlk, level, coeffs, lo, hi, technique=F.FIXEDMANUAL):
    N = prm.N
    rng = np.random.default_rng(140 + level)
    x = rng.uniform(lo, hi, N // 2)
    sc = 2.0 ** prm.scale_bits
    ct = keys.encrypt(C.encode(x.astype(complex), sc, N), level)
    s = Script(tmp_path, prm, technique)
    s.rlk(rlk)
    s.ct("x", ct, 1, sc, N // 2)
    s.op("cheb", "y", "x", float(lo), float(hi), *[float(c) for c in coeffs])
    s.store("y")
    res = s.run()
    # the same program on the oracle (OpenFHE's c_0 / 2 convention)
    cs = [coeffs[0] / 2] + list(coeffs[1:])
    P = S.chebyshev_program(prm, level, cs, lo, hi)  # its input scale is 2^scale_bits = sc
    want = SO.run_oracle(octx, P, [ct], rlk, {})
    assert np.array_equal(res["y"]["data"], want)
    assert res["y"]["limbs"] == P.output.level and res["y"]["deg"] == 1
    assert abs(res["y"]["sf"] - P.output.scale) <= 1e-12 * P.output.scale
    t = (2 * x - lo - hi) / (hi - lo)
    ref = np.polynomial.chebyshev.chebval(t, cs)
    dec = C.decode(keys.decrypt_coef(want, P.output.level), P.output.scale, N).real
    return C.precision_bits(dec, ref)


def test_chebyshev_series_small(tmp_path, small):
    prm, octx, keys, rlk = small
    f = lambda v: 1.0 / (1.0 + math.exp(-v))
    c = S.chebyshev_coefficients(f, 15, -8.0, 8.0)
    coeffs = [2 * c[0]] + list(c[1:])  # OpenFHE convention: c_0 not halved
    bits = _cheb_check(tmp_path, prm, octx, keys, rlk, 12, coeffs, -8.0, 8.0)
    say("Chebyshev degree 15: %.1f bits" % bits)
    assert bits > 18


def test_chebyshev_series_cfg1(tmp_path):
    """N = 2^13, 4 limbs: degree 5 on [-1, 1] (no input map)."""
    prm = params.config1()
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 92)
    coeffs = [0.5, 0.25, -0.125, 0.0625, 0.2, -0.1]
    bits = _cheb_check(tmp_path, prm, octx, keys, keys.relin_key(), 4, coeffs, -1.0, 1.0)
    assert bits > 15


def test_get_chebyshev_coefficients(tmp_path, small):
    prm = small[0]
    s = Script(tmp_path, prm, 0)
    s.op("chebcoef", "ce", "exp", -2.0, 3.0, 12)
    s.op("chebcoef", "cs", "sigmoid", -8.0, 8.0, 15)
    s.run()
    for name, f, a, b, d in (("ce", math.exp, -2.0, 3.0, 12), ("cs", lambda v: 1 / (1 + math.exp(-v)), -8.0, 8.0, 15)):
        got = np.array(s.read_text(name))
        # reference formula (src/PolyApprox.cu:11-32): nodes cos(pi (i+1/2)/n) mapped to [a, b], c_0 not halved
        n = d + 1
        fx = np.array([f(math.cos(math.pi / n * (i + 0.5)) * 0.5 * (b - a) + 0.5 * (b + a)) for i in range(n)])
        want = np.array([2.0 / n * sum(fx[j] * math.cos(math.pi / n * i * (j + 0.5)) for j in range(n))
                         for i in range(n)])
        assert np.allclose(got, want, rtol=0, atol=1e-13), name
        assert np.allclose(got[1:], S.chebyshev_coefficients(f, d, a, b)[1:], atol=1e-13)


def test_cfg3_api_surface(tmp_path):
    """N = 2^16, 30 limbs, K = 10, dnum 3: EvalChebyshevSeries (degree 15 on
    [-8, 8]), AccumulateSum over 16 slots, EvalAdd(ct, pt) and EvalAddMany,
    each against the oracle."""
    prm = params.config3()
    N, L = prm.N, prm.L
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 93)
    rlk = keys.relin_key()
    f = lambda v: 1.0 / (1.0 + math.exp(-v))
    c = S.chebyshev_coefficients(f, 15, -8.0, 8.0)
    say("cfg3 chebyshev")
    bits = _cheb_check(tmp_path, prm, octx, keys, rlk, 30, [2 * c[0]] + list(c[1:]), -8.0, 8.0)
    assert bits > 25
    say("cfg3 accumulate + plaintext add + add many")
    rng = np.random.default_rng(150)
    sc = 2.0 ** prm.scale_bits
    z = [rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2) for _ in range(4)]
    cts = [keys.encrypt(C.encode(v, sc, N), L) for v in z]
    hk = {pow(5, r, 2 * N): keys.rotation_key(pow(5, r, 2 * N)) for r in (1, 2, 3, 4, 8, 12)}
    coef = C.to_rns(C.encode(z[3], sc, N), prm.q)
    s = Script(tmp_path, prm, F.FIXEDMANUAL)
    s.rlk(rlk)
    for k, v in hk.items():
        s.autokey(k, v)
    for i in range(3):
        s.ct("c%d" % i, cts[i], 1, sc, N // 2)
    s.pt("p", coef, 0, 1, sc)
    s.op("accum", "acc", "c0", 16, 1)
    s.op("addpt", "ap", "c1", "p")
    s.op("addmany", "am", "c0", "c1", "c2")
    s.store("acc", "ap", "am")
    res = s.run()
    assert np.array_equal(res["acc"]["data"], X.accumulate(octx, cts[0], L, hk, 16, 1, N))
    fl = F.Flex(octx, F.FIXEDMANUAL, rlk)
    pe = octx.ntt_limbs(np.ascontiguousarray(coef), list(range(L)))
    same(res["ap"], fl.add_pt(F.Ct(cts[1], 1, sc), F.Ct(np.stack([pe, np.zeros_like(pe)]), 1, sc)), "ap")
    pid = list(range(L)) * 2
    am = octx.add_limbs(octx.add_limbs(cts[0].reshape(2 * L, N), cts[1].reshape(2 * L, N), pid),
                        cts[2].reshape(2 * L, N), pid)
    assert np.array_equal(res["am"]["data"], am.reshape(2, L, N))
