"""Pins the CPU oracle (oracle/ckks_oracle.c) independently of itself.

Each check is chosen so that a plausible mistake (dropped term, wrong
sign or index, transposed operand, off-by-one digit boundary, wrong key
component) fails it:
  - golden vectors from tests/golden/make_golden.py (pure-Python big-integer
    implementation written from the definitions, O(N^2) NTT);
  - closed forms: sympy's smallest primitive root, psi^N = -1, INTT o NTT = id,
    negacyclic convolution theorem vs schoolbook product;
  - invariants: ModUp output = x + u*Qhat with 0 <= u < #digit primes,
    ModDown(P*c) = c, key switch with the trivial key returns the input,
    decryption of HMult / Rescale / Rotate against an FP64 computation.
"""
import json
import os

import numpy as np
import pytest
import sympy

from oracle import ckks_codec as C
from oracle import oracle as O
from fideslib_amd import params as P

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden_small.json")))["cases"]
PRIMES = json.load(open(os.path.join(HERE, "golden", "reference_primes.json")))


def brev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


# ------------------------------------------------------------ scalar level
def test_modular_scalars():
    rng = np.random.default_rng(1)
    for q in PRIMES["p64"][:4] + PRIMES["sp64"][:2]:
        for a, b in [(0, 0), (1, q - 1), (q - 1, q - 1)] + [tuple(int(v) for v in rng.integers(0, q, 2, dtype=np.uint64)) for _ in range(50)]:
            assert O.mulmod(a, b, q) == a * b % q
        a = int(rng.integers(1, q, dtype=np.uint64))
        assert O.invmod(a, q) == pow(a, -1, q)
        assert O.powmod(a, 12345, q) == pow(a, 12345, q)


def test_generator_and_root_match_smallest_primitive_root():
    for q in PRIMES["p64"][:3] + PRIMES["sp64"][:2]:
        g = sympy.primitive_root(q)  # smallest primitive root
        assert O.lib().or_generator(q) == g
        for logN in (4, 13, 16):
            psi = O.root_2n(q, logN)
            assert psi == pow(g, (q - 1) >> (logN + 1), q)
            assert pow(psi, 1 << logN, q) == q - 1


def test_psi_tables_are_bit_reversed_powers():
    q = PRIMES["p64"][1]
    logN = 6
    psi = O.root_2n(q, logN)
    t, it = O.psi_tables(q, psi, logN)
    for j in range(1 << logN):
        assert int(t[j]) == pow(psi, brev(j, logN), q)
        assert int(it[j]) == pow(psi, -brev(j, logN), q)


# ------------------------------------------------------------------ NTT
@pytest.mark.parametrize("case", range(6))
def test_ntt_golden_definition(case):
    c = GOLD["ntt"][case]
    logN, q, psi = c["logN"], c["q"], c["psi"]
    assert O.root_2n(q, logN) == psi
    t, it = O.psi_tables(q, psi, logN)
    got = O.ntt(np.array(c["a"], np.uint64), logN, q, t)
    assert [int(x) for x in got] == c["ntt"]
    back = O.intt(got, logN, q, it)
    assert [int(x) for x in back] == c["a"]


def test_ntt_roundtrip_and_convolution_theorem():
    rng = np.random.default_rng(7)
    logN = 6
    N = 1 << logN
    for q in (PRIMES["p64"][0], PRIMES["p64"][2], PRIMES["sp64"][3]):
        psi = O.root_2n(q, logN)
        t, it = O.psi_tables(q, psi, logN)
        a = rng.integers(0, q, N, dtype=np.uint64)
        b = rng.integers(0, q, N, dtype=np.uint64)
        assert np.array_equal(O.intt(O.ntt(a, logN, q, t), logN, q, it), a)
        A = O.ntt(a, logN, q, t)
        B = O.ntt(b, logN, q, t)
        AB = np.array([(int(x) * int(y)) % q for x, y in zip(A, B)], np.uint64)
        got = O.intt(AB, logN, q, it)
        al, bl = [int(x) for x in a], [int(x) for x in b]
        want = [0] * N
        for i in range(N):
            for j in range(N):
                k = i + j
                if k < N:
                    want[k] += al[i] * bl[j]
                else:
                    want[k - N] -= al[i] * bl[j]
        assert [int(x) for x in got] == [w % q for w in want]


# ---------------------------------------------------------- conversions
def test_switch_basis_golden():
    c = GOLD["switch_basis"]
    src, dst = c["src"], c["dst"]
    ctx = O.Context(3, src, dst, 1)
    got = ctx.switch_basis(np.array(c["x"], np.uint64), range(len(src)), len(src) + np.arange(len(dst)))
    assert got.tolist() == c["y"]


def test_switch_basis_crt_invariant():
    """y = x + u * Qhat (mod p) with 0 <= u < #src, x the CRT value in [0, Qhat)."""
    rng = np.random.default_rng(3)
    src = PRIMES["p64"][1:6]
    dst = PRIMES["sp64"][:3] + PRIMES["p64"][10:12]
    ctx = O.Context(4, src, dst, 1)
    N = 16
    xs = np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in src])
    y = ctx.switch_basis(xs, range(5), 5 + np.arange(5))
    Qh = 1
    for q in src:
        Qh *= q
    for k in range(N):
        x = 0
        for i, q in enumerate(src):
            qh = Qh // q
            x += int(xs[i, k]) * qh * pow(qh, -1, q)
        x %= Qh
        for t, p in enumerate(dst):
            ok = [u for u in range(len(src)) if (x + u * Qh) % p == int(y[t, k])]
            assert ok, (k, t)


def _eval16_ctx():
    c = GOLD["eval16"]
    ctx = O.Context(c["logN"], c["Q"], c["P"], c["dnum"], roots=c["roots"])
    return c, ctx


def test_modup_golden():
    c, ctx = _eval16_ctx()
    x = np.array(c["x_eval"], np.uint64)
    for j in range(2):
        got = ctx.modup_digit(x, c["l"], j)
        assert got.tolist() == c["modup"][j], j


def test_moddown_golden():
    c, ctx = _eval16_ctx()
    got = ctx.moddown(np.array(c["y_eval"], np.uint64), c["l"])
    assert got.tolist() == c["moddown"]


def test_moddown_of_p_times_c_is_c():
    rng = np.random.default_rng(11)
    prm = P.config1()
    ctx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    l = 3
    Pp = 1
    for p in prm.p:
        Pp *= p
    c = np.stack([rng.integers(0, q, ctx.N, dtype=np.uint64) for q in prm.q[:l]])
    ext = np.zeros((l + ctx.K, ctx.N), np.uint64)
    for i in range(l):
        ext[i] = ctx.mul_limbs(c[i], np.full(ctx.N, Pp % prm.q[i], np.uint64), [i])[0]
    assert np.array_equal(ctx.moddown(ext, l), c)


def test_rescale_table_big_integer_definition():
    c, ctx = _eval16_ctx()
    Q = c["Q"]
    for l in range(2, len(Q) + 1):
        Qp = 1
        for q in Q[: l - 1]:
            Qp *= q
        qt = Q[l - 1]
        big = (Qp * pow(Qp, -1, qt)) // qt
        assert ctx.rescale_table(l) == [big % q for q in Q[: l - 1]]


def test_rescale_golden():
    c, ctx = _eval16_ctx()
    l = c["l"]
    x = np.array(c["x_eval"], np.uint64)
    ct = np.stack([x, x])
    got = ctx.rescale_ct(ct, l)
    assert got[0].tolist() == c["rescale"]
    assert got[1].tolist() == c["rescale"]


def test_switch_modulus_centred_lift():
    rng = np.random.default_rng(5)
    pairs = [(PRIMES["p64"][1], PRIMES["p64"][0]), (PRIMES["p64"][0], PRIMES["p64"][1]),
             (PRIMES["sp64"][0], PRIMES["p64"][0])]
    for om, nm in pairs:
        vals = [0, 1, om >> 1, (om >> 1) + 1, om - 1] + [int(v) for v in rng.integers(0, om, 200, dtype=np.uint64)]
        for a in vals:
            t = a - om if a > (om >> 1) else a
            assert O.switch_modulus(a, om, nm) == t % nm


def test_automorph_golden_and_substitution():
    c, ctx = _eval16_ctx()
    x = np.array(c["x_eval"], np.uint64)
    got = ctx.automorph(x[:1], c["auto_k"])
    assert got[0].tolist() == c["auto_limb0"]


# ------------------------------------------------------------ key switch
def test_keyswitch_trivial_key_is_identity():
    """a = 0, e = 0, s_old = 1: key b_j = P on digit j primes, 0 elsewhere,
    so the key switch must return (d, 0) exactly (component 0 pairs with b)."""
    rng = np.random.default_rng(9)
    prm = P.ckks_params(10, 60, 50, 6, dnum=3, K=2)
    ctx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    LK = ctx.L + ctx.K
    zeros = np.zeros((LK, ctx.N), np.uint64)
    ones = np.ones((LK, ctx.N), np.uint64)
    a = np.zeros((prm.dnum, LK, ctx.N), np.uint64)
    key = ctx.keygen_switch(zeros, ones, a, a)
    for l in (7, 5, 4, 2, 1):
        d = np.stack([rng.integers(0, q, ctx.N, dtype=np.uint64) for q in prm.q[:l]])
        o0, o1 = ctx.keyswitch(d, l, key)
        assert not o1.any(), l
        assert np.array_equal(o0, d), l


def _ckks_setup(seed=1):
    prm = P.config1()
    ctx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(ctx, seed)
    return prm, ctx, keys


def test_hmult_rescale_rotate_decrypt_fp64():
    prm, ctx, keys = _ckks_setup()
    N = ctx.N
    rng = np.random.default_rng(42)
    z1 = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    z2 = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    sc = 2.0 ** prm.scale_bits
    l = prm.L
    c1 = keys.encrypt(C.encode(z1, sc, N), l)
    c2 = keys.encrypt(C.encode(z2, sc, N), l)
    assert C.precision_bits(C.decode(keys.decrypt_coef(c1, l), sc, N), z1) > 25
    cm = ctx.hmult_relin(c1, c2, l, keys.relin_key())
    assert C.precision_bits(C.decode(keys.decrypt_coef(cm, l), sc * sc, N), z1 * z2) > 20
    cr = ctx.rescale_ct(cm, l)
    sc2 = sc * sc / prm.q[l - 1]
    assert C.precision_bits(C.decode(keys.decrypt_coef(cr, l - 1), sc2, N), z1 * z2) > 20
    k = 5  # rotate by +1
    crot = ctx.rotate(cr, l - 1, keys.rotation_key(k), k)
    got = C.decode(keys.decrypt_coef(crot, l - 1), sc2, N)
    assert C.precision_bits(got, np.roll(z1 * z2, -1)) > 18
    # a wrong rotation direction must fail the check
    assert C.precision_bits(got, np.roll(z1 * z2, 1)) < 2


def test_keyswitch_decrypts_to_product_with_old_key():
    prm, ctx, keys = _ckks_setup(3)
    rlk = keys.relin_key()
    rng = np.random.default_rng(4)
    l = prm.L
    d = np.stack([rng.integers(0, q, ctx.N, dtype=np.uint64) for q in prm.q[:l]])
    o0, o1 = ctx.keyswitch(d, l, rlk)
    pids = np.arange(l, dtype=np.uint32)
    s = keys.s_eval[:l]
    s2 = ctx.mul_limbs(s, s, pids)
    lhs = ctx.add_limbs(o0, ctx.mul_limbs(o1, s, pids), pids)
    rhs = ctx.mul_limbs(d, s2, pids)
    diff = ctx.intt_limbs(ctx.sub_limbs(lhs, rhs, pids), pids)
    err = C.crt_centered(diff, prm.q[:l])
    assert max(abs(e) for e in err) < 2 ** 24


def test_ext_composition_pins_on_the_oracle_keyswitch():
    """The extended-basis pieces composed in tests/ext_oracle.py: ModDown of
    the key switch without ModDown is the oracle's key switch; ModDown of the
    P-scaled embedding is the ciphertext itself; ModDown of an extended
    rotation decrypts to the rotated message (its automorphism is applied
    before the ModDown, so the bits may differ from a plain rotation)."""
    import ext_oracle as X
    from fideslib_amd import params
    from oracle import ckks_codec as CC
    prm = params.ckks_params(11, 60, 40, 5, dnum=3)
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = CC.Keys(octx, 11)
    N = prm.N
    rng = np.random.default_rng(5)
    for l in (prm.L, 3):
        z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
        ct = keys.encrypt(CC.encode(z, 2.0 ** 40, N), l)
        key = keys.rotation_key(pow(5, 3, 2 * N))
        k0, k1 = X.keyswitch_ext(octx, ct[1], l, key)
        w0, w1 = octx.keyswitch(ct[1], l, key)
        assert np.array_equal(octx.moddown(k0, l), w0) and np.array_equal(octx.moddown(k1, l), w1)
        assert np.array_equal(X.moddown_ext(octx, X.raise_ext(octx, ct, l), l), ct)
        r = X.moddown_ext(octx, X.rotate_ext(octx, ct, l, key, pow(5, 3, 2 * N)), l)
        dec = CC.decode(keys.decrypt_coef(r, l), 2.0 ** 40, N)
        assert CC.precision_bits(dec, np.roll(z, -3)) > 20
