"""GPU parity: every HIP operation, called through the C ABI, must equal the
CPU oracle limb-for-limb (bit-exact; the arithmetic is exact modular
integer arithmetic, so no tolerance applies).

Sizes: N = 2^10 .. 2^13 for the exhaustive cases, N = 2^16 for the
headline shapes (32 limbs of FIDESlib's 51-bit test primes; 30 Q limbs of
60/59-bit primes with K = 10, dnum = 3).
"""
import json
import os

import numpy as np
import pytest

from fideslib_amd import capi, params
from oracle import ckks_codec as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PRIMES = json.load(open(os.path.join(HERE, "golden", "reference_primes.json")))


def rand_limbs(rng, primes, N):
    return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])


def say(msg):
    print(msg, flush=True)


# ------------------------------------------------------------------ NTT
@pytest.mark.parametrize("logN", [10, 11, 12, 13, 14, 15, 16, 17])
def test_ntt_intt_bit_exact(logN):
    say("ntt logN=%d" % logN)
    rng = np.random.default_rng(logN)
    ok = [x for x in PRIMES["p64"] + PRIMES["sp64"] if (x - 1) % (2 << logN) == 0]
    q = ok[:6]
    p = ok[-3:]
    octx = O.Context(logN, q, p, 2)
    gctx = capi.Context(logN, q, p, 2)
    assert gctx.roots == octx.roots
    N = 1 << logN
    pids = [0, 1, 2, 3, 4, 5, 6, 7, 8, 2, 8, 0]  # mixed, repeated, non-consecutive
    prm = [octx.primes[i] for i in pids]
    x = rand_limbs(rng, prm, N)
    buf = capi.DeviceBuffer.from_array(x)
    gctx.ntt(buf, pids)
    got = buf.download(x.shape)
    want = octx.ntt_limbs(x, pids)
    assert np.array_equal(got, want)
    gctx.intt(buf, pids)
    assert np.array_equal(buf.download(x.shape), x)


@pytest.mark.parametrize("wide", [False, True])
def test_ntt_small_launch_boundary(wide):
    """N = 2^16: plain forward transforms of at most 32 limbs run the row
    pass that loads its twiddles up front, larger ones the regular row pass;
    1, 32 and 33 limbs (primes below 2^60, or with 62-bit primes among them)
    against the oracle, and back through the INTT."""
    logN = 16
    ok = [x for x in PRIMES["p64"] + PRIMES["sp64"] if (x - 1) % (2 << logN) == 0]
    narrow = [x for x in ok if x < 1 << 60]
    q = narrow[:3] if not wide else narrow[:2] + [x for x in ok if x >= 1 << 61][:1]
    p = [x for x in ok if x not in q][:1]
    octx = O.Context(logN, q, p, 3)
    gctx = capi.Context(logN, q, p, 3)
    rng = np.random.default_rng(33 + wide)
    for n in (1, 32, 33):
        say("ntt N=2^16 %d limbs wide=%s" % (n, wide))
        pids = [i % 4 for i in range(n)]
        x = rand_limbs(rng, [octx.primes[i] for i in pids], 1 << logN)
        buf = capi.DeviceBuffer.from_array(x)
        gctx.ntt(buf, pids)
        assert np.array_equal(buf.download(x.shape), octx.ntt_limbs(x, pids)), n
        gctx.intt(buf, pids)
        assert np.array_equal(buf.download(x.shape), x), n


def test_ntt_edge_values():
    """All-zero, all-(q-1) and single-impulse limbs."""
    logN = 12
    q = PRIMES["p64"][:2]
    p = PRIMES["sp64"][:1]
    octx = O.Context(logN, q, p, 1)
    gctx = capi.Context(logN, q, p, 1)
    N = 1 << logN
    x = np.zeros((3, N), np.uint64)
    x[1, :] = q[1] - 1
    x[2, 5] = 1
    pids = [0, 1, 2]
    buf = capi.DeviceBuffer.from_array(x)
    gctx.ntt(buf, pids)
    assert np.array_equal(buf.download(x.shape), octx.ntt_limbs(x, pids))


def test_ntt_many_segments_task_list():
    """More than 8 prime runs forces the device task-list path."""
    logN = 11
    q = PRIMES["p64"][:12]
    p = PRIMES["sp64"][:2]
    octx = O.Context(logN, q, p, 3)
    gctx = capi.Context(logN, q, p, 3)
    rng = np.random.default_rng(1)
    pids = [13, 0, 5, 2, 9, 11, 1, 7, 12, 3, 10, 4, 6, 8]
    x = rand_limbs(rng, [octx.primes[i] for i in pids], 1 << logN)
    buf = capi.DeviceBuffer.from_array(x)
    gctx.ntt(buf, pids)
    assert np.array_equal(buf.download(x.shape), octx.ntt_limbs(x, pids))


# ------------------------------------------------------- ModUp / ModDown (N = 2^16, 32 + 8 limbs)
@pytest.fixture(scope="module")
def cfg2():
    prm = params.config2()
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    return prm, octx, gctx


def test_cfg2_ntt_roundtrip(cfg2):
    prm, octx, gctx = cfg2
    say("cfg2 ntt")
    rng = np.random.default_rng(2)
    x = rand_limbs(rng, prm.q, prm.N)
    buf = capi.DeviceBuffer.from_array(x)
    pids = list(range(prm.L))
    gctx.ntt(buf, pids)
    assert np.array_equal(buf.download(x.shape), octx.ntt_limbs(x, pids))
    gctx.intt(buf, pids)
    assert np.array_equal(buf.download(x.shape), x)


@pytest.mark.parametrize("digit", [0, 3])
def test_cfg2_modup_moddown(cfg2, digit):
    prm, octx, gctx = cfg2
    say("cfg2 modup digit %d" % digit)
    rng = np.random.default_rng(10 + digit)
    l = prm.L
    x = rand_limbs(rng, prm.q, prm.N)
    inp = capi.DeviceBuffer.from_array(x)
    out = capi.DeviceBuffer((l + prm.K) * prm.N)
    gctx.modup(inp, l, digit, out)
    got = out.download((l + prm.K, prm.N))
    want = octx.modup_digit(x, l, digit)
    assert np.array_equal(got, want)
    # ModDown 40 -> 32 of the ModUp output and of a random extended polynomial
    y = rand_limbs(rng, prm.q + prm.p, prm.N)
    for src in (want, y):
        ib = capi.DeviceBuffer.from_array(src)
        ob = capi.DeviceBuffer(l * prm.N)
        gctx.moddown(ib, l, ob)
        assert np.array_equal(ob.download((l, prm.N)), octx.moddown(src, l))
        assert np.array_equal(ib.download(src.shape), src), "moddown must not modify its input"


@pytest.mark.parametrize("l", [1, 7, 8, 9, 20])
def test_partial_digits_modup_moddown(cfg2, l):
    """Levels whose last digit is partial, and single-limb levels."""
    prm, octx, gctx = cfg2
    rng = np.random.default_rng(100 + l)
    x = rand_limbs(rng, prm.q[:l], prm.N)
    inp = capi.DeviceBuffer.from_array(x)
    beta = gctx.digits(l)
    for j in range(beta):
        out = capi.DeviceBuffer((l + prm.K) * prm.N)
        gctx.modup(inp, l, j, out)
        assert np.array_equal(out.download((l + prm.K, prm.N)), octx.modup_digit(x, l, j)), (l, j)
    y = rand_limbs(rng, prm.q[:l] + prm.p, prm.N)
    ib = capi.DeviceBuffer.from_array(y)
    ob = capi.DeviceBuffer(l * prm.N)
    gctx.moddown(ib, l, ob)
    assert np.array_equal(ob.download((l, prm.N)), octx.moddown(y, l))


# ------------------------------------------------------- ciphertext ops, N = 2^13 (4 limbs)
@pytest.fixture(scope="module")
def cfg1():
    prm = params.config1()
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 11)
    return prm, octx, gctx, keys


def _enc_pair(prm, keys, seed, l):
    N = prm.N
    rng = np.random.default_rng(seed)
    z = [rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2) for _ in range(2)]
    sc = 2.0 ** prm.scale_bits
    return z, [keys.encrypt(C.encode(v, sc, N), l) for v in z]


def test_cfg1_keyswitch_hmult_rescale_rotate(cfg1):
    prm, octx, gctx, keys = cfg1
    say("cfg1 ct ops")
    N, l = prm.N, prm.L
    rlk = keys.relin_key()
    z, (c1, c2) = _enc_pair(prm, keys, 5, l)
    g_rlk = gctx.upload_key(rlk)
    # key switch alone
    d = capi.DeviceBuffer.from_array(c2[1])
    o0, o1 = capi.DeviceBuffer(l * N), capi.DeviceBuffer(l * N)
    gctx.keyswitch(d, l, g_rlk, o0, o1)
    w0, w1 = octx.keyswitch(c2[1], l, rlk)
    assert np.array_equal(o0.download((l, N)), w0)
    assert np.array_equal(o1.download((l, N)), w1)
    # HMult + relin
    a, b = capi.DeviceBuffer.from_array(c1), capi.DeviceBuffer.from_array(c2)
    m = capi.DeviceBuffer(2 * l * N)
    gctx.hmult(a, b, l, g_rlk, m)
    want_m = octx.hmult_relin(c1, c2, l, rlk)
    assert np.array_equal(m.download((2, l, N)), want_m)
    # in-place HMult (out aliases a)
    gctx.hmult(a, b, l, g_rlk, a)
    assert np.array_equal(a.download((2, l, N)), want_m)
    # Rescale
    r = capi.DeviceBuffer(2 * (l - 1) * N)
    gctx.rescale(m, l, r)
    want_r = octx.rescale_ct(want_m, l)
    assert np.array_equal(r.download((2, l - 1, N)), want_r)
    # Rotate by +1, -1 and conjugate at l - 1
    for k in (5, pow(5, -1, 2 * N), 2 * N - 1):
        key = keys.rotation_key(k)
        gk = gctx.upload_key(key)
        o = capi.DeviceBuffer(2 * (l - 1) * N)
        gctx.rotate(r, l - 1, gk, k, o)
        assert np.array_equal(o.download((2, l - 1, N)), octx.rotate(want_r, l - 1, key, k)), k
    # decrypted result against FP64
    sc2 = (2.0 ** prm.scale_bits) ** 2 / prm.q[l - 1]
    dec = C.decode(keys.decrypt_coef(r.download((2, l - 1, N)), l - 1), sc2, N)
    assert C.precision_bits(dec, z[0] * z[1]) > 20


def test_cfg1_hmult_square_extreme_residues(cfg1):
    """HMult and Square with residues at the ends of [0, q): every coefficient
    q - 1, 0 or 1 in the inputs (the largest products the tensor product's
    Montgomery reductions see), against the oracle."""
    prm, octx, gctx, keys = cfg1
    N, l = prm.N, prm.L
    rlk = keys.relin_key()
    g_rlk = gctx.upload_key(rlk)
    rng = np.random.default_rng(44)
    qs = np.asarray(prm.q[:l], np.uint64)[None, :, None]
    full = np.broadcast_to(qs - np.uint64(1), (2, l, N)).copy()
    mixed = np.where(rng.integers(0, 3, (2, l, N)) == 0, qs - np.uint64(1),
                     rng.integers(0, 2, (2, l, N)).astype(np.uint64))
    for x, y in ((full, full), (full, mixed), (mixed, full)):
        a, b = capi.DeviceBuffer.from_array(x), capi.DeviceBuffer.from_array(y)
        m = capi.DeviceBuffer(2 * l * N)
        gctx.hmult(a, b, l, g_rlk, m)
        assert np.array_equal(m.download((2, l, N)), octx.hmult_relin(x, y, l, rlk))
    for x in (full, mixed):
        a = capi.DeviceBuffer.from_array(x)
        m = capi.DeviceBuffer(2 * l * N)
        gctx.square(a, l, g_rlk, m)
        assert np.array_equal(m.download((2, l, N)), octx.hmult_relin(x, x, l, rlk))


@pytest.mark.parametrize("logN", [10, 11, 12, 14, 15, 17])
def test_key_switch_ops_every_ring_dimension(logN):
    """HMult+relin, Rescale, Rotate and a mixed-level batched HMult+Rescale at
    the ring dimensions the other tests do not use (each has its own NTT,
    fused row-pass and inner-product kernels): 4 Q limbs (60 + 3 x 50 bits),
    dnum 2, random residues and a random key, against the oracle."""
    prm = params.ckks_params(logN, 60, 50, 3, dnum=2)
    say("key switch logN=%d K=%d" % (logN, prm.K))
    N, L = prm.N, prm.L
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(100 + logN)
    key = np.stack([np.stack([rand_limbs(rng, prm.q + prm.p, N) for _ in range(prm.dnum)]) for _ in range(2)])
    gkey = gctx.upload_key(key)
    cts = [np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)]) for l in (L, L, 3, 2)]
    a, b = capi.DeviceBuffer.from_array(cts[0]), capi.DeviceBuffer.from_array(cts[1])
    m = capi.DeviceBuffer(2 * L * N)
    gctx.hmult(a, b, L, gkey, m)
    want_m = octx.hmult_relin(cts[0], cts[1], L, key)
    assert np.array_equal(m.download((2, L, N)), want_m)
    r = capi.DeviceBuffer(2 * (L - 1) * N)
    gctx.rescale(m, L, r)
    want_r = octx.rescale_ct(want_m, L)
    assert np.array_equal(r.download((2, L - 1, N)), want_r)
    k = pow(5, 3, 2 * N)
    o = capi.DeviceBuffer(2 * (L - 1) * N)
    gctx.rotate(r, L - 1, gkey, k, o)
    assert np.array_equal(o.download((2, L - 1, N)), octx.rotate(want_r, L - 1, key, k))
    levels = [L, 3, 2]
    A = [capi.DeviceBuffer.from_array(c) for c in (cts[0], cts[2], cts[3])]
    OB = [capi.DeviceBuffer(2 * (l - 1) * N) for l in levels]
    gctx.hmult_batch(A, A, levels, gkey, OB, rescale=True)
    for c, l, ob in zip((cts[0], cts[2], cts[3]), levels, OB):
        want = octx.rescale_ct(octx.hmult_relin(c, c, l, key), l)
        assert np.array_equal(ob.download((2, l - 1, N)), want), l


@pytest.mark.parametrize("dnum", [1, 6, 12])
def test_key_switch_digit_counts(dnum):
    """Digit counts outside the fused inner product's range (more than 4 active
    digits run the row pass and the inner product as separate launches) and a
    single digit: N = 2^12, 12 Q limbs, HMult+relin+Rescale and Rotate at
    l = 12 and l = 7, against the oracle."""
    prm = params.ckks_params(12, 60, 40, 11, dnum=dnum)
    say("key switch dnum=%d K=%d" % (dnum, prm.K))
    N, L = prm.N, prm.L
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(200 + dnum)
    key = np.stack([np.stack([rand_limbs(rng, prm.q + prm.p, N) for _ in range(prm.dnum)]) for _ in range(2)])
    gkey = gctx.upload_key(key)
    for l in (L, 7):
        x = np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])
        y = np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])
        a, b = capi.DeviceBuffer.from_array(x), capi.DeviceBuffer.from_array(y)
        m, r = capi.DeviceBuffer(2 * l * N), capi.DeviceBuffer(2 * (l - 1) * N)
        gctx.hmult(a, b, l, gkey, m)
        gctx.rescale(m, l, r)
        want_r = octx.rescale_ct(octx.hmult_relin(x, y, l, key), l)
        assert np.array_equal(r.download((2, l - 1, N)), want_r), l
        o = capi.DeviceBuffer(2 * l * N)
        gctx.rotate(a, l, gkey, 5, o)
        assert np.array_equal(o.download((2, l, N)), octx.rotate(x, l, key, 5)), l


@pytest.mark.parametrize("bits", [61, 62])
def test_key_switch_wide_primes(bits):
    """Primes at or above 2^60 (up to the 2^62 limit): the basis conversions
    switch from the carry-free split multiply-accumulate to the column
    accumulator with counted carries.  N = 2^12, 4 Q limbs + special primes of
    the same width, HMult+relin+Rescale and Rotate against the oracle."""
    prm = params.ckks_params(12, bits, bits, 3, dnum=2, special_bits=bits)
    assert max(prm.q + prm.p) >= 1 << 60
    say("key switch %d-bit primes" % bits)
    N, L = prm.N, prm.L
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(300 + bits)
    key = np.stack([np.stack([rand_limbs(rng, prm.q + prm.p, N) for _ in range(prm.dnum)]) for _ in range(2)])
    gkey = gctx.upload_key(key)
    x = np.stack([rand_limbs(rng, prm.q, N) for _ in range(2)])
    y = np.stack([rand_limbs(rng, prm.q, N) for _ in range(2)])
    a, b = capi.DeviceBuffer.from_array(x), capi.DeviceBuffer.from_array(y)
    m, r = capi.DeviceBuffer(2 * L * N), capi.DeviceBuffer(2 * (L - 1) * N)
    gctx.hmult(a, b, L, gkey, m)
    gctx.rescale(m, L, r)
    want_m = octx.hmult_relin(x, y, L, key)
    assert np.array_equal(m.download((2, L, N)), want_m)
    assert np.array_equal(r.download((2, L - 1, N)), octx.rescale_ct(want_m, L))
    o = capi.DeviceBuffer(2 * L * N)
    gctx.rotate(a, L, gkey, 5, o)
    assert np.array_equal(o.download((2, L, N)), octx.rotate(x, L, key, 5))


def test_lincomb_wide_primes_extreme_terms():
    """16 terms whose products are all (q-1)^2 under 62-bit primes just below
    2^62: the 128-bit column sum comes within 2^64 of overflowing, so the
    Montgomery reduction must first reduce its high word (a plain
    hi + floor(m q / 2^64) wraps).  17 terms are refused for such primes;
    beyond 16 digits the context refuses the parameters."""
    logN = 12
    q = params.ntt_primes(62, 4, logN)
    p = params.ntt_primes(62, 2, logN, exclude=q)
    assert min(q) > (1 << 62) - (1 << 40)
    say("lincomb 62-bit extremes")
    N, l = 1 << logN, 4
    octx = O.Context(logN, q, p, 2)
    gctx = capi.Context(logN, q, p, 2)
    x = np.stack([np.full(N, qi - 1, np.uint64) for qi in q] * 2)  # [2 l][N]
    dx = capi.DeviceBuffer.from_array(x)
    # constant residues m_j with m_j 2^64 = q_j - 1 (mod q_j): the kernel's
    # Montgomery-form constants are then q_j - 1, the largest possible
    res = [(qi - 1) * pow(pow(2, 64, qi), -1, qi) % qi for qi in q]
    Q = 1
    for qi in q:
        Q *= qi
    M = sum(r * (Q // qi) * pow(Q // qi, -1, qi) for r, qi in zip(res, q)) % Q
    out = capi.DeviceBuffer(2 * l * N)
    gctx.lincomb_batch([[(dx, l, M)] * 16], [l], [None], [out])
    pid = list(range(l)) * 2
    c = np.stack([np.full(N, M % q[j], np.uint64) for j in range(l)] * 2)
    t = octx.mul_limbs(x, c, pid)
    want = t
    for _ in range(15):
        want = octx.add_limbs(want, t, pid)
    assert np.array_equal(out.download((2 * l, N)), want)
    with pytest.raises(capi.FidesError):
        gctx.lincomb_batch([[(dx, l, M)] * 17], [l], [None], [out])


def test_cfg1_square_and_lower_levels(cfg1):
    prm, octx, gctx, keys = cfg1
    N = prm.N
    rlk = keys.relin_key()
    g_rlk = gctx.upload_key(rlk)
    for l in (1, 2, 3):
        z, (c1, c2) = _enc_pair(prm, keys, 20 + l, l)
        a = capi.DeviceBuffer.from_array(c1)
        o = capi.DeviceBuffer(2 * l * N)
        gctx.square(a, l, g_rlk, o)
        assert np.array_equal(o.download((2, l, N)), octx.hmult_relin(c1, c1, l, rlk)), l


def test_automorph_keyless(cfg1):
    prm, octx, gctx, keys = cfg1
    rng = np.random.default_rng(4)
    x = rand_limbs(rng, prm.q, prm.N)
    for k in (5, 25, 2 * prm.N - 1, 3):
        ib = capi.DeviceBuffer.from_array(x)
        ob = capi.DeviceBuffer(x.size)
        gctx.automorph(ib, prm.L, k, ob)
        assert np.array_equal(ob.download(x.shape), octx.automorph(x, k)), k


def test_error_paths(cfg1):
    prm, octx, gctx, keys = cfg1
    N = prm.N
    a = capi.DeviceBuffer(2 * prm.L * N)
    with pytest.raises(capi.FidesError):
        gctx.rescale(a, 1, capi.DeviceBuffer(2 * N))  # l < 2
    with pytest.raises(capi.FidesError):
        gctx.rescale(a, 2, a)  # aliasing
    with pytest.raises(capi.FidesError):
        gctx.automorph(a, 1, 4, capi.DeviceBuffer(N))  # even index
    with pytest.raises(capi.FidesError):
        gctx.ntt(a, [99])  # prime id out of range


# ------------------------------------------------------- headline shape, N = 2^16, L = 29 (30 limbs)
@pytest.fixture(scope="module")
def cfg3():
    prm = params.config3()
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 29)
    return prm, octx, gctx, keys


@pytest.mark.parametrize("l", [30, 16, 6])
def test_cfg3_hmult_rescale_bit_exact(cfg3, l):
    prm, octx, gctx, keys = cfg3
    say("cfg3 hmult l=%d" % l)
    N = prm.N
    rlk = keys.relin_key()
    z, (c1, c2) = _enc_pair(prm, keys, 300 + l, l)
    g_rlk = gctx.upload_key(rlk)
    a, b = capi.DeviceBuffer.from_array(c1), capi.DeviceBuffer.from_array(c2)
    m = capi.DeviceBuffer(2 * l * N)
    r = capi.DeviceBuffer(2 * (l - 1) * N)
    gctx.hmult(a, b, l, g_rlk, m)
    gctx.rescale(m, l, r)
    want = octx.rescale_ct(octx.hmult_relin(c1, c2, l, rlk), l)
    got = r.download((2, l - 1, N))
    assert np.array_equal(got, want)
    sc2 = (2.0 ** prm.scale_bits) ** 2 / prm.q[l - 1]
    dec = C.decode(keys.decrypt_coef(got, l - 1), sc2, N)
    assert C.precision_bits(dec, z[0] * z[1]) > 30


def test_cfg3_rotate_bit_exact(cfg3):
    prm, octx, gctx, keys = cfg3
    say("cfg3 rotate")
    N, l = prm.N, 30
    rng = np.random.default_rng(77)
    r = int(rng.integers(1, 1 << 15))
    k = pow(5, r, 2 * N)
    key = keys.rotation_key(k)
    z, (c1, _) = _enc_pair(prm, keys, 301, l)
    g = gctx.upload_key(key)
    a = capi.DeviceBuffer.from_array(c1)
    o = capi.DeviceBuffer(2 * l * N)
    gctx.rotate(a, l, g, k, o)
    got = o.download((2, l, N))
    assert np.array_equal(got, octx.rotate(c1, l, key, k))
    dec = C.decode(keys.decrypt_coef(got, l), 2.0 ** prm.scale_bits, N)
    assert C.precision_bits(dec, np.roll(z[0], -r)) > 30


@pytest.mark.parametrize("l", [30, 7])
def test_cfg3_conjugate_bit_exact(cfg3, l):
    """Conjugation (automorphism 2N - 1) at the headline shape against the
    oracle; decrypts to the conjugated slots."""
    prm, octx, gctx, keys = cfg3
    say("cfg3 conjugate l=%d" % l)
    N = prm.N
    k = 2 * N - 1
    key = keys.rotation_key(k)
    z, (c1, _) = _enc_pair(prm, keys, 303 + l, l)
    g = gctx.upload_key(key)
    a = capi.DeviceBuffer.from_array(c1)
    o = capi.DeviceBuffer(2 * l * N)
    gctx.rotate(a, l, g, k, o)
    got = o.download((2, l, N))
    assert np.array_equal(got, octx.rotate(c1, l, key, k))
    dec = C.decode(keys.decrypt_coef(got, l), 2.0 ** prm.scale_bits, N)
    assert C.precision_bits(dec, np.conj(z[0])) > 30


# ------------------------------------------------------- batched = plain = oracle
def test_cfg1_batch_equals_single_and_oracle(cfg1):
    prm, octx, gctx, keys = cfg1
    say("cfg1 batch")
    N = prm.N
    rlk = keys.relin_key()
    g_rlk = gctx.upload_key(rlk)
    levels = [4, 2, 3, 4, 2]
    pairs = [_enc_pair(prm, keys, 40 + i, l)[1] for i, l in enumerate(levels)]
    A = [capi.DeviceBuffer.from_array(p[0]) for p in pairs]
    B = [capi.DeviceBuffer.from_array(p[1]) for p in pairs]
    OB = [capi.DeviceBuffer(2 * (l - 1) * N) for l in levels]
    gctx.hmult_batch(A, B, levels, g_rlk, OB, rescale=True)
    for i, l in enumerate(levels):
        want = octx.rescale_ct(octx.hmult_relin(pairs[i][0], pairs[i][1], l, rlk), l)
        assert np.array_equal(OB[i].download((2, l - 1, N)), want), i
    # without rescale
    OM = [capi.DeviceBuffer(2 * l * N) for l in levels]
    gctx.hmult_batch(A, B, levels, g_rlk, OM, rescale=False)
    for i, l in enumerate(levels):
        assert np.array_equal(OM[i].download((2, l, N)), octx.hmult_relin(pairs[i][0], pairs[i][1], l, rlk)), i
    # rotations with distinct keys
    ks = [5, 25, 2 * N - 1, pow(5, 7, 2 * N), 5]
    hkeys = {k: keys.rotation_key(k) for k in set(ks)}
    gkeys = {k: gctx.upload_key(v) for k, v in hkeys.items()}
    lv2 = [l - 1 for l in levels]
    RO = [capi.DeviceBuffer(2 * l * N) for l in lv2]
    gctx.rotate_batch(OB, lv2, [gkeys[k] for k in ks], ks, RO)
    for i, l in enumerate(lv2):
        cin = OB[i].download((2, l, N))
        want = octx.rotate(cin, l, hkeys[ks[i]], ks[i])
        assert np.array_equal(RO[i].download((2, l, N)), want), i
        single = capi.DeviceBuffer(2 * l * N)
        gctx.rotate(OB[i], l, gkeys[ks[i]], ks[i], single)
        assert np.array_equal(single.download((2, l, N)), want), i


def test_cfg3_batch_mixed_levels_bit_exact(cfg3):
    """16 ciphertexts at levels drawn from U{6..30} in one batched call;
    each equals the oracle and the single-ciphertext GPU call."""
    prm, octx, gctx, keys = cfg3
    say("cfg3 batch")
    N = prm.N
    rng = np.random.default_rng(2026)
    levels = [int(x) for x in rng.integers(6, prm.L + 1, 16)]
    levels[0], levels[1] = 30, 6
    rlk = keys.relin_key()
    g_rlk = gctx.upload_key(rlk)
    cts = []
    for i, l in enumerate(levels):
        a = np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])
        b = np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])
        cts.append((a, b))
    A = [capi.DeviceBuffer.from_array(a) for a, _ in cts]
    B = [capi.DeviceBuffer.from_array(b) for _, b in cts]
    OB = [capi.DeviceBuffer(2 * (l - 1) * N) for l in levels]
    gctx.hmult_batch(A, B, levels, g_rlk, OB, rescale=True)
    for i in (0, 1, 7, 15):
        l = levels[i]
        want = octx.rescale_ct(octx.hmult_relin(cts[i][0], cts[i][1], l, rlk), l)
        assert np.array_equal(OB[i].download((2, l - 1, N)), want), (i, l)
    for i, l in enumerate(levels):
        m = capi.DeviceBuffer(2 * l * N)
        r = capi.DeviceBuffer(2 * (l - 1) * N)
        gctx.hmult(A[i], B[i], l, g_rlk, m)
        gctx.rescale(m, l, r)
        assert np.array_equal(r.download((2, l - 1, N)), OB[i].download((2, l - 1, N))), (i, l)


@pytest.mark.parametrize("bits,dnum", [(62, 4), (59, 4), (59, 3)])
def test_mixed_level_batch_every_digit_shape(bits, dnum):
    """N = 2^16, 8 Q limbs: a batched HMult+Rescale and Rotate at levels
    whose digit counts (1..4) and partial-digit sizes differ, so the ModUp
    conversions of every digit size share one launch and the fused inner
    products of every digit count share one launch; with 62-bit primes the
    conversion takes the counted-carry path and the NTTs the wide-prime
    butterflies.  Every ciphertext equals the single-ciphertext calls (one
    digit shape per launch), three of them the oracle."""
    prm = params.ckks_params(16, bits, bits, 7, dnum=dnum, special_bits=bits)
    say("mixed-level batch N=2^16 %d-bit primes dnum=%d K=%d" % (bits, dnum, prm.K))
    N, L = prm.N, prm.L
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(400 + bits + dnum)
    key = np.stack([np.stack([rand_limbs(rng, prm.q + prm.p, N) for _ in range(prm.dnum)]) for _ in range(2)])
    gkey = gctx.upload_key(key)
    levels = [8, 7, 5, 4, 3, 2]
    cts = [(np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)]),
            np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])) for l in levels]
    A = [capi.DeviceBuffer.from_array(a) for a, _ in cts]
    B = [capi.DeviceBuffer.from_array(b) for _, b in cts]
    OB = [capi.DeviceBuffer(2 * (l - 1) * N) for l in levels]
    gctx.hmult_batch(A, B, levels, gkey, OB, rescale=True)
    k = pow(5, 7, 2 * N)
    OR = [capi.DeviceBuffer(2 * l * N) for l in levels]
    gctx.rotate_batch(A, levels, [gkey] * len(levels), [k] * len(levels), OR)
    for i, l in enumerate(levels):
        m = capi.DeviceBuffer(2 * l * N)
        r = capi.DeviceBuffer(2 * (l - 1) * N)
        gctx.hmult(A[i], B[i], l, gkey, m)
        gctx.rescale(m, l, r)
        assert np.array_equal(r.download((2, l - 1, N)), OB[i].download((2, l - 1, N))), ("hmult", i, l)
        o = capi.DeviceBuffer(2 * l * N)
        gctx.rotate(A[i], l, gkey, k, o)
        assert np.array_equal(o.download((2, l, N)), OR[i].download((2, l, N))), ("rotate", i, l)
    for i in (0, 1, 4):
        l = levels[i]
        want = octx.rescale_ct(octx.hmult_relin(cts[i][0], cts[i][1], l, key), l)
        assert np.array_equal(OB[i].download((2, l - 1, N)), want), ("oracle hmult", i, l)
        assert np.array_equal(OR[i].download((2, l, N)), octx.rotate(cts[i][0], l, key, k)), ("oracle rotate", i, l)


def test_cfg3_rotate_batch_real_keys_mixed_levels(cfg3):
    """The headline rotation shape: 16 ciphertexts at limb counts U{6..30}
    (l = 30 and l = 6 included), each with its own real rotation key (index
    5^r mod 2N, r ~ U[1, 2^15)), in one fides_rotate_batch call.  Four of them
    (l = 30, l = 6 and two others) equal the oracle; every one equals a single
    fides_rotate; the l = 30 one is a real encryption and decrypts to the
    rotated message."""
    prm, octx, gctx, keys = cfg3
    N = prm.N
    rng = np.random.default_rng(2027)
    n = 16
    levels = [int(x) for x in rng.integers(6, prm.L + 1, n)]
    levels[0], levels[1] = 30, 6
    rots = [int(x) for x in rng.integers(1, 1 << 15, n)]
    ks = [pow(5, r, 2 * N) for r in rots]
    say("cfg3 rotate batch: levels %s" % levels)
    hkeys = [keys.rotation_key(k) for k in ks]
    gk = [gctx.upload_key(h) for h in hkeys]
    z, (c0, _) = _enc_pair(prm, keys, 303, 30)
    cts = [c0] + [np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)]) for l in levels[1:]]
    IN = [capi.DeviceBuffer.from_array(c) for c in cts]
    OUT = [capi.DeviceBuffer(2 * l * N) for l in levels]
    gctx.rotate_batch(IN, levels, gk, ks, OUT)
    got = [o.download((2, l, N)) for o, l in zip(OUT, levels)]
    for i in (0, 1, 5, 11):
        say("  oracle %d (l=%d)" % (i, levels[i]))
        assert np.array_equal(got[i], octx.rotate(cts[i], levels[i], hkeys[i], ks[i])), (i, levels[i])
    for i, l in enumerate(levels):
        single = capi.DeviceBuffer(2 * l * N)
        gctx.rotate(IN[i], l, gk[i], ks[i], single)
        assert np.array_equal(single.download((2, l, N)), got[i]), (i, l)
    dec = C.decode(keys.decrypt_coef(got[0], 30), 2.0 ** prm.scale_bits, N)
    assert C.precision_bits(dec, np.roll(z[0], -rots[0])) > 30
    # an output may not overlap any input of the batch (the ModDown store
    # scatters while other jobs still read their c0)
    with pytest.raises(capi.FidesError):
        gctx.rotate_batch(IN[:2], levels[:2], gk[:2], ks[:2], [OUT[0], IN[0]])


# ------------------------------------------------------- element-wise batch ops, limb drop, batched rescale
def test_cfg1_elementwise_batch_ops(cfg1):
    prm, octx, gctx, keys = cfg1
    say("cfg1 elementwise")
    N = prm.N
    rng = np.random.default_rng(55)
    levels = [4, 1, 3, 2]
    A = [np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)]) for l in levels]
    B = [np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)]) for l in levels]
    dA = [capi.DeviceBuffer.from_array(x) for x in A]
    dB = [capi.DeviceBuffer.from_array(x) for x in B]
    out = [capi.DeviceBuffer(x.size) for x in A]

    def flat(x, l):
        return x.reshape(2 * l, N)

    def pids(l):
        return list(range(l)) * 2

    for sub in (False, True):
        gctx.add_batch(dA, dB, levels, out, subtract=sub)
        for i, l in enumerate(levels):
            f = octx.sub_limbs if sub else octx.add_limbs
            assert np.array_equal(out[i].download((2 * l, N)), f(flat(A[i], l), flat(B[i], l), pids(l))), (sub, i)
    consts = [3, -7, (1 << 100) + 12345, -(1 << 70)]
    gctx.mul_const_batch(dA, levels, consts, out)
    for i, l in enumerate(levels):
        cres = np.stack([np.full(N, consts[i] % prm.q[j], np.uint64) for j in range(l)] * 2)
        assert np.array_equal(out[i].download((2 * l, N)), octx.mul_limbs(flat(A[i], l), cres, pids(l))), i
    gctx.add_const_batch(dA, levels, consts, out)
    for i, l in enumerate(levels):
        cres = np.stack([np.full(N, consts[i] % prm.q[j], np.uint64) for j in range(l)] + [np.zeros(N, np.uint64)] * l)
        assert np.array_equal(out[i].download((2 * l, N)), octx.add_limbs(flat(A[i], l), cres, pids(l))), i
    # in place: out aliases a
    gctx.add_batch(dA, dB, levels, dA)
    for i, l in enumerate(levels):
        assert np.array_equal(dA[i].download((2 * l, N)), octx.add_limbs(flat(A[i], l), flat(B[i], l), pids(l)))
    # limb drop
    lo = [2, 1, 1, 2]
    dout = [capi.DeviceBuffer(2 * x * N) for x in lo]
    gctx.drop_limbs_batch(dB, levels, lo, dout)
    for i in range(len(levels)):
        assert np.array_equal(dout[i].download((2, lo[i], N)), B[i][:, :lo[i]]), i
    with pytest.raises(capi.FidesError):
        gctx.drop_limbs_batch([dB[0]], [4], [2], [dB[0]])
    with pytest.raises(capi.FidesError):
        gctx.mul_const_batch([dA[0]], [5], [1], [out[0]])  # level above L
    # batched rescale at mixed levels
    lv = [4, 2, 3]
    rin = [capi.DeviceBuffer.from_array(A[0]), capi.DeviceBuffer.from_array(B[3]), capi.DeviceBuffer.from_array(B[2])]
    hin = [A[0], B[3], B[2]]
    rout = [capi.DeviceBuffer(2 * (l - 1) * N) for l in lv]
    gctx.rescale_batch(rin, lv, rout)
    for i, l in enumerate(lv):
        assert np.array_equal(rout[i].download((2, l - 1, N)), octx.rescale_ct(hin[i], l)), i


def test_cfg1_lin_batch_fused(cfg1):
    """out = a*m -/+ b + c with optional terms, against the oracle's limb ops."""
    prm, octx, gctx, keys = cfg1
    N = prm.N
    rng = np.random.default_rng(56)
    levels = [4, 3, 2, 4, 1]
    A = [np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)]) for l in levels]
    B = [np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)]) for l in levels]
    dA = [capi.DeviceBuffer.from_array(x) for x in A]
    dB = [capi.DeviceBuffer.from_array(x) for x in B]
    out = [capi.DeviceBuffer(x.size) for x in A]
    mul = [2, None, -(1 << 80), 5, 7]
    add = [-(1 << 59), 3, None, None, 1]
    useb = [True, True, False, True, False]
    sub = [True, False, False, True, False]
    gctx.lin_batch(dA, [dB[i] if useb[i] else None for i in range(5)], levels, mul, add, sub, out)
    for i, l in enumerate(levels):
        pid = list(range(l)) * 2
        x = A[i].reshape(2 * l, N)
        if mul[i] is not None:
            x = octx.mul_limbs(x, np.stack([np.full(N, mul[i] % prm.q[j], np.uint64) for j in pid]), pid)
        if useb[i]:
            x = (octx.sub_limbs if sub[i] else octx.add_limbs)(x, B[i].reshape(2 * l, N), pid)
        if add[i] is not None:
            c = np.stack([np.full(N, add[i] % prm.q[j], np.uint64) for j in range(l)] + [np.zeros(N, np.uint64)] * l)
            x = octx.add_limbs(x, c, pid)
        assert np.array_equal(out[i].download((2 * l, N)), x), i


# ------------------------------------------------------- hoisted rotations (one ModUp, many keys)
def test_cfg1_rotate_hoisted_equals_rotate(cfg1):
    prm, octx, gctx, keys = cfg1
    say("cfg1 hoisted")
    N, l = prm.N, 3
    _, (c1, _) = _enc_pair(prm, keys, 71, l)
    ks = [5, 1, 2 * N - 1, pow(5, 7, 2 * N), pow(5, -3, 2 * N)]
    hk = {k: keys.rotation_key(k) for k in ks if k != 1}
    gk = {k: gctx.upload_key(v) for k, v in hk.items()}
    inp = capi.DeviceBuffer.from_array(c1)
    outs = [capi.DeviceBuffer(2 * l * N) for _ in ks]
    gctx.rotate_hoisted(inp, l, [gk.get(k) for k in ks], ks, outs)
    for k, o in zip(ks, outs):
        got = o.download((2, l, N))
        if k == 1:
            assert np.array_equal(got, c1)
            continue
        assert np.array_equal(got, octx.rotate(c1, l, hk[k], k)), k
        single = capi.DeviceBuffer(2 * l * N)
        gctx.rotate(inp, l, gk[k], k, single)
        assert np.array_equal(single.download((2, l, N)), got), k
    assert np.array_equal(inp.download((2, l, N)), c1), "input must be unchanged"
    with pytest.raises(capi.FidesError):
        gctx.rotate_hoisted(inp, l, [gk[5]], [5], [inp])  # output overlaps input
    with pytest.raises(capi.FidesError):
        gctx.rotate_hoisted(inp, l, [gk[5], gk[5]], [5, 5], [outs[0], outs[0]])  # outputs overlap


def test_cfg3_rotate_hoisted_equals_batch(cfg3):
    prm, octx, gctx, keys = cfg3
    say("cfg3 hoisted")
    N, l = prm.N, 23
    rng = np.random.default_rng(72)
    c = np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])
    host_key = np.stack([np.stack([rand_limbs(rng, prm.q + prm.p, N) for _ in range(prm.dnum)]) for _ in range(2)])
    ks = [pow(5, int(r), 2 * N) for r in rng.integers(1, 1 << 15, 6)]
    gk = [gctx.upload_key(host_key) for _ in ks]
    inp = capi.DeviceBuffer.from_array(c)
    outs = [capi.DeviceBuffer(2 * l * N) for _ in ks]
    gctx.rotate_hoisted(inp, l, gk, ks, outs)
    ref = [capi.DeviceBuffer(2 * l * N) for _ in ks]
    gctx.rotate_batch([inp] * len(ks), [l] * len(ks), gk, ks, ref)
    for i in range(len(ks)):
        assert np.array_equal(outs[i].download((2, l, N)), ref[i].download((2, l, N))), i
    assert np.array_equal(outs[2].download((2, l, N)), octx.rotate(c, l, host_key, ks[2]))


def test_cfg1_lincomb_mixed_levels(cfg1):
    """out = sum_j m_j x_j + c with terms stored at levels >= the output's."""
    prm, octx, gctx, keys = cfg1
    N = prm.N
    rng = np.random.default_rng(57)
    lv_terms = [[4, 3, 2], [4] * 20, [2]]
    lv_out = [2, 4, 1]
    xs = [[np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)]) for l in t] for t in lv_terms]
    dx = [[capi.DeviceBuffer.from_array(x) for x in t] for t in xs]
    ms = [[int(rng.integers(-(1 << 62), 1 << 62)) * (1 << int(rng.integers(0, 40))) for _ in t] for t in lv_terms]
    adds = [5, None, -(1 << 70)]
    out = [capi.DeviceBuffer(2 * l * N) for l in lv_out]
    gctx.lincomb_batch([[(dx[i][j], lv_terms[i][j], ms[i][j]) for j in range(len(lv_terms[i]))] for i in range(3)],
                       lv_out, adds, out)
    for i, l in enumerate(lv_out):
        pid = list(range(l)) * 2
        y = None
        for x, m in zip(xs[i], ms[i]):
            c = np.stack([np.full(N, m % prm.q[j], np.uint64) for j in range(l)] * 2)
            t = octx.mul_limbs(np.ascontiguousarray(x[:, :l]).reshape(2 * l, N), c, pid)
            y = t if y is None else octx.add_limbs(y, t, pid)
        if adds[i] is not None:
            c = np.stack([np.full(N, adds[i] % prm.q[j], np.uint64) for j in range(l)] + [np.zeros(N, np.uint64)] * l)
            y = octx.add_limbs(y, c, pid)
        assert np.array_equal(out[i].download((2 * l, N)), y), i
