"""The extended basis Q_l u P on the GPU, through the C ABI, bit for bit
against the oracle composition of tests/ext_oracle.py: hoisted rotations
left in Q_l u P (with and without the P c0 term, the identity included),
the P-scaled embedding, additions and the two-component ModDown; the
radix-4 rotate-and-sum with one ModDown per round; and the plaintext
embedding (pt, 0).  N = 2^13 (4 limbs, dnum 2) and N = 2^16 (30 limbs,
K = 10, dnum 3)."""
import numpy as np
import pytest

import ext_oracle as X
from fideslib_amd import capi, params
from oracle import ckks_codec as C
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def cfg1():
    prm = params.config1()
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    return prm, octx, gctx, C.Keys(octx, 41)


def _ext_words(prm, l):
    return 2 * (l + prm.K) * prm.N


@pytest.mark.parametrize("add_first", [True, False])
def test_rotate_hoisted_ext_bit_exact(cfg1, add_first):
    prm, octx, gctx, keys = cfg1
    N, L, K = prm.N, prm.L, prm.K
    rng = np.random.default_rng(61)
    for l in (L, 2):
        say("ext rotations l=%d add_first=%s" % (l, add_first))
        z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
        ct = keys.encrypt(C.encode(z, 2.0 ** prm.scale_bits, N), l)
        ks = [5, 1, 2 * N - 1, pow(5, 9, 2 * N)]
        hk = {k: keys.rotation_key(k) for k in ks if k != 1}
        gk = {k: gctx.upload_key(v) for k, v in hk.items()}
        inp = capi.DeviceBuffer.from_array(ct)
        outs = [capi.DeviceBuffer(_ext_words(prm, l)) for _ in ks]
        gctx.rotate_hoisted_ext(inp, l, [gk.get(k) for k in ks], ks, outs, add_first=add_first)
        for k, o in zip(ks, outs):
            want = X.rotate_ext(octx, ct, l, hk.get(k), k, add_first)
            assert np.array_equal(o.download((2, l + K, N)), want), (l, k)
        if add_first:  # ModDown of an extended rotation decrypts to the rotation
            md = capi.DeviceBuffer(2 * l * N)
            gctx.ext_moddown_batch([outs[0]], [l], [md])
            got = md.download((2, l, N))
            assert np.array_equal(got, X.moddown_ext(octx, X.rotate_ext(octx, ct, l, hk[5], 5), l))
            dec = C.decode(keys.decrypt_coef(got, l), 2.0 ** prm.scale_bits, N)
            assert C.precision_bits(dec, np.roll(z, -1)) > 15


def test_ext_raise_add_moddown(cfg1):
    prm, octx, gctx, keys = cfg1
    N, K = prm.N, prm.K
    rng = np.random.default_rng(62)
    levels = [4, 1, 3]
    cts = [np.stack([np.stack([rng.integers(0, prm.q[j], N, dtype=np.uint64) for j in range(l)]) for _ in range(2)])
           for l in levels]
    din = [capi.DeviceBuffer.from_array(c) for c in cts]
    ext = [capi.DeviceBuffer(_ext_words(prm, l)) for l in levels]
    gctx.ext_raise_batch(din, levels, ext)
    for i, l in enumerate(levels):
        assert np.array_equal(ext[i].download((2, l + K, N)), X.raise_ext(octx, cts[i], l)), i
    # a + b, a - b over Q_l u P (b: random extended residues)
    bh = [np.stack([np.stack([rng.integers(0, p, N, dtype=np.uint64) for p in prm.q[:l] + prm.p]) for _ in range(2)])
          for l in levels]
    db = [capi.DeviceBuffer.from_array(b) for b in bh]
    for sub in (False, True):
        out = [capi.DeviceBuffer(_ext_words(prm, l)) for l in levels]
        gctx.ext_add_batch(ext, db, levels, out, subtract=sub)
        for i, l in enumerate(levels):
            want = X.add_ext(octx, X.raise_ext(octx, cts[i], l), bh[i], l, sub)
            assert np.array_equal(out[i].download((2, l + K, N)), want), (sub, i)
    # ModDown of random extended ciphertexts, and of the P-scaled embedding (identity)
    md = [capi.DeviceBuffer(2 * l * N) for l in levels]
    gctx.ext_moddown_batch(db, levels, md)
    for i, l in enumerate(levels):
        assert np.array_equal(md[i].download((2, l, N)), X.moddown_ext(octx, bh[i], l)), i
    gctx.ext_moddown_batch(ext, levels, md)
    for i, l in enumerate(levels):
        assert np.array_equal(md[i].download((2, l, N)), cts[i]), i


def test_accumulate_hoisted_moddown(cfg1):
    """sum_{j < 16} rot(ct, j) by radix-4 rounds: hoisted extended rotations,
    the P-scaled input added, one ModDown per round (reference GPU
    Accumulate); equals the oracle composition and decrypts to the slot sum."""
    prm, octx, gctx, keys = cfg1
    N, L, K = prm.N, prm.L, prm.K
    slots, stride = 16, 1
    rng = np.random.default_rng(63)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    ct = keys.encrypt(C.encode(z, 2.0 ** prm.scale_bits, N), L)
    rots = sorted({j * s for s in (1, 4) for j in (1, 2, 3)})
    hk = {pow(5, r, 2 * N): keys.rotation_key(pow(5, r, 2 * N)) for r in rots}
    gk = {k: gctx.upload_key(v) for k, v in hk.items()}
    x = capi.DeviceBuffer.from_array(ct)
    s = 1
    while s < slots:
        ks = [pow(5, j * s, 2 * N) for j in (1, 2, 3) if j * s < slots]
        exts = [capi.DeviceBuffer(_ext_words(prm, L)) for _ in ks]
        gctx.rotate_hoisted_ext(x, L, [gk[k] for k in ks], ks, exts)
        acc = capi.DeviceBuffer(_ext_words(prm, L))
        gctx.ext_raise_batch([x], [L], [acc])
        for e in exts:
            gctx.ext_add_batch([acc], [e], [L], [acc])
        nx = capi.DeviceBuffer(2 * L * N)
        gctx.ext_moddown_batch([acc], [L], [nx])
        x = nx
        s *= 4
    got = x.download((2, L, N))
    want = X.accumulate(octx, ct, L, hk, slots, stride, N)
    assert np.array_equal(got, want)
    dec = C.decode(keys.decrypt_coef(got, L), 2.0 ** prm.scale_bits, N)
    ref = sum(np.roll(z, -j) for j in range(slots))
    assert C.precision_bits(dec, ref) > 12


def test_plaintext_embed(cfg1):
    prm, octx, gctx, keys = cfg1
    N, l = prm.N, 3
    rng = np.random.default_rng(64)
    coef = np.stack([rng.integers(0, prm.q[j], N, dtype=np.uint64) for j in range(l)])
    pt = capi.DeviceBuffer.from_array(coef)
    gctx.plaintext_prepare(pt, l)
    out = capi.DeviceBuffer(2 * l * N)
    gctx.plaintext_embed(pt, l, out)
    got = out.download((2, l, N))
    assert np.array_equal(got[0], octx.ntt_limbs(coef, list(range(l))))
    assert not got[1].any()


def test_cfg3_rotate_hoisted_ext_bit_exact():
    """N = 2^16, l = 30, K = 10, dnum 3: two extended rotations with real keys."""
    prm = params.config3()
    N, l, K = prm.N, 30, prm.K
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    gctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 43)
    rng = np.random.default_rng(65)
    ct = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in prm.q[:l]]) for _ in range(2)])
    ks = [pow(5, int(r), 2 * N) for r in rng.integers(1, 1 << 15, 2)]
    say("cfg3 ext rotations %s" % ks)
    hk = [keys.rotation_key(k) for k in ks]
    gk = [gctx.upload_key(h) for h in hk]
    inp = capi.DeviceBuffer.from_array(ct)
    outs = [capi.DeviceBuffer(_ext_words(prm, l)) for _ in ks]
    gctx.rotate_hoisted_ext(inp, l, gk, ks, outs)
    for i, k in enumerate(ks):
        assert np.array_equal(outs[i].download((2, l + K, N)), X.rotate_ext(octx, ct, l, hk[i], k)), i
