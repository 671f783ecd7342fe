"""Host logic of the serving loop, on the CPU.

A mock backend stands in for the GPU: device memory is an array of limb
tokens, every batched call executes at once in issue order (the single-stream
order), and a call whose tasks read storage another task of the same call
writes is an error (those would race inside one launch).  Each op's result is
a token derived from the op and its input tokens, so a request served in any
mix must end with the token that plain program-order evaluation gives; a
scheduling or storage-reuse mistake changes it.
"""
import numpy as np
import pytest

from fideslib_amd import params
from fideslib_amd import serve as S


class MockBackend:
    def __init__(self, words):
        self.words = words
        self.base = 0
        self.mem = [None] * words
        self.calls = []

    def ptr(self, w):
        return w

    def _read(self, p, level):
        toks = self.mem[p:p + 2 * level]
        assert all(t is not None for t in toks), "read of unwritten storage"
        t0 = toks[0][0]
        assert all(t == (t0, j) for j, t in enumerate(toks)), "torn register"
        return t0

    def _write(self, p, level, tok):
        for j in range(2 * level):
            self.mem[p + j] = (tok, j)

    def _launch(self, name, reads, writes, fn):
        rset = [set(range(p, p + 2 * l)) for p, l in reads]
        wset = [set(range(p, p + 2 * l)) for p, l in writes]
        n = len(writes)
        per = len(reads) // n
        for i in range(n):
            for j in range(n):
                if i != j:
                    assert not (wset[i] & wset[j]), "%s: two tasks write the same storage" % name
                    for k in range(per):
                        assert not (rset[j * per + k] & wset[i]), "%s: task reads what another writes" % name
        vals = [self._read(p, l) for p, l in reads]
        outs = fn(vals)
        for (p, l), t in zip(writes, outs):
            self._write(p, l, t)
        self.calls.append((name, n))

    def hmul(self, a, b, lv, out, rescale):
        lo = [l - 1 if rescale else l for l in lv]
        reads = [x for i in range(len(a)) for x in ((a[i], lv[i]), (b[i], lv[i]))]
        self._launch("hmul", reads, list(zip(out, lo)),
                     lambda v: [hash(("hmul", rescale, v[2 * i], v[2 * i + 1], lv[i])) for i in range(len(a))])

    def rot(self, a, lv, k, out, keyid):
        self._launch("rot", list(zip(a, lv)), list(zip(out, lv)),
                     lambda v: [hash(("rot", k[i], keyid[i], v[i], lv[i])) for i in range(len(a))])

    def lin(self, a, b, lv, mul, add, sub, out):
        reads, idx = [], []
        for i in range(len(a)):
            reads.append((a[i], lv[i]))
            reads.append((b[i] if b[i] is not None else a[i], lv[i]))
        self._launch("lin", reads, list(zip(out, lv)),
                     lambda v: [hash(("lin", mul[i], add[i], sub[i], v[2 * i], v[2 * i + 1] if b[i] is not None
                                      else None, lv[i])) for i in range(len(a))])

    def lincomb(self, terms, lv, add, out):
        reads = [(p, xl) for t in terms for p, xl, _ in t]
        per = [len(t) for t in terms]

        def fn(v):
            res, o = [], 0
            for i, t in enumerate(terms):
                res.append(hash(("lincomb", tuple(m for _, _, m in t), add[i], tuple(v[o:o + per[i]]), lv[i])))
                o += per[i]
            return res
        self._launch_var("lincomb", reads, per, list(zip(out, lv)), fn)

    def _launch_var(self, name, reads, per, writes, fn):
        wset = [set(range(p, p + 2 * l)) for p, l in writes]
        rs, o = [], 0
        for c in per:
            rs.append([set(range(p, p + 2 * l)) for p, l in reads[o:o + c]])
            o += c
        for i in range(len(writes)):
            for j in range(len(writes)):
                if i != j:
                    assert not (wset[i] & wset[j]), "%s: two tasks write the same storage" % name
                    assert not any(r & wset[i] for r in rs[j]), "%s: task reads what another writes" % name
        vals = [self._read(p, l) for p, l in reads]
        for (p, l), t in zip(writes, fn(vals)):
            self._write(p, l, t)
        self.calls.append((name, len(writes)))

    def rescale(self, a, lv, out):
        self._launch("rescale", list(zip(a, lv)), [(o, l - 1) for o, l in zip(out, lv)],
                     lambda v: [hash(("rescale", v[i], lv[i])) for i in range(len(a))])

    def drop(self, a, lin, lout, out):
        self._launch("drop", list(zip(a, lin)), list(zip(out, lout)),
                     lambda v: [hash(("drop", v[i], lin[i], lout[i])) for i in range(len(a))])

    def upload(self, p, host):
        tok, level = host
        self._write(p, level, tok)

    def copy(self, dst, src, words):
        raise AssertionError("unused")

    def download(self, p, words):
        return self._read(p, words // 2)

    def record(self):
        return None

    def wait(self, ev):
        pass


def sequential(P: S.Program, input_tokens):
    val = {r.id: t for r, t in zip(P.inputs, input_tokens)}
    for o in P.ops:
        v = [val[s.id] for s in o.srcs]
        if o.kind in (S.HMUL, S.HMUL_NR):
            val[o.dst.id] = hash(("hmul", o.kind == S.HMUL, v[0], v[1], o.level))
        elif o.kind == S.ROT:
            val[o.dst.id] = hash(("rot", o.arg[0], o.arg[1], v[0], o.level))
        elif o.kind == S.LIN:
            val[o.dst.id] = hash(("lin", o.arg[0], o.arg[1], o.arg[2], v[0], v[1] if len(v) > 1 else None, o.level))
        elif o.kind == S.RESCALE:
            val[o.dst.id] = hash(("rescale", v[0], o.level))
        elif o.kind == S.DROP:
            val[o.dst.id] = hash(("drop", v[0], o.level, o.arg[0]))
        elif o.kind == S.LINCOMB:
            val[o.dst.id] = hash(("lincomb", o.arg[0], o.arg[1], tuple(v), o.level))
    return val[P.output.id]


class _MockServer(S.Server):
    """Results come back as tokens: words-per-limb is 1 in the mock."""

    def _retire(self, a, now):
        a.req.result = self.be._read(self.be.ptr(a.reg_ptr_words(a.req.program.output)), a.req.program.output.level)
        self.arena.release(a.off, a.comp.block)
        a.req.done_at = now
        a.req.status = "done"


PRM = params.config3()
COEF = S.chebyshev_coefficients(lambda x: 1 / (1 + np.exp(-x)), 15, -8, 8)


def make_requests(rng, n, t_scale=0.0):
    reqs = []
    for i in range(n):
        if rng.random() < 0.6:
            P = S.chebyshev_program(PRM, int(rng.integers(S.chebyshev_depth(15) + 1, PRM.L + 1)), COEF, -8, 8)
        else:
            P = S.rotsum_program(PRM, int(rng.integers(1, PRM.L + 1)), int(rng.integers(1, 16)))
        toks = [(hash(("input", i, j)), r.level) for j, r in enumerate(P.inputs)]
        reqs.append(S.Request(i, P, arrival=float(rng.random() * t_scale), inputs=toks, keep_result=True))
    return reqs


def test_chebyshev_program_shape():
    P = S.chebyshev_program(PRM, 30, COEF, -8, 8)
    assert P.output.level == 30 - S.chebyshev_depth(15)
    # baby steps T_2..T_4, giant T_8, three Chebyshev divisions (p = q T_8 + r,
    # q = q' T_4 + r', r = q'' T_4 + r'')
    assert P.count([S.HMUL]) == 7
    assert all(o.level >= 1 for o in P.ops)
    with pytest.raises(S.ServeError):
        S.chebyshev_program(PRM, S.chebyshev_depth(15), COEF, -8, 8)   # runs out of levels


def test_chebyshev_coefficients_match_numpy():
    f = lambda x: np.sin(x / 3)
    c = S.chebyshev_coefficients(f, 15, -8, 8)
    xs = np.linspace(-8, 8, 101)
    approx = np.polynomial.chebyshev.chebval(xs / 8, c)
    assert np.max(np.abs(approx - f(xs))) < 1e-9
    ref = np.polynomial.chebyshev.chebinterpolate(lambda y: f(8 * y), 15)
    assert np.allclose(c, ref, atol=1e-12)


def test_compile_storage_reuse_and_deps():
    P = S.chebyshev_program(PRM, 30, COEF, -8, 8)
    C = S._compile(P)
    # storage shared: far fewer slots than registers
    assert len(C.slot_size) < len(P.regs) // 2
    # every dependency is issued in a strictly earlier step (ASAP steps)
    assert all(all(C.asap[j] < C.asap[i] for j in d) for i, d in enumerate(C.deps))
    # registers sharing a slot have disjoint lifetimes in step space
    born = {r.id: 0 for r in P.inputs}
    last = {r.id: 0 for r in P.inputs}
    for i, o in enumerate(P.ops):
        born[o.dst.id] = C.asap[i]
        last.setdefault(o.dst.id, C.asap[i])
        for s in o.srcs:
            last[s.id] = max(last[s.id], C.asap[i])
    last[P.output.id] = 1 << 60
    regs = list(born)
    for x in regs:
        for y in regs:
            if x < y and C.slot_of[x] == C.slot_of[y]:
                assert last[x] < born[y] or last[y] < born[x], (x, y)
    # storage never delays an op: the schedule depth equals the data depth
    assert max(C.asap) <= 30


@pytest.mark.parametrize("seed,arena_limbs", [(1, 4000), (2, 1200), (3, 20000)])
def test_mix_equals_sequential_closed_batch(seed, arena_limbs):
    rng = np.random.default_rng(seed)
    reqs = make_requests(rng, 24)
    be = MockBackend(arena_limbs)
    srv = _MockServer(be, 1)
    srv.run(reqs, realtime=False)
    for r in reqs:
        assert r.status == "done"
        assert r.result == sequential(r.program, [t for t, _ in r.inputs]), r.rid
    # batching happened: fewer launches than ops (many fewer when the arena holds many requests)
    assert srv.stats["launches"] < srv.stats["ops"] / (3 if arena_limbs >= 4000 else 1.5)
    assert srv.arena.used == 0


def test_mix_equals_alone_realtime():
    rng = np.random.default_rng(7)
    reqs = make_requests(rng, 12, t_scale=0.02)
    be = MockBackend(3000)
    srv = _MockServer(be, 1)
    srv.run(reqs, realtime=True)
    for r in reqs:
        alone = S.Request(r.rid, r.program, 0.0, inputs=r.inputs, keep_result=True)
        _MockServer(MockBackend(3000), 1).run([alone], realtime=False)
        assert alone.result == r.result


def test_over_cap_request_rejected():
    rng = np.random.default_rng(3)
    reqs = make_requests(rng, 4)
    small = S.Request(99, S.rotsum_program(PRM, 2, 1), 0.0, inputs=[(5, 2)], keep_result=True)
    be = MockBackend(200)
    srv = _MockServer(be, 1)
    srv.run(reqs + [small], realtime=False)
    assert small.status == "done"
    assert all(r.status == "rejected" for r in reqs if r.compiled.block > 200)
    assert srv.stats["rejected"] >= 1


def test_arena_first_fit_coalesces():
    a = S.Arena(0, 100)
    x, y, z = a.alloc(30), a.alloc(30), a.alloc(40)
    assert (x, y, z) == (0, 30, 60) and a.alloc(1) is None
    a.release(y, 30)
    a.release(x, 30)
    assert a.free == [(0, 60)]
    a.release(z, 40)
    assert a.free == [(0, 100)] and a.used == 0


def test_mock_catches_storage_overlap(monkeypatch):
    """A planted allocation mistake (two registers whose lifetimes overlap
    forced into one slot) must be caught."""
    real = S._compile

    def overlapping(P):
        C = real(P)
        born = {r.id: 0 for r in P.inputs}
        last = {r.id: 0 for r in P.inputs}
        for i, o in enumerate(P.ops):
            born[o.dst.id] = C.asap[i]
            last.setdefault(o.dst.id, C.asap[i])
            for x in o.srcs:
                last[x.id] = max(last[x.id], C.asap[i])
        ids = sorted(born, key=lambda r: born[r])
        for a in ids:
            for b in ids:
                if (a != b and born[a] < born[b] <= last[a] and C.slot_of[a] != C.slot_of[b]
                        and C.slot_size[C.slot_of[a]] >= C.slot_size[C.slot_of[b]]):
                    C.slot_of[b] = C.slot_of[a]
                    return C
        return C

    monkeypatch.setattr(S, "_compile", overlapping)
    rng = np.random.default_rng(1)
    reqs = make_requests(rng, 8)
    caught = False
    try:
        _MockServer(MockBackend(20000), 1).run(reqs, realtime=False)
        caught = any(r.result != sequential(r.program, [t for t, _ in r.inputs]) for r in reqs)
    except AssertionError:
        caught = True
    assert caught


def test_programs_decrypt_to_fp64_on_the_oracle():
    """The request programs, run op by op on the CPU oracle, decrypt to the
    FP64 evaluation: Chebyshev series of the sigmoid on [-8, 8] and a
    rotate-and-sum over 8 slots."""
    from oracle import ckks_codec as C
    from oracle import oracle as O
    from serve_oracle import run_oracle
    prm = params.serve_small()
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 123)
    rlk = keys.relin_key()
    N = prm.N
    rng = np.random.default_rng(1)
    P = S.chebyshev_program(prm, 9, COEF, -8, 8)
    z = rng.uniform(-8, 8, N // 2)
    out = run_oracle(octx, P, [keys.encrypt(C.encode(z, 2.0 ** prm.scale_bits, N), 9)], rlk, {})
    dec = C.decode(keys.decrypt_coef(out, P.output.level), P.output.scale, N)
    assert C.precision_bits(dec.real, np.polynomial.chebyshev.chebval(z / 8, COEF)) > 25
    # a planted mistake in the oracle run (one coefficient off by 2^-10) is visible
    bad = S.chebyshev_program(prm, 9, COEF + np.eye(16)[5] * 2.0 ** -10, -8, 8)
    ct = keys.encrypt(C.encode(z, 2.0 ** prm.scale_bits, N), 9)
    dec_bad = C.decode(keys.decrypt_coef(run_oracle(octx, bad, [ct], rlk, {}), bad.output.level), bad.output.scale, N)
    assert C.precision_bits(dec_bad.real, np.polynomial.chebyshev.chebval(z / 8, COEF)) < 12
    k = 3
    rot = {pow(5, 1 << i, 2 * N): keys.rotation_key(pow(5, 1 << i, 2 * N)) for i in range(k)}
    P = S.rotsum_program(prm, 5, k)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    out = run_oracle(octx, P, [keys.encrypt(C.encode(z, 2.0 ** prm.scale_bits, N), 5)], rlk, rot)
    dec = C.decode(keys.decrypt_coef(out, 5), P.output.scale, N)
    assert C.precision_bits(dec, sum(np.roll(z, -t) for t in range(1 << k))) > 25
