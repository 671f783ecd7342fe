"""Continuous-batching request loop over the batched C ABI.

Requests are small CKKS programs (a degree-15 Chebyshev series, a
rotate-and-sum over 2^k slots, ...) compiled into a list of register ops.
The loop admits requests as they arrive, while the HBM arena has room for
their peak footprint, and runs them together: every step gathers the ops of
all admitted requests whose inputs exist and issues ONE batched launch per op
kind (all HMult+relin+Rescale of the step in one fides_hmult_batch, all
rotations in one fides_rotate_batch, all element-wise ops in one
fides_lin_batch, ...).  Requests that arrive while others are on the GPU join
at the next step, so their work shares launches with requests already in
flight.

Keys stay resident for the life of the server.  Ciphertexts live in one
device arena; each admitted request owns a block of it sized by a liveness
analysis of its program (registers whose values are dead share storage).

Everything runs on one HIP stream, so device order equals issue order.  The
host keeps at most two steps in flight: it builds step s+1 while the device
runs step s, and retires requests when the event after their last step has
completed.

Exact arithmetic makes each result a function of its own inputs only: the
same request gives the same bits alone and in any mix.  The scheduler's
ordering rules (below, `_compile`) are what keep that true when storage is
reused.

There is no serving loop in the reference; the operations it batches are the
reference's CryptoContext operations (api/CryptoContext.hpp:116-186).
"""
from __future__ import annotations

import bisect
import ctypes
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

# op kinds
HMUL = "hmul"          # HMult + relin + Rescale (levels l -> l-1)
HMUL_NR = "hmul_nr"    # HMult + relin, no rescale
ROT = "rot"            # key switch + automorphism
LIN = "lin"            # out = a*m (+/-) b + c
RESCALE = "rescale"
DROP = "drop"          # keep the first l_out limbs
PTDOT = "ptdot"        # sum_j ct_j (*) pt_j with device plaintexts
MODRAISE = "modraise"  # one limb at q_0 -> l_out limbs (centred lift)
LINCOMB = "lincomb"    # sum_j m_j x_j + c with integer constants (terms at any level >= the output's)
KIND_ORDER = (MODRAISE, HMUL, HMUL_NR, ROT, PTDOT, LINCOMB, LIN, RESCALE, DROP)
KEYSWITCH_KINDS = (HMUL, HMUL_NR, ROT)


class ServeError(RuntimeError):
    pass


# ----------------------------------------------------------------- programs
@dataclass
class Reg:
    id: int
    level: int            # number of Q limbs
    scale: float          # nominal scaling factor of the encoded value

    @property
    def words(self) -> int:
        return 2 * self.level


class Plaintext:
    """A device plaintext prepared for fides_ptdot_batch: evaluation form,
    Montgomery representation, `level` limbs, encoded at `scale`.  `host`
    optionally keeps the coefficient-form residues (for checking)."""

    def __init__(self, buf, level: int, scale: float, host=None):
        self.buf = buf
        self.level = level
        self.scale = scale
        self.host = host

    @property
    def ptr(self) -> int:
        return self.buf if isinstance(self.buf, int) else self.buf.ptr


@dataclass
class Op:
    kind: str
    dst: Reg
    srcs: tuple
    level: int            # level of the inputs
    arg: tuple = ()       # ROT: (autoIndex, key id), LIN: (mul, add, sub), DROP: (l_out,)


class Program:
    """A straight-line CKKS program over ciphertext registers.

    Levels are tracked in Q limbs; binary ops align their operands by
    dropping limbs of the higher one (as Ciphertext::add / mult do,
    src/CKKS/Ciphertext.cpp:229-233, 484-485).  Scales are tracked
    nominally (FIXEDMANUAL semantics: the rescaled scale is scale^2 / q_top).
    """

    def __init__(self, q, logN: int, scale_bits: int, name: str = ""):
        self.q = [int(x) for x in q]
        self.logN = logN
        self.N = 1 << logN
        self.delta = float(2 ** scale_bits)
        self.name = name
        self.regs: list[Reg] = []
        self.ops: list[Op] = []
        self.inputs: list[Reg] = []
        self.output: Reg | None = None

    # -- registers
    def _new(self, level: int, scale: float) -> Reg:
        if level < 1:
            raise ServeError("program runs out of levels")
        r = Reg(len(self.regs), level, scale)
        self.regs.append(r)
        return r

    def input(self, level: int, scale: float | None = None) -> Reg:
        if not 1 <= level <= len(self.q):
            raise ServeError("input level out of range")
        r = self._new(level, self.delta if scale is None else scale)
        self.inputs.append(r)
        return r

    def _emit(self, kind, dst, srcs, level, arg=()):
        self.ops.append(Op(kind, dst, tuple(srcs), level, tuple(arg)))
        return dst

    def drop(self, a: Reg, level: int) -> Reg:
        if level == a.level:
            return a
        if not 1 <= level < a.level:
            raise ServeError("drop needs 1 <= level < current level")
        return self._emit(DROP, self._new(level, a.scale), [a], a.level, (level,))

    def _align(self, a: Reg, b: Reg):
        l = min(a.level, b.level)
        return self.drop(a, l), self.drop(b, l)

    # -- key-switching ops
    def hmul(self, a: Reg, b: Reg, rescale: bool = True) -> Reg:
        a, b = self._align(a, b)
        l = a.level
        if rescale:
            if l < 2:
                raise ServeError("HMult+Rescale needs 2 limbs")
            out = self._new(l - 1, a.scale * b.scale / self.q[l - 1])
            return self._emit(HMUL, out, [a, b], l)
        return self._emit(HMUL_NR, self._new(l, a.scale * b.scale), [a, b], l)

    def rotate(self, a: Reg, r: int) -> Reg:
        k = pow(5, r % (self.N // 2), 2 * self.N)
        return self._emit(ROT, self._new(a.level, a.scale), [a], a.level, (k, k))

    def conjugate(self, a: Reg) -> Reg:
        return self._emit(ROT, self._new(a.level, a.scale), [a], a.level, (2 * self.N - 1, 2 * self.N - 1))

    def switch_key(self, a: Reg, key_id) -> Reg:
        """Key switch to another secret (identity automorphism) with the
        switching key registered under `key_id` (reference
        Ciphertext::keySwitch, src/CKKS/Ciphertext.cpp, as used by the
        sparse-secret encapsulation of src/CKKS/Bootstrap.cu:254-260,592-600)."""
        return self._emit(ROT, self._new(a.level, a.scale), [a], a.level, (1, key_id))

    # -- element-wise ops
    def lin(self, a: Reg, mul: int | None = None, b: Reg | None = None, sub: bool = False,
            add: int | None = None, scale: float | None = None) -> Reg:
        """out = a*mul (+/-) b + add for integers mul/add (exact)."""
        if b is not None:
            a, b = self._align(a, b)
        out = self._new(a.level, a.scale if scale is None else scale)
        srcs = [a] if b is None else [a, b]
        return self._emit(LIN, out, srcs, a.level, (mul, add, bool(sub)))

    def ptdot(self, terms) -> Reg:
        """sum_j a_j (*) pt_j for (register, Plaintext) pairs at one level."""
        if not 1 <= len(terms) <= 16:
            raise ServeError("a plaintext dot product takes 1..16 terms")
        l = min(a.level for a, _ in terms)
        regs = [self.drop(a, l) for a, _ in terms]
        pts = tuple(p for _, p in terms)
        if any(p.level < l for p in pts):
            raise ServeError("plaintext below the ciphertext level")
        scale = regs[0].scale * pts[0].scale
        return self._emit(PTDOT, self._new(l, scale), regs, l, (pts,))

    def lincomb(self, terms, add: int | None, scale: float) -> Reg:
        """sum_j m_j a_j + add for (register, integer) pairs; the result is at
        the lowest term level with nominal `scale` (terms read in place)."""
        if not 1 <= len(terms) <= 64:
            raise ServeError("a linear combination takes 1..64 terms")
        l = min(a.level for a, _ in terms)
        regs = [a for a, _ in terms]
        return self._emit(LINCOMB, self._new(l, scale), regs, l, (tuple(int(m) for _, m in terms), add))

    def mod_raise(self, a: Reg, level: int, scale: float) -> Reg:
        """One-limb ciphertext -> `level` limbs; `scale` is the nominal scale
        declared for the raised ciphertext."""
        if a.level != 1:
            a = self.drop(a, 1)
        if not 1 <= level <= len(self.q):
            raise ServeError("raise level out of range")
        return self._emit(MODRAISE, self._new(level, scale), [a], 1, (level,))

    def add(self, a: Reg, b: Reg) -> Reg:
        return self.lin(a, b=b)

    def sub(self, a: Reg, b: Reg) -> Reg:
        return self.lin(a, b=b, sub=True)

    def add_real(self, a: Reg, x: float) -> Reg:
        return self.lin(a, add=int(round(x * a.scale)))

    def mul_real(self, a: Reg, x: float) -> Reg:
        """a * x for a real constant: integer multiply by round(x*delta), then Rescale."""
        t = self.lin(a, mul=int(round(x * self.delta)), scale=a.scale * self.delta)
        return self.rescale(t)

    def match_scale(self, a: Reg, scale: float, level: int) -> Reg:
        """Same value at `scale` (to 2^-60 relative) and `level` < a.level:
        integer multiply by round(scale q / s_a) and Rescale.  In fixed
        scaling every Rescale squares a scale's deviation from q, so
        registers reached along different paths disagree; this aligns them
        at the cost of one of the source's spare levels."""
        if a.level <= level:
            raise ServeError("match_scale needs a spare level")
        q = self.q[a.level - 1]
        m = int(round(scale * q / a.scale))
        t = self.lin(a, mul=m, scale=a.scale * m)
        return self.drop(self.rescale(t), level)

    def rescale(self, a: Reg) -> Reg:
        if a.level < 2:
            raise ServeError("Rescale needs 2 limbs")
        return self._emit(RESCALE, self._new(a.level - 1, a.scale / self.q[a.level - 1]), [a], a.level)

    def sum_tree(self, regs: list) -> Reg:
        regs = list(regs)
        while len(regs) > 1:
            nxt = [self.add(regs[i], regs[i + 1]) for i in range(0, len(regs) - 1, 2)]
            if len(regs) % 2:
                nxt.append(regs[-1])
            regs = nxt
        return regs[0]

    def set_output(self, r: Reg) -> None:
        self.output = r

    # -- statistics
    def count(self, kinds) -> int:
        return sum(1 for o in self.ops if o.kind in kinds)


# ----------------------------------------------------------------- request programs
def chebyshev_coefficients(f, degree: int, lo: float, hi: float) -> np.ndarray:
    """Chebyshev interpolant of f on [lo, hi] (coefficients of T_k((2x-lo-hi)/(hi-lo)))."""
    k = np.arange(degree + 1)
    y = np.cos(np.pi * (k + 0.5) / (degree + 1))
    fx = np.array([f(0.5 * (hi - lo) * t + 0.5 * (hi + lo)) for t in y])
    c = np.array([2.0 / (degree + 1) * np.sum(fx * np.cos(np.pi * j * (k + 0.5) / (degree + 1)))
                  for j in range(degree + 1)])
    c[0] /= 2
    return c


def chebyshev_eval_depth(coeffs) -> int:
    """Levels consumed by `chebyshev_eval` for these coefficients (structural:
    a dry run on a program with 64 levels)."""
    P = Program([float(1 << 59)] * 64, 12, 59, "depth")
    y = P.input(64)
    return 64 - chebyshev_eval(P, y, coeffs).level


def chebyshev_depth(degree: int) -> int:
    """Levels consumed by `chebyshev_program` of this degree (input map and
    series; coefficients all non-zero, the deepest case)."""
    return 1 + chebyshev_eval_depth([1.0] * (degree + 1))


def chebyshev_program(prm, level: int, coeffs, lo: float, hi: float) -> Program:
    """Evaluate sum_k c_k T_k(y), y = (2x - lo - hi)/(hi - lo), on one ciphertext
    (see chebyshev_eval)."""
    P = Program(prm.q, prm.logN, prm.scale_bits, "chebyshev%d" % (len(coeffs) - 1))
    x = P.input(level)
    # y = a*x + b with a = 2/(hi-lo), b = -(hi+lo)/(hi-lo); none on [-1, 1]
    a = 2.0 / (hi - lo)
    b = -(hi + lo) / (hi - lo)
    y = x
    if not (a == 1.0 and b == 0.0):
        y = P.mul_real(x, a)
        if b != 0.0:
            y = P.add_real(y, b)
    P.set_output(chebyshev_eval(P, y, coeffs))
    return P


def _chebyshev_basis(P: Program, y: Reg, upto: int) -> dict:
    """T_1..T_upto by the product rule T_{m+n} = 2 T_m T_n - T_{|m-n|} with m
    the largest power of two below m+n (T_1..T_{2^j} ready after j HMult
    levels); each T costs one HMult+Rescale and one fused element-wise op."""
    T = {1: y}
    m = 1
    while m < upto:
        for k in range(m + 1, min(2 * m, upto) + 1):
            n = k - m
            prod = P.hmul(T[m], T[n])
            if n == m:  # T_2m = 2 T_m^2 - 1
                T[k] = P.lin(prod, mul=2, add=-int(round(prod.scale)))
            else:       # T_{m+n} = 2 T_m T_n - T_{m-n}, T_{m-n} brought to the product's scale
                b = T[m - n]
                if b.scale != prod.scale and b.level > prod.level:
                    b = P.match_scale(b, prod.scale, prod.level)
                T[k] = P.lin(prod, mul=2, b=b, sub=True)
        m *= 2
    return T


def chebyshev_eval(P: Program, y: Reg, coeffs) -> Reg:
    """sum_k c_k T_k(y) for a register whose values lie in [-1, 1], baby-step /
    giant-step (the reference's EvalChebyshevSeriesPS, src/PolyApprox.cu, and
    OpenFHE's Paterson-Stockmeyer variant for Chebyshev series): baby steps
    T_1..T_m (m the power of two near sqrt(deg)), giant steps T_m, T_2m, T_4m..
    by doubling, and a recursive Chebyshev division
        p = q T_M + r,  q = c_M + sum_{k>=1} 2 c_{M+k} T_k,  r_j = c_j - c_{2M-j}
    (from T_{M+k} = 2 T_M T_k - T_{M-k}) down to pieces of degree < m, each
    one fused integer linear combination of the baby steps.  deg = 55: 15
    HMults instead of 54 for the full product-rule basis, same depth
    (ceil(log2 deg) + 1 levels).

    Scales: every node is built for a required output scale, so products and
    remainders meet on exactly the same nominal scale; leaves fold each T_k's
    own scale into its integer constant."""
    c = [float(x) for x in coeffs]
    deg = len(c) - 1
    while deg > 0 and c[deg] == 0.0:
        deg -= 1
    c = c[:deg + 1]
    m = 2
    while m * m < deg + 1:
        m *= 2
    T = _chebyshev_basis(P, y, max(1, min(m, deg)))
    M = m
    while 2 * M <= deg:
        sq = P.hmul(T[M], T[M])
        T[2 * M] = P.lin(sq, mul=2, add=-int(round(sq.scale)))
        M *= 2
    q = P.q

    def giant(n):  # largest giant step M (m 2^i) with n <= 2M, n > M
        g = m
        while 2 * g < n:
            g *= 2
        return g

    def split(cs, Mg):
        qc = [cs[Mg]] + [2.0 * cs[Mg + k] for k in range(1, len(cs) - Mg)]
        rc = [cs[j] - (cs[2 * Mg - j] if 2 * Mg - j < len(cs) else 0.0) for j in range(Mg)]
        return qc, rc

    def leaf_terms(cs):
        ks = [k for k in range(1, len(cs)) if cs[k] != 0.0]
        return ks if ks else [1]

    def level(cs):
        """Level of the node's result (structure only)."""
        if len(cs) <= m:
            return min(T[k].level for k in leaf_terms(cs)) - 1
        Mg = giant(len(cs))
        qc, rc = split(cs, Mg)
        return min(min(level(qc), T[Mg].level) - 1, level(rc))

    def build(cs, S):
        """Register holding sum_k cs_k T_k at nominal scale S."""
        if len(cs) <= m:
            ks = leaf_terms(cs)
            l = min(T[k].level for k in ks)
            target = S * q[l - 1]
            terms = [(T[k], int(round(cs[k] * target / T[k].scale)) if k < len(cs) else 0) for k in ks]
            return P.rescale(P.lincomb(terms, add=int(round(cs[0] * target)), scale=target))
        Mg = giant(len(cs))
        qc, rc = split(cs, Mg)
        lp = min(level(qc), T[Mg].level)  # product level before its Rescale
        qr = build(qc, S * q[lp - 1] / T[Mg].scale)
        prod = P.hmul(qr, T[Mg])
        return P.add(prod, build(rc, S))

    if deg < 1:
        raise ServeError("a Chebyshev series needs degree >= 1")
    return build(c, P.delta)


def rotsum_program(prm, level: int, k: int) -> Program:
    """Sum of 2^k consecutive slots into every slot: k rounds of x += rot(x, 2^i)
    (the reference's Accumulate, src/CKKS/AccumulateBroadcast.cu)."""
    P = Program(prm.q, prm.logN, prm.scale_bits, "rotsum%d" % k)
    x = P.input(level)
    for i in range(k):
        x = P.add(x, P.rotate(x, 1 << i))
    P.set_output(x)
    return P


def rotation_indices_rotsum(kmax: int) -> list:
    return [1 << i for i in range(kmax)]


# ----------------------------------------------------------------- compile: storage and ordering
@dataclass
class Compiled:
    """Storage and ordering of one program.

    slot_of[reg] -> slot; slot_off/slot_size in limbs (units of N words);
    block = limbs the request owns.  deps[i] = ops that must be issued in an
    EARLIER step than op i: the producers of its inputs, and every op that
    wrote or read the previous value held in the slot op i writes (write
    after read / write after write on reused storage).  Ops of one step are
    issued grouped by kind, not in program order, so an op may never share a
    step with an op it depends on.
    """
    slot_of: dict
    slot_off: list
    slot_size: list
    block: int
    deps: list
    dependents: list
    asap: list = field(default_factory=list)


def _compile(P: Program) -> Compiled:
    """Storage by step-aware liveness.  Each op gets its ASAP step (one more
    than its latest producer; inputs are step 0), which is the step the loop
    issues it in when storage adds no ordering.  A value occupies its slot
    from its defining step through its last reader's step, and a slot is
    reused only by an op in a strictly later step, so the storage ordering
    recorded in deps never delays an op past its data-ready step."""
    if P.output is None:
        raise ServeError("program has no output")
    nops = len(P.ops)
    producer = {}
    asap = [0] * nops
    for i, o in enumerate(P.ops):
        st = 0
        for s_ in o.srcs:
            if s_.id in producer:
                st = max(st, asap[producer[s_.id]])
            elif s_.id not in {r.id for r in P.inputs}:
                raise ServeError("register %d read before it is written" % s_.id)
        asap[i] = st + 1
        producer[o.dst.id] = i
    INF = 1 << 60
    born = {r.id: 0 for r in P.inputs}
    last = {r.id: 0 for r in P.inputs}
    for i, o in enumerate(P.ops):
        born[o.dst.id] = asap[i]
        last.setdefault(o.dst.id, asap[i])
        for s_ in o.srcs:
            last[s_.id] = max(last[s_.id], asap[i])
    last[P.output.id] = INF
    reg = {r.id: r for r in P.inputs}
    for o in P.ops:
        reg[o.dst.id] = o.dst
    # allocation in step order (program order within a step)
    order = sorted(range(nops), key=lambda i: (asap[i], i))
    slot_size, slot_free, slot_of, hist = [], [], {}, []   # slot_free[s]: last step of its current value

    def alloc(rid, step):
        need = reg[rid].words
        best = None
        for sl, sz in enumerate(slot_size):
            if slot_free[sl] < step and sz >= need and (best is None or sz < slot_size[best]):
                best = sl
        if best is None:
            best = len(slot_size)
            slot_size.append(need)
            slot_free.append(0)
            hist.append([])
        slot_of[rid] = best
        slot_free[best] = last[rid]
        return best

    for r in P.inputs:
        alloc(r.id, 0)
        slot_free[slot_of[r.id]] = last[r.id]
    deps = [None] * nops
    for i in order:
        o = P.ops[i]
        d = {producer[s_.id] for s_ in o.srcs if s_.id in producer}
        sl = alloc(o.dst.id, asap[i])
        d.update(hist[sl])            # previous occupant's writer and readers
        hist[sl] = [i]
        d.discard(i)
        deps[i] = sorted(d)
        for s_ in o.srcs:
            hist[slot_of[s_.id]].append(i)
    dependents = [[] for _ in range(nops)]
    for i, d in enumerate(deps):
        for j in d:
            dependents[j].append(i)
    off, acc = [], 0
    for sz in slot_size:
        off.append(acc)
        acc += sz
    return Compiled(slot_of, off, slot_size, acc, deps, dependents, asap)


# ----------------------------------------------------------------- arena
class Arena:
    """First-fit allocator over one device allocation (units: words)."""

    def __init__(self, base_ptr: int, words: int):
        self.base = int(base_ptr)
        self.words = int(words)
        self.free = [(0, self.words)]  # sorted (offset, size)
        self.used = 0
        self.peak = 0

    def alloc(self, words: int):
        for i, (o, s) in enumerate(self.free):
            if s >= words:
                if s == words:
                    self.free.pop(i)
                else:
                    self.free[i] = (o + words, s - words)
                self.used += words
                self.peak = max(self.peak, self.used)
                return o
        return None

    def release(self, off: int, words: int) -> None:
        i = bisect.bisect_left(self.free, (off, 0))
        self.free.insert(i, (off, words))
        # coalesce with neighbours
        if i + 1 < len(self.free) and off + words == self.free[i + 1][0]:
            self.free[i] = (off, words + self.free[i + 1][1])
            self.free.pop(i + 1)
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == self.free[i][0]:
            self.free[i - 1] = (self.free[i - 1][0], self.free[i - 1][1] + self.free[i][1])
            self.free.pop(i)
        self.used -= words


# ----------------------------------------------------------------- backends
class GpuBackend:
    """Issues batched C ABI calls on one stream: the default stream, or a
    capi.Stream (needed for graph capture)."""

    def __init__(self, ctx, relin_key, rotation_keys: dict, arena_bytes: int, stream=None):
        from . import capi
        self.capi = capi
        self.ctx = ctx
        self.N = ctx.N
        self.rlk = relin_key
        self.keys = dict(rotation_keys)          # key id (autoIndex, or a switching-key name) -> capi.Key
        self.buf = capi.DeviceBuffer(arena_bytes // 8)
        self.base = self.buf.ptr
        self.words = self.buf.nwords
        self.stream = stream
        self.st = stream.h if stream is not None else None
        self.hip = _Hip()

    def ptr(self, word_off: int) -> int:
        return self.base + 8 * word_off

    def key_for(self, k):
        if k not in self.keys:
            raise ServeError("no key switching key %r" % (k,))
        return self.keys[k]

    def hmul(self, a, b, levels, out, rescale):
        self.ctx.hmult_batch(a, b, levels, self.rlk, out, rescale=rescale, stream=self.st)

    def rot(self, a, levels, kidx, out, keyid=None):
        keyid = kidx if keyid is None else keyid
        self.ctx.rotate_batch(a, levels, [self.key_for(k) for k in keyid], kidx, out, stream=self.st)

    def lin(self, a, b, levels, mul, add, sub, out):
        self.ctx.lin_batch(a, b, levels, mul, add, sub, out, stream=self.st)

    def rescale(self, a, levels, out):
        self.ctx.rescale_batch(a, levels, out, stream=self.st)

    def drop(self, a, l_in, l_out, out):
        self.ctx.drop_limbs_batch(a, l_in, l_out, out, stream=self.st)

    def ptdot(self, terms, levels, out):
        self.ctx.ptdot_batch(terms, levels, out, stream=self.st)

    def lincomb(self, terms, levels, add, out):
        self.ctx.lincomb_batch(terms, levels, add, out, stream=self.st)

    def modraise(self, a, l_out, out):
        self.ctx.mod_raise_batch(a, l_out, out, stream=self.st)

    def upload(self, ptr, host):
        arr = np.ascontiguousarray(host, np.uint64)
        self.capi.check(self.capi.load().fides_copy_to_device(ptr, arr.ctypes.data, arr.size * 8, self.st), "upload")

    def stage(self, host):
        """A device copy of a host array (graph-capture inputs)."""
        return self.capi.DeviceBuffer.from_array(np.ascontiguousarray(host, np.uint64), stream=self.st)

    def copy(self, dst, src, words):
        self.capi.check(self.capi.load().fides_copy_device(dst, src, words * 8, self.st), "copy")

    def download(self, ptr, words) -> np.ndarray:
        out = np.empty(words, np.uint64)
        self.capi.check(self.capi.load().fides_copy_to_host(out.ctypes.data, ptr, words * 8, self.st), "download")
        return out

    def record(self):
        return self.hip.record(self.st)

    def sync(self) -> None:
        self.capi.check(self.capi.load().fides_stream_sync(self.st), "sync")

    # -- graph capture of the issued operations (capi.Context.graph_begin)
    def begin_capture(self, scratch_bytes: int, task_bytes: int, scratch_ptr: int | None = None) -> None:
        if self.stream is None:
            raise ServeError("graph capture needs a backend created with a stream")
        self.ctx.graph_begin(self.stream, scratch_bytes, task_bytes, scratch_ptr)

    def end_capture(self):
        return self.ctx.graph_end()

    def launch(self, graph) -> None:
        graph.launch(self.stream)

    def wait(self, ev) -> None:
        self.hip.wait(ev)

    def query(self, ev) -> bool:
        return self.hip.query(ev)

    def release(self, ev) -> None:
        self.hip.release(ev)


class _Hip:
    """HIP events (plumbing for the step pipeline)."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")
        self.lib.hipEventCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
        self.lib.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.lib.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.lib.hipEventDestroy.argtypes = [ctypes.c_void_p]
        self.lib.hipEventQuery.argtypes = [ctypes.c_void_p]
        self.pool = []

    def record(self, stream=None):
        if self.pool:
            e = self.pool.pop()
        else:
            e = ctypes.c_void_p()
            if self.lib.hipEventCreateWithFlags(ctypes.byref(e), 2) != 0:  # hipEventDisableTiming
                raise ServeError("hipEventCreate failed")
        if self.lib.hipEventRecord(e, stream) != 0:
            raise ServeError("hipEventRecord failed")
        return e

    def wait(self, e) -> None:
        if self.lib.hipEventSynchronize(e) != 0:
            raise ServeError("hipEventSynchronize failed")
        self.pool.append(e)

    def query(self, e) -> bool:
        """True once the work recorded before `e` has finished."""
        rc = self.lib.hipEventQuery(e)
        if rc == 0:
            return True
        if rc == 600:  # hipErrorNotReady
            return False
        raise ServeError("hipEventQuery failed (%d)" % rc)

    def release(self, e) -> None:
        self.pool.append(e)


# ----------------------------------------------------------------- requests and the loop
@dataclass
class Request:
    rid: int
    program: Program
    arrival: float                      # seconds after the loop starts
    inputs: list = field(default_factory=list)   # host arrays [2][l][N] or ("dev", ptr) sources
    keep_result: bool = False
    # filled by the server
    admitted_at: float = -1.0
    done_at: float = -1.0
    result: np.ndarray | None = None
    status: str = "pending"
    compiled: Compiled | None = None
    graph: bool = False                 # served by a GraphCache replay
    signature: tuple | None = None      # program_signature, computed once when a GraphCache looks it up

    @property
    def latency(self) -> float:
        return self.done_at - self.arrival


class _Active:
    __slots__ = ("req", "comp", "off", "remaining", "ready", "left", "last_step", "N")

    def __init__(self, req, comp, off, N):
        self.req, self.comp, self.off, self.N = req, comp, off, N
        self.remaining = [len(d) for d in comp.deps]
        self.ready = [i for i, n in enumerate(self.remaining) if n == 0]
        self.left = len(self.remaining)
        self.last_step = -1

    def reg_ptr_words(self, reg) -> int:
        return (self.off + self.comp.slot_off[self.comp.slot_of[reg.id]]) * self.N


class Server:
    """Continuous-batching loop.  `backend` provides the batched ops, device
    memory (`base`, `words`, `ptr`) and events."""

    def __init__(self, backend, N: int, max_inflight_steps: int = 2):
        self.be = backend
        self.N = N
        self.arena = Arena(0, backend.words // N)   # units: limbs of N words
        self.max_inflight = max(1, max_inflight_steps)
        self.stats = dict(steps=0, launches=0, ops=0, keyswitch_ops=0, rejected=0)

    def footprint_bytes(self, program: Program) -> int:
        return _compile(program).block * self.N * 8

    # -- one scheduling step: every ready op of every active request, one launch per kind
    def _step(self, active, step_no):
        groups = {k: [] for k in KIND_ORDER}
        for a in active:
            if not a.ready:
                continue
            ops = a.req.program.ops
            issued, a.ready = a.ready, []
            for i in issued:
                groups[ops[i].kind].append((a, ops[i]))
            # dependents become ready for the NEXT step
            for i in issued:
                for j in a.comp.dependents[i]:
                    a.remaining[j] -= 1
                    if a.remaining[j] == 0:
                        a.ready.append(j)
            a.left -= len(issued)
            a.last_step = step_no
        be = self.be
        nops = 0
        for kind in KIND_ORDER:
            g = groups[kind]
            if not g:
                continue
            nops += len(g)
            ptr = lambda a, r: be.ptr(a.reg_ptr_words(r))
            s0 = [ptr(a, o.srcs[0]) for a, o in g]
            dst = [ptr(a, o.dst) for a, o in g]
            lv = [o.level for _, o in g]
            if kind in (HMUL, HMUL_NR):
                be.hmul(s0, [ptr(a, o.srcs[1]) for a, o in g], lv, dst, kind == HMUL)
            elif kind == ROT:
                be.rot(s0, lv, [o.arg[0] for _, o in g], dst, [o.arg[1] for _, o in g])
            elif kind == LIN:
                be.lin(s0, [ptr(a, o.srcs[1]) if len(o.srcs) > 1 else None for a, o in g], lv,
                       [o.arg[0] for _, o in g], [o.arg[1] for _, o in g], [o.arg[2] for _, o in g], dst)
            elif kind == RESCALE:
                be.rescale(s0, lv, dst)
            elif kind == DROP:
                be.drop(s0, lv, [o.arg[0] for _, o in g], dst)
            elif kind == LINCOMB:
                be.lincomb([[(ptr(a, r), r.level, m) for r, m in zip(o.srcs, o.arg[0])] for a, o in g], lv,
                           [o.arg[1] for _, o in g], dst)
            elif kind == PTDOT:
                be.ptdot([[(ptr(a, r), p.ptr) for r, p in zip(o.srcs, o.arg[0])] for a, o in g], lv, dst)
            elif kind == MODRAISE:
                be.modraise(s0, [o.arg[0] for _, o in g], dst)
            self.stats["launches"] += 1
            if kind in KEYSWITCH_KINDS:
                self.stats["keyswitch_ops"] += len(g)
        self.stats["ops"] += nops
        return nops

    def _admit(self, req: Request, now: float):
        comp = req.compiled
        off = self.arena.alloc(comp.block)
        if off is None:
            return None
        a = _Active(req, comp, off, self.N)
        for r, src in zip(req.program.inputs, req.inputs):
            p = self.be.ptr(a.reg_ptr_words(r))
            if isinstance(src, tuple) and src[0] == "dev":
                self.be.copy(p, src[1], r.words * self.N)
            else:
                self.be.upload(p, src)
        req.admitted_at = now
        req.status = "running"
        return a

    def _retire(self, a: _Active, now: float):
        req = a.req
        if req.keep_result:
            out = req.program.output
            req.result = self.be.download(self.be.ptr(a.reg_ptr_words(out)), out.words * self.N).reshape(
                2, out.level, self.N)
        self.arena.release(a.off, a.comp.block)
        req.done_at = now
        req.status = "done"

    def capture(self, requests, scratch_bytes: int = 4 << 30, task_bytes: int = 512 << 20,
                scratch_ptr: int | None = None) -> "CapturedBatch":
        """A closed batch recorded into one HIP graph: every request gets its
        arena storage now, then the copy of the inputs into it (one copy
        kernel) and every step of the same schedule `run` would issue are
        captured instead of run.  Device inputs ("dev", ptr) are read at
        every replay from where they are; host inputs are staged in device
        buffers owned by the batch.  `CapturedBatch.replay()` executes it;
        results are read after a replay.  All requests must fit the arena
        at once.  Requests with no inputs listed read their input registers
        in place (the caller writes them before each replay).  scratch_ptr:
        key-switch scratch shared with other graphs (fides_graph_begin_shared)."""
        active, staged, src, dst, lv = [], [], [], [], []
        try:
            for r in requests:
                r.compiled = _compile(r.program)
                off = self.arena.alloc(r.compiled.block)
                if off is None:
                    raise ServeError("captured requests do not fit the arena together")
                a = _Active(r, r.compiled, off, self.N)
                active.append(a)
                for reg, x in zip(r.program.inputs, r.inputs):
                    if isinstance(x, tuple) and x[0] == "dev":
                        src.append(x[1])
                    else:
                        buf = self.be.stage(x)
                        staged.append(buf)
                        src.append(buf.ptr)
                    dst.append(self.be.ptr(a.reg_ptr_words(reg)))
                    lv.append(reg.level)
                r.status = "running"
            self.be.sync()
            if scratch_ptr is None:
                self.be.begin_capture(scratch_bytes, task_bytes)
            else:
                self.be.begin_capture(scratch_bytes, task_bytes, scratch_ptr)
            try:
                if src:
                    self.be.drop(src, lv, lv, dst)    # the input copy, part of every replay
                step_no = 0
                while any(a.ready for a in active):
                    self._step(active, step_no)
                    step_no += 1
                    self.stats["steps"] += 1
            except BaseException:
                try:
                    self.be.end_capture().close()
                except Exception:
                    pass
                raise
            graph = self.be.end_capture()
        except BaseException:
            for a in active:
                self.arena.release(a.off, a.comp.block)
            raise
        return CapturedBatch(self, graph, active, staged)

    def run(self, requests, realtime: bool = True, clock=time.perf_counter, graphs: "GraphCache | None" = None):
        """Serve `requests` (arrival times relative to the start).  With
        realtime=False every request is available at once (closed batch).
        graphs: a GraphCache; a request whose program has a free captured
        graph runs as one replay of it (on the cache's stream, concurrent
        with this loop's steps) instead of joining the batched steps."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            r.compiled = _compile(r.program)
            # too large for the arena: rejected, unless a cached graph (with
            # storage of its own) can serve it
            if r.compiled.block > self.arena.words and (graphs is None or not graphs.has_shape(r)):
                r.status = "rejected"
                self.stats["rejected"] += 1
        pending = deque(r for r in reqs if r.status != "rejected")
        queue = deque()
        active = []
        inflight = deque()   # (event, step_no)
        step_no = 0
        t0 = clock()
        now = lambda: clock() - t0

        def drain_one():
            ev, s = inflight.popleft()
            self.be.wait(ev)
            t = now()
            for a in [x for x in active if x.left == 0 and x.last_step <= s]:
                self._retire(a, t)
                active.remove(a)

        while pending or queue or active or (graphs is not None and graphs.inflight):
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                queue.append(pending.popleft())
            if graphs is not None and graphs.inflight:
                graphs.poll(now)
            while queue:
                if graphs is not None:
                    e = graphs.free_entry(queue[0])
                    if e is not None:
                        graphs.start(e, queue.popleft(), t)
                        continue
                a = self._admit(queue[0], t)
                if a is None:
                    break
                active.append(a)
                queue.popleft()
            if any(a.ready for a in active):
                self._step(active, step_no)
                inflight.append((self.be.record(), step_no))
                step_no += 1
                self.stats["steps"] += 1
                while len(inflight) >= self.max_inflight:
                    drain_one()
            elif inflight:
                drain_one()
            elif graphs is not None and graphs.inflight:
                if pending or queue:
                    time.sleep(GRAPH_POLL_S)   # an arrival or a free slot may come first
                else:
                    graphs.wait_one(now)
            elif pending and not queue:
                wait = pending[0].arrival - now()
                if wait > 0:
                    time.sleep(min(wait, 0.01))
            elif queue and not active:
                raise ServeError("request %d does not fit the arena" % queue[0].rid)
        while inflight:
            drain_one()
        self.stats["wall_s"] = now()
        return reqs


GRAPH_POLL_S = 2e-4   # host poll interval while only captured graphs run


def program_signature(P: Program) -> tuple:
    """Structural identity of a request program: programs with equal
    signatures issue the same device work on the same register layout (the
    tracked scales are nominal and never reach the device; the integer
    constants they produce are part of the op arguments)."""
    def arg(o):
        return tuple(p.ptr for p in o.arg[0]) if o.kind == PTDOT else o.arg
    return (P.logN, tuple((r.id, r.level) for r in P.inputs), P.output.id if P.output else None,
            tuple((o.kind, o.dst.id, o.dst.level, tuple(x.id for x in o.srcs), o.level, arg(o)) for o in P.ops))


class _GraphEntry:
    __slots__ = ("batch", "active", "lane", "busy", "req", "event")

    def __init__(self, batch, lane):
        self.batch = batch
        self.active = batch.active[0]
        self.lane = lane
        self.busy = False
        self.req = None
        self.event = None


class _Lane:
    """One stream of the cache: a backend (its stream and register memory), a
    Server over that memory for capturing, and the key-switch scratch every
    graph of the lane shares."""

    def __init__(self, backend, N, scratch_ptr):
        self.be = backend
        self.srv = Server(backend, N)
        self.scratch_ptr = scratch_ptr


class GraphCache:
    """Request programs captured once as HIP graphs and replayed with new
    inputs (the reference caches CUDA graphs of its key-switch sequence,
    src/CKKS/LimbPartitionMGPU.cu:601-609; here a whole request program is
    one graph).

    The cache is warm: the caller captures the program shapes it expects
    (say every start level of its Chebyshev and bootstrap programs) before
    serving; a shape is identified by program_signature.  Each captured
    program owns a register block in the device memory of its lane and a
    task pool; both stay resident for the life of the cache, outside the
    serving loop's admission cap.  A lane is one stream with its own backend
    (register memory) and one key-switch scratch region that all of its
    graphs share, so a lane runs its graphs one at a time; different lanes,
    and the batched loop's steps, run concurrently.  Putting long programs
    (bootstraps) on a lane of their own keeps short ones from waiting behind
    them.  A replay first writes the request's inputs into the captured input
    registers, then launches the graph; an entry stays busy until its request
    is retired (the result read back), so `copies` entries of a shape serve
    that many requests at once."""

    def __init__(self, backend, N: int, scratch_bytes: int, scratch_ptr: int | None = None,
                 task_bytes: int = 16 << 20):
        self.N = N
        self.lanes = [_Lane(backend, N, scratch_ptr)]
        self.scratch_bytes = scratch_bytes
        self.task_bytes = task_bytes
        self.entries = {}        # signature -> [_GraphEntry]
        self.inflight = []       # entries running a request, in launch order
        self.stats = dict(shapes=0, graphs=0, replays=0, capture_s=0.0, task_bytes=0)

    @property
    def be(self):
        return self.lanes[0].be

    def add_lane(self, backend, scratch_ptr: int | None = None) -> int:
        """Another stream (backend with its own stream and register memory,
        and its own scratch); returns its index for add()."""
        self.lanes.append(_Lane(backend, self.N, scratch_ptr))
        return len(self.lanes) - 1

    @staticmethod
    def words_needed(programs, copies: int = 1) -> int:
        """Device words the register blocks of `programs` take (distinct
        signatures, `copies` each): the size of a lane backend's arena."""
        seen, total = set(), 0
        for P in programs:
            sig = program_signature(P)
            if sig not in seen:
                seen.add(sig)
                total += copies * _compile(P).block * P.N
        return total

    def add(self, program: Program, copies: int = 1, lane: int = 0) -> None:
        """Capture `program`'s shape on `lane` until it has `copies` entries."""
        sig = program_signature(program)
        lst = self.entries.setdefault(sig, [])
        if not lst:
            self.stats["shapes"] += 1
        L = self.lanes[lane]
        t0 = time.perf_counter()
        for _ in range(copies - len(lst)):
            req = Request(-1, program, 0.0, inputs=[])   # inputs are written in place before each replay
            tb = self.task_bytes
            while True:
                try:
                    batch = L.srv.capture([req], self.scratch_bytes, tb, L.scratch_ptr)
                    break
                except Exception as e:  # a task pool too small for the program: capture again, larger
                    if "task pool exhausted" not in str(e) or tb >= (1 << 31):
                        raise
                    tb *= 2
            self.stats["task_bytes"] += tb
            lst.append(_GraphEntry(batch, lane))
            self.stats["graphs"] += 1
        self.stats["capture_s"] += time.perf_counter() - t0

    def has_shape(self, req: Request) -> bool:
        if req.signature is None:
            req.signature = program_signature(req.program)
        return req.signature in self.entries

    def free_entry(self, req: Request):
        if req.signature is None:
            req.signature = program_signature(req.program)
        for e in self.entries.get(req.signature, ()):
            if not e.busy:
                return e
        return None

    def start(self, e: _GraphEntry, req: Request, now: float) -> None:
        be = self.lanes[e.lane].be
        a = e.active
        for r, src in zip(a.req.program.inputs, req.inputs):
            p = be.ptr(a.reg_ptr_words(r))
            if isinstance(src, tuple) and src[0] == "dev":
                be.copy(p, src[1], r.words * self.N)
            else:
                be.upload(p, src)
        be.launch(e.batch.graph)
        e.event = be.record()
        e.busy, e.req = True, req
        req.admitted_at = now
        req.status = "running"
        req.graph = True
        self.inflight.append(e)
        self.stats["replays"] += 1

    def _retire(self, e: _GraphEntry, t: float) -> None:
        req = e.req
        if req.keep_result:
            out = e.active.req.program.output
            req.result = self._read_output(self.lanes[e.lane].be, e.active, out)
        req.done_at = t
        req.status = "done"
        e.busy, e.req, e.event = False, None, None
        self.inflight.remove(e)

    def _read_output(self, be, a, out):
        return be.download(be.ptr(a.reg_ptr_words(out)), out.words * self.N).reshape(2, out.level, self.N)

    def poll(self, now) -> None:
        """Retire every request whose graph has finished (non-blocking)."""
        for e in list(self.inflight):
            be = self.lanes[e.lane].be
            if be.query(e.event):
                be.release(e.event)
                self._retire(e, now())

    def wait_one(self, now) -> None:
        """Block until the oldest running graph finishes, then retire it."""
        e = self.inflight[0]
        self.lanes[e.lane].be.wait(e.event)
        self._retire(e, now())

    def close(self) -> None:
        for lst in self.entries.values():
            for e in lst:
                e.batch.close()
        self.entries = {}
        self.inflight = []


class CapturedBatch:
    """Requests recorded into one HIP graph by Server.capture."""

    def __init__(self, server: Server, graph, active, staged):
        self.server, self.graph, self.active, self.staged = server, graph, active, staged

    def replay(self) -> None:
        self.server.be.launch(self.graph)

    def results(self) -> list:
        """Outputs of the last replay (waits for it), in request order."""
        be = self.server.be
        be.sync()
        out = []
        for a in self.active:
            o = a.req.program.output
            out.append(be.download(be.ptr(a.reg_ptr_words(o)), o.words * be.N).reshape(2, o.level, be.N))
        return out

    def close(self) -> None:
        if self.graph is not None:
            self.graph.close()
            self.graph = None
            for a in self.active:
                self.server.arena.release(a.off, a.comp.block)
            self.active = []
            self.staged = []


def latency_summary(reqs) -> dict:
    done = [r for r in reqs if r.status == "done"]
    if not done:
        return {}
    lat = np.array([r.latency for r in done]) * 1e3
    return dict(n=len(done), p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                mean_ms=float(lat.mean()), max_ms=float(lat.max()))
