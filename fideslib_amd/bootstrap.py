"""CKKS bootstrapping built from the batched device operations.

Algorithm (the reference's Bootstrap, src/CKKS/Bootstrap.cu:169-333, on
precomputation it takes from OpenFHE; here every table is derived from the
mathematics and computed on the host):

  1. prescale   x <- x * 2^-c, Rescale, keep one limb (emulates a larger q_0
                so the sine below is near-linear; undone by an exact integer
                multiply by 2^c at the end, as the reference's correction factor;
                under FLEXIBLEAUTO the same multiply also brings the input
                level's scaling factor to delta)
  2. ModRaise   one limb at q_0 -> L limbs: the ciphertext now encrypts
                t = m + q_0 I(X) with a small integer polynomial I.  With
                sparse-secret encapsulation (reference Bootstrap.cu:254-260,
                592-600) the one-limb ciphertext is first switched to a
                second, sparser secret s' and switched back to s right after
                ModRaise, so |I| is bounded by the weight of s', not of s
  3. CoeffToSlot  the inverse of the decoding map, so that slot k holds the
                complex-packed coefficients t_k + i t_{k+n} (bit-reversed slot
                order), scaled to y = t / (q_0 K')
  4. split      Re = ct + conj(ct), Im = (ct - conj(ct)) * (-i)  (conj = the
                automorphism 2N-1, -i = the monomial X^{3N/2})
  5. EvalMod    cos(2 pi (K' y - 1/4) / 2^r) by a Chebyshev series, then r
                double-angle steps c <- 2c^2 - 1, giving sin(2 pi t/q_0) ~
                2 pi m / q_0 on Re and Im (reference approxModReduction,
                src/CKKS/ApproxModEval.cu)
  6. recombine  Re + i Im  (i = the monomial X^{N/2})
  7. SlotToCoeff  the decoding map, scaled back by q_0 / (2 pi s)
  8. x 2^c      exact integer multiply

Sparse packing (cfg.slots = n_s < N/2, the reference's sparse path,
src/CKKS/Bootstrap.cu:100,116-143): the message lives in the subring of
X^(N / 2 n_s).  After ModRaise the partial trace (x += Rot(x, j n_s),
j = 1, 2, 4, ...) keeps that subring's part, times N / (2 n_s) (folded into
CoeffToSlot); the transforms are those of n_s slots on 2 n_s-periodic slot
vectors, CoeffToSlot leaves [2 Re w | 2 Im w] in one ciphertext so that one
EvalMod serves both halves, and SlotToCoeff recombines them (sparse_fold).

Decoding map: z_j = sum_k w_k zeta^(5^j k) for the complex-packed
coefficients w.  It factors as U = B_logn ... B_1 BR (BR the bit reversal,
B_s the butterfly stage of length 2^s with twiddles exp(2 pi i (5^j mod 2^(s+2))
/ 2^(s+2))).  The bit reversal is dropped on both sides (EvalMod is
slot-wise), and the stages are merged into `budget` groups (the reference's
level budget {3, 3}); a group of b stages has at most 2^(b+1) - 1 non-zero
diagonals, all multiples of one stride.  Each group is one level: a
baby-step / giant-step evaluation (hoisted baby rotations, plaintext dot
products of up to `g` terms, giant rotations, one Rescale), with plaintexts
encoded at scale q_top so the ciphertext scale is unchanged.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import codec
from .serve import Plaintext, Program, Reg, chebyshev_coefficients, chebyshev_eval, chebyshev_eval_depth


# ----------------------------------------------------------------- linear algebra on diagonals
def bitrev(n: int) -> np.ndarray:
    b = n.bit_length() - 1
    r = np.zeros(n, np.int64)
    for i in range(n):
        r[i] = int(format(i, "0%db" % b)[::-1], 2) if b else 0
    return r


def rot(v: np.ndarray, d: int) -> np.ndarray:
    """Slot rotation: rot(v, d)[p] = v[(p + d) mod n] (a Rotate by d)."""
    return np.roll(v, -d)


def stage_diags(n: int, N: int, s: int, inverse: bool) -> dict:
    """Butterfly stage of length 2^s as {offset mod n: diagonal}."""
    ln = 1 << s
    lenh = ln >> 1
    lenq = ln << 2
    M = 2 * N
    j = np.arange(n) % ln
    top = j < lenh
    jj = np.where(top, j, j - lenh)
    om = np.exp(2j * np.pi * (np.array([pow(5, int(x), M) for x in range(lenh)]) % lenq) / lenq)[jj]
    d0 = np.zeros(n, complex)
    dp = np.zeros(n, complex)
    dm = np.zeros(n, complex)
    if not inverse:
        # top: y = x + w x[+lenh];  bottom: y = x[-lenh] - w x
        d0[top], dp[top] = 1.0, om[top]
        d0[~top], dm[~top] = -om[~top], 1.0
    else:
        # top: x = (y + y[+lenh]) / 2;  bottom: x = conj(w) (y[-lenh] - y) / 2
        d0[top], dp[top] = 0.5, 0.5
        d0[~top], dm[~top] = -0.5 * np.conj(om[~top]), 0.5 * np.conj(om[~top])
    out = {}
    for off, v in ((0, d0), (lenh % n, dp), ((-lenh) % n, dm)):
        out[off] = out.get(off, 0) + v
    return out


def diag_mul(A: dict, B: dict, n: int) -> dict:
    """Diagonals of the product A B (apply B first)."""
    out = {}
    for e, a in A.items():
        for d, b in B.items():
            k = (d + e) % n
            out[k] = out.get(k, 0) + a * rot(b, e)
    return {k: v for k, v in out.items() if np.max(np.abs(v)) > 1e-300}


def diag_apply(D: dict, x: np.ndarray) -> np.ndarray:
    y = np.zeros_like(x, dtype=complex)
    for d, v in D.items():
        y = y + v * rot(x, d)
    return y


def split_stages(logn: int, budget: int) -> list:
    """Stages 1..logn into `budget` consecutive groups, larger groups first."""
    budget = max(1, min(budget, logn))
    base, extra = divmod(logn, budget)
    sizes = [base + (1 if i < extra else 0) for i in range(budget)]
    out, s = [], 1
    for sz in sizes:
        out.append(list(range(s, s + sz)))
        s += sz
    return out


def transform_groups(n: int, N: int, budget: int, inverse: bool) -> list:
    """Group matrices in application order.  Forward (SlotToCoeff without
    the bit reversal): B_1..B_logn; inverse (CoeffToSlot): B_logn^-1..B_1^-1."""
    logn = n.bit_length() - 1
    groups = split_stages(logn, budget)
    mats = []
    for g in groups:
        stages = g if not inverse else list(reversed(g))
        M = None
        for s in stages:
            S = stage_diags(n, N, s, inverse)
            M = S if M is None else diag_mul(S, M, n)
        mats.append(M)
    return mats if not inverse else list(reversed(mats))


# ----------------------------------------------------------------- baby-step / giant-step plan
@dataclass
class GroupPlan:
    stride: int
    g: int                                   # baby steps
    babies: list                             # b with rotation b*stride
    giants: dict = field(default_factory=dict)   # k -> [(b, diagonal for rotation k*g*stride)]
    n: int = 0

    def rotations(self) -> set:
        r = {(b * self.stride) % self.n for b in self.babies if b}
        r |= {(k * self.g * self.stride) % self.n for k in self.giants if k}
        return r


def plan_group(D: dict, n: int, factor: complex = 1.0) -> GroupPlan:
    offs = sorted(D)
    stride = 0
    for o in offs:
        stride = math.gcd(stride, o)
    stride = stride or n
    Mn = n // stride
    ms = sorted(((o // stride + Mn // 2) % Mn) - Mn // 2 for o in offs)   # signed multiples
    span = ms[-1] - ms[0] + 1
    g = 1 << max(0, math.ceil(math.log2(math.sqrt(span))))
    g = min(g, 16)
    P = GroupPlan(stride, g, [], {}, n)
    babies = set()
    for m in ms:
        k, b = divmod(m, g)
        diag = D[(m * stride) % n] * factor
        # y = sum_k rot(sum_b rot(diag_m, -k g s) * rot(x, b s), k g s)
        P.giants.setdefault(k, []).append((b, rot(diag, -k * g * stride)))
        babies.add(b)
    P.babies = sorted(babies)
    return P


def bsgs_apply(P: GroupPlan, x: np.ndarray) -> np.ndarray:
    """The plan evaluated on plain vectors (for checking)."""
    xb = {b: rot(x, b * P.stride) for b in P.babies}
    y = np.zeros_like(x, dtype=complex)
    for k, terms in P.giants.items():
        inner = sum(v * xb[b] for b, v in terms)
        y = y + rot(inner, k * P.g * P.stride)
    return y


# ----------------------------------------------------------------- the bootstrap plan
@dataclass
class BootConfig:
    budget: tuple = (3, 3)      # (CoeffToSlot, SlotToCoeff) levels
    degree: int = 55            # Chebyshev degree of the cosine
    double_angle: int = 3       # r
    K: float = 32.0             # |I| + |m/q_0| bound of the argument
    correction: int = 10        # prescale 2^-c
    encaps_h: int | None = None  # sparse-secret encapsulation: Hamming weight of s' (None: off)
    flexible: bool = False      # FLEXIBLEAUTO: the result carries its level's FLEXIBLEAUTO
                                # scaling factor (folded into the first SlotToCoeff group)
    slots: int | None = None    # sparse packing: n_s slots (a power of two < N/2); None: N/2

    @staticmethod
    def encapsulated(h: int = 32, flexible: bool = False, slots: int | None = None) -> "BootConfig":
        """Sparse-secret encapsulation with a weight-h secret: |I| stays
        below ~8 with overwhelming probability for h = 32 (I is a sum of h + 1
        terms uniform in [-1/2, 1/2]), so K' = 16."""
        return BootConfig(K=16.0, encaps_h=h, flexible=flexible, slots=slots)


def tile_diags(D: dict, n2: int) -> dict:
    """Diagonals of a matrix on n-periodic vectors, as a matrix on
    n2-periodic ones (n | n2): the same offsets, each diagonal tiled."""
    return {d: np.tile(v, n2 // len(v)) for d, v in D.items()}


def sparse_fold(cts: list, stc: list, ns: int):
    """The sparse-packing transforms on 2 ns-periodic slot vectors.

    A message in ns slots (replicated over N/2) is a polynomial in
    X^(N / 2 ns): its 2 ns coefficients t_k pack into ns complex slots
    w_k = t_k + i t_(k+ns), and the decoding map of ns slots (the group
    matrices of ns slots) takes w to the message.  On 2 ns-periodic vectors:
    CoeffToSlot ends with the mask (1 on [0, ns), -i on [ns, 2 ns)), so that
    ct + conj(ct) holds [2 Re w | 2 Im w] and one EvalMod treats both halves;
    SlotToCoeff starts with the recombination w_p = y_p + i y_(p+ns) (p < ns,
    and periodically), diagonals {0: a, ns: b}.  The mask is folded into the
    last CoeffToSlot group and the recombination into the first SlotToCoeff
    group (no extra level)."""
    n2 = 2 * ns
    cts = [tile_diags(D, n2) for D in cts]
    stc = [tile_diags(D, n2) for D in stc]
    low = np.arange(n2) < ns
    mask = np.where(low, 1.0 + 0j, -1j)
    cts[-1] = {d: mask * v for d, v in cts[-1].items()}
    comb = {0: np.where(low, 1.0 + 0j, 1j), ns: np.where(low, 1j, 1.0 + 0j)}
    stc[0] = diag_mul(stc[0], comb, n2)
    return cts, stc


class BootstrapPlan:
    """Tables and device plaintexts for bootstrapping ciphertexts of one
    parameter set (full slots, or cfg.slots < N/2 with sparse packing);
    `program()` builds the request program.

    ctx: a capi.Context (plaintexts are uploaded and prepared on it), or None
    to plan without a device (keep_host=True then keeps the plaintext
    coefficients for host-side checking)."""

    def __init__(self, prm, ctx=None, cfg: BootConfig = BootConfig(), keep_host: bool = False):
        self.prm = prm
        self.cfg = cfg
        self.N = prm.N
        self.n = prm.N // 2
        self.q = [int(x) for x in prm.q]
        self.L = len(self.q)
        self.delta = float(2 ** prm.scale_bits)
        self.ctx = ctx
        self.keep_host = keep_host
        N = self.N
        self.slots = cfg.slots or N // 2
        ns = self.slots
        if ns < 2 or ns > N // 2 or ns & (ns - 1):
            raise ValueError("slots must be a power of two in [2, N/2]")
        self.sparse = ns < N // 2
        cts = transform_groups(ns, N, cfg.budget[0], inverse=True)
        stc = transform_groups(ns, N, cfg.budget[1], inverse=False)
        if self.sparse:
            cts, stc = sparse_fold(cts, stc, ns)
            self.n = 2 * ns            # slot period the group plans work on
        n = self.n
        q0 = float(self.q[0])
        # CoeffToSlot output y = t / (q_0 K'), with the 1/2 of the Re/Im split
        # folded in (and 2 ns / N, the factor of the partial trace, when sparse)
        self.f_cts = self.delta / (2.0 * q0 * cfg.K)
        if self.sparse:
            self.f_cts *= 2.0 * ns / N
        # SlotToCoeff input sin(2 pi t/q_0) ~ 2 pi m/q_0  ->  m / s1 with s1 ~ delta
        self.f_stc = q0 / (2.0 * math.pi * self.delta)
        self.cts_plans = [plan_group(D, n, self.f_cts if i == 0 else 1.0) for i, D in enumerate(cts)]
        self.stc_plans = [plan_group(D, n, self.f_stc if i == 0 else 1.0) for i, D in enumerate(stc)]
        self.cheb = chebyshev_coefficients(
            lambda y: math.cos(2 * math.pi * (cfg.K * y - 0.25) / 2 ** cfg.double_angle), cfg.degree, -1.0, 1.0)
        # levels
        self.l_raise = self.L
        self.evalmod_depth = chebyshev_eval_depth(self.cheb) + cfg.double_angle
        # (fewer groups than the budget when the slots have fewer stages)
        self.depth = len(self.cts_plans) + len(self.stc_plans) + self.evalmod_depth
        if self.l_raise - self.depth < 1:
            raise ValueError("not enough levels for bootstrapping: need %d, have %d" % (self.depth + 1, self.L))
        self._pt_cache = {}
        self._pts = None
        self.out_scale = None
        if cfg.flexible:
            # the nominal output scale does not depend on the input (ModRaise
            # declares delta): fold target / nominal into SlotToCoeff's factor
            from .params import flexible_scaling_factors
            s_nom = self._nominal_output_scale()
            self.out_scale = flexible_scaling_factors(self.q)[self.output_level() - 1]
            self.f_stc *= self.out_scale / s_nom
            self.stc_plans = [plan_group(D, n, self.f_stc if i == 0 else 1.0) for i, D in enumerate(stc)]

    # -- rotation indices every program uses (slot rotations, mod N/2); conjugation separately
    def trace_rotations(self) -> list:
        """Sparse packing: the partial trace after ModRaise, x += Rot(x, j ns)
        for j = 1, 2, 4, ... below N / (2 ns) (the rotations that fix the
        subring of X^(N / 2 ns))."""
        out, j = [], 1
        while self.sparse and j * self.slots < self.N // 2:
            out.append(j * self.slots)
            j <<= 1
        return out

    def rotations(self) -> list:
        r = set(self.trace_rotations())
        for P in self.cts_plans + self.stc_plans:
            r |= P.rotations()
        r.discard(0)
        return sorted(r)

    def output_level(self) -> int:
        return self.l_raise - self.depth

    # -- switching keys of the sparse-secret encapsulation (key ids of the
    #    request loop's key table): s -> s' and s' -> s
    ENCAPS_DOWN = "encaps_down"
    ENCAPS_UP = "encaps_up"

    def switch_keys(self) -> list:
        return [self.ENCAPS_DOWN, self.ENCAPS_UP] if self.cfg.encaps_h else []

    # -- plaintexts
    def _encode(self, values, level: int, scale: float):
        coeffs = codec.encode(values, scale, self.N)
        rns = codec.to_rns(coeffs, self.q[:level])
        return self._upload(rns, level, scale)

    def _upload(self, rns, level, scale):
        host = rns if (self.keep_host or self.ctx is None) else None
        buf = None
        if self.ctx is not None:
            from . import capi
            buf = capi.DeviceBuffer.from_array(rns)
            self.ctx.plaintext_prepare(buf, level)
        return Plaintext(buf if buf is not None else 0, level, scale, host)

    def _monomial(self, power: int, level: int):
        """X^power mod X^N + 1 as an exact plaintext (scale 1)."""
        c = np.zeros(self.N, np.int64)
        sign = 1
        p = power % (2 * self.N)
        if p >= self.N:
            p -= self.N
            sign = -1
        c[p] = sign
        return self._upload(codec.to_rns(c, self.q[:level]), level, 1.0)

    def build_plaintexts(self, scale_only: bool = False) -> None:
        """Encode every diagonal of every group at the level it is used
        (scale_only: placeholders with level and scale, for structure)."""
        if self._pts is not None:
            return
        pts = {"cts": [], "stc": []}
        l = self.l_raise
        for P in self.cts_plans:
            pts["cts"].append(self._group_pts(P, l, scale_only))
            l -= 1
        l_split = l
        mono = (lambda p, lv: Plaintext(0, lv, 1.0)) if scale_only else self._monomial
        if not self.sparse:  # the Re/Im split and recombination of full slots
            self._minus_i = mono(3 * self.N // 2, l_split)
        l -= self.evalmod_depth
        if not self.sparse:
            self._plus_i = mono(self.N // 2, l)
        for P in self.stc_plans:
            pts["stc"].append(self._group_pts(P, l, scale_only))
            l -= 1
        self._pts = pts

    def _group_pts(self, P: GroupPlan, level: int, scale_only: bool = False) -> dict:
        scale = float(self.q[level - 1])      # Rescale after the products divides it out
        if scale_only:
            return {k: [(b, Plaintext(0, level, scale)) for b, _ in terms] for k, terms in P.giants.items()}
        return {k: [(b, self._encode(v, level, scale)) for b, v in terms] for k, terms in P.giants.items()}

    def _nominal_output_scale(self) -> float:
        """The tracked scale of program()'s result (structure only)."""
        self.build_plaintexts(scale_only=True)
        try:
            return self.program(self.L).output.scale
        finally:
            self._pts = None

    def export_precomputation(self) -> dict:
        """The plan's tables for the C++ API's EvalBootstrapSetup(const
        BootstrapPrecomputation&): per group (CoeffToSlot then SlotToCoeff,
        application order) the stride, baby-step count, babies and the terms
        (giant k, baby b, level, scale, coefficient-form residues), the
        EvalMod series and the output scale.  Needs keep_host=True."""
        self.build_plaintexts()

        def group(P, pts):
            terms = []
            for k in sorted(pts):
                for b, pt in pts[k]:
                    assert pt.host is not None, "export_precomputation needs keep_host=True"
                    terms.append((k, b, pt.level, pt.scale, np.ascontiguousarray(pt.host, np.uint64)))
            return dict(stride=P.stride, g=P.g, babies=list(P.babies), terms=terms)
        return dict(cts=[group(P, x) for P, x in zip(self.cts_plans, self._pts["cts"])],
                    stc=[group(P, x) for P, x in zip(self.stc_plans, self._pts["stc"])],
                    cheb=[float(c) for c in self.cheb], out_scale=self.out_scale or 0.0)

    # -- program pieces
    def _linear(self, Pg: Program, x: Reg, plan: GroupPlan, pts: dict) -> Reg:
        # rotation amounts modulo the plan's slot period (the data repeats with
        # it), the same indices rotations() lists
        babies = {b: (x if b == 0 else Pg.rotate(x, (b * plan.stride) % plan.n)) for b in plan.babies}
        parts = []
        for k in sorted(pts):
            inner = Pg.ptdot([(babies[b], pt) for b, pt in pts[k]])
            parts.append(inner if k == 0 else Pg.rotate(inner, (k * plan.g * plan.stride) % plan.n))
        # one fused sum (a linear combination with unit constants) instead of
        # a tree of pairwise additions: one launch and one step
        if len(parts) == 1:
            return Pg.rescale(parts[0])
        return Pg.rescale(Pg.lincomb([(p, 1) for p in parts], add=None, scale=parts[0].scale))

    def _evalmod(self, Pg: Program, y: Reg) -> Reg:
        c = chebyshev_eval(Pg, y, self.cheb)
        for _ in range(self.cfg.double_angle):
            sq = Pg.hmul(c, c)
            c = Pg.lin(sq, mul=2, add=-int(round(sq.scale)))
        return c

    def program(self, level_in: int, scale_in: float | None = None) -> Program:
        """Bootstrap of one ciphertext (self.slots slots) at `level_in` limbs (>= 2)."""
        self.build_plaintexts()
        cfg = self.cfg
        Pg = Program(self.q, self.prm.logN, self.prm.scale_bits, "bootstrap")
        x = Pg.input(level_in, scale_in)
        # prescale by 2^-c; with `flexible` also by q_{l-1} / s_in, so that
        # after the Rescale the message sits at delta whatever the input
        # level's factor (the reference's adjustment factor,
        # src/CKKS/Bootstrap.cu:441-470)
        pre = 2.0 ** -cfg.correction
        if cfg.flexible:
            pre = pre * float(self.q[level_in - 1]) / x.scale
        x = Pg.mul_real(x, pre)
        x = Pg.drop(x, 1)
        if cfg.encaps_h:
            x = Pg.switch_key(x, self.ENCAPS_DOWN)
        r = Pg.mod_raise(x, self.l_raise, self.delta)
        if cfg.encaps_h:
            r = Pg.switch_key(r, self.ENCAPS_UP)
        for amount in self.trace_rotations():
            r = Pg.add(r, Pg.rotate(r, amount))
        for plan, pts in zip(self.cts_plans, self._pts["cts"]):
            r = self._linear(Pg, r, plan, pts)
        cj = Pg.conjugate(r)
        if self.sparse:
            # [2 Re w | 2 Im w] in one ciphertext: one EvalMod
            w = self._evalmod(Pg, Pg.add(r, cj))
        else:
            re = Pg.add(r, cj)
            im = Pg.ptdot([(Pg.sub(r, cj), self._minus_i)])
            re = self._evalmod(Pg, re)
            im = self._evalmod(Pg, im)
            w = Pg.add(re, Pg.ptdot([(im, self._plus_i)]))
        for plan, pts in zip(self.stc_plans, self._pts["stc"]):
            w = self._linear(Pg, w, plan, pts)
        # value = the input message: the nominal scale of the result is the
        # tracked one (the constants above are folded relative to delta), or
        # with `flexible` the level's FLEXIBLEAUTO factor folded into SlotToCoeff
        out = Pg.lin(w, mul=1 << cfg.correction, scale=self.out_scale)
        Pg.set_output(out)
        return Pg
