"""Test-side restatement of the reference's level/scale management over the
oracle's exact operations (checker for the host C++ API under FIXEDAUTO and
FLEXIBLEAUTO(EXT)).  Each function follows the reference's control flow:

  scalar_residues      ContextData::ElemForEvalMult, src/CKKS/Context.cu:366-463
  add_scalar_residues  ContextData::ElemForEvalAddOrSub, Context.cu:468-560
  mult_scalar_noprecheck / mult_scalar   Ciphertext.cpp:747-777
  rescale              Ciphertext::rescale, Ciphertext.cpp:426-447
  adjust_scale_and_level   Ciphertext.cpp:1376-1455
  adjust_for_add_or_sub    Ciphertext.cpp:1457-1490
  adjust_for_mult          Ciphertext.cpp:1492-1508
  add / sub            Ciphertext.cpp:164-240 (recursion on the copied operand)
  mult                 Ciphertext.cpp:458-477 (the same recursion)
  add_pt               Ciphertext::addPt, Ciphertext.cpp:242-272, with
                       Plaintext::adjustPlaintextToCiphertext, Plaintext.cu:195-226

Floating-point steps are the reference's, in its order of operations, so the
integer constants (and with them every limb) must come out identical.
OpenFHE's constants LargeScalingFactorConstants::MAX_BITS_IN_WORD = 61 and
MAX_LOG_STEP = 60 are recalled (OpenFHE is not available here); the scaling
factor tables are OpenFHE's defaults as the reference copies them
(RawCiphertext.cu:260-272): FLEXIBLEAUTO sf[L-1] = q[L-1],
sf[l-1] = sf[l]^2 / q[l]; fixed techniques 2^round(log2 q[1])."""
import math
from dataclasses import dataclass

import numpy as np

FIXEDMANUAL, FIXEDAUTO, FLEXIBLEAUTO, FLEXIBLEAUTOEXT = 0, 1, 2, 3


@dataclass
class Ct:
    data: np.ndarray  # [2][l][N]
    deg: int          # NoiseLevel (scale degree)
    sf: float         # NoiseFactor

    @property
    def limbs(self):
        return self.data.shape[1]

    @property
    def level(self):
        return self.limbs - 1


def default_factors(q, technique):
    L = len(q)
    if technique in (FLEXIBLEAUTO, FLEXIBLEAUTOEXT):
        sf = [0.0] * L
        sf[L - 1] = float(q[L - 1])
        for l in range(L - 1, 0, -1):
            sf[l - 1] = sf[l] * sf[l] / float(q[l])
    else:
        p = 2.0 ** round(math.log2(float(q[1 if L > 1 else 0])))
        sf = [p] * L
    return sf, [x * x for x in sf]


class Flex:
    def __init__(self, octx, technique, rlk=None):
        self.o = octx
        self.q = octx.q
        self.N = octx.N
        self.L = len(self.q)
        self.tech = technique
        self.rlk = rlk
        self.sf, self.sfbig = default_factors(self.q, technique)
        self.mrf = [float(x) for x in self.q]  # ModReduceFactor default: the dropped prime

    def flexible(self):
        return self.tech in (FLEXIBLEAUTO, FLEXIBLEAUTOEXT)

    # ---- constants
    def scalar_residues(self, l, operand):
        sc = self.sf[l - 1]
        log_approx = 0
        res = abs(operand * sc)
        if res > 0:
            log_sf = int(math.ceil(math.log2(res)))
            log_approx = log_sf - min(log_sf, 125)
        approx = 2.0 ** log_approx
        large = int(operand / approx * sc + 0.5)  # C cast: truncation toward zero
        out = []
        for i in range(l):
            v = large % self.q[i]
            if log_approx > 0:
                v = v * pow(2, log_approx, self.q[i]) % self.q[i]
            out.append(v)
        return out

    def add_scalar_residues(self, l, c, deg):
        level = l - 1
        ext_top = self.tech == FLEXIBLEAUTOEXT and level == self.L - 1
        sc = self.sfbig[level] if ext_top else self.sf[level]
        log_approx = 0
        res = abs(abs(c) * sc)
        if res > 0:
            log_sf = int(math.ceil(math.log2(res)))
            log_approx = log_sf - min(log_sf, 61)
        base = int(abs(c) * sc / 2.0 ** log_approx + 0.5)
        isf = int(sc + 0.5)
        out = []
        for i in range(l):
            v = base % self.q[i]
            if log_approx > 0:
                v = v * pow(2, log_approx, self.q[i]) % self.q[i]
            if not ext_top:
                for _ in range(1, deg):
                    v = v * (isf % self.q[i]) % self.q[i]
            if c < 0 and v:
                v = self.q[i] - v
            out.append(v)
        return out

    # ---- limb operations
    def _mul_res(self, a: Ct, res):
        l, N = a.limbs, self.N
        pid = list(range(l)) * 2
        c = np.stack([np.full(N, res[j], np.uint64) for j in range(l)] * 2)
        return self.o.mul_limbs(a.data.reshape(2 * l, N), c, pid).reshape(2, l, N)

    def drop(self, a: Ct, limbs) -> Ct:
        return Ct(np.ascontiguousarray(a.data[:, :limbs]), a.deg, a.sf)

    def rescale(self, a: Ct) -> Ct:
        return Ct(self.o.rescale_ct(a.data, a.limbs), a.deg - 1, a.sf / self.mrf[a.limbs - 1])

    def mult_scalar_noprecheck(self, a: Ct, c) -> Ct:
        d = self._mul_res(a, self.scalar_residues(a.limbs, c))
        return Ct(d, a.deg + 1, a.sf * self.sf[a.level])

    def mult_scalar(self, a: Ct, c) -> Ct:
        if self.tech != FIXEDMANUAL and a.deg == 2:
            a = self.rescale(a)
        assert a.deg == 1
        return self.mult_scalar_noprecheck(a, c)

    def add_scalar(self, a: Ct, c) -> Ct:
        res = self.add_scalar_residues(a.limbs, abs(c), a.deg)
        if c < 0:
            res = [self.q[i] - r for i, r in enumerate(res)]
        l, N = a.limbs, self.N
        pid = list(range(l)) * 2
        cc = np.stack([np.full(N, res[j] % self.q[j], np.uint64) for j in range(l)] + [np.zeros(N, np.uint64)] * l)
        return Ct(self.o.add_limbs(a.data.reshape(2 * l, N), cc, pid).reshape(2, l, N), a.deg, a.sf)

    # ---- alignment
    def adjust_scale_and_level(self, a: Ct, d2, l2, sf2):
        """(ok, a'): a brought to level l2 (index), scale degree d2, factor sf2."""
        l1, d1 = a.level, a.deg
        if l1 > l2:
            if d1 == 2:
                if d2 == 2:
                    scf1, scf, q1 = a.sf, self.sf[l1], self.mrf[l2 + 1]
                    a = self.mult_scalar_noprecheck(a, sf2 * q1 / scf1 / scf)
                    if a.level > l2 + 1:
                        a = self.drop(a, l2 + 2)
                    a.sf = self.sfbig[l2] * self.mrf[l2 + 1]
                    a = self.rescale(a)
                    a.sf = sf2
                elif l1 - 1 == l2:
                    a = self.rescale(a)
                else:
                    scf1, scf2, scf, q1 = a.sf, self.sfbig[l2 + 1], self.sf[l1], self.mrf[l1]
                    a = self.mult_scalar_noprecheck(a, scf2 / scf1 * q1 / scf)
                    a.sf = self.sfbig[a.level - 1] * self.mrf[l1]
                    a = self.rescale(a)
                    if a.level - 1 > l2:
                        a = self.drop(a, l2 + 2)
                    a.sf = self.sfbig[a.level]
                    a = self.rescale(a)
                    a.sf = sf2
            else:
                if d2 == 2:
                    scf1, scf = a.sf, self.sf[l1]
                    a = self.mult_scalar_noprecheck(a, sf2 / scf1 / scf)
                    a = self.drop(a, l2 + 1)
                    a.sf = sf2
                else:
                    scf1, scf2, scf = a.sf, self.sfbig[l2 + 1], self.sf[l1]
                    a = self.mult_scalar_noprecheck(a, scf2 / scf1 / scf)
                    if l1 - 1 > l2:
                        a = self.drop(a, l2 + 2)
                    a = self.rescale(a)
                    a.sf = sf2
            return True, a
        if l1 < l2:
            return False, a
        if d1 < d2:
            a = self.mult_scalar(a, 1.0)
        elif d2 < d1:
            return False, a
        return True, a

    def adjust_for_add_or_sub(self, a: Ct, b: Ct):
        if self.flexible():
            return self.adjust_scale_and_level(a, b.deg, b.level, b.sf)
        if a.level - a.deg > b.level - b.deg:
            if b.deg == 1 and a.deg == 2:
                a = self.rescale(a)
            elif b.deg == 2 and a.deg == 1:
                a = self.mult_scalar(a, 1.0)
            return True, a
        if b.deg == 1 and a.deg == 2:
            return True, self.rescale(a)
        if a.deg == 1 and b.deg == 2:
            return False, a
        return True, a

    def adjust_for_mult(self, a: Ct, b: Ct):
        ok, a = self.adjust_for_add_or_sub(a, b)
        if a.deg == 2:
            a = self.rescale(a)
        return (ok and b.deg != 2), a

    # ---- operations
    def add(self, a: Ct, b: Ct, sub=False) -> Ct:
        if self.tech != FIXEDMANUAL:
            ok, a = self.adjust_for_add_or_sub(a, b)
            if not ok:
                ok2, b2 = self.adjust_for_add_or_sub(b, a)
                assert ok2
                return self.add(a, b2, sub)
        if not self.flexible() and a.level > b.level:
            a = self.drop(a, b.limbs)
        assert a.deg == b.deg
        l, N = a.limbs, self.N
        pid = list(range(l)) * 2
        f = self.o.sub_limbs if sub else self.o.add_limbs
        d = f(a.data.reshape(2 * l, N), np.ascontiguousarray(b.data[:, :l]).reshape(2 * l, N), pid)
        return Ct(d.reshape(2, l, N), a.deg, a.sf)

    def mult(self, a: Ct, b: Ct) -> Ct:
        if self.tech != FIXEDMANUAL:
            ok, a = self.adjust_for_mult(a, b)
            if not ok:
                ok2, b2 = self.adjust_for_mult(b, a)
                assert ok2
                return self.mult(a, b2)
        if not self.flexible():
            # the fixed rules leave the levels apart; the reference multiplies at
            # the first operand's level (reading the other's first limbs), OpenFHE
            # drops the higher operand: the latter where the former is undefined
            l = min(a.limbs, b.limbs)
            a, b = self.drop(a, l), self.drop(b, l)
        assert a.deg == 1 and b.deg == 1 and a.limbs == b.limbs
        d = self.o.hmult_relin(a.data, b.data, a.limbs, self.rlk)
        return Ct(d, 2, a.sf * b.sf)

    def add_pt(self, a: Ct, p: Ct, sub=False) -> Ct:
        """p is the plaintext as (pt, 0) in evaluation form."""
        if self.tech != FIXEDMANUAL:
            if p.deg == 1 and a.deg == 2 and p.level == a.level - 1:
                a = self.rescale(a)
            if p.level != a.level:
                if self.tech == FIXEDAUTO:
                    if p.level - p.deg > a.level - a.deg:
                        if a.deg == 1 and p.deg == 2:
                            p = self.rescale(self.drop(p, a.limbs + 1))
                        else:
                            p = self.drop(p, a.limbs)
                    elif a.deg == 1 and p.deg == 2:
                        p = self.rescale(p)
                else:
                    assert p.level >= a.level
                    ok, p = self.adjust_scale_and_level(p, a.deg, a.level, a.sf)
                    assert ok
        assert p.deg == a.deg and p.limbs >= a.limbs
        l, N = a.limbs, self.N
        pid = list(range(l)) * 2
        f = self.o.sub_limbs if sub else self.o.add_limbs
        d = f(a.data.reshape(2 * l, N), np.ascontiguousarray(p.data[:, :l]).reshape(2 * l, N), pid)
        return Ct(d.reshape(2, l, N), a.deg, a.sf)

    def set_level(self, a: Ct, level) -> Ct:
        """CryptoContext::SetLevel: keep L - level limbs (dropToLevel)."""
        t = self.L - level
        if t >= a.limbs:
            return a
        if self.flexible():
            ok, a = self.adjust_scale_and_level(a, a.deg, t - 1, self.sf[t - 1] if a.deg == 1 else self.sfbig[t - 1])
            assert ok
            return a
        return self.drop(a, t)
