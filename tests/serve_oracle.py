"""Runs a request-loop Program op by op on the CPU oracle (test helper)."""
import numpy as np

from fideslib_amd import serve as S


def run_oracle(octx, P: S.Program, inputs, rlk, rot_keys, keep=None):
    """The program, op by op, on the CPU oracle."""
    N = octx.N
    val = {r.id: np.asarray(x, np.uint64) for r, x in zip(P.inputs, inputs)}
    pt_eval = {}

    def plain(pt, l):
        if id(pt) not in pt_eval:
            pt_eval[id(pt)] = octx.ntt_limbs(np.ascontiguousarray(pt.host, np.uint64), list(range(pt.level)))
        return pt_eval[id(pt)][:l]
    for o in P.ops:
        x = [val[s.id] for s in o.srcs]
        l = o.level
        if o.kind == S.HMUL:
            y = octx.rescale_ct(octx.hmult_relin(x[0], x[1], l, rlk), l)
        elif o.kind == S.HMUL_NR:
            y = octx.hmult_relin(x[0], x[1], l, rlk)
        elif o.kind == S.ROT:
            y = octx.rotate(x[0], l, rot_keys[o.arg[1]], o.arg[0])
        elif o.kind == S.RESCALE:
            y = octx.rescale_ct(x[0], l)
        elif o.kind == S.DROP:
            y = np.ascontiguousarray(x[0][:, :o.arg[0]])
        elif o.kind == S.LINCOMB:
            ms, add = o.arg
            pid = list(range(l)) * 2
            y = None
            for xj, m in zip(x, ms):
                c = np.stack([np.full(N, m % octx.q[j], np.uint64) for j in range(l)] * 2)
                t = octx.mul_limbs(np.ascontiguousarray(xj[:, :l]).reshape(2 * l, N), c, pid)
                y = t if y is None else octx.add_limbs(y, t, pid)
            if add is not None:
                c = np.stack([np.full(N, add % octx.q[j], np.uint64) for j in range(l)] + [np.zeros(N, np.uint64)] * l)
                y = octx.add_limbs(y, c, pid)
            y = y.reshape(2, l, N)
        elif o.kind == S.MODRAISE:
            y = octx.mod_raise(x[0], o.arg[0])
        elif o.kind == S.PTDOT:
            pid = list(range(l)) * 2
            y = None
            for ct, pt in zip(x, o.arg[0]):
                m = plain(pt, l)
                t = octx.mul_limbs(ct.reshape(2 * l, N), np.concatenate([m, m]), pid)
                y = t if y is None else octx.add_limbs(y, t, pid)
            y = y.reshape(2, l, N)
        elif o.kind == S.LIN:
            mul, add, sub = o.arg
            pid = list(range(l)) * 2
            y = x[0].reshape(2 * l, N)
            if mul is not None:
                y = octx.mul_limbs(y, np.stack([np.full(N, mul % octx.q[j], np.uint64) for j in pid]), pid)
            if len(x) > 1:
                y = (octx.sub_limbs if sub else octx.add_limbs)(y, x[1].reshape(2 * l, N), pid)
            if add is not None:
                c = np.stack([np.full(N, add % octx.q[j], np.uint64) for j in range(l)] + [np.zeros(N, np.uint64)] * l)
                y = octx.add_limbs(y, c, pid)
            y = y.reshape(2, l, N)
        val[o.dst.id] = y
    if keep is not None:
        keep.update(val)
    return val[P.output.id]
