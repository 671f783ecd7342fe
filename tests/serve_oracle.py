"""Runs a request-loop Program op by op on the CPU oracle (test helper)."""
import numpy as np

from fideslib_amd import serve as S


def run_oracle(octx, P: S.Program, inputs, rlk, rot_keys):
    """The program, op by op, on the CPU oracle."""
    N = octx.N
    val = {r.id: np.asarray(x, np.uint64) for r, x in zip(P.inputs, inputs)}
    for o in P.ops:
        x = [val[s.id] for s in o.srcs]
        l = o.level
        if o.kind == S.HMUL:
            y = octx.rescale_ct(octx.hmult_relin(x[0], x[1], l, rlk), l)
        elif o.kind == S.HMUL_NR:
            y = octx.hmult_relin(x[0], x[1], l, rlk)
        elif o.kind == S.ROT:
            y = octx.rotate(x[0], l, rot_keys[o.arg[0]], o.arg[0])
        elif o.kind == S.RESCALE:
            y = octx.rescale_ct(x[0], l)
        elif o.kind == S.DROP:
            y = np.ascontiguousarray(x[0][:, :o.arg[0]])
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
    return val[P.output.id]
