"""Test-side composition of the extended-basis (Q_l u P) operations from the
oracle's pinned pieces: the key switch without its ModDown, rotations left
in Q_l u P, the P-scaled embedding of a ciphertext, and the hoisted-ModDown
rotate-and-sum.  Reference: Ciphertext::rotate_hoisted with ext = true
(src/CKKS/Ciphertext.cpp:940-1114) and Accumulate
(src/CKKS/AccumulateBroadcast.cu:79-106); OpenFHE's EvalFastRotationExt
adds P c0 before the automorphism.

Pinned on the CPU by tests/test_oracle.py: ModDown of the extended key
switch equals the oracle's own key switch, and the extended identity is
P times the ciphertext."""
import numpy as np


def qp_pids(octx, l):
    return list(range(l)) + list(range(octx.L, octx.L + octx.K))


def keyswitch_ext(octx, d, l, key):
    """(sum_j e_j b_j, sum_j e_j a_j) over Q_l u P, e_j the ModUp of digit j
    (src/CKKS/ElemenwiseBatchKernels.cu:202-273 without the ModDown)."""
    pids = qp_pids(octx, l)
    sel = np.asarray(pids)
    o0 = o1 = None
    for j in range(octx.num_digits(l)):
        e = octx.modup_digit(np.ascontiguousarray(d), l, j)
        t0 = octx.mul_limbs(e, np.ascontiguousarray(key[1, j][sel]), pids)
        t1 = octx.mul_limbs(e, np.ascontiguousarray(key[0, j][sel]), pids)
        o0 = t0 if o0 is None else octx.add_limbs(o0, t0, pids)
        o1 = t1 if o1 is None else octx.add_limbs(o1, t1, pids)
    return o0, o1


def p_mod(octx, l):
    P = 1
    for p in octx.p:
        P *= p
    return [P % octx.q[i] for i in range(l)]


def scale_by_p(octx, x, l):
    """[l][N] -> [l+K][N]: x P on the Q limbs, zero on the special limbs."""
    pm = p_mod(octx, l)
    c = np.stack([np.full(octx.N, pm[i], np.uint64) for i in range(l)])
    q = octx.mul_limbs(np.ascontiguousarray(x), c, list(range(l)))
    return np.concatenate([q, np.zeros((octx.K, octx.N), np.uint64)])


def raise_ext(octx, ct, l):
    return np.stack([scale_by_p(octx, ct[0], l), scale_by_p(octx, ct[1], l)])


def rotate_ext(octx, ct, l, key, k, add_first=True):
    """[2][l+K][N]: sigma_k(P c0 + KS0), sigma_k(KS1); k = 1 gives (P c0 or 0, P c1)."""
    if k == 1:
        r = raise_ext(octx, ct, l)
        if not add_first:
            r[0] = 0
        return r
    k0, k1 = keyswitch_ext(octx, ct[1], l, key)
    pids = qp_pids(octx, l)
    if add_first:
        k0 = octx.add_limbs(k0, scale_by_p(octx, ct[0], l), pids)
    return np.stack([octx.automorph(k0, k), octx.automorph(k1, k)])


def add_ext(octx, a, b, l, sub=False):
    pids = qp_pids(octx, l)
    f = octx.sub_limbs if sub else octx.add_limbs
    return np.stack([f(a[0], b[0], pids), f(a[1], b[1], pids)])


def moddown_ext(octx, x, l):
    return np.stack([octx.moddown(np.ascontiguousarray(x[0]), l), octx.moddown(np.ascontiguousarray(x[1]), l)])


def accumulate(octx, ct, l, keys_by_k, slots, stride, N, bstep=4, start=1):
    """sum_{j < slots} rot(ct, j stride), radix-bstep rounds of hoisted
    rotations whose sum takes one ModDown per round (the reference GPU
    Accumulate).  keys_by_k maps 5^r mod 2N to the rotation key."""
    logb = bstep.bit_length() - 1
    x = ct
    s = start
    while s < slots:
        acc = raise_ext(octx, x, l)
        idx = stride * s
        while idx < stride * slots and idx < bstep * stride * s:
            k = pow(5, idx % (N // 2), 2 * N)
            acc = add_ext(octx, acc, rotate_ext(octx, x, l, keys_by_k[k], k), l)
            idx += stride * s
        x = moddown_ext(octx, acc, l)
        s <<= logb
    return x
