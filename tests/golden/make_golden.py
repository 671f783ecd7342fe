"""Generates golden vectors with an independent pure-Python big-integer
implementation written straight from the definitions (no oracle code, no
fast algorithms):

  NTT(a)[k]   = sum_i a_i psi^(i (2 brev(k) + 1)) mod q         (O(N^2))
  INTT        = the inverse of that matrix
  basis switch y_t = sum_i [x_i (Qh/q_i)^-1]_{q_i} (Qh/q_i) mod p_t
  ModDown     = (x_Q - Conv_{P->Q}(x_P)) P^-1 in coefficient form
  Rescale     = (x - t) / q_top with t the centred residue of x mod q_top
                (OpenFHE SwitchModulus convention: t > q_top >> 1 -> t - q_top)
  Automorphism: a(X) -> a(X^k) by coefficient substitution
  psi         = g^((q-1)/2N) with g the smallest primitive root (sympy)

Primes come from reference_primes.json (FIDESlib's own test tables).
Writes golden_small.json next to this file.
"""
import json
import os
import random

import sympy

HERE = os.path.dirname(os.path.abspath(__file__))


def brev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


def root(q, N):
    g = sympy.primitive_root(q)
    return pow(g, (q - 1) // (2 * N), q)


def ntt_def(a, q, psi, logN):
    N = 1 << logN
    return [sum(a[i] * pow(psi, i * (2 * brev(k, logN) + 1), q) for i in range(N)) % q for k in range(N)]


def intt_def(A, q, psi, logN):
    N = 1 << logN
    ninv = pow(N, -1, q)
    ipsi = pow(psi, -1, q)
    return [ninv * sum(A[k] * pow(ipsi, i * (2 * brev(k, logN) + 1), q) for k in range(N)) % q for i in range(N)]


def crt(res, primes):
    Q = 1
    for q in primes:
        Q *= q
    x = 0
    for r, q in zip(res, primes):
        qh = Q // q
        x += r * qh * pow(qh, -1, q)
    return x % Q, Q


def switch_basis(xs, src, dst):
    """xs: list over source primes of coefficient lists."""
    N = len(xs[0])
    Qh = 1
    for q in src:
        Qh *= q
    out = []
    for p in dst:
        row = []
        for k in range(N):
            s = 0
            for i, q in enumerate(src):
                qh = Qh // q
                s += (xs[i][k] * pow(qh % q, -1, q) % q) * (qh % p)
            row.append(s % p)
        out.append(row)
    return out


def main():
    prim = json.load(open(os.path.join(HERE, "reference_primes.json")))
    p64, sp64 = prim["p64"], prim["sp64"]
    rnd = random.Random(20240611)
    g = {"note": "pure-Python big-integer golden vectors; see make_golden.py", "cases": {}}

    # 1. forward NTT by definition, N = 16 and 32, three primes
    ntt_cases = []
    for logN in (4, 5):
        N = 1 << logN
        for q in (p64[0], p64[1], sp64[0]):
            psi = root(q, N)
            a = [rnd.randrange(q) for _ in range(N)]
            ntt_cases.append({"logN": logN, "q": q, "psi": psi, "a": a, "ntt": ntt_def(a, q, psi, logN)})
    g["cases"]["ntt"] = ntt_cases

    # 2. basis switch (ModUp / ModDown conversion), N = 8
    src = p64[1:4]
    dst = p64[4:6] + sp64[:2]
    xs = [[rnd.randrange(q) for _ in range(8)] for q in src]
    g["cases"]["switch_basis"] = {"src": src, "dst": dst, "x": xs, "y": switch_basis(xs, src, dst)}

    # 3. full digit ModUp / ModDown / Rescale / automorphism in evaluation form, N = 16
    logN, N = 4, 16
    Q = [p64[0]] + p64[1:5]  # 5 Q primes
    P = sp64[:2]
    dnum = 2  # alpha = 3 -> digits {0,1,2}, {3,4}
    roots = {q: root(q, N) for q in Q + P}
    l = 5
    x_eval = [[rnd.randrange(q) for _ in range(N)] for q in Q[:l]]
    x_coef = [intt_def(x_eval[i], Q[i], roots[Q[i]], logN) for i in range(l)]
    alpha = (len(Q) + dnum - 1) // dnum
    modup = []
    for j in range(2):
        st, en = j * alpha, min((j + 1) * alpha, l)
        dsrc = Q[st:en]
        others = [Q[i] for i in range(l) if not (st <= i < en)] + P
        conv = switch_basis(x_coef[st:en], dsrc, others)
        conv_eval = [ntt_def(conv[t], others[t], roots[others[t]], logN) for t in range(len(others))]
        out = []
        it = iter(conv_eval)
        for i in range(l):
            out.append(x_eval[i] if st <= i < en else next(it))
        out += list(it)
        modup.append(out)
    # ModDown of a random extended polynomial
    y_eval = [[rnd.randrange(q) for _ in range(N)] for q in Q[:l] + P]
    yP_coef = [intt_def(y_eval[l + k], P[k], roots[P[k]], logN) for k in range(len(P))]
    conv = switch_basis(yP_coef, P, Q[:l])
    Pprod = P[0] * P[1]
    moddown = []
    for i in range(l):
        q = Q[i]
        ce = ntt_def(conv[i], q, roots[q], logN)
        pinv = pow(Pprod % q, -1, q)
        moddown.append([((y_eval[i][k] - ce[k]) * pinv) % q for k in range(N)])
    # Rescale of x (drop Q[l-1]) by big-integer rounding identity
    xs_big = [crt([x_coef[i][k] for i in range(l)], Q[:l])[0] for k in range(N)]
    qt = Q[l - 1]
    resc = []
    for i in range(l - 1):
        q = Q[i]
        coef = []
        for k in range(N):
            t = xs_big[k] % qt
            if t > (qt >> 1):
                t -= qt
            v = (xs_big[k] - t)
            assert v % qt == 0
            coef.append((v // qt) % q)
        resc.append(ntt_def(coef, q, roots[q], logN))
    # automorphism X -> X^k on limb 0, k = 5^3 mod 2N
    kidx = pow(5, 3, 2 * N)
    a0 = x_coef[0]
    sub = [0] * N
    for i in range(N):
        e = (i * kidx) % (2 * N)
        if e < N:
            sub[e] = (sub[e] + a0[i]) % Q[0]
        else:
            sub[e - N] = (sub[e - N] - a0[i]) % Q[0]
    g["cases"]["eval16"] = {
        "logN": logN, "Q": Q, "P": P, "dnum": dnum, "l": l, "roots": [roots[q] for q in Q + P],
        "x_eval": x_eval, "modup": modup, "y_eval": y_eval, "moddown": moddown,
        "rescale": resc, "auto_k": kidx, "auto_limb0": ntt_def(sub, Q[0], roots[Q[0]], logN),
    }
    json.dump(g, open(os.path.join(HERE, "golden_small.json"), "w"))


if __name__ == "__main__":
    main()
