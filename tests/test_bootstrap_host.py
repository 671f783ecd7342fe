"""Bootstrap tables and program, on the CPU.

* The merged special-FFT groups equal the decoding map z_j = sum_k w_k
  zeta^(5^j k) (checked against the definition) and its inverse.
* Each baby-step/giant-step plan evaluates its group exactly.
* The EvalMod polynomial (Chebyshev series of the shifted cosine plus
  double-angle steps) equals sin(2 pi u) in FP64.
* The whole bootstrap program, run op by op on the oracle with real keys
  (N = 2^12, sparse secret), decrypts to the input message; a plausible
  mistake (an argument bound K' below the actual |I|) is caught.
* Sparse packing (n_s < N/2 slots): the folded transforms on 2 n_s-periodic
  vectors equal the n_s-slot maps followed by the [Re | Im] layout (and
  back); the program decrypts on the oracle for 2, 64, 256 and 1024 slots
  (every message slot, replicated over N/2); leaving out the partial trace
  (a plausible mistake) is caught.
"""
import math

import numpy as np
import pytest

from fideslib_amd import bootstrap as B
from fideslib_amd import params
from fideslib_amd import serve as S


def decode_def(w, N):
    n = N // 2
    zeta = np.exp(1j * np.pi / N)
    return np.array([np.sum(w * zeta ** ((pow(5, j, 2 * N) * np.arange(n)) % (2 * N))) for j in range(n)])


@pytest.mark.parametrize("N,budget", [(64, 3), (256, 2), (4096, 3)])
def test_groups_equal_decoding_map(N, budget):
    n = N // 2
    rng = np.random.default_rng(N)
    w = rng.normal(size=n) + 1j * rng.normal(size=n)
    z = decode_def(w, N)
    br = B.bitrev(n)
    x = w[br]
    for D in B.transform_groups(n, N, budget, inverse=False):
        assert len(D) <= 2 ** (math.ceil(math.log2(n) / budget) + 1) - 1
        x = B.diag_apply(D, x)
    assert np.abs(x - z).max() < 1e-9 * np.abs(z).max()
    y = z
    for D in B.transform_groups(n, N, budget, inverse=True):
        y = B.diag_apply(D, y)
    assert np.abs(y - w[br]).max() < 1e-9 * np.abs(w).max()


def test_bsgs_plans_evaluate_groups():
    N = 4096
    n = N // 2
    rng = np.random.default_rng(1)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    for inv in (False, True):
        for D in B.transform_groups(n, N, 3, inv):
            P = B.plan_group(D, n, factor=0.25)
            assert max(len(t) for t in P.giants.values()) <= 16
            assert np.abs(B.bsgs_apply(P, x) - 0.25 * B.diag_apply(D, x)).max() < 1e-12 * np.abs(x).max() * 64


def test_evalmod_polynomial_fp64():
    cfg = B.BootConfig()
    bp = B.BootstrapPlan(params.boot_small(), None, cfg)
    eps = np.linspace(-1e-3, 1e-3, 201)
    worst = 0.0
    for I in range(-20, 21):
        u = I + eps
        v = np.polynomial.chebyshev.chebval(u / cfg.K, bp.cheb)
        for _ in range(cfg.double_angle):
            v = 2 * v * v - 1
        worst = max(worst, np.abs(v - np.sin(2 * np.pi * u)).max())
    assert worst < 1e-11


def test_plan_levels_and_rotations():
    bp = B.BootstrapPlan(params.config3(), None)
    # CoeffToSlot 3 + SlotToCoeff 3 + degree-55 series 6 (baby steps to T_8,
    # giants T_16 and T_32, three division levels) + 3 double angles
    assert bp.depth == 3 + 3 + 6 + 3
    assert bp.output_level() == 30 - bp.depth
    assert 0 < len(bp.rotations()) <= 48
    with pytest.raises(ValueError):
        B.BootstrapPlan(params.serve_small(), None)   # 12 limbs: too shallow


def _oracle_bootstrap(cfg, seed=3, h=64, patch=None):
    from oracle import ckks_codec as C
    from oracle import oracle as O
    from serve_oracle import run_oracle
    prm = params.boot_small()
    N = prm.N
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 5, h=h)
    bp = B.BootstrapPlan(prm, None, cfg, keep_host=True)
    rot = {}
    for r in bp.rotations() + [None]:
        k = pow(5, r, 2 * N) if r is not None else 2 * N - 1
        rot[k] = keys.rotation_key(k)
    if patch is not None:
        patch(bp)
    P = bp.program(3)
    if cfg.encaps_h:
        rot[bp.ENCAPS_DOWN], rot[bp.ENCAPS_UP] = keys.encapsulation_keys(cfg.encaps_h)
    rng = np.random.default_rng(seed)
    ns = bp.slots
    z = rng.uniform(-1, 1, ns) + 1j * rng.uniform(-1, 1, ns)
    ct = keys.encrypt(C.encode(z, 2.0 ** prm.scale_bits, N), 3)
    out = run_oracle(octx, P, [ct], keys.relin_key(), rot)
    dec = C.decode(keys.decrypt_coef(out, P.output.level), P.output.scale, N)
    # every slot: the message replicated over the N/2 slots
    return C.precision_bits(dec, np.tile(z, N // 2 // ns)), P


def test_bootstrap_on_the_oracle_decrypts():
    bits, P = _oracle_bootstrap(B.BootConfig())
    assert P.output.level == params.boot_small().L - B.BootstrapPlan(params.boot_small(), None).depth
    assert bits > 20, bits


def test_evalmod_polynomial_fp64_encapsulated():
    cfg = B.BootConfig.encapsulated()
    bp = B.BootstrapPlan(params.boot_small(), None, cfg)
    eps = np.linspace(-1e-3, 1e-3, 201)
    worst = 0.0
    for I in range(-15, 16):
        u = I + eps
        v = np.polynomial.chebyshev.chebval(u / cfg.K, bp.cheb)
        for _ in range(cfg.double_angle):
            v = 2 * v * v - 1
        worst = max(worst, np.abs(v - np.sin(2 * np.pi * u)).max())
    assert worst < 1e-12


def test_bootstrap_encapsulated_dense_secret():
    """Dense ternary secret (h ~ 2N/3): with sparse-secret encapsulation
    (weight-32 s') K' = 16 suffices; the program switches keys twice."""
    cfg = B.BootConfig.encapsulated()
    bits, P = _oracle_bootstrap(cfg, h=None)
    assert sum(1 for o in P.ops if o.kind == "rot" and o.arg[0] == 1) == 2
    assert bits > 20, bits


def test_bootstrap_dense_secret_without_encapsulation_fails():
    # the same K' = 16 without encapsulation: |I| of the dense secret reaches
    # several times K', the modular reduction is wrong
    bits, _ = _oracle_bootstrap(B.BootConfig(K=16.0), h=None)
    assert bits < 5, bits


def test_bootstrap_planted_mistake_is_caught():
    # K' = 4 cannot cover |I| up to ~8 for this secret: the modular reduction fails
    bits, _ = _oracle_bootstrap(B.BootConfig(K=4.0, degree=31))
    assert bits < 5, bits


def test_flexible_plan_lands_on_the_flexibleauto_factor():
    """BootConfig(flexible=True): the program's result carries the
    FLEXIBLEAUTO scaling factor of its level and decrypts to the message at
    it (oracle run of the whole program at N = 2^12, sparse-secret
    encapsulation, input at its own FLEXIBLEAUTO factor)."""
    import serve_oracle as SO
    from fideslib_amd import params as PR
    from oracle import ckks_codec as C
    from oracle import oracle as O
    prm = PR.boot_small()
    N = prm.N
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, 8, h=None)
    bp = B.BootstrapPlan(prm, None, B.BootConfig.encapsulated(flexible=True), keep_host=True)
    sf = PR.flexible_scaling_factors(prm.q)
    assert bp.out_scale == sf[bp.output_level() - 1]
    rot = {}
    for r in bp.rotations() + [None]:
        k = pow(5, r, 2 * N) if r is not None else 2 * N - 1
        rot[k] = keys.rotation_key(k)
    rot[bp.ENCAPS_DOWN], rot[bp.ENCAPS_UP] = keys.encapsulation_keys(32)
    rng = np.random.default_rng(21)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    lvl = 5
    ct = keys.encrypt(C.encode(z, sf[lvl - 1], N), lvl)
    P = bp.program(lvl, sf[lvl - 1])
    assert P.output.scale == sf[bp.output_level() - 1]
    out = SO.run_oracle(octx, P, [ct], keys.relin_key(), rot)
    dec = C.decode(keys.decrypt_coef(out, P.output.level), P.output.scale, N)
    assert C.precision_bits(dec, z) > 20


# ------------------------------------------------------------------ sparse packing
@pytest.mark.parametrize("ns", [2, 8, 64, 256])
def test_sparse_fold_matches_slot_maps(ns):
    """The folded transforms on 2 ns-periodic vectors: CoeffToSlot then
    x + conj(x) gives [2 Re w | 2 Im w] for w the ns-slot CoeffToSlot of the
    message; SlotToCoeff of [Re w | Im w] gives the ns-slot SlotToCoeff of w,
    replicated; every baby-step/giant-step plan evaluates its folded group."""
    N = 1 << 12
    rng = np.random.default_rng(ns)
    cts = B.transform_groups(ns, N, 3, inverse=True)
    stc = B.transform_groups(ns, N, 3, inverse=False)
    cts2, stc2 = B.sparse_fold(cts, stc, ns)
    z = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    w = z.copy()
    for D in cts:
        w = B.diag_apply(D, w)
    x = np.tile(z, 2)
    for D in cts2:
        P = B.plan_group(D, 2 * ns)
        x = B.bsgs_apply(P, x)
    assert np.allclose(x + np.conj(x), np.concatenate([2 * w.real, 2 * w.imag]), atol=1e-9)
    y = np.concatenate([w.real, w.imag])
    for D in stc2:
        y = B.bsgs_apply(B.plan_group(D, 2 * ns), y)
    back = w.copy()
    for D in stc:
        back = B.diag_apply(D, back)
    assert np.allclose(y, np.tile(back, 2), atol=1e-9)
    assert np.allclose(back, z, atol=1e-9)   # the two maps are inverse


def test_sparse_plan_levels_and_rotations():
    prm = params.boot_small()
    N = prm.N
    full = B.BootstrapPlan(prm, None)
    bp = B.BootstrapPlan(prm, None, B.BootConfig(slots=64))
    assert bp.sparse and bp.n == 128 and bp.depth == full.depth
    assert bp.trace_rotations() == [64, 128, 256, 512, 1024]   # j * 64 below N/2 = 2048
    assert set(bp.trace_rotations()) <= set(bp.rotations())
    assert all(0 < r < N // 2 for r in bp.rotations())
    small = B.BootstrapPlan(prm, None, B.BootConfig(slots=4))   # 2 stages: fewer groups than the budget
    assert len(small.cts_plans) == 2 and small.depth == full.depth - 2
    for bad in (1, 3, N):
        with pytest.raises(ValueError):
            B.BootstrapPlan(prm, None, B.BootConfig(slots=bad))


@pytest.mark.parametrize("ns,encaps", [(2, False), (64, False), (256, True), (1024, True)])
def test_sparse_bootstrap_on_the_oracle_decrypts(ns, encaps):
    cfg = B.BootConfig.encapsulated(slots=ns) if encaps else B.BootConfig(slots=ns)
    bits, P = _oracle_bootstrap(cfg, seed=ns, h=None if encaps else 64)
    # one EvalMod: half the full-slot bootstrap's HMults of the series
    full_hmul = sum(1 for o in _oracle_program(B.BootConfig()).ops if o.kind == S.HMUL)
    assert sum(1 for o in P.ops if o.kind == S.HMUL) * 2 == full_hmul
    assert bits > 20, bits


def _oracle_program(cfg):
    return B.BootstrapPlan(params.boot_small(), None, cfg, keep_host=True).program(3)


def test_sparse_bootstrap_without_the_trace_fails():
    # leaving out the partial trace keeps the non-subring part of q_0 I(X),
    # which CoeffToSlot of 64 slots folds into the slots: wrong result
    def no_trace(bp):
        bp.trace_rotations = lambda: []
    bits, _ = _oracle_bootstrap(B.BootConfig(slots=64), seed=7, patch=no_trace)
    assert bits < 5, bits
