"""Benchmark of the hybrid key-switching hot path on one MI355X.

Workload (the headline configuration): N = 2^16, 30 Q limbs (one 60-bit +
29 59-bit primes, "L = 29"), K = 10 special 60-bit primes, dnum = 3.  One
step processes a batch of 64 ciphertexts whose limb counts are drawn
uniformly from 6..30 (levels 5..29) with a fixed seed; each ciphertext gets
HMult + relinearize + Rescale and a Rotate by an index drawn uniformly from
[1, 2^15).  Every HMult and every Rotate is one key switch, so a step is
128 key-switch operations; `value` is key-switch ops/s with all inputs and
keys resident in HBM.  Ciphertext and key limbs are synthetic (seeded,
uniform mod q_i): throughput does not depend on their values.

The roofline entry is for the forward NTT (column + row pass), timed live
with HIP events on the stream it runs on, against HBM and against VALU issue
(weighted by measured per-opcode costs from profiles/); the CPU baseline is
the repo's C oracle (own CPU reference, not OpenFHE) on a bounded sample of
the same operation.

Further legs: 8 full-slot bootstraps (N = 2^16, FLEXIBLEAUTO, sparse-secret
encapsulation) captured in one HIP graph ("bootstrap"); Poisson arrivals of
Chebyshev-series, rotate-and-sum and bootstrap requests through the
continuous-batching loop (fideslib_amd/serve.py) with a warm per-shape graph
cache, reported as p50/p99 latency at several rates, the highest rate within
the p99 bound and the same arrivals through the batched steps alone ("serve").  With
the CPU baseline, the oracle also checks one GPU bootstrap per input level
with real keys and reports its precision in bits.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode plain|batched]
                       [--no-serve] [--no-boot] [--serve-requests R] [--serve-rate LAMBDA]
                       [--serve-load-points L1,L2] [--no-cpu] [--no-autotune]
"""
from __future__ import annotations

import argparse
import ctypes
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fideslib_amd import capi, params  # noqa: E402

HBM_PEAK_GBS = 8000.0
SEED = 20261203
NCT = 64


# ----------------------------------------------------------------- HIP events
class HipEvents:
    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def event(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream=None):
        assert self.hip.hipEventRecord(e, stream) == 0

    def elapsed_ms(self, a, b) -> float:
        assert self.hip.hipEventSynchronize(b) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b) == 0
        return float(ms.value)


def rand_limbs(rng, primes, N):
    return np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes])


# ----------------------------------------------------------------- workload
class Workload:
    def __init__(self, prm, ctx, rng):
        N, L = prm.N, prm.L
        self.N = N
        self.levels = [int(x) for x in rng.integers(6, L + 1, NCT)]
        self.rots = [int(x) for x in rng.integers(1, 1 << 15, NCT)]
        self.kidx = [pow(5, r, 2 * N) for r in self.rots]
        pool_a = np.stack([rand_limbs(rng, prm.q, N) for _ in range(2)])  # [2][L][N]
        pool_b = np.stack([rand_limbs(rng, prm.q, N) for _ in range(2)])
        self.pool_a, self.pool_b = pool_a, pool_b
        # batched outputs (r: HMult+relin+Rescale, o: its rotation) and the
        # plain path's own (m: HMult+relin, pr: Rescale, po: rotation), so the
        # two can be compared bit for bit after timing
        self.a, self.b, self.m, self.r, self.o, self.pr, self.po = [], [], [], [], [], [], []
        for l in self.levels:
            self.a.append(capi.DeviceBuffer.from_array(np.ascontiguousarray(pool_a[:, :l])))
            self.b.append(capi.DeviceBuffer.from_array(np.ascontiguousarray(pool_b[:, :l])))
            self.m.append(capi.DeviceBuffer(2 * l * N))
            for lst in (self.r, self.o, self.pr, self.po):
                lst.append(capi.DeviceBuffer(2 * (l - 1) * N))
        LK = prm.L + prm.K
        key_host = np.stack([np.stack([rand_limbs(rng, prm.q + prm.p, N) for _ in range(prm.dnum)])
                             for _ in range(2)])  # [2][dnum][L+K][N]
        assert key_host.shape == (2, prm.dnum, LK, N)
        self.key_host = key_host
        self.rlk = ctx.upload_key(key_host)
        # one resident key per rotation index (same synthetic content, distinct HBM copies)
        self.rot_keys = [ctx.upload_key(key_host) for _ in range(NCT)]

    def step_batched(self, ctx, stream=None):
        ctx.hmult_batch(self.a, self.b, self.levels, self.rlk, self.r, rescale=True, stream=stream)
        ctx.rotate_batch(self.r, [l - 1 for l in self.levels], self.rot_keys, self.kidx, self.o, stream=stream)

    def step_plain(self, ctx, stream=None):
        """Per-ciphertext API calls as an application issues them: EvalMult,
        Rescale, EvalRotate (three calls, no cross-call fusion)."""
        for i, l in enumerate(self.levels):
            ctx.hmult(self.a[i], self.b[i], l, self.rlk, self.m[i], stream)
            ctx.rescale(self.m[i], l, self.pr[i], stream)
            ctx.rotate(self.pr[i], l - 1, self.rot_keys[i], self.kidx[i], self.po[i], stream)

    def step_plain_fused(self, ctx, stream=None):
        """Per-ciphertext calls with the batched path's fusions: HMult+relin
        with Rescale fused into ModDown (a batch of one), then EvalRotate."""
        for i, l in enumerate(self.levels):
            ctx.hmult_batch([self.a[i]], [self.b[i]], [l], self.rlk, [self.pr[i]], rescale=True, stream=stream)
            ctx.rotate(self.pr[i], l - 1, self.rot_keys[i], self.kidx[i], self.po[i], stream)

    def outputs(self, which):
        """Host copies of one path's outputs: [(rescaled, rotated)] per ciphertext."""
        r, o = (self.r, self.o) if which == "batched" else (self.pr, self.po)
        return [(r[i].download((2, l - 1, self.N)), o[i].download((2, l - 1, self.N)))
                for i, l in enumerate(self.levels)]


ROOF_LIMBS = 960  # 32 limbs for each of the 30 Q primes: the size of the batched step's NTT launches
ROOF_REPS = 10
ROOF_LAUNCHES = 2 + ROOF_REPS  # the roofline's warm-up and timed launches, the first of this size in a run
ROOF_KERNELS = ("void k_ntt_fwd_col<16, 0, 1>", "void k_ntt_fwd_row<16, 0, 1>")  # primes below 2^60
# the two passes of a plain 32-limb transform at N = 2^16 (cfg2's 51-bit primes)
CFG2_KERNELS = ("void k_ntt_fwd_col8<1>", "void k_ntt_fwd_row_small<16, 1>")
VALU_CLOCK_GHZ = 2.4  # MI355X peak engine clock


def ntt_roofline(ctx, prm, ev, reps=ROOF_REPS):
    """Forward NTT (column + row pass) of ROOF_LIMBS limbs of N = 2^16,
    grouped by prime as the batched executor issues them, timed with HIP
    events on the stream it runs on.  Algorithmic bytes: one read and one
    write of every limb, 2 N 8 bytes per limb.

    A 256 MiB device-to-device copy runs first: its hardware FETCH/WRITE
    counters calibrate the counter scale when this command is profiled.
    The measured per-launch HBM traffic (from those counters) is read from
    profiles/ntt_fwd_960x65536.json when present."""
    rng = np.random.default_rng(SEED + 1)
    cal_words = 32 << 20
    ca, cb = capi.DeviceBuffer(cal_words), capi.DeviceBuffer(cal_words)
    capi.check(capi.load().fides_copy_device(cb.ptr, ca.ptr, cal_words * 8, None), "calibration copy")
    ctx.sync()
    del ca, cb
    nq = 30
    x = rand_limbs(rng, prm.q[:nq], prm.N)
    buf = capi.DeviceBuffer(ROOF_LIMBS * prm.N)
    pids = [i for i in range(nq) for _ in range(ROOF_LIMBS // nq)]
    # limb t holds residues of prime pids[t]
    host = np.empty((ROOF_LIMBS, prm.N), np.uint64)
    for t, p in enumerate(pids):
        host[t] = x[p]
    buf.upload(host)
    del host
    for _ in range(2):
        ctx.ntt(buf, pids)
    ctx.sync()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    for _ in range(reps):
        ctx.ntt(buf, pids)
    ev.record(e1)
    ms = ev.elapsed_ms(e0, e1) / reps
    alg_bytes = 2 * prm.N * 8 * ROOF_LIMBS
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    # the same bytes moved by a plain device-to-device copy (read + write of
    # every limb once): what the memory system delivers to a streaming kernel.
    # Two halves of 240 MiB, so the 256 MiB calibration copy above stays the
    # largest copy of the run (the counter summary calibrates on it).
    dst = capi.DeviceBuffer(ROOF_LIMBS * prm.N)
    half = ROOF_LIMBS // 2 * prm.N

    def cp():
        for o in (0, half):
            capi.check(capi.load().fides_copy_device(dst.ptr + 8 * o, buf.ptr + 8 * o, half * 8, None), "copy")
    cp()
    ctx.sync()
    ev.record(e0)
    for _ in range(reps):
        cp()
    ev.record(e1)
    copy_ms = ev.elapsed_ms(e0, e1) / reps
    del dst
    traffic, valu = None, None
    tf = os.path.join(ROOT, "profiles", "ntt_fwd_960x65536.json")
    if os.path.exists(tf):
        with open(tf) as f:
            prof = json.load(f)
        traffic = prof.get("hbm_bytes_per_launch")
        v = prof.get("valu")
        if v and all(k in v for k in ROOF_KERNELS):
            # VALU issue bound: every wave64 VALU instruction holds its SIMD
            # for 4 cycles; 1024 SIMDs at the 2.4 GHz peak engine clock
            inst = sum(v[k]["valu_inst_per_butterfly"] for k in ROOF_KERNELS)  # per butterfly of each pass
            wave_inst = inst * (prm.N // 2) * 8 * ROOF_LIMBS / 64
            t_min = wave_inst * 4 / (1024 * VALU_CLOCK_GHZ * 1e9)
            valu = {"bound": "valu issue (1024 SIMDs x %.1f GHz, 4 cycles per wave64 instruction)" % VALU_CLOCK_GHZ,
                    "inst_per_butterfly": {k: v[k]["valu_inst_per_butterfly"] for k in ROOF_KERNELS},
                    "min_launch_us": round(t_min * 1e6, 2), "frac": round(t_min / (ms * 1e-3), 4),
                    "profiled_busy": {k: v[k]["valu_busy"] for k in ROOF_KERNELS},
                    "profiled_clock_ghz": {k: v[k]["clock_ghz"] for k in ROOF_KERNELS}}
            w = weighted_valu_bound(prm, ms, {k: v[k]["clock_ghz"] for k in ROOF_KERNELS})
            if w:
                valu["weighted"] = w
    return {
        "kernel": "ntt_fwd = k_ntt_fwd_col + k_ntt_fwd_row, %d limbs x 2^16 (one fides_ntt call)" % ROOF_LIMBS,
        # 64-bit modular butterflies on 32-bit multipliers: VALU issue binds,
        # not HBM (achieved/peak/frac/traffic below are the HBM figures)
        "bound": "valu",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "launch_us": round(ms * 1e3, 2),
        "us_per_limb": round(ms * 1e3 / ROOF_LIMBS, 4),
        "alg_bytes_per_launch": alg_bytes,
        "copy_same_bytes": {"us": round(copy_ms * 1e3, 2),
                            "gbs": round(alg_bytes / (copy_ms * 1e-3) / 1e9, 1),
                            "note": "hipMemcpyAsync device-to-device of the 960 limbs in two halves (one read "
                                    "and one write of each): the streaming rate the memory system gives; the two "
                                    "NTT passes move every limb twice"},
        "valu": valu,
    }


def weighted_valu_bound(prm, ms, clocks):
    """VALU issue bound of the roofline launch with every opcode weighted by
    its measured issue cost: the passes' instruction mix
    (profiles/ntt_fwd_valu_mix.json, counted in this source's gfx950
    assembly) times cycles per wave64 instruction
    (profiles/valu_rates_gfx950.json, the valu_probe microbenchmark; opcodes
    it does not list take the median of those it does).  Per pass:
    waves x sum(count x cycles) / (1024 SIMDs x clock), at the clock the
    profiled run measured for that pass and at the 2.4 GHz peak; frac is the
    bound over the event-timed launch."""
    fr = os.path.join(ROOT, "profiles", "valu_rates_gfx950.json")
    fm = os.path.join(ROOT, "profiles", "ntt_fwd_valu_mix.json")
    if not (os.path.exists(fr) and os.path.exists(fm)):
        return None
    rates = json.load(open(fr))["ops"]
    mix = json.load(open(fm))["kernels"]
    if not all(k in mix for k in ROOF_KERNELS):
        return None
    default = float(np.median(list(rates.values())))
    butterflies = ROOF_LIMBS * (prm.N // 2) * (prm.logN // 2)  # per pass: N/2 butterflies x 8 stages
    cyc, t_prof, t_peak = {}, 0.0, 0.0
    for k in ROOF_KERNELS:
        m = mix[k]
        per_thread = sum(n * rates.get(op, default) for op, n in m["ops"].items())
        cyc[k] = round(per_thread / m["butterflies_per_thread"], 2)
        waves = butterflies / m["butterflies_per_thread"] / 64
        t_prof += waves * per_thread / (1024 * clocks[k] * 1e9)
        t_peak += waves * per_thread / (1024 * VALU_CLOCK_GHZ * 1e9)
    return {"cycles_per_butterfly": cyc, "default_cycles": round(default, 3),
            "min_launch_us_at_profiled_clock": round(t_prof * 1e6, 2),
            "frac_at_profiled_clock": round(t_prof / (ms * 1e-3), 4),
            "min_launch_us_at_peak_clock": round(t_peak * 1e6, 2),
            "frac_at_peak_clock": round(t_peak / (ms * 1e-3), 4),
            "sources": ["profiles/ntt_fwd_valu_mix.json", "profiles/valu_rates_gfx950.json",
                        "profiles/ntt_fwd_960x65536.json (clock per pass)"]}


def step_alg_bytes(prm, ctx, levels):
    """Algorithmic HBM bytes of one batched step (SURVEY's unit costs,
    one limb = N 8 B): the HMult batch reads 4 l and writes 2 (l - 1) limbs
    per ciphertext and streams the relinearization key once (2 beta (l + K)
    limbs at the largest l); every rotation reads and writes 2 (l - 1) limbs
    and its own key (2 beta (l - 1 + K) limbs)."""
    K, lmax = prm.K, max(levels)
    hm = sum(4 * l + 2 * (l - 1) for l in levels) + 2 * ctx.digits(lmax) * (lmax + K)
    rt = sum(4 * (l - 1) + 2 * ctx.digits(l - 1) * (l - 1 + K) for l in levels)
    return prm.N * 8 * (hm + rt)


def step_roofline(prm, ctx, levels, ms):
    """The batched step against HBM (algorithmic bytes over the event-timed
    step) and against VALU issue (the step's VALU wave instructions from the
    counter run in profiles/keyswitch_step_cfg3.json, at one 4-cycle slot
    each and at the valu_probe median cost, over 1024 SIMDs at the clock the
    counter run measured).  The *_counter_run fractions divide that bound by
    the counter run's own kernel time (same run, same clock); the others by
    the event-timed step of this run, whose clock is not measured."""
    alg = step_alg_bytes(prm, ctx, levels)
    out = {"alg_bytes": alg, "achieved_gbs": round(alg / (ms * 1e-3) / 1e9, 1),
           "hbm_frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    fp = os.path.join(ROOT, "profiles", "keyswitch_step_cfg3.json")
    fr = os.path.join(ROOT, "profiles", "valu_rates_gfx950.json")
    if os.path.exists(fp):
        st = json.load(open(fp))["step"]
        inst, ghz = st["valu_wave_instructions"], st["clock_ghz"]
        t4 = inst * 4 / (1024 * ghz * 1e9)
        kt = st["kernel_ns_counter_run"] * 1e-9
        out.update({"hbm_bytes_counted": st.get("hbm_bytes"),
                    "valu_wave_instructions": inst, "profiled_clock_ghz": ghz,
                    "profiled_valu_busy": st["valu_busy"], "counter_run_kernel_ms": round(kt * 1e3, 3),
                    "valu_issue_bound_ms": round(t4 * 1e3, 3), "valu_frac": round(t4 / (ms * 1e-3), 4),
                    "valu_frac_counter_run": round(t4 / kt, 4)})
        if os.path.exists(fr):
            c = float(np.median(list(json.load(open(fr))["ops"].values())))
            tw = inst * c / (1024 * ghz * 1e9)
            out.update({"valu_issue_bound_ms_weighted": round(tw * 1e3, 3),
                        "valu_frac_weighted": round(tw / (ms * 1e-3), 4),
                        "valu_frac_weighted_counter_run": round(tw / kt, 4), "cycles_per_instruction": round(c, 3)})
    return out


def ntt_roofline_cfg2(ev, reps=50):
    """The forward NTT at the shape the 70%-of-HBM target is quoted on: one
    fides_ntt call over 32 limbs of N = 2^16 (FIDESlib's 51-bit test primes),
    HIP-event timed over `reps` launches.  Algorithmic bytes 2 N 8 per limb."""
    p2 = params.config2()
    rng = np.random.default_rng(SEED + 11)
    c2 = capi.Context(p2.logN, p2.q, p2.p, p2.dnum)
    buf = capi.DeviceBuffer.from_array(rand_limbs(rng, p2.q, p2.N))
    pids = list(range(p2.L))
    for _ in range(3):
        c2.ntt(buf, pids)
    c2.sync()
    e0, e1 = ev.event(), ev.event()
    ev.record(e0)
    for _ in range(reps):
        c2.ntt(buf, pids)
    ev.record(e1)
    ms = ev.elapsed_ms(e0, e1) / reps
    alg = 2 * p2.N * 8 * p2.L
    gbs = alg / (ms * 1e-3) / 1e9
    out = {"kernel": "ntt_fwd = k_ntt_fwd_col8 + k_ntt_fwd_row_small, 32 limbs x 2^16 (cfg2, one fides_ntt "
                     "call)", "launch_us": round(ms * 1e3, 2),
           "us_per_limb": round(ms * 1e3 / p2.L, 4), "alg_bytes_per_launch": alg, "achieved": round(gbs, 1),
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
    pf = os.path.join(ROOT, "profiles", "ntt_fwd_32x65536.json")
    if os.path.exists(pf):
        with open(pf) as f:
            out["profiled_pass_us"] = json.load(f).get("pass_durations_us")
    return out


# ----------------------------------------------------------------- batch size
def batch_size_leg(ctx, w, ev, sizes=(1, 2, 4, 8, 16, 32, 64), reps=5):
    """Key-switch ops/s of the batched step on the first n ciphertexts of the
    workload (n HMult+relin+Rescale, then n Rotate), for each n."""
    out = []
    for n in sizes:
        def step():
            ctx.hmult_batch(w.a[:n], w.b[:n], w.levels[:n], w.rlk, w.r[:n], rescale=True)
            ctx.rotate_batch(w.r[:n], [l - 1 for l in w.levels[:n]], w.rot_keys[:n], w.kidx[:n], w.o[:n])
        step()
        ctx.sync()
        e0, e1 = ev.event(), ev.event()
        ev.record(e0)
        for _ in range(reps):
            step()
        ev.record(e1)
        ms = ev.elapsed_ms(e0, e1) / reps
        out.append({"ciphertexts": n, "ms_per_step": round(ms, 3), "ops_per_s": round(2 * n / (ms * 1e-3), 1)})
    return out


# ----------------------------------------------------------------- single primitives
def primitive_inputs():
    """Seeded inputs of the primitive timings, shared by the GPU leg and the
    CPU baseline: config 2 (N = 2^16, 32 Q limbs of FIDESlib's 51-bit test
    primes + 8 special, dnum 4: digits of 8 limbs) and config 1 (N = 2^13, 4
    limbs, dnum 2) with a synthetic key."""
    rng = np.random.default_rng(SEED + 5)
    p2, p1 = params.config2(), params.config1()
    x2 = rand_limbs(rng, p2.q, p2.N)                                    # [32][N]
    ext2 = rand_limbs(rng, p2.q + p2.p, p2.N)                           # [40][N]
    a1 = np.stack([rand_limbs(rng, p1.q, p1.N) for _ in range(2)])      # [2][4][N]
    b1 = np.stack([rand_limbs(rng, p1.q, p1.N) for _ in range(2)])
    k1 = np.stack([np.stack([rand_limbs(rng, p1.q + p1.p, p1.N) for _ in range(p1.dnum)]) for _ in range(2)])
    return p2, p1, x2, ext2, a1, b1, k1


def primitives_leg(ev, reps=20):
    """GPU time of single operations (HIP events over `reps` calls): NTT and
    INTT of 32 limbs, ModUp of one 8-limb digit to 40 limbs and ModDown 40 ->
    32 at N = 2^16; HMult+relin+Rescale and Rotate(+1) at N = 2^13, 4 limbs."""
    p2, p1, x2, ext2, a1, b1, k1 = primitive_inputs()
    c2 = capi.Context(p2.logN, p2.q, p2.p, p2.dnum)
    c1 = capi.Context(p1.logN, p1.q, p1.p, p1.dnum)
    L2, N2, L1, N1 = p2.L, p2.N, p1.L, p1.N
    x = capi.DeviceBuffer.from_array(x2)
    e = capi.DeviceBuffer.from_array(ext2)
    up = capi.DeviceBuffer((L2 + p2.K) * N2)
    dn = capi.DeviceBuffer(L2 * N2)
    a, b = capi.DeviceBuffer.from_array(a1), capi.DeviceBuffer.from_array(b1)
    m = capi.DeviceBuffer(2 * L1 * N1)
    r = capi.DeviceBuffer(2 * (L1 - 1) * N1)
    o = capi.DeviceBuffer(2 * (L1 - 1) * N1)
    key = c1.upload_key(k1)
    pids = list(range(L2))

    def t(ctx, fn):
        fn()
        ctx.sync()
        e0, e1 = ev.event(), ev.event()
        ev.record(e0)
        for _ in range(reps):
            fn()
        ev.record(e1)
        return round(ev.elapsed_ms(e0, e1) / reps * 1e3, 2)

    res = {
        "cfg2": {"ntt_32_limbs_us": t(c2, lambda: c2.ntt(x, pids)),
                 "intt_32_limbs_us": t(c2, lambda: c2.intt(x, pids)),
                 "modup_8_to_40_us": t(c2, lambda: c2.modup(x, L2, 0, up)),
                 "moddown_40_to_32_us": t(c2, lambda: c2.moddown(e, L2, dn))},
        "cfg1": {"hmult_relin_rescale_us": t(c1, lambda: (c1.hmult(a, b, L1, key, m), c1.rescale(m, L1, r))),
                 "rotate_us": t(c1, lambda: c1.rotate(r, L1 - 1, key, 5, o))},
    }
    res["note"] = ("single calls, inputs resident; cfg2 = N 2^16, 32+8 limbs of FIDESlib's 51-bit test primes; "
                   "cfg1 = N 2^13, 4 limbs, dnum 2")
    return res


def cpu_primitives(reps=2):
    """The oracle on the inputs of primitives_leg (best of `reps`, ms)."""
    from oracle import oracle as O
    p2, p1, x2, ext2, a1, b1, k1 = primitive_inputs()
    o2 = O.Context(p2.logN, p2.q, p2.p, p2.dnum)
    o1 = O.Context(p1.logN, p1.q, p1.p, p1.dnum)
    pids = list(range(p2.L))

    def t(fn):
        best = None
        for _ in range(reps):
            t0 = time.time()
            fn()
            d = (time.time() - t0) * 1e3
            best = d if best is None else min(best, d)
        return round(best, 3)

    return {
        "cfg2": {"ntt_32_limbs_ms": t(lambda: o2.ntt_limbs(x2, pids)),
                 "intt_32_limbs_ms": t(lambda: o2.intt_limbs(x2, pids)),
                 "modup_8_to_40_ms": t(lambda: o2.modup_digit(x2, p2.L, 0)),
                 "moddown_40_to_32_ms": t(lambda: o2.moddown(ext2, p2.L))},
        "cfg1": {"hmult_relin_rescale_ms": t(lambda: o1.rescale_ct(o1.hmult_relin(a1, b1, p1.L, k1), p1.L)),
                 "rotate_ms": t(lambda: o1.rotate(np.ascontiguousarray(a1[:, :p1.L - 1]), p1.L - 1, k1, 5))},
    }


def cpu_baseline(prm, budget_s=12.0):
    """Oracle HMult + relin + Rescale at the same parameters, levels drawn as
    in the workload, for about budget_s seconds (own CPU reference)."""
    from oracle import oracle as O
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(SEED + 2)
    N = prm.N
    key = np.stack([np.stack([rand_limbs(rng, prm.q + prm.p, N) for _ in range(prm.dnum)]) for _ in range(2)])
    levels = [int(x) for x in rng.integers(6, prm.L + 1, 64)]
    done, t0 = 0, time.time()
    ops = []
    for i in range(100000):
        l = levels[i % len(levels)]
        a = np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])
        b = np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])
        t = time.time()
        octx.rescale_ct(octx.hmult_relin(a, b, l, key), l)
        ops.append(time.time() - t)
        done += 1
        if time.time() - t0 > budget_s:
            break
    el = sum(ops)
    cores = O.lib().or_get_threads()
    # the same operation on one thread, the first levels of the same draw
    prev = O.set_threads(1)
    st, t1 = [], time.time()
    for i in range(len(levels)):
        l = levels[i]
        a = np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])
        b = np.stack([rand_limbs(rng, prm.q[:l], N) for _ in range(2)])
        t = time.time()
        octx.rescale_ct(octx.hmult_relin(a, b, l, key), l)
        st.append(time.time() - t)
        if time.time() - t1 > budget_s / 2:
            break
    O.set_threads(prev)
    prims = cpu_primitives()
    return {
        "value": round(done / el, 3),
        "unit": "key-switch ops/s (HMult+relin+Rescale)",
        "cores": cores,
        "kind": "port",
        "sample": "%d HMult+relin+Rescale at N=2^16, limbs drawn U{6..30} (first %s), %.1f s of oracle time; "
                  "own CPU oracle (OpenMP), not OpenFHE" % (done, levels[:4], el),
        "op_ms": {"median": round(float(np.median(ops)) * 1e3, 2), "p99": round(float(np.percentile(ops, 99)) * 1e3, 2),
                  "note": "per HMult+relin+Rescale, levels of the draw"},
        "single_thread": {"value": round(len(st) / sum(st), 3), "cores": 1,
                          "sample": "%d of the same operations (the first levels of the draw), %.1f s"
                                    % (len(st), sum(st))},
        "primitives": prims,
    }


def bootstrap_precision_check(prm, ctx, w, boot, levels=(3, 17), h=192, seed=SEED + 11):
    """Part of the CPU baseline leg, with the oracle as checker only: it
    generates a real secret (ternary, Hamming weight h), its relinearization,
    rotation and encapsulation keys, encrypts a random full-slot message at
    each level's FLEXIBLEAUTO factor, bootstraps it on the GPU with the same
    plan as the timed leg (one request each through the request loop), and
    decrypts.  Reports bits of precision: -log2 max |decoded - message|."""
    from fideslib_amd import serve as S
    from oracle import ckks_codec as C
    from oracle import oracle as O
    N = prm.N
    t0 = time.time()
    octx = O.Context(prm.logN, prm.q, prm.p, prm.dnum)
    keys = C.Keys(octx, seed, h=h)
    host = {}
    for r in boot.rotations() + [None]:
        k = pow(5, r, 2 * N) if r is not None else 2 * N - 1
        host[k] = keys.rotation_key(k)
    host[boot.ENCAPS_DOWN], host[boot.ENCAPS_UP] = keys.encapsulation_keys(boot.cfg.encaps_h)
    dev = {k: ctx.upload_key(v) for k, v in host.items()}
    rlk = ctx.upload_key(keys.relin_key())
    del host
    t_keys = time.time() - t0
    be = S.GpuBackend(ctx, rlk, dev, 8 << 30)
    rng = np.random.default_rng(seed)
    bits = {}
    for lvl in levels:
        z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
        sc = boot_input_scale(prm, boot, lvl)
        P = boot.program(lvl, sc)
        ct = keys.encrypt(C.encode(z, sc if sc is not None else 2.0 ** prm.scale_bits, N), lvl)
        req = S.Request(0, P, 0.0, inputs=[ct], keep_result=True)
        S.Server(be, N).run([req], realtime=False)
        if req.status != "done":
            raise RuntimeError("bootstrap precision check: request %s" % req.status)
        dec = C.decode(keys.decrypt_coef(req.result, P.output.level), P.output.scale, N)
        bits[lvl] = round(float(C.precision_bits(dec, z)), 2)
    del be
    for k in list(dev):
        dev.pop(k).close()
    rlk.close()
    return {"min_bits": min(bits.values()), "bits_by_input_limbs": bits,
            "secret": "ternary, Hamming weight %d; encapsulation h'=%d" % (h, boot.cfg.encaps_h),
            "message": "uniform real and imaginary parts in [-1, 1), 2^15 slots",
            "keys_s": round(t_keys, 1), "total_s": round(time.time() - t0, 1)}


def summarize(trace_dir, fetch_dir=None, write_dir=None, valu_dir=None):
    """Summarise a rocprofv3 run of this script: per-kernel time of the
    batched step, the roofline launch's kernels, (with the two --pmc runs) the
    HBM bytes per roofline launch scaled by the calibration copy, and (with a
    --pmc run of SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU, SQ_BUSY_CYCLES and
    GRBM_GUI_ACTIVE) the VALU work of the roofline launch's two passes."""
    import csv
    import glob

    def rows(d, pat):
        f = glob.glob(os.path.join(d, "**", pat), recursive=True)
        return list(csv.DictReader(open(f[0]))) if f else []

    out = []
    tr = rows(trace_dir, "*kernel_trace.csv")
    roof_threads = 4096 * ROOF_LIMBS
    small = {}
    if tr:
        per = {}
        for r in tr:
            d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            key = r["Kernel_Name"].split("(")[0]
            per.setdefault(key, []).append(d)
        tot = sum(sum(v) for v in per.values())
        out.append("| kernel | calls | total ms | avg us | share |")
        out.append("|---|---|---|---|---|")
        for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
            out.append("| %s | %d | %.3f | %.2f | %.1f%% |" % (k, len(v), sum(v) / 1e6, sum(v) / len(v) / 1e3,
                                                             100.0 * sum(v) / tot))
        # the last batched step: from its first launch (the HMult batch's
        # tensor product) to the first launch of the NTT roofline that follows
        seq = sorted(tr, key=lambda r: int(r["Start_Timestamp"]))
        name = lambda r: r["Kernel_Name"].split("(")[0]
        grid = lambda r: int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"])
        r0 = next((i for i, r in enumerate(seq) if grid(r) == roof_threads and name(r) in ROOF_KERNELS), None)
        s0 = max((i for i in range(r0 or 0) if name(seq[i]) == "k_tensor_batch"), default=None)
        if r0 is not None and s0 is not None:
            # without the roofline's calibration copy, which runs just before it
            win, agg = [r for r in seq[s0:r0] if "copyBuffer" not in name(r)], {}
            for r in win:
                k = name(r)
                if k.startswith("void k_conv_batch<") and not k.startswith("void k_conv_batch<10"):
                    k = "k_conv_batch<1..9, true> (ModUp of partial digits)"
                a = agg.setdefault(k, [0, 0.0, 0])
                a[0] += 1
                a[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
                a[2] += int(r["Grid_Size_Y"])
            busy = sum(a[1] for a in agg.values())
            out.append("")
            out.append("last batched step: %d launches, %.3f ms of kernel time" % (len(win), busy / 1e3))
            out.append("| kernel | launches | us | share | tasks |")
            out.append("|---|---|---|---|---|")
            for k, (n, us, tasks) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
                out.append("| %s | %d | %.0f | %.1f%% | %d |" % (k, n, us, 100.0 * us / busy, tasks))
        roof = {}
        for r in seq:
            k = r["Kernel_Name"].split("(")[0]
            if int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) == roof_threads and k in ROOF_KERNELS:
                if len(roof.get(k, [])) < ROOF_LAUNCHES:
                    roof.setdefault(k, []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        if roof:
            out.append("")
            out.append("roofline launch (%d limbs): " % ROOF_LIMBS + ", ".join(
                "%s avg %.2f us over %d" % (k, sum(v) / len(v) / 1e3, len(v)) for k, v in roof.items()))
            out.append("sum of pass averages: %.2f us" % sum(sum(v) / len(v) / 1e3 for v in roof.values()))
        for r in seq:
            k = name(r)
            if k in CFG2_KERNELS and int(r["Grid_Size_Y"]) == 32:
                small.setdefault(k, []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        if small:
            out.append("")
            out.append("32-limb transforms (cfg2): " + ", ".join(
                "%s avg %.2f us over %d" % (k, sum(v) / len(v) / 1e3, len(v)) for k, v in small.items()))
            out.append("sum of pass averages: %.2f us" % sum(sum(v) / len(v) / 1e3 for v in small.values()))
    res = {}
    if small:
        res["cfg2_32_limbs"] = {"pass_durations_us": {k: round(sum(v) / len(v) / 1e3, 2) for k, v in small.items()},
                                "launches": {k: len(v) for k, v in small.items()}}

    def step_dispatches(rws):
        """Dispatch ids of the last batched step of a counter run: from the
        HMult batch's tensor product to the first roofline launch, without
        the calibration copy."""
        first = {}
        for r in rws:
            d = r["Dispatch_Id"]
            if d not in first:
                first[d] = (int(r["Start_Timestamp"]), r["Kernel_Name"].split("(")[0], int(r["Grid_Size"]))
        seq = sorted(first.items(), key=lambda kv: kv[1][0])
        r0 = next((i for i, (_, (_, k, g)) in enumerate(seq) if g == roof_threads and k in ROOF_KERNELS), None)
        s0 = max((i for i in range(r0 or 0) if seq[i][1][1] == "k_tensor_batch"), default=None)
        if r0 is None or s0 is None:
            return set()
        return {d for d, (_, k, _) in seq[s0:r0] if "copyBuffer" not in k}
    if fetch_dir and write_dir:
        def counter(d, name):
            vals = {}
            cal = None
            for r in sorted(rows(d, "*counter_collection.csv"), key=lambda r: int(r["Start_Timestamp"])):
                if r["Counter_Name"] != name:
                    continue
                g = int(r["Grid_Size"])
                v = float(r["Counter_Value"]) * 1024.0
                if "copyBuffer" in r["Kernel_Name"] and (cal is None or g > cal[0]):
                    cal = (g, v)
                k = r["Kernel_Name"].split("(")[0]
                if g == roof_threads and k in ROOF_KERNELS and len(vals.get(k, [])) < ROOF_LAUNCHES:
                    vals.setdefault(k, []).append(v)
            return cal, {k: sum(v) / len(v) for k, v in vals.items()}
        cal_bytes = 256 << 20
        (cf, fetch), (cw, write) = counter(fetch_dir, "FETCH_SIZE"), counter(write_dir, "WRITE_SIZE")
        fs, ws = cal_bytes / cf[1], cal_bytes / cw[1]
        tot = sum(fetch.values()) * fs + sum(write.values()) * ws
        res = {"hbm_bytes_per_launch": int(tot), "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
               "fetch_bytes": {k: int(v * fs) for k, v in fetch.items()},
               "write_bytes": {k: int(v * ws) for k, v in write.items()},
               "alg_bytes_per_launch": 2 * 65536 * 8 * ROOF_LIMBS}
        out.append("")
        out.append("HBM traffic per roofline launch: %.1f MB (fetch scale %.3f, write scale %.3f "
                   "from a 256 MiB calibration copy); algorithmic %.1f MB" % (
                       tot / 1e6, fs, ws, res["alg_bytes_per_launch"] / 1e6))
        # the last batched step's counted traffic
        step_bytes = 0.0
        for d, name, sc in ((fetch_dir, "FETCH_SIZE", fs), (write_dir, "WRITE_SIZE", ws)):
            rws = rows(d, "*counter_collection.csv")
            ids = step_dispatches(rws)
            step_bytes += sum(float(r["Counter_Value"]) * 1024.0 * sc for r in rws
                              if r["Dispatch_Id"] in ids and r["Counter_Name"] == name)
        res.setdefault("step", {})["hbm_bytes"] = int(step_bytes)
        out.append("HBM traffic of the last batched step (FETCH_SIZE + WRITE_SIZE, scaled): %.1f MB"
                   % (step_bytes / 1e6))
    if valu_dir:
        # per pass of the roofline launch: VALU instructions per butterfly
        # (wave instructions x 64 lanes / butterflies of the pass) and VALU
        # busy as rocprofv3 derives it: SQ_ACTIVE_INST_VALU counts 4-cycle
        # issue slots, GRBM_GUI_ACTIVE is summed over the 8 XCDs, so
        # busy = 4 ACTIVE / (1024 SIMDs x GUI / 8); a multiply that needs more
        # than one pass counts once, so this is a lower bound
        names = ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE")
        per, allk = {}, {}
        for r in sorted(rows(valu_dir, "*counter_collection.csv"), key=lambda r: int(r["Start_Timestamp"])):
            k = r["Kernel_Name"].split("(")[0]
            if r["Counter_Name"] not in names:
                continue
            a = allk.setdefault(k, {}).setdefault(r["Dispatch_Id"], {})
            a[r["Counter_Name"]] = a.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            if int(r["Grid_Size"]) != roof_threads or k not in ROOF_KERNELS:
                continue
            d = per.setdefault(k, {}).setdefault(r["Dispatch_Id"], {})
            d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
            d["ns"] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        bfly = ROOF_LIMBS * (65536 // 2) * 8  # butterflies of one pass (8 of the 16 stages)
        simds, xcds = 256 * 4, 8
        vres = {}
        for k, ds in per.items():
            ds = [d for d in ds.values() if all(n in d for n in names)][:ROOF_LAUNCHES]
            if not ds:
                continue
            inst = sum(d["SQ_INSTS_VALU"] for d in ds) / len(ds)
            act = sum(d["SQ_ACTIVE_INST_VALU"] for d in ds) / len(ds)
            gui = sum(d["GRBM_GUI_ACTIVE"] for d in ds) / len(ds)
            ns = sum(d["ns"] for d in ds) / len(ds)
            vres[k] = {"valu_inst_per_butterfly": round(inst * 64 / bfly, 2),
                       "valu_busy": round(4 * act / (simds * gui / xcds), 3),
                       "clock_ghz": round(gui / xcds / ns, 3), "launches": len(ds)}
        res["valu"] = vres
        # the last batched step: VALU wave instructions, busy and clock over its dispatches
        vrows = rows(valu_dir, "*counter_collection.csv")
        ids = step_dispatches(vrows)
        st = {n: 0.0 for n in names}
        dur = {}
        for r in vrows:
            if r["Dispatch_Id"] in ids and r["Counter_Name"] in names:
                st[r["Counter_Name"]] += float(r["Counter_Value"])
                dur[r["Dispatch_Id"]] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
        if ids and st["GRBM_GUI_ACTIVE"] > 0:
            kns = sum(dur.values())
            step = res.setdefault("step", {})
            step.update({"dispatches": len(ids), "kernel_ns_counter_run": kns,
                         "valu_wave_instructions": int(st["SQ_INSTS_VALU"]),
                         "valu_busy": round(4 * st["SQ_ACTIVE_INST_VALU"] / (simds * st["GRBM_GUI_ACTIVE"] / xcds), 4),
                         "clock_ghz": round(st["GRBM_GUI_ACTIVE"] / xcds / kns, 4)})
            out.append("")
            out.append("last batched step in the counter run: %d dispatches, %.3f ms of kernel time, %.3g VALU "
                       "wave instructions, VALU busy %.1f%% at %.2f GHz" % (
                           len(ids), kns / 1e6, st["SQ_INSTS_VALU"], 100 * step["valu_busy"], step["clock_ghz"]))
        # every kernel of the counter run: VALU busy over all its dispatches
        out.append("")
        out.append("VALU busy per kernel over the whole counter run (weighted by GRBM_GUI_ACTIVE; short "
                   "dispatches can read slightly above 100%):")
        out.append("| kernel | dispatches | busy |")
        out.append("|---|---|---|")
        tab = []
        for k, ds in allk.items():
            ds = [d for d in ds.values() if all(n in d for n in names)]
            act = sum(d["SQ_ACTIVE_INST_VALU"] for d in ds)
            gui = sum(d["GRBM_GUI_ACTIVE"] for d in ds)
            if ds and gui > 0:
                tab.append((gui, k, len(ds), 4 * act / (simds * gui / xcds)))
        for gui, k, n, b in sorted(tab, reverse=True)[:14]:
            out.append("| %s | %d | %.1f%% |" % (k, n, 100 * b))
        out.append("")
        out.append("VALU per roofline pass (busy = 4 SQ_ACTIVE_INST_VALU / (1024 SIMDs x GRBM_GUI_ACTIVE / 8)): " + "; ".join(
            "%s %.2f VALU instructions per butterfly, busy %.1f%% at %.2f GHz" % (
                k, v["valu_inst_per_butterfly"], 100 * v["valu_busy"], v["clock_ghz"])
            for k, v in vres.items()))
    return "\n".join(out), res


# ----------------------------------------------------------------- hoisted rotations
def hoisted_leg(prm, ctx, w, ev, n_rot=16, reps=5):
    """n_rot rotations of one 30-limb ciphertext: fides_rotate_hoisted (one
    shared ModUp) against fides_rotate_batch of n_rot distinct copies (one
    ModUp each)."""
    N, l = prm.N, prm.L
    rng = np.random.default_rng(SEED + 7)
    inp = capi.DeviceBuffer.from_array(np.stack([rand_limbs(rng, prm.q, N) for _ in range(2)]))
    outs = [capi.DeviceBuffer(2 * l * N) for _ in range(n_rot)]
    ks = [pow(5, int(r), 2 * N) for r in rng.integers(1, 1 << 15, n_rot)]
    keys = w.rot_keys[:n_rot]

    def t(fn):
        fn()
        ctx.sync()
        e0, e1 = ev.event(), ev.event()
        ev.record(e0)
        for _ in range(reps):
            fn()
        ev.record(e1)
        return ev.elapsed_ms(e0, e1) / reps

    # separate rotations: distinct copies of the input, so no ModUp is shared
    copies = [capi.DeviceBuffer(2 * l * N) for _ in range(n_rot)]
    for c in copies:
        ctx.copy(c, inp, 2 * l * N)
    ms_h = t(lambda: ctx.rotate_hoisted(inp, l, keys, ks, outs))
    ms_b = t(lambda: ctx.rotate_batch(copies, [l] * n_rot, keys, ks, outs))
    return {"rotations": n_rot, "limbs": l, "hoisted_ms": round(ms_h, 3), "separate_batched_ms": round(ms_b, 3),
            "speedup": round(ms_b / ms_h, 3)}


# ----------------------------------------------------------------- serving leg
SERVE_KMAX = 15          # rotate-and-sum over 2^k slots, k ~ U{1..15}: rotation keys for 2^0 .. 2^14
SERVE_ARENA_GIB = 64     # admission cap: ciphertext storage of admitted requests
SERVE_P99_BOUND_MS = 50.0  # latency bound at the default rate (80 req/s, ~20% of closed-batch capacity):
                           # about three bootstraps' worth of device time


def bench_keys(prm, w, boot):
    """Resident synthetic keys for every automorphism the serving mix and the
    bootstrap use: rotations by 2^i (rotate-and-sum), the bootstrap's
    baby/giant steps and conjugation, and the two switching keys of the
    sparse-secret encapsulation (distinct HBM copies of one synthetic key)."""
    N = prm.N
    idx = {pow(5, 1 << i, 2 * N) for i in range(SERVE_KMAX)}
    idx |= {pow(5, r, 2 * N) for r in boot.rotations()}
    idx.add(2 * N - 1)
    idx = sorted(idx) + boot.switch_keys()
    if len(idx) > len(w.rot_keys):
        raise RuntimeError("need %d resident keys, have %d" % (len(idx), len(w.rot_keys)))
    return {k: w.rot_keys[i] for i, k in enumerate(idx)}


def boot_plan(prm, ctx):
    """The cfg4 bootstrap: FLEXIBLEAUTO, sparse-secret encapsulation (h'=32)."""
    from fideslib_amd import bootstrap as BT
    boot = BT.BootstrapPlan(prm, ctx, BT.BootConfig.encapsulated(flexible=True))
    boot.build_plaintexts()
    return boot


def boot_input_scale(prm, boot, level):
    """A fresh ciphertext's scale at `level` limbs (its FLEXIBLEAUTO factor)."""
    if not boot.cfg.flexible:
        return None
    return params.flexible_scaling_factors(prm.q)[level - 1]


def bootstrap_leg(prm, ctx, w, boot, keys, n_boot=8):
    """n_boot full-slot bootstraps (N=2^16, 30 limbs -> 14 limbs, FLEXIBLEAUTO
    factors, sparse-secret encapsulation) served together by the request
    loop; synthetic keys and ciphertexts (values do not affect timing).
    Precision is measured with real keys by bootstrap_precision_check."""
    from fideslib_amd import serve as S
    N = prm.N
    rng = np.random.default_rng(SEED + 9)
    be = S.GpuBackend(ctx, w.rlk, keys, 16 << 30)
    src = capi.DeviceBuffer.from_array(np.stack([rand_limbs(rng, prm.q[:3], N) for _ in range(2)]))

    lvl = 3
    P3 = boot.program(lvl, boot_input_scale(prm, boot, lvl))

    def reqs(n):
        return [S.Request(i, P3, 0.0, inputs=[("dev", src.ptr)]) for i in range(n)]

    S.Server(be, N).run(reqs(1), realtime=False)   # warm-up
    ctx.sync()
    srv = S.Server(be, N)
    srv.run(reqs(n_boot), realtime=False)
    wall = srv.stats["wall_s"]
    # the same batch captured into one HIP graph, replayed (timed with events on its stream)
    stream = capi.Stream()
    foot = S.Server(be, N).footprint_bytes(P3)
    gbe = S.GpuBackend(ctx, w.rlk, keys, n_boot * foot + (256 << 20), stream=stream)
    gsrv = S.Server(gbe, N)
    cap = gsrv.capture(reqs(n_boot), scratch_bytes=40 << 30, task_bytes=1 << 30)
    cap.replay()
    stream.sync()
    ev = HipEvents()
    e0, e1 = ev.event(), ev.event()
    reps = 3
    ev.record(e0, stream.h)
    for _ in range(reps):
        cap.replay()
    ev.record(e1, stream.h)
    graph_ms = ev.elapsed_ms(e0, e1) / reps
    launches = gsrv.stats["launches"] + 1
    cap.close()
    del gsrv, gbe
    stream.close()
    return {"workload": "cfg4: %d full-slot bootstraps (2^15 slots) captured in one HIP graph, N=2^16, 30 limbs "
                        "(60+29x59 bit), dnum=3, level budget {3,3}, %s scaling, sparse-secret encapsulation "
                        "(h'=%s, K'=%g), EvalMod Chebyshev degree %d (baby-step giant-step) + %d double angles, "
                        "input %d limbs, output %d limbs"
                        % (n_boot, "FLEXIBLEAUTO" if boot.cfg.flexible else "FIXEDMANUAL", boot.cfg.encaps_h,
                           boot.cfg.K, boot.cfg.degree, boot.cfg.double_angle, lvl, boot.output_level()),
            "bootstraps_per_s": round(n_boot * 1e3 / graph_ms, 3), "ms_per_bootstrap": round(graph_ms / n_boot, 3),
            "graph_ms": round(graph_ms, 3), "graph_launches": launches, "replays": reps,
            "request_loop_bootstraps_per_s": round(n_boot / wall, 3),
            "keyswitch_ops": srv.stats["keyswitch_ops"], "steps": srv.stats["steps"],
            "launches": srv.stats["launches"],
            "precision_bits": None}


SERVE_GRAPH_SCRATCH_GIB = 2   # key-switch scratch shared by the cached request graphs of one lane
SERVE_BOOT_LANES = 1          # lanes (streams) holding a copy of every bootstrap shape (2 measured: same p99)


def serve_graph_cache(prm, ctx, w, boot, keys, coef, dmin):
    """Warm per-shape graph cache of the serving mix: one captured graph per
    start level of the Chebyshev program (dmin..L) on one lane, one per start
    level of the bootstrap (2..L) on each of SERVE_BOOT_LANES further lanes
    (so short requests never wait behind a bootstrap replay), and one per
    (start level, k) of the rotate-and-sum program on a last lane; each lane
    has its own stream, register memory and key-switch scratch."""
    from fideslib_amd import serve as S
    cheb = [S.chebyshev_program(prm, l, coef, -8.0, 8.0) for l in range(dmin, prm.L + 1)]
    boots = [boot.program(l, boot_input_scale(prm, boot, l)) for l in range(2, prm.L + 1)]
    rotsums = [S.rotsum_program(prm, l, k) for l in range(1, prm.L + 1) for k in range(1, SERVE_KMAX + 1)]
    sb = SERVE_GRAPH_SCRATCH_GIB << 30
    words = []
    lanes = []
    for progs in [cheb] + [boots] * SERVE_BOOT_LANES + [rotsums]:
        nw = S.GraphCache.words_needed(progs)
        words.append(nw)
        lanes.append((S.GpuBackend(ctx, w.rlk, keys, nw * 8, stream=capi.Stream()), capi.DeviceBuffer(sb // 8)))
    cache = S.GraphCache(lanes[0][0], prm.N, sb, scratch_ptr=lanes[0][1].ptr)
    boot_lanes = [cache.add_lane(be, scratch_ptr=x.ptr) for be, x in lanes[1:-1]]
    rot_lane = cache.add_lane(lanes[-1][0], scratch_ptr=lanes[-1][1].ptr)
    cache._scratch = [x for _, x in lanes]
    t0 = time.time()
    for P in cheb:
        cache.add(P)
    for i, lane in enumerate(boot_lanes):   # one more copy of every bootstrap shape per lane
        for P in boots:
            cache.add(P, copies=i + 1, lane=lane)
    task_bytes, cache.task_bytes = cache.task_bytes, 1 << 20   # small programs: small task pools (doubled on demand)
    for P in rotsums:
        cache.add(P, lane=rot_lane)
    cache.task_bytes = task_bytes
    ctx.sync()
    info = dict(shapes=cache.stats["shapes"], graphs=cache.stats["graphs"], lanes=len(cache.lanes),
                registers_gib=round(sum(words) * 8 / 2 ** 30, 2), scratch_gib=len(lanes) * SERVE_GRAPH_SCRATCH_GIB,
                task_pools_gib=round(cache.stats["task_bytes"] / 2 ** 30, 2), capture_s=round(time.time() - t0, 2))
    return cache, info


def serve_leg(prm, ctx, w, boot, keys, n_req, rate, seed, more_rates=(), search=True):
    """Poisson(rate) arrivals of the serving mix through the continuous-
    batching loop with a warm per-shape graph cache (every Chebyshev and
    bootstrap start level captured once; a request whose shape has a free
    graph runs as one replay on the cache's stream, concurrent with the
    batched steps; the others join the steps); keys resident.  Mix: 60%
    degree-15 Chebyshev series of the sigmoid on slots in [-8, 8] (start
    limbs U{7..30}), 30% rotate-and-sum over 2^k slots (k ~ U{1..15}, start
    limbs U{1..30}), 10% full-slot bootstrap (start limbs U{2..30}).  Inputs
    are copied from resident ciphertexts at admission.  The same arrivals
    and programs also run through the batched steps alone
    ("batched_only")."""
    from fideslib_amd import serve as S
    N = prm.N
    rng = np.random.default_rng(seed)
    be = S.GpuBackend(ctx, w.rlk, keys, SERVE_ARENA_GIB << 30)
    top = np.stack([rand_limbs(rng, prm.q, N) for _ in range(2)])
    pool = {l: capi.DeviceBuffer.from_array(np.ascontiguousarray(top[:, :l])) for l in range(1, prm.L + 1)}
    coef = S.chebyshev_coefficients(lambda x: 1.0 / (1.0 + np.exp(-x)), 15, -8.0, 8.0)
    dmin = S.chebyshev_depth(15) + 1

    def make(n, arrivals):
        out = []
        for i in range(n):
            u = rng.random()
            if u < 0.6:
                P = S.chebyshev_program(prm, int(rng.integers(dmin, prm.L + 1)), coef, -8.0, 8.0)
            elif u < 0.9:
                P = S.rotsum_program(prm, int(rng.integers(1, prm.L + 1)), int(rng.integers(1, SERVE_KMAX + 1)))
            else:
                lb = int(rng.integers(2, prm.L + 1))
                P = boot.program(lb, boot_input_scale(prm, boot, lb))
            out.append(S.Request(i, P, float(arrivals[i]), inputs=[("dev", pool[P.inputs[0].level].ptr)]))
        return out

    def kinds_of(reqs):
        kinds = {}
        for name in ("chebyshev15", "rotsum", "bootstrap"):
            sub = [r for r in reqs if r.program.name.startswith(name)]
            if sub:
                kinds[name] = {k: round(v, 3) for k, v in S.latency_summary(sub).items()}
        return kinds

    # warm-up, then capacity of the batched steps: every request present at t = 0
    S.Server(be, N).run(make(16, np.zeros(16)), realtime=False)
    ctx.sync()
    cap_reqs = make(n_req, np.zeros(n_req))
    srv = S.Server(be, N)
    srv.run(cap_reqs, realtime=False)
    cap = dict(req_per_s=round(n_req / srv.stats["wall_s"], 2),
               keyswitch_ops_per_s=round(srv.stats["keyswitch_ops"] / srv.stats["wall_s"], 2),
               steps=srv.stats["steps"], launches=srv.stats["launches"], ops=srv.stats["ops"],
               arena_peak_gib=round(srv.arena.peak * N * 8 / 2 ** 30, 2))
    print("bench: serve capacity %s" % cap, file=sys.stderr, flush=True)
    cache, cinfo = serve_graph_cache(prm, ctx, w, boot, keys, coef, dmin)
    print("bench: serve graph cache %s" % cinfo, file=sys.stderr, flush=True)
    # open loop at the stated rate, with the graph cache
    arr = np.cumsum(rng.exponential(1.0 / rate, n_req))
    draw = rng.bit_generator.state
    reqs = make(n_req, arr)
    srv = S.Server(be, N)
    srv.run(reqs, realtime=True, graphs=cache)
    lat = S.latency_summary(reqs)
    # the same arrivals and programs through the batched steps alone
    after = rng.bit_generator.state
    rng.bit_generator.state = draw
    reqs_b = make(n_req, arr)
    rng.bit_generator.state = after
    srv_b = S.Server(be, N)
    srv_b.run(reqs_b, realtime=True)
    lat_b = S.latency_summary(reqs_b)
    batched_only = {
        "note": "the same arrivals and programs without the graph cache: every request joins the batched steps",
        "p50_ms": round(lat_b["p50_ms"], 3), "p99_ms": round(lat_b["p99_ms"], 3),
        "mean_ms": round(lat_b["mean_ms"], 3), "p99_within_bound": bool(lat_b["p99_ms"] <= SERVE_P99_BOUND_MS),
        "by_kind": kinds_of(reqs_b), "achieved_req_per_s": round(n_req / srv_b.stats["wall_s"], 2),
        "steps": srv_b.stats["steps"], "launches": srv_b.stats["launches"],
    }
    print("bench: serve batched only %s" % {k: v for k, v in batched_only.items() if k != "note"},
          file=sys.stderr, flush=True)
    res = {
        "workload": "cfg5: Poisson arrivals, N=2^16, 30 Q limbs, dnum=3; mix 60%% degree-15 Chebyshev (sigmoid "
                    "on [-8,8], start limbs U{%d..30}), 30%% rotate-and-sum 2^k slots (k~U{1..15}, start limbs "
                    "U{1..30}), 10%% bootstrap (start limbs U{2..30}); keys resident; admission cap %d GiB; "
                    "warm graph cache of every Chebyshev, bootstrap and rotate-and-sum shape (lanes: Chebyshev, %d bootstrap, rotate-and-sum)"
                    % (dmin, SERVE_ARENA_GIB, SERVE_BOOT_LANES),
        "rate_req_per_s": rate,
        "requests": n_req,
        "p50_ms": round(lat["p50_ms"], 3),
        "p99_ms": round(lat["p99_ms"], 3),
        "p99_bound_ms": SERVE_P99_BOUND_MS,
        "p99_within_bound": bool(lat["p99_ms"] <= SERVE_P99_BOUND_MS),
        "mean_ms": round(lat["mean_ms"], 3),
        "by_kind": kinds_of(reqs),
        "graph_replays": sum(r.graph for r in reqs),
        "achieved_req_per_s": round(n_req / srv.stats["wall_s"], 2),
        "keyswitch_ops_per_s_batched_steps": round(srv.stats["keyswitch_ops"] / srv.stats["wall_s"], 2),
        "steps": srv.stats["steps"],
        "launches": srv.stats["launches"],
        "arena_peak_gib": round(srv.arena.peak * N * 8 / 2 ** 30, 2),
        "graph_cache": cinfo,
        "batched_only": batched_only,
        "capacity_closed_batch": cap,
    }
    print("bench: serve %s" % {k: res[k] for k in ("p50_ms", "p99_ms", "graph_replays")}, file=sys.stderr,
          flush=True)

    # the same mix at higher arrival rates (fresh draws), with the graph cache
    def at_rate(r2):
        arr = np.cumsum(rng.exponential(1.0 / r2, n_req))
        rq = make(n_req, arr)
        s2 = S.Server(be, N)
        s2.run(rq, realtime=True, graphs=cache)
        l2 = S.latency_summary(rq)
        pt = {"rate_req_per_s": r2, "p50_ms": round(l2["p50_ms"], 3), "p99_ms": round(l2["p99_ms"], 3),
              "achieved_req_per_s": round(n_req / s2.stats["wall_s"], 2), "graph_replays": sum(r.graph for r in rq)}
        print("bench: serve at %g req/s %s" % (r2, pt), file=sys.stderr, flush=True)
        return pt

    pts = [at_rate(r2) for r2 in more_rates]
    if pts:
        res["load_points"] = pts
    if search:
        # highest arrival rate whose p99 met the bound, read conservatively:
        # only passing rates below the lowest failing one count (p99 of 400
        # requests is noisy near the bound); bisection to 10 req/s between
        # them, or upward probes (x1.5) while everything passes
        ok = [rate] if res["p99_within_bound"] else []
        bad = [] if res["p99_within_bound"] else [rate]
        for p in pts:
            (ok if p["p99_ms"] <= SERVE_P99_BOUND_MS else bad).append(p["rate_req_per_s"])
        probes = []
        while len(probes) < 6:
            hi = min(bad) if bad else None
            lo = max([r for r in ok if hi is None or r < hi], default=None)
            if lo is None or (hi is not None and hi - lo <= 10.0):
                break
            nxt = float(round(lo * 1.5)) if hi is None else float(round((lo + hi) / 2.0))
            p = at_rate(nxt)
            probes.append(p)
            (ok if p["p99_ms"] <= SERVE_P99_BOUND_MS else bad).append(nxt)
        hi = min(bad) if bad else None
        lo = max([r for r in ok if hi is None or r < hi], default=None)
        res["max_rate_within_bound"] = {
            "rate_req_per_s": lo, "p99_bound_ms": SERVE_P99_BOUND_MS, "next_failing_rate": hi,
            "probes": probes,
            "note": "highest measured Poisson rate (fresh arrivals per probe, graph cache on) with p99 <= the "
                    "bound and below every rate that failed it; bisection to 10 req/s, or upward probes while "
                    "all rates pass (at most 6 probes); p99 of 400 requests varies by several ms from run to run"}
    cache.close()
    return res


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize":
        a = sys.argv[2:]
        text, res = summarize(*(a + [None] * (4 - len(a)))[:4])
        print(text)
        if res:
            print(json.dumps(res))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="batched", choices=["plain", "batched"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-boot", action="store_true")
    ap.add_argument("--boot-only", action="store_true", help="only the bootstrap leg (for profiling)")
    ap.add_argument("--serve-only", action="store_true", help="only the serving leg (for profiling)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=80.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-load-points", default="160,240",
                    help="further arrival rates (comma-separated) reported as load_points; empty for none")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(batches are independent per GPU and shard by ciphertext).", file=sys.stderr)
        sys.exit(2)
    prm = params.config3()
    ctx = capi.Context(prm.logN, prm.q, prm.p, prm.dnum)
    rng = np.random.default_rng(SEED)
    print("bench: building workload (64 ciphertexts, 65 keys)", file=sys.stderr, flush=True)
    w = Workload(prm, ctx, rng)
    ev = HipEvents()
    tuning = None
    if not args.no_autotune:
        # launch shapes for the batched step's size classes, then for a single
        # ciphertext (the plain path); results are the same under every setting
        t0 = time.time()
        ctx.autotune(w.levels, reps=2)
        tuning = ctx.autotune([prm.L], reps=3)
        print("bench: autotune %s (%.1f s)" % (tuning, time.time() - t0), file=sys.stderr, flush=True)
    if args.boot_only:
        boot = boot_plan(prm, ctx)
        print(json.dumps({"bootstrap": bootstrap_leg(prm, ctx, w, boot, bench_keys(prm, w, boot))}))
        return
    if args.serve_only:
        boot = boot_plan(prm, ctx)
        more = [float(x) for x in args.serve_load_points.split(",") if x.strip()]
        print(json.dumps({"serve": serve_leg(prm, ctx, w, boot, bench_keys(prm, w, boot), args.serve_requests,
                                             args.serve_rate, SEED + 5, more)}))
        return

    def timed(step):
        for _ in range(args.warmup):
            step(ctx)
        ctx.sync()
        e0, e1 = ev.event(), ev.event()
        t0 = time.time()
        ev.record(e0)
        for _ in range(args.steps):
            step(ctx)
        ev.record(e1)
        ms = ev.elapsed_ms(e0, e1)
        ctx.sync()
        return ms / args.steps, time.time() - t0

    ms_plain_fused, _ = timed(w.step_plain_fused)
    print("bench: plain (fused rescale) %.3f ms/step" % ms_plain_fused, file=sys.stderr, flush=True)
    plain_fused_out = w.outputs("plain")
    ms_plain, wall_plain = timed(w.step_plain)
    print("bench: plain %.3f ms/step" % ms_plain, file=sys.stderr, flush=True)
    plain_out = w.outputs("plain")
    ms_batch, wall_batch = timed(w.step_batched)
    print("bench: batched %.3f ms/step" % ms_batch, file=sys.stderr, flush=True)
    batch_out = w.outputs("batched")
    # the timed runs' own outputs, compared bit for bit across the three paths
    eq_plain = all(np.array_equal(x, y) for p, q in zip(batch_out, plain_out) for x, y in zip(p, q))
    eq_fused = all(np.array_equal(x, y) for p, q in zip(batch_out, plain_fused_out) for x, y in zip(p, q))
    digest = hashlib.sha256()
    for pair in batch_out:
        for x in pair:
            digest.update(x.tobytes())
    del plain_out, plain_fused_out, batch_out
    print("bench: batched == plain %s, == plain fused %s" % (eq_plain, eq_fused), file=sys.stderr, flush=True)
    ops_per_step = 2 * NCT
    ms_step = ms_batch if args.mode == "batched" else ms_plain
    wall = wall_batch if args.mode == "batched" else wall_plain
    value = ops_per_step / (ms_step * 1e-3)
    roof = ntt_roofline(ctx, prm, ev)
    roof["cfg2_32_limbs"] = ntt_roofline_cfg2(ev)
    roof["batched_step"] = step_roofline(prm, ctx, w.levels, ms_batch)
    print("bench: roofline %s" % roof, file=sys.stderr, flush=True)
    out = {
        "metric": "KeySwitch ops/s N=2^16 L=29 batched vs plain, bit-exact; serve p50/p99 latency",
        "value": round(value, 2),
        "unit": "key-switch ops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "wall_s": round(wall, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded uniform residues; values do not affect timing)",
        "config": {
            "workload": "cfg3: N=2^16, 30 Q limbs (60+29x59 bit), K=10 x 60-bit special, dnum=3; "
                        "64 ciphertexts, limbs U{6..30}; per ct HMult+relin+Rescale and Rotate r~U[1,2^15)",
            "mode": args.mode,
            "ops_per_step": ops_per_step,
            "plain_ops_per_s": round(ops_per_step / (ms_plain * 1e-3), 2),
            "plain_fused_ops_per_s": round(ops_per_step / (ms_plain_fused * 1e-3), 2),
            "batched_ops_per_s": round(ops_per_step / (ms_batch * 1e-3), 2),
            "batched_over_plain": round(ms_plain / ms_batch, 3),
            "batched_over_plain_fused": round(ms_plain_fused / ms_batch, 3),
            "plain_path": "per ciphertext: fides_hmult, fides_rescale, fides_rotate (EvalMult, Rescale, EvalRotate)",
            "plain_fused_path": "per ciphertext: fides_hmult_batch of one with Rescale fused into ModDown, fides_rotate",
            "batched_equals_plain": bool(eq_plain),
            "batched_equals_plain_fused": bool(eq_fused),
            "outputs_sha256": digest.hexdigest(),
            "K": prm.K,
            "seed": SEED,
            "tuning": tuning if tuning is not None else ctx.tuning,
            "autotuned": tuning is not None,
        },
        "roofline": roof,
    }
    out["config"]["hoisted_rotations"] = hoisted_leg(prm, ctx, w, ev)
    print("bench: hoisted %s" % out["config"]["hoisted_rotations"], file=sys.stderr, flush=True)
    out["config"]["batch_sizes"] = batch_size_leg(ctx, w, ev)
    print("bench: batch sizes %s" % out["config"]["batch_sizes"], file=sys.stderr, flush=True)
    out["config"]["primitives"] = primitives_leg(ev)
    print("bench: primitives %s" % out["config"]["primitives"], file=sys.stderr, flush=True)
    if not (args.no_serve and args.no_boot):
        print("bench: bootstrap plaintexts", file=sys.stderr, flush=True)
        boot = boot_plan(prm, ctx)
        keys = bench_keys(prm, w, boot)
    if not args.no_boot:
        out["bootstrap"] = bootstrap_leg(prm, ctx, w, boot, keys)
        print("bench: bootstrap %s" % out["bootstrap"], file=sys.stderr, flush=True)
    if not args.no_serve:
        print("bench: serving leg", file=sys.stderr, flush=True)
        more = [float(x) for x in args.serve_load_points.split(",") if x.strip()]
        out["serve"] = serve_leg(prm, ctx, w, boot, keys, args.serve_requests, args.serve_rate, SEED + 5, more)
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(prm)
        if not args.no_boot:
            # the oracle as checker: real keys, one GPU bootstrap per level, decrypted
            print("bench: bootstrap precision check (real keys)", file=sys.stderr, flush=True)
            chk = bootstrap_precision_check(prm, ctx, w, boot)
            out["bootstrap"]["precision_bits"] = chk["min_bits"]
            out["bootstrap"]["precision_check"] = chk
            print("bench: bootstrap precision %s" % chk, file=sys.stderr, flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
