// Two-pass LDS negacyclic NTT / INTT kernels (see ntt.h for the design).
#include "kernels.h"

// Columns per workgroup of the forward and inverse column passes (build
// options).  16 is the measured best on MI355X at N = 2^16: 32 and 64 columns
// give the memory system longer row segments and, at 64, wave-uniform
// phase-B twiddles, but the passes issue VALU work at the same rate and the
// smaller launches lose workgroups.
#ifndef FIDES_FCOL_W
#define FIDES_FCOL_W 16
#endif
#ifndef FIDES_ICOL_W
#define FIDES_ICOL_W 16
#endif
constexpr int kFwdColW = FIDES_FCOL_W;
constexpr int kInvColW = FIDES_ICOL_W;

template <int LOGN, int PASS, int W = 16>
struct NttCfg {
  static constexpr int L1 = LOGN / 2;               // log #rows (column length)
  static constexpr int L2 = LOGN - L1;              // log row length
  static constexpr int LS = PASS == 0 ? L1 : L2;    // stages in this pass
  static constexpr int A = (LS + 1) / 2;            // stages in register phase A
  static constexpr int B = LS - A;                  // stages in register phase B
  static constexpr int E = 1 << A;                  // elements per thread
  static constexpr int G = 1 << (A - B);            // phase-B groups per thread
  static constexpr int TPL = 1 << (LS - A);         // threads per line
  static constexpr int M = 1 << LS;                 // line length
  static constexpr int NLINES = PASS == 0 ? (1 << L2) : (1 << L1);
  // row pass: one wave per workgroup (the LDS footprint, 8.5 KiB per wave at
  // N = 2^16, then allows 18 waves per CU instead of 16 in 4-wave groups)
  static constexpr int LPW0 = PASS == 0 ? W : 64 / TPL;
  static constexpr int LPW = LPW0 < NLINES ? LPW0 : NLINES;  // lines per workgroup
  static constexpr int THREADS = LPW * TPL;
  static constexpr int S0 = PASS == 0 ? 0 : L1;     // first global stage
  // LDS line layout.  Column pass: lines are columns, consecutive threads sit
  // in consecutive lines, stride M + 1.  Row pass: consecutive threads share a
  // line, so phase B (thread r reads 2^B consecutive elements) would put lanes
  // 2^B u64 apart, an 8-way bank conflict at B = 4; one pad u64 after every
  // 2^B elements (index i + (i >> B)) spreads them over all banks, and the
  // stride M + M/2^B (16 mod 32 u64 at N = 2^16) keeps neighbouring lines on
  // disjoint banks for both access patterns.
  static constexpr int LDS_LINE = PASS == 0 ? M + 1 : M + (M >> B);
  __device__ static constexpr u32 lx(u32 i) { return PASS == 0 ? i : i + (i >> B); }
};

// ---------------------------------------------------------------- butterflies
// Forward Cooley-Tukey, lazy: X' = x + t, Y' = x + 2q - t with t = Y w mod q
// in [0, 2q) (any Y < 2^64) and x = X, or with RED x = X - m if X >= m.  So
// the outputs exceed x's bound by 2q.  WS: twiddles are wave-uniform.
template <bool WS = false, bool RED = true>
__device__ __forceinline__ void ct_bfly(u64& X, u64& Y, u64 w, u64 wp, u64 q, u64 q2, u64 m) {
  const u64 x = RED ? csub_bf(X, m) : X;
  const u64 t = shoup_lazy_bf<WS>(Y, w, wp, (u64)0 - q);
  X = x + t;
  Y = x + q2 - t;
}

// Reduction schedule of the forward transform, a function of the global
// stage only (so the column pass, the row pass and the fused row pass of the
// key switch agree).  Values entering global stage s lie below bnd(s) q; the
// input of stage 0 below 4q.  A stage reduces its X operands only when its
// outputs could otherwise reach MAXB q, the largest multiple of q that stays
// below 2^64, by one conditional subtraction of SUB q (which maps
// [0, 2 SUB q) onto [0, SUB q)):
//   primes up to 2^62: MAXB = 4, SUB = 2 (Harvey's butterfly, every stage);
//   primes below 2^60 (NARROW): MAXB = 16, SUB = 8, every fourth stage.
// The last stage brings X below 2q (subtractions of 8q, 4q, 2q as the bound
// needs), so the transform ends below 4q and its final reduction to [0, q)
// takes two subtractions per element instead of four.
enum : int { RED_NONE = 0, RED_SUB = 1, RED_FULL = 2 };

template <int NARROW, int LOGN>
struct FwdLazy {
  static constexpr int MAXB = NARROW ? 16 : 4;
  static constexpr int SUB = NARROW ? 8 : 2;
  static constexpr int kind_for(int s, int b) {
    return (s == LOGN - 1 && b > 2) ? RED_FULL : (b + 2 > MAXB ? RED_SUB : RED_NONE);
  }
  static constexpr int after(int s, int b) {
    const int k = kind_for(s, b);
    return (k == RED_FULL ? 2 : k == RED_SUB ? (b - SUB > SUB ? b - SUB : SUB) : b) + 2;
  }
  static constexpr int bnd(int s) {
    int b = 4;
    for (int i = 0; i < s; ++i) b = after(i, b);
    return b;
  }
  // every stage's input fits (b <= MAXB), a single subtraction of SUB q
  // suffices where one is made (b <= 2 SUB), the outputs fit, and the
  // transform ends below 4q
  static constexpr bool valid() {
    for (int s = 0; s < LOGN; ++s) {
      const int b = bnd(s), k = kind_for(s, b);
      if (b > MAXB || (k == RED_SUB && b > 2 * SUB)) return false;
      const int bx = k == RED_FULL ? 2 : k == RED_SUB ? (b - SUB > SUB ? b - SUB : SUB) : b;
      if (bx + 2 > MAXB) return false;
    }
    return bnd(LOGN) <= 4;
  }
};
static_assert(FwdLazy<0, 16>::valid() && FwdLazy<1, 16>::valid() && FwdLazy<1, 10>::valid() &&
                  FwdLazy<1, 17>::valid() && FwdLazy<0, 10>::valid(),
              "reduction schedule out of range");

// One forward butterfly of global stage s under the schedule.  Callers pass
// the stage from fully unrolled loops, so s is a constant and the branches
// fold; the schedule enters as a table of booleans (the lookups of an int
// table of kinds were left to run time).
template <int NARROW, int LOGN>
struct FwdSubTable {
  struct T {
    bool v[18];
  };
  static constexpr T make() {
    using Lz = FwdLazy<NARROW, LOGN>;
    T t{};
    for (int s = 0; s < 18; ++s) t.v[s] = s < LOGN && Lz::kind_for(s, Lz::bnd(s)) == RED_SUB;
    return t;
  }
  static constexpr T tab = make();
};

template <bool WS, int NARROW, int LOGN>
__device__ __forceinline__ void fwd_bfly(int s, u64& X, u64& Y, u64 w, u64 wp, u64 q, u64 q2, u64 msub) {
  constexpr int bl = FwdLazy<NARROW, LOGN>::bnd(LOGN - 1);  // the last stage reduces fully
  if (s == LOGN - 1) {
    if (bl > 8) X = csub_bf(X, q << 3);
    if (bl > 4) X = csub_bf(X, q << 2);
    X = csub_bf(X, q2);
    ct_bfly<WS, false>(X, Y, w, wp, q, q2, msub);
  } else if (FwdSubTable<NARROW, LOGN>::tab.v[s]) {
    ct_bfly<WS, true>(X, Y, w, wp, q, q2, msub);
  } else {
    ct_bfly<WS, false>(X, Y, w, wp, q, q2, msub);
  }
}

// The transform's output (below 4q) -> [0, q).
__device__ __forceinline__ u64 fwd_final(u64 x, u64 q, u64 q2) { return csub_bf(csub_bf(x, q2), q); }

// Inverse Gentleman-Sande, lazy: X, Y in [0, 2q) -> [0, 2q).
template <bool WS = false>
__device__ __forceinline__ void gs_bfly(u64& X, u64& Y, u64 w, u64 wp, u64 q, u64 q2) {
  const u64 s = csub_bf(X + Y, q2);
  Y = shoup_lazy_bf<WS>(X + q2 - Y, w, wp, (u64)0 - q);
  X = s;
}

// ------------------------------------------------------------- prologue/epilogue

// Limb data is in device (global) memory.  Pointers taken from task records
// are generic, which makes the compiler emit flat loads and stores; those
// also count against the LDS wait counter, so a wait for LDS data would wait
// for outstanding memory traffic too.  Accesses through these helpers are
// global loads and stores.
__device__ __forceinline__ u64 ldg(const u64* p, u64 i) {
  return ((const __attribute__((address_space(1))) u64*)p)[i];
}
__device__ __forceinline__ void stg(u64* p, u64 i, u64 v) { ((__attribute__((address_space(1))) u64*)p)[i] = v; }

// Epilogue in two parts: the operands of coefficient idx are loaded first (a
// thread issues a batch of loads before it waits for them), then combined
// with the transform output v in [0, q).
// OpenFHE NativeVector::SwitchModulus (centred lift) of values a[k] < om to
// modulus nm (reference src/CKKS/Rescale.cuh:14-40 quoted, :47-68): values
// above om/2 are treated as negative.  diff = nm - (om mod nm) comes
// precomputed with the task; the final reduction is a Shoup product by 1
// (exact for any input below 2^64) instead of a division.  The comparison of
// the moduli is uniform, so it is made once for all n values.
template <int n>
__device__ __forceinline__ void switch_modulus_all(u64 (&a)[n], u64 om, const PrimeC& NP, u64 diff) {
  const u64 nm = NP.q, half = om >> 1;
  if (nm > om) {
#pragma unroll
    for (int k = 0; k < n; ++k) a[k] = a[k] > half ? a[k] + (nm - om) : a[k];
  } else {
#pragma unroll
    for (int k = 0; k < n; ++k) a[k] = shoup(a[k] > half ? a[k] + diff : a[k], 1, NP.one_sh, nm);
  }
}

template <int EPI>
__device__ __forceinline__ void epilogue_load(u32 idx, const NttTask& t, u64& a, u64& b) {
  if constexpr (EPI == EPI_RESCALE || EPI == EPI_MODDOWN) {
    a = ldg(t.x, idx);
  } else if constexpr (EPI == EPI_MODDOWN_ADD) {
    a = ldg(t.x, idx);
    b = t.acc ? ldg(t.acc, idx) : 0;
  } else if constexpr (EPI == EPI_MDRS) {
    a = ldg(t.acc, idx);
    b = ldg(t.x, idx);
  }
}
// LAZY: v is the transform's output before its final reduction (below 4q,
// primes below 2^60), folded into the epilogue's own arithmetic: Shoup's
// product takes any 64-bit operand, a + 4q - v stays positive, and the
// ModDown+Rescale sum (below 6q) needs three conditional subtractions where
// the reduction of v and two modular additions took four.
template <int EPI, bool LAZY = false>
__device__ __forceinline__ u64 epilogue_apply(u64 v, u64 a, u64 b, const NttTask& t, u64 q) {
  if constexpr (EPI == EPI_NONE) {
    return v;
  } else if constexpr (EPI == EPI_RESCALE) {
    return add_mod(shoup(a, t.c0, t.c0sh, q), shoup(v, t.c1, t.c1sh, q), q);
  } else if constexpr (EPI == EPI_MDRS) {
    if constexpr (LAZY) {
      const u64 s = shoup(a, t.c0, t.c0sh, q) + shoup(b, t.c3, t.c3sh, q) + v;
      return csub_bf(csub_bf(csub_bf(s, q << 2), 2 * q), q);
    }
    return add_mod(add_mod(shoup(a, t.c0, t.c0sh, q), shoup(b, t.c3, t.c3sh, q), q), v, q);
  } else {
    u64 r = shoup(a + (LAZY ? 4 * q : q) - v, t.c0, t.c0sh, q);
    if constexpr (EPI == EPI_MODDOWN_ADD) {
      if (t.acc) r = add_mod(r, b, q);
    }
    return r;
  }
}

// ------------------------------------------------------------------ forward
// Column pass: global stages [0, L1).  Output kept lazy (below bnd(L1) q of
// the reduction schedule).
template <int LOGN, int PRO, int NARROW>
__global__ __launch_bounds__((NttCfg<LOGN, 0, kFwdColW>::THREADS)) void k_ntt_fwd_col(LimbSpan span, NttTables T, NttFuse F) {
  using C = NttCfg<LOGN, 0, kFwdColW>;
  __shared__ u64 lds[C::LPW * C::LDS_LINE];
  const NttTask tk = resolve_task(span, F, blockIdx.y, LOGN);
  u64* data = tk.data;
  const u32 pid = tk.pid;
  const u64 q = T.prm[pid].q, q2 = 2 * q, msub = (u64)FwdLazy<NARROW, LOGN>::SUB * q;
  const u64* tw = T.psi + ((u64)pid << LOGN);
  const u64* tws = T.psi_sh + ((u64)pid << LOGN);
  const u32 tid = threadIdx.x;
  // with 64 or more columns per workgroup a wave shares one r0, so the
  // phase-B twiddles are wave-uniform (scalar loads)
  constexpr bool WSB = C::LPW >= 64;
  const u32 ll = tid % C::LPW, r0 = WSB ? __builtin_amdgcn_readfirstlane(tid / C::LPW) : tid / C::LPW;
  const u32 col = blockIdx.x * C::LPW + ll;
  u64 x[C::E];
  // The prologues load all their operands before the arithmetic, and the
  // workgroup-uniform branch of switch_modulus is taken once for all of them
  // (per element it would split the loop into blocks, and the compiler would
  // wait for each load in turn).
  if constexpr (PRO == PRO_SWITCH_MODULUS) {
#pragma unroll
    for (int k = 0; k < C::E; ++k) x[k] = ldg(tk.src, ((r0 + C::TPL * k) << C::L2) | col);
    switch_modulus_all<C::E>(x, T.prm[tk.src_pid].q, T.prm[pid], tk.swdiff);
  } else if constexpr (PRO == PRO_MDRS) {
    u64 y[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
      const u32 e = ((r0 + C::TPL * k) << C::L2) | col;
      x[k] = ldg(tk.src, e);
      y[k] = ldg(tk.src2, e);
    }
    switch_modulus_all<C::E>(x, T.prm[tk.src_pid].q, T.prm[pid], tk.swdiff);
#pragma unroll
    for (int k = 0; k < C::E; ++k)
      x[k] = sub_mod(shoup(x[k], tk.c2, tk.c2sh, q), shoup(y[k], tk.c3, tk.c3sh, q), q);
  } else {
    const u64* in = tk.src ? tk.src : data;
#pragma unroll
    for (int k = 0; k < C::E; ++k) x[k] = ldg(in, ((r0 + C::TPL * k) << C::L2) | col);
  }
#pragma unroll
  for (int u = 0; u < C::A; ++u) {
    const int half = C::E >> (u + 1);
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
      if (k & half) continue;
      const u32 idx = (1u << u) + (k >> (C::A - u));
      fwd_bfly<true, NARROW, LOGN>(u, x[k], x[k + half], tw[idx], tws[idx], q, q2, msub);
    }
  }
  constexpr int EB = 1 << C::B;
  u64* L = lds + ll * C::LDS_LINE;
#pragma unroll
  for (int k = 0; k < C::E; ++k) L[r0 + C::TPL * k] = x[k];
  __syncthreads();
#pragma unroll
  for (int g = 0; g < C::G; ++g) {
    const u32 r1 = r0 * C::G + g;
#pragma unroll
    for (int k = 0; k < EB; ++k) x[g * EB + k] = L[(r1 << C::B) + k];
#pragma unroll
    for (int v = 0; v < C::B; ++v) {
      const int u = C::A + v;
      const int half = EB >> (v + 1);
#pragma unroll
      for (int k = 0; k < EB; ++k) {
        if (k & half) continue;
        const u32 idx = (1u << u) + (r1 << v) + (k >> (C::B - v));
        fwd_bfly<WSB, NARROW, LOGN>(u, x[g * EB + k], x[g * EB + k + half], tw[idx], tws[idx], q, q2, msub);
      }
    }
#pragma unroll
    for (int k = 0; k < EB; ++k) stg(data, (((r1 << C::B) + k) << C::L2) | col, x[g * EB + k]);
  }
}

// Column pass at N = 2^16 with 8 elements per thread (32 threads per
// 256-point column, 16 columns per 512-thread workgroup): radix-8 register
// phases over stages 0-2, 3-5 and 6-7 with two LDS exchanges, twice the
// waves of the 16-element pass for launches too small to fill the GPU.
// Twiddle of stage s for the butterfly whose lower element is i:
// psi[2^s + (i >> (8 - s))], as in the 16-element pass.  Plain transforms of
// at most FIDES_COL8 limbs use it: 32 limbs 25.7-25.8 instead of
// 26.2-27.3 us, bit-identical; for every plain launch (960 limbs, the
// key-switch step's) it is no faster.
#ifndef FIDES_COL8
#define FIDES_COL8 32
#endif
template <int NARROW>
__global__ __launch_bounds__(512) void k_ntt_fwd_col8(LimbSpan span, NttTables T, NttFuse F) {
  constexpr int LOGN = 16, W = 16, LINE = 257;
  __shared__ u64 lds[W * LINE];
  const NttTask tk = resolve_task(span, F, blockIdx.y, LOGN);
  u64* data = tk.data;
  const u32 pid = tk.pid;
  const u64 q = T.prm[pid].q, q2 = 2 * q, msub = (u64)FwdLazy<NARROW, LOGN>::SUB * q;
  const u64* tw = T.psi + ((u64)pid << LOGN);
  const u64* tws = T.psi_sh + ((u64)pid << LOGN);
  const u32 tid = threadIdx.x;
  const u32 ll = tid % W, t = tid / W;  // column in the workgroup, thread in the column (0..31)
  const u32 col = blockIdx.x * W + ll;
  const u64* in = tk.src ? tk.src : data;
  u64* L = lds + ll * LINE;
  u64 x[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) x[k] = ldg(in, ((t + 32 * k) << 8) | col);
  // stages 0-2: elements t + 32k, pairs (k, k + half); wave-uniform twiddles
#pragma unroll
  for (int st = 0; st < 3; ++st) {
    const int half = 8 >> (st + 1);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if (k & half) continue;
      const u32 idx = (1u << st) + (k >> (3 - st));
      fwd_bfly<true, NARROW, LOGN>(st, x[k], x[k + half], tw[idx], tws[idx], q, q2, msub);
    }
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) L[t + 32 * k] = x[k];
  __syncthreads();
  // stages 3-5: elements 32c + b0 + 4j (c = t / 4, b0 = t % 4; a wave holds
  // t = 4w .. 4w + 3, so c and the twiddles are wave-uniform)
  const u32 c = __builtin_amdgcn_readfirstlane(t >> 2), b0 = t & 3;
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = L[32 * c + b0 + 4 * j];
#pragma unroll
  for (int v = 0; v < 3; ++v) {
    const int st = 3 + v, half = 8 >> (v + 1);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (j & half) continue;
      const u32 idx = (1u << st) + (c << v) + (j >> (3 - v));
      fwd_bfly<true, NARROW, LOGN>(st, x[j], x[j + half], tw[idx], tws[idx], q, q2, msub);
    }
  }
  __syncthreads();  // every thread has read its phase-2 elements
#pragma unroll
  for (int j = 0; j < 8; ++j) L[32 * c + b0 + 4 * j] = x[j];
  __syncthreads();
  // stages 6-7: two groups of 4 consecutive elements, 4g + m (g = 2t, 2t + 1)
#pragma unroll
  for (int e = 0; e < 8; ++e) x[e] = L[8 * t + e];
#pragma unroll
  for (int gi = 0; gi < 2; ++gi) {
    const u32 g = 2 * t + gi;
    fwd_bfly<false, NARROW, LOGN>(6, x[4 * gi], x[4 * gi + 2], tw[64 + g], tws[64 + g], q, q2, msub);
    fwd_bfly<false, NARROW, LOGN>(6, x[4 * gi + 1], x[4 * gi + 3], tw[64 + g], tws[64 + g], q, q2, msub);
    fwd_bfly<false, NARROW, LOGN>(7, x[4 * gi], x[4 * gi + 1], tw[128 + 2 * g], tws[128 + 2 * g], q, q2, msub);
    fwd_bfly<false, NARROW, LOGN>(7, x[4 * gi + 2], x[4 * gi + 3], tw[129 + 2 * g], tws[129 + 2 * g], q, q2, msub);
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) stg(data, ((8 * t + e) << 8) | col, x[e]);
}

// Row pass: global stages [L1, LOGN).  Input lazy as the column pass left
// it; output in [0, q).
// Occupancy target of the row pass at N = 2^16 per epilogue (waves per
// SIMD), set so the register allocation neither spills nor takes more than
// the kernel needs (left free, it grew to 4 waves where 7 fit).  Other ring
// dimensions keep the compiler's choice (1 = no constraint).
template <int LOGN, int EPI>
struct RowWaves {
  static constexpr int value = LOGN != 16 ? 1 : EPI == EPI_MODDOWN_ADD ? 5 : EPI == EPI_MDRS ? 6 : 7;
};
template <int LOGN, int EPI, int NARROW, int PF>
__device__ __forceinline__ void fwd_row_body(const LimbSpan& span, const NttTables& T, const NttFuse& F) {
  using C = NttCfg<LOGN, 1>;
  __shared__ u32 lds32[C::LPW * C::LDS_LINE];
  const NttTask tk = resolve_task(span, F, blockIdx.y, LOGN);
  u64* data = tk.data;
  const u32 pid = tk.pid;
  const u64 q = T.prm[pid].q, q2 = 2 * q, msub = (u64)FwdLazy<NARROW, LOGN>::SUB * q;
  const u64* tw = T.psi + ((u64)pid << LOGN);
  const u64* tws = T.psi_sh + ((u64)pid << LOGN);
  const u32 tid = threadIdx.x;
  const u32 r0 = tid % C::TPL, ll = tid / C::TPL;
  const u32 row = blockIdx.x * C::LPW + ll;
  u64* rowp = data + ((u64)row << C::L2);
  u64 x[C::E];
#pragma unroll
  for (int k = 0; k < C::E; ++k) x[k] = ldg(rowp, r0 + C::TPL * k);
  constexpr int EB = 1 << C::B;
  // PF: every twiddle of both phases (2^u per stage u) is loaded with the
  // row, so no stage waits for a twiddle load
  constexpr int NPA = PF ? C::E - 1 : 1, NPB = PF ? C::G * (EB - 1) : 1;
  [[maybe_unused]] u64 pa[NPA], pas[NPA], pb[NPB], pbs[NPB];
  if constexpr (PF) {
#pragma unroll
    for (int u = 0; u < C::A; ++u)
#pragma unroll
      for (int j = 0; j < (1 << u); ++j) {
        const u32 idx = (1u << (C::S0 + u)) + (row << u) + j;
        pa[(1 << u) - 1 + j] = tw[idx];
        pas[(1 << u) - 1 + j] = tws[idx];
      }
#pragma unroll
    for (int g = 0; g < C::G; ++g)
#pragma unroll
      for (int v = 0; v < C::B; ++v)
#pragma unroll
        for (int j = 0; j < (1 << v); ++j) {
          const int u = C::A + v;
          const u32 idx = (1u << (C::S0 + u)) + (row << u) + ((r0 * C::G + g) << v) + j;
          pb[g * (EB - 1) + (1 << v) - 1 + j] = tw[idx];
          pbs[g * (EB - 1) + (1 << v) - 1 + j] = tws[idx];
        }
  }
#pragma unroll
  for (int u = 0; u < C::A; ++u) {
    const int half = C::E >> (u + 1);
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
      if (k & half) continue;
      const u32 idx = (1u << (C::S0 + u)) + (row << u) + (k >> (C::A - u));
      [[maybe_unused]] const int pi = (1 << u) - 1 + (k >> (C::A - u));
      fwd_bfly<false, NARROW, LOGN>(C::S0 + u, x[k], x[k + half], PF ? pa[PF ? pi : 0] : tw[idx],
                                    PF ? pas[PF ? pi : 0] : tws[idx], q, q2, msub);
    }
  }
  u32* L = lds32 + ll * C::LDS_LINE;
  {  // strided -> groups of 2^B consecutive, low words then high words
    u32 lo[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) L[C::lx(r0 + C::TPL * k)] = (u32)x[k];
    __syncthreads();
#pragma unroll
    for (int g = 0; g < C::G; ++g)
#pragma unroll
      for (int k = 0; k < EB; ++k) lo[g * EB + k] = L[C::lx(((r0 * C::G + g) << C::B) + k)];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) L[C::lx(r0 + C::TPL * k)] = (u32)(x[k] >> 32);
    __syncthreads();
#pragma unroll
    for (int g = 0; g < C::G; ++g)
#pragma unroll
      for (int k = 0; k < EB; ++k)
        x[g * EB + k] = ((u64)L[C::lx(((r0 * C::G + g) << C::B) + k)] << 32) | lo[g * EB + k];
  }
#pragma unroll
  for (int g = 0; g < C::G; ++g) {
    const u32 r1 = r0 * C::G + g;
#pragma unroll
    for (int v = 0; v < C::B; ++v) {
      const int u = C::A + v;
      const int half = EB >> (v + 1);
#pragma unroll
      for (int k = 0; k < EB; ++k) {
        if (k & half) continue;
        const u32 idx = (1u << (C::S0 + u)) + (row << u) + (r1 << v) + (k >> (C::B - v));
        [[maybe_unused]] const int pi = g * (EB - 1) + (1 << v) - 1 + (k >> (C::B - v));
        fwd_bfly<false, NARROW, LOGN>(C::S0 + u, x[g * EB + k], x[g * EB + k + half], PF ? pb[PF ? pi : 0] : tw[idx],
                                      PF ? pbs[PF ? pi : 0] : tws[idx], q, q2, msub);
      }
    }
  }
  {  // groups -> strided
    u32 lo[C::E];
    __syncthreads();
#pragma unroll
    for (int g = 0; g < C::G; ++g)
#pragma unroll
      for (int k = 0; k < EB; ++k) L[C::lx(((r0 * C::G + g) << C::B) + k)] = (u32)x[g * EB + k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) lo[k] = L[C::lx(r0 + C::TPL * k)];
    __syncthreads();
#pragma unroll
    for (int g = 0; g < C::G; ++g)
#pragma unroll
      for (int k = 0; k < EB; ++k) L[C::lx(((r0 * C::G + g) << C::B) + k)] = (u32)(x[g * EB + k] >> 32);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) x[k] = ((u64)L[C::lx(r0 + C::TPL * k)] << 32) | lo[k];
  }
  u64* base = (EPI != EPI_NONE && tk.out) ? tk.out : data;
  constexpr int EPC = C::E < 8 ? C::E : 8;  // epilogue operands loaded per batch
  // the epilogues take the transform's lazy output (primes below 2^60:
  // batched key-switch step 0.15% faster, bit-identical)
  constexpr bool kLazyEpi = NARROW && EPI != EPI_NONE;
  if (EPI != EPI_NONE && tk.perm) {
    // store through the automorphism X -> X^k as a scatter with k^-1:
    // out[brev(((2 brev(j) + 1) k^-1 mod 2N) >> 1)] = v_j (reference
    // automorph_slot, src/Rotation.cuh:14-26)
    const u32 mask = (2u << LOGN) - 1;
#pragma unroll
    for (int k0 = 0; k0 < C::E; k0 += EPC) {
      u64 ea[EPC], eb[EPC];
#pragma unroll
      for (int k = 0; k < EPC; ++k)
        epilogue_load<EPI>((row << C::L2) | (r0 + C::TPL * (k0 + k)), tk, ea[k], eb[k]);
#pragma unroll
      for (int k = 0; k < EPC; ++k) {
        const u32 c = r0 + C::TPL * (k0 + k);
        const u32 j = (row << C::L2) | c;
        u64 v = kLazyEpi ? x[k0 + k] : fwd_final(x[k0 + k], q, q2);
        v = epilogue_apply<EPI, kLazyEpi>(v, ea[k], eb[k], tk, q);
        const u32 jr = __brev(j) >> (32 - LOGN);
        const u32 s = (((2u * jr + 1u) * tk.perm) & mask) >> 1;
        stg(base, __brev(s) >> (32 - LOGN), v);
      }
    }
    return;
  }
  u64* dst = base + ((u64)row << C::L2);
#pragma unroll
  for (int k0 = 0; k0 < C::E; k0 += EPC) {
    u64 ea[EPC], eb[EPC];
#pragma unroll
    for (int k = 0; k < EPC; ++k) epilogue_load<EPI>((row << C::L2) | (r0 + C::TPL * (k0 + k)), tk, ea[k], eb[k]);
#pragma unroll
    for (int k = 0; k < EPC; ++k) {
      const u32 c = r0 + C::TPL * (k0 + k);
      u64 v = kLazyEpi ? x[k0 + k] : fwd_final(x[k0 + k], q, q2);
      v = epilogue_apply<EPI, kLazyEpi>(v, ea[k], eb[k], tk, q);
      stg(dst, c, v);
    }
  }
}

template <int LOGN, int EPI, int NARROW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RowWaves<LOGN, EPI>::value))) void k_ntt_fwd_row(LimbSpan span, NttTables T, NttFuse F) {
  fwd_row_body<LOGN, EPI, NARROW, 0>(span, T, F);
}
// 2 waves per SIMD: room for all 30 twiddle pairs of a thread
template <int LOGN, int NARROW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_ntt_fwd_row_small(LimbSpan span, NttTables T, NttFuse F) {
  fwd_row_body<LOGN, EPI_NONE, NARROW, 1>(span, T, F);
}

// Inverse row stages (global stages LOGN-1 .. L1) of one row held by its TPL
// threads: in, thread r0 holds the 2^B consecutive elements (r1 << B) + k of
// each of its G groups (r1 = r0 G + g); out, it holds elements r0 + TPL k.
// One LDS exchange through the row's line L.  [0, q) -> [0, 2q).
template <int LOGN>
__device__ __forceinline__ void intt_row_stages(u32 row, u32 r0, u64* L, const u64* tw, const u64* tws, u64 q,
                                                u64 q2, u64 (&x)[NttCfg<LOGN, 1>::E]) {
  using C = NttCfg<LOGN, 1>;
  constexpr int EB = 1 << C::B;
#pragma unroll
  for (int g = 0; g < C::G; ++g) {
    const u32 r1 = r0 * C::G + g;
#pragma unroll
    for (int v = C::B - 1; v >= 0; --v) {
      const int u = C::A + v;
      const int half = EB >> (v + 1);
#pragma unroll
      for (int k = 0; k < EB; ++k) {
        if (k & half) continue;
        const u32 idx = (1u << (C::S0 + u)) + (row << u) + (r1 << v) + (k >> (C::B - v));
        gs_bfly(x[g * EB + k], x[g * EB + k + half], tw[idx], tws[idx], q, q2);
      }
    }
#pragma unroll
    for (int k = 0; k < EB; ++k) L[C::lx((r1 << C::B) + k)] = x[g * EB + k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < C::E; ++k) x[k] = L[C::lx(r0 + C::TPL * k)];
#pragma unroll
  for (int u = C::A - 1; u >= 0; --u) {
    const int half = C::E >> (u + 1);
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
      if (k & half) continue;
      const u32 idx = (1u << (C::S0 + u)) + (row << u) + (k >> (C::A - u));
      gs_bfly(x[k], x[k + half], tw[idx], tws[idx], q, q2);
    }
  }
}

// Row pass fused with the key-switch inner product (replaces the row pass of
// the ModUp outputs and k_dot_batch).  One 64-thread workgroup owns 64/TPL
// rows of one output limb t: for every digit j < BETA except the one owning t
// it loads that digit's column-pass output rows, runs the row stages and keeps
// the fully reduced values in registers; then per coefficient
//   o0 = sum_j e_j b_j,  o1 = sum_j e_j a_j   (e_jown = the input limb itself)
// with the key in Montgomery form, one column accumulation and one reduction
// (the arithmetic of k_dot_batch, so results are identical).  The ModUp
// outputs never go back to HBM in evaluation form.  For a special limb
// (t >= l), whose two outputs ModDown transforms back first, the inverse row
// stages follow in the same workgroup: the outputs are stored after the row
// pass of their INTT.
// One digit of k_ntt_row_dot: the row stages of digit j's rows (or the input
// limb itself for the digit owning t), fully reduced, into `out`.
template <int LOGN, int NARROW>
__device__ __forceinline__ void row_dot_digit(const DotTask& tk, u32 j, u64 es, u64 rbase, u32 row, u32 r0,
                                              u64* L, const u64* tw, const u64* tws, u64 q, u64 q2,
                                              u64 (&out)[NttCfg<LOGN, 1>::E]) {
  using C = NttCfg<LOGN, 1>;
  const u64 msub = (u64)FwdLazy<NARROW, LOGN>::SUB * q;
  if (j == tk.jown) {  // uniform per workgroup
#pragma unroll
    for (int k = 0; k < C::E; ++k) out[k] = ldg(tk.d, rbase + r0 + C::TPL * k);
    return;
  }
  const u64* rowp = tk.e + j * es + rbase;
  u64 x[C::E];
#pragma unroll
  for (int k = 0; k < C::E; ++k) x[k] = ldg(rowp, r0 + C::TPL * k);
#pragma unroll
  for (int u = 0; u < C::A; ++u) {
    const int half = C::E >> (u + 1);
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
      if (k & half) continue;
      const u32 idx = (1u << (C::S0 + u)) + (row << u) + (k >> (C::A - u));
      fwd_bfly<false, NARROW, LOGN>(C::S0 + u, x[k], x[k + half], tw[idx], tws[idx], q, q2, msub);
    }
  }
  __syncthreads();  // earlier readers of this workgroup's LDS are done
#pragma unroll
  for (int k = 0; k < C::E; ++k) L[C::lx(r0 + C::TPL * k)] = x[k];
  __syncthreads();
  constexpr int EB = 1 << C::B;
#pragma unroll
  for (int g = 0; g < C::G; ++g) {
    const u32 r1 = r0 * C::G + g;
#pragma unroll
    for (int k = 0; k < EB; ++k) x[g * EB + k] = L[C::lx((r1 << C::B) + k)];
#pragma unroll
    for (int v = 0; v < C::B; ++v) {
      const int u = C::A + v;
      const int half = EB >> (v + 1);
#pragma unroll
      for (int k = 0; k < EB; ++k) {
        if (k & half) continue;
        const u32 idx = (1u << (C::S0 + u)) + (row << u) + (r1 << v) + (k >> (C::B - v));
        fwd_bfly<false, NARROW, LOGN>(C::S0 + u, x[g * EB + k], x[g * EB + k + half], tw[idx], tws[idx], q, q2,
                                      msub);
      }
    }
#pragma unroll
    for (int k = 0; k < EB; ++k) L[C::lx((r1 << C::B) + k)] = x[g * EB + k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < C::E; ++k) out[k] = fwd_final(L[C::lx(r0 + C::TPL * k)], q, q2);
}

// Occupancy target of k_ntt_row_dot: at N = 2^16 with up to 3 digits, 3 waves
// per SIMD (168 VGPRs, 4-6 of them spilled to scratch) instead of the 2 the
// register allocator picks on its own (175-178 VGPRs): the batched key-switch
// step runs 4-5% faster, bit-identical.  A cap of 4 waves spills 27-70 VGPRs
// and is slower.  Other shapes keep the compiler's choice (1 = no target).
template <int LOGN, int BETA>
struct RowDotWaves {
  static constexpr int value = (LOGN == 16 && BETA <= 3) ? 3 : 1;
};
template <int LOGN, int BETA, int NARROW>
__device__ __forceinline__ void row_dot_task(const DotTask* __restrict__ tasks, const NttTables& T, u32 kstride,
                                             u64* lds) {
  using C = NttCfg<LOGN, 1>;
  constexpr int RPW = 64 / C::TPL;  // rows per workgroup
  const DotTask tk = tasks[blockIdx.y];
  const u32 pid = tk.pid;
  const PrimeC P = T.prm[pid];
  const u64 q = P.q, q2 = 2 * q;
  const u64* tw = T.psi + ((u64)pid << LOGN);
  const u64* tws = T.psi_sh + ((u64)pid << LOGN);
  const u32 tid = threadIdx.x;
  const u32 r0 = tid % C::TPL, ll = tid / C::TPL;
  const u32 row = blockIdx.x * RPW + ll;
  const u64 es = (u64)tk.estride << LOGN, ks = (u64)kstride << LOGN;
  const u64 rbase = (u64)row << C::L2;
  u64* L = lds + ll * C::LDS_LINE;
  u64 e0[C::E], e1[C::E], e2[C::E], e3[C::E];
  row_dot_digit<LOGN, NARROW>(tk, 0, es, rbase, row, r0, L, tw, tws, q, q2, e0);
  if constexpr (BETA > 1) row_dot_digit<LOGN, NARROW>(tk, 1, es, rbase, row, r0, L, tw, tws, q, q2, e1);
  if constexpr (BETA > 2) row_dot_digit<LOGN, NARROW>(tk, 2, es, rbase, row, r0, L, tw, tws, q, q2, e2);
  if constexpr (BETA > 3) row_dot_digit<LOGN, NARROW>(tk, 3, es, rbase, row, r0, L, tw, tws, q, q2, e3);
#pragma unroll
  for (int k = 0; k < C::E; ++k) {
    const u64 i = rbase + r0 + C::TPL * k;
    Acc128 a0, a1;
    acc_mac_v(a0, e0[k], ldg(tk.kb, i));
    acc_mac_v(a1, e0[k], ldg(tk.ka, i));
    if constexpr (BETA > 1) {
      acc_mac_v(a0, e1[k], ldg(tk.kb, ks + i));
      acc_mac_v(a1, e1[k], ldg(tk.ka, ks + i));
    }
    if constexpr (BETA > 2) {
      acc_mac_v(a0, e2[k], ldg(tk.kb, 2 * ks + i));
      acc_mac_v(a1, e2[k], ldg(tk.ka, 2 * ks + i));
    }
    if constexpr (BETA > 3) {
      acc_mac_v(a0, e3[k], ldg(tk.kb, 3 * ks + i));
      acc_mac_v(a1, e3[k], ldg(tk.ka, 3 * ks + i));
    }
    u64 h, l;
    acc_value(a0, h, l);
    e0[k] = mont_redc(h, l, P);
    acc_value(a1, h, l);
    e1[k] = mont_redc(h, l, P);
  }
  if (tk.jown != 0xffffffffu || tk.ext) {  // uniform per workgroup
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
      const u64 i = rbase + r0 + C::TPL * k;
      stg(tk.o0, i, e0[k]);
      stg(tk.o1, i, e1[k]);
    }
    return;
  }
  // Special limb (uniform per workgroup): ModDown starts with its INTT, so
  // the inverse row stages run here and only the column pass is left.
  const u64* itw = T.ipsi + ((u64)pid << LOGN);
  const u64* itws = T.ipsi_sh + ((u64)pid << LOGN);
  constexpr int EB = 1 << C::B;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    u64(&y)[C::E] = c == 0 ? e0 : e1;
    u64* o = c == 0 ? tk.o0 : tk.o1;
    __syncthreads();  // earlier readers of this workgroup's LDS are done
#pragma unroll
    for (int k = 0; k < C::E; ++k) L[C::lx(r0 + C::TPL * k)] = y[k];
    __syncthreads();
#pragma unroll
    for (int g = 0; g < C::G; ++g)
#pragma unroll
      for (int k = 0; k < EB; ++k) y[g * EB + k] = L[C::lx(((r0 * C::G + g) << C::B) + k)];
    intt_row_stages<LOGN>(row, r0, L, itw, itws, q, q2, y);
#pragma unroll
    for (int k = 0; k < C::E; ++k) stg(o, rbase + r0 + C::TPL * k, y[k]);
  }
}
template <int LOGN, int BETA, int NARROW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RowDotWaves<LOGN, BETA>::value))) void k_ntt_row_dot(const DotTask* __restrict__ tasks, NttTables T, u32 kstride) {
  using C = NttCfg<LOGN, 1>;
  __shared__ u64 lds[(64 / C::TPL) * C::LDS_LINE];
  row_dot_task<LOGN, BETA, NARROW>(tasks, T, kstride, lds);
}
// Tasks of different digit counts (1..MAXB) in one launch, each workgroup
// branching once, uniformly, to the body for its task's count: a batch at
// mixed levels has output limbs with 1, 2 and 3 active digits, and one launch
// (heaviest tasks first) drains better than one per count.  At N = 2^16 with
// MAXB = 3 it holds the 3-wave target with 18 VGPRs spilled (4-6 in the
// per-count kernels); the batched key-switch step still runs 0.7-0.8% faster
// than with three launches (2 waves, no spill: 5% slower).
template <int LOGN, int MAXB, int NARROW>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(RowDotWaves<LOGN, MAXB>::value))) void k_ntt_row_dot_mixed(const DotTask* __restrict__ tasks, NttTables T, u32 kstride) {
  using C = NttCfg<LOGN, 1>;
  __shared__ u64 lds[(64 / C::TPL) * C::LDS_LINE];
  switch (tasks[blockIdx.y].beta) {
    case 1: row_dot_task<LOGN, 1, NARROW>(tasks, T, kstride, lds); break;
    case 2: if constexpr (MAXB >= 2) row_dot_task<LOGN, 2, NARROW>(tasks, T, kstride, lds); break;
    case 3: if constexpr (MAXB >= 3) row_dot_task<LOGN, 3, NARROW>(tasks, T, kstride, lds); break;
    case 4: if constexpr (MAXB >= 4) row_dot_task<LOGN, 4, NARROW>(tasks, T, kstride, lds); break;
    default: break;
  }
}

// ------------------------------------------------------------------ inverse
// Row pass: global stages LOGN-1 .. L1 (descending).  [0, q) -> [0, 2q).
// AXPY (task lists only): the input is src + x c0 mod q, element-wise.
template <int LOGN, int AXPY = 0>
__global__ __launch_bounds__(256) void k_intt_row(LimbSpan span, NttTables T) {
  using C = NttCfg<LOGN, 1>;
  __shared__ u64 lds[C::LPW * C::LDS_LINE];
  u64* data;
  u32 pid;
  const u64* src;
  if (span.list) {
    const NttTask& t = span.list[blockIdx.y];
    data = t.data;
    pid = t.pid;
    src = t.src ? t.src : t.data;
  } else {
    span_get(span, blockIdx.y, LOGN, data, pid);
    src = data;
  }
  const u64 q = T.prm[pid].q, q2 = 2 * q;
  const u64* tw = T.ipsi + ((u64)pid << LOGN);
  const u64* tws = T.ipsi_sh + ((u64)pid << LOGN);
  const u32 tid = threadIdx.x;
  const u32 r0 = tid % C::TPL, ll = tid / C::TPL;
  const u32 row = blockIdx.x * C::LPW + ll;
  u64* rowp = data + ((u64)row << C::L2);
  const u64* srow = src + ((u64)row << C::L2);
  u64* L = lds + ll * C::LDS_LINE;
  u64 x[C::E];
  constexpr int EB = 1 << C::B;
#pragma unroll
  for (int g = 0; g < C::G; ++g) {
    const u32 r1 = r0 * C::G + g;
    // direct load of the thread's 2^B consecutive elements (no LDS staging;
    // measured faster than a coalesced load through LDS)
#pragma unroll
    for (int k = 0; k < EB; ++k) x[g * EB + k] = ldg(srow, (r1 << C::B) + k);
    if constexpr (AXPY) {
      const NttTask& t = span.list[blockIdx.y];
      const u64* xrow = t.x + ((u64)row << C::L2);
      const u64 c0 = t.c0, c0sh = t.c0sh;
#pragma unroll
      for (int k = 0; k < EB; ++k)
        x[g * EB + k] = add_mod(x[g * EB + k], shoup(ldg(xrow, (r1 << C::B) + k), c0, c0sh, q), q);
    }
  }
  intt_row_stages<LOGN>(row, r0, L, tw, tws, q, q2, x);
#pragma unroll
  for (int k = 0; k < C::E; ++k) stg(rowp, r0 + C::TPL * k, x[k]);
}

// Column pass: global stages L1-1 .. 0; N^-1 folded into stage 0.  [0, 2q) -> [0, q).
// AXPY (task lists only): the output is the transform + src2 c1 mod q.
template <int LOGN, int AXPY = 0>
__global__ __launch_bounds__((NttCfg<LOGN, 0, kInvColW>::THREADS)) void k_intt_col(LimbSpan span, NttTables T) {
  using C = NttCfg<LOGN, 0, kInvColW>;
  // The exchange moves 32-bit halves (low words, then high words) through a
  // line of M + 4 words per column: half the LDS of a 64-bit line, so LDS no
  // longer caps the pass at 4 waves per SIMD; the stride keeps the strided
  // writes of a wave on distinct banks.
  constexpr int S32 = C::M + 4;
  __shared__ u32 lds32[C::LPW * S32];
  u64* data;
  u32 pid;
  if (span.list) {
    data = span.list[blockIdx.y].data;
    pid = span.list[blockIdx.y].pid;
  } else {
    span_get(span, blockIdx.y, LOGN, data, pid);
  }
  const PrimeC P = T.prm[pid];
  const u64 q = P.q, q2 = 2 * q;
  const u64* tw = T.ipsi + ((u64)pid << LOGN);
  const u64* tws = T.ipsi_sh + ((u64)pid << LOGN);
  const u32 tid = threadIdx.x;
  constexpr bool WSB = C::LPW >= 64;  // one r0 per wave: uniform twiddles in the first phase
  const u32 ll = tid % C::LPW, r0 = WSB ? __builtin_amdgcn_readfirstlane(tid / C::LPW) : tid / C::LPW;
  const u32 col = blockIdx.x * C::LPW + ll;
  u32* L = lds32 + ll * S32;
  u64 x[C::E];
  constexpr int EB = 1 << C::B;
#pragma unroll
  for (int g = 0; g < C::G; ++g) {
    const u32 r1 = r0 * C::G + g;
#pragma unroll
    for (int k = 0; k < EB; ++k) x[g * EB + k] = ldg(data, (((r1 << C::B) + k) << C::L2) | col);
#pragma unroll
    for (int v = C::B - 1; v >= 0; --v) {
      const int u = C::A + v;
      const int half = EB >> (v + 1);
#pragma unroll
      for (int k = 0; k < EB; ++k) {
        if (k & half) continue;
        const u32 idx = (1u << u) + (r1 << v) + (k >> (C::B - v));
        gs_bfly<WSB>(x[g * EB + k], x[g * EB + k + half], tw[idx], tws[idx], q, q2);
      }
    }
  }
  {
    u32 lo[C::E];
#pragma unroll
    for (int g = 0; g < C::G; ++g)
#pragma unroll
      for (int k = 0; k < EB; ++k) L[((r0 * C::G + g) << C::B) + k] = (u32)x[g * EB + k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) lo[k] = L[r0 + C::TPL * k];
    __syncthreads();
#pragma unroll
    for (int g = 0; g < C::G; ++g)
#pragma unroll
      for (int k = 0; k < EB; ++k) L[((r0 * C::G + g) << C::B) + k] = (u32)(x[g * EB + k] >> 32);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) x[k] = ((u64)L[r0 + C::TPL * k] << 32) | lo[k];
  }
#pragma unroll
  for (int u = C::A - 1; u >= 1; --u) {
    const int half = C::E >> (u + 1);
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
      if (k & half) continue;
      const u32 idx = (1u << u) + (k >> (C::A - u));
      gs_bfly<true>(x[k], x[k + half], tw[idx], tws[idx], q, q2);
    }
  }
  {  // stage 0 with N^-1: X' = (X+Y) N^-1, Y' = (X-Y) psi^-brev(1) N^-1
    constexpr int half = C::E >> 1;
#pragma unroll
    for (int k = 0; k < half; ++k) {
      const u64 X = x[k], Y = x[k + half];
      x[k] = shoup(X + Y, P.ninv, P.ninv_sh, q);
      x[k + half] = shoup(X + q2 - Y, P.ninv_w, P.ninv_w_sh, q);
    }
  }
  if constexpr (AXPY) {
    const NttTask& t = span.list[blockIdx.y];
    const u64 c1 = t.c1, c1sh = t.c1sh;
#pragma unroll
    for (int k = 0; k < C::E; ++k)
      x[k] = add_mod(x[k], shoup(ldg(t.src2, ((r0 + C::TPL * k) << C::L2) | col), c1, c1sh, q), q);
  }
#pragma unroll
  for (int k = 0; k < C::E; ++k) stg(data, ((r0 + C::TPL * k) << C::L2) | col, x[k]);
}

// ------------------------------------------------------------------ launchers
// Plain forward row passes of at most this many limbs at N = 2^16 (at most
// two waves per SIMD, all resident at once, nothing to hide a twiddle load
// behind) run the variant that loads all its twiddles with the row: 14.4
// instead of 15.0 us at 32 limbs, bit-identical.  (The same for the column
// pass's phase-B twiddles measured no gain.)
#ifndef FIDES_SMALL_NTT
#define FIDES_SMALL_NTT 32
#endif
constexpr int kSmallNttLimbs = FIDES_SMALL_NTT;
template <int LOGN, int NARROW>
static void launch_col_n(const LimbSpan& s, const NttTables& T, int pro, const NttFuse& F, hipStream_t st) {
  using C0 = NttCfg<LOGN, 0, kFwdColW>;
  dim3 g0(C0::NLINES / C0::LPW, s.total);
  if constexpr (LOGN == 16 && FIDES_COL8 > 0) {
    if (pro == PRO_NONE && s.total <= FIDES_COL8) {
      k_ntt_fwd_col8<NARROW><<<dim3(256 / 16, s.total), 512, 0, st>>>(s, T, F);
      return;
    }
  }
  if (pro == PRO_SWITCH_MODULUS)
    k_ntt_fwd_col<LOGN, PRO_SWITCH_MODULUS, NARROW><<<g0, C0::THREADS, 0, st>>>(s, T, F);
  else if (pro == PRO_MDRS)
    k_ntt_fwd_col<LOGN, PRO_MDRS, NARROW><<<g0, C0::THREADS, 0, st>>>(s, T, F);
  else
    k_ntt_fwd_col<LOGN, PRO_NONE, NARROW><<<g0, C0::THREADS, 0, st>>>(s, T, F);
}

template <int LOGN, int NARROW>
static void launch_row_n(const LimbSpan& s, const NttTables& T, int epi, const NttFuse& F, hipStream_t st) {
  using C1 = NttCfg<LOGN, 1>;
  dim3 g1(C1::NLINES / C1::LPW, s.total);
  if constexpr (LOGN == 16 && kSmallNttLimbs > 0) {
    if (epi == EPI_NONE && s.total <= kSmallNttLimbs) {
      k_ntt_fwd_row_small<LOGN, NARROW><<<g1, C1::THREADS, 0, st>>>(s, T, F);
      return;
    }
  }
  switch (epi) {
    case EPI_MODDOWN: k_ntt_fwd_row<LOGN, EPI_MODDOWN, NARROW><<<g1, C1::THREADS, 0, st>>>(s, T, F); break;
    case EPI_MODDOWN_ADD:
      k_ntt_fwd_row<LOGN, EPI_MODDOWN_ADD, NARROW><<<g1, C1::THREADS, 0, st>>>(s, T, F);
      break;
    case EPI_RESCALE: k_ntt_fwd_row<LOGN, EPI_RESCALE, NARROW><<<g1, C1::THREADS, 0, st>>>(s, T, F); break;
    case EPI_MDRS: k_ntt_fwd_row<LOGN, EPI_MDRS, NARROW><<<g1, C1::THREADS, 0, st>>>(s, T, F); break;
    default: k_ntt_fwd_row<LOGN, EPI_NONE, NARROW><<<g1, C1::THREADS, 0, st>>>(s, T, F); break;
  }
}

// The column pass leaves its output lazy (below bnd(L1) q of the schedule
// it ran); the row pass that follows is launched with the same schedule.
template <int LOGN>
static void launch_col(const LimbSpan& s, const NttTables& T, int pro, const NttFuse& F, hipStream_t st) {
  if (T.narrow)
    launch_col_n<LOGN, 1>(s, T, pro, F, st);
  else
    launch_col_n<LOGN, 0>(s, T, pro, F, st);
}

template <int LOGN>
static void launch_fwd(const LimbSpan& s, const NttTables& T, int pro, int epi, const NttFuse& F, hipStream_t st) {
  launch_col<LOGN>(s, T, pro, F, st);
  if (T.narrow)
    launch_row_n<LOGN, 1>(s, T, epi, F, st);
  else
    launch_row_n<LOGN, 0>(s, T, epi, F, st);
}

template <int LOGN, int AXPY = 0>
static void launch_inv(const LimbSpan& s, const NttTables& T, hipStream_t st) {
  using C0 = NttCfg<LOGN, 0, kInvColW>;
  using C1 = NttCfg<LOGN, 1>;
  dim3 g0(C0::NLINES / C0::LPW, s.total), g1(C1::NLINES / C1::LPW, s.total);
  k_intt_row<LOGN, AXPY><<<g1, C1::THREADS, 0, st>>>(s, T);
  k_intt_col<LOGN, AXPY><<<g0, C0::THREADS, 0, st>>>(s, T);
}


template <int LOGN>
static void launch_row_dot(const DotTask* tasks, u32 n, int beta, const NttTables& T, u32 kstride, hipStream_t st) {
  using C1 = NttCfg<LOGN, 1>;
  dim3 g(C1::NLINES / (64 / C1::TPL), n);
#define ROWDOT(B)                                                 \
  if (T.narrow)                                                   \
    k_ntt_row_dot<LOGN, B, 1><<<g, 64, 0, st>>>(tasks, T, kstride); \
  else                                                            \
    k_ntt_row_dot<LOGN, B, 0><<<g, 64, 0, st>>>(tasks, T, kstride);
  switch (beta) {
    case 1: ROWDOT(1) break;
    case 2: ROWDOT(2) break;
    case 3: ROWDOT(3) break;
    default: ROWDOT(4) break;
  }
#undef ROWDOT
}
template <int LOGN>
static void launch_row_dot_mixed(const DotTask* tasks, u32 n, int maxb, const NttTables& T, u32 kstride,
                                 hipStream_t st) {
  using C1 = NttCfg<LOGN, 1>;
  dim3 g(C1::NLINES / (64 / C1::TPL), n);
#define ROWDOT(B)                                                       \
  if (T.narrow)                                                         \
    k_ntt_row_dot_mixed<LOGN, B, 1><<<g, 64, 0, st>>>(tasks, T, kstride); \
  else                                                                  \
    k_ntt_row_dot_mixed<LOGN, B, 0><<<g, 64, 0, st>>>(tasks, T, kstride);
  if (maxb <= 3) {
    ROWDOT(3)
  } else {
    ROWDOT(4)
  }
#undef ROWDOT
}

// Host entry points (called by the context code).  Return 0 or -1 for an
// unsupported ring dimension.
int fides_launch_ntt_col(int logN, const LimbSpan& s, const NttTables& T, int pro, const NttFuse& F,
                         hipStream_t st) {
  if (s.total == 0) return 0;
  switch (logN) {
    case 10: launch_col<10>(s, T, pro, F, st); return 0;
    case 11: launch_col<11>(s, T, pro, F, st); return 0;
    case 12: launch_col<12>(s, T, pro, F, st); return 0;
    case 13: launch_col<13>(s, T, pro, F, st); return 0;
    case 14: launch_col<14>(s, T, pro, F, st); return 0;
    case 15: launch_col<15>(s, T, pro, F, st); return 0;
    case 16: launch_col<16>(s, T, pro, F, st); return 0;
    case 17: launch_col<17>(s, T, pro, F, st); return 0;
  }
  return -1;
}

int fides_launch_intt_col(int logN, const LimbSpan& s, const NttTables& T, hipStream_t st) {
  if (s.total == 0) return 0;
  switch (logN) {
#define CASE(n)                                                                                         \
  case n: {                                                                                             \
    using C0 = NttCfg<n, 0, kInvColW>;                                                                  \
    k_intt_col<n><<<dim3(C0::NLINES / C0::LPW, s.total), C0::THREADS, 0, st>>>(s, T);                  \
    return 0;                                                                                           \
  }
    CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16) CASE(17)
#undef CASE
  }
  return -1;
}

int fides_launch_row_dot(int logN, const DotTask* tasks, u32 n, int beta, const NttTables& T, u32 kstride,
                         hipStream_t st) {
  if (n == 0) return 0;
  if (beta < 1 || beta > ROWDOT_MAXB) return -1;
  switch (logN) {
    case 10: launch_row_dot<10>(tasks, n, beta, T, kstride, st); return 0;
    case 11: launch_row_dot<11>(tasks, n, beta, T, kstride, st); return 0;
    case 12: launch_row_dot<12>(tasks, n, beta, T, kstride, st); return 0;
    case 13: launch_row_dot<13>(tasks, n, beta, T, kstride, st); return 0;
    case 14: launch_row_dot<14>(tasks, n, beta, T, kstride, st); return 0;
    case 15: launch_row_dot<15>(tasks, n, beta, T, kstride, st); return 0;
    case 16: launch_row_dot<16>(tasks, n, beta, T, kstride, st); return 0;
    case 17: launch_row_dot<17>(tasks, n, beta, T, kstride, st); return 0;
  }
  return -1;
}
// Instantiated for the large rings only (where batches at mixed levels are
// heavy enough for the launch count to matter); returns 1 for the others,
// whose callers then launch once per digit count.
int fides_launch_row_dot_mixed(int logN, const DotTask* tasks, u32 n, int maxb, const NttTables& T, u32 kstride,
                               hipStream_t st) {
  if (n == 0) return 0;
  if (maxb < 2 || maxb > ROWDOT_MAXB) return -1;
  switch (logN) {
    case 16: launch_row_dot_mixed<16>(tasks, n, maxb, T, kstride, st); return 0;
    case 17: launch_row_dot_mixed<17>(tasks, n, maxb, T, kstride, st); return 0;
  }
  return 1;
}
int fides_launch_ntt(int logN, const LimbSpan& s, const NttTables& T, int pro, int epi, const NttFuse& F,
                     hipStream_t st) {
  if (s.total == 0) return 0;
  switch (logN) {
    case 10: launch_fwd<10>(s, T, pro, epi, F, st); return 0;
    case 11: launch_fwd<11>(s, T, pro, epi, F, st); return 0;
    case 12: launch_fwd<12>(s, T, pro, epi, F, st); return 0;
    case 13: launch_fwd<13>(s, T, pro, epi, F, st); return 0;
    case 14: launch_fwd<14>(s, T, pro, epi, F, st); return 0;
    case 15: launch_fwd<15>(s, T, pro, epi, F, st); return 0;
    case 16: launch_fwd<16>(s, T, pro, epi, F, st); return 0;
    case 17: launch_fwd<17>(s, T, pro, epi, F, st); return 0;
    default: return -1;
  }
}

int fides_launch_intt_axpy(int logN, const LimbSpan& s, const NttTables& T, hipStream_t st) {
  if (s.total == 0) return 0;
  if (!s.list) return -1;
  switch (logN) {
    case 10: launch_inv<10, 1>(s, T, st); return 0;
    case 11: launch_inv<11, 1>(s, T, st); return 0;
    case 12: launch_inv<12, 1>(s, T, st); return 0;
    case 13: launch_inv<13, 1>(s, T, st); return 0;
    case 14: launch_inv<14, 1>(s, T, st); return 0;
    case 15: launch_inv<15, 1>(s, T, st); return 0;
    case 16: launch_inv<16, 1>(s, T, st); return 0;
    case 17: launch_inv<17, 1>(s, T, st); return 0;
    default: return -1;
  }
}
int fides_launch_intt(int logN, const LimbSpan& s, const NttTables& T, hipStream_t st) {
  if (s.total == 0) return 0;
  switch (logN) {
    case 10: launch_inv<10>(s, T, st); return 0;
    case 11: launch_inv<11>(s, T, st); return 0;
    case 12: launch_inv<12>(s, T, st); return 0;
    case 13: launch_inv<13>(s, T, st); return 0;
    case 14: launch_inv<14>(s, T, st); return 0;
    case 15: launch_inv<15>(s, T, st); return 0;
    case 16: launch_inv<16>(s, T, st); return 0;
    case 17: launch_inv<17>(s, T, st); return 0;
    default: return -1;
  }
}
