// Element-wise and basis-conversion kernels of the hybrid key switch.
//
// Every kernel maps one thread to one coefficient position (coalesced
// 64-bit accesses along a limb) and the grid's y dimension to a task record,
// so one launch covers any mix of ciphertexts and levels.
#include "kernels.h"

// ---------------------------------------------------------------- tensor
// (a0,a1) x (b0,b1) -> (a0 b0, a0 b1 + a1 b0, a1 b1) per limb (reference
// src/CKKS/ElemenwiseBatchKernels.cu:1363-1400 binomialMult_; squaring
// variant binomialSquare_).  Outputs may alias inputs (each element is read
// before it is written by the same thread).
__global__ __launch_bounds__(256) void k_tensor_batch(const TensorTask* __restrict__ tasks,
                                                      const PrimeC* __restrict__ prm, u32 logN) {
  const TensorTask t = tasks[blockIdx.y];
  const u32 k = blockIdx.x * 256u + threadIdx.x;
  const PrimeC P = prm[t.pid];
  const u64 x0 = t.a0[k], x1 = t.a1[k];
  // Montgomery products: the second factors are taken to Montgomery form
  // (y R mod q), so REDC(x * yR) = x y mod q; the cross term (< 2 q^2 < q R)
  // is reduced once.  Every result is the canonical residue.
  const u64 y0 = t.square ? x0 : t.b0[k], y1 = t.square ? x1 : t.b1[k];
  const u64 m0 = to_mont(y0, P), m1 = to_mont(y1, P);
  u64 h, l;
  mul64wide(x0, m0, h, l);
  const u64 r0 = redc_once(h, l, P);
  mul64wide(x0, m1, h, l);
  mac128(h, l, x1, m0);
  const u64 r1 = redc_once(h, l, P);
  mul64wide(x1, m1, h, l);
  const u64 r2 = redc_once(h, l, P);
  t.o0[k] = r0;
  t.o1[k] = r1;
  t.o2[k] = r2;
}

// ---------------------------------------------------------------- basis conversion
// y_t = sum_i [x_i * hatinv_i]_{q_i} * hatmod_{i,t}  mod p_t
// (reference src/CKKS/Conv.cu:258-402 DecompAndModUpConv and :26-125
// ModDown2; OpenFHE ApproxSwitchCRTBasis).  hatmod is stored pre-multiplied
// by R = 2^64 so that one Montgomery reduction of the 128-bit accumulator
// yields the exact residue.  Coefficient form in and out.  The NS sources
// are loaded and pre-scaled once per thread and reused for every target of
// the task's chunk.
// Descriptors and prime constants are read-only for the whole launch and
// uniform across a workgroup: reading them through the constant address
// space lets them come in through scalar loads.
#define CONSTANT_AS __attribute__((address_space(4)))

// SPLIT (every prime below 2^60): sources and matrix entries are split into
// 30-bit halves, so each of the four partial-product columns is a sum of at
// most 16 products below 2^60 and cannot overflow 64 bits -- four
// v_mad_u64_u32 per term and no carry tracking.  Matrix entries then hold
// (m mod 2^30) | (m >> 30) << 32.  Otherwise the column accumulator with
// counted carries handles primes up to 2^62.
template <int NS, bool SPLIT>
__device__ __forceinline__ void conv_task(const CONSTANT_AS ConvTask* tp, const CONSTANT_AS ConvDesc* d,
                                          const PrimeC* __restrict__ prm_, u32 logN) {
  const CONSTANT_AS PrimeC* prm = (const CONSTANT_AS PrimeC*)prm_;
  const u64* __restrict__ in = tp->in;
  u64* __restrict__ out = tp->out;
  const u32 t0 = tp->t0, t1 = tp->t1;
  const u32 k = blockIdx.x * 256u + threadIdx.x;
  if constexpr (SPLIT) {
    u32 x0[NS], x1[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const u64 q = prm[d->src_pid[i]].q;
      const u64 v = shoup(in[((u64)i << logN) + k], d->hatinv[i], d->hatinv_sh[i], q);
      x0[i] = (u32)v & 0x3fffffffu;
      x1[i] = (u32)(v >> 30);
    }
    for (u32 t = t0; t < t1; ++t) {
      PrimeC P;
      P.q = prm[d->dst_pid[t]].q;
      P.qneg_inv = prm[d->dst_pid[t]].qneg_inv;
      u64 L = 0, M1 = 0, M2 = 0, H = 0;
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        const u64 m = d->mat[t][i];
        const u32 m0 = (u32)m, m1 = (u32)(m >> 32);
        L += (u64)x0[i] * m0;
        M1 += (u64)x0[i] * m1;
        M2 += (u64)x1[i] * m0;
        H += (u64)x1[i] * m1;
      }
      // T = L + (M1 + M2) 2^30 + H 2^60
      const unsigned __int128 T = (unsigned __int128)L + (((unsigned __int128)M1 + M2) << 30) +
                                  ((unsigned __int128)H << 60);
      // T < NS q^2 <= 16 q^2 < q 2^64 (q < 2^60): one conditional subtraction
      out[((u64)d->dst_off[t] << logN) + k] = redc_once((u64)(T >> 64), (u64)T, P);
    }
    return;
  }
  u64 xs[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const u64 q = prm[d->src_pid[i]].q;
    xs[i] = shoup(in[((u64)i << logN) + k], d->hatinv[i], d->hatinv_sh[i], q);
  }
  for (u32 t = t0; t < t1; ++t) {
    PrimeC P;
    P.q = prm[d->dst_pid[t]].q;
    P.qneg_inv = prm[d->dst_pid[t]].qneg_inv;
    P.one_sh = prm[d->dst_pid[t]].one_sh;
    Acc128 acc;
#pragma unroll
    for (int i = 0; i < NS; ++i) acc_mac(acc, xs[i], d->mat[t][i]);
    u64 hi, lo;
    acc_value(acc, hi, lo);
    // up to 16 products of values below 2^62: the high word can exceed 4q
    out[((u64)d->dst_off[t] << logN) + k] = mont_redc_any(hi, lo, P);
  }
}
template <int NS, bool SPLIT>
__global__ __launch_bounds__(256) void k_conv_batch(const ConvTask* __restrict__ tasks,
                                                    const PrimeC* __restrict__ prm, u32 logN) {
  const CONSTANT_AS ConvTask* tp = (const CONSTANT_AS ConvTask*)(tasks + blockIdx.y);
  conv_task<NS, SPLIT>(tp, (const CONSTANT_AS ConvDesc*)tp->desc, prm, logN);
}
// Tasks with different source counts (1..MAXNS) in one launch: each
// workgroup row branches once, uniformly, to the body for its task's count.
// The ModUp of a batch at mixed levels has a partial last digit of a
// different size per level; one launch for all of them replaces one launch
// per size, each too small to fill the GPU.
template <int MAXNS, bool SPLIT>
__global__ __launch_bounds__(256) void k_conv_batch_mixed(const ConvTask* __restrict__ tasks,
                                                          const PrimeC* __restrict__ prm, u32 logN) {
  const CONSTANT_AS ConvTask* tp = (const CONSTANT_AS ConvTask*)(tasks + blockIdx.y);
  const CONSTANT_AS ConvDesc* d = (const CONSTANT_AS ConvDesc*)tp->desc;
  switch (d->ns) {
#define CASE(n) \
  case n:       \
    if constexpr (n <= MAXNS) conv_task<n, SPLIT>(tp, d, prm, logN); \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
    CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: break;
  }
}

// ---------------------------------------------------------------- key-switch inner product
// (o0, o1)[t] = sum_{j<beta} e_j[t] * (b_j, a_j)[t], where e_j is the ModUp
// output of digit j and, for the digit that owns limb t, the original
// evaluation-form limb (reference src/CKKS/ElemenwiseBatchKernels.cu:202-273
// fusedDotKSK_2_, :246-255).  The key is stored in Montgomery form (times
// R), so one reduction of the 128-bit sums gives the exact result.
__global__ __launch_bounds__(256) void k_dot_batch(const DotTask* __restrict__ tasks, const PrimeC* __restrict__ prm,
                                                   u32 logN, u32 kstride) {
  const DotTask t = tasks[blockIdx.y];
  const u64 k = blockIdx.x * 256u + threadIdx.x;
  const PrimeC P = prm[t.pid];
  const u64 es = (u64)t.estride << logN, ks = (u64)kstride << logN;
  Acc128 a0, a1;
  for (u32 j = 0; j < t.beta; ++j) {
    const u64 e = (j == t.jown) ? t.d[k] : t.e[j * es + k];
    acc_mac_v(a0, e, t.kb[j * ks + k]);
    acc_mac_v(a1, e, t.ka[j * ks + k]);
  }
  u64 h, l;
  acc_value(a0, h, l);
  t.o0[k] = mont_redc_any(h, l, P);
  acc_value(a1, h, l);
  t.o1[k] = mont_redc_any(h, l, P);
}

// ---------------------------------------------------------------- key conversion
// x <- x * R mod q (Montgomery form) over [2][dnum][L+K][N].
__global__ __launch_bounds__(256) void k_to_mont(u64* key, const PrimeC* prm, u32 logN, u32 LK) {
  const u32 limb = blockIdx.y;  // over 2 * dnum * LK
  const PrimeC P = prm[limb % LK];
  const u64 i = ((u64)limb << logN) + blockIdx.x * 256u + threadIdx.x;
  u64 hi, lo;
  mul64wide(key[i], P.r2, hi, lo);
  key[i] = mont_redc(hi, lo, P);
}

// ---------------------------------------------------------------- automorphism
// out[brev(j)] = in[brev(((2j+1) k mod 2N) >> 1)], evaluation form
// (OpenFHE PrecomputeAutoMap / AutomorphismTransform; the reference's
// scatter src/Rotation.cuh:14-26 with k^-1 is the same permutation).
__global__ __launch_bounds__(256) void k_automorph_batch(const AutoTask* __restrict__ tasks,
                                                         const PrimeC* __restrict__ prm, u32 logN) {
  const AutoTask t = tasks[blockIdx.y];
  const u32 i = blockIdx.x * 256u + threadIdx.x;
  const u32 j = __brev(i) >> (32 - logN);
  const u32 mask = (2u << logN) - 1;
  const u32 s = (((2u * j + 1u) * t.kidx) & mask) >> 1;
  const u32 src = __brev(s) >> (32 - logN);
  u64 v = t.in[src];
  if (t.add) {  // uniform per task
    const u64 q = prm[t.pid].q;
    v = add_mod(v, shoup(t.add[src], t.c, t.csh, q), q);
  }
  t.out[i] = v;
}

// ---------------------------------------------------------------- linear combination
// Element-wise ciphertext arithmetic: add/sub, multiplication by an integer
// constant, addition of a constant (reference AddSub.cu:10-123 add_/sub_/
// scalar_ kernels; ElemenwiseBatchKernels.cu:1109-1360).
__global__ __launch_bounds__(256) void k_lin_batch(const LinTask* __restrict__ tasks, const PrimeC* __restrict__ prm,
                                                   u32 logN) {
  const LinTask t = tasks[blockIdx.y];
  const u32 k = blockIdx.x * 256u + threadIdx.x;
  const u64 q = prm[t.pid].q;
  u64 v = t.a[k];
  if (t.flags & LIN_AMONT) {  // v R^-1 mod q: Montgomery reduction of (0, v)
    PrimeC P;
    P.q = q;
    P.qneg_inv = prm[t.pid].qneg_inv;
    v = redc_once(0, v, P);
  }
  if (t.flags & LIN_MUL) v = shoup(v, t.m, t.msh, q);
  if (t.flags & LIN_B) v = (t.flags & LIN_SUB) ? sub_mod(v, t.b[k], q) : add_mod(v, t.b[k], q);
  t.out[k] = add_mod(v, t.add, q);
}

// ---------------------------------------------------------------- plaintext products
// Sum of up to 16 element-wise ciphertext-limb x plaintext-limb products with
// one Montgomery reduction (the reference's plaintext dot products of the
// linear transforms, ElemenwiseBatchKernels.cu:1007-1107).  x < q and
// m < q: the 128-bit sum stays below 16 q^2 < q 2^64.  Both ciphertext
// components in one task: each plaintext limb is read once for the two.
__global__ __launch_bounds__(256) void k_ptdot_batch(const PtDotTask* __restrict__ tasks,
                                                     const PtTerm* __restrict__ terms,
                                                     const PrimeC* __restrict__ prm, u32 logN) {
  const PtDotTask t = tasks[blockIdx.y];
  const u32 k = blockIdx.x * 256u + threadIdx.x;
  const PrimeC P = prm[t.pid];
  Acc128 a0, a1;
  for (u32 j = 0; j < t.cnt; ++j) {
    const PtTerm tm = terms[t.first + j];
    const u64 m = tm.m[k];
    acc_mac_v(a0, tm.x0[k], m);
    acc_mac_v(a1, tm.x1[k], m);
  }
  u64 h, l;
  acc_value(a0, h, l);
  t.out0[k] = mont_redc_any(h, l, P);
  acc_value(a1, h, l);
  t.out1[k] = mont_redc_any(h, l, P);
}

// ---------------------------------------------------------------- linear combination
// sum_j m_j x_j + add for up to 64 ciphertext limbs and integer constants
// (the Chebyshev-series combination; the reference evaluates it with
// scalar multiplications and additions, src/CKKS/ApproxModEval.cu).  The
// constants are uniform per task and arrive through scalar loads; one
// column accumulation and one Montgomery reduction per coefficient.  The
// column sums stay exact for 64 terms below 2^61 and for 16 terms below 2^62
// (fides_lincomb_batch refuses more); the high word is reduced mod q first.
__global__ __launch_bounds__(256) void k_lincomb_batch(const LcTask* __restrict__ tasks_,
                                                       const LcTerm* __restrict__ terms_,
                                                       const PrimeC* __restrict__ prm_, u32 logN) {
  const CONSTANT_AS LcTask* t = (const CONSTANT_AS LcTask*)(tasks_ + blockIdx.y);
  const CONSTANT_AS LcTerm* terms = (const CONSTANT_AS LcTerm*)terms_;
  const CONSTANT_AS PrimeC* prm = (const CONSTANT_AS PrimeC*)prm_;
  const u32 k = blockIdx.x * 256u + threadIdx.x;
  PrimeC P;
  P.q = prm[t->pid].q;
  P.qneg_inv = prm[t->pid].qneg_inv;
  P.one_sh = prm[t->pid].one_sh;
  const u32 first = t->first, cnt = t->cnt;
  Acc128 acc;
  for (u32 j = 0; j < cnt; ++j) acc_mac(acc, terms[first + j].x[k], terms[first + j].m);
  u64 hi, lo;
  acc_value(acc, hi, lo);
  t->out[k] = add_mod(mont_redc_any(hi, lo, P), t->add, P.q);
}

// ---------------------------------------------------------------- launchers
void fides_launch_lincomb_batch(const LcTask* tasks, const LcTerm* terms, u32 ntasks, const PrimeC* prm, u32 logN,
                                hipStream_t st) {
  if (ntasks == 0) return;
  k_lincomb_batch<<<dim3((1u << logN) / 256, ntasks), 256, 0, st>>>(tasks, terms, prm, logN);
}

void fides_launch_ptdot_batch(const PtDotTask* tasks, const PtTerm* terms, u32 ntasks, const PrimeC* prm, u32 logN,
                              hipStream_t st) {
  if (ntasks == 0) return;
  k_ptdot_batch<<<dim3((1u << logN) / 256, ntasks), 256, 0, st>>>(tasks, terms, prm, logN);
}

void fides_launch_lin_batch(const LinTask* tasks, u32 ntasks, const PrimeC* prm, u32 logN, hipStream_t st) {
  if (ntasks == 0) return;
  k_lin_batch<<<dim3((1u << logN) / 256, ntasks), 256, 0, st>>>(tasks, prm, logN);
}

int fides_launch_conv_batch(const ConvTask* tasks, u32 ntasks, u32 ns, bool split, const PrimeC* prm, u32 logN,
                            hipStream_t st) {
  if (ntasks == 0) return 0;
  dim3 g((1u << logN) / 256, ntasks);
  switch (ns) {
#define CASE(n)                                                             \
  case n:                                                                   \
    if (split)                                                              \
      k_conv_batch<n, true><<<g, 256, 0, st>>>(tasks, prm, logN);           \
    else                                                                    \
      k_conv_batch<n, false><<<g, 256, 0, st>>>(tasks, prm, logN);          \
    return 0;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
    CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: return -1;
  }
}

int fides_launch_conv_mixed(const ConvTask* tasks, u32 ntasks, u32 maxns, bool split, const PrimeC* prm, u32 logN,
                            hipStream_t st) {
  if (ntasks == 0) return 0;
  dim3 g((1u << logN) / 256, ntasks);
  if (maxns <= 10) {
    if (split)
      k_conv_batch_mixed<10, true><<<g, 256, 0, st>>>(tasks, prm, logN);
    else
      k_conv_batch_mixed<10, false><<<g, 256, 0, st>>>(tasks, prm, logN);
  } else if (maxns <= (u32)CONV_MAXS) {
    if (split)
      k_conv_batch_mixed<CONV_MAXS, true><<<g, 256, 0, st>>>(tasks, prm, logN);
    else
      k_conv_batch_mixed<CONV_MAXS, false><<<g, 256, 0, st>>>(tasks, prm, logN);
  } else {
    return -1;
  }
  return 0;
}

void fides_launch_dot_batch(const DotTask* tasks, u32 ntasks, const PrimeC* prm, u32 logN, u32 kstride,
                            hipStream_t st) {
  if (ntasks == 0) return;
  k_dot_batch<<<dim3((1u << logN) / 256, ntasks), 256, 0, st>>>(tasks, prm, logN, kstride);
}

void fides_launch_tensor_batch(const TensorTask* tasks, u32 ntasks, const PrimeC* prm, u32 logN, hipStream_t st) {
  if (ntasks == 0) return;
  k_tensor_batch<<<dim3((1u << logN) / 256, ntasks), 256, 0, st>>>(tasks, prm, logN);
}

void fides_launch_automorph_batch(const AutoTask* tasks, u32 ntasks, const PrimeC* prm, u32 logN, hipStream_t st) {
  if (ntasks == 0) return;
  k_automorph_batch<<<dim3((1u << logN) / 256, ntasks), 256, 0, st>>>(tasks, prm, logN);
}

void fides_launch_to_mont(u64* key, const PrimeC* prm, u32 logN, u32 nlimbs, u32 LK, hipStream_t st) {
  k_to_mont<<<dim3((1u << logN) / 256, nlimbs), 256, 0, st>>>(key, prm, logN, LK);
}
