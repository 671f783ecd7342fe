// Element-wise and basis-conversion kernels of the hybrid key switch.
//
// All kernels map one thread to one coefficient position (coalesced 64-bit
// accesses along a limb) and the grid's y dimension to limbs or limb chunks.
#include "kernels.h"

// ---------------------------------------------------------------- tensor
// (a0,a1) x (b0,b1) -> (a0 b0, a0 b1 + a1 b0, a1 b1) per limb; limb i uses
// prime i (reference src/CKKS/ElemenwiseBatchKernels.cu:1363-1400,
// binomialMult_).  Outputs o0, o1, o2 may alias a0 / a1 (each element is
// read before it is written by the same thread).
__global__ __launch_bounds__(256) void k_tensor(const u64* a0, const u64* a1, const u64* b0, const u64* b1,
                                                u64* o0, u64* o1, u64* o2, const PrimeC* prm, u32 logN) {
  const u32 limb = blockIdx.y;
  const u64 i = ((u64)limb << logN) + blockIdx.x * 256u + threadIdx.x;
  const PrimeC P = prm[limb];
  const u64 x0 = a0[i], x1 = a1[i], y0 = b0[i], y1 = b1[i];
  const u64 r0 = barrett_mul(x0, y0, P);
  const u64 r1 = add_mod(barrett_mul(x0, y1, P), barrett_mul(x1, y0, P), P.q);
  const u64 r2 = barrett_mul(x1, y1, P);
  o0[i] = r0;
  o1[i] = r1;
  o2[i] = r2;
}

// Squaring variant: (a0^2, 2 a0 a1, a1^2).
__global__ __launch_bounds__(256) void k_square(const u64* a0, const u64* a1, u64* o0, u64* o1, u64* o2,
                                                const PrimeC* prm, u32 logN) {
  const u32 limb = blockIdx.y;
  const u64 i = ((u64)limb << logN) + blockIdx.x * 256u + threadIdx.x;
  const PrimeC P = prm[limb];
  const u64 x0 = a0[i], x1 = a1[i];
  const u64 r0 = barrett_mul(x0, x0, P);
  const u64 m = barrett_mul(x0, x1, P);
  const u64 r1 = add_mod(m, m, P.q);
  const u64 r2 = barrett_mul(x1, x1, P);
  o0[i] = r0;
  o1[i] = r1;
  o2[i] = r2;
}

// ---------------------------------------------------------------- basis conversion
// y_t = sum_i [x_i * hatinv_i]_{q_i} * hatmod_{i,t}  mod p_t
// (reference src/CKKS/Conv.cu:258-402 DecompAndModUpConv and :26-125
// ModDown2; OpenFHE ApproxSwitchCRTBasis).  hatmod is stored pre-multiplied
// by R = 2^64 so that one Montgomery reduction of the 128-bit accumulator
// yields the exact residue.  Coefficient form in and out.  The NS source
// limbs are contiguous at `in`; target t is written to out + dst_off[t] * N.
// grid: x = N / 256, y = ceil(nt / CONV_TCHUNK); the targets of one
// workgroup row share the loaded and pre-scaled inputs.
template <int NS>
__global__ __launch_bounds__(256) void k_conv(const u64* __restrict__ in, u64* __restrict__ out,
                                              const ConvDesc* __restrict__ d, const PrimeC* __restrict__ prm,
                                              u32 logN) {
  const u32 k = blockIdx.x * 256u + threadIdx.x;
  const u32 nt = d->nt;
  u64 xs[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const u64 q = prm[d->src_pid[i]].q;
    xs[i] = shoup(in[((u64)i << logN) + k], d->hatinv[i], d->hatinv_sh[i], q);
  }
  const u32 t0 = blockIdx.y * CONV_TCHUNK;
  const u32 t1 = min(nt, t0 + CONV_TCHUNK);
  for (u32 t = t0; t < t1; ++t) {
    const PrimeC P = prm[d->dst_pid[t]];
    const u64* m = d->mat[t];
    u64 hi = 0, lo = 0;
#pragma unroll
    for (int i = 0; i < NS; ++i) mac128(hi, lo, xs[i], m[i]);
    out[((u64)d->dst_off[t] << logN) + k] = mont_redc(hi, lo, P);
  }
}

// ---------------------------------------------------------------- key-switch inner product
// For every limb t of Q_l u P:  (o0, o1)[t] = sum_{j<beta} e_j[t] * (b_j, a_j)[t]
// where e_j = ModUp of digit j and, for the digit's own primes, the
// original evaluation-form limb (reference
// src/CKKS/ElemenwiseBatchKernels.cu:202-273 fusedDotKSK_2_, :246-255).
// The key is stored in Montgomery form (times R), so one reduction of the
// 128-bit sums gives the exact result.
__global__ __launch_bounds__(256) void k_ksk_dot(KsDotArgs A) {
  const u32 t = blockIdx.y;
  const u64 k = blockIdx.x * 256u + threadIdx.x;
  const u32 l = A.l;
  const u32 pid = t < l ? t : A.L + (t - l);
  const PrimeC P = A.prm[pid];
  const u64 N = 1ull << A.logN;
  u64 h0 = 0, l0 = 0, h1 = 0, l1 = 0;
  for (u32 j = 0; j < A.beta; ++j) {
    const bool own = t < l && t >= j * A.alpha && t < (j + 1) * A.alpha;
    const u64 e = own ? A.d[(u64)t * N + k] : A.ext[((u64)j * A.ext_stride + t) * N + k];
    const u64 kb = A.key_b[((u64)j * A.key_stride + pid) * N + k];
    const u64 ka = A.key_a[((u64)j * A.key_stride + pid) * N + k];
    mac128(h0, l0, e, kb);
    mac128(h1, l1, e, ka);
  }
  A.o0[(u64)t * N + k] = mont_redc(h0, l0, P);
  A.o1[(u64)t * N + k] = mont_redc(h1, l1, P);
}

// ---------------------------------------------------------------- key conversion
// x <- x * R mod q (Montgomery form) over [2][dnum][L+K][N].
__global__ __launch_bounds__(256) void k_to_mont(u64* key, const PrimeC* prm, u32 logN, u32 LK) {
  const u32 limb = blockIdx.y;  // over 2 * dnum * LK
  const u32 pid = limb % LK;
  const PrimeC P = prm[pid];
  const u64 i = ((u64)limb << logN) + blockIdx.x * 256u + threadIdx.x;
  u64 hi, lo;
  mul64wide(key[i], P.r2, hi, lo);
  key[i] = mont_redc(hi, lo, P);
}

// ---------------------------------------------------------------- automorphism
// out[brev(j)] = in[brev(((2j+1) k mod 2N) >> 1)], evaluation form
// (OpenFHE PrecomputeAutoMap / AutomorphismTransform; the reference's
// scatter src/Rotation.cuh:14-26 with k^-1 is the same permutation).
__global__ __launch_bounds__(256) void k_automorph(const u64* __restrict__ in, u64* __restrict__ out, u32 logN,
                                                   u32 kidx) {
  const u32 i = blockIdx.x * 256u + threadIdx.x;
  const u64 base = (u64)blockIdx.y << logN;
  const u32 j = __brev(i) >> (32 - logN);
  const u32 mask = (2u << logN) - 1;
  const u32 s = (((2u * j + 1u) * kidx) & mask) >> 1;
  const u32 src = __brev(s) >> (32 - logN);
  out[base + i] = in[base + src];
}

// ---------------------------------------------------------------- element-wise
__global__ __launch_bounds__(256) void k_add(const u64* a, const u64* b, u64* o, const PrimeC* prm, u32 logN,
                                             u32 pid0) {
  const u32 limb = blockIdx.y;
  const u64 i = ((u64)limb << logN) + blockIdx.x * 256u + threadIdx.x;
  o[i] = add_mod(a[i], b[i], prm[pid0 + limb].q);
}

// ---------------------------------------------------------------- launchers
void fides_launch_tensor(const u64* a0, const u64* a1, const u64* b0, const u64* b1, u64* o0, u64* o1, u64* o2,
                         const PrimeC* prm, u32 logN, u32 l, bool square, hipStream_t st) {
  dim3 g((1u << logN) / 256, l);
  if (square)
    k_square<<<g, 256, 0, st>>>(a0, a1, o0, o1, o2, prm, logN);
  else
    k_tensor<<<g, 256, 0, st>>>(a0, a1, b0, b1, o0, o1, o2, prm, logN);
}

int fides_launch_conv(const u64* in, u64* out, const ConvDesc* d, u32 ns, u32 nt, const PrimeC* prm, u32 logN,
                      hipStream_t st) {
  if (nt == 0) return 0;
  dim3 g((1u << logN) / 256, (nt + CONV_TCHUNK - 1) / CONV_TCHUNK);
  switch (ns) {
#define CASE(n) case n: k_conv<n><<<g, 256, 0, st>>>(in, out, d, prm, logN); return 0;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
    CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: return -1;
  }
}

void fides_launch_ksk_dot(const KsDotArgs& A, hipStream_t st) {
  dim3 g((1u << A.logN) / 256, A.l + A.K);
  k_ksk_dot<<<g, 256, 0, st>>>(A);
}

void fides_launch_to_mont(u64* key, const PrimeC* prm, u32 logN, u32 nlimbs, u32 LK, hipStream_t st) {
  dim3 g((1u << logN) / 256, nlimbs);
  k_to_mont<<<g, 256, 0, st>>>(key, prm, logN, LK);
}

void fides_launch_automorph(const u64* in, u64* out, u32 logN, u32 nlimbs, u32 kidx, hipStream_t st) {
  dim3 g((1u << logN) / 256, nlimbs);
  k_automorph<<<g, 256, 0, st>>>(in, out, logN, kidx);
}

void fides_launch_add(const u64* a, const u64* b, u64* o, const PrimeC* prm, u32 logN, u32 nlimbs, u32 pid0,
                      hipStream_t st) {
  dim3 g((1u << logN) / 256, nlimbs);
  k_add<<<g, 256, 0, st>>>(a, b, o, prm, logN, pid0);
}
