// Device-side modular arithmetic on 64-bit residues for gfx950.
//
// CDNA has no 64x64 multiplier; every 64-bit product is built from 32x32->64
// partial products (v_mad_u64_u32 / v_mul_lo_u32 / v_mul_hi_u32).  Three
// reductions are used, each picked for the operand pattern of its caller:
//
//   shoup_lazy   x * w for a constant w with precomputed w' = floor(w 2^64 / q)
//                (Shoup; reference src/ModMult.cuh:83-95).  x < 2^64, result
//                in [0, 2q).  Used by every NTT butterfly and constant scale.
//   mont_redc    reduction of a 128-bit accumulator T (sum of products with
//                one operand stored pre-multiplied by R = 2^64).  Used by the
//                basis conversions and the key-switch inner product, which
//                accumulate several 128-bit products before one reduction.
//   redc_once    the same for T < qR, with one conditional subtraction; with
//                to_mont (y -> yR mod q) it multiplies two variable operands
//                (the tensor product), where the reference uses Barrett
//                (src/ModMult.cuh:42-69): both give the canonical residue.
//
// All residues are kept below 4q (lazy reduction); q < 2^62 is required so
// that 4q fits in 64 bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint64_t u64;
typedef uint32_t u32;

// Per-prime constants, indexed by prime id (Q primes first, then P primes).
struct PrimeC {
  u64 q;         // the prime
  u64 q2;        // 2q
  u64 qneg_inv;  // -q^-1 mod 2^64 (Montgomery)
  u64 r2;        // R^2 mod q, R = 2^64
  u64 ninv;      // N^-1 mod q
  u64 ninv_sh;   // Shoup companion of ninv
  u64 ninv_w;    // ipsi[1] * N^-1 mod q (last inverse stage, scaled)
  u64 ninv_w_sh;
  u64 one_sh;    // floor(2^64 / q): Shoup companion of 1, reduces any x < 2^64
  u32 bits;
  u32 pad;
};

__device__ __forceinline__ void mul64wide(u64 a, u64 b, u64& hi, u64& lo) {
  const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
  const u64 p00 = (u64)a0 * b0;
  const u64 p01 = (u64)a0 * b1;
  const u64 p10 = (u64)a1 * b0;
  const u64 p11 = (u64)a1 * b1;
  const u64 mid = (p00 >> 32) + (u32)p01 + (u32)p10;
  lo = (mid << 32) | (u32)p00;
  hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}

// High 64 bits of a * b: one mul_hi plus three v_mad_u64_u32 whose 64-bit
// addends carry the partial sums (no intermediate can overflow 64 bits).
__device__ __forceinline__ u64 mulhi64(u64 a, u64 b) {
  const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
  const u64 t = (u64)a0 * b1 + (u64)__umulhi(a0, b0);  // < 2^64
  const u64 u = (u64)a1 * b0 + (u32)t;                 // < 2^64
  return (u64)a1 * b1 + (t >> 32) + (u >> 32);
}

// Low 64 bits of a * b + c: one v_mad_u64_u32 for the low word and the two
// cross products folded into the high word with 32-bit arithmetic.
__device__ __forceinline__ u64 mullo64_add(u64 a, u64 b, u64 c) {
  const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
  const u64 p = (u64)a0 * b0 + c;
  const u32 hi = (u32)(p >> 32) + a0 * b1 + a1 * b0;
  return ((u64)hi << 32) | (u32)p;
}

// x * w mod q in [0, 2q); any x < 2^64, w < q, wp = floor(w * 2^64 / q),
// qn = 2^64 - q.  r = x*w + h*(-q) mod 2^64 needs no final subtraction.
__device__ __forceinline__ u64 shoup_lazy(u64 x, u64 w, u64 wp, u64 q) {
  const u64 h = mulhi64(x, wp);
  return mullo64_add(h, (u64)0 - q, x * w);
}

// The same product for the NTT butterflies, with ALU-only inline assembly
// where the compiler's lowering costs extra instructions: the middle column
// of x * wp is summed by one v_mad_u64_u32 whose carry-out is added back with
// v_addc (instead of three register-pair moves), and the cross products of
// h * (2^64 - q) are two v_mul_lo_u32 and one v_add3_u32 (the compiler
// otherwise folds them into a v_mad_u64_u32 with two moves and a 64-bit add).
// WS: w and wp are wave-uniform (scalar registers).  qn = 2^64 - q, uniform.
template <bool WS>
__device__ __forceinline__ u64 shoup_lazy_bf(u64 x, u64 w, u64 wp, u64 qn) {
  const u32 x0 = (u32)x, x1 = (u32)(x >> 32);
  const u32 p0 = (u32)wp, p1 = (u32)(wp >> 32);
  u64 B = (u64)x0 * p1 + (u64)__umulhi(x0, p0);  // < 2^64
  u64 c;
  if constexpr (WS)
    asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(B), "=s"(c) : "v"(x1), "s"(p0));
  else
    asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(B), "=s"(c) : "v"(x1), "v"(p0));
  // floor(x wp / 2^64) = x1 p1 + floor(B / 2^32) + c 2^32 (fits 64 bits)
  const u64 H = (u64)x1 * p1 + (B >> 32);
  u32 h1 = (u32)(H >> 32);
  u64 junk;
  asm("v_addc_co_u32_e64 %0, %1, %0, 0, %2" : "+v"(h1), "=s"(junk) : "s"(c));
  const u32 h0 = (u32)H;
  const u64 P = (u64)h0 * (u32)qn + x * w;  // low word: x w + h0 qn0
  u32 t1, t2, hi;
  asm("v_mul_lo_u32 %0, %1, %2" : "=v"(t1) : "v"(h0), "s"((u32)(qn >> 32)));
  asm("v_mul_lo_u32 %0, %1, %2" : "=v"(t2) : "v"(h1), "s"((u32)qn));
  asm("v_add3_u32 %0, %1, %2, %3" : "=v"(hi) : "v"((u32)(P >> 32)), "v"(t1), "v"(t2));
  return ((u64)hi << 32) | (u32)P;
}

// x - m if x >= m else x: subtract with borrow, then select on the borrow
// (four instructions; m is kept in vector registers).
__device__ __forceinline__ u64 csub_bf(u64 x, u64 m) {
  const u32 x0 = (u32)x, x1 = (u32)(x >> 32);
  u32 r0, r1;
  asm("v_sub_co_u32_e32 %0, vcc, %2, %4\n\t"
      "v_subb_co_u32_e32 %1, vcc, %3, %5, vcc\n\t"
      "v_cndmask_b32_e32 %0, %0, %2, vcc\n\t"
      "v_cndmask_b32_e32 %1, %1, %3, vcc"
      : "=&v"(r0), "=&v"(r1)
      : "v"(x0), "v"(x1), "v"((u32)m), "v"((u32)(m >> 32))
      : "vcc");
  return ((u64)r1 << 32) | r0;
}

__device__ __forceinline__ u64 csub(u64 x, u64 m) {
  const u64 t = x + (0 - m);  // one 64-bit add, no carry chain
  return x >= m ? t : x;
}

// x * w mod q in [0, q).
__device__ __forceinline__ u64 shoup(u64 x, u64 w, u64 wp, u64 q) { return csub(shoup_lazy(x, w, wp, q), q); }


// 128-bit accumulate: (hi, lo) += a * b, with the low-word carry taken from
// an explicit add-with-carry.
__device__ __forceinline__ void mac128(u64& hi, u64& lo, u64 a, u64 b) {
  const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
  const u64 p00 = (u64)a0 * b0;
  const u64 t = (u64)a0 * b1 + (p00 >> 32);  // < 2^64
  const u64 u = (u64)a1 * b0 + (u32)t;       // < 2^64
  const u64 h = (u64)a1 * b1 + (t >> 32) + (u >> 32);
  const u64 l = (u << 32) | (u32)p00;
  unsigned long long c;
  lo = __builtin_addcll(lo, l, 0ull, &c);
  hi = hi + h + c;
}

// Column accumulator for sums of up to 16 products x * m with x, m < 2^62
// (or up to 64 with x, m < 2^61: H then stays below 64 * 2^58):
//   T = L + 2^32 M + 2^64 (H + cL) + 2^96 cM
// where L = sum x0 m0, M = sum (x0 m1 + x1 m0) and H = sum x1 m1 (32-bit
// halves), each kept modulo 2^64 with the carries out of L and M counted
// separately.  Every product is one v_mad_u64_u32 accumulating into its
// column; the carry-out of the instruction feeds a 32-bit counter, so no
// carry is propagated inside the loop.  (ALU-only inline assembly: the
// compiler cannot express a consumed carry-out of v_mad_u64_u32.)
struct Acc128 {
  u64 L = 0, M = 0, H = 0;
  u32 cL = 0, cM = 0;
};

__device__ __forceinline__ void acc_mac(Acc128& a, u64 x, u64 m) {
  const u32 x0 = (u32)x, x1 = (u32)(x >> 32), m0 = (u32)m, m1 = (u32)(m >> 32);
  u64 c0, c1, c2;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a.L), "=s"(c0) : "v"(x0), "s"(m0));
  asm("v_addc_co_u32_e64 %0, %1, %0, 0, %2" : "+v"(a.cL), "=s"(c2) : "s"(c0));
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a.M), "=s"(c1) : "v"(x0), "s"(m1));
  asm("v_addc_co_u32_e64 %0, %1, %0, 0, %2" : "+v"(a.cM), "=s"(c2) : "s"(c1));
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a.M), "=s"(c1) : "v"(x1), "s"(m0));
  asm("v_addc_co_u32_e64 %0, %1, %0, 0, %2" : "+v"(a.cM), "=s"(c2) : "s"(c1));
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a.H), "=s"(c1) : "v"(x1), "s"(m1));
}

// Same with both operands per lane.
__device__ __forceinline__ void acc_mac_v(Acc128& a, u64 x, u64 m) {
  const u32 x0 = (u32)x, x1 = (u32)(x >> 32), m0 = (u32)m, m1 = (u32)(m >> 32);
  u64 c0, c1, c2;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a.L), "=s"(c0) : "v"(x0), "v"(m0));
  asm("v_addc_co_u32_e64 %0, %1, %0, 0, %2" : "+v"(a.cL), "=s"(c2) : "s"(c0));
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a.M), "=s"(c1) : "v"(x0), "v"(m1));
  asm("v_addc_co_u32_e64 %0, %1, %0, 0, %2" : "+v"(a.cM), "=s"(c2) : "s"(c1));
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a.M), "=s"(c1) : "v"(x1), "v"(m0));
  asm("v_addc_co_u32_e64 %0, %1, %0, 0, %2" : "+v"(a.cM), "=s"(c2) : "s"(c1));
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a.H), "=s"(c1) : "v"(x1), "v"(m1));
}

// The 128-bit value (hi, lo) of a column accumulator.
__device__ __forceinline__ void acc_value(const Acc128& a, u64& hi, u64& lo) {
  const u64 ms = a.M << 32;
  unsigned long long c;
  lo = __builtin_addcll(a.L, ms, 0ull, &c);
  hi = a.H + (u64)a.cL + (a.M >> 32) + ((u64)a.cM << 32) + c;
}

// Montgomery reduction of T = hi * 2^64 + lo: returns T * 2^-64 mod q, fully
// reduced.  Valid for hi < c * q with a small c (the result before the final
// loop is below (c + 1) q); callers bound c by their accumulation length.
__device__ __forceinline__ u64 mont_redc(u64 hi, u64 lo, const PrimeC& P) {
  const u64 m = lo * P.qneg_inv;
  u64 u = hi + mulhi64(m, P.q) + (lo != 0);
  while (u >= P.q) u -= P.q;
  return u;
}

// Montgomery reduction of T = (hi, lo) < q R: T R^-1 mod q in [0, q)
// ((T + m q) / R < 2q, one conditional subtraction).
__device__ __forceinline__ u64 redc_once(u64 hi, u64 lo, const PrimeC& P) {
  const u64 m = lo * P.qneg_inv;
  return csub(hi + mulhi64(m, P.q) + (lo != 0), P.q);
}
// Montgomery reduction of any T = (hi, lo) < 2^128: the high word is first
// reduced mod q (a Shoup product by 1, exact for any 64-bit value), so the
// reduced T' = (hi mod q) R + lo is below q R and one conditional
// subtraction finishes.  Used by the accumulations whose high word can reach
// 4q and more (up to 16 products with q near 2^62, or 64 products below
// 2^61), where the sum hi + floor(m q / R) of mont_redc would overflow.
__device__ __forceinline__ u64 mont_redc_any(u64 hi, u64 lo, const PrimeC& P) {
  return redc_once(shoup(hi, 1, P.one_sh, P.q), lo, P);
}
// y R mod q for y < q: REDC(y (R^2 mod q)).
__device__ __forceinline__ u64 to_mont(u64 y, const PrimeC& P) {
  u64 hi, lo;
  mul64wide(y, P.r2, hi, lo);
  return redc_once(hi, lo, P);
}

__device__ __forceinline__ u64 add_mod(u64 a, u64 b, u64 q) { return csub(a + b, q); }
__device__ __forceinline__ u64 sub_mod(u64 a, u64 b, u64 q) { return a >= b ? a - b : a + q - b; }
