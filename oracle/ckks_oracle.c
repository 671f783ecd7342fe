/*
 * ckks_oracle.c -- CPU restatement of the CKKS-RNS hybrid key-switching path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library
 * (fideslib_amd/) never links, loads or calls it.
 *
 * Every function restates the algorithm of the reference path it cites
 * (paths relative to the FIDESlib source tree; OpenFHE routines the
 * reference mirrors are named where the reference quotes them).  All
 * arithmetic is exact modular integer arithmetic on 64-bit residues with
 * 128-bit intermediates, written for clarity, not speed (OpenMP over limbs
 * is the only concession to run time).
 *
 * Pinning (what fixes each function independently of this file) is listed
 * in tests/test_oracle.py; every function here is checked there against
 * golden vectors from an independent Python big-integer implementation
 * (tests/golden/), closed forms and brute force.
 *
 * Conventions
 *   - A limb is N = 2^logN contiguous uint64 residues in [0, q).
 *   - Evaluation (NTT) form uses OpenFHE's bit-reversed order:
 *       NTT(a)[k] = sum_i a_i * psi^(i * (2*brev(k) + 1))  mod q
 *     (reference: src/ConstantsGPU.cu:252-269 psi[j] = psi_no[brev(j)],
 *      test/cpuNTT_nega.cpp:24-103 Cooley-Tukey loop with psi[m+i]).
 *   - Prime indices: 0..L-1 are the Q primes (L = number of Q primes),
 *     L..L+K-1 are the special P primes.
 *   - A key-switching key is laid out [2][dnum][L+K][N] uint64, component
 *     0 = a, component 1 = b (reference RawKeySwitchKey r_key = {A, B},
 *     src/CKKS/openfhe-interface/RawCiphertext.cu:649-668).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef unsigned __int128 u128;

#include <omp.h>

#define OR_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Scalar modular arithmetic (reference: src/Math.cuh:15-64 modpow,    */
/* modinv; src/ModMult.cuh:42-297 is the GPU form of a*b mod q).       */
/* ------------------------------------------------------------------ */

static inline u64 mulmod(u64 a, u64 b, u64 q) { return (u64)(((u128)a * b) % q); }
static inline u64 addmod(u64 a, u64 b, u64 q) { u64 s = a + b; return s >= q ? s - q : s; }
static inline u64 submod(u64 a, u64 b, u64 q) { return a >= b ? a - b : a + q - b; }

OR_API u64 or_mulmod(u64 a, u64 b, u64 q) { return mulmod(a, b, q); }

OR_API u64 or_powmod(u64 a, u64 e, u64 q) {
  u64 r = 1 % q;
  a %= q;
  while (e) {
    if (e & 1) r = mulmod(r, a, q);
    a = mulmod(a, a, q);
    e >>= 1;
  }
  return r;
}

/* q prime: a^(q-2) */
OR_API u64 or_invmod(u64 a, u64 q) { return or_powmod(a, q - 2, q); }

OR_API uint32_t or_brev(uint32_t x, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

/* Smallest primitive root of prime q (reference: src/ConstantsGPU.cu:36-56
 * generator(): trial over 2,3,... testing g^((q-1)/f) != 1 for every prime
 * factor f of q-1). */
OR_API u64 or_generator(u64 q) {
  u64 fact[64];
  int nf = 0;
  u64 n = q - 1;
  for (u64 i = 2; i * i <= n; ++i) {
    if (n % i == 0) {
      fact[nf++] = i;
      while (n % i == 0) n /= i;
    }
  }
  if (n > 1) fact[nf++] = n;
  for (u64 g = 2; g < q; ++g) {
    int ok = 1;
    for (int i = 0; i < nf && ok; ++i) ok = or_powmod(g, (q - 1) / fact[i], q) != 1;
    if (ok) return g;
  }
  return 0;
}

/* Primitive 2N-th root of unity psi = g^((q-1)/2N) (reference:
 * src/ConstantsGPU.cu:58-67 unity_root(q, 2N), the fallback used when no
 * OpenFHE root is supplied, :207-208). */
OR_API u64 or_root_2n(u64 q, uint32_t logN) {
  u64 order = (u64)2 << logN;
  if ((q - 1) % order) return 0;
  return or_powmod(or_generator(q), (q - 1) / order, q);
}

/* psi_rev[j] = psi^brev(j), ipsi_rev[j] = psi^-brev(j)
 * (reference: src/ConstantsGPU.cu:226-269 psi_no[j] = root^j,
 *  psi[j] = psi_no[bit_reverse(j, logN)]). */
OR_API void or_psi_tables(u64 q, u64 psi, uint32_t logN, u64* psi_rev, u64* ipsi_rev) {
  const uint32_t N = 1u << logN;
  u64 ipsi = or_invmod(psi, q);
  u64* pw = (u64*)malloc(sizeof(u64) * N);
  u64* ipw = (u64*)malloc(sizeof(u64) * N);
  pw[0] = 1;
  ipw[0] = 1;
  for (uint32_t j = 1; j < N; ++j) {
    pw[j] = mulmod(pw[j - 1], psi, q);
    ipw[j] = mulmod(ipw[j - 1], ipsi, q);
  }
  for (uint32_t j = 0; j < N; ++j) {
    psi_rev[j] = pw[or_brev(j, logN)];
    ipsi_rev[j] = ipw[or_brev(j, logN)];
  }
  free(pw);
  free(ipw);
}

/* ------------------------------------------------------------------ */
/* Negacyclic NTT / INTT                                               */
/* ------------------------------------------------------------------ */

/* Forward: natural-order coefficients -> bit-reversed evaluations.
 * Cooley-Tukey loop of test/cpuNTT_nega.cpp:24-103 (without its final
 * bit-reversal swap, because evaluation form is kept bit-reversed as in
 * OpenFHE, reference src/CKKS/openfhe-interface/RawCiphertext.cu:47-73). */
OR_API void or_ntt(u64* a, uint32_t logN, u64 q, const u64* psi_rev) {
  const uint32_t N = 1u << logN;
  uint32_t t = N >> 1;
  for (uint32_t m = 1; m < N; m <<= 1, t >>= 1) {
    for (uint32_t i = 0; i < m; ++i) {
      const u64 w = psi_rev[m + i];
      const uint32_t j1 = 2 * i * t;
      for (uint32_t j = j1; j < j1 + t; ++j) {
        u64 u = a[j];
        u64 v = mulmod(a[j + t], w, q);
        a[j] = addmod(u, v, q);
        a[j + t] = submod(u, v, q);
      }
    }
  }
}

/* Inverse: Gentleman-Sande loop with psi^-brev table, then scaling by
 * N^-1 (reference src/NTT.cu:42-310 INTT_ with inv_psi_no pre-scaled by
 * N^-1, src/ConstantsGPU.cu:295-300; OpenFHE
 * InverseTransformFromBitReverseInPlace). */
OR_API void or_intt(u64* a, uint32_t logN, u64 q, const u64* ipsi_rev) {
  const uint32_t N = 1u << logN;
  uint32_t t = 1;
  for (uint32_t m = N >> 1; m >= 1; m >>= 1, t <<= 1) {
    for (uint32_t i = 0; i < m; ++i) {
      const u64 w = ipsi_rev[m + i];
      const uint32_t j1 = 2 * i * t;
      for (uint32_t j = j1; j < j1 + t; ++j) {
        u64 u = a[j];
        u64 v = a[j + t];
        a[j] = addmod(u, v, q);
        a[j + t] = mulmod(submod(u, v, q), w, q);
      }
    }
  }
  const u64 ninv = or_invmod(N % q, q);
  for (uint32_t j = 0; j < N; ++j) a[j] = mulmod(a[j], ninv, q);
}

/* ------------------------------------------------------------------ */
/* Context: primes, roots, digit partition, CRT tables                 */
/* ------------------------------------------------------------------ */

typedef struct {
  uint32_t logN, N, L, K, dnum, alpha;
  u64* primes;   /* L + K */
  u64* psi;      /* L + K roots */
  u64* psi_rev;  /* (L+K) * N */
  u64* ipsi_rev; /* (L+K) * N */
} or_ctx;

OR_API or_ctx* or_ctx_create(uint32_t logN, const u64* q, uint32_t L, const u64* p, uint32_t K,
                             uint32_t dnum, const u64* roots /* may be NULL */) {
  or_ctx* c = (or_ctx*)calloc(1, sizeof(or_ctx));
  c->logN = logN;
  c->N = 1u << logN;
  c->L = L;
  c->K = K;
  c->dnum = dnum;
  /* digit size alpha = ceil(L / dnum): OpenFHE m_numPartQ/numPerPartQ,
   * mirrored by reference src/CKKS/Context.cu:190-292 (generateMeta). */
  c->alpha = (L + dnum - 1) / dnum;
  const uint32_t np = L + K;
  c->primes = (u64*)malloc(sizeof(u64) * np);
  c->psi = (u64*)malloc(sizeof(u64) * np);
  c->psi_rev = (u64*)malloc(sizeof(u64) * np * c->N);
  c->ipsi_rev = (u64*)malloc(sizeof(u64) * np * c->N);
  for (uint32_t i = 0; i < L; ++i) c->primes[i] = q[i];
  for (uint32_t i = 0; i < K; ++i) c->primes[L + i] = p[i];
#pragma omp parallel for schedule(dynamic)
  for (uint32_t i = 0; i < np; ++i) {
    c->psi[i] = roots ? roots[i] : or_root_2n(c->primes[i], logN);
    or_psi_tables(c->primes[i], c->psi[i], logN, c->psi_rev + (size_t)i * c->N,
                  c->ipsi_rev + (size_t)i * c->N);
  }
  return c;
}

OR_API void or_ctx_destroy(or_ctx* c) {
  if (!c) return;
  free(c->primes);
  free(c->psi);
  free(c->psi_rev);
  free(c->ipsi_rev);
  free(c);
}

OR_API u64 or_ctx_root(const or_ctx* c, uint32_t pid) { return c->psi[pid]; }
OR_API uint32_t or_ctx_alpha(const or_ctx* c) { return c->alpha; }

/* NTT / INTT of n limbs; limb t of data uses prime pids[t]. */
OR_API void or_ntt_limbs(const or_ctx* c, u64* data, const uint32_t* pids, uint32_t n) {
#pragma omp parallel for schedule(static)
  for (uint32_t t = 0; t < n; ++t)
    or_ntt(data + (size_t)t * c->N, c->logN, c->primes[pids[t]], c->psi_rev + (size_t)pids[t] * c->N);
}
OR_API void or_intt_limbs(const or_ctx* c, u64* data, const uint32_t* pids, uint32_t n) {
#pragma omp parallel for schedule(static)
  for (uint32_t t = 0; t < n; ++t)
    or_intt(data + (size_t)t * c->N, c->logN, c->primes[pids[t]], c->ipsi_rev + (size_t)pids[t] * c->N);
}

/* Product of primes[ids[0..n)] except index skip, mod m. */
static u64 prod_except_mod(const u64* pr, const uint32_t* ids, uint32_t n, uint32_t skip, u64 m) {
  u64 r = 1 % m;
  for (uint32_t t = 0; t < n; ++t)
    if (t != skip) r = mulmod(r, pr[ids[t]] % m, m);
  return r;
}

/* Digit j of a ciphertext with l active Q limbs: first limb and size.
 * (OpenFHE m_paramsPartQ partition, reference
 *  src/CKKS/openfhe-interface/RawCiphertext.cu:412-420.) */
static void digit_range(const or_ctx* c, uint32_t l, uint32_t j, uint32_t* start, uint32_t* size) {
  *start = j * c->alpha;
  uint32_t end = (j + 1) * c->alpha;
  if (end > l) end = l;
  *size = end - *start;
}

OR_API uint32_t or_num_digits(const or_ctx* c, uint32_t l) { return (l + c->alpha - 1) / c->alpha; }

/* ------------------------------------------------------------------ */
/* Basis conversion (ModUp)                                            */
/* ------------------------------------------------------------------ */

/* Approximate CRT basis switch of one coefficient vector set:
 *   y_t = sum_i [x_i * (Qhat/q_i)^-1]_{q_i} * [(Qhat/q_i)]_{p_t}   mod p_t
 * where Qhat = prod of the source primes (reference src/CKKS/Conv.cu:258-402
 * DecompAndModUpConv with tables PartQlHatInvModq / PartQlHatModp,
 * src/ConstantsGPU.cu:528-621; OpenFHE ApproxSwitchCRTBasis).  Input and
 * output are in coefficient form. */
OR_API void or_switch_basis(const or_ctx* c, const u64* in, const uint32_t* src, uint32_t ns, u64* out,
                            const uint32_t* dst, uint32_t nd) {
  const uint32_t N = c->N;
  u64* hatinv = (u64*)malloc(sizeof(u64) * ns);
  u64* hatmod = (u64*)malloc(sizeof(u64) * ns * nd);
  for (uint32_t i = 0; i < ns; ++i) {
    const u64 qi = c->primes[src[i]];
    hatinv[i] = or_invmod(prod_except_mod(c->primes, src, ns, i, qi), qi);
    for (uint32_t t = 0; t < nd; ++t) hatmod[i * nd + t] = prod_except_mod(c->primes, src, ns, i, c->primes[dst[t]]);
  }
#pragma omp parallel for schedule(static)
  for (uint32_t k = 0; k < N; ++k) {
    u64 xs[64];
    for (uint32_t i = 0; i < ns; ++i) xs[i] = mulmod(in[(size_t)i * N + k], hatinv[i], c->primes[src[i]]);
    for (uint32_t t = 0; t < nd; ++t) {
      const u64 pt = c->primes[dst[t]];
      u128 acc = 0;
      for (uint32_t i = 0; i < ns; ++i) acc += (u128)xs[i] * hatmod[i * nd + t];
      out[(size_t)t * N + k] = (u64)(acc % pt);
    }
  }
  free(hatinv);
  free(hatmod);
}

/* ModUp of digit j of an evaluation-form polynomial with l active Q limbs
 * into the extended basis Q_l u P (l + K limbs, evaluation form).
 * Digit limbs are copied unchanged; the others are INTT -> basis switch
 * -> NTT (reference: src/CKKS/LimbPartitionMGPU.cu:849-876 INTT of the
 * digit, :1168-1208 DecompAndModUpConv, :1212-1228 / :1580-1605 NTT of
 * the outputs; src/CKKS/ElemenwiseBatchKernels.cu:246-255 uses the
 * original limb for the digit's own primes). */
OR_API void or_modup_digit(const or_ctx* c, const u64* in, uint32_t l, uint32_t j, u64* out) {
  const uint32_t N = c->N;
  uint32_t st, sz;
  digit_range(c, l, j, &st, &sz);
  uint32_t src[64], dst[128], dpos[128];
  uint32_t nd = 0;
  for (uint32_t i = 0; i < sz; ++i) src[i] = st + i;
  for (uint32_t i = 0; i < l; ++i)
    if (i < st || i >= st + sz) { dst[nd] = i; dpos[nd] = i; nd++; }
  for (uint32_t k = 0; k < c->K; ++k) { dst[nd] = c->L + k; dpos[nd] = l + k; nd++; }
  u64* coef = (u64*)malloc(sizeof(u64) * sz * N);
  memcpy(coef, in + (size_t)st * N, sizeof(u64) * sz * N);
  or_intt_limbs(c, coef, src, sz);
  u64* conv = (u64*)malloc(sizeof(u64) * nd * N);
  or_switch_basis(c, coef, src, sz, conv, dst, nd);
  or_ntt_limbs(c, conv, dst, nd);
  for (uint32_t t = 0; t < nd; ++t) memcpy(out + (size_t)dpos[t] * N, conv + (size_t)t * N, sizeof(u64) * N);
  for (uint32_t i = 0; i < sz; ++i) memcpy(out + (size_t)(st + i) * N, in + (size_t)(st + i) * N, sizeof(u64) * N);
  free(coef);
  free(conv);
}

/* ------------------------------------------------------------------ */
/* ModDown                                                             */
/* ------------------------------------------------------------------ */

/* ModDown of an evaluation-form polynomial over Q_l u P (l + K limbs) to
 * Q_l:  a_Q <- (a_Q - NTT(Conv_{P->Q}(INTT(a_P)))) * P^-1  mod q_i
 * (reference: src/CKKS/Conv.cu:26-125 ModDown2 with PHatInvModp /
 *  PHatModq, src/CKKS/NTTfusions.cuh:237-252 moddown epilogue with
 *  PInvModq, driver src/CKKS/LimbPartition.cu:1913-1990; OpenFHE
 *  ApproxModDown). */
OR_API void or_moddown(const or_ctx* c, const u64* in, uint32_t l, u64* out) {
  const uint32_t N = c->N, K = c->K;
  uint32_t src[64], dst[64];
  for (uint32_t k = 0; k < K; ++k) src[k] = c->L + k;
  for (uint32_t i = 0; i < l; ++i) dst[i] = i;
  u64* pc = (u64*)malloc(sizeof(u64) * K * N);
  memcpy(pc, in + (size_t)l * N, sizeof(u64) * K * N);
  or_intt_limbs(c, pc, src, K);
  u64* conv = (u64*)malloc(sizeof(u64) * l * N);
  or_switch_basis(c, pc, src, K, conv, dst, l);
  or_ntt_limbs(c, conv, dst, l);
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < l; ++i) {
    const u64 qi = c->primes[i];
    u64 pinv = 1;
    for (uint32_t k = 0; k < K; ++k) pinv = mulmod(pinv, c->primes[c->L + k] % qi, qi);
    pinv = or_invmod(pinv, qi);
    for (uint32_t k = 0; k < N; ++k) {
      size_t o = (size_t)i * N + k;
      out[o] = mulmod(submod(in[o], conv[o], qi), pinv, qi);
    }
  }
  free(pc);
  free(conv);
}

/* ------------------------------------------------------------------ */
/* Key switching                                                       */
/* ------------------------------------------------------------------ */

/* Hybrid key switch of one evaluation-form polynomial d (l limbs):
 *   for each digit j < beta: e_j = ModUp_j(d)                       (l+K limbs)
 *   (t0, t1) = sum_j e_j * (b_j, a_j)   over Q_l u P
 *   out0 = ModDown(t0), out1 = ModDown(t1)
 * key: [2][dnum][L+K][N], component 0 = a, 1 = b; Q_l limb i of the
 * key is limb i, special limb k is limb L+k.
 * (reference: src/CKKS/LimbPartitionMGPU.cu:588-1873
 *  modup_ksk_moddown_mgpu; inner product
 *  src/CKKS/ElemenwiseBatchKernels.cu:202-273 fusedDotKSK_2_ computing
 *  out1 = sum digit*ksk_a, out2 = sum digit*ksk_b; OpenFHE
 *  EvalFastKeySwitchCore + KeySwitchDown.) */
OR_API void or_keyswitch(const or_ctx* c, const u64* d, uint32_t l, const u64* key, u64* out0, u64* out1) {
  const uint32_t N = c->N, K = c->K, LK = c->L + c->K;
  const uint32_t beta = or_num_digits(c, l);
  const uint32_t ne = l + K;
  u64* ext = (u64*)malloc(sizeof(u64) * ne * N);
  u64* t0 = (u64*)calloc((size_t)ne * N, sizeof(u64));
  u64* t1 = (u64*)calloc((size_t)ne * N, sizeof(u64));
  for (uint32_t j = 0; j < beta; ++j) {
    or_modup_digit(c, d, l, j, ext);
    const u64* ka = key + ((size_t)0 * c->dnum + j) * LK * N;
    const u64* kb = key + ((size_t)1 * c->dnum + j) * LK * N;
#pragma omp parallel for schedule(static)
    for (uint32_t t = 0; t < ne; ++t) {
      const uint32_t pid = t < l ? t : c->L + (t - l);
      const u64 q = c->primes[pid];
      const u64* ea = ka + (size_t)pid * N;
      const u64* eb = kb + (size_t)pid * N;
      const u64* x = ext + (size_t)t * N;
      for (uint32_t k = 0; k < N; ++k) {
        size_t o = (size_t)t * N + k;
        t0[o] = addmod(t0[o], mulmod(x[k], eb[k], q), q);
        t1[o] = addmod(t1[o], mulmod(x[k], ea[k], q), q);
      }
    }
  }
  or_moddown(c, t0, l, out0);
  or_moddown(c, t1, l, out1);
  free(ext);
  free(t0);
  free(t1);
}

/* ------------------------------------------------------------------ */
/* Tensor product, rescale, automorphism                               */
/* ------------------------------------------------------------------ */

/* (c0,c1) x (d0,d1) -> (c0 d0, c0 d1 + c1 d0, c1 d1), l limbs each,
 * evaluation form (reference src/CKKS/ElemenwiseBatchKernels.cu:1363-1400
 * binomialMult_). Inputs [2][l][N], output [3][l][N]. */
OR_API void or_tensor(const or_ctx* c, const u64* a, const u64* b, uint32_t l, u64* out) {
  const uint32_t N = c->N;
  const size_t S = (size_t)l * N;
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < l; ++i) {
    const u64 q = c->primes[i];
    for (uint32_t k = 0; k < N; ++k) {
      size_t o = (size_t)i * N + k;
      u64 a0 = a[o], a1 = a[S + o], b0 = b[o], b1 = b[S + o];
      out[o] = mulmod(a0, b0, q);
      out[S + o] = addmod(mulmod(a0, b1, q), mulmod(a1, b0, q), q);
      out[2 * S + o] = mulmod(a1, b1, q);
    }
  }
}

/* OpenFHE NativeVector::SwitchModulus of one value from old modulus om to
 * new modulus nm (centred lift), as quoted at reference
 * src/CKKS/Rescale.cuh:14-40 and implemented :47-68. */
OR_API u64 or_switch_modulus(u64 a, u64 om, u64 nm) {
  const u64 halfQ = om >> 1;
  if (nm > om) {
    if (a > halfQ) a = a + (nm - om);
  } else {
    const u64 diff = nm - (om % nm);
    if (a > halfQ) a = a + diff;
    if (a >= nm) a = a % nm;
  }
  return a;
}

/* QlQlInvModqlDivqlModq for dropping q_{l-1}: for i < l-1,
 *   floor(Q' * [Q'^-1]_{q_{l-1}} / q_{l-1}) mod q_i,  Q' = prod_{i<l-1} q_i.
 * Computed exactly without big integers: u = Q' * [Q'^-1]_{q_{l-1}} is
 * congruent to 1 mod q_{l-1}, so floor(u / q_{l-1}) = (u - 1) / q_{l-1}, and
 * since q_i divides u this is -q_{l-1}^-1 mod q_i.  The tests compare this
 * against the big-integer definition. */
OR_API void or_rescale_table(const or_ctx* c, uint32_t l, u64* out) {
  const u64 qt = c->primes[l - 1];
  for (uint32_t i = 0; i + 1 < l; ++i) {
    const u64 qi = c->primes[i];
    out[i] = (qi - or_invmod(qt % qi, qi)) % qi;
  }
}

/* Rescale of an evaluation-form polynomial with l limbs (drop q_{l-1}):
 *   t = INTT(a_{l-1});  a_i <- a_i * q_{l-1}^-1 + NTT_i(SwitchModulus(t)) * QlQlInvModqlDivqlModq_i
 * with QlQlInvModqlDivqlModq_i = floor(Q' * [Q'^-1]_{q_{l-1}} / q_{l-1}) mod q_i,
 * Q' = prod_{i<l-1} q_i (reference: src/CKKS/LimbPartitionMGPU.cu:196-394
 * doubleRescaleMGPU, src/CKKS/NTTfusions.cuh:214-235 rescale_fusion,
 * table formula quoted at src/CKKS/openfhe-interface/RawCiphertext.cu:311-330;
 * OpenFHE DCRTPoly::DropLastElementAndScale).  In place; limb l-1 is left
 * untouched and becomes unused. */
OR_API void or_rescale_poly(const or_ctx* c, u64* a, uint32_t l) {
  const uint32_t N = c->N;
  const uint32_t top = l - 1;
  const u64 qt = c->primes[top];
  u64* t = (u64*)malloc(sizeof(u64) * N);
  memcpy(t, a + (size_t)top * N, sizeof(u64) * N);
  or_intt(t, c->logN, qt, c->ipsi_rev + (size_t)top * N);
  u64 big_tab[64];
  or_rescale_table(c, l, big_tab);
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < top; ++i) {
    const u64 qi = c->primes[i];
    const u64 qtinv = or_invmod(qt % qi, qi);
    const u64 big = big_tab[i]; /* QlQlInvModqlDivqlModq_i */
    u64* s = (u64*)malloc(sizeof(u64) * N);
    for (uint32_t k = 0; k < N; ++k) s[k] = or_switch_modulus(t[k], qt, qi);
    or_ntt(s, c->logN, qi, c->psi_rev + (size_t)i * N);
    u64* ai = a + (size_t)i * N;
    for (uint32_t k = 0; k < N; ++k) ai[k] = addmod(mulmod(ai[k], qtinv, qi), mulmod(s[k], big, qi), qi);
    free(s);
  }
  free(t);
}

/* ModRaise of a one-limb ciphertext [2][1][N] (modulus q_0, evaluation
 * form) to l_out limbs: INTT under q_0, SwitchModulus (centred lift) of each
 * coefficient to q_i, NTT under q_i (reference ModRaise,
 * src/CKKS/Bootstrap.cu:393-560, which raises c0/c1 the same way). */
OR_API void or_mod_raise(const or_ctx* c, const u64* in, uint32_t l_out, u64* out) {
  const uint32_t N = c->N;
  const u64 q0 = c->primes[0];
  for (uint32_t p = 0; p < 2; ++p) {
    u64* t = (u64*)malloc(sizeof(u64) * N);
    memcpy(t, in + (size_t)p * N, sizeof(u64) * N);
    or_intt(t, c->logN, q0, c->ipsi_rev);
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < l_out; ++i) {
      const u64 qi = c->primes[i];
      u64* o = out + ((size_t)p * l_out + i) * N;
      for (uint32_t k = 0; k < N; ++k) o[k] = or_switch_modulus(t[k], q0, qi);
      or_ntt(o, c->logN, qi, c->psi_rev + (size_t)i * N);
    }
    free(t);
  }
}

/* Ciphertext [2][l][N] -> [2][l-1][N] (packed), both components. */
OR_API void or_rescale_ct(const or_ctx* c, const u64* in, uint32_t l, u64* out) {
  const size_t N = c->N;
  u64* tmp = (u64*)malloc(sizeof(u64) * 2 * l * N);
  memcpy(tmp, in, sizeof(u64) * 2 * l * N);
  or_rescale_poly(c, tmp, l);
  or_rescale_poly(c, tmp + l * N, l);
  memcpy(out, tmp, sizeof(u64) * (l - 1) * N);
  memcpy(out + (l - 1) * N, tmp + l * N, sizeof(u64) * (l - 1) * N);
  free(tmp);
}

/* Evaluation-domain automorphism X -> X^k (k odd):
 *   out[brev(j)] = in[brev(((2j+1) k mod 2N) >> 1)]
 * a gather; the reference's scatter automorph_slot with k^-1
 * (src/Rotation.cuh:14-26, src/CKKS/RNSPoly.cpp:632-634) is the same
 * permutation (OpenFHE PrecomputeAutoMap + AutomorphismTransform). */
OR_API void or_automorph(const or_ctx* c, const u64* in, u64* out, uint32_t nlimbs, uint32_t k) {
  const uint32_t N = c->N, logN = c->logN;
  const uint32_t mask = 2 * N - 1;
  uint32_t* map = (uint32_t*)malloc(sizeof(uint32_t) * N);
  for (uint32_t j = 0; j < N; ++j) {
    uint32_t src = (uint32_t)((((u64)(2 * j + 1) * k) & mask) >> 1);
    map[or_brev(j, logN)] = or_brev(src, logN);
  }
  for (uint32_t t = 0; t < nlimbs; ++t)
    for (uint32_t j = 0; j < N; ++j) out[(size_t)t * N + j] = in[(size_t)t * N + map[j]];
  free(map);
}

/* ------------------------------------------------------------------ */
/* Ciphertext-level operations                                         */
/* ------------------------------------------------------------------ */

/* HMult + relinearize (no rescale), FIXEDMANUAL: tensor, key switch of the
 * c1*d1 term with the relinearization key, add (reference
 * src/CKKS/Ciphertext.cpp:652-674 mult active branch; OpenFHE EvalMult ->
 * KeySwitch).  a, b, out: [2][l][N]. */
OR_API void or_hmult_relin(const or_ctx* c, const u64* a, const u64* b, uint32_t l, const u64* rlk, u64* out) {
  const size_t N = c->N, S = (size_t)l * N;
  u64* t = (u64*)malloc(sizeof(u64) * 3 * S);
  or_tensor(c, a, b, l, t);
  u64* k0 = (u64*)malloc(sizeof(u64) * S);
  u64* k1 = (u64*)malloc(sizeof(u64) * S);
  or_keyswitch(c, t + 2 * S, l, rlk, k0, k1);
  for (uint32_t i = 0; i < l; ++i) {
    const u64 q = c->primes[i];
    for (size_t k = 0; k < N; ++k) {
      size_t o = i * N + k;
      out[o] = addmod(t[o], k0[o], q);
      out[S + o] = addmod(t[S + o], k1[o], q);
    }
  }
  free(t);
  free(k0);
  free(k1);
}

/* Rotate / automorphism with key for index k (OpenFHE EvalAutomorphism
 * order: key switch of c1 including ModDown, add c0, then permute both
 * components; the reference's fused hoisted path permutes before ModDown,
 * src/CKKS/Ciphertext.cpp:825-890, which is not bit-identical to this
 * order).  in, out: [2][l][N]. */
OR_API void or_rotate(const or_ctx* c, const u64* in, uint32_t l, const u64* key, uint32_t k, u64* out) {
  const size_t N = c->N, S = (size_t)l * N;
  u64* k0 = (u64*)malloc(sizeof(u64) * S);
  u64* k1 = (u64*)malloc(sizeof(u64) * S);
  or_keyswitch(c, in + S, l, key, k0, k1);
  for (uint32_t i = 0; i < l; ++i) {
    const u64 q = c->primes[i];
    for (size_t x = 0; x < N; ++x) k0[i * N + x] = addmod(k0[i * N + x], in[i * N + x], q);
  }
  or_automorph(c, k0, out, l, k);
  or_automorph(c, k1, out + S, l, k);
  free(k0);
  free(k1);
}

/* ------------------------------------------------------------------ */
/* Key generation helpers (test-side; OpenFHE KeySwitchHYBRID::         */
/* KeySwitchGenInternal structure).                                    */
/* ------------------------------------------------------------------ */

/* b_j = -a_j * s_new + e_j + [i in digit j] * (P mod q_i) * s_old  over Q u P.
 * s_new, s_old: evaluation form over all L+K primes ([L+K][N]);
 * a: [dnum][L+K][N] uniform; e: [dnum][L+K][N] evaluation form.
 * key out: [2][dnum][L+K][N] (a, b). */
OR_API void or_keygen_switch(const or_ctx* c, const u64* s_new, const u64* s_old, const u64* a, const u64* e,
                             u64* key) {
  const uint32_t N = c->N, LK = c->L + c->K;
  for (uint32_t j = 0; j < c->dnum; ++j) {
    uint32_t st = j * c->alpha, en = (j + 1) * c->alpha;
    if (en > c->L) en = c->L;
#pragma omp parallel for schedule(static)
    for (uint32_t i = 0; i < LK; ++i) {
      const u64 q = c->primes[i];
      u64 pmod = 1;
      for (uint32_t k = 0; k < c->K; ++k) pmod = mulmod(pmod, c->primes[c->L + k] % q, q);
      const int in_digit = (i >= st && i < en);
      const size_t off = ((size_t)j * LK + i) * N;
      u64* ka = key + off;
      u64* kb = key + (size_t)c->dnum * LK * N + off;
      for (uint32_t k = 0; k < N; ++k) {
        u64 av = a[off + k];
        u64 v = submod(e[off + k], mulmod(av, s_new[(size_t)i * N + k], q), q);
        if (in_digit) v = addmod(v, mulmod(pmod, s_old[(size_t)i * N + k], q), q);
        ka[k] = av;
        kb[k] = v;
      }
    }
  }
}

/* Element-wise helpers over limbs with prime ids (used by tests to build
 * encryptions/decryptions). out = x*y, out = x+y, out = x-y. */
OR_API void or_mul_limbs(const or_ctx* c, const u64* x, const u64* y, u64* out, const uint32_t* pids, uint32_t n) {
  for (uint32_t t = 0; t < n; ++t) {
    const u64 q = c->primes[pids[t]];
    for (uint32_t k = 0; k < c->N; ++k) out[(size_t)t * c->N + k] = mulmod(x[(size_t)t * c->N + k], y[(size_t)t * c->N + k], q);
  }
}
OR_API void or_add_limbs(const or_ctx* c, const u64* x, const u64* y, u64* out, const uint32_t* pids, uint32_t n) {
  for (uint32_t t = 0; t < n; ++t) {
    const u64 q = c->primes[pids[t]];
    for (uint32_t k = 0; k < c->N; ++k) out[(size_t)t * c->N + k] = addmod(x[(size_t)t * c->N + k], y[(size_t)t * c->N + k], q);
  }
}
OR_API void or_sub_limbs(const or_ctx* c, const u64* x, const u64* y, u64* out, const uint32_t* pids, uint32_t n) {
  for (uint32_t t = 0; t < n; ++t) {
    const u64 q = c->primes[pids[t]];
    for (uint32_t k = 0; k < c->N; ++k) out[(size_t)t * c->N + k] = submod(x[(size_t)t * c->N + k], y[(size_t)t * c->N + k], q);
  }
}

/* Worker threads of the parallel loops (timing of the CPU baseline). */
OR_API void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
OR_API int or_get_threads(void) { return omp_get_max_threads(); }
