// Negacyclic NTT / INTT over 64-bit RNS limbs for gfx950.
//
// Transform (OpenFHE evaluation order, reference src/NTT.cu:331-695 forward,
// :42-310 inverse; tables src/ConstantsGPU.cu:226-300):
//   NTT(a)[k] = sum_i a_i psi^(i (2 brev(k) + 1))  mod q
// computed by the Cooley-Tukey loop whose stage s (m = 2^s groups) uses the
// bit-reversed twiddle psi_rev[m + group]; the inverse runs the
// Gentleman-Sande loop with psi^-1 and folds N^-1 into its last stage.
//
// MI355X design: the N = 2^LOGN limb is viewed as a 2^L1 x 2^L2 matrix
// (row-major, L1 = LOGN/2).  The first L1 stages only pair elements of the
// same column, the last L2 stages only elements of the same row, so the
// transform is two launches, each holding whole lines (columns or rows) of
// one limb in LDS:
//   column pass: a workgroup owns 16 adjacent columns (128-byte row segments,
//                coalesced), all 2^L1 rows.
//   row pass:    a workgroup owns 256/TPL whole rows (contiguous).
// Inside a pass each thread keeps E = 2^A elements in VGPRs and runs A
// stages as a radix-2^A butterfly network, exchanges through LDS once, then
// runs the remaining B = LS - A stages on 2^(A-B) groups of 2^B elements.
// Butterflies use Harvey's lazy reduction (values in [0, 4q) forward,
// [0, 2q) inverse) with Shoup twiddle products, so a butterfly costs one
// modular multiply.  Limbs (any mix of primes, any number of polynomials)
// are the grid's y dimension, described by a LimbSpan.
#pragma once
#include "modarith.h"

// One limb of NTT work.  Device task lists carry everything a launch needs
// per limb, so one launch can cover any mix of ciphertexts, levels and
// primes (the cross-ciphertext batch).
struct NttTask {
  u64* data;       // limb transformed in place; final destination unless `out` is set
  const u64* src;  // input of the first pass (null: data).  PRO_SWITCH_MODULUS:
                   // coefficient-form source limb under prime src_pid
  const u64* x;    // epilogue operand
  const u64* acc;  // epilogue accumulator (null: none)
  u64* out;        // epilogue destination (null: data)
  const u64* src2; // PRO_MDRS: second prologue source (coefficient form, prime pid)
  u64 c0, c0sh, c1, c1sh;  // epilogue constants of this limb (with Shoup companions)
  u64 c2, c2sh, c3, c3sh;  // further constants of the fused ModDown+Rescale modes
  u32 pid;         // prime of the transform
  u32 src_pid;
  u32 perm;        // EPI_*: automorphism index applied to the store (0: none)
  u32 pad;
  u64 swdiff;      // SwitchModulus constant: q_pid - (q_src mod q_pid) when q_pid < q_src
};

// A set of limbs given inline as up to 8 segments of consecutive limbs
// (stride N) with consecutive prime ids, or a device task list.
struct LimbSpan {
  u64* seg_ptr[8];
  u32 seg_pid[8];
  u32 seg_end[8];  // cumulative limb count at the end of each segment
  u32 nseg;
  u32 total;
  const NttTask* list;  // if non-null, used instead of the segments
};

__device__ __forceinline__ void span_get(const LimbSpan& s, u32 y, u32 logN, u64*& ptr, u32& pid) {
  u32 g = 0, start = 0;
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    if (g + 1 < s.nseg && y >= s.seg_end[g]) {
      start = s.seg_end[g];
      ++g;
    }
  }
  ptr = s.seg_ptr[g] ? s.seg_ptr[g] + ((u64)(y - start) << logN) : nullptr;
  pid = s.seg_pid[g] + (y - start);
}

struct NttTables {
  const u64* psi;      // [P][N] psi^brev(j)
  const u64* psi_sh;   // Shoup companions
  const u64* ipsi;     // [P][N] psi^-brev(j)
  const u64* ipsi_sh;
  const PrimeC* prm;   // [P]
  // every prime of the context below 2^60: the forward butterflies may let
  // values grow to 16q between reductions (host-side launch selection)
  u32 narrow;
};

// Fused prologue of the forward column pass.
enum NttPro : int {
  PRO_NONE = 0,
  // input element = SwitchModulus(task.src[i], q_src_pid -> q_pid)
  // (OpenFHE NativeVector::SwitchModulus, reference src/CKKS/Rescale.cuh:47-68;
  //  fused as in src/CKKS/LimbPartitionMGPU.cu:339-373)
  PRO_SWITCH_MODULUS = 1,
  // fused ModDown + Rescale: input = SwitchModulus(src, q_src_pid -> q_pid) * c2
  //                                  - src2 * c3   (see EPI_MDRS)
  PRO_MDRS = 2,
};

// Fused epilogue of the forward row pass (final store).
enum NttEpi : int {
  EPI_NONE = 0,
  // out = (x - ntt) * P^-1 (+ acc)   ModDown, reference src/CKKS/NTTfusions.cuh:237-252
  EPI_MODDOWN = 1,
  EPI_MODDOWN_ADD = 2,
  // out = x * q_l^-1 + ntt * QlQlInvModqlDivqlModq   Rescale, src/CKKS/NTTfusions.cuh:214-235
  EPI_RESCALE = 3,
  // fused ModDown + Rescale: out = acc * c0 + x * c3 + ntt
  EPI_MDRS = 4,
};

// Inline (non-list) description of a fused launch: the epilogue operand,
// accumulator and destination as spans over the same limb order, and
// per-prime constant tables; a single prologue source limb.
struct NttFuse {
  const u64* pro_src;  // PRO_SWITCH_MODULUS: coefficient-form source limb
  u32 pro_src_pid;     //   and its prime id
  LimbSpan x;          // EPI_*: the x operand
  LimbSpan acc;        // EPI_MODDOWN_ADD: accumulator (segments may be null)
  LimbSpan out;        // EPI_*: destination (no segments: in place)
  const u64* c0;       // per-pid constant tables and Shoup companions
  const u64* c0sh;
  const u64* c1;
  const u64* c1sh;
};

// Resolve the task of workgroup row y from a list or from inline spans.
__device__ __forceinline__ NttTask resolve_task(const LimbSpan& s, const NttFuse& F, u32 y, u32 logN) {
  if (s.list) return s.list[y];
  NttTask t;
  span_get(s, y, logN, t.data, t.pid);
  t.src = F.pro_src;
  t.src_pid = F.pro_src_pid;
  t.x = nullptr;
  t.acc = nullptr;
  t.out = nullptr;
  t.src2 = nullptr;
  t.c0 = t.c0sh = t.c1 = t.c1sh = 0;
  t.c2 = t.c2sh = t.c3 = t.c3sh = 0;
  t.perm = 0;
  t.swdiff = 0;
  if (F.x.nseg) {
    u64* p;
    u32 q;
    span_get(F.x, y, logN, p, q);
    t.x = p;
    if (F.acc.nseg) {
      span_get(F.acc, y, logN, p, q);
      t.acc = p;
    }
    if (F.out.nseg) {
      span_get(F.out, y, logN, p, q);
      t.out = p;
    }
    t.c0 = F.c0[t.pid];
    t.c0sh = F.c0sh[t.pid];
    if (F.c1) {
      t.c1 = F.c1[t.pid];
      t.c1sh = F.c1sh[t.pid];
    }
  }
  return t;
}
