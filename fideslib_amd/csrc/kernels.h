// Shared declarations of the element-wise / conversion kernels.
#pragma once
#include "ntt.h"

constexpr int CONV_MAXS = 16;   // max source limbs of one conversion (digit size alpha, or K)
constexpr int CONV_MAXT = 96;   // max target limbs of one conversion
constexpr int CONV_TCHUNK = 8;  // targets per workgroup row of a single conversion (key-switch batches take 8-32)

// Tables of one approximate CRT basis switch (ModUp of one digit at one
// level, or ModDown at one level): reference PartQlHatInvModq /
// PartQlHatModp (src/ConstantsGPU.cu:528-621) and PHatInvModp / PHatModq
// (:469-526).
struct ConvDesc {
  u32 ns, nt;
  u32 src_pid[CONV_MAXS];
  u64 hatinv[CONV_MAXS];     // (Qhat / q_i)^-1 mod q_i
  u64 hatinv_sh[CONV_MAXS];  // Shoup companion
  u32 dst_pid[CONV_MAXT];
  u32 dst_off[CONV_MAXT];    // limb offset of target t in the output buffer
  u64 mat[CONV_MAXT][CONV_MAXS];  // (Qhat / q_i) mod p_t, times 2^64 mod p_t
};

// ---- task records of the batched kernels (one record per workgroup row)

// One (job, digit or polynomial, target chunk) of a basis conversion.
struct ConvTask {
  const u64* in;         // ns contiguous source limbs (coefficient form)
  u64* out;              // target t goes to out + desc->dst_off[t] * N
  const ConvDesc* desc;
  u32 t0, t1;            // target range of this workgroup row
};

// One output limb t of the key-switch inner product (for one job).
struct DotTask {
  const u64* d;     // original evaluation-form limb t (used by the digit owning t), or null
  const u64* e;     // ModUp output of digit 0 at limb t; digit j at e + j * estride * N
  const u64* ka;    // key a, digit 0, this limb's prime; digit j at + j * kstride * N
  const u64* kb;    // key b, likewise
  u64* o0;          // sum_j e_j * b_j
  u64* o1;          // sum_j e_j * a_j
  u32 beta;         // active digits
  u32 jown;         // digit owning limb t (Q limbs), 0xffffffff for special limbs
  u32 pid;
  u32 estride;      // limbs between consecutive digits' ModUp outputs
  u32 ext;          // 1: the result stays in Q u P (no ModDown follows): a special
                    // limb is stored in evaluation form, not row-inverted
  u32 pad;
};

// One limb of the tensor product (a0,a1) x (b0,b1).
struct TensorTask {
  const u64 *a0, *a1, *b0, *b1;
  u64 *o0, *o1, *o2;
  u32 pid;
  u32 square;
};

// One output limb of sum_j x_j * m_j (mod q_pid), m_j in Montgomery form
// (plaintext limbs prepared by fides_plaintext_prepare); terms[first ..
// first + cnt) hold the operand limbs, cnt <= 16.
// One term of a plaintext dot product at one limb: both ciphertext
// components' limbs and the plaintext limb they are multiplied by.
struct PtTerm {
  const u64* x0;
  const u64* x1;
  const u64* m;
};
// One limb of one output: out0/out1 = sum of the cnt terms from `first`.
struct PtDotTask {
  u64* out0;
  u64* out1;
  u32 pid;
  u32 cnt;
  u32 first;
  u32 pad;
};
constexpr u32 PTDOT_MAXT = 16;

// One output limb of sum_j m_j x_j + add (mod q_pid) with integer constants
// m_j < q given in Montgomery form (m_j 2^64 mod q); cnt <= 64.
struct LcTerm {
  const u64* x;
  u64 m;
};
struct LcTask {
  u64* out;
  u64 add;
  u32 pid;
  u32 cnt;
  u32 first;
  u32 pad;
};
constexpr u32 LINCOMB_MAXT = 64;

// One limb of out = (a * m or a) (+ or - b) + add  (mod q_pid); with
// LIN_AMONT, a is read in Montgomery form (a R mod q) and converted first.
struct LinTask {
  const u64* a;
  const u64* b;   // used when flags & LIN_B
  u64* out;
  u64 m, msh;     // used when flags & LIN_MUL
  u64 add;        // added constant residue (< q)
  u32 pid;
  u32 flags;
};
constexpr u32 LIN_B = 1, LIN_SUB = 2, LIN_MUL = 4, LIN_AMONT = 8;

// One limb of an evaluation-domain automorphism, optionally with a scaled
// addend permuted alongside: out = sigma_k(in + add * c)  (mod q_pid).
struct AutoTask {
  const u64* in;
  u64* out;
  const u64* add;  // null: plain permutation
  u64 c, csh;      // constant of the addend and its Shoup companion
  u32 kidx;
  u32 pid;
};

int fides_launch_conv_batch(const ConvTask* tasks, u32 ntasks, u32 ns, bool split, const PrimeC* prm, u32 logN,
                            hipStream_t st);
// Tasks of different source counts (desc->ns, all <= maxns <= CONV_MAXS) in one launch.
int fides_launch_conv_mixed(const ConvTask* tasks, u32 ntasks, u32 maxns, bool split, const PrimeC* prm, u32 logN,
                            hipStream_t st);
void fides_launch_dot_batch(const DotTask* tasks, u32 ntasks, const PrimeC* prm, u32 logN, u32 kstride,
                            hipStream_t st);
void fides_launch_tensor_batch(const TensorTask* tasks, u32 ntasks, const PrimeC* prm, u32 logN, hipStream_t st);
void fides_launch_ptdot_batch(const PtDotTask* tasks, const PtTerm* terms, u32 ntasks, const PrimeC* prm, u32 logN,
                              hipStream_t st);
void fides_launch_lincomb_batch(const LcTask* tasks, const LcTerm* terms, u32 ntasks, const PrimeC* prm, u32 logN,
                                hipStream_t st);
void fides_launch_lin_batch(const LinTask* tasks, u32 ntasks, const PrimeC* prm, u32 logN, hipStream_t st);
void fides_launch_automorph_batch(const AutoTask* tasks, u32 ntasks, const PrimeC* prm, u32 logN, hipStream_t st);
void fides_launch_to_mont(u64* key, const PrimeC* prm, u32 logN, u32 nlimbs, u32 LK, hipStream_t st);
int fides_launch_ntt(int logN, const LimbSpan& s, const NttTables& T, int pro, int epi, const NttFuse& F,
                     hipStream_t st);
int fides_launch_intt(int logN, const LimbSpan& s, const NttTables& T, hipStream_t st);
// INTT of task-list limbs with an element-wise prologue and epilogue:
// data = INTT(src + x c0) + src2 c1 (mod q), the top limb of a rescale fused
// into ModDown.
int fides_launch_intt_axpy(int logN, const LimbSpan& s, const NttTables& T, hipStream_t st);
// Column pass of the inverse NTT alone (after k_ntt_row_dot ran the row
// stages of the special limbs).
int fides_launch_intt_col(int logN, const LimbSpan& s, const NttTables& T, hipStream_t st);
// Column pass only (the row pass is fused into fides_launch_row_dot).
int fides_launch_ntt_col(int logN, const LimbSpan& s, const NttTables& T, int pro, const NttFuse& F,
                         hipStream_t st);
// Row pass of the ModUp outputs fused with the key-switch inner product; all
// tasks of one launch have `beta` digits (1..ROWDOT_MAXB).
constexpr int ROWDOT_MAXB = 4;
int fides_launch_row_dot(int logN, const DotTask* tasks, u32 n, int beta, const NttTables& T, u32 kstride,
                         hipStream_t st);
// Tasks with 1..maxb digits (2 <= maxb <= ROWDOT_MAXB) in one launch; logN
// 16 and 17 only (returns 1 for the others).
int fides_launch_row_dot_mixed(int logN, const DotTask* tasks, u32 n, int maxb, const NttTables& T, u32 kstride,
                               hipStream_t st);
