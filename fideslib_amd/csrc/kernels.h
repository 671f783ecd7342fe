// Shared declarations of the element-wise / conversion kernels.
#pragma once
#include "ntt.h"

constexpr int CONV_MAXS = 16;   // max source limbs of one conversion (digit size alpha, or K)
constexpr int CONV_MAXT = 96;   // max target limbs of one conversion
constexpr int CONV_TCHUNK = 8;  // targets per workgroup row

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

struct KsDotArgs {
  const u64* d;      // [l][N]  evaluation-form input (digit limbs)
  const u64* ext;    // [beta][ext_stride][N]  ModUp outputs per digit, limb t of Q_l u P
  const u64* key_a;  // [dnum][key_stride][N] Montgomery form
  const u64* key_b;
  u64* o0;           // [l+K][N]
  u64* o1;
  const PrimeC* prm;
  u32 logN, l, K, L, beta, alpha, ext_stride, key_stride;
};

void fides_launch_tensor(const u64* a0, const u64* a1, const u64* b0, const u64* b1, u64* o0, u64* o1, u64* o2,
                         const PrimeC* prm, u32 logN, u32 l, bool square, hipStream_t st);
int fides_launch_conv(const u64* in, u64* out, const ConvDesc* d, u32 ns, u32 nt, const PrimeC* prm, u32 logN,
                      hipStream_t st);
void fides_launch_ksk_dot(const KsDotArgs& A, hipStream_t st);
void fides_launch_to_mont(u64* key, const PrimeC* prm, u32 logN, u32 nlimbs, u32 LK, hipStream_t st);
void fides_launch_automorph(const u64* in, u64* out, u32 logN, u32 nlimbs, u32 kidx, hipStream_t st);
void fides_launch_add(const u64* a, const u64* b, u64* o, const PrimeC* prm, u32 logN, u32 nlimbs, u32 pid0,
                      hipStream_t st);
int fides_launch_ntt(int logN, const LimbSpan& s, const NttTables& T, int pro, int epi, const NttFuse& F,
                     hipStream_t st);
int fides_launch_intt(int logN, const LimbSpan& s, const NttTables& T, hipStream_t st);
