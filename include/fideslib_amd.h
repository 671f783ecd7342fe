/*
 * fideslib_amd.h -- C ABI of the MI355X (gfx950) CKKS key-switching core.
 *
 * The reference (FIDESlib) has no C ABI: its GPU seam is the C++ classes
 * FIDESlib::CKKS::{Context, Ciphertext, KeySwitchingKey} fed by the
 * interchange structs RawParams / RawCipherText / RawKeySwitchKey
 * (src/CKKS/openfhe-interface/RawCiphertext.cuh:19-93).  This header is that
 * seam as a C ABI: plain pointers and sizes, integer status returns, an
 * explicit HIP stream on every operation, no exceptions and no C++ or torch
 * types.  The host C++ API (fideslib::CryptoContext, ...) sits above it.
 *
 * Data formats
 *   limb        N = 2^logN contiguous uint64 residues in [0, q), device memory
 *   evaluation  OpenFHE bit-reversed NTT order:
 *               NTT(a)[k] = sum_i a_i psi^(i (2 brev(k) + 1)) mod q
 *               (reference src/CKKS/openfhe-interface/RawCiphertext.cu:47-73,
 *                REV = 1 default = no reordering)
 *   ciphertext  [2][l][N] uint64, evaluation form, l active Q limbs, limb i
 *               under Q prime i (FIDESlib level convention l = level + 1,
 *               api/Ciphertext.cpp:56-68)
 *   prime ids   0..L-1 = Q primes, L..L+K-1 = special (P) primes
 *   key         [2][dnum][L+K][N] uint64 host memory, component 0 = a,
 *               component 1 = b (RawKeySwitchKey r_key = {A, B},
 *               RawCiphertext.cu:649-668), evaluation form
 *
 * Threading: one context serves one stream at a time (its key-switch scratch
 * buffers are shared by all calls); calls on different streams with the
 * same context race.  Different contexts are independent.
 */
#ifndef FIDESLIB_AMD_H
#define FIDESLIB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (reference errors print and exit(0), src/CudaUtils.cuh;
 * here every failure is returned). */
#define FIDES_OK 0
#define FIDES_ERR_ARG (-1)         /* invalid argument / precondition */
#define FIDES_ERR_HIP (-2)         /* HIP runtime error */
#define FIDES_ERR_UNSUPPORTED (-3) /* parameter outside the compiled kernels */
#define FIDES_ERR_NOMEM (-4)       /* device allocation failed */

typedef struct fides_context fides_context;
typedef struct fides_key fides_key;
typedef struct fides_graph fides_graph;

/* Parameters (replaces RawParams, RawCiphertext.cuh:45-83, built by
 * GetRawParams RawCiphertext.cu:213-643 and Parameters::adaptTo
 * src/CKKS/Parameters.cu:113-140). */
typedef struct fides_params {
  uint32_t logN;       /* ring dimension N = 2^logN, 10 <= logN <= 17 */
  uint32_t L;          /* number of Q primes (max level + 1), <= 64 */
  uint32_t K;          /* number of special primes, 1..16 */
  uint32_t dnum;       /* number of digits; alpha = ceil(L / dnum) <= 16 */
  const uint64_t* q;   /* [L] Q primes, q = 1 mod 2N, q < 2^62 */
  const uint64_t* p;   /* [K] special primes, same conditions */
  const uint64_t* psi; /* [L+K] primitive 2N-th roots (RawParams
                          root_of_unity / SPECIALroot_of_unity), or NULL to
                          use g^((q-1)/2N) with g the smallest primitive root
                          (reference fallback src/ConstantsGPU.cu:36-67) */
} fides_params;

/* Last error message of the calling thread. */
const char* fides_last_error(void);

/* Context: tables, per-level conversion constants and key-switch scratch
 * (replaces GenCryptoContextGPU src/CKKS/Context.cu:972-990 +
 * SetupConstants src/ConstantsGPU.cu:125-686). */
int fides_context_create(const fides_params* params, fides_context** out);
void fides_context_destroy(fides_context* ctx);
int fides_context_info(const fides_context* ctx, uint32_t* logN, uint32_t* L, uint32_t* K, uint32_t* dnum,
                       uint32_t* alpha);
/* Roots actually used (psi, [L+K]) -- for callers that passed NULL. */
int fides_context_roots(const fides_context* ctx, uint64_t* psi_out);

/* Device memory and streams (stream = hipStream_t, NULL = default stream). */
int fides_device_alloc(size_t bytes, void** dptr);
int fides_device_free(void* dptr);
int fides_copy_to_device(void* dst, const void* src, size_t bytes, void* stream);
int fides_copy_to_host(void* dst, const void* src, size_t bytes, void* stream);
int fides_copy_device(void* dst, const void* src, size_t bytes, void* stream); /* async, stream-ordered */
int fides_stream_sync(void* stream);
/* A non-blocking stream at the default priority (graph capture needs a
 * stream other than the default one). */
int fides_stream_create(void** stream);
int fides_stream_destroy(void* stream);

/* ---- raw limb operations (config: NTT/INTT + ModUp/ModDown) ---------- */

/* In-place forward / inverse negacyclic NTT of nlimbs limbs at d_data,
 * limb t under prime id pids[t] (host array).  Replaces
 * LimbPartition::NTT / INTT launching NTT_ / INTT_ (src/NTT.cu:42-756). */
int fides_ntt(fides_context* ctx, uint64_t* d_data, const uint32_t* pids, uint32_t nlimbs, void* stream);
int fides_intt(fides_context* ctx, uint64_t* d_data, const uint32_t* pids, uint32_t nlimbs, void* stream);

/* ModUp of digit `digit` of an evaluation-form polynomial with l Q limbs
 * into Q_l u P: d_out receives l + K limbs (Q limbs 0..l-1, then P limbs),
 * evaluation form; the digit's own limbs are copied.  Replaces
 * RNSPoly::modup (src/CKKS/RNSPoly.cpp:283-322) ->
 * DecompAndModUpConv (src/CKKS/Conv.cu:258-402). */
int fides_modup(fides_context* ctx, const uint64_t* d_in, uint32_t l, uint32_t digit, uint64_t* d_out,
                void* stream);

/* ModDown of l + K evaluation-form limbs (Q_l u P) to l limbs.  Replaces
 * RNSPoly::moddown (src/CKKS/RNSPoly.cpp:584-624) -> ModDown2
 * (src/CKKS/Conv.cu:26-125) + NTT_MODDOWN epilogue
 * (src/CKKS/NTTfusions.cuh:237-252). */
int fides_moddown(fides_context* ctx, const uint64_t* d_in, uint32_t l, uint64_t* d_out, void* stream);

/* ---- keys ------------------------------------------------------------ */

/* Upload a key-switching key (relinearization, rotation or conjugation)
 * from host memory [2][dnum][L+K][N] (a, b).  Replaces
 * KeySwitchingKey::Initialize (src/CKKS/KeySwitchingKey.cu:20-35) fed by
 * GetEvalKeySwitchKey / GetRotationKeySwitchKey (RawCiphertext.cu:645-760). */
int fides_key_upload(fides_context* ctx, const uint64_t* h_key, fides_key** out);
/* Copies a key back to the host in its upload layout and form ([2][dnum][L+K][N],
 * standard residues): the inverse of fides_key_upload, for serialization
 * (reference CryptoContext::SerializeEvalMultKey, api/CryptoContext.hpp:90-95). */
int fides_key_download(const fides_key* key, uint64_t* h_key);
void fides_key_destroy(fides_key* key);

/* Hybrid key switch of one evaluation-form polynomial d (l limbs):
 * (out0, out1) with out0 + out1 * s_new = d * s_old + e.  Replaces
 * RNSPoly::modup_ksk_moddown_mgpu (src/CKKS/RNSPoly.cpp:1112-1244,
 * src/CKKS/LimbPartitionMGPU.cu:588-1873).  out0/out1: l limbs each. */
int fides_keyswitch(fides_context* ctx, const uint64_t* d_in, uint32_t l, const fides_key* key,
                    uint64_t* d_out0, uint64_t* d_out1, void* stream);

/* Allocation-free variant of fides_keyswitch, with the same result bits.  No
 * reference counterpart: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  Its scratch is
 * a file-scope device array of 160 MiB that every GPU running this library
 * holds until the process ends (a key switch needs
 * N (l + beta (l + K) + 2 (l + K) + 2 l + 2) words; larger requests return
 * FIDES_ERR_ARG), its task list goes through a task slot preallocated by
 * fides_context_create, and it calls no allocator.  All calls on one GPU share
 * that array: calls that overlap in time on one GPU (other streams, other
 * contexts, other threads) race and must be serialized by the caller. */
int fides_keyswitch_noalloc(fides_context* ctx, const uint64_t* d_in, uint32_t l, const fides_key* key,
                            uint64_t* d_out0, uint64_t* d_out1, void* stream);

/* ---- ciphertext operations ([2][l][N] device buffers) ---------------- */

/* HMult + relinearization (no rescale).  d_out may alias d_a.  Replaces
 * Ciphertext::mult (src/CKKS/Ciphertext.cpp:458-677) behind
 * CryptoContext::EvalMult (api/CryptoContext.cpp:1185-1210). */
int fides_hmult(fides_context* ctx, const uint64_t* d_a, const uint64_t* d_b, uint32_t l, const fides_key* rlk,
                uint64_t* d_out, void* stream);

/* Squaring + relinearization (Ciphertext::square, Ciphertext.cpp:679-760). */
int fides_square(fides_context* ctx, const uint64_t* d_a, uint32_t l, const fides_key* rlk, uint64_t* d_out,
                 void* stream);

/* Rescale: [2][l] -> [2][l-1] (OpenFHE DropLastElementAndScale).  d_out
 * must not alias d_in.  Replaces Ciphertext::rescale
 * (src/CKKS/Ciphertext.cpp:426-447) -> doubleRescaleMGPU
 * (src/CKKS/LimbPartitionMGPU.cu:196-394). */
int fides_rescale(fides_context* ctx, const uint64_t* d_in, uint32_t l, uint64_t* d_out, void* stream);

/* Automorphism X -> X^k (k odd) with its key: key switch of c1, add c0,
 * then permute both components (OpenFHE EvalAutomorphism order; bit-exact
 * with it).  Rotation by r uses k = 5^r mod 2N, conjugation k = 2N - 1.
 * d_out must not alias d_in.  Replaces Ciphertext::rotate / conjugate
 * (src/CKKS/Ciphertext.cpp:825-938) behind CryptoContext::EvalRotate
 * (api/CryptoContext.cpp:1406-1430). */
int fides_rotate(fides_context* ctx, const uint64_t* d_in, uint32_t l, const fides_key* key, uint32_t auto_index,
                 uint64_t* d_out, void* stream);

/* ---- batched operations (cross-ciphertext batching) ------------------ */

/* n ciphertext pairs at individual limb counts l[i] in one pass: every
 * pipeline stage (tensor, INTT, ModUp conversion, NTT, inner product,
 * ModDown, Rescale) is one launch over the limbs of all n ciphertexts.
 * With rescale != 0, out[i] is [2][l[i]-1] (HMult + relin + Rescale), else
 * [2][l[i]].  Results are identical to n calls of fides_hmult (+
 * fides_rescale).  Outputs must not overlap each other; without rescale an
 * output may coincide with its own a[i] or b[i] but must not overlap any
 * other input of the batch (refused with FIDES_ERR_ARG).  The reference
 * batches only inside linear transforms (src/CKKS/LimbPartitionBatch.cu:20-616);
 * this entry point has no single reference counterpart. */
int fides_hmult_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_a, const uint64_t* const* d_b,
                      const uint32_t* l, const fides_key* rlk, int rescale, uint64_t* const* d_out, void* stream);

/* n rotations/automorphisms, ciphertext i with its own key and index.
 * Identical to n calls of fides_rotate.  Entries that rotate the same input
 * (same pointer and level) share one ModUp, as fides_rotate_hoisted does.
 * No output may overlap any input of the batch or another output (refused
 * with FIDES_ERR_ARG). */
int fides_rotate_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_in, const uint32_t* l,
                       const fides_key* const* keys, const uint32_t* auto_index, uint64_t* const* d_out,
                       void* stream);

/* Hoisted rotations: n rotations of ONE ciphertext [2][l][N] sharing a single
 * ModUp of c1 (reference Ciphertext::rotate_hoisted,
 * src/CKKS/Ciphertext.cpp:940-1114, behind CryptoContext::EvalFastRotation /
 * EvalFastRotationPrecompute, api/CryptoContext.hpp:169-174).  keys[i] is the
 * rotation key of kidx[i] (5^r mod 2N, or 2N-1); kidx[i] == 1 copies the
 * input (its key may be NULL).  Each out[i] ([2][l][N]) is bit-identical to
 * fides_rotate(in, l, keys[i], kidx[i]).  Outputs must not overlap the input
 * or each other. */
int fides_rotate_hoisted(fides_context* ctx, const uint64_t* d_in, uint32_t l, uint32_t n,
                         const fides_key* const* keys, const uint32_t* kidx, uint64_t* const* d_out, void* stream);

/* ---- the extended basis Q_l u P (ModDown hoisted out of rotations) ----
 * An ext ciphertext is [2][l+K][N] uint64, evaluation form: limbs 0..l-1
 * under q_0..q_{l-1}, then the K special primes.  It holds P times a
 * ciphertext, so sums of rotations need one ModDown instead of one each
 * (reference Ciphertext::rotate_hoisted with ext = true,
 * src/CKKS/Ciphertext.cpp:940-1114, behind CryptoContext::EvalFastRotationExt,
 * api/CryptoContext.hpp:171,174; Accumulate, src/CKKS/AccumulateBroadcast.cu:79-106). */

/* Hoisted rotations of one ciphertext [2][l][N] left in Q_l u P: out[i] =
 * sigma_k(P c0 + KS_ext0(c1)), sigma_k(KS_ext1(c1)) with k = kidx[i] and the
 * key switch's inner product without its ModDown (with add_first == 0 the
 * P c0 term is left out, OpenFHE's addFirst).  kidx[i] == 1 gives
 * (P c0 or 0, P c1) (its key may be NULL).  One ModUp serves all n.  Outputs
 * must not overlap the input or each other. */
int fides_rotate_hoisted_ext(fides_context* ctx, const uint64_t* d_in, uint32_t l, uint32_t n,
                             const fides_key* const* keys, const uint32_t* kidx, int add_first, uint64_t* const* d_out,
                             void* stream);
/* out[i] = (P c0, P c1) in Q_l u P (zero special limbs) for [2][l[i]][N] inputs. */
int fides_ext_raise_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_in, const uint32_t* l,
                          uint64_t* const* d_out, void* stream);
/* out[i] = a[i] + b[i] (or - with subtract != 0) over [2][l[i]+K][N]; out may alias a or b. */
int fides_ext_add_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_a, const uint64_t* const* d_b,
                        const uint32_t* l, int subtract, uint64_t* const* d_out, void* stream);
/* ModDown of both components: [2][l[i]+K][N] -> [2][l[i]][N] (as fides_moddown). */
int fides_ext_moddown_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_in, const uint32_t* l,
                            uint64_t* const* d_out, void* stream);

/* Rescale of n ciphertexts ([2][l[i]] -> [2][l[i]-1]) in one pass. */
int fides_rescale_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_in, const uint32_t* l,
                        uint64_t* const* d_out, void* stream);

/* ---- element-wise ciphertext arithmetic (reference AddSub.cu:10-123 add_ /
 * sub_ / scalar_, behind CryptoContext::EvalAdd / EvalSub / EvalMult(scalar),
 * api/CryptoContext.hpp:118-160).  out[i] may alias a[i]. */

/* out = a + b, or a - b when subtract != 0 ([2][l[i]] each). */
int fides_add_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_a, const uint64_t* const* d_b,
                    const uint32_t* l, int subtract, uint64_t* const* d_out, void* stream);
/* out = a * c for an integer constant c given by its residues: c_res[i] is a
 * host array of l[i] values c mod q_j. */
int fides_mul_const_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_a, const uint32_t* l,
                          const uint64_t* const* c_res, uint64_t* const* d_out, void* stream);
/* out = a + c (the constant polynomial c added to component 0). */
int fides_add_const_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_a, const uint32_t* l,
                          const uint64_t* const* c_res, uint64_t* const* d_out, void* stream);

/* Fused form of the three above, one launch for n ciphertexts:
 * out = a * m  (+ or -) b  + c, where b[i] may be NULL (no b term), mul_res
 * or mul_res[i] NULL means m = 1, add_res or add_res[i] NULL means c = 0,
 * and sub[i] != 0 (sub may be NULL) selects subtraction of b.  Residue
 * arrays are host arrays of l[i] values.  Combines the reference's
 * mult_scalar / add_scalar / add / sub element-wise kernels
 * (ElemenwiseBatchKernels.cu:1281-1342). */
int fides_lin_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_a, const uint64_t* const* d_b,
                    const uint32_t* l, const uint64_t* const* mul_res, const uint64_t* const* add_res,
                    const uint32_t* sub, uint64_t* const* d_out, void* stream);

/* Linear combination with integer constants:
 * out[i] = sum_{j < cnt[i]} m_ij * x_ij + c_i  (c_i added to component 0),
 * 1 <= cnt[i] <= 64.  Terms are listed output by output (sum of cnt
 * entries): x_terms[k] a ciphertext stored with x_l[k] limbs per component
 * (x_l[k] >= l[i]: only its first l[i] limbs are read, so a term at a higher
 * level needs no limb drop), m_res[k] a host array of l[i] residues of the
 * constant, add_res[i] a host array of l[i] residues or NULL (add_res may be
 * NULL).  out[i] is [2][l[i]][N] and must not alias a term.  Fuses the
 * scalar multiplications and additions of a Chebyshev-series combination
 * (reference src/CKKS/ApproxModEval.cu, EvalChebyshevSeries). */
int fides_lincomb_batch(fides_context* ctx, uint32_t n, const uint32_t* cnt, const uint64_t* const* x_terms,
                        const uint32_t* x_l, const uint64_t* const* m_res, const uint64_t* const* add_res,
                        const uint32_t* l, uint64_t* const* d_out, void* stream);

/* Keep the first l_out[i] limbs of each component: [2][l_in] -> [2][l_out]
 * (Ciphertext::dropToLevel, src/CKKS/Ciphertext.cpp:1143, RNSPoly::dropToLevel
 * src/CKKS/RNSPoly.cpp:838; used for level alignment before add/mult,
 * src/CKKS/Ciphertext.cpp:229-233).  in == out only when l_out == l_in. */
int fides_drop_limbs_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_in, const uint32_t* l_in,
                           const uint32_t* l_out, uint64_t* const* d_out, void* stream);

/* ---- plaintexts and modulus raising (bootstrap building blocks) ---- */

/* In place: a plaintext [l][N] given as coefficient-form residues mod
 * q_0..q_{l-1} becomes evaluation form in Montgomery representation, the
 * form fides_ptdot_batch consumes (the reference keeps its linear-transform
 * plaintexts on the device in evaluation form, src/CKKS/Plaintext.cu). */
int fides_plaintext_prepare(fides_context* ctx, uint64_t* d_pt, uint32_t l, void* stream);

/* A prepared plaintext [l][N] as the ciphertext (pt, 0) [2][l][N] in standard
 * evaluation form (a trivial encryption), so plaintext additions and their
 * level/scale alignment run as ciphertext operations (reference
 * Ciphertext::addPt / subPt, src/CKKS/Ciphertext.cpp:242-300, which add the
 * plaintext limbs to c0). */
int fides_plaintext_embed(fides_context* ctx, const uint64_t* d_pt, uint32_t l, uint64_t* d_ct, void* stream);

/* Plaintext dot products: out[i] = sum_{j < cnt[i]} ct_ij (*) pt_ij, both
 * ciphertext components multiplied element-wise by the same plaintext
 * (reference dotProductLtBatchedPt, src/CKKS/ElemenwiseBatchKernels.cu:1007;
 * a single term is EvalMult(ct, pt), api/CryptoContext.hpp:150).
 * ct_terms / pt_terms list the terms of output 0, then output 1, ...
 * (sum of cnt entries); ct [2][l[i]][N], pt prepared [l[i]][N];
 * 1 <= cnt[i] <= 16.  out[i] may alias one of its own ct terms. */
int fides_ptdot_batch(fides_context* ctx, uint32_t n, const uint32_t* cnt, const uint64_t* const* ct_terms,
                      const uint64_t* const* pt_terms, const uint32_t* l, uint64_t* const* d_out, void* stream);

/* ModRaise: a ciphertext at one limb ([2][1][N], modulus q_0) is lifted to
 * l_out[i] limbs: each coefficient's centred representative mod q_0 is
 * reduced mod q_0..q_{l_out-1} (reference ModRaise, src/CKKS/Bootstrap.cu:393;
 * the lift is OpenFHE's SwitchModulus, src/CKKS/Rescale.cuh:113-133). */
int fides_mod_raise_batch(fides_context* ctx, uint32_t n, const uint64_t* const* d_in, const uint32_t* l_out,
                          uint64_t* const* d_out, void* stream);

/* Keyless evaluation-domain automorphism of nlimbs limbs (permutation only,
 * reference automorph_ src/Rotation.cu:10-28).  d_out must not alias. */
int fides_automorph(fides_context* ctx, const uint64_t* d_in, uint32_t nlimbs, uint32_t auto_index,
                    uint64_t* d_out, void* stream);

/* ---- graph capture ----
 * Every operation issued on `stream` between fides_graph_begin and
 * fides_graph_end is recorded into one HIP graph instead of running
 * (the reference caches CUDA graphs of its key-switch sequence,
 * src/CKKS/LimbPartitionMGPU.cu:601-609, and of RNSPoly operations,
 * src/CKKS/RNSPoly.cpp).  Task lists are copied to a device pool of
 * task_bytes owned by the graph while capturing; the key-switch scratch of
 * the captured operations is an arena of scratch_bytes owned by the graph
 * (an operation that needs more fails with FIDES_ERR_NOMEM and the capture
 * must be ended and discarded).  Issuing operations of the same context on
 * another stream while capturing is an error (refused by every operation
 * that builds task lists: the batched, key-switch and bootstrap calls).  A launch re-executes the recorded
 * kernels on the buffers named at capture time (update inputs in place to
 * run on new data); launches of one graph must not overlap each other. */
int fides_graph_begin(fides_context* ctx, void* stream, size_t scratch_bytes, size_t task_bytes);
/* fides_graph_begin with the key-switch scratch supplied by the caller:
 * `scratch` (device memory of scratch_bytes, 256-byte aligned) stays the
 * caller's and is not freed with the graph.  Graphs that share one scratch
 * region must never run at the same time (launch them all on one stream);
 * eager operations of the context use their own scratch and may overlap
 * them.  No counterpart in the reference: it serves many resident graphs,
 * one per request-program shape of a serving loop, without one scratch
 * arena each. */
int fides_graph_begin_shared(fides_context* ctx, void* stream, void* scratch, size_t scratch_bytes,
                             size_t task_bytes);
int fides_graph_end(fides_context* ctx, fides_graph** graph);
int fides_graph_launch(fides_graph* graph, void* stream);
/* Waits for the device, then frees the graph and its pools. */
int fides_graph_destroy(fides_graph* graph);

/* ---- tuning ----
 * Launch-shape settings of one context.  None of them changes a result bit:
 * they choose how the same arithmetic is split into tasks and launches.
 * The reference fixes these at compile time (block sizes and limb batches in
 * src/CKKS/LimbPartition*.cu); this interface has no reference counterpart.
 * Each setting is kept per size class c of a key-switch batch: a conversion
 * (ModUp of all digits, or ModDown) with < 512 (c = 0), < 1024 (1) or more
 * (2) target limbs in total; the flags follow the class of the ModUp.
 *   conv_chunk[c]      targets per basis-conversion task, 1..96
 *                      (defaults 8, 16, 32)
 *   group_by_prime[c]  order each NTT task list by prime so one prime's
 *                      twiddle tables stay in L2 (1, default) or keep
 *                      ciphertext order (0)
 *   fuse_row_dot[c]    run the row pass of the ModUp outputs inside the
 *                      key-switch inner product, and the row stages of the
 *                      special limbs' inverse NTT after it (1, default), or
 *                      as separate NTT, inner-product and INTT launches (0) */
typedef struct fides_tuning {
  uint32_t conv_chunk[3];
  uint32_t group_by_prime[3];
  uint32_t fuse_row_dot[3];
} fides_tuning;

int fides_tuning_get(const fides_context* ctx, fides_tuning* out);
int fides_tuning_set(fides_context* ctx, const fides_tuning* t);

/* Autotuner: times candidate settings on a synthetic batch shaped like the
 * caller's (n HMult+relin+Rescale and n Rotate at limb counts l[i], random
 * residues and a random key; a candidate's time is the best of 3 runs of
 * `reps` repetitions after one warm-up) by coordinate descent from the
 * current settings over the size classes the batch reaches (fusion,
 * grouping, then the conversion chunks); a candidate replaces the incumbent
 * only when it is at least 2% faster.  Settings of other classes are left
 * as they are.  Keeps the result in the context and writes it to *out (may
 * be NULL).  Runs on the
 * default stream and waits for the device; allocates its inputs, key and
 * outputs (about n + 3 full ciphertexts and one key) for the duration of the
 * call.  Refused during graph capture. */
int fides_autotune(fides_context* ctx, uint32_t n, const uint32_t* l, uint32_t reps, fides_tuning* out);

#ifdef __cplusplus
}
#endif

#endif /* FIDESLIB_AMD_H */
