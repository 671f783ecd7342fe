// fideslib.hpp -- host C++ API of the MI355X key-switching core.
//
// Mirrors the hot-path part of FIDESlib's public API
// (api/CryptoContext.hpp:24-251, api/Ciphertext.hpp:13-60) over the C ABI in
// fideslib_amd.h.  OpenFHE objects are not available here, so the data
// boundary is FIDESlib's own interchange format: RawParams,
// RawCipherText and RawKeySwitchKey (src/CKKS/openfhe-interface/
// RawCiphertext.cuh:19-93) without their OpenFHE members.  An OpenFHE
// adapter only has to fill these structs (GetRawParams, GetRawCipherText,
// GetEvalKeySwitchKey / GetRotationKeySwitchKey do exactly that in the
// reference).
//
// Errors throw std::runtime_error (the reference throws OPENFHE_THROW for
// precondition violations, api/CryptoContext.cpp:1922-1924).
#ifndef FIDESLIB_AMD_FIDESLIB_HPP
#define FIDESLIB_AMD_FIDESLIB_HPP

#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

extern "C" {
struct fides_context;
struct fides_key;
}

namespace fideslib {

enum ScalingTechnique { FIXEDMANUAL = 0, FIXEDAUTO = 1, FLEXIBLEAUTO = 2, FLEXIBLEAUTOEXT = 3 };

// Parameters (RawParams, RawCiphertext.cuh:45-83).  Roots may be left empty
// (the smallest-generator root is then used for every prime).
struct RawParams {
  int logN = 0;
  std::vector<uint64_t> moduli;                 // Q primes, level 0 first
  std::vector<uint64_t> root_of_unity;          // primitive 2N-th roots (optional)
  std::vector<uint64_t> SPECIALmoduli;          // P primes
  std::vector<uint64_t> SPECIALroot_of_unity;   // (optional)
  int dnum = 1;
  ScalingTechnique scalingTechnique = FIXEDMANUAL;
  std::vector<double> ModReduceFactor;          // per level (optional; default = the dropped prime)
  // OpenFHE's scaling factors per level (level = limbs - 1), as FIDESlib's
  // GetRawParams copies them (RawCiphertext.cu:260-272).  Optional: the
  // defaults follow OpenFHE's rule -- FLEXIBLEAUTO(EXT): sf[L-1] = q[L-1],
  // sf[l-1] = sf[l]^2 / q[l], Big[l] = sf[l]^2; fixed techniques: 2^p with p
  // the bit length of q[1] rounded (the scaling modulus size).
  std::vector<double> ScalingFactorReal;
  std::vector<double> ScalingFactorRealBig;
};

// Ciphertext in RNS evaluation form (RawCipherText, RawCiphertext.cuh:19-35).
struct RawCipherText {
  std::vector<std::vector<uint64_t>> sub_0;  // [numRes][N]
  std::vector<std::vector<uint64_t>> sub_1;
  int numRes = 0;
  int N = 0;
  double Noise = 0.0;  // scaling factor (NoiseFactor)
  int NoiseLevel = 1;  // scale degree
  int slots = 0;
  std::string keyid;
};

// Plaintext in RNS form (RawPlainText, RawCiphertext.cuh:37-48): evaluation
// form as FIDESlib keeps it on the device, or coefficient form.
enum Format { COEFFICIENT = 0, EVALUATION = 1 };
struct RawPlainText {
  std::vector<std::vector<uint64_t>> sub_0;  // [numRes][N]
  std::vector<uint64_t> moduli;              // optional check: q_0..q_{numRes-1}
  int numRes = 0;
  int N = 0;
  Format format = EVALUATION;
  double Noise = 0.0;  // scaling factor
  int slots = 0;
  int NoiseLevel = 1;
};

// Key-switching key (RawKeySwitchKey, RawCiphertext.cuh:85-90):
// r_key[0] = a, r_key[1] = b, each [dnum][L+K][N].
struct RawKeySwitchKey {
  std::vector<std::vector<std::vector<std::vector<uint64_t>>>> r_key;
  std::string keyid;
};

class CryptoContextImpl;
using CryptoContext = std::shared_ptr<CryptoContextImpl>;

// Bootstrapping parameters (reference EvalBootstrapSetup arguments,
// api/CryptoContext.hpp:183-186): level budget of CoeffToSlot and
// SlotToCoeff, the Chebyshev degree of the modular reduction and its double
// angles, the bound K' of t / q_0 it covers, and the prescale 2^-correction.
struct BootstrapConfig {
  uint32_t budgetCtS = 3, budgetStC = 3;
  uint32_t degree = 55;
  uint32_t doubleAngle = 3;
  double K = 32.0;
  uint32_t correction = 10;
  // Sparse-secret encapsulation (reference Bootstrap.cu:254-260, 592-600):
  // the one-limb ciphertext is switched to a sparser secret before ModRaise
  // and back right after it, with the keys of AddEncapsulationKeys; K can
  // then be 16 for a weight-32 sparse secret.
  bool encapsulation = false;
  // Message slots (a power of two up to N/2; 0 = N/2).  Fewer slots (sparse
  // packing, api/CryptoContext.hpp:80 `slots`, reference Bootstrap.cu:100,
  // 116-143): a partial trace after ModRaise, transforms of that many slots,
  // one EvalMod; the input holds its slots replicated over N/2.
  uint32_t slots = 0;
};

// Bootstrapping tables computed elsewhere (the reference takes them from
// OpenFHE, src/CKKS/BootstrapPrecomputation.cuh fed by RawCiphertext.cu:
// 488-631): for each linear-transform group, in application order, its
// baby-step / giant-step shape and plaintexts; the EvalMod Chebyshev series
// sum_k c_k T_k of the cosine; and the output's nominal scale.
struct BootstrapPrecomputation {
  BootstrapConfig cfg;  // budgets (= group counts), doubleAngle, correction, encapsulation
  struct Term {
    int64_t giant = 0, baby = 0;  // rotation k g stride of the product of baby b
    uint32_t level = 0;           // limbs of the plaintext
    double scale = 0;             // its encoding scale (q_top of the level: the Rescale divides it out)
    std::vector<uint64_t> coef;   // [level][N] coefficient-form residues mod q_0..q_{level-1}
  };
  struct Group {
    int64_t stride = 1, g = 1;
    std::vector<int64_t> babies;  // b with baby rotation b stride
    std::vector<Term> terms;      // giant ascending, then baby order
  };
  std::vector<Group> coeffToSlot, slotToCoeff;
  std::vector<double> chebyshev;
  double outScale = 0;  // nominal scale of the result (0: the tracked one)
};
struct BootPlan;

// A device-resident ciphertext: [2][l][N] evaluation form plus metadata
// (reference Ciphertext, src/CKKS/Ciphertext.cuh:34-54).
class CiphertextImpl {
 public:
  ~CiphertextImpl();
  CiphertextImpl(const CiphertextImpl&) = delete;
  CiphertextImpl& operator=(const CiphertextImpl&) = delete;

  uint32_t GetLevel() const { return limbs - 1; }  // FIDESlib level = limbs - 1
  uint32_t GetLimbs() const { return limbs; }
  int GetNoiseScaleDeg() const { return NoiseLevel; }
  double GetScalingFactor() const { return NoiseFactor; }
  int GetSlots() const { return slots; }
  uint64_t* DeviceData() const { return d; }
  // In the extended basis Q_l u P (an EvalFastRotationExt result or a sum of
  // them): [2][limbs + K][N] holding P times the ciphertext.  Every operation
  // other than addition and subtraction takes it back to Q_l first (one
  // ModDown per component), as the reference does lazily
  // (src/CKKS/Ciphertext.cpp:940-1114, isModUp()).
  bool IsExtended() const { return ext; }

  uint32_t limbs = 0;
  int NoiseLevel = 1;
  double NoiseFactor = 0.0;
  int slots = 0;
  std::string keyid;

 private:
  friend class CryptoContextImpl;
  CiphertextImpl(std::shared_ptr<CryptoContextImpl> cc, uint32_t limbs, bool ext = false);
  std::shared_ptr<CryptoContextImpl> cc;
  bool ext = false;
  uint64_t* d = nullptr;
  size_t words = 0;  // size of d (returned to the context's memory cache)
};
using Ciphertext = std::shared_ptr<CiphertextImpl>;

// A device-resident plaintext: [l][N] evaluation form in the Montgomery
// representation the plaintext products consume (reference Plaintext,
// src/CKKS/Plaintext.cuh).
class PlaintextImpl {
 public:
  ~PlaintextImpl();
  PlaintextImpl(const PlaintextImpl&) = delete;
  PlaintextImpl& operator=(const PlaintextImpl&) = delete;

  uint32_t GetLevel() const { return limbs - 1; }
  uint32_t GetLimbs() const { return limbs; }
  int GetNoiseScaleDeg() const { return NoiseLevel; }
  double GetScalingFactor() const { return NoiseFactor; }

  uint32_t limbs = 0;
  int NoiseLevel = 1;
  double NoiseFactor = 0.0;
  int slots = 0;

 private:
  friend class CryptoContextImpl;
  explicit PlaintextImpl(std::shared_ptr<CryptoContextImpl> cc, uint32_t limbs);
  std::shared_ptr<CryptoContextImpl> cc;
  uint64_t* d = nullptr;
  size_t words = 0;  // size of d (returned to the context's memory cache)
};
using Plaintext = std::shared_ptr<PlaintextImpl>;

class CryptoContextImpl : public std::enable_shared_from_this<CryptoContextImpl> {
 public:
  // GenCryptoContext + LoadContext (api/CryptoContext.cpp:199-249).
  static CryptoContext Create(const RawParams& params);
  ~CryptoContextImpl();

  uint32_t GetRingDimension() const { return N; }
  uint32_t GetNumSpecialPrimes() const { return K; }
  uint32_t GetCyclotomicOrder() const { return 2 * N; }
  ScalingTechnique GetScalingTechnique() const { return technique; }

  // ---- keys (LoadContext's key upload, api/CryptoContext.cpp:229-241)
  void AddEvalMultKey(const RawKeySwitchKey& key);
  void AddRotationKey(int32_t index, const RawKeySwitchKey& key);  // automorphism 5^index mod 2N
  void AddConjugateKey(const RawKeySwitchKey& key);                // automorphism 2N - 1
  bool HasRotationKey(int32_t index) const;

  // ---- data movement (LoadCiphertext / Decrypt's store, api/CryptoContext.cpp:270-288, 603-622)
  Ciphertext LoadCiphertext(const RawCipherText& raw);
  RawCipherText StoreCiphertext(const Ciphertext& ct);

  // ---- operations (api/CryptoContext.hpp:149-181)
  Ciphertext EvalMult(const Ciphertext& a, const Ciphertext& b);
  void EvalMultInPlace(Ciphertext& a, const Ciphertext& b);
  // Plaintexts (api/CryptoContext.hpp:60-66 LoadPlaintext, :150-159
  // EvalMult(ct, pt)): the plaintext needs at least as many limbs as the
  // ciphertext (its first limbs are used); with automatic rescaling a
  // degree-2 ciphertext is rescaled first (Ciphertext::multPt,
  // src/CKKS/Ciphertext.cpp:358-423).
  Plaintext LoadPlaintext(const RawPlainText& raw);
  Ciphertext EvalMult(const Ciphertext& a, const Plaintext& p);
  Ciphertext EvalMult(const Plaintext& p, const Ciphertext& a) { return EvalMult(a, p); }
  void EvalMultInPlace(Ciphertext& a, const Plaintext& p) { a = EvalMult(a, p); }
  Ciphertext EvalSquare(const Ciphertext& a);
  void EvalSquareInPlace(Ciphertext& a);
  // The reference's "Mutable" forms are aliases of the plain ones
  // (api/CryptoContext.cpp:864-883, 1321-1347, 1401-1404).
  Ciphertext EvalMultMutable(Ciphertext& a, Ciphertext& b) { return EvalMult(a, b); }
  Ciphertext EvalMultMutable(Ciphertext& a, Plaintext& p) { return EvalMult(a, p); }
  Ciphertext EvalMultMutable(Plaintext& p, Ciphertext& a) { return EvalMult(a, p); }
  void EvalMultMutableInPlace(Ciphertext& a, Ciphertext& b) { a = EvalMult(a, b); }
  Ciphertext EvalSquareMutable(Ciphertext& a) { return EvalSquare(a); }
  Ciphertext EvalRotate(const Ciphertext& a, int32_t index);
  void EvalRotateInPlace(Ciphertext& a, int32_t index);
  Ciphertext EvalConjugate(const Ciphertext& a);
  // Hoisted rotations (api/CryptoContext.hpp:169-174): one ModUp shared by
  // all indices; each result equals EvalRotate with the same index.
  std::shared_ptr<void> EvalFastRotationPrecompute(const Ciphertext& a);
  Ciphertext EvalFastRotation(const Ciphertext& a, int32_t index, uint32_t m, const std::shared_ptr<void>& precomp);
  std::vector<Ciphertext> EvalFastRotation(const Ciphertext& a, const std::vector<int32_t>& indices, uint32_t m,
                                           const std::shared_ptr<void>& precomp);
  // Hoisted rotations left in the extended basis (api/CryptoContext.hpp:171,
  // 174): sigma_k(P c0 + KS0), sigma_k(KS1) without the ModDown (addFirst ==
  // false leaves P c0 out, OpenFHE's meaning; the reference's GPU path always
  // adds it).  Sums of the results take one ModDown when used.  `digits` is
  // unused: the device path raises the digits itself (one ModUp per call).
  Ciphertext EvalFastRotationExt(const Ciphertext& a, int32_t index, const std::shared_ptr<void>& digits,
                                 bool addFirst);
  std::vector<Ciphertext> EvalFastRotationExt(const Ciphertext& a, const std::vector<int32_t>& indices,
                                              const std::shared_ptr<void>& digits, bool addFirst);
  // sum_{j < slots} Rotate(a, j stride) (api/CryptoContext.hpp:188-190): rounds
  // of three hoisted extended rotations whose sum takes one ModDown (the
  // reference GPU Accumulate with bStep 4, src/CKKS/AccumulateBroadcast.cu:
  // 79-106); with `start`, the rounds begin at rotation stride*start
  // (AccumulateCascadeImpl, :9-42).  Rotation keys for stride*s*j
  // (j = 1..3, s = 1, 4, 16, ...) must be loaded.
  Ciphertext AccumulateSum(const Ciphertext& a, int slots, int stride = 1);
  void AccumulateSumInPlace(Ciphertext& a, int slots, int stride = 1);
  void AccumulateSumInPlace(Ciphertext& a, int slots, int stride, int start);
  // Chebyshev series c_0 / 2 + sum_{k>=1} c_k T_k(y), y = (2x - a - b)/(b - a)
  // (api/CryptoContext.hpp:176-178, OpenFHE's coefficient convention; the
  // reference evalChebyshevSeries, src/CKKS/ApproxModEval.cu:371): the map to
  // [-1, 1] (skipped for [-1, 1]), baby steps T_1..T_m by the product rule,
  // giant steps by doubling and a recursive Chebyshev division into fused
  // linear combinations.  The result is at scale degree 1 with the nominal
  // scale 2^(bit length of q_1); it needs ceil(log2(deg)) + 2 levels.
  Ciphertext EvalChebyshevSeries(const Ciphertext& a, std::vector<double>& coeffs, double lo, double hi);
  void EvalChebyshevSeriesInPlace(Ciphertext& a, std::vector<double>& coeffs, double lo, double hi);
  // Interpolation coefficients of func on [a, b] at the degree + 1 Chebyshev
  // nodes, c_0 not halved (reference get_chebyshev_coefficients,
  // src/PolyApprox.cu:11-32).
  static std::vector<double> GetChebyshevCoefficients(std::function<double(double)>& func, double a, double b,
                                                      size_t degree);
  // OpenFHE's level convention (levels consumed): keep L - level limbs
  // (api/CryptoContext.cpp:1714-1717 -> Ciphertext::dropToLevel); with
  // FLEXIBLEAUTO(EXT) the scale follows the target level's factor
  // (adjustScaleAndLevel), otherwise limbs are dropped.
  static void SetLevel(Ciphertext& a, size_t level);
  Ciphertext Rescale(const Ciphertext& a);
  void RescaleInPlace(Ciphertext& a);
  // Drop limbs above `limbs` (OpenFHE LevelReduce / FIDESlib dropToLevel).
  Ciphertext DropToLimbs(const Ciphertext& a, uint32_t limbs);

  // ---- element-wise operations (api/CryptoContext.hpp:114-160), with the
  // level/scale alignment of the scaling technique (Ciphertext::add/sub,
  // adjustForAddOrSub, src/CKKS/Ciphertext.cpp:180-240, 1462-1500)
  Ciphertext EvalAdd(const Ciphertext& a, const Ciphertext& b);
  Ciphertext EvalSub(const Ciphertext& a, const Ciphertext& b);
  Ciphertext EvalAdd(const Ciphertext& a, double c);
  Ciphertext EvalSub(const Ciphertext& a, double c);
  Ciphertext EvalMult(const Ciphertext& a, double c);
  Ciphertext EvalAdd(double c, const Ciphertext& a) { return EvalAdd(a, c); }
  Ciphertext EvalSub(double c, const Ciphertext& a) { return EvalAdd(EvalNegate(a), c); }
  Ciphertext EvalMult(double c, const Ciphertext& a) { return EvalMult(a, c); }
  Ciphertext EvalNegate(const Ciphertext& a);
  void EvalNegateInPlace(Ciphertext& a) { a = EvalNegate(a); }
  void EvalAddInPlace(Ciphertext& a, const Ciphertext& b) { a = EvalAdd(a, b); }
  void EvalSubInPlace(Ciphertext& a, const Ciphertext& b) { a = EvalSub(a, b); }
  void EvalAddInPlace(Ciphertext& a, double c) { a = EvalAdd(a, c); }
  void EvalSubInPlace(Ciphertext& a, double c) { a = EvalSub(a, c); }
  void EvalAddInPlace(double c, Ciphertext& a) { a = EvalAdd(a, c); }
  void EvalSubInPlace(double c, Ciphertext& a) { a = EvalSub(c, a); }
  void EvalMultInPlace(Ciphertext& a, double c) { a = EvalMult(a, c); }
  void EvalMultInPlace(double c, Ciphertext& a) { a = EvalMult(a, c); }
  Ciphertext EvalAddMutable(Ciphertext& a, Ciphertext& b) { return EvalAdd(a, b); }
  Ciphertext EvalSubMutable(Ciphertext& a, Ciphertext& b) { return EvalSub(a, b); }
  void EvalAddMutableInPlace(Ciphertext& a, Ciphertext& b) { a = EvalAdd(a, b); }
  void EvalSubMutableInPlace(Ciphertext& a, Ciphertext& b) { a = EvalSub(a, b); }
  // Plaintext addition and subtraction (api/CryptoContext.hpp:119-146;
  // Ciphertext::addPt / subPt, src/CKKS/Ciphertext.cpp:242-300): with
  // automatic rescaling a degree-2 ciphertext one level above the plaintext
  // is rescaled, and a plaintext at another level is brought to the
  // ciphertext's (dropped for FIXEDAUTO, adjustScaleAndLevel for
  // FLEXIBLEAUTO(EXT); it must not be below the ciphertext).  FIXEDMANUAL
  // needs equal scale degrees and uses the plaintext's first limbs.
  Ciphertext EvalAdd(const Ciphertext& a, const Plaintext& p);
  Ciphertext EvalAdd(const Plaintext& p, const Ciphertext& a) { return EvalAdd(a, p); }
  Ciphertext EvalSub(const Ciphertext& a, const Plaintext& p);
  Ciphertext EvalSub(const Plaintext& p, const Ciphertext& a) { return EvalAdd(EvalNegate(a), p); }
  void EvalAddInPlace(Ciphertext& a, const Plaintext& p) { a = EvalAdd(a, p); }
  void EvalAddInPlace(const Plaintext& p, Ciphertext& a) { a = EvalAdd(a, p); }
  void EvalSubInPlace(Ciphertext& a, const Plaintext& p) { a = EvalSub(a, p); }
  Ciphertext EvalAddMutable(Ciphertext& a, Plaintext& p) { return EvalAdd(a, p); }
  Ciphertext EvalAddMutable(Plaintext& p, Ciphertext& a) { return EvalAdd(a, p); }
  Ciphertext EvalSubMutable(Ciphertext& a, Plaintext& p) { return EvalSub(a, p); }
  Ciphertext EvalSubMutable(Plaintext& p, Ciphertext& a) { return EvalSub(p, a); }
  // Sum of n ciphertexts (api/CryptoContext.hpp:133-134; the reference adds
  // them pairwise in a tree, api/CryptoContext.cpp:884-927).  Inputs at one
  // level and scale are summed in fused linear-combination launches (the same
  // bits: modular addition is exact); otherwise the pairwise tree of EvalAdd
  // with its level/scale alignment runs.
  Ciphertext EvalAddMany(const std::vector<Ciphertext>& cts);
  void EvalAddManyInPlace(std::vector<Ciphertext>& cts);

  double ScalingFactorReal(uint32_t level) const { return sfReal.at(level); }
  double ScalingFactorRealBig(uint32_t level) const { return sfBig.at(level); }

  // ---- cross-ciphertext batching (one launch per pipeline stage)
  std::vector<Ciphertext> EvalMultMany(const std::vector<Ciphertext>& a, const std::vector<Ciphertext>& b,
                                       bool rescale);
  std::vector<Ciphertext> EvalRotateMany(const std::vector<Ciphertext>& a, const std::vector<int32_t>& indices);

  void Synchronize() const;
  // Device memory of released ciphertexts, plaintexts and bootstrap
  // registers is kept by the context for reuse (no allocation, and no
  // device synchronization of a free, on the next operation of the same
  // size).  This returns the idle blocks to the device.
  void ReleaseCachedMemory();

  // Launch shapes (the C ABI's fides_tuning, per size class of a key-switch
  // batch); no setting changes a result bit.  No reference counterpart: the
  // reference fixes them at compile time.
  struct Tuning {
    uint32_t convChunk[3];
    uint32_t groupByPrime[3];
    uint32_t fuseRowDot[3];
  };
  Tuning GetTuning() const;
  void SetTuning(const Tuning& t);
  // Times candidate shapes on a synthetic batch of HMult+Rescale and Rotate
  // at these limb counts, keeps the fastest in the context and returns it.
  Tuning Autotune(const std::vector<uint32_t>& limbs, uint32_t reps = 2);

  // Bootstrapping of full-slot ciphertexts (api/CryptoContext.hpp:183-186;
  // reference Bootstrap.cu:169-333): the setup computes the linear-transform
  // plaintexts and the modular-reduction series and returns the rotation
  // indices whose keys EvalBootstrap needs (plus the conjugation key,
  // automorphism index 2N-1).  The result has limbs - depth limbs.
  // With FLEXIBLEAUTO(EXT) the result carries the FLEXIBLEAUTO scaling factor
  // of its level (folded into the first SlotToCoeff group), so it adds to
  // fresh ciphertexts of that level without a scale correction.
  std::vector<int32_t> EvalBootstrapSetup(const BootstrapConfig& cfg = BootstrapConfig());
  // The same from precomputed tables (the reference's path from OpenFHE's
  // precomputation); returns the rotation indices the groups need.
  std::vector<int32_t> EvalBootstrapSetup(const BootstrapPrecomputation& pre);
  // Switching keys of the sparse-secret encapsulation: s -> s' and s' -> s
  // (GetSecretSwitchingKey, reference Bootstrap.cu:259, 596).
  void AddEncapsulationKeys(const RawKeySwitchKey& toSparse, const RawKeySwitchKey& fromSparse);
  // numIterations > 1: iterated bootstrapping (OpenFHE's meaning of the
  // arguments; the reference's GPU path runs one iteration,
  // api/CryptoContext.cpp:1719-1745): each further iteration bootstraps the
  // error 2^precision (x - result) and adds it back; `precision` is the bits
  // one bootstrap keeps.  The result's scaling factor is the bootstrap's
  // output factor times 2^precision per extra iteration (exact; no level is
  // spent on a division).
  Ciphertext EvalBootstrap(const Ciphertext& ct, uint32_t numIterations = 1, uint32_t precision = 0);
  void EvalBootstrapInPlace(Ciphertext& ct, uint32_t numIterations = 1, uint32_t precision = 0) {
    ct = EvalBootstrap(ct, numIterations, precision);
  }
  // One bootstrap of each input (one iteration), the same results as
  // EvalBootstrap on each; after each input's prescale to one limb, every
  // step of the plan runs once for all of them (one batched call).  No
  // reference counterpart: the reference bootstraps one ciphertext per call
  // (api/CryptoContext.hpp:185).  The device memory of the plan's registers
  // is kept by the context between calls (no allocation once a batch size
  // has run) and released with it.
  std::vector<Ciphertext> EvalBootstrapMany(const std::vector<Ciphertext>& cts);
  // The key-switching key loaded for an automorphism index (throws if absent).
  const fides_key* RotationKeyFor(uint32_t autoIndex) const { return RotationKey(autoIndex); }

  fides_context* Handle() const { return ctx; }

  // ---- serialization support (fideslib/serialize.hpp)
  const RawParams& GetRawParams() const { return raw; }
  RawKeySwitchKey GetEvalMultKey() const;                       // downloads the resident key
  std::map<uint32_t, RawKeySwitchKey> GetRotationKeys() const;  // by automorphism index
  void AddAutomorphismKey(uint32_t autoIndex, const RawKeySwitchKey& key);

 private:
  CryptoContextImpl() = default;
  Ciphertext NewCiphertext(uint32_t limbs, bool ext = false);
  Ciphertext Regular(const Ciphertext& a);          // ModDown of an extended ciphertext (else a)
  Ciphertext Extended(const Ciphertext& a);         // (P c0, P c1) in Q_l u P (a if extended)
  Ciphertext AddSubExt(const Ciphertext& a, const Ciphertext& b, bool sub);
  Ciphertext EmbedPlaintext(const Plaintext& p);    // (pt, 0) as a ciphertext
  Ciphertext AddSubPt(const Ciphertext& a, const Plaintext& p, bool sub);
  void AccumulateImpl(Ciphertext& a, int stride, int size, int start);
  Ciphertext SetLevelImpl(const Ciphertext& a, size_t level);
  uint32_t AutoIndex(int32_t index) const;
  const fides_key* RotationKey(uint32_t autoIndex) const;
  void PrepareMult(Ciphertext& a, Ciphertext& b);  // FIXEDAUTO / FLEXIBLEAUTO level/scale alignment
  void MultMetadata(CiphertextImpl& out, const CiphertextImpl& a, const CiphertextImpl& b) const;
  bool Flexible() const { return technique == FLEXIBLEAUTO || technique == FLEXIBLEAUTOEXT; }
  std::vector<uint64_t> ScalarResidues(uint32_t limbs, double operand, double scale) const;
  std::vector<uint64_t> AddScalarResidues(uint32_t limbs, double operand, int noiseDeg) const;
  Ciphertext MulConst(const Ciphertext& a, const std::vector<uint64_t>& res);
  Ciphertext MultScalarNoPrecheck(const Ciphertext& a, double c);
  Ciphertext MultScalar(const Ciphertext& a, double c);
  bool AdjustScaleAndLevel(Ciphertext& a, int scaleDegree, uint32_t level, double scalingFactor);
  bool AdjustForAddOrSub(Ciphertext& a, const Ciphertext& b);
  bool AdjustForMult(Ciphertext& a, const Ciphertext& b);
  Ciphertext AddSub(const Ciphertext& a, const Ciphertext& b, bool sub);
  Ciphertext BootstrapOnce(const Ciphertext& a);

  fides_context* ctx = nullptr;
  RawParams raw;
  uint32_t logN = 0, N = 0, L = 0, K = 0, dnum = 0;
  ScalingTechnique technique = FIXEDMANUAL;
  std::vector<uint64_t> q;
  std::vector<double> modReduceFactor;
  std::vector<double> sfReal, sfBig;  // by level = limbs - 1
  fides_key* evalMultKey = nullptr;
  fides_key* encapDown = nullptr;  // s -> s' (sparse-secret encapsulation)
  fides_key* encapUp = nullptr;    // s' -> s
  std::map<uint32_t, fides_key*> rotationKeys;  // by automorphism index
  std::shared_ptr<BootPlan> boot;                // EvalBootstrapSetup
  std::shared_ptr<void> bootMem;                 // device blocks kept between bootstraps (a register pool)
 public:
  // (internal) the context's cache of device blocks; see ReleaseCachedMemory
  uint64_t* TakeBlock(size_t words);
  void GiveBlock(uint64_t* d, size_t words);

 private:
  std::shared_ptr<void> blockMem;                // idle ciphertext / plaintext blocks
  // the nominal scale of a fresh ciphertext: 2^(bit length of q_1, rounded)
  double boot_delta() const { return std::ldexp(1.0, (int)std::lround(std::log2((double)q[q.size() > 1 ? 1 : 0]))); }
};

}  // namespace fideslib

#endif
