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

  uint32_t limbs = 0;
  int NoiseLevel = 1;
  double NoiseFactor = 0.0;
  int slots = 0;
  std::string keyid;

 private:
  friend class CryptoContextImpl;
  explicit CiphertextImpl(std::shared_ptr<CryptoContextImpl> cc, uint32_t limbs);
  std::shared_ptr<CryptoContextImpl> cc;
  uint64_t* d = nullptr;
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
};
using Plaintext = std::shared_ptr<PlaintextImpl>;

class CryptoContextImpl : public std::enable_shared_from_this<CryptoContextImpl> {
 public:
  // GenCryptoContext + LoadContext (api/CryptoContext.cpp:199-249).
  static CryptoContext Create(const RawParams& params);
  ~CryptoContextImpl();

  uint32_t GetRingDimension() const { return N; }
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
  Ciphertext EvalRotate(const Ciphertext& a, int32_t index);
  void EvalRotateInPlace(Ciphertext& a, int32_t index);
  Ciphertext EvalConjugate(const Ciphertext& a);
  // Hoisted rotations (api/CryptoContext.hpp:169-174): one ModUp shared by
  // all indices; each result equals EvalRotate with the same index.
  std::shared_ptr<void> EvalFastRotationPrecompute(const Ciphertext& a);
  Ciphertext EvalFastRotation(const Ciphertext& a, int32_t index, uint32_t m, const std::shared_ptr<void>& precomp);
  std::vector<Ciphertext> EvalFastRotation(const Ciphertext& a, const std::vector<int32_t>& indices, uint32_t m,
                                           const std::shared_ptr<void>& precomp);
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
  Ciphertext EvalNegate(const Ciphertext& a);
  void EvalAddInPlace(Ciphertext& a, const Ciphertext& b) { a = EvalAdd(a, b); }
  void EvalSubInPlace(Ciphertext& a, const Ciphertext& b) { a = EvalSub(a, b); }
  void EvalMultInPlace(Ciphertext& a, double c) { a = EvalMult(a, c); }

  double ScalingFactorReal(uint32_t level) const { return sfReal.at(level); }
  double ScalingFactorRealBig(uint32_t level) const { return sfBig.at(level); }

  // ---- cross-ciphertext batching (one launch per pipeline stage)
  std::vector<Ciphertext> EvalMultMany(const std::vector<Ciphertext>& a, const std::vector<Ciphertext>& b,
                                       bool rescale);
  std::vector<Ciphertext> EvalRotateMany(const std::vector<Ciphertext>& a, const std::vector<int32_t>& indices);

  void Synchronize() const;

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
  std::vector<int32_t> EvalBootstrapSetup(const BootstrapConfig& cfg = BootstrapConfig());
  Ciphertext EvalBootstrap(const Ciphertext& ct);
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
  Ciphertext NewCiphertext(uint32_t limbs);
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

  fides_context* ctx = nullptr;
  RawParams raw;
  uint32_t logN = 0, N = 0, L = 0, K = 0, dnum = 0;
  ScalingTechnique technique = FIXEDMANUAL;
  std::vector<uint64_t> q;
  std::vector<double> modReduceFactor;
  std::vector<double> sfReal, sfBig;  // by level = limbs - 1
  fides_key* evalMultKey = nullptr;
  std::map<uint32_t, fides_key*> rotationKeys;  // by automorphism index
  std::shared_ptr<BootPlan> boot;                // EvalBootstrapSetup
  // the nominal scale of a fresh ciphertext: 2^(bit length of q_1, rounded)
  double boot_delta() const { return std::ldexp(1.0, (int)std::lround(std::log2((double)q[q.size() > 1 ? 1 : 0]))); }
};

}  // namespace fideslib

#endif
