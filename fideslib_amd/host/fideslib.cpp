// Host C++ API (include/fideslib/fideslib.hpp) over the C ABI.
//
// Metadata rules follow FIDESlib's Ciphertext (src/CKKS/Ciphertext.cpp):
//   mult:    NoiseLevel = a + b, NoiseFactor = a * b           (:131-143 multMetadata)
//   rescale: NoiseLevel -= 1, NoiseFactor /= ModReduceFactor   (:426-447)
//   FIXEDAUTO: operands at scale degree 2 are rescaled before a mult and
//   levels are aligned by dropping limbs (:1483-1518 adjustForAddOrSub /
//   adjustForMult, FIXED branch).
// FLEXIBLEAUTO(EXT) needs OpenFHE's per-level real scaling factors and the
// scalar-multiplication correction of adjustScaleAndLevel (:1376-1455); it is
// rejected with an error instead of being approximated.
#include "fideslib/fideslib.hpp"

#include <cmath>
#include <stdexcept>

#include "fideslib_amd.h"

namespace fideslib {

static void chk(int rc, const char* what) {
  if (rc != FIDES_OK) throw std::runtime_error(std::string(what) + ": " + fides_last_error());
}

CiphertextImpl::CiphertextImpl(std::shared_ptr<CryptoContextImpl> c, uint32_t l) : limbs(l), cc(std::move(c)) {
  void* p = nullptr;
  chk(fides_device_alloc((size_t)2 * l * cc->GetRingDimension() * 8, &p), "ciphertext allocation");
  d = (uint64_t*)p;
}

CiphertextImpl::~CiphertextImpl() {
  if (d) fides_device_free(d);
}

CryptoContext CryptoContextImpl::Create(const RawParams& P) {
  if (P.moduli.empty() || P.SPECIALmoduli.empty()) throw std::runtime_error("RawParams: empty moduli");
  if (P.scalingTechnique == FLEXIBLEAUTO || P.scalingTechnique == FLEXIBLEAUTOEXT)
    throw std::runtime_error("FLEXIBLEAUTO scaling needs OpenFHE's real scaling factors; not supported");
  std::shared_ptr<CryptoContextImpl> c(new CryptoContextImpl());
  c->logN = (uint32_t)P.logN;
  c->N = 1u << P.logN;
  c->L = (uint32_t)P.moduli.size();
  c->K = (uint32_t)P.SPECIALmoduli.size();
  c->dnum = (uint32_t)P.dnum;
  c->technique = P.scalingTechnique;
  c->q = P.moduli;
  std::vector<uint64_t> roots;
  const bool have_roots =
      P.root_of_unity.size() == P.moduli.size() && P.SPECIALroot_of_unity.size() == P.SPECIALmoduli.size();
  if (have_roots) {
    roots = P.root_of_unity;
    roots.insert(roots.end(), P.SPECIALroot_of_unity.begin(), P.SPECIALroot_of_unity.end());
  }
  fides_params fp{c->logN, c->L, c->K, c->dnum, P.moduli.data(), P.SPECIALmoduli.data(),
                  have_roots ? roots.data() : nullptr};
  chk(fides_context_create(&fp, &c->ctx), "context creation");
  c->modReduceFactor = P.ModReduceFactor;
  if (c->modReduceFactor.size() < c->L) {
    c->modReduceFactor.resize(c->L);
    for (uint32_t i = 0; i < c->L; ++i) c->modReduceFactor[i] = (double)c->q[i];
  }
  return c;
}

CryptoContextImpl::~CryptoContextImpl() {
  if (evalMultKey) fides_key_destroy(evalMultKey);
  for (auto& kv : rotationKeys) fides_key_destroy(kv.second);
  if (ctx) fides_context_destroy(ctx);
}

static fides_key* upload(fides_context* ctx, const RawKeySwitchKey& k, uint32_t dnum, uint32_t LK, uint32_t N) {
  if (k.r_key.size() != 2) throw std::runtime_error("RawKeySwitchKey: r_key must hold {a, b}");
  std::vector<uint64_t> flat((size_t)2 * dnum * LK * N);
  for (size_t c = 0; c < 2; ++c) {
    if (k.r_key[c].size() != dnum) throw std::runtime_error("RawKeySwitchKey: digit count != dnum");
    for (size_t j = 0; j < dnum; ++j) {
      if (k.r_key[c][j].size() != LK) throw std::runtime_error("RawKeySwitchKey: limb count != L + K");
      for (size_t i = 0; i < LK; ++i) {
        if (k.r_key[c][j][i].size() != N) throw std::runtime_error("RawKeySwitchKey: limb size != N");
        std::copy(k.r_key[c][j][i].begin(), k.r_key[c][j][i].end(), flat.begin() + ((c * dnum + j) * LK + i) * N);
      }
    }
  }
  fides_key* out = nullptr;
  chk(fides_key_upload(ctx, flat.data(), &out), "key upload");
  return out;
}

void CryptoContextImpl::AddEvalMultKey(const RawKeySwitchKey& k) {
  if (evalMultKey) fides_key_destroy(evalMultKey);
  evalMultKey = upload(ctx, k, dnum, L + K, N);
}

uint32_t CryptoContextImpl::AutoIndex(int32_t index) const {
  // 5 has order N/2 mod 2N; rotation by a negative index r is rotation by
  // r mod N/2 (reference index normalisation, src/CKKS/Context.cu:1118-1125;
  // autoIndex = 5^r mod 2N, RawCiphertext.cu:711).
  const int64_t half = N / 2;
  int64_t r = ((int64_t)index % half + half) % half;
  uint64_t k = 1, b = 5, m = 2 * (uint64_t)N;
  while (r) {
    if (r & 1) k = k * b % m;
    b = b * b % m;
    r >>= 1;
  }
  return (uint32_t)k;
}

void CryptoContextImpl::AddRotationKey(int32_t index, const RawKeySwitchKey& k) {
  const uint32_t a = AutoIndex(index);
  auto it = rotationKeys.find(a);
  if (it != rotationKeys.end()) fides_key_destroy(it->second);
  rotationKeys[a] = upload(ctx, k, dnum, L + K, N);
}

void CryptoContextImpl::AddConjugateKey(const RawKeySwitchKey& k) {
  const uint32_t a = 2 * N - 1;
  auto it = rotationKeys.find(a);
  if (it != rotationKeys.end()) fides_key_destroy(it->second);
  rotationKeys[a] = upload(ctx, k, dnum, L + K, N);
}

bool CryptoContextImpl::HasRotationKey(int32_t index) const { return rotationKeys.count(AutoIndex(index)) != 0; }

const fides_key* CryptoContextImpl::RotationKey(uint32_t a) const {
  auto it = rotationKeys.find(a);
  if (it == rotationKeys.end())
    throw std::runtime_error("rotation key for automorphism index " + std::to_string(a) + " is not loaded");
  return it->second;
}

Ciphertext CryptoContextImpl::NewCiphertext(uint32_t limbs) {
  return Ciphertext(new CiphertextImpl(shared_from_this(), limbs));
}

Ciphertext CryptoContextImpl::LoadCiphertext(const RawCipherText& raw) {
  const uint32_t l = (uint32_t)raw.sub_0.size();
  if (l == 0 || l > L || raw.sub_1.size() != l) throw std::runtime_error("RawCipherText: bad limb count");
  std::vector<uint64_t> flat((size_t)2 * l * N);
  for (uint32_t i = 0; i < l; ++i) {
    if (raw.sub_0[i].size() != N || raw.sub_1[i].size() != N) throw std::runtime_error("RawCipherText: limb size");
    std::copy(raw.sub_0[i].begin(), raw.sub_0[i].end(), flat.begin() + (size_t)i * N);
    std::copy(raw.sub_1[i].begin(), raw.sub_1[i].end(), flat.begin() + (size_t)(l + i) * N);
  }
  Ciphertext ct = NewCiphertext(l);
  chk(fides_copy_to_device(ct->d, flat.data(), flat.size() * 8, nullptr), "ciphertext upload");
  ct->NoiseLevel = raw.NoiseLevel;
  ct->NoiseFactor = raw.Noise;
  ct->slots = raw.slots;
  ct->keyid = raw.keyid;
  return ct;
}

RawCipherText CryptoContextImpl::StoreCiphertext(const Ciphertext& ct) {
  const uint32_t l = ct->limbs;
  std::vector<uint64_t> flat((size_t)2 * l * N);
  chk(fides_copy_to_host(flat.data(), ct->d, flat.size() * 8, nullptr), "ciphertext download");
  RawCipherText r;
  r.sub_0.resize(l);
  r.sub_1.resize(l);
  for (uint32_t i = 0; i < l; ++i) {
    r.sub_0[i].assign(flat.begin() + (size_t)i * N, flat.begin() + (size_t)(i + 1) * N);
    r.sub_1[i].assign(flat.begin() + (size_t)(l + i) * N, flat.begin() + (size_t)(l + i + 1) * N);
  }
  r.numRes = (int)l;
  r.N = (int)N;
  r.Noise = ct->NoiseFactor;
  r.NoiseLevel = ct->NoiseLevel;
  r.slots = ct->slots;
  r.keyid = ct->keyid;
  return r;
}

void CryptoContextImpl::MultMetadata(CiphertextImpl& o, const CiphertextImpl& a, const CiphertextImpl& b) const {
  o.NoiseLevel = a.NoiseLevel + b.NoiseLevel;
  o.NoiseFactor = a.NoiseFactor * b.NoiseFactor;
  o.slots = a.slots;
  o.keyid = a.keyid;
}

Ciphertext CryptoContextImpl::DropToLimbs(const Ciphertext& a, uint32_t limbs) {
  if (limbs == 0 || limbs > a->limbs) throw std::runtime_error("DropToLimbs: bad limb count");
  Ciphertext o = NewCiphertext(limbs);
  const size_t bytes = (size_t)limbs * N * 8;
  chk(fides_copy_device(o->d, a->d, bytes, nullptr), "DropToLimbs");
  chk(fides_copy_device(o->d + (size_t)limbs * N, a->d + (size_t)a->limbs * N, bytes, nullptr), "DropToLimbs");
  o->NoiseLevel = a->NoiseLevel;
  o->NoiseFactor = a->NoiseFactor;
  o->slots = a->slots;
  o->keyid = a->keyid;
  return o;
}

void CryptoContextImpl::PrepareMult(Ciphertext& a, Ciphertext& b) {
  if (technique == FIXEDMANUAL) {
    if (a->limbs != b->limbs) throw std::runtime_error("EvalMult: operands at different levels (FIXEDMANUAL)");
    return;
  }
  // FIXEDAUTO: rescale degree-2 operands, then align limb counts
  if (a->NoiseLevel == 2) a = Rescale(a);
  if (b->NoiseLevel == 2) b = Rescale(b);
  if (a->limbs > b->limbs) a = DropToLimbs(a, b->limbs);
  if (b->limbs > a->limbs) b = DropToLimbs(b, a->limbs);
}

Ciphertext CryptoContextImpl::EvalMult(const Ciphertext& a0, const Ciphertext& b0) {
  if (!evalMultKey) throw std::runtime_error("EvalMult: relinearization key not loaded");
  Ciphertext a = a0, b = b0;
  PrepareMult(a, b);
  Ciphertext o = NewCiphertext(a->limbs);
  chk(fides_hmult(ctx, a->d, b->d, a->limbs, evalMultKey, o->d, nullptr), "EvalMult");
  MultMetadata(*o, *a, *b);
  return o;
}

void CryptoContextImpl::EvalMultInPlace(Ciphertext& a, const Ciphertext& b) { a = EvalMult(a, b); }

Ciphertext CryptoContextImpl::EvalSquare(const Ciphertext& a0) {
  if (!evalMultKey) throw std::runtime_error("EvalSquare: relinearization key not loaded");
  Ciphertext a = a0;
  if (technique != FIXEDMANUAL && a->NoiseLevel == 2) a = Rescale(a);
  Ciphertext o = NewCiphertext(a->limbs);
  chk(fides_square(ctx, a->d, a->limbs, evalMultKey, o->d, nullptr), "EvalSquare");
  MultMetadata(*o, *a, *a);
  return o;
}

void CryptoContextImpl::EvalSquareInPlace(Ciphertext& a) { a = EvalSquare(a); }

Ciphertext CryptoContextImpl::Rescale(const Ciphertext& a) {
  if (a->limbs < 2) throw std::runtime_error("Rescale: no limb left to drop");
  Ciphertext o = NewCiphertext(a->limbs - 1);
  chk(fides_rescale(ctx, a->d, a->limbs, o->d, nullptr), "Rescale");
  o->NoiseLevel = a->NoiseLevel - 1;
  o->NoiseFactor = a->NoiseFactor / modReduceFactor[a->limbs - 1];
  o->slots = a->slots;
  o->keyid = a->keyid;
  return o;
}

void CryptoContextImpl::RescaleInPlace(Ciphertext& a) { a = Rescale(a); }

Ciphertext CryptoContextImpl::EvalRotate(const Ciphertext& a, int32_t index) {
  const uint32_t k = AutoIndex(index);
  if (k == 1) {
    Ciphertext o = DropToLimbs(a, a->limbs);
    return o;
  }
  Ciphertext o = NewCiphertext(a->limbs);
  chk(fides_rotate(ctx, a->d, a->limbs, RotationKey(k), k, o->d, nullptr), "EvalRotate");
  o->NoiseLevel = a->NoiseLevel;
  o->NoiseFactor = a->NoiseFactor;
  o->slots = a->slots;
  o->keyid = a->keyid;
  return o;
}

void CryptoContextImpl::EvalRotateInPlace(Ciphertext& a, int32_t index) { a = EvalRotate(a, index); }

Ciphertext CryptoContextImpl::EvalConjugate(const Ciphertext& a) {
  const uint32_t k = 2 * N - 1;
  Ciphertext o = NewCiphertext(a->limbs);
  chk(fides_rotate(ctx, a->d, a->limbs, RotationKey(k), k, o->d, nullptr), "EvalConjugate");
  o->NoiseLevel = a->NoiseLevel;
  o->NoiseFactor = a->NoiseFactor;
  o->slots = a->slots;
  o->keyid = a->keyid;
  return o;
}

std::vector<Ciphertext> CryptoContextImpl::EvalMultMany(const std::vector<Ciphertext>& a,
                                                        const std::vector<Ciphertext>& b, bool rescale) {
  if (!evalMultKey) throw std::runtime_error("EvalMultMany: relinearization key not loaded");
  if (a.size() != b.size()) throw std::runtime_error("EvalMultMany: operand counts differ");
  std::vector<Ciphertext> A(a), B(b), out;
  std::vector<const uint64_t*> pa, pb;
  std::vector<uint64_t*> po;
  std::vector<uint32_t> lv;
  for (size_t i = 0; i < a.size(); ++i) {
    PrepareMult(A[i], B[i]);
    if (rescale && A[i]->limbs < 2) throw std::runtime_error("EvalMultMany: cannot rescale a 1-limb product");
    Ciphertext o = NewCiphertext(rescale ? A[i]->limbs - 1 : A[i]->limbs);
    MultMetadata(*o, *A[i], *B[i]);
    if (rescale) {
      o->NoiseLevel -= 1;
      o->NoiseFactor /= modReduceFactor[A[i]->limbs - 1];
    }
    pa.push_back(A[i]->d);
    pb.push_back(B[i]->d);
    po.push_back(o->d);
    lv.push_back(A[i]->limbs);
    out.push_back(o);
  }
  chk(fides_hmult_batch(ctx, (uint32_t)a.size(), pa.data(), pb.data(), lv.data(), evalMultKey, rescale ? 1 : 0,
                        po.data(), nullptr),
      "EvalMultMany");
  return out;
}

std::shared_ptr<void> CryptoContextImpl::EvalFastRotationPrecompute(const Ciphertext&) {
  // The device path raises the digits inside EvalFastRotation (as the
  // reference's GPU path does, api/CryptoContext.cpp:1451-1465).
  return nullptr;
}

std::vector<Ciphertext> CryptoContextImpl::EvalFastRotation(const Ciphertext& a, const std::vector<int32_t>& indices,
                                                            uint32_t, const std::shared_ptr<void>&) {
  std::vector<uint64_t*> po;
  std::vector<uint32_t> kv;
  std::vector<const fides_key*> keys;
  std::vector<Ciphertext> out;
  for (int32_t idx : indices) {
    const uint32_t k = AutoIndex(idx);
    Ciphertext o = NewCiphertext(a->limbs);
    o->NoiseLevel = a->NoiseLevel;
    o->NoiseFactor = a->NoiseFactor;
    o->slots = a->slots;
    o->keyid = a->keyid;
    po.push_back(o->d);
    kv.push_back(k);
    keys.push_back(k == 1 ? nullptr : RotationKey(k));
    out.push_back(o);
  }
  chk(fides_rotate_hoisted(ctx, a->d, a->limbs, (uint32_t)indices.size(), keys.data(), kv.data(), po.data(), nullptr),
      "EvalFastRotation");
  return out;
}

Ciphertext CryptoContextImpl::EvalFastRotation(const Ciphertext& a, int32_t index, uint32_t m,
                                               const std::shared_ptr<void>& precomp) {
  return EvalFastRotation(a, std::vector<int32_t>{index}, m, precomp)[0];
}

std::vector<Ciphertext> CryptoContextImpl::EvalRotateMany(const std::vector<Ciphertext>& a,
                                                          const std::vector<int32_t>& idx) {
  if (a.size() != idx.size()) throw std::runtime_error("EvalRotateMany: counts differ");
  std::vector<const uint64_t*> pin;
  std::vector<uint64_t*> po;
  std::vector<uint32_t> lv, kv;
  std::vector<const fides_key*> keys;
  std::vector<Ciphertext> out;
  for (size_t i = 0; i < a.size(); ++i) {
    const uint32_t k = AutoIndex(idx[i]);
    if (k == 1) throw std::runtime_error("EvalRotateMany: rotation by 0");
    Ciphertext o = NewCiphertext(a[i]->limbs);
    o->NoiseLevel = a[i]->NoiseLevel;
    o->NoiseFactor = a[i]->NoiseFactor;
    o->slots = a[i]->slots;
    o->keyid = a[i]->keyid;
    pin.push_back(a[i]->d);
    po.push_back(o->d);
    lv.push_back(a[i]->limbs);
    kv.push_back(k);
    keys.push_back(RotationKey(k));
    out.push_back(o);
  }
  chk(fides_rotate_batch(ctx, (uint32_t)a.size(), pin.data(), lv.data(), keys.data(), kv.data(), po.data(), nullptr),
      "EvalRotateMany");
  return out;
}

void CryptoContextImpl::Synchronize() const { chk(fides_stream_sync(nullptr), "Synchronize"); }

}  // namespace fideslib
