// Host C++ API (include/fideslib/fideslib.hpp) over the C ABI.
//
// Metadata rules follow FIDESlib's Ciphertext (src/CKKS/Ciphertext.cpp):
//   mult:    NoiseLevel = a + b, NoiseFactor = a * b           (:131-143 multMetadata)
//   rescale: NoiseLevel -= 1, NoiseFactor /= ModReduceFactor   (:426-447)
//   FIXEDAUTO: operands at scale degree 2 are rescaled before a mult and
//   levels are aligned by dropping limbs (:1483-1518 adjustForAddOrSub /
//   adjustForMult, FIXED branch).
// FLEXIBLEAUTO(EXT): OpenFHE's per-level real scaling factors come in
// RawParams (or follow OpenFHE's rule by default) and operands are aligned by
// the scalar-multiply-and-rescale correction of adjustScaleAndLevel
// (:1376-1455).
#include "fideslib/fideslib.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "fideslib_amd.h"
#include "program.hpp"

namespace fideslib {

static void chk(int rc, const char* what) {
  if (rc != FIDES_OK) throw std::runtime_error(std::string(what) + ": " + fides_last_error());
}

// Device memory of ciphertexts and plaintexts comes from the context's cache
// of idle blocks (one allocation per size; a released block waits for the
// next operation of its size instead of being freed, which would
// synchronize the device).  Every operation of the C++ API runs on the
// default stream, so a block released on the host is reused only by work
// queued after everything that used it.
CiphertextImpl::CiphertextImpl(std::shared_ptr<CryptoContextImpl> c, uint32_t l, bool e)
    : limbs(l), cc(std::move(c)), ext(e) {
  const size_t nl = e ? l + cc->GetNumSpecialPrimes() : l;
  words = (size_t)2 * nl * cc->GetRingDimension();
  d = cc->TakeBlock(words);
}

CiphertextImpl::~CiphertextImpl() {
  if (d) cc->GiveBlock(d, words);
}

PlaintextImpl::PlaintextImpl(std::shared_ptr<CryptoContextImpl> c, uint32_t l) : limbs(l), cc(std::move(c)) {
  words = (size_t)l * cc->GetRingDimension();
  d = cc->TakeBlock(words);
}

PlaintextImpl::~PlaintextImpl() {
  if (d) cc->GiveBlock(d, words);
}

uint64_t* CryptoContextImpl::TakeBlock(size_t w) {
  if (!blockMem) blockMem = std::make_shared<detail::Pool>();
  return std::static_pointer_cast<detail::Pool>(blockMem)->get(w);
}

void CryptoContextImpl::GiveBlock(uint64_t* d, size_t w) {
  std::static_pointer_cast<detail::Pool>(blockMem)->free.emplace(w, d);
}

void CryptoContextImpl::ReleaseCachedMemory() {
  Synchronize();
  for (std::shared_ptr<void>* m : {&blockMem, &bootMem}) {
    if (!*m) continue;
    auto P = std::static_pointer_cast<detail::Pool>(*m);
    for (auto& kv : P->free) fides_device_free(kv.second);
    P->free.clear();
  }
}

CryptoContext CryptoContextImpl::Create(const RawParams& P) {
  if (P.moduli.empty() || P.SPECIALmoduli.empty()) throw std::runtime_error("RawParams: empty moduli");
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
  c->raw = P;
  c->modReduceFactor = P.ModReduceFactor;
  if (c->modReduceFactor.size() < c->L) {
    c->modReduceFactor.resize(c->L);
    for (uint32_t i = 0; i < c->L; ++i) c->modReduceFactor[i] = (double)c->q[i];
  }
  c->sfReal = P.ScalingFactorReal;
  c->sfBig = P.ScalingFactorRealBig;
  if (c->sfReal.size() < c->L || c->sfBig.size() < c->L) {
    c->sfReal.assign(c->L, 0.0);
    c->sfBig.assign(c->L, 0.0);
    if (c->Flexible()) {
      c->sfReal[c->L - 1] = (double)c->q[c->L - 1];
      for (uint32_t l = c->L - 1; l >= 1; --l) c->sfReal[l - 1] = c->sfReal[l] * c->sfReal[l] / (double)c->q[l];
    } else {
      const double p = std::ldexp(1.0, (int)std::lround(std::log2((double)c->q[c->L > 1 ? 1 : 0])));
      for (uint32_t l = 0; l < c->L; ++l) c->sfReal[l] = p;
    }
    for (uint32_t l = 0; l < c->L; ++l) c->sfBig[l] = c->sfReal[l] * c->sfReal[l];
  }
  return c;
}

CryptoContextImpl::~CryptoContextImpl() {
  if (evalMultKey) fides_key_destroy(evalMultKey);
  if (encapDown) fides_key_destroy(encapDown);
  if (encapUp) fides_key_destroy(encapUp);
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

static RawKeySwitchKey download(const fides_key* key, uint32_t dnum, uint32_t LK, uint32_t N) {
  std::vector<uint64_t> flat((size_t)2 * dnum * LK * N);
  chk(fides_key_download(key, flat.data()), "key download");
  RawKeySwitchKey k;
  k.r_key.resize(2);
  for (size_t c = 0; c < 2; ++c) {
    k.r_key[c].resize(dnum);
    for (size_t j = 0; j < dnum; ++j) {
      k.r_key[c][j].resize(LK);
      for (size_t i = 0; i < LK; ++i) {
        const auto b = flat.begin() + ((c * dnum + j) * LK + i) * N;
        k.r_key[c][j][i].assign(b, b + N);
      }
    }
  }
  return k;
}

RawKeySwitchKey CryptoContextImpl::GetEvalMultKey() const {
  if (!evalMultKey) throw std::runtime_error("no relinearization key loaded");
  return download(evalMultKey, dnum, L + K, N);
}

std::map<uint32_t, RawKeySwitchKey> CryptoContextImpl::GetRotationKeys() const {
  std::map<uint32_t, RawKeySwitchKey> out;
  for (const auto& kv : rotationKeys) out[kv.first] = download(kv.second, dnum, L + K, N);
  return out;
}

void CryptoContextImpl::AddAutomorphismKey(uint32_t a, const RawKeySwitchKey& k) {
  if ((a & 1) == 0 || a >= 2 * N) throw std::runtime_error("automorphism index must be odd and < 2N");
  auto it = rotationKeys.find(a);
  if (it != rotationKeys.end()) fides_key_destroy(it->second);
  rotationKeys[a] = upload(ctx, k, dnum, L + K, N);
}

void CryptoContextImpl::AddEncapsulationKeys(const RawKeySwitchKey& down, const RawKeySwitchKey& up) {
  fides_key* d = upload(ctx, down, dnum, L + K, N);
  fides_key* u = nullptr;
  try {
    u = upload(ctx, up, dnum, L + K, N);
  } catch (...) {
    fides_key_destroy(d);
    throw;
  }
  if (encapDown) fides_key_destroy(encapDown);
  if (encapUp) fides_key_destroy(encapUp);
  encapDown = d;
  encapUp = u;
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

Ciphertext CryptoContextImpl::NewCiphertext(uint32_t limbs, bool ext) {
  return Ciphertext(new CiphertextImpl(shared_from_this(), limbs, ext));
}

static void copy_meta(CiphertextImpl& o, const CiphertextImpl& a) {
  o.NoiseLevel = a.NoiseLevel;
  o.NoiseFactor = a.NoiseFactor;
  o.slots = a.slots;
  o.keyid = a.keyid;
}

Ciphertext CryptoContextImpl::Regular(const Ciphertext& a) {
  if (!a->ext) return a;
  Ciphertext o = NewCiphertext(a->limbs);
  const uint64_t* in = a->d;
  uint64_t* out = o->d;
  const uint32_t l = a->limbs;
  chk(fides_ext_moddown_batch(ctx, 1, &in, &l, &out, nullptr), "ModDown");
  copy_meta(*o, *a);
  return o;
}

Ciphertext CryptoContextImpl::Extended(const Ciphertext& a) {
  if (a->ext) return a;
  Ciphertext o = NewCiphertext(a->limbs, true);
  const uint64_t* in = a->d;
  uint64_t* out = o->d;
  const uint32_t l = a->limbs;
  chk(fides_ext_raise_batch(ctx, 1, &in, &l, &out, nullptr), "scale by P");
  copy_meta(*o, *a);
  return o;
}

// Sum in Q_l u P of two ciphertexts at one level and scale degree, at least
// one of them extended (the other is scaled by P first, as Ciphertext::add
// extends its operand in the reference's Accumulate).
Ciphertext CryptoContextImpl::AddSubExt(const Ciphertext& a0, const Ciphertext& b0, bool sub) {
  const Ciphertext a = Extended(a0), b = Extended(b0);
  Ciphertext o = NewCiphertext(a->limbs, true);
  const uint64_t* pa = a->d;
  const uint64_t* pb = b->d;
  uint64_t* po = o->d;
  const uint32_t l = a->limbs;
  chk(fides_ext_add_batch(ctx, 1, &pa, &pb, &l, sub ? 1 : 0, &po, nullptr), sub ? "EvalSub" : "EvalAdd");
  copy_meta(*o, *a);
  o->slots = std::max(a->slots, b->slots);
  return o;
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

Plaintext CryptoContextImpl::LoadPlaintext(const RawPlainText& raw) {
  const uint32_t l = (uint32_t)raw.sub_0.size();
  if (l == 0 || l > L) throw std::runtime_error("RawPlainText: bad limb count");
  if (!raw.moduli.empty()) {
    if (raw.moduli.size() != l) throw std::runtime_error("RawPlainText: moduli count");
    for (uint32_t i = 0; i < l; ++i)
      if (raw.moduli[i] != q[i]) throw std::runtime_error("RawPlainText: moduli differ from the context");
  }
  std::vector<uint64_t> flat((size_t)l * N);
  for (uint32_t i = 0; i < l; ++i) {
    if (raw.sub_0[i].size() != N) throw std::runtime_error("RawPlainText: limb size");
    for (uint64_t v : raw.sub_0[i])
      if (v >= q[i]) throw std::runtime_error("RawPlainText: residue not reduced");
    std::copy(raw.sub_0[i].begin(), raw.sub_0[i].end(), flat.begin() + (size_t)i * N);
  }
  Plaintext pt(new PlaintextImpl(shared_from_this(), l));
  chk(fides_copy_to_device(pt->d, flat.data(), flat.size() * 8, nullptr), "plaintext upload");
  if (raw.format == EVALUATION) {  // back to coefficients: the preparation transforms forward
    std::vector<uint32_t> pids(l);
    for (uint32_t i = 0; i < l; ++i) pids[i] = i;
    chk(fides_intt(ctx, pt->d, pids.data(), l, nullptr), "plaintext INTT");
  }
  chk(fides_plaintext_prepare(ctx, pt->d, l, nullptr), "plaintext preparation");
  pt->NoiseLevel = raw.NoiseLevel;
  pt->NoiseFactor = raw.Noise;
  pt->slots = raw.slots;
  return pt;
}

Ciphertext CryptoContextImpl::EvalMult(const Ciphertext& a0, const Plaintext& p) {
  Ciphertext a = Regular(a0);
  if (technique != FIXEDMANUAL && a->NoiseLevel == 2) a = Rescale(a);
  if (p->limbs < a->limbs) {
    if (technique == FIXEDMANUAL || Flexible())
      throw std::runtime_error("EvalMult(plaintext): plaintext below the ciphertext level");
    a = DropToLimbs(a, p->limbs);  // FIXEDAUTO: align down (Plaintext::adjustPlaintextToCiphertext)
  }
  const uint32_t l = a->limbs, one = 1;
  Ciphertext o = NewCiphertext(l);
  const uint64_t* ct = a->d;
  const uint64_t* pp = p->d;
  uint64_t* out = o->d;
  chk(fides_ptdot_batch(ctx, 1, &one, &ct, &pp, &l, &out, nullptr), "EvalMult(plaintext)");
  o->NoiseLevel = a->NoiseLevel + p->NoiseLevel;
  o->NoiseFactor = a->NoiseFactor * p->NoiseFactor;
  o->slots = a->slots;
  o->keyid = a->keyid;
  return o;
}

RawCipherText CryptoContextImpl::StoreCiphertext(const Ciphertext& ct0) {
  const Ciphertext ct = Regular(ct0);
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

Ciphertext CryptoContextImpl::DropToLimbs(const Ciphertext& a0, uint32_t limbs) {
  const Ciphertext a = Regular(a0);
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

// ---------------------------------------------------------------- scalars
// round(operand * scale) as residues mod q_0..q_{limbs-1}: rounded by adding
// 1/2 and truncating to a 128-bit integer, with magnitudes beyond 2^125
// divided by 2^k first and multiplied back in RNS (reference
// ContextData::ElemForEvalMult, src/CKKS/Context.cu:366-463).
static uint64_t mulmod_host(uint64_t a, uint64_t b, uint64_t q) {
  return (uint64_t)(((unsigned __int128)a * b) % q);
}
static uint64_t pow2mod(int k, uint64_t q) {
  uint64_t r = 1 % q, b = 2 % q;
  for (; k > 0; k >>= 1, b = mulmod_host(b, b, q))
    if (k & 1) r = mulmod_host(r, b, q);
  return r;
}

std::vector<uint64_t> CryptoContextImpl::ScalarResidues(uint32_t limbs, double operand, double scale) const {
  int logApprox = 0;
  const double mag = std::fabs(operand * scale);
  if (mag > 0) {
    const int logSF = (int)std::ceil(std::log2(mag));
    logApprox = logSF - std::min(logSF, 125);
  }
  const __int128 large = (__int128)(operand / std::ldexp(1.0, logApprox) * scale + 0.5);
  std::vector<uint64_t> r(limbs);
  for (uint32_t i = 0; i < limbs; ++i) {
    __int128 x = large % (__int128)q[i];
    if (x < 0) x += q[i];
    r[i] = (uint64_t)x;
    if (logApprox > 0) r[i] = mulmod_host(r[i], pow2mod(logApprox, q[i]), q[i]);
  }
  return r;
}

// The constant of an addition: round(|c| * sf) (64-bit, the reference's
// ElemForEvalAddOrSub, src/CKKS/Context.cu:468-560, with OpenFHE's
// MAX_BITS_IN_WORD = 61 (recalled; OpenFHE is not available here)), times
// round(sf) for every scale degree above one, negated for c < 0
// (Ciphertext::addScalar, src/CKKS/Ciphertext.cpp:780-798).
std::vector<uint64_t> CryptoContextImpl::AddScalarResidues(uint32_t limbs, double c, int noiseDeg) const {
  const uint32_t level = limbs - 1;
  const bool extTop = technique == FLEXIBLEAUTOEXT && level == L - 1;
  const double sf = extTop ? sfBig.at(level) : sfReal.at(level);
  int logApprox = 0;
  const double mag = std::fabs(c) * sf;
  if (mag > 0) {
    const int logSF = (int)std::ceil(std::log2(mag));
    logApprox = logSF - std::min(logSF, 61);
  }
  const uint64_t base = (uint64_t)(std::fabs(c) * sf / std::ldexp(1.0, logApprox) + 0.5);
  const uint64_t isf = (uint64_t)(sf + 0.5);
  std::vector<uint64_t> r(limbs);
  for (uint32_t i = 0; i < limbs; ++i) {
    uint64_t x = base % q[i];
    if (logApprox > 0) x = mulmod_host(x, pow2mod(logApprox, q[i]), q[i]);
    if (!extTop)
      for (int d = 1; d < noiseDeg; ++d) x = mulmod_host(x, isf % q[i], q[i]);
    if (c < 0 && x) x = q[i] - x;
    r[i] = x;
  }
  return r;
}

Ciphertext CryptoContextImpl::MulConst(const Ciphertext& a0, const std::vector<uint64_t>& res) {
  const Ciphertext a = Regular(a0);
  Ciphertext o = NewCiphertext(a->limbs);
  const uint64_t* in = a->d;
  uint64_t* out = o->d;
  const uint64_t* rp = res.data();
  const uint32_t l = a->limbs;
  chk(fides_mul_const_batch(ctx, 1, &in, &l, &rp, &out, nullptr), "scalar multiplication");
  o->NoiseLevel = a->NoiseLevel;
  o->NoiseFactor = a->NoiseFactor;
  o->slots = a->slots;
  o->keyid = a->keyid;
  return o;
}

// Ciphertext::multScalarNoPrecheck (src/CKKS/Ciphertext.cpp:747-762): the
// constant is encoded at the level's real scaling factor; scale degree + 1.
Ciphertext CryptoContextImpl::MultScalarNoPrecheck(const Ciphertext& a, double c) {
  const double sf = sfReal.at(a->limbs - 1);
  Ciphertext o = MulConst(a, ScalarResidues(a->limbs, c, sf));
  o->NoiseLevel = a->NoiseLevel + 1;
  o->NoiseFactor = a->NoiseFactor * sf;
  return o;
}

Ciphertext CryptoContextImpl::MultScalar(const Ciphertext& a0, double c) {
  Ciphertext a = Regular(a0);
  if (technique != FIXEDMANUAL && a->NoiseLevel == 2) a = Rescale(a);
  if (a->NoiseLevel != 1) throw std::runtime_error("EvalMult(scalar): operand at scale degree > 1 (FIXEDMANUAL)");
  return MultScalarNoPrecheck(a, c);
}

// ---------------------------------------------------------------- level / scale alignment
// Brings `a` (at a higher level) to `level` and to the given scale degree and
// scaling factor, as Ciphertext::adjustScaleAndLevel
// (src/CKKS/Ciphertext.cpp:1376-1455): a scalar multiplication by the ratio
// of the scaling factors merged with the Rescale that drops the level.
// Returns false when `a` is the lower of the two (the caller swaps roles).
bool CryptoContextImpl::AdjustScaleAndLevel(Ciphertext& a, int scaleDegree, uint32_t level, double scalingFactor) {
  const uint32_t l1 = a->limbs - 1, l2 = level;
  const int d1 = a->NoiseLevel, d2 = scaleDegree;
  if (l1 > l2) {
    if (d1 == 2) {
      if (d2 == 2) {
        const double scf1 = a->NoiseFactor, scf2 = scalingFactor, scf = sfReal.at(l1);
        const double q1 = modReduceFactor.at(l2 + 1);
        a = MultScalarNoPrecheck(a, scf2 * q1 / scf1 / scf);
        if (a->limbs - 1 > l2 + 1) a = DropToLimbs(a, l2 + 2);
        a->NoiseFactor = sfBig.at(l2) * modReduceFactor.at(l2 + 1);
        a = Rescale(a);
        a->NoiseFactor = scalingFactor;
      } else if (l1 - 1 == l2) {
        a = Rescale(a);
      } else {
        const double scf1 = a->NoiseFactor, scf2 = sfBig.at(l2 + 1), scf = sfReal.at(l1);
        const double q1 = modReduceFactor.at(l1);
        a = MultScalarNoPrecheck(a, scf2 / scf1 * q1 / scf);
        a->NoiseFactor = sfBig.at(a->limbs - 2) * modReduceFactor.at(l1);
        a = Rescale(a);
        if (a->limbs - 2 > l2) a = DropToLimbs(a, l2 + 2);
        a->NoiseFactor = sfBig.at(a->limbs - 1);
        a = Rescale(a);
        a->NoiseFactor = scalingFactor;
      }
    } else {
      if (d2 == 2) {
        const double scf1 = a->NoiseFactor, scf2 = scalingFactor, scf = sfReal.at(l1);
        a = MultScalarNoPrecheck(a, scf2 / scf1 / scf);
        a = DropToLimbs(a, l2 + 1);
        a->NoiseFactor = scf2;
      } else {
        const double scf1 = a->NoiseFactor, scf2 = sfBig.at(l2 + 1), scf = sfReal.at(l1);
        a = MultScalarNoPrecheck(a, scf2 / scf1 / scf);
        if (l1 - 1 > l2) a = DropToLimbs(a, l2 + 2);
        a = Rescale(a);
        a->NoiseFactor = scalingFactor;
      }
    }
    return true;
  }
  if (l1 < l2) return false;
  if (d1 < d2) {
    a = MultScalarNoPrecheck(a, 1.0);
  } else if (d2 < d1) {
    return false;
  }
  return true;
}

// Ciphertext::adjustForAddOrSub (src/CKKS/Ciphertext.cpp:1457-1490).
bool CryptoContextImpl::AdjustForAddOrSub(Ciphertext& a, const Ciphertext& b) {
  if (Flexible()) return AdjustScaleAndLevel(a, b->NoiseLevel, b->limbs - 1, b->NoiseFactor);
  const int la = (int)a->limbs - 1 - a->NoiseLevel, lb = (int)b->limbs - 1 - b->NoiseLevel;
  if (la > lb) {
    if (b->NoiseLevel == 1 && a->NoiseLevel == 2)
      a = Rescale(a);
    else if (b->NoiseLevel == 2 && a->NoiseLevel == 1)
      a = MultScalarNoPrecheck(a, 1.0);
    return true;
  }
  if (b->NoiseLevel == 1 && a->NoiseLevel == 2) {
    a = Rescale(a);
    return true;
  }
  return !(a->NoiseLevel == 1 && b->NoiseLevel == 2);
}

// Ciphertext::adjustForMult (src/CKKS/Ciphertext.cpp:1492-1508).
bool CryptoContextImpl::AdjustForMult(Ciphertext& a, const Ciphertext& b) {
  const bool ok = AdjustForAddOrSub(a, b);
  if (a->NoiseLevel == 2) a = Rescale(a);
  return ok && b->NoiseLevel != 2;
}

Ciphertext CryptoContextImpl::AddSub(const Ciphertext& a0, const Ciphertext& b0, bool sub) {
  if ((a0->ext || b0->ext) && a0->limbs == b0->limbs && a0->NoiseLevel == b0->NoiseLevel)
    return AddSubExt(a0, b0, sub);
  Ciphertext a = Regular(a0), b = Regular(b0);
  if (technique != FIXEDMANUAL) {
    if (!AdjustForAddOrSub(a, b)) {
      if (!AdjustForAddOrSub(b, a)) throw std::runtime_error("EvalAdd/EvalSub: cannot align operands");
    }
  } else if (a->NoiseLevel != b->NoiseLevel) {
    throw std::runtime_error("EvalAdd/EvalSub: operands at different scale degrees (FIXEDMANUAL)");
  }
  const uint32_t l = std::min(a->limbs, b->limbs);
  if (a->limbs > l) a = DropToLimbs(a, l);
  if (b->limbs > l) b = DropToLimbs(b, l);
  Ciphertext o = NewCiphertext(l);
  const uint64_t* pa = a->d;
  const uint64_t* pb = b->d;
  uint64_t* po = o->d;
  chk(fides_add_batch(ctx, 1, &pa, &pb, &l, sub ? 1 : 0, &po, nullptr), sub ? "EvalSub" : "EvalAdd");
  o->NoiseLevel = std::max(a->NoiseLevel, b->NoiseLevel);
  o->NoiseFactor = a->NoiseFactor;
  o->slots = a->slots;
  o->keyid = a->keyid;
  return o;
}

Ciphertext CryptoContextImpl::EvalAdd(const Ciphertext& a, const Ciphertext& b) { return AddSub(a, b, false); }
Ciphertext CryptoContextImpl::EvalSub(const Ciphertext& a, const Ciphertext& b) { return AddSub(a, b, true); }

Ciphertext CryptoContextImpl::EvalAdd(const Ciphertext& a0, double c) {
  const Ciphertext a = Regular(a0);
  Ciphertext o = NewCiphertext(a->limbs);
  const std::vector<uint64_t> res = AddScalarResidues(a->limbs, c, a->NoiseLevel);
  const uint64_t* in = a->d;
  uint64_t* out = o->d;
  const uint64_t* rp = res.data();
  const uint32_t l = a->limbs;
  chk(fides_add_const_batch(ctx, 1, &in, &l, &rp, &out, nullptr), "EvalAdd(scalar)");
  o->NoiseLevel = a->NoiseLevel;
  o->NoiseFactor = a->NoiseFactor;
  o->slots = a->slots;
  o->keyid = a->keyid;
  return o;
}

Ciphertext CryptoContextImpl::EvalSub(const Ciphertext& a, double c) { return EvalAdd(a, -c); }

// Ciphertext::multScalar(c) with its default rescale = false
// (src/CKKS/Ciphertext.cpp:764-777, api/CryptoContext.cpp:1244-1266): the
// product stays at scale degree 2 under every technique; automatic
// techniques rescale it before its next multiplication.
Ciphertext CryptoContextImpl::EvalMult(const Ciphertext& a, double c) { return MultScalar(a, c); }

Ciphertext CryptoContextImpl::EvalNegate(const Ciphertext& a) {
  std::vector<uint64_t> res(a->limbs);
  for (uint32_t i = 0; i < a->limbs; ++i) res[i] = q[i] - 1;
  return MulConst(a, res);
}

void CryptoContextImpl::PrepareMult(Ciphertext& a, Ciphertext& b) {
  if (technique == FIXEDMANUAL) {
    if (a->limbs != b->limbs) throw std::runtime_error("EvalMult: operands at different levels (FIXEDMANUAL)");
    return;
  }
  // Ciphertext::mult's alignment (src/CKKS/Ciphertext.cpp:458-477): adjust a
  // to b, or else b to the adjusted a
  if (!AdjustForMult(a, b)) {
    if (!AdjustForMult(b, a)) throw std::runtime_error("EvalMult: cannot align operands");
  }
  if (a->NoiseLevel == 2) a = Rescale(a);
  if (b->NoiseLevel == 2) b = Rescale(b);
  if (Flexible()) {
    if (a->limbs != b->limbs) throw std::runtime_error("EvalMult: alignment left different levels");
    return;
  }
  // FIXEDAUTO's rules leave the levels apart: the reference multiplies at the
  // first operand's level (reading the other's first limbs), OpenFHE drops
  // the higher operand; the latter, which agrees wherever the former is defined
  if (a->limbs > b->limbs) a = DropToLimbs(a, b->limbs);
  if (b->limbs > a->limbs) b = DropToLimbs(b, a->limbs);
}

Ciphertext CryptoContextImpl::EvalMult(const Ciphertext& a0, const Ciphertext& b0) {
  if (!evalMultKey) throw std::runtime_error("EvalMult: relinearization key not loaded");
  Ciphertext a = Regular(a0), b = Regular(b0);
  PrepareMult(a, b);
  Ciphertext o = NewCiphertext(a->limbs);
  chk(fides_hmult(ctx, a->d, b->d, a->limbs, evalMultKey, o->d, nullptr), "EvalMult");
  MultMetadata(*o, *a, *b);
  return o;
}

void CryptoContextImpl::EvalMultInPlace(Ciphertext& a, const Ciphertext& b) { a = EvalMult(a, b); }

Ciphertext CryptoContextImpl::EvalSquare(const Ciphertext& a0) {
  if (!evalMultKey) throw std::runtime_error("EvalSquare: relinearization key not loaded");
  Ciphertext a = Regular(a0);
  if (technique != FIXEDMANUAL && a->NoiseLevel == 2) a = Rescale(a);
  Ciphertext o = NewCiphertext(a->limbs);
  chk(fides_square(ctx, a->d, a->limbs, evalMultKey, o->d, nullptr), "EvalSquare");
  MultMetadata(*o, *a, *a);
  return o;
}

void CryptoContextImpl::EvalSquareInPlace(Ciphertext& a) { a = EvalSquare(a); }

Ciphertext CryptoContextImpl::Rescale(const Ciphertext& a0) {
  const Ciphertext a = Regular(a0);
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

Ciphertext CryptoContextImpl::EvalRotate(const Ciphertext& a0, int32_t index) {
  const Ciphertext a = Regular(a0);
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

Ciphertext CryptoContextImpl::EvalConjugate(const Ciphertext& a0) {
  const Ciphertext a = Regular(a0);
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
  std::vector<Ciphertext> A, B, out;
  for (const auto& x : a) A.push_back(Regular(x));
  for (const auto& x : b) B.push_back(Regular(x));
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

std::vector<Ciphertext> CryptoContextImpl::EvalFastRotation(const Ciphertext& a0, const std::vector<int32_t>& indices,
                                                            uint32_t, const std::shared_ptr<void>&) {
  const Ciphertext a = Regular(a0);
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

std::vector<Ciphertext> CryptoContextImpl::EvalRotateMany(const std::vector<Ciphertext>& a0,
                                                          const std::vector<int32_t>& idx) {
  std::vector<Ciphertext> a;
  for (const auto& x : a0) a.push_back(Regular(x));
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

static CryptoContextImpl::Tuning from_c(const fides_tuning& t) {
  CryptoContextImpl::Tuning r;
  for (int c = 0; c < 3; ++c) {
    r.convChunk[c] = t.conv_chunk[c];
    r.groupByPrime[c] = t.group_by_prime[c];
    r.fuseRowDot[c] = t.fuse_row_dot[c];
  }
  return r;
}

CryptoContextImpl::Tuning CryptoContextImpl::GetTuning() const {
  fides_tuning t;
  chk(fides_tuning_get(ctx, &t), "GetTuning");
  return from_c(t);
}

void CryptoContextImpl::SetTuning(const Tuning& r) {
  fides_tuning t;
  for (int c = 0; c < 3; ++c) {
    t.conv_chunk[c] = r.convChunk[c];
    t.group_by_prime[c] = r.groupByPrime[c];
    t.fuse_row_dot[c] = r.fuseRowDot[c];
  }
  chk(fides_tuning_set(ctx, &t), "SetTuning");
}

CryptoContextImpl::Tuning CryptoContextImpl::Autotune(const std::vector<uint32_t>& limbs, uint32_t reps) {
  fides_tuning t;
  chk(fides_autotune(ctx, (uint32_t)limbs.size(), limbs.data(), reps, &t), "Autotune");
  return from_c(t);
}

// ---------------------------------------------------------------- plaintext addition
Ciphertext CryptoContextImpl::EmbedPlaintext(const Plaintext& p) {
  Ciphertext o = NewCiphertext(p->limbs);
  chk(fides_plaintext_embed(ctx, p->d, p->limbs, o->d, nullptr), "plaintext embedding");
  o->NoiseLevel = p->NoiseLevel;
  o->NoiseFactor = p->NoiseFactor;
  o->slots = p->slots;
  return o;
}

// Ciphertext::addPt / subPt (src/CKKS/Ciphertext.cpp:242-300) with
// Plaintext::adjustPlaintextToCiphertext (src/CKKS/Plaintext.cu:195-226): the
// plaintext is handled as the ciphertext (pt, 0), so its rescale and scalar
// multiplication are the ciphertext kernels and c1 gets + 0.
Ciphertext CryptoContextImpl::AddSubPt(const Ciphertext& a0, const Plaintext& p, bool sub) {
  Ciphertext a = Regular(a0);
  Ciphertext pc = EmbedPlaintext(p);
  if (technique != FIXEDMANUAL) {
    if (pc->NoiseLevel == 1 && a->NoiseLevel == 2 && pc->limbs + 1 == a->limbs) a = Rescale(a);
    if (pc->limbs != a->limbs) {
      if (technique == FIXEDAUTO) {
        const int pl = (int)pc->limbs - 1 - pc->NoiseLevel, cl = (int)a->limbs - 1 - a->NoiseLevel;
        if (pl > cl) {
          if (a->NoiseLevel == 1 && pc->NoiseLevel == 2) {
            pc = Rescale(DropToLimbs(pc, a->limbs + 1));
          } else {
            pc = DropToLimbs(pc, a->limbs);
          }
        } else if (a->NoiseLevel == 1 && pc->NoiseLevel == 2) {
          pc = Rescale(pc);
        } else if (pc->NoiseLevel == 1 && a->NoiseLevel == 2) {
          throw std::runtime_error("EvalAdd(plaintext): cannot align the plaintext");
        }
      } else {
        if (pc->limbs < a->limbs || !AdjustScaleAndLevel(pc, a->NoiseLevel, a->limbs - 1, a->NoiseFactor))
          throw std::runtime_error("EvalAdd(plaintext): plaintext below the ciphertext level");
      }
    }
  }
  if (pc->NoiseLevel != a->NoiseLevel)
    throw std::runtime_error("EvalAdd(plaintext): plaintext and ciphertext at different scale degrees");
  if (pc->limbs < a->limbs) throw std::runtime_error("EvalAdd(plaintext): plaintext below the ciphertext level");
  if (pc->limbs > a->limbs) pc = DropToLimbs(pc, a->limbs);
  Ciphertext o = NewCiphertext(a->limbs);
  const uint64_t* pa = a->d;
  const uint64_t* pb = pc->d;
  uint64_t* po = o->d;
  const uint32_t l = a->limbs;
  chk(fides_add_batch(ctx, 1, &pa, &pb, &l, sub ? 1 : 0, &po, nullptr), sub ? "EvalSub(plaintext)" : "EvalAdd(plaintext)");
  copy_meta(*o, *a);
  return o;
}

Ciphertext CryptoContextImpl::EvalAdd(const Ciphertext& a, const Plaintext& p) { return AddSubPt(a, p, false); }
Ciphertext CryptoContextImpl::EvalSub(const Ciphertext& a, const Plaintext& p) { return AddSubPt(a, p, true); }

// ---------------------------------------------------------------- sums of many
Ciphertext CryptoContextImpl::EvalAddMany(const std::vector<Ciphertext>& cts) {
  if (cts.empty()) throw std::runtime_error("EvalAddMany: input ciphertext vector is empty");
  if (cts.size() == 1) return DropToLimbs(cts[0], cts[0]->limbs);
  bool uniform = true;
  for (const auto& c : cts)
    uniform = uniform && !c->ext && c->limbs == cts[0]->limbs && c->NoiseLevel == cts[0]->NoiseLevel &&
              c->NoiseFactor == cts[0]->NoiseFactor;
  if (!uniform) {  // the reference's pairwise tree (api/CryptoContext.cpp:909-926)
    const size_t n = cts.size();
    std::vector<Ciphertext> sums(n - 1);
    size_t k = 0;
    for (size_t i = 0; i < 2 * n - 2; i += 2)
      sums[k++] = EvalAdd(i < n ? cts[i] : sums[i - n], i + 1 < n ? cts[i + 1] : sums[i + 1 - n]);
    return sums.back();
  }
  // one level and scale: sums of up to 16 terms per fused launch
  const uint32_t l = cts[0]->limbs;
  const std::vector<uint64_t> one(l, 1);
  std::vector<Ciphertext> cur(cts.begin(), cts.end());
  while (cur.size() > 1) {
    std::vector<Ciphertext> next;
    for (size_t i = 0; i < cur.size(); i += 16) {
      const size_t m = std::min<size_t>(16, cur.size() - i);
      if (m == 1) {
        next.push_back(cur[i]);
        continue;
      }
      Ciphertext o = NewCiphertext(l);
      std::vector<const uint64_t*> xs, ms;
      std::vector<uint32_t> xl(m, l);
      for (size_t j = 0; j < m; ++j) {
        xs.push_back(cur[i + j]->d);
        ms.push_back(one.data());
      }
      const uint32_t cnt = (uint32_t)m;
      const uint64_t* add = nullptr;
      uint64_t* po = o->d;
      chk(fides_lincomb_batch(ctx, 1, &cnt, xs.data(), xl.data(), ms.data(), &add, &l, &po, nullptr), "EvalAddMany");
      copy_meta(*o, *cur[i]);
      next.push_back(o);
    }
    cur.swap(next);
  }
  int slots = 0;
  for (const auto& c : cts) slots = std::max(slots, c->slots);
  cur[0]->slots = slots;
  return cur[0];
}

void CryptoContextImpl::EvalAddManyInPlace(std::vector<Ciphertext>& cts) { cts.at(0) = EvalAddMany(cts); }

// ---------------------------------------------------------------- extended rotations, accumulation
std::vector<Ciphertext> CryptoContextImpl::EvalFastRotationExt(const Ciphertext& a0,
                                                               const std::vector<int32_t>& indices,
                                                               const std::shared_ptr<void>&, bool addFirst) {
  const Ciphertext a = Regular(a0);
  std::vector<uint64_t*> po;
  std::vector<uint32_t> kv;
  std::vector<const fides_key*> keys;
  std::vector<Ciphertext> out;
  for (int32_t idx : indices) {
    const uint32_t k = AutoIndex(idx);
    Ciphertext o = NewCiphertext(a->limbs, true);
    copy_meta(*o, *a);
    po.push_back(o->d);
    kv.push_back(k);
    keys.push_back(k == 1 ? nullptr : RotationKey(k));
    out.push_back(o);
  }
  chk(fides_rotate_hoisted_ext(ctx, a->d, a->limbs, (uint32_t)indices.size(), keys.data(), kv.data(),
                               addFirst ? 1 : 0, po.data(), nullptr),
      "EvalFastRotationExt");
  return out;
}

Ciphertext CryptoContextImpl::EvalFastRotationExt(const Ciphertext& a, int32_t index,
                                                  const std::shared_ptr<void>& digits, bool addFirst) {
  return EvalFastRotationExt(a, std::vector<int32_t>{index}, digits, addFirst)[0];
}

// Rounds s = start, 4 start, 16 start, ...: rotations by stride*s*j (j = 1..3,
// below stride*size) hoisted and left in Q_l u P, summed with P * a, one
// ModDown (src/CKKS/AccumulateBroadcast.cu:9-42 and :79-106, bStep = 4).
void CryptoContextImpl::AccumulateImpl(Ciphertext& a, int stride, int size, int start) {
  if (stride <= 0 || size <= 0 || start <= 0) throw std::runtime_error("AccumulateSum: stride, slots and start must be > 0");
  Ciphertext x = Regular(a);
  for (int64_t s = start; s < size; s <<= 2) {
    std::vector<int32_t> idx;
    for (int64_t i = stride * s; i < (int64_t)stride * size && i < 4 * stride * s; i += stride * s)
      idx.push_back((int32_t)i);
    if (idx.empty()) continue;
    const auto rots = EvalFastRotationExt(x, idx, nullptr, true);
    Ciphertext acc = Extended(x);
    const uint32_t l = acc->limbs;
    for (const auto& r : rots) {
      const uint64_t* pa = acc->d;
      const uint64_t* pb = r->d;
      uint64_t* po = acc->d;
      chk(fides_ext_add_batch(ctx, 1, &pa, &pb, &l, 0, &po, nullptr), "AccumulateSum");
    }
    x = Regular(acc);
  }
  if ((int64_t)size * stride == x->slots) x->slots = stride * start;
  a = x;
}

Ciphertext CryptoContextImpl::AccumulateSum(const Ciphertext& a, int slots, int stride) {
  Ciphertext x = a;
  AccumulateImpl(x, stride, slots, 1);
  return x;
}
void CryptoContextImpl::AccumulateSumInPlace(Ciphertext& a, int slots, int stride) { AccumulateImpl(a, stride, slots, 1); }
void CryptoContextImpl::AccumulateSumInPlace(Ciphertext& a, int slots, int stride, int start) {
  AccumulateImpl(a, stride, slots, start);
}

// ---------------------------------------------------------------- levels
Ciphertext CryptoContextImpl::SetLevelImpl(const Ciphertext& a0, size_t level) {
  const Ciphertext a = Regular(a0);
  if (level >= L) throw std::runtime_error("SetLevel: level beyond the modulus chain");
  const uint32_t target = L - (uint32_t)level;
  if (target >= a->limbs) return a;  // dropToLevel only drops
  if (Flexible()) {
    if (a->NoiseLevel != 1 && a->NoiseLevel != 2) throw std::runtime_error("SetLevel: scale degree must be 1 or 2");
    Ciphertext x = a;
    const uint32_t lev = target - 1;
    AdjustScaleAndLevel(x, x->NoiseLevel, lev, x->NoiseLevel == 1 ? sfReal.at(lev) : sfBig.at(lev));
    return x;
  }
  return DropToLimbs(a, target);
}

void CryptoContextImpl::SetLevel(Ciphertext& a, size_t level) { a = a->cc->SetLevelImpl(a, level); }

// ---------------------------------------------------------------- Chebyshev series
std::vector<double> CryptoContextImpl::GetChebyshevCoefficients(std::function<double(double)>& func, double a,
                                                                double b, size_t degree) {
  if (degree == 0) throw std::runtime_error("GetChebyshevCoefficients: the degree can not be zero");
  const size_t n = degree + 1;
  const double bma = 0.5 * (b - a), bpa = 0.5 * (b + a), pd = M_PI / (double)n;
  std::vector<double> fx(n), c(n, 0.0);
  for (size_t i = 0; i < n; ++i) fx[i] = func(std::cos(pd * ((double)i + 0.5)) * bma + bpa);
  const double f = 2.0 / (double)n;
  for (size_t i = 0; i < n; ++i) {
    for (size_t j = 0; j < n; ++j) c[i] += fx[j] * std::cos(pd * (double)i * ((double)j + 0.5));
    c[i] *= f;
  }
  return c;
}

Ciphertext CryptoContextImpl::EvalChebyshevSeries(const Ciphertext& a0, std::vector<double>& coeffs, double lo,
                                                  double hi) {
  if (!evalMultKey) throw std::runtime_error("EvalChebyshevSeries: relinearization key not loaded");
  if (coeffs.size() < 2) throw std::runtime_error("EvalChebyshevSeries: needs a degree of at least 1");
  if (!(hi > lo)) throw std::runtime_error("EvalChebyshevSeries: needs lo < hi");
  Ciphertext a = Regular(a0);
  if (a->NoiseLevel == 2) a = Rescale(a);
  if (a->NoiseLevel != 1) throw std::runtime_error("EvalChebyshevSeries: scale degree must be 1 or 2");
  detail::Exec X{*this, ctx, q, N, boot_delta(), evalMultKey};
  if (!bootMem) bootMem = std::make_shared<detail::Pool>();
  X.pool = std::static_pointer_cast<detail::Pool>(bootMem);  // registers from the context's cache
  detail::Reg x;
  x.buf = std::make_shared<detail::DevBuf>(a->d, false);
  x.level = a->limbs;
  x.scale = a->NoiseFactor;
  x.stride = (size_t)2 * a->limbs * N;
  // y = (2x - lo - hi) / (hi - lo): multiply by round(sa * delta) and Rescale,
  // then add round(sb * scale) (fideslib_amd/serve.py chebyshev_program)
  const double sa = 2.0 / (hi - lo), sb = -(hi + lo) / (hi - lo);
  detail::Reg y = x;
  if (!(sa == 1.0 && sb == 0.0)) {
    y = X.mul_real(x, sa);
    if (sb != 0.0) {
      const detail::i128 add = detail::Exec::round_i128(sb * y.scale);
      y = X.lin(y, nullptr, nullptr, false, &add, y.scale);
    }
  }
  std::vector<double> c = coeffs;
  c[0] *= 0.5;  // OpenFHE's series is c_0 / 2 + sum_k c_k T_k
  const detail::Reg r = detail::chebyshev_eval(X, y, c);
  Ciphertext o = NewCiphertext(r.level);
  chk(fides_copy_device(o->d, r.d(0), (size_t)2 * r.level * N * 8, nullptr), "EvalChebyshevSeries");
  o->NoiseLevel = 1;
  o->NoiseFactor = r.scale;
  o->slots = a->slots;
  o->keyid = a->keyid;
  return o;
}

void CryptoContextImpl::EvalChebyshevSeriesInPlace(Ciphertext& a, std::vector<double>& coeffs, double lo, double hi) {
  a = EvalChebyshevSeries(a, coeffs, lo, hi);
}

}  // namespace fideslib
