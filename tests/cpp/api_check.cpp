// Drives the host C++ API (fideslib::CryptoContext) the way an application
// does: Create -> AddEvalMultKey / AddRotationKey -> LoadCiphertext ->
// EvalMult -> Rescale -> EvalRotate / EvalFastRotation / EvalMultMany ->
// StoreCiphertext.  Inputs and outputs are raw little-endian uint64 files in
// the directory given as argv[1] (written and checked by
// tests/test_cpp_api_gpu.py against the CPU oracle).
//
//   params.bin : logN L K dnum l rot q[L] p[K]
//   rlk.bin, rot.bin : [2][dnum][L+K][N] (a then b)
//   c1.bin, c2.bin   : [2][l][N]
// Outputs: mult.bin ([2][l-1][N], Rescale(EvalMult(c1, c2))),
//          rot.bin  ([2][l-1][N], EvalRotate(mult, rot)),
//          fast.bin ([3][2][l-1][N], EvalFastRotation(mult, {rot, 0, rot})),
//          many.bin ([2][2][l-1][N], EvalMultMany({c1, c2}, {c2, c1}, rescale)),
//          meta.txt (NoiseLevel / NoiseFactor / limbs of the results).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <stdexcept>
#include <string>
#include <vector>

#include "fideslib/fideslib.hpp"
#include "fideslib/serialize.hpp"

static std::vector<uint64_t> read_u64(const std::string& path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open " + path);
  const std::streamsize n = f.tellg();
  f.seekg(0);
  std::vector<uint64_t> v((size_t)n / 8);
  f.read(reinterpret_cast<char*>(v.data()), n);
  return v;
}

static void write_u64(const std::string& path, const std::vector<uint64_t>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * 8));
}

static fideslib::RawKeySwitchKey key_of(const std::vector<uint64_t>& flat, uint32_t dnum, uint32_t LK, uint32_t N) {
  fideslib::RawKeySwitchKey k;
  k.r_key.resize(2);
  size_t o = 0;
  for (uint32_t c = 0; c < 2; ++c) {
    k.r_key[c].resize(dnum);
    for (uint32_t d = 0; d < dnum; ++d) {
      k.r_key[c][d].resize(LK);
      for (uint32_t t = 0; t < LK; ++t, o += N) k.r_key[c][d][t].assign(flat.begin() + o, flat.begin() + o + N);
    }
  }
  return k;
}

static fideslib::RawCipherText ct_of(const std::vector<uint64_t>& flat, uint32_t l, uint32_t N, double scale) {
  fideslib::RawCipherText c;
  c.sub_0.resize(l);
  c.sub_1.resize(l);
  for (uint32_t i = 0; i < l; ++i) {
    c.sub_0[i].assign(flat.begin() + (size_t)i * N, flat.begin() + (size_t)(i + 1) * N);
    c.sub_1[i].assign(flat.begin() + (size_t)(l + i) * N, flat.begin() + (size_t)(l + i + 1) * N);
  }
  c.numRes = (int)l;
  c.N = (int)N;
  c.Noise = scale;
  c.NoiseLevel = 1;
  c.slots = (int)N / 2;
  return c;
}

static void append(std::vector<uint64_t>& out, const fideslib::RawCipherText& r) {
  for (const auto& limb : r.sub_0) out.insert(out.end(), limb.begin(), limb.end());
  for (const auto& limb : r.sub_1) out.insert(out.end(), limb.begin(), limb.end());
}

static void store(const std::string& path, fideslib::CryptoContext& cc, const fideslib::Ciphertext& c,
                  std::ofstream& meta) {
  std::vector<uint64_t> out;
  append(out, cc->StoreCiphertext(c));
  write_u64(path, out);
  meta << c->GetLimbs() << " " << c->NoiseLevel << " " << c->NoiseFactor << "\n";
}

// FLEXIBLEAUTO: operands at different levels and scale degrees are aligned by
// the API (scalar-multiply-and-rescale), as FIDESlib does on OpenFHE's tables.
// Inputs c1 (L limbs), c2 (L-1 limbs), c3 (L limbs) at the default scaling
// factors of their levels.  Outputs m = c1*c2, a = m + c3, s = 0.5 a,
// t = s + 0.25, u = t - c2, v = -u, w = 3 c3, each with its metadata line.
static int flex(const std::string& dir) {
  const auto P = read_u64(dir + "/params.bin");
  const uint32_t logN = (uint32_t)P[0], L = (uint32_t)P[1], K = (uint32_t)P[2], dnum = (uint32_t)P[3];
  const uint32_t N = 1u << logN;
  fideslib::RawParams rp;
  rp.logN = (int)logN;
  rp.moduli.assign(P.begin() + 6, P.begin() + 6 + L);
  rp.SPECIALmoduli.assign(P.begin() + 6 + L, P.begin() + 6 + L + K);
  rp.dnum = (int)dnum;
  rp.scalingTechnique = fideslib::FLEXIBLEAUTO;
  auto cc = fideslib::CryptoContextImpl::Create(rp);
  cc->AddEvalMultKey(key_of(read_u64(dir + "/rlk.bin"), dnum, L + K, N));
  auto c1 = cc->LoadCiphertext(ct_of(read_u64(dir + "/c1.bin"), L, N, cc->ScalingFactorReal(L - 1)));
  auto c2 = cc->LoadCiphertext(ct_of(read_u64(dir + "/c2.bin"), L - 1, N, cc->ScalingFactorReal(L - 2)));
  auto c3 = cc->LoadCiphertext(ct_of(read_u64(dir + "/c3.bin"), L, N, cc->ScalingFactorReal(L - 1)));
  std::ofstream meta(dir + "/meta.txt");
  meta << std::setprecision(17);
  for (uint32_t l = 0; l < L; ++l) meta << cc->ScalingFactorReal(l) << (l + 1 < L ? " " : "\n");
  auto m = cc->EvalMult(c1, c2);
  auto a = cc->EvalAdd(m, c3);
  auto s = cc->EvalMult(a, 0.5);
  auto t = cc->EvalAdd(s, 0.25);
  auto u = cc->EvalSub(t, c2);
  auto v = cc->EvalNegate(u);
  auto w = cc->EvalMult(c3, 3.0);
  store(dir + "/m.bin", cc, m, meta);
  store(dir + "/a.bin", cc, a, meta);
  store(dir + "/s.bin", cc, s, meta);
  store(dir + "/t.bin", cc, t, meta);
  store(dir + "/u.bin", cc, u, meta);
  store(dir + "/v.bin", cc, v, meta);
  store(dir + "/w.bin", cc, w, meta);
  std::printf("api_check flex: ok\n");
  return 0;
}

static bool same(const fideslib::RawCipherText& a, const fideslib::RawCipherText& b) {
  return a.sub_0 == b.sub_0 && a.sub_1 == b.sub_1 && a.Noise == b.Noise && a.NoiseLevel == b.NoiseLevel &&
         a.slots == b.slots && a.keyid == b.keyid;
}

static fideslib::RawParams params_of(const std::vector<uint64_t>& P) {
  const uint32_t L = (uint32_t)P[1], K = (uint32_t)P[2];
  fideslib::RawParams rp;
  rp.logN = (int)P[0];
  rp.moduli.assign(P.begin() + 6, P.begin() + 6 + L);
  rp.SPECIALmoduli.assign(P.begin() + 6 + L, P.begin() + 6 + L + K);
  rp.dnum = (int)P[3];
  rp.scalingTechnique = fideslib::FIXEDMANUAL;
  rp.ModReduceFactor = {1.5, 2.5};
  rp.ScalingFactorReal = {3.25};
  return rp;
}

// Host structs through the binary format (no device needed).
static int serial_host(const std::string& dir) {
  using namespace fideslib;
  const auto P = read_u64(dir + "/params.bin");
  const uint32_t logN = (uint32_t)P[0], L = (uint32_t)P[1], K = (uint32_t)P[2], dnum = (uint32_t)P[3];
  const uint32_t l = (uint32_t)P[4], N = 1u << logN;
  RawParams rp = params_of(P), rp2;
  RawCipherText c = ct_of(read_u64(dir + "/c1.bin"), l, N, 1234.5), c2;
  c.keyid = "key-1";
  RawKeySwitchKey k = key_of(read_u64(dir + "/rlk.bin"), dnum, L + K, N), k2;
  k.keyid = "relin";
  bool ok = Serial::SerializeToFile(dir + "/p.fds", rp) && Serial::DeserializeFromFile(dir + "/p.fds", rp2);
  ok = ok && rp2.logN == rp.logN && rp2.moduli == rp.moduli && rp2.SPECIALmoduli == rp.SPECIALmoduli &&
       rp2.dnum == rp.dnum && rp2.ModReduceFactor == rp.ModReduceFactor && rp2.ScalingFactorReal == rp.ScalingFactorReal;
  ok = ok && Serial::SerializeToFile(dir + "/c.fds", c) && Serial::DeserializeFromFile(dir + "/c.fds", c2) && same(c, c2);
  ok = ok && Serial::SerializeToFile(dir + "/k.fds", k) && Serial::DeserializeFromFile(dir + "/k.fds", k2) &&
       k2.r_key == k.r_key && k2.keyid == k.keyid;
  // a file of another type, a truncated file and JSON are refused
  RawKeySwitchKey wrong;
  ok = ok && !Serial::DeserializeFromFile(dir + "/c.fds", wrong);
  {
    std::ifstream f(dir + "/c.fds", std::ios::binary);
    std::string all((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::ofstream g(dir + "/t.fds", std::ios::binary);
    g.write(all.data(), (std::streamsize)(all.size() / 2));
  }
  RawCipherText c3;
  ok = ok && !Serial::DeserializeFromFile(dir + "/t.fds", c3);
  ok = ok && !Serial::SerializeToFile(dir + "/j.fds", c, JSON);
  std::printf("api_check serial-host: %s\n", ok ? "ok" : "FAILED");
  return ok ? 0 : 1;
}

// A whole context (parameters + keys) and ciphertexts through files: the
// deserialized context computes the same bits.
static int serial(const std::string& dir) {
  using namespace fideslib;
  const auto P = read_u64(dir + "/params.bin");
  const uint32_t logN = (uint32_t)P[0], L = (uint32_t)P[1], K = (uint32_t)P[2], dnum = (uint32_t)P[3];
  const uint32_t l = (uint32_t)P[4];
  const int32_t rot = (int32_t)P[5];
  const uint32_t N = 1u << logN;
  RawParams rp = params_of(P);
  rp.ModReduceFactor.clear();
  rp.ScalingFactorReal.clear();
  auto cc = CryptoContextImpl::Create(rp);
  cc->AddEvalMultKey(key_of(read_u64(dir + "/rlk.bin"), dnum, L + K, N));
  cc->AddRotationKey(rot, key_of(read_u64(dir + "/rot.bin"), dnum, L + K, N));
  const double scale = (double)rp.moduli[l - 1];
  auto a = cc->LoadCiphertext(ct_of(read_u64(dir + "/c1.bin"), l, N, scale));
  auto b = cc->LoadCiphertext(ct_of(read_u64(dir + "/c2.bin"), l, N, scale));
  auto r1 = cc->StoreCiphertext(cc->EvalRotate(cc->Rescale(cc->EvalMult(a, b)), rot));
  bool ok = Serial::SerializeToFile(dir + "/ctx.fds", cc) &&
            Serial::SerializeToFile(dir + "/a.fds", cc->StoreCiphertext(a)) &&
            Serial::SerializeToFile(dir + "/b.fds", cc->StoreCiphertext(b));
  CryptoContext cc2;
  RawCipherText ra, rb;
  ok = ok && Serial::DeserializeFromFile(dir + "/ctx.fds", cc2) && Serial::DeserializeFromFile(dir + "/a.fds", ra) &&
       Serial::DeserializeFromFile(dir + "/b.fds", rb);
  if (!ok) {
    std::printf("api_check serial: FAILED (io)\n");
    return 1;
  }
  auto a2 = cc2->LoadCiphertext(ra), b2 = cc2->LoadCiphertext(rb);
  auto r2 = cc2->StoreCiphertext(cc2->EvalRotate(cc2->Rescale(cc2->EvalMult(a2, b2)), rot));
  ok = same(r1, r2) && cc2->HasRotationKey(rot) && cc2->GetEvalMultKey().r_key == cc->GetEvalMultKey().r_key;
  std::vector<uint64_t> out;
  append(out, r2);
  write_u64(dir + "/serial_result.bin", out);
  std::printf("api_check serial: %s\n", ok ? "ok" : "FAILED");
  return ok ? 0 : 1;
}

// Plaintext products: pt_eval.bin ([l][N] evaluation form) and pt_coef.bin
// (the same plaintext in coefficient form, one limb more) multiply c1.bin
// ([2][l][N]); outputs ptm_eval.bin, ptm_coef.bin, ptm_auto.bin (FIXEDAUTO,
// the product of a degree-2 ciphertext: rescaled first) and meta.txt.
static int plain(const std::string& dir) {
  const auto P = read_u64(dir + "/params.bin");
  const uint32_t logN = (uint32_t)P[0], L = (uint32_t)P[1], K = (uint32_t)P[2], dnum = (uint32_t)P[3];
  const uint32_t l = (uint32_t)P[4], N = 1u << logN;
  std::ofstream meta(dir + "/meta.txt");
  meta << std::setprecision(17);
  for (int tech = 0; tech < 2; ++tech) {
    fideslib::RawParams rp;
    rp.logN = (int)logN;
    rp.moduli.assign(P.begin() + 6, P.begin() + 6 + L);
    rp.SPECIALmoduli.assign(P.begin() + 6 + L, P.begin() + 6 + L + K);
    rp.dnum = (int)dnum;
    rp.scalingTechnique = tech == 0 ? fideslib::FIXEDMANUAL : fideslib::FIXEDAUTO;
    auto cc = fideslib::CryptoContextImpl::Create(rp);
    cc->AddEvalMultKey(key_of(read_u64(dir + "/rlk.bin"), dnum, L + K, N));
    const double sc = (double)(1ull << 40);
    auto c1 = cc->LoadCiphertext(ct_of(read_u64(dir + "/c1.bin"), l, N, sc));
    auto pe = read_u64(dir + "/pt_eval.bin"), pc = read_u64(dir + "/pt_coef.bin");
    fideslib::RawPlainText re, rc;
    for (uint32_t i = 0; i < l; ++i) re.sub_0.emplace_back(pe.begin() + (size_t)i * N, pe.begin() + (size_t)(i + 1) * N);
    for (uint32_t i = 0; i < l + 1; ++i)
      rc.sub_0.emplace_back(pc.begin() + (size_t)i * N, pc.begin() + (size_t)(i + 1) * N);
    re.format = fideslib::EVALUATION;
    rc.format = fideslib::COEFFICIENT;
    re.Noise = rc.Noise = sc;
    rc.moduli.assign(rp.moduli.begin(), rp.moduli.begin() + l + 1);
    auto pte = cc->LoadPlaintext(re);
    auto ptc = cc->LoadPlaintext(rc);
    if (tech == 0) {
      auto m1 = cc->EvalMult(c1, pte);
      auto m2 = cc->EvalMult(ptc, c1);
      store(dir + "/ptm_eval.bin", cc, m1, meta);
      store(dir + "/ptm_coef.bin", cc, m2, meta);
      bool threw = false;
      try {
        auto c2 = cc->LoadCiphertext(ct_of(read_u64(dir + "/c2.bin"), l, N, sc));
        fideslib::RawPlainText r1 = re;
        r1.sub_0.pop_back();
        cc->EvalMult(c2, cc->LoadPlaintext(r1));  // plaintext one limb short
      } catch (const std::runtime_error&) {
        threw = true;
      }
      meta << (threw ? "short-plaintext-throws" : "short-plaintext-accepted") << "\n";
    } else {
      auto sq = cc->EvalMult(c1, c1);            // degree 2, not rescaled yet (FIXEDAUTO)
      auto m3 = cc->EvalMult(sq, ptc);           // rescales sq first
      store(dir + "/ptm_auto.bin", cc, m3, meta);
    }
  }
  return 0;
}

// Bootstrapping: boot.bin = [scale bits, nkeys, auto indices...], rotkeys.bin
// = nkeys keys [2][dnum][L+K][N], c1.bin = [2][l][N].  Outputs boot_rots.txt
// (the rotation indices EvalBootstrapSetup asks for), boot_out.bin and
// meta.txt (limbs, NoiseLevel, NoiseFactor).
static int boot(const std::string& dir) {
  const auto P = read_u64(dir + "/params.bin");
  const uint32_t logN = (uint32_t)P[0], L = (uint32_t)P[1], K = (uint32_t)P[2], dnum = (uint32_t)P[3];
  const uint32_t l = (uint32_t)P[4], N = 1u << logN;
  fideslib::RawParams rp;
  rp.logN = (int)logN;
  rp.moduli.assign(P.begin() + 6, P.begin() + 6 + L);
  rp.SPECIALmoduli.assign(P.begin() + 6 + L, P.begin() + 6 + L + K);
  rp.dnum = (int)dnum;
  auto cc = fideslib::CryptoContextImpl::Create(rp);
  cc->AddEvalMultKey(key_of(read_u64(dir + "/rlk.bin"), dnum, L + K, N));
  const auto B = read_u64(dir + "/boot.bin");
  const auto keys = read_u64(dir + "/rotkeys.bin");
  const size_t ksz = (size_t)2 * dnum * (L + K) * N;
  for (uint64_t i = 0; i < B[1]; ++i) {
    std::vector<uint64_t> one(keys.begin() + i * ksz, keys.begin() + (i + 1) * ksz);
    cc->AddAutomorphismKey((uint32_t)B[2 + i], key_of(one, dnum, L + K, N));
  }
  const auto rots = cc->EvalBootstrapSetup();
  {
    std::ofstream f(dir + "/boot_rots.txt");
    for (int32_t r : rots) f << r << "\n";
  }
  auto c1 = cc->LoadCiphertext(ct_of(read_u64(dir + "/c1.bin"), l, N, std::ldexp(1.0, (int)B[0])));
  auto out = cc->EvalBootstrap(c1);
  std::ofstream meta(dir + "/meta.txt");
  meta << std::setprecision(17);
  store(dir + "/boot_out.bin", cc, out, meta);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: api_check DIR [flex|plain|serial|serial-host]\n");
    return 2;
  }
  const std::string dir = argv[1];
  if (argc > 2 && std::string(argv[2]) == "boot") {
    try {
      return boot(dir);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "api_check: %s\n", e.what());
      return 1;
    }
  }
  if (argc > 2 && std::string(argv[2]) == "plain") {
    try {
      return plain(dir);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "api_check: %s\n", e.what());
      return 1;
    }
  }
  if (argc > 2 && (std::string(argv[2]) == "serial-host" || std::string(argv[2]) == "serial")) {
    try {
      return std::string(argv[2]) == "serial" ? serial(dir) : serial_host(dir);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "api_check: %s\n", e.what());
      return 1;
    }
  }
  if (argc > 2 && std::string(argv[2]) == "flex") {
    try {
      return flex(dir);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "api_check: %s\n", e.what());
      return 1;
    }
  }
  try {
    const auto P = read_u64(dir + "/params.bin");
    const uint32_t logN = (uint32_t)P[0], L = (uint32_t)P[1], K = (uint32_t)P[2], dnum = (uint32_t)P[3];
    const uint32_t l = (uint32_t)P[4];
    const int32_t rot = (int32_t)P[5];
    const uint32_t N = 1u << logN;
    fideslib::RawParams rp;
    rp.logN = (int)logN;
    rp.moduli.assign(P.begin() + 6, P.begin() + 6 + L);
    rp.SPECIALmoduli.assign(P.begin() + 6 + L, P.begin() + 6 + L + K);
    rp.dnum = (int)dnum;
    rp.scalingTechnique = fideslib::FIXEDMANUAL;
    auto cc = fideslib::CryptoContextImpl::Create(rp);
    cc->AddEvalMultKey(key_of(read_u64(dir + "/rlk.bin"), dnum, L + K, N));
    cc->AddRotationKey(rot, key_of(read_u64(dir + "/rot.bin"), dnum, L + K, N));
    // the operations below run under autotuned launch shapes (same bits)
    const auto tuned = cc->Autotune({l, l, l - 1});
    const auto now = cc->GetTuning();
    for (int c = 0; c < 3; ++c)
      if (now.convChunk[c] != tuned.convChunk[c] || now.groupByPrime[c] != tuned.groupByPrime[c] ||
          now.fuseRowDot[c] != tuned.fuseRowDot[c])
        throw std::runtime_error("autotuned settings not kept");
    cc->SetTuning(now);
    const double scale = (double)rp.moduli[l - 1];
    auto a = cc->LoadCiphertext(ct_of(read_u64(dir + "/c1.bin"), l, N, scale));
    auto b = cc->LoadCiphertext(ct_of(read_u64(dir + "/c2.bin"), l, N, scale));
    auto m = cc->Rescale(cc->EvalMult(a, b));
    auto r = cc->EvalRotate(m, rot);
    auto f = cc->EvalFastRotation(m, std::vector<int32_t>{rot, 0, rot}, 2 * N, cc->EvalFastRotationPrecompute(m));
    auto many = cc->EvalMultMany({a, b}, {b, a}, true);
    std::vector<uint64_t> out;
    append(out, cc->StoreCiphertext(m));
    write_u64(dir + "/mult.bin", out);
    out.clear();
    append(out, cc->StoreCiphertext(r));
    write_u64(dir + "/rot.bin", out);
    out.clear();
    for (auto& x : f) append(out, cc->StoreCiphertext(x));
    write_u64(dir + "/fast.bin", out);
    out.clear();
    for (auto& x : many) append(out, cc->StoreCiphertext(x));
    write_u64(dir + "/many.bin", out);
    std::ofstream meta(dir + "/meta.txt");
    meta << std::setprecision(17);
    meta << m->GetLimbs() << " " << m->NoiseLevel << " " << m->NoiseFactor << "\n";
    meta << r->GetLimbs() << " " << r->NoiseLevel << " " << r->NoiseFactor << "\n";
    // a missing rotation key is an error, not a crash
    bool threw = false;
    try {
      cc->EvalRotate(m, rot + 1);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    meta << (threw ? "missing-key-throws" : "missing-key-accepted") << "\n";
    std::printf("api_check: ok\n");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "api_check: %s\n", e.what());
    return 1;
  }
  return 0;
}
