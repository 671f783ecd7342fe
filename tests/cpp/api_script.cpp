// Runs a script of fideslib::CryptoContext calls (the host C++ API, as an
// application issues them) over named registers; written and checked by the
// tests through tests/cpp_script.py against the CPU oracle.
//
// Usage: api_script DIR  (reads DIR/script.txt; inputs and outputs are raw
// little-endian uint64 files in DIR).  One command per line:
//   context <file> <technique 0..3>    params: logN L K dnum 0 0 q[L] p[K]
//   rlk <file>                          relinearization key [2][dnum][L+K][N]
//   autokey <autoIndex> <file>          rotation/conjugation key
//   ct <name> <file> <limbs> <deg> <noiseFactor> <slots>
//   pt <name> <file> <limbs> <format 0 coef|1 eval> <deg> <noiseFactor>
//   add|sub|mul|mulmut <out> <a> <b>    ct op ct
//   addpt|subpt|mulpt <out> <a> <p>     ct op pt;  ptsub <out> <p> <a> = p - a
//   adds|subs|muls <out> <a> <c>        ct op double;  rsubs <out> <c> <a> = c - a
//   sq|sqmut|neg|rescale|conj <out> <a>
//   rot <out> <a> <index>
//   fastrotext <out> <a> <index> <addFirst>
//   addmany <out> <a> <b> ...
//   accum <out> <a> <slots> <stride> [start]
//   cheb <out> <a> <lo> <hi> <c0> <c1> ...
//   setlevel <out> <a> <level>
//   chebcoef <name> <exp|sigmoid|sin> <a> <b> <degree>  -> name.txt (GetChebyshevCoefficients)
//   encapkeys <fileDown> <fileUp>       sparse-secret encapsulation keys
//   bootsetup <ctsBudget> <stcBudget> <degree> <doubleAngle> <K> <correction> <encaps>
//   bootsetup_pre <file>                EvalBootstrapSetup from a precomputation file
//                                       (both write "rots ..." to meta.txt)
//   boot <out> <a>                      EvalBootstrap
//   bootmany <prefix> <a> <b> ...       EvalBootstrapMany -> prefix0, prefix1, ...
//   releasemem                          ReleaseCachedMemory
//   synthkeys <seed>                    random keys (one shared value) for relinearization,
//                                       every rotation of the last bootstrap setup,
//                                       conjugation and encapsulation (timing only)
//   synthct <name> <limbs> <seed> <noiseFactor> <slots>   random residues (timing only)
//   boottime <reps> <n> <a>             ms for n EvalBootstrap calls, then for one
//                                       EvalBootstrapMany of n (line "boottime n one many")
//   store <name>                        -> name.bin [2][limbs][N], line in meta.txt:
//                                          name limbs deg noiseFactor slots extended
//   expect_throw <command...>           the command must throw (line "throws" in meta.txt)
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fideslib/fideslib.hpp"

using namespace fideslib;

static std::string g_dir;

static std::vector<uint64_t> read_u64(const std::string& name) {
  std::ifstream f(g_dir + "/" + name, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open " + name);
  const std::streamsize n = f.tellg();
  f.seekg(0);
  std::vector<uint64_t> v((size_t)n / 8);
  f.read(reinterpret_cast<char*>(v.data()), n);
  return v;
}

static void write_u64(const std::string& name, const std::vector<uint64_t>& v) {
  std::ofstream f(g_dir + "/" + name, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * 8));
}

struct State {
  CryptoContext cc;
  uint32_t N = 0, L = 0, K = 0, dnum = 0;
  std::map<std::string, Ciphertext> ct;
  std::map<std::string, Plaintext> pt;
  std::ofstream meta;

  std::vector<int32_t> last_rots;
  std::vector<uint64_t> moduli;  // Q then P

  static uint64_t rnd(uint64_t& st) {  // xorshift64*
    st ^= st >> 12;
    st ^= st << 25;
    st ^= st >> 27;
    return st * 0x2545f4914f6cdd1dull;
  }
  RawKeySwitchKey synth_key(uint64_t seed) {
    uint64_t st = seed * 0x9e3779b97f4a7c15ull + 7;
    RawKeySwitchKey k;
    k.r_key.resize(2);
    for (uint32_t c = 0; c < 2; ++c) {
      k.r_key[c].resize(dnum);
      for (uint32_t d = 0; d < dnum; ++d) {
        k.r_key[c][d].resize(L + K);
        for (uint32_t t = 0; t < L + K; ++t) {
          k.r_key[c][d][t].resize(N);
          for (uint32_t j = 0; j < N; ++j) k.r_key[c][d][t][j] = rnd(st) % moduli[t];
        }
      }
    }
    return k;
  }
  RawKeySwitchKey key(const std::string& file) {
    const auto flat = read_u64(file);
    RawKeySwitchKey k;
    k.r_key.resize(2);
    size_t o = 0;
    for (uint32_t c = 0; c < 2; ++c) {
      k.r_key[c].resize(dnum);
      for (uint32_t d = 0; d < dnum; ++d) {
        k.r_key[c][d].resize(L + K);
        for (uint32_t t = 0; t < L + K; ++t, o += N) k.r_key[c][d][t].assign(flat.begin() + o, flat.begin() + o + N);
      }
    }
    return k;
  }
  Ciphertext& C(const std::string& n) {
    auto it = ct.find(n);
    if (it == ct.end()) throw std::runtime_error("no ciphertext " + n);
    return it->second;
  }
  Plaintext& P(const std::string& n) {
    auto it = pt.find(n);
    if (it == pt.end()) throw std::runtime_error("no plaintext " + n);
    return it->second;
  }

  void run(const std::vector<std::string>& w) {
    const std::string& op = w.at(0);
    auto D = [&](size_t i) { return std::strtod(w.at(i).c_str(), nullptr); };
    auto I = [&](size_t i) { return std::strtoll(w.at(i).c_str(), nullptr, 10); };
    if (op == "context") {
      const auto p = read_u64(w.at(1));
      RawParams rp;
      rp.logN = (int)p[0];
      L = (uint32_t)p[1];
      K = (uint32_t)p[2];
      dnum = (uint32_t)p[3];
      N = 1u << p[0];
      rp.moduli.assign(p.begin() + 6, p.begin() + 6 + L);
      rp.SPECIALmoduli.assign(p.begin() + 6 + L, p.begin() + 6 + L + K);
      moduli.assign(p.begin() + 6, p.begin() + 6 + L + K);
      rp.dnum = (int)dnum;
      rp.scalingTechnique = (ScalingTechnique)I(2);
      cc = CryptoContextImpl::Create(rp);
      meta << "sf";
      for (uint32_t l = 0; l < L; ++l) meta << " " << cc->ScalingFactorReal(l);
      meta << "\n";
    } else if (op == "rlk") {
      cc->AddEvalMultKey(key(w.at(1)));
    } else if (op == "autokey") {
      cc->AddAutomorphismKey((uint32_t)I(1), key(w.at(2)));
    } else if (op == "ct") {
      const auto flat = read_u64(w.at(2));
      const uint32_t l = (uint32_t)I(3);
      RawCipherText r;
      r.sub_0.resize(l);
      r.sub_1.resize(l);
      for (uint32_t i = 0; i < l; ++i) {
        r.sub_0[i].assign(flat.begin() + (size_t)i * N, flat.begin() + (size_t)(i + 1) * N);
        r.sub_1[i].assign(flat.begin() + (size_t)(l + i) * N, flat.begin() + (size_t)(l + i + 1) * N);
      }
      r.numRes = (int)l;
      r.N = (int)N;
      r.NoiseLevel = (int)I(4);
      r.Noise = D(5);
      r.slots = (int)I(6);
      ct[w.at(1)] = cc->LoadCiphertext(r);
    } else if (op == "pt") {
      const auto flat = read_u64(w.at(2));
      const uint32_t l = (uint32_t)I(3);
      RawPlainText r;
      for (uint32_t i = 0; i < l; ++i) r.sub_0.emplace_back(flat.begin() + (size_t)i * N, flat.begin() + (size_t)(i + 1) * N);
      r.numRes = (int)l;
      r.N = (int)N;
      r.format = I(4) ? EVALUATION : COEFFICIENT;
      r.NoiseLevel = (int)I(5);
      r.Noise = D(6);
      pt[w.at(1)] = cc->LoadPlaintext(r);
    } else if (op == "add") {
      ct[w.at(1)] = cc->EvalAdd(C(w.at(2)), C(w.at(3)));
    } else if (op == "sub") {
      ct[w.at(1)] = cc->EvalSub(C(w.at(2)), C(w.at(3)));
    } else if (op == "mul") {
      ct[w.at(1)] = cc->EvalMult(C(w.at(2)), C(w.at(3)));
    } else if (op == "mulmut") {
      Ciphertext a = C(w.at(2)), b = C(w.at(3));
      cc->EvalMultMutableInPlace(a, b);
      ct[w.at(1)] = a;
    } else if (op == "addpt") {
      ct[w.at(1)] = cc->EvalAdd(C(w.at(2)), P(w.at(3)));
    } else if (op == "subpt") {
      ct[w.at(1)] = cc->EvalSub(C(w.at(2)), P(w.at(3)));
    } else if (op == "ptsub") {
      ct[w.at(1)] = cc->EvalSub(P(w.at(2)), C(w.at(3)));
    } else if (op == "mulpt") {
      ct[w.at(1)] = cc->EvalMult(C(w.at(2)), P(w.at(3)));
    } else if (op == "adds") {
      ct[w.at(1)] = cc->EvalAdd(C(w.at(2)), D(3));
    } else if (op == "subs") {
      ct[w.at(1)] = cc->EvalSub(C(w.at(2)), D(3));
    } else if (op == "rsubs") {
      ct[w.at(1)] = cc->EvalSub(D(2), C(w.at(3)));
    } else if (op == "muls") {
      ct[w.at(1)] = cc->EvalMult(C(w.at(2)), D(3));
    } else if (op == "sq") {
      ct[w.at(1)] = cc->EvalSquare(C(w.at(2)));
    } else if (op == "sqmut") {
      Ciphertext a = C(w.at(2));
      ct[w.at(1)] = cc->EvalSquareMutable(a);
    } else if (op == "neg") {
      ct[w.at(1)] = cc->EvalNegate(C(w.at(2)));
    } else if (op == "rescale") {
      ct[w.at(1)] = cc->Rescale(C(w.at(2)));
    } else if (op == "conj") {
      ct[w.at(1)] = cc->EvalConjugate(C(w.at(2)));
    } else if (op == "rot") {
      ct[w.at(1)] = cc->EvalRotate(C(w.at(2)), (int32_t)I(3));
    } else if (op == "fastrotext") {
      ct[w.at(1)] = cc->EvalFastRotationExt(C(w.at(2)), (int32_t)I(3), cc->EvalFastRotationPrecompute(C(w.at(2))),
                                            I(4) != 0);
    } else if (op == "addmany") {
      std::vector<Ciphertext> v;
      for (size_t i = 2; i < w.size(); ++i) v.push_back(C(w[i]));
      ct[w.at(1)] = cc->EvalAddMany(v);
    } else if (op == "accum") {
      Ciphertext a = C(w.at(2));
      if (w.size() > 5)
        cc->AccumulateSumInPlace(a, (int)I(3), (int)I(4), (int)I(5));
      else
        a = cc->AccumulateSum(a, (int)I(3), (int)I(4));
      ct[w.at(1)] = a;
    } else if (op == "cheb") {
      std::vector<double> c;
      for (size_t i = 5; i < w.size(); ++i) c.push_back(D(i));
      ct[w.at(1)] = cc->EvalChebyshevSeries(C(w.at(2)), c, D(3), D(4));
    } else if (op == "chebcoef") {
      const std::string fn = w.at(2);
      std::function<double(double)> f;
      if (fn == "exp") f = [](double x) { return std::exp(x); };
      else if (fn == "sigmoid") f = [](double x) { return 1.0 / (1.0 + std::exp(-x)); };
      else f = [](double x) { return std::sin(x); };
      const auto c = CryptoContextImpl::GetChebyshevCoefficients(f, D(3), D(4), (size_t)I(5));
      std::ofstream o(g_dir + "/" + w.at(1) + ".txt");
      o << std::setprecision(17);
      for (double v : c) o << v << "\n";
    } else if (op == "synthkeys") {  // synthkeys seed: random keys for rlk, every rotation of the last setup, conjugation, encapsulation
      const RawKeySwitchKey k = synth_key((uint64_t)I(1));
      cc->AddEvalMultKey(k);
      for (int32_t r : last_rots) {
        const uint64_t half = N / 2;
        const uint64_t e = (uint64_t)((((int64_t)r % (int64_t)half) + (int64_t)half) % (int64_t)half);
        uint64_t a = 1, b = 5;
        for (uint64_t x = e; x; x >>= 1, b = b * b % (2ull * N))
          if (x & 1) a = a * b % (2ull * N);
        if (a != 1) cc->AddAutomorphismKey((uint32_t)a, k);
      }
      cc->AddAutomorphismKey(2 * N - 1, k);
      cc->AddEncapsulationKeys(k, k);
    } else if (op == "synthct") {  // synthct name limbs seed noiseFactor slots: random residues
      const uint32_t l = (uint32_t)I(2);
      uint64_t st = (uint64_t)I(3) * 0x9e3779b97f4a7c15ull + 1;
      RawCipherText r;
      r.sub_0.resize(l);
      r.sub_1.resize(l);
      for (auto* sub : {&r.sub_0, &r.sub_1})
        for (uint32_t i = 0; i < l; ++i) {
          (*sub)[i].resize(N);
          for (uint32_t j = 0; j < N; ++j) (*sub)[i][j] = rnd(st) % moduli[i];
        }
      r.numRes = (int)l;
      r.N = (int)N;
      r.NoiseLevel = 1;
      r.Noise = D(4);
      r.slots = (int)I(5);
      ct[w.at(1)] = cc->LoadCiphertext(r);
    } else if (op == "releasemem") {
      cc->ReleaseCachedMemory();
    } else if (op == "encapkeys") {
      cc->AddEncapsulationKeys(key(w.at(1)), key(w.at(2)));
    } else if (op == "bootsetup") {
      BootstrapConfig c;
      c.budgetCtS = (uint32_t)I(1);
      c.budgetStC = (uint32_t)I(2);
      c.degree = (uint32_t)I(3);
      c.doubleAngle = (uint32_t)I(4);
      c.K = D(5);
      c.correction = (uint32_t)I(6);
      c.encapsulation = I(7) != 0;
      if (w.size() > 8) c.slots = (uint32_t)I(8);
      const auto rots = cc->EvalBootstrapSetup(c);
      last_rots = rots;
      meta << "rots";
      for (int32_t r : rots) meta << " " << r;
      meta << "\n";
    } else if (op == "bootsetup_pre") {
      // [ncts nstc ncheb correction doubleAngle encapsulation slots outScale] cheb[ncheb]
      // then per group: stride g nbabies babies[] nterms, per term: giant baby level scale coef[level N]
      const auto b = read_u64(w.at(1));
      size_t o = 0;
      auto u = [&]() { return b.at(o++); };
      auto d = [&]() {
        double x;
        const uint64_t v = b.at(o++);
        std::memcpy(&x, &v, 8);
        return x;
      };
      BootstrapPrecomputation pre;
      const uint64_t ncts = u(), nstc = u(), ncheb = u();
      pre.cfg.correction = (uint32_t)u();
      pre.cfg.doubleAngle = (uint32_t)u();
      pre.cfg.encapsulation = u() != 0;
      pre.cfg.slots = (uint32_t)u();
      pre.outScale = d();
      for (uint64_t i = 0; i < ncheb; ++i) pre.chebyshev.push_back(d());
      for (uint64_t gi = 0; gi < ncts + nstc; ++gi) {
        BootstrapPrecomputation::Group g;
        g.stride = (int64_t)u();
        g.g = (int64_t)u();
        const uint64_t nb = u();
        for (uint64_t i = 0; i < nb; ++i) g.babies.push_back((int64_t)u());
        const uint64_t nt = u();
        for (uint64_t i = 0; i < nt; ++i) {
          BootstrapPrecomputation::Term t;
          t.giant = (int64_t)u();
          t.baby = (int64_t)u();
          t.level = (uint32_t)u();
          t.scale = d();
          t.coef.assign(b.begin() + o, b.begin() + o + (size_t)t.level * N);
          o += (size_t)t.level * N;
          g.terms.push_back(std::move(t));
        }
        (gi < ncts ? pre.coeffToSlot : pre.slotToCoeff).push_back(std::move(g));
      }
      const auto rots = cc->EvalBootstrapSetup(pre);
      last_rots = rots;
      meta << "rots";
      for (int32_t r : rots) meta << " " << r;
      meta << "\n";
    } else if (op == "boot") {  // boot out in [iterations precision]
      ct[w.at(1)] = cc->EvalBootstrap(C(w.at(2)), w.size() > 3 ? (uint32_t)I(3) : 1, w.size() > 4 ? (uint32_t)I(4) : 0);
    } else if (op == "bootmany") {  // bootmany outprefix a b ...  -> outprefix0, outprefix1, ...
      std::vector<Ciphertext> in;
      for (size_t i = 2; i < w.size(); ++i) in.push_back(C(w[i]));
      const auto out = cc->EvalBootstrapMany(in);
      for (size_t i = 0; i < out.size(); ++i) ct[w.at(1) + std::to_string(i)] = out[i];
    } else if (op == "boottime") {  // boottime reps n a: n bootstraps one by one, then in one EvalBootstrapMany
      const int reps = (int)I(1), n = (int)I(2);
      const Ciphertext a = C(w.at(3));
      const std::vector<Ciphertext> in(n, a);
      auto ms = [&](const std::function<void()>& f) {
        f();  // warm-up
        cc->Synchronize();
        const auto t0 = std::chrono::steady_clock::now();
        for (int r = 0; r < reps; ++r) f();
        cc->Synchronize();
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / reps;
      };
      const double one = ms([&] {
        for (int i = 0; i < n; ++i) cc->EvalBootstrap(a);
      });
      const double many = ms([&] { cc->EvalBootstrapMany(in); });
      meta << "boottime " << n << " " << one << " " << many << "\n";
    } else if (op == "setlevel") {
      Ciphertext a = C(w.at(2));
      CryptoContextImpl::SetLevel(a, (size_t)I(3));
      ct[w.at(1)] = a;
    } else if (op == "store") {
      const Ciphertext& c = C(w.at(1));
      const RawCipherText r = cc->StoreCiphertext(c);
      std::vector<uint64_t> out;
      for (const auto& limb : r.sub_0) out.insert(out.end(), limb.begin(), limb.end());
      for (const auto& limb : r.sub_1) out.insert(out.end(), limb.begin(), limb.end());
      write_u64(w.at(1) + ".bin", out);
      meta << w.at(1) << " " << r.numRes << " " << r.NoiseLevel << " " << r.Noise << " " << r.slots << " "
           << (c->IsExtended() ? 1 : 0) << "\n";
    } else if (op == "expect_throw") {
      bool threw = false;
      try {
        run(std::vector<std::string>(w.begin() + 1, w.end()));
      } catch (const std::runtime_error&) {
        threw = true;
      }
      if (!threw) throw std::runtime_error("expected an exception: " + w.at(1));
      meta << "throws " << w.at(1) << "\n";
    } else {
      throw std::runtime_error("unknown command " + op);
    }
  }
};

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: api_script DIR\n");
    return 2;
  }
  g_dir = argv[1];
  State S;
  S.meta.open(g_dir + "/meta.txt");
  S.meta << std::setprecision(17);
  std::ifstream f(g_dir + "/script.txt");
  std::string line;
  int n = 0;
  try {
    while (std::getline(f, line)) {
      ++n;
      std::istringstream is(line);
      std::vector<std::string> w;
      for (std::string t; is >> t;) w.push_back(t);
      if (w.empty() || w[0][0] == '#') continue;
      S.run(w);
    }
    S.cc->Synchronize();
  } catch (const std::exception& e) {
    std::fprintf(stderr, "api_script: line %d (%s): %s\n", n, line.c_str(), e.what());
    return 1;
  }
  std::printf("api_script: %d lines ok\n", n);
  return 0;
}
