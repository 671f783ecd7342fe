// Binary serialization of the host API objects (format in serialize.hpp).
#include "fideslib/serialize.hpp"

#include <cstring>
#include <fstream>
#include <istream>
#include <ostream>
#include <stdexcept>

namespace fideslib {
namespace Serial {

namespace {

constexpr char MAGIC[8] = {'F', 'I', 'D', 'E', 'S', 'A', 'M', 'D'};
constexpr uint32_t VERSION = 1;
enum Tag : uint32_t { T_PARAMS = 1, T_CIPHERTEXT = 2, T_KEY = 3, T_CONTEXT = 4 };

struct W {
  std::ostream& os;
  void bytes(const void* p, size_t n) { os.write(static_cast<const char*>(p), (std::streamsize)n); }
  void u64(uint64_t v) {
    unsigned char b[8];
    for (int i = 0; i < 8; ++i) b[i] = (unsigned char)(v >> (8 * i));
    bytes(b, 8);
  }
  void u32(uint32_t v) {
    unsigned char b[4];
    for (int i = 0; i < 4; ++i) b[i] = (unsigned char)(v >> (8 * i));
    bytes(b, 4);
  }
  void f64(double d) {
    uint64_t v;
    std::memcpy(&v, &d, 8);
    u64(v);
  }
  void str(const std::string& s) {
    u64(s.size());
    bytes(s.data(), s.size());
  }
  void vec(const std::vector<uint64_t>& v) {
    u64(v.size());
    for (uint64_t x : v) u64(x);
  }
  void vecd(const std::vector<double>& v) {
    u64(v.size());
    for (double x : v) f64(x);
  }
  void header(Tag t) {
    bytes(MAGIC, 8);
    u32(VERSION);
    u32(t);
  }
};

struct R {
  std::istream& is;
  void bytes(void* p, size_t n) {
    is.read(static_cast<char*>(p), (std::streamsize)n);
    if (!is) throw std::runtime_error("truncated stream");
  }
  uint64_t u64() {
    unsigned char b[8];
    bytes(b, 8);
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= (uint64_t)b[i] << (8 * i);
    return v;
  }
  uint32_t u32() {
    unsigned char b[4];
    bytes(b, 4);
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) v |= (uint32_t)b[i] << (8 * i);
    return v;
  }
  double f64() {
    const uint64_t v = u64();
    double d;
    std::memcpy(&d, &v, 8);
    return d;
  }
  uint64_t count(uint64_t limit) {
    const uint64_t n = u64();
    if (n > limit) throw std::runtime_error("length field out of range");
    return n;
  }
  std::string str() {
    std::string s(count(1 << 20), '\0');
    if (!s.empty()) bytes(&s[0], s.size());
    return s;
  }
  std::vector<uint64_t> vec(uint64_t limit) {
    std::vector<uint64_t> v(count(limit));
    for (auto& x : v) x = u64();
    return v;
  }
  std::vector<double> vecd(uint64_t limit) {
    std::vector<double> v(count(limit));
    for (auto& x : v) x = f64();
    return v;
  }
  void header(Tag t) {
    char m[8];
    bytes(m, 8);
    if (std::memcmp(m, MAGIC, 8) != 0) throw std::runtime_error("not a serialized fideslib_amd object");
    if (u32() != VERSION) throw std::runtime_error("unsupported serialization version");
    if (u32() != t) throw std::runtime_error("serialized object has a different type");
  }
};

constexpr uint64_t MAXL = 4096, MAXN = 1u << 17;

void put_params(W& w, const RawParams& p) {
  w.u64((uint64_t)p.logN);
  w.u64((uint64_t)p.dnum);
  w.u64((uint64_t)p.scalingTechnique);
  w.vec(p.moduli);
  w.vec(p.root_of_unity);
  w.vec(p.SPECIALmoduli);
  w.vec(p.SPECIALroot_of_unity);
  w.vecd(p.ModReduceFactor);
  w.vecd(p.ScalingFactorReal);
  w.vecd(p.ScalingFactorRealBig);
}

void get_params(R& r, RawParams& p) {
  p.logN = (int)r.u64();
  p.dnum = (int)r.u64();
  const uint64_t t = r.u64();
  if (t > FLEXIBLEAUTOEXT) throw std::runtime_error("bad scaling technique");
  p.scalingTechnique = (ScalingTechnique)t;
  p.moduli = r.vec(MAXL);
  p.root_of_unity = r.vec(MAXL);
  p.SPECIALmoduli = r.vec(MAXL);
  p.SPECIALroot_of_unity = r.vec(MAXL);
  p.ModReduceFactor = r.vecd(MAXL);
  p.ScalingFactorReal = r.vecd(MAXL);
  p.ScalingFactorRealBig = r.vecd(MAXL);
}

void put_key(W& w, const RawKeySwitchKey& k) {
  if (k.r_key.size() != 2) throw std::runtime_error("RawKeySwitchKey: r_key must hold {a, b}");
  const uint64_t dnum = k.r_key[0].size();
  const uint64_t LK = dnum ? k.r_key[0][0].size() : 0;
  const uint64_t N = LK ? k.r_key[0][0][0].size() : 0;
  w.str(k.keyid);
  w.u64(dnum);
  w.u64(LK);
  w.u64(N);
  for (const auto& comp : k.r_key) {
    if (comp.size() != dnum) throw std::runtime_error("RawKeySwitchKey: ragged digits");
    for (const auto& dig : comp) {
      if (dig.size() != LK) throw std::runtime_error("RawKeySwitchKey: ragged limbs");
      for (const auto& limb : dig) {
        if (limb.size() != N) throw std::runtime_error("RawKeySwitchKey: ragged limb size");
        for (uint64_t x : limb) w.u64(x);
      }
    }
  }
}

void get_key(R& r, RawKeySwitchKey& k) {
  k.keyid = r.str();
  const uint64_t dnum = r.count(64), LK = r.count(MAXL), N = r.count(MAXN);
  k.r_key.assign(2, std::vector<std::vector<std::vector<uint64_t>>>(dnum, std::vector<std::vector<uint64_t>>(LK)));
  for (auto& comp : k.r_key)
    for (auto& dig : comp)
      for (auto& limb : dig) {
        limb.resize(N);
        for (auto& x : limb) x = r.u64();
      }
}

template <class F>
bool guarded(SerType t, F f) {
  if (t != BINARY) return false;
  try {
    f();
    return true;
  } catch (const std::exception&) {
    return false;
  }
}

}  // namespace

bool Serialize(std::ostream& os, const RawParams& obj, SerType t) {
  return guarded(t, [&] {
    W w{os};
    w.header(T_PARAMS);
    put_params(w, obj);
  }) && (bool)os;
}

bool Deserialize(std::istream& is, RawParams& obj, SerType t) {
  return guarded(t, [&] {
    R r{is};
    r.header(T_PARAMS);
    RawParams p;
    get_params(r, p);
    obj = p;
  });
}

bool Serialize(std::ostream& os, const RawCipherText& c, SerType t) {
  return guarded(t, [&] {
    if (c.sub_0.size() != c.sub_1.size()) throw std::runtime_error("RawCipherText: component sizes differ");
    W w{os};
    w.header(T_CIPHERTEXT);
    const uint64_t l = c.sub_0.size(), N = l ? c.sub_0[0].size() : 0;
    w.u64(l);
    w.u64(N);
    w.f64(c.Noise);
    w.u64((uint64_t)(int64_t)c.NoiseLevel);
    w.u64((uint64_t)(int64_t)c.slots);
    w.str(c.keyid);
    for (const auto* comp : {&c.sub_0, &c.sub_1})
      for (const auto& limb : *comp) {
        if (limb.size() != N) throw std::runtime_error("RawCipherText: ragged limbs");
        for (uint64_t x : limb) w.u64(x);
      }
  }) && (bool)os;
}

bool Deserialize(std::istream& is, RawCipherText& c, SerType t) {
  return guarded(t, [&] {
    R r{is};
    r.header(T_CIPHERTEXT);
    RawCipherText o;
    const uint64_t l = r.count(MAXL), N = r.count(MAXN);
    o.numRes = (int)l;
    o.N = (int)N;
    o.Noise = r.f64();
    o.NoiseLevel = (int)(int64_t)r.u64();
    o.slots = (int)(int64_t)r.u64();
    o.keyid = r.str();
    for (auto* comp : {&o.sub_0, &o.sub_1}) {
      comp->assign(l, std::vector<uint64_t>(N));
      for (auto& limb : *comp)
        for (auto& x : limb) x = r.u64();
    }
    c = o;
  });
}

bool Serialize(std::ostream& os, const RawKeySwitchKey& k, SerType t) {
  return guarded(t, [&] {
    W w{os};
    w.header(T_KEY);
    put_key(w, k);
  }) && (bool)os;
}

bool Deserialize(std::istream& is, RawKeySwitchKey& k, SerType t) {
  return guarded(t, [&] {
    R r{is};
    r.header(T_KEY);
    RawKeySwitchKey o;
    get_key(r, o);
    k = o;
  });
}

bool Serialize(std::ostream& os, const CryptoContext& cc, SerType t) {
  return guarded(t, [&] {
    if (!cc) throw std::runtime_error("null context");
    W w{os};
    w.header(T_CONTEXT);
    put_params(w, cc->GetRawParams());
    bool hasMult = true;
    RawKeySwitchKey mk;
    try {
      mk = cc->GetEvalMultKey();
    } catch (const std::runtime_error&) {
      hasMult = false;
    }
    w.u64(hasMult ? 1 : 0);
    if (hasMult) put_key(w, mk);
    const auto rot = cc->GetRotationKeys();
    w.u64(rot.size());
    for (const auto& kv : rot) {
      w.u64(kv.first);
      put_key(w, kv.second);
    }
  }) && (bool)os;
}

bool Deserialize(std::istream& is, CryptoContext& cc, SerType t) {
  return guarded(t, [&] {
    R r{is};
    r.header(T_CONTEXT);
    RawParams p;
    get_params(r, p);
    CryptoContext c = CryptoContextImpl::Create(p);
    if (r.u64()) {
      RawKeySwitchKey k;
      get_key(r, k);
      c->AddEvalMultKey(k);
    }
    const uint64_t n = r.count(1 << 20);
    for (uint64_t i = 0; i < n; ++i) {
      const uint32_t a = (uint32_t)r.u64();
      RawKeySwitchKey k;
      get_key(r, k);
      c->AddAutomorphismKey(a, k);
    }
    cc = c;
  });
}

template <class T>
bool SerializeToFile(const std::string& filename, const T& obj, SerType t) {
  std::ofstream f(filename, std::ios::binary);
  return f && Serialize(f, obj, t);
}

template <class T>
bool DeserializeFromFile(const std::string& filename, T& obj, SerType t) {
  std::ifstream f(filename, std::ios::binary);
  return f && Deserialize(f, obj, t);
}

template bool SerializeToFile(const std::string&, const RawParams&, SerType);
template bool SerializeToFile(const std::string&, const RawCipherText&, SerType);
template bool SerializeToFile(const std::string&, const RawKeySwitchKey&, SerType);
template bool SerializeToFile(const std::string&, const CryptoContext&, SerType);
template bool DeserializeFromFile(const std::string&, RawParams&, SerType);
template bool DeserializeFromFile(const std::string&, RawCipherText&, SerType);
template bool DeserializeFromFile(const std::string&, RawKeySwitchKey&, SerType);
template bool DeserializeFromFile(const std::string&, CryptoContext&, SerType);

}  // namespace Serial
}  // namespace fideslib
