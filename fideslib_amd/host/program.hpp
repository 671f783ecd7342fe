// program.hpp -- internal to the host C++ API (not installed): a straight-
// line program executor over device ciphertext registers and the Chebyshev
// series evaluator built on it.  It mirrors the Python program builder
// (fideslib_amd/serve.py Program, chebyshev_eval): every operation is one
// call of the batched C ABI over the executor's batch, levels are tracked in Q
// limbs and scales nominally (a rescaled scale is scale / q_top), and
// integer constants are rounded half to even exactly as Python's round().
#ifndef FIDESLIB_AMD_HOST_PROGRAM_HPP
#define FIDESLIB_AMD_HOST_PROGRAM_HPP

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fideslib/fideslib.hpp"
#include "fideslib_amd.h"

namespace fideslib {
namespace detail {

using i128 = __int128;

inline void chk(int rc, const char* what) {
  if (rc != 0) {
    const char* m = fides_last_error();
    throw std::runtime_error(std::string(what) + ": " + (m ? m : ""));
  }
}

inline uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  unsigned __int128 r = 1, x = b % m;
  while (e) {
    if (e & 1) r = r * x % m;
    x = x * x % m;
    e >>= 1;
  }
  return (uint64_t)r;
}


// Device memory of the registers.  A Pool keeps released blocks by size and
// hands them out again, so a program that runs many steps allocates each
// register shape once instead of once per step; blocks go back to the
// device when the pool is destroyed (after the last register using it).
struct Pool {
  std::multimap<size_t, uint64_t*> free;
  ~Pool() {
    for (auto& kv : free) fides_device_free(kv.second);
  }
  uint64_t* get(size_t words) {
    auto it = free.find(words);
    if (it != free.end()) {
      uint64_t* d = it->second;
      free.erase(it);
      return d;
    }
    void* p = nullptr;
    chk(fides_device_alloc(words * 8, &p), "device buffer");
    return (uint64_t*)p;
  }
};

// A device buffer owned by a register (own == false: a view of caller
// memory), returned to its pool (if any) when released.
struct DevBuf {
  uint64_t* d = nullptr;
  bool own = true;
  size_t words = 0;
  std::shared_ptr<Pool> pool;
  DevBuf(uint64_t* view, bool) : d(view), own(false) {}
  explicit DevBuf(size_t w, std::shared_ptr<Pool> p = nullptr) : words(w), pool(std::move(p)) {
    if (pool) {
      d = pool->get(w);
    } else {
      void* q = nullptr;
      chk(fides_device_alloc(w * 8, &q), "device buffer");
      d = (uint64_t*)q;
    }
  }
  ~DevBuf() {
    if (!d || !own) return;
    if (pool)
      pool->free.emplace(words, d);
    else
      fides_device_free(d);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// A register holds one ciphertext for each job of the executor's batch, all
// at the same level and nominal scale, in one buffer: job i's [2][level][N]
// starts `stride` words after job i-1's.
struct Reg {
  std::shared_ptr<DevBuf> buf;
  uint32_t level = 0;  // limbs
  double scale = 0;
  size_t stride = 0;   // words per job
  uint64_t* d(uint32_t i) const { return buf->d + (size_t)i * stride; }
};


// A prepared device plaintext (evaluation form, Montgomery representation).
struct DevPt {
  std::shared_ptr<DevBuf> buf;
  uint32_t level;
  double scale;
};

// The program builder/executor (mirrors serve.Program) over a batch of n
// ciphertexts running the same program in lockstep: every operation is one
// call of the batched C ABI with n jobs (the same constants for every job).
// With `dry` set no device work is issued: registers carry levels and
// nominal scales only (the structure of a program, e.g. its output scale).
struct Exec {
  CryptoContextImpl& cc;
  fides_context* ctx;
  const std::vector<uint64_t>& q;
  uint32_t N;
  double delta;
  const fides_key* rlk;
  bool dry = false;
  uint32_t n = 1;
  std::shared_ptr<Pool> pool = std::make_shared<Pool>();

  Reg fresh(uint32_t level, double scale) {
    if (level < 1) throw std::runtime_error("out of levels");
    Reg r;
    r.stride = (size_t)2 * level * N;
    if (!dry) r.buf = std::make_shared<DevBuf>(r.stride * n, pool);
    r.level = level;
    r.scale = scale;
    return r;
  }
  // one pointer per job
  std::vector<const uint64_t*> in_ptrs(const Reg& a) const {
    std::vector<const uint64_t*> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = a.d(i);
    return v;
  }
  std::vector<uint64_t*> out_ptrs(const Reg& a) const {
    std::vector<uint64_t*> v(n);
    for (uint32_t i = 0; i < n; ++i) v[i] = a.d(i);
    return v;
  }
  std::vector<uint32_t> same(uint32_t x) const { return std::vector<uint32_t>(n, x); }
  std::vector<uint64_t> residues(i128 x, uint32_t l) const {
    std::vector<uint64_t> r(l);
    for (uint32_t j = 0; j < l; ++j) {
      i128 v = x % (i128)q[j];
      if (v < 0) v += q[j];
      r[j] = (uint64_t)v;
    }
    return r;
  }
  static i128 round_i128(double x) { return (i128)std::nearbyint(x); }

  Reg drop(const Reg& a, uint32_t level) {
    if (level == a.level) return a;
    Reg o = fresh(level, a.scale);
    if (dry) return o;
    const auto in = in_ptrs(a);
    const auto out = out_ptrs(o);
    const auto li = same(a.level), lo = same(level);
    chk(fides_drop_limbs_batch(ctx, n, in.data(), li.data(), lo.data(), out.data(), nullptr), "drop");
    return o;
  }
  Reg hmul(Reg a, Reg b) {  // with Rescale
    const uint32_t l = std::min(a.level, b.level);
    a = drop(a, l);
    b = drop(b, l);
    if (l < 2) throw std::runtime_error("HMult+Rescale needs 2 limbs");
    Reg o = fresh(l - 1, a.scale * b.scale / (double)q[l - 1]);
    if (dry) return o;
    const auto pa = in_ptrs(a), pb = in_ptrs(b);
    const auto po = out_ptrs(o);
    const auto ls = same(l);
    chk(fides_hmult_batch(ctx, n, pa.data(), pb.data(), ls.data(), rlk, 1, po.data(), nullptr), "HMult");
    return o;
  }
  // automorphism k with `key` (the rotation key of k by default; k = 1 with
  // a switching key: a key switch to another secret)
  Reg automorph(const Reg& a, uint32_t k, const fides_key* key = nullptr) {
    Reg o = fresh(a.level, a.scale);
    if (dry) return o;
    if (!key) key = cc.RotationKeyFor(k);
    const auto in = in_ptrs(a);
    const auto out = out_ptrs(o);
    const auto ls = same(a.level), ks = same(k);
    const std::vector<const fides_key*> keys(n, key);
    chk(fides_rotate_batch(ctx, n, in.data(), ls.data(), keys.data(), ks.data(), out.data(), nullptr), "rotate");
    return o;
  }
  Reg rotate(const Reg& a, int64_t r) {
    const int64_t half = N / 2;
    const int64_t rr = ((r % half) + half) % half;
    if (rr == 0) return a;
    return automorph(a, (uint32_t)powmod(5, (uint64_t)rr, 2ull * N));
  }
  // Several rotations in one batched call (rotations of the same register
  // share one ModUp there); amount 0 returns the register itself.
  std::vector<Reg> rotate_many(const std::vector<std::pair<Reg, int64_t>>& items) {
    std::vector<Reg> outs(items.size());
    std::vector<const uint64_t*> in;
    std::vector<uint64_t*> out;
    std::vector<uint32_t> ls, ks;
    std::vector<const fides_key*> keys;
    const int64_t half = N / 2;
    for (size_t j = 0; j < items.size(); ++j) {
      const Reg& a = items[j].first;
      const int64_t rr = ((items[j].second % half) + half) % half;
      if (rr == 0) {
        outs[j] = a;
        continue;
      }
      outs[j] = fresh(a.level, a.scale);
      if (dry) continue;
      const uint32_t k = (uint32_t)powmod(5, (uint64_t)rr, 2ull * N);
      const fides_key* key = cc.RotationKeyFor(k);
      for (uint32_t i = 0; i < n; ++i) {
        in.push_back(a.d(i));
        out.push_back(outs[j].d(i));
        ls.push_back(a.level);
        ks.push_back(k);
        keys.push_back(key);
      }
    }
    if (!in.empty())
      chk(fides_rotate_batch(ctx, (uint32_t)in.size(), in.data(), ls.data(), keys.data(), ks.data(), out.data(),
                             nullptr),
          "rotate");
    return outs;
  }
  Reg conjugate(const Reg& a) { return automorph(a, 2 * N - 1); }
  Reg switch_key(const Reg& a, const fides_key* key) {
    if (!key && !dry) throw std::runtime_error("switching key not loaded");
    return automorph(a, 1, key);
  }
  // out = a*mul (+/-) b + add (integers)
  Reg lin(Reg a, const i128* mul, const Reg* b0, bool sub, const i128* add, double scale) {
    Reg b;
    if (b0) {
      const uint32_t l = std::min(a.level, b0->level);
      a = drop(a, l);
      b = drop(*b0, l);
    }
    Reg o = fresh(a.level, scale);
    if (dry) return o;
    const uint32_t l = a.level;
    std::vector<uint64_t> mr, ar;
    if (mul) mr = residues(*mul, l);
    if (add) ar = residues(*add, l);
    const auto pa = in_ptrs(a);
    std::vector<const uint64_t*> pb(n, nullptr);
    if (b0) pb = in_ptrs(b);
    const std::vector<const uint64_t*> pm(n, mul ? mr.data() : nullptr), pd(n, add ? ar.data() : nullptr);
    const auto ss = same(sub ? 1 : 0), ls = same(l);
    const auto po = out_ptrs(o);
    chk(fides_lin_batch(ctx, n, pa.data(), pb.data(), ls.data(), pm.data(), pd.data(), ss.data(), po.data(), nullptr),
        "linear");
    return o;
  }
  Reg add(const Reg& a, const Reg& b) { return lin(a, nullptr, &b, false, nullptr, a.scale); }
  Reg sub(const Reg& a, const Reg& b) { return lin(a, nullptr, &b, true, nullptr, a.scale); }
  Reg rescale(const Reg& a) {
    if (a.level < 2) throw std::runtime_error("Rescale needs 2 limbs");
    Reg o = fresh(a.level - 1, a.scale / (double)q[a.level - 1]);
    if (dry) return o;
    const auto in = in_ptrs(a);
    const auto out = out_ptrs(o);
    const auto ls = same(a.level);
    chk(fides_rescale_batch(ctx, n, in.data(), ls.data(), out.data(), nullptr), "Rescale");
    return o;
  }
  Reg mul_real(const Reg& a, double x) {
    const i128 m = round_i128(x * delta);
    return rescale(lin(a, &m, nullptr, false, nullptr, a.scale * delta));
  }
  Reg match_scale(const Reg& a, double scale, uint32_t level) {
    if (a.level <= level) throw std::runtime_error("match_scale needs a spare level");
    const double qq = (double)q[a.level - 1];
    const i128 m = round_i128(scale * qq / a.scale);
    Reg t = lin(a, &m, nullptr, false, nullptr, a.scale * (double)m);
    return drop(rescale(t), level);
  }
  Reg lincomb(const std::vector<std::pair<Reg, i128>>& terms, const i128* add, double scale) {
    uint32_t l = ~0u;
    for (const auto& t : terms) l = std::min(l, t.first.level);
    Reg o = fresh(l, scale);
    if (dry) return o;
    const auto cnt = same((uint32_t)terms.size());
    std::vector<std::vector<uint64_t>> mres;
    for (const auto& t : terms) mres.push_back(residues(t.second, l));
    std::vector<const uint64_t*> xs, mp;
    std::vector<uint32_t> xl;
    for (uint32_t i = 0; i < n; ++i)
      for (size_t j = 0; j < terms.size(); ++j) {
        xs.push_back(terms[j].first.d(i));
        xl.push_back(terms[j].first.level);
        mp.push_back(mres[j].data());
      }
    std::vector<uint64_t> ar;
    if (add) ar = residues(*add, l);
    const std::vector<const uint64_t*> pa(n, add ? ar.data() : nullptr);
    const auto ls = same(l);
    const auto po = out_ptrs(o);
    chk(fides_lincomb_batch(ctx, n, cnt.data(), xs.data(), xl.data(), mp.data(), pa.data(), ls.data(), po.data(),
                            nullptr),
        "lincomb");
    return o;
  }
  Reg ptdot(const std::vector<std::pair<Reg, const DevPt*>>& terms) {
    uint32_t l = ~0u;
    for (const auto& t : terms) l = std::min(l, t.first.level);
    std::vector<Reg> regs;
    for (const auto& t : terms) {
      regs.push_back(drop(t.first, l));
      if (t.second->level < l) throw std::runtime_error("plaintext below the ciphertext level");
    }
    Reg o = fresh(l, regs[0].scale * terms[0].second->scale);
    if (dry) return o;
    const auto cnt = same((uint32_t)terms.size());
    std::vector<const uint64_t*> cs, ps;
    for (uint32_t i = 0; i < n; ++i)
      for (size_t j = 0; j < terms.size(); ++j) {
        cs.push_back(regs[j].d(i));
        ps.push_back(terms[j].second->buf->d);
      }
    const auto ls = same(l);
    const auto po = out_ptrs(o);
    chk(fides_ptdot_batch(ctx, n, cnt.data(), cs.data(), ps.data(), ls.data(), po.data(), nullptr),
        "plaintext product");
    return o;
  }
  // Several plaintext dot products in one batched call.
  std::vector<Reg> ptdot_many(const std::vector<std::vector<std::pair<Reg, const DevPt*>>>& groups) {
    std::vector<Reg> outs;
    std::vector<uint32_t> cnt, ls;
    std::vector<const uint64_t*> cs, ps;
    std::vector<uint64_t*> po;
    std::vector<std::vector<Reg>> keep;
    for (const auto& terms : groups) {
      uint32_t l = ~0u;
      for (const auto& t : terms) l = std::min(l, t.first.level);
      std::vector<Reg> regs;
      for (const auto& t : terms) {
        regs.push_back(drop(t.first, l));
        if (t.second->level < l) throw std::runtime_error("plaintext below the ciphertext level");
      }
      outs.push_back(fresh(l, regs[0].scale * terms[0].second->scale));
      if (dry) continue;
      for (uint32_t i = 0; i < n; ++i) {
        cnt.push_back((uint32_t)terms.size());
        ls.push_back(l);
        po.push_back(outs.back().d(i));
        for (size_t j = 0; j < terms.size(); ++j) {
          cs.push_back(regs[j].d(i));
          ps.push_back(terms[j].second->buf->d);
        }
      }
      keep.push_back(std::move(regs));
    }
    if (!dry && !cnt.empty())
      chk(fides_ptdot_batch(ctx, (uint32_t)cnt.size(), cnt.data(), cs.data(), ps.data(), ls.data(), po.data(), nullptr),
          "plaintext product");
    return outs;
  }
  Reg mod_raise(Reg a, uint32_t level, double scale) {
    a = drop(a, 1);
    Reg o = fresh(level, scale);
    if (dry) return o;
    const auto in = in_ptrs(a);
    const auto out = out_ptrs(o);
    const auto ls = same(level);
    chk(fides_mod_raise_batch(ctx, n, in.data(), ls.data(), out.data(), nullptr), "ModRaise");
    return o;
  }
};

// ---- Chebyshev series by baby steps and giant steps (serve.chebyshev_eval)
struct Cheb {
  Exec* X;  // null: structural dry run (levels only)
  std::map<uint32_t, Reg> T;
  std::map<uint32_t, uint32_t> TL;  // levels (dry run and real)
  uint32_t m = 2;

  uint32_t lvl(uint32_t k) const { return TL.at(k); }
};

inline std::vector<uint32_t> leaf_terms(const std::vector<double>& cs) {
  std::vector<uint32_t> ks;
  for (uint32_t k = 1; k < cs.size(); ++k)
    if (cs[k] != 0.0) ks.push_back(k);
  if (ks.empty()) ks.push_back(1);
  return ks;
}

inline void split(const std::vector<double>& cs, uint32_t Mg, std::vector<double>& qc, std::vector<double>& rc) {
  qc.assign(1, cs[Mg]);
  for (size_t k = 1; k < cs.size() - Mg; ++k) qc.push_back(2.0 * cs[Mg + k]);
  rc.assign(Mg, 0.0);
  for (uint32_t j = 0; j < Mg; ++j) rc[j] = cs[j] - (2 * Mg - j < cs.size() ? cs[2 * Mg - j] : 0.0);
}

inline uint32_t giant_of(const Cheb& C, size_t n) {
  uint32_t g = C.m;
  while (2 * (size_t)g < n) g *= 2;
  return g;
}

inline uint32_t node_level(const Cheb& C, const std::vector<double>& cs) {
  if (cs.size() <= C.m) {
    uint32_t l = ~0u;
    for (uint32_t k : leaf_terms(cs)) l = std::min(l, C.lvl(k));
    return l - 1;
  }
  const uint32_t Mg = giant_of(C, cs.size());
  std::vector<double> qc, rc;
  split(cs, Mg, qc, rc);
  return std::min(std::min(node_level(C, qc), C.lvl(Mg)) - 1, node_level(C, rc));
}

inline Reg node_build(Cheb& C, const std::vector<double>& cs, double S) {
  Exec& X = *C.X;
  if (cs.size() <= C.m) {
    const auto ks = leaf_terms(cs);
    uint32_t l = ~0u;
    for (uint32_t k : ks) l = std::min(l, C.lvl(k));
    const double target = S * (double)X.q[l - 1];
    std::vector<std::pair<Reg, i128>> terms;
    for (uint32_t k : ks)
      terms.emplace_back(C.T.at(k), k < cs.size() ? Exec::round_i128(cs[k] * target / C.T.at(k).scale) : 0);
    const i128 add = Exec::round_i128(cs[0] * target);
    return X.rescale(X.lincomb(terms, &add, target));
  }
  const uint32_t Mg = giant_of(C, cs.size());
  std::vector<double> qc, rc;
  split(cs, Mg, qc, rc);
  const uint32_t lp = std::min(node_level(C, qc), C.lvl(Mg));
  Reg qr = node_build(C, qc, S * (double)X.q[lp - 1] / C.T.at(Mg).scale);
  Reg prod = X.hmul(qr, C.T.at(Mg));
  return X.add(prod, node_build(C, rc, S));
}

// Levels of the basis T_1..T_m and giants (the product rule and doubling)
// without a device: the structure of Cheb's basis construction.
inline void basis_levels(Cheb& C, uint32_t y_level, uint32_t deg) {
  C.TL.clear();
  C.TL[1] = y_level;
  const uint32_t upto = std::max(1u, std::min(C.m, deg));
  uint32_t m = 1;
  while (m < upto) {
    for (uint32_t k = m + 1; k <= std::min(2 * m, upto); ++k) {
      const uint32_t n = k - m;
      const uint32_t pl = std::min(C.TL[m], C.TL[n]) - 1;  // product level
      C.TL[k] = pl;  // the subtraction operand is aligned down to it
    }
    m *= 2;
  }
  uint32_t M = C.m;
  while (2 * M <= deg) {
    C.TL[2 * M] = C.TL[M] - 1;
    M *= 2;
  }
}

inline Reg chebyshev_eval(Exec& X, const Reg& y, std::vector<double> c) {
  size_t deg = c.size() - 1;
  while (deg > 0 && c[deg] == 0.0) --deg;
  c.resize(deg + 1);
  if (deg < 1) throw std::runtime_error("Chebyshev degree must be >= 1");
  Cheb C;
  C.X = &X;
  while ((size_t)C.m * C.m < deg + 1) C.m *= 2;
  const uint32_t upto = std::max<uint32_t>(1, std::min<uint32_t>(C.m, (uint32_t)deg));
  C.T[1] = y;
  uint32_t m = 1;
  while (m < upto) {
    for (uint32_t k = m + 1; k <= std::min(2 * m, upto); ++k) {
      const uint32_t n = k - m;
      Reg prod = X.hmul(C.T[m], C.T[n]);
      if (n == m) {
        const i128 two = 2, add = -Exec::round_i128(prod.scale);
        C.T[k] = X.lin(prod, &two, nullptr, false, &add, prod.scale);
      } else {
        Reg b = C.T[m - n];
        if (b.scale != prod.scale && b.level > prod.level) b = X.match_scale(b, prod.scale, prod.level);
        const i128 two = 2;
        C.T[k] = X.lin(prod, &two, &b, true, nullptr, prod.scale);
      }
    }
    m *= 2;
  }
  uint32_t M = C.m;
  while (2 * (size_t)M <= deg) {
    Reg sq = X.hmul(C.T[M], C.T[M]);
    const i128 two = 2, add = -Exec::round_i128(sq.scale);
    C.T[2 * M] = X.lin(sq, &two, nullptr, false, &add, sq.scale);
    M *= 2;
  }
  for (auto& kv : C.T) C.TL[kv.first] = kv.second.level;
  return node_build(C, c, X.delta);
}

inline uint32_t chebyshev_eval_depth(const std::vector<double>& c0) {
  std::vector<double> c = c0;
  size_t deg = c.size() - 1;
  while (deg > 0 && c[deg] == 0.0) --deg;
  c.resize(deg + 1);
  Cheb C;
  C.X = nullptr;
  while ((size_t)C.m * C.m < deg + 1) C.m *= 2;
  basis_levels(C, 64, (uint32_t)deg);
  return 64 - node_level(C, c);
}

}  // namespace detail
}  // namespace fideslib

#endif
