// bootstrap.cpp -- CKKS bootstrapping in the host C++ API.
//
// The same algorithm and plan as fideslib_amd/bootstrap.py (reference
// Bootstrap.cu:169-333 on OpenFHE's precomputation; here every table is
// derived from the mathematics):
//   prescale x 2^-c, one limb -> ModRaise -> CoeffToSlot (level budget groups,
//   baby-step / giant-step plaintext products) -> Re/Im split by conjugation
//   -> EvalMod (Chebyshev series of cos(2 pi (K'y - 1/4) / 2^r) by baby and
//   giant steps, r double angles) -> recombine -> SlotToCoeff -> x 2^c.
// Operations run one at a time through the C ABI on the context's default
// stream; levels and nominal scales are tracked exactly as the Python program
// builder does (FIXEDMANUAL semantics: a rescaled scale is scale / q_top).
#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "fideslib/fideslib.hpp"
#include "fideslib_amd.h"

namespace fideslib {

namespace {

using cd = std::complex<double>;
using Vec = std::vector<cd>;
using Diag = std::map<uint32_t, Vec>;  // offset mod n -> diagonal
using i128 = __int128;

void chk(int rc, const char* what) {
  if (rc != 0) {
    const char* m = fides_last_error();
    throw std::runtime_error(std::string(what) + ": " + (m ? m : ""));
  }
}

uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  unsigned __int128 r = 1, x = b % m;
  while (e) {
    if (e & 1) r = r * x % m;
    x = x * x % m;
    e >>= 1;
  }
  return (uint64_t)r;
}

Vec rot(const Vec& v, int64_t d) {  // rot(v, d)[p] = v[(p + d) mod n]
  const int64_t n = (int64_t)v.size();
  Vec o(v.size());
  for (int64_t p = 0; p < n; ++p) o[p] = v[(size_t)(((p + d) % n + n) % n)];
  return o;
}

// Butterfly stage of length 2^s (the decoding map's factor B_s) or its inverse.
Diag stage_diags(uint32_t n, uint32_t N, uint32_t s, bool inverse) {
  const uint32_t ln = 1u << s, lenh = ln >> 1, lenq = ln << 2;
  const uint64_t M = 2ull * N;
  std::vector<cd> om(lenh);
  for (uint32_t x = 0; x < lenh; ++x) {
    const double a = 2.0 * M_PI * (double)(powmod(5, x, M) % lenq) / (double)lenq;
    om[x] = cd(std::cos(a), std::sin(a));
  }
  Vec d0(n), dp(n), dm(n);
  for (uint32_t p = 0; p < n; ++p) {
    const uint32_t j = p % ln;
    const bool top = j < lenh;
    const cd w = om[top ? j : j - lenh];
    if (!inverse) {
      if (top) {
        d0[p] = 1.0;
        dp[p] = w;
      } else {
        d0[p] = -w;
        dm[p] = 1.0;
      }
    } else {
      if (top) {
        d0[p] = 0.5;
        dp[p] = 0.5;
      } else {
        d0[p] = -0.5 * std::conj(w);
        dm[p] = 0.5 * std::conj(w);
      }
    }
  }
  Diag out;
  auto add = [&](uint32_t off, const Vec& v) {
    auto it = out.find(off);
    if (it == out.end()) {
      out[off] = v;
    } else {
      for (uint32_t i = 0; i < n; ++i) it->second[i] += v[i];
    }
  };
  add(0, d0);
  add(lenh % n, dp);
  add((n - lenh % n) % n, dm);
  return out;
}

Diag diag_mul(const Diag& A, const Diag& B, uint32_t n) {  // A B (apply B first)
  Diag out;
  for (const auto& [e, a] : A)
    for (const auto& [d, b] : B) {
      const uint32_t k = (d + e) % n;
      const Vec rb = rot(b, e);
      auto it = out.find(k);
      if (it == out.end()) it = out.emplace(k, Vec(n)).first;
      for (uint32_t i = 0; i < n; ++i) it->second[i] += a[i] * rb[i];
    }
  Diag kept;
  for (auto& [k, v] : out) {
    double mx = 0;
    for (const cd& x : v) mx = std::max(mx, std::abs(x));
    if (mx > 1e-300) kept.emplace(k, std::move(v));
  }
  return kept;
}

std::vector<std::vector<uint32_t>> split_stages(uint32_t logn, uint32_t budget) {
  budget = std::max(1u, std::min(budget, logn));
  const uint32_t base = logn / budget, extra = logn % budget;
  std::vector<std::vector<uint32_t>> out;
  uint32_t s = 1;
  for (uint32_t i = 0; i < budget; ++i) {
    const uint32_t sz = base + (i < extra ? 1 : 0);
    std::vector<uint32_t> g;
    for (uint32_t t = s; t < s + sz; ++t) g.push_back(t);
    out.push_back(g);
    s += sz;
  }
  return out;
}

std::vector<Diag> transform_groups(uint32_t n, uint32_t N, uint32_t budget, bool inverse) {
  uint32_t logn = 0;
  while ((1u << logn) < n) ++logn;
  std::vector<Diag> mats;
  for (auto g : split_stages(logn, budget)) {
    if (inverse) std::reverse(g.begin(), g.end());
    Diag M;
    bool first = true;
    for (uint32_t s : g) {
      Diag S = stage_diags(n, N, s, inverse);
      M = first ? S : diag_mul(S, M, n);
      first = false;
    }
    mats.push_back(M);
  }
  if (inverse) std::reverse(mats.begin(), mats.end());
  return mats;
}

struct GroupPlan {
  uint32_t stride = 0, g = 1, n = 0;
  std::vector<int64_t> babies;
  std::map<int64_t, std::vector<std::pair<int64_t, Vec>>> giants;  // k -> [(b, diagonal)]
};

int64_t floordiv(int64_t a, int64_t b) { return a / b - ((a % b != 0) && ((a < 0) != (b < 0))); }

GroupPlan plan_group(const Diag& D, uint32_t n, cd factor) {
  uint64_t stride = 0;
  for (const auto& kv : D) stride = std::gcd<uint64_t, uint64_t>(stride, kv.first);
  if (stride == 0) stride = n;
  const int64_t Mn = (int64_t)(n / stride);
  std::vector<int64_t> ms;
  for (const auto& kv : D) ms.push_back((((int64_t)(kv.first / stride) + Mn / 2) % Mn) - Mn / 2);
  std::sort(ms.begin(), ms.end());
  const double span = (double)(ms.back() - ms.front() + 1);
  uint32_t g = 1;
  const int e = (int)std::ceil(std::log2(std::sqrt(span)));
  if (e > 0) g = 1u << e;
  g = std::min(g, 16u);
  GroupPlan P;
  P.stride = (uint32_t)stride;
  P.g = g;
  P.n = n;
  std::vector<int64_t> babies;
  for (int64_t m : ms) {
    const int64_t k = floordiv(m, g), b = m - k * g;
    const uint32_t off = (uint32_t)(((m * (int64_t)stride) % (int64_t)n + n) % n);
    Vec diag = D.at(off);
    for (cd& x : diag) x *= factor;
    P.giants[k].emplace_back(b, rot(diag, -k * (int64_t)g * (int64_t)stride));
    babies.push_back(b);
  }
  std::sort(babies.begin(), babies.end());
  babies.erase(std::unique(babies.begin(), babies.end()), babies.end());
  P.babies = babies;
  return P;
}

// In-place iterative radix-2 FFT, numpy.fft.fft convention (e^{-2 pi i jk/N}).
void fft(std::vector<cd>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = -2.0 * M_PI / (double)len;
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        const cd w(std::cos(ang * (double)k), std::sin(ang * (double)k));
        const cd u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

std::vector<double> chebyshev_coefficients(double (*f)(double, const double*), const double* fa, uint32_t deg) {
  std::vector<double> fx(deg + 1), c(deg + 1);
  for (uint32_t k = 0; k <= deg; ++k) fx[k] = f(std::cos(M_PI * (k + 0.5) / (deg + 1)), fa);
  for (uint32_t j = 0; j <= deg; ++j) {
    double s = 0;
    for (uint32_t k = 0; k <= deg; ++k) s += fx[k] * std::cos(M_PI * j * (k + 0.5) / (deg + 1));
    c[j] = 2.0 / (deg + 1) * s;
  }
  c[0] /= 2;
  return c;
}

double evalmod_fn(double y, const double* a) {  // cos(2 pi (K y - 1/4) / 2^r)
  return std::cos(2.0 * M_PI * (a[0] * y - 0.25) / a[1]);
}

}  // namespace

// ---------------------------------------------------------------- execution
namespace {

struct DevBuf {
  uint64_t* d = nullptr;
  explicit DevBuf(size_t words) {
    void* p = nullptr;
    chk(fides_device_alloc(words * 8, &p), "bootstrap buffer");
    d = (uint64_t*)p;
  }
  ~DevBuf() {
    if (d) fides_device_free(d);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

struct Reg {
  std::shared_ptr<DevBuf> buf;
  uint32_t level = 0;  // limbs
  double scale = 0;
};

}  // namespace

struct BootPlan {
  BootstrapConfig cfg;
  uint32_t n = 0, depth = 0, evalmod_depth = 0, l_raise = 0;
  double f_cts = 0, f_stc = 0;
  std::vector<GroupPlan> cts, stc;
  std::vector<double> cheb;
  // device plaintexts: per group, per giant k, (baby b, prepared plaintext)
  struct Pt {
    std::shared_ptr<DevBuf> buf;
    uint32_t level;
    double scale;
  };
  std::vector<std::map<int64_t, std::vector<std::pair<int64_t, Pt>>>> cts_pts, stc_pts;
  Pt minus_i, plus_i;
  std::vector<int32_t> rotations;
};

namespace {

// The program builder/executor over one ciphertext (mirrors serve.Program).
struct Exec {
  CryptoContextImpl& cc;
  fides_context* ctx;
  const std::vector<uint64_t>& q;
  uint32_t N;
  double delta;
  const fides_key* rlk;

  Reg fresh(uint32_t level, double scale) {
    if (level < 1) throw std::runtime_error("EvalBootstrap: out of levels");
    Reg r;
    r.buf = std::make_shared<DevBuf>((size_t)2 * level * N);
    r.level = level;
    r.scale = scale;
    return r;
  }
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
    const uint64_t* in = a.buf->d;
    uint64_t* out = o.buf->d;
    chk(fides_drop_limbs_batch(ctx, 1, &in, &a.level, &level, &out, nullptr), "drop");
    return o;
  }
  Reg hmul(Reg a, Reg b) {  // with Rescale
    const uint32_t l = std::min(a.level, b.level);
    a = drop(a, l);
    b = drop(b, l);
    if (l < 2) throw std::runtime_error("EvalBootstrap: HMult+Rescale needs 2 limbs");
    Reg o = fresh(l - 1, a.scale * b.scale / (double)q[l - 1]);
    const uint64_t* pa = a.buf->d;
    const uint64_t* pb = b.buf->d;
    uint64_t* po = o.buf->d;
    chk(fides_hmult_batch(ctx, 1, &pa, &pb, &l, rlk, 1, &po, nullptr), "HMult");
    return o;
  }
  Reg automorph(const Reg& a, uint32_t k) {
    Reg o = fresh(a.level, a.scale);
    const uint64_t* in = a.buf->d;
    uint64_t* out = o.buf->d;
    const fides_key* key = cc.RotationKeyFor(k);
    chk(fides_rotate_batch(ctx, 1, &in, &a.level, &key, &k, &out, nullptr), "rotate");
    return o;
  }
  Reg rotate(const Reg& a, int64_t r) {
    const int64_t half = N / 2;
    const int64_t rr = ((r % half) + half) % half;
    if (rr == 0) return a;
    return automorph(a, (uint32_t)powmod(5, (uint64_t)rr, 2ull * N));
  }
  Reg conjugate(const Reg& a) { return automorph(a, 2 * N - 1); }
  // out = a*mul (+/-) b + add (integers)
  Reg lin(Reg a, const i128* mul, const Reg* b0, bool sub, const i128* add, double scale) {
    Reg b;
    if (b0) {
      const uint32_t l = std::min(a.level, b0->level);
      a = drop(a, l);
      b = drop(*b0, l);
    }
    Reg o = fresh(a.level, scale);
    const uint32_t l = a.level;
    std::vector<uint64_t> mr, ar;
    if (mul) mr = residues(*mul, l);
    if (add) ar = residues(*add, l);
    const uint64_t* pa = a.buf->d;
    const uint64_t* pb = b0 ? b.buf->d : nullptr;
    const uint64_t* pm = mul ? mr.data() : nullptr;
    const uint64_t* pd = add ? ar.data() : nullptr;
    const uint32_t s = sub ? 1 : 0;
    uint64_t* po = o.buf->d;
    chk(fides_lin_batch(ctx, 1, &pa, &pb, &l, &pm, &pd, &s, &po, nullptr), "linear");
    return o;
  }
  Reg add(const Reg& a, const Reg& b) { return lin(a, nullptr, &b, false, nullptr, a.scale); }
  Reg sub(const Reg& a, const Reg& b) { return lin(a, nullptr, &b, true, nullptr, a.scale); }
  Reg rescale(const Reg& a) {
    if (a.level < 2) throw std::runtime_error("EvalBootstrap: Rescale needs 2 limbs");
    Reg o = fresh(a.level - 1, a.scale / (double)q[a.level - 1]);
    const uint64_t* in = a.buf->d;
    uint64_t* out = o.buf->d;
    chk(fides_rescale_batch(ctx, 1, &in, &a.level, &out, nullptr), "Rescale");
    return o;
  }
  Reg mul_real(const Reg& a, double x) {
    const i128 m = round_i128(x * delta);
    return rescale(lin(a, &m, nullptr, false, nullptr, a.scale * delta));
  }
  Reg match_scale(const Reg& a, double scale, uint32_t level) {
    if (a.level <= level) throw std::runtime_error("EvalBootstrap: match_scale needs a spare level");
    const double qq = (double)q[a.level - 1];
    const i128 m = round_i128(scale * qq / a.scale);
    Reg t = lin(a, &m, nullptr, false, nullptr, a.scale * (double)m);
    return drop(rescale(t), level);
  }
  Reg lincomb(const std::vector<std::pair<Reg, i128>>& terms, const i128* add, double scale) {
    uint32_t l = ~0u;
    for (const auto& t : terms) l = std::min(l, t.first.level);
    Reg o = fresh(l, scale);
    const uint32_t cnt = (uint32_t)terms.size();
    std::vector<const uint64_t*> xs;
    std::vector<uint32_t> xl;
    std::vector<std::vector<uint64_t>> mres;
    std::vector<const uint64_t*> mp;
    for (const auto& t : terms) {
      xs.push_back(t.first.buf->d);
      xl.push_back(t.first.level);
      mres.push_back(residues(t.second, l));
    }
    for (auto& v : mres) mp.push_back(v.data());
    std::vector<uint64_t> ar;
    if (add) ar = residues(*add, l);
    const uint64_t* pa = add ? ar.data() : nullptr;
    uint64_t* po = o.buf->d;
    chk(fides_lincomb_batch(ctx, 1, &cnt, xs.data(), xl.data(), mp.data(), &pa, &l, &po, nullptr), "lincomb");
    return o;
  }
  Reg ptdot(const std::vector<std::pair<Reg, const BootPlan::Pt*>>& terms) {
    uint32_t l = ~0u;
    for (const auto& t : terms) l = std::min(l, t.first.level);
    std::vector<Reg> regs;
    for (const auto& t : terms) {
      regs.push_back(drop(t.first, l));
      if (t.second->level < l) throw std::runtime_error("EvalBootstrap: plaintext below the ciphertext level");
    }
    Reg o = fresh(l, regs[0].scale * terms[0].second->scale);
    const uint32_t cnt = (uint32_t)terms.size();
    std::vector<const uint64_t*> cs, ps;
    for (size_t i = 0; i < terms.size(); ++i) {
      cs.push_back(regs[i].buf->d);
      ps.push_back(terms[i].second->buf->d);
    }
    uint64_t* po = o.buf->d;
    chk(fides_ptdot_batch(ctx, 1, &cnt, cs.data(), ps.data(), &l, &po, nullptr), "plaintext product");
    return o;
  }
  Reg mod_raise(Reg a, uint32_t level, double scale) {
    a = drop(a, 1);
    Reg o = fresh(level, scale);
    const uint64_t* in = a.buf->d;
    uint64_t* out = o.buf->d;
    chk(fides_mod_raise_batch(ctx, 1, &in, &level, &out, nullptr), "ModRaise");
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

std::vector<uint32_t> leaf_terms(const std::vector<double>& cs) {
  std::vector<uint32_t> ks;
  for (uint32_t k = 1; k < cs.size(); ++k)
    if (cs[k] != 0.0) ks.push_back(k);
  if (ks.empty()) ks.push_back(1);
  return ks;
}

void split(const std::vector<double>& cs, uint32_t Mg, std::vector<double>& qc, std::vector<double>& rc) {
  qc.assign(1, cs[Mg]);
  for (size_t k = 1; k < cs.size() - Mg; ++k) qc.push_back(2.0 * cs[Mg + k]);
  rc.assign(Mg, 0.0);
  for (uint32_t j = 0; j < Mg; ++j) rc[j] = cs[j] - (2 * Mg - j < cs.size() ? cs[2 * Mg - j] : 0.0);
}

uint32_t giant_of(const Cheb& C, size_t n) {
  uint32_t g = C.m;
  while (2 * (size_t)g < n) g *= 2;
  return g;
}

uint32_t node_level(const Cheb& C, const std::vector<double>& cs) {
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

Reg node_build(Cheb& C, const std::vector<double>& cs, double S) {
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
void basis_levels(Cheb& C, uint32_t y_level, uint32_t deg) {
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

Reg chebyshev_eval(Exec& X, const Reg& y, std::vector<double> c) {
  size_t deg = c.size() - 1;
  while (deg > 0 && c[deg] == 0.0) --deg;
  c.resize(deg + 1);
  if (deg < 1) throw std::runtime_error("EvalBootstrap: Chebyshev degree must be >= 1");
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

uint32_t chebyshev_eval_depth(const std::vector<double>& c0) {
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

}  // namespace

// ---------------------------------------------------------------- plan + API
std::vector<int32_t> CryptoContextImpl::EvalBootstrapSetup(const BootstrapConfig& cfg) {
  auto P = std::make_shared<BootPlan>();
  P->cfg = cfg;
  P->n = N / 2;
  const uint32_t n = P->n;
  const double q0 = (double)q[0];
  // CoeffToSlot output y = t / (q_0 K'), the 1/2 of the Re/Im split folded in;
  // SlotToCoeff input sin(2 pi t / q_0) ~ 2 pi m / q_0 -> m / delta
  P->f_cts = boot_delta() / (2.0 * q0 * cfg.K);
  P->f_stc = q0 / (2.0 * M_PI * boot_delta());
  const auto cts = transform_groups(n, N, cfg.budgetCtS, true);
  const auto stc = transform_groups(n, N, cfg.budgetStC, false);
  for (size_t i = 0; i < cts.size(); ++i) P->cts.push_back(plan_group(cts[i], n, i == 0 ? P->f_cts : 1.0));
  for (size_t i = 0; i < stc.size(); ++i) P->stc.push_back(plan_group(stc[i], n, i == 0 ? P->f_stc : 1.0));
  const double fa[2] = {cfg.K, std::ldexp(1.0, (int)cfg.doubleAngle)};
  P->cheb = chebyshev_coefficients(evalmod_fn, fa, cfg.degree);
  P->l_raise = L;
  P->evalmod_depth = chebyshev_eval_depth(P->cheb) + cfg.doubleAngle;
  P->depth = cfg.budgetCtS + cfg.budgetStC + P->evalmod_depth;
  if (P->l_raise <= P->depth) throw std::runtime_error("EvalBootstrapSetup: not enough levels for bootstrapping");
  // rotations
  std::vector<int32_t> rots;
  for (const auto* groups : {&P->cts, &P->stc})
    for (const GroupPlan& g : *groups) {
      for (int64_t b : g.babies)
        if (b) rots.push_back((int32_t)((b * (int64_t)g.stride) % n));
      for (const auto& kv : g.giants)
        if (kv.first) rots.push_back((int32_t)((((kv.first * (int64_t)g.g * (int64_t)g.stride) % n) + n) % n));
    }
  std::sort(rots.begin(), rots.end());
  rots.erase(std::unique(rots.begin(), rots.end()), rots.end());
  rots.erase(std::remove(rots.begin(), rots.end(), 0), rots.end());
  P->rotations = rots;
  // plaintexts: encode v at `level` limbs with scale `scale`
  const std::vector<uint64_t> slot_t = [&] {
    std::vector<uint64_t> t(n);
    uint64_t g = 1;
    for (uint32_t j = 0; j < n; ++j) {
      t[j] = (g - 1) / 2;
      g = g * 5 % (2ull * N);
    }
    return t;
  }();
  auto upload = [&](const std::vector<long double>& coef, uint32_t level, double scale) {
    BootPlan::Pt pt;
    pt.level = level;
    pt.scale = scale;
    pt.buf = std::make_shared<DevBuf>((size_t)level * N);
    std::vector<uint64_t> host((size_t)level * N);
    for (uint32_t j = 0; j < level; ++j)
      for (uint32_t i = 0; i < N; ++i) {
        long double r = std::fmod(coef[i], (long double)q[j]);
        if (r < 0) r += (long double)q[j];
        host[(size_t)j * N + i] = (uint64_t)r;
      }
    chk(fides_copy_to_device(pt.buf->d, host.data(), host.size() * 8, nullptr), "plaintext upload");
    chk(fides_plaintext_prepare(ctx, pt.buf->d, level, nullptr), "plaintext preparation");
    return pt;
  };
  auto encode = [&](const Vec& z, uint32_t level, double scale) {
    std::vector<cd> v(N);
    for (uint32_t j = 0; j < n; ++j) {
      v[slot_t[j]] = z[j];
      v[N - 1 - slot_t[j]] = std::conj(z[j]);
    }
    fft(v);
    std::vector<long double> coef(N);
    for (uint32_t i = 0; i < N; ++i) {
      const double a = -M_PI * (double)i / (double)N;
      const cd m = v[i] / (double)N * cd(std::cos(a), std::sin(a));
      coef[i] = std::nearbyint((long double)(m.real() * scale));
    }
    return upload(coef, level, scale);
  };
  auto monomial = [&](uint64_t power, uint32_t level) {
    std::vector<long double> c(N, 0.0L);
    uint64_t p = power % (2ull * N);
    long double sign = 1;
    if (p >= N) {
      p -= N;
      sign = -1;
    }
    c[p] = sign;
    return upload(c, level, 1.0);
  };
  auto group_pts = [&](const GroupPlan& G, uint32_t level) {
    std::map<int64_t, std::vector<std::pair<int64_t, BootPlan::Pt>>> out;
    const double scale = (double)q[level - 1];
    for (const auto& kv : G.giants)
      for (const auto& bv : kv.second) out[kv.first].emplace_back(bv.first, encode(bv.second, level, scale));
    return out;
  };
  uint32_t l = P->l_raise;
  for (const GroupPlan& G : P->cts) P->cts_pts.push_back(group_pts(G, l--));
  P->minus_i = monomial(3ull * N / 2, l);
  l -= P->evalmod_depth;
  P->plus_i = monomial(N / 2, l);
  for (const GroupPlan& G : P->stc) P->stc_pts.push_back(group_pts(G, l--));
  boot = P;
  return rots;
}

Ciphertext CryptoContextImpl::EvalBootstrap(const Ciphertext& ct) {
  if (!boot) throw std::runtime_error("EvalBootstrap: call EvalBootstrapSetup first");
  if (!evalMultKey) throw std::runtime_error("EvalBootstrap: relinearization key not loaded");
  const BootPlan& P = *boot;
  Exec X{*this, ctx, q, N, boot_delta(), evalMultKey};
  auto linear = [&](Reg x, const GroupPlan& G, const std::map<int64_t, std::vector<std::pair<int64_t, BootPlan::Pt>>>& pts) {
    std::map<int64_t, Reg> babies;
    for (int64_t b : G.babies) babies[b] = b == 0 ? x : X.rotate(x, b * (int64_t)G.stride);
    std::vector<Reg> parts;
    for (const auto& kv : pts) {
      std::vector<std::pair<Reg, const BootPlan::Pt*>> terms;
      for (const auto& bp : kv.second) terms.emplace_back(babies.at(bp.first), &bp.second);
      Reg inner = X.ptdot(terms);
      parts.push_back(kv.first == 0 ? inner : X.rotate(inner, kv.first * (int64_t)G.g * (int64_t)G.stride));
    }
    if (parts.size() == 1) return X.rescale(parts[0]);
    std::vector<std::pair<Reg, i128>> t;
    for (const Reg& p : parts) t.emplace_back(p, 1);
    return X.rescale(X.lincomb(t, nullptr, parts[0].scale));
  };
  auto evalmod = [&](Reg y) {
    Reg c = chebyshev_eval(X, y, P.cheb);
    for (uint32_t i = 0; i < P.cfg.doubleAngle; ++i) {
      Reg sq = X.hmul(c, c);
      const i128 two = 2, add = -Exec::round_i128(sq.scale);
      c = X.lin(sq, &two, nullptr, false, &add, sq.scale);
    }
    return c;
  };
  // the input as a register (its nominal scale is the tracked scaling factor)
  Reg x;
  x.buf = std::make_shared<DevBuf>((size_t)2 * ct->limbs * N);
  chk(fides_copy_device(x.buf->d, ct->DeviceData(), (size_t)2 * ct->limbs * N * 8, nullptr), "copy");
  x.level = ct->limbs;
  x.scale = ct->NoiseFactor > 0 ? ct->NoiseFactor : boot_delta();
  if (x.level < 2) throw std::runtime_error("EvalBootstrap: the input needs at least 2 limbs");
  x = X.mul_real(x, std::ldexp(1.0, -(int)P.cfg.correction));
  Reg r = X.mod_raise(x, P.l_raise, boot_delta());
  for (size_t i = 0; i < P.cts.size(); ++i) r = linear(r, P.cts[i], P.cts_pts[i]);
  Reg cj = X.conjugate(r);
  Reg re = X.add(r, cj);
  Reg d = X.sub(r, cj);
  Reg im = X.ptdot({{d, &P.minus_i}});
  re = evalmod(re);
  im = evalmod(im);
  Reg w = X.add(re, X.ptdot({{im, &P.plus_i}}));
  for (size_t i = 0; i < P.stc.size(); ++i) w = linear(w, P.stc[i], P.stc_pts[i]);
  const i128 corr = (i128)1 << P.cfg.correction;
  Reg out = X.lin(w, &corr, nullptr, false, nullptr, w.scale);
  Ciphertext o = NewCiphertext(out.level);
  chk(fides_copy_device(o->DeviceData(), out.buf->d, (size_t)2 * out.level * N * 8, nullptr), "copy");
  chk(fides_stream_sync(nullptr), "sync");
  o->NoiseLevel = 1;
  o->NoiseFactor = out.scale;
  o->slots = ct->slots;
  o->keyid = ct->keyid;
  return o;
}

}  // namespace fideslib
