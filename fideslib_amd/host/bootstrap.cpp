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
// stream (the executor of program.hpp); levels and nominal scales are tracked
// exactly as the Python program builder does (a rescaled scale is
// scale / q_top).
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
#include "program.hpp"

namespace fideslib {

namespace {

using namespace detail;
using cd = std::complex<double>;
using Vec = std::vector<cd>;
using Diag = std::map<uint32_t, Vec>;  // offset mod n -> diagonal
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

// Sparse packing (ns < N/2 slots): the ns-slot transforms on 2 ns-periodic
// slot vectors, CoeffToSlot ending with the mask (1 on [0, ns), -i on
// [ns, 2 ns)) so that ct + conj(ct) holds [2 Re w | 2 Im w], SlotToCoeff
// starting with the recombination w_p = y_p + i y_(p+ns) (diagonals
// {0: a, ns: b}); the mask folds into the last CoeffToSlot group, the
// recombination into the first SlotToCoeff group (the Python plan's
// sparse_fold).
void sparse_fold(std::vector<Diag>& cts, std::vector<Diag>& stc, uint32_t ns) {
  const uint32_t n2 = 2 * ns;
  auto tile = [&](std::vector<Diag>& groups) {
    for (Diag& D : groups)
      for (auto& kv : D) {
        Vec v(n2);
        for (uint32_t p = 0; p < n2; ++p) v[p] = kv.second[p % kv.second.size()];
        kv.second = std::move(v);
      }
  };
  tile(cts);
  tile(stc);
  for (auto& kv : cts.back())
    for (uint32_t p = ns; p < n2; ++p) kv.second[p] *= cd(0.0, -1.0);
  Vec a(n2), b(n2);
  for (uint32_t p = 0; p < n2; ++p) {
    a[p] = p < ns ? cd(1.0, 0.0) : cd(0.0, 1.0);
    b[p] = p < ns ? cd(0.0, 1.0) : cd(1.0, 0.0);
  }
  Diag comb;
  comb[0] = a;
  comb[ns] = b;
  stc.front() = diag_mul(stc.front(), comb, n2);
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

struct BootPlan {
  BootstrapConfig cfg;
  uint32_t n = 0, depth = 0, evalmod_depth = 0, l_raise = 0;
  double f_cts = 0, f_stc = 0;
  double out_scale = 0;  // nominal scale of the result (0: the tracked one)
  bool flexible = false; // FLEXIBLEAUTO: the prescale also brings the input's scale to delta
  uint32_t half = 0;     // N/2, the ring's slots
  uint32_t slots = 0;    // message slots (N/2: full packing)
  bool sparse = false;   // slots < N/2: partial trace, 2 slots-periodic transforms, one EvalMod
  std::vector<GroupPlan> cts, stc;
  std::vector<double> cheb;
  // device plaintexts: per group, per giant k, (baby b, prepared plaintext)
  using Pt = detail::DevPt;
  std::vector<std::map<int64_t, std::vector<std::pair<int64_t, Pt>>>> cts_pts, stc_pts;
  Pt minus_i, plus_i;
  std::vector<int32_t> rotations;
};

// ---------------------------------------------------------------- plan + API
namespace {

using GroupPts = std::map<int64_t, std::vector<std::pair<int64_t, BootPlan::Pt>>>;

// Sparse packing: the partial trace after ModRaise, x += Rot(x, j slots) for
// j = 1, 2, 4, ... below N / (2 slots) (the rotations that fix the subring of
// X^(N / 2 slots); reference src/CKKS/Bootstrap.cu:116-143).
std::vector<int64_t> trace_rotations(const BootPlan& P) {
  std::vector<int64_t> out;
  if (!P.sparse) return out;
  for (int64_t j = 1; j * (int64_t)P.slots < (int64_t)P.half; j <<= 1) out.push_back(j * (int64_t)P.slots);
  return out;
}

// Message slots of the plan (0: N/2) and the slot period its groups work on.
void set_slots(BootPlan& P, uint32_t slots, uint32_t N) {
  const uint32_t ns = slots ? slots : N / 2;
  if (ns < 2 || ns > N / 2 || (ns & (ns - 1)) != 0)
    throw std::runtime_error("EvalBootstrapSetup: slots must be a power of two in [2, N/2]");
  P.half = N / 2;
  P.slots = ns;
  P.sparse = ns < N / 2;
  P.n = P.sparse ? 2 * ns : N / 2;
}

// Rotation indices (slot rotations mod n, without 0) of the groups' baby
// and giant steps, and of the partial trace.
std::vector<int32_t> plan_rotations(const BootPlan& P) {
  const int64_t n = P.n;
  std::vector<int32_t> rots;
  auto add = [&](const GroupPlan& g, const GroupPts& pts) {
    for (int64_t b : g.babies)
      if (b) rots.push_back((int32_t)(((b * (int64_t)g.stride) % n + n) % n));
    for (const auto& kv : pts)
      if (kv.first) rots.push_back((int32_t)((((kv.first * (int64_t)g.g * (int64_t)g.stride) % n) + n) % n));
  };
  for (size_t i = 0; i < P.cts.size(); ++i) add(P.cts[i], P.cts_pts[i]);
  for (size_t i = 0; i < P.stc.size(); ++i) add(P.stc[i], P.stc_pts[i]);
  for (int64_t r : trace_rotations(P)) rots.push_back((int32_t)r);
  std::sort(rots.begin(), rots.end());
  rots.erase(std::unique(rots.begin(), rots.end()), rots.end());
  rots.erase(std::remove(rots.begin(), rots.end(), 0), rots.end());
  return rots;
}

// Plaintexts without data (levels and scales), for a dry run.
void placeholder_pts(BootPlan& P, const std::vector<uint64_t>& q) {
  auto group = [&](const GroupPlan& G, uint32_t level) {
    GroupPts out;
    for (const auto& kv : G.giants)
      for (const auto& bv : kv.second) out[kv.first].emplace_back(bv.first, BootPlan::Pt{nullptr, level, (double)q[level - 1]});
    return out;
  };
  P.cts_pts.clear();
  P.stc_pts.clear();
  uint32_t l = P.l_raise;
  for (const GroupPlan& G : P.cts) P.cts_pts.push_back(group(G, l--));
  if (!P.sparse) P.minus_i = BootPlan::Pt{nullptr, l, 1.0};
  l -= P.evalmod_depth;
  if (!P.sparse) P.plus_i = BootPlan::Pt{nullptr, l, 1.0};
  for (const GroupPlan& G : P.stc) P.stc_pts.push_back(group(G, l--));
}

// The bootstrap program (fideslib_amd/bootstrap.py BootstrapPlan.program):
// prescale by 2^-c, one limb, [switch to the sparse secret], ModRaise,
// [switch back], [partial trace (sparse packing)], CoeffToSlot, Re/Im split
// and two EvalMods (full packing) or one EvalMod on [Re | Im] (sparse),
// recombine, SlotToCoeff, x 2^c.
// The bootstrap program in two parts: the input-dependent prescale down to
// one limb, then everything from the encapsulation and ModRaise on, which
// no longer depends on the input's level or scaling factor (ModRaise
// declares delta), so a batch of inputs shares it in lockstep.
Reg boot_prescale(Exec& X, const BootPlan& P, Reg x) {
  // prescale by 2^-c; under FLEXIBLEAUTO also by q_{l-1} / s_in, so that after
  // the Rescale the message sits at delta whatever the input level's factor
  // (the reference's adjustment factor, src/CKKS/Bootstrap.cu:441-470)
  double pre = std::ldexp(1.0, -(int)P.cfg.correction);
  if (P.flexible) pre = pre * (double)X.q[x.level - 1] / x.scale;
  x = X.mul_real(x, pre);
  return X.drop(x, 1);
}
Reg boot_core(Exec& X, const BootPlan& P, Reg x, const fides_key* down, const fides_key* up) {
  auto linear = [&](Reg v, const GroupPlan& G, const GroupPts& pts) {
    // rotation amounts modulo the plan's slot period (the data repeats with
    // it), the indices plan_rotations lists
    const int64_t n = (int64_t)G.n;
    auto amount = [n](int64_t a) { return ((a % n) + n) % n; };
    // every baby rotation in one batched call (one shared ModUp), every
    // giant's plaintext products in one, every giant rotation in one
    std::vector<std::pair<Reg, int64_t>> brot;
    for (int64_t b : G.babies) brot.emplace_back(v, b == 0 ? 0 : amount(b * (int64_t)G.stride));
    const std::vector<Reg> bregs = X.rotate_many(brot);
    std::map<int64_t, Reg> babies;
    for (size_t i = 0; i < G.babies.size(); ++i) babies[G.babies[i]] = bregs[i];
    std::vector<std::vector<std::pair<Reg, const BootPlan::Pt*>>> groups;
    std::vector<std::pair<Reg, int64_t>> grot;
    for (const auto& kv : pts) {
      std::vector<std::pair<Reg, const BootPlan::Pt*>> terms;
      for (const auto& bp : kv.second) terms.emplace_back(babies.at(bp.first), &bp.second);
      groups.push_back(std::move(terms));
    }
    const std::vector<Reg> inners = X.ptdot_many(groups);
    size_t gi = 0;
    for (const auto& kv : pts) {
      grot.emplace_back(inners[gi++], kv.first == 0 ? 0 : amount(kv.first * (int64_t)G.g * (int64_t)G.stride));
    }
    const std::vector<Reg> parts = X.rotate_many(grot);
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
  if (P.cfg.encapsulation) x = X.switch_key(x, down);
  Reg r = X.mod_raise(x, P.l_raise, X.delta);
  if (P.cfg.encapsulation) r = X.switch_key(r, up);
  for (int64_t a : trace_rotations(P)) r = X.add(r, X.rotate(r, a));
  for (size_t i = 0; i < P.cts.size(); ++i) r = linear(r, P.cts[i], P.cts_pts[i]);
  Reg cj = X.conjugate(r);
  Reg w;
  if (P.sparse) {  // [2 Re w | 2 Im w] in one ciphertext: one EvalMod
    w = evalmod(X.add(r, cj));
  } else {
    Reg re = X.add(r, cj);
    Reg d = X.sub(r, cj);
    Reg im = X.ptdot({{d, &P.minus_i}});
    re = evalmod(re);
    im = evalmod(im);
    w = X.add(re, X.ptdot({{im, &P.plus_i}}));
  }
  for (size_t i = 0; i < P.stc.size(); ++i) w = linear(w, P.stc[i], P.stc_pts[i]);
  const i128 corr = (i128)1 << P.cfg.correction;
  return X.lin(w, &corr, nullptr, false, nullptr, P.out_scale > 0 ? P.out_scale : w.scale);
}
Reg boot_run(Exec& X, const BootPlan& P, Reg x, const fides_key* down, const fides_key* up) {
  return boot_core(X, P, boot_prescale(X, P, x), down, up);
}

}  // namespace

// Plaintext upload in coefficient form ([level][N] residues) and preparation.
static BootPlan::Pt upload_pt(fides_context* ctx, uint32_t N, const std::vector<uint64_t>& host, uint32_t level,
                              double scale) {
  BootPlan::Pt pt;
  pt.level = level;
  pt.scale = scale;
  pt.buf = std::make_shared<DevBuf>((size_t)level * N);
  chk(fides_copy_to_device(pt.buf->d, host.data(), (size_t)level * N * 8, nullptr), "plaintext upload");
  chk(fides_plaintext_prepare(ctx, pt.buf->d, level, nullptr), "plaintext preparation");
  return pt;
}

static BootPlan::Pt monomial_pt(fides_context* ctx, uint32_t N, const std::vector<uint64_t>& q, uint64_t power,
                                uint32_t level) {
  std::vector<uint64_t> host((size_t)level * N, 0);
  uint64_t p = power % (2ull * N);
  const bool neg = p >= N;
  if (neg) p -= N;
  for (uint32_t j = 0; j < level; ++j) host[(size_t)j * N + p] = neg ? q[j] - 1 : 1;
  return upload_pt(ctx, N, host, level, 1.0);
}

std::vector<int32_t> CryptoContextImpl::EvalBootstrapSetup(const BootstrapConfig& cfg) {
  auto P = std::make_shared<BootPlan>();
  P->cfg = cfg;
  set_slots(*P, cfg.slots, N);
  const uint32_t n = P->n;
  const double q0 = (double)q[0];
  // CoeffToSlot output y = t / (q_0 K'), the 1/2 of the Re/Im split folded in
  // (and 2 slots / N, the factor of the partial trace, when sparse);
  // SlotToCoeff input sin(2 pi t / q_0) ~ 2 pi m / q_0 -> m / delta
  P->f_cts = boot_delta() / (2.0 * q0 * cfg.K);
  if (P->sparse) P->f_cts *= 2.0 * (double)P->slots / (double)N;
  P->f_stc = q0 / (2.0 * M_PI * boot_delta());
  auto cts = transform_groups(P->slots, N, cfg.budgetCtS, true);
  auto stc = transform_groups(P->slots, N, cfg.budgetStC, false);
  if (P->sparse) sparse_fold(cts, stc, P->slots);
  for (size_t i = 0; i < cts.size(); ++i) P->cts.push_back(plan_group(cts[i], n, i == 0 ? P->f_cts : 1.0));
  for (size_t i = 0; i < stc.size(); ++i) P->stc.push_back(plan_group(stc[i], n, i == 0 ? P->f_stc : 1.0));
  const double fa[2] = {cfg.K, std::ldexp(1.0, (int)cfg.doubleAngle)};
  P->cheb = chebyshev_coefficients(evalmod_fn, fa, cfg.degree);
  P->l_raise = L;
  P->evalmod_depth = chebyshev_eval_depth(P->cheb) + cfg.doubleAngle;
  P->depth = (uint32_t)(P->cts.size() + P->stc.size()) + P->evalmod_depth;
  if (P->l_raise <= P->depth) throw std::runtime_error("EvalBootstrapSetup: not enough levels for bootstrapping");
  P->flexible = Flexible();
  if (Flexible()) {
    // the nominal output scale (independent of the input: ModRaise declares
    // delta) from a dry run; fold the level's FLEXIBLEAUTO factor into it
    placeholder_pts(*P, q);
    Exec X{*this, ctx, q, N, boot_delta(), evalMultKey};
    X.dry = true;
    Reg x;
    x.level = L;
    x.scale = boot_delta();
    const Reg out = boot_run(X, *P, x, nullptr, nullptr);
    P->out_scale = sfReal.at(out.level - 1);
    P->f_stc *= P->out_scale / out.scale;
    P->stc.clear();
    for (size_t i = 0; i < stc.size(); ++i) P->stc.push_back(plan_group(stc[i], n, i == 0 ? P->f_stc : 1.0));
  }
  // plaintexts: encode v at `level` limbs with scale `scale`
  const uint32_t half = N / 2;
  const std::vector<uint64_t> slot_t = [&] {
    std::vector<uint64_t> t(half);
    uint64_t g = 1;
    for (uint32_t j = 0; j < half; ++j) {
      t[j] = (g - 1) / 2;
      g = g * 5 % (2ull * N);
    }
    return t;
  }();
  // z: a diagonal of period n (n | N/2), replicated over the N/2 slots
  auto encode = [&](const Vec& z, uint32_t level, double scale) {
    std::vector<cd> v(N);
    for (uint32_t j = 0; j < half; ++j) {
      const cd zj = z[j % z.size()];
      v[slot_t[j]] = zj;
      v[N - 1 - slot_t[j]] = std::conj(zj);
    }
    fft(v);
    std::vector<uint64_t> host((size_t)level * N);
    for (uint32_t i = 0; i < N; ++i) {
      const double a = -M_PI * (double)i / (double)N;
      const cd m = v[i] / (double)N * cd(std::cos(a), std::sin(a));
      const long double c = std::nearbyint((long double)(m.real() * scale));
      for (uint32_t j = 0; j < level; ++j) {
        long double r = std::fmod(c, (long double)q[j]);
        if (r < 0) r += (long double)q[j];
        host[(size_t)j * N + i] = (uint64_t)r;
      }
    }
    return upload_pt(ctx, N, host, level, scale);
  };
  auto group_pts = [&](const GroupPlan& G, uint32_t level) {
    GroupPts out;
    const double scale = (double)q[level - 1];
    for (const auto& kv : G.giants)
      for (const auto& bv : kv.second) out[kv.first].emplace_back(bv.first, encode(bv.second, level, scale));
    return out;
  };
  P->cts_pts.clear();
  P->stc_pts.clear();
  uint32_t l = P->l_raise;
  for (const GroupPlan& G : P->cts) P->cts_pts.push_back(group_pts(G, l--));
  if (!P->sparse) P->minus_i = monomial_pt(ctx, N, q, 3ull * N / 2, l);
  l -= P->evalmod_depth;
  if (!P->sparse) P->plus_i = monomial_pt(ctx, N, q, N / 2, l);
  for (const GroupPlan& G : P->stc) P->stc_pts.push_back(group_pts(G, l--));
  P->rotations = plan_rotations(*P);
  boot = P;
  return P->rotations;
}

std::vector<int32_t> CryptoContextImpl::EvalBootstrapSetup(const BootstrapPrecomputation& pre) {
  auto P = std::make_shared<BootPlan>();
  P->cfg = pre.cfg;
  P->cfg.budgetCtS = (uint32_t)pre.coeffToSlot.size();
  P->cfg.budgetStC = (uint32_t)pre.slotToCoeff.size();
  set_slots(*P, pre.cfg.slots, N);
  P->cheb = pre.chebyshev;
  if (P->cheb.size() < 2) throw std::runtime_error("EvalBootstrapSetup: the EvalMod series needs degree >= 1");
  P->l_raise = L;
  P->evalmod_depth = chebyshev_eval_depth(P->cheb) + P->cfg.doubleAngle;
  P->depth = P->cfg.budgetCtS + P->cfg.budgetStC + P->evalmod_depth;
  if (P->l_raise <= P->depth) throw std::runtime_error("EvalBootstrapSetup: not enough levels for bootstrapping");
  P->out_scale = pre.outScale;
  P->flexible = pre.outScale > 0;
  auto load = [&](const std::vector<BootstrapPrecomputation::Group>& groups, uint32_t level0,
                  std::vector<GroupPlan>& plans, std::vector<GroupPts>& pts) {
    uint32_t level = level0;
    for (const auto& G : groups) {
      GroupPlan g;
      g.stride = (uint32_t)G.stride;
      g.g = (uint32_t)G.g;
      g.babies = G.babies;
      g.n = P->n;
      GroupPts gp;
      for (const auto& t : G.terms) {
        if (t.level != level || t.coef.size() != (size_t)level * N)
          throw std::runtime_error("EvalBootstrapSetup: plaintext level or size does not match its group");
        for (uint32_t j = 0; j < level; ++j)
          for (uint32_t i = 0; i < N; ++i)
            if (t.coef[(size_t)j * N + i] >= q[j]) throw std::runtime_error("EvalBootstrapSetup: residue not reduced");
        gp[t.giant].emplace_back(t.baby, upload_pt(ctx, N, t.coef, level, t.scale));
      }
      plans.push_back(g);
      pts.push_back(std::move(gp));
      --level;
    }
  };
  load(pre.coeffToSlot, P->l_raise, P->cts, P->cts_pts);
  const uint32_t l_split = P->l_raise - P->cfg.budgetCtS;
  if (!P->sparse) P->minus_i = monomial_pt(ctx, N, q, 3ull * N / 2, l_split);
  const uint32_t l_stc = l_split - P->evalmod_depth;
  if (!P->sparse) P->plus_i = monomial_pt(ctx, N, q, N / 2, l_stc);
  load(pre.slotToCoeff, l_stc, P->stc, P->stc_pts);
  P->rotations = plan_rotations(*P);
  boot = P;
  return P->rotations;
}

// numIterations > 1 (OpenFHE's iterated bootstrapping; the reference's GPU
// path runs one iteration, api/CryptoContext.cpp:1719-1745): with
// bt = EvalBootstrap(x, numIterations - 1), the error 2^p (x - bt) is formed
// at the lower of the two levels (FIXEDAUTO/FLEXIBLEAUTO: EvalSub's
// alignment; FIXEDMANUAL: both cut to the fewer limbs), bootstrapped, and added to
// 2^p bt; the sum is 2^p x up to the second bootstrap's error, returned with
// the scaling factor s_out 2^p (exact, no level spent on a division).  p is
// `precision`, the bits one bootstrap keeps (2^p (x - bt) must stay below 1).
Ciphertext CryptoContextImpl::EvalBootstrap(const Ciphertext& ct0, uint32_t numIterations, uint32_t precision) {
  if (numIterations <= 1) return BootstrapOnce(ct0);
  if (precision == 0 || precision > 50)
    throw std::runtime_error("EvalBootstrap: numIterations > 1 needs the precision of one bootstrap (1..50 bits)");
  Ciphertext x = Regular(ct0);
  if (x->NoiseLevel == 2) x = Rescale(x);
  const Ciphertext bt = EvalBootstrap(x, numIterations - 1, precision);
  auto pow2 = [&](uint32_t limbs) {
    std::vector<uint64_t> r(limbs);
    for (uint32_t i = 0; i < limbs; ++i) r[i] = detail::powmod(2, precision, q[i]);
    return r;
  };
  Ciphertext e;
  if (technique == FIXEDMANUAL) {
    const uint32_t l = std::min(x->limbs, bt->limbs);
    e = EvalSub(DropToLimbs(x, l), DropToLimbs(bt, l));
  } else {
    e = EvalSub(x, bt);
  }
  e = MulConst(e, pow2(e->limbs));
  const Ciphertext be = BootstrapOnce(e);
  Ciphertext r = EvalAdd(MulConst(bt, pow2(bt->limbs)), be);
  r->NoiseFactor = bt->NoiseFactor * std::ldexp(1.0, (int)precision);
  return r;
}

Ciphertext CryptoContextImpl::BootstrapOnce(const Ciphertext& ct0) { return EvalBootstrapMany({ct0}).at(0); }

// Each input is prescaled to one limb on its own (a batch of one: the
// constant depends on its level and scaling factor); the rest of the
// program then runs once for all inputs, every step one batched call.
std::vector<Ciphertext> CryptoContextImpl::EvalBootstrapMany(const std::vector<Ciphertext>& cts0) {
  if (!boot) throw std::runtime_error("EvalBootstrap: call EvalBootstrapSetup first");
  if (!evalMultKey) throw std::runtime_error("EvalBootstrap: relinearization key not loaded");
  const BootPlan& P = *boot;
  if (P.cfg.encapsulation && (!encapDown || !encapUp))
    throw std::runtime_error("EvalBootstrap: sparse-secret encapsulation keys not loaded (AddEncapsulationKeys)");
  const uint32_t n = (uint32_t)cts0.size();
  if (n == 0) return {};
  std::vector<Ciphertext> cts;
  for (const Ciphertext& c0 : cts0) {
    Ciphertext ct = Regular(c0);
    if (ct->NoiseLevel == 2) ct = Rescale(ct);  // as the reference's ModRaise (Bootstrap.cu:423)
    if (ct->limbs < 2) throw std::runtime_error("EvalBootstrap: the input needs at least 2 limbs");
    cts.push_back(ct);
  }
  Exec X{*this, ctx, q, N, boot_delta(), evalMultKey};
  X.n = n;
  // the register pool outlives the call, so repeated bootstraps of one batch
  // size allocate no device memory after the first
  if (!bootMem) bootMem = std::make_shared<Pool>();
  X.pool = std::static_pointer_cast<Pool>(bootMem);
  Reg x1 = X.fresh(1, 0.0);  // one limb per input
  // inputs of one level and scaling factor share their prescale (a batch)
  std::map<std::pair<uint32_t, double>, std::vector<uint32_t>> groups;
  for (uint32_t i = 0; i < n; ++i) {
    const double sc = cts[i]->NoiseFactor > 0 ? cts[i]->NoiseFactor : boot_delta();
    groups[{cts[i]->limbs, sc}].push_back(i);
  }
  for (const auto& g : groups) {
    const std::vector<uint32_t>& idx = g.second;
    Exec Xg{*this, ctx, q, N, boot_delta(), evalMultKey};
    Xg.n = (uint32_t)idx.size();
    Xg.pool = X.pool;
    // the inputs as a register (its nominal scale is the tracked scaling factor)
    Reg x = Xg.fresh(g.first.first, g.first.second);
    for (uint32_t j = 0; j < Xg.n; ++j)
      chk(fides_copy_device(x.d(j), cts[idx[j]]->DeviceData(), (size_t)2 * g.first.first * N * 8, nullptr), "copy");
    const Reg p = boot_prescale(Xg, P, x);
    for (uint32_t j = 0; j < Xg.n; ++j)
      chk(fides_copy_device(x1.d(idx[j]), p.d(j), (size_t)2 * N * 8, nullptr), "copy");
    x1.scale = p.scale;
  }
  const Reg out = boot_core(X, P, x1, encapDown, encapUp);
  std::vector<Ciphertext> res;
  for (uint32_t i = 0; i < n; ++i) {
    Ciphertext o = NewCiphertext(out.level);
    chk(fides_copy_device(o->DeviceData(), out.d(i), (size_t)2 * out.level * N * 8, nullptr), "copy");
    o->NoiseLevel = 1;
    o->NoiseFactor = out.scale;
    o->slots = cts[i]->slots;
    o->keyid = cts[i]->keyid;
    res.push_back(o);
  }
  chk(fides_stream_sync(nullptr), "sync");
  return res;
}

}  // namespace fideslib
