// Host side of the C ABI: context tables, key storage and the key-switch
// pipelines that sequence the HIP kernels.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fideslib_amd.h"
#include "kernels.h"

typedef unsigned __int128 u128;

// ======================================================================
// errors
// ======================================================================
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess)                                                                \
      return fail(FIDES_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));       \
  } while (0)

extern "C" const char* fides_last_error(void) { return g_err.c_str(); }

// ======================================================================
// host number theory (independent of any test code)
// ======================================================================
static inline u64 mulm(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }
static u64 powm(u64 a, u64 e, u64 m) {
  u64 r = 1 % m;
  a %= m;
  while (e) {
    if (e & 1) r = mulm(r, a, m);
    a = mulm(a, a, m);
    e >>= 1;
  }
  return r;
}
static u64 invm(u64 a, u64 m) { return powm(a % m, m - 2, m); }  // m prime
static u32 brev32(u32 x, u32 bits) {
  u32 r = 0;
  for (u32 i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}
static bool is_prime64(u64 n) {
  if (n < 2) return false;
  static const u64 bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (u64 p : bases)
    if (n % p == 0) return n == p;
  u64 d = n - 1;
  int s = 0;
  while (!(d & 1)) { d >>= 1; ++s; }
  for (u64 a : bases) {
    u64 x = powm(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int i = 1; i < s; ++i) {
      x = mulm(x, x, n);
      if (x == n - 1) { comp = false; break; }
    }
    if (comp) return false;
  }
  return true;
}
static u64 gcd64(u64 a, u64 b) {
  while (b) { u64 t = a % b; a = b; b = t; }
  return a;
}
// Pollard-Brent rho: a non-trivial factor of composite n.
static u64 rho(u64 n) {
  if (n % 2 == 0) return 2;
  for (u64 c = 1;; ++c) {
    u64 y = 2, x = 2, g = 1, q = 1, ys = 2;
    u64 r = 1, m = 128;
    auto f = [&](u64 v) { return (mulm(v, v, n) + c) % n; };
    do {
      x = y;
      for (u64 i = 0; i < r; ++i) y = f(y);
      u64 k = 0;
      do {
        ys = y;
        for (u64 i = 0; i < std::min(m, r - k); ++i) {
          y = f(y);
          q = mulm(q, x > y ? x - y : y - x, n);
        }
        g = gcd64(q, n);
        k += m;
      } while (k < r && g == 1);
      r <<= 1;
    } while (g == 1);
    if (g == n) {
      do {
        ys = f(ys);
        g = gcd64(x > ys ? x - ys : ys - x, n);
      } while (g == 1);
    }
    if (g != n) return g;
  }
}
static void factor(u64 n, std::vector<u64>& out) {
  if (n == 1) return;
  for (u64 p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull}) {
    if (n % p == 0) {
      out.push_back(p);
      while (n % p == 0) n /= p;
    }
  }
  if (n == 1) return;
  if (is_prime64(n)) { out.push_back(n); return; }
  u64 d = rho(n);
  std::vector<u64> a, b;
  factor(d, a);
  factor(n / d, b);
  for (u64 x : a) out.push_back(x);
  for (u64 x : b) out.push_back(x);
}
// Smallest primitive root of prime q, then psi = g^((q-1)/2N)
// (same rule as the reference fallback, src/ConstantsGPU.cu:36-67).
static u64 default_root(u64 q, u32 logN) {
  std::vector<u64> f;
  factor(q - 1, f);
  std::sort(f.begin(), f.end());
  f.erase(std::unique(f.begin(), f.end()), f.end());
  for (u64 g = 2; g < q; ++g) {
    bool ok = true;
    for (u64 p : f)
      if (powm(g, (q - 1) / p, q) == 1) { ok = false; break; }
    if (ok) return powm(g, (q - 1) >> (logN + 1), q);
  }
  return 0;
}
static u64 shoup_of(u64 w, u64 q) { return (u64)(((u128)w << 64) / q); }
static u32 bitlen(u64 x) { return 64 - __builtin_clzll(x); }

// ======================================================================
// context
// ======================================================================
struct fides_key {
  fides_context* ctx;
  u64* d;  // [2][dnum][L+K][N] Montgomery form
};

struct fides_context {
  u32 logN, N, L, K, dnum, alpha;
  std::vector<u64> primes, roots;
  std::vector<PrimeC> hprm;
  // device tables
  PrimeC* prm = nullptr;
  u64* tw = nullptr;  // psi | psi_sh | ipsi | ipsi_sh, each [(L+K) N]
  NttTables T{};
  u64* pinv = nullptr;     // [L] P^-1 mod q_i, and Shoup at [L..2L)
  u64* resc = nullptr;     // [(L+1)][4][L]: qlinv, qlinv_sh, QlQlInv, QlQlInv_sh rows per level l
  ConvDesc* modup = nullptr;    // [(L+1) * dnum]
  ConvDesc* moddown = nullptr;  // [(L+1)]
  std::vector<ConvDesc> hmodup, hmoddown;
  LimbTask* ext_tasks = nullptr;  // per level: NTT task list of the ModUp outputs
  std::vector<u32> ext_task_off, ext_task_cnt;
  LimbTask* tmp_tasks = nullptr;  // scratch task list for fides_ntt with many segments
  u32 tmp_tasks_cap = 0;
  // key-switch scratch (sized for l = L)
  u64 *Dc = nullptr, *EXT = nullptr, *T0 = nullptr, *T1 = nullptr, *Y = nullptr, *D = nullptr,
      *TMP = nullptr, *RS = nullptr;
  std::vector<void*> allocs;

  u64* limb(u64* base, u64 i) const { return base + (i << logN); }
  const u64* limb(const u64* base, u64 i) const { return base + (i << logN); }
};

static int dalloc(fides_context* c, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(FIDES_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  c->allocs.push_back(*p);
  return FIDES_OK;
}

static LimbSpan span0() {
  LimbSpan s;
  memset(&s, 0, sizeof(s));
  return s;
}
static void span_add(LimbSpan& s, u64* ptr, u32 pid, u32 n) {
  if (n == 0) return;
  s.seg_ptr[s.nseg] = ptr;
  s.seg_pid[s.nseg] = pid;
  s.total += n;
  s.seg_end[s.nseg] = s.total;
  s.nseg++;
}

// Conversion descriptor for source primes src -> targets dst (dst_off given).
static void make_conv(const fides_context* c, const std::vector<u32>& src, const std::vector<u32>& dst,
                      const std::vector<u32>& off, ConvDesc& d) {
  memset(&d, 0, sizeof(d));
  d.ns = (u32)src.size();
  d.nt = (u32)dst.size();
  for (u32 i = 0; i < d.ns; ++i) {
    const u64 qi = c->primes[src[i]];
    u64 prod = 1;
    for (u32 j = 0; j < d.ns; ++j)
      if (j != i) prod = mulm(prod, c->primes[src[j]] % qi, qi);
    d.src_pid[i] = src[i];
    d.hatinv[i] = invm(prod, qi);
    d.hatinv_sh[i] = shoup_of(d.hatinv[i], qi);
  }
  for (u32 t = 0; t < d.nt; ++t) {
    const u64 pt = c->primes[dst[t]];
    d.dst_pid[t] = dst[t];
    d.dst_off[t] = off[t];
    const u64 R = (u64)(((u128)1 << 64) % pt);
    for (u32 i = 0; i < d.ns; ++i) {
      u64 prod = 1;
      for (u32 j = 0; j < d.ns; ++j)
        if (j != i) prod = mulm(prod, c->primes[src[j]] % pt, pt);
      d.mat[t][i] = mulm(prod, R, pt);
    }
  }
}

static void digit_range(const fides_context* c, u32 l, u32 j, u32& st, u32& en) {
  st = j * c->alpha;
  en = std::min((j + 1) * c->alpha, l);
}
static u32 num_digits(const fides_context* c, u32 l) { return (l + c->alpha - 1) / c->alpha; }

extern "C" int fides_context_create(const fides_params* P, fides_context** out) {
  if (!P || !out) return fail(FIDES_ERR_ARG, "null argument");
  *out = nullptr;
  if (P->logN < 10 || P->logN > 17) return fail(FIDES_ERR_UNSUPPORTED, "logN must be in [10, 17]");
  if (P->L == 0 || P->L > 64 || P->K == 0 || P->K > CONV_MAXS || P->dnum == 0)
    return fail(FIDES_ERR_ARG, "need 1 <= L <= 64, 1 <= K <= 16, dnum >= 1");
  const u32 alpha = (P->L + P->dnum - 1) / P->dnum;
  if (alpha > CONV_MAXS) return fail(FIDES_ERR_UNSUPPORTED, "digit size alpha = ceil(L/dnum) must be <= 16");
  if (P->L + P->K > CONV_MAXT) return fail(FIDES_ERR_UNSUPPORTED, "L + K too large");
  const u64 twoN = 2ull << P->logN;
  std::vector<u64> pr;
  for (u32 i = 0; i < P->L; ++i) pr.push_back(P->q[i]);
  for (u32 i = 0; i < P->K; ++i) pr.push_back(P->p[i]);
  for (size_t i = 0; i < pr.size(); ++i) {
    if (pr[i] >= (1ull << 62) || pr[i] % twoN != 1 || !is_prime64(pr[i]))
      return fail(FIDES_ERR_ARG, "prime " + std::to_string(i) + " is not a prime < 2^62 with q = 1 mod 2N");
    for (size_t j = 0; j < i; ++j)
      if (pr[j] == pr[i]) return fail(FIDES_ERR_ARG, "duplicate prime");
  }
  fides_context* c = new fides_context();
  c->logN = P->logN;
  c->N = 1u << P->logN;
  c->L = P->L;
  c->K = P->K;
  c->dnum = P->dnum;
  c->alpha = alpha;
  c->primes = pr;
  const u32 NP = P->L + P->K, N = c->N;
  c->roots.resize(NP);
  for (u32 i = 0; i < NP; ++i) {
    u64 psi = P->psi ? P->psi[i] : default_root(pr[i], P->logN);
    if (powm(psi, N, pr[i]) != pr[i] - 1) {
      delete c;
      return fail(FIDES_ERR_ARG, "psi[" + std::to_string(i) + "] is not a primitive 2N-th root of unity");
    }
    c->roots[i] = psi;
  }
  // ---- per-prime constants and twiddle tables
  std::vector<u64> tw((size_t)4 * NP * N);
  c->hprm.resize(NP);
  auto work = [&](u32 i) {
    const u64 q = pr[i];
    PrimeC& p = c->hprm[i];
    memset(&p, 0, sizeof(p));
    p.q = q;
    p.q2 = 2 * q;
    u64 inv = q;  // Newton: q^-1 mod 2^64
    for (int it = 0; it < 6; ++it) inv *= 2 - q * inv;
    p.qneg_inv = (u64)0 - inv;
    p.bits = bitlen(q);
    p.mu = (u64)(((u128)1 << (2 * p.bits)) / q);
    const u64 r1 = (u64)(((u128)1 << 64) % q);
    p.r2 = mulm(r1, r1, q);
    const u64 ipsi = invm(c->roots[i], q);
    p.ninv = invm(N % q, q);
    p.ninv_sh = shoup_of(p.ninv, q);
    // psi_rev[1] = psi^brev(1) = psi^(N/2); inverse table entry ipsi^(N/2)
    const u64 w1 = powm(ipsi, N / 2, q);
    p.ninv_w = mulm(w1, p.ninv, q);
    p.ninv_w_sh = shoup_of(p.ninv_w, q);
    u64* t_psi = tw.data() + ((size_t)0 * NP + i) * N;
    u64* t_psish = tw.data() + ((size_t)1 * NP + i) * N;
    u64* t_ipsi = tw.data() + ((size_t)2 * NP + i) * N;
    u64* t_ipsish = tw.data() + ((size_t)3 * NP + i) * N;
    std::vector<u64> pw(N), ipw(N);
    pw[0] = ipw[0] = 1;
    for (u32 j = 1; j < N; ++j) {
      pw[j] = mulm(pw[j - 1], c->roots[i], q);
      ipw[j] = mulm(ipw[j - 1], ipsi, q);
    }
    for (u32 j = 0; j < N; ++j) {
      const u32 r = brev32(j, P->logN);
      t_psi[j] = pw[r];
      t_psish[j] = shoup_of(pw[r], q);
      t_ipsi[j] = ipw[r];
      t_ipsish[j] = shoup_of(ipw[r], q);
    }
  };
  {
    std::vector<std::thread> th;
    const u32 nth = std::max(1u, std::min(NP, std::thread::hardware_concurrency()));
    for (u32 w = 0; w < nth; ++w)
      th.emplace_back([&, w] {
        for (u32 i = w; i < NP; i += nth) work(i);
      });
    for (auto& t : th) t.join();
  }
  int rc;
#define ALLOC(ptr, bytes)                                          \
  do {                                                             \
    if ((rc = dalloc(c, (void**)&(ptr), (bytes))) != FIDES_OK) {   \
      fides_context_destroy(c);                                    \
      return rc;                                                   \
    }                                                              \
  } while (0)
#define H2D(dst, src, bytes)                                                         \
  do {                                                                               \
    hipError_t e_ = hipMemcpy((dst), (src), (bytes), hipMemcpyHostToDevice);         \
    if (e_ != hipSuccess) {                                                          \
      fides_context_destroy(c);                                                      \
      return fail(FIDES_ERR_HIP, std::string("upload: ") + hipGetErrorString(e_));   \
    }                                                                                \
  } while (0)
  ALLOC(c->prm, sizeof(PrimeC) * NP);
  H2D(c->prm, c->hprm.data(), sizeof(PrimeC) * NP);
  ALLOC(c->tw, tw.size() * 8);
  H2D(c->tw, tw.data(), tw.size() * 8);
  c->T.psi = c->tw;
  c->T.psi_sh = c->tw + (size_t)1 * NP * N;
  c->T.ipsi = c->tw + (size_t)2 * NP * N;
  c->T.ipsi_sh = c->tw + (size_t)3 * NP * N;
  c->T.prm = c->prm;
  // ---- ModDown / rescale scalars
  const u32 L = c->L, K = c->K;
  {
    std::vector<u64> pinv(2 * L);
    for (u32 i = 0; i < L; ++i) {
      const u64 qi = pr[i];
      u64 Pm = 1;
      for (u32 k = 0; k < K; ++k) Pm = mulm(Pm, pr[L + k] % qi, qi);
      pinv[i] = invm(Pm, qi);
      pinv[L + i] = shoup_of(pinv[i], qi);
    }
    ALLOC(c->pinv, pinv.size() * 8);
    H2D(c->pinv, pinv.data(), pinv.size() * 8);
    // rescale rows for l = 2..L: dropping q_{l-1}
    std::vector<u64> rs((size_t)(L + 1) * 4 * L, 0);
    for (u32 l = 2; l <= L; ++l) {
      const u64 qt = pr[l - 1];
      for (u32 i = 0; i + 1 < l; ++i) {
        const u64 qi = pr[i];
        const u64 qtinv = invm(qt % qi, qi);
        // QlQlInvModqlDivqlModq_i = floor(Q' [Q'^-1]_qt / qt) mod q_i = -qt^-1 mod q_i
        // (Q' = prod_{i<l-1} q_i; Q' [Q'^-1]_qt = 1 mod qt and = 0 mod q_i)
        const u64 big = (qi - qtinv) % qi;
        u64* row = rs.data() + (size_t)l * 4 * L;
        row[0 * L + i] = qtinv;
        row[1 * L + i] = shoup_of(qtinv, qi);
        row[2 * L + i] = big;
        row[3 * L + i] = shoup_of(big, qi);
      }
    }
    ALLOC(c->resc, rs.size() * 8);
    H2D(c->resc, rs.data(), rs.size() * 8);
  }
  // ---- conversion descriptors
  c->hmodup.resize((size_t)(L + 1) * c->dnum);
  c->hmoddown.resize(L + 1);
  std::vector<LimbTask> tasks;
  c->ext_task_off.assign(L + 1, 0);
  c->ext_task_cnt.assign(L + 1, 0);
  {
    // scratch first (task lists point into EXT)
    ALLOC(c->Dc, (size_t)L * N * 8);
    ALLOC(c->EXT, (size_t)c->dnum * (L + K) * N * 8);
    ALLOC(c->T0, (size_t)(L + K) * N * 8);
    ALLOC(c->T1, (size_t)(L + K) * N * 8);
    ALLOC(c->Y, (size_t)2 * L * N * 8);
    ALLOC(c->D, (size_t)L * N * 8);
    ALLOC(c->TMP, (size_t)2 * L * N * 8);
    ALLOC(c->RS, (size_t)2 * N * 8);
  }
  for (u32 l = 1; l <= L; ++l) {
    const u32 beta = num_digits(c, l);
    c->ext_task_off[l] = (u32)tasks.size();
    for (u32 j = 0; j < beta; ++j) {
      u32 st, en;
      digit_range(c, l, j, st, en);
      std::vector<u32> src, dst, off;
      for (u32 i = st; i < en; ++i) src.push_back(i);
      for (u32 i = 0; i < l; ++i)
        if (i < st || i >= en) { dst.push_back(i); off.push_back(i); }
      for (u32 k = 0; k < K; ++k) { dst.push_back(L + k); off.push_back(l + k); }
      make_conv(c, src, dst, off, c->hmodup[(size_t)l * c->dnum + j]);
      u64* ext = c->EXT + (size_t)j * (L + K) * N;
      for (size_t t = 0; t < dst.size(); ++t) tasks.push_back(LimbTask{ext + (size_t)off[t] * N, dst[t], 0});
    }
    c->ext_task_cnt[l] = (u32)tasks.size() - c->ext_task_off[l];
    std::vector<u32> src, dst, off;
    for (u32 k = 0; k < K; ++k) src.push_back(L + k);
    for (u32 i = 0; i < l; ++i) { dst.push_back(i); off.push_back(i); }
    make_conv(c, src, dst, off, c->hmoddown[l]);
  }
  ALLOC(c->modup, sizeof(ConvDesc) * c->hmodup.size());
  H2D(c->modup, c->hmodup.data(), sizeof(ConvDesc) * c->hmodup.size());
  ALLOC(c->moddown, sizeof(ConvDesc) * c->hmoddown.size());
  H2D(c->moddown, c->hmoddown.data(), sizeof(ConvDesc) * c->hmoddown.size());
  ALLOC(c->ext_tasks, sizeof(LimbTask) * tasks.size());
  H2D(c->ext_tasks, tasks.data(), sizeof(LimbTask) * tasks.size());
  c->tmp_tasks_cap = 4096;
  ALLOC(c->tmp_tasks, sizeof(LimbTask) * c->tmp_tasks_cap);
#undef ALLOC
#undef H2D
  *out = c;
  return FIDES_OK;
}

extern "C" void fides_context_destroy(fides_context* c) {
  if (!c) return;
  for (void* p : c->allocs) (void)hipFree(p);
  delete c;
}

extern "C" int fides_context_info(const fides_context* c, uint32_t* logN, uint32_t* L, uint32_t* K, uint32_t* dnum,
                                  uint32_t* alpha) {
  if (!c) return fail(FIDES_ERR_ARG, "null context");
  if (logN) *logN = c->logN;
  if (L) *L = c->L;
  if (K) *K = c->K;
  if (dnum) *dnum = c->dnum;
  if (alpha) *alpha = c->alpha;
  return FIDES_OK;
}

extern "C" int fides_context_roots(const fides_context* c, uint64_t* psi) {
  if (!c || !psi) return fail(FIDES_ERR_ARG, "null argument");
  for (size_t i = 0; i < c->roots.size(); ++i) psi[i] = c->roots[i];
  return FIDES_OK;
}

// ======================================================================
// memory helpers
// ======================================================================
extern "C" int fides_device_alloc(size_t bytes, void** p) {
  if (!p) return fail(FIDES_ERR_ARG, "null argument");
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(FIDES_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  return FIDES_OK;
}
extern "C" int fides_device_free(void* p) {
  HIPCHK(hipFree(p));
  return FIDES_OK;
}
extern "C" int fides_copy_to_device(void* dst, const void* src, size_t bytes, void* stream) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return FIDES_OK;
}
extern "C" int fides_copy_to_host(void* dst, const void* src, size_t bytes, void* stream) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return FIDES_OK;
}
extern "C" int fides_stream_sync(void* stream) {
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return FIDES_OK;
}

// ======================================================================
// raw limb operations
// ======================================================================
static int build_span(fides_context* c, u64* data, const uint32_t* pids, uint32_t n, hipStream_t st,
                      LimbSpan& s) {
  s = span0();
  const u32 NP = c->L + c->K;
  for (u32 t = 0; t < n; ++t)
    if (pids[t] >= NP) return fail(FIDES_ERR_ARG, "prime id out of range");
  // group consecutive prime ids into segments
  std::vector<std::pair<u32, u32>> segs;  // (start limb, count)
  for (u32 t = 0; t < n; ++t) {
    if (!segs.empty() && pids[t] == pids[segs.back().first] + segs.back().second) {
      segs.back().second++;
    } else {
      segs.push_back({t, 1});
    }
  }
  if (segs.size() <= 8) {
    for (auto& sg : segs) span_add(s, c->limb(data, sg.first), pids[sg.first], sg.second);
    return FIDES_OK;
  }
  if (n > c->tmp_tasks_cap) return fail(FIDES_ERR_ARG, "too many limbs in one call");
  std::vector<LimbTask> tasks(n);
  for (u32 t = 0; t < n; ++t) tasks[t] = LimbTask{c->limb(data, t), pids[t], t};
  HIPCHK(hipMemcpyAsync(c->tmp_tasks, tasks.data(), n * sizeof(LimbTask), hipMemcpyHostToDevice, st));
  s.list = c->tmp_tasks;
  s.total = n;
  return FIDES_OK;
}

extern "C" int fides_ntt(fides_context* c, uint64_t* d, const uint32_t* pids, uint32_t n, void* stream) {
  if (!c || (!d && n) || (!pids && n)) return fail(FIDES_ERR_ARG, "null argument");
  if (n == 0) return FIDES_OK;
  LimbSpan s;
  int rc = build_span(c, d, pids, n, (hipStream_t)stream, s);
  if (rc) return rc;
  NttFuse F;
  memset(&F, 0, sizeof(F));
  if (fides_launch_ntt(c->logN, s, c->T, PRO_NONE, EPI_NONE, F, (hipStream_t)stream))
    return fail(FIDES_ERR_UNSUPPORTED, "ring dimension");
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_intt(fides_context* c, uint64_t* d, const uint32_t* pids, uint32_t n, void* stream) {
  if (!c || (!d && n) || (!pids && n)) return fail(FIDES_ERR_ARG, "null argument");
  if (n == 0) return FIDES_OK;
  LimbSpan s;
  int rc = build_span(c, d, pids, n, (hipStream_t)stream, s);
  if (rc) return rc;
  if (fides_launch_intt(c->logN, s, c->T, (hipStream_t)stream)) return fail(FIDES_ERR_UNSUPPORTED, "ring dimension");
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// ModUp of digit j into ext (l+K limbs, Q_l u P order), target limbs only
// (the digit's own limbs are not written).  Input d: l limbs, evaluation.
static int modup_digit(fides_context* c, const u64* d, u32 l, u32 j, u64* ext, bool intt_done, hipStream_t st) {
  const u32 N = c->N;
  u32 stt, en;
  digit_range(c, l, j, stt, en);
  const u32 sz = en - stt;
  if (!intt_done) {
    HIPCHK(hipMemcpyAsync(c->limb(c->Dc, stt), c->limb(d, stt), (size_t)sz * N * 8, hipMemcpyDeviceToDevice, st));
    LimbSpan s = span0();
    span_add(s, c->limb(c->Dc, stt), stt, sz);
    fides_launch_intt(c->logN, s, c->T, st);
  }
  const ConvDesc& hd = c->hmodup[(size_t)l * c->dnum + j];
  if (fides_launch_conv(c->limb(c->Dc, stt), ext, c->modup + (size_t)l * c->dnum + j, hd.ns, hd.nt, c->prm,
                        c->logN, st))
    return fail(FIDES_ERR_UNSUPPORTED, "digit size");
  return FIDES_OK;
}

extern "C" int fides_modup(fides_context* c, const uint64_t* d, uint32_t l, uint32_t j, uint64_t* out,
                           void* stream) {
  if (!c || !d || !out) return fail(FIDES_ERR_ARG, "null argument");
  if (l == 0 || l > c->L || j >= num_digits(c, l)) return fail(FIDES_ERR_ARG, "level / digit out of range");
  hipStream_t st = (hipStream_t)stream;
  int rc = modup_digit(c, d, l, j, out, false, st);
  if (rc) return rc;
  u32 stt, en;
  digit_range(c, l, j, stt, en);
  LimbSpan s = span0();
  span_add(s, out, 0, stt);
  span_add(s, c->limb(out, en), en, l - en);
  span_add(s, c->limb(out, l), c->L, c->K);
  NttFuse F;
  memset(&F, 0, sizeof(F));
  fides_launch_ntt(c->logN, s, c->T, PRO_NONE, EPI_NONE, F, st);
  HIPCHK(hipMemcpyAsync(c->limb(out, stt), c->limb(d, stt), (size_t)(en - stt) * c->N * 8, hipMemcpyDeviceToDevice,
                        st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// ModDown of npoly polynomials src[p] (each l+K limbs, evaluation form) into
// dst[p] (l limbs); with acc[p] != null the result is added to acc[p]
// (dst[p] may alias acc[p]).  src special limbs are clobbered.
static int moddown_multi(fides_context* c, u64* const* src, u32 npoly, u32 l, u64* const* dst, const u64* const* acc,
                         hipStream_t st) {
  const u32 N = c->N, K = c->K;
  LimbSpan sp = span0();
  for (u32 p = 0; p < npoly; ++p) span_add(sp, c->limb(src[p], l), c->L, K);
  fides_launch_intt(c->logN, sp, c->T, st);
  const ConvDesc& hd = c->hmoddown[l];
  for (u32 p = 0; p < npoly; ++p)
    fides_launch_conv(c->limb(src[p], l), c->Y + (size_t)p * c->L * N, c->moddown + l, hd.ns, hd.nt, c->prm, c->logN,
                      st);
  LimbSpan ys = span0();
  NttFuse F;
  memset(&F, 0, sizeof(F));
  F.x = span0();
  F.out = span0();
  F.acc = span0();
  bool any_acc = false;
  for (u32 p = 0; p < npoly; ++p) {
    span_add(ys, c->Y + (size_t)p * c->L * N, 0, l);
    span_add(F.x, src[p], 0, l);
    span_add(F.out, dst[p], 0, l);
    span_add(F.acc, (u64*)(acc ? acc[p] : nullptr), 0, l);
    if (acc && acc[p]) any_acc = true;
  }
  F.c0 = c->pinv;
  F.c0sh = c->pinv + c->L;
  fides_launch_ntt(c->logN, ys, c->T, PRO_NONE, any_acc ? EPI_MODDOWN_ADD : EPI_MODDOWN, F, st);
  (void)N;
  return FIDES_OK;
}

extern "C" int fides_moddown(fides_context* c, const uint64_t* d, uint32_t l, uint64_t* out, void* stream) {
  if (!c || !d || !out) return fail(FIDES_ERR_ARG, "null argument");
  if (l == 0 || l > c->L) return fail(FIDES_ERR_ARG, "level out of range");
  hipStream_t st = (hipStream_t)stream;
  // work on a copy so the caller's input is preserved
  HIPCHK(hipMemcpyAsync(c->T0, d, (size_t)(l + c->K) * c->N * 8, hipMemcpyDeviceToDevice, st));
  u64* src[1] = {c->T0};
  u64* dst[1] = {out};
  int rc = moddown_multi(c, src, 1, l, dst, nullptr, st);
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// ======================================================================
// keys and key switching
// ======================================================================
extern "C" int fides_key_upload(fides_context* c, const uint64_t* h, fides_key** out) {
  if (!c || !h || !out) return fail(FIDES_ERR_ARG, "null argument");
  const size_t LK = c->L + c->K, nl = 2 * (size_t)c->dnum * LK;
  const size_t bytes = nl * c->N * 8;
  fides_key* k = new fides_key();
  k->ctx = c;
  hipError_t e = hipMalloc((void**)&k->d, bytes);
  if (e != hipSuccess) {
    delete k;
    return fail(FIDES_ERR_NOMEM, "key allocation");
  }
  e = hipMemcpy(k->d, h, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    fides_launch_to_mont(k->d, c->prm, c->logN, (u32)nl, (u32)LK, nullptr);
    e = hipDeviceSynchronize();
  }
  if (e != hipSuccess) {
    (void)hipFree(k->d);
    delete k;
    return fail(FIDES_ERR_HIP, std::string("key upload: ") + hipGetErrorString(e));
  }
  *out = k;
  return FIDES_OK;
}

extern "C" void fides_key_destroy(fides_key* k) {
  if (!k) return;
  (void)hipFree(k->d);
  delete k;
}

// Key switch of d (l limbs, evaluation) -> out0/out1 (l limbs), optionally
// accumulating acc0/acc1 (null = none).  Pipeline of the reference
// modup_ksk_moddown_mgpu (src/CKKS/LimbPartitionMGPU.cu:588-1873):
// INTT -> per-digit ModUp conversion -> NTT of all conversion outputs ->
// inner product with the key -> ModDown of both components (INTT of the
// special limbs, conversion, NTT fused with (x - conv) P^-1 [+ acc]).
static int keyswitch_core(fides_context* c, const u64* d, u32 l, const fides_key* key, u64* out0, u64* out1,
                          const u64* acc0, const u64* acc1, hipStream_t st) {
  const u32 N = c->N, K = c->K, L = c->L;
  const u32 beta = num_digits(c, l);
  // 1. coefficient-form copy of the input
  HIPCHK(hipMemcpyAsync(c->Dc, d, (size_t)l * N * 8, hipMemcpyDeviceToDevice, st));
  {
    LimbSpan s = span0();
    span_add(s, c->Dc, 0, l);
    fides_launch_intt(c->logN, s, c->T, st);
  }
  // 2. ModUp conversions of every digit
  for (u32 j = 0; j < beta; ++j) {
    int rc = modup_digit(c, d, l, j, c->EXT + (size_t)j * (L + K) * N, true, st);
    if (rc) return rc;
  }
  // 3. NTT of all conversion outputs (one launch, precomputed task list)
  {
    LimbSpan s = span0();
    s.list = c->ext_tasks + c->ext_task_off[l];
    s.total = c->ext_task_cnt[l];
    NttFuse F;
    memset(&F, 0, sizeof(F));
    fides_launch_ntt(c->logN, s, c->T, PRO_NONE, EPI_NONE, F, st);
  }
  // 4. inner product with the key
  {
    KsDotArgs A;
    A.d = d;
    A.ext = c->EXT;
    A.key_a = key->d;
    A.key_b = key->d + (size_t)c->dnum * (L + K) * N;
    A.o0 = c->T0;
    A.o1 = c->T1;
    A.prm = c->prm;
    A.logN = c->logN;
    A.l = l;
    A.K = K;
    A.L = L;
    A.beta = beta;
    A.alpha = c->alpha;
    A.ext_stride = L + K;
    A.key_stride = L + K;
    fides_launch_ksk_dot(A, st);
  }
  // 5. ModDown of both components
  u64* src[2] = {c->T0, c->T1};
  u64* dst[2] = {out0, out1};
  const u64* acc[2] = {acc0, acc1};
  return moddown_multi(c, src, 2, l, dst, (acc0 || acc1) ? acc : nullptr, st);
}

static int check_ks(fides_context* c, const fides_key* key, u32 l) {
  if (!c || !key) return fail(FIDES_ERR_ARG, "null argument");
  if (key->ctx != c) return fail(FIDES_ERR_ARG, "key belongs to another context");
  if (l == 0 || l > c->L) return fail(FIDES_ERR_ARG, "level out of range");
  return FIDES_OK;
}

extern "C" int fides_keyswitch(fides_context* c, const uint64_t* d, uint32_t l, const fides_key* key, uint64_t* o0,
                               uint64_t* o1, void* stream) {
  int rc = check_ks(c, key, l);
  if (rc) return rc;
  if (!d || !o0 || !o1) return fail(FIDES_ERR_ARG, "null buffer");
  rc = keyswitch_core(c, d, l, key, o0, o1, nullptr, nullptr, (hipStream_t)stream);
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_hmult(fides_context* c, const uint64_t* a, const uint64_t* b, uint32_t l, const fides_key* rlk,
                           uint64_t* out, void* stream) {
  int rc = check_ks(c, rlk, l);
  if (rc) return rc;
  if (!a || !b || !out) return fail(FIDES_ERR_ARG, "null buffer");
  hipStream_t st = (hipStream_t)stream;
  const size_t S = (size_t)l * c->N;
  fides_launch_tensor(a, a + S, b, b + S, out, out + S, c->D, c->prm, c->logN, l, false, st);
  rc = keyswitch_core(c, c->D, l, rlk, out, out + S, out, out + S, st);
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_square(fides_context* c, const uint64_t* a, uint32_t l, const fides_key* rlk, uint64_t* out,
                            void* stream) {
  int rc = check_ks(c, rlk, l);
  if (rc) return rc;
  if (!a || !out) return fail(FIDES_ERR_ARG, "null buffer");
  hipStream_t st = (hipStream_t)stream;
  const size_t S = (size_t)l * c->N;
  fides_launch_tensor(a, a + S, nullptr, nullptr, out, out + S, c->D, c->prm, c->logN, l, true, st);
  rc = keyswitch_core(c, c->D, l, rlk, out, out + S, out, out + S, st);
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_rescale(fides_context* c, const uint64_t* in, uint32_t l, uint64_t* out, void* stream) {
  if (!c || !in || !out) return fail(FIDES_ERR_ARG, "null argument");
  if (l < 2 || l > c->L) return fail(FIDES_ERR_ARG, "rescale needs 2 <= l <= L");
  if (in == out) return fail(FIDES_ERR_ARG, "rescale output must not alias its input");
  hipStream_t st = (hipStream_t)stream;
  const u32 N = c->N, L = c->L;
  const size_t S = (size_t)l * N, So = (size_t)(l - 1) * N;
  // top limbs of both components -> coefficient form
  HIPCHK(hipMemcpyAsync(c->RS, in + (size_t)(l - 1) * N, (size_t)N * 8, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(c->RS + N, in + S + (size_t)(l - 1) * N, (size_t)N * 8, hipMemcpyDeviceToDevice, st));
  {
    LimbSpan s = span0();
    span_add(s, c->RS, l - 1, 1);
    span_add(s, c->RS + N, l - 1, 1);
    fides_launch_intt(c->logN, s, c->T, st);
  }
  // per component: NTT(SwitchModulus(top)) fused with x * q^-1 + . * QlQlInv
  const u64* row = c->resc + (size_t)l * 4 * L;
  for (u32 p = 0; p < 2; ++p) {
    LimbSpan s = span0();
    span_add(s, out + p * So, 0, l - 1);
    NttFuse F;
    memset(&F, 0, sizeof(F));
    F.pro_src = c->RS + (size_t)p * N;
    F.pro_src_pid = l - 1;
    F.x = span0();
    span_add(F.x, (u64*)in + p * S, 0, l - 1);
    F.out = span0();
    F.acc = span0();
    F.c0 = row;
    F.c0sh = row + L;
    F.c1 = row + 2 * L;
    F.c1sh = row + 3 * L;
    fides_launch_ntt(c->logN, s, c->T, PRO_SWITCH_MODULUS, EPI_RESCALE, F, st);
  }
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_automorph(fides_context* c, const uint64_t* in, uint32_t n, uint32_t k, uint64_t* out,
                               void* stream) {
  if (!c || !in || !out) return fail(FIDES_ERR_ARG, "null argument");
  if ((k & 1) == 0 || k >= 2 * c->N) return fail(FIDES_ERR_ARG, "automorphism index must be odd and < 2N");
  if (in == out) return fail(FIDES_ERR_ARG, "automorphism output must not alias its input");
  fides_launch_automorph(in, out, c->logN, n, k, (hipStream_t)stream);
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_rotate(fides_context* c, const uint64_t* in, uint32_t l, const fides_key* key, uint32_t k,
                            uint64_t* out, void* stream) {
  int rc = check_ks(c, key, l);
  if (rc) return rc;
  if (!in || !out) return fail(FIDES_ERR_ARG, "null buffer");
  if ((k & 1) == 0 || k >= 2 * c->N) return fail(FIDES_ERR_ARG, "automorphism index must be odd and < 2N");
  if (in == out) return fail(FIDES_ERR_ARG, "rotate output must not alias its input");
  hipStream_t st = (hipStream_t)stream;
  const size_t S = (size_t)l * c->N;
  rc = keyswitch_core(c, in + S, l, key, c->TMP, c->TMP + S, in, nullptr, st);
  if (rc) return rc;
  fides_launch_automorph(c->TMP, out, c->logN, 2 * l, k, st);
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}
