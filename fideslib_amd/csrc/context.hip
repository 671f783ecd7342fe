// Host side of the C ABI: context tables, key storage and the executor that
// sequences the HIP kernels of the key-switching pipelines.
//
// Every ciphertext operation (single or batched) is a list of "jobs"; the
// executor builds the per-limb task records of every pipeline stage on the
// host, uploads them in one copy through a pinned ring of task buffers, and
// issues one launch per stage for the whole list.  A single-ciphertext call
// is a list of one job, so both paths run the same kernels.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

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

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(FIDES_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_));  \
  } while (0)

// Tracing: every compute entry point is one ROCTx range named after it, so
// `rocprofv3 --marker-trace` shows each call beside the kernels it launched
// (the reference wraps its API calls in NVTX ranges, CudaNvtxRange in
// src/CudaUtils.cuh:26-45, e.g. api/CryptoContext.cpp:127).  Without a
// profiler attached a range is two calls into the ROCTx library.
namespace {
struct TraceRange {
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};
}  // namespace
#define FIDES_TRACE() TraceRange fides_trace_range_(__func__)

#define RCHK(x)               \
  do {                        \
    int r_ = (x);             \
    if (r_ != FIDES_OK) return r_; \
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
// Inverse of an odd k modulo a power of two m.
static u32 inv_odd(u32 k, u32 m) {
  u64 x = k;  // Newton iteration, doubling correct bits
  for (int i = 0; i < 5; ++i) x = x * (2 - (u64)k * x);
  return (u32)(x & (m - 1));
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

constexpr int RING = 16;  // task-buffer slots in flight

struct TaskSlot {
  char* host = nullptr;  // pinned
  char* dev = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;    // the operation that read this slot has finished
  hipEvent_t copied = nullptr;  // the slot's upload has landed on the device
  bool used = false;
};
constexpr size_t SLOT_BYTES = 8u << 20;  // initial capacity of every task slot

// Capture of a sequence of operations into one HIP graph.  While capturing,
// task lists go to a pool owned by the graph (copied eagerly on the copy
// stream, so the graph holds kernel nodes only) and the key-switch scratch
// comes from an arena owned by the graph, sized up front.
struct Capture {
  hipStream_t st = nullptr;
  char* pool = nullptr;
  size_t pool_cap = 0, pool_used = 0;
  u64* arena = nullptr;
  size_t arena_words = 0;
  bool own_arena = true;  // false: the caller's scratch (fides_graph_begin_shared)
  u64* eager_arena = nullptr;  // the context's own arena, restored at the end
  size_t eager_words = 0;
};

struct fides_context {
  u32 logN, N, L, K, dnum, alpha;
  int device = 0;  // HIP device current at creation: tables, keys and scratch live there
  std::vector<u64> primes, roots;
  std::vector<PrimeC> hprm;
  // device tables
  PrimeC* prm = nullptr;
  u64* tw = nullptr;  // psi | psi_sh | ipsi | ipsi_sh, each [(L+K) N]
  NttTables T{};
  ConvDesc* modup = nullptr;    // [(L+1) * dnum]
  ConvDesc* moddown = nullptr;  // [(L+1)]
  std::vector<ConvDesc> hmodup, hmoddown;
  // host copies of per-limb scalars
  std::vector<u64> pinv, pinv_sh;                    // [L]
  std::vector<u64> pmod, pmod_sh;                    // [L] P mod q_i (the ext basis scales by P)
  std::vector<u64> qlinv, qlinv_sh, qlql, qlql_sh;   // [(L+1) * L], row l = dropping q_{l-1}
  // scratch arena (grows on demand) and task ring
  u64* arena = nullptr;
  size_t arena_words = 0;
  TaskSlot ring[RING];
  int ring_next = 0;
  hipStream_t copy_stream = nullptr;  // task-list uploads, overlapped with the previous operation's kernels
  bool conv_split = false;            // every prime < 2^60: carry-free basis conversion (k_conv_batch)
  bool wide = false;                  // some prime >= 2^61: at most 16 products per 128-bit accumulation
  Capture* cap = nullptr;             // graph capture in progress
  fides_tuning tune{{8, 16, 32}, {1, 1, 1}, {1, 1, 1}};  // launch shapes per size class (fides_tuning_set)
  std::vector<void*> allocs;

  u64* limb(u64* base, u64 i) const { return base + (i << logN); }
  const u64* limb(const u64* base, u64 i) const { return base + (i << logN); }
};

// ----------------------------------------------------------------------
// Device scratch of the allocation-free key switch (fides_keyswitch_noalloc):
// a file-scope array that every GPU running this library holds, at its own
// address, until the process ends.  Host code never uses the array's name as
// an address: each device's address comes from hipGetSymbolAddress, taken
// once per device (with that device current) when a context is created there.
constexpr size_t NOALLOC_SCRATCH_WORDS = (size_t)20 << 20;  // 160 MiB
static __device__ __align__(256) u64 scratch[NOALLOC_SCRATCH_WORDS];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static u64* g_scratch_ptr[MAX_DEVICES];

// Address of `scratch` on the current device `dev`, resolved on first use.
static int scratch_init(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(FIDES_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return FIDES_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(FIDES_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[dev] = (u64*)p;
  return FIDES_OK;
}

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
static NttFuse fuse0() {
  NttFuse F;
  memset(&F, 0, sizeof(F));
  return F;
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
      const u64 m = mulm(prod, R, pt);
      d.mat[t][i] = c->conv_split ? ((m & 0x3fffffffull) | ((m >> 30) << 32)) : m;
    }
  }
}

static void digit_range(const fides_context* c, u32 l, u32 j, u32& st, u32& en) {
  st = j * c->alpha;
  en = std::min((j + 1) * c->alpha, l);
}
static u32 num_digits(const fides_context* c, u32 l) { return (l + c->alpha - 1) / c->alpha; }

extern "C" void fides_context_destroy(fides_context* c) {
  if (!c) return;
  if (c->cap) {  // capture left open: discard it
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(c->cap->st, &g);
    if (g) (void)hipGraphDestroy(g);
    if (c->cap->pool) (void)hipFree(c->cap->pool);
    if (c->cap->arena) (void)hipFree(c->cap->arena);
    c->arena = c->cap->eager_arena;
    delete c->cap;
    c->cap = nullptr;
  }
  (void)hipDeviceSynchronize();
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->arena) (void)hipFree(c->arena);
  for (auto& s : c->ring) {
    if (s.host) (void)hipHostFree(s.host);
    if (s.dev) (void)hipFree(s.dev);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.copied) (void)hipEventDestroy(s.copied);
  }
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  delete c;
}

extern "C" int fides_context_create(const fides_params* P, fides_context** out) {
  FIDES_TRACE();
  if (!P || !out) return fail(FIDES_ERR_ARG, "null argument");
  *out = nullptr;
  if (P->logN < 10 || P->logN > 17) return fail(FIDES_ERR_UNSUPPORTED, "logN must be in [10, 17]");
  if (!P->q || !P->p || P->L == 0 || P->L > 64 || P->K == 0 || P->K > CONV_MAXS || P->dnum == 0)
    return fail(FIDES_ERR_ARG, "need 1 <= L <= 64, 1 <= K <= 16, dnum >= 1 and prime arrays");
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
  // The inner product accumulates one 128-bit product per active digit (up to
  // dnum); the column sums of Acc128 hold 16 products of values below 2^62
  // (64 below 2^61).
  const bool wide = std::any_of(pr.begin(), pr.end(), [](u64 q) { return q >= (1ull << 61); });
  if (wide && P->dnum > 16) return fail(FIDES_ERR_UNSUPPORTED, "dnum > 16 needs every prime below 2^61");
  fides_context* c = new fides_context();
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return fail(FIDES_ERR_HIP, "hipGetDevice");
  }
  if (const int rc = scratch_init(c->device)) {
    delete c;
    return rc;
  }
  c->logN = P->logN;
  c->N = 1u << P->logN;
  c->L = P->L;
  c->K = P->K;
  c->dnum = P->dnum;
  c->alpha = alpha;
  c->primes = pr;
  c->conv_split = std::all_of(pr.begin(), pr.end(), [](u64 q) { return q < (1ull << 60); });
  c->wide = wide;
  const u32 NP = P->L + P->K, N = c->N, L = c->L, K = c->K;
  c->roots.resize(NP);
  for (u32 i = 0; i < NP; ++i) {
    const u64 psi = P->psi ? P->psi[i] : default_root(pr[i], P->logN);
    if (psi == 0 || powm(psi, N, pr[i]) != pr[i] - 1) {
      delete c;
      return fail(FIDES_ERR_ARG, "psi[" + std::to_string(i) + "] is not a primitive 2N-th root of unity");
    }
    c->roots[i] = psi;
  }
  // ---- per-prime constants and twiddle tables (psi^brev(j), psi^-brev(j), Shoup companions)
  std::vector<u64> tw((size_t)4 * NP * N);
  c->hprm.resize(NP);
  auto work = [&](u32 i) {
    const u64 q = pr[i];
    PrimeC& p = c->hprm[i];
    memset(&p, 0, sizeof(p));
    p.q = q;
    p.q2 = 2 * q;
    u64 inv = q;  // Newton iteration for q^-1 mod 2^64
    for (int it = 0; it < 6; ++it) inv *= 2 - q * inv;
    p.qneg_inv = (u64)0 - inv;
    p.bits = bitlen(q);
    const u64 r1 = (u64)(((u128)1 << 64) % q);
    p.r2 = mulm(r1, r1, q);
    const u64 ipsi = invm(c->roots[i], q);
    p.ninv = invm(N % q, q);
    p.ninv_sh = shoup_of(p.ninv, q);
    const u64 w1 = powm(ipsi, N / 2, q);  // psi^-brev(1)
    p.ninv_w = mulm(w1, p.ninv, q);
    p.ninv_w_sh = shoup_of(p.ninv_w, q);
    p.one_sh = shoup_of(1, q);
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
#define ALLOC(ptr, bytes)                                        \
  do {                                                           \
    if ((rc = dalloc(c, (void**)&(ptr), (bytes))) != FIDES_OK) { \
      fides_context_destroy(c);                                  \
      return rc;                                                 \
    }                                                            \
  } while (0)
#define H2D(dst, src, bytes)                                                       \
  do {                                                                             \
    hipError_t e_ = hipMemcpy((dst), (src), (bytes), hipMemcpyHostToDevice);       \
    if (e_ != hipSuccess) {                                                        \
      fides_context_destroy(c);                                                    \
      return fail(FIDES_ERR_HIP, std::string("upload: ") + hipGetErrorString(e_)); \
    }                                                                              \
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
  c->T.narrow = 1;
  for (u32 i = 0; i < NP; ++i)
    if (c->hprm[i].q >> 60) c->T.narrow = 0;
  // ---- ModDown and rescale scalars
  c->pinv.resize(L);
  c->pinv_sh.resize(L);
  c->pmod.resize(L);
  c->pmod_sh.resize(L);
  for (u32 i = 0; i < L; ++i) {
    const u64 qi = pr[i];
    u64 Pm = 1;
    for (u32 k = 0; k < K; ++k) Pm = mulm(Pm, pr[L + k] % qi, qi);
    c->pmod[i] = Pm;
    c->pmod_sh[i] = shoup_of(Pm, qi);
    c->pinv[i] = invm(Pm, qi);
    c->pinv_sh[i] = shoup_of(c->pinv[i], qi);
  }
  c->qlinv.assign((size_t)(L + 1) * L, 0);
  c->qlinv_sh = c->qlql = c->qlql_sh = c->qlinv;
  for (u32 l = 2; l <= L; ++l) {
    const u64 qt = pr[l - 1];
    for (u32 i = 0; i + 1 < l; ++i) {
      const u64 qi = pr[i];
      const u64 qtinv = invm(qt % qi, qi);
      // QlQlInvModqlDivqlModq_i = floor(Q' [Q'^-1]_qt / qt) mod q_i with
      // Q' = prod_{i<l-1} q_i: Q' [Q'^-1]_qt is 1 mod qt and 0 mod q_i, so the
      // floor is (Q' [Q'^-1]_qt - 1) / qt = -qt^-1 mod q_i.
      const u64 big = (qi - qtinv) % qi;
      const size_t o = (size_t)l * L + i;
      c->qlinv[o] = qtinv;
      c->qlinv_sh[o] = shoup_of(qtinv, qi);
      c->qlql[o] = big;
      c->qlql_sh[o] = shoup_of(big, qi);
    }
  }
  // ---- conversion descriptors per level (ModUp per digit, ModDown)
  c->hmodup.resize((size_t)(L + 1) * c->dnum);
  c->hmoddown.resize(L + 1);
  for (u32 l = 1; l <= L; ++l) {
    const u32 beta = num_digits(c, l);
    for (u32 j = 0; j < beta; ++j) {
      u32 st, en;
      digit_range(c, l, j, st, en);
      std::vector<u32> src, dst, off;
      for (u32 i = st; i < en; ++i) src.push_back(i);
      for (u32 i = 0; i < l; ++i)
        if (i < st || i >= en) {
          dst.push_back(i);
          off.push_back(i);
        }
      for (u32 k = 0; k < K; ++k) {
        dst.push_back(L + k);
        off.push_back(l + k);
      }
      make_conv(c, src, dst, off, c->hmodup[(size_t)l * c->dnum + j]);
    }
    std::vector<u32> src, dst, off;
    for (u32 k = 0; k < K; ++k) src.push_back(L + k);
    for (u32 i = 0; i < l; ++i) {
      dst.push_back(i);
      off.push_back(i);
    }
    make_conv(c, src, dst, off, c->hmoddown[l]);
  }
  ALLOC(c->modup, sizeof(ConvDesc) * c->hmodup.size());
  H2D(c->modup, c->hmodup.data(), sizeof(ConvDesc) * c->hmodup.size());
  ALLOC(c->moddown, sizeof(ConvDesc) * c->hmoddown.size());
  H2D(c->moddown, c->hmoddown.data(), sizeof(ConvDesc) * c->hmoddown.size());
  // Task slots are allocated up front: an allocation inside a stream of
  // operations would stall the host while the device drains.
  if (hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess) {
    fides_context_destroy(c);
    return fail(FIDES_ERR_HIP, "copy stream creation");
  }
  for (auto& s : c->ring) {
    if (hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.copied, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&s.host, SLOT_BYTES, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&s.dev, SLOT_BYTES) != hipSuccess) {
      fides_context_destroy(c);
      return fail(FIDES_ERR_NOMEM, "task slot allocation");
    }
    s.cap = SLOT_BYTES;
  }
#undef ALLOC
#undef H2D
  *out = c;
  return FIDES_OK;
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
extern "C" int fides_copy_device(void* dst, const void* src, size_t bytes, void* stream) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return FIDES_OK;
}
extern "C" int fides_stream_sync(void* stream) {
  HIPCHK(hipStreamSynchronize((hipStream_t)stream));
  return FIDES_OK;
}
extern "C" int fides_stream_create(void** stream) {
  if (!stream) return fail(FIDES_ERR_ARG, "null argument");
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return FIDES_OK;
}
extern "C" int fides_stream_destroy(void* stream) {
  if (stream) HIPCHK(hipStreamDestroy((hipStream_t)stream));
  return FIDES_OK;
}

// ======================================================================
// graph capture
// ======================================================================
struct fides_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<void*> mem;  // task pool and scratch arena of the captured operations
};

static void capture_free(Capture* g) {
  if (g->pool) (void)hipFree(g->pool);
  if (g->arena && g->own_arena) (void)hipFree(g->arena);
  delete g;
}

// scratch: the caller's key-switch scratch (not owned), or null for one the
// graph allocates and owns.
static int graph_begin(fides_context* c, void* stream, void* scratch, size_t scratch_bytes, size_t task_bytes) {
  if (!c || !stream) return fail(FIDES_ERR_ARG, "graph capture needs a context and a (non-default) stream");
  if (c->cap) return fail(FIDES_ERR_ARG, "a graph capture is already in progress on this context");
  if (task_bytes == 0 || scratch_bytes == 0) return fail(FIDES_ERR_ARG, "empty graph pools");
  if (scratch && ((uintptr_t)scratch & 255)) return fail(FIDES_ERR_ARG, "shared scratch must be 256-byte aligned");
  Capture* g = new Capture();
  g->st = (hipStream_t)stream;
  g->own_arena = scratch == nullptr;
  if (scratch) g->arena = (u64*)scratch;
  if (hipMalloc((void**)&g->pool, task_bytes) != hipSuccess ||
      (!scratch && hipMalloc((void**)&g->arena, scratch_bytes) != hipSuccess)) {
    capture_free(g);
    return fail(FIDES_ERR_NOMEM, "graph pool allocation");
  }
  g->pool_cap = task_bytes;
  g->arena_words = scratch_bytes / 8;
  hipError_t e = hipStreamBeginCapture(g->st, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    capture_free(g);
    return fail(FIDES_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  }
  g->eager_arena = c->arena;
  g->eager_words = c->arena_words;
  c->arena = g->arena;
  c->arena_words = g->arena_words;
  c->cap = g;
  return FIDES_OK;
}

extern "C" int fides_graph_begin(fides_context* c, void* stream, size_t scratch_bytes, size_t task_bytes) {
  return graph_begin(c, stream, nullptr, scratch_bytes, task_bytes);
}

extern "C" int fides_graph_begin_shared(fides_context* c, void* stream, void* scratch, size_t scratch_bytes,
                                        size_t task_bytes) {
  if (!scratch) return fail(FIDES_ERR_ARG, "null shared scratch");
  return graph_begin(c, stream, scratch, scratch_bytes, task_bytes);
}

extern "C" int fides_graph_end(fides_context* c, fides_graph** out) {
  if (!c || !out) return fail(FIDES_ERR_ARG, "null argument");
  if (!c->cap) return fail(FIDES_ERR_ARG, "no graph capture in progress");
  Capture* g = c->cap;
  c->cap = nullptr;
  c->arena = g->eager_arena;
  c->arena_words = g->eager_words;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(g->st, &graph);
  if (e != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    capture_free(g);
    return fail(FIDES_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    capture_free(g);
    return fail(FIDES_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  fides_graph* G = new fides_graph();
  G->graph = graph;
  G->exec = exec;
  G->mem = {g->pool};
  if (g->own_arena) G->mem.push_back(g->arena);
  delete g;
  *out = G;
  return FIDES_OK;
}

extern "C" int fides_graph_launch(fides_graph* g, void* stream) {
  FIDES_TRACE();
  if (!g) return fail(FIDES_ERR_ARG, "null graph");
  HIPCHK(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return FIDES_OK;
}

extern "C" int fides_graph_destroy(fides_graph* g) {
  if (!g) return FIDES_OK;
  (void)hipDeviceSynchronize();  // no launch of the graph may still read its pools
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (void* p : g->mem) (void)hipFree(p);
  delete g;
  return FIDES_OK;
}

// ======================================================================
// executor: task blob + arena
// ======================================================================
// Host-side builder of all task lists of one operation.  Lists are appended
// to one byte blob; `off` values are byte offsets resolved against the
// device copy after upload.
struct Blob {
  std::vector<char> b;
  template <class T>
  size_t add(const std::vector<T>& v) {
    size_t o = (b.size() + 255) & ~(size_t)255;
    b.resize(o + v.size() * sizeof(T));
    if (!v.empty()) memcpy(b.data() + o, v.data(), v.size() * sizeof(T));
    return o;
  }
};

// Upload a blob through the next ring slot; returns its device base.
// During a graph capture the blob goes to the graph's task pool instead.
static int upload_blob(fides_context* c, const Blob& blob, hipStream_t st, char** dev, int* slot_out,
                       bool may_grow = true) {
  if (c->cap) {
    Capture& g = *c->cap;
    if (st != g.st) return fail(FIDES_ERR_ARG, "operation on another stream during a graph capture");
    const size_t off = (g.pool_used + 255) & ~(size_t)255;
    if (off + blob.b.size() > g.pool_cap)
      return fail(FIDES_ERR_NOMEM, "graph task pool exhausted (task_bytes of fides_graph_begin)");
    HIPCHK(hipMemcpyAsync(g.pool + off, blob.b.data(), blob.b.size(), hipMemcpyHostToDevice, c->copy_stream));
    HIPCHK(hipStreamSynchronize(c->copy_stream));
    g.pool_used = off + blob.b.size();
    *dev = g.pool + off;
    *slot_out = -1;
    return FIDES_OK;
  }
  TaskSlot& s = c->ring[c->ring_next];
  *slot_out = c->ring_next;
  c->ring_next = (c->ring_next + 1) % RING;
  if (s.used) HIPCHK(hipEventSynchronize(s.done));  // slot's previous operation has finished
  const size_t need = std::max<size_t>(blob.b.size(), 1);
  if (need > s.cap && !may_grow)
    return fail(FIDES_ERR_ARG, "task list of " + std::to_string(need) + " bytes exceeds the preallocated task slot");
  if (need > s.cap) {
    if (s.host) HIPCHK(hipHostFree(s.host));
    if (s.dev) HIPCHK(hipFree(s.dev));
    s.host = nullptr;
    s.dev = nullptr;
    const size_t cap = std::max<size_t>(need * 2, 1 << 20);
    HIPCHK(hipHostMalloc((void**)&s.host, cap, hipHostMallocDefault));
    HIPCHK(hipMalloc((void**)&s.dev, cap));
    s.cap = cap;
  }
  memcpy(s.host, blob.b.data(), blob.b.size());
  // The upload runs on the copy stream while `st` still executes earlier
  // work; `st` waits for it only before this operation's first kernel.
  HIPCHK(hipMemcpyAsync(s.dev, s.host, blob.b.size(), hipMemcpyHostToDevice, c->copy_stream));
  HIPCHK(hipEventRecord(s.copied, c->copy_stream));
  HIPCHK(hipStreamWaitEvent(st, s.copied, 0));
  *dev = s.dev;
  return FIDES_OK;
}

static int finish_slot(fides_context* c, int slot, hipStream_t st) {
  if (slot < 0) return FIDES_OK;  // graph pool
  TaskSlot& s = c->ring[slot];
  HIPCHK(hipEventRecord(s.done, st));
  s.used = true;
  return FIDES_OK;
}

static int ensure_arena(fides_context* c, size_t words, hipStream_t st) {
  if (c->cap) {
    if (st != c->cap->st) return fail(FIDES_ERR_ARG, "operation on another stream during a graph capture");
    if (words > c->arena_words)
      return fail(FIDES_ERR_NOMEM, "graph scratch arena too small: " + std::to_string(words * 8) +
                                       " bytes needed (scratch_bytes of fides_graph_begin)");
    return FIDES_OK;
  }
  if (words <= c->arena_words) return FIDES_OK;
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipDeviceSynchronize());
  if (c->arena) HIPCHK(hipFree(c->arena));
  c->arena = nullptr;
  c->arena_words = 0;
  const size_t w = words + words / 4;
  hipError_t e = hipMalloc((void**)&c->arena, w * 8);
  if (e != hipSuccess) return fail(FIDES_ERR_NOMEM, "scratch arena allocation");
  c->arena_words = w;
  return FIDES_OK;
}

template <class T>
static const T* at(char* dev, size_t off) {
  return (const T*)(dev + off);
}

static int launch_ntt_list(fides_context* c, const NttTask* list, u32 n, int pro, int epi, hipStream_t st) {
  for (u32 o = 0; o < n; o += 65535) {
    LimbSpan s = span0();
    s.list = list + o;
    s.total = std::min<u32>(65535, n - o);
    if (fides_launch_ntt(c->logN, s, c->T, pro, epi, fuse0(), st)) return fail(FIDES_ERR_UNSUPPORTED, "logN");
  }
  return FIDES_OK;
}
static int launch_intt_list(fides_context* c, const NttTask* list, u32 n, hipStream_t st) {
  for (u32 o = 0; o < n; o += 65535) {
    LimbSpan s = span0();
    s.list = list + o;
    s.total = std::min<u32>(65535, n - o);
    if (fides_launch_intt(c->logN, s, c->T, st)) return fail(FIDES_ERR_UNSUPPORTED, "logN");
  }
  return FIDES_OK;
}
static int launch_conv_list(fides_context* c, const ConvTask* list, u32 n, u32 ns, hipStream_t st) {
  for (u32 o = 0; o < n; o += 65535)
    if (fides_launch_conv_batch(list + o, std::min<u32>(65535, n - o), ns, c->conv_split, c->prm, c->logN, st))
      return fail(FIDES_ERR_UNSUPPORTED, "conversion size");
  return FIDES_OK;
}

static NttTask ntask(u64* data, u32 pid) {
  NttTask t;
  memset(&t, 0, sizeof(t));
  t.data = data;
  t.pid = pid;
  return t;
}

// ======================================================================
// key-switch jobs
// ======================================================================
struct KsJob {
  const u64* d;  // l limbs, evaluation form
  u32 l;
  const fides_key* key;
  u64* out0;        // l limbs each (or, with rescale, [2][l-1] at out0)
  u64* out1;
  const u64* acc0;  // nullable
  const u64* acc1;  // nullable
  u32 perm = 0;     // store the result through the automorphism with this k^-1 (0: none)
  bool rescale = false;  // fuse Rescale into the ModDown: out0 = [2][l-1] result of
                         // Rescale(acc + KeySwitch(d)); acc0/acc1 required
  int share = -1;        // index of an earlier job with the same d and l whose ModUp
                         // (decomposed, extended digits) this job reuses (hoisting)
  bool ext = false;      // no ModDown: out0 receives [2][l+K] limbs in Q_l u P,
                         // sigma_kidx(T0 + P acc0) and sigma_kidx(T1) (acc0 nullable)
  u32 kidx = 1;          // automorphism of an ext job
};

// Arena words needed by the key switch of one job.
static size_t ks_words(const fides_context* c, u32 l, bool shared_modup = false) {
  const size_t N = c->N;
  const u32 beta = num_digits(c, l);
  const size_t modup = shared_modup ? 0 : (size_t)l + (size_t)beta * (l + c->K);
  return N * (modup + 2 * (size_t)(l + c->K) + 2 * (size_t)l + 2);
}

// Appends the task lists of the batched key switch (reference pipeline
// modup_ksk_moddown_mgpu, src/CKKS/LimbPartitionMGPU.cu:588-1873):
//   1 INTT of the input (out of place into Dc)
//   2 ModUp conversion of every digit (one launch for all digit sizes)
//   3 NTT of every conversion output
//   4 inner product with the key
//   5 INTT of the special limbs of both products
//   6 ModDown conversion P -> Q_l of both products
//   7 NTT fused with (x - conv) * P^-1 [+ acc] into the outputs
//   (fused Rescale jobs) top limb: acc + T P^-1 -> INTT -> - conv P^-1 (one
//   INTT launch pair with the two element-wise steps fused in), then
//   one NTT per remaining limb with PRO_MDRS / EPI_MDRS (see ntt.h).
struct KsLists {
  size_t intt_in, conv_up[CONV_MAXS + 1], ntt_ext, dot, intt_sp, conv_down, ntt_md, intt_top, ntt_mdrs,
      auto_ext;
  u32 n_auto_ext = 0;
  // dot tasks grouped by digit count for the fused row pass: [start, count) per beta
  u32 dot_start[ROWDOT_MAXB + 1] = {}, dot_count[ROWDOT_MAXB + 1] = {};
  bool fused_dot = false;
  u32 conv_up_maxns = 0;  // largest source count in the mixed list conv_up[0]
  u32 n_intt_in = 0, n_conv_up[CONV_MAXS + 1] = {}, n_ntt_ext = 0, n_dot = 0, n_intt_sp = 0, n_conv_down = 0,
      n_ntt_md = 0, n_intt_top = 0, n_ntt_mdrs = 0;
  bool any_acc = false;
};

// Targets per conversion task: the sources of a task are read and pre-scaled
// once for all its targets, so large batches take long chunks; a small
// operation keeps short ones to give the GPU enough workgroups.
static u32 conv_class(size_t total_targets) { return total_targets >= 1024 ? 2 : total_targets >= 512 ? 1 : 0; }
static u32 conv_chunk(const fides_context* c, size_t total_targets) {
  return c->tune.conv_chunk[conv_class(total_targets)];
}

static void ks_build(fides_context* c, const std::vector<KsJob>& jobs, u64* scratch, Blob& blob, KsLists& K_) {
  const u32 N = c->N, K = c->K, L = c->L;
  std::vector<NttTask> intt_in, ntt_ext, intt_sp, ntt_md, intt_top, ntt_mdrs;
  std::vector<ConvTask> conv_up[CONV_MAXS + 1], conv_down;
  std::vector<DotTask> dot;
  std::vector<AutoTask> auto_ext;
  u64* w = scratch;
  std::vector<u64*> ext_of(jobs.size(), nullptr);
  size_t up_targets = 0, down_targets = 0;
  for (const KsJob& J : jobs) {
    if (J.share < 0)
      for (u32 j = 0; j < num_digits(c, J.l); ++j) up_targets += c->hmodup[(size_t)J.l * c->dnum + j].nt;
    down_targets += 2 * J.l;
  }
  const u32 up_chunk = conv_chunk(c, up_targets), down_chunk = conv_chunk(c, down_targets);
  const u32 cls = conv_class(up_targets);
  for (size_t jn = 0; jn < jobs.size(); ++jn) {
    const KsJob& J = jobs[jn];
    const u32 l = J.l, beta = num_digits(c, l), ext_stride = l + K;
    const bool shared = J.share >= 0;
    u64* Dc = w;
    u64* EXT = shared ? ext_of[J.share] : Dc + (size_t)l * N;
    u64* T0 = shared ? w : EXT + (size_t)beta * ext_stride * N;
    u64* T1 = T0 + (size_t)(l + K) * N;
    u64* Y = T1 + (size_t)(l + K) * N;
    u64* RS = Y + (size_t)2 * l * N;
    w = RS + (size_t)2 * N;
    ext_of[jn] = EXT;
    for (u32 i = 0; i < l && !shared; ++i) {
      NttTask t = ntask(c->limb(Dc, i), i);
      t.src = c->limb(J.d, i);
      intt_in.push_back(t);
    }
    for (u32 j = 0; j < beta && !shared; ++j) {
      u32 st, en;
      digit_range(c, l, j, st, en);
      const u32 ns = en - st;
      const ConvDesc& hd = c->hmodup[(size_t)l * c->dnum + j];
      u64* slab = EXT + (size_t)j * ext_stride * N;
      for (u32 t0 = 0; t0 < hd.nt; t0 += up_chunk)
        conv_up[ns].push_back(ConvTask{c->limb(Dc, st), slab, c->modup + (size_t)l * c->dnum + j, t0,
                                       std::min(hd.nt, t0 + up_chunk)});
      for (u32 t = 0; t < hd.nt; ++t) ntt_ext.push_back(ntask(c->limb(slab, hd.dst_off[t]), hd.dst_pid[t]));
    }
    const u64* ka = J.key->d;
    const u64* kb = J.key->d + (size_t)c->dnum * (L + K) * N;
    for (u32 t = 0; t < l + K; ++t) {
      const u32 pid = t < l ? t : L + (t - l);
      DotTask d;
      d.d = t < l ? c->limb(J.d, t) : nullptr;
      d.e = c->limb(EXT, t);
      d.ka = c->limb(ka, pid);
      d.kb = c->limb(kb, pid);
      d.o0 = c->limb(T0, t);
      d.o1 = c->limb(T1, t);
      d.beta = beta;
      d.jown = t < l ? t / c->alpha : 0xffffffffu;
      d.pid = pid;
      d.estride = ext_stride;
      d.ext = J.ext ? 1u : 0u;
      d.pad = 0;
      dot.push_back(d);
    }
    u64* Tp[2] = {T0, T1};
    if (J.ext) {  // Q_l u P result, permuted, with P c0 added to component 0
      for (u32 p = 0; p < 2; ++p)
        for (u32 t = 0; t < l + K; ++t) {
          AutoTask a;
          a.in = c->limb(Tp[p], t);
          a.out = c->limb(J.out0, (size_t)p * (l + K) + t);
          a.add = (p == 0 && J.acc0 && t < l) ? c->limb(J.acc0, t) : nullptr;
          a.c = t < l ? c->pmod[t] : 0;
          a.csh = t < l ? c->pmod_sh[t] : 0;
          a.kidx = J.kidx;
          a.pid = t < l ? t : L + (t - l);
          auto_ext.push_back(a);
        }
      continue;
    }
    u64* outp[2] = {J.out0, J.out1};
    const u64* accp[2] = {J.acc0, J.acc1};
    const ConvDesc& md = c->hmoddown[l];
    for (u32 p = 0; p < 2; ++p) {
      for (u32 k = 0; k < K; ++k) intt_sp.push_back(ntask(c->limb(Tp[p], l + k), L + k));
      u64* Yp = Y + (size_t)p * l * N;
      for (u32 t0 = 0; t0 < md.nt; t0 += down_chunk)
        conv_down.push_back(ConvTask{c->limb(Tp[p], l), Yp, c->moddown + l, t0, std::min(md.nt, t0 + down_chunk)});
      if (!J.rescale) {
        for (u32 i = 0; i < l; ++i) {
          NttTask t = ntask(c->limb(Yp, i), i);
          t.x = c->limb(Tp[p], i);
          t.acc = accp[p] ? c->limb(accp[p], i) : nullptr;
          t.out = c->limb(outp[p], i);
          t.c0 = c->pinv[i];
          t.c0sh = c->pinv_sh[i];
          t.perm = J.perm;
          ntt_md.push_back(t);
          if (accp[p]) K_.any_acc = true;
        }
        continue;
      }
      // ModDown fused with Rescale (l -> l-1).  With R = acc + (T - NTT(conv)) P^-1
      // the rescaled limb i is R_i ql^-1 + NTT(SwitchModulus(INTT(R_top))) QlQlInv_i,
      // which equals acc_i ql^-1 + T_i (P^-1 ql^-1) + NTT(z_i) with
      // z_i = SwitchModulus(top) QlQlInv_i - conv_i (P^-1 ql^-1) and
      // top = INTT(acc_top + T_top P^-1) - conv_top P^-1 (all maps are linear mod q_i).
      const u32 tp = l - 1;
      const u64 qt = c->primes[tp];
      u64* rs = RS + (size_t)p * N;
      // top = INTT(acc_top + T_top P^-1) - conv_top P^-1 in one fused INTT
      NttTask top = ntask(rs, tp);
      top.src = c->limb(accp[p], tp);
      top.x = c->limb(Tp[p], tp);
      top.c0 = c->pinv[tp];
      top.c0sh = c->pinv_sh[tp];
      const u64 npinv = (qt - c->pinv[tp]) % qt;
      top.src2 = c->limb(Yp, tp);
      top.c1 = npinv;
      top.c1sh = shoup_of(npinv, qt);
      intt_top.push_back(top);
      u64* dst = J.out0 + (size_t)p * (l - 1) * N;
      for (u32 i = 0; i + 1 < l; ++i) {
        const u64 qi = c->primes[i];
        const size_t o = (size_t)l * L + i;
        const u64 pq = mulm(c->pinv[i], c->qlinv[o], qi);
        NttTask t = ntask(c->limb(dst, i), i);
        t.src = rs;
        t.src_pid = tp;
        t.swdiff = qi - qt % qi;
        t.src2 = c->limb(Yp, i);
        t.x = c->limb(Tp[p], i);
        t.acc = c->limb(accp[p], i);
        t.c0 = c->qlinv[o];
        t.c0sh = c->qlinv_sh[o];
        t.c2 = c->qlql[o];
        t.c2sh = c->qlql_sh[o];
        t.c3 = pq;
        t.c3sh = shoup_of(pq, qi);
        ntt_mdrs.push_back(t);
      }
    }
  }
  // Group limbs of the same prime so a prime's twiddle tables stay in L2
  // while its limbs are transformed (order does not affect results).
  auto by_prime = [c, cls](std::vector<NttTask>& v) {
    if (c->tune.group_by_prime[cls])
      std::stable_sort(v.begin(), v.end(), [](const NttTask& a, const NttTask& b) { return a.pid < b.pid; });
  };
  by_prime(intt_in);
  by_prime(ntt_ext);
  by_prime(intt_sp);
  by_prime(ntt_md);
  by_prime(ntt_mdrs);
  // Key limbs of the same prime (shared by every job using the same key) next
  // to each other; with the fused row pass, tasks are also grouped by digit
  // count (one launch per count).
  u32 maxb = 0;
  for (const DotTask& d : dot) maxb = std::max(maxb, d.beta);
  K_.fused_dot = c->tune.fuse_row_dot[cls] && maxb <= (u32)ROWDOT_MAXB;
  std::stable_sort(dot.begin(), dot.end(), [&](const DotTask& a, const DotTask& b) {
    if (K_.fused_dot && a.beta != b.beta) return a.beta > b.beta;
    return a.pid < b.pid;
  });
  if (K_.fused_dot)
    for (u32 i = 0; i < dot.size(); ++i) {
      const u32 b = dot[i].beta;
      if (K_.dot_count[b] == 0) K_.dot_start[b] = i;
      K_.dot_count[b]++;
    }
  K_.intt_in = blob.add(intt_in);
  K_.n_intt_in = (u32)intt_in.size();
  // Digits of different sizes (the partial last digit of each level) go into
  // one launch (list 0, largest digits first) instead of one small launch
  // per size.
  u32 sizes = 0;
  for (u32 ns = 1; ns <= CONV_MAXS; ++ns) sizes += !conv_up[ns].empty();
  if (sizes > 1)
    for (u32 ns = CONV_MAXS; ns >= 1; --ns) {
      if (conv_up[ns].empty()) continue;
      if (K_.conv_up_maxns == 0) K_.conv_up_maxns = ns;
      conv_up[0].insert(conv_up[0].end(), conv_up[ns].begin(), conv_up[ns].end());
      conv_up[ns].clear();
    }
  for (u32 ns = 0; ns <= CONV_MAXS; ++ns) {
    K_.conv_up[ns] = blob.add(conv_up[ns]);
    K_.n_conv_up[ns] = (u32)conv_up[ns].size();
  }
  K_.ntt_ext = blob.add(ntt_ext);
  K_.n_ntt_ext = (u32)ntt_ext.size();
  K_.dot = blob.add(dot);
  K_.n_dot = (u32)dot.size();
  K_.intt_sp = blob.add(intt_sp);
  K_.n_intt_sp = (u32)intt_sp.size();
  K_.conv_down = blob.add(conv_down);
  K_.n_conv_down = (u32)conv_down.size();
  K_.ntt_md = blob.add(ntt_md);
  K_.n_ntt_md = (u32)ntt_md.size();
  K_.intt_top = blob.add(intt_top);
  K_.n_intt_top = (u32)intt_top.size();
  K_.ntt_mdrs = blob.add(ntt_mdrs);
  K_.n_ntt_mdrs = (u32)ntt_mdrs.size();
  K_.auto_ext = blob.add(auto_ext);
  K_.n_auto_ext = (u32)auto_ext.size();
}

static int ks_launch(fides_context* c, char* dev, const KsLists& K_, hipStream_t st) {
  RCHK(launch_intt_list(c, at<NttTask>(dev, K_.intt_in), K_.n_intt_in, st));
  for (u32 o = 0; o < K_.n_conv_up[0]; o += 65535)
    if (fides_launch_conv_mixed(at<ConvTask>(dev, K_.conv_up[0]) + o, std::min<u32>(65535, K_.n_conv_up[0] - o),
                                K_.conv_up_maxns, c->conv_split, c->prm, c->logN, st))
      return fail(FIDES_ERR_UNSUPPORTED, "conversion size");
  for (u32 ns = 1; ns <= CONV_MAXS; ++ns)
    if (K_.n_conv_up[ns]) RCHK(launch_conv_list(c, at<ConvTask>(dev, K_.conv_up[ns]), K_.n_conv_up[ns], ns, st));
  if (K_.fused_dot) {
    // column pass of the ModUp outputs, then row pass fused with the inner product
    for (u32 o = 0; o < K_.n_ntt_ext; o += 65535) {
      LimbSpan sp = span0();
      sp.list = at<NttTask>(dev, K_.ntt_ext) + o;
      sp.total = std::min<u32>(65535, K_.n_ntt_ext - o);
      if (fides_launch_ntt_col(c->logN, sp, c->T, PRO_NONE, fuse0(), st)) return fail(FIDES_ERR_UNSUPPORTED, "logN");
    }
    int counts = 0, maxb = 0;
    for (int b = 1; b <= ROWDOT_MAXB; ++b)
      if (K_.dot_count[b]) ++counts, maxb = b;
    // every digit count in one launch (tasks sorted heaviest first) where the
    // ring has that kernel
    const bool mixed = counts > 1 && (c->logN == 16 || c->logN == 17);
    if (mixed) {
      for (u32 o = 0; o < K_.n_dot; o += 65535)
        if (fides_launch_row_dot_mixed(c->logN, at<DotTask>(dev, K_.dot) + o, std::min<u32>(65535, K_.n_dot - o),
                                       maxb, c->T, c->L + c->K, st))
          return fail(FIDES_ERR_UNSUPPORTED, "logN");
    } else {
      for (int b = 1; b <= ROWDOT_MAXB; ++b)
        for (u32 o = 0; o < K_.dot_count[b]; o += 65535)
          if (fides_launch_row_dot(c->logN, at<DotTask>(dev, K_.dot) + K_.dot_start[b] + o,
                                   std::min<u32>(65535, K_.dot_count[b] - o), b, c->T, c->L + c->K, st))
            return fail(FIDES_ERR_UNSUPPORTED, "logN");
    }
  } else {
    RCHK(launch_ntt_list(c, at<NttTask>(dev, K_.ntt_ext), K_.n_ntt_ext, PRO_NONE, EPI_NONE, st));
    for (u32 o = 0; o < K_.n_dot; o += 65535)
      fides_launch_dot_batch(at<DotTask>(dev, K_.dot) + o, std::min<u32>(65535, K_.n_dot - o), c->prm, c->logN,
                             c->L + c->K, st);
  }
  if (K_.fused_dot) {
    // the row stages of the special limbs' INTT ran inside the inner product
    for (u32 o = 0; o < K_.n_intt_sp; o += 65535) {
      LimbSpan sp = span0();
      sp.list = at<NttTask>(dev, K_.intt_sp) + o;
      sp.total = std::min<u32>(65535, K_.n_intt_sp - o);
      if (fides_launch_intt_col(c->logN, sp, c->T, st)) return fail(FIDES_ERR_UNSUPPORTED, "logN");
    }
  } else {
    RCHK(launch_intt_list(c, at<NttTask>(dev, K_.intt_sp), K_.n_intt_sp, st));
  }
  RCHK(launch_conv_list(c, at<ConvTask>(dev, K_.conv_down), K_.n_conv_down, c->K, st));
  if (K_.n_ntt_md)
    RCHK(launch_ntt_list(c, at<NttTask>(dev, K_.ntt_md), K_.n_ntt_md, PRO_NONE,
                         K_.any_acc ? EPI_MODDOWN_ADD : EPI_MODDOWN, st));
  if (K_.n_ntt_mdrs) {
    for (u32 o = 0; o < K_.n_intt_top; o += 65535) {
      LimbSpan sp = span0();
      sp.list = at<NttTask>(dev, K_.intt_top) + o;
      sp.total = std::min<u32>(65535, K_.n_intt_top - o);
      if (fides_launch_intt_axpy(c->logN, sp, c->T, st)) return fail(FIDES_ERR_UNSUPPORTED, "logN");
    }
    RCHK(launch_ntt_list(c, at<NttTask>(dev, K_.ntt_mdrs), K_.n_ntt_mdrs, PRO_MDRS, EPI_MDRS, st));
  }
  for (u32 o = 0; o < K_.n_auto_ext; o += 65535)
    fides_launch_automorph_batch(at<AutoTask>(dev, K_.auto_ext) + o, std::min<u32>(65535, K_.n_auto_ext - o), c->prm,
                                 c->logN, st);
  return FIDES_OK;
}

// Rescale jobs: in [2][l] -> out [2][l-1] (OpenFHE DropLastElementAndScale;
// reference doubleRescaleMGPU, src/CKKS/LimbPartitionMGPU.cu:196-394):
//   INTT of both top limbs (out of place), then per remaining limb
//   NTT(SwitchModulus(top)) fused with x * q_top^-1 + . * QlQlInvModqlDivqlModq.
struct RsJob {
  const u64* in;
  u32 l;
  u64* out;
};
struct RsLists {
  size_t intt, ntt;
  u32 n_intt = 0, n_ntt = 0;
};
static void rs_build(fides_context* c, const std::vector<RsJob>& jobs, u64* scratch, Blob& blob, RsLists& R) {
  const u32 N = c->N, L = c->L;
  std::vector<NttTask> intt, ntt;
  u64* w = scratch;
  for (const RsJob& J : jobs) {
    const u32 l = J.l;
    u64* RS = w;
    w += 2 * (size_t)N;
    for (u32 p = 0; p < 2; ++p) {
      const u64* inp = J.in + (size_t)p * l * N;
      u64* outp = J.out + (size_t)p * (l - 1) * N;
      NttTask t = ntask(RS + (size_t)p * N, l - 1);
      t.src = c->limb(inp, l - 1);
      intt.push_back(t);
      for (u32 i = 0; i + 1 < l; ++i) {
        NttTask u = ntask(c->limb(outp, i), i);
        u.src = RS + (size_t)p * N;
        u.src_pid = l - 1;
        u.swdiff = c->primes[i] - c->primes[l - 1] % c->primes[i];
        u.x = c->limb(inp, i);
        const size_t o = (size_t)l * L + i;
        u.c0 = c->qlinv[o];
        u.c0sh = c->qlinv_sh[o];
        u.c1 = c->qlql[o];
        u.c1sh = c->qlql_sh[o];
        ntt.push_back(u);
      }
    }
  }
  std::stable_sort(ntt.begin(), ntt.end(), [](const NttTask& a, const NttTask& b) { return a.pid < b.pid; });
  R.intt = blob.add(intt);
  R.n_intt = (u32)intt.size();
  R.ntt = blob.add(ntt);
  R.n_ntt = (u32)ntt.size();
}
static int rs_launch(fides_context* c, char* dev, const RsLists& R, hipStream_t st) {
  RCHK(launch_intt_list(c, at<NttTask>(dev, R.intt), R.n_intt, st));
  RCHK(launch_ntt_list(c, at<NttTask>(dev, R.ntt), R.n_ntt, PRO_SWITCH_MODULUS, EPI_RESCALE, st));
  return FIDES_OK;
}

// ======================================================================
// raw limb operations
// ======================================================================
// In-place NTT/INTT of n limbs with arbitrary prime ids: runs of
// consecutive ids become inline segments; more than 8 runs go through a
// task list uploaded via the task ring.
static int raw_transform(fides_context* c, u64* data, const uint32_t* pids, uint32_t n, bool inverse,
                         hipStream_t st) {
  const u32 NP = c->L + c->K;
  for (u32 t = 0; t < n; ++t)
    if (pids[t] >= NP) return fail(FIDES_ERR_ARG, "prime id out of range");
  std::vector<std::pair<u32, u32>> segs;  // (first limb, count) of runs of consecutive prime ids
  for (u32 t = 0; t < n; ++t) {
    if (!segs.empty() && pids[t] == pids[segs.back().first] + segs.back().second)
      segs.back().second++;
    else
      segs.push_back({t, 1});
  }
  if (segs.size() <= 8) {
    LimbSpan s = span0();
    for (auto& sg : segs) span_add(s, c->limb(data, sg.first), pids[sg.first], sg.second);
    const int rc = inverse ? fides_launch_intt(c->logN, s, c->T, st)
                           : fides_launch_ntt(c->logN, s, c->T, PRO_NONE, EPI_NONE, fuse0(), st);
    if (rc) return fail(FIDES_ERR_UNSUPPORTED, "ring dimension");
    HIPCHK(hipGetLastError());
    return FIDES_OK;
  }
  std::vector<NttTask> tasks(n);
  for (u32 t = 0; t < n; ++t) tasks[t] = ntask(c->limb(data, t), pids[t]);
  Blob blob;
  const size_t o = blob.add(tasks);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  if (inverse)
    RCHK(launch_intt_list(c, at<NttTask>(dev, o), n, st));
  else
    RCHK(launch_ntt_list(c, at<NttTask>(dev, o), n, PRO_NONE, EPI_NONE, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_ntt(fides_context* c, uint64_t* d, const uint32_t* pids, uint32_t n, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!d || !pids))) return fail(FIDES_ERR_ARG, "null argument");
  if (n == 0) return FIDES_OK;
  return raw_transform(c, d, pids, n, false, (hipStream_t)stream);
}

extern "C" int fides_intt(fides_context* c, uint64_t* d, const uint32_t* pids, uint32_t n, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!d || !pids))) return fail(FIDES_ERR_ARG, "null argument");
  if (n == 0) return FIDES_OK;
  return raw_transform(c, d, pids, n, true, (hipStream_t)stream);
}

extern "C" int fides_modup(fides_context* c, const uint64_t* d, uint32_t l, uint32_t j, uint64_t* out,
                           void* stream) {
  FIDES_TRACE();
  if (!c || !d || !out) return fail(FIDES_ERR_ARG, "null argument");
  if (l == 0 || l > c->L || j >= num_digits(c, l)) return fail(FIDES_ERR_ARG, "level / digit out of range");
  hipStream_t st = (hipStream_t)stream;
  const u32 N = c->N;
  u32 stt, en;
  digit_range(c, l, j, stt, en);
  const u32 sz = en - stt;
  RCHK(ensure_arena(c, (size_t)sz * N, st));
  Blob blob;
  std::vector<NttTask> intt, ntt;
  for (u32 i = 0; i < sz; ++i) {
    NttTask t = ntask(c->limb(c->arena, i), stt + i);
    t.src = c->limb(d, stt + i);
    intt.push_back(t);
  }
  const ConvDesc& hd = c->hmodup[(size_t)l * c->dnum + j];
  std::vector<ConvTask> conv;
  for (u32 t0 = 0; t0 < hd.nt; t0 += CONV_TCHUNK)
    conv.push_back(ConvTask{c->arena, out, c->modup + (size_t)l * c->dnum + j, t0, std::min(hd.nt, t0 + CONV_TCHUNK)});
  for (u32 t = 0; t < hd.nt; ++t) ntt.push_back(ntask(c->limb(out, hd.dst_off[t]), hd.dst_pid[t]));
  const size_t o_intt = blob.add(intt), o_conv = blob.add(conv), o_ntt = blob.add(ntt);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(launch_intt_list(c, at<NttTask>(dev, o_intt), sz, st));
  RCHK(launch_conv_list(c, at<ConvTask>(dev, o_conv), (u32)conv.size(), sz, st));
  RCHK(launch_ntt_list(c, at<NttTask>(dev, o_ntt), (u32)ntt.size(), PRO_NONE, EPI_NONE, st));
  HIPCHK(hipMemcpyAsync(c->limb(out, stt), c->limb(d, stt), (size_t)sz * N * 8, hipMemcpyDeviceToDevice, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_moddown(fides_context* c, const uint64_t* d, uint32_t l, uint64_t* out, void* stream) {
  FIDES_TRACE();
  if (!c || !d || !out) return fail(FIDES_ERR_ARG, "null argument");
  if (l == 0 || l > c->L) return fail(FIDES_ERR_ARG, "level out of range");
  hipStream_t st = (hipStream_t)stream;
  const u32 N = c->N, K = c->K, L = c->L;
  // scratch: copy of the special limbs (INTT in place) + conversion output
  RCHK(ensure_arena(c, (size_t)(K + l) * N, st));
  u64* sp = c->arena;
  u64* Y = c->arena + (size_t)K * N;
  Blob blob;
  std::vector<NttTask> intt, ntt;
  for (u32 k = 0; k < K; ++k) {
    NttTask t = ntask(c->limb(sp, k), L + k);
    t.src = c->limb(d, l + k);
    intt.push_back(t);
  }
  const ConvDesc& md = c->hmoddown[l];
  std::vector<ConvTask> conv;
  for (u32 t0 = 0; t0 < md.nt; t0 += CONV_TCHUNK)
    conv.push_back(ConvTask{sp, Y, c->moddown + l, t0, std::min(md.nt, t0 + CONV_TCHUNK)});
  for (u32 i = 0; i < l; ++i) {
    NttTask t = ntask(c->limb(Y, i), i);
    t.x = c->limb(d, i);
    t.out = c->limb(out, i);
    t.c0 = c->pinv[i];
    t.c0sh = c->pinv_sh[i];
    ntt.push_back(t);
  }
  const size_t o_intt = blob.add(intt), o_conv = blob.add(conv), o_ntt = blob.add(ntt);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(launch_intt_list(c, at<NttTask>(dev, o_intt), K, st));
  RCHK(launch_conv_list(c, at<ConvTask>(dev, o_conv), (u32)conv.size(), K, st));
  RCHK(launch_ntt_list(c, at<NttTask>(dev, o_ntt), l, PRO_NONE, EPI_MODDOWN, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// ======================================================================
// keys
// ======================================================================
extern "C" int fides_key_upload(fides_context* c, const uint64_t* h, fides_key** out) {
  FIDES_TRACE();
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

extern "C" int fides_key_download(const fides_key* k, uint64_t* h) {
  if (!k || !h) return fail(FIDES_ERR_ARG, "null argument");
  const fides_context* c = k->ctx;
  const size_t LK = c->L + c->K, nl = 2 * (size_t)c->dnum * LK, N = c->N;
  HIPCHK(hipMemcpy(h, k->d, nl * N * 8, hipMemcpyDeviceToHost));
  // out of Montgomery form: x * 2^-64 mod q
  for (size_t limb = 0; limb < nl; ++limb) {
    const u64 q = c->primes[limb % LK];
    const u64 r = (u64)(((unsigned __int128)1 << 64) % q);
    const u64 rinv = powm(r, q - 2, q);
    u64* x = h + limb * N;
    for (size_t i = 0; i < N; ++i) x[i] = mulm(x[i], rinv, q);
  }
  return FIDES_OK;
}

extern "C" void fides_key_destroy(fides_key* k) {
  if (!k) return;
  (void)hipFree(k->d);
  delete k;
}

// ======================================================================
// ciphertext operations (single = batch of one)
// ======================================================================
// [a, a + na) and [b, b + nb) share a word.
static bool overlaps(const u64* a, size_t na, const u64* b, size_t nb) { return a < b + nb && b < a + na; }

static int check_key(fides_context* c, const fides_key* key, u32 l) {
  if (!key) return fail(FIDES_ERR_ARG, "null key");
  if (key->ctx != c) return fail(FIDES_ERR_ARG, "key belongs to another context");
  if (l == 0 || l > c->L) return fail(FIDES_ERR_ARG, "level out of range");
  return FIDES_OK;
}

static int run_keyswitch(fides_context* c, const std::vector<KsJob>& jobs, hipStream_t st) {
  size_t words = 0;
  for (const KsJob& J : jobs) words += ks_words(c, J.l);
  RCHK(ensure_arena(c, words, st));
  Blob blob;
  KsLists K_;
  ks_build(c, jobs, c->arena, blob, K_);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(ks_launch(c, dev, K_, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_keyswitch(fides_context* c, const uint64_t* d, uint32_t l, const fides_key* key, uint64_t* o0,
                               uint64_t* o1, void* stream) {
  FIDES_TRACE();
  if (!c) return fail(FIDES_ERR_ARG, "null context");
  RCHK(check_key(c, key, l));
  if (!d || !o0 || !o1) return fail(FIDES_ERR_ARG, "null buffer");
  return run_keyswitch(c, {KsJob{d, l, key, o0, o1, nullptr, nullptr}}, (hipStream_t)stream);
}

// The same key switch with its scratch in the file-scope device array and
// its task list in a preallocated task slot: no device or pinned allocation.
extern "C" int fides_keyswitch_noalloc(fides_context* c, const uint64_t* d, uint32_t l, const fides_key* key,
                                       uint64_t* o0, uint64_t* o1, void* stream) {
  if (!c) return fail(FIDES_ERR_ARG, "null context");
  RCHK(check_key(c, key, l));
  if (!d || !o0 || !o1) return fail(FIDES_ERR_ARG, "null buffer");
  const size_t words = ks_words(c, l);
  if (words > NOALLOC_SCRATCH_WORDS)
    return fail(FIDES_ERR_ARG, "key switch needs " + std::to_string(words * 8) +
                                   " bytes of scratch; the allocation-free entry point holds " +
                                   std::to_string(NOALLOC_SCRATCH_WORDS * 8));
  RCHK(scratch_init(c->device));
  u64* scr = g_scratch_ptr[c->device];
  hipStream_t st = (hipStream_t)stream;
  Blob blob;
  KsLists K_;
  ks_build(c, {KsJob{d, l, key, o0, o1, nullptr, nullptr}}, scr, blob, K_);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot, false));
  RCHK(ks_launch(c, dev, K_, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// HMult (+ relinearize [+ rescale]) of n ciphertext pairs.
static int hmult_batch(fides_context* c, u32 n, const u64* const* a, const u64* const* b, const u32* lv,
                       const fides_key* rlk, bool square, bool rescale, u64* const* out, hipStream_t st) {
  const u32 N = c->N;
  for (u32 i = 0; i < n; ++i) {
    RCHK(check_key(c, rlk, lv[i]));
    if (!a[i] || (!square && !b[i]) || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (rescale && lv[i] < 2) return fail(FIDES_ERR_ARG, "rescale needs l >= 2");
  }
  // Without rescale the tensor product writes out[i] in the launch that reads
  // every job's inputs: out[i] may only coincide with its own a[i] or b[i]
  // (each element is read before the same thread writes it).  With rescale
  // the outputs are written after the last read of any input.  Outputs never
  // overlap each other.
  for (u32 i = 0; i < n; ++i) {
    const size_t so = (size_t)2 * (rescale ? lv[i] - 1 : lv[i]) * N;
    for (u32 j = 0; j < n; ++j) {
      const size_t sj = (size_t)2 * lv[j] * N, sjo = (size_t)2 * (rescale ? lv[j] - 1 : lv[j]) * N;
      if (j < i && overlaps(out[i], so, out[j], sjo)) return fail(FIDES_ERR_ARG, "outputs overlap");
      if (rescale) continue;
      for (const u64* in : {a[j], square ? nullptr : b[j]}) {
        if (!in || (j == i && in == out[i])) continue;
        if (overlaps(out[i], so, in, sj)) return fail(FIDES_ERR_ARG, "output overlaps an input of the batch");
      }
    }
  }
  // arena: [key-switch scratch | per job: c1d1 (l) | tensor product [2][l] if rescaling]
  size_t ksw = 0, extra = 0;
  for (u32 i = 0; i < n; ++i) {
    ksw += ks_words(c, lv[i]);
    extra += (size_t)N * (lv[i] + (rescale ? 2 * lv[i] : 0));
  }
  RCHK(ensure_arena(c, ksw + extra, st));
  u64* w = c->arena + ksw;
  std::vector<TensorTask> tens;
  std::vector<KsJob> jobs;
  for (u32 i = 0; i < n; ++i) {
    const u32 l = lv[i];
    u64* D = w;
    w += (size_t)l * N;
    u64* prod = out[i];
    if (rescale) {
      prod = w;
      w += (size_t)2 * l * N;
    }
    const size_t S = (size_t)l * N;
    for (u32 t = 0; t < l; ++t) {
      TensorTask tt;
      tt.a0 = c->limb(a[i], t);
      tt.a1 = c->limb(a[i] + S, t);
      tt.b0 = square ? nullptr : c->limb(b[i], t);
      tt.b1 = square ? nullptr : c->limb(b[i] + S, t);
      tt.o0 = c->limb(prod, t);
      tt.o1 = c->limb(prod + S, t);
      tt.o2 = c->limb(D, t);
      tt.pid = t;
      tt.square = square ? 1u : 0u;
      tens.push_back(tt);
    }
    KsJob J{D, l, rlk, prod, prod + S, prod, prod + S};
    if (rescale) {
      J.rescale = true;
      J.out0 = out[i];
      J.out1 = nullptr;
    }
    jobs.push_back(J);
  }
  Blob blob;
  const size_t o_tens = blob.add(tens);
  KsLists K_;
  ks_build(c, jobs, c->arena, blob, K_);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  for (u32 o = 0; o < tens.size(); o += 65535)
    fides_launch_tensor_batch(at<TensorTask>(dev, o_tens) + o, std::min<u32>(65535, (u32)tens.size() - o), c->prm,
                              c->logN, st);
  RCHK(ks_launch(c, dev, K_, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_hmult(fides_context* c, const uint64_t* a, const uint64_t* b, uint32_t l, const fides_key* rlk,
                           uint64_t* out, void* stream) {
  FIDES_TRACE();
  if (!c) return fail(FIDES_ERR_ARG, "null context");
  return hmult_batch(c, 1, &a, &b, &l, rlk, false, false, &out, (hipStream_t)stream);
}

extern "C" int fides_square(fides_context* c, const uint64_t* a, uint32_t l, const fides_key* rlk, uint64_t* out,
                            void* stream) {
  FIDES_TRACE();
  if (!c) return fail(FIDES_ERR_ARG, "null context");
  const uint64_t* b = nullptr;
  return hmult_batch(c, 1, &a, &b, &l, rlk, true, false, &out, (hipStream_t)stream);
}

extern "C" int fides_hmult_batch(fides_context* c, uint32_t n, const uint64_t* const* a, const uint64_t* const* b,
                                 const uint32_t* l, const fides_key* rlk, int rescale, uint64_t* const* out,
                                 void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!a || !b || !l || !out))) return fail(FIDES_ERR_ARG, "null argument");
  if (n == 0) return FIDES_OK;
  return hmult_batch(c, n, a, b, l, rlk, false, rescale != 0, out, (hipStream_t)stream);
}

extern "C" int fides_rescale(fides_context* c, const uint64_t* in, uint32_t l, uint64_t* out, void* stream) {
  FIDES_TRACE();
  if (!c || !in || !out) return fail(FIDES_ERR_ARG, "null argument");
  if (l < 2 || l > c->L) return fail(FIDES_ERR_ARG, "rescale needs 2 <= l <= L");
  if (in == out) return fail(FIDES_ERR_ARG, "rescale output must not alias its input");
  hipStream_t st = (hipStream_t)stream;
  RCHK(ensure_arena(c, 2 * (size_t)c->N, st));
  Blob blob;
  RsLists R;
  rs_build(c, {RsJob{in, l, out}}, c->arena, blob, R);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(rs_launch(c, dev, R, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_automorph(fides_context* c, const uint64_t* in, uint32_t n, uint32_t k, uint64_t* out,
                               void* stream) {
  FIDES_TRACE();
  if (!c || !in || !out) return fail(FIDES_ERR_ARG, "null argument");
  if ((k & 1) == 0 || k >= 2 * c->N) return fail(FIDES_ERR_ARG, "automorphism index must be odd and < 2N");
  if (in == out) return fail(FIDES_ERR_ARG, "automorphism output must not alias its input");
  hipStream_t st = (hipStream_t)stream;
  Blob blob;
  std::vector<AutoTask> at_;
  for (u32 t = 0; t < n; ++t) at_.push_back(AutoTask{c->limb(in, t), c->limb(out, t), nullptr, 0, 0, k, 0});
  const size_t o = blob.add(at_);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  for (u32 x = 0; x < n; x += 65535)
    fides_launch_automorph_batch(at<AutoTask>(dev, o) + x, std::min<u32>(65535, n - x), c->prm, c->logN, st);
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// Rotations of n ciphertexts: key switch of c1 with ModDown fused with the
// c0 add, then the permutation of both components (OpenFHE EvalAutomorphism
// order).
static int rotate_batch(fides_context* c, u32 n, const u64* const* in, const u32* lv, const fides_key* const* keys,
                        const u32* kidx, u64* const* out, hipStream_t st) {
  const u32 N = c->N;
  size_t ksw = 0;
  // Rotations of the same ciphertext share one ModUp (hoisting); the results
  // are bit-identical to separate rotations.
  std::vector<int> share(n, -1);
  for (u32 i = 0; i < n; ++i) {
    RCHK(check_key(c, keys[i], lv[i]));
    if (!in[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if ((kidx[i] & 1) == 0 || kidx[i] >= 2 * N) return fail(FIDES_ERR_ARG, "automorphism index must be odd and < 2N");
    // the ModDown store scatters out[i] while other workgroups still read
    // every job's c0: no output may overlap any input or another output
    const size_t si = (size_t)2 * lv[i] * N;
    for (u32 j = 0; j < n; ++j) {
      if (overlaps(out[i], si, in[j], (size_t)2 * lv[j] * N))
        return fail(FIDES_ERR_ARG, "rotate output overlaps an input of the batch");
      if (j < i && overlaps(out[i], si, out[j], (size_t)2 * lv[j] * N)) return fail(FIDES_ERR_ARG, "outputs overlap");
    }
    for (u32 j = 0; j < i; ++j)
      if (in[j] == in[i] && lv[j] == lv[i] && share[j] < 0) {
        share[i] = (int)j;
        break;
      }
    ksw += ks_words(c, lv[i], share[i] >= 0);
  }
  RCHK(ensure_arena(c, ksw, st));
  std::vector<KsJob> jobs;
  for (u32 i = 0; i < n; ++i) {
    const size_t S = (size_t)lv[i] * N;
    KsJob J{in[i] + S, lv[i], keys[i], out[i], out[i] + S, in[i], nullptr};
    // the ModDown store applies the permutation (scatter with k^-1)
    J.perm = inv_odd(kidx[i], 2 * N);
    J.share = share[i];
    jobs.push_back(J);
  }
  Blob blob;
  KsLists K_;
  ks_build(c, jobs, c->arena, blob, K_);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(ks_launch(c, dev, K_, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// n rotations of one ciphertext sharing one ModUp (reference
// Ciphertext::rotate_hoisted, src/CKKS/Ciphertext.cpp:940-1114: the digits are
// raised once, then each index runs the inner product with its own rotation
// key, ModDown, the c0 addition and the automorphism).  The result of each
// index is bit-identical to fides_rotate with the same key; index 1 (the
// identity) copies.
extern "C" int fides_rotate_hoisted(fides_context* c, const uint64_t* in, uint32_t l, uint32_t n,
                                    const fides_key* const* keys, const uint32_t* kidx, uint64_t* const* out,
                                    void* stream) {
  FIDES_TRACE();
  if (!c || !in || (n && (!keys || !kidx || !out))) return fail(FIDES_ERR_ARG, "null argument");
  if (l == 0 || l > c->L) return fail(FIDES_ERR_ARG, "level out of range");
  hipStream_t st = (hipStream_t)stream;
  const u32 N = c->N;
  const size_t S = (size_t)l * N;
  for (u32 i = 0; i < n; ++i) {
    if (!out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if ((kidx[i] & 1) == 0 || kidx[i] >= 2 * N) return fail(FIDES_ERR_ARG, "automorphism index must be odd and < 2N");
    if (out[i] + 2 * S > in && in + 2 * S > out[i]) return fail(FIDES_ERR_ARG, "rotate output must not overlap its input");
    for (u32 j = 0; j < i; ++j)
      if (out[i] + 2 * S > out[j] && out[j] + 2 * S > out[i]) return fail(FIDES_ERR_ARG, "outputs overlap");
    if (kidx[i] != 1) RCHK(check_key(c, keys[i], l));
  }
  std::vector<KsJob> jobs;
  size_t ksw = 0;
  for (u32 i = 0; i < n; ++i) {
    if (kidx[i] == 1) {
      HIPCHK(hipMemcpyAsync(out[i], in, 2 * S * sizeof(u64), hipMemcpyDeviceToDevice, st));
      continue;
    }
    KsJob J{in + S, l, keys[i], out[i], out[i] + S, in, nullptr};
    J.perm = inv_odd(kidx[i], 2 * N);
    J.share = jobs.empty() ? -1 : 0;
    ksw += ks_words(c, l, !jobs.empty());
    jobs.push_back(J);
  }
  if (jobs.empty()) return FIDES_OK;
  RCHK(ensure_arena(c, ksw, st));
  Blob blob;
  KsLists K_;
  ks_build(c, jobs, c->arena, blob, K_);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(ks_launch(c, dev, K_, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_rotate(fides_context* c, const uint64_t* in, uint32_t l, const fides_key* key, uint32_t k,
                            uint64_t* out, void* stream) {
  FIDES_TRACE();
  if (!c) return fail(FIDES_ERR_ARG, "null context");
  return rotate_batch(c, 1, &in, &l, &key, &k, &out, (hipStream_t)stream);
}

extern "C" int fides_rotate_batch(fides_context* c, uint32_t n, const uint64_t* const* in, const uint32_t* l,
                                  const fides_key* const* keys, const uint32_t* k, uint64_t* const* out,
                                  void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!in || !l || !keys || !k || !out))) return fail(FIDES_ERR_ARG, "null argument");
  if (n == 0) return FIDES_OK;
  return rotate_batch(c, n, in, l, keys, k, out, (hipStream_t)stream);
}

// ======================================================================
// extended basis Q_l u P (rotations whose ModDown is hoisted)
// ======================================================================
// An ext ciphertext is [2][l+K][N]: limbs 0..l-1 under q_0..q_{l-1}, then the
// K special primes; it holds P times a ciphertext (OpenFHE's
// EvalFastRotationExt / the reference's Ciphertext::rotate_hoisted with
// ext = true, src/CKKS/Ciphertext.cpp:940-1114).  Sums of ext ciphertexts
// take one ModDown (fides_ext_moddown_batch) instead of one per rotation.
static u32 ext_pid(const fides_context* c, u32 l, u32 t) { return t < l ? t : c->L + (t - l); }
static int run_lin(fides_context* c, const std::vector<LinTask>& tasks, hipStream_t st);

extern "C" int fides_rotate_hoisted_ext(fides_context* c, const uint64_t* in, uint32_t l, uint32_t n,
                                        const fides_key* const* keys, const uint32_t* kidx, int add_first,
                                        uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || !in || (n && (!keys || !kidx || !out))) return fail(FIDES_ERR_ARG, "null argument");
  if (l == 0 || l > c->L) return fail(FIDES_ERR_ARG, "level out of range");
  hipStream_t st = (hipStream_t)stream;
  const u32 N = c->N, K = c->K;
  const size_t S = (size_t)l * N, SE = (size_t)2 * (l + K) * N;
  for (u32 i = 0; i < n; ++i) {
    if (!out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if ((kidx[i] & 1) == 0 || kidx[i] >= 2 * N) return fail(FIDES_ERR_ARG, "automorphism index must be odd and < 2N");
    if (overlaps(out[i], SE, in, 2 * S)) return fail(FIDES_ERR_ARG, "rotate output must not overlap its input");
    for (u32 j = 0; j < i; ++j)
      if (overlaps(out[i], SE, out[j], SE)) return fail(FIDES_ERR_ARG, "outputs overlap");
    if (kidx[i] != 1) RCHK(check_key(c, keys[i], l));
  }
  std::vector<LinTask> ident;  // index 1: (P c0 or 0, P c1)
  std::vector<KsJob> jobs;
  size_t ksw = 0;
  for (u32 i = 0; i < n; ++i) {
    if (kidx[i] == 1) {
      for (u32 p = 0; p < 2; ++p)
        for (u32 t = 0; t < l + K; ++t) {
          LinTask x;
          memset(&x, 0, sizeof(x));
          x.a = c->limb(in, (size_t)p * l + (t < l ? t : 0));
          x.out = c->limb(out[i], (size_t)p * (l + K) + t);
          x.pid = ext_pid(c, l, t);
          x.flags = LIN_MUL;
          x.m = (t < l && (p == 1 || add_first)) ? c->pmod[t] : 0;
          x.msh = (t < l && (p == 1 || add_first)) ? c->pmod_sh[t] : 0;
          ident.push_back(x);
        }
      continue;
    }
    KsJob J{in + S, l, keys[i], out[i], nullptr, add_first ? in : nullptr, nullptr};
    J.ext = true;
    J.kidx = kidx[i];
    J.share = jobs.empty() ? -1 : 0;
    ksw += ks_words(c, l, !jobs.empty());
    jobs.push_back(J);
  }
  if (!ident.empty()) RCHK(run_lin(c, ident, st));
  if (jobs.empty()) return FIDES_OK;
  RCHK(ensure_arena(c, ksw, st));
  Blob blob;
  KsLists K_;
  ks_build(c, jobs, c->arena, blob, K_);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(ks_launch(c, dev, K_, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_ext_raise_batch(fides_context* c, uint32_t n, const uint64_t* const* in, const uint32_t* l,
                                     uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!in || !l || !out))) return fail(FIDES_ERR_ARG, "null argument");
  const u32 K = c->K;
  std::vector<LinTask> tasks;
  for (u32 i = 0; i < n; ++i) {
    if (!in[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l[i] == 0 || l[i] > c->L) return fail(FIDES_ERR_ARG, "level out of range");
    if (overlaps(out[i], (size_t)2 * (l[i] + K) * c->N, in[i], (size_t)2 * l[i] * c->N))
      return fail(FIDES_ERR_ARG, "output must not overlap its input");
    for (u32 p = 0; p < 2; ++p)
      for (u32 t = 0; t < l[i] + K; ++t) {
        LinTask x;
        memset(&x, 0, sizeof(x));
        x.a = c->limb(in[i], (size_t)p * l[i] + (t < l[i] ? t : 0));
        x.out = c->limb(out[i], (size_t)p * (l[i] + K) + t);
        x.pid = ext_pid(c, l[i], t);
        x.flags = LIN_MUL;
        x.m = t < l[i] ? c->pmod[t] : 0;  // P = 0 mod every special prime
        x.msh = t < l[i] ? c->pmod_sh[t] : 0;
        tasks.push_back(x);
      }
  }
  return run_lin(c, tasks, (hipStream_t)stream);
}

extern "C" int fides_ext_add_batch(fides_context* c, uint32_t n, const uint64_t* const* a, const uint64_t* const* b,
                                   const uint32_t* l, int subtract, uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!a || !b || !l || !out))) return fail(FIDES_ERR_ARG, "null argument");
  const u32 K = c->K;
  std::vector<LinTask> tasks;
  for (u32 i = 0; i < n; ++i) {
    if (!a[i] || !b[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l[i] == 0 || l[i] > c->L) return fail(FIDES_ERR_ARG, "level out of range");
    for (u32 p = 0; p < 2; ++p)
      for (u32 t = 0; t < l[i] + K; ++t) {
        const size_t off = (size_t)p * (l[i] + K) + t;
        LinTask x;
        memset(&x, 0, sizeof(x));
        x.a = c->limb(a[i], off);
        x.b = c->limb(b[i], off);
        x.out = c->limb(out[i], off);
        x.pid = ext_pid(c, l[i], t);
        x.flags = LIN_B | (subtract ? LIN_SUB : 0);
        tasks.push_back(x);
      }
  }
  return run_lin(c, tasks, (hipStream_t)stream);
}

// ModDown of both components (src/CKKS/Conv.cu:26-125 ModDown2 + the
// NTT_MODDOWN epilogue, as in fides_moddown) for n ext ciphertexts.
extern "C" int fides_ext_moddown_batch(fides_context* c, uint32_t n, const uint64_t* const* in, const uint32_t* l,
                                       uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!in || !l || !out))) return fail(FIDES_ERR_ARG, "null argument");
  if (n == 0) return FIDES_OK;
  hipStream_t st = (hipStream_t)stream;
  const u32 N = c->N, K = c->K, L = c->L;
  size_t words = 0;
  for (u32 i = 0; i < n; ++i) {
    if (!in[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l[i] == 0 || l[i] > L) return fail(FIDES_ERR_ARG, "level out of range");
    if (overlaps(out[i], (size_t)2 * l[i] * N, in[i], (size_t)2 * (l[i] + K) * N))
      return fail(FIDES_ERR_ARG, "output must not overlap its input");
    words += (size_t)2 * (K + l[i]) * N;
  }
  RCHK(ensure_arena(c, words, st));
  std::vector<NttTask> intt, ntt;
  std::vector<ConvTask> conv;
  u64* w = c->arena;
  for (u32 i = 0; i < n; ++i)
    for (u32 p = 0; p < 2; ++p) {
      const u32 li = l[i];
      const u64* src = in[i] + (size_t)p * (li + K) * N;
      u64* sp = w;
      u64* Y = w + (size_t)K * N;
      w = Y + (size_t)li * N;
      for (u32 k = 0; k < K; ++k) {
        NttTask t = ntask(c->limb(sp, k), L + k);
        t.src = c->limb(src, li + k);
        intt.push_back(t);
      }
      const ConvDesc& md = c->hmoddown[li];
      for (u32 t0 = 0; t0 < md.nt; t0 += CONV_TCHUNK)
        conv.push_back(ConvTask{sp, Y, c->moddown + li, t0, std::min(md.nt, t0 + CONV_TCHUNK)});
      for (u32 j = 0; j < li; ++j) {
        NttTask t = ntask(c->limb(Y, j), j);
        t.x = c->limb(src, j);
        t.out = c->limb(out[i], (size_t)p * li + j);
        t.c0 = c->pinv[j];
        t.c0sh = c->pinv_sh[j];
        ntt.push_back(t);
      }
    }
  std::stable_sort(ntt.begin(), ntt.end(), [](const NttTask& a, const NttTask& b) { return a.pid < b.pid; });
  Blob blob;
  const size_t o_intt = blob.add(intt), o_conv = blob.add(conv), o_ntt = blob.add(ntt);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(launch_intt_list(c, at<NttTask>(dev, o_intt), (u32)intt.size(), st));
  RCHK(launch_conv_list(c, at<ConvTask>(dev, o_conv), (u32)conv.size(), K, st));
  RCHK(launch_ntt_list(c, at<NttTask>(dev, o_ntt), (u32)ntt.size(), PRO_NONE, EPI_MODDOWN, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// ======================================================================
// tuning
// ======================================================================
extern "C" int fides_tuning_get(const fides_context* c, fides_tuning* t) {
  if (!c || !t) return fail(FIDES_ERR_ARG, "null argument");
  *t = c->tune;
  return FIDES_OK;
}

extern "C" int fides_tuning_set(fides_context* c, const fides_tuning* t) {
  if (!c || !t) return fail(FIDES_ERR_ARG, "null argument");
  for (u32 b = 0; b < 3; ++b) {
    if (t->conv_chunk[b] < 1 || t->conv_chunk[b] > CONV_MAXT)
      return fail(FIDES_ERR_ARG, "conv_chunk must be in 1..96");
    if (t->group_by_prime[b] > 1 || t->fuse_row_dot[b] > 1) return fail(FIDES_ERR_ARG, "flags must be 0 or 1");
  }
  c->tune = *t;
  return FIDES_OK;
}

// Synthetic batch of the autotuner: random residues (xorshift, reduced per
// prime), one random key, n HMult+relin+Rescale and n Rotate (index 5).
namespace {
struct TuneBatch {
  fides_context* c;
  fides_key* key = nullptr;
  std::vector<u64*> in, out;
  std::vector<u32> lv, kidx;
  std::vector<const fides_key*> keys;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~TuneBatch() {
    for (u64* p : in) (void)hipFree(p);
    for (u64* p : out) (void)hipFree(p);
    if (key) fides_key_destroy(key);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  int step() {
    const u32 n = (u32)lv.size();
    std::vector<const u64*> a(n), b(n);
    for (u32 i = 0; i < n; ++i) a[i] = in[i], b[i] = in[(i + 1) % n];
    // inputs of a pair must have the same limb count: pair i with itself when they differ
    for (u32 i = 0; i < n; ++i)
      if (lv[(i + 1) % n] != lv[i]) b[i] = in[i];
    RCHK(hmult_batch(c, n, a.data(), b.data(), lv.data(), key, false, true, out.data(), nullptr));
    RCHK(rotate_batch(c, n, a.data(), lv.data(), keys.data(), kidx.data(), out.data(), nullptr));
    return FIDES_OK;
  }
  // best of 3 runs of `reps` steps, after one warm-up step
  int time(u32 reps, float* ms) {
    RCHK(step());
    HIPCHK(hipDeviceSynchronize());
    *ms = 0;
    for (int run = 0; run < 3; ++run) {
      HIPCHK(hipEventRecord(e0, nullptr));
      for (u32 r = 0; r < reps; ++r) RCHK(step());
      HIPCHK(hipEventRecord(e1, nullptr));
      HIPCHK(hipEventSynchronize(e1));
      float t;
      HIPCHK(hipEventElapsedTime(&t, e0, e1));
      t /= reps;
      if (run == 0 || t < *ms) *ms = t;
    }
    return FIDES_OK;
  }
};
}  // namespace

static void fill_residues(std::vector<u64>& h, const fides_context* c, u32 nlimbs, u32 LK, u64& st) {
  const size_t N = c->N;
  h.resize((size_t)nlimbs * N);
  for (u32 m = 0; m < nlimbs; ++m) {
    const u64 q = c->primes[m % LK];
    for (size_t i = 0; i < N; ++i) {
      st ^= st << 13, st ^= st >> 7, st ^= st << 17;
      h[m * N + i] = st % q;
    }
  }
}

extern "C" int fides_autotune(fides_context* c, uint32_t n, const uint32_t* l, uint32_t reps, fides_tuning* out) {
  FIDES_TRACE();
  if (!c || !l || n == 0) return fail(FIDES_ERR_ARG, "null argument or empty batch");
  if (c->cap) return fail(FIDES_ERR_ARG, "autotune during graph capture");
  for (u32 i = 0; i < n; ++i)
    if (l[i] < 2 || l[i] > c->L) return fail(FIDES_ERR_ARG, "autotune needs 2 <= l[i] <= L");
  if (reps == 0) reps = 1;
  const u32 N = c->N, LK = c->L + c->K;
  TuneBatch B{c};
  u64 seed = 0x9e3779b97f4a7c15ull;
  std::vector<u64> h;
  fill_residues(h, c, 2 * c->dnum * LK, LK, seed);
  RCHK(fides_key_upload(c, h.data(), &B.key));
  // one random ciphertext at L limbs (component c, limb m at c * L + m, prime m); input i
  // takes the first l[i] limbs of each component
  fill_residues(h, c, 2 * c->L, c->L, seed);
  for (u32 i = 0; i < n; ++i) {
    u64 *pi = nullptr, *po = nullptr;
    const size_t S = (size_t)l[i] * N;
    HIPCHK(hipMalloc((void**)&pi, 2 * S * 8));
    B.in.push_back(pi);
    HIPCHK(hipMalloc((void**)&po, 2 * S * 8));
    B.out.push_back(po);
    HIPCHK(hipMemcpy(pi, h.data(), S * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pi + S, h.data() + (size_t)c->L * N, S * 8, hipMemcpyHostToDevice));
    B.lv.push_back(l[i]);
    B.kidx.push_back(5);
    B.keys.push_back(B.key);
  }
  HIPCHK(hipEventCreate(&B.e0));
  HIPCHK(hipEventCreate(&B.e1));
  // size classes of the batch's conversions (ModUp and ModDown of one call)
  size_t up = 0, down = 0;
  for (u32 i = 0; i < n; ++i) {
    for (u32 j = 0; j < num_digits(c, l[i]); ++j) up += c->hmodup[(size_t)l[i] * c->dnum + j].nt;
    down += 2 * l[i];
  }
  const u32 cu = conv_class(up);  // class of the flags
  std::vector<u32> classes{cu};    // classes of the conversion chunks
  if (conv_class(down) != cu) classes.push_back(conv_class(down));
  const fides_tuning start = c->tune;
  fides_tuning best = start;
  float tb;
  int rc = B.time(reps, &tb);
  auto trial = [&](const fides_tuning& t) -> int {
    c->tune = t;
    float ms;
    const int r = B.time(reps, &ms);
    if (r != FIDES_OK) return r;
    if (ms < 0.98f * tb) best = t, tb = ms;
    return FIDES_OK;
  };
  if (rc == FIDES_OK) {
    fides_tuning t = best;
    t.fuse_row_dot[cu] ^= 1;
    rc = trial(t);
  }
  if (rc == FIDES_OK) {
    fides_tuning t = best;
    t.group_by_prime[cu] ^= 1;
    rc = trial(t);
  }
  for (u32 cl : classes)
    for (u32 v : {4u, 8u, 16u, 24u, 32u, 48u, 64u}) {
      if (rc != FIDES_OK) break;
      fides_tuning t = best;
      t.conv_chunk[cl] = v;
      if (v != best.conv_chunk[cl]) rc = trial(t);
    }
  c->tune = rc == FIDES_OK ? best : start;
  if (rc != FIDES_OK) return rc;
  if (out) *out = best;
  return FIDES_OK;
}

// ======================================================================
// element-wise ciphertext arithmetic and batched rescale
// ======================================================================
static int run_lin(fides_context* c, const std::vector<LinTask>& tasks, hipStream_t st) {
  if (tasks.empty()) return FIDES_OK;
  Blob blob;
  const size_t o = blob.add(tasks);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  for (u32 x = 0; x < tasks.size(); x += 65535)
    fides_launch_lin_batch(at<LinTask>(dev, o) + x, std::min<u32>(65535, (u32)tasks.size() - x), c->prm, c->logN, st);
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

static LinTask lin0(const u64* a, u64* out, u32 pid) {
  LinTask t;
  memset(&t, 0, sizeof(t));
  t.a = a;
  t.out = out;
  t.pid = pid;
  return t;
}

extern "C" int fides_add_batch(fides_context* c, uint32_t n, const uint64_t* const* a, const uint64_t* const* b,
                               const uint32_t* l, int subtract, uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!a || !b || !l || !out))) return fail(FIDES_ERR_ARG, "null argument");
  std::vector<LinTask> tasks;
  for (u32 i = 0; i < n; ++i) {
    if (!a[i] || !b[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l[i] == 0 || l[i] > c->L) return fail(FIDES_ERR_ARG, "level out of range");
    for (u32 p = 0; p < 2; ++p)
      for (u32 j = 0; j < l[i]; ++j) {
        const size_t off = (size_t)p * l[i] + j;
        LinTask t = lin0(c->limb(a[i], off), c->limb(out[i], off), j);
        t.b = c->limb(b[i], off);
        t.flags = LIN_B | (subtract ? LIN_SUB : 0);
        tasks.push_back(t);
      }
  }
  return run_lin(c, tasks, (hipStream_t)stream);
}

extern "C" int fides_mul_const_batch(fides_context* c, uint32_t n, const uint64_t* const* a, const uint32_t* l,
                                     const uint64_t* const* res, uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!a || !l || !res || !out))) return fail(FIDES_ERR_ARG, "null argument");
  std::vector<LinTask> tasks;
  for (u32 i = 0; i < n; ++i) {
    if (!a[i] || !res[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l[i] == 0 || l[i] > c->L) return fail(FIDES_ERR_ARG, "level out of range");
    for (u32 j = 0; j < l[i]; ++j)
      if (res[i][j] >= c->primes[j]) return fail(FIDES_ERR_ARG, "constant residue not reduced");
    for (u32 p = 0; p < 2; ++p)
      for (u32 j = 0; j < l[i]; ++j) {
        const size_t off = (size_t)p * l[i] + j;
        LinTask t = lin0(c->limb(a[i], off), c->limb(out[i], off), j);
        t.m = res[i][j];
        t.msh = shoup_of(res[i][j], c->primes[j]);
        t.flags = LIN_MUL;
        tasks.push_back(t);
      }
  }
  return run_lin(c, tasks, (hipStream_t)stream);
}

extern "C" int fides_add_const_batch(fides_context* c, uint32_t n, const uint64_t* const* a, const uint32_t* l,
                                     const uint64_t* const* res, uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!a || !l || !res || !out))) return fail(FIDES_ERR_ARG, "null argument");
  std::vector<LinTask> tasks;
  for (u32 i = 0; i < n; ++i) {
    if (!a[i] || !res[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l[i] == 0 || l[i] > c->L) return fail(FIDES_ERR_ARG, "level out of range");
    for (u32 j = 0; j < l[i]; ++j)
      if (res[i][j] >= c->primes[j]) return fail(FIDES_ERR_ARG, "constant residue not reduced");
    for (u32 p = 0; p < 2; ++p)
      for (u32 j = 0; j < l[i]; ++j) {
        const size_t off = (size_t)p * l[i] + j;
        LinTask t = lin0(c->limb(a[i], off), c->limb(out[i], off), j);
        // a constant polynomial evaluates to the same value at every point
        t.add = p == 0 ? res[i][j] : 0;
        tasks.push_back(t);
      }
  }
  return run_lin(c, tasks, (hipStream_t)stream);
}

extern "C" int fides_lin_batch(fides_context* c, uint32_t n, const uint64_t* const* a, const uint64_t* const* b,
                               const uint32_t* l, const uint64_t* const* mul_res, const uint64_t* const* add_res,
                               const uint32_t* sub, uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!a || !l || !out))) return fail(FIDES_ERR_ARG, "null argument");
  std::vector<LinTask> tasks;
  for (u32 i = 0; i < n; ++i) {
    if (!a[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l[i] == 0 || l[i] > c->L) return fail(FIDES_ERR_ARG, "level out of range");
    const uint64_t* m = mul_res ? mul_res[i] : nullptr;
    const uint64_t* k = add_res ? add_res[i] : nullptr;
    const uint64_t* bb = b ? b[i] : nullptr;
    for (u32 j = 0; j < l[i]; ++j)
      if ((m && m[j] >= c->primes[j]) || (k && k[j] >= c->primes[j]))
        return fail(FIDES_ERR_ARG, "constant residue not reduced");
    for (u32 p = 0; p < 2; ++p)
      for (u32 j = 0; j < l[i]; ++j) {
        const size_t off = (size_t)p * l[i] + j;
        LinTask t = lin0(c->limb(a[i], off), c->limb(out[i], off), j);
        if (m) {
          t.m = m[j];
          t.msh = shoup_of(m[j], c->primes[j]);
          t.flags |= LIN_MUL;
        }
        if (bb) {
          t.b = c->limb(bb, off);
          t.flags |= LIN_B | (sub && sub[i] ? LIN_SUB : 0);
        }
        t.add = (k && p == 0) ? k[j] : 0;
        tasks.push_back(t);
      }
  }
  return run_lin(c, tasks, (hipStream_t)stream);
}

extern "C" int fides_drop_limbs_batch(fides_context* c, uint32_t n, const uint64_t* const* in, const uint32_t* l_in,
                                      const uint32_t* l_out, uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!in || !l_in || !l_out || !out))) return fail(FIDES_ERR_ARG, "null argument");
  std::vector<LinTask> tasks;
  for (u32 i = 0; i < n; ++i) {
    if (!in[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l_in[i] == 0 || l_in[i] > c->L || l_out[i] == 0 || l_out[i] > l_in[i])
      return fail(FIDES_ERR_ARG, "need 1 <= l_out <= l_in <= L");
    if (in[i] == out[i] && l_out[i] != l_in[i]) return fail(FIDES_ERR_ARG, "limb drop must not run in place");
    if (in[i] == out[i]) continue;
    for (u32 p = 0; p < 2; ++p)
      for (u32 j = 0; j < l_out[i]; ++j)
        tasks.push_back(lin0(c->limb(in[i], (size_t)p * l_in[i] + j), c->limb(out[i], (size_t)p * l_out[i] + j), j));
  }
  return run_lin(c, tasks, (hipStream_t)stream);
}

extern "C" int fides_rescale_batch(fides_context* c, uint32_t n, const uint64_t* const* in, const uint32_t* l,
                                   uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!in || !l || !out))) return fail(FIDES_ERR_ARG, "null argument");
  if (n == 0) return FIDES_OK;
  hipStream_t st = (hipStream_t)stream;
  std::vector<RsJob> jobs;
  for (u32 i = 0; i < n; ++i) {
    if (!in[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l[i] < 2 || l[i] > c->L) return fail(FIDES_ERR_ARG, "rescale needs 2 <= l <= L");
    if (in[i] == out[i]) return fail(FIDES_ERR_ARG, "rescale output must not alias its input");
    jobs.push_back(RsJob{in[i], l[i], out[i]});
  }
  RCHK(ensure_arena(c, 2 * (size_t)c->N * n, st));
  Blob blob;
  RsLists R;
  rs_build(c, jobs, c->arena, blob, R);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(rs_launch(c, dev, R, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// ======================================================================
// plaintexts and modulus raising (bootstrap building blocks)
// ======================================================================
extern "C" int fides_plaintext_prepare(fides_context* c, uint64_t* pt, uint32_t l, void* stream) {
  FIDES_TRACE();
  if (!c || !pt) return fail(FIDES_ERR_ARG, "null argument");
  if (l == 0 || l > c->L) return fail(FIDES_ERR_ARG, "level out of range");
  hipStream_t st = (hipStream_t)stream;
  std::vector<uint32_t> pids(l);
  for (u32 i = 0; i < l; ++i) pids[i] = i;
  RCHK(raw_transform(c, pt, pids.data(), l, false, st));
  fides_launch_to_mont(pt, c->prm, c->logN, l, l, st);
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_plaintext_embed(fides_context* c, const uint64_t* pt, uint32_t l, uint64_t* ct, void* stream) {
  FIDES_TRACE();
  if (!c || !pt || !ct) return fail(FIDES_ERR_ARG, "null argument");
  if (l == 0 || l > c->L) return fail(FIDES_ERR_ARG, "level out of range");
  if (overlaps(ct, (size_t)2 * l * c->N, pt, (size_t)l * c->N)) return fail(FIDES_ERR_ARG, "output overlaps the plaintext");
  std::vector<LinTask> tasks;
  for (u32 p = 0; p < 2; ++p)
    for (u32 j = 0; j < l; ++j) {
      LinTask t = lin0(c->limb(pt, j), c->limb(ct, (size_t)p * l + j), j);
      t.flags = p == 0 ? LIN_AMONT : LIN_MUL;  // c0 = pt in standard form, c1 = pt * 0
      tasks.push_back(t);
    }
  return run_lin(c, tasks, (hipStream_t)stream);
}

extern "C" int fides_ptdot_batch(fides_context* c, uint32_t n, const uint32_t* cnt, const uint64_t* const* ct,
                                 const uint64_t* const* pt, const uint32_t* l, uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!cnt || !ct || !pt || !l || !out))) return fail(FIDES_ERR_ARG, "null argument");
  std::vector<PtDotTask> tasks;
  std::vector<PtTerm> terms;
  size_t o = 0;
  for (u32 i = 0; i < n; ++i) {
    if (cnt[i] == 0 || cnt[i] > PTDOT_MAXT) return fail(FIDES_ERR_ARG, "term count must be 1..16");
    if (l[i] == 0 || l[i] > c->L) return fail(FIDES_ERR_ARG, "level out of range");
    if (!out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    for (u32 j = 0; j < cnt[i]; ++j)
      if (!ct[o + j] || !pt[o + j]) return fail(FIDES_ERR_ARG, "null term");
    for (u32 t = 0; t < l[i]; ++t) {
      PtDotTask tk;
      tk.out0 = c->limb(out[i], t);
      tk.out1 = c->limb(out[i], (size_t)l[i] + t);
      tk.pid = t;
      tk.cnt = cnt[i];
      tk.first = (u32)terms.size();
      tk.pad = 0;
      for (u32 j = 0; j < cnt[i]; ++j)
        terms.push_back(
            PtTerm{c->limb(ct[o + j], t), c->limb(ct[o + j], (size_t)l[i] + t), c->limb(pt[o + j], t)});
      tasks.push_back(tk);
    }
    o += cnt[i];
  }
  if (tasks.empty()) return FIDES_OK;
  // Tasks reading the same plaintext limbs (the same linear transform applied
  // to several ciphertexts) next to each other, so the plaintext stays in L2.
  std::stable_sort(tasks.begin(), tasks.end(), [&](const PtDotTask& a, const PtDotTask& b) {
    if (a.pid != b.pid) return a.pid < b.pid;
    return terms[a.first].m < terms[b.first].m;
  });
  hipStream_t st = (hipStream_t)stream;
  Blob blob;
  const size_t ot = blob.add(tasks);
  const size_t om = blob.add(terms);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  for (u32 x = 0; x < tasks.size(); x += 65535)
    fides_launch_ptdot_batch(at<PtDotTask>(dev, ot) + x, at<PtTerm>(dev, om), std::min<u32>(65535, (u32)tasks.size() - x),
                             c->prm, c->logN, st);
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

extern "C" int fides_mod_raise_batch(fides_context* c, uint32_t n, const uint64_t* const* in, const uint32_t* l_out,
                                     uint64_t* const* out, void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!in || !l_out || !out))) return fail(FIDES_ERR_ARG, "null argument");
  if (n == 0) return FIDES_OK;
  hipStream_t st = (hipStream_t)stream;
  const u32 N = c->N;
  for (u32 i = 0; i < n; ++i) {
    if (!in[i] || !out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    if (l_out[i] == 0 || l_out[i] > c->L) return fail(FIDES_ERR_ARG, "level out of range");
    if (out[i] < in[i] + 2 * (size_t)N && in[i] < out[i] + 2 * (size_t)l_out[i] * N)
      return fail(FIDES_ERR_ARG, "mod raise output must not overlap its input");
  }
  RCHK(ensure_arena(c, 2 * (size_t)N * n, st));
  std::vector<NttTask> intt, ntt;
  const u64 q0 = c->primes[0];
  for (u32 i = 0; i < n; ++i) {
    u64* tmp = c->arena + (size_t)2 * N * i;
    for (u32 p = 0; p < 2; ++p) {
      NttTask t = ntask(tmp + (size_t)p * N, 0);
      t.src = c->limb(in[i], p);
      intt.push_back(t);
      for (u32 j = 0; j < l_out[i]; ++j) {
        NttTask u = ntask(c->limb(out[i], (size_t)p * l_out[i] + j), j);
        u.src = tmp + (size_t)p * N;
        u.src_pid = 0;
        u.swdiff = c->primes[j] - q0 % c->primes[j];
        ntt.push_back(u);
      }
    }
  }
  std::stable_sort(ntt.begin(), ntt.end(), [](const NttTask& a, const NttTask& b) { return a.pid < b.pid; });
  Blob blob;
  const size_t oi = blob.add(intt);
  const size_t on = blob.add(ntt);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  RCHK(launch_intt_list(c, at<NttTask>(dev, oi), (u32)intt.size(), st));
  RCHK(launch_ntt_list(c, at<NttTask>(dev, on), (u32)ntt.size(), PRO_SWITCH_MODULUS, EPI_NONE, st));
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}

// ======================================================================
// linear combination of ciphertexts with integer constants
// ======================================================================
// R = 2^64 mod q times m (Montgomery form of m), on the host.
static u64 to_mont_host(u64 m, u64 q) {
  const unsigned __int128 r = ((unsigned __int128)1 << 64) % q;
  return (u64)(((unsigned __int128)m * r) % q);
}

extern "C" int fides_lincomb_batch(fides_context* c, uint32_t n, const uint32_t* cnt, const uint64_t* const* x,
                                   const uint32_t* x_l, const uint64_t* const* m_res,
                                   const uint64_t* const* add_res, const uint32_t* l, uint64_t* const* out,
                                   void* stream) {
  FIDES_TRACE();
  if (!c || (n && (!cnt || !x || !x_l || !m_res || !l || !out))) return fail(FIDES_ERR_ARG, "null argument");
  std::vector<LcTask> tasks;
  std::vector<LcTerm> terms;
  size_t o = 0;
  for (u32 i = 0; i < n; ++i) {
    if (cnt[i] == 0 || cnt[i] > LINCOMB_MAXT) return fail(FIDES_ERR_ARG, "term count must be 1..64");
    if (c->wide && cnt[i] > 16) return fail(FIDES_ERR_ARG, "term count must be 1..16 when a prime is >= 2^61");
    if (l[i] == 0 || l[i] > c->L) return fail(FIDES_ERR_ARG, "level out of range");
    if (!out[i]) return fail(FIDES_ERR_ARG, "null buffer");
    const uint64_t* a = add_res ? add_res[i] : nullptr;
    for (u32 j = 0; j < cnt[i]; ++j) {
      if (!x[o + j] || !m_res[o + j]) return fail(FIDES_ERR_ARG, "null term");
      if (x_l[o + j] < l[i] || x_l[o + j] > c->L) return fail(FIDES_ERR_ARG, "term below the output level");
      for (u32 t = 0; t < l[i]; ++t)
        if (m_res[o + j][t] >= c->primes[t]) return fail(FIDES_ERR_ARG, "constant residue not reduced");
    }
    for (u32 t = 0; t < l[i]; ++t)
      if (a && a[t] >= c->primes[t]) return fail(FIDES_ERR_ARG, "constant residue not reduced");
    for (u32 p = 0; p < 2; ++p)
      for (u32 t = 0; t < l[i]; ++t) {
        LcTask tk;
        tk.out = c->limb(out[i], (size_t)p * l[i] + t);
        tk.add = (a && p == 0) ? a[t] : 0;
        tk.pid = t;
        tk.cnt = cnt[i];
        tk.first = (u32)terms.size();
        tk.pad = 0;
        for (u32 j = 0; j < cnt[i]; ++j)
          terms.push_back(
              LcTerm{c->limb(x[o + j], (size_t)p * x_l[o + j] + t), to_mont_host(m_res[o + j][t], c->primes[t])});
        tasks.push_back(tk);
      }
    o += cnt[i];
  }
  if (tasks.empty()) return FIDES_OK;
  hipStream_t st = (hipStream_t)stream;
  Blob blob;
  const size_t ot = blob.add(tasks);
  const size_t om = blob.add(terms);
  char* dev;
  int slot;
  RCHK(upload_blob(c, blob, st, &dev, &slot));
  for (u32 x0 = 0; x0 < tasks.size(); x0 += 65535)
    fides_launch_lincomb_batch(at<LcTask>(dev, ot) + x0, at<LcTerm>(dev, om),
                               std::min<u32>(65535, (u32)tasks.size() - x0), c->prm, c->logN, st);
  RCHK(finish_slot(c, slot, st));
  HIPCHK(hipGetLastError());
  return FIDES_OK;
}
