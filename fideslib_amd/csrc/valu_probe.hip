// Issue cost of the VALU opcodes the NTT passes execute, on the GPU it runs
// on (measurement support for the NTT roofline; not part of the library).
//
// For each opcode a kernel runs ITERS x 8 independent instances per lane (no
// dependency between consecutive instructions of a lane) with 8 waves per
// SIMD, so issue, not latency, limits it.  The engine clock under that load
// comes from clock64() (shader clock) against wall_clock64() (constant rate)
// sampled by every workgroup, and the cost is reported as cycles per wave64
// instruction per SIMD:  1024 SIMDs x clock x 64 lanes / lane-ops per second.
// A full-rate instruction on a 16-lane SIMD takes 4 cycles.
//
// Output: one JSON object {"device", "clock_ghz", "ops": {opcode: cycles}}.
// Usage: valu_probe [iterations]
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

typedef uint64_t u64;
typedef uint32_t u32;

#define CHECK(x)                                                                \
  do {                                                                          \
    hipError_t e_ = (x);                                                        \
    if (e_ != hipSuccess) {                                                     \
      fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                   \
      exit(1);                                                                  \
    }                                                                           \
  } while (0)

enum Op {
  MAD_U64_U32, MUL_LO_U32, MUL_HI_U32, LSHL_ADD_U64, ADD3_U32, CNDMASK_B32, SUB_CO_U32, SUBB_CO_U32, MOV_B32,
  ADDC_CO_U32, CMP_LE_U64, OR_B32, LSHLREV_B32, MOV_B64, NOPS
};
static const char* NAMES[NOPS] = {"v_mad_u64_u32", "v_mul_lo_u32", "v_mul_hi_u32", "v_lshl_add_u64", "v_add3_u32",
                                  "v_cndmask_b32", "v_sub_co_u32", "v_subb_co_u32", "v_mov_b32", "v_addc_co_u32",
                                  "v_cmp_le_u64", "v_or_b32", "v_lshlrev_b32", "v_mov_b64"};

// Every asm statement below is ALU only: no memory access, no loads.
template <int OP>
__global__ __launch_bounds__(256) void k_rate(u64* out, u64* clk, u32 seed, int iters) {
  u32 a[8];
  u64 b[8];
  for (int i = 0; i < 8; ++i) {
    a[i] = seed * (threadIdx.x + 7 * i + 1) + i;
    b[i] = ((u64)a[i] << 32) | (a[i] ^ 0x9e3779b9u);
  }
  const u32 m = seed | 1u;
  const u64 mask = 0x5555555555555555ull ^ seed;
  const u64 c0 = clock64(), w0 = wall_clock64();
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if constexpr (OP == MAD_U64_U32) {
        u64 c;
        asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(b[i]), "=s"(c) : "v"(a[i]), "v"(m));
      } else if constexpr (OP == MUL_LO_U32) {
        asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a[i]) : "v"(m));
      } else if constexpr (OP == MUL_HI_U32) {
        asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(a[i]) : "v"(m));
      } else if constexpr (OP == LSHL_ADD_U64) {
        asm volatile("v_lshl_add_u64 %0, %0, 1, %1" : "+v"(b[i]) : "v"(b[(i + 1) & 7]));
      } else if constexpr (OP == ADD3_U32) {
        asm volatile("v_add3_u32 %0, %0, %1, %2" : "+v"(a[i]) : "v"(m), "v"(a[(i + 3) & 7]));
      } else if constexpr (OP == CNDMASK_B32) {
        asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[i]) : "v"(m), "s"(mask));
      } else if constexpr (OP == SUB_CO_U32) {
        asm volatile("v_sub_co_u32_e32 %0, vcc, %0, %1" : "+v"(a[i]) : "v"(m) : "vcc");
      } else if constexpr (OP == SUBB_CO_U32) {
        asm volatile("v_subb_co_u32_e32 %0, vcc, %0, %1, vcc" : "+v"(a[i]) : "v"(m) : "vcc");
      } else if constexpr (OP == MOV_B32) {
        asm volatile("v_mov_b32_e32 %0, %1" : "=v"(a[i]) : "v"(a[(i + 1) & 7]));
      } else if constexpr (OP == ADDC_CO_U32) {
        u64 c;
        asm volatile("v_addc_co_u32_e64 %0, %1, %0, 0, %2" : "+v"(a[i]), "=s"(c) : "s"(mask));
      } else if constexpr (OP == CMP_LE_U64) {
        asm volatile("v_cmp_le_u64_e32 vcc, %0, %1" : : "v"(b[i]), "v"(b[(i + 1) & 7]) : "vcc");
      } else if constexpr (OP == OR_B32) {
        asm volatile("v_or_b32_e32 %0, %0, %1" : "+v"(a[i]) : "v"(a[(i + 5) & 7]));
      } else if constexpr (OP == LSHLREV_B32) {
        asm volatile("v_lshlrev_b32_e32 %0, 3, %0" : "+v"(a[i]));
      } else if constexpr (OP == MOV_B64) {
        asm volatile("v_mov_b64_e32 %0, %1" : "=v"(b[i]) : "v"(b[(i + 1) & 7]));
      }
    }
  }
  const u64 c1 = clock64(), w1 = wall_clock64();
  u64 acc = 0;
  for (int i = 0; i < 8; ++i) acc += a[i] + b[i];
  out[blockIdx.x * 256 + threadIdx.x] = acc;
  if (threadIdx.x == 0) {
    clk[blockIdx.x * 2] = c1 - c0;
    clk[blockIdx.x * 2 + 1] = w1 - w0;
  }
}

struct Result {
  double cycles, ghz;
};

template <int OP>
static Result run(u64* out, u64* clk, int blocks, int iters, double wall_mhz, int simds) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  k_rate<OP><<<blocks, 256>>>(out, clk, 12345u, iters);  // warm-up
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(e0));
  k_rate<OP><<<blocks, 256>>>(out, clk, 12345u, iters);
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  std::vector<u64> h(2 * (size_t)blocks);
  CHECK(hipMemcpy(h.data(), clk, 16 * (size_t)blocks, hipMemcpyDeviceToHost));
  double cs = 0, ws = 0;
  for (int i = 0; i < blocks; ++i) cs += (double)h[2 * i], ws += (double)h[2 * i + 1];
  const double ghz = cs / ws * wall_mhz / 1e3;
  const double lane_ops = (double)blocks * 256 * iters * 8;
  const double rate = lane_ops / (ms * 1e-3);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return {(double)simds * ghz * 1e9 * 64 / rate, ghz};
}

template <int OP>
static void run_all(Result* r, u64* out, u64* clk, int blocks, int iters, double wall_mhz, int simds) {
  if constexpr (OP < NOPS) {
    r[OP] = run<OP>(out, clk, blocks, iters, wall_mhz, simds);
    run_all<OP + 1>(r, out, clk, blocks, iters, wall_mhz, simds);
  }
}

int main(int argc, char** argv) {
  const int iters = argc > 1 ? atoi(argv[1]) : 2048;
  if (iters < 1 || iters > (1 << 20)) {
    fprintf(stderr, "iterations must be in [1, 2^20]\n");
    return 2;
  }
  int wall_khz = 0, cus = 0;
  CHECK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, 0));
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  hipDeviceProp_t prop;
  CHECK(hipGetDeviceProperties(&prop, 0));
  const int simds = 4 * cus;
  const int blocks = simds * 8 / 4;  // 8 waves per SIMD, 4 waves per workgroup
  u64 *out = nullptr, *clk = nullptr;
  CHECK(hipMalloc(&out, (size_t)blocks * 256 * 8));
  CHECK(hipMalloc(&clk, (size_t)blocks * 16));
  Result r[NOPS];
  run_all<0>(r, out, clk, blocks, iters, wall_khz / 1000.0, simds);
  double gsum = 0;
  for (int i = 0; i < NOPS; ++i) gsum += r[i].ghz;
  printf("{\"device\": \"%s\", \"compute_units\": %d, \"waves_per_simd\": 8, \"iterations\": %d, "
         "\"clock_ghz\": %.3f, \"ops\": {",
         prop.gcnArchName, cus, iters, gsum / NOPS);
  for (int i = 0; i < NOPS; ++i) printf("%s\"%s\": %.3f", i ? ", " : "", NAMES[i], r[i].cycles);
  printf("}, \"clock_ghz_by_op\": {");
  for (int i = 0; i < NOPS; ++i) printf("%s\"%s\": %.3f", i ? ", " : "", NAMES[i], r[i].ghz);
  printf("}}\n");
  CHECK(hipFree(out));
  CHECK(hipFree(clk));
  return 0;
}
