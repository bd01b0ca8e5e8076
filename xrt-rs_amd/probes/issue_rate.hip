// VALU issue rates of gfx950: FP32 FMA, v_sin_f32, v_cos_f32, v_rsq_f32 and
// v_rcp_f32 in independent chains at full occupancy (8 waves per SIMD).  The
// pair kernels' rooflines are read against these rates.
//
// Every launch writes one float per thread into a buffer of grid x block
// floats; trip counts are fixed.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kChains = 16;
constexpr int kBlock = 256;

enum Op { OP_FMA, OP_FMAC, OP_SIN, OP_COS, OP_RSQ, OP_RCP };

template <Op OP>
__global__ __launch_bounds__(kBlock) void k_rate(float* out, int iters, float seed) {
  float x[kChains], y[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) {
    x[c] = seed + 1e-3f * (float)(threadIdx.x + 7 * c);
    y[c] = 1e-6f * (float)(c + 1);
  }
  const float z = seed * 1e-3f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      if constexpr (OP == OP_FMA) x[c] = fmaf(x[c], 0.999f, 1e-3f);     // v_fmamk (literal)
      if constexpr (OP == OP_FMAC) x[c] = fmaf(y[c], z, x[c]);           // v_fmac (registers)
      if constexpr (OP == OP_SIN) x[c] = __builtin_amdgcn_sinf(x[c]);
      if constexpr (OP == OP_COS) x[c] = __builtin_amdgcn_cosf(x[c]);
      if constexpr (OP == OP_RSQ) x[c] = __builtin_amdgcn_rsqf(x[c]);
      if constexpr (OP == OP_RCP) x[c] = __builtin_amdgcn_rcpf(x[c]);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  out[blockIdx.x * kBlock + threadIdx.x] = s;
}

template <Op OP>
int run(const char* name, float* d, int blocks, int iters, int n_cu, double ghz) {
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1;
  hipLaunchKernelGGL(k_rate<OP>, dim3(blocks), dim3(kBlock), 0, 0, d, 8, 0.25f);
  (void)hipEventRecord(a);
  hipLaunchKernelGGL(k_rate<OP>, dim3(blocks), dim3(kBlock), 0, 0, d, iters, 0.25f);
  (void)hipEventRecord(b);
  if (hipEventSynchronize(b) != hipSuccess) return 1;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  const double waves_inst = (double)blocks * (kBlock / 64) * (double)iters * kChains;
  const double per_cu_clk = waves_inst / (ms * 1e-3) / n_cu / (ghz * 1e9);
  std::printf("%-4s %9.3f ms  %.3f wave-instructions per CU per clock (at %.2f GHz)\n", name, ms,
              per_cu_clk, ghz);
  std::fflush(stdout);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return 0;
}

}  // namespace

int main() {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
  const int n_cu = prop.multiProcessorCount;
  const double ghz = prop.clockRate / 1e6;  // clockRate is in kHz
  const int blocks = n_cu * 8, iters = 4000;
  float* d = nullptr;
  if (hipMalloc(&d, sizeof(float) * blocks * kBlock) != hipSuccess) return 1;
  std::printf("%s, %d CUs, %d blocks x %d threads, %d chains x %d iterations (fma: %d)\n",
              prop.gcnArchName, n_cu, blocks, kBlock, kChains, iters, 8 * iters);
  int rc = run<OP_FMA>("fma", d, blocks, 8 * iters, n_cu, ghz);
  rc |= run<OP_FMAC>("fmac", d, blocks, 8 * iters, n_cu, ghz);
  rc |= run<OP_SIN>("sin", d, blocks, iters, n_cu, ghz);
  rc |= run<OP_COS>("cos", d, blocks, iters, n_cu, ghz);
  rc |= run<OP_RSQ>("rsq", d, blocks, iters, n_cu, ghz);
  rc |= run<OP_RCP>("rcp", d, blocks, iters, n_cu, ghz);
  (void)hipFree(d);
  return rc;
}
