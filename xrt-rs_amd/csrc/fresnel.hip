// Fresnel propagation (src/waves/fresnel.rs:53-106) on gfx950.
//
//   U(x'_ix, z'_iz) = sum_j Es_j exp(i K ((x'_ix - x_j)^2 + (z'_iz - z_j)^2)),
//   K = k / (2 dist), k = E / CHBAR * 1e7.
//
// The kernel factorises exactly, exp(i K (dx^2 + dz^2)) = exp(i K dx^2)
// exp(i K dz^2), and the screen is a grid, so
//
//   U = X Z'^T,  X[ix][j] = exp(i K (x'_ix - x_j)^2),
//                Z'[iz][j] = Es_j exp(i K (z'_iz - z_j)^2):
//
// a complex matrix product over the rays.  The two factor matrices
// ((nx + nz) n_rays entries, phases reduced and evaluated in FP64) are built
// by k_fresnel_factor; the product, nx nz n_rays complex multiply-adds, runs
// on the matrix cores through rocBLAS (cgemm with FP32 factors, or zgemm with
// FP64 factors), chunk by chunk over the rays; k_accum_c64 adds each FP32
// chunk product into the FP64 result.  Output index iz * nx + ix: the
// column-major nx x nz result.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "runtime.h"

namespace xrtk {
namespace {

constexpr double kPi = 3.14159265358979323846;

// out[i + j * n] = (amp_j or 1) * exp(i 2 pi krev (p_i - r_j)^2), i < n, j < nc.
template <class T>
__global__ __launch_bounds__(256) void k_fresnel_factor(const double* __restrict__ p, uint32_t n,
                                                        const double* __restrict__ r,
                                                        const double2* __restrict__ amp,
                                                        uint32_t nc, double krev,
                                                        T* __restrict__ out) {
  const size_t total = (size_t)n * nc;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const uint32_t i = (uint32_t)(e % n), j = (uint32_t)(e / n);
    double s, c;
    {
#pragma clang fp contract(off)
      const double d = p[i] - r[j];
      const double ph = krev * (d * d);  // revolutions
      sincospi(2.0 * (ph - rint(ph)), &s, &c);
    }
    if (amp) {
      const double2 a = amp[j];
      const double re = a.x * c - a.y * s, im = a.x * s + a.y * c;
      c = re;
      s = im;
    }
    T v;
    v.x = c;
    v.y = s;
    out[e] = v;
  }
}

// acc[i] += c[i] (FP32 chunk product into the FP64 result).
__global__ __launch_bounds__(256) void k_accum_c64(const float2* __restrict__ c, size_t n,
                                                   double2* __restrict__ acc) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float2 v = c[i];
    double2 a = acc[i];
    a.x += (double)v.x;
    a.y += (double)v.y;
    acc[i] = a;
  }
}

__global__ void k_zero_c64(double2* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = make_double2(0.0, 0.0);
}

unsigned grid_for(size_t n) {
  return (unsigned)std::min<size_t>((n + 255) / 256, 256u * 64u);
}

// rocBLAS, loaded on first use (the Kirchhoff path never needs it).
struct Blas {
  using create_t = int (*)(void**);
  using destroy_t = int (*)(void*);
  using stream_t = int (*)(void*, hipStream_t);
  using cgemm_t = int (*)(void*, int, int, int, int, int, const float2*, const float2*, int,
                          const float2*, int, const float2*, float2*, int);
  using zgemm_t = int (*)(void*, int, int, int, int, int, const double2*, const double2*, int,
                          const double2*, int, const double2*, double2*, int);
  void* lib = nullptr;
  create_t create = nullptr;
  destroy_t destroy = nullptr;
  stream_t set_stream = nullptr;
  cgemm_t cgemm = nullptr;
  zgemm_t zgemm = nullptr;
  bool ok = false;
};
constexpr int kOpNone = 111, kOpTrans = 112;  // rocblas_operation

Blas& blas() {
  static Blas b;
  static std::once_flag once;
  std::call_once(once, [] {
    for (const char* name : {"librocblas.so.5", "librocblas.so", "/opt/rocm/lib/librocblas.so"}) {
      b.lib = dlopen(name, RTLD_NOW | RTLD_LOCAL);
      if (b.lib) break;
    }
    if (!b.lib) return;
    b.create = reinterpret_cast<Blas::create_t>(dlsym(b.lib, "rocblas_create_handle"));
    b.destroy = reinterpret_cast<Blas::destroy_t>(dlsym(b.lib, "rocblas_destroy_handle"));
    b.set_stream = reinterpret_cast<Blas::stream_t>(dlsym(b.lib, "rocblas_set_stream"));
    b.cgemm = reinterpret_cast<Blas::cgemm_t>(dlsym(b.lib, "rocblas_cgemm"));
    b.zgemm = reinterpret_cast<Blas::zgemm_t>(dlsym(b.lib, "rocblas_zgemm"));
    b.ok = b.create && b.destroy && b.set_stream && b.cgemm && b.zgemm;
  });
  return b;
}

}  // namespace

void fresnel_release(xrtk_ctx* ctx) {
  if (ctx->blas) {
    Blas& b = blas();
    if (b.ok) (void)b.destroy(ctx->blas);
    ctx->blas = nullptr;
  }
  ctx->fresnel.release();
}

}  // namespace xrtk

using namespace xrtk;

namespace {

struct DeviceScope {
  int prev = -1;
  explicit DeviceScope(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceScope() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Screen axis of src/waves/fresnel.rs:66-73.
void fresnel_axis(double lo, double hi, uint32_t n, double* v) {
  const double d = (double)(n > 1 ? n - 1 : 1);
  for (uint32_t i = 0; i < n; ++i) v[i] = lo + (hi - lo) * (double)i / d;
}

}  // namespace

extern "C" int xrtk_fresnel_propagate(xrtk_ctx* ctx, const double* ray_x, const double* ray_z,
                                      const double* es_re, const double* es_im, size_t n_rays,
                                      double energy, double dist,
                                      const XrtkFresnelScreen* screen, double* out2,
                                      double* x_pixels, double* z_pixels,
                                      const XrtkOpts* opts) {
  if (!ctx) return XRTK_ENODEV;
  if (!screen || !out2 || screen->nx == 0 || screen->nz == 0) return XRTK_EINVAL;
  if (n_rays && (!ray_x || !ray_z || !es_re || !es_im)) return XRTK_EINVAL;
  const uint32_t nx = screen->nx, nz = screen->nz;
  if ((uint64_t)nx * nz > (1ull << 31) || n_rays > 0x7fffffffull) return XRTK_EINVAL;
  const bool f64 = opts && opts->mode == XRTK_MODE_F64;
  try {
    std::vector<double> xp(nx), zp(nz);
    fresnel_axis(screen->x_min, screen->x_max, nx, xp.data());
    fresnel_axis(screen->z_min, screen->z_max, nz, zp.data());
    if (x_pixels) std::memcpy(x_pixels, xp.data(), sizeof(double) * nx);
    if (z_pixels) std::memcpy(z_pixels, zp.data(), sizeof(double) * nz);
    const size_t np = (size_t)nx * nz;
    if (n_rays == 0) {
      std::memset(out2, 0, sizeof(double) * 2 * np);
      return XRTK_OK;
    }
    Blas& b = blas();
    if (!b.ok) return XRTK_ELAUNCH;
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceScope scope(ctx->device);
    hipStream_t st = ctx->stream;
    if (!ctx->blas) {
      if (b.create(&ctx->blas) != 0) { ctx->blas = nullptr; return XRTK_ELAUNCH; }
    }
    if (b.set_stream(ctx->blas, st) != 0) return XRTK_ELAUNCH;

    // ray chunk: each factor block at most ~256 MiB
    const size_t esz = f64 ? sizeof(double2) : sizeof(float2);
    const size_t cap = (256ull << 20) / (esz * std::max(nx, nz));
    size_t nc = std::min<size_t>(n_rays, std::max<size_t>(256, cap / 256 * 256));
    // device buffers: pixel axes, rays (x, z, Es), factors, chunk product, result
    DBuf& W = ctx->fresnel;
    const size_t o_xp = 0, o_zp = o_xp + 8 * nx, o_rx = (o_zp + 8 * nz + 255) / 256 * 256;
    const size_t o_rz = o_rx + 8 * n_rays, o_es = (o_rz + 8 * n_rays + 255) / 256 * 256;
    const size_t o_fx = (o_es + 16 * n_rays + 255) / 256 * 256;
    const size_t o_fz = (o_fx + esz * nx * nc + 255) / 256 * 256;
    const size_t o_c = (o_fz + esz * nz * nc + 255) / 256 * 256;
    const size_t o_u = (o_c + (f64 ? 0 : sizeof(float2) * np) + 255) / 256 * 256;
    const size_t total = o_u + sizeof(double2) * np;
    if (W.ensure(total) != hipSuccess) return XRTK_ENOMEM;
    char* base = W.as<char>();
    std::vector<double> es(2 * n_rays);
    for (size_t j = 0; j < n_rays; ++j) { es[2 * j] = es_re[j]; es[2 * j + 1] = es_im[j]; }
    hipError_t e = hipMemcpyAsync(base + o_xp, xp.data(), 8 * nx, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(base + o_zp, zp.data(), 8 * nz, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(base + o_rx, ray_x, 8 * n_rays, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(base + o_rz, ray_z, 8 * n_rays, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(base + o_es, es.data(), 16 * n_rays, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return map_err(e);
    const double k = energy / kCHBAR_H * 1e7;
    const double krev = k / (2.0 * dist) / (2.0 * kPi);
    const double* dxp = reinterpret_cast<const double*>(base + o_xp);
    const double* dzp = reinterpret_cast<const double*>(base + o_zp);
    const double* drx = reinterpret_cast<const double*>(base + o_rx);
    const double* drz = reinterpret_cast<const double*>(base + o_rz);
    const double2* des = reinterpret_cast<const double2*>(base + o_es);
    double2* U = reinterpret_cast<double2*>(base + o_u);
    hipLaunchKernelGGL(k_zero_c64, dim3(grid_for(np)), dim3(256), 0, st, U, np);
    (void)hipEventRecord(ctx->ev[0], st);
    for (size_t j0 = 0; j0 < n_rays; j0 += nc) {
      const uint32_t c = (uint32_t)std::min<size_t>(nc, n_rays - j0);
      if (f64) {
        double2* FX = reinterpret_cast<double2*>(base + o_fx);
        double2* FZ = reinterpret_cast<double2*>(base + o_fz);
        hipLaunchKernelGGL(k_fresnel_factor<double2>, dim3(grid_for((size_t)nx * c)), dim3(256),
                           0, st, dxp, nx, drx + j0, nullptr, c, krev, FX);
        hipLaunchKernelGGL(k_fresnel_factor<double2>, dim3(grid_for((size_t)nz * c)), dim3(256),
                           0, st, dzp, nz, drz + j0, des + j0, c, krev, FZ);
        const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0);
        if (b.zgemm(ctx->blas, kOpNone, kOpTrans, (int)nx, (int)nz, (int)c, &one, FX, (int)nx,
                    FZ, (int)nz, j0 ? &one : &zero, U, (int)nx) != 0)
          return XRTK_ELAUNCH;
      } else {
        float2* FX = reinterpret_cast<float2*>(base + o_fx);
        float2* FZ = reinterpret_cast<float2*>(base + o_fz);
        float2* Cc = reinterpret_cast<float2*>(base + o_c);
        hipLaunchKernelGGL(k_fresnel_factor<float2>, dim3(grid_for((size_t)nx * c)), dim3(256),
                           0, st, dxp, nx, drx + j0, nullptr, c, krev, FX);
        hipLaunchKernelGGL(k_fresnel_factor<float2>, dim3(grid_for((size_t)nz * c)), dim3(256),
                           0, st, dzp, nz, drz + j0, des + j0, c, krev, FZ);
        const float2 one = make_float2(1.f, 0.f), zero = make_float2(0.f, 0.f);
        if (b.cgemm(ctx->blas, kOpNone, kOpTrans, (int)nx, (int)nz, (int)c, &one, FX, (int)nx,
                    FZ, (int)nz, &zero, Cc, (int)nx) != 0)
          return XRTK_ELAUNCH;
        hipLaunchKernelGGL(k_accum_c64, dim3(grid_for(np)), dim3(256), 0, st, Cc, np, U);
      }
    }
    (void)hipEventRecord(ctx->ev[3], st);
    e = hipMemcpyAsync(out2, U, sizeof(double2) * np, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return map_err(e);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]) != hipSuccess) ms = 0.f;
    ctx->last_pair_ms = ms;
    ctx->last_solve_ms = ms;
    return XRTK_OK;
  } catch (...) {
    return XRTK_ENOMEM;
  }
}
