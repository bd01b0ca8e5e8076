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
// by k_fresnel_factor*; the product, nx nz n_rays complex multiply-adds, runs
// on the matrix cores through rocBLAS, chunk by chunk over the rays:
//   FP64  zgemm on the complex factors;
//   FP32  one real sgemm on the block forms
//           [Xr -Xi; Xi Xr] (2 nx x 2 nc) times [Zr Zi]^T (2 nc x nz)
//         = [Ur; Ui] (2 nx x nz), the same 8 flops per pair, which gfx950's
//         FP32 matrix path runs at about twice the rate of cgemm;
//         k_accum_blk adds each chunk product into the FP64 result.
// Output index iz * nx + ix: the column-major nx x nz result.
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

// Block factors for the real product, FP32 (column-major):
//   kind 0 (X):  A is 2n x 2nc, A[i][j] = A[n+i][nc+j] = Re, A[n+i][j] = Im,
//                A[i][nc+j] = -Im;
//   kind 1 (Z'): B is n x 2nc,  B[i][j] = Re, B[i][nc+j] = Im;
// with Re + i Im = (amp_j or 1) exp(i 2 pi krev (p_i - r_j)^2).
__global__ __launch_bounds__(256) void k_fresnel_factor_blk(const double* __restrict__ p,
                                                            uint32_t n,
                                                            const double* __restrict__ r,
                                                            const double2* __restrict__ amp,
                                                            uint32_t nc, double krev, int kind,
                                                            float* __restrict__ out) {
  const size_t total = (size_t)n * nc;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t i = e % n, j = e / n;
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
    if (kind == 0) {
      const size_t ld = 2 * (size_t)n;
      out[i + j * ld] = (float)c;
      out[i + (nc + j) * ld] = (float)-s;
      out[n + i + j * ld] = (float)s;
      out[n + i + (nc + j) * ld] = (float)c;
    } else {
      out[i + j * (size_t)n] = (float)c;
      out[i + (nc + j) * (size_t)n] = (float)s;
    }
  }
}

// acc[ix + iz nx] += (C[ix + iz 2nx], C[nx + ix + iz 2nx]): the FP32 chunk
// product [Ur; Ui] into the FP64 result.
__global__ __launch_bounds__(256) void k_accum_blk(const float* __restrict__ c, uint32_t nx,
                                                   size_t np, double2* __restrict__ acc) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < np;
       e += (size_t)gridDim.x * blockDim.x) {
    const size_t ix = e % nx, iz = e / nx;
    const float* col = c + iz * 2 * (size_t)nx;
    double2 a = acc[e];
    a.x += (double)col[ix];
    a.y += (double)col[nx + ix];
    acc[e] = a;
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
  using sgemm_t = int (*)(void*, int, int, int, int, int, const float*, const float*, int,
                          const float*, int, const float*, float*, int);
  using zgemm_t = int (*)(void*, int, int, int, int, int, const double2*, const double2*, int,
                          const double2*, int, const double2*, double2*, int);
  void* lib = nullptr;
  create_t create = nullptr;
  destroy_t destroy = nullptr;
  stream_t set_stream = nullptr;
  sgemm_t sgemm = nullptr;
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
    b.sgemm = reinterpret_cast<Blas::sgemm_t>(dlsym(b.lib, "rocblas_sgemm"));
    b.zgemm = reinterpret_cast<Blas::zgemm_t>(dlsym(b.lib, "rocblas_zgemm"));
    b.ok = b.create && b.destroy && b.set_stream && b.sgemm && b.zgemm;
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
  // rocBLAS takes int dimensions; the FP32 product has 2 nx rows and 2 nc columns
  if ((uint64_t)nx * nz > (1ull << 31) || nx > (1u << 29) || nz > (1u << 30) ||
      n_rays > 0x7fffffffull)
    return XRTK_EINVAL;
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

    // ray chunk: each factor block at most ~256 MiB (an FP32 X entry takes
    // four floats in the block form)
    const size_t esz = f64 ? sizeof(double2) : 4 * sizeof(float);
    const size_t cap = (256ull << 20) / (esz * std::max(nx, nz));
    size_t nc = std::min<size_t>(n_rays, std::max<size_t>(256, cap / 256 * 256));
    // device buffers: pixel axes, rays (x, z, Es), factors, chunk product, result
    DBuf& W = ctx->fresnel;
    const size_t o_xp = 0, o_zp = o_xp + 8 * nx, o_rx = (o_zp + 8 * nz + 255) / 256 * 256;
    const size_t o_rz = o_rx + 8 * n_rays, o_es = (o_rz + 8 * n_rays + 255) / 256 * 256;
    const size_t o_fx = (o_es + 16 * n_rays + 255) / 256 * 256;
    const size_t o_fz = (o_fx + esz * nx * nc + 255) / 256 * 256;
    const size_t o_c = (o_fz + esz * nz * nc + 255) / 256 * 256;
    const size_t o_u = (o_c + (f64 ? 0 : 2 * sizeof(float) * np) + 255) / 256 * 256;
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
        float* FX = reinterpret_cast<float*>(base + o_fx);
        float* FZ = reinterpret_cast<float*>(base + o_fz);
        float* Cc = reinterpret_cast<float*>(base + o_c);
        hipLaunchKernelGGL(k_fresnel_factor_blk, dim3(grid_for((size_t)nx * c)), dim3(256), 0,
                           st, dxp, nx, drx + j0, nullptr, c, krev, 0, FX);
        hipLaunchKernelGGL(k_fresnel_factor_blk, dim3(grid_for((size_t)nz * c)), dim3(256), 0,
                           st, dzp, nz, drz + j0, des + j0, c, krev, 1, FZ);
        const float one = 1.f, zero = 0.f;
        if (b.sgemm(ctx->blas, kOpNone, kOpTrans, 2 * (int)nx, (int)nz, 2 * (int)c, &one, FX,
                    2 * (int)nx, FZ, (int)nz, &zero, Cc, 2 * (int)nx) != 0)
          return XRTK_ELAUNCH;
        hipLaunchKernelGGL(k_accum_blk, dim3(grid_for(np)), dim3(256), 0, st, Cc, nx, np, U);
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
