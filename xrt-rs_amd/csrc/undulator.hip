// Undulator intensity map (src/sources/undulator.rs:120-244, build_i_map;
// src/gpu/undulator.rs + shaders/undulator.wgsl, undulator_gpu) on gfx950.
//
// For each observation point (E, theta, psi) the single-period radiation
// integral over n_steps midpoint samples of the electron trajectory, times the
// N-period sum.  Everything that depends only on the step (velocity,
// position and longitudinal wiggle of the electron) is evaluated once on the
// host with the reference's expressions and read by every point through
// scalar loads; each point evaluates its retarded phase and one FP64
// sin/cos pair per step.  FP64 throughout with contraction off, so the map
// matches the f64 CPU path to rounding (the reference's shader runs in f32).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "runtime.h"

#ifdef __HIP_DEVICE_COMPILE__
#define XRTK_UCONST __attribute__((address_space(4)))
#else
#define XRTK_UCONST
#endif

namespace xrtk {
namespace {

constexpr double kUPi = 3.14159265358979323846;
constexpr double kUPi2 = 2.0 * kUPi;
constexpr double kSIE0 = 1.602176565e-19;        // src/core/consts.rs:18
constexpr double kSIC = 2.99792458e10 * 1e-2;     // consts.rs:21,51
constexpr double kFineStr = 1.0 / 137.03599976;   // consts.rs:54
constexpr double kE2W = 1519267514747457.9;       // consts.rs:57

// Per-map constants.
struct UndConst {
  double q;           // 1 - mean beta: (1 + K^2/2) / (2 gamma^2)
  double gamma2;
  double dt;          // 1 / n_steps
  double lambda_u;    // period [m]
  double amp2flux;    // FINE_STR I / e * N
  double n_periods;
  uint32_t n_steps, pad;
};

// Step table: 6 doubles per step (beta_x, beta_z, x_e, z_e, s, wiggle).
constexpr int kTabW = 6;

__device__ __forceinline__ double sinc_d(double x) {
#pragma clang fp contract(off)
  if (fabs(x) < 0.1) {
    const double x2 = x * x;
    return 1.0 - x2 / 6.0 * (1.0 - x2 / 20.0 * (1.0 - x2 / 42.0));
  }
  return sin(x) / x;
}

// One thread per point; out3: intensity[n], amp_s[2n], amp_p[2n].
__global__ __launch_bounds__(256) void k_undulator(UndConst K, const double* __restrict__ tab,
                                                   const double* __restrict__ e_in,
                                                   const double* __restrict__ th_in,
                                                   const double* __restrict__ ps_in, uint32_t n,
                                                   double* __restrict__ inten,
                                                   double2* __restrict__ amp_s,
                                                   double2* __restrict__ amp_p) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double e = e_in[i], theta = th_in[i], psi = ps_in[i];
  const double omega = e * kE2W;
  const double off_axis = theta * theta + psi * psi;
  const double dirz = sqrt(1.0 - off_axis);
  const double wc = omega / (kSIC * (1.0 - K.q));
  const double slip = off_axis / (1.0 + dirz) + K.q * dirz;
  const XRTK_UCONST double* T = (const XRTK_UCONST double*)tab;
  double axr = 0.0, axi = 0.0, azr = 0.0, azi = 0.0;
  const double dt = K.dt;
  for (uint32_t j = 0; j < K.n_steps; ++j) {
    const double beta_x = T[kTabW * j], beta_z = T[kTabW * j + 1];
    const double x_e = T[kTabW * j + 2], z_e = T[kTabW * j + 3];
    const double s = T[kTabW * j + 4], wiggle = T[kTabW * j + 5];
    const double tp = wc * (s * slip - dirz * wiggle - x_e * theta - z_e * psi);
    double sn, c;
    sincos(tp, &sn, &c);
    const double bx = beta_x - theta, bz = beta_z - psi;
    axr += bx * c * dt;
    axi += bx * sn * dt;
    azr += bz * c * dt;
    azi += bz * sn * dt;
  }
  // N-period sum: exp(i (N-1) d) N sinc(N d) / sinc(d), half-phase folded
  // onto +-pi/2 (round() rounds half away from zero, as f64::round)
  const double nper = K.n_periods;
  const double half = 0.5 * (wc * K.lambda_u * slip);
  const double d = half - kUPi * round(half / kUPi);
  const double arg = (nper - 1.0) * d;
  const double m = nper * sinc_d(nper * d) / sinc_d(d);
  double sa, ca;
  sincos(arg, &sa, &ca);
  const double rr = ca * m, ri = sa * m;
  const double xr = (axr * rr - axi * ri) * K.gamma2, xi = (axr * ri + axi * rr) * K.gamma2;
  const double zr = (azr * rr - azi * ri) * K.gamma2, zi = (azr * ri + azi * rr) * K.gamma2;
  const double is_val = xr * xr - xi * -xi;
  const double ip_val = zr * zr - zi * -zi;
  const double inv_e = 1.0 / e;
  inten[i] = K.amp2flux * inv_e * (is_val + ip_val);
  const double sf = sqrt(K.amp2flux * inv_e);
  amp_s[i] = make_double2(xr * sf, xi * sf);
  amp_p[i] = make_double2(zr * sf, zi * sf);
}

// GpuObsPoint (f32 x 4) -> f64 columns.
__global__ void k_obs_from_f32(const float4* __restrict__ obs, uint32_t n, double* __restrict__ e,
                               double* __restrict__ th, double* __restrict__ ps) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 o = obs[i];
  e[i] = o.x;
  th[i] = o.y;
  ps[i] = o.z;
}

struct DeviceScopeU {
  int prev = -1;
  explicit DeviceScopeU(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceScopeU() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Step table and constants with the reference's expressions
// (undulator.rs:133-147,164-195).
void build_tables(const XrtkUndulator& u, UndConst* K, std::vector<double>& tab) {
  const double gamma = u.gamma, gamma2 = u.gamma2;
  const double lambda_u = u.period_mm * 1e-3;
  const double phase_rad = u.phase_deg * (kUPi / 180.0);
  const double k2 = u.kx * u.kx + u.ky * u.ky;
  const double q = (1.0 + 0.5 * k2) / (2.0 * gamma2);
  const double wiggle_amp = lambda_u / (8.0 * kUPi2 * gamma2);
  const uint32_t ns = u.n_steps;
  const double dt = 1.0 / (double)ns;
  tab.assign((size_t)kTabW * ns, 0.0);
  for (uint32_t j = 0; j < ns; ++j) {
    const double t = ((double)j + 0.5) * dt;
    const double phi_t = kUPi2 * t;
    const double beta_x = u.ky / gamma * std::sin(phi_t);
    const double beta_z = -u.kx / gamma * std::sin(phi_t + phase_rad);
    const double amp_x = u.ky * lambda_u / (kUPi2 * gamma);
    const double amp_z = u.kx * lambda_u / (kUPi2 * gamma);
    const double x_e = -amp_x * std::cos(phi_t);
    const double z_e = amp_z * std::cos(phi_t + phase_rad);
    const double s = lambda_u * t;
    const double wiggle = wiggle_amp * (u.ky * u.ky * std::sin(2.0 * phi_t) +
                                        u.kx * u.kx * std::sin(2.0 * phi_t + 2.0 * phase_rad));
    double* r = &tab[(size_t)kTabW * j];
    r[0] = beta_x; r[1] = beta_z; r[2] = x_e; r[3] = z_e; r[4] = s; r[5] = wiggle;
  }
  K->q = q;
  K->gamma2 = gamma2;
  K->dt = dt;
  K->lambda_u = lambda_u;
  K->amp2flux = kFineStr * u.beam_current / kSIE0 * (double)u.n_periods;
  K->n_periods = (double)u.n_periods;
  K->n_steps = ns;
  K->pad = 0;
}

// Runs the map on columns already on the device (d_e, d_th, d_ps); results
// copied to the host buffers, through `pin` (a page-locked block of at least
// 5 n + 1 doubles, or null) when they are large.
int run_map(xrtk_ctx* ctx, const XrtkUndulator& u, size_t n, const double* d_e,
            const double* d_th, const double* d_ps, char* work, double* inten, double* amp_s2,
            double* amp_p2, char* pin = nullptr) {
  hipStream_t st = ctx->stream;
  UndConst K;
  std::vector<double> tab;
  build_tables(u, &K, tab);
  double* d_tab = reinterpret_cast<double*>(work);
  double* d_int = d_tab + tab.size();
  double2* d_s = reinterpret_cast<double2*>(d_int + ((n + 1) / 2) * 2);
  double2* d_p = d_s + n;
  hipError_t e = hipMemcpyAsync(d_tab, tab.data(), sizeof(double) * tab.size(),
                                hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return map_err(e);
  (void)hipEventRecord(ctx->ev[0], st);
  hipLaunchKernelGGL(k_undulator, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, K, d_tab,
                     d_e, d_th, d_ps, (uint32_t)n, d_int, d_s, d_p);
  e = hipGetLastError();
  if (e != hipSuccess) return map_err(e);
  (void)hipEventRecord(ctx->ev[3], st);
  // the three results lie back to back on the device (intensity padded to
  // an even count)
  const Segment segs[4] = {{inten, sizeof(double) * n},
                           {nullptr, sizeof(double) * (((n + 1) / 2) * 2 - n)},
                           {amp_s2, sizeof(double2) * n},
                           {amp_p2, sizeof(double2) * n}};
  e = copy_d2h(st, pin, d_int, segs, 4);
  if (e != hipSuccess) return map_err(e);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]) != hipSuccess) ms = 0.f;
  ctx->last_pair_ms = ms;
  ctx->last_solve_ms = ms;
  return XRTK_OK;
}

size_t map_bytes(const XrtkUndulator& u, size_t n) {
  // inputs (3 columns) + step table + intensity + two complex outputs
  return sizeof(double) * (3 * n + kTabW * (size_t)u.n_steps + 2 * ((n + 1) / 2) + 4 * n) + 1024;
}

bool valid(const XrtkUndulator* u) {
  return u && u->n_steps > 0 && u->n_steps <= 65536 && u->n_periods > 0;
}

}  // namespace
}  // namespace xrtk

using namespace xrtk;

extern "C" int xrtk_undulator_i_map(xrtk_ctx* ctx, const XrtkUndulator* und,
                                    const double* energies, const double* thetas,
                                    const double* psis, size_t n, double* intensity,
                                    double* amp_s2, double* amp_p2) {
  if (!ctx) return XRTK_ENODEV;
  if (!valid(und) || n > 0x7fffffffull) return XRTK_EINVAL;
  if (!n) return XRTK_OK;
  if (!energies || !thetas || !psis || !intensity || !amp_s2 || !amp_p2) return XRTK_EINVAL;
  try {
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceScopeU scope(ctx->device);
    DBuf& W = ctx->undulator;
    const size_t in_bytes = (sizeof(double) * 3 * n + 255) / 256 * 256;
    if (W.ensure(in_bytes + map_bytes(*und, n)) != hipSuccess) return XRTK_ENOMEM;
    double* d_e = W.as<double>();
    double* d_th = d_e + n;
    double* d_ps = d_th + n;
    hipStream_t st = ctx->stream;
    // inputs and results through the context's page-locked block (3 n then
    // 5 n (+ 1 pad) doubles; the map's synchronisation frees it for the next
    // call)
    char* pin = nullptr;
    if (ctx->ws.hpin.ensure(sizeof(double) * (5 * n + 1)) == hipSuccess)
      pin = static_cast<char*>(ctx->ws.hpin.p);
    hipError_t e = copy_h2d(st, pin, d_e, energies, 8 * n);
    if (e == hipSuccess) e = copy_h2d(st, pin ? pin + 8 * n : nullptr, d_th, thetas, 8 * n);
    if (e == hipSuccess) e = copy_h2d(st, pin ? pin + 16 * n : nullptr, d_ps, psis, 8 * n);
    if (e != hipSuccess) return map_err(e);
    return run_map(ctx, *und, n, d_e, d_th, d_ps, W.as<char>() + in_bytes, intensity, amp_s2,
                   amp_p2, pin);
  } catch (...) {
    return XRTK_ENOMEM;
  }
}

extern "C" int xrtk_undulator_gpu_f32(xrtk_ctx* ctx, const XrtkObsPoint* obs, size_t n,
                                      double kx, double ky, double period_mm, size_t n_periods,
                                      double gamma, double beam_current, uint32_t n_steps,
                                      double phase_deg, double* out5) {
  if (!ctx) return XRTK_ENODEV;
  XrtkUndulator u;
  u.gamma = gamma;
  u.gamma2 = gamma * gamma;
  u.beam_current = beam_current;
  u.kx = kx;
  u.ky = ky;
  u.period_mm = period_mm;
  u.phase_deg = phase_deg;
  u.n_periods = (uint32_t)n_periods;
  u.n_steps = n_steps;
  if (!valid(&u) || n_periods > 0xffffffffull || n > 0x7fffffffull) return XRTK_EINVAL;
  if (!n) return XRTK_OK;
  if (!obs || !out5) return XRTK_EINVAL;
  try {
    std::vector<double> inten(n), as(2 * n), ap(2 * n);
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      DeviceScopeU scope(ctx->device);
      DBuf& W = ctx->undulator;
      const size_t obs_bytes = (sizeof(float4) * n + 255) / 256 * 256;
      const size_t in_bytes = (sizeof(double) * 3 * n + 255) / 256 * 256;
      if (W.ensure(obs_bytes + in_bytes + map_bytes(u, n)) != hipSuccess) return XRTK_ENOMEM;
      char* base = W.as<char>();
      float4* d_obs = reinterpret_cast<float4*>(base);
      double* d_e = reinterpret_cast<double*>(base + obs_bytes);
      double* d_th = d_e + n;
      double* d_ps = d_th + n;
      hipStream_t st = ctx->stream;
      hipError_t e = hipMemcpyAsync(d_obs, obs, sizeof(float4) * n, hipMemcpyHostToDevice, st);
      if (e != hipSuccess) return map_err(e);
      hipLaunchKernelGGL(k_obs_from_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                         d_obs, (uint32_t)n, d_e, d_th, d_ps);
      const int rc = run_map(ctx, u, n, d_e, d_th, d_ps, base + obs_bytes + in_bytes,
                             inten.data(), as.data(), ap.data());
      if (rc != XRTK_OK) return rc;
    }
    for (size_t i = 0; i < n; ++i) {
      out5[5 * i] = as[2 * i];
      out5[5 * i + 1] = as[2 * i + 1];
      out5[5 * i + 2] = ap[2 * i];
      out5[5 * i + 3] = ap[2 * i + 1];
      out5[5 * i + 4] = inten[i];
    }
    return XRTK_OK;
  } catch (...) {
    return XRTK_ENOMEM;
  }
}
