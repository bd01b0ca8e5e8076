// gfx950 kernels of the Kirchhoff diffraction engine.
//
// The physics is the per-pair term of src/waves/diffraction.rs:75-111
//   U = i k/(4 pi) (nl + n.d/path) exp(i k path) / path,  Es += es U, Ep += ep U
// evaluated two ways:
//   * k_pair_f64   plain FP64 with Kahan sums, the operation order of
//                  diffraction.rs:76-110 (the FP64 device oracle and the FP64
//                  speed baseline);
//   * k_pair_f32r  FP32 with the phase reduced against the FP64 ray centroid
//                  c — the method of src/gpu/kirchhoff.rs:96-104,276-297 and
//                  src/gpu/shaders/kirchhoff.wgsl:95-113, redesigned:
//                  - o = r - c and q = p - c are formed in FP64, rounded once;
//                  - phases are in revolutions for the hardware v_sin/v_cos;
//                  - |p - r|^2 = |q|^2 + kr^-1 * kr(|o|^2 - 2 o.q) reuses the
//                    phase numerator (no difference vector per pair);
//                  - TILE variant: each wave's 64*P pixels are one tile with
//                    an FP64 reference point q_t; the per-(ray, tile) phase
//                    kr(|o|^2 - 2 o.q_t)/(2|q_t|) is folded into the ray
//                    amplitudes in FP64 (lane-parallel, staged in LDS), so the
//                    FP32 pair phase only carries the small in-tile residual;
//                  - FP32 partial sums are folded into FP64 every `tile` rays.
// Work is split into items (instance, block of pixel slots, ray chunk); each
// chunk's FP64 partial goes to HBM and k_merge sums the chunks in a fixed
// order with TwoSum, then applies k/(4 pi) exp(i k |p - c|) in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "internal.h"

namespace xrtk {

// Wave-uniform reads through the constant address space become scalar loads
// (SMEM) into SGPRs.  The host pass of the single-source compile only parses
// device code, so the qualifier is dropped there.
#if defined(__HIP_DEVICE_COMPILE__)
#define XRTK_CONST __attribute__((address_space(4)))
#else
#define XRTK_CONST
#endif

static constexpr double kCHBAR = 1973.2697177417986;  // src/core/consts.rs:74
static constexpr double kPI = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float vcopy(float s) {
  // Copies a wave-uniform value into a VGPR once, so the VALU ops that use it
  // together with another scalar operand do not each re-materialise it.
  float v;
  asm("v_mov_b32 %0, %1" : "=v"(v) : "s"(s));
  return v;
}

__device__ __forceinline__ uint32_t uniform(uint32_t x) {
  return __builtin_amdgcn_readfirstlane(x);
}

// Workgroup i runs on XCD i % 8.  Maps launch order to item order so that
// each XCD receives a contiguous run of items (items of one ray chunk share
// that XCD's L2).  Bijective on [0, n).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t n) {
  const uint32_t xcd = bid & 7u, local = bid >> 3;
  const uint32_t per = n >> 3, rem = n & 7u;
  return xcd < rem ? xcd * (per + 1u) + local
                   : rem * (per + 1u) + (xcd - rem) * per + local;
}

// exp(i 2 pi (a_hi + a_lo) * x) for a large phase: the product is formed as a
// double-double so the fractional revolution keeps full FP64 accuracy.
__device__ __forceinline__ void cis_rev_dd(double a_hi, double a_lo, double x,
                                           double* c, double* s) {
  #pragma clang fp contract(off)
  const double hi = a_hi * x;
  const double lo = fma(a_hi, x, -hi) + a_lo * x;
  const double n = rint(hi);
  const double f = (hi - n) + lo;  // |f| <= 0.5 + tiny
  sincospi(2.0 * f, s, c);
}

// Double-double helpers (Knuth TwoSum, FMA TwoProduct).
__device__ __forceinline__ void two_sum(double a, double b, double* s, double* e) {
  #pragma clang fp contract(off)
  *s = a + b;
  const double bp = *s - a;
  *e = (a - (*s - bp)) + (b - bp);
}
__device__ __forceinline__ void two_prod(double a, double b, double* p, double* e) {
  #pragma clang fp contract(off)
  *p = a * b;
  *e = fma(a, b, -*p);
}

// |q| of q = (x_hi + x_lo, ...) as a double-double.
__device__ __forceinline__ void norm_dd(const double (&h)[3], const double (&l)[3], double* rh,
                                        double* rl) {
  #pragma clang fp contract(off)
  double sh = 0.0, sl = 0.0;
  for (int i = 0; i < 3; ++i) {
    double p, e;
    two_prod(h[i], h[i], &p, &e);
    e += 2.0 * h[i] * l[i];
    double t, te;
    two_sum(sh, p, &t, &te);
    sl += te + e;
    sh = t;
  }
  double s2, e2;
  two_sum(sh, sl, &s2, &e2);  // renormalise
  const double r0 = sqrt(s2);
  double p, e;
  two_prod(r0, r0, &p, &e);
  const double corr = r0 > 0.0 ? ((s2 - p) - e + e2) / (2.0 * r0) : 0.0;
  two_sum(r0, corr, rh, rl);
}

// exp(i 2 pi (a_hi + a_lo)(x_hi + x_lo)) for a large phase.
__device__ __forceinline__ void cis_rev_dd2(double a_hi, double a_lo, double x_hi, double x_lo,
                                            double* c, double* s) {
  #pragma clang fp contract(off)
  double hi, e;
  two_prod(a_hi, x_hi, &hi, &e);
  const double lo = e + a_hi * x_lo + a_lo * x_hi;
  const double f = (hi - rint(hi)) + lo;
  sincospi(2.0 * f, s, c);
}

__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_maxd(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_mind(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// Axis order of the FP32 records (o, n, m) and of the tile reference points:
// an axis along which every ray has o = r - c == 0 goes last, so that the
// LIN2 variant of the pair kernel can drop it from o.m (fma(0, m, x) == x,
// so the result is the same).  Read after k_ray_bounds has reduced |o|.
template <class T>
__device__ __forceinline__ T pick3(const T (&v)[3], int a) {
  return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]);
}
// Order-preserving unsigned keys of floats (for atomicMin / atomicMax).
__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
__device__ __forceinline__ void kernel_axes(const InstScratch* scr, int (&ax)[3]) {
  if (scr->omax_bits[2] == 0u)      { ax[0] = 0; ax[1] = 1; ax[2] = 2; }
  else if (scr->omax_bits[1] == 0u) { ax[0] = 0; ax[1] = 2; ax[2] = 1; }
  else if (scr->omax_bits[0] == 0u) { ax[0] = 1; ax[1] = 2; ax[2] = 0; }
  else                              { ax[0] = 0; ax[1] = 1; ax[2] = 2; }
}

// ---------------------------------------------------------------------------
// staging: caller layouts -> resident f64 records
// ---------------------------------------------------------------------------
__global__ void k_rays_from_f32aos(const float* __restrict__ src,
                                   RayRec64* __restrict__ dst, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* s = src + 12ull * i;  // GpuRay field order
  RayRec64 r;
  r.x = s[0]; r.y = s[1]; r.z = s[2];
  r.nx = s[3]; r.ny = s[4]; r.nz = s[5];
  r.nl = s[6]; r.energy = s[7];
  r.es_re = s[8]; r.es_im = s[9]; r.ep_re = s[10]; r.ep_im = s[11];
  r.path_up = 0.0; r.pad0 = r.pad1 = r.pad2 = 0.0;
  dst[i] = r;
}

// cols: 13 columns of n doubles (x y z nx ny nz nl es_re es_im ep_re ep_im
// energy path_up).
__global__ void k_rays_from_cols(const double* __restrict__ cols,
                                 RayRec64* __restrict__ dst, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t N = n;
  RayRec64 r;
  r.x = cols[0 * N + i]; r.y = cols[1 * N + i]; r.z = cols[2 * N + i];
  r.nx = cols[3 * N + i]; r.ny = cols[4 * N + i]; r.nz = cols[5 * N + i];
  r.nl = cols[6 * N + i];
  r.es_re = cols[7 * N + i]; r.es_im = cols[8 * N + i];
  r.ep_re = cols[9 * N + i]; r.ep_im = cols[10 * N + i];
  r.energy = cols[11 * N + i];
  r.path_up = cols[12 * N + i];
  r.pad0 = r.pad1 = r.pad2 = 0.0;
  dst[i] = r;
}

// beam_to_diffraction_rays (src/waves/prepare_wave.rs:78-111) as a gather
// on the device: ray k is beam ray good[k], its normal normals[3 good[k]],
// nl = n . (a, b, c) in the host helper's operation order (no contraction,
// so the records hold the host gather's bits), Es / Ep from the beam's
// amplitude columns or (1, 0) without them (prepare_wave.rs:89-95), and the
// upstream path of the beam ray when a path column is given.
// cols: x y z a b c e [es_re es_im ep_re ep_im] [path], n_beam doubles each.
__global__ void k_rays_from_beam(const double* __restrict__ cols, uint32_t n_beam,
                                 const double* __restrict__ normals,
                                 const uint64_t* __restrict__ good, uint32_t n_good,
                                 uint32_t has_amps, uint32_t has_path,
                                 RayRec64* __restrict__ dst) {
#pragma clang fp contract(off)
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_good) return;
  const size_t N = n_beam;
  const size_t i = good[k];  // checked against n_beam on the host
  const double* n = normals + 3 * i;
  RayRec64 r;
  r.x = cols[0 * N + i]; r.y = cols[1 * N + i]; r.z = cols[2 * N + i];
  r.nx = n[0]; r.ny = n[1]; r.nz = n[2];
  r.nl = n[0] * cols[3 * N + i] + n[1] * cols[4 * N + i] + n[2] * cols[5 * N + i];
  r.energy = cols[6 * N + i];
  r.es_re = has_amps ? cols[7 * N + i] : 1.0;
  r.es_im = has_amps ? cols[8 * N + i] : 0.0;
  r.ep_re = has_amps ? cols[9 * N + i] : 0.0;
  r.ep_im = has_amps ? cols[10 * N + i] : 0.0;
  r.path_up = has_path ? cols[(has_amps ? 11 : 7) * N + i] : 0.0;
  r.pad0 = r.pad1 = r.pad2 = 0.0;
  dst[k] = r;
}

// Explicit pixel lists: within each tile of 64 P slots, lane l's P slots
// (slot t*64P + i*64 + l) take the list entries t*64P + l*P + i, so that a
// lane holds P consecutive entries (neighbours along a row of a grid-ordered
// list); perm[slot] is the entry, or ~0u for a padding slot, which repeats
// entry n-1.
__device__ __forceinline__ uint32_t list_entry(uint32_t s, uint32_t P) {
  const uint32_t tsz = 64u * P, t = s / tsz, r = s - t * tsz;
  return t * tsz + (r & 63u) * P + (r >> 6);
}

__global__ void k_pix_from_f32aos(const float4* __restrict__ src, PixF64* __restrict__ dst,
                                  uint32_t* __restrict__ perm, uint32_t n, uint32_t npad,
                                  uint32_t P) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npad) return;
  const uint32_t j = list_entry(s, P);
  const float4 p = src[j < n ? j : n - 1u];
  dst[s] = PixF64{(double)p.x, (double)p.y, (double)p.z, 0.0};
  perm[s] = j < n ? j : ~0u;
}

__global__ void k_pix_from_cols(const double* __restrict__ cols, PixF64* __restrict__ dst,
                                uint32_t* __restrict__ perm, uint32_t n, uint32_t npad,
                                uint32_t P) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npad) return;
  const uint32_t e = list_entry(s, P);
  const uint32_t j = e < n ? e : n - 1u;
  const size_t N = n;
  dst[s] = PixF64{cols[j], cols[N + j], cols[2 * N + j], 0.0};
  perm[s] = e < n ? e : ~0u;
}

// prepare_wave (src/waves/prepare_wave.rs:40-68), evaluated in the same
// operation order without contraction so the points equal the host's, and
// laid out tile by tile: tiles of th x tw pixels (th rows of ix, tw columns of
// iz), tiles in row-major order over (ix, iz).  perm[slot] is the
// prepare_wave index ix*nz + iz, or ~0u for a padding slot, which repeats the
// nearest pixel of the grid (so a tile's mean stays inside the tile).
__global__ void k_pix_grid(double cx, double cy, double cz, double dx, double dz, uint32_t nx,
                           uint32_t nz, uint32_t th, uint32_t tw, uint32_t npad,
                           PixF64* __restrict__ dst, uint32_t* __restrict__ perm) {
#pragma clang fp contract(off)
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npad) return;
  const uint32_t tsz = th * tw;
  const uint32_t ntz = (nz + tw - 1) / tw;
  const uint32_t t = s / tsz, r = s - t * tsz;
  const uint32_t tix = t / ntz, tiz = t - tix * ntz;
  uint32_t ix = tix * th + r / tw, iz = tiz * tw + r % tw;
  const bool real = ix < nx && iz < nz;
  if (!real) { ix = min(ix, nx - 1u); iz = min(iz, nz - 1u); }
  const double fx = nx > 1 ? (double)ix / (double)(nx - 1) - 0.5 : 0.0;
  const double fz = nz > 1 ? (double)iz / (double)(nz - 1) - 0.5 : 0.0;
  const double px = cx + fx * 2.0 * dx;
  const double pz = cz + fz * 2.0 * dz;
  dst[s] = PixF64{px, cy, pz, 0.0};
  perm[s] = real ? ix * nz + iz : ~0u;
}

// The slots of a float32 pixel list that lies on an nx x nz grid in the
// prepare_wave order (entry ix * nz + iz): k_pix_grid's layout, each slot
// holding the list's own point.
__global__ void k_pix_grid_list(const float4* __restrict__ list, uint32_t nx, uint32_t nz,
                                uint32_t th, uint32_t tw, uint32_t npad,
                                PixF64* __restrict__ dst, uint32_t* __restrict__ perm) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npad) return;
  const uint32_t tsz = th * tw;
  const uint32_t ntz = (nz + tw - 1) / tw;
  const uint32_t t = s / tsz, r = s - t * tsz;
  const uint32_t tix = t / ntz, tiz = t - tix * ntz;
  uint32_t ix = tix * th + r / tw, iz = tiz * tw + r % tw;
  const bool real = ix < nx && iz < nz;
  if (!real) { ix = min(ix, nx - 1u); iz = min(iz, nz - 1u); }
  const float4 v = list[ix * nz + iz];
  dst[s] = PixF64{v.x, v.y, v.z, 0.0};
  perm[s] = real ? ix * nz + iz : ~0u;
}

// ---------------------------------------------------------------------------
// FP64 centroid (fixed-order tree, deterministic) and bounds
// ---------------------------------------------------------------------------
static constexpr int kRedBlocks = 240;  // partial sums of the centroid

__global__ __launch_bounds__(256) void k_centroid_partial(
    const RayRec64* __restrict__ rays, uint32_t n, double* __restrict__ part) {
  __shared__ double sx[256], sy[256], sz[256];
  double ax = 0, ay = 0, az = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    ax += rays[i].x; ay += rays[i].y; az += rays[i].z;
  }
  sx[threadIdx.x] = ax; sy[threadIdx.x] = ay; sz[threadIdx.x] = az;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sx[threadIdx.x] += sx[threadIdx.x + s];
      sy[threadIdx.x] += sy[threadIdx.x + s];
      sz[threadIdx.x] += sz[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x * 4 + 0] = sx[0];
    part[blockIdx.x * 4 + 1] = sy[0];
    part[blockIdx.x * 4 + 2] = sz[0];
    part[blockIdx.x * 4 + 3] = 0.0;
  }
}

// Also resets the atomically reduced bounds of the instance.
__global__ __launch_bounds__(256) void k_centroid_final(
    const double* __restrict__ part, uint32_t nblk, uint32_t n,
    const RayRec64* __restrict__ rays, InstScratch* __restrict__ scr) {
  __shared__ double sx[256], sy[256], sz[256];
  double ax = 0, ay = 0, az = 0;
  for (uint32_t b = threadIdx.x; b < nblk; b += 256) {
    ax += part[b * 4 + 0]; ay += part[b * 4 + 1]; az += part[b * 4 + 2];
  }
  sx[threadIdx.x] = ax; sy[threadIdx.x] = ay; sz[threadIdx.x] = az;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sx[threadIdx.x] += sx[threadIdx.x + s];
      sy[threadIdx.x] += sy[threadIdx.x + s];
      sz[threadIdx.x] += sz[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double inv = n ? 1.0 / (double)n : 0.0;
    scr->cx = sx[0] * inv; scr->cy = sy[0] * inv; scr->cz = sz[0] * inv;
    scr->blk_rev = 0.0;
    scr->amax_bits = 0u;
    for (int k = 0; k < 4; ++k) scr->omax_bits[k] = 0u;
    for (int k = 0; k < 3; ++k) scr->qmax_bits[k] = scr->dqmax_bits[k] = 0u;
    for (int k = 0; k < 3; ++k) { scr->qlo_key[k] = 0xffffffffu; scr->qhi_key[k] = 0u; }
    scr->nl_lo_key = scr->c_lo_key = 0xffffffffu;
    scr->nl_hi_key = scr->c_hi_key = 0u;
    scr->nu_max_bits = 0u;
    scr->hmax_bits = scr->sigma_bits = scr->ldev_bits = 0u;
    scr->rmin_bits = 0x7f800000u;  // +inf
    scr->rmax_bits = 0u;
    scr->flags = 0u;
    scr->ep_nonzero = 0u;
    scr->n_differs = 0u;
    scr->n0x = n ? rays[0].nx : 0.0;
    scr->n0y = n ? rays[0].ny : 0.0;
    scr->n0z = n ? rays[0].nz : 0.0;
    scr->v0_rev = scr->v2_rev = 0.0;
  }
}

// Per-axis max |o| (o rounded to FP32 as the pair kernel sees it), max |o|,
// whether any Ep != 0 and whether any normal differs from ray 0's; the range
// of nl and max |n0.o| (rounded outwards) for the separable amplitude.
__global__ __launch_bounds__(256) void k_ray_bounds(
    const RayRec64* __restrict__ rays, uint32_t n, InstScratch* __restrict__ scr) {
  const double cx = scr->cx, cy = scr->cy, cz = scr->cz;
  const double n0x = scr->n0x, n0y = scr->n0y, n0z = scr->n0z;
  float mx = 0.f, my = 0.f, mz = 0.f, mo = 0.f;
  float nlo = __uint_as_float(0x7f800000u), nhi = -__uint_as_float(0x7f800000u), nu = 0.f;
  uint32_t ep = 0u, nd = 0u;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const RayRec64 r = rays[i];
    const double oe[3] = {r.x - cx, r.y - cy, r.z - cz};
    const float ox = (float)oe[0], oy = (float)oe[1], oz = (float)oe[2];
    mx = fmaxf(mx, fabsf(ox)); my = fmaxf(my, fabsf(oy)); mz = fmaxf(mz, fabsf(oz));
    const double o2 = (double)ox * ox + (double)oy * oy + (double)oz * oz;
    mo = fmaxf(mo, __double2float_ru(sqrt(o2)));
    ep |= (r.ep_re != 0.0 || r.ep_im != 0.0) ? 1u : 0u;
    nd |= (r.nx != n0x || r.ny != n0y || r.nz != n0z) ? 1u : 0u;
    nlo = fminf(nlo, __double2float_rd(r.nl));
    nhi = fmaxf(nhi, __double2float_ru(r.nl));
    nu = fmaxf(nu, __double2float_ru(fabs(n0x * oe[0] + n0y * oe[1] + n0z * oe[2])));
  }
  mx = wave_max(mx); my = wave_max(my); mz = wave_max(mz); mo = wave_max(mo);
  nlo = wave_min(nlo); nhi = wave_max(nhi); nu = wave_max(nu);
  const bool epw = __any(ep), ndw = __any(nd);
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&scr->omax_bits[0], __float_as_uint(mx));
    atomicMax(&scr->omax_bits[1], __float_as_uint(my));
    atomicMax(&scr->omax_bits[2], __float_as_uint(mz));
    atomicMax(&scr->omax_bits[3], __float_as_uint(mo));
    if (epw) atomicOr(&scr->ep_nonzero, 1u);
    if (ndw) atomicOr(&scr->n_differs, 1u);
    atomicMin(&scr->nl_lo_key, fkey(nlo));
    atomicMax(&scr->nl_hi_key, fkey(nhi));
    atomicMax(&scr->nu_max_bits, __float_as_uint(nu));
  }
}

// Reference value of the rays' nl (midpoint of its range; 0 without rays).
__device__ __forceinline__ double nl_ref(const InstScratch* scr) {
  if (scr->nl_lo_key > scr->nl_hi_key) return 0.0;
  return 0.5 * ((double)fkey_inv(scr->nl_lo_key) + (double)fkey_inv(scr->nl_hi_key));
}

// One wave per tile of tsz = 64 * P slots (P slots per lane, slot = t*tsz +
// i*64 + lane).  q = p - c is kept as a double-double and |q| likewise (the
// base phase k|q| reaches ~1e12 rad, where one FP64 rounding of |q| is
// already ~1e-4 rad); the tile reference point q_t is the centre of the
// tile's bounding box.  Writes the FP32 slot records, |q| per slot (two
// doubles) and the tile record, and reduces the bounds of |q|, |dq|, R and
// of the lanes' slot lines (real slots only: perm[s] != ~0u, or s < np
// without perm; a lane's real slots are a prefix of its P slots).
// Last slot of a lane that holds a pixel (0 if no slot past slot 0 does);
// s0 is the lane's slot 0.  The lane's slot line runs from slot 0 to it.
template <int P>
__device__ __forceinline__ int line_last(const uint32_t* __restrict__ perm, uint32_t np,
                                         uint32_t s0) {
#pragma unroll 1
  for (int i = P - 1; i >= 1; --i) {
    const uint32_t s = s0 + i * 64u;
    if (perm ? perm[s] != ~0u : s < np) return i;
  }
  return 0;
}

// The lane's slot-line step h: the secant from its slot 0 (slot s0) to its
// last real slot (to slot 1 when none past slot 0 is real).
template <int P>
__device__ __forceinline__ void line_secant(const PixF64* __restrict__ pix,
                                            const uint32_t* __restrict__ perm, uint32_t np,
                                            uint32_t s0, double (&hv)[3]) {
  const int kl = line_last<P>(perm, np, s0);
  const PixF64 p0 = pix[s0];
  const PixF64 pk = pix[s0 + 64u * (uint32_t)(kl > 0 ? kl : 1)];
  const double dk = kl > 0 ? (double)kl : 1.0;
  hv[0] = (pk.x - p0.x) / dk; hv[1] = (pk.y - p0.y) / dk; hv[2] = (pk.z - p0.z) / dk;
}

// The line's step constants at the anchor q_a (relative to the centroid,
// |q_a| = Ra) for the step h: -2 kr h in kernel axis order, and r0 and c4 at
// q_a + h minus at q_a, from |q_a + h|^2 - |q_a|^2 = 2 q_a.h + |h|^2 without
// cancellation, and c4 at q_a + h.
__device__ __forceinline__ LineRec line_step(const double (&qa)[3], double Ra,
                                             const double (&hv)[3], double kr,
                                             const int (&ax)[3]) {
  LineRec r;
  r.hx = (float)(-2.0 * kr * pick3(hv, ax[0]));
  r.hy = (float)(-2.0 * kr * pick3(hv, ax[1]));
  r.hz = (float)(-2.0 * kr * pick3(hv, ax[2]));
  const double t = 2.0 * (qa[0] * hv[0] + qa[1] * hv[1] + qa[2] * hv[2]) +
                   (hv[0] * hv[0] + hv[1] * hv[1] + hv[2] * hv[2]);
  const double Rb = sqrt(Ra * Ra + t);
  const double ia = 1.0 / Ra, ib = 1.0 / Rb;
  const double di = -(t / (Ra + Rb)) * ia * ib;  // 1/|q_a + h| - 1/|q_a|
  r.dr0 = (float)(0.5 * di);
  r.dc4 = (float)(0.125 / kr * di * (ib * ib + ib * ia + ia * ia));
  r.c4n = (float)(0.125 * (ib * ib * ib / kr));
  r.pad0 = r.pad1 = 0.f;
  return r;
}

template <int P>
__global__ __launch_bounds__(64) void k_tile_prep(const PixF64* __restrict__ pix, double kr,
                                                  InstScratch* __restrict__ scr,
                                                  PixRec32* __restrict__ out32,
                                                  TileRec* __restrict__ tiles,
                                                  double* __restrict__ R64,
                                                  const uint32_t* __restrict__ perm,
                                                  uint32_t np) {
  #pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x;
  const uint32_t base = blockIdx.x * 64u * P;
  const double c[3] = {scr->cx, scr->cy, scr->cz};
  const bool uni = scr->n_differs == 0u;
  const double n0x = scr->n0x, n0y = scr->n0y, n0z = scr->n0z;
  double qh[P][3], ql[P][3];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const PixF64 p = pix[base + i * 64u + lane];
    const double pv[3] = {p.x, p.y, p.z};
    for (int k = 0; k < 3; ++k) {
      two_sum(pv[k], -c[k], &qh[i][k], &ql[i][k]);
      lo[k] = fmin(lo[k], qh[i][k]);
      hi[k] = fmax(hi[k], qh[i][k]);
    }
  }
  double t[3], wlo[3], whi[3];
  for (int k = 0; k < 3; ++k) {
    wlo[k] = wave_mind(lo[k]);
    whi[k] = wave_maxd(hi[k]);
    t[k] = 0.5 * (wlo[k] + whi[k]);
  }
  const double Rt = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
  int ax[3];
  kernel_axes(scr, ax);
  if (lane == 0) tiles[blockIdx.x] = TileRec{pick3(t, ax[0]), pick3(t, ax[1]), pick3(t, ax[2]), Rt};
  float aq[3] = {0.f, 0.f, 0.f}, ad[3] = {0.f, 0.f, 0.f};
  float rmin = __uint_as_float(0x7f800000u), rmax = 0.f;
  float clo = __uint_as_float(0x7f800000u), chi = -__uint_as_float(0x7f800000u);
#pragma unroll
  for (int i = 0; i < P; ++i) {
    double Rh, Rl;
    norm_dd(qh[i], ql[i], &Rh, &Rl);
    double dq[3];
    for (int k = 0; k < 3; ++k) dq[k] = (qh[i][k] - t[k]) + ql[i][k];
    PixRec32 r;
    r.mx = (float)(-2.0 * kr * pick3(dq, ax[0]));
    r.my = (float)(-2.0 * kr * pick3(dq, ax[1]));
    r.mz = (float)(-2.0 * kr * pick3(dq, ax[2]));
    r.R2 = (float)(Rh * Rh);
    r.R = (float)Rh;
    r.g = (float)((Rt - Rh) / Rt);
    r.nq = uni ? (float)(n0x * qh[i][0] + n0y * qh[i][1] + n0z * qh[i][2]) : 0.f;
    const double ri = 1.0 / Rh, ri3 = ri * ri * ri / kr;
    r.r0 = (float)(0.5 * ri);
    r.c4 = (float)(0.125 * ri3);
    r.rinv = (float)ri;
    r.c2 = (float)(0.5 * ri3);
    // separable amplitude: range of C = n0.q / |q| (the slot factor
    // (nl_ref + C) / |q| is applied by k_merge)
    r.pad = 0.f;
    if (uni && Rh > 0.0) {
      const double C = (n0x * qh[i][0] + n0y * qh[i][1] + n0z * qh[i][2]) * ri;
      clo = fminf(clo, __double2float_rd(C));
      chi = fmaxf(chi, __double2float_ru(C));
    }
    const uint32_t s = base + i * 64u + lane;
    out32[s] = r;
    R64[2 * s] = Rh;
    R64[2 * s + 1] = Rl;
    for (int k = 0; k < 3; ++k) {
      aq[k] = fmaxf(aq[k], __double2float_ru(fabs(qh[i][k])));
      ad[k] = fmaxf(ad[k], __double2float_ru(fabs(dq[k])));
    }
    rmin = fminf(rmin, __double2float_rd(Rh));
    rmax = fmaxf(rmax, __double2float_ru(Rh));
  }
  for (int k = 0; k < 3; ++k) {
    aq[k] = wave_max(aq[k]);
    ad[k] = wave_max(ad[k]);
  }
  rmin = wave_min(rmin); rmax = wave_max(rmax);
  clo = wave_min(clo); chi = wave_max(chi);
  // the lane's slot line q_i = q_0 + i h, h the secant to the lane's last
  // real slot (so rounded positions depart from it by their own rounding,
  // not by the rounding of one step times i)
  float hm = 0.f, sg = 0.f, dv = 0.f;
  auto real = [&](int i) {
    const uint32_t s = base + i * 64u + lane;
    return perm ? perm[s] != ~0u : s < np;
  };
  const int kl = line_last<P>(perm, np, base + lane);
  double h[3];
  line_secant<P>(pix, perm, np, base + lane, h);
  {
    // the lane's line record for the slot-line kernel (anchored at slot 0)
    const PixF64 p0 = pix[base + lane];
    const double qa[3] = {p0.x - c[0], p0.y - c[1], p0.z - c[2]};
    LineRec* lines = reinterpret_cast<LineRec*>(tiles + gridDim.x);
    lines[blockIdx.x * 64u + lane] = line_step(qa, R64[2 * (base + lane)], h, kr, ax);
  }
  if (P >= 2 && real(0) && kl >= 1) {
    const double hn = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    const double r0 = sqrt(qh[0][0] * qh[0][0] + qh[0][1] * qh[0][1] + qh[0][2] * qh[0][2]);
    hm = __double2float_ru(hn);
    if (hn > 0.0 && r0 > 0.0)
      sg = __double2float_ru(fabs(qh[0][0] * h[0] + qh[0][1] * h[1] + qh[0][2] * h[2]) / (r0 * hn));
#pragma unroll
    for (int i = 1; i < P; ++i) {
      if (i > kl || !real(i)) continue;
      for (int k = 0; k < 3; ++k) {
        const double d = ((qh[i][k] - qh[0][k]) + (ql[i][k] - ql[0][k])) - (double)i * h[k];
        dv = fmaxf(dv, __double2float_ru(fabs(d)));
      }
    }
  }
  hm = wave_max(hm); sg = wave_max(sg); dv = wave_max(dv);
  if (lane == 0) {
    atomicMax(&scr->hmax_bits, __float_as_uint(hm));
    atomicMax(&scr->sigma_bits, __float_as_uint(sg));
    atomicMax(&scr->ldev_bits, __float_as_uint(dv));
    atomicMin(&scr->c_lo_key, fkey(clo));
    atomicMax(&scr->c_hi_key, fkey(chi));
    for (int k = 0; k < 3; ++k) {
      atomicMax(&scr->qmax_bits[k], __float_as_uint(aq[k]));
      atomicMax(&scr->dqmax_bits[k], __float_as_uint(ad[k]));
      atomicMin(&scr->qlo_key[k], fkey((float)wlo[k]));
      atomicMax(&scr->qhi_key[k], fkey((float)whi[k]));
    }
    atomicMin(&scr->rmin_bits, __float_as_uint(rmin));
    atomicMax(&scr->rmax_bits, __float_as_uint(rmax));
  }
}

// F64 instances only need min/max |q| (for the skip bound).
__global__ __launch_bounds__(256) void k_pix_bounds(const PixF64* __restrict__ pix, uint32_t n,
                                                    InstScratch* __restrict__ scr) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float rmin = __uint_as_float(0x7f800000u), rmax = 0.f;
  if (i < n) {
    const double qx = pix[i].x - scr->cx, qy = pix[i].y - scr->cy, qz = pix[i].z - scr->cz;
    const double R = sqrt(qx * qx + qy * qy + qz * qz);
    rmin = __double2float_rd(R);
    rmax = __double2float_ru(R);
  }
  rmin = wave_min(rmin);
  rmax = wave_max(rmax);
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&scr->rmin_bits, __float_as_uint(rmin));
    atomicMax(&scr->rmax_bits, __float_as_uint(rmax));
  }
}

// Per-instance flags from the device-computed bounds.
//   F64:   CHECK when some pair may fall below the skip threshold
//          (min |q| - max |o| <= skip).
//   F32R:  the bound of the centroid-referenced phase,
//            v0 = kr (|o|^2 + 2 sum_i |o_i| |q_i|) / (2 min|q| - |o|),
//          and of the tile-referenced residual,
//            v2 = kr (2 sum_i |o_i| |dq_i| + |o|^3 / (2 min|q| - |o|)) / (...)
//          TILE is used unless v0 is already small; NOFRACT when the phase
//          the pair kernel sees stays well inside v_sin's +-256 rev domain.
//          LIN2 when o is zero along one axis (dropped from o.m).
//          PAX (tile variant) when the paraxial expansions of 1/path and
//          1/(path + R) are exact to FP32.
//          SEPAMP (PAX, uniform normal) when the amplitude
//            am (nl + n.d / path) / path
//          is a ray factor am (1 + (nl - nl_ref - n.o / R_ref) / (nl_ref + C_ref))
//          times a slot factor (nl_ref + C) / |q| to a relative 1e-6, C =
//          n.q / |q|.  With path = |q| (1 + e), |e| <= rho, the difference
//          of the two, times |q| / am, is
//            nl (1/(1+e) - 1) + (C - n.o/|q|) (1/(1+e)^2 - 1)
//            - n.o (1/|q| - 1/R_ref) - (nl - nl_ref - n.o/R_ref) (C - C_ref) / (nl_ref + C_ref),
//          bounded term by term from the ranges of nl, C, |q| and max |n.o|.
//          LINE (SEPAMP, P >= 4) when every lane's real slots lie on a line
//          q_i = q_0 + i h (h the secant from slot 0 to the lane's last real
//          slot) and the ray's reduced phase along it,
//            psi(t) = n(t) (r0(t) - n(t) c4(t)),  n(t) = kr (|o|^2 - 2 o.q(t)),
//            r0 = 1/(2|q|), c4 = 1/(8 kr |q|^3)  (the PAX form),
//          departs from its linear extrapolation from q_0 and q_0 + h by at most
//            (P-1)(P-2)/2 max |psi''| + kr |q_i - q_0 - i h| 2 |o| / min|q|
//          <= 1e-7 rev, with |n'| <= 2 kr |o| |h|, |q|' <= sigma |h|, |q|'' <= |h|^2 / |q|:
//            |psi''| <= 2 kr |o| |h|^2 s / R^2 + kr rho |h|^2 (1 + 2 s^2) / (2 R)
//                       + kr |o|^2 |h|^2 / R^3 + 3 kr rho |o| |h|^2 s / R^2
//                       + 15/8 kr rho^2 |h|^2 / R,
//          R = min |q|, s = sigma + (P-1) |h| / R bounds |q^.h^| along the line.
//          (A float32 pixel list departs from the secant by its own rounding,
//          ~2e-9 mm: up to ~1.6e-5 rev at 1 m, 10 keV and a 0.2 mm aperture,
//          so such lists take LINEDEV.)
//          LINEDEV: the same with the slots' offsets delta_i = q_i - q_0 - i h
//          from the line (a list whose points stray further from it) taken
//          into the phase to first order, kr (d^ - q^).delta_i ~ o.w_i with
//          w_i = -kr (delta_i - q^ (q^.delta_i)) / |q_i|; what is left is
//          bounded by kr sqrt(3) max|delta| (2 |o| / min|q|)^2, and the first-
//          order rotation by that angle needs it below 1e-4 rev.
// force_flags: 0x100 forces TILE to (force & TILE), 0x200 forces CHECK,
// 0x400 forbids LIN2, 0x800 forbids PAX, 0x1000 forbids SEPAMP, 0x2000
// forbids LINE, 0x4000 forbids LINEDEV, 0x8000 forbids UNIFORM_N (every ray
// then takes its own normal, as rays off a curved optic do), 0x10000 forbids
// EPZERO.
__global__ void k_plan(InstScratch* __restrict__ scr, double skip, double kr, uint32_t mode,
                       uint32_t force_flags, uint32_t static_flags, uint32_t P, double gsx,
                       double gsz, double gemax) {
  const double ox = __uint_as_float(scr->omax_bits[0]), oy = __uint_as_float(scr->omax_bits[1]);
  const double oz = __uint_as_float(scr->omax_bits[2]), om = __uint_as_float(scr->omax_bits[3]);
  const double qx = __uint_as_float(scr->qmax_bits[0]), qy = __uint_as_float(scr->qmax_bits[1]);
  const double qz = __uint_as_float(scr->qmax_bits[2]);
  const double dx = __uint_as_float(scr->dqmax_bits[0]), dy = __uint_as_float(scr->dqmax_bits[1]);
  const double dz = __uint_as_float(scr->dqmax_bits[2]);
  const double rmin = __uint_as_float(scr->rmin_bits);
  uint32_t f = static_flags;
  if (scr->n_differs == 0u && !(force_flags & 0x8000u)) f |= INST_F_UNIFORM_N;
  if (!(rmin - om > skip * (1.0 + 1e-9))) f |= INST_F_CHECK;
  const double den = 2.0 * rmin - om;
  double v0 = 1e300, v2 = 1e300;
  if (den > 0.0) {
    v0 = kr * (om * om + 2.0 * (ox * qx + oy * qy + oz * qz)) / den;
    v2 = kr * (2.0 * (ox * dx + oy * dy + oz * dz) + om * om * (om / den)) / den;
  }
  if (mode == MODE_F32R) {
    // no ray has Ep != 0: the Ep sums are exact zeros
    if (scr->ep_nonzero == 0u && !(force_flags & 0x10000u)) f |= INST_F_EPZERO;
    bool tile = v0 > 2.0;
    if (force_flags & 0x100u) tile = (force_flags & INST_F_TILE) != 0u;
    if (tile) f |= INST_F_TILE;
    if ((tile ? v2 : v0) < 128.0) f |= INST_F_NOFRACT;
    const bool zero_axis = scr->omax_bits[0] == 0u || scr->omax_bits[1] == 0u ||
                           scr->omax_bits[2] == 0u;
    if (zero_axis && !(force_flags & 0x400u)) f |= INST_F_LIN2;
    // rho = |path^2 - R^2| / R^2 <= (|o|^2 + 2 sum_i |o_i| |q_i|) / min R^2;
    // the expansions' first-order errors are 3 rho^2 / 8 (amplitude) and
    // rho^2 / 8 (relative, phase)
    if (tile && rmin > 0.0) {
      const double rho = (om * om + 2.0 * (ox * qx + oy * qy + oz * qz)) / (rmin * rmin);
      if (rho <= 1e-4 && !(force_flags & 0x800u)) f |= INST_F_PAX;
      const bool keys = scr->nl_lo_key <= scr->nl_hi_key && scr->c_lo_key <= scr->c_hi_key;
      if ((f & INST_F_PAX) && (f & INST_F_UNIFORM_N) && keys && !(force_flags & 0x1000u)) {
        const double nlo = fkey_inv(scr->nl_lo_key), nhi = fkey_inv(scr->nl_hi_key);
        const double clo = fkey_inv(scr->c_lo_key), chi = fkey_inv(scr->c_hi_key);
        const double nu = __uint_as_float(scr->nu_max_bits);
        const double rmax = __uint_as_float(scr->rmax_bits);
        const double nlb = 0.5 * (nlo + nhi), cb = 0.5 * (clo + chi), rb = 0.5 * (rmin + rmax);
        const double e1 = rho / (1.0 - rho), e2 = rho * (2.0 + rho) / ((1.0 - rho) * (1.0 - rho));
        const double nla = fmax(fabs(nlo), fabs(nhi)), ca = fmax(fabs(clo), fabs(chi));
        const double num = nla * e1 + (ca + nu / rmin) * e2 +
                           nu * fmax(1.0 / rmin - 1.0 / rb, 1.0 / rb - 1.0 / rmax) +
                           (0.5 * (nhi - nlo) + nu / rb) * 0.5 * (chi - clo) / (nlb + cb);
        const double den = (nlo + clo - nu / rmin) - nla * e1 - (ca + nu / rmin) * e2;
        if (nlb + cb > 0.0 && den > 0.0 && num <= 1e-6 * den) f |= INST_F_SEPAMP;
      }
      if ((f & INST_F_SEPAMP) && P >= 4u && !(force_flags & 0x2000u)) {
        const double hm = __uint_as_float(scr->hmax_bits);
        const double dv = __uint_as_float(scr->ldev_bits);
        const double R = rmin, h2 = hm * hm;
        const double s = fmin(1.0, (double)__uint_as_float(scr->sigma_bits) + (P - 1.0) * hm / R);
        const double d2 = 2.0 * kr * om * h2 * s / (R * R) + 0.5 * kr * rho * h2 * (1.0 + 2.0 * s * s) / R +
                          kr * om * om * h2 / (R * R * R) + 3.0 * kr * rho * om * h2 * s / (R * R) +
                          1.875 * kr * rho * rho * h2 / R;
        const double bend = 0.5 * (P - 1.0) * (P - 2.0) * d2;
        const double dev1 = kr * dv * 1.7320508075688772 * 2.0 * om / R;  // first order
        const double dev2 = dev1 * 2.0 * om / R;                           // after it
        // LINEDEV runs in passes of at most 8 slots, each anchored at its
        // first slot: past P = 8 the offsets from an anchor's line reach
        // twice the offsets from the lane's line
        const double fd = P > 8u ? 2.0 : 1.0;
        if (bend + dev1 <= 1e-7) {
          f |= INST_F_LINE;
        } else if (bend + fd * dev2 <= 1e-7 && fd * dev1 <= 1e-4 && !(force_flags & 0x4000u)) {
          f |= INST_F_LINE | INST_F_LINEDEV;
        }
      }
    }
    // MFMA (a grid screen in the x-z plane, steps gsx, gsz): over a block of
    // 64 x 64 pixels the reduced phase f(q) = kr (|q - o| - |q|) is taken
    // linear about the block centre a, f(a) + grad f(a).(q - a).  The
    // remainder is at most 1/2 max|f_ee| |q - a|^2, |q - a| <= h = 32 |(gsx,
    // gsz)|, with e a unit vector in the screen plane, d = q - o, and
    //   f_ee = kr [(1/|d| - 1/|q|) - ((d.e)^2/|d|^3 - (q.e)^2/|q|^3)],
    //   |1/|d| - 1/|q|| <= dR / Rl^2,  dR = (|o|^2 + 2 sum |o_i| Q_i) / (2 Rl)
    //     (a bound of ||d| - |q||),
    //   (d.e)^2/|d|^3 - (q.e)^2/|q|^3 = (q.e)^2 (1/|d|^3 - 1/|q|^3)
    //     + ((o.e)^2 - 2 (q.e)(o.e)) / |d|^3,
    // where Rl = min|q| - |o|max - h bounds |q| and |d| from below over the
    // blocks, Q_i = max|q_i| (+ h in the plane), and |q.e| <= |(Q_x, Q_z)|,
    // |o.e| <= |(o_x, o_z)|.  A phase error of at most E rev per term moves
    // the field by at most 2 pi E relative (rel-L2 included), so the variant
    // is allowed when E <= 1e-6 rev (6.3e-6 rad).  The phase step per pixel
    // along x is |df . step| = kr |gsx| |q_x (1/|d| - 1/|q|) - o_x/|d||
    // <= smax = kr |gsx| (Q_x dR / Rl^2 + o_x / Rl) (likewise z).  The
    // kernel takes the steps mod 1, so at any slope the factors' FP32 phases
    // stay within 1/2 + 15.5/2 rev (~5e-7 rev of rounding).
    // A float32 pixel list on the grid (offsets |e| <= gemax pixels from it)
    // widens the block to h = (32 + gemax) |(gsx, gsz)|, and its offset term
    // e sx needs the step itself, so its steps stay below 1/2 rev
    // (16 smax + 1/2 <= 8).  The kernel applies the offsets of the second
    // row and column tile to first order, which leaves (2 pi smax 2 gemax)^2
    // / 2 per term, held below 2 pi 1e-7 by smax gemax <= 8.9e-5.
    const uint32_t need_m = INST_F_TILE | INST_F_PAX | INST_F_UNIFORM_N | INST_F_SEPAMP;
    if ((static_flags & INST_F_MFMA_OK) && (f & need_m) == need_m && !(force_flags & 0x20000u)) {
      const double ax = fabs(gsx), az = fabs(gsz);
      const double h = (32.0 + gemax) * sqrt(ax * ax + az * az);
      const double Rl = rmin - om - h;
      if (Rl > 0.0) {
        const double Qx = qx + h, Qy = qy, Qz = qz + h;
        const double dR = (om * om + 2.0 * (ox * Qx + oy * Qy + oz * Qz)) / (2.0 * Rl);
        const double R2 = Rl * Rl, R3 = R2 * Rl;
        const double Qp2 = Qx * Qx + Qz * Qz, op = sqrt(ox * ox + oz * oz);
        const double t1 = dR / R2;
        const double t2 = Qp2 * 3.0 * dR / (R2 * R2) + (2.0 * sqrt(Qp2) * op + op * op) / R3;
        const double E = 0.5 * kr * (t1 + t2) * h * h;
        const double smax = kr * fmax(ax * (Qx * t1 + ox / Rl), az * (Qz * t1 + oz / Rl));
        scr->blk_rev = E;
        const bool slope_ok = gemax == 0.0 || 16.0 * smax + 0.5 <= 8.0;
        if (E <= 1e-6 && slope_ok && smax * gemax <= 8.9e-5) f |= INST_F_MFMA;
      }
    }
  }
  if (force_flags & 0x200u) f |= INST_F_CHECK;
  scr->flags = f;
  scr->v0_rev = v0;
  scr->v2_rev = v2;
}

// FP32 ray records of the phase-reduced kernel (layout and meaning: RayRec32).
// Rounding o = r - c to FP32 moves the ray by delta = o_f32 - o (~1e-9 mm),
// a phase error kr delta.d at a pixel in direction d from the ray: first order
// along the line of sight (an aperture not aligned with an axis), and across
// it wherever the ray-to-pixel directions fan out (near field).  The
// tile-referenced variant takes kr delta.(q_t - o)/|q_t - o| per (ray, tile)
// from delta kept in the record; the centroid-referenced variant, used only
// when the whole phase is small, takes kr delta.u once per ray, u from the
// centroid to the centre of the screen's bounding box.
// Returns max(|cs|, |pr| + |pi|) of the record, a bound of the moduli of
// both amplitudes (the MFMA kernel scales its FP16 factors by the largest).
__device__ __forceinline__ float ray_prep_one(const RayRec64* __restrict__ rays, uint32_t i,
                                              const InstScratch* __restrict__ scr, double kr,
                                              double kr_lo, uint32_t have_pix,
                                              RayRec32* __restrict__ out) {
  const double cx = scr->cx, cy = scr->cy, cz = scr->cz;
  const uint32_t flags = scr->flags;
  int ax[3];
  kernel_axes(scr, ax);
  double u[3] = {0.0, 0.0, 0.0};  // unit direction to the screen
  if (have_pix) {
    for (int a = 0; a < 3; ++a)
      u[a] = 0.5 * ((double)fkey_inv(scr->qlo_key[a]) + (double)fkey_inv(scr->qhi_key[a]));
    const double m = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double inv = m > 0.0 ? 1.0 / m : 0.0;
    for (int a = 0; a < 3; ++a) u[a] *= inv;
  }
  const RayRec64 r = rays[i];
  const double oe[3] = {r.x - cx, r.y - cy, r.z - cz};
  const float o3[3] = {(float)oe[0], (float)oe[1], (float)oe[2]};
  const double dl[3] = {(double)o3[0] - oe[0], (double)o3[1] - oe[1], (double)o3[2] - oe[2]};
  const bool tile = (flags & INST_F_TILE) != 0u;
  const double los = tile ? 0.0 : kr * (dl[0] * u[0] + dl[1] * u[1] + dl[2] * u[2]);
  const double oo = (double)o3[0] * o3[0] + (double)o3[1] * o3[1] + (double)o3[2] * o3[2];
  double er = r.es_re, ei = r.es_im, pr = r.ep_re, pi = r.ep_im;
  if (flags & INST_F_PATHUP) {
    #pragma clang fp contract(off)
    const double hi = kr * r.path_up;
    const double lo = fma(kr, r.path_up, -hi) + kr_lo * r.path_up;
    double s, c;
    sincospi(2.0 * ((hi - rint(hi)) + lo), &s, &c);
    const double er2 = er * c - ei * s, ei2 = er * s + ei * c;
    const double pr2 = pr * c - pi * s, pi2 = pr * s + pi * c;
    er = er2; ei = ei2; pr = pr2; pi = pi2;
  }
  // polarisation form: Es = am e^{i th} cs, Ep = am e^{i th} (qr + i qi)
  const double es2 = er * er + ei * ei, ep2 = pr * pr + pi * pi;
  const double am = sqrt(es2 + ep2);
  double th = 0.0, cs = 0.0, qr = 0.0, qi = 0.0;
  if (am > 0.0) {
    th = es2 > 0.0 ? atan2(ei, er) : atan2(pi, pr);
    double s, c;
    sincos(th, &s, &c);
    cs = sqrt(es2) / am;
    qr = (pr * c + pi * s) / am;
    qi = (pi * c - pr * s) / am;
  }
  double off = th * (0.5 / kPI) + 0.25 + los;  // + 1/4 rev: the factor i
  off -= rint(off);
  const double n3[3] = {r.nx, r.ny, r.nz};
  RayRec32 o;
  o.ox = pick3(o3, ax[0]); o.oy = pick3(o3, ax[1]); o.oz = pick3(o3, ax[2]);
  o.oo = tile ? (float)pick3(dl, ax[2]) : (float)(kr * oo);
  o.dx = (float)pick3(dl, ax[0]);
  o.dy = (float)pick3(dl, ax[1]);
  // n = -am n_ray / (2 kr): n_ray . d = n . m + n_ray . (q_t - o) with m = -2 kr dq
  const double nf = -0.5 * am / kr;
  o.nx = (float)(pick3(n3, ax[0]) * nf);
  o.ny = (float)(pick3(n3, ax[1]) * nf);
  o.nz = (float)(pick3(n3, ax[2]) * nf);
  o.nl = (float)(am * r.nl);
  if (flags & INST_F_SEPAMP) {
    // the ray factor of the separable amplitude joins the polarisation terms
    const double nlb = nl_ref(scr);
    const double cb = 0.5 * ((double)fkey_inv(scr->c_lo_key) + (double)fkey_inv(scr->c_hi_key));
    const double rb = 0.5 * ((double)__uint_as_float(scr->rmin_bits) +
                             (double)__uint_as_float(scr->rmax_bits));
    const double nu = scr->n0x * oe[0] + scr->n0y * oe[1] + scr->n0z * oe[2];
    const double A = am * (1.0 + (r.nl - nlb - nu / rb) / (nlb + cb));
    cs *= A;
    qr *= A;
    qi *= A;
  }
  o.cs = (float)cs;
  o.pr = (float)qr;
  o.pi = (float)qi;
  o.no = (float)(am * (r.nx * (double)o3[0] + r.ny * (double)o3[1] + r.nz * (double)o3[2]));
  o.am = (float)am;
  o.off = (float)off;
  out[i] = o;
  return fmaxf(fabsf(o.cs), fabsf(o.pr) + fabsf(o.pi));
}

__global__ __launch_bounds__(256) void k_ray_prep32(const RayRec64* __restrict__ rays,
                                                    uint32_t n, InstScratch* __restrict__ scr,
                                                    double kr, double kr_lo, uint32_t have_pix,
                                                    RayRec32* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  float a = 0.f;
  if (i < n) a = ray_prep_one(rays, i, scr, kr, kr_lo, have_pix, out);
  if (scr->flags & INST_F_MFMA) {
    // every lane of the wave takes part in the reduction (lanes past n hold 0)
    a = wave_max(a);
    if ((threadIdx.x & 63u) == 0u) atomicMax(&scr->amax_bits, __float_as_uint(a));
  }
}

// ---------------------------------------------------------------------------
// FP32 phase-reduced pair kernel
// ---------------------------------------------------------------------------
// Per-(ray, tile) values staged in LDS by the TILE variant.
struct TileRay {
  float B;    // kr (|o|^2 - 2 o.q_t)
  float off;  // fractional revolution of kr (|o|^2 - 2 o.q_t) / (2 R_t), plus the ray's offset
  float n;    // UNIFORM_N: -am n.o; else am n.(q_t - o)
  float pad;
};
static_assert(sizeof(TileRay) == 16, "TileRay is 16 B");

static constexpr uint32_t kBlockRays = 64;  // rays staged per LDS block (one per lane)

template <int P>
struct SlotRegs {
  float mx[P], my[P], mz[P], R2[P], R[P], g[P], nq[P];
  float r0[P], c4[P], rinv[P], c2[P];
};

// Accumulates one ray into a[P][4] for the P slots of this lane.
//   TILE:  B, the phase offset and the normal term come from the
//          per-(ray, tile) LDS record; otherwise from the ray record.
//   UNIN:  every ray has the same normal, n.d = am (n.q - n.o).
//   LIN2:  o is zero along the last kernel axis.
//   PAX:   (tile variant) 1/path and 1/(path + R) from the paraxial
//          expansion: one FMA each instead of rsq and rcp.
//   SEP:   (PAX variant) the amplitude is the ray factor already in the
//          ray's polarisation terms times the slot factor that k_merge
//          applies to the summed partials: no per-pair amplitude.
// Per pair (TILE, UNIN, NOFRACT, LIN2, PAX): 21 FP32 add/mul/FMA + sin, cos;
// with SEP: 14 + sin, cos; without PAX: 20 + rsq, rcp, sin, cos.
// EP: some ray has Ep != 0; without it the Ep sums (a[i][2], a[i][3]) stay
// exact zeros and are not formed (4 of the 6 sum operations per slot).
template <int P, bool TILE, bool UNIN, bool NOFRACT, bool LIN2, bool PAX, bool SEP,
          bool EP = true>
__device__ __forceinline__ void pair_ray(const RayRec32& r, float B, float off, float nt,
                                         const SlotRegs<P>& s, float ikr, float h,
                                         float (&a)[P][4]) {
  float oo = r.oo, mno = -r.no;
  if constexpr (!TILE) {
    oo = vcopy(oo);
    off = r.off;
    mno = vcopy(mno);
  }
#pragma unroll
  for (int i = 0; i < P; ++i) {
    // kr (|o|^2 - 2 o.q) split as base + o.m
    float lin = r.ox * s.mx[i];
    lin = fmaf(r.oy, s.my[i], lin);
    if constexpr (!LIN2) lin = fmaf(r.oz, s.mz[i], lin);
    const float num = lin + (TILE ? B : oo);
    float inv, rden;  // 1 / path, 1 / (path + R)
    if constexpr (PAX) {
      inv = fmaf(-num, s.c2[i], s.rinv[i]);
      rden = fmaf(-num, s.c4[i], s.r0[i]);
    } else {
      const float p2 = fmaf(num, ikr, s.R2[i]);  // |p - r|^2
      inv = __builtin_amdgcn_rsqf(p2);
      rden = __builtin_amdgcn_rcpf(fmaf(p2, inv, s.R[i]));
    }
    float rev;
    if constexpr (TILE) {
      const float eps = num * rden;            // kr (path - R)
      const float u = fmaf(-eps, h, s.g[i]);   // (S_t - S) / S_t
      rev = fmaf(fmaf(B, u, lin), rden, off);
    } else {
      rev = fmaf(num, rden, off);
    }
    float nd;  // am n.d
    if constexpr (UNIN) {
      nd = fmaf(s.nq[i], r.am, TILE ? nt : mno);
    } else {
      nd = fmaf(r.nz, s.mz[i], fmaf(r.ny, s.my[i], fmaf(r.nx, s.mx[i], TILE ? nt : mno)));
    }
    const float f = NOFRACT ? rev : __builtin_amdgcn_fractf(rev);
    const float sn = __builtin_amdgcn_sinf(f);
    const float cs = __builtin_amdgcn_cosf(f);
    float wr = cs, wi = sn;
    if constexpr (!SEP) {
      const float obl = fmaf(nd, inv, r.nl);   // am (nl + n.d / path)
      const float amp = obl * inv;
      wr = amp * cs;
      wi = amp * sn;
    }
    a[i][0] = fmaf(r.cs, wr, a[i][0]);
    a[i][1] = fmaf(r.cs, wi, a[i][1]);
    if constexpr (EP) {
      a[i][2] = fmaf(r.pr, wr, a[i][2]);
      a[i][2] = fmaf(-r.pi, wi, a[i][2]);
      a[i][3] = fmaf(r.pr, wi, a[i][3]);
      a[i][3] = fmaf(r.pi, wr, a[i][3]);
    }
  }
}

// Per-lane constants of the LINE variant: the anchor slot's record and the
// step along the lane's line to the point q + h after it (-2 kr h, and r0,
// c4 there minus at the anchor, c4 there), from FP64.
struct LineRegs {
  float mx, my, mz, g, r0, c4;
  float hx, hy, hz, dr0, dc4, c4n;
};

// Accumulates one ray into the P slots of a lane on a slot line (TILE, UNIN,
// PAX, SEP): the reduced phase at slot 0 (the anchor) as pair_ray computes
// it, the step to the next point of the line
//   psi_1 - psi_0 = dn rden_1 + num_0 (rden_1 - rden_0),  dn = o.(m_1 - m_0),
//   rden_1 - rden_0 = (r0_1 - r0_0) - dn c4_1 - num_0 (c4_1 - c4_0),
// evaluated directly (small, so to full FP32 relative precision); slot 1
// takes e^{2 pi i rev} by one rotation by the step and each further slot by
// the three-term recurrence x_{i+1} = 2 cos(step) x_i - x_{i-1} (two FMAs;
// its rounding grows with i, ~4e-7 rms at slot 7, and it keeps the
// magnitude of slots 0 and 1).
// Per ray and lane: 15 FP32 operations + 2 sin + 2 cos, 5 for slot 1, 2 per
// further slot, and 6 per slot (sums).
// DEV: slot i lies off the line by its own offset; its phase gains
// o.w_i radians (w_i from the item prologue), applied to the slot's
// e^{i phase} to first order: 5 more operations per slot (4 with LIN2).
template <int P, bool NOFRACT, bool LIN2, bool DEV = false, bool EP = true>
__device__ __forceinline__ void pair_line(const RayRec32& r, float B, float off,
                                          const LineRegs& L, float h, float (&a)[P][4],
                                          const float (&dw)[P][3]) {
  float lin = r.ox * L.mx;
  lin = fmaf(r.oy, L.my, lin);
  if constexpr (!LIN2) lin = fmaf(r.oz, L.mz, lin);
  const float num = lin + B;
  const float rden = fmaf(-num, L.c4, L.r0);
  const float eps = num * rden;
  const float u = fmaf(-eps, h, L.g);
  const float rev = fmaf(fmaf(B, u, lin), rden, off);
  float dn = r.ox * L.hx;
  dn = fmaf(r.oy, L.hy, dn);
  if constexpr (!LIN2) dn = fmaf(r.oz, L.hz, dn);
  const float dr = fmaf(-dn, L.c4n, fmaf(-num, L.dc4, L.dr0));
  const float step = fmaf(num, dr, dn * (rden + dr));
  const float f0 = NOFRACT ? rev : __builtin_amdgcn_fractf(rev);
  const float f1 = NOFRACT ? step : __builtin_amdgcn_fractf(step);
  float c = __builtin_amdgcn_cosf(f0), s = __builtin_amdgcn_sinf(f0);
  const float rc = __builtin_amdgcn_cosf(f1), rs = __builtin_amdgcn_sinf(f1);
  const float t2 = rc + rc;
  float cp = 0.f, sp = 0.f;  // slot i - 1
#pragma unroll
  for (int i = 0; i < P; ++i) {
    if (i == 1) {
      cp = c; sp = s;
      c = fmaf(cp, rc, -(sp * rs));
      s = fmaf(cp, rs, sp * rc);
    } else if (i > 1) {
      const float cn = fmaf(t2, c, -cp), sn = fmaf(t2, s, -sp);
      cp = c; sp = s;
      c = cn; s = sn;
    }
    float cc = c, ss = s;
    if constexpr (DEV) {
      float th = r.ox * dw[i][0];
      th = fmaf(r.oy, dw[i][1], th);
      if constexpr (!LIN2) th = fmaf(r.oz, dw[i][2], th);
      cc = fmaf(-th, s, c);
      ss = fmaf(th, c, s);
    }
    a[i][0] = fmaf(r.cs, cc, a[i][0]);
    a[i][1] = fmaf(r.cs, ss, a[i][1]);
    if constexpr (EP) {
      a[i][2] = fmaf(r.pr, cc, a[i][2]);
      a[i][2] = fmaf(-r.pi, ss, a[i][2]);
      a[i][3] = fmaf(r.pr, ss, a[i][3]);
      a[i][3] = fmaf(r.pi, cc, a[i][3]);
    }
  }
}

// Slots first .. first + PH - 1 of each lane of one work item (the wave's
// tile holds P slots per lane).
template <int P, int PH, int W, bool TILE, bool UNIN, bool NOFRACT, bool LIN2, bool PAX,
          bool SEP, bool LINE, bool DEV, bool EP>
__device__ __forceinline__ void f32r_pass(const KInst& I, const WorkItem& w, TileRay* lds,
                                          uint32_t lane, uint32_t wave, uint32_t first) {
  const uint32_t npad = I.npad;
  const uint32_t base = w.pix0 + wave * 64u * P;  // first slot of this wave's tile
  const uint32_t sbase = base + first * 64u;      // first slot of this pass
  const double kr = I.kr_hi;
  const TileRec T = ((const XRTK_CONST TileRec*)I.tiles)[base / (64u * P)];
  const float ikr = (float)(1.0 / kr);
  const float h = (float)(1.0 / (2.0 * T.R * kr));
  SlotRegs<PH> s;
  float mtx = 0.f, mty = 0.f, mtz = 0.f;
  if constexpr (!TILE) {
    mtx = (float)(-2.0 * kr * T.qx); mty = (float)(-2.0 * kr * T.qy);
    mtz = (float)(-2.0 * kr * T.qz);
  }
#pragma unroll
  for (int i = 0; i < (LINE ? 1 : PH); ++i) {  // the line form reads slot 0 only
    const PixRec32 pr = I.pix32[sbase + i * 64u + lane];
    s.mx[i] = TILE ? pr.mx : pr.mx + mtx;
    s.my[i] = TILE ? pr.my : pr.my + mty;
    s.mz[i] = TILE ? pr.mz : pr.mz + mtz;
    s.R2[i] = pr.R2; s.R[i] = pr.R; s.g[i] = pr.g; s.nq[i] = pr.nq;
    s.r0[i] = pr.r0; s.c4[i] = pr.c4; s.rinv[i] = pr.rinv; s.c2[i] = pr.c2;
  }
  LineRegs L{};
  double hv[3] = {0.0, 0.0, 0.0};  // the lane's step along its slot line
  if constexpr (LINE) {
    // the line: the secant from the lane's slot 0 to its last real slot
    // (the statistics tile prep bounded), anchored at the pass's first slot
    // a; slot a's record and the step from q_a to q_a + h
    L.mx = s.mx[0]; L.my = s.my[0]; L.mz = s.mz[0];
    L.g = s.g[0]; L.r0 = s.r0[0]; L.c4 = s.c4[0];
    LineRec lr;
    if constexpr (!DEV) {
      // one pass: anchored at slot 0, the record tile prep wrote
      const uint32_t nt = npad / (64u * P);
      lr = ((const XRTK_CONST LineRec*)(I.tiles + nt))[(base / (64u * P)) * 64u + lane];
    } else {
      const XRTK_CONST PixF64* px = (const XRTK_CONST PixF64*)I.pix64;
      const InstScratch* sc = (const InstScratch*)I.scr;
      line_secant<P>(px, I.perm, I.np, base + lane, hv);
      int ax[3];
      kernel_axes(sc, ax);
      const uint32_t sa = sbase + lane;
      const PixF64 pa = px[sa];
      const double qa[3] = {pa.x - sc->cx, pa.y - sc->cy, pa.z - sc->cz};
      lr = line_step(qa, ((const XRTK_CONST double*)I.R64)[2 * sa], hv, kr, ax);
    }
    L.hx = lr.hx; L.hy = lr.hy; L.hz = lr.hz;
    L.dr0 = lr.dr0; L.dc4 = lr.dc4; L.c4n = lr.c4n;
  }
  // DEV: each slot's offset from the line and its phase gradient (radians
  // per unit o, kernel axis order): w_i = -2 pi kr (d_i - q^ (q^.d_i)) / |q_i|
  float dw[PH][3];
#pragma unroll
  for (int i = 0; i < PH; ++i) dw[i][0] = dw[i][1] = dw[i][2] = 0.f;
  if constexpr (LINE && DEV) {
    const InstScratch* sc = (const InstScratch*)I.scr;
    const uint32_t s0 = sbase + lane;
    const PixF64 p0 = ((const XRTK_CONST PixF64*)I.pix64)[s0];
    int ax[3];
    kernel_axes(sc, ax);
#pragma unroll
    for (int i = 1; i < PH; ++i) {
      const uint32_t si = s0 + 64u * i;
      const PixF64 pi = ((const XRTK_CONST PixF64*)I.pix64)[si];
      const double d[3] = {(pi.x - p0.x) - i * hv[0], (pi.y - p0.y) - i * hv[1],
                           (pi.z - p0.z) - i * hv[2]};
      const double R = ((const XRTK_CONST double*)I.R64)[2 * si];
      const double q[3] = {(pi.x - sc->cx) / R, (pi.y - sc->cy) / R, (pi.z - sc->cz) / R};
      const double qd = q[0] * d[0] + q[1] * d[1] + q[2] * d[2];
      const double f = -2.0 * kPI * kr / R;
      const double w[3] = {f * (d[0] - q[0] * qd), f * (d[1] - q[1] * qd), f * (d[2] - q[2] * qd)};
      dw[i][0] = (float)pick3(w, ax[0]);
      dw[i][1] = (float)pick3(w, ax[1]);
      dw[i][2] = (float)pick3(w, ax[2]);
    }
  }
  // FP32 sums over this item's ray chunk (at most I.tile rays); the chunk
  // partials are summed in FP64 by k_merge.
  float a[PH][4];
#pragma unroll
  for (int i = 0; i < PH; ++i) a[i][0] = a[i][1] = a[i][2] = a[i][3] = 0.f;

  const XRTK_CONST RayRec32* rays = (const XRTK_CONST RayRec32*)I.rays32;
  const uint32_t r0 = w.chunk * I.chunk_len;
  const uint32_t r1 = min(r0 + I.chunk_len, I.nr);
  if constexpr (TILE) {
    TileRay* my = lds + wave * kBlockRays;
    for (uint32_t b0 = r0; b0 < r1; b0 += kBlockRays) {
      const uint32_t nb = min(kBlockRays, r1 - b0);
      // lane-parallel per-(ray, tile) values, one ray per lane: the phase
      // A = kr (|o|^2 - 2 o.q_t) / (2 |q_t|) is reduced in FP64 and its
      // fractional revolution joins the ray's phase offset
      __syncthreads();
      if (lane < nb) {
        const RayRec32 rr = I.rays32[b0 + lane];
        const double ox = rr.ox, oy = rr.oy, oz = rr.oz;
        const double Bt = (ox * ox + oy * oy + oz * oz) -
                          2.0 * (ox * T.qx + oy * T.qy + oz * T.qz);
        const double A = kr * Bt / (2.0 * T.R);  // revolutions
        // phase of the FP32 rounding delta of o towards this tile
        const double dx = rr.dx, dy = rr.dy, dz = rr.oo;
        const double corr = kr * (dx * (T.qx - ox) + dy * (T.qy - oy) + dz * (T.qz - oz)) /
                            sqrt(T.R * T.R + Bt);
        TileRay tr;
        tr.B = (float)(kr * Bt);
        tr.off = (float)((A - rint(A)) + ((double)rr.off + corr));
        // am n.(q_t - o), with am n = -2 kr n' (the record stores n' = -am n/(2 kr))
        tr.n = UNIN ? -rr.no
                    : (float)(-2.0 * kr * ((double)rr.nx * T.qx + (double)rr.ny * T.qy +
                                           (double)rr.nz * T.qz) - (double)rr.no);
        tr.pad = 0.f;
        my[lane] = tr;
      }
      __syncthreads();
      if constexpr (LINE) {
        // the ray loop starts on a 64-byte instruction line: a backward
        // branch to the last bytes of a line costs a fetch on every trip
        // (~1.5 % of this loop's time, measured)
        asm volatile(".p2align 6");
#pragma unroll 2
        for (uint32_t jj = 0; jj < nb; ++jj) {
          const RayRec32 r = rays[b0 + jj];
          const TileRay tr = my[jj];
          pair_line<PH, NOFRACT, LIN2, DEV, EP>(r, tr.B, tr.off, L, h, a, dw);
        }
      } else {
        asm volatile(".p2align 6");  // (as above)
#pragma unroll 2
        for (uint32_t jj = 0; jj < nb; ++jj) {
          const RayRec32 r = rays[b0 + jj];
          const TileRay tr = my[jj];
          pair_ray<PH, true, UNIN, NOFRACT, LIN2, PAX, SEP, EP>(r, tr.B, tr.off, tr.n, s, ikr, h,
                                                                a);
        }
      }
    }
  } else {
#pragma unroll 2
    for (uint32_t j = r0; j < r1; ++j) {
      const RayRec32 r = rays[j];
      pair_ray<PH, false, UNIN, NOFRACT, LIN2, false, false, EP>(r, 0.f, 0.f, 0.f, s, ikr, h, a);
    }
  }
  // (SEP: k_merge applies the slot factor of the amplitude)
  float4* part = reinterpret_cast<float4*>(I.partial) + (size_t)w.chunk * npad;
#pragma unroll
  for (int i = 0; i < PH; ++i) {
    const uint32_t p = sbase + i * 64u + lane;
    part[p] = make_float4(a[i][0], a[i][1], EP ? a[i][2] : 0.f, EP ? a[i][3] : 0.f);
  }
}

// One work item: the wave's 64 P slots, in one pass for the slot-line form
// of the lean kernel and in passes of at most 8 slots per lane for the other
// forms (so that their registers at P = 16 are those of P = 8; each pass
// reads the ray chunk again).  (The slot line with per-slot offsets in one
// pass of 16 would need ~256 VGPRs.)
template <int P, int W, bool TILE, bool UNIN, bool NOFRACT, bool LIN2, bool PAX, bool SEP,
          bool LINE = false, bool DEV = false, bool EP = true>
__device__ __forceinline__ void f32r_item(const KInst& I, const WorkItem& w, TileRay* lds,
                                          uint32_t lane, uint32_t wave) {
  constexpr int PH = ((LINE && !DEV) || P <= 8) ? P : 8;
  static_assert(P % PH == 0, "passes must tile the lane's slots");
#pragma unroll 1
  for (uint32_t first = 0; first < (uint32_t)P; first += PH)
    f32r_pass<P, PH, W, TILE, UNIN, NOFRACT, LIN2, PAX, SEP, LINE, DEV, EP>(I, w, lds, lane, wave,
                                                                             first);
}

// Whether an instance runs the slot-line form at this P.
template <int P>
__device__ __forceinline__ bool line_variant(uint32_t f) {
  constexpr uint32_t need = INST_F_TILE | INST_F_PAX | INST_F_UNIFORM_N | INST_F_SEPAMP |
                            INST_F_LINE;
  return P >= 4 && (f & need) == need && !(f & INST_F_LINEDEV);
}

// The slot-line form with per-slot offsets (run by the other kernel, whose
// register budget already holds P slot records).
template <int P>
__device__ __forceinline__ bool linedev_variant(uint32_t f) {
  constexpr uint32_t need = INST_F_TILE | INST_F_PAX | INST_F_UNIFORM_N | INST_F_SEPAMP |
                            INST_F_LINE | INST_F_LINEDEV;
  return P >= 4 && (f & need) == need;
}

// Runs one work item with the variant the instance's flags select (PAX only
// exists in the tile-referenced form, SEPAMP only with PAX and a uniform
// normal).  LINE: the slot-line variants only (the lean kernel); otherwise
// every other variant.
template <int P, int W, bool LINE>
__device__ __forceinline__ void f32r_dispatch(uint32_t f, const KInst& I, const WorkItem& w,
                                              TileRay* lds, uint32_t lane, uint32_t wave) {
  // bit 4 of the slot-line variant index: Ep sums present
  const uint32_t ep = (f & INST_F_EPZERO) ? 0u : 4u;
  if constexpr (LINE) {
    const uint32_t v = ep | ((f & INST_F_NOFRACT) ? 2u : 0u) | ((f & INST_F_LIN2) ? 1u : 0u);
    switch (v) {
#define XRTK_L(V_)                                                                               \
  case V_:                                                                                      \
    f32r_item<P, W, true, true, (V_ & 2) != 0, (V_ & 1) != 0, true, true, true, false,           \
              (V_ & 4) != 0>(I, w, lds, lane, wave);                                             \
    break;
      XRTK_L(0) XRTK_L(1) XRTK_L(2) XRTK_L(3) XRTK_L(4) XRTK_L(5) XRTK_L(6) XRTK_L(7)
#undef XRTK_L
    }
    return;
  }
  if constexpr (P >= 4) {
    if (linedev_variant<P>(f)) {
      const uint32_t v = ep | ((f & INST_F_NOFRACT) ? 2u : 0u) | ((f & INST_F_LIN2) ? 1u : 0u);
      switch (v) {
#define XRTK_D(V_)                                                                               \
  case V_:                                                                                      \
    f32r_item<P, W, true, true, (V_ & 2) != 0, (V_ & 1) != 0, true, true, true, true,            \
              (V_ & 4) != 0>(I, w, lds, lane, wave);                                             \
    break;
        XRTK_D(0) XRTK_D(1) XRTK_D(2) XRTK_D(3) XRTK_D(4) XRTK_D(5) XRTK_D(6) XRTK_D(7)
#undef XRTK_D
      }
      return;
    }
  }
  const bool tile = (f & INST_F_TILE) != 0u;
  const bool pax = tile && (f & INST_F_PAX);
  const bool sep = pax && (f & INST_F_UNIFORM_N) && (f & INST_F_SEPAMP);
  const uint32_t v = (tile ? 8u : 0u) | ((f & INST_F_UNIFORM_N) ? 4u : 0u) |
                     ((f & INST_F_NOFRACT) ? 2u : 0u) | ((f & INST_F_LIN2) ? 1u : 0u) |
                     (pax ? 16u : 0u) | (sep ? 32u : 0u);
  // The tile-referenced paraxial forms (per-ray normals: 24-27; separable
  // amplitude: 60-63) also come without Ep sums (bit 64); the rest always
  // form them.
  const uint32_t vx = v | ((ep == 0u && (v & 24u) == 24u) ? 64u : 0u);
  switch (vx) {
#define XRTK_V(V_)                                                                      \
  case V_:                                                                             \
    f32r_item<P, W, (V_ & 8) != 0, (V_ & 4) != 0, (V_ & 2) != 0, (V_ & 1) != 0,         \
              (V_ & 16) != 0, (V_ & 32) != 0, false, false, (V_ & 64) == 0>(I, w, lds,  \
                                                                            lane, wave); \
    break;
    XRTK_V(0) XRTK_V(1) XRTK_V(2) XRTK_V(3) XRTK_V(4) XRTK_V(5) XRTK_V(6) XRTK_V(7)
    XRTK_V(8) XRTK_V(9) XRTK_V(10) XRTK_V(11) XRTK_V(12) XRTK_V(13) XRTK_V(14) XRTK_V(15)
    XRTK_V(24) XRTK_V(25) XRTK_V(26) XRTK_V(27) XRTK_V(28) XRTK_V(29) XRTK_V(30) XRTK_V(31)
    XRTK_V(60) XRTK_V(61) XRTK_V(62) XRTK_V(63)
    XRTK_V(88) XRTK_V(89) XRTK_V(90) XRTK_V(91) XRTK_V(92) XRTK_V(93) XRTK_V(94) XRTK_V(95)
    XRTK_V(124) XRTK_V(125) XRTK_V(126) XRTK_V(127)
#undef XRTK_V
  }
}

// LINE: the lean kernel, which runs the slot-line items and skips the rest
// (its register allocation is that of the slot-line code alone); the other
// kernel runs every other item.  Each has its own queue counter.
template <int P, int W, bool PERSIST, bool LINE>
__global__ __launch_bounds__(64 * W) void k_pair_f32r(
    const KInst* __restrict__ insts, const WorkItem* __restrict__ items,
    uint32_t n_items, uint32_t* __restrict__ queue, uint32_t remap) {
  __shared__ TileRay lds[W * kBlockRays];
  __shared__ uint32_t s_item;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  for (;;) {
    uint32_t it;
    if constexpr (PERSIST) {
      __syncthreads();
      if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
      __syncthreads();
      it = uniform(s_item);
    } else {
      it = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    }
    if (it >= n_items) break;
    const WorkItem w = ((const XRTK_CONST WorkItem*)items)[it];
    const KInst I = ((const XRTK_CONST KInst*)insts)[w.inst];
    const uint32_t f = ((const XRTK_CONST InstScratch*)I.scr)->flags;
    // (instances on the matrix cores are k_pair_mfma's)
    if (!(f & INST_F_MFMA) && line_variant<P>(f) == LINE)
      f32r_dispatch<P, W, LINE>(f, I, w, lds, lane, wave);
    if constexpr (!PERSIST) break;
  }
}

// ---------------------------------------------------------------------------
// Block pair kernel on the matrix cores (MFMA variant of the FP32 kernel)
// ---------------------------------------------------------------------------
// Over a block of 64 x 64 grid pixels with centre a, a ray's reduced phase
// is linear in the pixel offsets (u, v) from a, up to a remainder that
// k_plan bounds (<= 1e-6 rev):
//   f(u, v) = ph + u sx + v sz,
// so its term factors into a row factor and a column factor,
//   A e^{2 pi i f} = [A e^{2 pi i (ph + u sx)}] [e^{2 pi i v sz}],
// and the block's field is a complex product of a 64 x K row-factor matrix
// and a K x 64 column-factor matrix over the K rays of the item's chunk.
// ph (with the ray's phase offset and the FP64 rounding correction of o),
// sx and sz are formed per (ray, block) in FP64, by one lane per ray; the
// factors are then generated in FP32 in the layout of the MFMA operands and
// split into FP16 pairs x = hi + lo (both rounded to nearest, so x - hi - lo
// is ~2^-22 |x|); the three products hi.hi + hi.lo + lo.hi run on
// v_mfma_f32_32x32x16_f16 with FP32 accumulation, leaving ~2^-21 per term.
// Amplitudes are scaled by a power of two so the largest is in [1/2, 1).
// Each wave owns one block (2 x 2 tiles of 32 x 32 pixels), one ray chunk
// and one of the two sums: the Es pass (real amplitude cs) writes the
// chunk's partial Es of the block's real pixels, the Ep pass (complex
// amplitude pr + i pi in the row factors, the same column factors) their
// partial Ep, in the slots the F32R kernel would write (k_merge is
// shared).  Rays that all have Ep = 0 take the Es pass alone, which then
// writes Ep = 0.
typedef _Float16 xh8 __attribute__((ext_vector_type(8)));
typedef _Float16 xh2 __attribute__((ext_vector_type(2)));
typedef float xf2 __attribute__((ext_vector_type(2)));
typedef float xf16 __attribute__((ext_vector_type(16)));

static constexpr uint32_t kMfmaBlk = 64;  // block side in pixels

// Per-(ray, block) values, one array each, 64 rays per LDS block.
enum : int {
  MR_PH = 0,  // phase at the centre of the first row tile, revolutions in [-1/2, 1/2]
  MR_SX,      // phase step per row (grid x), mod 1
  MR_GZ,      // phase at the centre of the first column tile relative to a
  MR_SZ,      // phase step per column (grid z), mod 1
  MR_AMP,     // scaled amplitude: cs (Es pass) or pr (Ep pass)
  MR_RXC, MR_RXS,  // e^{2 pi i 32 sx}: first row tile to the second
  MR_RZC, MR_RZS,  // e^{2 pi i 32 sz}
  MR_AMPI,    // Ep pass: scaled pi (the imaginary part of the amplitude)
  MR_N
};

// hi = RNE(x), lo = RNE(x - hi) of two values, packed in one register each;
// m1 is -1 held in a register, so fma(hi, m1, x) rounded to FP16 is one
// v_fma_mixlo/hi_f16.
__device__ __forceinline__ void split2(float a, float b, float m1, uint32_t& hi, uint32_t& lo) {
  const xf2 p = {a, b};
  const xh2 h = __builtin_convertvector(p, xh2);
  const xh2 l = {(_Float16)__builtin_fmaf((float)h.x, m1, a),
                 (_Float16)__builtin_fmaf((float)h.y, m1, b)};
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, l);
}

// The MFMA operands of one 16-ray step: row factors (A, 2 row tiles) as
// re, im and -im, column factors (B, 2 column tiles) as re and im, each as
// an FP16 hi and lo part; 4 registers (8 rays) per part and tile.
struct MFrag {
  uint32_t a[2][6][4];  // [tile][re hi, re lo, im hi, im lo, -im hi, -im lo][pair]
  uint32_t b[2][4][4];  // [tile][re hi, re lo, im hi, im lo][pair]
};

__device__ __forceinline__ xh8 as_h8(const uint32_t (&w)[4]) {
  const uint32_t v[4] __attribute__((aligned(16))) = {w[0], w[1], w[2], w[3]};
  return __builtin_bit_cast(xh8, v);
}

// Offsets of a float32 pixel list from the grid, for this lane's row and
// column (in pixels): ex, ez of tile 0; cx, cz = 2 pi (offset of tile 1 -
// offset of tile 0).
struct LaneOff {
  float ex, ez, cx, cz;
};

// Factors of this lane for the 8 rays kb .. kb + 7 of the LDS block: row
// (grid x) u0 of row tile 0 and, by one rotation, of row tile 1; column u0
// likewise.  OFF: rows and columns at u0 + offset, the rotation to tile 1
// followed by a first-order turn by the difference of the two offsets.
template <bool EPP, bool OFF>
__device__ __forceinline__ void mfma_gen(const float (*lds)[kBlockRays], uint32_t kb, float u0,
                                         float m1, const LaneOff& lo, MFrag& f) {
#pragma unroll
  for (int jp = 0; jp < 4; ++jp) {
    float fr[2][2], fi[2][2], gr[2][2], gi[2][2];  // [tile][ray of the pair]
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const uint32_t j = kb + 2 * jp + q;
      const float ph = lds[MR_PH][j], sx = lds[MR_SX][j];
      const float gz = lds[MR_GZ][j], sz = lds[MR_SZ][j];
      const float amp = lds[MR_AMP][j];
      const float rxc = lds[MR_RXC][j], rxs = lds[MR_RXS][j];
      const float rzc = lds[MR_RZC][j], rzs = lds[MR_RZS][j];
      const float a = fmaf(u0, sx, OFF ? fmaf(lo.ex, sx, ph) : ph);
      float c, s;
      if constexpr (EPP) {
        // (pr + i pi) e^{2 pi i a}
        const float ampi = lds[MR_AMPI][j];
        const float ca = __builtin_amdgcn_cosf(a), sa = __builtin_amdgcn_sinf(a);
        c = fmaf(amp, ca, -(ampi * sa));
        s = fmaf(amp, sa, ampi * ca);
      } else {
        c = __builtin_amdgcn_cosf(a) * amp;
        s = __builtin_amdgcn_sinf(a) * amp;
      }
      fr[0][q] = c; fi[0][q] = s;
      fr[1][q] = fmaf(c, rxc, -(s * rxs));
      fi[1][q] = fmaf(c, rxs, s * rxc);
      const float b = fmaf(u0, sz, OFF ? fmaf(lo.ez, sz, gz) : gz);
      const float gc = __builtin_amdgcn_cosf(b), gs = __builtin_amdgcn_sinf(b);
      gr[0][q] = gc; gi[0][q] = gs;
      gr[1][q] = fmaf(gc, rzc, -(gs * rzs));
      gi[1][q] = fmaf(gc, rzs, gs * rzc);
      if constexpr (OFF) {
        const float tx = lo.cx * sx, tz = lo.cz * sz;
        const float r1 = fr[1][q], g1 = gr[1][q];
        fr[1][q] = fmaf(-tx, fi[1][q], r1);
        fi[1][q] = fmaf(tx, r1, fi[1][q]);
        gr[1][q] = fmaf(-tz, gi[1][q], g1);
        gi[1][q] = fmaf(tz, g1, gi[1][q]);
      }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      split2(fr[t][0], fr[t][1], m1, f.a[t][0][jp], f.a[t][1][jp]);
      split2(fi[t][0], fi[t][1], m1, f.a[t][2][jp], f.a[t][3][jp]);
      f.a[t][4][jp] = f.a[t][2][jp] ^ 0x80008000u;
      f.a[t][5][jp] = f.a[t][3][jp] ^ 0x80008000u;
      split2(gr[t][0], gr[t][1], m1, f.b[t][0][jp], f.b[t][1][jp]);
      split2(gi[t][0], gi[t][1], m1, f.b[t][2][jp], f.b[t][3][jp]);
    }
  }
}

__device__ __forceinline__ xf16 mfma3(xh8 ah, xh8 al, xh8 bh, xh8 bl, xf16 c) {
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, c, 0, 0, 0);
  return c;
}

// The 48 products of one step: re += Fr Gr - Fi Gi, im += Fr Gi + Fi Gr
// per pair of tiles, each real product as hi.hi + hi.lo + lo.hi.
__device__ __forceinline__ void mfma_step(const MFrag& f, xf16 (&acc)[2][2][2]) {
#pragma unroll
  for (int tr = 0; tr < 2; ++tr) {
    const xh8 rh = as_h8(f.a[tr][0]), rl = as_h8(f.a[tr][1]);
    const xh8 ih = as_h8(f.a[tr][2]), il = as_h8(f.a[tr][3]);
    const xh8 nh = as_h8(f.a[tr][4]), nl = as_h8(f.a[tr][5]);
#pragma unroll
    for (int tc = 0; tc < 2; ++tc) {
      const xh8 grh = as_h8(f.b[tc][0]), grl = as_h8(f.b[tc][1]);
      const xh8 gih = as_h8(f.b[tc][2]), gil = as_h8(f.b[tc][3]);
      acc[tr][tc][0] = mfma3(rh, rl, grh, grl, acc[tr][tc][0]);
      acc[tr][tc][0] = mfma3(nh, nl, gih, gil, acc[tr][tc][0]);
      acc[tr][tc][1] = mfma3(rh, rl, gih, gil, acc[tr][tc][1]);
      acc[tr][tc][1] = mfma3(ih, il, grh, grl, acc[tr][tc][1]);
    }
  }
}

template <bool EPP, bool OFF>
__device__ __forceinline__ void mfma_item(const KInst& I, const WorkItem& w,
                                          float (*lds)[MR_N][kBlockRays], uint32_t lane,
                                          float m1) {
  const InstScratch* sc = (const InstScratch*)I.scr;
  const uint32_t nbz = (I.gnz + kMfmaBlk - 1u) / kMfmaBlk;
  const uint32_t bx = w.pix0 / nbz, bz = w.pix0 - bx * nbz;
  const uint32_t ix0 = bx * kMfmaBlk, iz0 = bz * kMfmaBlk;
  const double kr = I.kr_hi;
  // block centre relative to the centroid (world axes)
  const double qa[3] = {(I.gx0 + (ix0 + 31.5) * I.gsx) - sc->cx, I.gy - sc->cy,
                        (I.gz0 + (iz0 + 31.5) * I.gsz) - sc->cz};
  const double Ra = sqrt(qa[0] * qa[0] + qa[1] * qa[1] + qa[2] * qa[2]);
  const double iRa = 1.0 / Ra;
  int ax[3];
  kernel_axes(sc, ax);
  const float amax = __uint_as_float(sc->amax_bits);
  int ex = 0;
  if (amax > 0.f) (void)frexpf(amax, &ex);
  const float scale = ldexpf(1.f, -ex), unscale = ldexpf(1.f, ex);

  const uint32_t half = lane >> 5;
  const float u0 = (float)(lane & 31u) - 15.5f;  // offset from a tile's centre
  LaneOff lo = {0.f, 0.f, 0.f, 0.f};
  if (OFF && I.gex) {
    // this lane's row ix0 + m and column iz0 + m (m = lane & 31) of tile 0
    // and of tile 1 (32 further; rows past the screen take offset 0)
    const uint32_t m = lane & 31u;
    auto off = [](const float* e, uint32_t i, uint32_t n) { return i < n ? e[i] : 0.f; };
    const float x0 = off(I.gex, ix0 + m, I.gnx), x1 = off(I.gex, ix0 + 32u + m, I.gnx);
    const float z0 = off(I.gez, iz0 + m, I.gnz), z1 = off(I.gez, iz0 + 32u + m, I.gnz);
    lo = LaneOff{x0, z0, 6.283185307179586f * (x1 - x0), 6.283185307179586f * (z1 - z0)};
  }
  xf16 acc[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[a][b][c][i] = 0.f;

  const RayRec32* rays = I.rays32;  // one ray per lane: vector loads
  const uint32_t r0 = w.chunk * I.chunk_len;
  const uint32_t r1 = min(r0 + I.chunk_len, I.nr);
  // The fields of one ray record that the block values need, loaded one LDS
  // block ahead of their use so the load has four steps to arrive.
  struct RawRay {
    float ox, oy, oz, oo, amp, ampi, off, dx, dy;
    bool ok;
  };
  auto load = [&](uint32_t b0) {
    RawRay q;
    q.ok = b0 + lane < r1;
    const RayRec32* p = rays + (q.ok ? b0 + lane : r0);  // in bounds either way
    const float4 o = *reinterpret_cast<const float4*>(&p->ox);
    q.ox = o.x; q.oy = o.y; q.oz = o.z; q.oo = o.w;
    q.amp = EPP ? p->pr : p->cs;
    q.ampi = EPP ? p->pi : 0.f;
    q.off = p->off; q.dx = p->dx; q.dy = p->dy;
    return q;
  };
  // per-(ray, block) values of the 64 rays of `rr`, one ray per lane, in
  // FP64, into LDS buffer `buf` (zero amplitude past r1)
  auto prologue = [&](const RawRay& rr, uint32_t buf) {
    float v[MR_N] = {0.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 1.f, 0.f, 0.f};
    if (rr.ok) {
      const float ok[3] = {rr.ox, rr.oy, rr.oz};
      const float dk[3] = {rr.dx, rr.dy, rr.oo};  // TILE: oo holds the third delta
      double o[3], dl[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) { o[ax[k]] = ok[k]; dl[ax[k]] = dk[k]; }
      const double Bq = (o[0] * o[0] + o[1] * o[1] + o[2] * o[2]) -
                        2.0 * (o[0] * qa[0] + o[1] * qa[1] + o[2] * qa[2]);
      const double path = sqrt(Ra * Ra + Bq);
      const double ip = 1.0 / path, iden = 1.0 / (path + Ra);
      // kr (|a - o| - |a|), the phase of the FP32 rounding delta of o and
      // the ray's own offset
      const double corr = kr * (dl[0] * (qa[0] - o[0]) + dl[1] * (qa[1] - o[1]) +
                                dl[2] * (qa[2] - o[2])) * ip;
      const double g = -Bq * iden * ip * iRa;  // 1/|a - o| - 1/|a|
      // d/dt of |a + t s - o| - |a + t s| = (a.s) g - (o.s) / |a - o|
      const double sx = kr * I.gsx * (qa[0] * g - o[0] * ip);
      const double sz = kr * I.gsz * (qa[2] * g - o[2] * ip);
      // The factors take the steps mod 1, fx = sx - nx: row m of a tile
      // (u0 = m - 15.5 from its centre) has phase u0 sx = u0 fx + m nx -
      // 15.5 nx, m nx an integer, so -15.5 nx goes into the anchor phase.
      // The FP32 phase then stays within 1/2 + 15.5/2 rev at any slope.
      const double nx = rint(sx), nz = rint(sz);
      const double fx = sx - nx, fz = sz - nz;
      double ph = kr * Bq * iden + ((double)rr.off + corr) - 16.0 * fx - 15.5 * nx;
      ph -= rint(ph);
      double gz = -16.0 * fz - 15.5 * nz;
      gz -= rint(gz);
      double tx = 32.0 * fx, tz = 32.0 * fz;
      tx -= rint(tx);
      tz -= rint(tz);
      v[MR_PH] = (float)ph; v[MR_SX] = (float)fx;
      v[MR_GZ] = (float)gz; v[MR_SZ] = (float)fz;
      v[MR_AMP] = rr.amp * scale;
      v[MR_AMPI] = rr.ampi * scale;
      v[MR_RXC] = __builtin_amdgcn_cosf((float)tx); v[MR_RXS] = __builtin_amdgcn_sinf((float)tx);
      v[MR_RZC] = __builtin_amdgcn_cosf((float)tz); v[MR_RZS] = __builtin_amdgcn_sinf((float)tz);
    }
#pragma unroll
    for (int k = 0; k < MR_N; ++k) lds[buf][k][lane] = v[k];
  };
  // steps of 16 rays over the chunk (4 per LDS block of 64; the last block
  // rounds up to an even count, its missing rays with zero amplitude),
  // software-pipelined in pairs: the factors of the next step are generated
  // (VALU) while this step's products run (MFMA), one MFMA then five VALU
  // instructions at a time.  The LDS blocks are double-buffered: block b + 1
  // is prepared while block b's steps run.
  const uint32_t nblk = (r1 - r0 + kBlockRays - 1u) / kBlockRays;
  const uint32_t nlast = r1 - r0 - (nblk - 1u) * kBlockRays;
  const uint32_t nsteps = 4u * (nblk - 1u) + (nlast + 31u) / 32u * 2u;
  auto step_ptr = [&](uint32_t k) {
    return (const float(*)[kBlockRays])lds[(k >> 2) & 1u];
  };
  auto step_kb = [&](uint32_t k) { return (k & 3u) * 16u + 8u * half; };
  prologue(load(r0), 0u);
  RawRay next = load(r0 + kBlockRays);
  __syncthreads();
  MFrag f0, f1;
  mfma_gen<EPP, OFF>(step_ptr(0u), step_kb(0u), u0, m1, lo, f0);
  for (uint32_t ks = 0; ks < nsteps; ks += 2u) {
    if ((ks & 3u) == 0u && (ks >> 2) + 1u < nblk) {
      prologue(next, ((ks >> 2) + 1u) & 1u);
      next = load(r0 + ((ks >> 2) + 2u) * kBlockRays);
      __syncthreads();
    }
    mfma_gen<EPP, OFF>(step_ptr(ks + 1u), step_kb(ks + 1u), u0, m1, lo, f1);
    mfma_step(f0, acc);
    __builtin_amdgcn_sched_group_barrier(0x100, 18, 0);  // the LDS reads first
#pragma unroll
    for (int k = 0; k < 48; ++k) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
      __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);  // then VALU
    }
    const uint32_t kn = ks + 2u < nsteps ? ks + 2u : ks + 1u;  // (the last is unused)
    mfma_gen<EPP, OFF>(step_ptr(kn), step_kb(kn), u0, m1, lo, f0);
    mfma_step(f1, acc);
    __builtin_amdgcn_sched_group_barrier(0x100, 18, 1);
#pragma unroll
    for (int k = 0; k < 48; ++k) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 1);
      __builtin_amdgcn_sched_group_barrier(0x002, 5, 1);
    }
  }
  // partial sums of the block's real pixels, in the grid's slot order
  const uint32_t th = I.gth, tw = I.gtw, tsz = th * tw;
  const uint32_t ntz = (I.gnz + tw - 1u) / tw;
  float4* part = reinterpret_cast<float4*>(I.partial) + (size_t)w.chunk * I.npad;
  const uint32_t n = lane & 31u;
  // the Es pass writes Ep = 0 as well when every ray has Ep = 0 (no Ep pass
  // runs); otherwise each pass writes its own half of the slot's partial
  const bool ep0 = (sc->flags & INST_F_EPZERO) != 0u;
#pragma unroll
  for (int tr = 0; tr < 2; ++tr)
#pragma unroll
    for (int tc = 0; tc < 2; ++tc)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const uint32_t m = 8u * (i >> 2) + 4u * half + (i & 3);
        const uint32_t ix = ix0 + 32u * tr + m, iz = iz0 + 32u * tc + n;
        if (ix < I.gnx && iz < I.gnz) {
          const uint32_t slot = ((ix / th) * ntz + iz / tw) * tsz + (ix % th) * tw + iz % tw;
          const float re = acc[tr][tc][0][i] * unscale, im = acc[tr][tc][1][i] * unscale;
          float2* p2 = reinterpret_cast<float2*>(part + slot);
          if constexpr (EPP) {
            p2[1] = make_float2(re, im);
          } else if (ep0) {
            part[slot] = make_float4(re, im, 0.f, 0.f);
          } else {
            p2[0] = make_float2(re, im);
          }
        }
      }
}

// Persistent: one wave per group takes items from the queue until it runs
// dry; items of instances without the MFMA flag are skipped.  EPP: the Ep
// pass (its items, pad = 1), which instances with Ep = 0 everywhere skip;
// each pass is its own kernel, so each has the registers of its own code.
// OFF: the kernel of float32 pixel lists on a grid (instances with offsets;
// null offsets are zeros).
template <bool EPP, bool OFF>
__global__ __launch_bounds__(64) void k_pair_mfma(const KInst* __restrict__ insts,
                                                  const WorkItem* __restrict__ items,
                                                  uint32_t n_items, uint32_t* __restrict__ queue,
                                                  float m1) {
  __shared__ float lds[2][MR_N][kBlockRays];
  const uint32_t lane = threadIdx.x;
  for (;;) {
    const uint32_t v = lane == 0u ? atomicAdd(queue, 1u) : 0u;
    const uint32_t it = __builtin_amdgcn_readfirstlane(v);
    if (it >= n_items) break;
    const WorkItem w = ((const XRTK_CONST WorkItem*)items)[it];
    const KInst I = ((const XRTK_CONST KInst*)insts)[w.inst];
    const uint32_t f = ((const XRTK_CONST InstScratch*)I.scr)->flags;
    if (!(f & INST_F_MFMA)) continue;
    if constexpr (EPP) {
      if (!(f & INST_F_EPZERO)) mfma_item<true, OFF>(I, w, lds, lane, m1);
    } else {
      mfma_item<false, OFF>(I, w, lds, lane, m1);
    }
  }
}

// ---------------------------------------------------------------------------
// plain FP64 pair kernel (the device oracle and the FP64 speed baseline)
// ---------------------------------------------------------------------------
template <int P, bool CHECK>
__device__ __forceinline__ void f64_item(const KInst& I, const WorkItem& w,
                                         uint32_t lane, uint32_t wave) {
#pragma clang fp contract(off)
  const uint32_t npad = I.npad;
  const uint32_t base = w.pix0 + wave * 64u * P;
  double px[P], py[P], pz[P];
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const PixF64 q = I.pix64[base + i * 64u + lane];
    px[i] = q.x; py[i] = q.y; pz[i] = q.z;
  }
  double sum[P][4], cmp[P][4];
#pragma unroll
  for (int i = 0; i < P; ++i)
    for (int c = 0; c < 4; ++c) sum[i][c] = cmp[i][c] = 0.0;

  const uint32_t r0 = w.chunk * I.chunk_len;
  const uint32_t r1 = min(r0 + I.chunk_len, I.nr);
  const double skip = I.skip;
  const double eov = I.energy;
  const bool pathup = ((const XRTK_CONST InstScratch*)I.scr)->flags & INST_F_PATHUP;
  const XRTK_CONST RayRec64* rays = (const XRTK_CONST RayRec64*)I.rays64;
  for (uint32_t j = r0; j < r1; ++j) {
    const RayRec64 ray = rays[j];
    const double energy = eov > 0.0 ? eov : ray.energy;
    // k = E / CHBAR * 1e7 [mm^-1]                      (diffraction.rs:89)
    const double k = energy / kCHBAR * 1e7;
    const double kf = k / (4.0 * kPI);
    double er = ray.es_re, ei = ray.es_im, qr = ray.ep_re, qi = ray.ep_im;
    if (pathup) {
      double s, c;
      sincos(k * ray.path_up, &s, &c);
      const double er2 = er * c - ei * s, ei2 = er * s + ei * c;
      const double qr2 = qr * c - qi * s, qi2 = qr * s + qi * c;
      er = er2; ei = ei2; qr = qr2; qi = qi2;
    }
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const double dx = px[i] - ray.x;
      const double dy = py[i] - ray.y;
      const double dz = pz[i] - ray.z;
      const double path = sqrt(dx * dx + dy * dy + dz * dz);
      if (CHECK && path < skip) continue;                // diffraction.rs:81
      const double ns = (ray.nx * dx + ray.ny * dy + ray.nz * dz) / path;
      const double phase = k * path;
      double sp, cp;
      sincos(phase, &sp, &cp);
      const double obl = ray.nl + ns;
      const double amp = kf * obl / path;
      // u = i * amp * exp(i phase)                      (diffraction.rs:98)
      const double ur = -(amp * sp), ui = amp * cp;
      // Kahan, diffraction.rs:100-110
      const double t0 = (er * ur - ei * ui) - cmp[i][0];
      const double t1 = (er * ui + ei * ur) - cmp[i][1];
      const double t2 = (qr * ur - qi * ui) - cmp[i][2];
      const double t3 = (qr * ui + qi * ur) - cmp[i][3];
      const double n0 = sum[i][0] + t0, n1 = sum[i][1] + t1;
      const double n2 = sum[i][2] + t2, n3 = sum[i][3] + t3;
      cmp[i][0] = (n0 - sum[i][0]) - t0; cmp[i][1] = (n1 - sum[i][1]) - t1;
      cmp[i][2] = (n2 - sum[i][2]) - t2; cmp[i][3] = (n3 - sum[i][3]) - t3;
      sum[i][0] = n0; sum[i][1] = n1; sum[i][2] = n2; sum[i][3] = n3;
    }
  }
  double* part = reinterpret_cast<double*>(I.partial) + (size_t)w.chunk * npad * 4u;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const uint32_t p = base + i * 64u + lane;
    double2* d = reinterpret_cast<double2*>(part + (size_t)p * 4u);
    d[0] = make_double2(sum[i][0], sum[i][1]);
    d[1] = make_double2(sum[i][2], sum[i][3]);
  }
}

template <int P, int W, bool PERSIST>
__global__ __launch_bounds__(64 * W) void k_pair_f64(
    const KInst* __restrict__ insts, const WorkItem* __restrict__ items,
    uint32_t n_items, uint32_t* __restrict__ queue, uint32_t remap) {
  __shared__ uint32_t s_item;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  for (;;) {
    uint32_t it;
    if constexpr (PERSIST) {
      __syncthreads();
      if (threadIdx.x == 0) s_item = atomicAdd(queue, 1u);
      __syncthreads();
      it = uniform(s_item);
    } else {
      it = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    }
    if (it >= n_items) break;
    const WorkItem w = ((const XRTK_CONST WorkItem*)items)[it];
    const KInst I = ((const XRTK_CONST KInst*)insts)[w.inst];
    if (((const XRTK_CONST InstScratch*)I.scr)->flags & INST_F_CHECK)
      f64_item<P, true>(I, w, lane, wave);
    else
      f64_item<P, false>(I, w, lane, wave);
    if constexpr (!PERSIST) break;
  }
}

// ---------------------------------------------------------------------------
// merge: compensated FP64 sum of the chunk partials (FP32 partials of the
// reduced kernel, FP64 of the plain kernel) in chunk order, then the
// base phase (FP32R): out = k/(4 pi) exp(i k |p - c|) * sum (k|p - c| as a
// double-double product), written to the
// output index of the slot.  grid = (ceil(max npad / 256), n_list).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_merge(const KInst* __restrict__ insts,
                                               const uint32_t* __restrict__ list) {
  #pragma clang fp contract(off)
  const KInst& I = insts[list[blockIdx.y]];
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= I.npad) return;
  const uint32_t dst = I.perm ? I.perm[p] : (p < I.np ? p : ~0u);
  if (dst == ~0u) return;
  double s[4] = {0, 0, 0, 0}, e[4] = {0, 0, 0, 0};
  const bool f32 = I.mode == MODE_F32R;
  for (uint32_t c = 0; c < I.n_chunks; ++c) {
    double v[4];
    if (f32) {
      const float4 q = reinterpret_cast<const float4*>(I.partial)[(size_t)c * I.npad + p];
      v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
      const double2* q = reinterpret_cast<const double2*>(
          reinterpret_cast<const double*>(I.partial) + ((size_t)c * I.npad + p) * 4u);
      const double2 v01 = q[0], v23 = q[1];
      v[0] = v01.x; v[1] = v01.y; v[2] = v23.x; v[3] = v23.y;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {  // TwoSum (Knuth)
      const double t = s[k] + v[k];
      const double bp = t - s[k];
      e[k] += (s[k] - (t - bp)) + (v[k] - bp);
      s[k] = t;
    }
  }
  double o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) o[k] = s[k] + e[k];
  if (I.mode == MODE_F32R) {
    double c, sn;
    cis_rev_dd2(I.kr_hi, I.kr_lo, I.R64[2 * p], I.R64[2 * p + 1], &c, &sn);
    double kf = I.k / (4.0 * kPI);
    const InstScratch* scr = (const InstScratch*)I.scr;
    if (scr->flags & INST_F_SEPAMP) {
      // the slot factor of the separable amplitude, (nl_ref + C) / |q| with
      // C = n0.q / |q| (the ray factor is in the sums)
      const PixF64 px = I.pix64[p];
      const double R = I.R64[2 * p];
      const double C = (scr->n0x * (px.x - scr->cx) + scr->n0y * (px.y - scr->cy) +
                        scr->n0z * (px.z - scr->cz)) / R;
      kf *= (nl_ref(scr) + C) / R;
    }
    c *= kf; sn *= kf;
    const double a0 = o[0] * c - o[1] * sn, a1 = o[0] * sn + o[1] * c;
    const double a2 = o[2] * c - o[3] * sn, a3 = o[2] * sn + o[3] * c;
    o[0] = a0; o[1] = a1; o[2] = a2; o[3] = a3;
  }
  double2* d = reinterpret_cast<double2*>(I.out + (size_t)dst * 4u);
  if (I.accumulate) {
    const double2 d0 = d[0], d1 = d[1];
    o[0] += d0.x; o[1] += d0.y; o[2] += d1.x; o[3] += d1.y;
  }
  d[0] = make_double2(o[0], o[1]);
  d[1] = make_double2(o[2], o[3]);
}

__global__ void k_reset_u32(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__global__ void k_zero_f64(double* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// host-side launchers (called by the runtime)
// ---------------------------------------------------------------------------
static inline unsigned cdiv(size_t a, unsigned b) { return (unsigned)((a + b - 1) / b); }

hipError_t launch_rays_from_f32aos(const float* src, RayRec64* dst, uint32_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_rays_from_f32aos, dim3(cdiv(n, 256)), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}
hipError_t launch_rays_from_cols(const double* cols, RayRec64* dst, uint32_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_rays_from_cols, dim3(cdiv(n, 256)), dim3(256), 0, st, cols, dst, n);
  return hipGetLastError();
}
hipError_t launch_rays_from_beam(const double* cols, uint32_t n_beam, const double* normals,
                                 const uint64_t* good, uint32_t n_good, bool has_amps,
                                 bool has_path, RayRec64* dst, hipStream_t st) {
  if (!n_good) return hipSuccess;
  hipLaunchKernelGGL(k_rays_from_beam, dim3(cdiv(n_good, 256)), dim3(256), 0, st, cols, n_beam,
                     normals, good, n_good, has_amps ? 1u : 0u, has_path ? 1u : 0u, dst);
  return hipGetLastError();
}
hipError_t launch_pix_from_f32aos(const void* src, PixF64* dst, uint32_t* perm, uint32_t n,
                                  uint32_t npad, uint32_t P, hipStream_t st) {
  if (!n) return hipSuccess;
  if (npad % (64u * P)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pix_from_f32aos, dim3(cdiv(npad, 256)), dim3(256), 0, st,
                     (const float4*)src, dst, perm, n, npad, P);
  return hipGetLastError();
}
hipError_t launch_pix_from_cols(const double* cols, PixF64* dst, uint32_t* perm, uint32_t n,
                                uint32_t npad, uint32_t P, hipStream_t st) {
  if (!n) return hipSuccess;
  if (npad % (64u * P)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pix_from_cols, dim3(cdiv(npad, 256)), dim3(256), 0, st, cols, dst, perm,
                     n, npad, P);
  return hipGetLastError();
}
hipError_t launch_pix_grid(const double c[3], double dx, double dz, uint32_t nx, uint32_t nz,
                           uint32_t th, uint32_t tw, uint32_t npad, PixF64* dst, uint32_t* perm,
                           hipStream_t st) {
  if (!npad) return hipSuccess;
  hipLaunchKernelGGL(k_pix_grid, dim3(cdiv(npad, 256)), dim3(256), 0, st, c[0], c[1], c[2], dx,
                     dz, nx, nz, th, tw, npad, dst, perm);
  return hipGetLastError();
}

hipError_t launch_pix_grid_list(const void* list, uint32_t nx, uint32_t nz, uint32_t th,
                                uint32_t tw, uint32_t npad, PixF64* dst, uint32_t* perm,
                                hipStream_t st) {
  if (!npad) return hipSuccess;
  hipLaunchKernelGGL(k_pix_grid_list, dim3(cdiv(npad, 256)), dim3(256), 0, st,
                     static_cast<const float4*>(list), nx, nz, th, tw, npad, dst, perm);
  return hipGetLastError();
}

// Centroid, bounds, plan and the per-solve FP32 records of one instance.
// part: kRedBlocks * 4 doubles of scratch.  P is the pixels-per-lane of the
// F32R launch (tile = 64 * P slots; npad a multiple of it).
hipError_t launch_prep(const RayRec64* rays, uint32_t nr, const PixF64* pix, uint32_t npad,
                       const uint32_t* perm, uint32_t np, InstScratch* scr, double* part,
                       uint32_t mode, double kr, double kr_lo, double skip, uint32_t force_flags,
                       uint32_t static_flags, int P, double gsx, double gsz, double gemax,
                       RayRec32* rays32, PixRec32* pix32, TileRec* tiles, double* R64,
                       hipStream_t st) {
  const unsigned nb = nr ? std::min<unsigned>(kRedBlocks, cdiv(nr, 256)) : 1u;
  hipLaunchKernelGGL(k_centroid_partial, dim3(nb), dim3(256), 0, st, rays, nr, part);
  hipLaunchKernelGGL(k_centroid_final, dim3(1), dim3(256), 0, st, part, nb, nr, rays, scr);
  if (nr) hipLaunchKernelGGL(k_ray_bounds, dim3(nb), dim3(256), 0, st, rays, nr, scr);
  if (npad) {
    if (mode == MODE_F32R) {
      const unsigned ntile = npad / (64u * P);
      if (P == 2)
        hipLaunchKernelGGL(k_tile_prep<2>, dim3(ntile), dim3(64), 0, st, pix, kr, scr, pix32,
                           tiles, R64, perm, np);
      else if (P == 4)
        hipLaunchKernelGGL(k_tile_prep<4>, dim3(ntile), dim3(64), 0, st, pix, kr, scr, pix32,
                           tiles, R64, perm, np);
      else if (P == 8)
        hipLaunchKernelGGL(k_tile_prep<8>, dim3(ntile), dim3(64), 0, st, pix, kr, scr, pix32,
                           tiles, R64, perm, np);
      else if (P == 16)
        hipLaunchKernelGGL(k_tile_prep<16>, dim3(ntile), dim3(64), 0, st, pix, kr, scr, pix32,
                           tiles, R64, perm, np);
      else
        return hipErrorInvalidValue;
    } else {
      hipLaunchKernelGGL(k_pix_bounds, dim3(cdiv(npad, 256)), dim3(256), 0, st, pix, npad, scr);
    }
  }
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(1), 0, st, scr, skip, kr, mode, force_flags,
                     static_flags, (uint32_t)P, gsx, gsz, gemax);
  if (mode == MODE_F32R && nr)
    hipLaunchKernelGGL(k_ray_prep32, dim3(cdiv(nr, 256)), dim3(256), 0, st, rays, nr, scr, kr,
                       kr_lo, npad ? 1u : 0u, rays32);
  return hipGetLastError();
}

// F32R: the lean slot-line kernel (P >= 4) on queue[1], then the kernel of
// every other variant on queue[0].
template <int P, int W, bool PERSIST>
static void launch_f32r_t(const KInst* insts, const WorkItem* items, uint32_t n_items,
                          uint32_t* queue, uint32_t grid, uint32_t remap, hipStream_t st) {
  if constexpr (P >= 4)
    hipLaunchKernelGGL((k_pair_f32r<P, W, PERSIST, true>), dim3(grid), dim3(64 * W), 0, st,
                       insts, items, n_items, queue + 1, remap);
  hipLaunchKernelGGL((k_pair_f32r<P, W, PERSIST, false>), dim3(grid), dim3(64 * W), 0, st, insts,
                     items, n_items, queue, remap);
}
template <int P, int W, bool PERSIST>
static void launch_f64_t(const KInst* insts, const WorkItem* items, uint32_t n_items,
                         uint32_t* queue, uint32_t grid, uint32_t remap, hipStream_t st) {
  hipLaunchKernelGGL((k_pair_f64<P, W, PERSIST>), dim3(grid), dim3(64 * W), 0, st, insts, items,
                     n_items, queue, remap);
}

// Returns hipErrorInvalidValue for a configuration that is not instantiated.
std::atomic<uint32_t> g_fail_launches{0};

hipError_t launch_pair(uint32_t mode, const LaunchCfg& cfg, const KInst* insts,
                       const WorkItem* items, uint32_t n_items, uint32_t* queue, uint32_t grid,
                       uint32_t remap, hipStream_t st) {
  if (!n_items) return hipSuccess;
  for (uint32_t f = g_fail_launches.load(); f > 0;)
    if (g_fail_launches.compare_exchange_weak(f, f - 1)) return hipErrorLaunchFailure;
  // a non-persistent launch runs one item per group: every item needs one
  if (!cfg.persistent && grid < n_items) return hipErrorInvalidValue;
  if (cfg.persistent) hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(64), 0, st, queue, 2u);
  const bool ps = cfg.persistent != 0;
#define XRTK_F32(P_, W_)                                                                \
  if (cfg.P == P_ && cfg.W == W_) {                                                     \
    if (ps) launch_f32r_t<P_, W_, true>(insts, items, n_items, queue, grid, remap, st); \
    else launch_f32r_t<P_, W_, false>(insts, items, n_items, queue, grid, remap, st);   \
    return hipGetLastError();                                                           \
  }
#define XRTK_F64(P_, W_)                                                               \
  if (cfg.P == P_ && cfg.W == W_) {                                                    \
    if (ps) launch_f64_t<P_, W_, true>(insts, items, n_items, queue, grid, remap, st); \
    else launch_f64_t<P_, W_, false>(insts, items, n_items, queue, grid, remap, st);   \
    return hipGetLastError();                                                          \
  }
  if (mode == MODE_F32R) {
    XRTK_F32(2, 1) XRTK_F32(2, 4) XRTK_F32(4, 1) XRTK_F32(4, 4) XRTK_F32(8, 1) XRTK_F32(8, 4)
    XRTK_F32(16, 1) XRTK_F32(16, 4)
  } else {
    XRTK_F64(1, 1) XRTK_F64(1, 4) XRTK_F64(2, 1) XRTK_F64(2, 4)
  }
#undef XRTK_F32
#undef XRTK_F64
  return hipErrorInvalidValue;
}

hipError_t launch_mfma(const KInst* insts, const WorkItem* items, uint32_t n_items,
                       uint32_t* queue, uint32_t grid, bool ep_pass, bool offsets,
                       hipStream_t st) {
  if (!n_items) return hipSuccess;
  for (uint32_t f = g_fail_launches.load(); f > 0;)
    if (g_fail_launches.compare_exchange_weak(f, f - 1)) return hipErrorLaunchFailure;
  hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(64), 0, st, queue, 1u);
  const dim3 g(std::max<uint32_t>(1u, grid));
  auto* k = ep_pass ? (offsets ? k_pair_mfma<true, true> : k_pair_mfma<true, false>)
                    : (offsets ? k_pair_mfma<false, true> : k_pair_mfma<false, false>);
  hipLaunchKernelGGL(k, g, dim3(64), 0, st, insts, items, n_items, queue, -1.0f);
  return hipGetLastError();
}

bool pair_cfg_supported(uint32_t mode, const LaunchCfg& c) {
  if (c.ray_src != 1) return false;
  if (mode == MODE_F32R)
    return (c.P == 2 || c.P == 4 || c.P == 8 || c.P == 16) && (c.W == 1 || c.W == 4);
  return (c.P == 1 || c.P == 2) && (c.W == 1 || c.W == 4);
}

hipError_t launch_merge(const KInst* insts, const uint32_t* list, uint32_t n_list,
                        uint32_t max_npad, hipStream_t st) {
  if (!n_list || !max_npad) return hipSuccess;
  hipLaunchKernelGGL(k_merge, dim3(cdiv(max_npad, 256), n_list), dim3(256), 0, st, insts, list);
  return hipGetLastError();
}

hipError_t launch_zero_f64(double* p, size_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_zero_f64, dim3(cdiv(n, 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace xrtk
