// gfx950 kernels of the Kirchhoff diffraction engine.
//
// The physics is the per-pair term of src/waves/diffraction.rs:75-111
//   U = i k/(4 pi) (nl + n.d/path) exp(i k path) / path,  Es += es U, Ep += ep U
// evaluated two ways:
//   * k_pair_f64   plain FP64 with Kahan sums, the operation order of
//                  diffraction.rs:76-110 (the FP64 device oracle);
//   * k_pair_f32r  FP32 with the phase reduced against the FP64 ray centroid
//                  c (the method of src/gpu/kirchhoff.rs:96-104,276-297 and
//                  src/gpu/shaders/kirchhoff.wgsl:95-113), redesigned: the
//                  offsets o = r - c and q = p - c are formed in FP64 and
//                  rounded once, the phase is kept in revolutions for the
//                  hardware v_sin/v_cos, FP32 partial sums are folded into
//                  FP64 every `tile` rays, and an optional quadratic fold
//                  moves the per-ray term k|o|^2/(2 R0) into the amplitudes.
// Work is split into items (instance, pixel block, ray chunk); each chunk's
// FP64 partial goes to HBM and k_merge sums the chunks in a fixed order with
// TwoSum, then applies k/(4 pi) exp(i k |p - c|) in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>

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
  const double hi = a_hi * x;
  const double lo = fma(a_hi, x, -hi) + a_lo * x;
  const double n = rint(hi);
  const double f = (hi - n) + lo;  // |f| <= 0.5 + tiny
  sincospi(2.0 * f, s, c);
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

__global__ void k_pix_from_f32aos(const float4* __restrict__ src,
                                  PixF64* __restrict__ dst, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = src[i];
  dst[i] = PixF64{(double)p.x, (double)p.y, (double)p.z, 0.0};
}

__global__ void k_pix_from_cols(const double* __restrict__ cols,
                                PixF64* __restrict__ dst, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t N = n;
  dst[i] = PixF64{cols[i], cols[N + i], cols[2 * N + i], 0.0};
}

// prepare_wave (src/waves/prepare_wave.rs:40-68), evaluated in the same
// operation order without contraction so the points equal the host's.
__global__ void k_pix_grid(double cx, double cy, double cz, double dx,
                           double dz, uint32_t nx, uint32_t nz,
                           PixF64* __restrict__ dst) {
#pragma clang fp contract(off)
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx * nz) return;
  const uint32_t ix = i / nz, iz = i - ix * nz;
  const double fx = nx > 1 ? (double)ix / (double)(nx - 1) - 0.5 : 0.0;
  const double fz = nz > 1 ? (double)iz / (double)(nz - 1) - 0.5 : 0.0;
  const double px = cx + fx * 2.0 * dx;
  const double pz = cz + fz * 2.0 * dz;
  dst[i] = PixF64{px, cy, pz, 0.0};
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
    InstScratch* __restrict__ scr) {
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
    scr->R0 = 0.0;
    scr->omax_bits = 0u;
    scr->rmin_bits = 0x7f800000u;  // +inf
    scr->rmax_bits = 0u;
    scr->flags = 0u;
    scr->ep_nonzero = 0u;
  }
}

__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

// max |r - c| over the rays (rounded up to float) and whether any Ep != 0.
__global__ __launch_bounds__(256) void k_ray_bounds(
    const RayRec64* __restrict__ rays, uint32_t n, InstScratch* __restrict__ scr) {
  const double cx = scr->cx, cy = scr->cy, cz = scr->cz;
  float m = 0.f;
  uint32_t ep = 0u;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double ox = rays[i].x - cx, oy = rays[i].y - cy, oz = rays[i].z - cz;
    const float d = __double2float_ru(sqrt(ox * ox + oy * oy + oz * oz) * (1.0 + 1e-12));
    m = fmaxf(m, d);
    ep |= (rays[i].ep_re != 0.0 || rays[i].ep_im != 0.0) ? 1u : 0u;
  }
  m = wave_max(m);
  const uint32_t epw = __any(ep) ? 1u : 0u;
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&scr->omax_bits, __float_as_uint(m));
    if (epw) atomicOr(&scr->ep_nonzero, 1u);
  }
}

// q = p - c and R = |q| in FP64, the FP32 pixel record, and min/max R.
__global__ __launch_bounds__(256) void k_pix_prep(
    const PixF64* __restrict__ pix, uint32_t n, double kr,
    InstScratch* __restrict__ scr, PixRec32* __restrict__ out32,
    double* __restrict__ R64) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const double cx = scr->cx, cy = scr->cy, cz = scr->cz;
  float rmin = __uint_as_float(0x7f800000u), rmax = 0.f;
  if (i < n) {
    const double qx = pix[i].x - cx, qy = pix[i].y - cy, qz = pix[i].z - cz;
    const double R = sqrt(qx * qx + qy * qy + qz * qz);
    PixRec32 r;
    r.qx = (float)qx; r.qy = (float)qy; r.qz = (float)qz; r.R = (float)R;
    r.mx = (float)(-2.0 * kr * qx);
    r.my = (float)(-2.0 * kr * qy);
    r.mz = (float)(-2.0 * kr * qz);
    r.pad = 0.f;
    if (out32) out32[i] = r;
    R64[i] = R;
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

// Decides the per-instance flags from the device-computed bounds.
//  CHECK: a pair could have path below the skip threshold (or be close enough
//         to singular that FP32 loses the offset), so the skip test runs.
//  QFOLD: the reduced phase would exceed qfold_rev revolutions, so the
//         per-ray quadratic term is folded into the amplitudes.
__global__ void k_plan(InstScratch* __restrict__ scr, double skip, double kr,
                       uint32_t mode, double qfold_rev, uint32_t force_flags,
                       uint32_t static_flags) {
  const double omax = (double)__uint_as_float(scr->omax_bits);
  const double rmin = (double)__uint_as_float(scr->rmin_bits);
  const double rmax = (double)__uint_as_float(scr->rmax_bits);
  uint32_t f = static_flags;
  const double slack = mode == MODE_F32R ? 1e-5 * (rmax + omax) : 0.0;
  if (!(rmin - omax > skip * (1.0 + 1e-9) + slack)) f |= INST_F_CHECK;
  if (mode == MODE_F32R && rmin > 0.0 && kr * omax * omax / (2.0 * rmin) > qfold_rev)
    f |= INST_F_QFOLD;
  if (force_flags & 0x100u) f = (f & ~INST_F_QFOLD) | (force_flags & INST_F_QFOLD);
  if (force_flags & 0x200u) f |= INST_F_CHECK;
  scr->flags = f;
  scr->R0 = rmin > 0.0 ? rmin : 1.0;
}

// FP32 ray records of the phase-reduced kernel.
__global__ __launch_bounds__(256) void k_ray_prep32(
    const RayRec64* __restrict__ rays, uint32_t n, const InstScratch* __restrict__ scr,
    double kr, double kr_lo, RayRec32* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double cx = scr->cx, cy = scr->cy, cz = scr->cz;
  const uint32_t flags = scr->flags;
  const RayRec64 r = rays[i];
  const double ox = r.x - cx, oy = r.y - cy, oz = r.z - cz;
  const double oo = ox * ox + oy * oy + oz * oz;
  double er = r.es_re, ei = r.es_im, pr = r.ep_re, pi = r.ep_im;
  // Phase carried by the amplitude, in revolutions.
  double rev = 0.0;
  if (flags & INST_F_QFOLD) rev += kr * oo / (2.0 * scr->R0);
  if (flags & INST_F_PATHUP) {
    const double hi = kr * r.path_up;
    const double lo = fma(kr, r.path_up, -hi) + kr_lo * r.path_up;
    rev += (hi - rint(hi)) + lo;
  }
  if (rev != 0.0) {
    double s, c;
    sincospi(2.0 * rev, &s, &c);
    const double er2 = er * c - ei * s, ei2 = er * s + ei * c;
    const double pr2 = pr * c - pi * s, pi2 = pr * s + pi * c;
    er = er2; ei = ei2; pr = pr2; pi = pi2;
  }
  RayRec32 o;
  o.ox = (float)ox; o.oy = (float)oy; o.oz = (float)oz;
  o.oo = (float)(kr * oo);
  o.nx = (float)r.nx; o.ny = (float)r.ny; o.nz = (float)r.nz; o.nl = (float)r.nl;
  // i * Es, i * Ep
  o.ar = (float)(-ei); o.ai = (float)er;
  o.br = (float)(-pi); o.bi = (float)pr;
  out[i] = o;
}

// ---------------------------------------------------------------------------
// FP32 phase-reduced pair kernel
// ---------------------------------------------------------------------------
struct SmemRays {
  const XRTK_CONST RayRec32* p;
  __device__ __forceinline__ RayRec32 get(uint32_t j) const { return p[j]; }
  static constexpr bool kScalar = true;
};
struct LdsRays {
  const RayRec32* p;
  __device__ __forceinline__ RayRec32 get(uint32_t j) const { return p[j]; }
  static constexpr bool kScalar = false;
};

template <int P>
struct PixRegs {
  float qx[P], qy[P], qz[P], R[P], mx[P], my[P], mz[P], g[P];
};

// Accumulates rays [j0, j1) of `src` into a[P][4] (FP32).
template <int P, bool CHECK, bool QFOLD, class Src>
__device__ __forceinline__ void pair_span(const Src& src, uint32_t j0, uint32_t j1,
                                          const PixRegs<P>& px, float h,
                                          float skip2, float (&a)[P][4]) {
#pragma unroll 2
  for (uint32_t j = j0; j < j1; ++j) {
    const RayRec32 r = src.get(j);
    float oo = r.oo;
    if constexpr (Src::kScalar && !QFOLD) oo = vcopy(oo);
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const float dx = px.qx[i] - r.ox;
      const float dy = px.qy[i] - r.oy;
      const float dz = px.qz[i] - r.oz;
      const float p2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
      const float inv = __builtin_amdgcn_rsqf(p2);
      const float nd = fmaf(r.nz, dz, fmaf(r.ny, dy, r.nx * dx));
      const float obl = fmaf(nd, inv, r.nl);
      const float rden = __builtin_amdgcn_rcpf(fmaf(p2, inv, px.R[i]));
      float rev;
      if constexpr (QFOLD) {
        // kr [|o|^2 t - 2 o.q] / (path + R),  t = (R0 - R)/R0 - eps/(2 R0)
        const float lin = fmaf(r.oz, px.mz[i], fmaf(r.oy, px.my[i], r.ox * px.mx[i]));
        const float eps = (lin + oo) * rden;  // kr (path - R)
        const float t = fmaf(-eps, h, px.g[i]);
        rev = fmaf(oo, t, lin) * rden;
      } else {
        // kr (|o|^2 - 2 o.q) / (path + R) = kr (path - R)
        const float num = fmaf(r.oz, px.mz[i], fmaf(r.oy, px.my[i], fmaf(r.ox, px.mx[i], oo)));
        rev = num * rden;
      }
      const float f = __builtin_amdgcn_fractf(rev);
      const float s = __builtin_amdgcn_sinf(f);
      const float c = __builtin_amdgcn_cosf(f);
      const float amp = obl * inv;
      float wr = amp * c, wi = amp * s;
      if constexpr (CHECK) {
        const bool skip = !(p2 > skip2);
        wr = skip ? 0.f : wr;
        wi = skip ? 0.f : wi;
      }
      a[i][0] = fmaf(r.ar, wr, a[i][0]);
      a[i][0] = fmaf(-r.ai, wi, a[i][0]);
      a[i][1] = fmaf(r.ar, wi, a[i][1]);
      a[i][1] = fmaf(r.ai, wr, a[i][1]);
      a[i][2] = fmaf(r.br, wr, a[i][2]);
      a[i][2] = fmaf(-r.bi, wi, a[i][2]);
      a[i][3] = fmaf(r.br, wi, a[i][3]);
      a[i][3] = fmaf(r.bi, wr, a[i][3]);
    }
  }
}

static constexpr uint32_t kLdsRays = 512;  // rays staged per LDS tile

template <int P, int W, int SRC, bool CHECK, bool QFOLD>
__device__ __forceinline__ void f32r_item(const KInst& I, const WorkItem& w,
                                          RayRec32* lds, uint32_t lane,
                                          uint32_t wave) {
  const uint32_t np = I.np;
  const uint32_t base = w.pix0 + wave * 64u * P;
  const double R0 = ((const XRTK_CONST InstScratch*)I.scr)->R0;
  const float h = (float)(1.0 / (2.0 * R0 * I.kr_hi));
  const float skip2 = (float)(I.skip * I.skip);
  PixRegs<P> px;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    uint32_t p = base + i * 64u + lane;
    p = p < np ? p : np - 1u;
    const PixRec32 pr = I.pix32[p];
    px.qx[i] = pr.qx; px.qy[i] = pr.qy; px.qz[i] = pr.qz; px.R[i] = pr.R;
    px.mx[i] = pr.mx; px.my[i] = pr.my; px.mz[i] = pr.mz;
    px.g[i] = QFOLD ? (float)((R0 - I.R64[p]) / R0) : 0.f;
  }
  double acc[P][4];
#pragma unroll
  for (int i = 0; i < P; ++i) acc[i][0] = acc[i][1] = acc[i][2] = acc[i][3] = 0.0;

  const uint32_t r0 = w.chunk * I.chunk_len;
  const uint32_t r1 = min(r0 + I.chunk_len, I.nr);
  const uint32_t tile = I.tile;
  for (uint32_t t0 = r0; t0 < r1; t0 += tile) {
    const uint32_t t1 = min(t0 + tile, r1);
    float a[P][4];
#pragma unroll
    for (int i = 0; i < P; ++i) a[i][0] = a[i][1] = a[i][2] = a[i][3] = 0.f;
    if constexpr (SRC == 1) {
      SmemRays src{(const XRTK_CONST RayRec32*)I.rays32};
      pair_span<P, CHECK, QFOLD>(src, t0, t1, px, h, skip2, a);
    } else {
      for (uint32_t s0 = t0; s0 < t1; s0 += kLdsRays) {
        const uint32_t ns = min(kLdsRays, t1 - s0);
        __syncthreads();
        const float4* g4 = reinterpret_cast<const float4*>(I.rays32 + s0);
        float4* l4 = reinterpret_cast<float4*>(lds);
        for (uint32_t k = threadIdx.x; k < ns * 3u; k += 64u * W) l4[k] = g4[k];
        __syncthreads();
        LdsRays src{lds};
        pair_span<P, CHECK, QFOLD>(src, 0u, ns, px, h, skip2, a);
      }
    }
#pragma unroll
    for (int i = 0; i < P; ++i) {
      acc[i][0] += (double)a[i][0]; acc[i][1] += (double)a[i][1];
      acc[i][2] += (double)a[i][2]; acc[i][3] += (double)a[i][3];
    }
  }
  double* part = I.partial + (size_t)w.chunk * np * 4u;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const uint32_t p = base + i * 64u + lane;
    if (p < np) {
      double2* d = reinterpret_cast<double2*>(part + (size_t)p * 4u);
      d[0] = make_double2(acc[i][0], acc[i][1]);
      d[1] = make_double2(acc[i][2], acc[i][3]);
    }
  }
}

template <int P, int W, int SRC, bool PERSIST>
__global__ __launch_bounds__(64 * W) void k_pair_f32r(
    const KInst* __restrict__ insts, const WorkItem* __restrict__ items,
    uint32_t n_items, uint32_t* __restrict__ queue, uint32_t remap) {
  __shared__ RayRec32 lds[SRC == 2 ? kLdsRays : 1];
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
    const uint32_t flags = ((const XRTK_CONST InstScratch*)I.scr)->flags;
    const bool chk = flags & INST_F_CHECK, qf = flags & INST_F_QFOLD;
    if (chk) {
      if (qf) f32r_item<P, W, SRC, true, true>(I, w, lds, lane, wave);
      else    f32r_item<P, W, SRC, true, false>(I, w, lds, lane, wave);
    } else {
      if (qf) f32r_item<P, W, SRC, false, true>(I, w, lds, lane, wave);
      else    f32r_item<P, W, SRC, false, false>(I, w, lds, lane, wave);
    }
    if constexpr (!PERSIST) break;
  }
}

// ---------------------------------------------------------------------------
// plain FP64 pair kernel (the device oracle and the FP64 speed baseline)
// ---------------------------------------------------------------------------
template <int P, bool CHECK>
__device__ __forceinline__ void f64_item(const KInst& I, const WorkItem& w,
                                         uint32_t lane, uint32_t wave) {
#pragma clang fp contract(off)
  const uint32_t np = I.np;
  const uint32_t base = w.pix0 + wave * 64u * P;
  double px[P], py[P], pz[P];
#pragma unroll
  for (int i = 0; i < P; ++i) {
    uint32_t p = base + i * 64u + lane;
    p = p < np ? p : np - 1u;
    const PixF64 q = I.pix64[p];
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
  double* part = I.partial + (size_t)w.chunk * np * 4u;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    const uint32_t p = base + i * 64u + lane;
    if (p < np) {
      double2* d = reinterpret_cast<double2*>(part + (size_t)p * 4u);
      d[0] = make_double2(sum[i][0], sum[i][1]);
      d[1] = make_double2(sum[i][2], sum[i][3]);
    }
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
    if (((const XRTK_CONST InstScratch*)I.scr)->flags & INST_F_CHECK) f64_item<P, true>(I, w, lane, wave);
    else                             f64_item<P, false>(I, w, lane, wave);
    if constexpr (!PERSIST) break;
  }
}

// ---------------------------------------------------------------------------
// merge: compensated FP64 sum of the chunk partials in chunk order, then the
// base phase (FP32R): out = k/(4 pi) exp(i k |p - c|) * sum.
// grid = (ceil(max np / 256), n_list); list[y] is an instance index.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_merge(const KInst* __restrict__ insts,
                                               const uint32_t* __restrict__ list) {
  const KInst& I = insts[list[blockIdx.y]];
  const uint32_t p = blockIdx.x * 256u + threadIdx.x;
  if (p >= I.np) return;
  double s[4] = {0, 0, 0, 0}, e[4] = {0, 0, 0, 0};
  const size_t stride = (size_t)I.np * 4u;
  const double* src = I.partial + (size_t)p * 4u;
  for (uint32_t c = 0; c < I.n_chunks; ++c) {
    const double2 v01 = reinterpret_cast<const double2*>(src + c * stride)[0];
    const double2 v23 = reinterpret_cast<const double2*>(src + c * stride)[1];
    const double v[4] = {v01.x, v01.y, v23.x, v23.y};
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
    cis_rev_dd(I.kr_hi, I.kr_lo, I.R64[p], &c, &sn);
    const double kf = I.k / (4.0 * kPI);
    c *= kf; sn *= kf;
    const double a0 = o[0] * c - o[1] * sn, a1 = o[0] * sn + o[1] * c;
    const double a2 = o[2] * c - o[3] * sn, a3 = o[2] * sn + o[3] * c;
    o[0] = a0; o[1] = a1; o[2] = a2; o[3] = a3;
  }
  double2* d = reinterpret_cast<double2*>(I.out + (size_t)p * 4u);
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
hipError_t launch_pix_from_f32aos(const void* src, PixF64* dst, uint32_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_pix_from_f32aos, dim3(cdiv(n, 256)), dim3(256), 0, st,
                     (const float4*)src, dst, n);
  return hipGetLastError();
}
hipError_t launch_pix_from_cols(const double* cols, PixF64* dst, uint32_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_pix_from_cols, dim3(cdiv(n, 256)), dim3(256), 0, st, cols, dst, n);
  return hipGetLastError();
}
hipError_t launch_pix_grid(const double c[3], double dx, double dz, uint32_t nx,
                           uint32_t nz, PixF64* dst, hipStream_t st) {
  const size_t n = (size_t)nx * nz;
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_pix_grid, dim3(cdiv(n, 256)), dim3(256), 0, st, c[0], c[1], c[2],
                     dx, dz, nx, nz, dst);
  return hipGetLastError();
}

// Centroid, bounds, plan and the per-solve FP32 records of one instance.
// part: kRedBlocks * 4 doubles of scratch.
hipError_t launch_prep(const RayRec64* rays, uint32_t nr, const PixF64* pix, uint32_t np,
                       InstScratch* scr, double* part, uint32_t mode, double kr,
                       double kr_lo, double skip, double qfold_rev, uint32_t force_flags,
                       uint32_t static_flags, RayRec32* rays32, PixRec32* pix32,
                       double* R64, hipStream_t st) {
  const unsigned nb = nr ? std::min<unsigned>(kRedBlocks, cdiv(nr, 256)) : 1u;
  hipLaunchKernelGGL(k_centroid_partial, dim3(nb), dim3(256), 0, st, rays, nr, part);
  hipLaunchKernelGGL(k_centroid_final, dim3(1), dim3(256), 0, st, part, nb, nr, scr);
  if (nr)
    hipLaunchKernelGGL(k_ray_bounds, dim3(nb), dim3(256), 0, st, rays, nr, scr);
  if (np)
    hipLaunchKernelGGL(k_pix_prep, dim3(cdiv(np, 256)), dim3(256), 0, st, pix, np, kr, scr,
                       mode == MODE_F32R ? pix32 : (PixRec32*)nullptr, R64);
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(1), 0, st, scr, skip, kr, mode, qfold_rev,
                     force_flags, static_flags);
  if (mode == MODE_F32R && nr)
    hipLaunchKernelGGL(k_ray_prep32, dim3(cdiv(nr, 256)), dim3(256), 0, st, rays, nr, scr, kr,
                       kr_lo, rays32);
  return hipGetLastError();
}

template <int P, int W, int SRC, bool PERSIST>
static void launch_f32r_t(const KInst* insts, const WorkItem* items, uint32_t n_items,
                          uint32_t* queue, uint32_t grid, uint32_t remap, hipStream_t st) {
  hipLaunchKernelGGL((k_pair_f32r<P, W, SRC, PERSIST>), dim3(grid), dim3(64 * W), 0, st, insts,
                     items, n_items, queue, remap);
}
template <int P, int W, bool PERSIST>
static void launch_f64_t(const KInst* insts, const WorkItem* items, uint32_t n_items,
                         uint32_t* queue, uint32_t grid, uint32_t remap, hipStream_t st) {
  hipLaunchKernelGGL((k_pair_f64<P, W, PERSIST>), dim3(grid), dim3(64 * W), 0, st, insts,
                     items, n_items, queue, remap);
}

// Returns hipErrorInvalidValue for a configuration that is not instantiated.
hipError_t launch_pair(uint32_t mode, const LaunchCfg& cfg, const KInst* insts,
                       const WorkItem* items, uint32_t n_items, uint32_t* queue,
                       uint32_t grid, uint32_t remap, hipStream_t st) {
  if (!n_items) return hipSuccess;
  if (cfg.persistent) hipLaunchKernelGGL(k_reset_u32, dim3(1), dim3(64), 0, st, queue, 1u);
  const bool ps = cfg.persistent != 0;
#define XRTK_F32(P_, W_, S_)                                                              \
  if (cfg.P == P_ && cfg.W == W_ && cfg.ray_src == S_) {                                  \
    if (ps) launch_f32r_t<P_, W_, S_, true>(insts, items, n_items, queue, grid, remap, st); \
    else launch_f32r_t<P_, W_, S_, false>(insts, items, n_items, queue, grid, remap, st);   \
    return hipGetLastError();                                                             \
  }
#define XRTK_F64(P_, W_)                                                               \
  if (cfg.P == P_ && cfg.W == W_) {                                                    \
    if (ps) launch_f64_t<P_, W_, true>(insts, items, n_items, queue, grid, remap, st); \
    else launch_f64_t<P_, W_, false>(insts, items, n_items, queue, grid, remap, st);   \
    return hipGetLastError();                                                          \
  }
  if (mode == MODE_F32R) {
    XRTK_F32(2, 1, 1) XRTK_F32(2, 4, 1) XRTK_F32(4, 1, 1) XRTK_F32(4, 4, 1)
    XRTK_F32(8, 1, 1) XRTK_F32(8, 4, 1)
    XRTK_F32(2, 4, 2) XRTK_F32(4, 4, 2) XRTK_F32(8, 4, 2)
  } else {
    XRTK_F64(1, 1) XRTK_F64(1, 4) XRTK_F64(2, 1) XRTK_F64(2, 4)
  }
#undef XRTK_F32
#undef XRTK_F64
  return hipErrorInvalidValue;
}

bool pair_cfg_supported(uint32_t mode, const LaunchCfg& c) {
  if (mode == MODE_F32R) {
    if (c.ray_src == 1) return (c.P == 2 || c.P == 4 || c.P == 8) && (c.W == 1 || c.W == 4);
    if (c.ray_src == 2) return (c.P == 2 || c.P == 4 || c.P == 8) && c.W == 4;
    return false;
  }
  return (c.P == 1 || c.P == 2) && (c.W == 1 || c.W == 4);
}

hipError_t launch_merge(const KInst* insts, const uint32_t* list, uint32_t n_list,
                        uint32_t max_np, hipStream_t st) {
  if (!n_list || !max_np) return hipSuccess;
  hipLaunchKernelGGL(k_merge, dim3(cdiv(max_np, 256), n_list), dim3(256), 0, st, insts, list);
  return hipGetLastError();
}

hipError_t launch_zero_f64(double* p, size_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_zero_f64, dim3(cdiv(n, 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

}  // namespace xrtk
