// Fresnel propagation (src/waves/fresnel.rs:53-106) on gfx950.
//
//   U(x'_ix, z'_iz) = sum_j Es_j exp(i K ((x'_ix - x_j)^2 + (z'_iz - z_j)^2)),
//   K = k / (2 dist), k = E / CHBAR * 1e7.
//
// The kernel factorises exactly, exp(i K (dx^2 + dz^2)) = exp(i K dx^2)
// exp(i K dz^2), and the screen is a grid, so
//
//   U[iz][ix] = sum_j Z_j[iz] X_j[ix],  Z_j[iz] = Es_j exp(i K (z'_iz - z_j)^2),
//                                       X_j[ix] = exp(i K (x'_ix - x_j)^2):
//
// a complex matrix product over the rays, which runs on the matrix cores.
// The factors are never stored in HBM: each workgroup owns a tile of the
// screen and one slice of the rays, and generates the factors of its tile's
// rows and columns block by block into LDS (each phase K d^2 evaluated in
// FP64 and reduced to a fractional revolution first), then multiplies them
// there:
//   FP32 (XRTK_MODE_F32R): 128 x 128 pixels per workgroup of 4 waves
//        (64 x 64 each), sin/cos of the reduced phase by v_sin/v_cos; each
//        workgroup's FP32 sums cover at most 8,192 rays (a ray slice; 16,384
//        with FP32 operands) and go to HBM as partial sums.  The products run on
//        v_mfma_f32_32x32x16_f16 with each FP32 factor split into FP16
//        hi = RNE(x) and lo = RNE(x - hi) (x - hi - lo ~ 2^-22 |x|), three
//        products hi.hi + hi.lo + lo.hi per real product, FP32 accumulation
//        (~2^-21 per term); the amplitudes are scaled by a power of two so
//        the largest |Re| + |Im| lies in [1/2, 1), and the merge scales back
//        exactly.  XrtkOpts.engine = 1 takes the FP32 operands on
//        v_mfma_f32_32x32x2f32 instead (the comparison kernel);
//   FP64 (XRTK_MODE_F64): v_mfma_f64_16x16x4f64, 64 x 64 pixels per
//        workgroup (32 x 32 per wave); e^{2 pi i f} of the reduced phase f
//        from a 256-entry table and a degree-7 Taylor series (cis_rev_tab).
// The complex product takes four real products per pair of factors:
//   Re U += Zr Xr + Zi (-Xi),  Im U += Zr Xi + Zi Xr   (8 flops per pair;
//   24 MFMA flops with the FP16 parts).
// k_fresnel_merge* sums the ray slices' partials in FP64, in slice order.
// Output index iz * nx + ix (fresnel.rs:78-99).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "runtime.h"

#if defined(__HIP_DEVICE_COMPILE__)
#define XRTK_FCONST __attribute__((address_space(4)))
#else
#define XRTK_FCONST
#endif

namespace xrtk {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

constexpr uint32_t kT32 = 128;     // FP32: tile side in pixels
constexpr uint32_t kKB32 = 16;     // FP32: rays per LDS block
constexpr uint32_t kSlice32 = 16384;  // FP32: most rays in one FP32 sum
constexpr uint32_t kT16 = 128;     // FP16 parts: tile side in pixels
constexpr uint32_t kKB16 = 16;     // FP16 parts: rays per LDS block (one K of the MFMA)
constexpr uint32_t kSlice16 = 8192;   // FP16 parts: most rays in one FP32 sum
static_assert(kT16 == kT32 && kKB16 == kKB32, "both FP32-sum kernels share one launch shape");
constexpr uint32_t kT64 = 64;      // FP64: tile side
constexpr uint32_t kKB64 = 8;      // FP64: rays per LDS block
constexpr uint32_t kTab = 256;     // FP64: entries of the e^{2 pi i j / 256} table

// Factor planes of one block of rays: Z (rows) and X (columns), real and
// imaginary parts, [ray][pixel], rows LD elements apart.
template <class T, uint32_t KB, uint32_t N, uint32_t LD = N>
struct Planes {
  T zr[KB][LD], zi[KB][LD], xr[KB][LD], xi[KB][LD];
};


// ---- FP32 -------------------------------------------------------------------------------
// Generates the factors of rays kb .. kb + KB - 1 for column / row g (thread
// t: g = t % 128, rays (t / 128) * 8 .. + 7 of the block): the phase
// K d^2 / (2 pi) in FP64, its fractional revolution, v_sin / v_cos.
__device__ __forceinline__ void gen32(Planes<float, kKB32, kT32>& P, uint32_t kb, uint32_t g,
                                      uint32_t h, double xg, double zg, double krev,
                                      const XRTK_FCONST double* crx,
                                      const XRTK_FCONST double* crz,
                                      const XRTK_FCONST double2* ces) {
#pragma unroll
  for (uint32_t i = 0; i < kKB32 / 2; ++i) {
    const uint32_t kl = h * (kKB32 / 2) + i, k = kb + kl;
    const double xk = crx[k], zk = crz[k];
    const double2 e = ces[k];
    float fx, fz;
    {
#pragma clang fp contract(off)
      // revolutions, reduced to [-1/2, 1/2] (FP32 resolves the phase twice
      // as finely there as on [0, 1))
      const double dx = xg - xk, dz = zg - zk;
      const double px = krev * (dx * dx), pz = krev * (dz * dz);
      fx = (float)(px - rint(px));
      fz = (float)(pz - rint(pz));
    }
    P.xr[kl][g] = __builtin_amdgcn_cosf(fx);
    P.xi[kl][g] = __builtin_amdgcn_sinf(fx);
    const float c = __builtin_amdgcn_cosf(fz), sn = __builtin_amdgcn_sinf(fz);
    const float er = (float)e.x, ei = (float)e.y;
    P.zr[kl][g] = er * c - ei * sn;
    P.zi[kl][g] = er * sn + ei * c;
  }
}

// The products of one block: each wave's 2 x 2 blocks of 32 x 32, four real
// products per pair of complex factors (v_mfma_f32_32x32x2f32: lane l
// supplies A[l % 32][l / 32] and B[l / 32][l % 32]).
// A wave's 2 x NB blocks of 32 x 32 (z rows zb.., x columns xb..).
template <int NB>
__device__ __forceinline__ void mma32(const Planes<float, kKB32, kT32>& P, uint32_t lane,
                                      uint32_t zb, uint32_t xb, floatx16 (&ur)[2][NB],
                                      floatx16 (&ui)[2][NB]) {
#pragma unroll
  for (uint32_t kk = 0; kk < kKB32 / 2; ++kk) {
    const uint32_t kr = 2u * kk + (lane >> 5), c = lane & 31u;
    float ar[2], ai[2], br[NB], bi[NB], nbi[NB];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      ar[a] = P.zr[kr][zb + 32u * a + c];
      ai[a] = P.zi[kr][zb + 32u * a + c];
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      br[b] = P.xr[kr][xb + 32u * b + c];
      bi[b] = P.xi[kr][xb + 32u * b + c];
      nbi[b] = -bi[b];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        ur[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[a], br[b], ur[a][b], 0, 0, 0);
        ui[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ar[a], bi[b], ui[a][b], 0, 0, 0);
      }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        ur[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ai[a], nbi[b], ur[a][b], 0, 0, 0);
        ui[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(ai[a], br[b], ui[a][b], 0, 0, 0);
      }
  }
}

// FP32 kernel.  grid = (tiles, ray slices); block = 256 (4 waves), two
// workgroups per CU.  Wave w owns z rows (w & 1) * 64 .. + 63 and x columns
// (w >> 1) * 64 .. + 63 of the tile (D element r of lane l is row
// 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32 of a 32 x 32 block).  Two
// LDS buffers: the waves generate the next block of rays' factors into one
// while they multiply the block in the other (one barrier per block).  A
// slice is a whole number of block pairs; rays are padded with zero
// amplitudes to whole slices plus one block (the block generated after a
// slice's last is never used).
__global__ __launch_bounds__(256, 2) void k_fresnel_mfma32(
    const double* __restrict__ xp, uint32_t nx, const double* __restrict__ zp, uint32_t nz,
    const double* __restrict__ rx, const double* __restrict__ rz, const double2* __restrict__ es,
    uint32_t slice, double krev, uint32_t n_tx, float2* __restrict__ partial) {
  __shared__ Planes<float, kKB32, kT32> P0, P1;
  const uint32_t t = threadIdx.x, lane = t & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const uint32_t tx = blockIdx.x % n_tx, tz = blockIdx.x / n_tx;
  const uint32_t x0 = tx * kT32, z0 = tz * kT32;
  // factor generation: column / row g of the tile, half h of each ray block
  const uint32_t g = t & 127u, h = __builtin_amdgcn_readfirstlane(t >> 7);
  const double xg = xp[min(x0 + g, nx - 1u)];
  const double zg = zp[min(z0 + g, nz - 1u)];
  const uint32_t zb = (wave & 1u) * 64u, xb = (wave >> 1) * 64u;
  floatx16 ur[2][2], ui[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) ur[a][b][r] = ui[a][b][r] = 0.f;
  const XRTK_FCONST double* crx = (const XRTK_FCONST double*)rx;
  const XRTK_FCONST double* crz = (const XRTK_FCONST double*)rz;
  const XRTK_FCONST double2* ces = (const XRTK_FCONST double2*)es;
  const uint32_t k0 = blockIdx.y * slice;
  gen32(P0, k0, g, h, xg, zg, krev, crx, crz, ces);
  __syncthreads();
  for (uint32_t kb = k0; kb < k0 + slice; kb += 2u * kKB32) {
    gen32(P1, kb + kKB32, g, h, xg, zg, krev, crx, crz, ces);
    mma32<2>(P0, lane, zb, xb, ur, ui);
    __syncthreads();
    gen32(P0, kb + 2u * kKB32, g, h, xg, zg, krev, crx, crz, ces);
    mma32<2>(P1, lane, zb, xb, ur, ui);
    __syncthreads();
  }
  float2* Pp = partial + (size_t)blockIdx.y * nz * nx;
  const uint32_t xl = x0 + xb + (lane & 31u);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const uint32_t x = xl + 32u * b;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const uint32_t z = z0 + zb + 32u * a + 8u * (r / 4) + 4u * (lane >> 5) + (r % 4);
        if (z < nz && x < nx) Pp[(size_t)z * nx + x] = make_float2(ur[a][b][r], ui[a][b][r]);
      }
    }
}

// ---- FP32 factors in FP16 parts ------------------------------------------------------
// hi = RNE(x), lo = RNE(x - hi) of two values, packed in one register each;
// m1 is -1 held in a register, so fma(hi, m1, x) rounded to FP16 is one
// v_fma_mix instruction.
__device__ __forceinline__ void split_pair(float a, float b, float m1, uint32_t& hi,
                                           uint32_t& lo) {
  const floatx2 p = {a, b};
  const halfx2 h = __builtin_convertvector(p, halfx2);
  const halfx2 l = {(_Float16)__builtin_fmaf((float)h.x, m1, a),
                    (_Float16)__builtin_fmaf((float)h.y, m1, b)};
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, l);
}

// Factors of one block of 16 rays in the operand layout of
// v_mfma_f32_32x32x16_f16 (lane l takes row / column l % 32 and rays
// 8 (l / 32) .. + 7): part (re hi, re lo, im hi, im lo), pixel, half (8 rays,
// 16 bytes).  The two halves of a pixel are stored swapped when bits 2 and 3
// of the pixel index differ, so that 8 or 16 consecutive lanes of one
// ds_read_b128 or ds_write_b128 (pixels 32 bytes apart) cover distinct banks.
struct Planes16 {
  uint4 z[4][kT16][2];
  uint4 x[4][kT16][2];
};

__device__ __forceinline__ uint32_t half_slot(uint32_t pix, uint32_t h) {
  return h ^ (((pix >> 2) ^ (pix >> 3)) & 1u);
}

// Thread t: column / row g = t % 128, rays 8 h .. 8 h + 7 of the block
// (h = t / 128): the phases as gen32, the factors split into FP16 parts.
__device__ __forceinline__ void gen16(Planes16& P, uint32_t kb, uint32_t g, uint32_t h,
                                      double xg, double zg, double krev, float m1,
                                      const XRTK_FCONST double* crx,
                                      const XRTK_FCONST double* crz,
                                      const XRTK_FCONST double2* ces) {
  float xr[8], xi[8], zr[8], zi[8];
#pragma unroll
  for (uint32_t i = 0; i < 8; ++i) {
    const uint32_t k = kb + 8u * h + i;
    const double xk = crx[k], zk = crz[k];
    const double2 e = ces[k];
    float fx, fz;
    {
#pragma clang fp contract(off)
      const double dx = xg - xk, dz = zg - zk;
      const double px = krev * (dx * dx), pz = krev * (dz * dz);
      fx = (float)(px - rint(px));
      fz = (float)(pz - rint(pz));
    }
    xr[i] = __builtin_amdgcn_cosf(fx);
    xi[i] = __builtin_amdgcn_sinf(fx);
    const float c = __builtin_amdgcn_cosf(fz), sn = __builtin_amdgcn_sinf(fz);
    const float er = (float)e.x, ei = (float)e.y;
    zr[i] = er * c - ei * sn;
    zi[i] = er * sn + ei * c;
  }
  uint32_t q[8][4];  // [xr hi, xr lo, xi hi, xi lo, zr hi, zr lo, zi hi, zi lo][pair]
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    split_pair(xr[2 * j], xr[2 * j + 1], m1, q[0][j], q[1][j]);
    split_pair(xi[2 * j], xi[2 * j + 1], m1, q[2][j], q[3][j]);
    split_pair(zr[2 * j], zr[2 * j + 1], m1, q[4][j], q[5][j]);
    split_pair(zi[2 * j], zi[2 * j + 1], m1, q[6][j], q[7][j]);
  }
  const uint32_t hs = half_slot(g, h);
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    P.x[p][g][hs] = make_uint4(q[p][0], q[p][1], q[p][2], q[p][3]);
    P.z[p][g][hs] = make_uint4(q[4 + p][0], q[4 + p][1], q[4 + p][2], q[4 + p][3]);
  }
}

__device__ __forceinline__ halfx8 as_h8(uint4 v) { return __builtin_bit_cast(halfx8, v); }

__device__ __forceinline__ floatx16 mfma3(halfx8 ah, halfx8 al, halfx8 bh, halfx8 bl,
                                          floatx16 c) {
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh, c, 0, 0, 0);
  return c;
}

// The products of one block of 16 rays: each wave's 2 x 2 blocks of 32 x 32,
// four real products per pair of complex factors, three MFMAs each.
__device__ __forceinline__ void mma16(const Planes16& P, uint32_t lane, uint32_t zb, uint32_t xb,
                                      floatx16 (&ur)[2][2], floatx16 (&ui)[2][2]) {
  const uint32_t c = lane & 31u, kh = lane >> 5;
  uint4 a[2][4], b[2][4], nb[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const uint32_t r = zb + 32u * t + c, x = xb + 32u * t + c;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      a[t][p] = P.z[p][r][half_slot(r, kh)];
      b[t][p] = P.x[p][x][half_slot(x, kh)];
    }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const uint4 v = b[t][2 + p];
      nb[t][p] = make_uint4(v.x ^ 0x80008000u, v.y ^ 0x80008000u, v.z ^ 0x80008000u,
                            v.w ^ 0x80008000u);
    }
  }
#pragma unroll
  for (int tr = 0; tr < 2; ++tr)
#pragma unroll
    for (int tc = 0; tc < 2; ++tc) {
      const halfx8 zrh = as_h8(a[tr][0]), zrl = as_h8(a[tr][1]);
      const halfx8 zih = as_h8(a[tr][2]), zil = as_h8(a[tr][3]);
      const halfx8 xrh = as_h8(b[tc][0]), xrl = as_h8(b[tc][1]);
      const halfx8 xih = as_h8(b[tc][2]), xil = as_h8(b[tc][3]);
      ur[tr][tc] = mfma3(zrh, zrl, xrh, xrl, ur[tr][tc]);
      ur[tr][tc] = mfma3(zih, zil, as_h8(nb[tc][0]), as_h8(nb[tc][1]), ur[tr][tc]);
      ui[tr][tc] = mfma3(zrh, zrl, xih, xil, ui[tr][tc]);
      ui[tr][tc] = mfma3(zih, zil, xrh, xrl, ui[tr][tc]);
    }
}

// FP16-part kernel: the shape, LDS double buffering, slices and padding of
// k_fresnel_mfma32, 16 rays per block.
__global__ __launch_bounds__(256, 2) void k_fresnel_mfma16(
    const double* __restrict__ xp, uint32_t nx, const double* __restrict__ zp, uint32_t nz,
    const double* __restrict__ rx, const double* __restrict__ rz, const double2* __restrict__ es,
    uint32_t slice, double krev, uint32_t n_tx, float2* __restrict__ partial, float m1) {
  __shared__ Planes16 P0, P1;
  const uint32_t t = threadIdx.x, lane = t & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const uint32_t tx = blockIdx.x % n_tx, tz = blockIdx.x / n_tx;
  const uint32_t x0 = tx * kT16, z0 = tz * kT16;
  const uint32_t g = t & 127u, h = __builtin_amdgcn_readfirstlane(t >> 7);
  const double xg = xp[min(x0 + g, nx - 1u)];
  const double zg = zp[min(z0 + g, nz - 1u)];
  const uint32_t zb = (wave & 1u) * 64u, xb = (wave >> 1) * 64u;
  floatx16 ur[2][2], ui[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 16; ++r) ur[a][b][r] = ui[a][b][r] = 0.f;
  const XRTK_FCONST double* crx = (const XRTK_FCONST double*)rx;
  const XRTK_FCONST double* crz = (const XRTK_FCONST double*)rz;
  const XRTK_FCONST double2* ces = (const XRTK_FCONST double2*)es;
  const uint32_t k0 = blockIdx.y * slice;
  gen16(P0, k0, g, h, xg, zg, krev, m1, crx, crz, ces);
  __syncthreads();
  for (uint32_t kb = k0; kb < k0 + slice; kb += 2u * kKB16) {
    gen16(P1, kb + kKB16, g, h, xg, zg, krev, m1, crx, crz, ces);
    mma16(P0, lane, zb, xb, ur, ui);
    __syncthreads();
    gen16(P0, kb + 2u * kKB16, g, h, xg, zg, krev, m1, crx, crz, ces);
    mma16(P1, lane, zb, xb, ur, ui);
    __syncthreads();
  }
  float2* Pp = partial + (size_t)blockIdx.y * nz * nx;
  const uint32_t xl = x0 + xb + (lane & 31u);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const uint32_t x = xl + 32u * b;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const uint32_t z = z0 + zb + 32u * a + 8u * (r / 4) + 4u * (lane >> 5) + (r % 4);
        if (z < nz && x < nx) Pp[(size_t)z * nx + x] = make_float2(ur[a][b][r], ui[a][b][r]);
      }
    }
}

// ---- FP64 -------------------------------------------------------------------------------
// e^{2 pi i f} for f in [0, 1), to FP64 rounding: the nearest table entry
// e^{2 pi i j / 256} (tab, in LDS, from sincospi) times e^{i x},
// x = 2 pi (f - j / 256), |x| <= pi / 256, by its Taylor series through x^6
// (cos) and x^7 (sin): the first omitted terms are below 6e-21.
__device__ __forceinline__ void cis_rev_tab(double f, const double2* tab, double* c, double* s) {
  const double jd = rint(f * (double)kTab);
  const double x = fma(-jd, 1.0 / kTab, f) * (2.0 * 3.14159265358979323846);
  const double x2 = x * x;
  const double cx = fma(x2, fma(x2, fma(x2, -1.0 / 720.0, 1.0 / 24.0), -0.5), 1.0);
  const double sx = x * fma(x2, fma(x2, fma(x2, -1.0 / 5040.0, 1.0 / 120.0), -1.0 / 6.0), 1.0);
  const double2 t = tab[(uint32_t)jd & (kTab - 1u)];
  *c = fma(t.x, cx, -(t.y * sx));
  *s = fma(t.y, cx, t.x * sx);
}

// thread t: column / row g = t % 64, rays (t / 64) * 2 .. + 1 of the block
__device__ __forceinline__ void gen64(Planes<double, kKB64, kT64>& P, const double2* tab,
                                      uint32_t kb, uint32_t g, uint32_t q, double xg, double zg,
                                      double krev, const XRTK_FCONST double* crx,
                                      const XRTK_FCONST double* crz,
                                      const XRTK_FCONST double2* ces) {
#pragma unroll
  for (uint32_t i = 0; i < kKB64 / 4; ++i) {
    const uint32_t kl = q * (kKB64 / 4) + i, k = kb + kl;
    const double xk = crx[k], zk = crz[k];
    const double2 e = ces[k];
    double fx, fz;
    {
#pragma clang fp contract(off)
      const double dx = xg - xk, dz = zg - zk;
      fx = __builtin_amdgcn_fract(krev * (dx * dx));  // revolutions, [0, 1)
      fz = __builtin_amdgcn_fract(krev * (dz * dz));
    }
    double cx, sx, cz, sz;
    cis_rev_tab(fx, tab, &cx, &sx);
    cis_rev_tab(fz, tab, &cz, &sz);
    P.xr[kl][g] = cx;
    P.xi[kl][g] = sx;
    P.zr[kl][g] = e.x * cz - e.y * sz;
    P.zi[kl][g] = e.x * sz + e.y * cz;
  }
}

// v_mfma_f64_16x16x4f64: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; D element r of lane l is row 4 r + l / 16, column l % 16.
__device__ __forceinline__ void mma64(const Planes<double, kKB64, kT64>& P, uint32_t lane,
                                      uint32_t zb, uint32_t xb, doublex4 (&ur)[2][2],
                                      doublex4 (&ui)[2][2]) {
#pragma unroll
  for (uint32_t kk = 0; kk < kKB64 / 4; ++kk) {
    const uint32_t kr = 4u * kk + (lane >> 4), c = lane & 15u;
    double ar[2], ai[2], br[2], bi[2], nbi[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      ar[b] = P.zr[kr][zb + 16u * b + c];
      ai[b] = P.zi[kr][zb + 16u * b + c];
      br[b] = P.xr[kr][xb + 16u * b + c];
      bi[b] = P.xi[kr][xb + 16u * b + c];
      nbi[b] = -bi[b];
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        ur[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], ur[a][b], 0, 0, 0);
        ui[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ui[a][b], 0, 0, 0);
      }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        ur[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], nbi[b], ur[a][b], 0, 0, 0);
        ui[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ui[a][b], 0, 0, 0);
      }
  }
}

// FP64 kernel.  grid = (tiles, ray slices); block = 256.  Wave w owns z rows
// (w & 1) * 32 .. + 31 and x columns (w >> 1) * 32 .. + 31 of the 64 x 64
// tile, as 2 x 2 blocks of 16 x 16; double-buffered like the FP32 kernel.
__global__ __launch_bounds__(256, 2) void k_fresnel_mfma64(
    const double* __restrict__ xp, uint32_t nx, const double* __restrict__ zp, uint32_t nz,
    const double* __restrict__ rx, const double* __restrict__ rz, const double2* __restrict__ es,
    uint32_t slice, double krev, uint32_t n_tx, double2* __restrict__ partial) {
  __shared__ Planes<double, kKB64, kT64> P0, P1;
  __shared__ double2 tab[kTab];
  const uint32_t t = threadIdx.x, lane = t & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(t >> 6);
  {
    double sj, cj;
    sincospi((double)t / (kTab / 2), &sj, &cj);  // e^{2 pi i t / 256}
    tab[t] = make_double2(cj, sj);
  }
  const uint32_t tx = blockIdx.x % n_tx, tz = blockIdx.x / n_tx;
  const uint32_t x0 = tx * kT64, z0 = tz * kT64;
  const uint32_t g = t & 63u, q = wave;
  const double xg = xp[min(x0 + g, nx - 1u)];
  const double zg = zp[min(z0 + g, nz - 1u)];
  const uint32_t zb = (wave & 1u) * 32u, xb = (wave >> 1) * 32u;
  doublex4 ur[2][2], ui[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) ur[a][b][r] = ui[a][b][r] = 0.0;
  const XRTK_FCONST double* crx = (const XRTK_FCONST double*)rx;
  const XRTK_FCONST double* crz = (const XRTK_FCONST double*)rz;
  const XRTK_FCONST double2* ces = (const XRTK_FCONST double2*)es;
  const uint32_t k0 = blockIdx.y * slice;
  __syncthreads();  // the table
  gen64(P0, tab, k0, g, q, xg, zg, krev, crx, crz, ces);
  __syncthreads();
  for (uint32_t kb = k0; kb < k0 + slice; kb += 2u * kKB64) {
    gen64(P1, tab, kb + kKB64, g, q, xg, zg, krev, crx, crz, ces);
    mma64(P0, lane, zb, xb, ur, ui);
    __syncthreads();
    gen64(P0, tab, kb + 2u * kKB64, g, q, xg, zg, krev, crx, crz, ces);
    mma64(P1, lane, zb, xb, ur, ui);
    __syncthreads();
  }
  double2* Pp = partial + (size_t)blockIdx.y * nz * nx;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const uint32_t x = x0 + xb + 16u * b + (lane & 15u);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t z = z0 + zb + 16u * a + 4u * r + (lane >> 4);
        if (z < nz && x < nx) Pp[(size_t)z * nx + x] = make_double2(ur[a][b][r], ui[a][b][r]);
      }
    }
}

// U[p] = sum over the ray slices of partial[s][p], in FP64 and slice order,
// times `back` (a power of two: the amplitudes' scale undone).
template <class T>
__global__ __launch_bounds__(256) void k_fresnel_merge(const T* __restrict__ partial,
                                                       uint32_t n_slices, size_t np,
                                                       double back, double2* __restrict__ out) {
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < np;
       p += (size_t)gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (uint32_t s = 0; s < n_slices; ++s) {
      const T v = partial[(size_t)s * np + p];
      re += (double)v.x;
      im += (double)v.y;
    }
    out[p] = make_double2(re * back, im * back);
  }
}

unsigned grid_for(size_t n) {
  return (unsigned)std::min<size_t>((n + 255) / 256, 256u * 64u);
}

}  // namespace

void fresnel_release(xrtk_ctx* ctx) { ctx->fresnel.release(); }

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
  if ((uint64_t)nx * nz > 0x7fff0000ull || n_rays > 0x7fffffffull) return XRTK_EINVAL;
  const bool f64 = opts && opts->mode == XRTK_MODE_F64;
  // FP32 factors in FP16 parts, unless engine = 1 asks for FP32 operands
  const bool f16 = !f64 && !(opts && opts->engine == XRTK_ENGINE_VALU);
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
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceScope scope(ctx->device);
    hipStream_t st = ctx->stream;
    Range whole("xrtk.fresnel");
    Phases ph;
    ph.next("xrtk.upload");
    // launch shape: tiles of the screen x ray slices; slices of at most
    // kSlice16 (FP16 parts) or kSlice32 (FP32 operands) rays for the FP32
    // sums, and enough of them for ~4 workgroups
    // per CU over the launch
    const uint32_t T = f64 ? kT64 : kT32, KB = f64 ? kKB64 : kKB32;
    const uint32_t n_tx = (nx + T - 1) / T, n_tz = (nz + T - 1) / T;
    const uint64_t tiles = (uint64_t)n_tx * n_tz;
    const uint64_t want = std::max<uint64_t>(1, (4ull * ctx->n_cu + tiles - 1) / tiles);
    const uint64_t cap = f16 ? kSlice16 : kSlice32;
    uint64_t n_sl = std::max<uint64_t>(want, f64 ? 1 : (n_rays + cap - 1) / cap);
    n_sl = std::min<uint64_t>(n_sl, (n_rays + 2 * KB - 1) / (2 * KB));
    // a slice is a whole number of block pairs (the kernels alternate two LDS
    // buffers); the rays are padded with zeros to whole slices plus the one
    // block the kernels generate past a slice's end
    uint64_t slice = (n_rays + n_sl - 1) / n_sl;
    slice = (slice + 2 * KB - 1) / (2 * KB) * (2 * KB);
    n_sl = (n_rays + slice - 1) / slice;
    const uint64_t nr_pad = n_sl * slice + KB;
    if (n_sl > 65535 || tiles > 0x7fffffffull) return XRTK_EINVAL;
    // device buffers: pixel axes, rays (x, z, Es; zero-padded to whole
    // slices), partial sums per slice, result
    const size_t psz = f64 ? sizeof(double2) : sizeof(float2);
    DBuf& W = ctx->fresnel;
    const size_t o_xp = 0, o_zp = (8 * (size_t)nx + 255) / 256 * 256;
    const size_t o_rx = o_zp + (8 * (size_t)nz + 255) / 256 * 256;
    const size_t o_rz = o_rx + (8 * nr_pad + 255) / 256 * 256;
    const size_t o_es = o_rz + (8 * nr_pad + 255) / 256 * 256;
    const size_t o_p = o_es + (16 * nr_pad + 255) / 256 * 256;
    const size_t o_u = o_p + (psz * n_sl * np + 255) / 256 * 256;
    const size_t total = o_u + sizeof(double2) * np;
    if (W.ensure(total) != hipSuccess) return XRTK_ENOMEM;
    char* base = W.as<char>();
    // The inputs are laid out on the host as they lie on the device (pixel
    // axes, ray x, ray z, scaled Es; zero padding past the rays) and go up
    // in one copy, from the context's page-locked block when it can hold
    // them (else from a pageable buffer); the result comes back through the
    // same block.
    char* pin = nullptr;
    if (ctx->ws.hpin.ensure(std::max<size_t>(o_p, sizeof(double2) * np)) == hipSuccess)
      pin = static_cast<char*>(ctx->ws.hpin.p);
    std::vector<char> pageable;
    if (!pin) pageable.resize(o_p);
    char* hs = pin ? pin : pageable.data();
    std::memset(hs, 0, o_p);
    std::memcpy(hs + o_xp, xp.data(), 8 * (size_t)nx);
    std::memcpy(hs + o_zp, zp.data(), 8 * (size_t)nz);
    std::memcpy(hs + o_rx, ray_x, 8 * n_rays);
    std::memcpy(hs + o_rz, ray_z, 8 * n_rays);
    double* es = reinterpret_cast<double*>(hs + o_es);
    // FP16 parts: amplitudes scaled by 2^-ex so the largest |Re| + |Im| is
    // in [1/2, 1) (exact; the merge multiplies by 2^ex)
    double scale = 1.0, back = 1.0;
    if (f16) {
      double amax = 0.0;
      for (size_t j = 0; j < n_rays; ++j)
        amax = std::max(amax, std::fabs(es_re[j]) + std::fabs(es_im[j]));
      if (amax > 0.0 && std::isfinite(amax)) {
        int ex = 0;
        (void)std::frexp(amax, &ex);
        scale = std::ldexp(1.0, -ex);
        back = std::ldexp(1.0, ex);
      }
    }
    for (size_t j = 0; j < n_rays; ++j) {
      es[2 * j] = es_re[j] * scale;
      es[2 * j + 1] = es_im[j] * scale;
    }
    hipError_t e = hipMemcpyAsync(base, hs, o_p, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return map_err(e);
    const double k = energy / kCHBAR_H * 1e7;
    const double krev = k / (2.0 * dist) / (2.0 * M_PI);
    const double* dxp = reinterpret_cast<const double*>(base + o_xp);
    const double* dzp = reinterpret_cast<const double*>(base + o_zp);
    const double* drx = reinterpret_cast<const double*>(base + o_rx);
    const double* drz = reinterpret_cast<const double*>(base + o_rz);
    const double2* des = reinterpret_cast<const double2*>(base + o_es);
    double2* U = reinterpret_cast<double2*>(base + o_u);
    ph.next("xrtk.pair");
    (void)hipEventRecord(ctx->ev[0], st);
    (void)hipEventRecord(ctx->ev[1], st);
    const dim3 grid((unsigned)tiles, (unsigned)n_sl);
    if (f64) {
      double2* Pd = reinterpret_cast<double2*>(base + o_p);
      hipLaunchKernelGGL(k_fresnel_mfma64, grid, dim3(256), 0, st, dxp, nx, dzp, nz, drx, drz, des,
                         (uint32_t)slice, krev, n_tx, Pd);
      e = hipGetLastError();
      (void)hipEventRecord(ctx->ev[2], st);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_fresnel_merge<double2>, dim3(grid_for(np)), dim3(256), 0, st, Pd,
                           (uint32_t)n_sl, np, 1.0, U);
    } else {
      float2* Pf = reinterpret_cast<float2*>(base + o_p);
      if (f16)
        hipLaunchKernelGGL(k_fresnel_mfma16, grid, dim3(256), 0, st, dxp, nx, dzp, nz, drx, drz,
                           des, (uint32_t)slice, krev, n_tx, Pf, -1.0f);
      else
        hipLaunchKernelGGL(k_fresnel_mfma32, grid, dim3(256), 0, st, dxp, nx, dzp, nz, drx, drz,
                           des, (uint32_t)slice, krev, n_tx, Pf);
      e = hipGetLastError();
      (void)hipEventRecord(ctx->ev[2], st);
      if (e == hipSuccess)
        hipLaunchKernelGGL(k_fresnel_merge<float2>, dim3(grid_for(np)), dim3(256), 0, st, Pf,
                           (uint32_t)n_sl, np, back, U);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return map_err(e);
    (void)hipEventRecord(ctx->ev[3], st);
    ph.next("xrtk.readback");
    const Segment seg{out2, sizeof(double2) * np};
    e = copy_d2h(st, pin, U, &seg, 1);
    if (e != hipSuccess) return map_err(e);
    float ms = 0.f, kms = 0.f;
    if (hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[3]) != hipSuccess) ms = 0.f;
    if (hipEventElapsedTime(&kms, ctx->ev[1], ctx->ev[2]) != hipSuccess) kms = 0.f;
    (void)hipGetLastError();
    ctx->last_pair_ms = kms;
    ctx->last_solve_ms = ms;
    return XRTK_OK;
  } catch (...) {
    return XRTK_ENOMEM;
  }
}
