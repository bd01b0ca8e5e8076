// Internal layouts shared by the host runtime and the gfx950 kernels.
//
// HBM layout of one diffraction instance (one ray set at one energy on one
// screen):
//   RayRec64[nr]   resident f64 rays, 128 B each (AoS so that one wave reads a
//                  whole record with two scalar loads)
//   PixF64[npad]   f64 screen points in slot order (tiles of 64*P slots; a
//                  grid screen is laid out tile by tile, padding slots repeat a
//                  real pixel), 32 B each, with perm[slot] -> output index
//   RayRec32[nr]   per-solve FP32 ray records of the reduced kernel, 64 B
//   PixRec32[npad] per-solve FP32 slot records, 48 B
//   TileRec[ntile] per-solve FP64 tile reference points, 32 B
//   double R64[2 * npad]           |p - c| per slot as a double-double
//   partial[n_chunks][npad]        per-chunk partial sums (float4 from the
//                                  FP32 kernel, double4 from the FP64 one)
//   double4 out[np]                final {Es.re, Es.im, Ep.re, Ep.im}
#pragma once
#include <stdint.h>

namespace xrtk {

// Resident ray record: the fields of DiffractionRay (src/waves/diffraction.rs
// :15-33) plus an upstream optical path (the beam.path term that
// beam_to_diffraction_rays leaves to the caller, src/waves/prepare_wave.rs
// :78-111).
struct RayRec64 {
  double x, y, z;
  double nx, ny, nz;
  double nl;
  double energy;
  double es_re, es_im, ep_re, ep_im;
  double path_up;
  double pad0, pad1, pad2;
};
static_assert(sizeof(RayRec64) == 128, "RayRec64 is 128 B");

struct PixF64 {
  double x, y, z, pad;
};

// FP32 record of one ray for the phase-reduced pair kernel.
//   o  = r - c, formed in FP64 and rounded once; every later use of the ray
//        position (phase, path, fold) uses this rounded o, so the geometry is
//        exactly that of a ray at c + o.  Stored in the kernel axis order
//        (kernel_axes): an axis along which every o is zero comes last.
//   oo = kr |o|^2 (kr = k / 2 pi: phases are in revolutions), read by the
//        centroid-referenced variant; the tile-referenced variant keeps dz
//        there instead.
//   (dx, dy, dz) = o_f32 - o in FP64 rounded to FP32: the rounding of the
//        position, whose phase kr delta.(q_t - o)/|q_t - o| the
//        tile-referenced variant adds per (ray, tile).
//   Amplitudes in polarisation form, after the upstream path phase:
//        Es = am e^{i th} cs,  Ep = am e^{i th} (pr + i pi),
//        am = sqrt(|Es|^2 + |Ep|^2), cs = |Es| / am >= 0.
//   am scales the geometric factors: n = -am n_ray / (2 kr) (kernel axis
//   order), nl = am nl_ray, no = am n_ray . o.
//   off = th / 2 pi + 1/4 (the quarter turn is the factor i of the
//        integrand), reduced to [-1/2, 1/2] revolutions; the
//        centroid-referenced variant also adds kr delta.u here (u: direction
//        to the screen centre).
// The factor k/(4 pi) is applied once per pixel in the FP64 merge.
struct RayRec32 {
  float ox, oy, oz, oo;
  float nx, ny, nz, nl;
  float cs, pr, pi, no;
  float am, off, dx, dy;
};
static_assert(sizeof(RayRec32) == 64, "RayRec32 is 64 B");

// FP32 record of one pixel slot.  With the slot's tile reference point t
// (TileRec) and q = p - c, dq = q - q_t:
//   m  = -2 kr dq   (FP64, rounded once; kernel axis order)
//   R2 = |q|^2, R = |q|
//   g  = (R_t - R) / R_t
//   nq = n . q   (only when every ray has the same normal n)
//   paraxial expansion constants (rho = (path^2 - R^2) / R^2 small):
//     1 / path       = rinv - num c2,  rinv = 1 / R,     c2 = 1 / (2 kr R^3)
//     1 / (path + R) = r0 - num c4,    r0 = 1 / (2 R),   c4 = 1 / (8 kr R^3)
//   with num = kr (path^2 - R^2).
struct PixRec32 {
  float mx, my, mz, R2;
  float R, g, nq, r0;
  float c4, rinv, c2, pad;
};
static_assert(sizeof(PixRec32) == 48, "PixRec32 is 48 B");

// FP64 reference point of one tile of 64*P slots (its bounding-box centre),
// in the kernel axis order.
struct TileRec {
  double qx, qy, qz, R;
};

// One lane's slot line of one tile, anchored at the lane's slot 0 (from
// k_tile_prep): the step to q_0 + h, h the secant to the lane's last real
// slot: -2 kr h in the kernel axis order, and r0 = 1/(2|q|), c4 =
// 1/(8 kr |q|^3) at q_0 + h minus at q_0, and c4 at q_0 + h.
struct LineRec {
  float hx, hy, hz, dr0;
  float dc4, c4n, pad0, pad1;
};
static_assert(sizeof(LineRec) == 32, "LineRec is 32 B");

// A tile table: the TileRecs of npad / (64 P) tiles, then 64 LineRecs per
// tile (tile-major).
inline size_t tile_table_bytes(uint32_t npad, uint32_t P) {
  const size_t nt = npad / (64u * P);
  return nt * (sizeof(TileRec) + 64 * sizeof(LineRec));
}

// Per-instance values computed on the device (centroid, bounds, flags).
struct InstScratch {
  double cx, cy, cz;       // FP64 ray centroid
  double blk_rev;          // MFMA: bound of the block phase model's remainder [rev]
  uint32_t omax_bits[4];   // max |o_x|, |o_y|, |o_z|, |o| as float bits
  uint32_t qmax_bits[3];   // max |q_x|, |q_y|, |q_z|
  uint32_t dqmax_bits[3];  // max |dq_x|, |dq_y|, |dq_z| (slot to tile reference)
  uint32_t rmin_bits;      // min |q|
  uint32_t rmax_bits;      // max |q|
  uint32_t flags;          // INST_F_* decided by the plan kernel
  uint32_t ep_nonzero;     // 1 if any ray has Ep != 0
  uint32_t n_differs;      // 1 if some ray's normal differs from ray 0's
  uint32_t amax_bits;      // max of max(|cs|, |pr| + |pi|) over the FP32 ray
                           // records (float bits)
  double n0x, n0y, n0z;    // normal of ray 0
  double v0_rev, v2_rev;   // bounds of the reduced phase of the two variants
  uint32_t qlo_key[3];     // min q_x, q_y, q_z (order-preserving float keys)
  uint32_t qhi_key[3];     // max q_x, q_y, q_z
  // separable amplitude (uniform normal n0): ranges of the rays' nl, of
  // |n0.o| and of the slots' C = n0.q / |q|
  uint32_t nl_lo_key, nl_hi_key;
  uint32_t nu_max_bits;
  uint32_t c_lo_key, c_hi_key;
  // slot lines (a lane's slots q_0 + i h): max |h|, max |q_0.h| / (|q_0| |h|)
  // and max |q_i - q_0 - i h| (any axis) over the real slots
  uint32_t hmax_bits, sigma_bits, ldev_bits;
};

enum : uint32_t {
  INST_F_CHECK = 1u,      // F64 kernel: some pair may fall below the skip threshold
  INST_F_TILE = 2u,       // F32R: per-(ray, tile) phase folded into the phase offset
  INST_F_PATHUP = 4u,     // rays carry an upstream path phase
  INST_F_UNIFORM_N = 8u,  // every ray has the same normal
  INST_F_NOFRACT = 16u,   // |reduced phase| < 128 rev: v_sin/v_cos take it as is
  INST_F_LIN2 = 32u,      // F32R: o is zero along one axis (the last kernel axis)
  INST_F_PAX = 64u,       // F32R tile variant: 1/path and 1/(path + R) from the
                          // first-order paraxial expansion (rho <= 1e-4)
  INST_F_SEPAMP = 128u,   // F32R PAX variant: the amplitude as a ray factor times a
                          // slot factor (bound <= 1e-6)
  INST_F_LINE = 256u,     // F32R SEPAMP variant: each lane's slots lie on a line, the
                          // phase steps along it by one rotation per slot
                          // (bound <= 1e-7 rev)
  INST_F_LINEDEV = 1024u, // LINE with each slot's own offset from the line applied
                          // to first order (pixel lists off a line by rounding)
  INST_F_EPZERO = 2048u,  // F32R: every ray has Ep == 0, so the Ep sums are exact
                          // zeros and the pair kernel skips them
  INST_F_MFMA = 4096u,    // F32R grid screen: blocks of 64 x 64 pixels run on the
                          // matrix cores (phase linear over a block, bound <= 1e-6
                          // rev); the VALU pair kernels skip the instance
  INST_F_MFMA_OK = 8192u, // static flag from the host: this launch runs the MFMA
                          // items of the instance (a grid screen)
};

enum : uint32_t { MODE_F32R = 0, MODE_F64 = 1 };

// One instance as the pair and merge kernels see it.  Read through scalar
// loads: keep it a multiple of 16 B.
struct KInst {
  const RayRec64* rays64;   // resident rays (F64 kernel, prep kernels)
  const RayRec32* rays32;   // prepared FP32 rays (F32R kernel)
  const PixF64* pix64;      // f64 pixels in slot order
  const PixRec32* pix32;    // prepared FP32 slots
  const TileRec* tiles;     // tile reference points
  const uint32_t* perm;     // slot -> output index (~0u: padding); null = identity
  const double* R64;        // |p - c| per slot, double-double (hi, lo)
  void* partial;            // [n_chunks][npad] float4 (F32R) or double4 (F64)
  double* out;              // [np][4]
  InstScratch* scr;         // device-computed values
  uint32_t nr, np, npad, n_chunks;
  uint32_t chunk_len, mode, tile, accumulate;  // tile: unused by the kernels
  uint32_t tile_slots;      // slots per tile (64 * P of the F32R launch)
  uint32_t pad0, pad1, pad2;
  double energy;            // > 0: overrides the rays' energy column (F64)
  double k;                 // E / CHBAR * 1e7 of this instance [mm^-1]
  double kr_hi, kr_lo;      // k / (2 pi) as a double-double [rev / mm]
  double skip;              // skip pairs with path < skip [mm]
  double pad3;
  // prepare_wave grid of the screen (MFMA items): pixel (ix, iz) at
  // (gx0 + ix gsx, gy, gz0 + iz gsz); slots in tiles of gth x gtw pixels.
  // A float32 pixel list on such a grid: pixel (ix, iz) at
  // (gx0 + (ix + gex[ix]) gsx, gy, gz0 + (iz + gez[iz]) gsz), the offsets in
  // pixels, |offset| <= gemax (null offsets: the grid itself).
  double gx0, gy, gz0, gsx, gsz, gemax;
  uint32_t gnx, gnz, gth, gtw;
  const float* gex;
  const float* gez;
};
static_assert(sizeof(KInst) % 16 == 0, "KInst size");

// One unit of pair-kernel work: a block of slots of one instance against one
// chunk of its rays.  A workgroup of W waves handles W * 64 * P slots.  MFMA
// items: pix0 is the 64 x 64-pixel block and pad the sum (0 = Es, 1 = Ep).
struct WorkItem {
  uint32_t inst, pix0, chunk, pad;
};

// Kernel launch parameters chosen by the planner / autotuner.
struct LaunchCfg {
  int P = 4;          // pixels per lane
  int W = 4;          // waves per workgroup
  int ray_src = 1;    // 1 = SMEM (the only source of this build)
  int persistent = 1; // 1 = persistent groups on an atomic queue
  int tile = 8192;    // F32R: max rays per FP32 chunk partial
  int split = 0;      // ray chunks per pixel block (0 = planner)
};

}  // namespace xrtk
