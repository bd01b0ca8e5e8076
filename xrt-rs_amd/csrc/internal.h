// Internal layouts shared by the host runtime and the gfx950 kernels.
//
// HBM layout of one diffraction instance (one ray set at one energy on one
// screen):
//   RayRec64[nr]   resident f64 rays, 128 B each (AoS so that one wave reads a
//                  whole record with two scalar loads)
//   PixF64[np]     f64 screen points, 32 B each
//   RayRec32[nr]   per-solve FP32 records of the phase-reduced kernel, 48 B
//   PixRec32[np]   per-solve FP32 pixel records, 32 B
//   double R64[np] FP64 |p - c| for the base phase
//   double4 partial[n_chunks][np]  FP64 partial sums of each ray chunk
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

// FP32 record of one ray for the phase-reduced pair kernel.  o = r - c is
// formed in FP64 and rounded once.  oo = kr |o|^2 (kr = k / 2 pi, so phases
// are in revolutions).  (ar, ai) = i * Es' and (br, bi) = i * Ep', where Es'
// and Ep' carry the upstream phase and, in the quadratic-fold variant, the
// per-ray phase exp(i 2 pi kr |o|^2 / (2 R0)).  The factor k / (4 pi) is
// applied once per pixel in the FP64 merge.
struct RayRec32 {
  float ox, oy, oz, oo;
  float nx, ny, nz, nl;
  float ar, ai, br, bi;
};
static_assert(sizeof(RayRec32) == 48, "RayRec32 is 48 B");

// FP32 record of one pixel: q = p - c (rounded once from FP64), R = |q| in
// FP32, m = -2 kr q (formed in FP64, rounded once).
struct PixRec32 {
  float qx, qy, qz, R;
  float mx, my, mz, pad;
};
static_assert(sizeof(PixRec32) == 32, "PixRec32 is 32 B");

// Per-instance values computed on the device (centroid, bounds, flags).
struct InstScratch {
  double cx, cy, cz;       // FP64 ray centroid
  double R0;               // reference distance of the quadratic fold
  uint32_t omax_bits;      // max |r - c| as float bits (atomicMax)
  uint32_t rmin_bits;      // min |p - c| as float bits (atomicMin)
  uint32_t rmax_bits;      // max |p - c| as float bits (atomicMax)
  uint32_t flags;          // INST_F_* decided by the plan kernel
  uint32_t ep_nonzero;     // 1 if any ray has Ep != 0
  uint32_t pad[5];
};

enum : uint32_t {
  INST_F_CHECK = 1u,   // some pair may be near-singular: run the skip test
  INST_F_QFOLD = 2u,   // quadratic phase folded into the ray amplitudes
  INST_F_PATHUP = 4u,  // rays carry an upstream path phase
};

enum : uint32_t { MODE_F32R = 0, MODE_F64 = 1 };

// One instance as the pair and merge kernels see it.  Read through scalar
// loads: keep it a multiple of 16 B.
struct KInst {
  const RayRec64* rays64;   // resident rays (F64 kernel, prep kernels)
  const RayRec32* rays32;   // prepared FP32 rays (F32R kernel)
  const PixF64* pix64;      // f64 pixels
  const PixRec32* pix32;    // prepared FP32 pixels
  const double* R64;        // FP64 |p - c|
  double* partial;          // [n_chunks][np][4]
  double* out;              // [np][4]
  InstScratch* scr;         // device-computed values
  uint32_t nr, np, n_chunks, chunk_len;
  uint32_t mode;            // MODE_*
  uint32_t tile;            // rays per FP32 partial (F32R)
  uint32_t accumulate;      // merge adds into out (energy groups)
  uint32_t pad0;
  double energy;            // > 0: overrides the rays' energy column
  double k;                 // E / CHBAR * 1e7 of this instance [mm^-1]
  double kr_hi, kr_lo;      // k / (2 pi) as a double-double [rev / mm]
  double skip;              // skip pairs with path < skip [mm]
  double pad1;
};
static_assert(sizeof(KInst) % 16 == 0, "KInst size");

// One unit of pair-kernel work: a block of pixels of one instance against one
// chunk of its rays.  A workgroup of W waves handles W * 64 * P pixels.
struct WorkItem {
  uint32_t inst, pix0, chunk, pad;
};

// Kernel launch parameters chosen by the planner / autotuner.
struct LaunchCfg {
  int P = 4;          // pixels per lane
  int W = 4;          // waves per workgroup
  int ray_src = 1;    // 1 = SMEM, 2 = LDS
  int persistent = 0; // 1 = persistent groups on an atomic queue
  int tile = 1024;    // rays per FP32 partial
  int split = 0;      // ray chunks per pixel block (0 = planner)
};

}  // namespace xrtk
