// Host-side helpers of the C ABI that need no device: prepare_wave's grid
// and beam_to_diffraction_rays' gather (src/waves/prepare_wave.rs:40-111).
// Plain C++ with no HIP dependency, so that the same source builds into the
// sanitizer checks of the CPU code.  Built without FP contraction, as the
// reference's Rust is.
#include <cstddef>
#include <cstdint>

#include "../../include/xrtk.h"

extern "C" {

int xrtk_prepare_wave(const XrtkWaveParams* w, double* px, double* py, double* pz) {
  if (!w) return XRTK_EINVAL;
  const size_t n = (size_t)w->nx * w->nz;
  if (n && (!px || !py || !pz)) return XRTK_EINVAL;
  // prepare_wave.rs:43-65, same operation order (built without contraction)
  size_t k = 0;
  for (uint32_t ix = 0; ix < w->nx; ++ix) {
    const double fx = w->nx > 1 ? (double)ix / (double)(w->nx - 1) - 0.5 : 0.0;
    const double x = w->center[0] + fx * 2.0 * w->dx;
    for (uint32_t iz = 0; iz < w->nz; ++iz) {
      const double fz = w->nz > 1 ? (double)iz / (double)(w->nz - 1) - 0.5 : 0.0;
      px[k] = x;
      py[k] = w->center[1];
      pz[k] = w->center[2] + fz * 2.0 * w->dz;
      ++k;
    }
  }
  return XRTK_OK;
}

int xrtk_beam_to_rays(const double* bx, const double* by, const double* bz, const double* ba,
                      const double* bb, const double* bc, const double* be, const double* es_re,
                      const double* es_im, const double* ep_re, const double* ep_im,
                      size_t n_beam, const double* normals3, const size_t* good, size_t n_good,
                      double* const out_cols[12]) {
  if (!out_cols) return XRTK_EINVAL;
  if (n_good && (!bx || !by || !bz || !ba || !bb || !bc || !be || !normals3 || !good))
    return XRTK_EINVAL;
  for (int c = 0; c < 12; ++c)
    if (n_good && !out_cols[c]) return XRTK_EINVAL;
  const bool amps = es_re && es_im && ep_re && ep_im;
  for (size_t k = 0; k < n_good; ++k) {
    if (good[k] >= n_beam) return XRTK_EINVAL;
  }
  for (size_t k = 0; k < n_good; ++k) {
    const size_t i = good[k];
    const double* n = normals3 + 3 * i;
    // prepare_wave.rs:86-87: nl = n . (a, b, c)
    const double nl = n[0] * ba[i] + n[1] * bb[i] + n[2] * bc[i];
    out_cols[0][k] = bx[i]; out_cols[1][k] = by[i]; out_cols[2][k] = bz[i];
    out_cols[3][k] = n[0]; out_cols[4][k] = n[1]; out_cols[5][k] = n[2];
    out_cols[6][k] = nl;
    out_cols[7][k] = amps ? es_re[i] : 1.0;  // prepare_wave.rs:89-95
    out_cols[8][k] = amps ? es_im[i] : 0.0;
    out_cols[9][k] = amps ? ep_re[i] : 0.0;
    out_cols[10][k] = amps ? ep_im[i] : 0.0;
    out_cols[11][k] = be[i];
  }
  return XRTK_OK;
}

}  // extern "C"
