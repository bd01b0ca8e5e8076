/*
 * Oracle: CPU restatement of the reference's Kirchhoff diffraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check or
 * time against the device engine.  The engine (libxrtk) never calls it.
 *
 * Functions and the reference code they restate (physwkim/xrt-rs):
 *   oracle_diffraction_f64   src/waves/diffraction.rs:56-117  (f64, complex
 *                            Kahan sums, skip path < 1e-30, rayon over pixels
 *                            -> OpenMP over pixels; output order = pixel order)
 *   oracle_diffraction_f64_sincos  the same with one glibc sincos() call per
 *                            pair instead of sin() and cos() (the lowering a
 *                            compiler may pick for Rust's sin_cos); the same
 *                            bits, timed beside it for the CPU baseline
 *   oracle_fallback_f32      src/gpu/fallback.rs:29-71 and the test helper
 *                            tests/golden_gpu_kirchhoff.rs:68-99 (f32 inputs
 *                            promoted to f64, skip < 1e-20, no Kahan)
 *   oracle_reduced_f32       the phase-reduced method of src/gpu/kirchhoff.rs
 *                            :96-115,276-297 + src/gpu/shaders/kirchhoff.wgsl
 *                            :65-129, in float32 (the reference GPU path)
 *   oracle_prepare_wave      src/waves/prepare_wave.rs:40-68
 *   oracle_fresnel_grid      the pixel grid of src/waves/fresnel.rs:66-73
 *   oracle_fresnel_f64       src/waves/fresnel.rs:53-106 (f64, plain complex
 *                            sum in ray order, rayon over pixels -> OpenMP;
 *                            output index iz * nx + ix)
 *   oracle_diffraction_ld    the same sum as oracle_diffraction_f64 with the
 *                            path and phase in x87 extended precision (64-bit
 *                            significand, sinl/cosl): the checker for phases
 *                            k*path ~ 1e12 rad, where any FP64 evaluation
 *                            rounds each term's phase by ~1e-4 rad
 *
 * Arithmetic follows the reference's operation order.  Build with
 * -ffp-contract=off and without fast-math (Rust neither contracts nor
 * reassociates).  Rust's f64::sin_cos calls libm sin and cos; glibc's are
 * used here.
 *
 * Extension (no counterpart in the reference): path_up, an upstream path per
 * ray; Es and Ep are multiplied by exp(i k path_up) before the pixel sum, the
 * term beam_to_diffraction_rays leaves to the caller (prepare_wave.rs:78-111).
 */
#define _GNU_SOURCE /* sincos */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#define CHBAR 1973.2697177417986 /* src/core/consts.rs:74 */
#define PI_F64 3.14159265358979323846

typedef struct {
  const double *x, *y, *z, *nx, *ny, *nz, *nl, *es_re, *es_im, *ep_re, *ep_im, *energy;
} oracle_rays;

/* diffraction_at_pixel, diffraction.rs:67-117 */
static void pixel_f64(const oracle_rays *r, const double *path_up, size_t nr, double px,
                      double py, double pz, int use_sincos, double *o4) {
  double es_re = 0, es_im = 0, ep_re = 0, ep_im = 0;     /* sums         */
  double ces_re = 0, ces_im = 0, cep_re = 0, cep_im = 0; /* compensators */
  for (size_t j = 0; j < nr; ++j) {
    const double dx = px - r->x[j];
    const double dy = py - r->y[j];
    const double dz = pz - r->z[j];
    const double path = sqrt(dx * dx + dy * dy + dz * dz);
    if (path < 1e-30) continue; /* :81 */
    const double ns = (r->nx[j] * dx + r->ny[j] * dy + r->nz[j] * dz) / path; /* :86 */
    const double k = r->energy[j] / CHBAR * 1e7;                               /* :89 */
    const double phase = k * path;                                             /* :92 */
    double sp, cp;                                                             /* :93 */
    if (use_sincos) {
      sincos(phase, &sp, &cp);
    } else {
      sp = sin(phase);
      cp = cos(phase);
    }
    const double obl = r->nl[j] + ns;                                          /* :96 */
    const double amp = k / (4.0 * PI_F64) * obl / path;                       /* :97 */
    /* u = i * amp * (cos + i sin) = (-amp sin, amp cos)                         :98 */
    const double ur = -(amp * sp), ui = amp * cp;
    double er = r->es_re[j], ei = r->es_im[j], qr = r->ep_re[j], qi = r->ep_im[j];
    if (path_up) {
      const double ph = k * path_up[j];
      const double s = sin(ph), c = cos(ph);
      const double er2 = er * c - ei * s, ei2 = er * s + ei * c;
      const double qr2 = qr * c - qi * s, qi2 = qr * s + qi * c;
      er = er2; ei = ei2; qr = qr2; qi = qi2;
    }
    /* Kahan for Es, :100-104 */
    {
      const double tr = (er * ur - ei * ui) - ces_re, ti = (er * ui + ei * ur) - ces_im;
      const double nr_ = es_re + tr, ni = es_im + ti;
      ces_re = (nr_ - es_re) - tr;
      ces_im = (ni - es_im) - ti;
      es_re = nr_;
      es_im = ni;
    }
    /* Kahan for Ep, :106-110 */
    {
      const double tr = (qr * ur - qi * ui) - cep_re, ti = (qr * ui + qi * ur) - cep_im;
      const double nr_ = ep_re + tr, ni = ep_im + ti;
      cep_re = (nr_ - ep_re) - tr;
      cep_im = (ni - ep_im) - ti;
      ep_re = nr_;
      ep_im = ni;
    }
  }
  o4[0] = es_re; o4[1] = es_im; o4[2] = ep_re; o4[3] = ep_im;
}

/* diffraction_integral, diffraction.rs:56-64.  Pixels are independent, so the
 * result does not depend on the thread count.  pix_stride > 1 evaluates every
 * pix_stride-th pixel only (output packed). */
static void diffraction_f64(const double *const cols[12], const double *path_up, size_t nr,
                            const double *px, const double *py, const double *pz, size_t np,
                            size_t pix_stride, int threads, int use_sincos, double *out4) {
  oracle_rays r = {cols[0], cols[1], cols[2], cols[3], cols[4], cols[5],
                   cols[6], cols[7], cols[8], cols[9], cols[10], cols[11]};
  if (pix_stride < 1) pix_stride = 1;
  const long long nout = (long long)((np + pix_stride - 1) / pix_stride);
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : 1)
  for (long long q = 0; q < nout; ++q) {
    const size_t p = (size_t)q * pix_stride;
    pixel_f64(&r, path_up, nr, px[p], py[p], pz[p], use_sincos, out4 + 4 * q);
  }
}

void oracle_diffraction_f64(const double *const cols[12], const double *path_up, size_t nr,
                            const double *px, const double *py, const double *pz, size_t np,
                            size_t pix_stride, int threads, double *out4) {
  diffraction_f64(cols, path_up, nr, px, py, pz, np, pix_stride, threads, 0, out4);
}

void oracle_diffraction_f64_sincos(const double *const cols[12], const double *path_up,
                                   size_t nr, const double *px, const double *py,
                                   const double *pz, size_t np, size_t pix_stride, int threads,
                                   double *out4) {
  diffraction_f64(cols, path_up, nr, px, py, pz, np, pix_stride, threads, 1, out4);
}

/* fallback.rs:29-71 over GpuRay-typed (f32) inputs: 12 floats per ray in the
 * GpuRay field order, 4 floats per pixel. */
void oracle_fallback_f32(const float *rays, size_t nr, const float *pix, size_t np,
                         int threads, double *out4) {
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads > 0 ? threads : 1)
  for (long long p = 0; p < (long long)np; ++p) {
    double es_re = 0, es_im = 0, ep_re = 0, ep_im = 0;
    const double px = pix[4 * p], py = pix[4 * p + 1], pz = pix[4 * p + 2];
    for (size_t j = 0; j < nr; ++j) {
      const float *r = rays + 12 * j;
      const double dx = px - (double)r[0];
      const double dy = py - (double)r[1];
      const double dz = pz - (double)r[2];
      const double path = sqrt(dx * dx + dy * dy + dz * dz);
      if (path < 1e-20) continue;                                           /* :48 */
      const double inv_path = 1.0 / path;
      const double ns = ((double)r[3] * dx + (double)r[4] * dy + (double)r[5] * dz) * inv_path;
      const double k = (double)r[7] / CHBAR * 1e7;
      const double phase = k * path;
      const double sp = sin(phase), cp = cos(phase);
      const double obl = (double)r[6] + ns;
      const double amp = k / (4.0 * PI_F64) * obl * inv_path;               /* :59 */
      const double ur = -amp * sp, ui = amp * cp;                           /* :60-61 */
      es_re += (double)r[8] * ur - (double)r[9] * ui;
      es_im += (double)r[8] * ui + (double)r[9] * ur;
      ep_re += (double)r[10] * ur - (double)r[11] * ui;
      ep_im += (double)r[10] * ui + (double)r[11] * ur;
    }
    out4[4 * p] = es_re; out4[4 * p + 1] = es_im;
    out4[4 * p + 2] = ep_re; out4[4 * p + 3] = ep_im;
  }
}

/* The reference GPU path in float32: host f64 centroid and mean energy
 * (kirchhoff.rs:96-104), f32-cast centroid in the shader (:106-115), the
 * shader's reduced phase (kirchhoff.wgsl:65-129, with sin/cos of the f32
 * argument evaluated in f64 and rounded), and the host f64 base phase
 * exp(i kbar |p - c64|) (kirchhoff.rs:276-297). */
void oracle_reduced_f32(const float *rays, size_t nr, const float *pix, size_t np,
                        double *out4) {
  double cx = 0, cy = 0, cz = 0, me = 0;
  for (size_t j = 0; j < nr; ++j) {
    cx += (double)rays[12 * j];
    cy += (double)rays[12 * j + 1];
    cz += (double)rays[12 * j + 2];
    me += (double)rays[12 * j + 7];
  }
  const double n = (double)nr;
  cx /= n; cy /= n; cz /= n; me /= n;
  const double k64 = me / CHBAR * 1e7;
  const float fcx = (float)cx, fcy = (float)cy, fcz = (float)cz;
  const float chb = (float)(1.0 / CHBAR * 1e7);
  const float PI32 = 3.14159265358979323846f;
  for (size_t p = 0; p < np; ++p) {
    const float px = pix[4 * p], py = pix[4 * p + 1], pz = pix[4 * p + 2];
    const float rdx = px - fcx, rdy = py - fcy, rdz = pz - fcz;
    const float ref = sqrtf(rdx * rdx + rdy * rdy + rdz * rdz);
    float es_re = 0, es_im = 0, ep_re = 0, ep_im = 0;
    for (size_t j = 0; j < nr; ++j) {
      const float *r = rays + 12 * j;
      const float dx = px - r[0], dy = py - r[1], dz = pz - r[2];
      const float path = sqrtf(dx * dx + dy * dy + dz * dz);
      if (path < 1e-20f) continue;
      const float inv = 1.0f / path;
      const float ns = (r[3] * dx + r[4] * dy + r[5] * dz) * inv;
      const float k = r[7] * chb;
      const float ox = fcx - r[0], oy = fcy - r[1], oz = fcz - r[2];
      const float dsq = ox * (2.0f * px - fcx - r[0]) + oy * (2.0f * py - fcy - r[1]) +
                        oz * (2.0f * pz - fcz - r[2]);
      const float dpath = dsq / (path + ref);
      const float dph = k * dpath;
      const float cs = (float)cos((double)dph), sn = (float)sin((double)dph);
      const float obl = r[6] + ns;
      const float amp = k / (4.0f * PI32) * obl * inv;
      const float ur = -amp * sn, ui = amp * cs;
      es_re += r[8] * ur - r[9] * ui;
      es_im += r[8] * ui + r[9] * ur;
      ep_re += r[10] * ur - r[11] * ui;
      ep_im += r[10] * ui + r[11] * ur;
    }
    const double ddx = (double)px - cx, ddy = (double)py - cy, ddz = (double)pz - cz;
    const double refp = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    const double bp = k64 * refp;
    const double c = cos(bp), s = sin(bp);
    out4[4 * p] = (double)es_re * c - (double)es_im * s;
    out4[4 * p + 1] = (double)es_re * s + (double)es_im * c;
    out4[4 * p + 2] = (double)ep_re * c - (double)ep_im * s;
    out4[4 * p + 3] = (double)ep_re * s + (double)ep_im * c;
  }
}

/* prepare_wave.rs:40-68 */
void oracle_prepare_wave(const double center[3], double dx, double dz, size_t nx, size_t nz,
                         double *px, double *py, double *pz) {
  size_t k = 0;
  for (size_t ix = 0; ix < nx; ++ix) {
    const double fx = nx > 1 ? (double)ix / (double)(nx - 1) - 0.5 : 0.0;
    const double x = center[0] + fx * 2.0 * dx;
    for (size_t iz = 0; iz < nz; ++iz) {
      const double fz = nz > 1 ? (double)iz / (double)(nz - 1) - 0.5 : 0.0;
      px[k] = x;
      py[k] = center[1];
      pz[k] = center[2] + fz * 2.0 * dz;
      ++k;
    }
  }
}

/* Extended-precision restatement for large phases: path and k*path in long
 * double (rounding ~ phase * 2^-64), the amplitude in long double, Kahan sums
 * in long double.  Inputs are the same f64 columns. */
static void pixel_ld(const oracle_rays *r, const double *path_up, size_t nr, double px,
                     double py, double pz, double *o4) {
  long double es_re = 0, es_im = 0, ep_re = 0, ep_im = 0;
  const long double PI_L = 3.141592653589793238462643383279502884L;
  for (size_t j = 0; j < nr; ++j) {
    const long double dx = (long double)px - r->x[j];
    const long double dy = (long double)py - r->y[j];
    const long double dz = (long double)pz - r->z[j];
    const long double path = sqrtl(dx * dx + dy * dy + dz * dz);
    if (path < 1e-30L) continue;
    const long double ns = (r->nx[j] * dx + r->ny[j] * dy + r->nz[j] * dz) / path;
    const long double k = (long double)(r->energy[j] / CHBAR * 1e7);  /* the f64 k */
    const long double phase = k * path;
    const long double sp = sinl(phase), cp = cosl(phase);
    const long double amp = k / (4.0L * PI_L) * (r->nl[j] + ns) / path;
    const long double ur = -(amp * sp), ui = amp * cp;
    long double er = r->es_re[j], ei = r->es_im[j], qr = r->ep_re[j], qi = r->ep_im[j];
    if (path_up) {
      const long double ph = k * path_up[j];
      const long double s = sinl(ph), c = cosl(ph);
      const long double er2 = er * c - ei * s, ei2 = er * s + ei * c;
      const long double qr2 = qr * c - qi * s, qi2 = qr * s + qi * c;
      er = er2; ei = ei2; qr = qr2; qi = qi2;
    }
    es_re += er * ur - ei * ui;
    es_im += er * ui + ei * ur;
    ep_re += qr * ur - qi * ui;
    ep_im += qr * ui + qi * ur;
  }
  o4[0] = (double)es_re; o4[1] = (double)es_im; o4[2] = (double)ep_re; o4[3] = (double)ep_im;
}

void oracle_diffraction_ld(const double *const cols[12], const double *path_up, size_t nr,
                           const double *px, const double *py, const double *pz, size_t np,
                           size_t pix_stride, int threads, double *out4) {
  oracle_rays r = {cols[0], cols[1], cols[2], cols[3], cols[4], cols[5],
                   cols[6], cols[7], cols[8], cols[9], cols[10], cols[11]};
  if (pix_stride < 1) pix_stride = 1;
  const long long nout = (long long)((np + pix_stride - 1) / pix_stride);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
  for (long long q = 0; q < nout; ++q) {
    const size_t p = (size_t)q * pix_stride;
    pixel_ld(&r, path_up, nr, px[p], py[p], pz[p], out4 + 4 * q);
  }
}

/* Fresnel screen axis (src/waves/fresnel.rs:66-73):
 * v[i] = lo + (hi - lo) * i / max(n - 1, 1), evaluated as
 * lo + (((hi - lo) * i) / d) like the reference expression. */
void oracle_fresnel_grid(double lo, double hi, size_t n, double *v) {
  const double d = (double)(n > 1 ? n - 1 : 1);
  for (size_t i = 0; i < n; ++i) v[i] = lo + (hi - lo) * (double)i / d;
}

/* fresnel_propagate (src/waves/fresnel.rs:53-106):
 *   U(x', z') = sum_j Es_j exp(prefactor * ((x' - x_j)^2 + (z' - z_j)^2)),
 *   prefactor = i k / (2 dist), k = E / CHBAR * 1e7.
 * prefactor * r2 is (0, (k / (2 dist)) * r2); its exp is (cos, sin) of that
 * angle (num-complex from_polar(exp(0), im)); the product with Es and the
 * sum follow num-complex's textbook formulas in ray order.  Pixels are
 * evaluated for q = 0, stride, 2 stride, ... of the iz * nx + ix order. */
void oracle_fresnel_f64(const double *rx, const double *rz, const double *es_re,
                        const double *es_im, size_t nr, double energy, double dist,
                        const double *xp, size_t nx, const double *zp, size_t nz,
                        size_t pix_stride, int threads, double *out2) {
  const double k = energy / CHBAR * 1e7;
  const double pim = k / (2.0 * dist);
  if (pix_stride < 1) pix_stride = 1;
  const size_t np = nx * nz;
  const long long nout = (long long)((np + pix_stride - 1) / pix_stride);
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads > 0 ? threads : 1)
  for (long long q = 0; q < nout; ++q) {
    const size_t p = (size_t)q * pix_stride;
    const size_t iz = p / nx, ix = p % nx;
    const double x = xp[ix], z = zp[iz];
    double sr = 0.0, si = 0.0;
    for (size_t j = 0; j < nr; ++j) {
      const double dx = x - rx[j];
      const double dz = z - rz[j];
      const double r2 = dx * dx + dz * dz;
      const double th = pim * r2;
      const double c = cos(th), s = sin(th);
      sr += es_re[j] * c - es_im[j] * s;
      si += es_re[j] * s + es_im[j] * c;
    }
    out2[2 * q] = sr;
    out2[2 * q + 1] = si;
  }
}

/* Timing helper for the CPU baseline: pairs per second of
 * oracle_diffraction_f64 is measured by the caller; this returns the number
 * of OpenMP threads the runtime grants. */
int oracle_max_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}
