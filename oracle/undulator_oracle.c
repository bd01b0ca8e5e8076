/*
 * Oracle: CPU restatement of the reference's undulator intensity map.
 *
 * TEST INFRASTRUCTURE ONLY, like the rest of liboracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline legs load it.
 *
 * Functions and the reference code they restate (physwkim/xrt-rs):
 *   oracle_undulator_gamma        src/sources/bending_magnet.rs:38-53
 *                                 (gamma = E[GeV] 1e3 / M0C2)
 *   oracle_undulator_e1           src/sources/undulator.rs:101-110
 *                                 (fundamental_energy)
 *   oracle_undulator_i_map        src/sources/undulator.rs:120-244
 *                                 (build_i_map, f64, with sinc :439-445 and
 *                                 period_sum :457-463); n_steps is 64 there,
 *                                 a parameter here (undulator_gpu passes it)
 *   oracle_undulator_wgsl_f32     src/gpu/shaders/undulator.wgsl:60-169 in
 *                                 float32 (the reference GPU path), as
 *                                 src/gpu/undulator.rs:93-275 feeds it
 * Constants: src/core/consts.rs.  Operation order follows the reference;
 * build with -ffp-contract=off.
 */
#include <math.h>
#include <stddef.h>

#define U_PI 3.14159265358979323846
#define U_PI2 (2.0 * U_PI)
#define U_SIE0 1.602176565e-19          /* consts.rs:18 */
#define U_SIHPLANCK 6.626069573e-34     /* consts.rs:48 */
#define U_SIC (2.99792458e10 * 1e-2)    /* consts.rs:21,51 */
#define U_FINE_STR (1.0 / 137.03599976) /* consts.rs:54 */
#define U_E2W 1519267514747457.9        /* consts.rs:57 */
#define U_M0C2 0.510998928              /* consts.rs:33 */

double oracle_undulator_gamma(double electron_energy_gev) {
  return electron_energy_gev * 1e3 / U_M0C2;
}

double oracle_undulator_e1(double gamma, double kx, double ky, double period_mm) {
  const double lambda_u = period_mm * 1e-3;
  const double k2 = kx * kx + ky * ky;
  const double e1 = 2.0 * gamma * gamma * U_SIHPLANCK * U_SIC / (lambda_u * (1.0 + k2 / 2.0));
  return e1 / U_SIE0;
}

static double sinc_f64(double x) {
  if (fabs(x) < 0.1) {
    const double x2 = x * x;
    return 1.0 - x2 / 6.0 * (1.0 - x2 / 20.0 * (1.0 - x2 / 42.0));
  }
  return sin(x) / x;
}

/* period_sum: exp(i (N-1) d) N sinc(N d) / sinc(d), d = half-phase folded
 * onto +-pi/2 (Rust's f64::round rounds half away from zero, as C round). */
static void period_sum_f64(size_t n_periods, double d_phase, double *re, double *im) {
  const double n = (double)n_periods;
  const double half = 0.5 * d_phase;
  const double d = half - U_PI * round(half / U_PI);
  const double arg = (n - 1.0) * d;
  const double m = n * sinc_f64(n * d) / sinc_f64(d);
  *re = cos(arg) * m;
  *im = sin(arg) * m;
}

/* build_i_map.  gamma2 is passed as the reference keeps it (gamma * gamma).
 * Outputs: intensity[n], amp_s[2n], amp_p[2n] (re, im). */
void oracle_undulator_i_map(double gamma, double gamma2, double beam_current, double kx,
                            double ky, double period_mm, size_t n_periods, double phase_deg,
                            int n_steps, const double *energies, const double *thetas,
                            const double *psis, size_t n, int threads, double *intensity,
                            double *amp_s, double *amp_p) {
  const double lambda_u = period_mm * 1e-3;
  const double phase_rad = phase_deg * (U_PI / 180.0); /* f64::to_radians */
  const double k2 = kx * kx + ky * ky;
  const double q = (1.0 + 0.5 * k2) / (2.0 * gamma2);
  const double wiggle_amp = lambda_u / (8.0 * U_PI2 * gamma2);
  const double amp2flux = U_FINE_STR * beam_current / U_SIE0 * (double)n_periods;
  const double dt = 1.0 / (double)n_steps;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads > 0 ? threads : 1)
  for (long long ii = 0; ii < (long long)n; ++ii) {
    const size_t i = (size_t)ii;
    const double e = energies[i], theta = thetas[i], psi = psis[i];
    const double omega = e * U_E2W;
    const double off_axis = theta * theta + psi * psi;
    const double dirz = sqrt(1.0 - off_axis);
    const double wc = omega / (U_SIC * (1.0 - q));
    const double slip = off_axis / (1.0 + dirz) + q * dirz;
    double axr = 0.0, axi = 0.0, azr = 0.0, azi = 0.0;
    for (int j = 0; j < n_steps; ++j) {
      const double t = ((double)j + 0.5) * dt;
      const double phi_t = U_PI2 * t;
      const double beta_x = ky / gamma * sin(phi_t);
      const double beta_z = -kx / gamma * sin(phi_t + phase_rad);
      const double amp_x = ky * lambda_u / (U_PI2 * gamma);
      const double amp_z = kx * lambda_u / (U_PI2 * gamma);
      const double x_e = -amp_x * cos(phi_t);
      const double z_e = amp_z * cos(phi_t + phase_rad);
      const double s = lambda_u * t;
      const double wiggle =
          wiggle_amp * (ky * ky * sin(2.0 * phi_t) + kx * kx * sin(2.0 * phi_t + 2.0 * phase_rad));
      const double tp = wc * (s * slip - dirz * wiggle - x_e * theta - z_e * psi);
      const double c = cos(tp), sn = sin(tp);
      const double bx = beta_x - theta, bz = beta_z - psi;
      axr += bx * c * dt;
      axi += bx * sn * dt;
      azr += bz * c * dt;
      azi += bz * sn * dt;
    }
    double rr, ri;
    period_sum_f64(n_periods, wc * lambda_u * slip, &rr, &ri);
    /* ax = ax_sum * resonance * gamma2 */
    double xr = (axr * rr - axi * ri) * gamma2, xi = (axr * ri + axi * rr) * gamma2;
    double zr = (azr * rr - azi * ri) * gamma2, zi = (azr * ri + azi * rr) * gamma2;
    const double is_val = xr * xr - xi * -xi;
    const double ip_val = zr * zr - zi * -zi;
    const double inv_e = 1.0 / e;
    intensity[i] = amp2flux * inv_e * (is_val + ip_val);
    const double sf = sqrt(amp2flux * inv_e);
    amp_s[2 * i] = xr * sf;
    amp_s[2 * i + 1] = xi * sf;
    amp_p[2 * i] = zr * sf;
    amp_p[2 * i + 1] = zi * sf;
  }
}

/* --- the reference GPU path in float32 (undulator.wgsl:60-169) ---------- */
static float sinc_f32(float x) {
  if (fabsf(x) < 0.1f) {
    const float x2 = x * x;
    return 1.0f - x2 / 6.0f * (1.0f - x2 / 20.0f * (1.0f - x2 / 42.0f));
  }
  return sinf(x) / x;
}

/* Parameters are rounded to f32 as undulator.rs:111-124 does; WGSL round()
 * rounds half to even (rintf). out5 per point: ax_re, ax_im, az_re, az_im,
 * intensity. */
void oracle_undulator_wgsl_f32(double gamma_d, double beam_current, double kx_d, double ky_d,
                               double period_mm, size_t n_periods, double phase_deg,
                               int n_steps, const float *obs4, size_t n, double *out5) {
  const double amp2flux_d = U_FINE_STR * beam_current / U_SIE0 * (double)n_periods;
  const float kx = (float)kx_d, ky = (float)ky_d, gamma = (float)gamma_d;
  const float gamma2 = (float)(gamma_d * gamma_d);
  const float period_m = (float)(period_mm * 1e-3);
  const float phase_rad = (float)(phase_deg * (U_PI / 180.0));
  const float amp2flux = (float)amp2flux_d;
  const float PI = 3.14159265358979323846f, TWO_PI = 6.28318530717958647692f;
  const float E2W = 1519267514747457.9f, SIC = 299792458.0f;
  for (size_t idx = 0; idx < n; ++idx) {
    const float energy = obs4[4 * idx], theta = obs4[4 * idx + 1], psi = obs4[4 * idx + 2];
    const float omega = energy * E2W;
    const float dt = 1.0f / (float)n_steps;
    float ax_re = 0.f, ax_im = 0.f, az_re = 0.f, az_im = 0.f;
    const float k2 = kx * kx + ky * ky;
    const float q = (1.0f + 0.5f * k2) / (2.0f * gamma2);
    const float off_axis = theta * theta + psi * psi;
    const float dirz = sqrtf(1.0f - off_axis);
    const float slip = off_axis / (1.0f + dirz) + q * dirz;
    const float wc = omega / (SIC * (1.0f - q));
    const float wiggle_amp = period_m / (8.0f * TWO_PI * gamma2);
    for (int j = 0; j < n_steps; ++j) {
      const float t = ((float)j + 0.5f) * dt;
      const float phi_t = TWO_PI * t;
      const float beta_x = ky / gamma * sinf(phi_t);
      const float beta_z = -kx / gamma * sinf(phi_t + phase_rad);
      const float amp_x = ky * period_m / (TWO_PI * gamma);
      const float amp_z = kx * period_m / (TWO_PI * gamma);
      const float x_e = -amp_x * cosf(phi_t);
      const float z_e = amp_z * cosf(phi_t + phase_rad);
      const float s = period_m * t;
      const float wiggle = wiggle_amp * (ky * ky * sinf(2.0f * phi_t) +
                                         kx * kx * sinf(2.0f * phi_t + 2.0f * phase_rad));
      const float tp = wc * (s * slip - dirz * wiggle - x_e * theta - z_e * psi);
      const float cp = cosf(tp), sp = sinf(tp);
      const float bx = beta_x - theta, bz = beta_z - psi;
      ax_re += bx * cp * dt;
      ax_im += bx * sp * dt;
      az_re += bz * cp * dt;
      az_im += bz * sp * dt;
    }
    const float n_per = (float)n_periods;
    const float half = 0.5f * wc * period_m * slip;
    const float d = half - PI * rintf(half / PI);
    const float mag = n_per * sinc_f32(n_per * d) / sinc_f32(d);
    const float arg = (n_per - 1.0f) * d;
    const float rr = cosf(arg) * mag, ri = sinf(arg) * mag;
    float xr = (ax_re * rr - ax_im * ri) * gamma2, xi = (ax_re * ri + ax_im * rr) * gamma2;
    float zr = (az_re * rr - az_im * ri) * gamma2, zi = (az_re * ri + az_im * rr) * gamma2;
    const float is_val = xr * xr + xi * xi, ip_val = zr * zr + zi * zi;
    const float inv_e = 1.0f / energy;
    const float total = amp2flux * inv_e * (is_val + ip_val);
    const float sf = sqrtf(amp2flux * inv_e);
    out5[5 * idx] = xr * sf;
    out5[5 * idx + 1] = xi * sf;
    out5[5 * idx + 2] = zr * sf;
    out5[5 * idx + 3] = zi * sf;
    out5[5 * idx + 4] = total;
  }
}
