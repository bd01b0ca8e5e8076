/*
 * xrtk.h — C ABI of the MI355X (gfx950) Kirchhoff diffraction engine.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never lets
 * an exception cross the boundary, and returns an int status (XRTK_OK == 0).
 * The caller owns every input and output buffer; inputs are borrowed for the
 * duration of the call only.  Ray sets uploaded with xrtk_rayset_* live in the
 * library until released.
 *
 * Reference interfaces replaced (physwkim/xrt-rs):
 *   GpuRay / GpuPixel layouts ......... src/gpu/kirchhoff.rs:19-45
 *   kirchhoff_gpu(ctx, rays, pixels) .. src/gpu/kirchhoff.rs:77-83
 *   kirchhoff_auto(rays, pixels) ...... src/gpu/fallback.rs:19-26
 *   GpuContext::new / ::shared ........ src/gpu/context.rs:20-31
 *   diffraction_integral(rays, pixels)  src/waves/diffraction.rs:56-64
 *   prepare_wave(WaveParams) .......... src/waves/prepare_wave.rs:40-68
 *   CHBAR ............................. src/core/consts.rs:74
 */
#ifndef XRTK_H
#define XRTK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define XRTK_OK        0
#define XRTK_ENODEV    1  /* no usable gfx950 device                        */
#define XRTK_EINVAL    2  /* NULL pointer, bad length or bad option         */
#define XRTK_ENOMEM    3  /* device allocation failed or HBM budget exceeded */
#define XRTK_ELAUNCH   4  /* a HIP launch or copy failed                     */
#define XRTK_EAGAIN    5  /* serving: result not ready yet                   */
#define XRTK_ETIMEDOUT 6  /* serving: wait timed out                         */
#define XRTK_ESHUTDOWN 7  /* serving: server is shutting down                */

/* Physical constant used for k = E / CHBAR * 1e7 [mm^-1] (eV * Angstrom).
 * Same value as src/core/consts.rs:74. */
#define XRTK_CHBAR 1973.2697177417986

/* ---- data layouts ------------------------------------------------------ */

/* Byte-identical to GpuRay (src/gpu/kirchhoff.rs:19-35), 48 bytes. */
typedef struct XrtkGpuRay {
    float x, y, z;        /* position on the optical element [mm] */
    float nx, ny, nz;     /* surface normal                       */
    float nl;             /* obliquity factor n . beam direction  */
    float energy;         /* [eV]                                 */
    float es_re, es_im;   /* s-polarised amplitude                */
    float ep_re, ep_im;   /* p-polarised amplitude                */
} XrtkGpuRay;

/* Byte-identical to GpuPixel (src/gpu/kirchhoff.rs:37-45), 16 bytes. */
typedef struct XrtkGpuPixel {
    float x, y, z, _pad;
} XrtkGpuPixel;

/* f64 rays as twelve structure-of-arrays columns, the field order of
 * DiffractionRay (src/waves/diffraction.rs:15-33) and of the Python binding
 * diffraction_integral_rs (crates/xrt-python/src/lib.rs:321-338).  Any
 * column may be shared between fields (for instance one zero column). */
typedef struct XrtkRaysF64 {
    const double *x, *y, *z;
    const double *nx, *ny, *nz;
    const double *nl;
    const double *es_re, *es_im;
    const double *ep_re, *ep_im;
    const double *energy;
} XrtkRaysF64;

/* Regular screen exactly as prepare_wave builds it
 * (src/waves/prepare_wave.rs:40-68): pixel (ix, iz) lies at
 * x = cx + (ix/(nx-1) - 0.5)*2*dx, y = cy, z = cz + (iz/(nz-1) - 0.5)*2*dz,
 * index ix*nz + iz; a dimension of 1 maps to the centre. */
typedef struct XrtkWaveParams {
    double center[3];
    double dx, dz;
    uint32_t nx, nz;
} XrtkWaveParams;

/* ---- options ----------------------------------------------------------- */

/* Arithmetic used on the device. */
#define XRTK_MODE_F32R 0  /* FP32 phase-reduced pair kernel, FP64 centroid,
                             base phase and merge (the default)            */
#define XRTK_MODE_F64  1  /* plain FP64 pair kernel with Kahan sums        */

/* Where the pair kernel reads ray records from: scalar loads into SGPRs
 * (wave-uniform), with the per-(ray, pixel tile) values staged in LDS. */
#define XRTK_RAYSRC_AUTO 0
#define XRTK_RAYSRC_SMEM 1

typedef struct XrtkOpts {
    int32_t mode;             /* XRTK_MODE_*                                */
    int32_t pixels_per_lane;  /* 0 = autotune, else 2, 4, 8 or 16 (F32R), 1
                                 or 2 (FP64)                                 */
    int32_t waves_per_group;  /* 0 = autotune, else 1, 2 or 4               */
    int32_t tile_rays;        /* 0 = autotune: rays per FP32 partial sum    */
    int32_t split;            /* 0 = autotune: ray chunks per pixel block   */
    int32_t ray_src;          /* XRTK_RAYSRC_*                              */
    int32_t persistent;       /* -1 = autotune, 0 = one group per item,
                                 1 = persistent groups pulling a queue       */
    int32_t engine;           /* XRTK_ENGINE_*                              */
} XrtkOpts;

/* Which pair kernels an FP32-reduced solve may use.  AUTO: an instance on a
 * prepare_wave grid whose phase is linear over 64 x 64-pixel blocks to
 * within 1e-6 rev (bounded by the device planner) runs on the matrix cores
 * (FP16-split MFMA products, FP32 block sums); every other instance, and
 * VALU always, on the FP32 VALU pair kernels. */
#define XRTK_ENGINE_AUTO 0
#define XRTK_ENGINE_VALU 1

/* Fills *o with the defaults (mode F32R, everything autotuned). */
void xrtk_opts_default(XrtkOpts *o);

/* ---- context (replaces GpuContext, src/gpu/context.rs:6-31) ------------ */

typedef struct xrtk_ctx xrtk_ctx;

/* Process-wide context, created on first use, thread-safe.  NULL when no
 * gfx950 device is present (the "is there a GPU?" answer kirchhoff_auto
 * uses, src/gpu/context.rs:24-31).  Never destroy it. */
xrtk_ctx *xrtk_ctx_shared(void);

/* Private context on device `device` (GpuContext::new, context.rs:20-22). */
int xrtk_ctx_create(int device, xrtk_ctx **out);
void xrtk_ctx_destroy(xrtk_ctx *ctx);

/* Adapter name (GpuContext::adapter_name, context.rs:70-72). */
int xrtk_ctx_device_name(const xrtk_ctx *ctx, char *buf, size_t len);

/* ---- one-shot solves --------------------------------------------------- */

/* kirchhoff_gpu (src/gpu/kirchhoff.rs:77-298): f32 rays and pixels in,
 * np x {es_re, es_im, ep_re, ep_im} f64 out (the memory order of
 * Vec<(Complex64, Complex64)>).  Pairs with path < 1e-20 mm are skipped as
 * in kirchhoff.wgsl:85-87.  0 rays gives zeros; 0 pixels writes nothing.
 * opts may be NULL. */
int xrtk_kirchhoff_gpu_f32(xrtk_ctx *ctx,
                           const XrtkGpuRay *rays, size_t n_rays,
                           const XrtkGpuPixel *pixels, size_t n_pixels,
                           double *out4, const XrtkOpts *opts);

/* The same solve as xrtk_kirchhoff_gpu_f32 (same arguments, same result bit
 * for bit), with every device buffer taken from a fixed 64 MiB block that
 * the library holds on each device for the life of the process: the call
 * makes no device allocation (no hipMalloc / hipFree).  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * A solve that needs more than the block is refused with XRTK_ENOMEM; a
 * monochromatic solve of 65,536 rays on 65,536 pixels needs ~43 MiB
 * (about 240 B per ray plus 150 B per pixel and 16 B per pixel per 4096-ray
 * chunk).  Calls on one device share the block: calls through one context
 * are serialised, calls through different contexts on one device must not
 * overlap. */
int xrtk_kirchhoff_gpu_f32_noalloc(xrtk_ctx *ctx,
                                   const XrtkGpuRay *rays, size_t n_rays,
                                   const XrtkGpuPixel *pixels, size_t n_pixels,
                                   double *out4, const XrtkOpts *opts);

/* kirchhoff_auto (src/gpu/fallback.rs:19-26): the shared context when a
 * device exists, else the f64 CPU loop of fallback.rs:29-71 (no Kahan, skip
 * below 1e-20 mm).  Setting XRTK_REQUIRE_GPU=1 in the environment turns the
 * CPU branch into XRTK_ENODEV.  Never fails otherwise. */
int xrtk_kirchhoff_auto_f32(const XrtkGpuRay *rays, size_t n_rays,
                            const XrtkGpuPixel *pixels, size_t n_pixels,
                            double *out4);

/* diffraction_integral (src/waves/diffraction.rs:56-117) on the device:
 * f64 rays (SoA) and pixel coordinates in, np x 4 f64 out, pairs with
 * path < 1e-30 mm skipped.  opts->mode picks FP32-reduced or plain FP64. */
int xrtk_diffraction_f64(xrtk_ctx *ctx, const XrtkRaysF64 *rays, size_t n_rays,
                         const double *pix_x, const double *pix_y,
                         const double *pix_z, size_t n_pixels,
                         double *out4, const XrtkOpts *opts);

/* diffraction_integral (src/waves/diffraction.rs:56-117) with the dispatch
 * of kirchhoff_auto (src/gpu/fallback.rs:19-26): the plain FP64 device kernel
 * on the shared context when a device exists, else the reference's own f64
 * CPU loop (complex Kahan sums, skip below 1e-30 mm), which is what the
 * Python caller diffraction_integral_rs (crates/xrt-python/src/lib.rs
 * :321-394) always runs.  XRTK_REQUIRE_GPU=1 turns the CPU branch into
 * XRTK_ENODEV.  xrtk_auto_last_backend reports which branch ran. */
int xrtk_diffraction_auto_f64(const XrtkRaysF64 *rays, size_t n_rays,
                              const double *pix_x, const double *pix_y,
                              const double *pix_z, size_t n_pixels,
                              double *out4);

/* Same, with the screen generated on the device from a prepare_wave grid
 * (nx*nz pixels, output in prepare_wave's order). */
int xrtk_diffraction_grid_f64(xrtk_ctx *ctx, const XrtkRaysF64 *rays,
                              size_t n_rays, const XrtkWaveParams *screen,
                              double *out4, const XrtkOpts *opts);

/* prepare_wave (src/waves/prepare_wave.rs:40-68) on the host: writes
 * nx*nz pixel coordinates into px, py, pz. */
int xrtk_prepare_wave(const XrtkWaveParams *w, double *px, double *py,
                      double *pz);

/* beam_to_diffraction_rays (src/waves/prepare_wave.rs:78-111): gathers rays
 * good[0..n_good) from a Beam's SoA columns with normals n (3 per beam ray),
 * computing nl = n . (a, b, c).  es/ep columns may be NULL (then Es = 1,
 * Ep = 0, prepare_wave.rs:89-95).  The 12 output columns hold n_good each. */
int xrtk_beam_to_rays(const double *bx, const double *by, const double *bz,
                      const double *ba, const double *bb, const double *bc,
                      const double *be, const double *es_re,
                      const double *es_im, const double *ep_re,
                      const double *ep_im, size_t n_beam,
                      const double *normals3, const size_t *good,
                      size_t n_good, double *const out_cols[12]);

/* ---- Fresnel propagation ------------------------------------------------ */

/* FresnelScreen (src/waves/fresnel.rs:13-27): pixel (ix, iz) lies at
 * x = x_min + (x_max - x_min) ix / max(nx - 1, 1), likewise z; the output
 * index is iz * nx + ix (fresnel.rs:78-99). */
typedef struct XrtkFresnelScreen {
    double x_min, x_max, z_min, z_max;
    uint32_t nx, nz;
} XrtkFresnelScreen;

/* fresnel_propagate (src/waves/fresnel.rs:53-106): from source rays on the
 * z = 0 plane (x_j, z_j, complex Es_j) to the screen at distance dist,
 *   U(x', z') = sum_j Es_j exp(i k/(2 dist) ((x' - x_j)^2 + (z' - z_j)^2)),
 * k = E / CHBAR * 1e7.  out2 receives nx*nz complex values (re, im) in the
 * index order above; x_pixels (nx) and z_pixels (nz) receive the axes and may
 * be NULL.  opts->mode XRTK_MODE_F32R (default) multiplies FP32 factor
 * matrices, XRTK_MODE_F64 FP64 ones; both evaluate every phase in FP64.
 * nx or nz of 0 is XRTK_EINVAL (the reference's max(n - 1, 1) underflows
 * there); 0 rays gives zeros. */
int xrtk_fresnel_propagate(xrtk_ctx *ctx, const double *ray_x, const double *ray_z,
                           const double *es_re, const double *es_im, size_t n_rays,
                           double energy, double dist, const XrtkFresnelScreen *screen,
                           double *out2, double *x_pixels, double *z_pixels,
                           const XrtkOpts *opts);

/* ---- undulator intensity map ------------------------------------------- */

/* The parameters Undulator::build_i_map reads (src/sources/undulator.rs
 * :36-60 and SynchrotronParams, src/sources/bending_magnet.rs:38-53):
 * gamma = E[GeV] 1e3 / 0.510998928, gamma2 = gamma * gamma. */
typedef struct XrtkUndulator {
    double gamma, gamma2;
    double beam_current;   /* [A]                                          */
    double kx, ky;         /* deflection parameters of B_x and B_y          */
    double period_mm;
    double phase_deg;      /* K_x - K_y phase                              */
    uint32_t n_periods;
    uint32_t n_steps;      /* integration steps per period (64 in build_i_map) */
} XrtkUndulator;

/* Undulator::build_i_map (src/sources/undulator.rs:120-244) on the device,
 * FP64: intensity[n], amp_s and amp_p as n complex values (re, im) each.
 * build_i_map_auto (:251-266) sends maps of 1024 points or more here. */
int xrtk_undulator_i_map(xrtk_ctx *ctx, const XrtkUndulator *und, const double *energies,
                         const double *thetas, const double *psis, size_t n,
                         double *intensity, double *amp_s2, double *amp_p2);

/* Byte-identical to GpuObsPoint (src/gpu/undulator.rs:14-22), 16 bytes. */
typedef struct XrtkObsPoint {
    float energy, theta, psi, _pad;
} XrtkObsPoint;

/* undulator_gpu (src/gpu/undulator.rs:66-91,93-275): out5 receives, per
 * point, ax_re, ax_im, az_re, az_im, intensity (the fields of
 * GpuUndulatorResult, undulator.rs:24-36), computed in FP64 from the f32
 * inputs. */
int xrtk_undulator_gpu_f32(xrtk_ctx *ctx, const XrtkObsPoint *obs, size_t n, double kx,
                           double ky, double period_mm, size_t n_periods, double gamma,
                           double beam_current, uint32_t n_steps, double phase_deg,
                           double *out5);

/* ---- resident ray sets and batched solves ------------------------------ */

typedef uint64_t xrtk_rayset;

/* Copies a ray set to HBM once and computes its FP64 centroid there.
 * path_up (may be NULL) is an upstream optical path per ray [mm]: each solve
 * multiplies Es and Ep by exp(i k path_up) in FP64 on the device, the term
 * beam_to_diffraction_rays leaves to the caller (prepare_wave.rs:78-111). */
int xrtk_rayset_upload(xrtk_ctx *ctx, const XrtkRaysF64 *rays,
                       const double *path_up, size_t n_rays,
                       xrtk_rayset *out);
/* beam_to_diffraction_rays (src/waves/prepare_wave.rs:78-111) on the
 * device, straight into a resident ray set: the Beam's SoA columns (x, y, z,
 * direction a, b, c, energy e, optional amplitudes es / ep: all four or
 * NULL, then Es = 1, Ep = 0 as prepare_wave.rs:89-95), its normals (3 per
 * beam ray) and the good indices go up once, and a gather kernel writes ray
 * k from beam ray good[k] with nl = n . (a, b, c): the same records as
 * xrtk_beam_to_rays followed by xrtk_rayset_upload.  path (may be NULL) is
 * the beam's optical path per beam ray [mm], folded into the phase as
 * xrtk_rayset_upload's path_up (the term the reference leaves to the
 * caller). */
int xrtk_rayset_from_beam(xrtk_ctx *ctx, const double *bx, const double *by,
                          const double *bz, const double *ba, const double *bb,
                          const double *bc, const double *be, const double *es_re,
                          const double *es_im, const double *ep_re,
                          const double *ep_im, const double *path, size_t n_beam,
                          const double *normals3, const size_t *good,
                          size_t n_good, xrtk_rayset *out);
int xrtk_rayset_release(xrtk_ctx *ctx, xrtk_rayset rs);
/* Bytes of HBM the ray set holds. */
int xrtk_rayset_bytes(xrtk_ctx *ctx, xrtk_rayset rs, size_t *bytes);

/* One diffraction instance of a batch: a resident ray set at one energy
 * (overrides the rays' own energy column when energy > 0) on a
 * prepare_wave screen.  out4 receives nx*nz x 4 f64. */
typedef struct XrtkJob {
    xrtk_rayset rays;
    double energy;
    XrtkWaveParams screen;
    double *out4;
} XrtkJob;

/* Runs every job of the batch in one persistent launch of the pair kernel
 * (captured with its prep and merge kernels in one hipGraph, replayed while
 * the batch shape repeats).  Each job's result equals its standalone
 * xrtk_solve_jobs() result bit for bit. */
int xrtk_solve_jobs(xrtk_ctx *ctx, const XrtkJob *jobs, size_t n_jobs,
                    const XrtkOpts *opts);

/* Device time in ms of the last pair-kernel launch of this context (HIP
 * events around the launch on the context's stream) and of the whole last
 * solve (prep + pair + merge, excluding host copies). */
int xrtk_last_timing(xrtk_ctx *ctx, double *pair_ms, double *solve_ms);

/* The plan the last solve used. */
typedef struct XrtkPlan {
    int32_t mode, pixels_per_lane, waves_per_group, tile_rays;
    int32_t split, ray_src, persistent, n_items;
    int32_t n_fp64;     /* instances the bound routed to the FP64 kernel    */
    int32_t backend;    /* 1 = device, 0 = CPU fallback                     */
    int32_t inst_flags; /* variant flags the device planner chose for the
                           first instance (XRTK_VAR_*)                      */
    int32_t n_mfma;     /* instances on the matrix cores (XRTK_VAR_MFMA)    */
} XrtkPlan;
/* Variant flags reported in XrtkPlan.inst_flags. */
#define XRTK_VAR_CHECK     1   /* FP64: a pair may fall below the skip threshold */
#define XRTK_VAR_TILE      2   /* tile-referenced reduced phase                  */
#define XRTK_VAR_PATHUP    4   /* upstream path phase                            */
#define XRTK_VAR_UNIFORM_N 8   /* every ray has the same normal                  */
#define XRTK_VAR_NOFRACT   16  /* reduced phase inside v_sin's direct domain     */
#define XRTK_VAR_LIN2      32  /* o is zero along one axis                       */
#define XRTK_VAR_PAX       64  /* paraxial 1/path and 1/(path + R)               */
#define XRTK_VAR_SEPAMP    128 /* amplitude = ray factor x pixel factor          */
#define XRTK_VAR_LINE      256 /* phase along a lane's pixel line by rotation    */
#define XRTK_VAR_LINEDEV   1024 /* LINE with per-pixel offsets from the line     */
#define XRTK_VAR_EPZERO    2048 /* every ray has Ep = 0: no Ep sums are formed    */
#define XRTK_VAR_MFMA      4096 /* 64 x 64-pixel blocks on the matrix cores       */
int xrtk_last_plan(xrtk_ctx *ctx, XrtkPlan *plan);
/* Backend used by the last xrtk_kirchhoff_auto_f32 call of this process:
 * 1 = device, 0 = CPU fallback, -1 = none yet. */
int xrtk_auto_last_backend(void);

/* ---- autotuner ---------------------------------------------------------- */

/* Times the candidate plans for an (n_rays, n_pixels) bucket on the device
 * and records the fastest in the context's table (and in the JSON cache file
 * named by XRTK_TUNE_CACHE, when set).  mode is XRTK_MODE_*. */
int xrtk_autotune(xrtk_ctx *ctx, size_t n_rays, size_t n_pixels, int mode,
                  XrtkOpts *best, double *best_ms);

/* Times the candidate launch shapes (pixels per lane, waves per group,
 * persistent or not) on the given batch, each instance cut to its first
 * 65536 rays, and records the fastest as this context's configuration for
 * batched solves: xrtk_solve_jobs and servers created afterwards use it
 * (explicit XrtkOpts fields still override).  best_ms is the device time of
 * the winner on the cut batch.  No reference counterpart: the reference
 * builds one fixed 256-thread WGSL pipeline (kirchhoff.rs:151-218). */
int xrtk_autotune_jobs(xrtk_ctx *ctx, const XrtkJob *jobs, size_t n_jobs, int mode,
                       XrtkOpts *best, double *best_ms);

/* ---- serving ------------------------------------------------------------ */

typedef struct xrtk_server xrtk_server;
typedef uint64_t xrtk_ticket;

typedef struct XrtkServerCfg {
    size_t hbm_budget_bytes;   /* admission budget (resident rays excluded) */
    int32_t max_active;        /* concurrent requests on the device         */
    double quantum_ms;         /* target device time of one batched launch  */
    int32_t mode;              /* XRTK_MODE_*                               */
    int32_t policy;            /* XRTK_SCHED_*                              */
    const XrtkOpts *opts;      /* F32R launch shape (the FP64 kernel takes
                                  it where it has that shape); NULL or 0
                                  fields: the context's batch configuration
                                  (xrtk_autotune_jobs)                      */
} XrtkServerCfg;

/* Which requests fill a launch, and which queued request is admitted next
 * (the one with the fewest planned pairs, or the oldest).  Every active
 * request with work left gets at least one work item in every launch under
 * either policy; the results do not depend on it. */
#define XRTK_SCHED_SHORTEST 0  /* fewest remaining pairs first (default)    */
#define XRTK_SCHED_ARRIVAL  1  /* oldest request first                      */

typedef struct XrtkServerStats {
    uint64_t submitted, admitted, completed, rejected;
    uint64_t launches;          /* batched pair-kernel launches             */
    uint64_t max_concurrent;    /* most requests sharing one launch          */
    double pairs_done;          /* ray-pixel pairs evaluated                */
    double busy_ms;             /* device time inside pair kernels          */
    size_t hbm_in_use, hbm_peak;
} XrtkServerStats;

int xrtk_server_create(xrtk_ctx *ctx, const XrtkServerCfg *cfg,
                       xrtk_server **out);
/* Queues one solve (non-blocking).  out4 must stay valid until the ticket
 * completes.  Returns XRTK_ENOMEM at once when the request alone exceeds the
 * budget.  The request holds its ray set until the ticket is collected by
 * xrtk_server_wait (xrtk_rayset_release returns XRTK_EAGAIN meanwhile). */
int xrtk_server_submit(xrtk_server *s, const XrtkJob *job,
                       xrtk_ticket *ticket);
/* XRTK_OK when done, XRTK_EAGAIN while pending. */
int xrtk_server_poll(xrtk_server *s, xrtk_ticket t);
/* Blocks until done or timeout_ms elapses (negative: no limit).  On
 * success fills the request's timestamps (seconds, steady clock):
 * t[0] = submit, t[1] = admit, t[2] = first launch, t[3] = done. */
int xrtk_server_wait(xrtk_server *s, xrtk_ticket t, double timeout_ms,
                     double t_out[4]);
int xrtk_server_stats(xrtk_server *s, XrtkServerStats *st);
/* Completes the queued work, then stops the scheduler thread. */
void xrtk_server_destroy(xrtk_server *s);

/* Steady-clock seconds, the clock of xrtk_server_wait's timestamps. */
double xrtk_now(void);

/* Page-locked host memory for result buffers that a caller reuses across
 * solves (device-to-host copies into it run at the link's rate instead of
 * through a staging copy).  No reference counterpart: the reference returns
 * a fresh Vec per call (src/gpu/kirchhoff.rs:276-297).  The memory belongs
 * to the process, not to a context; free it with xrtk_host_free. */
int xrtk_host_alloc(size_t bytes, void **out);
void xrtk_host_free(void *p);

/* ---- testing ------------------------------------------------------------ */

/* Makes the next n pair-kernel launches of this process fail before anything
 * is enqueued, as a failed HIP launch would: the calls that own them return
 * XRTK_ELAUNCH (served requests in such a launch complete with XRTK_ELAUNCH)
 * and the context stays usable.  n = 0 clears it.  No reference counterpart:
 * it exercises the error path that the reference reaches by panicking on a
 * failed buffer map (src/gpu/kirchhoff.rs:268-271). */
void xrtk_debug_fail_launches(uint32_t n);

#ifdef __cplusplus
}
#endif
#endif /* XRTK_H */
