/*
 * pmpic_hip.h -- C ABI of libpmpic_hip.so, the MI355X (gfx950) implementation of PMPIC's
 * per-timestep parcel-in-cell path.
 *
 * Plain C types only (pointers, sizes, int status); no HIP or torch types cross this boundary,
 * so the Fortran host binds it with iso_c_binding (pmpic_amd/fortran/pmpic_hip_mod.F90) and
 * Python with ctypes (pmpic_amd/lib.py).
 *
 * State lives on the GPU between calls.  Host arrays cross the boundary only in the set/get
 * calls and the standalone operator calls, always in the reference's layouts:
 *   - parcel arrays: 1-D FP64, n entries, the 17 parcel_type properties in writebinary order
 *     x y z p q r dxdt dydt dzdt dpdt dqdt drdt h b vol stretch tag
 *     (model_core/src/parcels/parcels.F90:19-26, components/writebinary/src/writebinary.F90:182-204);
 *   - grids: prognostic_field_type%data(z,y,x), z fastest, (nz+1) x (ny+4) x (nx+4) with x/y
 *     halo 2 (model_core/src/grid/prognostics.F90, components/setup_grid/src/setup_grid.F90:130-142);
 *     exported halos hold the periodic copies of the interior;
 *   - spectral arrays: (nz+1) x ((ny/2+1)*2) x ((nx/2+1)*2), z fastest, (Re,Im) interleaved
 *     (model_core/src/fft/pencilfft.F90:766-899, components/vort2vel/src/vort2vel.F90:75-84).
 *
 * Every entry point returns an int status from pmpic_status.h (0 = success); the Fortran shim
 * turns a non-zero status into `error stop pmpic_status_message(status)`, the reference's own
 * failure mode.  A context is not re-entrant; call it from one host thread.
 */
#ifndef PMPIC_HIP_H
#define PMPIC_HIP_H

#include <stdint.h>

#include "pmpic_status.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmpic_ctx pmpic_ctx;

/* Grid and run configuration: the config.mcf keys the path reads (setup_grid.F90:48-58,
 * parcelsetup_base.F90:49, rk4_integrator.F90:38-39, vorticity_tendency.F90:96), reals already
 * converted the way the reference's parser converts them (conversions.F90:111-116, 445-453). */
typedef struct pmpic_config {
  int32_t nx, ny, nz;                 /* cells per direction ("nx", "ny", "nz") */
  double xmin, xmax, ymin, ymax, zmin, zmax;
  int64_t max_parcels;                /* "max_parcels" */
  double dtm, dtmax;                  /* "dtm", "dtmax" */
  double timestep_prefactor;          /* "timestep_prefactor" */
  int32_t precision;                  /* 0 = plain FP64 (the reference's arithmetic, held to the
                                         oracle); 1 = FP32-bulk (cell index + FP32 offsets, FP32
                                         parcel fields, interpolation and RK4, FP32 semi-spectral
                                         transforms and spectral arrays; FP64 mean flow, boundary
                                         planes, column solves, timestep and conservation sums) */
  int32_t splitting_enabled;          /* "parcel_splitting_enabled" */
  int32_t mixing_enabled;             /* "parcel_mixing_enabled" */
  int32_t device;                     /* HIP device ordinal */
} pmpic_config;

/* Context lifecycle.  Replaces the initialisation callbacks of setup_grid
 * (setup_grid.F90:30-43), parcelsetup_base (parcelsetup_base.F90:36-93),
 * rk4_integrator (rk4_integrator.F90:33-78), vort2vel (vort2vel.F90:46-90),
 * vorticity_tendency (vorticity_tendency.F90:52-99) and parcel_mixing (parcel_mixing.F90:46-64),
 * and their finalisation callbacks. */
int pmpic_create(const pmpic_config *cfg, pmpic_ctx **out);
int pmpic_destroy(pmpic_ctx *ctx);
const char *pmpic_status_message(int status);

/* Parcel state transfer (parcelsetup / readfromfile_parcelsetup.F90:257-316 into the device;
 * writebinary.F90:149-209 out of it).  fields[f] may be NULL in get (field skipped). */
int pmpic_set_parcels(pmpic_ctx *ctx, int64_t n, const double *const *fields17);
int pmpic_get_parcels(pmpic_ctx *ctx, int64_t *n, double *const *fields17);
/* the 12 RK arrays xo yo zo xf yf zf po qo ro pf qf rf (parcels.F90:30-34) */
int pmpic_set_rk_arrays(pmpic_ctx *ctx, const double *const *fields12);
int pmpic_get_rk_arrays(pmpic_ctx *ctx, double *const *fields12);
/* model time, current dtm and parcel count (state.F90:44-47, parcels.F90:11-14) */
int pmpic_get_state(pmpic_ctx *ctx, int64_t *n, double *time, double *dtm);
int pmpic_set_time(pmpic_ctx *ctx, double time, double dtm);

/* Component timestep callbacks, called in the integrator group order of global_config:43 once
 * per RK stage (timestepper.F90:89-106).  Each keeps its own iteration counter like the
 * reference component.  Parcel tendencies (dxdt..drdt) are interpolated to the parcels inside
 * pmpic_rk4_integrator, fused with the stage update; the grids are complete after the
 * component that produces them. */
int pmpic_vort2vel(pmpic_ctx *ctx);             /* vort2vel.F90:98-400 */
int pmpic_vorticity_tendency(pmpic_ctx *ctx);   /* vorticity_tendency.F90:107-334 */
int pmpic_rk4_integrator(pmpic_ctx *ctx);       /* rk4_integrator.F90:95-319 + haloswap.F90:109-289 */
int pmpic_parcel_splitting(pmpic_ctx *ctx);     /* parcel_splitting.F90:47-168 */
int pmpic_parcel_mixing(pmpic_ctx *ctx);        /* parcel_mixing.F90:68-431 */
/* nsteps full timesteps: the integrator group rksteps times, then time += dtm
 * (monc.F90:216-228, timestepper.F90:89-106) */
int pmpic_step(pmpic_ctx *ctx, int32_t nsteps);
/* Ensemble batching: nsteps timesteps of `count` independent contexts (members of any grid size
 * and either precision) at once, one host thread per member, each on its own HIP stream, so
 * the members' kernels overlap on the GPU.  Returns the first non-zero status in member order.
 * The contexts must be distinct.  No reference counterpart: the reference runs one model per
 * MPI job (monc.F90:200-235). */
int pmpic_step_concurrent(pmpic_ctx *const *ctxs, int32_t count, int32_t nsteps);

/* parcel_interpolation_mod entry points used by other components
 * (parcel_interpolation.F90:186-262, 267-332, 336-481); host arrays */
int pmpic_cache_parcel_interp_weights(pmpic_ctx *ctx, int32_t *is, int32_t *js, int32_t *ks,
                                      double *delxs, double *delys, double *delzs);
int pmpic_par2grid(pmpic_ctx *ctx, const double *var, double *grid);
int pmpic_grid2par(pmpic_ctx *ctx, const double *grid, double *var);

/* pmpic_par2grid on device arrays, allocating nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  var: the context's current parcel count of FP64
 * values in device memory of the context's device; grid: (nz+1) x (ny+4) x (nx+4) FP64 in device
 * memory, written in pmpic_par2grid's layout with the same values, bit for bit.  The working
 * memory is one static device array of PMPIC_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends, so the sizes are capped: at most PMPIC_NOALLOC_MAX_PARCELS parcels and
 * PMPIC_NOALLOC_MAX_CELLS cells, else PMPIC_ERR_ARG.  Calls that overlap on one GPU (from two
 * threads or two contexts' streams) race on that array.  Plain FP64 form only. */
#define PMPIC_NOALLOC_MAX_PARCELS (1 << 20)
#define PMPIC_NOALLOC_MAX_CELLS (1 << 15)
#define PMPIC_NOALLOC_SCRATCH_BYTES 17826304
int pmpic_par2grid_noalloc(pmpic_ctx *ctx, const double *var, double *grid);
/* parcel_haloswap (haloswap.F90:109-289) on the current device parcels */
int pmpic_parcel_haloswap(pmpic_ctx *ctx);

/* grids for writebinary (writebinary.F90:287-296): id 0 u,1 v,2 w,3 p,4 q,5 r,6 b,7 hg,
 * 8 hgliq, 9 vol, 10 dp, 11 dq, 12 dr; spectral id 0 u_s, 1 v_s, 2 w_s.  A grid holds what the
 * reference's state grid of that name holds at this point: p, q, r and b (total buoyancy) from
 * vort2vel / vorticity_tendency, the kept vol, b, hg, p, q, r after parcel_mixing, and hgliq (0
 * until then) plus writebinary's b, hg, p, q, r after pmpic_writebinary_grids. */
int pmpic_get_grid(pmpic_ctx *ctx, int32_t id, double *grid);
/* the grid computation of write_grids_to_file (writebinary.F90:217-233): p, q, r, b and hg by
 * par2grid of the current parcels' p, q, r, b, h; hgliq = max(0, hg - exp(-z)); b += 12.5 hgliq */
int pmpic_writebinary_grids(pmpic_ctx *ctx);
int pmpic_get_spectral(pmpic_ctx *ctx, int32_t id, double *spec);

/* standalone semi-spectral operators on host arrays (pencilfft.F90:151-212, fft_ops.F90) */
int pmpic_fft_forward(pmpic_ctx *ctx, const double *grid, double *spec);
int pmpic_fft_backward(pmpic_ctx *ctx, const double *spec, double *grid);
int pmpic_diffx(pmpic_ctx *ctx, const double *in, double *out);
int pmpic_diffy(pmpic_ctx *ctx, const double *in, double *out);
int pmpic_spectral_filter(pmpic_ctx *ctx, const double *in, double *out);
int pmpic_diffz(pmpic_ctx *ctx, const double *in, double *out, const double *bot,
                const double *top);
int pmpic_laplinv(pmpic_ctx *ctx, double *f, int32_t df_zero_on_boundary);

/* FP64 volume integrals sum(vol), sum(vol*b), sum(vol*h), sum(vol*p), sum(vol*q), sum(vol*r)
 * over the current parcels, compensated (Neumaier) and in a fixed order; replaces the
 * commented-out conservation check of parcel_mixing.F90:99-100, 412-422 */
int pmpic_conservation_sums(pmpic_ctx *ctx, double out6[6]);

/* last step's event counts: split parcels, removed parcels, added parcels
 * (parcel_splitting.F90:149-155, parcel_mixing.F90:408-410) */
int pmpic_last_counts(pmpic_ctx *ctx, int64_t out3[3]);

/* HIP-event timing on the context's stream: enable, then read the accumulated milliseconds and
 * counts of timer `which`: 0 = per-cell par2grid accumulation kernel, 1 = fused grid2par + RK
 * stage kernel, 2..6 = the component callbacks vort2vel, vorticity_tendency, rk4_integrator,
 * parcel_splitting (active stages), parcel_mixing (active stages).  enable: 0 = off, 1 = every
 * timer, 2 = timers 0 and 1 only.  Timers 0 and 1 are taken by the kernel dispatches themselves;
 * timers 2..6 record events around the component on the stream.  pmpic_timing_name gives a
 * timer's region name (the reference's timer names for the components, timer.F90:247-406),
 * NULL past the last. */
int pmpic_timing_enable(pmpic_ctx *ctx, int32_t enable);
/* HBM bandwidth probe (no reference counterpart; a measurement aid): a STREAM-style triad
 * a = b + s*c over three device arrays of bytes_per_array bytes on `device`, the best of `reps`
 * timed launches; gbps counts 3 * bytes_per_array per launch.  Allocates and frees its arrays. */
int pmpic_hbm_triad(int32_t device, int64_t bytes_per_array, int32_t reps, double *best_ms,
                    double *gbps);
int pmpic_timing_read(pmpic_ctx *ctx, int32_t which, double *total_ms, int64_t *launches);
/* the same timer's statistics for the reference's timer summary (timer.F90:247-303): total,
 * number of intervals, longest interval and the sum of squared intervals (ms, ms^2) */
int pmpic_timing_stats(pmpic_ctx *ctx, int32_t which, double *total_ms, int64_t *calls,
                       double *max_ms, double *sumsq_ms2);
const char *pmpic_timing_name(int32_t which);
/* HIP-graph use of pmpic_step (no reference counterpart; a diagnostic).  With PMPIC_GRAPH=1 in
 * the environment at pmpic_create, an FP32-bulk context with timing off launches RK stages 1-3
 * of each step after its first as one captured graph, captured again when the parcel count
 * leaves the range a capture covers (the results are the same bits as eager launches; off by
 * default: it measured no faster).  `captures` graphs captured, `launches` graph launches. */
int pmpic_graph_stats(pmpic_ctx *ctx, int64_t *captures, int64_t *launches);
/* Launch-configuration autotuning of an FP32-bulk context (no reference counterpart; for the
 * batched ensemble, whose members differ in grid size).  A configuration is four numbers:
 * config[0] the binning's slot capacity (16 or 32; 0 where the grid has no slot table),
 * config[1] the binning fused into RK stages 1-3 (1) or a pass of its own (0), config[2] RK
 * stages 1-3 launched as one HIP graph (1, only with timing off) or eagerly (0), config[3] the
 * RK stage's workgroup size (128 or 256).  Every configuration gives the same bits.  The first
 * context of a grid size (nx, ny, nz) and occupancy class (the slot capacity the mean rule
 * picks) runs each configuration twice (the list forward, then backward, so that a drift of
 * the step time with the model state cancels) for one settling step and `steps` timed steps,
 * all of them ordinary model steps, and keeps the fastest, for itself and for every later context of that
 * size and class in the process (`measured` 0: applied from that table without stepping).
 * Tune members one after another, before stepping them concurrently.  Errors: PMPIC_ERR_ARG
 * for an FP64 context, no parcels or steps < 1; a step's error as pmpic_step returns it. */
int pmpic_autotune(pmpic_ctx *ctx, int32_t steps, int32_t *config, float *ms_per_step,
                   int32_t *measured);
/* The launch configuration an FP32-bulk context uses now, as pmpic_autotune reports it. */
int pmpic_launch_config(pmpic_ctx *ctx, int32_t *config);
/* Sets an FP32-bulk context's launch configuration (for tuners that time members stepped
 * together); config[0] 0 leaves the slot capacity to the mean-occupancy rule.  PMPIC_ERR_ARG
 * for a value outside those pmpic_autotune lists or a capacity the slot table cannot hold. */
int pmpic_set_launch_config(pmpic_ctx *ctx, const int32_t *config);

#ifdef __cplusplus
}
#endif
#endif
