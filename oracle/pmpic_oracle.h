/*
 * pmpic_oracle.h -- CPU FP64 restatement of the PMPIC per-timestep parcel-in-cell path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The product
 * (pmpic_amd/, libpmpic_hip.so) never links, loads or calls it.
 *
 * Every routine restates one reference routine in plain sequential C (the order of the
 * reference run with OMP_NUM_THREADS=1 on one MPI rank), citing the reference file:line it
 * follows.  Grids use the reference's own Fortran layout data(z,y,x): z fastest, x/y halo of 2,
 * no z halo, 1-based indices in the macros of pmpic_oracle.c.
 */
#ifndef PMPIC_ORACLE_H
#define PMPIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* parcel field ids: the 17 properties of parcel_type (model_core/src/parcels/parcels.F90:19-26),
 * in writebinary record order (components/writebinary/src/writebinary.F90:182-204), followed by
 * the 12 RK arrays (parcels.F90:30-34) and the per-parcel total buoyancy btot. */
enum {
  ORC_X = 0, ORC_Y, ORC_Z, ORC_P, ORC_Q, ORC_R, ORC_DXDT, ORC_DYDT, ORC_DZDT, ORC_DPDT,
  ORC_DQDT, ORC_DRDT, ORC_H, ORC_B, ORC_VOL, ORC_STRETCH, ORC_TAG,
  ORC_XO, ORC_YO, ORC_ZO, ORC_XF, ORC_YF, ORC_ZF, ORC_PO, ORC_QO, ORC_RO, ORC_PF, ORC_QF,
  ORC_RF, ORC_BTOT, ORC_NFIELDS
};

/* halo'd physical grids ((n_z+1) x (n_y+4) x (n_x+4), z fastest) */
enum {
  ORC_GU = 0, ORC_GV, ORC_GW, ORC_GP, ORC_GQ, ORC_GR, ORC_GB, ORC_GHG, ORC_GHGLIQ, ORC_GVOL,
  ORC_GDP, ORC_GDQ, ORC_GDR, ORC_NGRIDS
};

/* spectral arrays ((n_z+1) x (n_y/2+1)*2 x (n_x/2+1)*2, z fastest) */
enum { ORC_SU = 0, ORC_SV, ORC_SW, ORC_NSPEC };

orc_t *orc_create(int nx, int ny, int nz_cells, double xmin, double xmax, double ymin,
                  double ymax, double zmin, double zmax, int64_t maxparcels, double dtm,
                  double dtmax, double timestep_prefactor);
void orc_destroy(orc_t *o);
/* OpenMP threads for later calls (0 or less: unchanged).  Results do not depend on the count:
 * every threaded loop either touches disjoint data or adds into each grid node in the
 * sequential parcel order. */
void orc_set_threads(int nthreads);
int orc_max_threads(void);  /* threads a parallel region of the next call uses */

double *orc_field(orc_t *o, int id);     /* capacity-sized parcel array */
double *orc_grid(orc_t *o, int id);      /* halo'd grid */
double *orc_spec(orc_t *o, int id);      /* spectral u_s, v_s, w_s */
int64_t orc_get_n(const orc_t *o);
void orc_set_n(orc_t *o, int64_t n);
double orc_get_dtm(const orc_t *o);
void orc_set_dtm(orc_t *o, double dtm);
double orc_get_time(const orc_t *o);
void orc_geometry(const orc_t *o, double out[16]);
int32_t *orc_cache_ijk(orc_t *o, int axis); /* is/js/ks (1-based, halo'd) */
double *orc_cache_del(orc_t *o, int axis);  /* delxs/delys/delzs */

/* parcel_interpolation_mod */
int orc_cache_weights(orc_t *o);
int orc_par2grid(orc_t *o, const double *var, double *grid);
int orc_grid2par(orc_t *o, double *grid, double *var);
void orc_fold(orc_t *o, double *grid);   /* par2grid_haloswap on one rank */
void orc_fill(orc_t *o, double *grid);   /* grid2par_haloswap on one rank */

/* pencil_fft_mod / fftops_mod, acting on halo'd physical grids and spectral arrays */
void orc_fft_forward(orc_t *o, const double *grid, double *spec);
void orc_fft_backward(orc_t *o, const double *spec, double *grid);
void orc_diffx(orc_t *o, const double *in, double *out);
void orc_diffy(orc_t *o, const double *in, double *out);
int orc_diffz(orc_t *o, const double *f, double *dfdz, const double *bot, const double *top);
int orc_laplinv(orc_t *o, double *f, int df_zero_on_boundary);
void orc_filter(orc_t *o, const double *in, double *out);
double orc_wavenumber(orc_t *o, int which, int jj, int ii); /* 0 kx, 1 ky, 2 k2, 3 filter */
int orc_tridiagonal(int n, const double *a, const double *b, const double *c, const double *d,
                    double *f);
/* 1-D transforms in the ffte_mod wrapper convention (ffte_mod.F90:66-124) */
void orc_r2c(int n, const double *in, double *out_interleaved);
void orc_c2r(int n, const double *in_interleaved, double *out);

/* component timestep callbacks (each keeps its own iteration counter like the reference) */
int orc_vort2vel(orc_t *o);
int orc_vorticity_tendency(orc_t *o);
int orc_rk4_integrator(orc_t *o);
int orc_parcel_splitting(orc_t *o);
int orc_parcel_mixing(orc_t *o);
int orc_parcel_haloswap(orc_t *o);
int orc_step(orc_t *o);
/* the grids writebinary dumps (p, q, r, b, hg, hgliq) from the current parcels */
int orc_writebinary_grids(orc_t *o);  /* one timestep: 4 passes of the integrator group, then time += dtm */
int64_t orc_last_counts(const orc_t *o, int which); /* 0 split, 1 removed, 2 added */

#ifdef __cplusplus
}
#endif
#endif
