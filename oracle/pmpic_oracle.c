/*
 * pmpic_oracle.c -- CPU FP64 restatement of the PMPIC per-timestep parcel-in-cell path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing shipped or measured as the product):
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * It follows the reference as run on one MPI rank with OMP_NUM_THREADS=1: every loop runs in
 * the reference's sequential order, every expression keeps the reference's operand order and
 * the single-precision literals the Fortran source carries (1./6., 1./3., 1.E-15, ...).  Build
 * with -ffp-contract=off so no multiply-add is fused.  Grid arrays use the reference layout
 * data(z,y,x): z fastest, x/y halo of 2 cells, no z halo (model_core/src/grid/prognostics.F90,
 * components/setup_grid/src/setup_grid.F90:116-142).
 *
 * Parity pins (tests/): FFT vs numpy.fft and vs the reference's own FFTE built in oracle/_ref;
 * tridiagonal vs dense solves; diffz/laplinv 4th-order convergence on manufactured columns;
 * grid2par/par2grid linear-field and constant-field known answers (grid2partest.F90:119-265);
 * vort2vel invariants (divergence-free vorticity, curl(u) = filtered vorticity); RK4
 * exactness under constant velocity; haloswap wrap/reflect closed forms and the one-rank
 * slot permutation against an independent model; split closed form; mixing conservation.
 */
#include "pmpic_oracle.h"
#include "../include/pmpic_status.h"

#include <complex.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

struct orc {
  /* geometry: setup_grid.F90:45-104, parcel_interpolation.F90:51-157 */
  int ncx, ncy, ncz;          /* cells in x, y, z (config nx, ny, nz) */
  int nxh, nyh, nzn;          /* halo'd x/y sizes (n+4) and z nodes (nz+1) */
  int sx, sy;                 /* spectral sizes (n/2+1)*2 */
  double xbot, ybot, zbot, xtop_cfg, ytop_cfg, ztop_cfg;
  double dx, dy, dz;
  double xmin_h, ymin_h, zmin_h, zmax_h;
  double *xc, *yc, *zc;       /* x_coords, y_coords, z_coords */
  double hs_xmin, hs_xmax, hs_ymin, hs_ymax; /* haloswap.F90:314-318 */
  double xshift_e, xshift_w, yshift_n, yshift_s; /* haloswap.F90:529-552 */
  /* parcels */
  int64_t maxp, np;
  double *f[ORC_NFIELDS];
  int32_t *is, *js, *ks;
  double *delxs, *delys, *delzs;
  int *index;   /* haloswap.F90:21 */
  /* x-slabs of halo'd node columns for the threaded scatters: slab t owns node x indices
   * [slab_lo[t], slab_lo[t+1]) and lists, in ascending parcel order, the parcels whose cell
   * column is slab_lo[t]-1 .. slab_lo[t+1]-1 (those that deposit into its nodes) */
  int nslab;
  int *slab_lo;
  int64_t *slab_n, *slab_cap;
  int64_t **slab_list;
  int *keep;    /* parcel_mixing.F90:29 */
  /* grids */
  size_t gsz, ssz;
  double *g[ORC_NGRIDS];
  double *s[ORC_NSPEC];
  double *sa, *sb, *sc, *sd, *se, *sf; /* vort2vel scratch a..f */
  double *as, *bs;                     /* vorticity_tendency scratch */
  double *ap, *dudx, *dwdx, *dvdy, *dwdy;
  double *wgt, *dat;                   /* par2grid accumulators */
  double *vres, *bres, *hres, *pres, *qres, *rres;
  int *npercell;
  /* spectral operators: fft_ops.F90:48-168 */
  double *kx, *ky, *k2, *filt; /* per (jj,ii) column, sy*sx */
  /* time stepping */
  double time, dtm, dtmax, prefactor;
  int rksteps;
  int it_v2v, it_vt, rkstep, it_split, it_mix;
  int64_t last_split, last_removed, last_added;
};

#define GI(o, k, j, i) ((((size_t)(i)-1) * (size_t)(o)->nyh + ((size_t)(j)-1)) * (size_t)(o)->nzn + ((size_t)(k)-1))
#define SI(o, k, j, i) ((((size_t)(i)-1) * (size_t)(o)->sy + ((size_t)(j)-1)) * (size_t)(o)->nzn + ((size_t)(k)-1))
#define CI(o, j, i) (((size_t)(i)-1) * (size_t)(o)->sy + ((size_t)(j)-1))
#define NPC(o, k, j, i) ((((size_t)(i)-1) * (size_t)((o)->nyh - 1) + ((size_t)(j)-1)) * (size_t)((o)->nzn - 1) + ((size_t)(k)-1))

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static const double complex *twiddles(int n);
static double complex *make_twiddles(int n);

/* ffte_mod.F90:137-161 */
static int check_factors(int n) {
  int m = n;
  if (m < 1) return 0;
  while (m % 5 == 0) m /= 5;
  while (m % 3 == 0) m /= 3;
  while (m % 2 == 0) m /= 2;
  return m == 1;
}

orc_t *orc_create(int nx, int ny, int nz_cells, double xmin, double xmax, double ymin,
                  double ymax, double zmin, double zmax, int64_t maxparcels, double dtm,
                  double dtmax, double timestep_prefactor) {
  if (nx < 2 || ny < 2 || nz_cells < 2 || maxparcels < 1) return NULL;
  if (!check_factors(nx) || !check_factors(ny)) return NULL;
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->ncx = nx; o->ncy = ny; o->ncz = nz_cells;
  o->nxh = nx + 4; o->nyh = ny + 4; o->nzn = nz_cells + 1;
  o->sx = (nx / 2 + 1) * 2; o->sy = (ny / 2 + 1) * 2;
  o->xbot = xmin; o->ybot = ymin; o->zbot = zmin;
  o->xtop_cfg = xmax; o->ytop_cfg = ymax; o->ztop_cfg = zmax;
  /* setup_grid.F90:73-75 (nz there is the node count, so nz-1 = cells) */
  o->dx = (xmax - xmin) / nx;
  o->dy = (ymax - ymin) / ny;
  o->dz = (zmax - zmin) / nz_cells;
  /* parcel_interpolation.F90:94-115: xstart = 1-2, zstart = 1, zstop = nz nodes */
  o->xmin_h = (double)(-2) * o->dx + xmin;
  o->ymin_h = (double)(-2) * o->dy + ymin;
  o->zmin_h = (double)0 * o->dz + zmin;
  o->zmax_h = (double)(o->nzn - 1) * o->dz;
  o->xc = dalloc(o->nxh); o->yc = dalloc(o->nyh); o->zc = dalloc(o->nzn);
  for (int n = 1; n <= o->nxh; n++) o->xc[n - 1] = o->xmin_h + (double)(n - 1) * o->dx + xmin;
  for (int n = 1; n <= o->nyh; n++) o->yc[n - 1] = o->ymin_h + (double)(n - 1) * o->dy + ymin;
  for (int n = 1; n <= o->nzn; n++) o->zc[n - 1] = o->zmin_h + (double)(n - 1) * o->dz + zmin;
  /* haloswap.F90:314-318: ymax=y_coords(ny-hy+1), ymin=y_coords(hy+1) */
  o->hs_xmin = o->xc[2]; o->hs_xmax = o->xc[o->nxh - 2];
  o->hs_ymin = o->yc[2]; o->hs_ymax = o->yc[o->nyh - 2];
  /* haloswap.F90:531-552 with top = xmax - dx (setup_grid.F90:92-93) */
  {
    double xtop = xmax - o->dx, ytop = ymax - o->dy;
    o->xshift_e = xmin - xtop - o->dx;
    o->xshift_w = xtop - xmin + o->dx;
    o->yshift_n = ymin - ytop - o->dy;
    o->yshift_s = ytop - ymin + o->dy;
  }
  o->maxp = maxparcels;
  for (int f = 0; f < ORC_NFIELDS; f++) o->f[f] = dalloc((size_t)maxparcels);
  o->is = (int32_t *)calloc(maxparcels, sizeof(int32_t));
  o->js = (int32_t *)calloc(maxparcels, sizeof(int32_t));
  o->ks = (int32_t *)calloc(maxparcels, sizeof(int32_t));
  o->delxs = dalloc(maxparcels); o->delys = dalloc(maxparcels); o->delzs = dalloc(maxparcels);
  o->index = (int *)calloc(maxparcels, sizeof(int));
  o->keep = (int *)calloc(maxparcels, sizeof(int));
  o->gsz = (size_t)o->nzn * o->nyh * o->nxh;
  o->ssz = (size_t)o->nzn * o->sy * o->sx;
  for (int g = 0; g < ORC_NGRIDS; g++) o->g[g] = dalloc(o->gsz);
  for (int s = 0; s < ORC_NSPEC; s++) o->s[s] = dalloc(o->ssz);
  o->sa = dalloc(o->ssz); o->sb = dalloc(o->ssz); o->sc = dalloc(o->ssz);
  o->sd = dalloc(o->ssz); o->se = dalloc(o->ssz); o->sf = dalloc(o->ssz);
  o->as = dalloc(o->ssz); o->bs = dalloc(o->ssz);
  o->ap = dalloc(o->gsz); o->dudx = dalloc(o->gsz); o->dwdx = dalloc(o->gsz);
  o->dvdy = dalloc(o->gsz); o->dwdy = dalloc(o->gsz);
  o->wgt = dalloc(o->gsz); o->dat = dalloc(o->gsz);
  o->vres = dalloc(o->gsz); o->bres = dalloc(o->gsz); o->hres = dalloc(o->gsz);
  o->pres = dalloc(o->gsz); o->qres = dalloc(o->gsz); o->rres = dalloc(o->gsz);
  o->npercell = (int *)calloc((size_t)(o->nzn - 1) * (o->nyh - 1) * (o->nxh - 1), sizeof(int));
  /* fft_ops.F90:111-139 on one rank (x_start = y_start = 1) */
  o->kx = dalloc((size_t)o->sy * o->sx); o->ky = dalloc((size_t)o->sy * o->sx);
  o->k2 = dalloc((size_t)o->sy * o->sx); o->filt = dalloc((size_t)o->sy * o->sx);
  {
    const double twopi = 8.0 * atan(1.0);
    double lx = o->dx * (double)nx, ly = o->dy * (double)ny;
    /* size/2. is a default-real (FP32) quotient */
    double hx = (double)((float)nx / 2.0f), hy = (double)((float)ny / 2.0f);
    double ksqmax = (hx / lx) * (hx / lx) + (hy / ly) * (hy / ly);
    ksqmax = ksqmax * 2. / 9.;
    for (int i = 1; i <= o->sx; i++) {
      double xval = (double)((1 + i - 2) / 2) / lx;
      for (int j = 1; j <= o->sy; j++) {
        double yval = (double)((1 + j - 2) / 2) / ly;
        double k2 = xval * xval + yval * yval;
        size_t c = CI(o, j, i);
        o->filt[c] = (k2 > ksqmax) ? 0. : 1.;
        o->kx[c] = xval * twopi;
        o->ky[c] = yval * twopi;
        o->k2[c] = k2 * twopi * twopi;
      }
    }
  }
  o->time = 0.0;
  o->dtm = dtm; o->dtmax = dtmax; o->prefactor = timestep_prefactor;
  o->rksteps = 4;    /* rk4_integrator.F90:51 */
  o->rkstep = 1;     /* rk4_integrator.F90:69 */
  /* node-column slabs: one per thread, at least two node columns each */
  {
    int t = omp_get_max_threads();
    if (t > o->nxh / 2) t = o->nxh / 2;
    if (t < 1) t = 1;
    o->nslab = t;
    o->slab_lo = (int *)calloc((size_t)t + 1, sizeof(int));
    o->slab_n = (int64_t *)calloc((size_t)t, sizeof(int64_t));
    o->slab_cap = (int64_t *)calloc((size_t)t, sizeof(int64_t));
    o->slab_list = (int64_t **)calloc((size_t)t, sizeof(int64_t *));
    for (int s = 0; s <= t; s++) o->slab_lo[s] = 1 + (int)(((int64_t)o->nxh * s) / t);
  }
  /* the transform lengths' twiddle tables, built here so threads only read them */
  twiddles(nx);
  twiddles(ny);
  return o;
}

void orc_set_threads(int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
}

int orc_max_threads(void) { return omp_get_max_threads(); }

void orc_destroy(orc_t *o) {
  if (!o) return;
  free(o->xc); free(o->yc); free(o->zc);
  for (int f = 0; f < ORC_NFIELDS; f++) free(o->f[f]);
  free(o->is); free(o->js); free(o->ks);
  free(o->delxs); free(o->delys); free(o->delzs);
  free(o->index); free(o->keep);
  for (int g = 0; g < ORC_NGRIDS; g++) free(o->g[g]);
  for (int s = 0; s < ORC_NSPEC; s++) free(o->s[s]);
  free(o->sa); free(o->sb); free(o->sc); free(o->sd); free(o->se); free(o->sf);
  free(o->as); free(o->bs);
  free(o->ap); free(o->dudx); free(o->dwdx); free(o->dvdy); free(o->dwdy);
  free(o->wgt); free(o->dat);
  free(o->vres); free(o->bres); free(o->hres); free(o->pres); free(o->qres); free(o->rres);
  free(o->npercell);
  free(o->kx); free(o->ky); free(o->k2); free(o->filt);
  for (int s = 0; s < o->nslab; s++) free(o->slab_list[s]);
  free(o->slab_list); free(o->slab_n); free(o->slab_cap); free(o->slab_lo);
  free(o);
}

double *orc_field(orc_t *o, int id) { return (id >= 0 && id < ORC_NFIELDS) ? o->f[id] : NULL; }
double *orc_grid(orc_t *o, int id) { return (id >= 0 && id < ORC_NGRIDS) ? o->g[id] : NULL; }
double *orc_spec(orc_t *o, int id) { return (id >= 0 && id < ORC_NSPEC) ? o->s[id] : NULL; }
int64_t orc_get_n(const orc_t *o) { return o->np; }
void orc_set_n(orc_t *o, int64_t n) { o->np = n; }
double orc_get_dtm(const orc_t *o) { return o->dtm; }
void orc_set_dtm(orc_t *o, double dtm) { o->dtm = dtm; }
double orc_get_time(const orc_t *o) { return o->time; }
int32_t *orc_cache_ijk(orc_t *o, int axis) { return axis == 0 ? o->is : axis == 1 ? o->js : o->ks; }
double *orc_cache_del(orc_t *o, int axis) { return axis == 0 ? o->delxs : axis == 1 ? o->delys : o->delzs; }
int64_t orc_last_counts(const orc_t *o, int which) {
  return which == 0 ? o->last_split : which == 1 ? o->last_removed : o->last_added;
}

void orc_geometry(const orc_t *o, double out[16]) {
  out[0] = o->dx; out[1] = o->dy; out[2] = o->dz;
  out[3] = o->xmin_h; out[4] = o->ymin_h; out[5] = o->zmin_h; out[6] = o->zmax_h;
  out[7] = o->hs_xmin; out[8] = o->hs_xmax; out[9] = o->hs_ymin; out[10] = o->hs_ymax;
  out[11] = o->xshift_e; out[12] = o->xshift_w; out[13] = o->yshift_n; out[14] = o->yshift_s;
  out[15] = o->dtm;
}

/* ---------------------------------------------------------------------------------------- */
/* parcel_interpolation_mod                                                                  */
/* ---------------------------------------------------------------------------------------- */

/* the slab lists of the current cache (see struct orc): each thread scans the cached cell
 * columns and keeps, in ascending order, the parcels that deposit into its slab's nodes */
static void build_slabs(orc_t *o) {
  const int64_t np = o->np;
#pragma omp parallel num_threads(o->nslab)
  {
    for (int t = omp_get_thread_num(); t < o->nslab; t += omp_get_num_threads()) {
      const int lo = o->slab_lo[t] - 1, hi = o->slab_lo[t + 1] - 1;  /* cell columns lo..hi */
      int64_t cnt = 0;
      for (int64_t n = 0; n < np; n++) cnt += (o->is[n] >= lo && o->is[n] <= hi);
      if (cnt > o->slab_cap[t]) {
        free(o->slab_list[t]);
        o->slab_cap[t] = cnt + cnt / 8 + 16;
        o->slab_list[t] = (int64_t *)malloc(sizeof(int64_t) * (size_t)o->slab_cap[t]);
      }
      int64_t m = 0;
      for (int64_t n = 0; n < np; n++)
        if (o->is[n] >= lo && o->is[n] <= hi) o->slab_list[t][m++] = n;
      o->slab_n[t] = m;
    }
  }
}

/* cache_parcel_interp_weights, parcel_interpolation.F90:186-262.  Parcels are independent;
 * the status returned is that of the first failing parcel, as the sequential loop's. */
int orc_cache_weights(orc_t *o) {
  const double *X = o->f[ORC_X], *Y = o->f[ORC_Y], *Z = o->f[ORC_Z];
  int64_t bad = INT64_MAX;
  int badst = PMPIC_OK;
#pragma omp parallel
  {
    int64_t mybad = INT64_MAX;
    int myst = PMPIC_OK;
#pragma omp for schedule(static)
    for (int64_t n = 0; n < o->np; n++) {
      double xp = X[n], yp = Y[n], zp = Z[n];
      int i = (int)floor((xp - o->xmin_h) / o->dx) + 1;
      int j = (int)floor((yp - o->ymin_h) / o->dy) + 1;
      int k = (int)floor((zp - o->zmin_h) / o->dz) + 1;
      int st = PMPIC_OK;
      if (i < 1 || i >= o->nxh) st = PMPIC_ERR_OUT_OF_BOX_X;
      else if (j < 1 || j >= o->nyh) st = PMPIC_ERR_OUT_OF_BOX_Y;
      else if (k < 1 || k >= o->nzn) st = PMPIC_ERR_OUT_OF_BOX_Z;
      double delx = 0, dely = 0, delz = 0;
      if (st == PMPIC_OK) {
        delx = ((xp - o->xmin_h) - (double)(i - 1) * o->dx) / o->dx;
        dely = ((yp - o->ymin_h) - (double)(j - 1) * o->dy) / o->dy;
        delz = ((zp - o->zmin_h) - (double)(k - 1) * o->dz) / o->dz;
        if (delx > 1. || delx < 0) st = PMPIC_ERR_DELTA;
        else if (dely > 1. || dely < 0) st = PMPIC_ERR_DELTA;
        else if (delz > 1. || delz < 0) st = PMPIC_ERR_DELTA;
      }
      if (st != PMPIC_OK) {
        if (n < mybad) { mybad = n; myst = st; }
        continue;
      }
      o->is[n] = i; o->js[n] = j; o->ks[n] = k;
      o->delxs[n] = delx; o->delys[n] = dely; o->delzs[n] = delz;
    }
#pragma omp critical(orc_cache)
    if (mybad < bad) { bad = mybad; badst = myst; }
  }
  if (badst != PMPIC_OK) return badst;
  build_slabs(o);
  return PMPIC_OK;
}

/* par2grid_haloswap on one rank (MPIC_Haloswap.F90:188-290): the three messages are sent to
 * self in tag order x, y, corner and MPI_Probe(ANY_TAG) returns them in that order. */
void orc_fold(orc_t *o, double *A) {
  const int xi = 3, xf = o->ncx + 2, yi = 3, yf = o->ncy + 2, nz = o->nzn;
  double *right = dalloc((size_t)nz * o->ncy), *up = dalloc((size_t)nz * o->ncx), *cor = dalloc(nz);
  for (int j = yi; j <= yf; j++)
    for (int k = 1; k <= nz; k++) right[(size_t)(j - yi) * nz + k - 1] = A[GI(o, k, j, xf + 1)];
  for (int i = xi; i <= xf; i++)
    for (int k = 1; k <= nz; k++) up[(size_t)(i - xi) * nz + k - 1] = A[GI(o, k, yf + 1, i)];
  for (int k = 1; k <= nz; k++) cor[k - 1] = A[GI(o, k, yf + 1, xf + 1)];
  for (int j = yi; j <= yf; j++)
    for (int k = 1; k <= nz; k++) A[GI(o, k, j, xi)] = A[GI(o, k, j, xi)] + right[(size_t)(j - yi) * nz + k - 1];
  for (int i = xi; i <= xf; i++)
    for (int k = 1; k <= nz; k++) A[GI(o, k, yi, i)] = A[GI(o, k, yi, i)] + up[(size_t)(i - xi) * nz + k - 1];
  for (int k = 1; k <= nz; k++) A[GI(o, k, yi, xi)] = A[GI(o, k, yi, xi)] + cor[k - 1];
  free(right); free(up); free(cor);
}

/* grid2par_haloswap on one rank (MPIC_Haloswap.F90:95-184) */
void orc_fill(orc_t *o, double *A) {
  const int xi = 3, xf = o->ncx + 2, yi = 3, yf = o->ncy + 2, nz = o->nzn;
  for (int j = yi; j <= yf; j++)
    for (int k = 1; k <= nz; k++) A[GI(o, k, j, xf + 1)] = A[GI(o, k, j, xi)];
  for (int i = xi; i <= xf; i++)
    for (int k = 1; k <= nz; k++) A[GI(o, k, yf + 1, i)] = A[GI(o, k, yi, i)];
  for (int k = 1; k <= nz; k++) A[GI(o, k, yf + 1, xf + 1)] = A[GI(o, k, yi, xi)];
}

/* grid2par, parcel_interpolation.F90:267-332 */
int orc_grid2par(orc_t *o, double *grid, double *var) {
  orc_fill(o, grid);
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < o->np; n++) {
    double delx = o->delxs[n], dely = o->delys[n], delz = o->delzs[n];
    int i = o->is[n], j = o->js[n], k = o->ks[n];
    double c000 = grid[GI(o, k, j, i)];
    double c001 = grid[GI(o, k, j, i + 1)];
    double c010 = grid[GI(o, k, j + 1, i)];
    double c011 = grid[GI(o, k, j + 1, i + 1)];
    double c100 = grid[GI(o, k + 1, j, i)];
    double c101 = grid[GI(o, k + 1, j, i + 1)];
    double c110 = grid[GI(o, k + 1, j + 1, i)];
    double c111 = grid[GI(o, k + 1, j + 1, i + 1)];
    double c00 = c000 * (1 - delz) + c100 * delz;
    double c01 = c001 * (1 - delz) + c101 * delz;
    double c10 = c010 * (1 - delz) + c110 * delz;
    double c11 = c011 * (1 - delz) + c111 * delz;
    double c0 = c00 * (1 - dely) + c10 * dely;
    double c1 = c01 * (1 - dely) + c11 * dely;
    var[n] = c0 * (1 - delx) + c1 * delx;
  }
  return PMPIC_OK;
}

/* par2grid, parcel_interpolation.F90:336-481 (the z-level doubling at :458-459 is overwritten
 * by the division at :463-469, so it has no effect).  Each thread owns a slab of node columns
 * and walks its slab list in ascending parcel order, adding only into its own nodes: every
 * node receives the same terms in the same order as in the sequential loop. */
int orc_par2grid(orc_t *o, const double *var, double *grid) {
  double *W = o->wgt, *D = o->dat;
  memset(W, 0, o->gsz * sizeof(double));
  memset(D, 0, o->gsz * sizeof(double));
  const double *VOL = o->f[ORC_VOL];
#pragma omp parallel num_threads(o->nslab)
  {
    for (int t = omp_get_thread_num(); t < o->nslab; t += omp_get_num_threads()) {
      const int lo = o->slab_lo[t], hi = o->slab_lo[t + 1];  /* owned node columns [lo, hi) */
      const int64_t *list = o->slab_list[t];
      for (int64_t m = 0; m < o->slab_n[t]; m++) {
        const int64_t n = list[m];
        double delx = o->delxs[n], dely = o->delys[n], delz = o->delzs[n];
        int i = o->is[n], j = o->js[n], k = o->ks[n];
        double v = VOL[n], w = var[n];
        const int own0 = (i >= lo && i < hi), own1 = (i + 1 >= lo && i + 1 < hi);
        double w000 = (1 - delz) * (1 - dely) * (1 - delx) * v;
        if (own0) { D[GI(o, k, j, i)] += w000 * w; W[GI(o, k, j, i)] += w000; }
        double w001 = (1 - delz) * (1 - dely) * (delx)*v;
        if (own1) { D[GI(o, k, j, i + 1)] += w001 * w; W[GI(o, k, j, i + 1)] += w001; }
        double w010 = (1 - delz) * (dely) * (1 - delx) * v;
        if (own0) { D[GI(o, k, j + 1, i)] += w010 * w; W[GI(o, k, j + 1, i)] += w010; }
        double w011 = (1 - delz) * (dely) * (delx)*v;
        if (own1) { D[GI(o, k, j + 1, i + 1)] += w011 * w; W[GI(o, k, j + 1, i + 1)] += w011; }
        double w100 = (delz) * (1 - dely) * (1 - delx) * v;
        if (own0) { D[GI(o, k + 1, j, i)] += w100 * w; W[GI(o, k + 1, j, i)] += w100; }
        double w101 = (delz) * (1 - dely) * (delx)*v;
        if (own1) { D[GI(o, k + 1, j, i + 1)] += w101 * w; W[GI(o, k + 1, j, i + 1)] += w101; }
        double w110 = (delz) * (dely) * (1 - delx) * v;
        if (own0) { D[GI(o, k + 1, j + 1, i)] += w110 * w; W[GI(o, k + 1, j + 1, i)] += w110; }
        double w111 = (delz) * (dely) * (delx)*v;
        if (own1) { D[GI(o, k + 1, j + 1, i + 1)] += w111 * w; W[GI(o, k + 1, j + 1, i + 1)] += w111; }
      }
    }
  }
  orc_fold(o, W);
  orc_fold(o, D);
#pragma omp parallel for schedule(static)
  for (size_t c = 0; c < o->gsz; c++) {
    if (W[c] == 0.) W[c] = 1.;
    grid[c] = D[c] / W[c];
  }
  return PMPIC_OK;
}

/* ---------------------------------------------------------------------------------------- */
/* FFT: ffte_mod.F90 r2c/c2r over a mixed-radix DFT with FFTE's convention                    */
/* (forward sum x_j exp(-2 pi i jk/N), unnormalised; inverse = conj . forward . conj / N,      */
/* vzfft1d.f:41-69, fft235.f sign).                                                          */
/* ---------------------------------------------------------------------------------------- */

/* twiddle table W_N^e = exp(-2 pi i e/N), e = 0..N-1, cached per transform length */
typedef struct { int n; double complex *w; } tw_t;
#define TW_SLOTS 64
static tw_t tw_cache[TW_SLOTS];

/* tables are built under a lock and never freed, so a pointer handed out stays valid while
 * other threads transform */
static const double complex *twiddles(int n) {
  const double complex *found = NULL;
#pragma omp critical(orc_twiddles)
  {
    for (int c = 0; c < TW_SLOTS && !found; c++)
      if (tw_cache[c].n == n) found = tw_cache[c].w;
    if (!found) {
      int slot = -1;
      for (int c = 0; c < TW_SLOTS; c++)
        if (tw_cache[c].n == 0) { slot = c; break; }
      double complex *w = make_twiddles(n);
      if (slot >= 0) { tw_cache[slot].n = n; tw_cache[slot].w = w; }
      found = w;  /* past TW_SLOTS lengths a table is not cached (and not freed) */
    }
  }
  return found;
}

static double complex *make_twiddles(int n) {
  const double twopi = 8.0 * atan(1.0);
  double complex *w = (double complex *)malloc(sizeof(double complex) * n);
  for (int e = 0; e < n; e++) {
    if (e == 0) { w[e] = 1.0; continue; }
    double ang = -twopi * (double)e / (double)n;
    w[e] = cos(ang) + I * sin(ang);
  }
  return w;
}

/* decimation in time over the smallest factor p in {2,3,5}:
 * X[k] = sum_r W_n^{rk} DFT_{n/p}(x[r::p])[k mod n/p]; tw holds W_N, step = N/n */
static void cfft_level(int n, const double complex *in, int stride, double complex *out,
                       double complex *scratch, const double complex *tw, int step) {
  if (n == 1) { out[0] = in[0]; return; }
  int p = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : 5;
  int m = n / p;
  for (int r = 0; r < p; r++)
    cfft_level(m, in + (size_t)r * stride, stride * p, scratch + (size_t)r * m, out, tw, step * p);
  for (int k = 0; k < n; k++) {
    double complex acc = scratch[k % m];
    for (int r = 1; r < p; r++) {
      int e = (int)(((long)r * k) % n);
      acc += tw[(size_t)e * step] * scratch[(size_t)r * m + (k % m)];
    }
    out[k] = acc;
  }
}

/* work buffers of one thread's transforms of length n */
typedef struct { int n; const double complex *tw; double complex *a, *in, *scr; } fftws_t;

static void fftws_init(fftws_t *w, int n) {
  w->n = n;
  w->tw = twiddles(n);
  w->a = (double complex *)malloc(sizeof(double complex) * n);
  w->in = (double complex *)malloc(sizeof(double complex) * n);
  w->scr = (double complex *)malloc(sizeof(double complex) * n);
}

static void fftws_free(fftws_t *w) { free(w->a); free(w->in); free(w->scr); }

/* forward transform of w->a in place */
static void cfft_forward_ws(fftws_t *w) {
  memcpy(w->in, w->a, sizeof(double complex) * w->n);
  cfft_level(w->n, w->in, 1, w->a, w->scr, w->tw, 1);
}

/* ffte_r2c, ffte_mod.F90:66-89 */
static void r2c_ws(fftws_t *w, const double *in, double *out) {
  const int n = w->n;
  for (int i = 0; i < n; i++) w->a[i] = in[i];
  cfft_forward_ws(w);
  for (int i = 0; i <= n / 2; i++) { out[2 * i] = creal(w->a[i]); out[2 * i + 1] = cimag(w->a[i]); }
}

/* ffte_c2r, ffte_mod.F90:97-124 with ZFFT1D(IOPT=+1), vzfft1d.f:41-69 */
static void c2r_ws(fftws_t *w, const double *in, double *out) {
  const int n = w->n;
  double complex *a = w->a;
  for (int i = 0; i <= n / 2; i++) a[i] = in[2 * i] + I * in[2 * i + 1];
  for (int i = n / 2 + 1; i < n; i++) a[i] = conj(a[n - i]);
  for (int i = 0; i < n; i++) a[i] = conj(a[i]);
  cfft_forward_ws(w);
  double dn = 1.0 / (double)n;
  for (int i = 0; i < n; i++) out[i] = creal(conj(a[i]) * dn);
}

void orc_r2c(int n, const double *in, double *out) {
  fftws_t w;
  fftws_init(&w, n);
  r2c_ws(&w, in, out);
  fftws_free(&w);
}

void orc_c2r(int n, const double *in, double *out) {
  fftws_t w;
  fftws_init(&w, n);
  c2r_ws(&w, in, out);
  fftws_free(&w);
}

/* perform_forward_3dfft, pencilfft.F90:151-178 with transpose_and_forward_fft_in_y (:329-342)
 * and _in_x (:372-385): r2c along y of the interior, (Re,Im) interleaved (:838-864), then r2c
 * along x of each of those real y-rows.  Rows are independent; threads take whole rows. */
void orc_fft_forward(orc_t *o, const double *grid, double *spec) {
  const int nx = o->ncx, ny = o->ncy, nz = o->nzn, sy = o->sy, sx = o->sx;
  double *rb = dalloc((size_t)nz * sy * nx); /* (k, jj, x) */
#pragma omp parallel
  {
    double *rowy = dalloc(ny), *outy = dalloc(sy), *rowx = dalloc(nx), *outx = dalloc(sx);
    fftws_t wy, wx;
    fftws_init(&wy, ny);
    fftws_init(&wx, nx);
#pragma omp for schedule(static)
    for (int64_t kx = 0; kx < (int64_t)nz * nx; kx++) {
      const int k = (int)(kx / nx) + 1, x = (int)(kx % nx);
      for (int y = 0; y < ny; y++) rowy[y] = grid[GI(o, k, y + 3, x + 3)];
      r2c_ws(&wy, rowy, outy);
      for (int jj = 0; jj < sy; jj++) rb[((size_t)(k - 1) * sy + jj) * nx + x] = outy[jj];
    }
#pragma omp for schedule(static)
    for (int64_t kj = 0; kj < (int64_t)nz * sy; kj++) {
      const int k = (int)(kj / sy) + 1, jj = (int)(kj % sy);
      for (int x = 0; x < nx; x++) rowx[x] = rb[((size_t)(k - 1) * sy + jj) * nx + x];
      r2c_ws(&wx, rowx, outx);
      for (int ii = 0; ii < sx; ii++) spec[SI(o, k, jj + 1, ii + 1)] = outx[ii];
    }
    fftws_free(&wy); fftws_free(&wx);
    free(rowy); free(outy); free(rowx); free(outx);
  }
  free(rb);
}

/* perform_backwards_3dfft, pencilfft.F90:187-212: c2r along x of each (Re,Im)-interleaved
 * row (:351-364), then c2r along y (:394-408), written to the interior of grid. */
void orc_fft_backward(orc_t *o, const double *spec, double *grid) {
  const int nx = o->ncx, ny = o->ncy, nz = o->nzn, sy = o->sy, sx = o->sx;
  double *rb = dalloc((size_t)nz * sy * nx);
#pragma omp parallel
  {
    double *inx = dalloc(sx), *rowx = dalloc(nx), *iny = dalloc(sy), *rowy = dalloc(ny);
    fftws_t wy, wx;
    fftws_init(&wy, ny);
    fftws_init(&wx, nx);
#pragma omp for schedule(static)
    for (int64_t kj = 0; kj < (int64_t)nz * sy; kj++) {
      const int k = (int)(kj / sy) + 1, jj = (int)(kj % sy);
      for (int ii = 0; ii < sx; ii++) inx[ii] = spec[SI(o, k, jj + 1, ii + 1)];
      c2r_ws(&wx, inx, rowx);
      for (int x = 0; x < nx; x++) rb[((size_t)(k - 1) * sy + jj) * nx + x] = rowx[x];
    }
#pragma omp for schedule(static)
    for (int64_t kx = 0; kx < (int64_t)nz * nx; kx++) {
      const int k = (int)(kx / nx) + 1, x = (int)(kx % nx);
      for (int jj = 0; jj < sy; jj++) iny[jj] = rb[((size_t)(k - 1) * sy + jj) * nx + x];
      c2r_ws(&wy, iny, rowy);
      for (int y = 0; y < ny; y++) grid[GI(o, k, y + 3, x + 3)] = rowy[y];
    }
    fftws_free(&wy); fftws_free(&wx);
    free(inx); free(rowx); free(iny); free(rowy);
  }
  free(rb);
}

/* ---------------------------------------------------------------------------------------- */
/* fftops_mod (model_core/src/fft/fft_ops.F90)                                               */
/* ---------------------------------------------------------------------------------------- */

double orc_wavenumber(orc_t *o, int which, int jj, int ii) {
  size_t c = CI(o, jj, ii);
  return which == 0 ? o->kx[c] : which == 1 ? o->ky[c] : which == 2 ? o->k2[c] : o->filt[c];
}

/* spectral_filter, fft_ops.F90:176-197 */
void orc_filter(orc_t *o, const double *in, double *out) {
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= o->sx; i++)
    for (int j = 1; j <= o->sy; j++) {
      double fl = o->filt[CI(o, j, i)];
      for (int k = 1; k <= o->nzn; k++) out[SI(o, k, j, i)] = in[SI(o, k, j, i)] * fl;
    }
}

/* diffx, fft_ops.F90:204-302 (one rank: no pair split across processes) */
void orc_diffx(orc_t *o, const double *in, double *out) {
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= o->sx; i += 2)
    for (int j = 1; j <= o->sy; j++)
      for (int k = 1; k <= o->nzn; k++) {
        out[SI(o, k, j, i)] = -1. * in[SI(o, k, j, i + 1)];
        out[SI(o, k, j, i + 1)] = in[SI(o, k, j, i)];
      }
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= o->sx; i++)
    for (int j = 1; j <= o->sy; j++) {
      double kx = o->kx[CI(o, j, i)];
      for (int k = 1; k <= o->nzn; k++) out[SI(o, k, j, i)] = out[SI(o, k, j, i)] * kx;
    }
}

/* diffy, fft_ops.F90:305-401 */
void orc_diffy(orc_t *o, const double *in, double *out) {
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= o->sx; i++)
    for (int j = 1; j <= o->sy; j += 2)
      for (int k = 1; k <= o->nzn; k++) {
        out[SI(o, k, j, i)] = -1. * in[SI(o, k, j + 1, i)];
        out[SI(o, k, j + 1, i)] = in[SI(o, k, j, i)];
      }
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= o->sx; i++)
    for (int j = 1; j <= o->sy; j++) {
      double ky = o->ky[CI(o, j, i)];
      for (int k = 1; k <= o->nzn; k++) out[SI(o, k, j, i)] = out[SI(o, k, j, i)] * ky;
    }
}

/* tridiagonal, fft_ops.F90:620-656 (Thomas algorithm, no pivoting) */
int orc_tridiagonal(int n, const double *a, const double *b, const double *c, const double *d,
                    double *f) {
  double *cp = dalloc((size_t)n + 1);
  double denominator = b[0];
  f[0] = d[0] / denominator;
  for (int i = 2; i <= n; i++) {
    cp[i] = c[i - 2] / denominator;
    denominator = b[i - 1] - a[i - 1] * cp[i];
    if (denominator == 0) { free(cp); return PMPIC_ERR_ZERO_PIVOT; }
    f[i - 1] = (d[i - 1] - a[i - 1] * f[i - 2]) / (denominator);
  }
  for (int i = n - 1; i >= 1; i--) f[i - 1] = f[i - 1] - cp[i + 1] * f[i];
  free(cp);
  return PMPIC_OK;
}

/* diffz, fft_ops.F90:419-483 (literals 1./6., 2./3., 1./3. are default-real FP32).  Columns
 * are independent: threads take x-pair columns; a failing solve reports the first failing
 * column in the sequential (i, j) order. */
int orc_diffz(orc_t *o, const double *f, double *dfdz, const double *bot, const double *top) {
  const int nz = o->nzn;
  const double a0 = 0., b0 = 1., c0 = 0.;
  const double a1 = (double)(1.f / 6.f), b1 = (double)(2.f / 3.f), c1 = (double)(1.f / 6.f);
  const double an = 0., bn = 1., cn = 0.;
  const double al = (double)(1.f / 3.f), cl = (double)(1.f / 3.f);
  const double dz = o->dz;
  int bt = (bot != NULL && top != NULL);
  int64_t bad = INT64_MAX;
  int badst = PMPIC_OK;
#pragma omp parallel
  {
    double *a = dalloc(nz), *b = dalloc(nz), *c = dalloc(nz), *d = dalloc(nz), *col = dalloc(nz);
    a[0] = a0; b[0] = b0; c[0] = c0;
    for (int k = 2; k <= nz - 1; k++) { a[k - 1] = a1; b[k - 1] = b1; c[k - 1] = c1; }
    a[nz - 1] = an; b[nz - 1] = bn; c[nz - 1] = cn;
    if (bt) { c[0] = cl; a[nz - 1] = al; }
    int64_t mybad = INT64_MAX;
    int myst = PMPIC_OK;
#pragma omp for schedule(static)
    for (int i = 1; i <= o->sx; i++)
      for (int j = 1; j <= o->sy; j++) {
        if (bt) {
          d[0] = f[SI(o, 2, j, i)] / dz - a1 * bot[CI(o, j, i)] * dz;
          d[nz - 1] = -f[SI(o, nz - 1, j, i)] / dz + a1 * top[CI(o, j, i)] * dz;
        } else {
          d[0] = 0.0; d[nz - 1] = 0.0;
        }
        for (int k = 2; k <= nz - 1; k++) d[k - 1] = (f[SI(o, k + 1, j, i)] - f[SI(o, k - 1, j, i)]) / 2. / dz;
        int st = orc_tridiagonal(nz, a, b, c, d, col);
        if (st) {
          int64_t key = (int64_t)i * o->sy + j;
          if (key < mybad) { mybad = key; myst = st; }
          continue;
        }
        for (int k = 1; k <= nz; k++) dfdz[SI(o, k, j, i)] = col[k - 1];
      }
#pragma omp critical(orc_diffz)
    if (mybad < bad) { bad = mybad; badst = myst; }
    free(a); free(b); free(c); free(d); free(col);
  }
  return badst;
}

/* laplinv, fft_ops.F90:505-601; columns with k2 == 0 are left unchanged (:567).  Threads take
 * x-pair columns, as in orc_diffz. */
int orc_laplinv(orc_t *o, double *f, int df0) {
  const int nz = o->nzn;
  const double f16 = (double)(1.f / 6.f), f13 = (double)(1.f / 3.f);
  const double f112 = (double)(1.f / 12.f), f56 = (double)(5.f / 6.f);
  const double dz = o->dz;
  const double i2dz2 = 2. / dz / dz, idz2 = 1. / dz / dz;
  int64_t bad = INT64_MAX;
  int badst = PMPIC_OK;
#pragma omp parallel
  {
    double *a = dalloc(nz), *b = dalloc(nz), *c = dalloc(nz), *d = dalloc(nz), *col = dalloc(nz);
    int64_t mybad = INT64_MAX;
    int myst = PMPIC_OK;
#pragma omp for schedule(static)
    for (int i = 1; i <= o->sx; i++)
      for (int j = 1; j <= o->sy; j++) {
        double k2 = o->k2[CI(o, j, i)];
        if (k2 == 0) continue;
        if (!df0) {
          a[0] = 0.; b[0] = 1.; c[0] = 0.; d[0] = 0.;
          a[nz - 1] = 0.; b[nz - 1] = 1.; c[nz - 1] = 0.; d[nz - 1] = 0.;
        } else {
          a[0] = 0.;
          b[0] = -idz2 - f13 * k2;
          c[0] = idz2 - f16 * k2;
          d[0] = f13 * f[SI(o, 1, j, i)] + f16 * f[SI(o, 2, j, i)];
          a[nz - 1] = idz2 - f16 * k2;
          b[nz - 1] = -idz2 - f13 * k2;
          c[nz - 1] = 0.;
          d[nz - 1] = f13 * f[SI(o, nz, j, i)] + f16 * f[SI(o, nz - 1, j, i)];
        }
        for (int k = 2; k <= nz - 1; k++) {
          a[k - 1] = idz2 - f112 * k2;
          b[k - 1] = -i2dz2 - f56 * k2;
          c[k - 1] = idz2 - f112 * k2;
          d[k - 1] = f112 * f[SI(o, k - 1, j, i)] + f56 * f[SI(o, k, j, i)] + f112 * f[SI(o, k + 1, j, i)];
        }
        int st = orc_tridiagonal(nz, a, b, c, d, col);
        if (st) {
          int64_t key = (int64_t)i * o->sy + j;
          if (key < mybad) { mybad = key; myst = st; }
          continue;
        }
        for (int k = 1; k <= nz; k++) f[SI(o, k, j, i)] = col[k - 1];
      }
#pragma omp critical(orc_laplinv)
    if (mybad < bad) { bad = mybad; badst = myst; }
    free(a); free(b); free(c); free(d); free(col);
  }
  return badst;
}

/* ---------------------------------------------------------------------------------------- */
/* components                                                                                */
/* ---------------------------------------------------------------------------------------- */

static void interior_copy_spec(orc_t *o, const double *src, double *dst) { memcpy(dst, src, o->ssz * sizeof(double)); }

/* vort2vel timestep_callback, components/vort2vel/src/vort2vel.F90:98-400 */
int orc_vort2vel(orc_t *o) {
  const int nz = o->nzn, sy = o->sy, sx = o->sx;
  double *a = o->sa, *b = o->sb, *c = o->sc, *d = o->sd, *e = o->se, *f = o->sf;
  double *P = o->g[ORC_GP], *Q = o->g[ORC_GQ], *R = o->g[ORC_GR];
  const double dz = o->dz;
  const double hdzi = 1. / 2. / dz;
  int st;
  double *atop = dalloc((size_t)sy * sx), *abot = dalloc((size_t)sy * sx);
  double *btop = dalloc((size_t)sy * sx), *bbot = dalloc((size_t)sy * sx);
  double *ubar = dalloc(nz), *vbar = dalloc(nz);

  if ((st = orc_cache_weights(o))) goto done;
  orc_par2grid(o, o->f[ORC_P], P);
  orc_par2grid(o, o->f[ORC_Q], Q);
  orc_par2grid(o, o->f[ORC_R], R);

  orc_fft_forward(o, P, a);
  orc_fft_forward(o, Q, b);
  orc_fft_forward(o, R, c);

  orc_diffx(o, a, d);
  orc_diffy(o, b, e);
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= sx; i++)
    for (int j = 1; j <= sy; j++) {
      f[SI(o, 1, j, i)] = hdzi * (4. * c[SI(o, 2, j, i)] - c[SI(o, 3, j, i)] - 3. * c[SI(o, 1, j, i)]);
      for (int k = 2; k <= nz - 1; k++) f[SI(o, k, j, i)] = (c[SI(o, k + 1, j, i)] - c[SI(o, k - 1, j, i)]) * hdzi;
      f[SI(o, nz, j, i)] = hdzi * (c[SI(o, nz - 2, j, i)] + 3. * c[SI(o, nz, j, i)] - 4. * c[SI(o, nz - 1, j, i)]);
    }
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < o->ssz; q++) f[q] = d[q] + e[q] + f[q];
  for (int k = 1; k <= nz; k++) f[SI(o, k, 1, 1)] = 0.0;

  if ((st = orc_laplinv(o, f, 1))) goto done;
  orc_filter(o, f, f);

  orc_diffx(o, f, d);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < o->ssz; q++) a[q] = a[q] - d[q];
  orc_diffy(o, f, d);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < o->ssz; q++) b[q] = b[q] - d[q];
  if ((st = orc_diffz(o, f, d, NULL, NULL))) goto done;
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < o->ssz; q++) c[q] = c[q] - d[q];

#pragma omp parallel for schedule(static)
  for (int i = 1; i <= sx; i++)
    for (int j = 1; j <= sy; j++) c[SI(o, 1, j, i)] = 0.;

  orc_filter(o, a, d);
  orc_filter(o, b, e);
  orc_filter(o, c, f);

#pragma omp parallel for schedule(static)
  for (int i = 1; i <= sx; i++)
    for (int j = 1; j <= sy; j++) {
      atop[CI(o, j, i)] = a[SI(o, nz, j, i)];
      abot[CI(o, j, i)] = a[SI(o, 1, j, i)];
      btop[CI(o, j, i)] = b[SI(o, nz, j, i)];
      bbot[CI(o, j, i)] = b[SI(o, 1, j, i)];
    }

  orc_fft_backward(o, d, P);
  orc_fft_backward(o, e, Q);
  orc_fft_backward(o, f, R);

  /* mean flow, vort2vel.F90:230-241 (1./nz is a default-real quotient) */
  {
    ubar[0] = 0; vbar[0] = 0;
    for (int k = 1; k <= nz - 1; k++) {
      ubar[k] = ubar[k - 1] + dz / 2. * (e[SI(o, k, 1, 1)] + e[SI(o, k + 1, 1, 1)]);
      vbar[k] = vbar[k - 1] - dz / 2. * (d[SI(o, k, 1, 1)] + d[SI(o, k + 1, 1, 1)]);
    }
    double rnz = (double)(1.f / (float)nz);
    double su = 0, sv = 0;
    for (int k = 1; k <= nz - 1; k++) {
      su += ubar[k - 1] + 0.5 * ubar[nz - 1];
      sv += vbar[k - 1] + 0.5 * vbar[nz - 1];
    }
    double uavg = rnz * su, vavg = rnz * sv;
    for (int k = 0; k < nz; k++) { ubar[k] = ubar[k] - uavg; vbar[k] = vbar[k] - vavg; }
  }

  if ((st = orc_laplinv(o, a, 0))) goto done;
  if ((st = orc_laplinv(o, b, 0))) goto done;
  if ((st = orc_laplinv(o, c, 1))) goto done;

  /* u = db/dz - dc/dy */
  if ((st = orc_diffz(o, b, e, bbot, btop))) goto done;
  orc_diffy(o, c, d);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < o->ssz; q++) f[q] = e[q] - d[q];
  for (int k = 1; k <= nz; k++) f[SI(o, k, 1, 1)] = ubar[k - 1];
  orc_filter(o, f, o->s[ORC_SU]);
  orc_fft_backward(o, f, o->g[ORC_GU]);

  /* v = dc/dx - da/dz */
  if ((st = orc_diffz(o, a, e, abot, atop))) goto done;
  orc_diffx(o, c, d);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < o->ssz; q++) f[q] = d[q] - e[q];
  for (int k = 1; k <= nz; k++) f[SI(o, k, 1, 1)] = vbar[k - 1];
  orc_filter(o, f, o->s[ORC_SV]);
  orc_fft_backward(o, f, o->g[ORC_GV]);

  /* w = da/dy - db/dx */
  orc_diffx(o, b, d);
  orc_diffy(o, a, e);
#pragma omp parallel for schedule(static)
  for (size_t q = 0; q < o->ssz; q++) f[q] = e[q] - d[q];
  orc_filter(o, f, o->s[ORC_SW]);
  orc_fft_backward(o, f, o->g[ORC_GW]);

  orc_grid2par(o, o->g[ORC_GU], o->f[ORC_DXDT]);
  orc_grid2par(o, o->g[ORC_GV], o->f[ORC_DYDT]);
  orc_grid2par(o, o->g[ORC_GW], o->f[ORC_DZDT]);

  if (o->it_v2v % o->rksteps == 0) {
    /* maxval of the squares: the maximum does not depend on the order it is taken in */
    double umax = 0, vmax = 0, wmax = 0;
#pragma omp parallel for schedule(static) reduction(max : umax, vmax, wmax)
    for (int i = 3; i <= o->ncx + 2; i++)
      for (int j = 3; j <= o->ncy + 2; j++)
        for (int k = 1; k <= nz; k++) {
          double u = o->g[ORC_GU][GI(o, k, j, i)], v = o->g[ORC_GV][GI(o, k, j, i)], w = o->g[ORC_GW][GI(o, k, j, i)];
          double u2 = u * u, v2 = v * v, w2 = w * w;
          if (u2 > umax) umax = u2;
          if (v2 > vmax) vmax = v2;
          if (w2 > wmax) wmax = w2;
        }
    umax = sqrt(umax); vmax = sqrt(vmax); wmax = sqrt(wmax);
    const double tiny = (double)1.E-15f;
    if (umax == 0) umax = tiny;
    if (vmax == 0) vmax = tiny;
    if (wmax == 0) wmax = tiny;
    double Lx = (double)o->ncx * o->dx, Ly = (double)o->ncy * o->dy, Lz = (double)nz * o->dz;
    double t1 = Lx / umax, t2 = Ly / vmax, t3 = Lz / wmax;
    double dtmaxg = t1;
    if (t2 < dtmaxg) dtmaxg = t2;
    if (t3 < dtmaxg) dtmaxg = t3;
    if (o->dtmax > dtmaxg) o->dtm = dtmaxg; else o->dtm = o->dtmax;
  }
  o->it_v2v++;
  st = PMPIC_OK;
done:
  free(atop); free(abot); free(btop); free(bbot); free(ubar); free(vbar);
  return st;
}

/* vorticity_tendency timestep_callback, vorticity_tendency.F90:107-334 */
int orc_vorticity_tendency(orc_t *o) {
  double *P = o->g[ORC_GP], *Q = o->g[ORC_GQ], *R = o->g[ORC_GR];
  double *DP = o->g[ORC_GDP], *DQ = o->g[ORC_GDQ], *DR = o->g[ORC_GDR];
  double *bt = o->f[ORC_BTOT];
  const double *B = o->f[ORC_B], *H = o->f[ORC_H], *Z = o->f[ORC_Z];
#pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < o->np; n++) {
    double hl = H[n] - exp(-Z[n]);
    bt[n] = B[n] + 12.5 * (hl > 0.0 ? hl : 0.0);
  }
  orc_par2grid(o, bt, o->g[ORC_GB]);
  orc_fft_forward(o, o->g[ORC_GB], o->bs);

  orc_diffy(o, o->bs, o->as);
  orc_fft_backward(o, o->as, DP);
  orc_diffx(o, o->s[ORC_SU], o->as);
  orc_fft_backward(o, o->as, o->dudx);
#pragma omp parallel for schedule(static)
  for (int i = 3; i <= o->ncx + 2; i++)
    for (int j = 3; j <= o->ncy + 2; j++)
      for (int k = 1; k <= o->nzn; k++) { size_t q = GI(o, k, j, i); DP[q] = DP[q] + P[q] * o->dudx[q]; }
  orc_diffy(o, o->s[ORC_SU], o->as);
  orc_fft_backward(o, o->as, o->ap);
#pragma omp parallel for schedule(static)
  for (int i = 3; i <= o->ncx + 2; i++)
    for (int j = 3; j <= o->ncy + 2; j++)
      for (int k = 1; k <= o->nzn; k++) { size_t q = GI(o, k, j, i); DP[q] = DP[q] + Q[q] * o->ap[q]; }
  orc_diffx(o, o->s[ORC_SW], o->as);
  orc_fft_backward(o, o->as, o->dwdx);
#pragma omp parallel for schedule(static)
  for (int i = 3; i <= o->ncx + 2; i++)
    for (int j = 3; j <= o->ncy + 2; j++)
      for (int k = 1; k <= o->nzn; k++) { size_t q = GI(o, k, j, i); DP[q] = DP[q] + R[q] * (Q[q] + o->dwdx[q]); }

  orc_diffx(o, o->bs, o->as);
  orc_fft_backward(o, o->as, DQ);
  orc_diffx(o, o->s[ORC_SV], o->as);
  orc_fft_backward(o, o->as, o->ap);
#pragma omp parallel for schedule(static)
  for (int i = 3; i <= o->ncx + 2; i++)
    for (int j = 3; j <= o->ncy + 2; j++)
      for (int k = 1; k <= o->nzn; k++) { size_t q = GI(o, k, j, i); DQ[q] = -DQ[q] + P[q] * o->ap[q]; }
  orc_diffy(o, o->s[ORC_SV], o->as);
  orc_fft_backward(o, o->as, o->dvdy);
#pragma omp parallel for schedule(static)
  for (int i = 3; i <= o->ncx + 2; i++)
    for (int j = 3; j <= o->ncy + 2; j++)
      for (int k = 1; k <= o->nzn; k++) { size_t q = GI(o, k, j, i); DQ[q] = DQ[q] + Q[q] * o->dvdy[q]; }
  orc_diffy(o, o->s[ORC_SW], o->as);
  orc_fft_backward(o, o->as, o->dwdy);
#pragma omp parallel for schedule(static)
  for (int i = 3; i <= o->ncx + 2; i++)
    for (int j = 3; j <= o->ncy + 2; j++)
      for (int k = 1; k <= o->nzn; k++) { size_t q = GI(o, k, j, i); DQ[q] = DQ[q] + R[q] * (-P[q] + o->dwdy[q]); }

#pragma omp parallel for schedule(static)
  for (int i = 3; i <= o->ncx + 2; i++)
    for (int j = 3; j <= o->ncy + 2; j++)
      for (int k = 1; k <= o->nzn; k++) {
        size_t q = GI(o, k, j, i);
        DR[q] = P[q] * o->dwdx[q];
        DR[q] = DR[q] + Q[q] * o->dwdy[q];
        DR[q] = DR[q] + R[q] * (-o->dudx[q]);
        DR[q] = DR[q] + R[q] * (-o->dvdy[q]);
      }

  orc_grid2par(o, DP, o->f[ORC_DPDT]);
  orc_grid2par(o, DQ, o->f[ORC_DQDT]);
  orc_grid2par(o, DR, o->f[ORC_DRDT]);

  if (o->it_vt % o->rksteps == 0) {
    double omax = 0;
#pragma omp parallel for schedule(static) reduction(max : omax)
    for (int i = 3; i <= o->ncx + 2; i++)
      for (int j = 3; j <= o->ncy + 2; j++)
        for (int k = 1; k <= o->nzn; k++) {
          size_t q = GI(o, k, j, i);
          double w2 = P[q] * P[q] + Q[q] * Q[q] + R[q] * R[q];
          if (w2 > omax) omax = w2;
        }
    omax = sqrt(omax);
    double dtmax;
    if (omax > 0.) dtmax = o->prefactor / omax; else dtmax = o->dtmax;
    if (o->dtm > dtmax) o->dtm = dtmax;
  }
  o->it_vt++;
  return PMPIC_OK;
}

/* parcels_reflect_vertical_boundaries + one-rank parcel_haloswap, haloswap.F90:109-289,
 * 296-456, 489-607, 678-771: parcels that leave through x/y are "sent" to self in direction
 * order N,NE,E,SE,S,SW,W,NW (1..8), shifted by one period, and unpacked into the vacated
 * slots in ascending slot order; the count is unchanged so backfill never runs. */
int orc_parcel_haloswap(orc_t *o) {
  const int64_t n0 = o->np;
  double *X = o->f[ORC_X], *Y = o->f[ORC_Y], *Z = o->f[ORC_Z];
  const double xmin = o->hs_xmin, xmax = o->hs_xmax, ymin = o->hs_ymin, ymax = o->hs_ymax;
  int64_t total = 0;
  /* reflection, then each parcel's direction (the eight tests are mutually exclusive, so
   * testing them per parcel marks what the direction-by-direction sweeps mark) */
#pragma omp parallel for schedule(static) reduction(+ : total)
  for (int64_t i = 0; i < n0; i++) {
    double z = Z[i];
    if (z < o->zmin_h) Z[i] = o->zmin_h + (o->zmin_h - z);
    else if (z > o->zmax_h) Z[i] = o->zmax_h - (z - o->zmax_h);
    double x = X[i], y = Y[i];
    int dir = 0;
    if ((y >= ymax) && (x >= xmin && x < xmax)) dir = 1;          /* N */
    else if (x >= xmax && y >= ymax) dir = 2;                     /* NE */
    else if ((x >= xmax) && (y >= ymin && y < ymax)) dir = 3;     /* E */
    else if (x >= xmax && y < ymin) dir = 4;                      /* SE */
    else if ((y < ymin) && (x >= xmin && x < xmax)) dir = 5;      /* S */
    else if (x < xmin && y < ymin) dir = 6;                       /* SW */
    else if ((x < xmin) && (y >= ymin && y < ymax)) dir = 7;      /* W */
    else if (x < xmin && y >= ymax) dir = 8;                      /* NW */
    o->index[i] = dir;
    total += (dir != 0);
  }
  if (total == 0) return PMPIC_OK;
  /* the crossing parcels in (direction, ascending slot) order */
  int64_t nsend[9] = {0};
  for (int64_t i = 0; i < n0; i++) nsend[o->index[i]]++;
  int64_t off[9];
  off[1] = 0;
  for (int d = 1; d < 8; d++) off[d + 1] = off[d] + nsend[d];
  int64_t *src = (int64_t *)malloc(sizeof(int64_t) * total);
  for (int64_t i = 0; i < n0; i++)
    if (o->index[i] != 0) src[off[o->index[i]]++] = i;
  /* pack, with the periodic shift */
  const int NF = ORC_BTOT; /* 29 fields travel: 17 properties + 12 RK arrays */
  double *buf = dalloc((size_t)total * NF);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < total; t++) {
    const int64_t i = src[t];
    const int d = o->index[i];
    double xs = 0, ys = 0;
    if (d == 3 || d == 2 || d == 4) xs = o->xshift_e;
    else if (d == 7 || d == 8 || d == 6) xs = o->xshift_w;
    if (d == 1 || d == 2 || d == 8) ys = o->yshift_n;
    else if (d == 5 || d == 6 || d == 4) ys = o->yshift_s;
    for (int fld = 0; fld < NF; fld++) {
      double v = o->f[fld][i];
      if (fld == ORC_X || fld == ORC_XO || fld == ORC_XF) v = v + xs;
      if (fld == ORC_Y || fld == ORC_YO || fld == ORC_YF) v = v + ys;
      buf[(size_t)t * NF + fld] = v;
    }
  }
  /* unpack (haloswap.F90:699-716): free slot = beyond n_initial or index != 0 */
  int64_t istart = 0;
  for (int64_t m = 0; m < total; m++) {
    int64_t slot = -1;
    for (int64_t i = istart; i < o->maxp; i++)
      if (i >= n0 || o->index[i] != 0) { slot = i; o->index[i] = 0; istart = i + 1; break; }
    if (slot < 0) { free(buf); free(src); return PMPIC_ERR_MAXPARCELS; }
    for (int fld = 0; fld < NF; fld++) o->f[fld][slot] = buf[(size_t)m * NF + fld];
  }
  free(buf); free(src);
  return PMPIC_OK;
}

/* rk4_integrator timestep_callback, rk4_integrator.F90:95-319 */
int orc_rk4_integrator(orc_t *o) {
  const double dt = o->dtm, dt6 = dt / 6, dt2 = dt / 2, dt3 = dt / 3;
  const int64_t n = o->np;
  double *Yv[6] = {o->f[ORC_X], o->f[ORC_Y], o->f[ORC_Z], o->f[ORC_P], o->f[ORC_Q], o->f[ORC_R]};
  double *Yo[6] = {o->f[ORC_XO], o->f[ORC_YO], o->f[ORC_ZO], o->f[ORC_PO], o->f[ORC_QO], o->f[ORC_RO]};
  double *Yf[6] = {o->f[ORC_XF], o->f[ORC_YF], o->f[ORC_ZF], o->f[ORC_PF], o->f[ORC_QF], o->f[ORC_RF]};
  double *K[6] = {o->f[ORC_DXDT], o->f[ORC_DYDT], o->f[ORC_DZDT], o->f[ORC_DPDT], o->f[ORC_DQDT], o->f[ORC_DRDT]};
  const int stage = o->rkstep;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++)
    for (int v = 0; v < 6; v++) {
      switch (stage) {
        case 1:
          Yo[v][i] = Yv[v][i];
          Yf[v][i] = Yo[v][i] + dt6 * K[v][i];
          Yv[v][i] = Yo[v][i] + dt2 * K[v][i];
          break;
        case 2:
          Yf[v][i] = Yf[v][i] + dt3 * K[v][i];
          Yv[v][i] = Yo[v][i] + dt2 * K[v][i];
          break;
        case 3:
          Yf[v][i] = Yf[v][i] + dt3 * K[v][i];
          Yv[v][i] = Yo[v][i] + dt * K[v][i];
          break;
        default:
          Yf[v][i] = Yf[v][i] + dt6 * K[v][i];
          Yv[v][i] = Yf[v][i];
          break;
      }
    }
  o->rkstep = (o->rkstep < 4) ? o->rkstep + 1 : 1;
  return orc_parcel_haloswap(o);
}

/* parcel_splitting timestep_callback, parcel_splitting.F90:47-168 */
int orc_parcel_splitting(orc_t *o) {
  int st = PMPIC_OK;
  if (o->it_split % o->rksteps == o->rksteps - 1) {
    const double dt = o->dtm;
    const double third = (double)(1.f / 3.f);
    const double pi = 4. * atan(1.0);
    const double stretchmax = 4.;
    int64_t n = o->np;
    const int64_t norig = o->np;
    double *X = o->f[ORC_X], *Y = o->f[ORC_Y], *Z = o->f[ORC_Z];
    double *P = o->f[ORC_P], *Q = o->f[ORC_Q], *R = o->f[ORC_R];
    double *S = o->f[ORC_STRETCH], *V = o->f[ORC_VOL];
    o->last_split = 0;
    /* the stretch increments touch only parcel i, so they are taken for every parcel first;
     * the splits then run in ascending parcel order, appending children as the loop does */
    int *hit = o->keep;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < norig; i++) {
      double ds = P[i] * o->f[ORC_DPDT][i] + Q[i] * o->f[ORC_DQDT][i] + R[i] * o->f[ORC_DRDT][i];
      ds = pow(fabs(ds), third);
      S[i] = S[i] + ds * dt;
      hit[i] = S[i] > stretchmax;
    }
    for (int64_t i = 0; i < norig; i++) {
      if (hit[i]) {
        double r = V[i] / 4. / pi;
        r = pow(r, third);
        double absw = P[i] * P[i] + Q[i] * Q[i] + R[i] * R[i];
        absw = sqrt(absw);
        if (n == o->maxp) return PMPIC_ERR_MAXPARCELS;
        int64_t c = n;  /* new slot n+1 (1-based) */
        n = n + 1;
        X[c] = X[i] + r * P[i] / absw;
        Y[c] = Y[i] + r * Q[i] / absw;
        Z[c] = Z[i] + r * R[i] / absw;
        P[c] = P[i]; Q[c] = Q[i]; R[c] = R[i];
        o->f[ORC_DXDT][c] = o->f[ORC_DXDT][i];
        o->f[ORC_DYDT][c] = o->f[ORC_DYDT][i];
        o->f[ORC_DZDT][c] = o->f[ORC_DZDT][i];
        o->f[ORC_DPDT][c] = o->f[ORC_DPDT][i];
        o->f[ORC_DQDT][c] = o->f[ORC_DQDT][i];
        o->f[ORC_DRDT][c] = o->f[ORC_DRDT][i];
        o->f[ORC_H][c] = o->f[ORC_H][i];
        o->f[ORC_B][c] = o->f[ORC_B][i];
        V[c] = V[i] / 2;
        S[c] = 0.;
        o->f[ORC_TAG][c] = o->f[ORC_TAG][i];
        X[i] = X[i] - r * P[i] / absw;
        Y[i] = Y[i] - r * Q[i] / absw;
        Z[i] = Z[i] - r * R[i] / absw;
        V[i] = V[i] / 2;
        S[i] = 0.;
        o->last_split++;
      }
    }
    o->np = n;
    st = orc_parcel_haloswap(o);
  }
  o->it_split++;
  return st;
}

/* parcel_mixing timestep_callback + construct_grids, parcel_mixing.F90:68-431, 440-624 */
int orc_parcel_mixing(orc_t *o) {
  int st = PMPIC_OK;
  if (o->it_mix % o->rksteps == o->rksteps - 1) {
    const double dx = o->dx, dy = o->dy, dz = o->dz;
    const double vmin = dx * dy * dz / 6. / 6. / 6.;
    const int nx = o->nxh, ny = o->nyh, nz = o->nzn;
    if ((st = orc_cache_weights(o))) goto out;
    const int64_t norig = o->np;
    int *keep = o->keep;
    double *GV = o->g[ORC_GVOL], *GB = o->g[ORC_GB], *GH = o->g[ORC_GHG];
    double *GP = o->g[ORC_GP], *GQ = o->g[ORC_GQ], *GR = o->g[ORC_GR];
    double *vres = o->vres, *bres = o->bres, *hres = o->hres, *pres = o->pres, *qres = o->qres, *rres = o->rres;
    memset(o->npercell, 0, sizeof(int) * (size_t)(nz - 1) * (ny - 1) * (nx - 1));
    double *Gs[6] = {GV, GB, GH, GP, GQ, GR};
    double *Rs[6] = {vres, bres, hres, pres, qres, rres};
    for (int g = 0; g < 6; g++) { memset(Gs[g], 0, o->gsz * sizeof(double)); memset(Rs[g], 0, o->gsz * sizeof(double)); }
    int64_t nremove = 0;
    const double *PB = o->f[ORC_B], *PH = o->f[ORC_H], *PP = o->f[ORC_P], *PQ = o->f[ORC_Q], *PR = o->f[ORC_R];
#pragma omp parallel for schedule(static) reduction(+ : nremove)
    for (int64_t n = 0; n < norig; n++) {
      keep[n] = (o->f[ORC_VOL][n] < vmin) ? 0 : 1;
      nremove += (keep[n] == 0);
    }
    /* the scatter, by node-column slab as in orc_par2grid: each node gets its terms in
     * ascending parcel order; a slab counts the kept parcels of the cells in its own columns */
#pragma omp parallel num_threads(o->nslab)
    {
      for (int t = omp_get_thread_num(); t < o->nslab; t += omp_get_num_threads()) {
        const int lo = o->slab_lo[t], hi = o->slab_lo[t + 1];
        const int64_t *list = o->slab_list[t];
        for (int64_t mm = 0; mm < o->slab_n[t]; mm++) {
          const int64_t n = list[mm];
          if (n >= norig) continue;
          int i = o->is[n], j = o->js[n], k = o->ks[n];
          double delx = o->delxs[n], dely = o->delys[n], delz = o->delzs[n];
          double v = o->f[ORC_VOL][n];
          double w[8];
          w[0] = (1 - delz) * (1 - dely) * (1 - delx) * v;
          w[1] = (1 - delz) * (1 - dely) * (delx)*v;
          w[2] = (1 - delz) * (dely) * (1 - delx) * v;
          w[3] = (1 - delz) * (dely) * (delx)*v;
          w[4] = (delz) * (1 - dely) * (1 - delx) * v;
          w[5] = (delz) * (1 - dely) * (delx)*v;
          w[6] = (delz) * (dely) * (1 - delx) * v;
          w[7] = (delz) * (dely) * (delx)*v;
          size_t q[8] = {GI(o, k, j, i), GI(o, k, j, i + 1), GI(o, k, j + 1, i), GI(o, k, j + 1, i + 1),
                         GI(o, k + 1, j, i), GI(o, k + 1, j, i + 1), GI(o, k + 1, j + 1, i), GI(o, k + 1, j + 1, i + 1)};
          /* corners 0, 2, 4, 6 are in node column i, the others in i + 1 */
          const int own[2] = {i >= lo && i < hi, i + 1 >= lo && i + 1 < hi};
          double **T;
          if (keep[n] == 0) T = Rs;
          else {
            if (own[0]) o->npercell[NPC(o, k, j, i)] += 1;
            T = Gs;
          }
          const double val[5] = {PB[n], PH[n], PP[n], PQ[n], PR[n]};
          for (int c = 0; c < 8; c++)
            if (own[c & 1]) T[0][q[c]] = T[0][q[c]] + w[c];
          for (int g = 1; g < 6; g++)
            for (int c = 0; c < 8; c++)
              if (own[c & 1]) T[g][q[c]] = T[g][q[c]] + w[c] * val[g - 1];
        }
      }
    }
    for (int g = 0; g < 6; g++) { orc_fold(o, Gs[g]); orc_fill(o, Gs[g]); }
    for (int g = 0; g < 6; g++)
      for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++) {
          Gs[g][GI(o, 1, j, i)] = Gs[g][GI(o, 1, j, i)] * 2.;
          Gs[g][GI(o, nz, j, i)] = Gs[g][GI(o, nz, j, i)] * 2;
        }
    int64_t nadd = 0;
    int64_t n = 1; /* 1-based slot cursor, parcel_mixing.F90:158 */
    double *X = o->f[ORC_X], *Y = o->f[ORC_Y], *Z = o->f[ORC_Z];
    for (int i = 3; i <= nx - 2; i++)
      for (int j = 3; j <= ny - 2; j++)
        for (int k = 1; k <= nz - 1; k++) {
          if (o->npercell[NPC(o, k, j, i)] < 3) {
            nadd++;
            if (n <= o->np) {
              while (keep[n - 1] != 0) {
                n = n + 1;
                if (n > o->np) break;
              }
            }
            size_t q[8] = {GI(o, k, j, i), GI(o, k + 1, j, i), GI(o, k, j + 1, i), GI(o, k + 1, j + 1, i),
                           GI(o, k, j, i + 1), GI(o, k + 1, j, i + 1), GI(o, k, j + 1, i + 1), GI(o, k + 1, j + 1, i + 1)};
            double v = 0;
            for (int c = 0; c < 8; c++) v += GV[q[c]];
            double portion = 2 * vmin / v;
            int64_t s = n - 1;
            X[s] = 0.5 * (o->xc[i - 1] + o->xc[i]);
            Y[s] = 0.5 * (o->yc[j - 1] + o->yc[j]);
            Z[s] = 0.5 * (o->zc[k - 1] + o->zc[k]);
            o->f[ORC_VOL][s] = 2 * vmin;
            o->f[ORC_STRETCH][s] = 0;
            double sb = 0, sh = 0, sp = 0, sq = 0, sr = 0;
            for (int c = 0; c < 8; c++) sb += GB[q[c]];
            for (int c = 0; c < 8; c++) sh += GH[q[c]];
            for (int c = 0; c < 8; c++) sp += GP[q[c]];
            for (int c = 0; c < 8; c++) sq += GQ[q[c]];
            for (int c = 0; c < 8; c++) sr += GR[q[c]];
            o->f[ORC_B][s] = sb / v;
            o->f[ORC_H][s] = sh / v;
            o->f[ORC_P][s] = sp / v;
            o->f[ORC_Q][s] = sq / v;
            o->f[ORC_R][s] = sr / v;
            if (n < o->np) keep[n - 1] = 2;
            n = n + 1;
            if (n > o->maxp) { st = PMPIC_ERR_MAXPARCELS; goto out; }
            for (int g = 0; g < 6; g++)
              for (int c = 0; c < 8; c++) Rs[g][q[c]] = Rs[g][q[c]] - portion * Gs[g][q[c]];
          }
        }
    o->np = (n - 1 > o->np) ? n - 1 : o->np;
    for (int g = 0; g < 6; g++) { orc_fold(o, Rs[g]); orc_fill(o, Rs[g]); }
    for (int g = 0; g < 6; g++)
      for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++) {
          Rs[g][GI(o, 1, j, i)] = Rs[g][GI(o, 1, j, i)] * 2;
          Rs[g][GI(o, nz, j, i)] = Rs[g][GI(o, nz, j, i)] * 2;
        }
    double *PV = o->f[ORC_VOL];
    double *PBw = o->f[ORC_B], *PHw = o->f[ORC_H], *PPw = o->f[ORC_P], *PQw = o->f[ORC_Q], *PRw = o->f[ORC_R];
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < norig; m++) {
      if (keep[m] != 1) continue;
      int i = o->is[m], j = o->js[m], k = o->ks[m];
      double delx = o->delxs[m], dely = o->delys[m], delz = o->delzs[m];
      double v = PV[m];
      double w[8];
      w[0] = (1 - delz) * (1 - dely) * (1 - delx) * v;
      w[1] = (1 - delz) * (1 - dely) * (delx)*v;
      w[2] = (1 - delz) * (dely) * (1 - delx) * v;
      w[3] = (1 - delz) * (dely) * (delx)*v;
      w[4] = (delz) * (1 - dely) * (1 - delx) * v;
      w[5] = (delz) * (1 - dely) * (delx)*v;
      w[6] = (delz) * (dely) * (1 - delx) * v;
      w[7] = (delz) * (dely) * (delx)*v;
      size_t q[8] = {GI(o, k, j, i), GI(o, k, j, i + 1), GI(o, k, j + 1, i), GI(o, k, j + 1, i + 1),
                     GI(o, k + 1, j, i), GI(o, k + 1, j, i + 1), GI(o, k + 1, j + 1, i), GI(o, k + 1, j + 1, i + 1)};
      double vn = PV[m];
      for (int c = 0; c < 8; c++) vn = vn + w[c] * vres[q[c]] / GV[q[c]];
      PV[m] = vn;
      double *Pw[5] = {PBw, PHw, PPw, PQw, PRw};
      double *Rw[5] = {bres, hres, pres, qres, rres};
      for (int g = 0; g < 5; g++) {
        double acc = Pw[g][m] * v;
        for (int c = 0; c < 8; c++) acc = acc + w[c] / GV[q[c]] * Rw[g][q[c]];
        Pw[g][m] = acc;
        Pw[g][m] = Pw[g][m] / PV[m];
      }
    }
    o->last_removed = nremove;
    o->last_added = nadd;
    if (o->np == norig) {
      int64_t last = norig; /* 1-based */
      for (int64_t m = 1; m <= norig; m++) {
        if (keep[m - 1] == 0) {
          for (;;) {
            if (last < 1) { st = PMPIC_ERR_BACKFILL; goto out; }
            if (keep[last - 1] == 1) break;
            last = last - 1;
          }
          static const int cp[11] = {ORC_X, ORC_Y, ORC_Z, ORC_P, ORC_Q, ORC_R, ORC_B, ORC_H, ORC_VOL, ORC_TAG, ORC_STRETCH};
          for (int c = 0; c < 11; c++) o->f[cp[c]][m - 1] = o->f[cp[c]][last - 1];
          last = last - 1;
          o->np = o->np - 1;
        }
      }
    }
  }
out:
  o->it_mix++;
  return st;
}

/* write_grids_to_file's grid computation, writebinary.F90:217-233: par2grid of b, p, q, r and
 * h, then over every element hgliq = max(0., hg - exp(-z_coords(k))) and b = b + 12.5*hgliq */
int orc_writebinary_grids(orc_t *o) {
  int st;
  if ((st = orc_cache_weights(o))) return st;
  orc_par2grid(o, o->f[ORC_B], o->g[ORC_GB]);
  orc_par2grid(o, o->f[ORC_P], o->g[ORC_GP]);
  orc_par2grid(o, o->f[ORC_Q], o->g[ORC_GQ]);
  orc_par2grid(o, o->f[ORC_R], o->g[ORC_GR]);
  orc_par2grid(o, o->f[ORC_H], o->g[ORC_GHG]);
  double *B = o->g[ORC_GB], *HG = o->g[ORC_GHG], *HL = o->g[ORC_GHGLIQ];
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= o->nxh; i++)
    for (int j = 1; j <= o->nyh; j++)
      for (int k = 1; k <= o->nzn; k++) {
        size_t q = GI(o, k, j, i);
        double hl = HG[q] - exp(-o->zc[k - 1]);
        HL[q] = hl > 0. ? hl : 0.;
        B[q] = B[q] + 12.5 * HL[q];
      }
  return PMPIC_OK;
}

/* one timestep: the integrator group (global_config:43) run rksteps times
 * (timestepper.F90:89-106), then time += dtm (monc.F90:223-227) */
int orc_step(orc_t *o) {
  int st;
  for (int s = 0; s < o->rksteps; s++) {
    if ((st = orc_vort2vel(o))) return st;
    if ((st = orc_vorticity_tendency(o))) return st;
    if ((st = orc_rk4_integrator(o))) return st;
    if ((st = orc_parcel_splitting(o))) return st;
    if ((st = orc_parcel_mixing(o))) return st;
  }
  o->time = o->time + o->dtm;
  return PMPIC_OK;
}
