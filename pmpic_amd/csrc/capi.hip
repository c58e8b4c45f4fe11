// C ABI of libpmpic_hip.so (declared in include/pmpic_hip.h): context set-up, host transfers in
// the reference's layouts, the component callbacks and the standalone operators.
#include "common.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <array>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/pmpic_hip.h"

using namespace pmpic;

namespace pmpic {

int timing_begin(Ctx *c, int which) {
  if (c->timing != 1) return PMPIC_OK;
  if (c->nev[which] >= 64) PMPIC_TRY(timing_harvest(c));
  PMPIC_CHECK(hipEventRecord(c->ev[which][c->nev[which]][0], c->stream));
  return PMPIC_OK;
}

int timing_end(Ctx *c, int which) {
  if (c->timing != 1) return PMPIC_OK;
  PMPIC_CHECK(hipEventRecord(c->ev[which][c->nev[which]][1], c->stream));
  c->nev[which]++;
  return PMPIC_OK;
}

int timing_harvest(Ctx *c) {
  for (int w = 0; w < NTIMERS; w++) {
    for (int e = 0; e < c->nev[w]; e++) {
      PMPIC_CHECK(hipEventSynchronize(c->ev[w][e][1]));
      float ms = 0.f;
      PMPIC_CHECK(hipEventElapsedTime(&ms, c->ev[w][e][0], c->ev[w][e][1]));
      c->t_ms[w] += ms;
      c->t_calls[w]++;
      if (ms > c->t_max[w]) c->t_max[w] = ms;
      c->t_sq[w] += (double)ms * (double)ms;
    }
    c->nev[w] = 0;
  }
  return PMPIC_OK;
}

// the device staging buffer of host-array transfers, allocated at its first use: a parcel field
// or a parcel array followed by a grid (max(cap, gsz) + max(gsz, ssz) doubles)
int ensure_fbuf(Ctx *c) {
  if (c->fbuf) return PMPIC_OK;
  const int64_t a = c->cap > c->gsz ? c->cap : c->gsz;
  const int64_t b = c->gsz > c->ssz ? c->gsz : c->ssz;
  PMPIC_CHECK(hipMalloc((void **)&c->fbuf, sizeof(double) * (size_t)(a + b)));
  return PMPIC_OK;
}

// status raised by kernels since the last check; clears it
int take_status(Ctx *c) {
  int st = sync_scalars(c);
  if (st != PMPIC_OK) {
    int zero = 0;
    hipMemcpy(&c->dsc->status, &zero, sizeof(int), hipMemcpyHostToDevice);
  }
  return st;
}

}  // namespace pmpic

namespace {

template <class T>
int dalloc(T **p, int64_t n) {
  *p = nullptr;
  if (n < 1) n = 1;
  PMPIC_CHECK(hipMalloc((void **)p, sizeof(T) * (size_t)n));
  PMPIC_CHECK(hipMemset(*p, 0, sizeof(T) * (size_t)n));
  return PMPIC_OK;
}

// a transform length of the semi-spectral passes: factors 2, 3 and 5 only, and at most 512
// (the twiddle table a workgroup stages)
int check_factors(int n) {
  int m = n;
  if (m < 1 || m > 512) return 0;
  while (m % 5 == 0) m /= 5;
  while (m % 3 == 0) m /= 3;
  while (m % 2 == 0) m /= 2;
  return m == 1;
}

// geometry exactly as setup_grid.F90:73-95, parcel_interpolation.F90:94-136 and
// haloswap.F90:314-318, 529-552 compute it on one rank
void make_geometry(const pmpic_config *cf, Geometry &g, std::vector<double> &coords) {
  g.nx = cf->nx; g.ny = cf->ny; g.ncz = cf->nz; g.nzn = cf->nz + 1;
  g.sx = (cf->nx / 2 + 1) * 2; g.sy = (cf->ny / 2 + 1) * 2;
  g.xbot = cf->xmin; g.ybot = cf->ymin; g.zbot = cf->zmin;
  g.xtop = cf->xmax; g.ytop = cf->ymax; g.ztop = cf->zmax;
  g.dx = (cf->xmax - cf->xmin) / cf->nx;
  g.dy = (cf->ymax - cf->ymin) / cf->ny;
  g.dz = (cf->zmax - cf->zmin) / cf->nz;
  g.xmin_h = (double)(-2) * g.dx + cf->xmin;
  g.ymin_h = (double)(-2) * g.dy + cf->ymin;
  g.zmin_h = (double)0 * g.dz + cf->zmin;
  g.zmax_h = (double)(g.nzn - 1) * g.dz;
  const int nxh = cf->nx + 4, nyh = cf->ny + 4, nzn = g.nzn;
  coords.assign(nxh + nyh + nzn, 0.0);
  for (int n = 1; n <= nxh; n++) coords[n - 1] = g.xmin_h + (double)(n - 1) * g.dx + cf->xmin;
  for (int n = 1; n <= nyh; n++) coords[nxh + n - 1] = g.ymin_h + (double)(n - 1) * g.dy + cf->ymin;
  for (int n = 1; n <= nzn; n++) coords[nxh + nyh + n - 1] = g.zmin_h + (double)(n - 1) * g.dz + cf->zmin;
  g.hs_xmin = coords[2];
  g.hs_xmax = coords[nxh - 2];
  g.hs_ymin = coords[nxh + 2];
  g.hs_ymax = coords[nxh + nyh - 2];
  const double xtop = cf->xmax - g.dx, ytop = cf->ymax - g.dy;
  g.xshift_e = cf->xmin - xtop - g.dx;
  g.xshift_w = xtop - cf->xmin + g.dx;
  g.yshift_n = cf->ymin - ytop - g.dy;
  g.yshift_s = ytop - cf->ymin + g.dy;
  g.vmin = g.dx * g.dy * g.dz / 6. / 6. / 6.;
  g.lx_cfl = (double)cf->nx * g.dx;
  g.ly_cfl = (double)cf->ny * g.dy;
  g.lz_cfl = (double)g.nzn * g.dz;
}

// reference grid layout (z fastest, x/y halo 2) <-> device interior [k][j][i]
void grid_to_device_layout(const Geometry &g, const double *ref, std::vector<double> &dev) {
  const int nyh = g.ny + 4, nzn = g.nzn;
  dev.resize((size_t)nzn * g.ny * g.nx);
  for (int k = 0; k < nzn; k++)
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++)
        dev[((size_t)k * g.ny + j) * g.nx + i] = ref[((size_t)(i + 2) * nyh + (j + 2)) * nzn + k];
}

void grid_from_device_layout(const Geometry &g, const std::vector<double> &dev, double *ref) {
  const int nxh = g.nx + 4, nyh = g.ny + 4, nzn = g.nzn;
  for (int ih = 0; ih < nxh; ih++) {
    int i = ((ih - 2) % g.nx + g.nx) % g.nx;
    for (int jh = 0; jh < nyh; jh++) {
      int j = ((jh - 2) % g.ny + g.ny) % g.ny;
      for (int k = 0; k < nzn; k++)
        ref[((size_t)ih * nyh + jh) * nzn + k] = dev[((size_t)k * g.ny + j) * g.nx + i];
    }
  }
}

void spec_to_device_layout(const Geometry &g, const double *ref, std::vector<double> &dev) {
  dev.resize((size_t)g.nzn * g.sy * g.sx);
  for (int k = 0; k < g.nzn; k++)
    for (int jj = 0; jj < g.sy; jj++)
      for (int ii = 0; ii < g.sx; ii++)
        dev[((size_t)k * g.sy + jj) * g.sx + ii] = ref[((size_t)ii * g.sy + jj) * g.nzn + k];
}

void spec_from_device_layout(const Geometry &g, const std::vector<double> &dev, double *ref) {
  for (int k = 0; k < g.nzn; k++)
    for (int jj = 0; jj < g.sy; jj++)
      for (int ii = 0; ii < g.sx; ii++)
        ref[((size_t)ii * g.sy + jj) * g.nzn + k] = dev[((size_t)k * g.sy + jj) * g.sx + ii];
}

void col_to_device_layout(const Geometry &g, const double *ref, std::vector<double> &dev) {
  dev.resize((size_t)g.sy * g.sx);
  for (int jj = 0; jj < g.sy; jj++)
    for (int ii = 0; ii < g.sx; ii++) dev[(size_t)jj * g.sx + ii] = ref[(size_t)ii * g.sy + jj];
}

int upload(Ctx *c, double *dst, const std::vector<double> &h) {
  PMPIC_CHECK(hipMemcpyAsync(dst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                             c->stream));
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  return PMPIC_OK;
}

int download(Ctx *c, std::vector<double> &h, const double *src, size_t n) {
  h.resize(n);
  PMPIC_CHECK(hipMemcpyAsync(h.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  return PMPIC_OK;
}


int set_dev_scalars(Ctx *c) {
  c->hsc.dtm = c->dtm;
  c->hsc.dtmax = c->dtmax;
  c->hsc.prefactor = c->prefactor;
  c->hsc.n = c->n;
  c->hsc.status = 0;
  PMPIC_CHECK(hipMemcpy(c->dsc, &c->hsc, sizeof(DevScalars), hipMemcpyHostToDevice));
  return PMPIC_OK;
}

void graph_drop(Ctx *c) {
  if (c->g123x) hipGraphExecDestroy(c->g123x);
  if (c->g123) hipGraphDestroy(c->g123);
  c->g123x = nullptr;
  c->g123 = nullptr;
  c->g123_n = 0;
}

int free_all(Ctx *c) {
  graph_drop(c);
  b32_free(c);
  hipFree(c->par); hipFree(c->tmp);
  hipFree(c->cell); hipFree(c->rank); hipFree(c->sorted); hipFree(c->flag);
  hipFree(c->list_a); hipFree(c->list_b);
  hipFree(c->cell_count); hipFree(c->cell_start); hipFree(c->partial); hipFree(c->partial2);
  hipFree(c->npercell); hipFree(c->cellflag);
  for (int g = 0; g < NGRIDS; g++) hipFree(c->grid[g]);
  for (int g = 0; g < 8; g++) hipFree(c->gtmp[g]);
  for (int g = 0; g < 6; g++) hipFree(c->kres[g]);
  double *sp[] = {c->sa, c->sb, c->sc, c->sd, c->se, c->sf, c->su, c->sv, c->sw, c->sbs};
  for (double *p : sp) hipFree(p);
  for (int g = 0; g < 8; g++) hipFree(c->sder[g]);
  hipFree(c->atop); hipFree(c->abot); hipFree(c->btop); hipFree(c->bbot);
  hipFree(c->ubar); hipFree(c->vbar);
  hipFree(c->kx); hipFree(c->ky); hipFree(c->k2); hipFree(c->filt); hipFree(c->coords);
  hipFree(c->dz_den); hipFree(c->dz_cp);
  hipFree(c->fbuf);
  hipFree(c->cbuf); hipFree(c->tw_x); hipFree(c->tw_y); hipFree(c->tw_x32); hipFree(c->tw_y32);
  hipFree(c->scan_tmp); hipFree(c->blk); hipFree(c->maxblk); hipFree(c->dsc);
  for (int w = 0; w < NTIMERS; w++)
    for (int e = 0; e < 64; e++) {
      if (c->ev[w][e][0]) hipEventDestroy(c->ev[w][e][0]);
      if (c->ev[w][e][1]) hipEventDestroy(c->ev[w][e][1]);
    }
  if (c->sync_ev) hipEventDestroy(c->sync_ev);
  if (c->pin) hipHostFree(c->pin);
  if (c->stream) hipStreamDestroy(c->stream);
  return PMPIC_OK;
}

int create(const pmpic_config *cf, Ctx *c) {
  if (!cf || cf->nx < 2 || cf->ny < 2 || cf->nz < 2 || cf->max_parcels < 1) return PMPIC_ERR_ARG;
  if (cf->precision != 0 && cf->precision != 1) return PMPIC_ERR_ARG;
  if (!check_factors(cf->nx) || !check_factors(cf->ny)) return PMPIC_ERR_FFT_SIZE;
  if (cf->max_parcels > (int64_t)INT32_MAX) return PMPIC_ERR_ARG;
  PMPIC_CHECK(hipSetDevice(cf->device));
  std::vector<double> coords;
  make_geometry(cf, c->g, coords);
  const Geometry &g = c->g;
  c->cap = cf->max_parcels;
  c->n = 0;
  c->time = 0;
  c->dtm = cf->dtm;
  c->dtmax = cf->dtmax;
  c->prefactor = cf->timestep_prefactor;
  c->rkstep = 1;
  c->precision = cf->precision;
  c->device = cf->device;
  c->splitting_enabled = cf->splitting_enabled;
  c->mixing_enabled = cf->mixing_enabled;
  PMPIC_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->ncell = (int64_t)g.ncz * g.ny * g.nx;
  if (c->ncell > (int64_t)INT32_MAX) return PMPIC_ERR_ARG;
  c->gsz = (int64_t)g.nzn * g.ny * g.nx;
  c->ssz = (int64_t)g.nzn * g.sy * g.sx;
  c->csz = (int64_t)g.nzn * ((int64_t)g.ny * (g.nx / 2 + 1) > (int64_t)(g.ny / 2 + 1) * g.nx
                                  ? (int64_t)g.ny * (g.nx / 2 + 1)
                                  : (int64_t)(g.ny / 2 + 1) * g.nx);
  if (c->gsz > (int64_t)INT32_MAX) return PMPIC_ERR_ARG;
  // the FP32-bulk form packs a cell as k << 20 | j << 10 | i
  if (c->precision == 1 && (g.nx > 1024 || g.ny > 1024 || g.ncz > 2047)) return PMPIC_ERR_ARG;
  // ... and addresses its FP32 parcel fields with 32-bit byte offsets
  if (c->precision == 1 && c->cap >= ((int64_t)1 << 30)) return PMPIC_ERR_ARG;
  const int64_t big = c->cap > c->ncell ? c->cap : c->ncell;
  if (c->precision == 1) PMPIC_TRY(b32_alloc(c));
  else PMPIC_TRY(dalloc(&c->par, (int64_t)NFIELDS * c->cap));
  PMPIC_TRY(dalloc(&c->cell, c->cap));
  PMPIC_TRY(dalloc(&c->rank, c->cap));
  PMPIC_TRY(dalloc(&c->sorted, big));
  PMPIC_TRY(dalloc(&c->flag, c->cap));
  PMPIC_TRY(dalloc(&c->list_a, c->cap));
  PMPIC_TRY(dalloc(&c->list_b, c->cap));
  PMPIC_TRY(dalloc(&c->cell_count, c->ncell + 1));
  PMPIC_TRY(dalloc(&c->cell_start, c->ncell + 1));
  // corner partial sums: 8 corners x up to 6 values per cell, FP64 or FP32 by the form
  const int64_t pdoubles = (c->precision == 1 ? 24 : 48) * c->ncell;
  PMPIC_TRY(dalloc(&c->partial, pdoubles));
  if (c->precision == 1) PMPIC_TRY(dalloc(&c->partial2, pdoubles));
  PMPIC_TRY(dalloc(&c->npercell, c->ncell));
  PMPIC_TRY(dalloc(&c->cellflag, c->ncell));
  for (int i = 0; i < NGRIDS; i++) PMPIC_TRY(dalloc(&c->grid[i], c->gsz));
  for (int i = 0; i < 8; i++) PMPIC_TRY(dalloc(&c->gtmp[i], c->gsz));
  for (int i = 0; i < 6; i++) PMPIC_TRY(dalloc(&c->kres[i], c->gsz));
  double **sp[] = {&c->sa, &c->sb, &c->sc, &c->sd, &c->se, &c->sf, &c->su, &c->sv, &c->sw, &c->sbs};
  for (double **p : sp) PMPIC_TRY(dalloc(p, c->ssz));
  for (int i = 0; i < 8; i++) PMPIC_TRY(dalloc(&c->sder[i], c->ssz));
  const int64_t plane = (int64_t)g.sy * g.sx;
  PMPIC_TRY(dalloc(&c->atop, plane));
  PMPIC_TRY(dalloc(&c->abot, plane));
  PMPIC_TRY(dalloc(&c->btop, plane));
  PMPIC_TRY(dalloc(&c->bbot, plane));
  PMPIC_TRY(dalloc(&c->ubar, g.nzn));
  PMPIC_TRY(dalloc(&c->vbar, g.nzn));
  PMPIC_TRY(dalloc(&c->kx, plane));
  PMPIC_TRY(dalloc(&c->ky, plane));
  PMPIC_TRY(dalloc(&c->k2, plane));
  PMPIC_TRY(dalloc(&c->filt, plane));
  PMPIC_TRY(dalloc(&c->dz_den, 2 * (int64_t)g.nzn));
  PMPIC_TRY(dalloc(&c->dz_cp, 2 * (int64_t)g.nzn));
  PMPIC_TRY(dalloc(&c->coords, (int64_t)coords.size()));
  PMPIC_CHECK(hipMemcpy(c->coords, coords.data(), sizeof(double) * coords.size(),
                        hipMemcpyHostToDevice));
  PMPIC_TRY(dalloc(&c->cbuf, 8 * c->csz));
  c->nblk_max = cdiv(big, 256) + 1;
  PMPIC_TRY(dalloc(&c->blk, 20 * c->nblk_max + 64));
  PMPIC_TRY(dalloc(&c->maxblk, 4 * 1024));
  PMPIC_TRY(dalloc(&c->dsc, 1));
  PMPIC_CHECK(hipHostMalloc(&c->pin, 256));
  PMPIC_CHECK(hipEventCreateWithFlags(&c->sync_ev, hipEventDisableTiming));
  for (int w = 0; w < NTIMERS; w++)
    for (int e = 0; e < 64; e++) {
      PMPIC_CHECK(hipEventCreate(&c->ev[w][e][0]));
      PMPIC_CHECK(hipEventCreate(&c->ev[w][e][1]));
    }
  PMPIC_TRY(fft_setup(c));
  PMPIC_TRY(spectral_setup(c));
  PMPIC_TRY(set_dev_scalars(c));
  // PMPIC_GRAPH=1: the FP32-bulk form launches stages 1-3 of a step as a HIP graph
  const char *genv = getenv("PMPIC_GRAPH");
  c->graph_on = (genv && genv[0] == '1') ? 1 : 0;
  // The FP64 form's residual partials of mixing get their own buffer when the device still has
  // room for it (and a margin for the haloswap staging and scan storage allocated later);
  // otherwise, or with PMPIC_LEAN_MIXING=1, mixing sums the kept and the residual parcels in two
  // passes over one buffer, with the same sums.
  if (c->precision == 0) {
    const char *lean = getenv("PMPIC_LEAN_MIXING");
    size_t fr = 0, tot = 0;
    PMPIC_CHECK(hipMemGetInfo(&fr, &tot));
    const size_t want = sizeof(double) * (size_t)pdoubles;
    if (!(lean && lean[0] == '1') && fr > want + ((size_t)8 << 30))
      PMPIC_TRY(dalloc(&c->partial2, pdoubles));
  }
  return PMPIC_OK;
}

// the stage number (1..4) a component acting now belongs to
inline int stage_of(int it) { return (it % 4) + 1; }

}  // namespace

// ---- launch-configuration autotuning (FP32-bulk form) ----
// A configuration: the binning's slot capacity (16 or 32; 0 without a slot table), the binning
// fused into RK stages 1-3 or run as a pass of its own, RK stages 1-3 as a graph or eagerly,
// and the RK stage's workgroup size.  Every configuration gives the same bits, so the tuning
// steps are ordinary steps of the model.  The choice is kept per grid size and occupancy class
// for the life of the process, so later contexts of that size take it without measuring.
namespace {
struct TuneEntry {
  int nx, ny, ncz, capmax, capmean;
  int32_t cfg[4];
  float ms;
};
std::mutex tune_mu;
std::vector<TuneEntry> tune_table;

void tune_apply(Ctx *c, const int32_t cfg[4]) {
  c->b32.capfix = cfg[0];
  c->b32.fused_bin = cfg[1] != 0;
  c->graph_on = cfg[2];
  c->b32.rk_tpb = cfg[3];
  c->b32.captune = false;
  graph_drop(c);
}

void tune_current(const Ctx *c, int32_t cfg[4]) {
  cfg[0] = b32_choose_capc(c, c->n);
  cfg[1] = c->b32.fused_bin ? 1 : 0;
  cfg[2] = c->graph_on;
  cfg[3] = c->b32.rk_tpb;
}
}  // namespace

extern "C" {

const char *pmpic_status_message(int s) {
  switch (s) {
    case PMPIC_OK: return "ok";
    case PMPIC_ERR_OUT_OF_BOX_X: return "parcel out of box (x direction)";
    case PMPIC_ERR_OUT_OF_BOX_Y: return "parcel out of box (y direction)";
    case PMPIC_ERR_OUT_OF_BOX_Z: return "parcel out of box (z direction)";
    case PMPIC_ERR_DELTA: return "delx/dely/delz wrong size";
    case PMPIC_ERR_MAXPARCELS: return "Aborting: maxparcels reached";
    case PMPIC_ERR_ZERO_PIVOT: return "Error in tridiagonal solver: divide by zero";
    case PMPIC_ERR_FFT_SIZE: return "NX and/or NY are the wrong sizes for FFTE";
    case PMPIC_ERR_COUNT: return "parcel number not conserved";
    case PMPIC_ERR_BACKFILL: return "parcel mixing backfill ran out of parcels to move";
    case PMPIC_ERR_ARG: return "invalid argument";
    case PMPIC_ERR_HIP: return "HIP runtime error";
    case PMPIC_ERR_NOMEM: return "host allocation failed";
    case PMPIC_ERR_BC: return "laplinv needs exactly one boundary condition";
    default: return "unknown status";
  }
}

int pmpic_create(const pmpic_config *cfg, pmpic_ctx **out) {
  if (!out) return PMPIC_ERR_ARG;
  *out = nullptr;
  pmpic_ctx *x = (pmpic_ctx *)calloc(1, sizeof(pmpic_ctx));
  if (!x) return PMPIC_ERR_NOMEM;
  int st = create(cfg, &x->c);
  if (st != PMPIC_OK) {
    free_all(&x->c);
    free(x);
    return st;
  }
  *out = x;
  return PMPIC_OK;
}

int pmpic_destroy(pmpic_ctx *ctx) {
  if (!ctx) return PMPIC_OK;
  hipStreamSynchronize(ctx->c.stream);
  free_all(&ctx->c);
  free(ctx);
  return PMPIC_OK;
}

int pmpic_set_parcels(pmpic_ctx *ctx, int64_t n, const double *const *fields) {
  if (!ctx || !fields || n < 0 || n > ctx->c.cap) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  for (int f = 0; f < 17; f++)
    if (!fields[f]) return PMPIC_ERR_ARG;
  c->split_flags_fresh = 0;
  if (c->precision == 1) {
    PMPIC_TRY(b32_set_parcels(c, n, fields));
    c->n = n;
    PMPIC_TRY(b32_publish_n(c));
    return take_status(c);
  }
  for (int f = 0; f < 17; f++) {
    if (n > 0)
      PMPIC_CHECK(hipMemcpyAsync(c->par + (int64_t)f * c->cap, fields[f], sizeof(double) * n,
                                 hipMemcpyHostToDevice, c->stream));
  }
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  c->n = n;
  return PMPIC_OK;
}

int pmpic_get_parcels(pmpic_ctx *ctx, int64_t *n, double *const *fields) {
  if (!ctx) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  if (n) *n = c->n;
  if (fields && c->precision == 1) return b32_get_parcels(c, fields);
  if (fields)
    for (int f = 0; f < 17; f++)
      if (fields[f] && c->n > 0)
        PMPIC_CHECK(hipMemcpyAsync(fields[f], c->par + (int64_t)f * c->cap, sizeof(double) * c->n,
                                   hipMemcpyDeviceToHost, c->stream));
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  return PMPIC_OK;
}

int pmpic_set_rk_arrays(pmpic_ctx *ctx, const double *const *fields) {
  if (!ctx || !fields || ctx->c.precision != 0) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  for (int f = 0; f < 12; f++) {
    if (!fields[f]) return PMPIC_ERR_ARG;
    if (c->n > 0)
      PMPIC_CHECK(hipMemcpyAsync(c->par + (int64_t)(FXO + f) * c->cap, fields[f],
                                 sizeof(double) * c->n, hipMemcpyHostToDevice, c->stream));
  }
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  return PMPIC_OK;
}

int pmpic_get_rk_arrays(pmpic_ctx *ctx, double *const *fields) {
  if (!ctx || !fields || ctx->c.precision != 0) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  for (int f = 0; f < 12; f++)
    if (fields[f] && c->n > 0)
      PMPIC_CHECK(hipMemcpyAsync(fields[f], c->par + (int64_t)(FXO + f) * c->cap,
                                 sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  return PMPIC_OK;
}

int pmpic_get_state(pmpic_ctx *ctx, int64_t *n, double *time, double *dtm) {
  if (!ctx) return PMPIC_ERR_ARG;
  if (n) *n = ctx->c.n;
  if (time) *time = ctx->c.time;
  if (dtm) *dtm = ctx->c.dtm;
  return PMPIC_OK;
}

int pmpic_set_time(pmpic_ctx *ctx, double time, double dtm) {
  if (!ctx) return PMPIC_ERR_ARG;
  ctx->c.time = time;
  ctx->c.dtm = dtm;
  return set_dev_scalars(&ctx->c);
}

// vort2vel: par2grid of p, q, r (and, fused into the same pass, the total buoyancy that
// vorticity_tendency grids from the same cached weights), the spectral solve, and at stage 1
// the grid maxima for the timestep limits
int pmpic_vort2vel(pmpic_ctx *ctx) {
  if (!ctx) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  const int stage = stage_of(c->it_v2v);
  PMPIC_TRY(timing_begin(c, T_VORT2VEL));
  if (c->precision == 1) {
    PMPIC_TRY(b32_par2grid(c));
    PMPIC_TRY(vort2vel_spectral(c, stage));  // also writes the FP32 u, v, w
  } else {
    PMPIC_TRY(par2grid_fused(c, c->n));
    PMPIC_TRY(vort2vel_spectral(c, stage));
  }
  PMPIC_TRY(timing_end(c, T_VORT2VEL));
  c->it_v2v++;
  return PMPIC_OK;
}

int pmpic_vorticity_tendency(pmpic_ctx *ctx) {
  if (!ctx) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  const int stage = stage_of(c->it_vt);
  PMPIC_TRY(timing_begin(c, T_VORT_TEND));
  PMPIC_TRY(vorticity_tendency_spectral(c, stage));  // and the FP32 tendencies
  if (stage == 1) {
    PMPIC_TRY(grid_max_reduce(c));
  }
  PMPIC_TRY(timing_end(c, T_VORT_TEND));
  c->it_vt++;
  return PMPIC_OK;
}

int pmpic_rk4_integrator(pmpic_ctx *ctx) {
  if (!ctx) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  const int stage = c->rkstep;
  PMPIC_TRY(timing_begin(c, T_RK4));
  if (c->precision == 1) PMPIC_TRY(b32_rk_stage(c, stage, c->splitting_enabled));
  else PMPIC_TRY(rk_stage_update(c, stage, c->n, c->splitting_enabled));
  PMPIC_TRY(timing_end(c, T_RK4));
  c->rkstep = (c->rkstep < 4) ? c->rkstep + 1 : 1;
  if (c->defer_status) return PMPIC_OK;  // pmpic_step checks once per step
  int st = take_status(c);
  c->dtm = c->hsc.dtm;
  return st;
}

int pmpic_parcel_splitting(pmpic_ctx *ctx) {
  if (!ctx) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  int st = PMPIC_OK;
  if (c->it_split % 4 == 3) {
    const int64_t n0 = c->n;
    PMPIC_TRY(timing_begin(c, T_SPLIT));
    st = (c->precision == 1) ? b32_split(c) : parcel_split(c);
    if (st == PMPIC_OK) st = timing_end(c, T_SPLIT);
    c->last_split = c->n - n0;
    if (st == PMPIC_OK && !c->defer_status) st = take_status(c);
  }
  c->it_split++;
  return st;
}

int pmpic_parcel_mixing(pmpic_ctx *ctx) {
  if (!ctx) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  int st = PMPIC_OK;
  if (c->it_mix % 4 == 3) {
    PMPIC_TRY(timing_begin(c, T_MIX));
    st = (c->precision == 1) ? b32_mix(c) : parcel_mix(c);
    if (st == PMPIC_OK) st = timing_end(c, T_MIX);
    if (st == PMPIC_OK && !c->defer_status) st = take_status(c);
  }
  c->it_mix++;
  return st;
}

// one RK stage of a timestep: the components in the order of the reference's model loop
static int one_stage(pmpic_ctx *ctx) {
  Ctx *c = &ctx->c;
  PMPIC_TRY(pmpic_vort2vel(ctx));
  PMPIC_TRY(pmpic_vorticity_tendency(ctx));
  PMPIC_TRY(pmpic_rk4_integrator(ctx));
  if (c->splitting_enabled) PMPIC_TRY(pmpic_parcel_splitting(ctx));
  else c->it_split++;
  if (c->mixing_enabled) PMPIC_TRY(pmpic_parcel_mixing(ctx));
  else c->it_mix++;
  return PMPIC_OK;
}

// The FP32-bulk form's stages 1-3 as a HIP graph.  Their launches depend on the host only
// through the slot capacity of the binning and the grid sizes of the per-parcel launches; the
// kernels read the parcel count from device memory, so one capture serves every step whose
// count lies within the captured bound (and not far below it) and whose capacity is the
// captured one.  The graph
// holds kernel launches only.  Stage 4 (splitting and mixing, whose launches depend on counts
// read back) runs eagerly.
static bool graph_eligible(const Ctx *c) {
  return c->precision == 1 && c->graph_on && c->timing == 0 && !c->b32.captune &&
         c->eager_steps > 0 && c->n > 0 && c->rkstep == 1 && c->it_v2v % 4 == 0 &&
         c->it_vt % 4 == 0 && c->it_split % 4 == 0 && c->it_mix % 4 == 0;
}

// the parcel count a capture covers: the count plus 1/16 and 4096 (rounded to 4096), so that
// the growth by splitting of a few steps stays inside it; at most the capacity
static int64_t graph_bound(const Ctx *c) {
  int64_t bound = c->n + c->n / 16 + 4096;
  bound = (bound + 4095) / 4096 * 4096;
  return bound > c->cap ? c->cap : bound;
}

static int graph_capture(pmpic_ctx *ctx, int capc) {
  Ctx *c = &ctx->c;
  const int64_t bound = graph_bound(c);
  PMPIC_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->graph_n = bound;
  int st = PMPIC_OK;
  for (int r = 0; r < 3 && st == PMPIC_OK; r++) st = one_stage(ctx);
  c->graph_n = 0;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &g);
  if (st == PMPIC_OK && ec != hipSuccess) st = PMPIC_ERR_HIP;
  hipGraphExec_t x = nullptr;
  if (st == PMPIC_OK && hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess)
    st = PMPIC_ERR_HIP;
  if (st != PMPIC_OK) {
    if (g) hipGraphDestroy(g);
    return st;
  }
  c->g123 = g;
  c->g123x = x;
  c->g123_n = bound;
  c->g123_capc = capc;
  c->g123_prebinned = c->b32.prebinned;
  c->g123_captures++;
  return PMPIC_OK;
}

// stages 1-3 of a step by the graph (captured first where none fits); the host counters and
// state advance as the eager calls advance them
static int graph_stages123(pmpic_ctx *ctx) {
  Ctx *c = &ctx->c;
  const int capc = b32_choose_capc(c, c->n);
  // a new capture for another capacity, a count past the bound, or a bound fallen below 7/8
  // of the captured one (the launches' idle groups kept few)
  if (c->g123x && (capc != c->g123_capc || c->n > c->g123_n || 8 * graph_bound(c) < 7 * c->g123_n))
    graph_drop(c);
  b32_graph_reuse(c);
  if (!c->g123x) {
    // the capture itself advanced the counters; the launch below does the work
    PMPIC_TRY(graph_capture(ctx, capc));
  } else {
    c->it_v2v += 3;
    c->it_vt += 3;
    c->it_split += 3;
    c->it_mix += 3;
    c->rkstep = 4;
    c->split_flags_fresh = 0;
    c->b32.capc = capc;
    c->b32.kept_pending[1] = c->b32.kept_pending[3] = false;
    c->b32.kept_pending[4] = c->b32.kept_pending[5] = false;
    for (int i = 0; i < 4; i++) c->b32.p2g_pending[i] = true;
    c->b32.prebinned = c->g123_prebinned;
    c->b32.p2g_reuse = false;  // consumed by stage 1's par2grid, as the eager call does
    if (c->b32.prebinned) c->b32.capc = capc;
  }
  PMPIC_CHECK(hipGraphLaunch(c->g123x, c->stream));
  c->g123_replays++;
  return PMPIC_OK;
}

// one timestep with the kernels' status checked once at its end (the components called one
// by one check after each call); no kernel indexes memory from data a failed kernel left, so
// the later launches of a failing step are harmless
static int step_once(pmpic_ctx *ctx) {
  Ctx *c = &ctx->c;
  if (graph_eligible(c)) {
    PMPIC_TRY(graph_stages123(ctx));
    return one_stage(ctx);
  }
  for (int r = 0; r < 4; r++) PMPIC_TRY(one_stage(ctx));
  c->eager_steps++;
  return PMPIC_OK;
}

int pmpic_step(pmpic_ctx *ctx, int32_t nsteps) {
  if (!ctx || nsteps < 0) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  for (int s = 0; s < nsteps; s++) {
    c->defer_status = 1;
    int st = step_once(ctx);
    c->defer_status = 0;
    const int ks = take_status(c);
    if (st == PMPIC_OK) st = ks;
    if (st != PMPIC_OK) return st;
    c->dtm = c->hsc.dtm;
    c->time = c->time + c->dtm;
  }
  return PMPIC_OK;
}

int pmpic_step_concurrent(pmpic_ctx *const *ctxs, int32_t count, int32_t nsteps) {
  if (!ctxs || count < 0 || nsteps < 0) return PMPIC_ERR_ARG;
  for (int32_t i = 0; i < count; i++) {
    if (!ctxs[i]) return PMPIC_ERR_ARG;
    for (int32_t j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) return PMPIC_ERR_ARG;
  }
  std::vector<int> st(count, PMPIC_OK);
  std::vector<std::thread> th;
  th.reserve(count);
  for (int32_t i = 0; i < count; i++)
    th.emplace_back([&, i]() {
      if (hipSetDevice(ctxs[i]->c.device) != hipSuccess) {
        st[i] = PMPIC_ERR_HIP;
        return;
      }
      st[i] = pmpic_step(ctxs[i], nsteps);
    });
  for (auto &t : th) t.join();
  for (int32_t i = 0; i < count; i++)
    if (st[i] != PMPIC_OK) return st[i];
  return PMPIC_OK;
}

int pmpic_parcel_haloswap(pmpic_ctx *ctx) {
  if (!ctx || ctx->c.precision != 0) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  PMPIC_TRY(parcel_haloswap(c, c->n, 1));
  return take_status(c);
}

int pmpic_cache_parcel_interp_weights(pmpic_ctx *ctx, int32_t *is, int32_t *js, int32_t *ks,
                                      double *delxs, double *delys, double *delzs) {
  if (!ctx || ctx->c.precision != 0) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  const Geometry &g = c->g;
  std::vector<double> x, y, z;
  PMPIC_TRY(download(c, x, c->par + FX * c->cap, c->n));
  PMPIC_TRY(download(c, y, c->par + FY * c->cap, c->n));
  PMPIC_TRY(download(c, z, c->par + FZ * c->cap, c->n));
  // the same arithmetic the kernels use (locate), evaluated here for the host copy
  for (int64_t p = 0; p < c->n; p++) {
    const double ax = x[p] - g.xmin_h, ay = y[p] - g.ymin_h, az = z[p] - g.zmin_h;
    const int i = (int)floor(ax / g.dx) + 1, j = (int)floor(ay / g.dy) + 1,
              k = (int)floor(az / g.dz) + 1;
    if (i < 3 || i > g.nx + 3) return PMPIC_ERR_OUT_OF_BOX_X;
    if (j < 3 || j > g.ny + 3) return PMPIC_ERR_OUT_OF_BOX_Y;
    if (k < 1 || k >= g.nzn) return PMPIC_ERR_OUT_OF_BOX_Z;
    const double dlx = (ax - (double)(i - 1) * g.dx) / g.dx;
    const double dly = (ay - (double)(j - 1) * g.dy) / g.dy;
    const double dlz = (az - (double)(k - 1) * g.dz) / g.dz;
    if (dlx > 1. || dlx < 0 || dly > 1. || dly < 0 || dlz > 1. || dlz < 0) return PMPIC_ERR_DELTA;
    if (is) is[p] = i;
    if (js) js[p] = j;
    if (ks) ks[p] = k;
    if (delxs) delxs[p] = dlx;
    if (delys) delys[p] = dly;
    if (delzs) delzs[p] = dlz;
  }
  return PMPIC_OK;
}

int pmpic_par2grid(pmpic_ctx *ctx, const double *var, double *grid) {
  if (!ctx || !var || !grid || ctx->c.precision != 0) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  PMPIC_TRY(ensure_fbuf(c));
  if (c->n > 0)
    PMPIC_CHECK(hipMemcpyAsync(c->fbuf, var, sizeof(double) * c->n, hipMemcpyHostToDevice,
                               c->stream));
  double *dg = c->fbuf + (c->cap > c->gsz ? c->cap : c->gsz);
  PMPIC_TRY(par2grid_single(c, c->n, c->fbuf, dg));
  PMPIC_TRY(take_status(c));
  std::vector<double> h;
  PMPIC_TRY(download(c, h, dg, c->gsz));
  grid_from_device_layout(c->g, h, grid);
  return PMPIC_OK;
}

int pmpic_grid2par(pmpic_ctx *ctx, const double *grid, double *var) {
  if (!ctx || !var || !grid || ctx->c.precision != 0) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  PMPIC_TRY(ensure_fbuf(c));
  std::vector<double> h;
  grid_to_device_layout(c->g, grid, h);
  double *dg = c->fbuf + (c->cap > c->gsz ? c->cap : c->gsz);
  PMPIC_TRY(upload(c, dg, h));
  PMPIC_TRY(grid2par_single(c, c->n, dg, c->fbuf));
  PMPIC_TRY(take_status(c));
  if (c->n > 0)
    PMPIC_CHECK(hipMemcpyAsync(var, c->fbuf, sizeof(double) * c->n, hipMemcpyDeviceToHost,
                               c->stream));
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  return PMPIC_OK;
}

int pmpic_get_grid(pmpic_ctx *ctx, int32_t id, double *grid) {
  if (!ctx || !grid || id < 0 || id >= NGRIDS) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  std::vector<double> h;
  const bool vel = id >= GU && id <= GW, tend = id >= GDP && id <= GDR;
  if (c->precision == 1 && (vel || tend)) {
    // the FP32-bulk form keeps u, v, w and the tendencies only in the FP32 grids its gathers
    // read (b32.grid: U, V, W, DP, DQ, DR)
    std::vector<float> hf(c->gsz);
    const float *src32 = c->b32.grid[vel ? id - GU : 3 + (id - GDP)];
    PMPIC_CHECK(hipMemcpyAsync(hf.data(), src32, sizeof(float) * c->gsz,
                               hipMemcpyDeviceToHost, c->stream));
    PMPIC_CHECK(hipStreamSynchronize(c->stream));
    h.assign(hf.begin(), hf.end());
  } else {
    if (c->precision == 1) PMPIC_TRY(b32_export_grid(c, id));
    PMPIC_TRY(download(c, h, c->grid[id], c->gsz));
  }
  grid_from_device_layout(c->g, h, grid);
  return PMPIC_OK;
}

int pmpic_writebinary_grids(pmpic_ctx *ctx) {
  if (!ctx) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  if (c->precision == 1) PMPIC_TRY(b32_dump_grids(c));
  else PMPIC_TRY(dump_grids_f64(c, c->n));
  PMPIC_TRY(hgliq_grid(c));
  return take_status(c);
}

int pmpic_get_spectral(pmpic_ctx *ctx, int32_t id, double *spec) {
  if (!ctx || !spec || id < 0 || id > 2) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  double *src = id == 0 ? c->su : id == 1 ? c->sv : c->sw;
  std::vector<double> h;
  if (c->precision == 1) {
    // the FP32-bulk form's spectral arrays hold FP32 values; widened exactly
    std::vector<float> h32((size_t)c->ssz);
    PMPIC_CHECK(hipMemcpyAsync(h32.data(), src, sizeof(float) * (size_t)c->ssz,
                               hipMemcpyDeviceToHost, c->stream));
    PMPIC_CHECK(hipStreamSynchronize(c->stream));
    h.assign(h32.begin(), h32.end());
  } else {
    PMPIC_TRY(download(c, h, src, c->ssz));
  }
  spec_from_device_layout(c->g, h, spec);
  return PMPIC_OK;
}

int pmpic_fft_forward(pmpic_ctx *ctx, const double *grid, double *spec) {
  if (!ctx || !grid || !spec) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  std::vector<double> h;
  grid_to_device_layout(c->g, grid, h);
  PMPIC_TRY(upload(c, c->gtmp[0], h));
  const double *src[1] = {c->gtmp[0]};
  double *dst[1] = {c->sder[0]};
  PMPIC_TRY(fft_forward(c, 1, src, dst));
  PMPIC_TRY(download(c, h, c->sder[0], c->ssz));
  spec_from_device_layout(c->g, h, spec);
  return PMPIC_OK;
}

int pmpic_fft_backward(pmpic_ctx *ctx, const double *spec, double *grid) {
  if (!ctx || !grid || !spec) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  std::vector<double> h;
  spec_to_device_layout(c->g, spec, h);
  PMPIC_TRY(upload(c, c->sder[0], h));
  const double *src[1] = {c->sder[0]};
  double *dst[1] = {c->gtmp[0]};
  PMPIC_TRY(fft_backward(c, 1, src, dst));
  PMPIC_TRY(download(c, h, c->gtmp[0], c->gsz));
  grid_from_device_layout(c->g, h, grid);
  return PMPIC_OK;
}

static int spectral_unary(pmpic_ctx *ctx, const double *in, double *out, int which,
                          const double *bot, const double *top) {
  if (!ctx || !in || !out) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  std::vector<double> h, hb, ht;
  spec_to_device_layout(c->g, in, h);
  PMPIC_TRY(upload(c, c->sder[0], h));
  switch (which) {
    case 0: PMPIC_TRY(op_diffx(c, c->sder[0], c->sder[1])); break;
    case 1: PMPIC_TRY(op_diffy(c, c->sder[0], c->sder[1])); break;
    case 2: PMPIC_TRY(op_filter(c, c->sder[0], c->sder[1])); break;
    default: {
      const double *db = nullptr, *dt = nullptr;
      if (bot && top) {
        col_to_device_layout(c->g, bot, hb);
        col_to_device_layout(c->g, top, ht);
        PMPIC_TRY(upload(c, c->atop, ht));
        PMPIC_TRY(upload(c, c->abot, hb));
        db = c->abot;
        dt = c->atop;
      }
      PMPIC_TRY(op_diffz(c, c->sder[0], c->sder[1], db, dt));
    }
  }
  PMPIC_TRY(take_status(c));
  PMPIC_TRY(download(c, h, c->sder[1], c->ssz));
  spec_from_device_layout(c->g, h, out);
  return PMPIC_OK;
}

int pmpic_diffx(pmpic_ctx *ctx, const double *in, double *out) { return spectral_unary(ctx, in, out, 0, nullptr, nullptr); }
int pmpic_diffy(pmpic_ctx *ctx, const double *in, double *out) { return spectral_unary(ctx, in, out, 1, nullptr, nullptr); }
int pmpic_spectral_filter(pmpic_ctx *ctx, const double *in, double *out) { return spectral_unary(ctx, in, out, 2, nullptr, nullptr); }
int pmpic_diffz(pmpic_ctx *ctx, const double *in, double *out, const double *bot, const double *top) {
  return spectral_unary(ctx, in, out, 3, bot, top);
}

int pmpic_laplinv(pmpic_ctx *ctx, double *f, int32_t df0) {
  if (!ctx || !f) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  std::vector<double> h;
  spec_to_device_layout(c->g, f, h);
  PMPIC_TRY(upload(c, c->sder[0], h));
  PMPIC_TRY(op_laplinv(c, c->sder[0], df0 ? 1 : 0));
  PMPIC_TRY(take_status(c));
  PMPIC_TRY(download(c, h, c->sder[0], c->ssz));
  spec_from_device_layout(c->g, h, f);
  return PMPIC_OK;
}

int pmpic_conservation_sums(pmpic_ctx *ctx, double out6[6]) {
  if (!ctx || !out6) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  if (c->precision == 1) return b32_conservation_sums(c, out6);
  std::vector<double> v, f;
  PMPIC_TRY(download(c, v, c->par + FVOL * c->cap, c->n));
  const int fl[6] = {-1, FB, FH, FP, FQ, FR};
  for (int k = 0; k < 6; k++) {
    if (fl[k] >= 0) PMPIC_TRY(download(c, f, c->par + (int64_t)fl[k] * c->cap, c->n));
    double s = 0.0, comp = 0.0;  // Neumaier summation in parcel order
    for (int64_t p = 0; p < c->n; p++) {
      const double term = fl[k] < 0 ? v[p] : v[p] * f[p];
      const double t = s + term;
      if (fabs(s) >= fabs(term)) comp += (s - t) + term;
      else comp += (term - t) + s;
      s = t;
    }
    out6[k] = s + comp;
  }
  return PMPIC_OK;
}

int pmpic_last_counts(pmpic_ctx *ctx, int64_t out3[3]) {
  if (!ctx || !out3) return PMPIC_ERR_ARG;
  out3[0] = ctx->c.last_split;
  out3[1] = ctx->c.last_removed;
  out3[2] = ctx->c.last_added;
  return PMPIC_OK;
}

int pmpic_timing_enable(pmpic_ctx *ctx, int32_t enable) {
  if (!ctx) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  PMPIC_TRY(timing_harvest(c));
  c->timing = (enable == 2) ? 2 : (enable ? 1 : 0);
  for (int w = 0; w < NTIMERS; w++) {
    c->t_ms[w] = 0;
    c->t_calls[w] = 0;
    c->t_max[w] = 0;
    c->t_sq[w] = 0;
  }
  return PMPIC_OK;
}

const char *pmpic_timing_name(int32_t which) {
  // the kernels, then the components under the reference's timer region names
  // (vort2vel.F90:88, vorticity_tendency.F90:97, rk4_integrator.F90:71,
  // parcel_splitting.F90:40, parcel_mixing.F90:49)
  static const char *names[NTIMERS] = {"par2grid_cells", "grid2par_rk_stage", "vort2vel",
                                       "vort_tend", "RK4_all_steps", "splitting", "mixing"};
  return (which >= 0 && which < NTIMERS) ? names[which] : nullptr;
}

int pmpic_timing_stats(pmpic_ctx *ctx, int32_t which, double *total_ms, int64_t *calls,
                       double *max_ms, double *sumsq_ms2) {
  if (!ctx || which < 0 || which >= NTIMERS) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  PMPIC_TRY(timing_harvest(c));
  if (total_ms) *total_ms = c->t_ms[which];
  if (calls) *calls = c->t_calls[which];
  if (max_ms) *max_ms = c->t_max[which];
  if (sumsq_ms2) *sumsq_ms2 = c->t_sq[which];
  return PMPIC_OK;
}


// every launch configuration timed over `steps` steps (after one step to settle), the fastest
// kept for this context and for later ones of its grid size
int pmpic_autotune(pmpic_ctx *ctx, int32_t steps, int32_t *config, float *ms_per_step,
                   int32_t *measured) {
  if (!ctx || steps < 1 || ctx->c.precision != 1 || ctx->c.n == 0) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  const int capmean = b32_mean_capc(c, c->n);
  auto same = [&](const TuneEntry &e) {
    return e.nx == c->g.nx && e.ny == c->g.ny && e.ncz == c->g.ncz && e.capmax == c->b32.capmax &&
           e.capmean == capmean;
  };
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    for (const TuneEntry &e : tune_table)
      if (same(e)) {
        tune_apply(c, e.cfg);
        if (config) memcpy(config, e.cfg, sizeof(e.cfg));
        if (ms_per_step) *ms_per_step = e.ms;
        if (measured) *measured = 0;
        return PMPIC_OK;
      }
  }
  std::vector<int> caps;
  if (c->b32.capmax == 0) caps.push_back(0);
  if (c->b32.capmax >= 16) caps.push_back(16);
  if (c->b32.capmax >= 32) caps.push_back(32);
  const int graphs = c->timing == 0 ? 2 : 1;  // a graph needs the timers off
  hipEvent_t e0, e1;
  PMPIC_CHECK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return PMPIC_ERR_HIP;
  }
  std::vector<std::array<int32_t, 4>> cand;
  for (int cap : caps)
    for (int fused = 0; fused < 2; fused++)
      for (int graph = 0; graph < graphs; graph++)
        for (int tpb = 128; tpb <= 256; tpb *= 2) cand.push_back({cap, fused, graph, tpb});
  // The candidates run in order and then in reverse, and each is scored by the mean of its two
  // timings: the model state (and with it the speed of a step) drifts over the tuning steps,
  // and a linear drift cancels in that mean.
  const int nc = (int)cand.size();
  std::vector<float> total(nc, 0.f);
  int st = PMPIC_OK;
  for (int r = 0; r < 2 * nc && st == PMPIC_OK; r++) {
    const int i = r < nc ? r : 2 * nc - 1 - r;
    tune_apply(c, cand[i].data());
    // one step to settle (a capture, a new binning capacity), then the timed steps
    st = pmpic_step(ctx, 1);
    float ms = 0.f;
    if (st == PMPIC_OK && hipEventRecord(e0, c->stream) != hipSuccess) st = PMPIC_ERR_HIP;
    if (st == PMPIC_OK) st = pmpic_step(ctx, steps);
    if (st == PMPIC_OK && (hipEventRecord(e1, c->stream) != hipSuccess ||
                           hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      st = PMPIC_ERR_HIP;
    total[i] += ms;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (st != PMPIC_OK) return st;
  int bi = 0;
  for (int i = 1; i < nc; i++)
    if (total[i] < total[bi]) bi = i;
  int32_t best[4];
  memcpy(best, cand[bi].data(), sizeof(best));
  const float best_ms = total[bi] / (2.f * steps);
  tune_apply(c, best);
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    bool have = false;
    for (const TuneEntry &e : tune_table) have = have || same(e);
    if (!have) {
      TuneEntry e{c->g.nx, c->g.ny, c->g.ncz, c->b32.capmax, capmean, {0, 0, 0, 0}, best_ms};
      memcpy(e.cfg, best, sizeof(best));
      tune_table.push_back(e);
    }
  }
  if (config) memcpy(config, best, sizeof(best));
  if (ms_per_step) *ms_per_step = best_ms;
  if (measured) *measured = 1;
  return PMPIC_OK;
}

int pmpic_set_launch_config(pmpic_ctx *ctx, const int32_t *config) {
  if (!ctx || !config || ctx->c.precision != 1) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  const int32_t cap = config[0];
  if (!(cap == 0 || cap == 16 || cap == 32) || cap > c->b32.capmax ||
      (config[1] != 0 && config[1] != 1) || (config[2] != 0 && config[2] != 1) ||
      (config[3] != 128 && config[3] != 256))
    return PMPIC_ERR_ARG;
  tune_apply(c, config);
  return PMPIC_OK;
}

int pmpic_launch_config(pmpic_ctx *ctx, int32_t *config) {
  if (!ctx || !config || ctx->c.precision != 1) return PMPIC_ERR_ARG;
  tune_current(&ctx->c, config);
  return PMPIC_OK;
}

int pmpic_graph_stats(pmpic_ctx *ctx, int64_t *captures, int64_t *launches) {
  if (!ctx) return PMPIC_ERR_ARG;
  if (captures) *captures = ctx->c.g123_captures;
  if (launches) *launches = ctx->c.g123_replays;
  return PMPIC_OK;
}

int pmpic_timing_read(pmpic_ctx *ctx, int32_t which, double *total_ms, int64_t *launches) {
  if (!ctx || which < 0 || which >= NTIMERS) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  PMPIC_TRY(timing_harvest(c));
  if (total_ms) *total_ms = c->t_ms[which];
  if (launches) *launches = c->t_calls[which];
  return PMPIC_OK;
}

}  // extern "C"
