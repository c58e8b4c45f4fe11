// Parcel <-> grid interpolation on MI355X: deterministic cell-binned par2grid, grid2par.
//
// Reference: model_core/src/parcels/parcel_interpolation.F90:186-481 (cache weights, grid2par,
// par2grid) and model_core/src/communication/MPIC_Haloswap.F90:95-307 (halo fold / fill, which
// on one GPU is periodic index arithmetic), components/parcel_mixing/src/parcel_mixing.F90:440-624
// (kept / residual grids).
//
// par2grid is a scatter; here it is a gather over a cell binning so that no floating-point
// atomics are used and every node sum has a fixed order:
//   1. bin:   each parcel finds its cell and takes a slot in the cell's bucket (integer atomics
//             only), an exclusive scan turns bucket sizes into offsets, parcel indices are
//             scattered into buckets and each bucket is sorted by parcel index;
//   2. cell:  one thread per cell walks its parcels in index order and accumulates the 8 corner
//             weights (and weight * value) into per-cell corner partial sums;
//   3. node:  one thread per node adds the matching corner partials of its 8 adjacent cells
//             (periodic in x, y) in a fixed order and normalises by the summed weights.
// The result differs from the reference's sequential sum only by the order of additions.
#include "common.h"

#include <hipcub/hipcub.hpp>

namespace pmpic {

namespace {

constexpr int TPB = 256;

__global__ void k_fill_i32(int32_t *a, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_fill_f64(double *a, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_bin_count(Geometry g, const double *__restrict__ X, const double *__restrict__ Y,
                            const double *__restrict__ Z, int64_t n, int32_t *__restrict__ cell,
                            int32_t *__restrict__ rank, int32_t *__restrict__ count,
                            DevScalars *sc) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i0, j0, k0;
  double dx_, dy_, dz_;
  int st = locate(g, X[p], Y[p], Z[p], i0, j0, k0, dx_, dy_, dz_);
  if (st != PMPIC_OK) {
    raise_status(sc, st);
    cell[p] = -1;
    return;
  }
  int32_t c = (int32_t)(((int64_t)k0 * g.ny + j0) * g.nx + i0);
  cell[p] = c;
  rank[p] = atomicAdd(&count[c], 1);
}

__global__ void k_bin_scatter(const int32_t *__restrict__ cell, const int32_t *__restrict__ rank,
                              const int32_t *__restrict__ start, int64_t n,
                              int32_t *__restrict__ sorted) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t c = cell[p];
  if (c < 0) return;
  sorted[start[c] + rank[p]] = (int32_t)p;
}

// insertion sort of each bucket by parcel index (buckets hold a few tens of parcels)
__global__ void k_bin_sort(const int32_t *__restrict__ start, int64_t ncell,
                           int32_t *__restrict__ sorted) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int32_t b = start[c], e = start[c + 1];
  for (int32_t a = b + 1; a < e; a++) {
    int32_t v = sorted[a];
    int32_t q = a - 1;
    while (q >= b && sorted[q] > v) {
      sorted[q + 1] = sorted[q];
      q--;
    }
    sorted[q + 1] = v;
  }
}

// value sources for the per-cell accumulation
enum AccMode { ACC_P2G = 0, ACC_SINGLE = 1, ACC_MIX_KEEP = 2, ACC_MIX_RES = 3 };

template <int MODE, int NV>
__global__ void k_cell_accum(Geometry g, const double *__restrict__ par, int64_t cap,
                             const double *__restrict__ var, const int32_t *__restrict__ start,
                             const int32_t *__restrict__ sorted, int64_t ncell,
                             double *__restrict__ partial, int32_t *__restrict__ npercell) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const double *X = par + FX * cap, *Y = par + FY * cap, *Z = par + FZ * cap;
  const double *VOL = par + FVOL * cap;
  double acc[8][NV + 1];
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int f = 0; f <= NV; f++) acc[q][f] = 0.0;
  int count = 0;
  const int32_t b = start[c], e = start[c + 1];
  for (int32_t s = b; s < e; s++) {
    const int64_t p = sorted[s];
    const double v = VOL[p];
    if (MODE == ACC_MIX_KEEP && v < g.vmin) continue;
    if (MODE == ACC_MIX_RES && !(v < g.vmin)) continue;
    int i0, j0, k0;
    double delx, dely, delz;
    locate(g, X[p], Y[p], Z[p], i0, j0, k0, delx, dely, delz);
    double val[NV > 0 ? NV : 1];
    if (MODE == ACC_P2G) {
      val[0] = par[FP * cap + p];
      val[1] = par[FQ * cap + p];
      val[2] = par[FR * cap + p];
      // total buoyancy, vorticity_tendency.F90:121-126
      const double hl = par[FH * cap + p] - exp(-Z[p]);
      val[3] = par[FB * cap + p] + 12.5 * (hl > 0.0 ? hl : 0.0);
    } else if (MODE == ACC_SINGLE) {
      val[0] = var[p];
    } else {
      val[0] = par[FB * cap + p];
      val[1] = par[FH * cap + p];
      val[2] = par[FP * cap + p];
      val[3] = par[FQ * cap + p];
      val[4] = par[FR * cap + p];
    }
    double w[8];
    corner_weights(delx, dely, delz, v, w);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      acc[q][0] += w[q];
#pragma unroll
      for (int f = 0; f < NV; f++) acc[q][f + 1] += w[q] * val[f];
    }
    count++;
  }
#pragma unroll
  for (int q = 0; q < 8; q++)
#pragma unroll
    for (int f = 0; f <= NV; f++) partial[(int64_t)(q * (NV + 1) + f) * ncell + c] = acc[q][f];
  if (MODE == ACC_MIX_KEEP) npercell[c] = count;
}

// node sums: NORM = divide by summed weights (par2grid, parcel_interpolation.F90:463-469);
// otherwise raw sums with the mixing z-boundary doubling (parcel_mixing.F90:141-152).
struct OutTable {
  double *p[6];
};

template <int NV, bool NORM, bool DBL = true>
__global__ void k_node_sum(Geometry g, const double *__restrict__ partial, int64_t ncell,
                           OutTable out) {
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode) return;
  const int i = (int)(t % g.nx);
  const int j = (int)((t / g.nx) % g.ny);
  const int k = (int)(t / ((int64_t)g.nx * g.ny));
  double s[NV + 1];
#pragma unroll
  for (int f = 0; f <= NV; f++) s[f] = 0.0;
#pragma unroll
  for (int bx = 0; bx < 2; bx++) {
    int ci = i - bx;
    if (ci < 0) ci += g.nx;
#pragma unroll
    for (int by = 0; by < 2; by++) {
      int cj = j - by;
      if (cj < 0) cj += g.ny;
#pragma unroll
      for (int bz = 0; bz < 2; bz++) {
        const int ck = k - bz;
        if (ck < 0 || ck >= g.ncz) continue;
        const int64_t c = ((int64_t)ck * g.ny + cj) * g.nx + ci;
        const int q = (bz << 2) | (by << 1) | bx;
#pragma unroll
        for (int f = 0; f <= NV; f++) s[f] += partial[(int64_t)(q * (NV + 1) + f) * ncell + c];
      }
    }
  }
  if (NORM) {
    double w = s[0];
    if (w == 0.) w = 1.;
#pragma unroll
    for (int f = 1; f <= NV; f++) out.p[f - 1][t] = s[f] / w;
  } else {
    const double m = (DBL && (k == 0 || k == g.nzn - 1)) ? 2. : 1.;
#pragma unroll
    for (int f = 0; f <= NV; f++) out.p[f][t] = s[f] * m;
  }
}

// grid2par of one field (parcel_interpolation.F90:287-327)
__global__ void k_grid2par(Geometry g, const double *__restrict__ par, int64_t cap, int64_t n,
                           const double *__restrict__ G, double *__restrict__ var,
                           DevScalars *sc) {
  int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i0, j0, k0;
  double delx, dely, delz;
  int st = locate(g, par[FX * cap + p], par[FY * cap + p], par[FZ * cap + p], i0, j0, k0, delx,
                  dely, delz);
  if (st != PMPIC_OK) {
    raise_status(sc, st);
    return;
  }
  const int i1 = (i0 + 1 == g.nx) ? 0 : i0 + 1;
  const int j1 = (j0 + 1 == g.ny) ? 0 : j0 + 1;
  const double c000 = G[gidx(g, k0, j0, i0)], c001 = G[gidx(g, k0, j0, i1)];
  const double c010 = G[gidx(g, k0, j1, i0)], c011 = G[gidx(g, k0, j1, i1)];
  const double c100 = G[gidx(g, k0 + 1, j0, i0)], c101 = G[gidx(g, k0 + 1, j0, i1)];
  const double c110 = G[gidx(g, k0 + 1, j1, i0)], c111 = G[gidx(g, k0 + 1, j1, i1)];
  const double c00 = c000 * (1 - delz) + c100 * delz;
  const double c01 = c001 * (1 - delz) + c101 * delz;
  const double c10 = c010 * (1 - delz) + c110 * delz;
  const double c11 = c011 * (1 - delz) + c111 * delz;
  const double c0 = c00 * (1 - dely) + c10 * dely;
  const double c1 = c01 * (1 - dely) + c11 * dely;
  var[p] = c0 * (1 - delx) + c1 * delx;
}

}  // namespace

int scan_exclusive_i32(Ctx *c, const int32_t *in, int32_t *out, int64_t n) {
  size_t need = 0;
  PMPIC_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)n, c->stream));
  if (need > c->scan_tmp_bytes) {
    if (c->scan_tmp) PMPIC_CHECK(hipFree(c->scan_tmp));
    c->scan_tmp = nullptr;
    PMPIC_CHECK(hipMalloc(&c->scan_tmp, need));
    c->scan_tmp_bytes = need;
  }
  size_t have = c->scan_tmp_bytes;
  PMPIC_CHECK(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp, have, in, out, (int)n, c->stream));
  return PMPIC_OK;
}

int bin_parcels(Ctx *c, int64_t n) {
  const Geometry &g = c->g;
  const int64_t nc1 = c->ncell + 1;
  hipLaunchKernelGGL(k_fill_i32, dim3(cdiv(nc1, TPB)), dim3(TPB), 0, c->stream, c->cell_count,
                     nc1, 0);
  if (n > 0)
    hipLaunchKernelGGL(k_bin_count, dim3(cdiv(n, TPB)), dim3(TPB), 0, c->stream, g,
                       c->par + FX * c->cap, c->par + FY * c->cap, c->par + FZ * c->cap, n,
                       c->cell, c->rank, c->cell_count, c->dsc);
  PMPIC_TRY(scan_exclusive_i32(c, c->cell_count, c->cell_start, nc1));
  if (n > 0)
    hipLaunchKernelGGL(k_bin_scatter, dim3(cdiv(n, TPB)), dim3(TPB), 0, c->stream, c->cell,
                       c->rank, c->cell_start, n, c->sorted);
  hipLaunchKernelGGL(k_bin_sort, dim3(cdiv(c->ncell, TPB)), dim3(TPB), 0, c->stream,
                     c->cell_start, c->ncell, c->sorted);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int par2grid_fused(Ctx *c, int64_t n) {
  PMPIC_TRY(bin_parcels(c, n));
  const int64_t ncell = c->ncell;
  PMPIC_TRY(timing_begin(c, 0));
  hipLaunchKernelGGL((k_cell_accum<ACC_P2G, 4>), dim3(cdiv(ncell, TPB)), dim3(TPB), 0, c->stream,
                     c->g, c->par, c->cap, (const double *)nullptr, c->cell_start, c->sorted,
                     ncell, c->partial, (int32_t *)nullptr);
  PMPIC_TRY(timing_end(c, 0));
  OutTable table = {{c->grid[GP], c->grid[GQ], c->grid[GR], c->grid[GB], nullptr, nullptr}};
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  hipLaunchKernelGGL((k_node_sum<4, true>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream, c->g,
                     c->partial, ncell, table);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int par2grid_single(Ctx *c, int64_t n, const double *d_var, double *d_grid) {
  PMPIC_TRY(bin_parcels(c, n));
  const int64_t ncell = c->ncell;
  hipLaunchKernelGGL((k_cell_accum<ACC_SINGLE, 1>), dim3(cdiv(ncell, TPB)), dim3(TPB), 0,
                     c->stream, c->g, c->par, c->cap, d_var, c->cell_start, c->sorted, ncell,
                     c->partial, (int32_t *)nullptr);
  OutTable table = {{d_grid, nullptr, nullptr, nullptr, nullptr, nullptr}};
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  hipLaunchKernelGGL((k_node_sum<1, true>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream, c->g,
                     c->partial, ncell, table);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int grid2par_single(Ctx *c, int64_t n, const double *d_grid, double *d_var) {
  if (n > 0)
    hipLaunchKernelGGL(k_grid2par, dim3(cdiv(n, TPB)), dim3(TPB), 0, c->stream, c->g, c->par,
                       c->cap, n, d_grid, d_var, c->dsc);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

// kept and residual grids of parcel_mixing (construct_grids, parcel_mixing.F90:440-624, then
// the fold/fill of :133-138): grid[GVOL], GB, GHG, GP, GQ, GR hold the kept grids with the
// z-boundary doubling of :141-152; kres[0..5] hold the residual grids before the new-parcel
// correction of :202-207 and before their own doubling (:229-240).
int mixing_grids(Ctx *c, int64_t n) {
  PMPIC_TRY(bin_parcels(c, n));
  const int64_t ncell = c->ncell;
  hipLaunchKernelGGL((k_cell_accum<ACC_MIX_KEEP, 5>), dim3(cdiv(ncell, TPB)), dim3(TPB), 0,
                     c->stream, c->g, c->par, c->cap, (const double *)nullptr, c->cell_start,
                     c->sorted, ncell, c->partial, c->npercell);
  hipLaunchKernelGGL((k_cell_accum<ACC_MIX_RES, 5>), dim3(cdiv(ncell, TPB)), dim3(TPB), 0,
                     c->stream, c->g, c->par, c->cap, (const double *)nullptr, c->cell_start,
                     c->sorted, ncell, c->partial2, (int32_t *)nullptr);
  OutTable kept = {{c->grid[GVOL], c->grid[GB], c->grid[GHG], c->grid[GP], c->grid[GQ],
                    c->grid[GR]}};
  OutTable res = {{c->kres[0], c->kres[1], c->kres[2], c->kres[3], c->kres[4], c->kres[5]}};
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  hipLaunchKernelGGL((k_node_sum<5, false>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream,
                     c->g, c->partial, ncell, kept);
  hipLaunchKernelGGL((k_node_sum<5, false, false>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0,
                     c->stream, c->g, c->partial2, ncell, res);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

}  // namespace pmpic
