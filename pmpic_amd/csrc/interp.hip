// Parcel <-> grid interpolation on MI355X: deterministic cell-binned par2grid, grid2par.
//
// Reference: model_core/src/parcels/parcel_interpolation.F90:186-481 (cache weights, grid2par,
// par2grid) and model_core/src/communication/MPIC_Haloswap.F90:95-307 (halo fold / fill, which
// on one GPU is periodic index arithmetic), components/parcel_mixing/src/parcel_mixing.F90:440-624
// (kept / residual grids).
//
// par2grid is a scatter; here it is a gather over a cell binning so that no floating-point
// atomics are used and every node sum has a fixed order:
//   1. bin:   each parcel finds its cell and takes a slot in the cell's bucket (integer atomics,
//             one per run of equal cells within a wavefront), an exclusive scan turns bucket
//             sizes into offsets and parcel indices are scattered into the buckets;
//   2. cell:  a subgroup of lanes per cell loads the cell's parcels (coalesced), sorts their
//             indices across the lanes, forms the 8 corner weights (and weight * value) per
//             parcel and reduces them over the subgroup in a fixed butterfly;
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

// cell of every parcel and its slot in the cell's bucket.  Parcels that are neighbours in slot
// order usually share a cell, so lanes are grouped into runs of equal cell and one lane per run
// takes the whole run's slots with a single integer atomic.
__global__ void k_bin_count(Geometry g, const double *__restrict__ X, const double *__restrict__ Y,
                            const double *__restrict__ Z, int64_t n, int32_t *__restrict__ cell,
                            int32_t *__restrict__ rank, int32_t *__restrict__ count,
                            DevScalars *sc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int32_t c = -1;
  if (p < n) {
    int i0, j0, k0;
    double dx_, dy_, dz_;
    const int st = locate(g, X[p], Y[p], Z[p], i0, j0, k0, dx_, dy_, dz_);
    if (st != PMPIC_OK) raise_status(sc, st);
    else c = (int32_t)(((int64_t)k0 * g.ny + j0) * g.nx + i0);
  }
  const int32_t cprev = __shfl_up(c, 1);
  const bool head = (lane == 0) || (c != cprev);
  const unsigned long long hm = __ballot(head);
  const unsigned long long upto = hm & ((2ull << lane) - 1ull);
  const int hl = 63 - __clzll(upto);
  const unsigned long long above = hm & ~((2ull << hl) - 1ull);
  const int next = above ? (__ffsll(above) - 1) : 64;
  int base = 0;
  if (head && c >= 0) base = atomicAdd(&count[c], next - hl);
  base = __shfl(base, hl);
  if (p < n) {
    cell[p] = c;
    rank[p] = base + (lane - hl);
  }
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

// buckets larger than one subgroup are sorted by parcel index here (insertion sort; such cells
// are few); smaller buckets are sorted inside the accumulation kernel
__global__ void k_bin_sort_big(const int32_t *__restrict__ start, int64_t ncell, int sgsize,
                               int32_t *__restrict__ sorted) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int32_t b = start[c], e = start[c + 1];
  if (e - b <= sgsize) return;
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

// Per-cell corner partial sums.  A subgroup of S lanes owns one cell: each lane takes one of the
// cell's parcels (consecutive bucket entries, which are mostly consecutive parcel slots, so the
// loads coalesce), the subgroup sorts its parcel indices (bitonic network over the lanes) so the
// result does not depend on the order the bucket was filled in, every lane forms the 8 corner
// weights times (1, values), and a butterfly over the subgroup adds them in a fixed tree.
// Cells with more than S parcels are processed in chunks of S (bucket pre-sorted) whose sums are
// added in chunk order.
template <int MODE, int NV, int S>
__global__ void __launch_bounds__(256) k_cell_accum(Geometry g, const double *__restrict__ par,
                                                    int64_t cap, const double *__restrict__ var,
                                                    const int32_t *__restrict__ start,
                                                    const int32_t *__restrict__ sorted,
                                                    int64_t ncell, double *__restrict__ partial,
                                                    int32_t *__restrict__ npercell) {
  const int lane = threadIdx.x & 63;
  const int sl = lane & (S - 1);
  const int64_t c = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / S;
  const bool valid = c < ncell;
  const int32_t b = valid ? start[c] : 0;
  const int32_t e = valid ? start[c + 1] : 0;
  const int nch = (e - b + S - 1) / S;
  int wmax = nch;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) wmax = max(wmax, __shfl_xor(wmax, o));
  const double *X = par + FX * cap, *Y = par + FY * cap, *Z = par + FZ * cap;
  const double *VOL = par + FVOL * cap;
  int kept = 0;
  for (int t_ = 0; t_ < wmax; t_++) {
    const int32_t s = b + t_ * S + sl;
    int32_t p = (s < e) ? sorted[s] : INT32_MAX;
#pragma unroll
    for (int k = 2; k <= S; k <<= 1)
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        const int32_t q = __shfl_xor(p, j);
        const bool up = (sl & k) == 0, lower = (sl & j) == 0;
        p = (lower == up) ? min(p, q) : max(p, q);
      }
    // per-lane factors; lanes without a (selected) parcel contribute exact zeros
    double val[NV];
    double fx[2] = {0.0, 0.0}, fy[2] = {0.0, 0.0}, fz[2] = {0.0, 0.0}, vol = 0.0;
#pragma unroll
    for (int f = 0; f < NV; f++) val[f] = 0.0;
    if (p != INT32_MAX) {
      const double pv = VOL[p];
      bool use = true;
      if (MODE == ACC_MIX_KEEP) use = !(pv < g.vmin);
      if (MODE == ACC_MIX_RES) use = (pv < g.vmin);
      if (use) {
        int i0, j0, k0;
        double delx, dely, delz;
        locate(g, X[p], Y[p], Z[p], i0, j0, k0, delx, dely, delz);
        fx[0] = 1 - delx; fx[1] = delx;
        fy[0] = 1 - dely; fy[1] = dely;
        fz[0] = 1 - delz; fz[1] = delz;
        vol = pv;
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
        kept++;
      }
    }
    // one corner at a time: weight (parcel_interpolation.F90:390-434 association) and
    // weight * value, reduced over the subgroup
#pragma unroll 1
    for (int q = 0; q < 8; q++) {
      const double w = fz[(q >> 2) & 1] * fy[(q >> 1) & 1] * fx[q & 1] * vol;
      double t[NV + 1];
      t[0] = w;
#pragma unroll
      for (int f = 0; f < NV; f++) t[f + 1] = w * val[f];
#pragma unroll
      for (int o = S / 2; o > 0; o >>= 1)
#pragma unroll
        for (int f = 0; f <= NV; f++) t[f] += __shfl_xor(t[f], o);
      if (valid && sl == 0 && t_ < nch) {
#pragma unroll
        for (int f = 0; f <= NV; f++) {
          double *d = partial + (int64_t)(q * (NV + 1) + f) * ncell + c;
          *d = (t_ == 0) ? t[f] : *d + t[f];
        }
      }
    }
  }
  if (valid && sl == 0 && nch == 0) {
#pragma unroll
    for (int q = 0; q < 8; q++)
#pragma unroll
      for (int f = 0; f <= NV; f++) partial[(int64_t)(q * (NV + 1) + f) * ncell + c] = 0.0;
  }
  if (MODE == ACC_MIX_KEEP) {
#pragma unroll
    for (int o = S / 2; o > 0; o >>= 1) kept += __shfl_xor(kept, o);
    if (valid && sl == 0) npercell[c] = kept;
  }
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

// subgroup size for the cell accumulation: the smallest of 16, 32, 64 holding a typical cell
static int subgroup_size(const Ctx *c, int64_t n) {
  const double ppc = (double)n / (double)(c->ncell > 0 ? c->ncell : 1);
  return ppc <= 12.0 ? 16 : (ppc <= 28.0 ? 32 : 64);
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
  hipLaunchKernelGGL(k_bin_sort_big, dim3(cdiv(c->ncell, TPB)), dim3(TPB), 0, c->stream,
                     c->cell_start, c->ncell, subgroup_size(c, n), c->sorted);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

template <int MODE, int NV>
static void launch_accum(Ctx *c, int S, const double *var, double *partial, int32_t *npc) {
  const int64_t threads = c->ncell * S;
  const dim3 grid(cdiv(threads, 256)), block(256);
  if (S == 16)
    hipLaunchKernelGGL((k_cell_accum<MODE, NV, 16>), grid, block, 0, c->stream, c->g, c->par,
                       c->cap, var, c->cell_start, c->sorted, c->ncell, partial, npc);
  else if (S == 32)
    hipLaunchKernelGGL((k_cell_accum<MODE, NV, 32>), grid, block, 0, c->stream, c->g, c->par,
                       c->cap, var, c->cell_start, c->sorted, c->ncell, partial, npc);
  else
    hipLaunchKernelGGL((k_cell_accum<MODE, NV, 64>), grid, block, 0, c->stream, c->g, c->par,
                       c->cap, var, c->cell_start, c->sorted, c->ncell, partial, npc);
}

int par2grid_fused(Ctx *c, int64_t n) {
  PMPIC_TRY(bin_parcels(c, n));
  const int64_t ncell = c->ncell;
  PMPIC_TRY(timing_begin(c, 0));
  launch_accum<ACC_P2G, 4>(c, subgroup_size(c, n), nullptr, c->partial, nullptr);
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
  launch_accum<ACC_SINGLE, 1>(c, subgroup_size(c, n), d_var, c->partial, nullptr);
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
  const int S = subgroup_size(c, n);
  launch_accum<ACC_MIX_KEEP, 5>(c, S, nullptr, c->partial, c->npercell);
  launch_accum<ACC_MIX_RES, 5>(c, S, nullptr, c->partial2, nullptr);
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
