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
//   2. cell:  a workgroup stages a tile of cells' parcels in LDS (coalesced loads, one
//             location per parcel), then one thread per cell adds the 8 corner weights (and
//             weight * value) of its parcels in ascending parcel order;
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

// cell of every parcel and its slot in the cell's bucket.  Parcels that are neighbours in slot
// order usually share a cell, so lanes are grouped into runs of equal cell and one lane per run
// takes the whole run's slots with a single integer atomic.  Each wavefront walks BIN_SEQ
// consecutive groups of 64 parcels in order, so a cell whose parcels are consecutive in slot
// order gets ascending bucket slots unless it straddles two wavefronts' ranges.
constexpr int BIN_SEQ = 16;

__global__ void k_bin_count(Geometry g, const double *__restrict__ X, const double *__restrict__ Y,
                            const double *__restrict__ Z, int64_t n, int32_t *__restrict__ cell,
                            int32_t *__restrict__ rank, int32_t *__restrict__ count,
                            DevScalars *sc) {
  const int lane = threadIdx.x & 63;
  const int64_t w0 = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64 * BIN_SEQ;
  // all atomics of the wavefront are issued before any result is used, in group order
  int base[BIN_SEQ], hls[BIN_SEQ];
  // the positions of the next BIN_AHEAD groups are loaded before the current group is located
  // (clamped index, n >= 1), so the wavefront does not wait on memory once per group
  constexpr int BIN_AHEAD = 4;
  double px[BIN_SEQ], py[BIN_SEQ], pz[BIN_SEQ];
  auto load_group = [&](int it) {
    const int64_t p = min(w0 + (int64_t)it * 64 + lane, n - 1);
    px[it] = X[p];
    py[it] = Y[p];
    pz[it] = Z[p];
  };
#pragma unroll
  for (int it = 0; it < BIN_AHEAD; it++) load_group(it);
#pragma unroll
  for (int it = 0; it < BIN_SEQ; it++) {
    if (it + BIN_AHEAD < BIN_SEQ) load_group(it + BIN_AHEAD);
    const int64_t p = w0 + (int64_t)it * 64 + lane;
    int32_t c = -1;
    if (p < n) {
      int i0, j0, k0;
      double dx_, dy_, dz_;
      const int st = locate(g, px[it], py[it], pz[it], i0, j0, k0, dx_, dy_, dz_);
      if (st != PMPIC_OK) raise_status(sc, st);
      else c = (int32_t)(((int64_t)k0 * g.ny + j0) * g.nx + i0);
      cell[p] = c;
    }
    const int32_t cprev = __shfl_up(c, 1);
    const bool head = (lane == 0) || (c != cprev);
    const unsigned long long hm = __ballot(head);
    const unsigned long long upto = hm & ((2ull << lane) - 1ull);
    const int hl = 63 - __clzll(upto);
    const unsigned long long above = hm & ~((2ull << hl) - 1ull);
    const int next = above ? (__ffsll(above) - 1) : 64;
    base[it] = (head && c >= 0) ? atomicAdd(&count[c], next - hl) : 0;
    hls[it] = hl;
  }
#pragma unroll
  for (int it = 0; it < BIN_SEQ; it++) {
    const int64_t p = w0 + (int64_t)it * 64 + lane;
    const int b = __shfl(base[it], hls[it]);
    if (p < n) rank[p] = b + (lane - hls[it]);
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

// buckets larger than the LDS tile are sorted by parcel index here (insertion sort; such cells
// do not occur at the target workloads); smaller buckets are ordered inside the accumulation
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
enum AccMode { ACC_P2G = 0, ACC_SINGLE = 1, ACC_MIX_KEEP = 2, ACC_MIX_RES = 3, ACC_DUMP = 4 };

// Per-cell corner partial sums.
// A workgroup owns TILE consecutive cells.  Their buckets are consecutive in the bucket array, so
// the workgroup loads them with coalesced reads (bucket entries are mostly consecutive parcel
// slots), locates each parcel once and stages (offsets, volume, values, parcel index) in LDS as
// 16-byte pairs.  The 8 threads of a cell put its rows in ascending parcel order (a pass over
// adjacent pairs, and a parallel ranking only when a pair is out of order); then one thread per
// (cell, corner) adds that corner's weight and weight * value over the cell's rows in that
// order, so within a cell the additions run in the reference's order.
// MIX mode sums kept parcels into `partial` (and counts them) and residual parcels into
// `partial2`.  Cells with more than ROWS parcels (never at the target workloads) are summed
// straight from global memory after k_bin_sort_big ordered their buckets.
constexpr int TILE = 16;  // cells per workgroup: 8 threads (corners) each
constexpr int ROWS_P2G = 512, ROWS_MIX = 448;

template <int MODE, int NV>
__device__ __forceinline__ void load_values(const double *__restrict__ par, int64_t cap,
                                            const double *__restrict__ var, int64_t p, double z,
                                            double *val) {
  if constexpr (MODE == ACC_P2G) {
    val[0] = ld_stream(&par[FP * cap + p]);
    val[1] = ld_stream(&par[FQ * cap + p]);
    val[2] = ld_stream(&par[FR * cap + p]);
    // total buoyancy, vorticity_tendency.F90:121-126
    const double hl = ld_stream(&par[FH * cap + p]) - exp(-z);
    val[3] = ld_stream(&par[FB * cap + p]) + 12.5 * (hl > 0.0 ? hl : 0.0);
  } else if constexpr (MODE == ACC_SINGLE) {
    val[0] = ld_stream(&var[p]);
  } else {
    val[0] = ld_stream(&par[FB * cap + p]);
    val[1] = ld_stream(&par[FH * cap + p]);
    val[2] = ld_stream(&par[FP * cap + p]);
    val[3] = ld_stream(&par[FQ * cap + p]);
    val[4] = ld_stream(&par[FR * cap + p]);
  }
}

template <int MODE, int NV, int ROWS>
__global__ void __launch_bounds__(256) k_cell_accum_lds(Geometry g, const double *__restrict__ par,
                                                        int64_t cap, const double *__restrict__ var,
                                                        const int32_t *__restrict__ start,
                                                        const int32_t *__restrict__ sorted,
                                                        int64_t ncell, double *__restrict__ partial,
                                                        double *__restrict__ partial2,
                                                        int32_t *__restrict__ npercell) {
  constexpr int NF = NV + 1;
  constexpr int NVS = (NV + 1) / 2;  // value pairs per row
  constexpr bool MIX = (MODE == ACC_MIX_KEEP);
  __shared__ double2 sgeo[2][ROWS];    // (delx, dely), (delz, vol)
  __shared__ double2 sval[NVS][ROWS];  // values, two per slot
  __shared__ int32_t sidx[ROWS];
  __shared__ int16_t sperm[ROWS];
  __shared__ uint8_t skeep[MIX ? ROWS : 1];
  __shared__ int32_t sstart[TILE + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int cl = tid >> 3;          // cell within the tile
  const int q = tid & 7;            // corner: bits z<<2 | y<<1 | x
  const int bx = q & 1, by = (q >> 1) & 1, bz = (q >> 2) & 1;
  const int64_t c0 = (int64_t)blockIdx.x * TILE;
  const int ntile = (int)min((int64_t)TILE, ncell - c0);
  for (int t = tid; t <= ntile; t += blockDim.x) sstart[t] = start[c0 + t];
  __syncthreads();
  const double *X = par + FX * cap, *Y = par + FY * cap, *Z = par + FZ * cap;
  const double *VOL = par + FVOL * cap;
  double acc[NF], res[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) { acc[f] = 0.0; res[f] = 0.0; }
  int kept = 0;
  bool big = false;
  int first = 0;
  while (first < ntile) {
    // the group of cells [first, last) whose rows fit in LDS (at least one cell)
    int last = first + 1;
    if (sstart[ntile] - sstart[first] <= ROWS) last = ntile;  // the usual case: the rest fits
    else
      while (last < ntile && sstart[last + 1] - sstart[first] <= ROWS) last++;
    const int32_t r0 = sstart[first];
    const int nrows = sstart[last] - r0;
    const bool fits = nrows <= ROWS;
    if (fits) {
      // all of this thread's rows are fetched before any is used, so the dependent
      // bucket -> parcel loads of different rows overlap
      constexpr int RPT = (ROWS + TILE * 8 - 1) / (TILE * 8);
      int32_t pr[RPT];
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const int r = tid + k * TILE * 8;
        pr[k] = (r < nrows) ? sorted[r0 + r] : -1;
      }
      double px[RPT], py[RPT], pz[RPT], pv[RPT], pval[RPT][2 * NVS];
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        pval[k][2 * NVS - 1] = 0.0;
        if (pr[k] >= 0) {
          const int64_t pp = pr[k];
          px[k] = ld_stream(&X[pp]);
          py[k] = ld_stream(&Y[pp]);
          pz[k] = ld_stream(&Z[pp]);
          pv[k] = ld_stream(&VOL[pp]);
          load_values<MODE, NV>(par, cap, var, pp, pz[k], pval[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        const int r = tid + k * TILE * 8;
        if (pr[k] < 0) continue;
        int i0, j0, k0;
        double delx, dely, delz;
        locate(g, px[k], py[k], pz[k], i0, j0, k0, delx, dely, delz);
        sgeo[0][r] = make_double2(delx, dely);
        sgeo[1][r] = make_double2(delz, pv[k]);
#pragma unroll
        for (int f = 0; f < NVS; f++) sval[f][r] = make_double2(pval[k][2 * f], pval[k][2 * f + 1]);
        if constexpr (MIX) skeep[r] = (pv[k] < g.vmin) ? 0 : 1;
        sidx[r] = pr[k];
      }
    }
    __syncthreads();
    const bool mine = fits && cl >= first && cl < last;
    const int lo = mine ? sstart[cl] - r0 : 0, hi = mine ? sstart[cl + 1] - r0 : 0;
    // Canonical order: a cell's rows by ascending parcel index.  Bucket slots come from atomics,
    // so runs of a cell's parcels filled by different wavefronts may arrive out of order.  The 8
    // corner lanes of a cell check their share of adjacent pairs; if any pair is out of order,
    // each lane ranks its share of rows by counting smaller parcel indices in the cell.
    bool ord = true;
    for (int a = lo + 1 + q; a < hi; a += 8) ord &= (sidx[a - 1] < sidx[a]);
    const unsigned long long bad = __ballot(!ord);
    const bool sort_cell = ((bad >> (lane & ~7)) & 0xffull) != 0ull;
    if (sort_cell) {
      for (int a = lo + q; a < hi; a += 8) {
        const int32_t key = sidx[a];
        int rk = 0;
        for (int b = lo; b < hi; b++) rk += (sidx[b] < key) ? 1 : 0;
        sperm[lo + rk] = (int16_t)a;
      }
    } else {
      for (int a = lo + q; a < hi; a += 8) sperm[a] = (int16_t)a;
    }
    __syncthreads();
#pragma unroll 2
    for (int a = lo; a < hi; a++) {
      const int r = sperm[a];
      const double2 g0 = sgeo[0][r], g1 = sgeo[1][r];
      double v[2 * NVS];
#pragma unroll
      for (int f = 0; f < NVS; f++) {
        const double2 t = sval[f][r];
        v[2 * f] = t.x;
        v[2 * f + 1] = t.y;
      }
      const double fx = bx ? g0.x : 1 - g0.x;
      const double fy = by ? g0.y : 1 - g0.y;
      const double fz = bz ? g1.x : 1 - g1.x;
      const double w = fz * fy * fx * g1.y;
      if (!MIX || skeep[r]) {
        acc[0] += w;
#pragma unroll
        for (int f = 0; f < NV; f++) acc[f + 1] += w * v[f];
        kept++;
      } else {
        res[0] += w;
#pragma unroll
        for (int f = 0; f < NV; f++) res[f + 1] += w * v[f];
      }
    }
    if (fits == false && cl == first) big = true;
    __syncthreads();
    first = last;
  }
  if (cl >= ntile) return;
  const int64_t c = c0 + cl;
  if (big) {
    // bucket already sorted by k_bin_sort_big; the same sums straight from global memory
    for (int32_t s = sstart[cl]; s < sstart[cl + 1]; s++) {
      const int32_t p = sorted[s];
      const double v = VOL[p];
      int i0, j0, k0;
      double delx, dely, delz;
      const double zp = Z[p];
      locate(g, X[p], Y[p], zp, i0, j0, k0, delx, dely, delz);
      double val[NV];
      load_values<MODE, NV>(par, cap, var, p, zp, val);
      const double fx = bx ? delx : 1 - delx;
      const double fy = by ? dely : 1 - dely;
      const double fz = bz ? delz : 1 - delz;
      const double w = fz * fy * fx * v;
      if (!MIX || !(v < g.vmin)) {
        acc[0] += w;
#pragma unroll
        for (int f = 0; f < NV; f++) acc[f + 1] += w * val[f];
        kept++;
      } else {
        res[0] += w;
#pragma unroll
        for (int f = 0; f < NV; f++) res[f + 1] += w * val[f];
      }
    }
  }
  // MIX: either output may be null (the two-pass form of mixing_grids); the sums are the same
  if (!MIX || partial) {
#pragma unroll
    for (int f = 0; f < NF; f++) partial[((int64_t)q * ncell + c) * NF + f] = acc[f];
  }
  if constexpr (MIX) {
    if (partial2) {
#pragma unroll
      for (int f = 0; f < NF; f++) partial2[((int64_t)q * ncell + c) * NF + f] = res[f];
    }
    if (npercell && q == 0) npercell[c] = kept;
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
  // every corner's partials are loaded first (a corner below the bottom or above the top cell
  // reads cell 0 and is left out of the sum), then summed in the fixed corner order
  double v[8][NV + 1];
  bool ok[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int bx = q & 1, by = (q >> 1) & 1, bz = q >> 2;
    int ci = i - bx;
    if (ci < 0) ci += g.nx;
    int cj = j - by;
    if (cj < 0) cj += g.ny;
    const int ck = k - bz;
    ok[q] = ck >= 0 && ck < g.ncz;
    const int64_t c = ok[q] ? ((int64_t)ck * g.ny + cj) * g.nx + ci : 0;
#pragma unroll
    for (int f = 0; f <= NV; f++) v[q][f] = partial[((int64_t)q * ncell + c) * (NV + 1) + f];
  }
  double s[NV + 1];
#pragma unroll
  for (int f = 0; f <= NV; f++) s[f] = 0.0;
#pragma unroll
  for (int bx = 0; bx < 2; bx++)
#pragma unroll
    for (int by = 0; by < 2; by++)
#pragma unroll
      for (int bz = 0; bz < 2; bz++) {
        const int q = (bz << 2) | (by << 1) | bx;
#pragma unroll
        for (int f = 0; f <= NV; f++) s[f] = ok[q] ? s[f] + v[q][f] : s[f];
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

size_t scan_temp_bytes(Ctx *c, int64_t n) {
  size_t need = 0;
  if (hipcub::DeviceScan::ExclusiveSum(nullptr, need, (const int32_t *)nullptr, (int32_t *)nullptr,
                                       (int)n, c->stream) != hipSuccess)
    return 0;
  return need;
}

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

// the context's own binning buffers (the scan's temporary storage is grown on demand)
static BinWork ctx_work(Ctx *c) {
  BinWork w;
  w.cell = c->cell;
  w.rank = c->rank;
  w.count = c->cell_count;
  w.start = c->cell_start;
  w.sorted = c->sorted;
  w.partial = c->partial;
  w.scan_tmp = nullptr;  // c->scan_tmp, grown by scan_exclusive_i32
  w.scan_bytes = 0;
  return w;
}

int bin_parcels_ws(Ctx *c, int64_t n, const BinWork &w) {
  const Geometry &g = c->g;
  const int64_t nc1 = c->ncell + 1;
  hipLaunchKernelGGL(k_fill_i32, dim3(cdiv(nc1, TPB)), dim3(TPB), 0, c->stream, w.count, nc1, 0);
  if (n > 0)
    hipLaunchKernelGGL(k_bin_count, dim3(cdiv(n, (int64_t)TPB * BIN_SEQ)), dim3(TPB), 0, c->stream, g,
                       c->par + FX * c->cap, c->par + FY * c->cap, c->par + FZ * c->cap, n,
                       w.cell, w.rank, w.count, c->dsc);
  if (w.scan_tmp) {
    size_t have = w.scan_bytes;
    PMPIC_CHECK(hipcub::DeviceScan::ExclusiveSum(w.scan_tmp, have, w.count, w.start, (int)nc1,
                                                 c->stream));
  } else {
    PMPIC_TRY(scan_exclusive_i32(c, w.count, w.start, nc1));
  }
  if (n > 0)
    hipLaunchKernelGGL(k_bin_scatter, dim3(cdiv(n, TPB)), dim3(TPB), 0, c->stream, w.cell, w.rank,
                       w.start, n, w.sorted);
  hipLaunchKernelGGL(k_bin_sort_big, dim3(cdiv(c->ncell, TPB)), dim3(TPB), 0, c->stream, w.start,
                     c->ncell, ROWS_MIX, w.sorted);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int bin_parcels(Ctx *c, int64_t n) { return bin_parcels_ws(c, n, ctx_work(c)); }

template <int MODE, int NV, int ROWS>
static void launch_accum_lds(Ctx *c, const double *var, double *partial, double *partial2,
                             int32_t *npc) {
  hipLaunchKernelGGL((k_cell_accum_lds<MODE, NV, ROWS>), dim3(cdiv(c->ncell, TILE)), dim3(TILE * 8), 0,
                     c->stream, c->g, c->par, c->cap, var, c->cell_start, c->sorted, c->ncell,
                     partial, partial2, npc);
}

int par2grid_fused(Ctx *c, int64_t n) {
  PMPIC_TRY(bin_parcels(c, n));
  const int64_t ncell = c->ncell;
  PMPIC_TRY(launch_timed(c, T_ACCUM, k_cell_accum_lds<ACC_P2G, 4, ROWS_P2G>,
                         dim3(cdiv(c->ncell, TILE)), dim3(TILE * 8), 0, c->g, c->par, c->cap,
                         nullptr, c->cell_start, c->sorted, c->ncell, c->partial, nullptr,
                         nullptr));
  OutTable table = {{c->grid[GP], c->grid[GQ], c->grid[GR], c->grid[GB], nullptr, nullptr}};
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  hipLaunchKernelGGL((k_node_sum<4, true>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream, c->g,
                     c->partial, ncell, table);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int par2grid_single_ws(Ctx *c, int64_t n, const double *d_var, double *d_grid,
                       const BinWork &w) {
  PMPIC_TRY(bin_parcels_ws(c, n, w));
  const int64_t ncell = c->ncell;
  hipLaunchKernelGGL((k_cell_accum_lds<ACC_SINGLE, 1, ROWS_P2G>), dim3(cdiv(ncell, TILE)),
                     dim3(TILE * 8), 0, c->stream, c->g, c->par, c->cap, d_var, w.start, w.sorted,
                     ncell, w.partial, nullptr, nullptr);
  OutTable table = {{d_grid, nullptr, nullptr, nullptr, nullptr, nullptr}};
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  hipLaunchKernelGGL((k_node_sum<1, true>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream, c->g,
                     w.partial, ncell, table);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int par2grid_single(Ctx *c, int64_t n, const double *d_var, double *d_grid) {
  return par2grid_single_ws(c, n, d_var, d_grid, ctx_work(c));
}

// interior [nzn][ny][nx] -> the reference layout data(z, y, x) with x/y halo 2 holding the
// periodic copies of the interior (the layout pmpic_par2grid returns)
__global__ void k_to_ref_layout(Geometry g, const double *__restrict__ in, double *__restrict__ out) {
  const int nxh = g.nx + 4, nyh = g.ny + 4;
  const int64_t total = (int64_t)nxh * nyh * g.nzn;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int k = (int)(t % g.nzn);
  const int jh = (int)((t / g.nzn) % nyh);
  const int ih = (int)(t / ((int64_t)g.nzn * nyh));
  int i = ih - 2, j = jh - 2;
  if (i < 0) i += g.nx;
  else if (i >= g.nx) i -= g.nx;
  if (j < 0) j += g.ny;
  else if (j >= g.ny) j -= g.ny;
  out[t] = in[gidx(g, k, j, i)];
}

int to_ref_layout(Ctx *c, const double *d_interior, double *d_ref) {
  const int64_t total = (int64_t)(c->g.nx + 4) * (c->g.ny + 4) * c->g.nzn;
  hipLaunchKernelGGL(k_to_ref_layout, dim3(cdiv(total, TPB)), dim3(TPB), 0, c->stream, c->g,
                     d_interior, d_ref);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

// writebinary's grids (writebinary.F90:217-233): b, hg, p, q, r by par2grid of the parcels'
// b, h, p, q, r (one binning, the five sums of each node in parcel order as five par2grid
// calls would take them), then hgliq = max(0, hg - exp(-z)) and b += 12.5 hgliq per node
__global__ void k_hgliq(Geometry g, double *__restrict__ B, const double *__restrict__ HG,
                        double *__restrict__ HGLIQ) {
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode) return;
  const int k = (int)(t / ((int64_t)g.nx * g.ny));
  const double z = g.zmin_h + (double)k * g.dz + g.zbot;  // z_coords(k+1), setup as in create
  const double hl = HG[t] - exp(-z);
  const double liq = hl > 0. ? hl : 0.;
  HGLIQ[t] = liq;
  B[t] = B[t] + 12.5 * liq;
}

int dump_grids_f64(Ctx *c, int64_t n) {
  PMPIC_TRY(bin_parcels(c, n));
  launch_accum_lds<ACC_DUMP, 5, ROWS_MIX>(c, nullptr, c->partial, nullptr, nullptr);
  OutTable out = {{c->grid[GB], c->grid[GHG], c->grid[GP], c->grid[GQ], c->grid[GR], nullptr}};
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  hipLaunchKernelGGL((k_node_sum<5, true>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream, c->g,
                     c->partial, c->ncell, out);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int hgliq_grid(Ctx *c) {
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  hipLaunchKernelGGL(k_hgliq, dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream, c->g, c->grid[GB],
                     c->grid[GHG], c->grid[GHGLIQ]);
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
// Without a residual buffer (partial2 null: a context too large for both) the kept and the
// residual sums take one pass each over `partial`, with the same sums.
int mixing_grids(Ctx *c, int64_t n) {
  PMPIC_TRY(bin_parcels(c, n));
  const int64_t ncell = c->ncell;
  OutTable kept = {{c->grid[GVOL], c->grid[GB], c->grid[GHG], c->grid[GP], c->grid[GQ],
                    c->grid[GR]}};
  OutTable res = {{c->kres[0], c->kres[1], c->kres[2], c->kres[3], c->kres[4], c->kres[5]}};
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  if (c->partial2) {
    launch_accum_lds<ACC_MIX_KEEP, 5, ROWS_MIX>(c, nullptr, c->partial, c->partial2, c->npercell);
    hipLaunchKernelGGL((k_node_sum<5, false>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream,
                       c->g, c->partial, ncell, kept);
    hipLaunchKernelGGL((k_node_sum<5, false, false>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0,
                       c->stream, c->g, c->partial2, ncell, res);
  } else {
    launch_accum_lds<ACC_MIX_KEEP, 5, ROWS_MIX>(c, nullptr, c->partial, nullptr, c->npercell);
    hipLaunchKernelGGL((k_node_sum<5, false>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream,
                       c->g, c->partial, ncell, kept);
    launch_accum_lds<ACC_MIX_KEEP, 5, ROWS_MIX>(c, nullptr, nullptr, c->partial, nullptr);
    hipLaunchKernelGGL((k_node_sum<5, false, false>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0,
                       c->stream, c->g, c->partial, ncell, res);
  }
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

}  // namespace pmpic
