// Parcel dynamics on MI355X: fused grid2par + RK4 stage update, the one-GPU parcel haloswap,
// parcel splitting and parcel mixing.
//
// Reference: components/rk4_integrator/src/rk4_integrator.F90:95-319,
// model_core/src/parcels/parcel_interpolation.F90:267-332 (grid2par),
// model_core/src/parcels/haloswap.F90:109-289, 296-456, 489-607, 678-771 (parcel haloswap),
// components/parcel_splitting/src/parcel_splitting.F90:47-168,
// components/parcel_mixing/src/parcel_mixing.F90:68-431.
//
// Slot order is kept exactly as the reference's one-rank run leaves it, because later steps
// depend on it (split children are appended in parent order, new mixing parcels fill removed
// slots in ascending order, backfill moves parcels from the end):
//  * haloswap: on one rank every parcel leaving through x or y is sent to self in direction
//    order N, NE, E, SE, S, SW, W, NW, shifted by one period, and unpacked into the vacated
//    slots in ascending slot order.  Here the stage kernel tags each parcel with its direction,
//    block-level ballot ranks plus one exclusive scan give every crossing parcel its rank in
//    (direction, slot) order and the rank of its own slot among the vacated ones, and two
//    gather/scatter kernels move the crossing parcels (all 29 fields) to their new slots.
//  * ranks within a block always come from wavefront ballots in lane order, so every
//    compaction and permutation is deterministic.
#include "common.h"

#include <string.h>

#include <hipcub/hipcub.hpp>
#include <math.h>

namespace pmpic {

namespace {

constexpr int TPB = 256;
constexpr int NWAVE = TPB / 64;

// exclusive rank of `pred` among the block's threads (thread order) and the block total
__device__ __forceinline__ void block_rank(bool pred, int &rank, int &total, int *lds) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(pred);
  const int below = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) lds[wave] = __popcll(m);
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NWAVE; w++) {
    const int v = lds[w];
    if (w < wave) off += v;
    tot += v;
  }
  __syncthreads();
  rank = off + below;
  total = tot;
}

struct GridSet {
  const double *g[6];
};

// haloswap direction of a parcel (haloswap.F90:296-437); 0 = stays
__device__ __forceinline__ int direction(const Geometry &g, double x, double y) {
  const double xmin = g.hs_xmin, xmax = g.hs_xmax, ymin = g.hs_ymin, ymax = g.hs_ymax;
  if ((y >= ymax) && (x >= xmin && x < xmax)) return 1;   // N
  if (x >= xmax && y >= ymax) return 2;                    // NE
  if ((x >= xmax) && (y >= ymin && y < ymax)) return 3;   // E
  if (x >= xmax && y < ymin) return 4;                     // SE
  if ((y < ymin) && (x >= xmin && x < xmax)) return 5;    // S
  if (x < xmin && y < ymin) return 6;                      // SW
  if ((x < xmin) && (y >= ymin && y < ymax)) return 7;    // W
  if (x < xmin && y >= ymax) return 8;                     // NW
  return 0;
}

// per-block direction counts: rows 0..7 = directions 1..8, row 8 = all crossing parcels
__device__ __forceinline__ void write_dir_counts(int dir, bool active, int32_t *blk, int64_t nblk,
                                                 int *lds) {
  for (int d = 1; d <= 8; d++) {
    int r, t;
    block_rank(active && dir == d, r, t, lds);
    if (threadIdx.x == 0) blk[(int64_t)(d - 1) * nblk + blockIdx.x] = t;
  }
  int r, t;
  block_rank(active && dir != 0, r, t, lds);
  if (threadIdx.x == 0) blk[(int64_t)8 * nblk + blockIdx.x] = t;
}

// grid2par of the six tendencies (vort2vel.F90:325-327, vorticity_tendency.F90:273-275) at the
// current positions, then the RK4 stage update (rk4_integrator.F90:109-296), the vertical
// reflection (haloswap.F90:440-456) and the haloswap direction of each parcel.  At stage 4 the
// stretch increment of parcel_splitting.F90:64-73 is applied too (it only needs the updated
// vorticity and the stage-4 tendencies of the same parcel).
template <int STAGE>
__global__ void __launch_bounds__(TPB) k_rk_stage(Geometry g, double *__restrict__ par, int64_t cap,
                                                  int64_t n, GridSet gs, int stretch,
                                                  DevScalars *sc,
                                                  uint8_t *__restrict__ flag,
                                                  int32_t *__restrict__ blk, int64_t nblk) {
  __shared__ int lds[NWAVE];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < n;
  int dir = 0;
  if (active) {
    // every parcel value the stage reads is loaded up front, so the loads overlap each other
    // and the grid gathers instead of waiting behind the stores
    const double x = ld_stream(&par[FX * cap + p]), y = ld_stream(&par[FY * cap + p]), z0 = ld_stream(&par[FZ * cap + p]);
    double yo[6], yf[6];
    if (STAGE == 1) {
      yo[0] = x;
      yo[1] = y;
      yo[2] = z0;
#pragma unroll
      for (int f = 3; f < 6; f++) yo[f] = ld_stream(&par[(FX + f) * cap + p]);
    } else {
#pragma unroll
      for (int f = 0; f < 6; f++) {
        if (STAGE != 4) yo[f] = ld_stream(&par[(f < 3 ? FXO + f : FPO + f - 3) * cap + p]);
        yf[f] = ld_stream(&par[(f < 3 ? FXF + f : FPF + f - 3) * cap + p]);
      }
    }
    const double s_old = (STAGE == 4 && stretch) ? ld_stream(&par[FSTRETCH * cap + p]) : 0.0;
    int i0, j0, k0;
    double delx, dely, delz;
    const int st = locate(g, x, y, z0, i0, j0, k0, delx, dely, delz);
    if (st != PMPIC_OK) {
      raise_status(sc, st);
    } else {
      const int i1 = (i0 + 1 == g.nx) ? 0 : i0 + 1;
      const int j1 = (j0 + 1 == g.ny) ? 0 : j0 + 1;
      const int64_t o000 = gidx(g, k0, j0, i0), o001 = gidx(g, k0, j0, i1);
      const int64_t o010 = gidx(g, k0, j1, i0), o011 = gidx(g, k0, j1, i1);
      const int64_t o100 = gidx(g, k0 + 1, j0, i0), o101 = gidx(g, k0 + 1, j0, i1);
      const int64_t o110 = gidx(g, k0 + 1, j1, i0), o111 = gidx(g, k0 + 1, j1, i1);
      double kv[6];
#pragma unroll
      for (int f = 0; f < 6; f++) {
        const double *G = gs.g[f];
        const double c00 = G[o000] * (1 - delz) + G[o100] * delz;
        const double c01 = G[o001] * (1 - delz) + G[o101] * delz;
        const double c10 = G[o010] * (1 - delz) + G[o110] * delz;
        const double c11 = G[o011] * (1 - delz) + G[o111] * delz;
        const double c0 = c00 * (1 - dely) + c10 * dely;
        const double c1 = c01 * (1 - dely) + c11 * dely;
        kv[f] = c0 * (1 - delx) + c1 * delx;
      }
      const double dt = sc->dtm;
      const double dt6 = dt / 6, dt2 = dt / 2, dt3 = dt / 3;
      double yv[6], yfn[6];
#pragma unroll
      for (int f = 0; f < 6; f++) {
        if (STAGE == 1) {
          yfn[f] = yo[f] + dt6 * kv[f];
          yv[f] = yo[f] + dt2 * kv[f];
        } else if (STAGE == 2) {
          yfn[f] = yf[f] + dt3 * kv[f];
          yv[f] = yo[f] + dt2 * kv[f];
        } else if (STAGE == 3) {
          yfn[f] = yf[f] + dt3 * kv[f];
          yv[f] = yo[f] + dt * kv[f];
        } else {
          yfn[f] = yf[f] + dt6 * kv[f];
          yv[f] = yfn[f];
        }
      }
      // reflect the current z only (haloswap.F90:447-454)
      double z = yv[2];
      if (z < g.zmin_h) z = g.zmin_h + (g.zmin_h - z);
      else if (z > g.zmax_h) z = g.zmax_h - (z - g.zmax_h);
      yv[2] = z;
#pragma unroll
      for (int f = 0; f < 6; f++) {
        st_stream(&par[(FDXDT + f) * cap + p], kv[f]);
        if (STAGE == 1) st_stream(&par[(f < 3 ? FXO + f : FPO + f - 3) * cap + p], yo[f]);
        st_stream(&par[(f < 3 ? FXF + f : FPF + f - 3) * cap + p], yfn[f]);
        st_stream(&par[(FX + f) * cap + p], yv[f]);
      }
      if (STAGE == 4 && stretch) {
        const double third = (double)(1.f / 3.f);
        double ds = yv[3] * kv[3] + yv[4] * kv[4] + yv[5] * kv[5];
        ds = pow(fabs(ds), third);
        st_stream(&par[FSTRETCH * cap + p], s_old + ds * dt);
      }
      dir = direction(g, yv[0], yv[1]);
    }
    flag[p] = (uint8_t)dir;
  }
  write_dir_counts(dir, active, blk, nblk, lds);
}

// direction classification only (haloswap after splitting): reflect, then tag
__global__ void __launch_bounds__(TPB) k_reflect_classify(Geometry g, double *__restrict__ par,
                                                          int64_t cap, int64_t n,
                                                          uint8_t *__restrict__ flag,
                                                          int32_t *__restrict__ blk, int64_t nblk) {
  __shared__ int lds[NWAVE];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < n;
  int dir = 0;
  if (active) {
    double z = par[FZ * cap + p];
    if (z < g.zmin_h) z = g.zmin_h + (g.zmin_h - z);
    else if (z > g.zmax_h) z = g.zmax_h - (z - g.zmax_h);
    par[FZ * cap + p] = z;
    dir = direction(g, par[FX * cap + p], par[FY * cap + p]);
    flag[p] = (uint8_t)dir;
  }
  write_dir_counts(dir, active, blk, nblk, lds);
}

__global__ void k_cross_total(const int32_t *__restrict__ pre, const int32_t *__restrict__ blk,
                              int64_t nblk, DevScalars *sc) {
  if (threadIdx.x || blockIdx.x) return;
  const int64_t last = 9 * nblk - 1;
  const int64_t total9 = (int64_t)pre[last] + blk[last];
  const int64_t base8 = pre[8 * nblk];
  sc->cross = total9 - base8;  // sum of row 8 = number of crossing parcels
}

// crossing parcels: record their slot at its rank among vacated slots and copy their fields,
// shifted by one period, to the staging slot given by their (direction, slot) rank
__global__ void __launch_bounds__(TPB) k_hs_gather(Geometry g, const double *__restrict__ par,
                                                   int64_t cap, int64_t n,
                                                   const uint8_t *__restrict__ flag,
                                                   const int32_t *__restrict__ pre, int64_t nblk,
                                                   int32_t *__restrict__ vacated,
                                                   double *__restrict__ tmp, int64_t tcap) {
  __shared__ int lds[NWAVE];
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = p < n;
  const int dir = active ? flag[p] : 0;
  int wr = 0, wc = 0, t_;
  for (int d = 1; d <= 8; d++) {
    int r;
    block_rank(active && dir == d, r, t_, lds);
    if (dir == d) wr = r;
  }
  block_rank(active && dir != 0, wc, t_, lds);
  if (!active || dir == 0) return;
  const int64_t t = (int64_t)pre[(int64_t)(dir - 1) * nblk + blockIdx.x] + wr;
  const int64_t v = (int64_t)pre[(int64_t)8 * nblk + blockIdx.x] - pre[(int64_t)8 * nblk] + wc;
  vacated[v] = (int32_t)p;
  double xs = 0, ys = 0;
  if (dir == 3 || dir == 2 || dir == 4) xs = g.xshift_e;
  else if (dir == 7 || dir == 8 || dir == 6) xs = g.xshift_w;
  if (dir == 1 || dir == 2 || dir == 8) ys = g.yshift_n;
  else if (dir == 5 || dir == 6 || dir == 4) ys = g.yshift_s;
#pragma unroll 1
  for (int f = 0; f < NFIELDS; f++) {
    double val = par[(int64_t)f * cap + p];
    if (f == FX || f == FXO || f == FXF) val = val + xs;
    if (f == FY || f == FYO || f == FYF) val = val + ys;
    tmp[(int64_t)f * tcap + t] = val;
  }
}

__global__ void k_hs_scatter(double *__restrict__ par, int64_t cap, const double *__restrict__ tmp,
                             int64_t tcap, const int32_t *__restrict__ vacated,
                             const DevScalars *sc) {
  const int64_t K = sc->cross;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < K;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = vacated[t];
#pragma unroll 1
    for (int f = 0; f < NFIELDS; f++) par[(int64_t)f * cap + s] = tmp[(int64_t)f * tcap + t];
  }
}

// ---- generic compaction: ascending indices i < n with flags[i] == want ----
// A thread takes FPT consecutive flags (one 16-byte load when all are in range), so a block
// covers FPB flags; the flags arrays start at allocation boundaries, so the loads are aligned.
constexpr int FPT = 16;
constexpr int FPB = FPT * TPB;

// bit q set when flags[i0 + q] == want (i0 < n)
__device__ __forceinline__ uint32_t flag_bits(const uint8_t *__restrict__ flags, uint8_t want,
                                              int64_t i0, int64_t n) {
  uint32_t m = 0;
  if (i0 + FPT <= n) {
    const uint4 v = *reinterpret_cast<const uint4 *>(flags + i0);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int b = 0; b < 4; b++)
        if (((w[k] >> (8 * b)) & 0xffu) == want) m |= 1u << (4 * k + b);
  } else {
    for (int q = 0; q < FPT && i0 + q < n; q++)
      if (flags[i0 + q] == want) m |= 1u << q;
  }
  return m;
}

// exclusive prefix sum of v over the block in thread order; total = the block's sum
__device__ __forceinline__ int block_excl_scan(int v, int &total, int *lds) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(incl, d);
    if (lane >= d) incl += u;
  }
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NWAVE; w++) {
    const int sw = lds[w];
    if (w < wave) off += sw;
    tot += sw;
  }
  __syncthreads();
  total = tot;
  return off + incl - v;
}

__global__ void __launch_bounds__(TPB) k_flag_count(const uint8_t *__restrict__ flags,
                                                    uint8_t want, int64_t n,
                                                    int32_t *__restrict__ blk) {
  __shared__ int lds[NWAVE];
  const int64_t i0 = ((int64_t)blockIdx.x * TPB + threadIdx.x) * FPT;
  const uint32_t m = (i0 < n) ? flag_bits(flags, want, i0, n) : 0u;
  int t;
  block_excl_scan(__popc(m), t, lds);
  if (threadIdx.x == 0) blk[blockIdx.x] = t;
}

__global__ void __launch_bounds__(TPB) k_flag_scatter(const uint8_t *__restrict__ flags,
                                                      uint8_t want, int64_t n,
                                                      const int32_t *__restrict__ pre,
                                                      const int32_t *__restrict__ blk,
                                                      int64_t nblk, int32_t *__restrict__ out,
                                                      int64_t *__restrict__ count) {
  __shared__ int lds[NWAVE];
  const int64_t i0 = ((int64_t)blockIdx.x * TPB + threadIdx.x) * FPT;
  uint32_t m = (i0 < n) ? flag_bits(flags, want, i0, n) : 0u;
  int t;
  int64_t o = (int64_t)pre[blockIdx.x] + block_excl_scan(__popc(m), t, lds);
  while (m) {
    out[o++] = (int32_t)(i0 + (__ffs(m) - 1));
    m &= m - 1u;
  }
  if (blockIdx.x == nblk - 1 && threadIdx.x == 0 && count)
    *count = (int64_t)pre[nblk - 1] + blk[nblk - 1];
}

// ---- splitting (parcel_splitting.F90:76-131) ----
__global__ void k_split_flags(const double *__restrict__ par, int64_t cap, int64_t n,
                              uint8_t *__restrict__ flag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  flag[p] = (par[FSTRETCH * cap + p] > 4.) ? 1 : 0;
}

__global__ void k_split_apply(double *__restrict__ par, int64_t cap, int64_t n,
                              const int32_t *__restrict__ parents, int64_t nsplit,
                              DevScalars *sc) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nsplit) return;
  const int64_t i = parents[m];
  const int64_t c = n + m;  // child slot n+1+m (1-based)
  if (c >= cap) {
    raise_status(sc, PMPIC_ERR_MAXPARCELS);
    return;
  }
  const double third = (double)(1.f / 3.f);
  const double pi = 4. * atan(1.0);
  const double p = par[FP * cap + i], q = par[FQ * cap + i], r = par[FR * cap + i];
  const double vol = par[FVOL * cap + i];
  double rr = vol / 4. / pi;
  rr = pow(rr, third);
  double absw = p * p + q * q + r * r;
  absw = sqrt(absw);
  const double x = par[FX * cap + i], y = par[FY * cap + i], z = par[FZ * cap + i];
  par[FX * cap + c] = x + rr * p / absw;
  par[FY * cap + c] = y + rr * q / absw;
  par[FZ * cap + c] = z + rr * r / absw;
  par[FP * cap + c] = p;
  par[FQ * cap + c] = q;
  par[FR * cap + c] = r;
  for (int f = FDXDT; f <= FDRDT; f++) par[(int64_t)f * cap + c] = par[(int64_t)f * cap + i];
  par[FH * cap + c] = par[FH * cap + i];
  par[FB * cap + c] = par[FB * cap + i];
  par[FVOL * cap + c] = vol / 2;
  par[FSTRETCH * cap + c] = 0.;
  par[FTAG * cap + c] = par[FTAG * cap + i];
  par[FX * cap + i] = x - rr * p / absw;
  par[FY * cap + i] = y - rr * q / absw;
  par[FZ * cap + i] = z - rr * r / absw;
  par[FVOL * cap + i] = vol / 2;
  par[FSTRETCH * cap + i] = 0.;
}

// ---- mixing (parcel_mixing.F90:157-389) ----
// KPT parcels per thread, at stride TPB (each a coalesced pass), the loads issued together
constexpr int KPT = 4;
__global__ void k_keep_flags(const double *__restrict__ par, int64_t cap, int64_t n, double vmin,
                             uint8_t *__restrict__ keep) {
  const int64_t p0 = (int64_t)blockIdx.x * (TPB * KPT) + threadIdx.x;
  const double *vol = par + (int64_t)FVOL * cap;
  double v[KPT];
#pragma unroll
  for (int k = 0; k < KPT; k++) v[k] = vol[min(p0 + (int64_t)k * TPB, n - 1)];
#pragma unroll
  for (int k = 0; k < KPT; k++)
    if (p0 + (int64_t)k * TPB < n) keep[p0 + (int64_t)k * TPB] = (v[k] < vmin) ? 0 : 1;
}

// flags over cells in the reference's loop order (i outer, j, k inner; parcel_mixing.F90:160-162)
__global__ void k_newcell_flags(Geometry g, const int32_t *__restrict__ npercell,
                                uint8_t *__restrict__ cellflag, int64_t ncell) {
  const int64_t ord = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ord >= ncell) return;
  const int k = (int)(ord % g.ncz);
  const int j = (int)((ord / g.ncz) % g.ny);
  const int i = (int)(ord / ((int64_t)g.ncz * g.ny));
  const int64_t c = ((int64_t)k * g.ny + j) * g.nx + i;
  cellflag[ord] = (npercell[c] < 3) ? 1 : 0;
}

struct MixGrids {
  const double *G[6];  // kept vol, b, hg, p, q, r (folded, doubled)
  double *R[6];        // residual vres..rres
};

struct Coords {
  const double *xc, *yc, *zc;  // x_coords, y_coords, z_coords (reference values)
};

// one new parcel per sparse cell (parcel_mixing.F90:163-207 minus the residual update, which
// k_mix_residual applies per node in the same order)
__global__ void k_mix_new(Geometry g, double *__restrict__ par, int64_t cap, int64_t norig,
                          const int32_t *__restrict__ newcells, int64_t nadd,
                          const int32_t *__restrict__ zlist, int64_t nz_,
                          uint8_t *__restrict__ keep, MixGrids mg, Coords cc,
                          double *__restrict__ portion_cell, DevScalars *sc) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nadd) return;
  const int64_t ord = newcells[m];
  const int k = (int)(ord % g.ncz);
  const int j = (int)((ord / g.ncz) % g.ny);
  const int i = (int)(ord / ((int64_t)g.ncz * g.ny));
  const int64_t s = (m < nz_) ? (int64_t)zlist[m] : norig + (m - nz_);
  if (s + 2 > cap) {  // n > maxparcels after placing (parcel_mixing.F90:196-199)
    raise_status(sc, PMPIC_ERR_MAXPARCELS);
    return;
  }
  const int i1 = (i + 1 == g.nx) ? 0 : i + 1;
  const int j1 = (j + 1 == g.ny) ? 0 : j + 1;
  // sum over the 2x2x2 block in Fortran array-element order (z fastest, then y, then x)
  const int64_t q[8] = {gidx(g, k, j, i),   gidx(g, k + 1, j, i),   gidx(g, k, j1, i),
                        gidx(g, k + 1, j1, i), gidx(g, k, j, i1),  gidx(g, k + 1, j, i1),
                        gidx(g, k, j1, i1),  gidx(g, k + 1, j1, i1)};
  double s6[6];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    double a = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) a += mg.G[f][q[c]];
    s6[f] = a;
  }
  const double v = s6[0];
  const double portion = 2 * g.vmin / v;
  par[FX * cap + s] = 0.5 * (cc.xc[i + 2] + cc.xc[i + 3]);
  par[FY * cap + s] = 0.5 * (cc.yc[j + 2] + cc.yc[j + 3]);
  par[FZ * cap + s] = 0.5 * (cc.zc[k] + cc.zc[k + 1]);
  par[FVOL * cap + s] = 2 * g.vmin;
  par[FSTRETCH * cap + s] = 0;
  par[FB * cap + s] = s6[1] / v;
  par[FH * cap + s] = s6[2] / v;
  par[FP * cap + s] = s6[3] / v;
  par[FQ * cap + s] = s6[4] / v;
  par[FR * cap + s] = s6[5] / v;
  if (s + 1 < norig) keep[s] = 2;
  portion_cell[((int64_t)k * g.ny + j) * g.nx + i] = portion;
}

// residual minus the share of every adjacent new parcel, visiting cells in the reference's
// loop order, then the z-boundary doubling (parcel_mixing.F90:202-207, 221-240)
__global__ void k_mix_residual(Geometry g, MixGrids mg, const int32_t *__restrict__ npercell,
                               const double *__restrict__ portion_cell) {
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode) return;
  const int i = (int)(t % g.nx);
  const int j = (int)((t / g.nx) % g.ny);
  const int k = (int)(t / ((int64_t)g.nx * g.ny));
  // candidate cells in ascending halo'd index: the wrapped neighbour of node 0 comes last
  int ci[2], cj[2];
  if (i == 0) { ci[0] = 0; ci[1] = g.nx - 1; } else { ci[0] = i - 1; ci[1] = i; }
  if (j == 0) { cj[0] = 0; cj[1] = g.ny - 1; } else { cj[0] = j - 1; cj[1] = j; }
  double r[6], gv[6];
#pragma unroll
  for (int f = 0; f < 6; f++) { r[f] = mg.R[f][t]; gv[f] = mg.G[f][t]; }
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int kc = k - 1; kc <= k; kc++) {
        if (kc < 0 || kc >= g.ncz) continue;
        // a cell that received a new parcel: the test of the new-cell flags (npercell < 3),
        // read in cell order so consecutive nodes read consecutive cells
        const int64_t cn = ((int64_t)kc * g.ny + cj[b]) * g.nx + ci[a];
        if (npercell[cn] >= 3) continue;
        const double portion = portion_cell[cn];
#pragma unroll
        for (int f = 0; f < 6; f++) r[f] = r[f] - portion * gv[f];
      }
  const double mlt = (k == 0 || k == g.nzn - 1) ? 2. : 1.;
#pragma unroll
  for (int f = 0; f < 6; f++) mg.R[f][t] = r[f] * mlt;
}

// kept parcels absorb the residuals (parcel_mixing.F90:247-340)
__global__ void k_mix_absorb(Geometry g, double *__restrict__ par, int64_t cap, int64_t norig,
                             const uint8_t *__restrict__ keep, MixGrids mg) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= norig || keep[p] != 1) return;
  // all loads first (parcel values, then the gathers), all stores last
  const int fld[5] = {FB, FH, FP, FQ, FR};
  const double x = ld_stream(&par[FX * cap + p]), y = ld_stream(&par[FY * cap + p]),
               z = ld_stream(&par[FZ * cap + p]);
  const double v = ld_stream(&par[FVOL * cap + p]);
  double fv[5];
#pragma unroll
  for (int f = 0; f < 5; f++) fv[f] = ld_stream(&par[(int64_t)fld[f] * cap + p]);
  int i0, j0, k0;
  double delx, dely, delz;
  locate(g, x, y, z, i0, j0, k0, delx, dely, delz);
  const int i1 = (i0 + 1 == g.nx) ? 0 : i0 + 1;
  const int j1 = (j0 + 1 == g.ny) ? 0 : j0 + 1;
  const int64_t q[8] = {gidx(g, k0, j0, i0),     gidx(g, k0, j0, i1),
                        gidx(g, k0, j1, i0),     gidx(g, k0, j1, i1),
                        gidx(g, k0 + 1, j0, i0), gidx(g, k0 + 1, j0, i1),
                        gidx(g, k0 + 1, j1, i0), gidx(g, k0 + 1, j1, i1)};
  double w[8], g0[8];
  corner_weights(delx, dely, delz, v, w);
#pragma unroll
  for (int c = 0; c < 8; c++) g0[c] = mg.G[0][q[c]];
  double vn = v;
#pragma unroll
  for (int c = 0; c < 8; c++) vn = vn + w[c] * mg.R[0][q[c]] / g0[c];
  double out[5];
#pragma unroll
  for (int f = 0; f < 5; f++) {
    double acc = fv[f] * v;
#pragma unroll
    for (int c = 0; c < 8; c++) acc = acc + w[c] / g0[c] * mg.R[f + 1][q[c]];
    out[f] = acc / vn;
  }
  st_stream(&par[FVOL * cap + p], vn);
#pragma unroll
  for (int f = 0; f < 5; f++) st_stream(&par[(int64_t)fld[f] * cap + p], out[f]);
}

// backfill (parcel_mixing.F90:347-389): j-th remaining removed slot <- j-th kept slot from the
// end; only x, y, z, p, q, r, b, h, vol, tag and stretch are copied
__global__ void k_mix_backfill(double *__restrict__ par, int64_t cap,
                               const int32_t *__restrict__ dst, int64_t R,
                               const int32_t *__restrict__ k1, int64_t K1) {
  const int64_t jx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (jx >= R) return;
  const int64_t s = k1[K1 - 1 - jx], d = dst[jx];
  const int fl[11] = {FX, FY, FZ, FP, FQ, FR, FB, FH, FVOL, FTAG, FSTRETCH};
#pragma unroll
  for (int c = 0; c < 11; c++) par[(int64_t)fl[c] * cap + d] = par[(int64_t)fl[c] * cap + s];
}

}  // namespace

// The step's small read-backs go through pinned staging and wait on an event without the
// blocking-sync flag, which the host spins on: a pageable copy plus a stream synchronisation
// costs several times the latency.
int read_back(Ctx *c, void *host, const void *dev, size_t bytes) {
  if (bytes > 256) return PMPIC_ERR_ARG;
  PMPIC_CHECK(hipMemcpyAsync(c->pin, dev, bytes, hipMemcpyDeviceToHost, c->stream));
  PMPIC_CHECK(hipEventRecord(c->sync_ev, c->stream));
  PMPIC_CHECK(hipEventSynchronize(c->sync_ev));
  memcpy(host, c->pin, bytes);
  return PMPIC_OK;
}

int sync_scalars(Ctx *c) {
  PMPIC_TRY(read_back(c, &c->hsc, c->dsc, sizeof(DevScalars)));
  c->sync_seq++;
  return c->hsc.status;
}

int compact_flags(Ctx *c, const uint8_t *flags, uint8_t want, int64_t n, int32_t *out,
                  int64_t *d_count) {
  const int64_t nblk = cdiv(n > 0 ? n : 1, FPB);
  if (nblk > c->nblk_max) return PMPIC_ERR_ARG;
  int32_t *blk = c->blk, *pre = c->blk + nblk;
  hipLaunchKernelGGL(k_flag_count, dim3(nblk), dim3(TPB), 0, c->stream, flags, want, n, blk);
  PMPIC_TRY(scan_exclusive_i32(c, blk, pre, nblk));
  hipLaunchKernelGGL(k_flag_scatter, dim3(nblk), dim3(TPB), 0, c->stream, flags, want, n, pre, blk,
                     nblk, out, d_count);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

static int ensure_tmp(Ctx *c, int64_t need) {
  if (need <= c->tmp_cap) return PMPIC_OK;
  int64_t cap = need + need / 2 + 1024;
  if (cap > c->cap) cap = c->cap;
  if (c->tmp) PMPIC_CHECK(hipFree(c->tmp));
  c->tmp = nullptr;
  PMPIC_CHECK(hipMalloc(&c->tmp, sizeof(double) * NFIELDS * cap));
  c->tmp_cap = cap;
  return PMPIC_OK;
}

// permutation part of the haloswap; expects flag[] and the per-block direction counts
static int haloswap_permute(Ctx *c, int64_t n, int64_t nblk) {
  int32_t *blk = c->blk, *pre = c->blk + 9 * nblk;
  PMPIC_TRY(scan_exclusive_i32(c, blk, pre, 9 * nblk));
  hipLaunchKernelGGL(k_cross_total, dim3(1), dim3(1), 0, c->stream, pre, blk, nblk, c->dsc);
  // the staging buffer must hold every crossing parcel; read the count back
  PMPIC_TRY(read_back(c, &c->hsc.cross, &c->dsc->cross, sizeof(int64_t)));
  const int64_t K = c->hsc.cross;
  if (K == 0) return PMPIC_OK;
  PMPIC_TRY(ensure_tmp(c, K));
  hipLaunchKernelGGL(k_hs_gather, dim3(nblk), dim3(TPB), 0, c->stream, c->g, c->par, c->cap, n,
                     c->flag, pre, nblk, c->list_a, c->tmp, c->tmp_cap);
  int64_t sb = cdiv(K, TPB);
  if (sb > 4096) sb = 4096;
  hipLaunchKernelGGL(k_hs_scatter, dim3(sb), dim3(TPB), 0, c->stream, c->par, c->cap, c->tmp,
                     c->tmp_cap, c->list_a, c->dsc);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int rk_stage_update(Ctx *c, int stage, int64_t n, int stretch) {
  const int64_t nblk = cdiv(n > 0 ? n : 1, TPB);
  if (nblk > c->nblk_max) return PMPIC_ERR_ARG;
  GridSet gs = {{c->grid[GU], c->grid[GV], c->grid[GW], c->grid[GDP], c->grid[GDQ], c->grid[GDR]}};
  switch (stage) {
    case 1:
      PMPIC_TRY(launch_timed(c, T_RK, k_rk_stage<1>, dim3(nblk), dim3(TPB), 0, c->g, c->par,
                             c->cap, n, gs, stretch, c->dsc, c->flag, c->blk, nblk));
      break;
    case 2:
      PMPIC_TRY(launch_timed(c, T_RK, k_rk_stage<2>, dim3(nblk), dim3(TPB), 0, c->g, c->par,
                             c->cap, n, gs, stretch, c->dsc, c->flag, c->blk, nblk));
      break;
    case 3:
      PMPIC_TRY(launch_timed(c, T_RK, k_rk_stage<3>, dim3(nblk), dim3(TPB), 0, c->g, c->par,
                             c->cap, n, gs, stretch, c->dsc, c->flag, c->blk, nblk));
      break;
    default:
      PMPIC_TRY(launch_timed(c, T_RK, k_rk_stage<4>, dim3(nblk), dim3(TPB), 0, c->g, c->par,
                             c->cap, n, gs, stretch, c->dsc, c->flag, c->blk, nblk));
      break;
  }
  PMPIC_CHECK(hipGetLastError());
  return haloswap_permute(c, n, nblk);
}

int parcel_haloswap(Ctx *c, int64_t n, int reflect) {
  (void)reflect;
  const int64_t nblk = cdiv(n > 0 ? n : 1, TPB);
  if (nblk > c->nblk_max) return PMPIC_ERR_ARG;
  hipLaunchKernelGGL(k_reflect_classify, dim3(nblk), dim3(TPB), 0, c->stream, c->g, c->par, c->cap,
                     n, c->flag, c->blk, nblk);
  PMPIC_CHECK(hipGetLastError());
  return haloswap_permute(c, n, nblk);
}

// stage-4 splitting after the stretch update; then the second haloswap.  Updates c->n.
int parcel_split(Ctx *c) {
  const int64_t n = c->n;
  if (n > 0) {
    hipLaunchKernelGGL(k_split_flags, dim3(cdiv(n, TPB)), dim3(TPB), 0, c->stream, c->par, c->cap,
                       n, c->flag);
    PMPIC_TRY(compact_flags(c, c->flag, 1, n, c->list_b, &c->dsc->nsplit));
  } else {
    c->hsc.nsplit = 0;
    PMPIC_CHECK(hipMemcpyAsync(&c->dsc->nsplit, &c->hsc.nsplit, sizeof(int64_t),
                               hipMemcpyHostToDevice, c->stream));
  }
  PMPIC_TRY(sync_scalars(c));
  const int64_t ns = c->hsc.nsplit;
  if (ns > 0) {
    if (n + ns > c->cap) return PMPIC_ERR_MAXPARCELS;
    hipLaunchKernelGGL(k_split_apply, dim3(cdiv(ns, TPB)), dim3(TPB), 0, c->stream, c->par, c->cap,
                       n, c->list_b, ns, c->dsc);
    PMPIC_CHECK(hipGetLastError());
  }
  c->n = n + ns;
  return parcel_haloswap(c, c->n, 1);
}

// stage-4 mixing; updates c->n
int parcel_mix(Ctx *c) {
  const Geometry &g = c->g;
  const int64_t norig = c->n;
  PMPIC_TRY(mixing_grids(c, norig));
  if (norig > 0)
    hipLaunchKernelGGL(k_keep_flags, dim3(cdiv(norig, TPB * KPT)), dim3(TPB), 0, c->stream, c->par,
                       c->cap, norig, g.vmin, c->flag);
  // removed slots, ascending
  PMPIC_TRY(compact_flags(c, c->flag, 0, norig, c->list_a, &c->dsc->nzero));
  // sparse cells in the reference's loop order
  hipLaunchKernelGGL(k_newcell_flags, dim3(cdiv(c->ncell, TPB)), dim3(TPB), 0, c->stream, g,
                     c->npercell, c->cellflag, c->ncell);
  PMPIC_TRY(compact_flags(c, c->cellflag, 1, c->ncell, c->sorted, &c->dsc->nadd));
  PMPIC_TRY(sync_scalars(c));
  const int64_t Z = c->hsc.nzero, nadd = c->hsc.nadd;
  MixGrids mg = {{c->grid[GVOL], c->grid[GB], c->grid[GHG], c->grid[GP], c->grid[GQ], c->grid[GR]},
                 {c->kres[0], c->kres[1], c->kres[2], c->kres[3], c->kres[4], c->kres[5]}};
  Coords cc = {c->coords, c->coords + (g.nx + 4), c->coords + (g.nx + 4) + (g.ny + 4)};
  double *portion = c->partial;
  if (nadd > 0)
    hipLaunchKernelGGL(k_mix_new, dim3(cdiv(nadd, TPB)), dim3(TPB), 0, c->stream, g, c->par,
                       c->cap, norig, c->sorted, nadd, c->list_a, Z, c->flag, mg, cc, portion,
                       c->dsc);
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  hipLaunchKernelGGL(k_mix_residual, dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream, g, mg,
                     c->npercell, portion);
  if (norig > 0)
    hipLaunchKernelGGL(k_mix_absorb, dim3(cdiv(norig, TPB)), dim3(TPB), 0, c->stream, g, c->par,
                       c->cap, norig, c->flag, mg);
  PMPIC_CHECK(hipGetLastError());
  // new count (parcel_mixing.F90:217)
  int64_t np = norig;
  if (nadd > 0) {
    int64_t slast;
    if (nadd - 1 < Z) {
      int32_t v;
      PMPIC_TRY(read_back(c, &v, c->list_a + (nadd - 1), sizeof(int32_t)));
      slast = v;
    } else {
      slast = norig + (nadd - 1 - Z);
    }
    if (slast + 1 > np) np = slast + 1;
  }
  PMPIC_TRY(sync_scalars(c));
  c->last_removed = Z;
  c->last_added = nadd;
  if (np == norig) {
    // removed slots still marked 0 after placement
    int64_t R = 0;
    const int32_t *dst = nullptr;
    if (nadd < Z) {
      R = Z - nadd;
      dst = c->list_a + nadd;
    } else if (nadd == Z && Z > 0) {
      int32_t v;
      PMPIC_TRY(read_back(c, &v, c->list_a + (Z - 1), sizeof(int32_t)));
      if (v == norig - 1) {
        R = 1;
        dst = c->list_a + (Z - 1);
      }
    }
    if (R > 0) {
      PMPIC_TRY(compact_flags(c, c->flag, 1, norig, c->list_b, &c->dsc->nkeep1));
      PMPIC_TRY(sync_scalars(c));
      const int64_t K1 = c->hsc.nkeep1;
      if (R > K1) return PMPIC_ERR_BACKFILL;
      hipLaunchKernelGGL(k_mix_backfill, dim3(cdiv(R, TPB)), dim3(TPB), 0, c->stream, c->par,
                         c->cap, dst, R, c->list_b, K1);
      PMPIC_CHECK(hipGetLastError());
      np -= R;
    }
  }
  c->n = np;
  return PMPIC_OK;
}

}  // namespace pmpic
