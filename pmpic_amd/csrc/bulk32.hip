// FP32-bulk form of the per-timestep parcel path on MI355X (pmpic_config.precision = 1).
//
// Reference path: components/rk4_integrator/src/rk4_integrator.F90:95-319,
// model_core/src/parcels/parcel_interpolation.F90:186-481, model_core/src/parcels/haloswap.F90,
// components/parcel_splitting/src/parcel_splitting.F90:47-168,
// components/parcel_mixing/src/parcel_mixing.F90:68-624.
//
// The algorithm is the plain FP64 form's, step for step (same RK4 stage formulas, split and
// mix rules, slot placement of children and new parcels, same deterministic cell-binned
// par2grid); the representation is what changes:
//  * a position is the index of the cell holding the parcel (int32, c = (k*ny + j)*nx + i) plus
//    three FP32 offsets in [0,1) inside the cell, so it carries 24 bits relative to the cell
//    rather than to the whole domain; vorticity, buoyancy, humidity, volume and stretch are
//    FP32, and interpolation and the par2grid sums run in FP32;
//  * the RK4 origin of the position is (cell_o, offset_o) and the RK4 accumulator
//    dt/6 k1 + dt/3 k2 + dt/3 k3 + dt/6 k4 is kept in cells relative to it;
//  * the haloswap becomes index arithmetic: leaving through x or y wraps the cell index and the
//    z reflection mirrors (cell, offset) about the boundary, so parcels keep their slots (the
//    plain form reproduces the reference's send-to-self slot permutation instead);
//  * parcel tendencies dxdt..drdt are stored at stage 4 only (the values that survive a step).
// The spectral inversion, the timestep selection and the conservation sums stay FP64: the
// grids are a few percent of the parcel bytes.
#include "common.h"

#include <hipcub/hipcub.hpp>
#include <rocprim/iterator/transform_iterator.hpp>
#include <algorithm>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

// The FP32-bulk form is not held to the reference's operation order (its bar is the velocity
// error against the FP64 form), so multiply-adds may fuse here.
#pragma clang fp contract(fast)

namespace pmpic {

namespace {

constexpr int TPB = 256;
constexpr int NWAVE = TPB / 64;

// A parcel's cell is stored packed as k << 20 | j << 10 | i (nx, ny <= 1024, nz < 2048), so
// unpacking is shifts and masks; the linear cell index (k*ny + j)*nx + i is formed only where
// buckets are indexed.
__device__ __forceinline__ int32_t pack_cell(int i, int j, int k) { return (k << 20) | (j << 10) | i; }
__device__ __forceinline__ void cell_ijk(const Geometry &, int32_t c, int &i, int &j, int &k) {
  i = c & 1023;
  j = (c >> 10) & 1023;
  k = c >> 20;
}
__device__ __forceinline__ int32_t linear_cell(const Geometry &g, int32_t c) {
  return ((c >> 20) * g.ny + ((c >> 10) & 1023)) * g.nx + (c & 1023);
}

// Loads and stores at a uniform base plus a 32-bit byte offset (the global_load/store form
// with a scalar base), valid while index * sizeof(T) < 2^32: capacities are checked at set-up.
template <typename T>
__device__ __forceinline__ T ld(const T *base, uint32_t i) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                      i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st(T *base, uint32_t i, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(T)) = v;
}
// streaming forms for parcel fields that one kernel reads or writes once: they bypass the
// caches' retention so the grids being gathered stay resident
template <typename T>
__device__ __forceinline__ T ldnt(const T *base, uint32_t i) {
  return __builtin_nontemporal_load(reinterpret_cast<const T *>(
      reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void stnt(T *base, uint32_t i, T v) {
  __builtin_nontemporal_store(v, reinterpret_cast<T *>(reinterpret_cast<char *>(base) +
                                                       i * (uint32_t)sizeof(T)));
}

// fold an offset outside [0,1) into the cell index
__device__ __forceinline__ void fold1(int &i, float &o) {
  const float f = floorf(o);
  i += (int)f;
  o = o - f;
  if (o >= 1.0f) {  // o was a tiny negative number
    o -= 1.0f;
    i += 1;
  }
}

// cell (i, j, k) + offsets (tx, ty, tz) in cells, offsets anywhere -> cell index and offsets in
// [0,1): periodic in x and y, z mirrored about 0 and nz (haloswap.F90:447-454 on the cell
// scale; a position exactly on the top boundary stays in the top cell with offset 1).  Returns
// false if the position is still outside in z.
__device__ __forceinline__ bool place(const Geometry &g, int i, int j, int k, float tx, float ty,
                                      float tz, int32_t &cell, float &ox, float &oy, float &oz) {
  fold1(i, tx);
  fold1(j, ty);
  fold1(k, tz);
  // a stage moves a parcel by far less than a period, so one correction is the common case
  if (i < 0 || i >= g.nx) {
    i %= g.nx;
    if (i < 0) i += g.nx;
  }
  if (j < 0 || j >= g.ny) {
    j %= g.ny;
    if (j < 0) j += g.ny;
  }
  const int nz = g.ncz;
  if (k < 0) {  // z' = -z
    if (tz == 0.0f) {
      k = -k;
    } else {
      k = -k - 1;
      tz = 1.0f - tz;
      if (tz >= 1.0f) { tz -= 1.0f; k += 1; }
    }
  } else if (k >= nz && !(k == nz && tz == 0.0f)) {  // z' = 2 nz - z
    if (tz == 0.0f) {
      k = 2 * nz - k;
    } else {
      k = 2 * nz - k - 1;
      tz = 1.0f - tz;
      if (tz >= 1.0f) { tz -= 1.0f; k += 1; }
    }
  }
  if (k == nz && tz == 0.0f) {
    k = nz - 1;
    tz = 1.0f;
  }
  if (k < 0 || k >= nz) return false;
  cell = pack_cell(i, j, k);
  ox = tx;
  oy = ty;
  oz = tz;
  return true;
}

// node offsets of the 8 corners of cell (i, j, k), corner bits z<<2 | y<<1 | x
__device__ __forceinline__ void corner_nodes(const Geometry &g, int i, int j, int k, int32_t o[8]) {
  const int i1 = (i + 1 == g.nx) ? 0 : i + 1;
  const int j1 = (j + 1 == g.ny) ? 0 : j + 1;
  const int32_t b00 = (k * g.ny + j) * g.nx, b01 = (k * g.ny + j1) * g.nx;
  const int32_t b10 = ((k + 1) * g.ny + j) * g.nx, b11 = ((k + 1) * g.ny + j1) * g.nx;
  o[0] = b00 + i; o[1] = b00 + i1; o[2] = b01 + i; o[3] = b01 + i1;
  o[4] = b10 + i; o[5] = b10 + i1; o[6] = b11 + i; o[7] = b11 + i1;
}

// trilinear interpolation in the reference's order (parcel_interpolation.F90:309-321)
__device__ __forceinline__ float interp(const float *__restrict__ G, const int32_t o[8], float ox,
                                        float oy, float oz) {
  const float c00 = ld(G, o[0]) * (1 - oz) + ld(G, o[4]) * oz;
  const float c01 = ld(G, o[1]) * (1 - oz) + ld(G, o[5]) * oz;
  const float c10 = ld(G, o[2]) * (1 - oz) + ld(G, o[6]) * oz;
  const float c11 = ld(G, o[3]) * (1 - oz) + ld(G, o[7]) * oz;
  const float c0 = c00 * (1 - oy) + c10 * oy;
  const float c1 = c01 * (1 - oy) + c11 * oy;
  return c0 * (1 - ox) + c1 * ox;
}

// ---- host <-> device conversion of the parcel state ----
// axis a of the position: (x - bottom)/d -> cell index part and offset
__global__ void k32_pos_in(Geometry g, int axis, const double *__restrict__ x, int64_t n,
                           int32_t *__restrict__ cell, float *__restrict__ off, DevScalars *sc) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double bot = axis == 0 ? g.xbot : (axis == 1 ? g.ybot : g.zbot);
  const double d = axis == 0 ? g.dx : (axis == 1 ? g.dy : g.dz);
  const int ncells = axis == 0 ? g.nx : (axis == 1 ? g.ny : g.ncz);
  const double u = (x[p] - bot) / d;
  const double fu = floor(u);
  int i = (int)fu;
  float o = (float)(u - fu);
  if (o >= 1.0f) { o = 0.0f; i += 1; }
  if (axis < 2) {
    i %= ncells;
    if (i < 0) i += ncells;
  } else {
    if (i == ncells && o == 0.0f) { i = ncells - 1; o = 1.0f; }
    if (i < 0 || i >= ncells) {
      raise_status(sc, PMPIC_ERR_OUT_OF_BOX_Z);
      i = 0;
    }
  }
  off[p] = o;
  cell[p] = (axis == 0 ? 0 : cell[p]) | (i << (10 * axis));
}

__global__ void k32_pos_out(Geometry g, int axis, const int32_t *__restrict__ cell,
                            const float *__restrict__ off, int64_t n, double *__restrict__ x) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i, j, k;
  cell_ijk(g, cell[p], i, j, k);
  const int idx = axis == 0 ? i : (axis == 1 ? j : k);
  const double bot = axis == 0 ? g.xbot : (axis == 1 ? g.ybot : g.zbot);
  const double d = axis == 0 ? g.dx : (axis == 1 ? g.dy : g.dz);
  x[p] = bot + ((double)idx + (double)off[p]) * d;
}

__global__ void k32_from_f64(const double *__restrict__ a, int64_t n, float *__restrict__ b) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) b[p] = (float)a[p];
}

__global__ void k32_to_f64(const float *__restrict__ a, int64_t n, double *__restrict__ b) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) b[p] = (double)a[p];
}

// ---- cell binning ----
// A wavefront walks BIN_SEQ consecutive groups of 64 parcels in order and takes one atomic per
// run of equal cells, so a cell whose parcels are consecutive in slot order gets its ranks in
// ascending parcel order unless it straddles two wavefronts' ranges; the accumulation only has
// to reorder those cells.  A parcel of rank < capc is listed at slot[cell * capc + rank], so a
// typical cell needs no scan and no scatter pass.  A parcel of higher rank is appended, with
// its rank, to an overflow list; every cell holding more than capc parcels is then laid out
// densely, all its parcels in rank order, in `sorted` from ostart[cell] (an exclusive scan of
// those cells' counts).  capc = 0 lays out every cell that way.  capc is chosen per binning:
// CAPC_LOW while the mean is at most LOW_PPC parcels per cell, so a sparse grid does not read
// mostly empty entries, else CAPC.
constexpr int BIN_SEQ = 16;
constexpr int CAPC = 32;      // slot entries per cell of the table, and the full capacity
constexpr int CAPC_LOW = 16;
constexpr int LOW_PPC = 10;

__global__ void k32_bin_reset(int32_t *__restrict__ count, int64_t nc1, DevScalars *dsc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nc1) count[i] = 0;
  if (i == 0) dsc->novf = 0;
}

// capc = 0 (no slot table): every parcel's rank within its cell goes to rankv instead, and
// k32_dense_scatter lays all cells out densely after the scan
__global__ void k32_bin_slot(Geometry g, const int32_t *__restrict__ cellv, int64_t n, int capc,
                             int32_t *__restrict__ slot, int2 *__restrict__ ovf,
                             int32_t *__restrict__ rankv, int32_t *__restrict__ count,
                             DevScalars *dsc) {
  // parcel indices fit 32 bits (capacity < 2^30)
  const int lane = threadIdx.x & 63;
  const int32_t w0 = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64 * BIN_SEQ;
  // the parcel count from device memory: a graph replays this launch, sized for a bound on
  // the count, whatever the count is then (n is the launch's size only)
  (void)n;
  const int32_t n32 = (int32_t)dsc->n;
  if (n32 == 0 || w0 >= n32) return;
  // all count atomics of the wavefront are issued before any result is used, in group order;
  // ch[it] = cell << 6 | head lane of the parcel's run (the cell part is only used while the
  // slot table exists, ncell * capc < 2^30)
  int base[BIN_SEQ], ch[BIN_SEQ];
  int32_t cv[BIN_SEQ];  // every load issued before the first is used (clamped index, n >= 1)
#pragma unroll
  for (int it = 0; it < BIN_SEQ; it++)
    cv[it] = ld(cellv, (uint32_t)min(w0 + it * 64 + lane, n32 - 1));
#pragma unroll
  for (int it = 0; it < BIN_SEQ; it++) {
    const int32_t p = w0 + it * 64 + lane;
    const int32_t c = (p < n32) ? linear_cell(g, cv[it]) : -1;
    const int32_t cprev = __shfl_up(c, 1);
    const bool head = (lane == 0) || (c != cprev);
    const unsigned long long hm = __ballot(head);
    const unsigned long long upto = hm & ((2ull << lane) - 1ull);
    const int hl = 63 - __clzll(upto);
    const unsigned long long above = hm & ~((2ull << hl) - 1ull);
    const int next = above ? (__ffsll(above) - 1) : 64;
    base[it] = (head && c >= 0) ? atomicAdd(&count[c], next - hl) : 0;
    ch[it] = (int)((uint32_t)c << 6) | hl;
  }
  if (capc == 0) {
#pragma unroll
    for (int it = 0; it < BIN_SEQ; it++) {
      const int32_t p = w0 + it * 64 + lane;
      const int hl = ch[it] & 63;
      const int32_t rk = __shfl(base[it], hl) + (lane - hl);
      if (p < n32) st(rankv, (uint32_t)p, rk);
    }
    return;
  }
  bool anyover = false;
#pragma unroll
  for (int it = 0; it < BIN_SEQ; it++) {
    const int32_t p = w0 + it * 64 + lane;
    const int hl = ch[it] & 63;
    const int32_t rk = __shfl(base[it], hl) + (lane - hl);
    if (p < n32) {
      if (rk < capc) st(slot, (uint32_t)((ch[it] >> 6) * capc + rk), p);
      else anyover = true;
    }
  }
  if (__ballot(anyover) == 0ull) return;
#pragma unroll
  for (int it = 0; it < BIN_SEQ; it++) {
    const int32_t p = w0 + it * 64 + lane;
    const int hl = ch[it] & 63;
    const int32_t rk = __shfl(base[it], hl) + (lane - hl);
    const bool over = (p < n32) && rk >= capc;
    const unsigned long long om = __ballot(over);
    if (om) {
      const int lead = __ffsll(om) - 1;
      int32_t ob = 0;
      if (lane == lead) ob = atomicAdd(&dsc->novf, (int32_t)__popcll(om));
      ob = __shfl(ob, lead);
      if (over) ovf[ob + __popcll(om & ((1ull << lane) - 1ull))] = make_int2(p, rk);
    }
  }
}

struct OvfCount {
  int capc;
  __host__ __device__ int32_t operator()(int32_t k) const { return k > capc ? k : 0; }
};

// the overflow list into the dense region; the parcel of rank capc also copies its cell's
// capc slot entries there, so the region holds ranks 0 .. count-1
__global__ void k32_ovf_scatter(Geometry g, const int32_t *__restrict__ cellv,
                                const int2 *__restrict__ ovf, const DevScalars *dsc,
                                const int32_t *__restrict__ ostart,
                                const int32_t *__restrict__ slot, int capc,
                                int32_t *__restrict__ sorted) {
  const int32_t no = dsc->novf;
  const int32_t stride = gridDim.x * blockDim.x;
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < no; i += stride) {
    const int2 e = ovf[i];
    const int32_t c = linear_cell(g, cellv[e.x]);
    const int32_t b0 = ostart[c];
    sorted[b0 + e.y] = e.x;
    if (e.y == capc)
      for (int r = 0; r < capc; r++) sorted[b0 + r] = slot[c * capc + r];
  }
}

// without a slot table: parcel p at its cell's dense offset plus its rank
__global__ void k32_dense_scatter(Geometry g, const int32_t *__restrict__ cellv,
                                  const int32_t *__restrict__ rankv, const DevScalars *dsc,
                                  const int32_t *__restrict__ start, int32_t *__restrict__ sorted) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= dsc->n) return;
  const int32_t c = linear_cell(g, ld(cellv, (uint32_t)p));
  st(sorted, (uint32_t)(start[c] + ld(rankv, (uint32_t)p)), (int32_t)p);
}

__global__ void k32_sort_big(const int32_t *__restrict__ start, int64_t ncell, int rows,
                             int32_t *__restrict__ sorted) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int32_t b = start[c], e = start[c + 1];
  if (e - b <= rows) return;
  for (int32_t a = b + 1; a < e; a++) {
    const int32_t v = sorted[a];
    int32_t q = a - 1;
    while (q >= b && sorted[q] > v) {
      sorted[q + 1] = sorted[q];
      q--;
    }
    sorted[q + 1] = v;
  }
}

// ---- per-cell corner partial sums (the plain form's LDS kernel in FP32) ----
// A workgroup of 8 threads per cell takes a tile of TILE32 cells.  The tile's slot entries are
// loaded together with the cell counts (their addresses do not depend on the counts), the
// parcels' values are staged in LDS in dense rows, cell after cell, and thread q of a cell sums
// corner q over the cell's parcels in ascending parcel order, so the sums are deterministic.
// P2G: p, q, r and the total buoyancy (vort2vel + vorticity_tendency); MIX: kept / residual
// vol, b, h, p, q, r (parcel_mixing); DUMP: b, h, p, q, r (writebinary's grids)
// P2GX, MIXX: P2G and MIX whose tiles are 16 cells of one x row (nx a multiple of 16), each
// tile writing its corner sums combined along x into the 17 node positions it touches (see
// k32_node_sum_x); MIXX sums vol, p, q, r, the total buoyancy, b and h, so that its first five
// planes are the ones P2GX writes (the next par2grid keeps them where mixing changed nothing)
enum Acc32 { A32_P2G = 0, A32_MIX = 1, A32_DUMP = 2, A32_P2GX = 3, A32_MIXX = 4 };
constexpr int TILE32 = 16;
constexpr int ROWS32 = 512;     // LDS rows with the full capacity
constexpr int ROWS_LOW = 256;   // LDS rows with CAPC_LOW
__host__ __device__ constexpr int rows_of(int capc) { return capc == CAPC_LOW ? ROWS_LOW : ROWS32; }
constexpr int ATPB = TILE32 * 8;  // threads of the accumulation workgroup

struct Zmap {
  float zbot, dz;  // z of a parcel = zbot + (k + oz) dz
};

// A parcel's weight at corner (bx, by, bz) of its cell, in the form's one order of operations
// (every accumulation kernel and path uses it, so a cell's sums do not depend on which of them
// ran): the factor of an axis is |t - offset|, t = 1 for the lower corner (1 - offset) and 0 for
// the upper (the offset, which lies in [0, 1]); the weight is ((fz fy) vol) fx.
__device__ __forceinline__ float corner_weight(float ox, float oy, float oz, float vol, float tx,
                                               float ty, float tz) {
#pragma clang fp contract(off)
  const float fy = fabsf(ty - oy);
  const float fz = fabsf(tz - oz);
  const float wv = (fz * fy) * vol;
  return wv * fabsf(tx - ox);
}

// acc[0] += w, acc[1 + f] += w * v[f], fused.  No contraction here or in corner_weight: the
// compiler would otherwise fuse the weight's last product into the first sum in some of the
// kernels that inline these and not in others, and every accumulation path must give the same
// bits.
template <int NV>
__device__ __forceinline__ void add_parcel(float *acc, float w, const float *v) {
#pragma clang fp contract(off)
  acc[0] += w;
#pragma unroll
  for (int f = 0; f < NV; f++) acc[f + 1] = fmaf(w, v[f], acc[f + 1]);
}

// CAP: the binning's slot capacity (any value when capc = 0); ROWS: LDS rows of a pass
// The kernel is bound by memory latency, so waves in flight set its speed: with the 16-entry
// table the LDS rows allow 16 workgroups (32 waves) per CU, and holding the registers to 72
// allows 7 waves per SIMD (28 per CU) instead of 5 (4-5 % faster at 8 parcels per cell); the
// x-row form fits 64 registers without spilling, so all 8 (0.95 -> 0.90 ms at 256x256x128).
// The x-row mixing form's seven sums and three staged values per row allow 12 workgroups (6
// waves per SIMD) by LDS, so it takes the registers of 6 waves.
// The 32-entry form is bound by its LDS rows (8 workgroups per CU) instead.
template <int MODE, int CAP, int ROWS>
__global__ void __launch_bounds__(ATPB)
__attribute__((amdgpu_waves_per_eu(CAP == CAPC_LOW ? (MODE == A32_P2GX ? 8 : MODE == A32_MIXX ? 6 : 7) : 1))) k32_cell_accum(Bulk32 b, int64_t cap, Geometry g, Zmap zm,
                                                       float vmin, const int32_t *__restrict__ count,
                                                       const int32_t *__restrict__ ostart,
                                                       const int32_t *__restrict__ sorted,
                                                       int64_t ncell, float *__restrict__ partial,
                                                       float *__restrict__ partial2,
                                                       int32_t *__restrict__ npercell,
                                                       uint8_t *__restrict__ tileres,
                                                       const uint8_t *__restrict__ dirty) {
  constexpr bool MIX = (MODE == A32_MIX || MODE == A32_MIXX);
  constexpr bool XC = (MODE == A32_P2GX || MODE == A32_MIXX);
  // BT: the values start p, q, r and the total buoyancy (par2grid's four; the x-row mixing adds
  // b and h after them, so its first five planes are the ones the x-row par2grid writes);
  // FIVE: b, h, p, q, r (the other mixing form and the dump)
  constexpr bool BT = (MODE == A32_P2G || MODE == A32_P2GX || MODE == A32_MIXX);
  constexpr bool FIVE = !BT;
  constexpr int NV = MODE == A32_MIXX ? 6 : (FIVE ? 5 : 4);
  constexpr int NF = NV + 1;
  constexpr int SPT = TILE32 * CAP / ATPB;  // slot entries per thread
  constexpr int RPT = ROWS / ATPB;         // dense rows per thread
  __shared__ float4 sgeo[ROWS];  // ox, oy, oz, vol
  __shared__ float4 sva[ROWS];   // values 0..3
  __shared__ float svb[NV >= 5 ? ROWS : 1];
  __shared__ float svc[NV >= 6 ? ROWS : 1];
  __shared__ int32_t sidx[ROWS];
  __shared__ int16_t sperm[ROWS];  // the rows of a cell in parcel order, as byte offsets 16 r
  __shared__ uint8_t skeep[MIX ? ROWS : 1];
  __shared__ int32_t soff[TILE32 + 1];  // first dense row of each cell of the tile
  __shared__ int32_t scnt[TILE32], sost[TILE32];
  __shared__ uint32_t sover;            // bit cl: cell cl is listed in the dense overflow region
  __shared__ float skz[TILE32];         // z level of each cell of the tile
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int cl = tid >> 3;
  const int q = tid & 7;
  const int bx = q & 1, by = (q >> 1) & 1, bz = (q >> 2) & 1;
  // corner factor f = s * offset + t: offset for the upper corner, 1 - offset for the lower
  // corner factors |t - offset|: t = 1 at the lower corner, 0 at the upper
  const float tx = bx ? 0.f : 1.f, ty = by ? 0.f : 1.f, tz = bz ? 0.f : 1.f;
  const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
  // a par2grid whose partials the last x-row mixing sums left current: only the tiles mixing
  // changed (dirty) are summed again
  if (dirty && !dirty[tile]) return;
  const int64_t c0 = (int64_t)tile * TILE32;
  // cell, tile and row numbers are below 2^31, so their divisions are 32-bit; an x-row tile
  // has ntx tiles per row
  const uint32_t ntx = XC ? (uint32_t)g.nx / TILE32 : 1u;
  const int ntile = (int)min((int64_t)TILE32, ncell - c0);
  const int capc = b.capc;
  int32_t sl[SPT];
#pragma unroll
  for (int k = 0; k < SPT; k++) {
    const int e = tid + k * ATPB;
    sl[k] = (capc > 0 && e < ntile * CAP) ? ld(b.slot, (uint32_t)(c0 * CAP + e)) : -1;
  }
  if (tid < 64) {
    const int32_t v = (tid < ntile) ? count[c0 + tid] : 0;
    const int32_t o = (tid < ntile) ? ostart[c0 + tid] : 0;
    int32_t incl = v;
#pragma unroll
    for (int d = 1; d < TILE32; d <<= 1) {
      const int32_t u = __shfl_up(incl, d);
      if (lane >= d) incl += u;
    }
    const unsigned long long om = __ballot(tid < TILE32 && v > capc);
    if (tid < TILE32) {
      scnt[tid] = v;
      sost[tid] = o;
      soff[tid + 1] = incl;
      const uint32_t kc = XC ? tile / ntx / (uint32_t)g.ny
                             : (uint32_t)(c0 + tid) / ((uint32_t)g.nx * (uint32_t)g.ny);
      skz[tid] = (tid < ntile) ? (float)kc : 0.f;
    }
    if (tid == 0) {
      soff[0] = 0;
      sover = (uint32_t)om;
    }
  }
  __syncthreads();
  const float *F = b.f;
  auto fld = [&](int f) { return b.fp[f]; };
  // one parcel's staged values
  auto fetch = [&](uint32_t pp, float4 &geo, float4 &va, float &vb) {
    geo = make_float4(ldnt(fld(B_OX), pp), ldnt(fld(B_OY), pp), ldnt(fld(B_OZ), pp),
                      ldnt(fld(B_VOL), pp));
    if constexpr (BT) {
      va = make_float4(ldnt(fld(B_P), pp), ldnt(fld(B_Q), pp), ldnt(fld(B_R), pp),
                       ldnt(fld(B_B), pp));
      vb = ldnt(fld(B_H), pp);  // humidity, for the total buoyancy below
    } else {
      va = make_float4(ldnt(fld(B_B), pp), ldnt(fld(B_H), pp), ldnt(fld(B_P), pp),
                       ldnt(fld(B_Q), pp));
      vb = ldnt(fld(B_R), pp);
    }
  };
  // kz: the z level of the parcel's cell (known from the tile, so the cell is not read)
  auto put = [&](int r, int32_t pp, float4 geo, float4 v, float vb, float kz) {
    if constexpr (BT) {
      // total buoyancy (vorticity_tendency.F90:121-126)
      const float z = fmaf(kz + geo.z, zm.dz, zm.zbot);
      const float hl = vb - expf(-z);
      const float bv = v.w;
      v.w = fmaf(12.5f, hl > 0.f ? hl : 0.f, bv);
      if constexpr (NV >= 6) {
        svb[r] = bv;
        svc[r] = vb;
      }
    } else {
      svb[r] = vb;
    }
    if constexpr (MIX) skeep[r] = (geo.w < vmin) ? 0 : 1;
    sgeo[r] = geo;
    sva[r] = v;
    sidx[r] = pp;
  };
  float acc[NF], res[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) { acc[f] = 0.f; res[f] = 0.f; }
  int kept = 0, nres = 0;
  bool big = false;
  int first = 0;
  while (first < ntile) {
    int last = first + 1;
    if (soff[ntile] - soff[first] <= ROWS) last = ntile;  // the usual case: the rest fits
    else
      while (last < ntile && soff[last + 1] - soff[first] <= ROWS) last++;
    const int32_t r0 = soff[first];
    const int nrows = soff[last] - r0;
    const bool fits = nrows <= ROWS;
    if (fits) {
      // rows of the cells listed in the slot table, from the entries already loaded
      int32_t pr[SPT], row[SPT];
#pragma unroll
      for (int k = 0; k < SPT; k++) {
        const int e = tid + k * ATPB;
        const int ce = e / CAP, j = e % CAP;
        pr[k] = -1;
        row[k] = 0;
        if (sl[k] >= 0 && ce >= first && ce < last && j < scnt[ce] && scnt[ce] <= capc) {
          pr[k] = sl[k];
          row[k] = soff[ce] - r0 + j;
        }
      }
      float4 geo[SPT], va[SPT];
      float vb[SPT];
#pragma unroll
      for (int k = 0; k < SPT; k++)
        if (pr[k] >= 0) fetch((uint32_t)pr[k], geo[k], va[k], vb[k]);
#pragma unroll
      for (int k = 0; k < SPT; k++)
        if (pr[k] >= 0) put(row[k], pr[k], geo[k], va[k], vb[k], skz[(tid + k * ATPB) / CAP]);
      // rows of the cells listed in the dense overflow region
      const uint32_t pass = (last - first >= 32 ? ~0u : ((1u << (last - first)) - 1u)) << first;
      if (sover & pass) {
        for (int k = 0; k < RPT; k++) {
          const int r = tid + k * ATPB;
          if (r >= nrows) break;
          // the cell of dense row r: the last cell of the pass starting at or before it
          int lo_ = first, hi_ = last - 1;
          while (lo_ < hi_) {
            const int mid = (lo_ + hi_ + 1) >> 1;
            if (soff[mid] - r0 <= r) lo_ = mid;
            else hi_ = mid - 1;
          }
          if (!((sover >> lo_) & 1u)) continue;
          const int32_t pp = ld(sorted, (uint32_t)(sost[lo_] + (r + r0 - soff[lo_])));
          float4 g1, v1;
          float b1;
          fetch((uint32_t)pp, g1, v1, b1);
          put(r, pp, g1, v1, b1, skz[lo_]);
        }
      }
    }
    __syncthreads();
    const bool mine = fits && cl >= first && cl < last;
    const int lo = mine ? soff[cl] - r0 : 0, hi = mine ? soff[cl + 1] - r0 : 0;
    bool ord = true;
    for (int a = lo + 1 + q; a < hi; a += 8) ord &= (sidx[a - 1] < sidx[a]);
    const unsigned long long bad = __ballot(!ord);
    const bool sort_cell = ((bad >> (lane & ~7)) & 0xffull) != 0ull;
    if (sort_cell) {
      for (int a = lo + q; a < hi; a += 8) {
        const int32_t key = sidx[a];
        int rk = 0;
        for (int bb = lo; bb < hi; bb++) rk += (sidx[bb] < key) ? 1 : 0;
        sperm[lo + rk] = (int16_t)(a * 16);
      }
    } else {
      for (int a = lo + q; a < hi; a += 8) sperm[a] = (int16_t)(a * 16);
    }
    __syncthreads();
#pragma unroll 2
    for (int a = lo; a < hi; a++) {
      const int off = sperm[a];
      const float4 gr = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(sgeo) + off);
      const float4 v = *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(sva) + off);
      const float w = corner_weight(gr.x, gr.y, gr.z, gr.w, tx, ty, tz);
      float vals[6] = {v.x, v.y, v.z, v.w, 0.f, 0.f};
      if constexpr (NV >= 5) vals[4] = svb[off >> 4];
      if constexpr (NV >= 6) vals[5] = svc[off >> 4];
      if (!MIX || skeep[off >> 4]) {
        add_parcel<NV>(acc, w, vals);
        kept++;
      } else {
        nres++;
        add_parcel<NV>(res, w, vals);
      }
    }
    if (fits == false && cl == first) big = true;
    __syncthreads();
    first = last;
  }
  const bool live = cl < ntile;
  const int64_t c = c0 + cl;
  if (live && big) {
    // more parcels than LDS rows: always in the overflow region, sorted by parcel there
    const int32_t s0 = sost[cl], s1 = s0 + scnt[cl];
    for (int32_t s = s0; s < s1; s++) {
      const int64_t pp = sorted[s];
      const float ox = F[B_OX * cap + pp], oy = F[B_OY * cap + pp], oz = F[B_OZ * cap + pp];
      const float vol = F[B_VOL * cap + pp];
      float vals[6];
      if constexpr (BT) {
        const float z = fmaf(skz[cl] + oz, zm.dz, zm.zbot);
        const float hv = F[B_H * cap + pp], bv = F[B_B * cap + pp];
        const float hl = hv - expf(-z);
        vals[0] = F[B_P * cap + pp];
        vals[1] = F[B_Q * cap + pp];
        vals[2] = F[B_R * cap + pp];
        vals[3] = fmaf(12.5f, hl > 0.f ? hl : 0.f, bv);
        vals[4] = bv;
        vals[5] = hv;
      } else {
        vals[0] = F[B_B * cap + pp];
        vals[1] = F[B_H * cap + pp];
        vals[2] = F[B_P * cap + pp];
        vals[3] = F[B_Q * cap + pp];
        vals[4] = F[B_R * cap + pp];
      }
      const float w = corner_weight(ox, oy, oz, vol, tx, ty, tz);
      if (!MIX || !(vol < vmin)) {
        add_parcel<NV>(acc, w, vals);
        kept++;
      } else {
        nres++;
        add_parcel<NV>(res, w, vals);
      }
    }
  }
  if constexpr (XC) {
    // the tile's 16 cells lie in one x row: node position pos (0..16) of corner pair
    // bb = (by, bz) sums the bx = 1 corner of cell pos - 1 and the bx = 0 corner of cell pos
    // (left + right); positions 0 and 16 hold one side, completed by k32_node_sum_x.  Value f
    // of entry (bb, row, tx, pos) sits at ((f 4 + bb) nrows + row) W + 17 tx + pos, W = 17 ntx:
    // each value's plane is read coalesced by the node sums.
    static_assert(TILE32 == 16, "x-combined partials of 16-cell tiles");
    float *cor = reinterpret_cast<float *>(sgeo);  // free after the passes' last barrier
    static_assert(ATPB * NF <= ROWS * 4, "the corner sums fit the staging rows");
    const uint32_t xrow = tile / ntx;
    const int64_t row = xrow;
    const int tx = (int)(tile - xrow * ntx);
    const int64_t nrows = (int64_t)g.ny * g.ncz;
    const int64_t fstride = 4 * nrows * (int64_t)(ntx * 17);
    auto write_x = [&](const float *v, float *out) {
      __syncthreads();  // cor free (the passes' staging, or the previous write_x's reads)
#pragma unroll
      for (int f = 0; f < NF; f++) cor[tid * NF + f] = v[f];  // tid = cell * 8 + corner
      __syncthreads();
      if (tid < 4 * 17) {
        const int bb = tid / 17, pos = tid - bb * 17;
        float *dst = out + ((int64_t)bb * nrows + row) * (ntx * 17) + tx * 17 + pos;
        // cell pos - 1, bx = 1 and cell pos, bx = 0 (indices clamped into the tile; the ends
        // use one side only)
        const float *L = cor + ((pos > 0 ? pos - 1 : 0) * 8 + 1 + 2 * bb) * NF;
        const float *R = cor + ((pos < 16 ? pos : 15) * 8 + 2 * bb) * NF;
#pragma unroll
        for (int f = 0; f < NF; f++)
          dst[f * fstride] = pos == 0 ? R[f] : pos == 16 ? L[f] : L[f] + R[f];
      }
    };
    write_x(acc, partial);
    if constexpr (MIX) {
      // residual partials only from tiles that have residual parcels (tileres)
      const bool anyres = __syncthreads_or(nres > 0);
      if (tid == 0) tileres[tile] = anyres ? 1 : 0;
      if (q == 0) npercell[c0 + cl] = kept;
      if (anyres) write_x(res, partial2);
    }
    return;
  }
  // mixing: a tile without residual parcels writes no residual partials; tileres tells the
  // residual node sums which tiles have them
  bool anyres = false;
  if constexpr (MIX) {
    anyres = __syncthreads_or(live && nres > 0);
    if (tid == 0) tileres[tile] = anyres ? 1 : 0;
  }
  if (!live) return;
#pragma unroll
  for (int f = 0; f < NF; f++) partial[((int64_t)q * ncell + c) * NF + f] = acc[f];
  if constexpr (MIX) {
    if (anyres) {
#pragma unroll
      for (int f = 0; f < NF; f++) partial2[((int64_t)q * ncell + c) * NF + f] = res[f];
    }
    if (q == 0) npercell[c] = kept;
  }
}

// node sums of the FP32 corner partials (the plain form's k_node_sum): NORM divides by the
// summed weight (par2grid) into FP64 grids; otherwise raw FP32 sums, doubled on the z
// boundaries when DBL (mixing kept grids)
template <typename T>
struct Out32 {
  T *p[7];  // null: a plane not written
};

// tileres (may be null): per accumulation tile, whether its cells' partials were written; the
// cells of a tile without them contribute zero
template <int NV, bool NORM, bool DBL, typename T>
__global__ void k32_node_sum(Geometry g, const float *__restrict__ partial, int64_t ncell,
                             Out32<T> out, const uint8_t *__restrict__ tileres) {
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode) return;
  const int i = (int)(t % g.nx);
  const int j = (int)((t / g.nx) % g.ny);
  const int k = (int)(t / ((int64_t)g.nx * g.ny));
  // every corner's partials are loaded first (a corner below the bottom or above the top cell
  // reads cell 0 and is left out of the sum), then summed in the fixed corner order
  float v[8][NV + 1];
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
    if (ok[q] && tileres && !tileres[c / TILE32]) ok[q] = false;
#pragma unroll
    for (int f = 0; f <= NV; f++)
      v[q][f] = ok[q] ? partial[((int64_t)q * ncell + c) * (NV + 1) + f] : 0.f;
  }
  float s[NV + 1];
#pragma unroll
  for (int f = 0; f <= NV; f++) s[f] = 0.f;
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
    float w = s[0];
    if (w == 0.f) w = 1.f;
#pragma unroll
    for (int f = 1; f <= NV; f++) out.p[f - 1][t] = (T)(s[f] / w);
  } else {
    const float m = (DBL && (k == 0 || k == g.nzn - 1)) ? 2.f : 1.f;
#pragma unroll
    for (int f = 0; f <= NV; f++) out.p[f][t] = (T)(s[f] * m);
  }
}

// node sums of the x-combined partials of A32_P2GX / A32_MIXX (NF values: the weight and NF - 1
// weighted sums): node (i, j, k) adds, for each corner pair (by, bz) in a fixed order, the
// entry of cell row (j - by, k - bz) at its node position, and at a tile boundary the left
// tile's last position before the right tile's first (the same left + right order as inside a
// tile).  NORM divides by the summed weight into NF - 1 grids; otherwise the NF raw sums,
// doubled on the z boundaries when DBL.  tileres (may be null): per tile, whether its entries
// were written; the entries of a tile without them contribute zero.
template <int NF, bool NORM, bool DBL>
__global__ void k32_node_sum_x(Geometry g, const float *__restrict__ partial, Out32<float> out,
                               const uint8_t *__restrict__ tileres) {
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode) return;
  const int i = (int)(t % g.nx);
  const int j = (int)((t / g.nx) % g.ny);
  const int k = (int)(t / ((int64_t)g.nx * g.ny));
  const int ntx = g.nx / TILE32, tx = i / TILE32, pos = i - tx * TILE32;
  const int txl = tx == 0 ? ntx - 1 : tx - 1;  // the tile left of a boundary (periodic in x)
  const int64_t nrows = (int64_t)g.ny * g.ncz, w = (int64_t)ntx * 17;
  const int64_t fstride = 4 * nrows * w;  // value planes (the layout k32_cell_accum writes)
  float v[4][NF], vl[4][NF];
  bool ok[4], okl[4];
#pragma unroll
  for (int bb = 0; bb < 4; bb++) {
    const int by = bb & 1, bz = bb >> 1;
    int cj = j - by;
    if (cj < 0) cj += g.ny;
    const int ck = k - bz;
    const bool in = ck >= 0 && ck < g.ncz;
    const int64_t crow = in ? (int64_t)ck * g.ny + cj : 0;
    ok[bb] = in && (!tileres || tileres[crow * ntx + tx]);
    okl[bb] = in && pos == 0 && (!tileres || tileres[crow * ntx + txl]);
    const int64_t base = ((int64_t)bb * nrows + crow) * w;
    const float *e = partial + base + tx * 17 + pos;
    const float *el = partial + base + txl * 17 + 16;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      v[bb][f] = ok[bb] ? e[f * fstride] : 0.f;
      vl[bb][f] = okl[bb] ? el[f * fstride] : 0.f;
    }
  }
  float s[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) s[f] = 0.f;
#pragma unroll
  for (int bb = 0; bb < 4; bb++)
#pragma unroll
    for (int f = 0; f < NF; f++) {
      if (pos == 0) {
        if (okl[bb] || ok[bb]) s[f] = s[f] + (vl[bb][f] + v[bb][f]);
      } else if (ok[bb]) {
        s[f] = s[f] + v[bb][f];
      }
    }
  if constexpr (NORM) {
    float wsum = s[0];
    if (wsum == 0.f) wsum = 1.f;
#pragma unroll
    for (int f = 1; f < NF; f++)
      if (out.p[f - 1]) out.p[f - 1][t] = s[f] / wsum;
  } else {
    const float m = (DBL && (k == 0 || k == g.nzn - 1)) ? 2.f : 1.f;
#pragma unroll
    for (int f = 0; f < NF; f++)
      if (out.p[f]) out.p[f][t] = s[f] * m;
  }
}

__global__ void k32_grid_to_f64(const float *__restrict__ a, double *__restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (double)a[t];
}

// ---- RK4 stage: gather the six tendencies, update, place (wrap x, y; reflect z) ----
struct Grid32Set {
  const float *g[6];
};

struct Scale32 {
  float idx, idy, idz;  // cells per unit length
};

// binning fused into an RK stage: the cell lists k32_bin_slot would build from the cells the
// stage writes (on = 0: no binning)
struct BinOut {
  int on;
  int capc;              // slot capacity of the binning, 0 for the dense ranks
  int32_t *count;        // per cell, zeroed before the stage
  int32_t *slot;         // [cell * capc + rank]
  int2 *ovf;             // (parcel, rank) past the capacity; their number in DevScalars::novf
  int32_t *rank;         // per parcel, when capc = 0
};

// k32_bin_slot's binning of one parcel per lane (c: its cell as stored, or -1 for none): the
// lanes of a run of equal cells share one count atomic, issued by the run's first lane
__device__ __forceinline__ void bin_lane(const Geometry &g, int32_t cell, uint32_t p,
                                         const BinOut &bo, DevScalars *dsc) {
  const int lane = threadIdx.x & 63;
  const int32_t c = cell >= 0 ? linear_cell(g, cell) : -1;
  const int32_t cprev = __shfl_up(c, 1);
  const bool head = (lane == 0) || (c != cprev);
  const unsigned long long hm = __ballot(head);
  const unsigned long long upto = hm & ((2ull << lane) - 1ull);
  const int hl = 63 - __clzll(upto);
  const unsigned long long above = hm & ~((2ull << hl) - 1ull);
  const int next = above ? (__ffsll(above) - 1) : 64;
  const int base = (head && c >= 0) ? atomicAdd(&bo.count[c], next - hl) : 0;
  const int32_t rk = __shfl(base, hl) + (lane - hl);
  if (bo.capc == 0) {
    if (c >= 0) st(bo.rank, p, rk);
    return;
  }
  const bool over = c >= 0 && rk >= bo.capc;
  if (c >= 0 && !over) st(bo.slot, (uint32_t)(c * bo.capc + rk), (int32_t)p);
  const unsigned long long om = __ballot(over);
  if (om) {
    const int lead = __ffsll(om) - 1;
    int32_t ob = 0;
    if (lane == lead) ob = atomicAdd(&dsc->novf, (int32_t)__popcll(om));
    ob = __shfl(ob, lead);
    if (over) bo.ovf[ob + __popcll(om & ((1ull << lane) - 1ull))] = make_int2((int32_t)p, rk);
  }
}

// the RK stage of parcel p; returns its new cell, or -1 when it left the box (status raised)
template <int STAGE>
__device__ __forceinline__ int32_t rk_update(const Geometry &g, const Bulk32 &b, int64_t cap,
                                             const Grid32Set &gs, const Scale32 &sc3,
                                             int stretch, uint8_t *__restrict__ split_flag,
                                             DevScalars *sc, uint32_t p) {
  float *F = b.f;
  auto fld = [&](int f) { return F + (size_t)f * cap; };
  const int32_t cell = ldnt(b.cell, p);
  const float ox = ldnt(fld(B_OX), p), oy = ldnt(fld(B_OY), p), oz = ldnt(fld(B_OZ), p);
  int32_t cello;
  float oo[3], df[3] = {0.f, 0.f, 0.f}, vo[3] = {0.f, 0.f, 0.f}, vf[3] = {0.f, 0.f, 0.f};
  if (STAGE == 1) {
    cello = cell;
    oo[0] = ox;
    oo[1] = oy;
    oo[2] = oz;
#pragma unroll
    for (int a = 0; a < 3; a++) vo[a] = ldnt(fld(B_P + a), p);
  } else {
    cello = ldnt(b.cello, p);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      oo[a] = ldnt(fld(B_OXO + a), p);
      df[a] = ldnt(fld(B_DFX + a), p);
      if (STAGE != 4) vo[a] = ldnt(fld(B_PO + a), p);
      vf[a] = ldnt(fld(B_PF + a), p);
    }
  }
  const float s_old = (STAGE == 4 && stretch) ? ldnt(fld(B_STRETCH), p) : 0.f;
  int i, j, k;
  cell_ijk(g, cell, i, j, k);
  int32_t o[8];
  corner_nodes(g, i, j, k, o);
  float kv[6];
#pragma unroll
  for (int f = 0; f < 6; f++) kv[f] = interp(gs.g[f], o, ox, oy, oz);
  const double dtd = sc->dtm;
  const float dt = (float)dtd, dt6 = (float)(dtd / 6), dt2 = (float)(dtd / 2),
              dt3 = (float)(dtd / 3);
  const float kp[3] = {kv[0] * sc3.idx, kv[1] * sc3.idy, kv[2] * sc3.idz};
  float tpos[3], dfn[3], vn[3], vfn[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (STAGE == 1) {
      dfn[a] = dt6 * kp[a];
      tpos[a] = oo[a] + dt2 * kp[a];
      vfn[a] = vo[a] + dt6 * kv[3 + a];
      vn[a] = vo[a] + dt2 * kv[3 + a];
    } else if (STAGE == 2) {
      dfn[a] = df[a] + dt3 * kp[a];
      tpos[a] = oo[a] + dt2 * kp[a];
      vfn[a] = vf[a] + dt3 * kv[3 + a];
      vn[a] = vo[a] + dt2 * kv[3 + a];
    } else if (STAGE == 3) {
      dfn[a] = df[a] + dt3 * kp[a];
      tpos[a] = oo[a] + dt * kp[a];
      vfn[a] = vf[a] + dt3 * kv[3 + a];
      vn[a] = vo[a] + dt * kv[3 + a];
    } else {
      dfn[a] = df[a] + dt6 * kp[a];
      tpos[a] = oo[a] + dfn[a];
      vfn[a] = vf[a] + dt6 * kv[3 + a];
      vn[a] = vfn[a];
    }
  }
  int io, jo, ko;
  cell_ijk(g, cello, io, jo, ko);
  int32_t ncell_;
  float nox, noy, noz;
  if (!place(g, io, jo, ko, tpos[0], tpos[1], tpos[2], ncell_, nox, noy, noz)) {
    raise_status(sc, PMPIC_ERR_OUT_OF_BOX_Z);
    return -1;
  }
  stnt(b.cell, p, ncell_);
  stnt(fld(B_OX), p, nox);
  stnt(fld(B_OY), p, noy);
  stnt(fld(B_OZ), p, noz);
#pragma unroll
  for (int a = 0; a < 3; a++) stnt(fld(B_P + a), p, vn[a]);
  if (STAGE == 1) {
    stnt(b.cello, p, cello);
#pragma unroll
    for (int a = 0; a < 3; a++) {
      stnt(fld(B_OXO + a), p, oo[a]);
      stnt(fld(B_PO + a), p, vo[a]);
    }
  }
  if (STAGE != 4) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      stnt(fld(B_DFX + a), p, dfn[a]);
      stnt(fld(B_PF + a), p, vfn[a]);
    }
  } else {
#pragma unroll
    for (int f = 0; f < 6; f++) stnt(fld(B_DXDT + f), p, kv[f]);
    if (stretch) {
      const float third = 1.f / 3.f;
      float ds = vn[0] * kv[3] + vn[1] * kv[4] + vn[2] * kv[5];
      ds = powf(fabsf(ds), third);
      const float snew = s_old + ds * dt;
      stnt(fld(B_STRETCH), p, snew);
      split_flag[p] = (snew > 4.f) ? 1 : 0;  // the test of parcel_splitting.F90:76
    }
  }
  return ncell_;
}

template <int STAGE>
__global__ void __launch_bounds__(TPB) k32_rk_stage(Geometry g, Bulk32 b, int64_t cap, int64_t n,
                                                    Grid32Set gs, Scale32 sc3, int stretch,
                                                    uint8_t *__restrict__ split_flag,
                                                    DevScalars *sc, BinOut bo) {
  const int64_t p64 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  (void)n;  // launch size only: the count comes from device memory (see k32_bin_slot)
  const bool valid = p64 < sc->n;
  const uint32_t p = (uint32_t)p64;
  // with bo.on, the parcel's new cell is binned for the next stage's par2grid as k32_bin_slot
  // bins it; every lane of the wavefront then takes part, parcels past the count as empty
  if (bo.on) {
    const int32_t nc = valid ? rk_update<STAGE>(g, b, cap, gs, sc3, stretch, split_flag, sc, p)
                             : -1;
    bin_lane(g, nc, p, bo, sc);
    return;
  }
  if (!valid) return;
  rk_update<STAGE>(g, b, cap, gs, sc3, stretch, split_flag, sc, p);
}

// ---- splitting (parcel_splitting.F90:76-131, then the haloswap's wrap and reflection) ----
__global__ void k32_split_flags(const float *__restrict__ F, int64_t cap, int64_t n,
                                uint8_t *__restrict__ flag) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  flag[p] = (F[B_STRETCH * cap + p] > 4.f) ? 1 : 0;
}

__global__ void k32_split_apply(Geometry g, Bulk32 b, int64_t cap, int64_t n, Scale32 sc3,
                                const int32_t *__restrict__ parents, int64_t nsplit,
                                DevScalars *sc) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nsplit) return;
  const int64_t i = parents[m];
  const int64_t c = n + m;
  if (c >= cap) {
    raise_status(sc, PMPIC_ERR_MAXPARCELS);
    return;
  }
  float *F = b.f;
  const float third = 1.f / 3.f;
  const float pi = 3.14159265358979f;
  const float p = F[B_P * cap + i], q = F[B_Q * cap + i], r = F[B_R * cap + i];
  const float vol = F[B_VOL * cap + i];
  const float rr = powf(vol / 4.f / pi, third);
  const float absw = sqrtf(p * p + q * q + r * r);
  const float d[3] = {rr * p / absw * sc3.idx, rr * q / absw * sc3.idy, rr * r / absw * sc3.idz};
  int ci, cj, ck;
  cell_ijk(g, b.cell[i], ci, cj, ck);
  const float ox = F[B_OX * cap + i], oy = F[B_OY * cap + i], oz = F[B_OZ * cap + i];
  int32_t cc, cp;
  float cx, cy, cz, px, py, pz;
  const bool okc = place(g, ci, cj, ck, ox + d[0], oy + d[1], oz + d[2], cc, cx, cy, cz);
  const bool okp = place(g, ci, cj, ck, ox - d[0], oy - d[1], oz - d[2], cp, px, py, pz);
  if (!okc || !okp) {
    raise_status(sc, PMPIC_ERR_OUT_OF_BOX_Z);
    return;
  }
  b.cell[c] = cc;
  F[B_OX * cap + c] = cx;
  F[B_OY * cap + c] = cy;
  F[B_OZ * cap + c] = cz;
  F[B_P * cap + c] = p;
  F[B_Q * cap + c] = q;
  F[B_R * cap + c] = r;
  for (int f = B_DXDT; f <= B_DRDT; f++) F[(int64_t)f * cap + c] = F[(int64_t)f * cap + i];
  F[B_H * cap + c] = F[B_H * cap + i];
  F[B_B * cap + c] = F[B_B * cap + i];
  F[B_VOL * cap + c] = vol / 2;
  F[B_STRETCH * cap + c] = 0.f;
  b.tag[c] = b.tag[i];
  b.cell[i] = cp;
  F[B_OX * cap + i] = px;
  F[B_OY * cap + i] = py;
  F[B_OZ * cap + i] = pz;
  F[B_VOL * cap + i] = vol / 2;
  F[B_STRETCH * cap + i] = 0.f;
}

// ---- mixing (parcel_mixing.F90:157-389) ----
// KPT parcels per thread, at stride TPB (each a coalesced pass), the loads issued together
constexpr int KPT = 4;
__global__ void k32_keep_flags(const float *__restrict__ F, int64_t cap, int64_t n, float vmin,
                               uint8_t *__restrict__ keep) {
  const int64_t p0 = (int64_t)blockIdx.x * (TPB * KPT) + threadIdx.x;
  const float *vol = F + (int64_t)B_VOL * cap;
  float v[KPT];
#pragma unroll
  for (int k = 0; k < KPT; k++) v[k] = vol[min(p0 + (int64_t)k * TPB, n - 1)];
#pragma unroll
  for (int k = 0; k < KPT; k++)
    if (p0 + (int64_t)k * TPB < n) keep[p0 + (int64_t)k * TPB] = (v[k] < vmin) ? 0 : 1;
}

__global__ void k32_newcell_flags(Geometry g, const int32_t *__restrict__ npercell,
                                  uint8_t *__restrict__ cellflag, int64_t ncell) {
  const int64_t ord = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ord >= ncell) return;
  const int k = (int)(ord % g.ncz);
  const int j = (int)((ord / g.ncz) % g.ny);
  const int i = (int)(ord / ((int64_t)g.ncz * g.ny));
  const int64_t c = ((int64_t)k * g.ny + j) * g.nx + i;
  cellflag[ord] = (npercell[c] < 3) ? 1 : 0;
}

struct Mix32 {
  const float *G[6];  // kept vol, b, h, p, q, r (folded, doubled)
  float *R[6];        // residual vol .. r; after k32_mix_residual, divided by the kept volume
};

__global__ void k32_mix_new(Geometry g, Bulk32 b, int64_t cap, int64_t norig,
                            const int32_t *__restrict__ newcells, int64_t nadd,
                            const int32_t *__restrict__ zlist, int64_t nz_,
                            uint8_t *__restrict__ keep, Mix32 mg, float vmin,
                            float *__restrict__ portion_cell, DevScalars *sc) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nadd) return;
  const int64_t ord = newcells[m];
  const int k = (int)(ord % g.ncz);
  const int j = (int)((ord / g.ncz) % g.ny);
  const int i = (int)(ord / ((int64_t)g.ncz * g.ny));
  const int64_t s = (m < nz_) ? (int64_t)zlist[m] : norig + (m - nz_);
  if (s + 2 > cap) {
    raise_status(sc, PMPIC_ERR_MAXPARCELS);
    return;
  }
  const int i1 = (i + 1 == g.nx) ? 0 : i + 1;
  const int j1 = (j + 1 == g.ny) ? 0 : j + 1;
  // 2x2x2 block in Fortran array-element order (z fastest, then y, then x)
  const int64_t q[8] = {gidx(g, k, j, i),      gidx(g, k + 1, j, i),  gidx(g, k, j1, i),
                        gidx(g, k + 1, j1, i), gidx(g, k, j, i1),     gidx(g, k + 1, j, i1),
                        gidx(g, k, j1, i1),    gidx(g, k + 1, j1, i1)};
  float s6[6];
#pragma unroll
  for (int f = 0; f < 6; f++) {
    float a = 0.f;
#pragma unroll
    for (int c = 0; c < 8; c++) a += mg.G[f][q[c]];
    s6[f] = a;
  }
  const float v = s6[0];
  float *F = b.f;
  b.cell[s] = pack_cell(i, j, k);
  F[B_OX * cap + s] = 0.5f;  // the cell centre
  F[B_OY * cap + s] = 0.5f;
  F[B_OZ * cap + s] = 0.5f;
  F[B_VOL * cap + s] = 2 * vmin;
  F[B_STRETCH * cap + s] = 0.f;
  F[B_B * cap + s] = s6[1] / v;
  F[B_H * cap + s] = s6[2] / v;
  F[B_P * cap + s] = s6[3] / v;
  F[B_Q * cap + s] = s6[4] / v;
  F[B_R * cap + s] = s6[5] / v;
  if (s + 1 < norig) keep[s] = 2;
  portion_cell[((int64_t)k * g.ny + j) * g.nx + i] = 2 * vmin / v;
}

__global__ void k32_mix_residual(Geometry g, Mix32 mg, const int32_t *__restrict__ npercell,
                                 const float *__restrict__ portion_cell,
                                 uint8_t *__restrict__ touch, uint8_t *__restrict__ dirty) {
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnode) return;
  const int i = (int)(t % g.nx);
  const int j = (int)((t / g.nx) % g.ny);
  const int k = (int)(t / ((int64_t)g.nx * g.ny));
  int ci[2], cj[2];
  if (i == 0) { ci[0] = 0; ci[1] = g.nx - 1; } else { ci[0] = i - 1; ci[1] = i; }
  if (j == 0) { cj[0] = 0; cj[1] = g.ny - 1; } else { cj[0] = j - 1; cj[1] = j; }
  float r[6], gv[6];
#pragma unroll
  for (int f = 0; f < 6; f++) { r[f] = mg.R[f][t]; gv[f] = mg.G[f][t]; }
  for (int a = 0; a < 2; a++)
    for (int bb = 0; bb < 2; bb++)
      for (int kc = k - 1; kc <= k; kc++) {
        if (kc < 0 || kc >= g.ncz) continue;
        // a cell that received a new parcel: the test of the new-cell flags (npercell < 3),
        // read in cell order so consecutive nodes read consecutive cells
        const int64_t cn = ((int64_t)kc * g.ny + cj[bb]) * g.nx + ci[a];
        if (npercell[cn] >= 3) continue;
        const float portion = portion_cell[cn];
#pragma unroll
        for (int f = 0; f < 6; f++) r[f] = r[f] - portion * gv[f];
      }
  const float mlt = (k == 0 || k == g.nzn - 1) ? 2.f : 1.f;
  // stored divided by the node's kept volume: the absorption weights each node's residual by
  // w / vol_g (parcel_mixing.F90:247-340), so the division is done once per node here
  bool any = false;
#pragma unroll
  for (int f = 0; f < 6; f++) {
    const float v = (r[f] * mlt) / gv[0];
    mg.R[f][t] = v;
    any = any || v != 0.f;  // a NaN counts as not zero
  }
  touch[t] = any ? 1 : 0;
  if (any && dirty) {
    // the parcels of the eight cells around the node absorb: their tiles are summed again
    const int ntx = g.nx / TILE32;
    for (int kc = k - 1; kc <= k; kc++) {
      if (kc < 0 || kc >= g.ncz) continue;
      for (int bb = 0; bb < 2; bb++)
        for (int a = 0; a < 2; a++)
          dirty[((int64_t)kc * g.ny + cj[bb]) * ntx + ci[a] / TILE32] = 1;
    }
  }
}

// dirty tiles (x-row grids): the tile of each listed parcel's cell (parcels != 0) or of each
// listed new cell, given in the new-cell flags' order (z fastest, then y, then x)
__global__ void k32_mark_tiles(Geometry g, const int32_t *__restrict__ cellv,
                               const int32_t *__restrict__ list, int64_t n, int parcels,
                               uint8_t *__restrict__ dirty) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t e = list[t];
  int64_t lc;
  if (parcels) {
    lc = linear_cell(g, cellv[e]);
  } else {
    const int k = (int)(e % g.ncz);
    const int j = (int)((e / g.ncz) % g.ny);
    const int i = (int)(e / ((int64_t)g.ncz * g.ny));
    lc = ((int64_t)k * g.ny + j) * g.nx + i;
  }
  dirty[lc / TILE32] = 1;
}

// *any = 1 if a byte of a[0, n) is set (a is read a 32-bit word at a time; n a multiple of 4
// or the tail read as bytes)
__global__ void k32_any_u8(const uint8_t *__restrict__ a, int64_t n, int64_t *any) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool set = false;
  if (4 * t + 3 < n) {
    set = *reinterpret_cast<const uint32_t *>(a + 4 * t) != 0u;
  } else {
    for (int64_t e = 4 * t; e < n; e++) set = set || a[e] != 0;
  }
  if (__ballot(set) != 0ull && (threadIdx.x & 63) == 0) *any = 1;
}

__global__ void k32_zero_i64(int64_t *a) {
  if (threadIdx.x == 0) *a = 0;
}

__global__ void k32_fill_u8(uint8_t *__restrict__ a, int64_t n, uint8_t v) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// A parcel none of whose corner nodes carries a residual (almost all of them: residuals arise
// only around removed parcels and the cells that received new ones) keeps its values as they
// are, without the reads, the arithmetic (which would return them to within rounding) and the
// writes; `touch` marks the nodes with a residual.
__global__ void k32_mix_absorb(Geometry g, Bulk32 b, int64_t cap, int64_t norig,
                               const uint8_t *__restrict__ keep, Mix32 mg,
                               const uint8_t *__restrict__ touch) {
  const int64_t p64 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p64 >= norig || keep[p64] != 1) return;
  const uint32_t p = (uint32_t)p64;
  const int32_t cell = ldnt(b.cell, p);
  int i, j, k;
  cell_ijk(g, cell, i, j, k);
  int32_t o[8];
  corner_nodes(g, i, j, k, o);
  int any = 0;
#pragma unroll
  for (int c = 0; c < 8; c++) any |= touch[o[c]];
  if (!any) return;
  float *F = b.f;
  auto fp = [&](int f) { return F + (size_t)f * cap; };
  const int fld[5] = {B_B, B_H, B_P, B_Q, B_R};
  const float ox = ldnt(fp(B_OX), p), oy = ldnt(fp(B_OY), p), oz = ldnt(fp(B_OZ), p);
  const float v = ldnt(fp(B_VOL), p);
  float fv[5];
#pragma unroll
  for (int f = 0; f < 5; f++) fv[f] = ldnt(fp(fld[f]), p);
  float w[8];
  w[0] = (1 - oz) * (1 - oy) * (1 - ox) * v;
  w[1] = (1 - oz) * (1 - oy) * (ox) * v;
  w[2] = (1 - oz) * (oy) * (1 - ox) * v;
  w[3] = (1 - oz) * (oy) * (ox) * v;
  w[4] = (oz) * (1 - oy) * (1 - ox) * v;
  w[5] = (oz) * (1 - oy) * (ox) * v;
  w[6] = (oz) * (oy) * (1 - ox) * v;
  w[7] = (oz) * (oy) * (ox) * v;
  // the residual grids already hold residual / vol_g; 1 / vol_new once per parcel
  float vn = v;
#pragma unroll
  for (int c = 0; c < 8; c++) vn = vn + w[c] * ld((const float *)mg.R[0], (uint32_t)o[c]);
  const float ivn = 1.f / vn;
  float out[5];
#pragma unroll
  for (int f = 0; f < 5; f++) {
    float acc = fv[f] * v;
#pragma unroll
    for (int c = 0; c < 8; c++) acc = acc + w[c] * ld((const float *)mg.R[f + 1], (uint32_t)o[c]);
    out[f] = acc * ivn;
  }
  stnt(fp(B_VOL), p, vn);
#pragma unroll
  for (int f = 0; f < 5; f++) stnt(fp(fld[f]), p, out[f]);
}

// backfill: the state fields of the j-th kept slot from the end into the j-th removed slot
// (the reference copies x, y, z, p, q, r, b, h, vol, tag and stretch)
__global__ void k32_mix_backfill(Geometry g, Bulk32 b, int64_t cap,
                                 const int32_t *__restrict__ dst, int64_t R,
                                 const int32_t *__restrict__ k1, int64_t K1,
                                 uint8_t *__restrict__ dirty) {
  const int64_t jx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (jx >= R) return;
  const int64_t s = k1[K1 - 1 - jx], d = dst[jx];
  // a moved parcel changes the parcel order of its cell's sums
  if (dirty) dirty[linear_cell(g, b.cell[s]) / TILE32] = 1;
  float *F = b.f;
  const int fl[10] = {B_OX, B_OY, B_OZ, B_P, B_Q, B_R, B_B, B_H, B_VOL, B_STRETCH};
#pragma unroll
  for (int c = 0; c < 10; c++) F[(int64_t)fl[c] * cap + d] = F[(int64_t)fl[c] * cap + s];
  b.cell[d] = b.cell[s];
  b.tag[d] = b.tag[s];
}

Zmap zmap_of(const Geometry &g) { return Zmap{(float)g.zbot, (float)g.dz}; }
Scale32 scale_of(const Geometry &g) {
  return Scale32{(float)(1.0 / g.dx), (float)(1.0 / g.dy), (float)(1.0 / g.dz)};
}

template <typename T>
int dalloc32(T **p, int64_t n) {
  *p = nullptr;
  if (n <= 0) n = 1;
  if (hipMalloc((void **)p, sizeof(T) * (size_t)n) != hipSuccess) return PMPIC_ERR_NOMEM;
  return PMPIC_OK;
}

int bin32(Ctx *c, int64_t n) {
  const int64_t nc1 = c->ncell + 1;
  Bulk32 &b = c->b32;
  // the capacity of this binning: 16 entries while the mean is at most LOW_PPC parcels per
  // cell (fewer empty entries to read), else 32
  const int capc = b32_choose_capc(c, n);
  // the launches of a captured graph cover a parcel bound (the kernels read the count)
  const int64_t ln = c->graph_n ? c->graph_n : n;
  if (b.prebinned && b.capc == capc) {
    // the last RK stage binned the cells it wrote (same count, same capacity)
    b.prebinned = false;
  } else {
    b.prebinned = false;
    b.capc = capc;
    hipLaunchKernelGGL(k32_bin_reset, dim3(cdiv(nc1, TPB)), dim3(TPB), 0, c->stream,
                       c->cell_count, nc1, c->dsc);
    if (ln > 0)
      hipLaunchKernelGGL(k32_bin_slot, dim3(cdiv(ln, (int64_t)TPB * BIN_SEQ)), dim3(TPB), 0,
                         c->stream, c->g, b.cell, ln, b.capc, b.slot, b.ovf, c->rank,
                         c->cell_count, c->dsc);
  }
  // dense-region offsets of the cells past the slot capacity
  auto it = rocprim::make_transform_iterator(static_cast<const int32_t *>(c->cell_count),
                                             OvfCount{b.capc});
  size_t need = 0;
  PMPIC_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, need, it, c->cell_start, (int)nc1,
                                               c->stream));
  if (need > c->scan_tmp_bytes) {
    if (c->scan_tmp) PMPIC_CHECK(hipFree(c->scan_tmp));
    c->scan_tmp = nullptr;
    PMPIC_CHECK(hipMalloc(&c->scan_tmp, need));
    c->scan_tmp_bytes = need;
  }
  size_t have = c->scan_tmp_bytes;
  PMPIC_CHECK(hipcub::DeviceScan::ExclusiveSum(c->scan_tmp, have, it, c->cell_start, (int)nc1,
                                               c->stream));
  if (ln > 0 && b.capc == 0) {
    hipLaunchKernelGGL(k32_dense_scatter, dim3(cdiv(ln, TPB)), dim3(TPB), 0, c->stream, c->g,
                       b.cell, c->rank, c->dsc, c->cell_start, c->sorted);
  } else if (ln > 0) {
    const int64_t nb = std::min<int64_t>(cdiv(ln, TPB), 1024);
    hipLaunchKernelGGL(k32_ovf_scatter, dim3(nb), dim3(TPB), 0, c->stream, c->g, b.cell, b.ovf,
                       c->dsc, c->cell_start, b.slot, b.capc, c->sorted);
  }
  hipLaunchKernelGGL(k32_sort_big, dim3(cdiv(c->ncell, TPB)), dim3(TPB), 0, c->stream,
                     c->cell_start, c->ncell, rows_of(b.capc), c->sorted);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

// FP32 corner partials live in the plain form's partial buffers (48 doubles per cell hold
// 96 floats)
float *partial32(Ctx *c) { return reinterpret_cast<float *>(c->partial); }
float *partial32b(Ctx *c) { return reinterpret_cast<float *>(c->partial2); }

}  // namespace

__global__ void k32_set_n(DevScalars *sc, int64_t n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) sc->n = n;
}

int b32_publish_n(Ctx *c) {
  hipLaunchKernelGGL(k32_set_n, dim3(1), dim3(64), 0, c->stream, c->dsc, c->n);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

// 16 entries while the mean is at most LOW_PPC parcels per cell (fewer empty entries to read),
// else 32; a capacity fixed from the environment or picked by the tuning wins; never more than
// the table holds
int b32_mean_capc(const Ctx *c, int64_t n) {
  const int capc = (n <= (int64_t)LOW_PPC * c->ncell) ? CAPC_LOW : CAPC;
  return capc > c->b32.capmax ? c->b32.capmax : capc;
}

int b32_choose_capc(const Ctx *c, int64_t n) {
  const Bulk32 &b = c->b32;
  int capc;
  if (b.capmax == 0) capc = 0;
  else if (b.capfix) capc = b.capfix;
  else if (b.captuned) capc = b.captuned;
  else capc = (n <= (int64_t)LOW_PPC * c->ncell) ? CAPC_LOW : CAPC;
  return capc > b.capmax ? b.capmax : capc;
}

int b32_alloc(Ctx *c) {
  Bulk32 &b = c->b32;
  for (int f = 0; f < 6; f++) b.kept_pending[f] = false;
  PMPIC_TRY(dalloc32(&b.f, (int64_t)NF32 * c->cap));
  for (int k = 0; k < NF32; k++) b.fp[k] = b.f + (int64_t)k * c->cap;
  PMPIC_TRY(dalloc32(&b.cell, c->cap));
  PMPIC_TRY(dalloc32(&b.cello, c->cap));
  PMPIC_TRY(dalloc32(&b.tag, c->cap));
  // a parcel created by mixing takes neither a tag nor tendencies (parcel_mixing.F90 copies
  // tags only when it backfills, :377, and never writes dxdt..drdt): a slot past every earlier
  // count reads 0 for them, as in the FP64 form's arrays, which are cleared at allocation too
  PMPIC_CHECK(hipMemset(b.f, 0, sizeof(float) * (size_t)NF32 * (size_t)c->cap));
  PMPIC_CHECK(hipMemset(b.tag, 0, sizeof(double) * (size_t)c->cap));
  for (int i = 0; i < 6; i++) PMPIC_TRY(dalloc32(&b.grid[i], c->gsz));
  for (int i = 0; i < 12; i++) PMPIC_TRY(dalloc32(&b.mgrid[i], c->gsz));
  PMPIC_TRY(dalloc32(&b.mtouch, c->gsz));
  PMPIC_TRY(dalloc32(&b.dirty, c->ncell / TILE32 + 1));
  b.p2g_reuse = false;
  b.dirty_seq = 0;
  const char *renv = getenv("PMPIC_P2G_REUSE");
  b.reuse_off = renv && renv[0] == '0';
  for (int i = 0; i < 4; i++) {
    PMPIC_TRY(dalloc32(&b.pgrid[i], c->gsz));
    b.p2g_pending[i] = false;
  }
  // 32-bit byte offsets into the slot table.  PMPIC_SLOT_CAPACITY in the environment: 0 drops
  // the table (every cell then binned through the dense region, as grids beyond that range
  // are); 16 or 32 fixes the capacity instead of choosing it by the mean parcels per cell;
  // "tune" lets the first par2grid time the accumulation at both and keep the faster
  // (a table of 16 entries per cell reaches 2^30 entries at 512 x 512 x 256 cells)
  b.capmax = (c->ncell * CAPC <= ((int64_t)1 << 30))       ? CAPC
             : (c->ncell * CAPC_LOW <= ((int64_t)1 << 30)) ? CAPC_LOW
                                                            : 0;
  b.capfix = 0;
  b.captune = false;
  b.captuned = 0;
  b.rk_tpb = TPB;
  // binning fused into RK stages 1-3: it costs the RK kernel more than the separate binning
  // pass on small grids (the 32^3..96^3 ensemble members: 1% slower) and saves on large ones
  // (256 x 256 x 128: 1.6% faster), so from 2^21 cells on; PMPIC_FUSED_BINNING=0 or 1 forces it
  const char *fenv = getenv("PMPIC_FUSED_BINNING");
  b.fused_bin = fenv ? fenv[0] == '1' : c->ncell >= ((int64_t)1 << 21);
  b.prebinned = false;
  const char *env = getenv("PMPIC_SLOT_CAPACITY");
  if (env && strcmp(env, "tune") == 0) {
    b.captune = b.capmax == CAPC;
  } else if (env) {
    const int v = atoi(env);
    if (v == 0) b.capmax = 0;
    else if (v == CAPC_LOW || v == CAPC) b.capfix = v;
  }
  b.capc = b.capmax;
  b.slot = nullptr;
  if (b.capmax > 0) PMPIC_TRY(dalloc32(&b.slot, c->ncell * b.capmax));
  PMPIC_TRY(dalloc32(&b.ovf, c->cap));
  return PMPIC_OK;
}

void b32_free(Ctx *c) {
  Bulk32 &b = c->b32;
  hipFree(b.f);
  hipFree(b.cell);
  hipFree(b.cello);
  hipFree(b.tag);
  for (int i = 0; i < 6; i++) hipFree(b.grid[i]);
  for (int i = 0; i < 12; i++) hipFree(b.mgrid[i]);
  hipFree(b.mtouch);
  hipFree(b.dirty);
  for (int i = 0; i < 4; i++) hipFree(b.pgrid[i]);
  hipFree(b.slot);
  hipFree(b.ovf);
}

int b32_set_parcels(Ctx *c, int64_t n, const double *const *fields) {
  PMPIC_TRY(ensure_fbuf(c));
  Bulk32 &b = c->b32;
  b.prebinned = false;
  b.p2g_reuse = false;
  const dim3 gb(cdiv(n > 0 ? n : 1, TPB)), tb(TPB);
  for (int a = 0; a < 3; a++) {
    if (n > 0) {
      PMPIC_CHECK(hipMemcpyAsync(c->fbuf, fields[a], sizeof(double) * n, hipMemcpyHostToDevice,
                                 c->stream));
      hipLaunchKernelGGL(k32_pos_in, gb, tb, 0, c->stream, c->g, a, c->fbuf, n, b.cell,
                         b.f + (int64_t)(B_OX + a) * c->cap, c->dsc);
    }
  }
  for (int f = FP; f <= FSTRETCH; f++) {
    if (n > 0) {
      PMPIC_CHECK(hipMemcpyAsync(c->fbuf, fields[f], sizeof(double) * n, hipMemcpyHostToDevice,
                                 c->stream));
      hipLaunchKernelGGL(k32_from_f64, gb, tb, 0, c->stream, c->fbuf, n,
                         b.f + (int64_t)f * c->cap);
    }
  }
  if (n > 0)
    PMPIC_CHECK(hipMemcpyAsync(b.tag, fields[FTAG], sizeof(double) * n, hipMemcpyHostToDevice,
                               c->stream));
  PMPIC_CHECK(hipGetLastError());
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  return PMPIC_OK;
}

int b32_get_parcels(Ctx *c, double *const *fields) {
  PMPIC_TRY(ensure_fbuf(c));
  Bulk32 &b = c->b32;
  const int64_t n = c->n;
  if (n == 0) return PMPIC_OK;
  const dim3 gb(cdiv(n, TPB)), tb(TPB);
  for (int f = 0; f < 17; f++) {
    if (!fields[f]) continue;
    if (f == FTAG) {
      PMPIC_CHECK(hipMemcpyAsync(fields[f], b.tag, sizeof(double) * n, hipMemcpyDeviceToHost,
                                 c->stream));
      continue;
    }
    if (f < 3)
      hipLaunchKernelGGL(k32_pos_out, gb, tb, 0, c->stream, c->g, f, b.cell,
                         b.f + (int64_t)(B_OX + f) * c->cap, n, c->fbuf);
    else
      hipLaunchKernelGGL(k32_to_f64, gb, tb, 0, c->stream, b.f + (int64_t)f * c->cap, n, c->fbuf);
    PMPIC_CHECK(hipMemcpyAsync(fields[f], c->fbuf, sizeof(double) * n, hipMemcpyDeviceToHost,
                               c->stream));
  }
  PMPIC_CHECK(hipGetLastError());
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  return PMPIC_OK;
}

// the x-combined partials need tiles within one x row
static bool p2g_xrows(const Ctx *c) { return c->g.nx % TILE32 == 0; }

// A graph of stages 1-3 always captures stage 1's par2grid reading the dirty-tile flags (on
// x-row grids with the reuse on); where the last mixing's sums are not current, every tile is
// marked dirty first, so the captured par2grid sums them all.
void b32_graph_reuse(Ctx *c) {
  Bulk32 &b = c->b32;
  if (!p2g_xrows(c) || b.reuse_off) return;
  if (!b.p2g_reuse) {
    const int64_t nt = c->ncell / TILE32 + 1;
    hipLaunchKernelGGL(k32_fill_u8, dim3(cdiv(nt, TPB)), dim3(TPB), 0, c->stream, b.dirty, nt,
                       (uint8_t)1);
  }
  b.p2g_reuse = true;
}

// the par2grid accumulation; with `dirty`, only the tiles it marks (the others keep the
// partials the last mixing summed): that launch is left out of the kernel timer, whose
// per-launch figures are those of full launches
static int p2g_accum(Ctx *c, const uint8_t *dirty = nullptr) {
  const dim3 grid(cdiv(c->ncell, TILE32)), blk(ATPB);
  const Bulk32 &b = c->b32;
  if (p2g_xrows(c)) {
    auto k = b.capc == CAPC_LOW ? k32_cell_accum<A32_P2GX, CAPC_LOW, ROWS_LOW>
                                : k32_cell_accum<A32_P2GX, CAPC, ROWS32>;
    if (dirty) {
      hipLaunchKernelGGL(k, grid, blk, 0, c->stream, b, c->cap, c->g, zmap_of(c->g),
                         (float)c->g.vmin, c->cell_count, c->cell_start, c->sorted, c->ncell,
                         partial32(c), nullptr, nullptr, nullptr, dirty);
      return PMPIC_OK;
    }
    return launch_timed(c, T_ACCUM, k, grid, blk, 0, b, c->cap, c->g, zmap_of(c->g),
                        (float)c->g.vmin, c->cell_count, c->cell_start, c->sorted, c->ncell,
                        partial32(c), nullptr, nullptr, nullptr, nullptr);
  }
  auto k = b.capc == CAPC_LOW ? k32_cell_accum<A32_P2G, CAPC_LOW, ROWS_LOW>
                              : k32_cell_accum<A32_P2G, CAPC, ROWS32>;
  return launch_timed(c, T_ACCUM, k, grid, blk, 0, b, c->cap, c->g, zmap_of(c->g),
                      (float)c->g.vmin, c->cell_count, c->cell_start, c->sorted, c->ncell,
                      partial32(c), nullptr, nullptr, nullptr, nullptr);
}

// The capacity tuning: bin at 16 and at 32 slots, time the accumulation of each (its second
// launch, after one to warm the caches) and keep the faster for every later binning of this
// context.  Both capacities sum each cell's parcels in ascending parcel order, so the choice
// changes no result; the partials of the last run are this par2grid's.
static int tune_capacity(Ctx *c, int64_t n) {
  Bulk32 &b = c->b32;
  b.prebinned = false;
  const int caps[2] = {CAPC_LOW, CAPC};
  float ms[2] = {0.f, 0.f};
  hipEvent_t e0, e1;
  PMPIC_CHECK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    return PMPIC_ERR_HIP;
  }
  int rc = PMPIC_OK;
  for (int t = 0; t < 2 && rc == PMPIC_OK; t++) {
    b.capfix = caps[t];
    rc = bin32(c, n);
    if (rc == PMPIC_OK) rc = p2g_accum(c);
    if (rc == PMPIC_OK && hipEventRecord(e0, c->stream) != hipSuccess) rc = PMPIC_ERR_HIP;
    if (rc == PMPIC_OK) rc = p2g_accum(c);
    if (rc == PMPIC_OK && (hipEventRecord(e1, c->stream) != hipSuccess ||
                           hipEventSynchronize(e1) != hipSuccess ||
                           hipEventElapsedTime(&ms[t], e0, e1) != hipSuccess))
      rc = PMPIC_ERR_HIP;
  }
  b.capfix = 0;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc != PMPIC_OK) return rc;
  b.captuned = (ms[0] < ms[1]) ? caps[0] : caps[1];
  b.captune = false;
  return PMPIC_OK;
}

int b32_par2grid(Ctx *c) {
  const int64_t n = c->n;
  // this par2grid writes the FP32 p, q, r and total-buoyancy grids, current for the exports of
  // GP..GR and GB: the kept p, q, r, b of the last mixing are superseded
  Bulk32 &b = c->b32;
  b.kept_pending[1] = b.kept_pending[3] = false;
  b.kept_pending[4] = b.kept_pending[5] = false;
  for (int i = 0; i < 4; i++) b.p2g_pending[i] = true;
  // the last mixing's kept sums, where it left them current, stand for this accumulation
  // except on the tiles it changed
  const uint8_t *dirty = (b.p2g_reuse && p2g_xrows(c) && !b.captune) ? b.dirty : nullptr;
  b.p2g_reuse = false;
  // a mixing that changed no tile (known to the host once it has read the device scalars back,
  // and not inside a graph capture): the partials are all current, so no binning or sums
  const bool none = dirty && c->graph_n == 0 && b.dirty_seq < c->sync_seq && c->hsc.anydirty == 0;
  if (c->b32.captune && n > 0) {
    PMPIC_TRY(tune_capacity(c, n));
  } else if (!none) {
    PMPIC_TRY(bin32(c, n));
    PMPIC_TRY(p2g_accum(c, dirty));
  }
  Out32<float> out = {{b.pgrid[0], b.pgrid[1], b.pgrid[2], b.pgrid[3], nullptr, nullptr}};
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  if (p2g_xrows(c))
    hipLaunchKernelGGL((k32_node_sum_x<5, true, false>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0,
                       c->stream, c->g, partial32(c), out, nullptr);
  else
    hipLaunchKernelGGL((k32_node_sum<4, true, true, float>), dim3(cdiv(nnode, TPB)), dim3(TPB),
                       0, c->stream, c->g, partial32(c), c->ncell, out, nullptr);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int b32_rk_stage(Ctx *c, int stage, int stretch) {
  const int64_t n = c->n;
  if (n == 0) return PMPIC_OK;
  Grid32Set gs = {{c->b32.grid[0], c->b32.grid[1], c->b32.grid[2], c->b32.grid[3],
                   c->b32.grid[4], c->b32.grid[5]}};
  const Scale32 s3 = scale_of(c->g);
  Bulk32 &b = c->b32;
  b.p2g_reuse = false;
  const dim3 gb(cdiv(c->graph_n ? c->graph_n : n, b.rk_tpb)), tb(b.rk_tpb);
  // stages 1-3 bin the cells they write for the next stage's par2grid (stage 4 is followed by
  // splitting and mixing, which change the parcels); not while the capacity is being tuned
  BinOut bo{0, 0, c->cell_count, b.slot, b.ovf, c->rank};
  b.prebinned = false;
  if (stage < 4 && b.fused_bin && !b.captune) {
    bo.on = 1;
    bo.capc = b32_choose_capc(c, n);
    const int64_t nc1 = c->ncell + 1;
    hipLaunchKernelGGL(k32_bin_reset, dim3(cdiv(nc1, TPB)), dim3(TPB), 0, c->stream,
                       c->cell_count, nc1, c->dsc);
  }
  switch (stage) {
    case 1:
      PMPIC_TRY(launch_timed(c, T_RK, k32_rk_stage<1>, gb, tb, 0, c->g, c->b32, c->cap, n, gs,
                             s3, stretch, c->flag, c->dsc, bo));
      break;
    case 2:
      PMPIC_TRY(launch_timed(c, T_RK, k32_rk_stage<2>, gb, tb, 0, c->g, c->b32, c->cap, n, gs,
                             s3, stretch, c->flag, c->dsc, bo));
      break;
    case 3:
      PMPIC_TRY(launch_timed(c, T_RK, k32_rk_stage<3>, gb, tb, 0, c->g, c->b32, c->cap, n, gs,
                             s3, stretch, c->flag, c->dsc, bo));
      break;
    default:
      PMPIC_TRY(launch_timed(c, T_RK, k32_rk_stage<4>, gb, tb, 0, c->g, c->b32, c->cap, n, gs,
                             s3, stretch, c->flag, c->dsc, bo));
      break;
  }
  PMPIC_CHECK(hipGetLastError());
  if (bo.on) {
    b.prebinned = true;
    b.capc = bo.capc;
  }
  c->split_flags_fresh = (stage == 4 && stretch) ? 1 : 0;
  return PMPIC_OK;
}

int b32_split(Ctx *c) {
  const int64_t n = c->n;
  c->b32.prebinned = false;
  c->b32.p2g_reuse = false;
  if (n > 0) {
    // the stage-4 RK kernel already flagged the parcels to split
    if (!c->split_flags_fresh)
      hipLaunchKernelGGL(k32_split_flags, dim3(cdiv(n, TPB)), dim3(TPB), 0, c->stream,
                         c->b32.f, c->cap, n, c->flag);
    c->split_flags_fresh = 0;
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
    hipLaunchKernelGGL(k32_split_apply, dim3(cdiv(ns, TPB)), dim3(TPB), 0, c->stream, c->g,
                       c->b32, c->cap, n, scale_of(c->g), c->list_b, ns, c->dsc);
    PMPIC_CHECK(hipGetLastError());
  }
  c->n = n + ns;
  return b32_publish_n(c);
}

int b32_mix(Ctx *c) {
  const Geometry &g = c->g;
  Bulk32 &b = c->b32;
  b.prebinned = false;
  const int64_t norig = c->n;
  const float vmin = (float)g.vmin;
  // kept and residual grids (construct_grids, parcel_mixing.F90:440-624)
  PMPIC_TRY(bin32(c, norig));
  const bool xr = p2g_xrows(c);
  // x-row grids: the kept sums' first five planes (weight, p, q, r, total buoyancy) are what
  // the next par2grid would sum on every tile this mixing leaves unchanged; the tiles it changes
  // are marked below (removed, new, absorbing and moved parcels)
  b.p2g_reuse = false;
  const bool reuse = xr && !b.reuse_off;
  uint8_t *dirty = reuse ? b.dirty : nullptr;
  if (reuse) {
    const int64_t nt = c->ncell / TILE32 + 1;
    hipLaunchKernelGGL(k32_fill_u8, dim3(cdiv(nt, TPB)), dim3(TPB), 0, c->stream, b.dirty, nt,
                       (uint8_t)0);
  }
  if (xr && b.capc == CAPC_LOW)
    hipLaunchKernelGGL((k32_cell_accum<A32_MIXX, CAPC_LOW, ROWS_LOW>), dim3(cdiv(c->ncell, TILE32)),
                       dim3(ATPB), 0, c->stream, b, c->cap, g, zmap_of(g), vmin, c->cell_count,
                       c->cell_start, c->sorted, c->ncell, partial32(c), partial32b(c),
                       c->npercell, c->cellflag, nullptr);
  else if (xr)
    hipLaunchKernelGGL((k32_cell_accum<A32_MIXX, CAPC, ROWS32>), dim3(cdiv(c->ncell, TILE32)),
                       dim3(ATPB), 0, c->stream, b, c->cap, g, zmap_of(g), vmin, c->cell_count,
                       c->cell_start, c->sorted, c->ncell, partial32(c), partial32b(c),
                       c->npercell, c->cellflag, nullptr);
  else if (b.capc == CAPC_LOW)
    hipLaunchKernelGGL((k32_cell_accum<A32_MIX, CAPC_LOW, ROWS_LOW>), dim3(cdiv(c->ncell, TILE32)),
                       dim3(ATPB), 0, c->stream, b, c->cap, g, zmap_of(g), vmin, c->cell_count,
                       c->cell_start, c->sorted, c->ncell, partial32(c), partial32b(c),
                       c->npercell, c->cellflag, nullptr);
  else
    hipLaunchKernelGGL((k32_cell_accum<A32_MIX, CAPC, ROWS32>), dim3(cdiv(c->ncell, TILE32)),
                       dim3(ATPB), 0, c->stream, b, c->cap, g, zmap_of(g), vmin, c->cell_count,
                       c->cell_start, c->sorted, c->ncell, partial32(c), partial32b(c),
                       c->npercell, c->cellflag, nullptr);
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  Out32<float> kept = {{b.mgrid[0], b.mgrid[1], b.mgrid[2], b.mgrid[3], b.mgrid[4], b.mgrid[5]}};
  Out32<float> res = {{b.mgrid[6], b.mgrid[7], b.mgrid[8], b.mgrid[9], b.mgrid[10], b.mgrid[11]}};
  // cellflag holds the per-tile residual flags until the new-parcel flags reuse it below
  if (xr) {
    // planes weight, p, q, r, total buoyancy (not kept), b, h
    const Out32<float> keptx = {{b.mgrid[0], b.mgrid[3], b.mgrid[4], b.mgrid[5], nullptr,
                                 b.mgrid[1], b.mgrid[2]}};
    const Out32<float> resx = {{b.mgrid[6], b.mgrid[9], b.mgrid[10], b.mgrid[11], nullptr,
                                b.mgrid[7], b.mgrid[8]}};
    hipLaunchKernelGGL((k32_node_sum_x<7, false, true>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0,
                       c->stream, g, partial32(c), keptx, nullptr);
    hipLaunchKernelGGL((k32_node_sum_x<7, false, false>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0,
                       c->stream, g, partial32b(c), resx, c->cellflag);
  } else {
    hipLaunchKernelGGL((k32_node_sum<5, false, true, float>), dim3(cdiv(nnode, TPB)), dim3(TPB),
                       0, c->stream, g, partial32(c), c->ncell, kept, nullptr);
    hipLaunchKernelGGL((k32_node_sum<5, false, false, float>), dim3(cdiv(nnode, TPB)), dim3(TPB),
                       0, c->stream, g, partial32b(c), c->ncell, res, c->cellflag);
  }
  // the reference builds the kept grids in the model's vol, b, hg, p, q, r grids
  // (parcel_mixing.F90:131), which is what a grid export shows after mixing; nothing in the
  // step reads them there, so b32_export_grid widens them when one is exported
  for (int f = 0; f < 6; f++) b.kept_pending[f] = true;
  for (int f = 0; f < 4; f++) b.p2g_pending[f] = false;
  if (norig > 0)
    hipLaunchKernelGGL(k32_keep_flags, dim3(cdiv(norig, TPB * KPT)), dim3(TPB), 0, c->stream, b.f,
                       c->cap, norig, vmin, c->flag);
  PMPIC_TRY(compact_flags(c, c->flag, 0, norig, c->list_a, &c->dsc->nzero));
  hipLaunchKernelGGL(k32_newcell_flags, dim3(cdiv(c->ncell, TPB)), dim3(TPB), 0, c->stream, g,
                     c->npercell, c->cellflag, c->ncell);
  PMPIC_TRY(compact_flags(c, c->cellflag, 1, c->ncell, c->sorted, &c->dsc->nadd));
  PMPIC_TRY(sync_scalars(c));
  const int64_t Z = c->hsc.nzero, nadd = c->hsc.nadd;
  Mix32 mg = {{b.mgrid[0], b.mgrid[1], b.mgrid[2], b.mgrid[3], b.mgrid[4], b.mgrid[5]},
              {b.mgrid[6], b.mgrid[7], b.mgrid[8], b.mgrid[9], b.mgrid[10], b.mgrid[11]}};
  float *portion = partial32b(c);  // the residual partials are consumed by now
  if (reuse) {
    // the removed parcels' cells (read before new parcels take their slots) and the new cells
    if (Z > 0)
      hipLaunchKernelGGL(k32_mark_tiles, dim3(cdiv(Z, TPB)), dim3(TPB), 0, c->stream, g, b.cell,
                         c->list_a, Z, 1, b.dirty);
    if (nadd > 0)
      hipLaunchKernelGGL(k32_mark_tiles, dim3(cdiv(nadd, TPB)), dim3(TPB), 0, c->stream, g,
                         b.cell, c->sorted, nadd, 0, b.dirty);
  }
  if (nadd > 0)
    hipLaunchKernelGGL(k32_mix_new, dim3(cdiv(nadd, TPB)), dim3(TPB), 0, c->stream, g, b, c->cap,
                       norig, c->sorted, nadd, c->list_a, Z, c->flag, mg, vmin, portion, c->dsc);
  hipLaunchKernelGGL(k32_mix_residual, dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream, g, mg,
                     c->npercell, portion, b.mtouch, dirty);
  if (norig > 0)
    hipLaunchKernelGGL(k32_mix_absorb, dim3(cdiv(norig, TPB)), dim3(TPB), 0, c->stream, g, b,
                       c->cap, norig, c->flag, mg, b.mtouch);
  PMPIC_CHECK(hipGetLastError());
  // new count (parcel_mixing.F90:217), then the backfill of removed slots (:347-389)
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
      hipLaunchKernelGGL(k32_mix_backfill, dim3(cdiv(R, TPB)), dim3(TPB), 0, c->stream, g, b,
                         c->cap, dst, R, c->list_b, K1, dirty);
      PMPIC_CHECK(hipGetLastError());
      np -= R;
    }
  }
  c->n = np;
  b.p2g_reuse = reuse;
  if (reuse) {
    // whether any tile changed; the host learns it with the next read-back of the device
    // scalars (the status check that ends the component or the step)
    const int64_t nt = c->ncell / TILE32;
    hipLaunchKernelGGL(k32_zero_i64, dim3(1), dim3(64), 0, c->stream, &c->dsc->anydirty);
    hipLaunchKernelGGL(k32_any_u8, dim3(cdiv(cdiv(nt, 4), TPB)), dim3(TPB), 0, c->stream, b.dirty,
                       nt, &c->dsc->anydirty);
    b.dirty_seq = c->sync_seq;
  }
  return b32_publish_n(c);
}

// writebinary's grids (writebinary.F90:217-228): b, hg, p, q, r by par2grid of the FP32
// parcels' b, h, p, q, r, FP32 sums normalised into the FP64 grids
int b32_dump_grids(Ctx *c) {
  Bulk32 &b = c->b32;
  const Geometry &g = c->g;
  b.p2g_reuse = false;  // the dump's sums overwrite the partials
  PMPIC_TRY(bin32(c, c->n));
  if (b.capc == CAPC_LOW)
    hipLaunchKernelGGL((k32_cell_accum<A32_DUMP, CAPC_LOW, ROWS_LOW>), dim3(cdiv(c->ncell, TILE32)),
                       dim3(ATPB), 0, c->stream, b, c->cap, g, zmap_of(g), (float)g.vmin,
                       c->cell_count, c->cell_start, c->sorted, c->ncell, partial32(c), nullptr,
                       nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k32_cell_accum<A32_DUMP, CAPC, ROWS32>), dim3(cdiv(c->ncell, TILE32)),
                       dim3(ATPB), 0, c->stream, b, c->cap, g, zmap_of(g), (float)g.vmin,
                       c->cell_count, c->cell_start, c->sorted, c->ncell, partial32(c), nullptr,
                       nullptr, nullptr, nullptr);
  Out32<double> out = {{c->grid[GB], c->grid[GHG], c->grid[GP], c->grid[GQ], c->grid[GR], nullptr}};
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  hipLaunchKernelGGL((k32_node_sum<5, true, true, double>), dim3(cdiv(nnode, TPB)), dim3(TPB), 0,
                     c->stream, g, partial32(c), c->ncell, out, nullptr);
  PMPIC_CHECK(hipGetLastError());
  // these grids now hold writebinary's values, not the kept grids of the last mixing or the
  // step's FP32 grids
  for (int f = 1; f < 6; f++) b.kept_pending[f] = false;
  for (int f = 0; f < 4; f++) b.p2g_pending[f] = false;
  return PMPIC_OK;
}

// the FP64 grid `id` as a grid export sees it in the FP32-bulk form: a kept grid of the last
// mixing still in mgrid, or the step's FP32 p, q, r or total buoyancy, is widened into it first
int b32_export_grid(Ctx *c, int id) {
  const GridId kept_ids[6] = {GVOL, GB, GHG, GP, GQ, GR};
  const GridId p2g_ids[4] = {GP, GQ, GR, GB};
  Bulk32 &b = c->b32;
  const int64_t nnode = (int64_t)c->g.nzn * c->g.ny * c->g.nx;
  for (int f = 0; f < 6; f++)
    if (kept_ids[f] == id && b.kept_pending[f]) {
      hipLaunchKernelGGL(k32_grid_to_f64, dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream,
                         b.mgrid[f], c->grid[id], nnode);
      PMPIC_CHECK(hipGetLastError());
      b.kept_pending[f] = false;
    }
  for (int f = 0; f < 4; f++)
    if (p2g_ids[f] == id && b.p2g_pending[f]) {
      hipLaunchKernelGGL(k32_grid_to_f64, dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream,
                         b.pgrid[f], c->grid[id], nnode);
      PMPIC_CHECK(hipGetLastError());
      b.p2g_pending[f] = false;
    }
  return PMPIC_OK;
}

int b32_conservation_sums(Ctx *c, double out6[6]) {
  const int64_t n = c->n;
  std::vector<float> v(n), f(n);
  if (n > 0)
    PMPIC_CHECK(hipMemcpy(v.data(), c->b32.f + (int64_t)B_VOL * c->cap, sizeof(float) * n,
                          hipMemcpyDeviceToHost));
  const int fl[6] = {-1, B_B, B_H, B_P, B_Q, B_R};
  for (int k = 0; k < 6; k++) {
    if (fl[k] >= 0 && n > 0)
      PMPIC_CHECK(hipMemcpy(f.data(), c->b32.f + (int64_t)fl[k] * c->cap, sizeof(float) * n,
                            hipMemcpyDeviceToHost));
    double s = 0.0, comp = 0.0;  // Neumaier summation in parcel order, FP64
    for (int64_t p = 0; p < n; p++) {
      const double term = fl[k] < 0 ? (double)v[p] : (double)v[p] * (double)f[p];
      const double t = s + term;
      if (fabs(s) >= fabs(term)) comp += (s - t) + term;
      else comp += (term - t) + s;
      s = t;
    }
    out6[k] = s + comp;
  }
  return PMPIC_OK;
}

}  // namespace pmpic
