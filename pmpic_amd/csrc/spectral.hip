// Semi-spectral operators and the vort2vel / vorticity_tendency compositions on MI355X.
//
// Reference: model_core/src/fft/fft_ops.F90 (fftops_init :48-168, spectral_filter :176-197,
// diffx :204-302, diffy :305-401, diffz :419-483, laplinv :505-601, tridiagonal :620-656),
// components/vort2vel/src/vort2vel.F90:98-400 and
// components/vorticity_tendency/src/vorticity_tendency.F90:107-334.
//
// Every operator is memory bound: element-wise kernels over the [nzn][sy][sx] spectral arrays
// (fused where the reference's sequence allows it without changing any operation) and column
// kernels with one thread per (kx,ky) column walking z; with z slowest in memory the column
// walk is coalesced across the threads of a wavefront.  The tridiagonal solves keep the
// reference's Thomas recurrence operation for operation (FP32-rounded literal coefficients,
// skipped k^2 = 0 columns).
#include "common.h"

#include <math.h>
#include <type_traits>
#include <vector>

namespace pmpic {

namespace {

constexpr int TPB = 256;

struct SpecDims {
  int sx, sy, nzn;
  int64_t ssz;
};

__device__ __forceinline__ SpecDims sdims(const Geometry &g) {
  SpecDims d;
  d.sx = g.sx; d.sy = g.sy; d.nzn = g.nzn;
  d.ssz = (int64_t)g.nzn * g.sy * g.sx;
  return d;
}

// The spectral arrays are of type S: FP64, or FP32 in the FP32-bulk form, whose spectral
// kernels then read and write FP32 values and compute in FP64 between.
//
// diffx of element q (fft_ops.F90:266-269, :295): pair swap along x, times kx
template <typename S>
__device__ __forceinline__ double dx_at(const S *in, const double *kx, int64_t q, int ii,
                                        int64_t col) {
  return ((ii & 1) == 0 ? (-1. * (double)in[q + 1]) : (double)in[q - 1]) * kx[col];
}
// diffy of element q (fft_ops.F90:366-369, :394)
template <typename S>
__device__ __forceinline__ double dy_at(const S *in, const double *ky, int64_t q, int jj,
                                        int sx, int64_t col) {
  return ((jj & 1) == 0 ? (-1. * (double)in[q + sx]) : (double)in[q - sx]) * ky[col];
}

__global__ void k_diffx(Geometry g, const double *__restrict__ in, double *__restrict__ out,
                        const double *__restrict__ kx) {
  const int64_t ssz = (int64_t)g.nzn * g.sy * g.sx;
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ssz) return;
  const int64_t col = q % ((int64_t)g.sy * g.sx);
  const int ii = (int)(q % g.sx);
  out[q] = dx_at(in, kx, q, ii, col);
}

__global__ void k_diffy(Geometry g, const double *__restrict__ in, double *__restrict__ out,
                        const double *__restrict__ ky) {
  const int64_t ssz = (int64_t)g.nzn * g.sy * g.sx;
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ssz) return;
  const int64_t col = q % ((int64_t)g.sy * g.sx);
  const int jj = (int)(col / g.sx);
  out[q] = dy_at(in, ky, q, jj, g.sx, col);
}

// The element-wise spectral kernels take one thread per element of a z level (blockIdx.y): the
// column col = (jj, ii) of the level's sy x sx plane and the element q = k plane + col, with no
// 64-bit division.  SPEC_ELEMENT declares plane, k, col, q, ii, jj and leaves threads past the
// plane.
#define SPEC_ELEMENT(g)                                                                      \
  const int64_t plane = (int64_t)(g).sy * (g).sx;                                            \
  const int col32 = (int)(blockIdx.x * blockDim.x + threadIdx.x);                            \
  if (col32 >= plane) return;                                                                \
  const int k = (int)blockIdx.y;                                                             \
  const int64_t col = col32;                                                                 \
  const int64_t q = (int64_t)k * plane + col;                                                \
  const int ii = col32 % (g).sx, jj = col32 / (g).sx;                                        \
  (void)k; (void)ii; (void)jj
inline dim3 spec_grid(const Geometry &g) {
  return dim3((unsigned)cdiv((int64_t)g.sy * g.sx, TPB), (unsigned)g.nzn);
}

template <typename S>
__global__ void k_filter(Geometry g, const S *__restrict__ in, S *__restrict__ out,
                         const double *__restrict__ filt) {
  SPEC_ELEMENT(g);
  out[q] = (S)((double)in[q] * filt[col]);
}

// vort2vel.F90:158-174: f = (diffx(a) + diffy(b)) + dc/dz (explicit 2nd order, one-sided at
// the ends), then the constant column zeroed
template <typename S>
__global__ void k_divergence(Geometry g, const S *__restrict__ a, const S *__restrict__ b,
                             const S *__restrict__ cs, S *__restrict__ f,
                             const double *__restrict__ kx, const double *__restrict__ ky,
                             double hdzi) {
  auto c = [&](int64_t i) { return (double)cs[i]; };
  SPEC_ELEMENT(g);
  const int nz = g.nzn;
  const double d = dx_at(a, kx, q, ii, col);
  const double e = dy_at(b, ky, q, jj, g.sx, col);
  double fz;
  if (k == 0)
    fz = hdzi * (4. * c(q + plane) - c(q + 2 * plane) - 3. * c(q));
  else if (k == nz - 1)
    fz = hdzi * (c(q - 2 * plane) + 3. * c(q) - 4. * c(q - plane));
  else
    fz = (c(q + plane) - c(q - plane)) * hdzi;
  double r = d + e + fz;
  if (col == 0) r = 0.0;
  f[q] = (S)r;
}

// vort2vel.F90:184-194: a -= diffx(f); b -= diffy(f)
template <typename S>
__global__ void k_correct_ab(Geometry g, S *__restrict__ a, S *__restrict__ b,
                             const S *__restrict__ f, const double *__restrict__ kx,
                             const double *__restrict__ ky) {
  SPEC_ELEMENT(g);
  a[q] = (S)((double)a[q] - dx_at(f, kx, q, ii, col));
  b[q] = (S)((double)b[q] - dy_at(f, ky, q, jj, g.sx, col));
}

// vort2vel.F90:197-218: c -= dz(f); c(level 1) = 0; d,e,f = filtered a,b,c; cache the top and
// bottom levels of the unfiltered a, b
template <typename S>
__global__ void k_correct_c_filter(Geometry g, const S *__restrict__ a,
                                   const S *__restrict__ b, S *__restrict__ c,
                                   const S *__restrict__ dzf, S *__restrict__ d,
                                   S *__restrict__ e, S *__restrict__ f,
                                   const double *__restrict__ filt, double *__restrict__ atop,
                                   double *__restrict__ abot, double *__restrict__ btop,
                                   double *__restrict__ bbot) {
  SPEC_ELEMENT(g);
  double cv = (double)c[q] - (double)dzf[q];
  if (k == 0) cv = 0.;
  c[q] = (S)cv;
  const double fl = filt[col];
  const double av = a[q], bv = b[q];
  d[q] = (S)(av * fl);
  e[q] = (S)(bv * fl);
  f[q] = (S)((double)(S)cv * fl);
  if (k == g.nzn - 1) { atop[col] = av; btop[col] = bv; }
  if (k == 0) { abot[col] = av; bbot[col] = bv; }
}

// mean flow, vort2vel.F90:230-241 (sequential over z; 1./nz is a default-real quotient)
template <typename S>
__global__ void k_mean_flow(Geometry g, const S *__restrict__ d, const S *__restrict__ e,
                            double *__restrict__ ubar, double *__restrict__ vbar) {
  // the (1,1) columns are staged in LDS by the whole block; one thread then runs the
  // reference's sequential recurrences and sums on them, and the block writes the result
  extern __shared__ double sm[];
  const int64_t plane = (int64_t)g.sy * g.sx;
  const int nz = g.nzn;
  double *se = sm, *sd = sm + nz, *ub = sm + 2 * nz, *vb = sm + 3 * nz, *avg = sm + 4 * nz;
  for (int k = threadIdx.x; k < nz; k += blockDim.x) {
    se[k] = e[(int64_t)k * plane];
    sd[k] = d[(int64_t)k * plane];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double dz = g.dz;
    ub[0] = 0;
    vb[0] = 0;
    for (int k = 1; k <= nz - 1; k++) {
      ub[k] = ub[k - 1] + dz / 2. * (se[k - 1] + se[k]);
      vb[k] = vb[k - 1] - dz / 2. * (sd[k - 1] + sd[k]);
    }
    const double rnz = (double)(1.f / (float)nz);
    double su = 0, sv = 0;
    for (int k = 1; k <= nz - 1; k++) {
      su += ub[k - 1] + 0.5 * ub[nz - 1];
      sv += vb[k - 1] + 0.5 * vb[nz - 1];
    }
    avg[0] = rnz * su;
    avg[1] = rnz * sv;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nz; k += blockDim.x) {
    ubar[k] = ub[k] - avg[0];
    vbar[k] = vb[k] - avg[1];
  }
}

// Column stage-in with many loads in flight: a chunk of LVL levels is read into registers
// (independent loads, issued back to back) before any is written to the LDS slice.
constexpr int LVL = 16;
template <typename S>
__device__ __forceinline__ void column_to_lds(const S *__restrict__ f, int64_t plane, int nz,
                                              double *F, int nc) {
  for (int k0 = 0; k0 < nz; k0 += LVL) {
    double t[LVL];
#pragma unroll
    for (int u = 0; u < LVL; u++)
      if (k0 + u < nz) t[u] = f[(int64_t)(k0 + u) * plane];
#pragma unroll
    for (int u = 0; u < LVL; u++)
      if (k0 + u < nz) F[(k0 + u) * nc] = t[u];
  }
}

// Column solvers.  One thread owns one (kx,ky) column and runs the reference's Thomas
// recurrence on it.  The column (and the elimination coefficients cp) sit either in HBM
// (stride = one spectral plane) or, when they fit, in an LDS slice private to the thread
// (stride = threads per block): the whole column is loaded with independent coalesced reads,
// solved in LDS, and stored back, so the level-by-level recurrence never waits on HBM.  Both
// forms run the same arithmetic.
template <typename T>
struct Col {
  T *p;
  int64_t s;
  __device__ __forceinline__ T &operator[](int k) const { return p[(int64_t)k * s]; }
};

// laplinv (fft_ops.F90:505-601) of one column, in place.  The right-hand side of row k reads
// the original f(k-1), f(k), f(k+1); a sliding window keeps the originals while the forward
// sweep overwrites f with d'.  Returns PMPIC_ERR_ZERO_PIVOT on a zero pivot.
template <typename CF, typename CC>
__device__ __forceinline__ int laplinv_column(const Geometry &g, CF F, CC CP, double k2, int df0) {
  const int nz = g.nzn;
  const double f16 = (double)(1.f / 6.f), f13 = (double)(1.f / 3.f);
  const double f112 = (double)(1.f / 12.f), f56 = (double)(5.f / 6.f);
  const double dz = g.dz;
  const double i2dz2 = 2. / dz / dz, idz2 = 1. / dz / dz;
  // interior coefficients (constant along the column)
  const double ai = idz2 - f112 * k2, bi = -i2dz2 - f56 * k2, ci = idz2 - f112 * k2;
  double fm = 0.0, f0 = F[0], fp = F[1];  // originals f(k-1), f(k), f(k+1)
  // row 1
  double b1, c1, d1;
  if (!df0) { b1 = 1.; c1 = 0.; d1 = 0.; }
  else { b1 = -idz2 - f13 * k2; c1 = idz2 - f16 * k2; d1 = f13 * f0 + f16 * fp; }
  double denominator = b1;
  double fprev = d1 / denominator;
  F[0] = fprev;
  double cprev = c1;  // c(i-1)
#pragma unroll 4
  for (int k = 2; k <= nz; k++) {
    fm = f0;
    f0 = fp;
    fp = (k < nz) ? F[k] : 0.0;
    double a, b, c, d;
    if (k < nz) {
      a = ai; b = bi; c = ci;
      d = f112 * fm + f56 * f0 + f112 * fp;
    } else if (!df0) {
      a = 0.; b = 1.; c = 0.; d = 0.;
    } else {
      a = idz2 - f16 * k2; b = -idz2 - f13 * k2; c = 0.;
      d = f13 * f0 + f16 * fm;
    }
    const double cpk = cprev / denominator;
    CP[k - 1] = cpk;
    denominator = b - a * cpk;
    if (denominator == 0) return PMPIC_ERR_ZERO_PIVOT;
    fprev = (d - a * fprev) / (denominator);
    F[k - 1] = fprev;
    cprev = c;
  }
  double fnext = fprev;
#pragma unroll 4
  for (int k = nz - 1; k >= 1; k--) {
    const double v = F[k - 1] - CP[k] * fnext;
    F[k - 1] = v;
    fnext = v;
  }
  return PMPIC_OK;
}

// diffz (fft_ops.F90:419-483) of one column; O may alias F (the original f(k-2), f(k-1) are
// kept in registers, and f(k) is read before d'(k-1) is written).
template <typename CI, typename CO, typename CC>
__device__ __forceinline__ int diffz_column(const Geometry &g, CI F, CO O, CC CP, bool bt,
                                            double bot, double top) {
  const int nz = g.nzn;
  const double a1 = (double)(1.f / 6.f), b1 = (double)(2.f / 3.f), c1 = (double)(1.f / 6.f);
  const double al = (double)(1.f / 3.f), cl = (double)(1.f / 3.f);
  const double dz = g.dz;
  double fkm2 = F[0], fkm1 = F[1];  // f(k-2), f(k-1) of the next row
  // row 1
  double d = bt ? fkm1 / dz - a1 * bot * dz : 0.0;
  double denominator = 1.;  // b(1) = b0
  double fprev = d / denominator;
  O[0] = fprev;
  double cprev = bt ? cl : 0.;  // c(1)
#pragma unroll 4
  for (int k = 2; k <= nz; k++) {
    double a, b, c, fk = 0.0;
    if (k < nz) {
      a = a1; b = b1; c = c1;
      fk = F[k];
      d = (fk - fkm2) / 2. / dz;
    } else {
      a = bt ? al : 0.; b = 1.; c = 0.;
      d = bt ? -fkm2 / dz + a1 * top * dz : 0.0;
    }
    const double cpk = cprev / denominator;
    CP[k - 1] = cpk;
    denominator = b - a * cpk;
    if (denominator == 0) return PMPIC_ERR_ZERO_PIVOT;
    fprev = (d - a * fprev) / (denominator);
    O[k - 1] = fprev;
    cprev = c;
    fkm2 = fkm1;
    fkm1 = fk;
  }
  double fnext = fprev;
#pragma unroll 4
  for (int k = nz - 1; k >= 1; k--) {
    const double v = O[k - 1] - CP[k] * fnext;
    O[k - 1] = v;
    fnext = v;
  }
  return PMPIC_OK;
}

// up to 4 fields per launch (blockIdx.y); the boundary planes are FP64 in both forms
template <typename S>
struct ColFields {
  const S *in[4];
  S *out[4];
  const double *bot[4];
  const double *top[4];
  int df0[4];
};

template <typename S>
__global__ void k_laplinv(Geometry g, ColFields<S> cf, double *__restrict__ cp,
                          const double *__restrict__ k2a, DevScalars *sc) {
  const int64_t plane = (int64_t)g.sy * g.sx;
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane) return;
  const double k2 = k2a[col];
  if (k2 == 0) return;
  Col<S> F{cf.out[blockIdx.y] + col, plane};
  Col<double> CP{cp + (int64_t)blockIdx.y * g.nzn * plane + col, plane};
  const int st = laplinv_column(g, F, CP, k2, cf.df0[blockIdx.y]);
  if (st != PMPIC_OK) raise_status(sc, st);
}

template <typename S>
__global__ void __launch_bounds__(64) k_laplinv_lds(Geometry g, ColFields<S> cf,
                                                    const double *__restrict__ k2a,
                                                    DevScalars *sc) {
  extern __shared__ double sh[];
  const int nc = blockDim.x, t = threadIdx.x, nz = g.nzn;
  const int64_t plane = (int64_t)g.sy * g.sx;
  const int64_t col = (int64_t)blockIdx.x * nc + t;
  if (col >= plane) return;  // no barriers below: each thread touches only its own slice
  const double k2 = k2a[col];
  if (k2 == 0) return;
  S *f = cf.out[blockIdx.y] + col;
  Col<double> F{sh + t, nc}, CP{sh + (int64_t)nz * nc + t, nc};
  column_to_lds(f, plane, nz, sh + t, nc);
  const int st = laplinv_column(g, F, CP, k2, cf.df0[blockIdx.y]);
  if (st != PMPIC_OK) {
    raise_status(sc, st);
    return;
  }
#pragma unroll 8
  for (int k = 0; k < nz; k++) f[(int64_t)k * plane] = (S)F[k];
}

template <typename S>
__global__ void k_diffz(Geometry g, ColFields<S> cf, double *__restrict__ cp, DevScalars *sc) {
  const int64_t plane = (int64_t)g.sy * g.sx;
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= plane) return;
  const int fi = blockIdx.y;
  const bool bt = cf.bot[fi] != nullptr && cf.top[fi] != nullptr;
  Col<const S> F{cf.in[fi] + col, plane};
  Col<S> O{cf.out[fi] + col, plane};
  Col<double> CP{cp + (int64_t)fi * g.nzn * plane + col, plane};
  const int st = diffz_column(g, F, O, CP, bt, bt ? cf.bot[fi][col] : 0.0,
                              bt ? cf.top[fi][col] : 0.0);
  if (st != PMPIC_OK) raise_status(sc, st);
}

template <typename S>
__global__ void __launch_bounds__(64) k_diffz_lds(Geometry g, ColFields<S> cf, DevScalars *sc) {
  extern __shared__ double sh[];
  const int nc = blockDim.x, t = threadIdx.x, nz = g.nzn;
  const int64_t plane = (int64_t)g.sy * g.sx;
  const int64_t col = (int64_t)blockIdx.x * nc + t;
  if (col >= plane) return;  // no barriers below: each thread touches only its own slice
  const int fi = blockIdx.y;
  const bool bt = cf.bot[fi] != nullptr && cf.top[fi] != nullptr;
  const S *f = cf.in[fi] + col;
  S *o = cf.out[fi] + col;
  Col<double> F{sh + t, nc}, CP{sh + (int64_t)nz * nc + t, nc};
  column_to_lds(f, plane, nz, sh + t, nc);
  const int st = diffz_column(g, F, F, CP, bt, bt ? cf.bot[fi][col] : 0.0,
                              bt ? cf.top[fi][col] : 0.0);
  if (st != PMPIC_OK) {
    raise_status(sc, st);
    return;
  }
#pragma unroll 8
  for (int k = 0; k < nz; k++) o[(int64_t)k * plane] = (S)F[k];
}

// Register-resident column solvers: the column stays in registers (NZ levels, unrolled) and
// the level-by-level recurrence never waits on memory.  diffz's Thomas denominators and c' are
// the same for every column, so they are built once (k_dz_tables, with the same operations as
// diffz_column) and each call runs only the data recurrence.  laplinv's depend on the column's
// k^2: they are computed alongside the data recurrence (an independent chain).  Results are
// bit-identical to the LDS and HBM forms.
__global__ void k_dz_tables(Geometry g, double *__restrict__ den, double *__restrict__ cp) {
  if (threadIdx.x || blockIdx.x) return;
  const int nz = g.nzn;
  const double a1 = (double)(1.f / 6.f), b1 = (double)(2.f / 3.f), c1 = (double)(1.f / 6.f);
  const double al = (double)(1.f / 3.f), cl = (double)(1.f / 3.f);
  for (int bt = 0; bt < 2; bt++) {
    double *D = den + bt * nz, *C = cp + bt * nz;
    double denominator = 1.;
    D[0] = denominator;
    C[0] = 0.;
    double cprev = bt ? cl : 0.;
    for (int k = 2; k <= nz; k++) {
      double a, b, c;
      if (k < nz) { a = a1; b = b1; c = c1; }
      else { a = bt ? al : 0.; b = 1.; c = 0.; }
      const double cpk = cprev / denominator;
      C[k - 1] = cpk;
      denominator = b - a * cpk;
      D[k - 1] = denominator;
      cprev = c;
    }
  }
}

// laplinv with Q lanes per column (1, 2 or 4): lane q * CPW + c holds levels [q L, q L + L) of
// column c of the wavefront's CPW = 64 / Q, both the column values and the forward sweep's c',
// in registers.  The sweeps run lane group after lane group (the others masked), the recurrence
// state crossing to the next group by a lane shuffle, so no level waits on memory and every
// operation is the one of laplinv_column in the same order: results are bit-identical.  A zero
// pivot is flagged and raised after the sweeps (the reference stops there; the column's values
// are then moot).
template <int Q, int NZ, typename S>
__global__ void __launch_bounds__(64) k_laplinv_reg(Geometry g, ColFields<S> cf,
                                                    const double *__restrict__ k2a,
                                                    DevScalars *sc) {
  static_assert(Q == 1 || Q == 2 || Q == 4, "one, two or four lanes per column");
  constexpr int CPW = 64 / Q;  // columns per wavefront
  constexpr int L = (NZ + Q - 1) / Q;
  const int lane = threadIdx.x;
  const int q = lane / CPW;
  const int64_t plane = (int64_t)g.sy * g.sx;
  const int64_t col = (int64_t)blockIdx.x * CPW + (lane % CPW);
  if (col >= plane) return;
  const double k2 = k2a[col];
  if (k2 == 0) return;
  const int df0 = cf.df0[blockIdx.y];
  S *f = cf.out[blockIdx.y] + col;
  // the FP32-bulk form (S float) forms one reciprocal per pivot and multiplies by it, and
  // holds the column and the c' values in FP32; the FP64 form divides as the reference does
  constexpr bool F32 = std::is_same<S, float>::value;
  using FT = typename std::conditional<F32, float, double>::type;
  FT F[L], CPl[L];
#pragma unroll
  for (int i = 0; i < L; i++) {
    F[i] = 0.0;
    CPl[i] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < L; i++)
    if (q * L + i < NZ) F[i] = f[(int64_t)(q * L + i) * plane];
  const double f16 = (double)(1.f / 6.f), f13 = (double)(1.f / 3.f);
  const double f112 = (double)(1.f / 12.f), f56 = (double)(5.f / 6.f);
  const double dz = g.dz;
  const double i2dz2 = 2. / dz / dz, idz2 = 1. / dz / dz;
  const double ai = idz2 - f112 * k2, bi = -i2dz2 - f56 * k2, ci = idz2 - f112 * k2;
  // the original of level q L + L, the next group's first (unused with one lane per column)
  const double nxt0 = __shfl((double)F[0], (lane + CPW) & 63);
  constexpr int PTOP = (NZ - 1) / L;  // group holding level NZ - 1
  double denominator = 0.0, rden = 0.0, fprev = 0.0, cprev = 0.0, f0 = 0.0, fp = 0.0;
  bool bad = false;
#pragma unroll
  for (int p = 0; p <= PTOP; p++) {
    if (p > 0) {
      denominator = __shfl(denominator, lane - CPW);
      if (F32) rden = __shfl(rden, lane - CPW);
      fprev = __shfl(fprev, lane - CPW);
      cprev = __shfl(cprev, lane - CPW);
      f0 = __shfl(f0, lane - CPW);
      fp = __shfl(fp, lane - CPW);
    }
    if (q == p) {
#pragma unroll
      for (int i = 0; i < L; i++) {
        const int j = p * L + i;  // level (0-based); row j + 1 of the reference
        if (j >= NZ) break;
        const double onext = (i + 1 < L) ? F[i + 1 < L ? i + 1 : 0] : nxt0;  // original j + 1
        if (j == 0) {
          f0 = F[0];
          fp = onext;
          double b1, c1, d1;
          if (!df0) { b1 = 1.; c1 = 0.; d1 = 0.; }
          else { b1 = -idz2 - f13 * k2; c1 = idz2 - f16 * k2; d1 = f13 * f0 + f16 * fp; }
          denominator = b1;
          if constexpr (F32) {
            rden = 1.0 / denominator;
            fprev = d1 * rden;
          } else {
            fprev = d1 / denominator;
          }
          F[0] = (FT)fprev;
          cprev = c1;
          continue;
        }
        const double fm = f0;
        f0 = fp;
        fp = (j + 1 < NZ) ? onext : 0.0;
        double a, b, c, d;
        if (j + 1 < NZ) {
          a = ai; b = bi; c = ci;
          d = f112 * fm + f56 * f0 + f112 * fp;
        } else if (!df0) {
          a = 0.; b = 1.; c = 0.; d = 0.;
        } else {
          a = idz2 - f16 * k2; b = -idz2 - f13 * k2; c = 0.;
          d = f13 * f0 + f16 * fm;
        }
        const double cpk = F32 ? cprev * rden : cprev / denominator;
        CPl[i] = (FT)cpk;
        denominator = b - a * cpk;
        bad |= (denominator == 0);
        if constexpr (F32) {
          rden = 1.0 / denominator;
          fprev = (d - a * fprev) * rden;
        } else {
          fprev = (d - a * fprev) / (denominator);
        }
        F[i] = (FT)fprev;
        cprev = c;
      }
    }
  }
  // backward sweep from the top group down: level j takes c' of level j + 1
  const double cpn = __shfl((double)CPl[0], (lane + CPW) & 63);
  double fnext = fprev;
#pragma unroll
  for (int p = PTOP; p >= 0; p--) {
    if (p < PTOP) fnext = __shfl(fnext, lane + CPW);
    if (q == p) {
#pragma unroll
      for (int i = L - 1; i >= 0; i--) {
        const int j = p * L + i;
        if (j >= NZ - 1) continue;
        const double cpj1 = (i + 1 < L) ? (double)CPl[i + 1 < L ? i + 1 : 0] : cpn;
        const double v = (double)F[i] - cpj1 * fnext;
        F[i] = (FT)v;
        fnext = v;
      }
    }
  }
  if (bad) raise_status(sc, PMPIC_ERR_ZERO_PIVOT);
#pragma unroll
  for (int i = 0; i < L; i++)
    if (q * L + i < NZ) f[(int64_t)(q * L + i) * plane] = (S)F[i];
}

template <int NZ, typename S>
__global__ void __launch_bounds__(64) k_diffz_reg(Geometry g, ColFields<S> cf,
                                                  const double *__restrict__ tden,
                                                  const double *__restrict__ tcp,
                                                  DevScalars *sc) {
  // the Thomas tables are the same for every column: staged once per block in LDS, so the
  // recurrence reads them without a memory round trip per level.  The FP32-bulk form (S float)
  // stages the pivots' reciprocals and multiplies by them, takes 1/dz and 1/(2 dz) as factors
  // and holds the column in FP32 between the sweeps (the sweeps themselves run in FP64); the
  // FP64 form divides as the reference does
  constexpr bool F32 = std::is_same<S, float>::value;
  using FT = typename std::conditional<F32, float, double>::type;
  __shared__ double den[NZ], cpt[NZ];
  const int64_t plane = (int64_t)g.sy * g.sx;
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int fi = blockIdx.y;
  const bool bt = cf.bot[fi] != nullptr && cf.top[fi] != nullptr;
  for (int k = threadIdx.x; k < NZ; k += blockDim.x) {
    const double dn = tden[(bt ? NZ : 0) + k];
    if (F32 && dn == 0 && blockIdx.x == 0) raise_status(sc, PMPIC_ERR_ZERO_PIVOT);
    den[k] = F32 ? 1.0 / dn : dn;
    cpt[k] = tcp[(bt ? NZ : 0) + k];
  }
  __syncthreads();
  if (col >= plane) return;
  const S *fin = cf.in[fi] + col;
  S *o = cf.out[fi] + col;
  FT F[NZ];
#pragma unroll
  for (int k = 0; k < NZ; k++) F[k] = fin[(int64_t)k * plane];
  const double bot = bt ? cf.bot[fi][col] : 0.0, top = bt ? cf.top[fi][col] : 0.0;
  const double a1 = (double)(1.f / 6.f), al = (double)(1.f / 3.f);
  const double dz = g.dz, idz = 1.0 / dz, ih2dz = 0.5 / dz;
  double fkm2 = F[0], fkm1 = F[1];  // originals f(k-2), f(k-1) of the next row
  const double d0 = bt ? (F32 ? fkm1 * idz : fkm1 / dz) - a1 * bot * dz : 0.0;
  double fprev = F32 ? d0 * den[0] : d0 / den[0];
  F[0] = (FT)fprev;
#pragma unroll
  for (int k = 2; k <= NZ; k++) {
    double a, d, fk = 0.0;
    if (k < NZ) {
      a = a1;
      fk = F[k < NZ ? k : 0];
      d = F32 ? (fk - fkm2) * ih2dz : (fk - fkm2) / 2. / dz;
    } else {
      a = bt ? al : 0.;
      d = bt ? (F32 ? -fkm2 * idz : -fkm2 / dz) + a1 * top * dz : 0.0;
    }
    const double dn = den[k - 1];
    if constexpr (F32) {
      fprev = (d - a * fprev) * dn;
    } else {
      if (dn == 0) { raise_status(sc, PMPIC_ERR_ZERO_PIVOT); return; }
      fprev = (d - a * fprev) / (dn);
    }
    F[k - 1] = (FT)fprev;
    fkm2 = fkm1;
    fkm1 = fk;
  }
  double fnext = fprev;
#pragma unroll
  for (int k = NZ - 1; k >= 1; k--) {
    const double v = (double)F[k - 1] - cpt[k] * fnext;
    F[k - 1] = (FT)v;
    fnext = v;
  }
#pragma unroll
  for (int k = 0; k < NZ; k++) o[(int64_t)k * plane] = (S)F[k];
}

// Long columns (more z nodes than the instances above): Q lanes per column, lane q holding
// levels [q L, q L + L) of the column in registers, for any column length nz <= Q L chosen at
// run time.  The sweeps run lane group after lane group as in k_laplinv_reg, the recurrence
// state crossing by lane shuffles, every operation the one of laplinv_column / k_diffz_reg in
// the same order, so results are bit-identical to them.
template <int Q, int L, typename S>
__global__ void __launch_bounds__(64) k_laplinv_qn(Geometry g, ColFields<S> cf,
                                                   const double *__restrict__ k2a,
                                                   DevScalars *sc) {
  constexpr int CPW = 64 / Q;  // columns per wavefront
  const int NZ = g.nzn;
  const int lane = threadIdx.x;
  const int q = lane / CPW;
  const int64_t plane = (int64_t)g.sy * g.sx;
  const int64_t col = (int64_t)blockIdx.x * CPW + (lane % CPW);
  if (col >= plane) return;  // all Q lanes of a column leave together
  const double k2 = k2a[col];
  if (k2 == 0) return;
  const int df0 = cf.df0[blockIdx.y];
  S *f = cf.out[blockIdx.y] + col;
  // the FP32-bulk form as in k_laplinv_reg: one reciprocal per pivot, FP32 column and c'
  constexpr bool F32 = std::is_same<S, float>::value;
  using FT = typename std::conditional<F32, float, double>::type;
  FT F[L], CPl[L];
#pragma unroll
  for (int i = 0; i < L; i++) {
    F[i] = 0.0;
    CPl[i] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < L; i++)
    if (q * L + i < NZ) F[i] = f[(int64_t)(q * L + i) * plane];
  const double f16 = (double)(1.f / 6.f), f13 = (double)(1.f / 3.f);
  const double f112 = (double)(1.f / 12.f), f56 = (double)(5.f / 6.f);
  const double dz = g.dz;
  const double i2dz2 = 2. / dz / dz, idz2 = 1. / dz / dz;
  const double ai = idz2 - f112 * k2, bi = -i2dz2 - f56 * k2, ci = idz2 - f112 * k2;
  const double nxt0 = __shfl((double)F[0], lane + CPW);  // the next group's first original
  const int PTOP = (NZ - 1) / L;                         // group holding level NZ - 1
  double denominator = 0.0, rden = 0.0, fprev = 0.0, cprev = 0.0, f0 = 0.0, fp = 0.0;
  bool bad = false;
  for (int p = 0; p <= PTOP; p++) {
    if (p > 0) {
      denominator = __shfl(denominator, lane - CPW);
      if (F32) rden = __shfl(rden, lane - CPW);
      fprev = __shfl(fprev, lane - CPW);
      cprev = __shfl(cprev, lane - CPW);
      f0 = __shfl(f0, lane - CPW);
      fp = __shfl(fp, lane - CPW);
    }
    if (q == p) {
#pragma unroll
      for (int i = 0; i < L; i++) {
        const int j = p * L + i;
        if (j >= NZ) break;
        const double onext = (i + 1 < L) ? F[i + 1 < L ? i + 1 : 0] : nxt0;
        if (j == 0) {
          f0 = F[0];
          fp = onext;
          double b1, c1, d1;
          if (!df0) { b1 = 1.; c1 = 0.; d1 = 0.; }
          else { b1 = -idz2 - f13 * k2; c1 = idz2 - f16 * k2; d1 = f13 * f0 + f16 * fp; }
          denominator = b1;
          if constexpr (F32) {
            rden = 1.0 / denominator;
            fprev = d1 * rden;
          } else {
            fprev = d1 / denominator;
          }
          F[0] = (FT)fprev;
          cprev = c1;
          continue;
        }
        const double fm = f0;
        f0 = fp;
        fp = (j + 1 < NZ) ? onext : 0.0;
        double a, b, c, d;
        if (j + 1 < NZ) {
          a = ai; b = bi; c = ci;
          d = f112 * fm + f56 * f0 + f112 * fp;
        } else if (!df0) {
          a = 0.; b = 1.; c = 0.; d = 0.;
        } else {
          a = idz2 - f16 * k2; b = -idz2 - f13 * k2; c = 0.;
          d = f13 * f0 + f16 * fm;
        }
        const double cpk = F32 ? cprev * rden : cprev / denominator;
        CPl[i] = (FT)cpk;
        denominator = b - a * cpk;
        bad |= (denominator == 0);
        if constexpr (F32) {
          rden = 1.0 / denominator;
          fprev = (d - a * fprev) * rden;
        } else {
          fprev = (d - a * fprev) / (denominator);
        }
        F[i] = (FT)fprev;
        cprev = c;
      }
    }
  }
  const double cpn = __shfl((double)CPl[0], lane + CPW);
  double fnext = fprev;
  for (int p = PTOP; p >= 0; p--) {
    if (p < PTOP) fnext = __shfl(fnext, lane + CPW);
    if (q == p) {
#pragma unroll
      for (int i = L - 1; i >= 0; i--) {
        const int j = p * L + i;
        if (j >= NZ - 1) continue;
        const double cpj1 = (i + 1 < L) ? (double)CPl[i + 1 < L ? i + 1 : 0] : cpn;
        const double v = (double)F[i] - cpj1 * fnext;
        F[i] = (FT)v;
        fnext = v;
      }
    }
  }
  if (bad) raise_status(sc, PMPIC_ERR_ZERO_PIVOT);
#pragma unroll
  for (int i = 0; i < L; i++)
    if (q * L + i < NZ) f[(int64_t)(q * L + i) * plane] = (S)F[i];
}

template <int Q, int L, typename S>
__global__ void __launch_bounds__(64) k_diffz_qn(Geometry g, ColFields<S> cf,
                                                 const double *__restrict__ tden,
                                                 const double *__restrict__ tcp,
                                                 DevScalars *sc) {
  constexpr int CPW = 64 / Q;
  // the FP32-bulk form as in k_diffz_reg: pivot reciprocals, 1/dz factors, FP32 column
  constexpr bool F32 = std::is_same<S, float>::value;
  using FT = typename std::conditional<F32, float, double>::type;
  __shared__ double den[Q * L], cpt[Q * L];
  const int NZ = g.nzn;
  const int lane = threadIdx.x;
  const int q = lane / CPW;
  const int64_t plane = (int64_t)g.sy * g.sx;
  const int64_t col = (int64_t)blockIdx.x * CPW + (lane % CPW);
  const int fi = blockIdx.y;
  const bool bt = cf.bot[fi] != nullptr && cf.top[fi] != nullptr;
  for (int k = threadIdx.x; k < NZ; k += blockDim.x) {
    const double dn = tden[(bt ? NZ : 0) + k];
    den[k] = F32 ? 1.0 / dn : dn;
    cpt[k] = tcp[(bt ? NZ : 0) + k];
  }
  __syncthreads();
  if (col >= plane) return;  // all Q lanes of a column leave together
  const S *fin = cf.in[fi] + col;
  S *o = cf.out[fi] + col;
  FT F[L];
#pragma unroll
  for (int i = 0; i < L; i++) F[i] = (q * L + i < NZ) ? fin[(int64_t)(q * L + i) * plane] : (S)0;
  const double bot = bt ? cf.bot[fi][col] : 0.0, top = bt ? cf.top[fi][col] : 0.0;
  const double a1 = (double)(1.f / 6.f), al = (double)(1.f / 3.f);
  const double dz = g.dz, idz = 1.0 / dz, ih2dz = 0.5 / dz;
  const double nxt0 = __shfl((double)F[0], lane + CPW);  // the original of the next group's first level
  const int PTOP = (NZ - 1) / L;
  double fprev = 0.0, fkm2 = 0.0, fkm1 = 0.0;  // originals f(k-2), f(k-1) of the next row
  bool bad = false;
  for (int p = 0; p <= PTOP; p++) {
    if (p > 0) {
      fprev = __shfl(fprev, lane - CPW);
      fkm2 = __shfl(fkm2, lane - CPW);
      fkm1 = __shfl(fkm1, lane - CPW);
    }
    if (q == p) {
#pragma unroll
      for (int i = 0; i < L; i++) {
        const int j = p * L + i;  // level written: row j + 1 of the reference
        if (j >= NZ) break;
        const double onext = (i + 1 < L) ? F[i + 1 < L ? i + 1 : 0] : nxt0;  // original j + 1
        if (j == 0) {
          fkm2 = F[0];
          fkm1 = onext;
          const double d0 = bt ? (F32 ? fkm1 * idz : fkm1 / dz) - a1 * bot * dz : 0.0;
          if constexpr (F32) {
            bad |= !__builtin_isfinite(den[0]);
            fprev = d0 * den[0];
          } else {
            bad |= (den[0] == 0);
            fprev = d0 / den[0];
          }
          F[0] = (FT)fprev;
          continue;
        }
        double a, d, fk = 0.0;
        if (j + 1 < NZ) {
          a = a1;
          fk = onext;
          d = F32 ? (fk - fkm2) * ih2dz : (fk - fkm2) / 2. / dz;
        } else {
          a = bt ? al : 0.;
          d = bt ? (F32 ? -fkm2 * idz : -fkm2 / dz) + a1 * top * dz : 0.0;
        }
        const double dn = den[j];
        if constexpr (F32) {
          bad |= !__builtin_isfinite(dn);
          fprev = (d - a * fprev) * dn;
        } else {
          bad |= (dn == 0);
          fprev = (d - a * fprev) / (dn);
        }
        F[i] = (FT)fprev;
        fkm2 = fkm1;
        fkm1 = fk;
      }
    }
  }
  double fnext = fprev;
  for (int p = PTOP; p >= 0; p--) {
    if (p < PTOP) fnext = __shfl(fnext, lane + CPW);
    if (q == p) {
#pragma unroll
      for (int i = L - 1; i >= 0; i--) {
        const int j = p * L + i;
        if (j >= NZ - 1) continue;
        const double v = (double)F[i] - cpt[j + 1] * fnext;
        F[i] = (FT)v;
        fnext = v;
      }
    }
  }
  if (bad) raise_status(sc, PMPIC_ERR_ZERO_PIVOT);
#pragma unroll
  for (int i = 0; i < L; i++)
    if (q * L + i < NZ) o[(int64_t)(q * L + i) * plane] = (S)F[i];
}

// u, v, w spectra (vort2vel.F90:251-320): fu = dzb - diffy(c), column (1,1) = ubar;
// fv = diffx(c) - dza, column (1,1) = vbar; fw = diffy(a) - diffx(b); u_s.. = filtered
template <typename S>
__global__ void k_velocity_spectra(Geometry g, const S *__restrict__ a,
                                   const S *__restrict__ b, const S *__restrict__ c,
                                   const S *__restrict__ dzb, const S *__restrict__ dza,
                                   const double *__restrict__ ubar, const double *__restrict__ vbar,
                                   S *__restrict__ fu, S *__restrict__ fv,
                                   S *__restrict__ fw, S *__restrict__ us,
                                   S *__restrict__ vs, S *__restrict__ ws,
                                   const double *__restrict__ kx, const double *__restrict__ ky,
                                   const double *__restrict__ filt) {
  SPEC_ELEMENT(g);
  const double fl = filt[col];
  double u = (double)dzb[q] - dy_at(c, ky, q, jj, g.sx, col);
  if (col == 0) u = ubar[k];
  double v = dx_at(c, kx, q, ii, col) - (double)dza[q];
  if (col == 0) v = vbar[k];
  const double w = dy_at(a, ky, q, jj, g.sx, col) - dx_at(b, kx, q, ii, col);
  fu[q] = (S)u; fv[q] = (S)v; fw[q] = (S)w;
  us[q] = (S)(u * fl); vs[q] = (S)(v * fl); ws[q] = (S)(w * fl);
}

// vorticity_tendency.F90:169-265, physical interior, accumulation order as written
// T: the type of the input grids (FP64, or FP32 in the FP32-bulk form, widened exactly); the
// products are formed in FP64 in both forms
template <typename T>
__global__ void k_tendency(int64_t nnode, const T *__restrict__ P, const T *__restrict__ Q,
                           const T *__restrict__ R, const T *__restrict__ dbdy,
                           const T *__restrict__ dudx, const T *__restrict__ dudy,
                           const T *__restrict__ dwdx, const T *__restrict__ dbdx,
                           const T *__restrict__ dvdx, const T *__restrict__ dvdy,
                           const T *__restrict__ dwdy, double *__restrict__ DP,
                           double *__restrict__ DQ, double *__restrict__ DR,
                           float *__restrict__ DP32, float *__restrict__ DQ32,
                           float *__restrict__ DR32) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nnode) return;
  const double p = P[q], qq = Q[q], r = R[q];
  double dp = dbdy[q];
  dp = dp + p * (double)dudx[q];
  dp = dp + qq * (double)dudy[q];
  dp = dp + r * (qq + (double)dwdx[q]);
  double dq = dbdx[q];
  dq = -dq + p * (double)dvdx[q];
  dq = dq + qq * (double)dvdy[q];
  dq = dq + r * (-p + (double)dwdy[q]);
  double dr = p * (double)dwdx[q];
  dr = dr + qq * (double)dwdy[q];
  dr = dr + r * (-(double)dudx[q]);
  dr = dr + r * (-(double)dvdy[q]);
  if (DP) {
    DP[q] = dp;
    DQ[q] = dq;
    DR[q] = dr;
  }
  if (DP32) {
    DP32[q] = (float)dp;
    DQ32[q] = (float)dq;
    DR32[q] = (float)dr;
  }
}

// maxima of u^2, v^2, w^2 and p^2+q^2+r^2 over the interior (vort2vel.F90:331-337,
// vorticity_tendency.F90:279-282), per block; a maximum is exact in any order, so the result
// is deterministic.
// TU: the grids' type (double, or float for the FP32-bulk form's grids)
template <typename TU>
__global__ void k_max4(int64_t nnode, const TU *__restrict__ U, const TU *__restrict__ V,
                       const TU *__restrict__ W, const TU *__restrict__ P,
                       const TU *__restrict__ Q, const TU *__restrict__ R,
                       double *__restrict__ blkmax) {
  __shared__ double sm[4][TPB];
  double m[4] = {0, 0, 0, 0};
  // MPT nodes per pass, their loads issued together; an index past the end is clamped to the
  // last node, which the maximum tolerates
  constexpr int MPT = 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q0 < nnode;
       q0 += MPT * stride) {
    double u[MPT], v[MPT], w[MPT], pp[MPT], qq[MPT], rr[MPT];  // velocities widened exactly
#pragma unroll
    for (int k = 0; k < MPT; k++) {
      const int64_t q = min(q0 + k * stride, nnode - 1);
      u[k] = U[q]; v[k] = V[q]; w[k] = W[q];
      pp[k] = P[q]; qq[k] = Q[q]; rr[k] = R[q];
    }
#pragma unroll
    for (int k = 0; k < MPT; k++) {
      const double o = pp[k] * pp[k] + qq[k] * qq[k] + rr[k] * rr[k];
      m[0] = fmax(m[0], u[k] * u[k]);
      m[1] = fmax(m[1], v[k] * v[k]);
      m[2] = fmax(m[2], w[k] * w[k]);
      m[3] = fmax(m[3], o);
    }
  }
  for (int f = 0; f < 4; f++) sm[f][threadIdx.x] = m[f];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int f = 0; f < 4; f++) sm[f][threadIdx.x] = fmax(sm[f][threadIdx.x], sm[f][threadIdx.x + s]);
    __syncthreads();
  }
  // one value per block and quantity; k_max4_final reduces them (no contended atomics)
  if (threadIdx.x == 0)
    for (int f = 0; f < 4; f++) blkmax[(int64_t)f * gridDim.x + blockIdx.x] = sm[f][0];
}

// the maxima over the nb blocks of k_max4 (one workgroup); a maximum is exact in any order
__global__ void k_max4_final(int nb, const double *__restrict__ blkmax, DevScalars *sc) {
  __shared__ double sm[4][TPB];
  double m[4] = {0, 0, 0, 0};
  for (int i = threadIdx.x; i < nb; i += blockDim.x)
#pragma unroll
    for (int f = 0; f < 4; f++) m[f] = fmax(m[f], blkmax[(int64_t)f * nb + i]);
  for (int f = 0; f < 4; f++) sm[f][threadIdx.x] = m[f];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int f = 0; f < 4; f++) sm[f][threadIdx.x] = fmax(sm[f][threadIdx.x], sm[f][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sc->umax2 = sm[0][0];
    sc->vmax2 = sm[1][0];
    sc->wmax2 = sm[2][0];
    sc->omax2 = sm[3][0];
  }
}

// timestep selection at RK stage 1: vort2vel.F90:329-381 then vorticity_tendency.F90:277-321
__global__ void k_select_dt(Geometry g, DevScalars *sc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double umax = sqrt(sc->umax2), vmax = sqrt(sc->vmax2), wmax = sqrt(sc->wmax2);
  const double tiny = (double)1.E-15f;
  if (umax == 0) umax = tiny;
  if (vmax == 0) vmax = tiny;
  if (wmax == 0) wmax = tiny;
  const double t1 = g.lx_cfl / umax, t2 = g.ly_cfl / vmax, t3 = g.lz_cfl / wmax;
  double dtmaxg = t1;
  if (t2 < dtmaxg) dtmaxg = t2;
  if (t3 < dtmaxg) dtmaxg = t3;
  double dtm = (sc->dtmax > dtmaxg) ? dtmaxg : sc->dtmax;
  const double omax = sqrt(sc->omax2);
  double dtv;
  if (omax > 0.) dtv = sc->prefactor / omax;
  else dtv = sc->dtmax;
  if (dtm > dtv) dtm = dtv;
  sc->dtm = dtm;
}

inline int64_t ssz_of(const Geometry &g) { return (int64_t)g.nzn * g.sy * g.sx; }
inline int64_t plane_of(const Geometry &g) { return (int64_t)g.sy * g.sx; }

}  // namespace

// wavenumbers and 2/3-rule filter, fft_ops.F90:111-139 on one rank (x_start = y_start = 1)
int spectral_setup(Ctx *c) {
  const Geometry &g = c->g;
  const int sx = g.sx, sy = g.sy;
  std::vector<double> kx(sx * sy), ky(sx * sy), k2(sx * sy), fl(sx * sy);
  const double twopi = 8.0 * atan(1.0);
  const double lx = g.dx * (double)g.nx, ly = g.dy * (double)g.ny;
  const double hx = (double)((float)g.nx / 2.0f), hy = (double)((float)g.ny / 2.0f);
  double ksqmax = (hx / lx) * (hx / lx) + (hy / ly) * (hy / ly);
  ksqmax = ksqmax * 2. / 9.;
  for (int jj = 0; jj < sy; jj++)
    for (int ii = 0; ii < sx; ii++) {
      const double xval = (double)(ii / 2) / lx;
      const double yval = (double)(jj / 2) / ly;
      const double kk = xval * xval + yval * yval;
      const int q = jj * sx + ii;
      fl[q] = (kk > ksqmax) ? 0. : 1.;
      kx[q] = xval * twopi;
      ky[q] = yval * twopi;
      k2[q] = kk * twopi * twopi;
    }
  const size_t b = sizeof(double) * sx * sy;
  PMPIC_CHECK(hipMemcpy(c->kx, kx.data(), b, hipMemcpyHostToDevice));
  PMPIC_CHECK(hipMemcpy(c->ky, ky.data(), b, hipMemcpyHostToDevice));
  PMPIC_CHECK(hipMemcpy(c->k2, k2.data(), b, hipMemcpyHostToDevice));
  PMPIC_CHECK(hipMemcpy(c->filt, fl.data(), b, hipMemcpyHostToDevice));
  const int64_t plane = (int64_t)sx * sy;
  hipLaunchKernelGGL(k_dz_tables, dim3(1), dim3(1), 0, c->stream, g, c->dz_den, c->dz_cp);
  PMPIC_CHECK(hipGetLastError());
  PMPIC_CHECK(hipStreamSynchronize(c->stream));
  return PMPIC_OK;
}

int op_diffx(Ctx *c, const double *in, double *out) {
  hipLaunchKernelGGL(k_diffx, dim3(cdiv(ssz_of(c->g), TPB)), dim3(TPB), 0, c->stream, c->g, in,
                     out, c->kx);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}
int op_diffy(Ctx *c, const double *in, double *out) {
  hipLaunchKernelGGL(k_diffy, dim3(cdiv(ssz_of(c->g), TPB)), dim3(TPB), 0, c->stream, c->g, in,
                     out, c->ky);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}
// threads per block of the LDS column kernels: the most columns (64 down to 8) whose two
// slices (f and cp) fit in 64 KiB of LDS; 0 = columns too long, solve straight from HBM
static int lds_cols(const Geometry &g) {
  for (int nc = 64; nc >= 8; nc /= 2)
    if (2 * (int64_t)nc * g.nzn * (int64_t)sizeof(double) <= 65536) return nc;
  return 0;
}

// column lengths (z nodes) with a register-resident solver instance
#define PMPIC_REG_COLUMNS(X) X(9) X(17) X(33) X(65) X(129)
// longer columns, up to LONG_Q * LONG_L z nodes: LONG_Q lanes per column, LONG_L levels each
constexpr int LONG_Q = 8, LONG_L = 33;
// other lengths up to 129 z nodes (48^3 and 96^3 grids, say) in the same lane-group form: four
// lanes per column with 13 levels each up to 52 nodes, 33 each up to 129
constexpr int MID_Q = 4, MID_L_SHORT = 13, MID_L = 33;
// the FP32-bulk form, whose columns sit in registers as FP32, takes long columns up to
// F32_Q * F32_L z nodes with four lanes per column (sixteen columns a wavefront, twice the FP64
// form's), and the 257-node columns of 256-cell-deep grids in the register-resident solvers
constexpr int F32_Q = 4, F32_L = 65, F32_REG_LONG = 257;
// lanes per column of the register-resident laplinv: four, but two (32 columns a wavefront) for
// the FP32-bulk form's columns of up to F32_LAP_Q2 nodes, whose column and c' then still fit in
// registers at three wavefronts per SIMD (18.8 against 24.7 us a call at 65 nodes, 128^2
// columns).  One lane per column, or two on longer columns, gains nothing (the sweeps are then
// held at one or two wavefronts per SIMD by their registers)
constexpr int F32_LAP_Q2 = 65;
template <typename S>
constexpr int lap_q(int nz) {
  return std::is_same<S, float>::value && nz <= F32_LAP_Q2 ? 2 : 4;
}

template <typename S>
int op_filter_t(Ctx *c, const S *in, S *out) {
  hipLaunchKernelGGL(k_filter<S>, spec_grid(c->g), dim3(TPB), 0, c->stream, c->g,
                     in, out, c->filt);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}
int op_filter(Ctx *c, const double *in, double *out) { return op_filter_t<double>(c, in, out); }

template <typename S>
int op_diffz_n(Ctx *c, int nf, const S *const *in, S *const *out, const double *const *bot,
               const double *const *top) {
  if (nf < 1 || nf > 4) return PMPIC_ERR_ARG;
  const int64_t plane = plane_of(c->g);
  const int nc = lds_cols(c->g);
  ColFields<S> rf{};
  for (int f = 0; f < nf; f++) {
    rf.in[f] = in[f];
    rf.out[f] = out[f];
    rf.bot[f] = bot ? bot[f] : nullptr;
    rf.top[f] = top ? top[f] : nullptr;
  }
  switch (c->g.nzn) {
#define PMPIC_DZ_CASE(NZ)                                                                     \
  case NZ:                                                                                     \
    hipLaunchKernelGGL((k_diffz_reg<NZ, S>), dim3(cdiv(plane, 64), nf), dim3(64), 0, c->stream, \
                       c->g, rf, c->dz_den, c->dz_cp, c->dsc);                                 \
    PMPIC_CHECK(hipGetLastError());                                                            \
    return PMPIC_OK;
    PMPIC_REG_COLUMNS(PMPIC_DZ_CASE)
#undef PMPIC_DZ_CASE
    default:
      break;
  }
  if (c->g.nzn <= MID_Q * MID_L_SHORT) {
    hipLaunchKernelGGL((k_diffz_qn<MID_Q, MID_L_SHORT, S>), dim3(cdiv(plane, 64 / MID_Q), nf),
                       dim3(64), 0, c->stream, c->g, rf, c->dz_den, c->dz_cp, c->dsc);
    PMPIC_CHECK(hipGetLastError());
    return PMPIC_OK;
  }
  if (c->g.nzn <= 129) {
    hipLaunchKernelGGL((k_diffz_qn<MID_Q, MID_L, S>), dim3(cdiv(plane, 64 / MID_Q), nf),
                       dim3(64), 0, c->stream, c->g, rf, c->dz_den, c->dz_cp, c->dsc);
    PMPIC_CHECK(hipGetLastError());
    return PMPIC_OK;
  }
  if constexpr (std::is_same<S, float>::value) {
    if (c->g.nzn == F32_REG_LONG) {
      hipLaunchKernelGGL((k_diffz_reg<F32_REG_LONG, S>), dim3(cdiv(plane, 64), nf), dim3(64), 0,
                         c->stream, c->g, rf, c->dz_den, c->dz_cp, c->dsc);
      PMPIC_CHECK(hipGetLastError());
      return PMPIC_OK;
    }
    if (c->g.nzn <= F32_Q * F32_L) {
      hipLaunchKernelGGL((k_diffz_qn<F32_Q, F32_L, S>), dim3(cdiv(plane, 64 / F32_Q), nf),
                         dim3(64), 0, c->stream, c->g, rf, c->dz_den, c->dz_cp, c->dsc);
      PMPIC_CHECK(hipGetLastError());
      return PMPIC_OK;
    }
  }
  if (c->g.nzn <= LONG_Q * LONG_L) {
    hipLaunchKernelGGL((k_diffz_qn<LONG_Q, LONG_L, S>), dim3(cdiv(plane, 64 / LONG_Q), nf),
                       dim3(64), 0, c->stream, c->g, rf, c->dz_den, c->dz_cp, c->dsc);
    PMPIC_CHECK(hipGetLastError());
    return PMPIC_OK;
  }
  if (nc > 0) {
    hipLaunchKernelGGL(k_diffz_lds<S>, dim3(cdiv(plane, nc), nf), dim3(nc),
                       (size_t)2 * nc * c->g.nzn * sizeof(double), c->stream, c->g, rf, c->dsc);
  } else {
    for (int f = 0; f < nf; f++) {  // one field at a time: sder[7] holds one field's cp
      ColFields<S> one{};
      one.in[0] = in[f];
      one.out[0] = out[f];
      one.bot[0] = bot ? bot[f] : nullptr;
      one.top[0] = top ? top[f] : nullptr;
      hipLaunchKernelGGL(k_diffz<S>, dim3(cdiv(plane, 64), 1), dim3(64), 0, c->stream, c->g, one,
                         c->sder[7], c->dsc);
    }
  }
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}
int op_diffz(Ctx *c, const double *in, double *out, const double *bot, const double *top) {
  return op_diffz_n<double>(c, 1, &in, &out, bot ? &bot : nullptr, top ? &top : nullptr);
}

template <typename S>
int op_laplinv_n(Ctx *c, int nf, S *const *f, const int *df0) {
  if (nf < 1 || nf > 4) return PMPIC_ERR_ARG;
  const int64_t plane = plane_of(c->g);
  const int nc = lds_cols(c->g);
  ColFields<S> rf{};
  for (int q = 0; q < nf; q++) {
    rf.out[q] = f[q];
    rf.df0[q] = df0[q] ? 1 : 0;
  }
  switch (c->g.nzn) {
#define PMPIC_LAP_CASE(NZ)                                                                    \
  case NZ:                                                                                     \
    hipLaunchKernelGGL((k_laplinv_reg<lap_q<S>(NZ), NZ, S>), dim3(cdiv(plane, 64 / lap_q<S>(NZ)), \
                       nf), dim3(64), 0, c->stream, c->g, rf, c->k2, c->dsc);                  \
    PMPIC_CHECK(hipGetLastError());                                                            \
    return PMPIC_OK;
    PMPIC_REG_COLUMNS(PMPIC_LAP_CASE)
#undef PMPIC_LAP_CASE
    default:
      break;
  }
  if (c->g.nzn <= MID_Q * MID_L_SHORT) {
    hipLaunchKernelGGL((k_laplinv_qn<MID_Q, MID_L_SHORT, S>), dim3(cdiv(plane, 64 / MID_Q), nf),
                       dim3(64), 0, c->stream, c->g, rf, c->k2, c->dsc);
    PMPIC_CHECK(hipGetLastError());
    return PMPIC_OK;
  }
  if (c->g.nzn <= 129) {
    hipLaunchKernelGGL((k_laplinv_qn<MID_Q, MID_L, S>), dim3(cdiv(plane, 64 / MID_Q), nf),
                       dim3(64), 0, c->stream, c->g, rf, c->k2, c->dsc);
    PMPIC_CHECK(hipGetLastError());
    return PMPIC_OK;
  }
  if constexpr (std::is_same<S, float>::value) {
    if (c->g.nzn == F32_REG_LONG) {
      constexpr int Q = lap_q<S>(F32_REG_LONG);
      hipLaunchKernelGGL((k_laplinv_reg<Q, F32_REG_LONG, S>), dim3(cdiv(plane, 64 / Q), nf),
                         dim3(64), 0, c->stream, c->g, rf, c->k2, c->dsc);
      PMPIC_CHECK(hipGetLastError());
      return PMPIC_OK;
    }
    if (c->g.nzn <= F32_Q * F32_L) {
      hipLaunchKernelGGL((k_laplinv_qn<F32_Q, F32_L, S>), dim3(cdiv(plane, 64 / F32_Q), nf),
                         dim3(64), 0, c->stream, c->g, rf, c->k2, c->dsc);
      PMPIC_CHECK(hipGetLastError());
      return PMPIC_OK;
    }
  }
  if (c->g.nzn <= LONG_Q * LONG_L) {
    hipLaunchKernelGGL((k_laplinv_qn<LONG_Q, LONG_L, S>), dim3(cdiv(plane, 64 / LONG_Q), nf),
                       dim3(64), 0, c->stream, c->g, rf, c->k2, c->dsc);
    PMPIC_CHECK(hipGetLastError());
    return PMPIC_OK;
  }
  if (nc > 0) {
    hipLaunchKernelGGL(k_laplinv_lds<S>, dim3(cdiv(plane, nc), nf), dim3(nc),
                       (size_t)2 * nc * c->g.nzn * sizeof(double), c->stream, c->g, rf, c->k2,
                       c->dsc);
  } else {
    for (int q = 0; q < nf; q++) {
      ColFields<S> one{};
      one.out[0] = f[q];
      one.df0[0] = df0[q];
      hipLaunchKernelGGL(k_laplinv<S>, dim3(cdiv(plane, 64), 1), dim3(64), 0, c->stream, c->g,
                         one, c->sder[7], c->k2, c->dsc);
    }
  }
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}
int op_laplinv(Ctx *c, double *f, int df0) { return op_laplinv_n<double>(c, 1, &f, &df0); }

namespace {
// the spectral scratch arrays as arrays of the form's spectral type (the FP32-bulk form uses
// the first half of each FP64 allocation)
template <typename S> S *sp(double *a) { return reinterpret_cast<S *>(a); }
}  // namespace

// vort2vel timestep_callback (vort2vel.F90:134-320) after par2grid of p, q, r into the grids:
// corrected and filtered vorticity back into grid[GP..GR], velocities into grid[GU..GW], the
// filtered velocity spectra into su, sv, sw.  S: the spectral arrays' type (FP32 in the
// FP32-bulk form: its transforms and spectral arrays are FP32, the mean flow, the boundary
// planes and the column solvers' arithmetic FP64).
// the physical p, q, r and total-buoyancy grids (k = 0..3) in the transforms' precision: the
// FP64 grids GP, GQ, GR, GB, or the FP32-bulk form's FP32 grids
static_assert(GQ == GP + 1 && GR == GP + 2 && GB == GP + 3, "grid ids p, q, r, b in sequence");
template <typename S> S *pqrb(Ctx *c, int k);
template <> double *pqrb<double>(Ctx *c, int k) { return c->grid[GP + k]; }
template <> float *pqrb<float>(Ctx *c, int k) { return c->b32.pgrid[k]; }

template <typename S>
static int vort2vel_spectral_t(Ctx *c) {
  const Geometry &g = c->g;
  const dim3 eb = spec_grid(g), et(TPB);
  S *sa = sp<S>(c->sa), *sb = sp<S>(c->sb), *sc = sp<S>(c->sc), *sd = sp<S>(c->sd);
  S *se = sp<S>(c->se), *sf = sp<S>(c->sf);
  S *sder[5] = {sp<S>(c->sder[0]), sp<S>(c->sder[1]), sp<S>(c->sder[2]), sp<S>(c->sder[3]),
                sp<S>(c->sder[4])};
  {
    const S *src[3] = {pqrb<S>(c, 0), pqrb<S>(c, 1), pqrb<S>(c, 2)};
    S *dst[3] = {sa, sb, sc};
    PMPIC_TRY(fft_forward_t<S>(c, 3, src, dst));
  }
  const double hdzi = 1. / 2. / g.dz;
  hipLaunchKernelGGL(k_divergence<S>, eb, et, 0, c->stream, g, sa, sb, sc, sf, c->kx, c->ky,
                     hdzi);
  {
    const int df0[1] = {1};
    PMPIC_TRY(op_laplinv_n<S>(c, 1, &sf, df0));
  }
  PMPIC_TRY(op_filter_t<S>(c, sf, sf));
  hipLaunchKernelGGL(k_correct_ab<S>, eb, et, 0, c->stream, g, sa, sb, sf, c->kx, c->ky);
  {
    const S *in[1] = {sf};
    PMPIC_TRY(op_diffz_n<S>(c, 1, in, &sd, nullptr, nullptr));
  }
  hipLaunchKernelGGL(k_correct_c_filter<S>, eb, et, 0, c->stream, g, sa, sb, sc, sd, sd, se, sf,
                     c->filt, c->atop, c->abot, c->btop, c->bbot);
  hipLaunchKernelGGL(k_mean_flow<S>, dim3(1), dim3(256), (size_t)(4 * g.nzn + 2) * sizeof(double),
                     c->stream, g, sd, se, c->ubar, c->vbar);
  {
    S *fl[3] = {sa, sb, sc};
    const int df0[3] = {0, 0, 1};
    PMPIC_TRY(op_laplinv_n<S>(c, 3, fl, df0));
  }
  {
    // dz(b) with b's cached top/bottom into sder[0], dz(a) into sder[1]
    const S *in[2] = {sb, sa};
    S *out[2] = {sder[0], sder[1]};
    const double *bot[2] = {c->bbot, c->abot}, *top[2] = {c->btop, c->atop};
    PMPIC_TRY(op_diffz_n<S>(c, 2, in, out, bot, top));
  }
  hipLaunchKernelGGL(k_velocity_spectra<S>, eb, et, 0, c->stream, g, sa, sb, sc, sder[0],
                     sder[1], c->ubar, c->vbar, sder[2], sder[3], sder[4], sp<S>(c->su),
                     sp<S>(c->sv), sp<S>(c->sw), c->kx, c->ky, c->filt);
  {
    const S *src[6] = {sd, se, sf, sder[2], sder[3], sder[4]};
    // the FP32-bulk form keeps p, q, r in its FP32 grids and u, v, w only as the FP32 grids its
    // gathers read
    const bool b32 = c->precision == 1;
    double *dst[6] = {c->grid[GP], c->grid[GQ], c->grid[GR], c->grid[GU], c->grid[GV],
                      c->grid[GW]};
    float *dst32[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (b32)
      for (int a = 0; a < 3; a++) {
        dst[a] = dst[3 + a] = nullptr;
        dst32[a] = c->b32.pgrid[a];
        dst32[3 + a] = c->b32.grid[a];
      }
    PMPIC_TRY(fft_backward_t<S>(c, 6, src, dst, dst32, nullptr));
  }
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int vort2vel_spectral(Ctx *c, int stage) {
  (void)stage;
  return c->precision == 1 ? vort2vel_spectral_t<float>(c) : vort2vel_spectral_t<double>(c);
}

// vorticity_tendency timestep_callback (vorticity_tendency.F90:147-265) after par2grid of btot
// into grid[GB]; tendencies into grid[GDP..GDR] (plain FP64 form) or only into the FP32 grids
// the FP32-bulk gathers read
template <typename S>
static int vorticity_tendency_spectral_t(Ctx *c) {
  const Geometry &g = c->g;
  const bool b32 = c->precision == 1;  // the FP32-bulk form: FP32 grids, FP32 tendencies only
  S *sbs = sp<S>(c->sbs);
  {
    const S *src[1] = {pqrb<S>(c, 3)};
    S *dst[1] = {sbs};
    PMPIC_TRY(fft_forward_t<S>(c, 1, src, dst));
  }
  {
    // the eight spectral derivatives of vorticity_tendency.F90:155-232 (db/dy, du/dx, du/dy,
    // dw/dx, db/dx, dv/dx, dv/dy, dw/dy), each taken while its inverse transform reads it
    S *su = sp<S>(c->su), *sv = sp<S>(c->sv), *sw = sp<S>(c->sw);
    const S *src[8] = {sbs, su, su, sw, sbs, sv, sv, sw};
    const int op[8] = {SPEC_DY, SPEC_DX, SPEC_DY, SPEC_DX, SPEC_DX, SPEC_DX, SPEC_DY, SPEC_DY};
    // the derivative grids in the transforms' precision (the FP32-bulk form's in the first
    // half of each FP64 scratch grid)
    double *dst[8];
    float *dst32[8];
    for (int f = 0; f < 8; f++) {
      dst[f] = b32 ? nullptr : c->gtmp[f];
      dst32[f] = b32 ? reinterpret_cast<float *>(c->gtmp[f]) : nullptr;
    }
    PMPIC_TRY(fft_backward_t<S>(c, 8, src, dst, dst32, op));
  }
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  const S *d[8];
  for (int f = 0; f < 8; f++) d[f] = reinterpret_cast<const S *>(c->gtmp[f]);
  hipLaunchKernelGGL(k_tendency<S>, dim3(cdiv(nnode, TPB)), dim3(TPB), 0, c->stream, nnode,
                     pqrb<S>(c, 0), pqrb<S>(c, 1), pqrb<S>(c, 2), d[0], d[1], d[2], d[3], d[4],
                     d[5], d[6], d[7], b32 ? nullptr : c->grid[GDP], b32 ? nullptr : c->grid[GDQ],
                     b32 ? nullptr : c->grid[GDR], b32 ? c->b32.grid[3] : nullptr,
                     b32 ? c->b32.grid[4] : nullptr, b32 ? c->b32.grid[5] : nullptr);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

int vorticity_tendency_spectral(Ctx *c, int stage) {
  (void)stage;
  return c->precision == 1 ? vorticity_tendency_spectral_t<float>(c)
                           : vorticity_tendency_spectral_t<double>(c);
}

int grid_max_reduce(Ctx *c) {
  const Geometry &g = c->g;
  const int64_t nnode = (int64_t)g.nzn * g.ny * g.nx;
  int64_t nb = cdiv(nnode, TPB);
  if (nb > 1024) nb = 1024;  // c->maxblk holds 4 * 1024 block maxima
  if (c->precision == 1)  // the FP32-bulk form's grids are FP32
    hipLaunchKernelGGL(k_max4<float>, dim3(nb), dim3(TPB), 0, c->stream, nnode, c->b32.grid[0],
                       c->b32.grid[1], c->b32.grid[2], c->b32.pgrid[0], c->b32.pgrid[1],
                       c->b32.pgrid[2], c->maxblk);
  else
    hipLaunchKernelGGL(k_max4<double>, dim3(nb), dim3(TPB), 0, c->stream, nnode, c->grid[GU],
                       c->grid[GV], c->grid[GW], c->grid[GP], c->grid[GQ], c->grid[GR],
                       c->maxblk);
  hipLaunchKernelGGL(k_max4_final, dim3(1), dim3(TPB), 0, c->stream, (int)nb, c->maxblk, c->dsc);
  hipLaunchKernelGGL(k_select_dt, dim3(1), dim3(1), 0, c->stream, g, c->dsc);
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

}  // namespace pmpic
