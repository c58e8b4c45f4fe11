// Semi-spectral 2-D transforms on MI355X (periodic x, y; z stays physical).
//
// Reference: model_core/src/fft/pencilfft.F90:151-212 (forward: r2c along y, (Re,Im) pairs
// interleaved, then r2c along x of each real y-row; backward: the reverse), over FFTE's 1-D
// transforms wrapped by model_core/src/fft/ffte_mod.F90:66-124 (forward sum x_j e^{-2 pi i jk/N},
// unnormalised; inverse = conj . forward . conj, scaled by 1/N, real part kept, so the imaginary
// parts of the DC and Nyquist bins are dropped).
//
// The spectral element S[k][2m+a][2l+b] is component b (Re/Im) at x-wavenumber l of the x-DFT
// of component a of the y-DFT at y-wavenumber m.  With F(m,l) the full 2-D DFT,
//   x-DFT of Re Y_m = (F(m,l) + F(-m,l))/2,   x-DFT of Im Y_m = (F(m,l) - F(-m,l))/(2i),
// so the layout follows from one real-to-complex pass along the contiguous x rows and one
// complex pass along y.  Each pass is one kernel: a workgroup stages a batch of rows (or of
// adjacent columns, for coalesced strided access) in LDS and runs a mixed radix-4/2/3/5
// Stockham FFT there, so each pass reads and writes HBM once.  The inverse builds, per (z, m),
// the complex row H = A_m + i B_m of the two real x-spectra, transforms it along x, and finishes
// with a complex-to-real pass along y.
#include "common.h"

#include <math.h>
#include <type_traits>
#include <vector>

namespace pmpic {

namespace {

// radix schedule: radix of stage s in bits [4 s, 4 s + 4) of `code` (held in registers; no
// per-stage memory read)
struct Plan {
  int n;
  int nrad;
  uint64_t code;
};

// radix 4 stages first, then 2, 3, 5; nrad = -1 when n has another prime factor
__host__ __device__ constexpr Plan make_plan(int n) {
  Plan p{n, 0, 0};
  int m = n;
  const int radix[4] = {4, 2, 3, 5};
  for (int a = 0; a < 4; a++)
    while (m % radix[a] == 0 && p.nrad < 16) {
      p.code |= (uint64_t)radix[a] << (4 * p.nrad);
      p.nrad++;
      m /= radix[a];
    }
  if (m != 1) p.nrad = -1;
  return p;
}

// Transform lengths with kernels compiled for them (every index computation then divides by
// constants); other lengths run the same kernels with the length read at run time (NT = 0).
constexpr int FIXED_N[] = {16, 32, 48, 64, 96, 128, 256, 512};

constexpr int FFT_TPB = 256;
constexpr int MAXF = 8;

// The transforms run in the precision S of the spectral arrays: FP64 (double2 complex
// elements), or FP32 (float2) in the FP32-bulk form, whose LDS buffers then hold twice as many
// elements, so a column pass covers twice as many adjacent columns.  The forward transform
// reads physical grids of the same precision; the inverse writes FP64 grids, FP32 grids or
// both.
template <typename S> struct Cx;
template <> struct Cx<double> { using T = double2; };
template <> struct Cx<float> { using T = float2; };
template <typename S> using cx_t = typename Cx<S>::T;
template <typename C> struct Re;
template <> struct Re<double2> { using T = double; };
template <> struct Re<float2> { using T = float; };
template <typename S>
__host__ __device__ constexpr int lds_cplx() { return 16384 / (int)sizeof(cx_t<S>); }  // 16 KiB

template <typename C, typename S>
__device__ __forceinline__ C mkc(S a, S b) {
  C r;
  r.x = a;
  r.y = b;
  return r;
}

struct FieldPtrs {
  const void *in[MAXF];  // forward: physical grids of type S; inverse: spectral arrays of type S
  void *out[MAXF];       // forward: spectral arrays of type S; inverse: FP64 physical grids
  float *out32[MAXF];    // optional FP32 copy of a real output (the FP32-bulk form's gathers)
  int op[MAXF];          // SpecOp applied to the input of the inverse (k_rows_inv)
  const double *kx, *ky;  // wavenumber planes for SPEC_DX / SPEC_DY
};

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// DFT of length R on registers: v[o] <- sum_q v[q] W^(o q), W = exp(-2 pi i / R) for the
// forward transform and its conjugate for the inverse (INV); products with +-1 and +-i are
// exact swaps and sign changes.  C is the complex type, S its real type.
template <int R, bool INV, typename C>
__device__ __forceinline__ void dft(C *v) {
  using S = typename Re<C>::T;
  if constexpr (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = mkc<C, S>(a.x + b.x, a.y + b.y);
    v[1] = mkc<C, S>(a.x - b.x, a.y - b.y);
  } else if constexpr (R == 4) {
    const C s02 = mkc<C, S>(v[0].x + v[2].x, v[0].y + v[2].y);
    const C d02 = mkc<C, S>(v[0].x - v[2].x, v[0].y - v[2].y);
    const C s13 = mkc<C, S>(v[1].x + v[3].x, v[1].y + v[3].y);
    const C d13 = mkc<C, S>(v[1].x - v[3].x, v[1].y - v[3].y);
    // forward: y1 = d02 - i d13, y3 = d02 + i d13 (inverse swaps them)
    const C mi = mkc<C, S>(d13.y, -d13.x);  // -i * d13
    v[0] = mkc<C, S>(s02.x + s13.x, s02.y + s13.y);
    v[2] = mkc<C, S>(s02.x - s13.x, s02.y - s13.y);
    if (!INV) {
      v[1] = mkc<C, S>(d02.x + mi.x, d02.y + mi.y);
      v[3] = mkc<C, S>(d02.x - mi.x, d02.y - mi.y);
    } else {
      v[1] = mkc<C, S>(d02.x - mi.x, d02.y - mi.y);
      v[3] = mkc<C, S>(d02.x + mi.x, d02.y + mi.y);
    }
  } else if constexpr (R == 3) {
    constexpr S c = (S)-0.5;
    constexpr S s = INV ? (S)-0.86602540378443864676 : (S)0.86602540378443864676;
    const C t = mkc<C, S>(v[1].x + v[2].x, v[1].y + v[2].y);
    const C d = mkc<C, S>(v[1].x - v[2].x, v[1].y - v[2].y);
    const C a = mkc<C, S>(v[0].x + c * t.x, v[0].y + c * t.y);
    v[0] = mkc<C, S>(v[0].x + t.x, v[0].y + t.y);
    // y1 = a - i s d, y2 = a + i s d;  -i (x + i y) = y - i x
    v[1] = mkc<C, S>(a.x + s * d.y, a.y - s * d.x);
    v[2] = mkc<C, S>(a.x - s * d.y, a.y + s * d.x);
  } else {
    static_assert(R == 5, "radix 2, 3, 4 or 5");
    constexpr S c1 = (S)0.30901699437494742410, c2 = (S)-0.80901699437494742410;
    constexpr S s1 = INV ? (S)-0.95105651629515357212 : (S)0.95105651629515357212;
    constexpr S s2 = INV ? (S)-0.58778525229247312917 : (S)0.58778525229247312917;
    const C t1 = mkc<C, S>(v[1].x + v[4].x, v[1].y + v[4].y);
    const C t2 = mkc<C, S>(v[2].x + v[3].x, v[2].y + v[3].y);
    const C t3 = mkc<C, S>(v[1].x - v[4].x, v[1].y - v[4].y);
    const C t4 = mkc<C, S>(v[2].x - v[3].x, v[2].y - v[3].y);
    const C a1 = mkc<C, S>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
    const C a2 = mkc<C, S>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
    const C b1 = mkc<C, S>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    const C b2 = mkc<C, S>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    v[0] = mkc<C, S>(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
    // y1 = a1 - i b1, y4 = a1 + i b1, y2 = a2 - i b2, y3 = a2 + i b2
    v[1] = mkc<C, S>(a1.x + b1.y, a1.y - b1.x);
    v[4] = mkc<C, S>(a1.x - b1.y, a1.y + b1.x);
    v[2] = mkc<C, S>(a2.x + b2.y, a2.y - b2.x);
    v[3] = mkc<C, S>(a2.x - b2.y, a2.y + b2.x);
  }
}

// One Stockham stage of radix R over nb transforms of length n (t-major in x), Ns = product of
// the radices already applied: item (t, i) takes x[i + q n/R], q < R, twiddled by
// W_n^(q k n/(Ns R)) with k = i mod Ns, and writes the R-point DFT to
// y[(i / Ns) Ns R + k + o Ns].  tw[e] = W_n^e staged in LDS (conjugated for the inverse).
template <int R, bool INV, typename C>
__device__ __forceinline__ void stockham_stage(const C *x, C *y, const C *tw, int n, int Ns,
                                               int nb) {
  const int nq = n / R;
  const int tstep = n / (Ns * R);
  const int tot = nb * nq;
  for (int bi = threadIdx.x; bi < tot; bi += FFT_TPB) {
    const int t = bi / nq, i = bi - t * nq;
    const int j = i / Ns, k = i - j * Ns;
    const C *xs = x + t * n + i;
    C v[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
      const C e = xs[q * nq];
      v[q] = (q == 0 || k == 0) ? e : cmul(e, tw[q * k * tstep]);
    }
    dft<R, INV>(v);
    C *ys = y + t * n + j * Ns * R + k;
#pragma unroll
    for (int o = 0; o < R; o++) ys[o * Ns] = v[o];
  }
}

// the stages from stage ST of a fixed length N (NS = product of the earlier radices)
template <bool INV, int N, int ST, int NS, typename C>
__device__ __forceinline__ C *fixed_stages(C *a, C *b, const C *tw, int nb) {
  constexpr Plan P = make_plan(N);
  static_assert(P.nrad > 0, "a fixed length has a radix schedule");
  if constexpr (ST == P.nrad) {
    return a;
  } else {
    constexpr int R = (int)((P.code >> (4 * ST)) & 15u);
    stockham_stage<R, INV>(a, b, tw, N, NS, nb);
    __syncthreads();
    return fixed_stages<INV, N, ST + 1, NS * R>(b, a, tw, nb);
  }
}

// nb transforms of length p.n (NT when NT > 0) stored t-major in `a`; result in the returned
// buffer.
template <bool INV, int NT, typename C>
__device__ C *lds_fft(C *a, C *b, const C *tw, const Plan &p, int nb) {
  if constexpr (NT > 0) return fixed_stages<INV, NT, 0, 1>(a, b, tw, nb);
  int Ns = 1;
  for (int s = 0; s < p.nrad; s++) {
    const int r = (int)((p.code >> (4 * s)) & 15u);
    if (r == 4) stockham_stage<4, INV>(a, b, tw, p.n, Ns, nb);
    else if (r == 2) stockham_stage<2, INV>(a, b, tw, p.n, Ns, nb);
    else if (r == 3) stockham_stage<3, INV>(a, b, tw, p.n, Ns, nb);
    else stockham_stage<5, INV>(a, b, tw, p.n, Ns, nb);
    __syncthreads();
    C *sw = a;
    a = b;
    b = sw;
    Ns *= r;
  }
  return a;
}

// Loads into LDS are issued as a batch (each thread's elements loaded into registers, then
// stored), so a workgroup waits on memory once per pass rather than once per element.
template <typename S>
__host__ __device__ constexpr int ept() { return lds_cplx<S>() / FFT_TPB; }  // per thread
// The row passes read contiguous rows, so a smaller batch costs them no coalescing: in FP32
// they stage half as many elements (8 KiB per buffer), so twice the workgroups fit a CU's LDS
template <typename S>
__host__ __device__ constexpr int lds_rows() {
  return sizeof(S) == 4 ? lds_cplx<S>() / 2 : lds_cplx<S>();
}
template <typename S>
__host__ __device__ constexpr int ept_rows() { return lds_rows<S>() / FFT_TPB; }
constexpr int TPT = 512 / FFT_TPB;  // twiddles per thread (n <= 512)

// The column passes of FP32 columns of at most 128 points stage half the batch (8 KiB per
// buffer), so twice the workgroups fit a CU; that still takes at least 8 adjacent columns, 64
// bytes of each row.  Longer columns keep the full batch (fewer columns would cut the rows' runs
// to 32 bytes: 256-point columns measured 160 -> 264 us per inverse pass), and the inverse pass
// of 512-point FP32 columns takes twice the batch (eight column pairs, 64-byte runs: 1.97 ->
// 1.64 ms at 512 x 512 x 257 nodes, with half the workgroups per CU; the forward pass, whose
// registers then hold twice the elements, measured slower that way).  n = 0: a length known
// only at run time; inv: the inverse (complex-to-real) pass.
template <typename S>
__host__ __device__ constexpr int lds_cols(int n, bool inv) {
  return (sizeof(S) == 4 && n > 0 && n <= 128)   ? lds_cplx<S>() / 2
         : (sizeof(S) == 4 && inv && n == 512)   ? lds_cplx<S>() * 2
                                                 : lds_cplx<S>();
}

// adjacent columns (column pairs for the inverse) a column pass takes per workgroup
template <typename S>
__host__ __device__ constexpr int col_batch(int n, bool inv) {
  return lds_cols<S>(n, inv) / n > 32 ? 32 : lds_cols<S>(n, inv) / n;
}

// q = a / d and r = a - q d for 0 <= a < 2^22, 0 < d, with rd = 1/d rounded to FP32: the
// product's relative error is below 2^-23, so the truncated estimate is off by at most one
__device__ __forceinline__ int div_small(int a, int d, float rd, int &r) {
  int q = (int)((float)a * rd);
  r = a - q * d;
  if (r < 0) {
    q--;
    r += d;
  } else if (r >= d) {
    q++;
    r -= d;
  }
  return q;
}

template <typename C>
__device__ __forceinline__ void load_twiddles(C (&w)[TPT], const C *tw, int n) {
#pragma unroll
  for (int i = 0; i < TPT; i++) {
    const int e = threadIdx.x + i * FFT_TPB;
    if (e < n) w[i] = tw[e];
  }
}
template <typename C>
__device__ __forceinline__ void put_twiddles(C *tws, const C (&w)[TPT], int n, bool conj_) {
#pragma unroll
  for (int i = 0; i < TPT; i++) {
    const int e = threadIdx.x + i * FFT_TPB;
    if (e < n) tws[e] = conj_ ? mkc<C>(w[i].x, -w[i].y) : w[i];
  }
}

// forward pass 1: r2c along x of every row (k, y) -> C1[k][y][l], l <= nx/2.  Two real rows
// a, b share one complex transform Z = FFT(a + i b): A[l] = (Z[l] + conj Z[n-l]) / 2 and
// B[l] = (Z[l] - conj Z[n-l]) / (2i).
template <typename S, int NT>
__global__ void __launch_bounds__(FFT_TPB) k_rows_r2c(Plan p, const cx_t<S> *__restrict__ tw,
                                                      FieldPtrs fp, cx_t<S> *__restrict__ c1,
                                                      int64_t c1_field, int nrows, int nb) {
  using C = cx_t<S>;
  constexpr int EPT = ept_rows<S>();
  __shared__ C A[lds_rows<S>()], B[lds_rows<S>()], T[512];
  const int f = blockIdx.y;
  const S *src = static_cast<const S *>(fp.in[f]);
  C *dst = c1 + (int64_t)f * c1_field;
  const int n = NT ? NT : p.n, nl = n / 2 + 1;
  const int r0 = blockIdx.x * 2 * nb;                  // first row of this workgroup
  const int nbb = min(nb, (nrows - r0 + 1) / 2);       // transforms (row pairs)
  C w[TPT], v[EPT];
  load_twiddles(w, tw, n);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int e = threadIdx.x + i * FFT_TPB;
    if (e < nbb * n) {
      const int t = e / n, x = e - t * n;
      const int ra = r0 + 2 * t, rb = ra + 1;
      v[i].x = src[(int64_t)ra * n + x];
      v[i].y = (rb < nrows) ? src[(int64_t)rb * n + x] : (S)0;
    }
  }
  put_twiddles(T, w, n, false);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int e = threadIdx.x + i * FFT_TPB;
    if (e < nbb * n) A[e] = v[i];
  }
  __syncthreads();
  C *R = lds_fft<false, NT>(A, B, T, p, nbb);
  for (int e = threadIdx.x; e < nbb * nl; e += FFT_TPB) {
    const int t = e / nl, l = e - t * nl;
    const int ra = r0 + 2 * t, rb = ra + 1;
    const C z = R[t * n + l];
    const C w = R[t * n + (l == 0 ? 0 : n - l)];  // Z[n-l]
    dst[(int64_t)ra * nl + l] = mkc<C, S>((z.x + w.x) / 2, (z.y - w.y) / 2);
    if (rb < nrows) dst[(int64_t)rb * nl + l] = mkc<C, S>((z.y + w.y) / 2, (w.x - z.x) / 2);
  }
}

// forward pass 2: complex FFT along y of nb adjacent columns l of one z level, then the
// (Re,Im)-pair split into the semi-spectral layout
template <typename S, int NT>
__global__ void __launch_bounds__(FFT_TPB) k_cols_fwd(Plan p, const cx_t<S> *__restrict__ tw,
                                                      FieldPtrs fp, const cx_t<S> *__restrict__ c1,
                                                      int64_t c1_field, int nl, int nb, int sx,
                                                      int sy) {
  using C = cx_t<S>;
  constexpr int EPT = lds_cols<S>(NT, false) / FFT_TPB;
  __shared__ C A[lds_cols<S>(NT, false)], B[lds_cols<S>(NT, false)], T[512];
  const int f = blockIdx.z;
  const C *src = c1 + (int64_t)f * c1_field;
  S *dst = static_cast<S *>(fp.out[f]);
  const int ny = NT ? NT : p.n;
  const int nbd = NT ? col_batch<S>(NT, false) : nb;  // the batch every workgroup but the last takes
  const int k = blockIdx.y;
  const int l0 = blockIdx.x * nb;
  const int nbb = min(nb, nl - l0);
  const C *lev = src + (int64_t)k * ny * nl;
  C w[TPT];
  S vx[EPT], vy[EPT];
  int at[EPT];
  load_twiddles(w, tw, ny);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int e = threadIdx.x + i * FFT_TPB;
    const int y = e / nbd, t = e - y * nbd;
    at[i] = (y < ny && t < nbb) ? t * ny + y : -1;
    vx[i] = 0;
    vy[i] = 0;
    if (at[i] >= 0) {
      const C u = lev[(int64_t)y * nl + l0 + t];
      vx[i] = u.x;
      vy[i] = u.y;
    }
  }
  put_twiddles(T, w, ny, false);
#pragma unroll
  for (int i = 0; i < EPT; i++)
    if (at[i] >= 0) A[at[i]] = mkc<C, S>(vx[i], vy[i]);
  __syncthreads();
  C *R = lds_fft<false, NT>(A, B, T, p, nbb);
  const int nm = ny / 2 + 1;
  S *out = dst + (int64_t)k * sy * sx;
  for (int e = threadIdx.x; e < nbd * nm; e += FFT_TPB) {
    const int m = e / nbd, t = e - m * nbd;
    if (t >= nbb) continue;
    const int l = l0 + t;
    C g = R[t * ny + m];
    C h = R[t * ny + ((ny - m) % ny)];
    C s = mkc<C, S>(g.x + h.x, g.y + h.y);
    C d = mkc<C, S>(g.x - h.x, g.y - h.y);
    out[(int64_t)(2 * m) * sx + 2 * l] = s.x / 2;
    out[(int64_t)(2 * m) * sx + 2 * l + 1] = s.y / 2;
    out[(int64_t)(2 * m + 1) * sx + 2 * l] = d.y / 2;
    out[(int64_t)(2 * m + 1) * sx + 2 * l + 1] = -d.x / 2;
  }
}

// inverse pass 1: per (k, m) build H = A_m + i B_m over l in [0, nx) from the two real
// x-spectra (imaginary parts of the DC and Nyquist bins dropped as c2r drops them), inverse
// FFT along x, scale 1/nx -> C2[k][m][x]
template <typename S, int NT>
__global__ void __launch_bounds__(FFT_TPB) k_rows_inv(Plan p, const cx_t<S> *__restrict__ tw,
                                                      FieldPtrs fp, cx_t<S> *__restrict__ c2,
                                                      int64_t c2_field, int nrows, int nm,
                                                      float rnm, int nb, int sx) {
  using C = cx_t<S>;
  constexpr int EPT = ept_rows<S>();
  __shared__ C A[lds_rows<S>()], B[lds_rows<S>()], T[512];
  const int f = blockIdx.y;
  const S *src = static_cast<const S *>(fp.in[f]);
  const int op = fp.op[f];
  C *dst = c2 + (int64_t)f * c2_field;
  const int nx = NT ? NT : p.n;
  const int r0 = blockIdx.x * nb;
  const int nbb = min(nb, nrows - r0);
  const int half = nx / 2;
  const bool even = (nx % 2) == 0;
  C w[TPT], va[EPT], vb[EPT];
  load_twiddles(w, tw, nx);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int e = threadIdx.x + i * FFT_TPB;
    if (e < nbb * nx) {
      const int t = e / nx, l = e - t * nx;
      const int row = r0 + t;  // row = k * nm + m
      int m;
      const int k = div_small(row, nm, rnm, m);
      const S *sa = src + ((int64_t)k * (2 * nm) + 2 * m) * sx;
      const int ll = (l <= half) ? l : nx - l;
      va[i] = *reinterpret_cast<const C *>(sa + 2 * ll);       // (ar, ai)
      vb[i] = *reinterpret_cast<const C *>(sa + sx + 2 * ll);  // (br, bi)
      if (op != SPEC_PLAIN) {
        // the derivative of spectral rows 2m (a) and 2m+1 (b), element for element as the
        // standalone diffx / diffy compute it (fft_ops.F90:266-269, :366-369); the
        // wavenumber planes are indexed by (row, x) like the spectral plane
        const double *kp = (op == SPEC_DX ? fp.kx : fp.ky) + (int64_t)(2 * m) * sx + 2 * ll;
        const double2 kad = *reinterpret_cast<const double2 *>(kp);
        const double2 kbd = *reinterpret_cast<const double2 *>(kp + sx);
        const S kax = (S)kad.x, kay = (S)kad.y, kbx = (S)kbd.x, kby = (S)kbd.y;
        const C a0 = va[i], b0 = vb[i];
        if (op == SPEC_DX) {  // pair swap along x within each row
          va[i] = mkc<C, S>(((S)-1. * a0.y) * kax, a0.x * kay);
          vb[i] = mkc<C, S>(((S)-1. * b0.y) * kbx, b0.x * kby);
        } else {  // pair swap along y between the two rows
          va[i] = mkc<C, S>(((S)-1. * b0.x) * kax, ((S)-1. * b0.y) * kay);
          vb[i] = mkc<C, S>(a0.x * kbx, a0.y * kby);
        }
      }
    }
  }
  put_twiddles(T, w, nx, true);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int e = threadIdx.x + i * FFT_TPB;
    if (e < nbb * nx) {
      const int t = e / nx, l = e - t * nx;
      const int ll = (l <= half) ? l : nx - l;
      const S ar = va[i].x, br = vb[i].x;
      S ai = va[i].y, bi = vb[i].y;
      if (ll == 0 || (even && ll == half)) { ai = 0; bi = 0; }
      C h;
      if (l <= half) h = mkc<C, S>(ar - bi, ai + br);
      else h = mkc<C, S>(ar + bi, -ai + br);
      A[t * nx + l] = h;
    }
  }
  __syncthreads();
  C *R = lds_fft<true, NT>(A, B, T, p, nbb);
  const S dn = (S)1 / (S)nx;
  for (int e = threadIdx.x; e < nbb * nx; e += FFT_TPB) {
    int t = e / nx, x = e - t * nx;
    C v = R[t * nx + x];
    dst[(int64_t)(r0 + t) * nx + x] = mkc<C, S>(v.x * dn, v.y * dn);
  }
}

// inverse pass 2: c2r along y for 2 nb adjacent x columns of one z level.  The Hermitian
// spectra of two columns a, b share one complex inverse transform of H_a + i H_b, whose real
// and imaginary parts are the two real columns.
template <typename S, int NT>
__global__ void __launch_bounds__(FFT_TPB) k_cols_c2r(Plan p, const cx_t<S> *__restrict__ tw,
                                                      FieldPtrs fp, const cx_t<S> *__restrict__ c2,
                                                      int64_t c2_field, int nx, int nb) {
  using C = cx_t<S>;
  constexpr int EPT = lds_cols<S>(NT, true) / FFT_TPB;
  __shared__ C A[lds_cols<S>(NT, true)], B[lds_cols<S>(NT, true)], T[512];
  const int f = blockIdx.z;
  const C *src = c2 + (int64_t)f * c2_field;
  double *dst = static_cast<double *>(fp.out[f]);
  const int ny = NT ? NT : p.n, nm = ny / 2 + 1;
  const int nbd = NT ? col_batch<S>(NT, true) : nb;  // the batch every workgroup but the last takes
  const int k = blockIdx.y;
  const int x0 = blockIdx.x * 2 * nb;
  const int nbb = min(nb, (nx - x0 + 1) / 2);
  const C *lev = src + (int64_t)k * nm * nx;
  const bool even = (ny % 2) == 0;
  C w[TPT], va[EPT], vb[EPT];
  load_twiddles(w, tw, ny);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int e = threadIdx.x + i * FFT_TPB;
    const int y = e / nbd, t = e - y * nbd;
    if (y < ny && t < nbb) {
      const int xa = x0 + 2 * t, xb = xa + 1;
      const int m = (y < nm) ? y : ny - y;
      va[i] = lev[(int64_t)m * nx + xa];
      vb[i] = (xb < nx) ? lev[(int64_t)m * nx + xb] : mkc<C, S>(0, 0);
    }
  }
  put_twiddles(T, w, ny, true);
#pragma unroll
  for (int i = 0; i < EPT; i++) {
    const int e = threadIdx.x + i * FFT_TPB;
    const int y = e / nbd, t = e - y * nbd;
    if (y < ny && t < nbb) {
      const int m = (y < nm) ? y : ny - y;
      C wa = va[i], wb = vb[i];
      if (m == 0 || (even && m == ny / 2)) { wa.y = 0; wb.y = 0; }
      if (y >= nm) { wa.y = -wa.y; wb.y = -wb.y; }
      A[t * ny + y] = mkc<C, S>(wa.x - wb.y, wa.y + wb.x);
    }
  }
  __syncthreads();
  C *R = lds_fft<true, NT>(A, B, T, p, nbb);
  const S dn = (S)1 / (S)ny;
  // either output may be absent: the FP32-bulk form keeps some fields only in FP32
  double *out = dst ? dst + (int64_t)k * ny * nx : nullptr;
  float *o32 = fp.out32[f] ? fp.out32[f] + (int64_t)k * ny * nx : nullptr;
  for (int e = threadIdx.x; e < nbd * ny; e += FFT_TPB) {
    const int y = e / nbd, t = e - y * nbd;
    if (t >= nbb) continue;
    const int xa = x0 + 2 * t, xb = xa + 1;
    const C v = R[t * ny + y];
    if (out) {
      out[(int64_t)y * nx + xa] = v.x * dn;
      if (xb < nx) out[(int64_t)y * nx + xb] = v.y * dn;
    }
    if (o32) {
      o32[(int64_t)y * nx + xa] = (float)(v.x * dn);
      if (xb < nx) o32[(int64_t)y * nx + xb] = (float)(v.y * dn);
    }
  }
}

// W_n^e for e in [0, n), in FP64 and rounded to FP32
int upload_twiddles(int n, double2 **dptr, float2 **dptr32) {
  std::vector<double2> h(n);
  std::vector<float2> h32(n);
  for (int e = 0; e < n; e++) {
    // quarter turns are exact so that purely real or imaginary inputs stay so
    if ((4L * e) % n == 0) {
      const int q = (int)((4L * e) / n);
      const double re[4] = {1.0, 0.0, -1.0, 0.0}, im[4] = {0.0, -1.0, 0.0, 1.0};
      h[e] = make_double2(re[q], im[q]);
    } else {
      long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)e / (long double)n;
      h[e] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
    h32[e] = make_float2((float)h[e].x, (float)h[e].y);
  }
  PMPIC_CHECK(hipMalloc(dptr, sizeof(double2) * n));
  PMPIC_CHECK(hipMemcpy(*dptr, h.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
  PMPIC_CHECK(hipMalloc(dptr32, sizeof(float2) * n));
  PMPIC_CHECK(hipMemcpy(*dptr32, h32.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  return PMPIC_OK;
}

template <typename S> const cx_t<S> *twiddles_x(const Ctx *c);
template <> const double2 *twiddles_x<double>(const Ctx *c) { return c->tw_x; }
template <> const float2 *twiddles_x<float>(const Ctx *c) { return c->tw_x32; }
template <typename S> const cx_t<S> *twiddles_y(const Ctx *c);
template <> const double2 *twiddles_y<double>(const Ctx *c) { return c->tw_y; }
template <> const float2 *twiddles_y<float>(const Ctx *c) { return c->tw_y32; }

}  // namespace

int fft_setup(Ctx *c) {
  if (c->g.nx > 512 || c->g.ny > 512) return PMPIC_ERR_FFT_SIZE;
  if (make_plan(c->g.nx).nrad < 0 || make_plan(c->g.ny).nrad < 0) return PMPIC_ERR_FFT_SIZE;
  PMPIC_TRY(upload_twiddles(c->g.nx, &c->tw_x, &c->tw_x32));
  PMPIC_TRY(upload_twiddles(c->g.ny, &c->tw_y, &c->tw_y32));
  return PMPIC_OK;
}

static int64_t c1_field_size(const Geometry &g) { return (int64_t)g.nzn * g.ny * (g.nx / 2 + 1); }
static int64_t c2_field_size(const Geometry &g) { return (int64_t)g.nzn * (g.ny / 2 + 1) * g.nx; }

// f(std::integral_constant<int, N>) with N = n for a length in FIXED_N, else N = 0
template <class F>
int with_length(int n, F &&f) {
  switch (n) {
    case 16: return f(std::integral_constant<int, 16>{});
    case 32: return f(std::integral_constant<int, 32>{});
    case 48: return f(std::integral_constant<int, 48>{});
    case 64: return f(std::integral_constant<int, 64>{});
    case 96: return f(std::integral_constant<int, 96>{});
    case 128: return f(std::integral_constant<int, 128>{});
    case 256: return f(std::integral_constant<int, 256>{});
    case 512: return f(std::integral_constant<int, 512>{});
    default: return f(std::integral_constant<int, 0>{});
  }
}
static_assert(sizeof(FIXED_N) / sizeof(int) == 8, "with_length lists every fixed length");

template <typename S>
int fft_forward_t(Ctx *c, int nf, const S *const *src, S *const *dst) {
  const Geometry &g = c->g;
  if (nf < 1 || nf > MAXF) return PMPIC_ERR_ARG;
  FieldPtrs fp{};
  for (int f = 0; f < nf; f++) { fp.in[f] = src[f]; fp.out[f] = dst[f]; }
  const Plan px = make_plan(g.nx), py = make_plan(g.ny);
  const int64_t c1f = c1_field_size(g);
  const int nrows = g.nzn * g.ny;
  cx_t<S> *cb = reinterpret_cast<cx_t<S> *>(c->cbuf);
  const int nbx = lds_rows<S>() / g.nx;  // transforms per workgroup, two rows each
  with_length(g.nx, [&](auto N) {
    hipLaunchKernelGGL((k_rows_r2c<S, decltype(N)::value>), dim3(cdiv(nrows, 2 * nbx), nf),
                       dim3(FFT_TPB), 0, c->stream, px, twiddles_x<S>(c), fp, cb, c1f, nrows, nbx);
    return 0;
  });
  const int nl = g.nx / 2 + 1;
  const int nby = col_batch<S>(g.ny, false);
  with_length(g.ny, [&](auto N) {
    hipLaunchKernelGGL((k_cols_fwd<S, decltype(N)::value>), dim3(cdiv(nl, nby), g.nzn, nf),
                       dim3(FFT_TPB), 0, c->stream, py, twiddles_y<S>(c), fp, cb, c1f, nl, nby,
                       g.sx, g.sy);
    return 0;
  });
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

template <typename S>
int fft_backward_t(Ctx *c, int nf, const S *const *src, double *const *dst, float *const *dst32,
                   const int *deriv) {
  const Geometry &g = c->g;
  if (nf < 1 || nf > MAXF) return PMPIC_ERR_ARG;
  FieldPtrs fp{};
  fp.kx = c->kx;
  fp.ky = c->ky;
  for (int f = 0; f < nf; f++) {
    fp.in[f] = src[f];
    fp.out[f] = dst[f];
    fp.out32[f] = dst32 ? dst32[f] : nullptr;
    fp.op[f] = deriv ? deriv[f] : SPEC_PLAIN;
    if (fp.op[f] < SPEC_PLAIN || fp.op[f] > SPEC_DY) return PMPIC_ERR_ARG;
  }
  const Plan px = make_plan(g.nx), py = make_plan(g.ny);
  const int64_t c2f = c2_field_size(g);
  const int nm = g.ny / 2 + 1;
  const int nrows = g.nzn * nm;
  if (nrows >= (1 << 22)) return PMPIC_ERR_ARG;  // the row split's reciprocal division
  const float rnm = (float)(1.0 / nm);
  cx_t<S> *cb = reinterpret_cast<cx_t<S> *>(c->cbuf);
  const int nbx = lds_rows<S>() / g.nx;
  with_length(g.nx, [&](auto N) {
    hipLaunchKernelGGL((k_rows_inv<S, decltype(N)::value>), dim3(cdiv(nrows, nbx), nf),
                       dim3(FFT_TPB), 0, c->stream, px, twiddles_x<S>(c), fp, cb, c2f, nrows, nm,
                       rnm, nbx, g.sx);
    return 0;
  });
  const int nby = col_batch<S>(g.ny, true);  // transforms per workgroup, two x columns each
  with_length(g.ny, [&](auto N) {
    hipLaunchKernelGGL((k_cols_c2r<S, decltype(N)::value>), dim3(cdiv(g.nx, 2 * nby), g.nzn, nf),
                       dim3(FFT_TPB), 0, c->stream, py, twiddles_y<S>(c), fp, cb, c2f, g.nx, nby);
    return 0;
  });
  PMPIC_CHECK(hipGetLastError());
  return PMPIC_OK;
}

template int fft_forward_t<double>(Ctx *, int, const double *const *, double *const *);
template int fft_forward_t<float>(Ctx *, int, const float *const *, float *const *);
template int fft_backward_t<double>(Ctx *, int, const double *const *, double *const *,
                                    float *const *, const int *);
template int fft_backward_t<float>(Ctx *, int, const float *const *, double *const *,
                                   float *const *, const int *);

int fft_forward(Ctx *c, int nf, const double *const *src, double *const *dst) {
  return fft_forward_t<double>(c, nf, src, dst);
}

int fft_backward(Ctx *c, int nf, const double *const *src, double *const *dst,
                 float *const *dst32, const int *deriv) {
  return fft_backward_t<double>(c, nf, src, dst, dst32, deriv);
}

}  // namespace pmpic
