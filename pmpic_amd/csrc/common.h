// Shared definitions of the MI355X PMPIC hot-path library (libpmpic_hip.so).
//
// Device data layout (all device-resident between calls; see the C-ABI in include/pmpic_hip.h):
//  * parcels: structure of arrays, FP64, capacity `cap`, field f at parcels + f*cap, fields in
//    writebinary order x..tag then the 12 RK arrays (model_core/src/parcels/parcels.F90:19-34);
//  * physical grids: interior nodes only, [nzn][ny][nx] (x fastest, z slowest); the reference's
//    x/y halos (decomposition.F90:449-451) become periodic index arithmetic, so the halo fold of
//    par2grid (MPIC_Haloswap.F90:188-290) and the fill of grid2par (:95-184) are implicit;
//  * spectral arrays: [nzn][sy][sx] with sy = (ny/2+1)*2, sx = (nx/2+1)*2 holding the
//    reference's (Re,Im)-interleaved semi-spectral layout (pencilfft.F90:838-899), z slowest so
//    that one thread per (kx,ky) column walks z with unit-stride, coalesced across threads;
//  * cells: c = (k*ny + j)*nx + i over ncz*ny*nx cells.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/pmpic_status.h"

namespace pmpic {

enum Field {
  FX = 0, FY, FZ, FP, FQ, FR, FDXDT, FDYDT, FDZDT, FDPDT, FDQDT, FDRDT, FH, FB, FVOL,
  FSTRETCH, FTAG, FXO, FYO, FZO, FXF, FYF, FZF, FPO, FQO, FRO, FPF, FQF, FRF, NFIELDS
};

// FP32-bulk form: per-parcel FP32 fields (the position is an int32 cell index plus the three
// offsets B_OX..B_OZ in [0,1) inside that cell; B_OXO..B_OZO and the cell_o array hold the RK4
// origin, B_DFX..B_DFZ the accumulated RK4 increment in cells relative to it).  Fields 3..15
// share their numbers with Field.
enum Field32 {
  B_OX = 0, B_OY, B_OZ, B_P, B_Q, B_R, B_DXDT, B_DYDT, B_DZDT, B_DPDT, B_DQDT, B_DRDT, B_H, B_B,
  B_VOL, B_STRETCH, B_OXO, B_OYO, B_OZO, B_DFX, B_DFY, B_DFZ, B_PO, B_QO, B_RO, B_PF, B_QF, B_RF,
  NF32
};

struct Bulk32 {
  float *f;          // NF32 * cap
  float *fp[NF32];   // fp[k] = f + k * cap: field k (kernels take these from their arguments
                     // instead of forming them)
  int32_t *cell;     // cap: cell holding the parcel
  int32_t *cello;    // cap: cell of the RK4 origin
  double *tag;       // cap
  float *grid[6];    // gathered grids U, V, W, DP, DQ, DR [nzn][ny][nx]
  float *mgrid[12];  // mixing: kept vol, b, h, p, q, r (0..5), residuals (6..11)
  uint8_t *mtouch;   // mixing, per node: a residual grid is not zero there
  // cell binning: slot[c * capc + rank] lists the parcels of cell c while rank < capc; a cell
  // with more parcels lists all of them in the dense overflow region.  The table holds 32
  // entries per cell (none, and capc = 0, when its byte offsets would exceed 32 bits); each
  // binning uses capc = 16 or 32 of them by the mean parcels per cell.
  int32_t *slot;
  int2 *ovf;         // cap: (parcel, rank) of the parcels past the slot capacity
  int capc;          // capacity of the current binning (0, 16 or 32)
  int capmax;        // entries per cell allocated (0 or 32)
  int capfix;        // capacity forced by PMPIC_SLOT_CAPACITY (16 or 32), or 0 for the mean rule
  bool captune;      // PMPIC_SLOT_CAPACITY=tune and the first par2grid has not timed both yet
  int captuned;      // the capacity that timing picked (16 or 32), or 0 while none has
  int rk_tpb;        // workgroup size of the RK stage launches (128 or 256)
  // kept grid k (vol, b, hg, p, q, r) of the last mixing is current in mgrid[k] and not yet
  // widened into the FP64 grid a grid export reads
  bool kept_pending[6];
  // the step's p, q, r and total-buoyancy grids in FP32 (par2grid's node sums, then vort2vel's
  // filtered p, q, r), read by the spectral passes; pgrid[k] is current and not yet widened
  // into FP64 grid GP, GQ, GR, GB (k = 0..3) while p2g_pending[k]
  float *pgrid[4];
  bool p2g_pending[4];
  // RK stages 1-3 bin the cells they write for the next par2grid (fused_bin: grids of 2^21
  // cells or more, or as PMPIC_FUSED_BINNING says); prebinned: the cell lists of the current
  // cells are built
  bool fused_bin;
  bool prebinned;
  // x-row grids: the last mixing's kept sums are the next par2grid's partials except on the
  // tiles mixing changed (dirty: one byte per 16-cell tile); p2g_reuse while that holds (set by
  // mixing, cleared by the par2grid that uses it and by anything else that changes the
  // parcels or the partials); reuse_off: PMPIC_P2G_REUSE=0
  uint8_t *dirty;
  int64_t dirty_seq;  // Ctx::sync_seq when the last mixing left its flags: the host knows
                      // whether any tile changed once a later read-back has happened
  bool p2g_reuse;
  bool reuse_off;
};

// grid ids exported through the C-ABI (writebinary.F90:287-296 order, then the tendencies)
enum GridId { GU = 0, GV, GW, GP, GQ, GR, GB, GHG, GHGLIQ, GVOL, GDP, GDQ, GDR, NGRIDS };

struct Geometry {
  int nx, ny, ncz, nzn;          // cells x, y, z; z nodes
  int sx, sy;                    // spectral extents
  double xbot, ybot, zbot, xtop, ytop, ztop;  // config bounds
  double dx, dy, dz;
  double xmin_h, ymin_h, zmin_h, zmax_h;      // parcel_interpolation.F90:104-115
  double hs_xmin, hs_xmax, hs_ymin, hs_ymax;  // haloswap.F90:314-318
  double xshift_e, xshift_w, yshift_n, yshift_s;  // haloswap.F90:529-552
  double vmin;                   // parcel_mixing.F90:54
  double lx_cfl, ly_cfl, lz_cfl; // vort2vel.F90:346-348
};

// Per-launch scalars living in device memory (no host round trip inside a stage).
struct DevScalars {
  double dtm;        // current timestep
  double dtmax;      // configured dtmax
  double prefactor;  // timestep_prefactor
  double umax2, vmax2, wmax2, omax2;  // reductions (stage 1)
  int64_t n;         // parcel count
  int64_t cross;     // parcels crossing x/y this haloswap
  int64_t nsplit;    // parcels split this step
  int64_t nremove, nadd, nzero, nkeep1, nbackfill;
  int64_t anydirty;  // the last x-row mixing changed some tile (Bulk32::dirty)
  int status;        // first error code raised by a kernel
  int32_t novf;      // parcels past their cell's slot capacity in the last FP32 binning
};

enum TimerId { T_ACCUM = 0, T_RK = 1, T_VORT2VEL, T_VORT_TEND, T_RK4, T_SPLIT, T_MIX, NTIMERS };

struct Ctx {
  Geometry g;
  int64_t cap;
  int64_t n;             // host copy of the parcel count (valid between calls)
  double time, dtm, dtmax, prefactor;
  int it_v2v, it_vt, rkstep, it_split, it_mix;  // per-component iteration counters
  int precision;
  int device;
  int defer_status;      // set inside pmpic_step: kernel status checked once per step
  int split_flags_fresh; // flag[] holds the split test of the last stage-4 RK update
  int splitting_enabled, mixing_enabled;
  int64_t last_split, last_removed, last_added;
  int64_t sync_seq;      // device scalars read back so far (sync_scalars)
  hipStream_t stream;
  Bulk32 b32;            // FP32-bulk form state (precision == 1)
  // parcels
  double *par;           // NFIELDS * cap
  double *tmp;           // NFIELDS * cap staging for permutations (allocated lazily)
  int64_t tmp_cap;
  int32_t *cell, *rank, *sorted;  // cap
  uint8_t *flag;                  // cap (haloswap direction / keep code)
  int32_t *list_a, *list_b;       // cap compaction outputs
  // cells
  int64_t ncell;
  int32_t *cell_count, *cell_start;   // ncell + 1
  double *partial;                    // corner partial sums: 48 * ncell doubles (FP32 form: 24)
  double *partial2;                   // mixing residual partials, same size; may be null in the
                                      // FP64 form (then mixing takes two passes over partial)
  int32_t *npercell;                  // ncell
  uint8_t *cellflag;                  // ncell
  // grids (interior [nzn][ny][nx])
  int64_t gsz, ssz, csz;
  double *grid[NGRIDS];
  double *gtmp[8];                    // vorticity tendency derivative grids
  double *kres[6];                    // mixing residual grids
  // spectral [nzn][sy][sx]
  double *sa, *sb, *sc, *sd, *se, *sf, *su, *sv, *sw, *sbs;
  double *sder[8];
  double *atop, *abot, *btop, *bbot;  // [sy][sx]
  double *ubar, *vbar;                // [nzn]
  double *kx, *ky, *k2, *filt;        // [sy][sx]
  // Thomas elimination: diffz's denominators and c' [bt][k] (the same for every column)
  double *dz_den, *dz_cp;
  double *coords;                     // x_coords (nx+4), y_coords (ny+4), z_coords (nzn)
  double *fbuf;                       // host-array staging, allocated at first use (ensure_fbuf)
  double2 *cbuf;                      // complex FFT intermediates, 8 fields
  double2 *tw_x, *tw_y;               // twiddles W_n^e, e in [0,n)
  float2 *tw_x32, *tw_y32;            // the same rounded to FP32
  // scan scratch
  void *scan_tmp;
  size_t scan_tmp_bytes;
  int32_t *blk;                       // per-block counters (10 * nblk_max + 16)
  double *maxblk;                     // per-block maxima of the timestep reduction (4 * MAXBLK)
  int64_t nblk_max;
  DevScalars *dsc;                    // device scalars
  DevScalars hsc;                     // host mirror
  void *pin;                          // pinned host staging for small read-backs (256 B)
  hipEvent_t sync_ev;                 // read-back completion (spin wait)
  // FP32-bulk form with PMPIC_GRAPH=1: RK stages 1-3 of a step (vort2vel, vorticity_tendency,
  // rk4_integrator; the splitting and mixing callbacks do nothing there) captured once as a HIP
  // graph and replayed by later steps; the stage-4 work, whose launches depend on data read
  // back, stays eager
  hipGraph_t g123;
  hipGraphExec_t g123x;
  int g123_capc;        // slot capacity the captured binning used
  int64_t g123_n;       // parcel count the captured launches cover (their grids' size)
  int64_t graph_n;      // while capturing: g123_n, the size of the per-parcel launches; else 0
  int graph_on;         // PMPIC_GRAPH=1 in the environment at creation
  int64_t eager_steps;  // steps run without the graph (the first sets up the scan storage)
  int64_t g123_captures, g123_replays;  // graphs captured and launched
  bool g123_prebinned;                  // the binning state the captured stages leave
  // HIP-event timing: 0 = par2grid cell accumulation kernel, 1 = gather + RK stage kernel,
  // 2..6 = the component callbacks (timer regions named as the reference's timers)
  int timing;
  hipEvent_t ev[NTIMERS][64][2];
  int nev[NTIMERS];
  double t_ms[NTIMERS];
  int64_t t_calls[NTIMERS];
  double t_max[NTIMERS];  // longest single interval (ms)
  double t_sq[NTIMERS];   // sum of squared intervals (ms^2), for the standard deviation
};

#define PMPIC_CHECK(call)                                   \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return PMPIC_ERR_HIP;             \
  } while (0)

#define PMPIC_TRY(expr)                                     \
  do {                                                      \
    int s_ = (expr);                                        \
    if (s_ != PMPIC_OK) return s_;                          \
  } while (0)

static inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// XCD-aware block order.  MI355X dispatches consecutive workgroups round-robin over its 8
// XCDs, each with its own L2.  This maps the workgroup to the piece of work it takes so that
// every XCD gets one contiguous run of pieces: neighbouring pieces, which share cache lines
// (the ends of parcel runs, gathered grid nodes, partial-sum lines), meet in the same L2.  A
// bijection of [0, nb).
#ifndef PMPIC_XCD_SWIZZLE
#define PMPIC_XCD_SWIZZLE 1
#endif
constexpr uint32_t NXCD = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
#if PMPIC_XCD_SWIZZLE
  const uint32_t per = nb / NXCD, rem = nb % NXCD;
  const uint32_t x = b % NXCD, i = b / NXCD;
  return x * per + (x < rem ? x : rem) + i;
#else
  (void)nb;
  return b;
#endif
}

__device__ __forceinline__ void raise_status(DevScalars *s, int code) {
  atomicCAS(&s->status, 0, code);
}

// cache_parcel_interp_weights for one parcel (parcel_interpolation.F90:206-246), returning the
// periodic interior cell (i0, j0, k0) and the offsets.  Arithmetic is the reference's, operand
// for operand.  Returns a status code.
__device__ __forceinline__ int locate(const Geometry &g, double xp, double yp, double zp,
                                      int &i0, int &j0, int &k0, double &delx, double &dely,
                                      double &delz) {
  const double ax = xp - g.xmin_h, ay = yp - g.ymin_h, az = zp - g.zmin_h;
  const int i = (int)floor(ax / g.dx) + 1;
  const int j = (int)floor(ay / g.dy) + 1;
  const int k = (int)floor(az / g.dz) + 1;
  // the reference accepts 1 <= i < nx+4; halo columns 1,2 only arise for a parcel the
  // haloswap failed to wrap, which is rejected here instead of silently mis-deposited
  if (i < 3 || i > g.nx + 3) return PMPIC_ERR_OUT_OF_BOX_X;
  if (j < 3 || j > g.ny + 3) return PMPIC_ERR_OUT_OF_BOX_Y;
  if (k < 1 || k >= g.nzn) return PMPIC_ERR_OUT_OF_BOX_Z;
  delx = (ax - (double)(i - 1) * g.dx) / g.dx;
  dely = (ay - (double)(j - 1) * g.dy) / g.dy;
  delz = (az - (double)(k - 1) * g.dz) / g.dz;
  if (delx > 1. || delx < 0) return PMPIC_ERR_DELTA;
  if (dely > 1. || dely < 0) return PMPIC_ERR_DELTA;
  if (delz > 1. || delz < 0) return PMPIC_ERR_DELTA;
  i0 = i - 3;
  if (i0 >= g.nx) i0 -= g.nx;
  j0 = j - 3;
  if (j0 >= g.ny) j0 -= g.ny;
  k0 = k - 1;
  return PMPIC_OK;
}

// the 8 trilinear corner weights times volume, in the reference's order and association
// (parcel_interpolation.F90:390-434): corner bits z<<2 | y<<1 | x
__device__ __forceinline__ void corner_weights(double delx, double dely, double delz, double v,
                                               double w[8]) {
  w[0] = (1 - delz) * (1 - dely) * (1 - delx) * v;
  w[1] = (1 - delz) * (1 - dely) * (delx) * v;
  w[2] = (1 - delz) * (dely) * (1 - delx) * v;
  w[3] = (1 - delz) * (dely) * (delx) * v;
  w[4] = (delz) * (1 - dely) * (1 - delx) * v;
  w[5] = (delz) * (1 - dely) * (delx) * v;
  w[6] = (delz) * (dely) * (1 - delx) * v;
  w[7] = (delz) * (dely) * (delx) * v;
}

// streaming access for parcel fields a kernel touches once: the caches do not retain them, so
// the grids being gathered stay resident
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v) { __builtin_nontemporal_store(v, p); }

__device__ __forceinline__ int64_t gidx(const Geometry &g, int k, int j, int i) {
  return ((int64_t)k * g.ny + j) * g.nx + i;
}

// Buffers of one FP64 binning + single-field par2grid: the context's own, or caller-provided
// memory (pmpic_par2grid_noalloc).  cell, rank, sorted: n entries; count, start: ncell + 1;
// partial: 16 * ncell doubles; scan_tmp (scan_bytes) the exclusive scan's temporary storage,
// or null to use the context's (grown on demand).
struct BinWork {
  int32_t *cell, *rank, *count, *start, *sorted;
  double *partial;
  void *scan_tmp;
  size_t scan_bytes;
};

// ---- launchers implemented in the translation units ----
// interp.hip
int bin_parcels(Ctx *c, int64_t n);
int bin_parcels_ws(Ctx *c, int64_t n, const BinWork &w);
int par2grid_single_ws(Ctx *c, int64_t n, const double *d_var, double *d_grid, const BinWork &w);
size_t scan_temp_bytes(Ctx *c, int64_t n);  // temporary storage of an exclusive scan of n
int to_ref_layout(Ctx *c, const double *d_interior, double *d_ref);
int par2grid_fused(Ctx *c, int64_t n);  // P,Q,R,B grids from p,q,r,btot
int par2grid_single(Ctx *c, int64_t n, const double *d_var, double *d_grid);
int grid2par_single(Ctx *c, int64_t n, const double *d_grid, double *d_var);
int mixing_grids(Ctx *c, int64_t n);
int dump_grids_f64(Ctx *c, int64_t n);  // writebinary's b, hg, p, q, r grids
int hgliq_grid(Ctx *c);                 // hgliq and b += 12.5 hgliq (writebinary.F90:224-233)
// fft.hip
int fft_setup(Ctx *c);
int fft_forward(Ctx *c, int nf, const double *const *src, double *const *dst);
// dst32 (optional, entries may be NULL): an FP32 copy of each real output.
// deriv (optional): per field, SPEC_PLAIN, or SPEC_DX / SPEC_DY to transform diffx / diffy of
// the input (applied while the input is read, so the derivative is never stored)
enum SpecOp { SPEC_PLAIN = 0, SPEC_DX = 1, SPEC_DY = 2 };
int fft_backward(Ctx *c, int nf, const double *const *src, double *const *dst,
                 float *const *dst32 = nullptr, const int *deriv = nullptr);
// the same transforms with spectral arrays (and the transforms' arithmetic) of type S; the
// forward transform reads physical grids of type S, the inverse writes dst (FP64) and dst32
// (FP32), either of which may be NULL
template <typename S>
int fft_forward_t(Ctx *c, int nf, const S *const *src, S *const *dst);
template <typename S>
int fft_backward_t(Ctx *c, int nf, const S *const *src, double *const *dst, float *const *dst32,
                   const int *deriv);
// spectral.hip
int spectral_setup(Ctx *c);
int vort2vel_spectral(Ctx *c, int stage);
int vorticity_tendency_spectral(Ctx *c, int stage);
int op_diffx(Ctx *c, const double *in, double *out);
int op_diffy(Ctx *c, const double *in, double *out);
int op_filter(Ctx *c, const double *in, double *out);
int op_diffz(Ctx *c, const double *in, double *out, const double *bot, const double *top);
int op_laplinv(Ctx *c, double *f, int df0);
int grid_max_reduce(Ctx *c);
// parcels.hip
int rk_stage_update(Ctx *c, int stage, int64_t n, int stretch);
// Component timers (timing mode 1) bracket a component with events on the stream.
int timing_begin(Ctx *c, int which);
int timing_end(Ctx *c, int which);
int timing_harvest(Ctx *c);
// device-to-host read of a few bytes (<= 256) in stream order, through pinned staging
int read_back(Ctx *c, void *host, const void *dev, size_t bytes);

// A hot kernel (timer T_ACCUM or T_RK, timing modes 1 and 2): when timing is on, the start
// and stop timestamps are taken by the dispatch itself (hipExtLaunchKernelGGL), so the timer
// puts no packet between kernels on the stream.
template <typename... KArgs, typename... Args>
int launch_timed(Ctx *c, int which, void (*kernel)(KArgs...), dim3 grid, dim3 block,
                 uint32_t shmem, Args... args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  if (!c->timing) {
    hipLaunchKernelGGL(kernel, grid, block, shmem, c->stream, static_cast<KArgs>(args)...);
    return PMPIC_OK;
  }
  if (c->nev[which] >= 64) {
    const int st = timing_harvest(c);
    if (st != PMPIC_OK) return st;
  }
  hipExtLaunchKernelGGL(kernel, grid, block, shmem, c->stream, c->ev[which][c->nev[which]][0],
                        c->ev[which][c->nev[which]][1], 0u, static_cast<KArgs>(args)...);
  c->nev[which]++;
  return PMPIC_OK;
}
int parcel_haloswap(Ctx *c, int64_t n, int reflect);
int parcel_split(Ctx *c);
int parcel_mix(Ctx *c);
int compact_flags(Ctx *c, const uint8_t *flags, uint8_t want, int64_t n, int32_t *out,
                  int64_t *d_count);
// bulk32.hip (FP32-bulk form)
int b32_alloc(Ctx *c);
void b32_free(Ctx *c);
int b32_set_parcels(Ctx *c, int64_t n, const double *const *fields);
int b32_get_parcels(Ctx *c, double *const *fields);
int b32_par2grid(Ctx *c);                 // P, Q, R, B grids (FP64) from the FP32 parcels
int b32_rk_stage(Ctx *c, int stage, int stretch);
int b32_split(Ctx *c);
int b32_mix(Ctx *c);
int b32_conservation_sums(Ctx *c, double out6[6]);
int b32_choose_capc(const Ctx *c, int64_t n);  // the slot capacity a binning of n parcels takes
int b32_mean_capc(const Ctx *c, int64_t n);    // ... by the mean-occupancy rule alone
void b32_graph_reuse(Ctx *c);  // stage 1's par2grid reading the dirty-tile flags, all set if stale
int b32_publish_n(Ctx *c);                     // the parcel count into device memory
int b32_export_grid(Ctx *c, int id);  // brings grid[id] up to date for a grid export
int b32_dump_grids(Ctx *c);             // writebinary's b, hg, p, q, r grids
int scan_exclusive_i32(Ctx *c, const int32_t *in, int32_t *out, int64_t n);
int sync_scalars(Ctx *c);
// capi.hip: status raised by kernels since the last check, cleared once read
int take_status(Ctx *c);
int ensure_fbuf(Ctx *c);  // allocates the host-transfer staging buffer at its first use

}  // namespace pmpic

// the opaque context handle of the C ABI
struct pmpic_ctx {
  pmpic::Ctx c;
};
