// pmpic_par2grid_noalloc: par2grid of a device array that allocates nothing.
//
// The same binning and per-cell sums as pmpic_par2grid (k_bin_count, the exclusive scan,
// k_bin_scatter, k_cell_accum_lds, k_node_sum), so the grid it returns is bit-identical to that
// entry point's.  Its working memory (cell and rank of each parcel, the per-cell counts and
// offsets, the cell-sorted parcel list, the corner partial sums, the interior grid and the
// scan's temporary storage) is carved out of `scratch`, one static device array per GPU that
// lives until the process ends, instead of the context's buffers or hipMalloc.  Inputs and the
// output are the caller's device arrays.
//
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two host threads,
// or two streams) race on it.  The per-cell counts are cleared by a fill kernel on each call;
// every other region is written in full before it is read, so nothing from an earlier call
// leaks into a later one.
#include "common.h"

#include <mutex>

#include "../../include/pmpic_hip.h"

namespace pmpic {
namespace {

constexpr int64_t NA_MAX_PARCELS = PMPIC_NOALLOC_MAX_PARCELS;
constexpr int64_t NA_MAX_CELLS = PMPIC_NOALLOC_MAX_CELLS;
constexpr size_t NA_ALIGN = 256;
constexpr size_t na_round(size_t b) { return (b + NA_ALIGN - 1) / NA_ALIGN * NA_ALIGN; }

// byte offsets of the regions inside `scratch`
constexpr size_t NA_CELL = 0;
constexpr size_t NA_RANK = NA_CELL + na_round(4 * NA_MAX_PARCELS);
constexpr size_t NA_SORTED = NA_RANK + na_round(4 * NA_MAX_PARCELS);
constexpr size_t NA_COUNT = NA_SORTED + na_round(4 * NA_MAX_PARCELS);
constexpr size_t NA_START = NA_COUNT + na_round(4 * (NA_MAX_CELLS + 1));
// 8 corners x (weight, weight * value) per cell
constexpr size_t NA_PARTIAL = NA_START + na_round(4 * (NA_MAX_CELLS + 1));
// interior grid [nzn][ny][nx]: nzn * ny * nx = ncell + nx * ny <= 2 ncell since nz >= 2
constexpr size_t NA_GRID = NA_PARTIAL + na_round(8 * 16 * NA_MAX_CELLS);
constexpr size_t NA_SCAN = NA_GRID + na_round(8 * 2 * NA_MAX_CELLS);
constexpr size_t NA_SCAN_BYTES = 256 * 1024;
constexpr size_t NA_TOTAL = NA_SCAN + NA_SCAN_BYTES;
static_assert(NA_TOTAL == PMPIC_NOALLOC_SCRATCH_BYTES, "scratch size in pmpic_hip.h");

// Host code must not pass `scratch` itself to a kernel: on the host the name is a stand-in, not
// the device address.  scratch_addr() gives each device's address of it.
static __device__ __align__(256) unsigned char scratch[NA_TOTAL];

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mutex;
unsigned char *na_addr[NA_MAX_DEVICES];  // device address of `scratch` per device, once looked up

// The address of `scratch` on `device`, looked up once per device with that device current.
int scratch_addr(int device, unsigned char **out) {
  if (device < 0 || device >= NA_MAX_DEVICES) return PMPIC_ERR_ARG;
  std::lock_guard<std::mutex> lock(na_mutex);
  if (!na_addr[device]) {
    int prev = 0;
    PMPIC_CHECK(hipGetDevice(&prev));
    if (prev != device) PMPIC_CHECK(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) PMPIC_CHECK(hipSetDevice(prev));
    if (e != hipSuccess || p == nullptr) return PMPIC_ERR_HIP;
    na_addr[device] = static_cast<unsigned char *>(p);
  }
  *out = na_addr[device];
  return PMPIC_OK;
}

}  // namespace
}  // namespace pmpic

using namespace pmpic;

extern "C" int pmpic_par2grid_noalloc(pmpic_ctx *ctx, const double *var, double *grid) {
  if (!ctx || !var || !grid || ctx->c.precision != 0) return PMPIC_ERR_ARG;
  Ctx *c = &ctx->c;
  if (c->n > NA_MAX_PARCELS || c->ncell > NA_MAX_CELLS) return PMPIC_ERR_ARG;
  if (scan_temp_bytes(c, c->ncell + 1) > NA_SCAN_BYTES) return PMPIC_ERR_ARG;
  unsigned char *base = nullptr;
  PMPIC_TRY(scratch_addr(c->device, &base));
  BinWork w;
  w.cell = reinterpret_cast<int32_t *>(base + NA_CELL);
  w.rank = reinterpret_cast<int32_t *>(base + NA_RANK);
  w.sorted = reinterpret_cast<int32_t *>(base + NA_SORTED);
  w.count = reinterpret_cast<int32_t *>(base + NA_COUNT);
  w.start = reinterpret_cast<int32_t *>(base + NA_START);
  w.partial = reinterpret_cast<double *>(base + NA_PARTIAL);
  w.scan_tmp = base + NA_SCAN;
  w.scan_bytes = NA_SCAN_BYTES;
  double *interior = reinterpret_cast<double *>(base + NA_GRID);
  PMPIC_TRY(par2grid_single_ws(c, c->n, var, interior, w));
  PMPIC_TRY(to_ref_layout(c, interior, grid));
  return take_status(c);
}
