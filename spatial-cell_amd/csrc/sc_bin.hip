// Binning of caller-owned device arrays: the reference's updateGridAndSort
// (spatial_cell_sim/src/grid.cuh:111-139) as a routine -- cell keys
// (grid.cuh:13-44), their stable LSD sort (grid.cuh:76-108, CUB SortPairs in
// the reference; the library's onesweep here) and the cell table
// (grid.cuh:46-73, as CSR starts) -- over arrays the caller owns, with the
// same kernels the simulation's sort stage runs.  The reference sorts whole
// particle structs; this routine hands back the permutation instead.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "sc_device.h"
#include "sc_setup.h"

using namespace scd;

namespace scd {
void set_handleless_error(const std::string& e);
}

namespace {

constexpr size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace of one binning call: counters, the sort's unsorted and
// intermediate keys and values, digit histograms, look-back words, tile
// counters and the queue of long empty-cell runs.
struct BinLayout {
  size_t ctr, keys_in, keys_tmp, vals_tmp, hist, status, tile_ctr, gap_ctr, gaps, total;
  int max_tiles;
};

constexpr BinLayout bin_layout(size_t n, size_t cells) {
  BinLayout L{};
  const size_t slots = n < 1 ? 1 : n;
  L.max_tiles = (int)((slots + kSortTile - 1) / kSortTile);
  size_t o = 0;
  L.ctr = o;      o = up256(o + sizeof(Counters));
  L.keys_in = o;  o = up256(o + slots * 4);
  L.keys_tmp = o; o = up256(o + slots * 4);
  L.vals_tmp = o; o = up256(o + 2 * slots * 4);
  L.hist = o;     o = up256(o + (size_t)kMaxPasses * kRadix * 4);
  L.status = o;   o = up256(o + (size_t)kMaxPasses * L.max_tiles * kRadix * 4);
  L.tile_ctr = o; o = up256(o + (size_t)kMaxPasses * 4);
  L.gap_ctr = o;  o = up256(o + 4);
  L.gaps = o;     o = up256(o + (cells / 64 + 2) * sizeof(uint4));   // runs longer than 64 cells
  L.total = o;
  return L;
}

sc_status bin_fail(sc_status s, const std::string& msg) {
  set_handleless_error(msg);
  return s;
}

struct DeviceGuard {   // makes `device` current for the call and restores the caller's device
  int prev = -1;
  bool ok = false;
  explicit DeviceGuard(int device) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    ok = hipSetDevice(device) == hipSuccess;
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};

// Validation shared by the entry points; fills the grid of the call.
sc_status bin_check(int device, const float* pos, const int32_t* hdr, int n, float sim_size, float grid_cell_size,
                    int boundary, uint32_t* keys_out, int32_t* perm_out, uint32_t* cell_start, GridP& g) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bin_fail(SC_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return bin_fail(SC_ERR_INVALID, "device ordinal out of range");
  if (n < 0 || n >= (1 << 30)) return bin_fail(SC_ERR_INVALID, "particle count must be in [0, 2^30)");
  if (!cell_start || (n > 0 && (!pos || !hdr || !keys_out || !perm_out)))
    return bin_fail(SC_ERR_INVALID, "null array");
  if (((uintptr_t)pos | (uintptr_t)hdr) & 15u) return bin_fail(SC_ERR_INVALID, "pos and hdr must be 16-byte aligned");
  sc_params P{};
  P.sim_size = sim_size;
  P.grid_cell_size = grid_cell_size;
  P.boundary = boundary;
  int n_cells = 0, bits = 0, n_per = 0;
  std::string err;
  if (!scs::derive_grid(P, n_cells, bits, n_per, err)) return bin_fail(SC_ERR_INVALID, "grid: " + err);
  std::memset(&g, 0, sizeof(g));
  g.cell = grid_cell_size; g.L = sim_size; g.n_cells = n_cells; g.bits = bits;
  g.boundary = boundary; g.n_per = n_per;
  const uint64_t cells = (uint64_t)1 << (3 * bits);
  g.sentinel = (uint32_t)cells;
  g.key_bits = 3 * bits + 1;                      // the cells plus the inactive marker
  g.passes = (g.key_bits + kRadixBits - 1) / kRadixBits;
  return SC_OK;
}

// The kernels of the call on stream s, with the workspace at ws.
sc_status bin_enqueue(char* ws, const BinLayout& L, const float* pos, const int32_t* hdr, int n, const GridP& g,
                      uint32_t* keys_out, int32_t* perm_out, uint32_t* cell_start, hipStream_t s) {
  Pop p;
  std::memset(&p, 0, sizeof(p));
  p.buf[F_POS][0] = (void*)pos;
  p.buf[F_HDR][0] = (void*)hdr;
  p.cap = n < 1 ? 1 : n;
  p.launch = p.cap;
  p.keys_in = (uint32_t*)(ws + L.keys_in);
  p.keys_tmp = (uint32_t*)(ws + L.keys_tmp);
  p.vals_tmp = (uint32_t*)(ws + L.vals_tmp);
  // the last pass writes the caller's arrays (nothing is written when n == 0)
  p.keys_sorted = n > 0 ? keys_out : p.keys_tmp;
  p.vals_sorted = n > 0 ? (uint32_t*)perm_out : p.vals_tmp;
  p.cs = cell_start;
  p.gaps = (uint4*)(ws + L.gaps);
  p.gap_ctr = (uint32_t*)(ws + L.gap_ctr);
  p.hist = (uint32_t*)(ws + L.hist);
  p.status = (uint32_t*)(ws + L.status);
  p.tile_ctr = (uint32_t*)(ws + L.tile_ctr);
  p.max_tiles = L.max_tiles;
  Counters* ctr = (Counters*)(ws + L.ctr);
  (void)hipGetLastError();
  launch_bin_init(ctr, n, s);
  launch_keys(p, 0, g, ctr, s);
  launch_sort_passes(p, 0, g, ctr, s);
  launch_cell_starts(p, 0, g, ctr, s);
  launch_bin_marker(keys_out, n, g.sentinel, s);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bin_fail(SC_ERR_DEVICE, std::string("launch failed: ") + hipGetErrorString(e));
  return SC_OK;
}

// ---- the no-allocation variant --------------------------------------------
// Its workspace is one file-scope device array sized for the largest call it
// accepts.  Each GPU holds its own copy at its own address, found once per
// device with hipGetSymbolAddress; in host code the array's name is not a
// device address, so only the per-device pointer below is ever handed to a
// kernel.  Every call on a GPU shares that GPU's copy: overlapping calls on
// one GPU race.  What must start at zero (counters, digit histograms, tile
// counters, look-back words, the empty-run queue's count) is cleared by the
// call's own kernels, so a call never depends on what the last one left.
constexpr BinLayout kScratchLayout = bin_layout(SC_BIN_NOALLOC_MAX_SLOTS, SC_BIN_NOALLOC_MAX_CELLS);
static __device__ __align__(256) unsigned char scratch[kScratchLayout.total];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (which is `device`).
sc_status scratch_of(int device, char** out) {
  if (device >= kMaxDevices) return bin_fail(SC_ERR_INVALID, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return bin_fail(SC_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = (char*)g_scratch_ptr[device];
  return SC_OK;
}

}  // namespace

extern "C" {

sc_status sc_bin_noalloc(int device, const float* pos, const int32_t* hdr, int n, float sim_size,
                         float grid_cell_size, int boundary, uint32_t* keys_out, int32_t* perm_out,
                         uint32_t* cell_start, void* stream) {
  GridP g;
  sc_status st = bin_check(device, pos, hdr, n, sim_size, grid_cell_size, boundary, keys_out, perm_out, cell_start, g);
  if (st != SC_OK) return st;
  if (n > SC_BIN_NOALLOC_MAX_SLOTS)
    return bin_fail(SC_ERR_BUDGET, "sc_bin_noalloc: more than SC_BIN_NOALLOC_MAX_SLOTS (2^20) slots: " +
                                       std::to_string(n));
  if ((uint64_t)g.sentinel > (uint64_t)SC_BIN_NOALLOC_MAX_CELLS)
    return bin_fail(SC_ERR_BUDGET, "sc_bin_noalloc: grid of more than SC_BIN_NOALLOC_MAX_CELLS (2^24) cells: " +
                                       std::to_string(g.sentinel));
  DeviceGuard dg(device);
  if (!dg.ok) return bin_fail(SC_ERR_DEVICE, "hipSetDevice failed");
  char* ws = nullptr;
  if ((st = scratch_of(device, &ws)) != SC_OK) return st;
  // the scratch layout is the one for the largest call; a smaller call uses
  // the same offsets (every region is at least as large as it needs)
  return bin_enqueue(ws, kScratchLayout, pos, hdr, n, g, keys_out, perm_out, cell_start, (hipStream_t)stream);
}

sc_status sc_bin(int device, const float* pos, const int32_t* hdr, int n, float sim_size, float grid_cell_size,
                 int boundary, uint32_t* keys_out, int32_t* perm_out, uint32_t* cell_start, void* stream) {
  GridP g;
  sc_status st = bin_check(device, pos, hdr, n, sim_size, grid_cell_size, boundary, keys_out, perm_out, cell_start, g);
  if (st != SC_OK) return st;
  DeviceGuard dg(device);
  if (!dg.ok) return bin_fail(SC_ERR_DEVICE, "hipSetDevice failed");
  const BinLayout L = bin_layout((size_t)n, (size_t)g.sentinel);
  void* ws = nullptr;
  hipError_t e = hipMalloc(&ws, L.total);
  if (e != hipSuccess) return bin_fail(SC_ERR_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
  hipStream_t s = (hipStream_t)stream;
  st = bin_enqueue((char*)ws, L, pos, hdr, n, g, keys_out, perm_out, cell_start, s);
  e = hipStreamSynchronize(s);
  (void)hipFree(ws);
  if (st != SC_OK) return st;
  if (e != hipSuccess) return bin_fail(SC_ERR_DEVICE, std::string("binning failed: ") + hipGetErrorString(e));
  return SC_OK;
}

}  // extern "C"
