// gfx950 kernels of the per-step particle pipeline.
//
// Every neighbour kernel runs one lane per sorted slot and visits its
// neighbours in the reference order (cells by ascending gx, gy, gz, slots
// ascending inside a cell, dynamics.cuh:56-62), so each particle's floating
// point accumulation happens in exactly the reference's order.  With the
// clamp boundary the three cells of one (gx, gy) column are consecutive keys,
// hence one contiguous slot range: nine ranges per particle instead of 27.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include "sc_device.h"
#include "sc_vec.h"

#pragma clang fp contract(off)

namespace scd {

using namespace scv;

#define SC_GLOBAL __global__ __launch_bounds__(kBlock)

// Every launch of the library goes through here; the count numbers the
// nodes of a captured step (a captured stream records one node per launch).
thread_local uint64_t g_launches = 0;
thread_local PassTimer* g_pass_timer = nullptr;
#define SC_LAUNCH(...)              \
  do {                              \
    ++g_launches;                   \
    hipLaunchKernelGGL(__VA_ARGS__); \
  } while (0)

__device__ __forceinline__ v3 xyz(float4 a) { return m3(a.x, a.y, a.z); }
__device__ __forceinline__ v4 q4(float4 a) { return m4(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ float4 f4(v4 a) { return make_float4(a.x, a.y, a.z, a.w); }
__device__ __forceinline__ float4 f4w(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// grid.cuh:13-17 -- truncate, then clamp to the last cell
__device__ __forceinline__ uint32_t grid_idx(const GridP& g, float c) {
  uint32_t v = (c >= 0.0f) ? (uint32_t)(c / g.cell) : 0u;
  uint32_t lim = (g.boundary == SC_BOUNDARY_PERIODIC) ? (uint32_t)(g.n_per - 1) : (uint32_t)(g.n_cells - 1);
  return v < lim ? v : lim;
}
// grid.cuh:19-25, offset by the instance's first key (aligned, so | and + agree)
__device__ __forceinline__ uint32_t make_key(const GridP& g, uint32_t x, uint32_t y, uint32_t z) {
  return g.key_base + ((x << (2 * g.bits)) | (y << g.bits) | z);
}
// active flag and hydrophobicity of a neighbour, as the list builders test
// them, and the low 16 bits of its id (a bonded particle's walk compares
// them with its partners' first and loads the id only on a match)
__device__ __forceinline__ uint32_t pack_flags(int flags, int hydrophobic, int id) {
  return ((uint32_t)id << 16) | ((uint32_t)(hydrophobic & 0xff) << 8) | (uint32_t)(flags & 1);
}
__device__ __forceinline__ bool outside(v3 d, float lim) {
  return scm::fabs_(d.x) > lim || scm::fabs_(d.y) > lim || scm::fabs_(d.z) > lim;
}

// The c-th slot range of for_neighbors' walk around the cell (cx, cy, cz), c
// in [0, 27): false when it has none (clamped box: c >= 9, or a column
// outside the box).  For a loop that must not unroll the walk.
__device__ __forceinline__ bool stencil_range(const GridP& g, const uint32_t* __restrict__ cs, int cx, int cy, int cz,
                                              int c, uint32_t& lo, uint32_t& hi, float& sx, float& sy, float& sz) {
  sx = sy = sz = 0.0f;
  if (g.boundary != SC_BOUNDARY_PERIODIC) {
    if (c >= 9) return false;
    const int gx = cx - 1 + c / 3, gy = cy - 1 + c % 3;
    if (gx < 0 || gx >= g.n_cells || gy < 0 || gy >= g.n_cells) return false;
    const int z0 = max(cz - 1, 0), z1 = min(cz + 1, g.n_cells - 1);
    const uint32_t kb = make_key(g, gx, gy, 0);
    lo = cs[kb | (uint32_t)z0];
    hi = cs[(kb | (uint32_t)z1) + 1u];
    return true;
  }
  int gx = cx + c / 9 - 1, gy = cy + (c / 3) % 3 - 1, gz = cz + c % 3 - 1;
  if (gx < 0) { gx += g.n_per; sx = -g.L; } else if (gx >= g.n_per) { gx -= g.n_per; sx = g.L; }
  if (gy < 0) { gy += g.n_per; sy = -g.L; } else if (gy >= g.n_per) { gy -= g.n_per; sy = g.L; }
  if (gz < 0) { gz += g.n_per; sz = -g.L; } else if (gz >= g.n_per) { gz -= g.n_per; sz = g.L; }
  const uint32_t k = make_key(g, gx, gy, gz);
  lo = cs[k];
  hi = cs[k + 1u];
  return true;
}

// The periodic image shift of stencil_range's c-th range (no shift in a
// clamped box).
__device__ __forceinline__ void stencil_shift(const GridP& g, int cx, int cy, int cz, int c, float& sx, float& sy,
                                              float& sz) {
  sx = sy = sz = 0.0f;
  if (g.boundary != SC_BOUNDARY_PERIODIC) return;
  const int gx = cx + c / 9 - 1, gy = cy + (c / 3) % 3 - 1, gz = cz + c % 3 - 1;
  if (gx < 0) sx = -g.L; else if (gx >= g.n_per) sx = g.L;
  if (gy < 0) sy = -g.L; else if (gy >= g.n_per) sy = g.L;
  if (gz < 0) sz = -g.L; else if (gz >= g.n_per) sz = g.L;
}

// Visits the stencil cells one by one (for kernels that interleave two arrays
// per cell): f(key, sx, sy, sz).
template <class F>
__device__ __forceinline__ void for_cells(const GridP& g, v3 pos, F f) {
  int cx = (int)grid_idx(g, pos.x), cy = (int)grid_idx(g, pos.y), cz = (int)grid_idx(g, pos.z);
  if (g.boundary != SC_BOUNDARY_PERIODIC) {
    int x0 = max(cx - 1, 0), x1 = min(cx + 1, g.n_cells - 1);
    int y0 = max(cy - 1, 0), y1 = min(cy + 1, g.n_cells - 1);
    int z0 = max(cz - 1, 0), z1 = min(cz + 1, g.n_cells - 1);
    for (int gx = x0; gx <= x1; gx++)
      for (int gy = y0; gy <= y1; gy++)
        for (int gz = z0; gz <= z1; gz++) f(make_key(g, gx, gy, gz), 0.0f, 0.0f, 0.0f);
    return;
  }
  for (int dx = -1; dx <= 1; dx++) {
    int gx = cx + dx; float sx = 0.0f;
    if (gx < 0) { gx += g.n_per; sx = -g.L; } else if (gx >= g.n_per) { gx -= g.n_per; sx = g.L; }
    for (int dy = -1; dy <= 1; dy++) {
      int gy = cy + dy; float sy = 0.0f;
      if (gy < 0) { gy += g.n_per; sy = -g.L; } else if (gy >= g.n_per) { gy -= g.n_per; sy = g.L; }
      for (int dz = -1; dz <= 1; dz++) {
        int gz = cz + dz; float sz = 0.0f;
        if (gz < 0) { gz += g.n_per; sz = -g.L; } else if (gz >= g.n_per) { gz -= g.n_per; sz = g.L; }
        f(make_key(g, gx, gy, gz), sx, sy, sz);
      }
    }
  }
}

__device__ __forceinline__ float4 shift4(float4 p, float sx, float sy, float sz) {
  if (sx != 0.0f || sy != 0.0f || sz != 0.0f) { p.x = p.x + sx; p.y = p.y + sy; p.z = p.z + sz; }
  return p;
}

// Constants of slot i's cell: the instance table entry in a batch, the uniform
// ones for a single cell.
__device__ __forceinline__ GridP grid_of(const GridP& g, int i) { return g.table ? g.table[g.inst[i]] : g; }
// box length of slot i's cell (periodic image shifts)
__device__ __forceinline__ float box_of(const GridP& g, int i) { return g.table ? g.table[g.inst[i]].L : g.L; }
// step counter of slot i's cell (Philox counter, state-change steps)
__device__ __forceinline__ int local_step(const GridP& g, int step) { return step - g.step0; }

// Tile (kBlock slots of the sorted array) of this block in the neighbour
// kernels.  The dispatcher hands blocks to the 8 XCDs in turn (block b to
// XCD b % 8), and each XCD has its own L2.  Renumbered, runs of `run`
// consecutive tiles go to one XCD each, the runs to the XCDs in turn: tiles
// next to each other share neighbour cells, which then come from one L2,
// while the dense parts of a cell, which take contiguous ranges of tiles,
// are still spread over all XCDs.  Blocks past the last whole round of runs
// keep their own index, so the map is a bijection of [0, gridDim.x).
constexpr int kXcds = 8;
// (b: the block's index among nb blocks that take tiles; a multiple of
// kXcds blocks in front of them keeps their XCDs)
__device__ __forceinline__ int xcd_tile_of(int b, int nb, int run) {
  if (run <= 0 || b >= nb / (kXcds * run) * (kXcds * run)) return b;
  const int x = b % kXcds, r = b / kXcds;   // XCD, and the block's rank among that XCD's blocks
  return ((r / run) * kXcds + x) * run + r % run;
}
__device__ __forceinline__ int xcd_tile(int run) { return xcd_tile_of((int)blockIdx.x, (int)gridDim.x, run); }

// slot of this pass thread: a worklist entry of its tile, or its own slot when the pass visits every slot
__device__ __forceinline__ int pass_slot_at(int tile, const int* work, const int* cnt, int n_active) {
  const int t = tile * kBlock + threadIdx.x;
  if (work) {
    // both in flight together: the list slot exists whether or not it is in use
    const int c = cnt[tile];
    const int w = work[t];
    return (int)threadIdx.x < c ? w : -1;
  }
  return t < n_active ? t : -1;
}
__device__ __forceinline__ int pass_slot(const int* work, const int* cnt, int n_active, int run) {
  return pass_slot_at(xcd_tile(run), work, cnt, n_active);
}

__device__ __forceinline__ bool participant_order(int ptype, int pid, int ttype, int tid) {
  return ptype < ttype || (ptype == ttype && pid < tid);   // interactions.cuh:186-189
}

__device__ __forceinline__ sc_interaction_info info_at(const Tables& t, int id) {
  if (id >= 0 && id < t.n_flat) return t.flat[id];
  sc_interaction_info z;
  z.id = 0; z.group = 0; z.first_partner_type = 0; z.first_partner_state = 0; z.second_partner_type = 0;
  z.second_partner_state = 0; z.propagate_state = 0; z.set_state = 0; z.wait_for_state = 0;
  z.wait_for_alignment = 0; z.only_via_transition = 0; z.transition_to = 0; z.polymerize = 0;
  z.break_after_transition = 0;
  z.relative_orientation[0] = z.relative_orientation[1] = z.relative_orientation[2] = z.relative_orientation[3] = 0.0f;
  z.orientation_alignment_threshold = 0.0f;
  z.relative_position[0] = z.relative_position[1] = z.relative_position[2] = 0.0f;
  z.position_alignment_threshold = 0.0f;
  z.bond_probability = 0.0f; z.contact_distance = 0.0f;
  return z;
}
__device__ __forceinline__ v4 ro_of(const sc_interaction_info& i) {
  return m4(i.relative_orientation[0], i.relative_orientation[1], i.relative_orientation[2], i.relative_orientation[3]);
}
__device__ __forceinline__ v3 rp_of(const sc_interaction_info& i) {
  return m3(i.relative_position[0], i.relative_position[1], i.relative_position[2]);
}

// ============================================================================
// Binning: keys + histogram, onesweep LSD radix sort, permute, cell starts
// ============================================================================

struct PrepArgs2 { uint32_t* hist; uint32_t* status; size_t status_stride; uint32_t* tile_ctr; Counters* ctr; int pop; int passes;
                   int tile_keys; };

// Zeroes the global digit histograms, tile counters and look-back words of
// the passes that will run.
SC_GLOBAL void k_sort_zero(PrepArgs2 a) {
  int n = a.ctr->n_active[a.pop];
  size_t words = (size_t)((n + a.tile_keys - 1) / a.tile_keys) * kRadix;
  size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  size_t stride = (size_t)gridDim.x * kBlock;
  for (int p = 0; p < a.passes; p++)
    for (size_t k = i; k < words; k += stride) a.status[p * a.status_stride + k] = 0u;
  if (i < (size_t)a.passes * kRadix) a.hist[i] = 0u;
  if (i < (size_t)kMaxPasses) a.tile_ctr[i] = 0u;
  if (i == 0) a.ctr->rmax_bits[a.pop] = 0u;
}

struct KeysArgs {
  const float4* pos; const int4* hdr;
  uint32_t* keys; uint32_t* hist; uint32_t* gap_ctr;
  Counters* ctr; GridP g; int pop;
  int rounds;   // tiles of 256 slots per block
};

// grid.cuh:34-44 plus the digit histogram of every radix pass.  A block
// takes `rounds` x 256 consecutive slots, so that the block histograms reach
// the global one with few atomics (every block adds to the same bins); small
// arrays keep one tile per block, for parallelism.
constexpr int kKeysRounds = 8, kKeysMinBlocks = 512;
SC_GLOBAL void k_keys(KeysArgs a) {
  __shared__ uint32_t lh[kMaxPasses * kRadix];
  for (int b = threadIdx.x; b < a.g.passes * kRadix; b += kBlock) lh[b] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.gap_ctr = 0u;
  __syncthreads();
  const int n = a.ctr->n_active[a.pop];
  const int lane = (int)(threadIdx.x & 63u);
  const uint64_t above = lane == 63 ? 0ull : ~((2ull << lane) - 1ull);
  uint32_t rmax = 0u;   // largest radius as bits (non-negative floats order as their bits)
  for (int r = 0; r < a.rounds; r++) {
    const int i = (blockIdx.x * a.rounds + r) * kBlock + threadIdx.x;
    if ((blockIdx.x * a.rounds + r) * kBlock >= n) break;   // the block's remaining rounds are empty
    const bool valid = i < n;
    uint32_t key = 0u;
    if (valid) {
      float4 p = a.pos[i];
      const uint32_t rb = __float_as_uint(p.w);
      rmax = max(rmax, (rb >> 31) ? 0x7f800000u : rb);
      int flags = a.hdr[i].z;
      const GridP g = grid_of(a.g, i);
      if (a.g.drop_evicted && g.evicted) {
        key = a.g.sentinel + 1u;          // sorts after every live slot; cut off after the sort
      } else if (flags & SC_FLAG_ACTIVE) {
        key = make_key(g, grid_idx(g, p.x), grid_idx(g, p.y), grid_idx(g, p.z));
      } else {
        key = a.g.sentinel;
      }
      a.keys[i] = key;
    }
    // The slots are still in last step's order, so neighbouring lanes mostly
    // share digits: each run of equal digits within a wave adds its length
    // with one LDS atomic (instead of one atomic per lane on the same bin).
    for (int ps = 0; ps < a.g.passes; ps++) {
      const uint32_t d = valid ? (key >> (ps * kRadixBits)) & (kRadix - 1) : 0xffffffffu;
      const uint32_t prev = (uint32_t)__shfl_up((int)d, 1u);
      const bool head = lane == 0 || prev != d;
      const uint64_t heads = __ballot(head);
      if (head && valid) {
        const uint64_t later = heads & above;
        const int next = later ? __ffsll((unsigned long long)later) - 1 : 64;
        atomicAdd(&lh[ps * kRadix + d], (uint32_t)(next - lane));
      }
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < a.g.passes * kRadix; b += kBlock)
    if (lh[b]) atomicAdd(&a.hist[b], lh[b]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) rmax = max(rmax, (uint32_t)__shfl_xor((int)rmax, o));
  if (lane == 0 && rmax) atomicMax(&a.ctr->rmax_bits[a.pop], rmax);
}

// Exclusive prefix sum of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp /* >= 4 */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_tmp[w] = x;
  __syncthreads();
  uint32_t wofs = 0;
  for (int k = 0; k < w; k++) wofs += s_tmp[k];
  __syncthreads();
  return wofs + x - v;
}

#define SC_FLAG_AGG 0x40000000u
#define SC_FLAG_INC 0x80000000u
#define SC_COUNT_MASK 0x3fffffffu

struct SweepArgs {
  const uint32_t* keys_in; const uint32_t* vals_in;   // vals_in == nullptr: identity
  uint32_t* keys_out; uint32_t* vals_out;
  const uint32_t* hist;      // this pass, kRadix counts
  uint32_t* status;          // this pass, [tile][kRadix]
  uint32_t* tile_ctr;        // this pass
  Counters* ctr; int pop; int shift;
};

// One stable LSD pass over 8 key bits with decoupled look-back (onesweep).
// Tile = 4 waves x ITEMS rounds x 64 lanes, loaded wave-striped so that the
// (wave, round, lane) order is the input order; keys with equal digits are
// ranked with 8 ballots per round, which keeps the pass stable.  ITEMS
// (GridP::sort_items) trades tiles for parallelism: fewer keys per tile, more
// tiles to fill the GPU on small arrays.
template <int ITEMS>
SC_GLOBAL void k_onesweep(SweepArgs a) {
  constexpr int kTile = kBlock * ITEMS;
  __shared__ uint32_t s_keys[kTile];
  __shared__ uint32_t s_vals[kTile];
  __shared__ uint32_t s_wh[4][kRadix];
  __shared__ uint32_t s_base[kRadix];
  __shared__ uint32_t s_start[kRadix];
  __shared__ uint32_t s_tmp[8];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int n = a.ctr->n_active[a.pop];
  const uint32_t ntiles = (uint32_t)((n + kTile - 1) / kTile);
  if (tid == 0) s_tile = atomicAdd(a.tile_ctr, 1u);
  for (int b = tid; b < 4 * kRadix; b += kBlock) (&s_wh[0][0])[b] = 0u;
  __syncthreads();
  const uint32_t t = s_tile;
  if (t >= ntiles) return;

  const int base = (int)t * kTile + w * 64 * ITEMS;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t rk[ITEMS], rv[ITEMS], rr[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    int idx = base + r * 64 + lane;
    bool valid = idx < n;
    uint32_t key = valid ? a.keys_in[idx] : 0xffffffffu;
    uint32_t val = valid ? (a.vals_in ? a.vals_in[idx] : (uint32_t)idx) : 0u;
    uint32_t d = (key >> a.shift) & (kRadix - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; b++) {
      bool bit = (d >> b) & 1u;
      uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t before = s_wh[w][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && (peers & lt) == 0ull) s_wh[w][d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rk[r] = key;
    rv[r] = val;
    rr[r] = before + (uint32_t)__popcll(peers & lt);
  }
  __syncthreads();

  // per digit: wave offsets, tile total, global base via look-back
  const int d = tid;
  uint32_t c0 = s_wh[0][d], c1 = s_wh[1][d], c2 = s_wh[2][d], c3 = s_wh[3][d];
  uint32_t total = c0 + c1 + c2 + c3;
  uint32_t* st = a.status + (size_t)t * kRadix + d;
  if (t == 0) __hip_atomic_store(st, SC_FLAG_INC | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else __hip_atomic_store(st, SC_FLAG_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t gscan = block_exclusive_scan(a.hist[d], s_tmp);
  uint32_t tscan = block_exclusive_scan(total, s_tmp);
  uint32_t excl = 0;
  if (t > 0) {
    int tt = (int)t - 1;
    for (;;) {
      const uint32_t* q = a.status + (size_t)tt * kRadix + d;
      uint32_t v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      while ((v & (SC_FLAG_AGG | SC_FLAG_INC)) == 0u) {
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      excl += v & SC_COUNT_MASK;
      if (v & SC_FLAG_INC) break;
      tt--;
    }
    __hip_atomic_store(st, SC_FLAG_INC | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  s_wh[0][d] = 0u;
  s_wh[1][d] = c0;
  s_wh[2][d] = c0 + c1;
  s_wh[3][d] = c0 + c1 + c2;
  s_start[d] = tscan;
  s_base[d] = gscan + excl;
  __syncthreads();

#pragma unroll
  for (int r = 0; r < ITEMS; r++) {
    int idx = base + r * 64 + lane;
    if (idx < n) {
      uint32_t dd = (rk[r] >> a.shift) & (kRadix - 1);
      uint32_t lp = s_start[dd] + s_wh[w][dd] + rr[r];
      s_keys[lp] = rk[r];
      s_vals[lp] = rv[r];
    }
  }
  __syncthreads();
  const int cnt = min(kTile, n - (int)t * kTile);
  for (int i = tid; i < cnt; i += kBlock) {
    uint32_t key = s_keys[i];
    uint32_t dd = (key >> a.shift) & (kRadix - 1);
    uint32_t o = s_base[dd] + (uint32_t)i - s_start[dd];
    a.keys_out[o] = key;
    a.vals_out[o] = s_vals[i];
  }
}
template __global__ void k_onesweep<4>(SweepArgs);
template __global__ void k_onesweep<8>(SweepArgs);
template __global__ void k_onesweep<16>(SweepArgs);

struct PermArgs {
  Pop p;                     // by value: current/next pointers resolved on the host
  Counters* ctr; GridP g; int pop;
};

// Gather of every field into sorted order (the reference sorts whole structs
// as CUB values, grid.cuh:101-102), plus the {start, end} of each run of
// equal keys (grid.cuh:46-73).
SC_GLOBAL void k_permute(PermArgs a) {
  const Pop& P = a.p;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int n = a.ctr->n_active[a.pop];
  if (i >= n) return;
  const int j = (int)P.vals_sorted[i];
  const int cap = P.cap;
#define SC_GATHER(F, T) P.n<T>(F)[i] = P.c<T>(F)[j]
  SC_GATHER(F_POS, float4); SC_GATHER(F_VEL, float4); SC_GATHER(F_ROT, float4); SC_GATHER(F_AVEL, float4);
  SC_GATHER(F_PNOISE, float4); SC_GATHER(F_ANOISE, float4); SC_GATHER(F_DBG, float4);
  SC_GATHER(F_HDR, int4); SC_GATHER(F_INST, int);
  int4 st = P.c<int4>(F_ST)[j];
  P.n<int4>(F_ST)[i] = st;
  if (!P.metabolic) {   // the packed flags the stencil walks of the step read
    const int4 h = P.c<int4>(F_HDR)[j];
    P.flw[i] = pack_flags(h.z, h.w, h.x);
  }
  if (P.metabolic) {
    SC_GATHER(F_MET0, float4); SC_GATHER(F_MET1, float4); SC_GATHER(F_MET2, float2);
  } else {
    const int4* bc = P.c<int4>(F_BONDS);
    int4* bn = P.n<int4>(F_BONDS);
    int nb = min(max(st.z, 0), SC_MAX_BONDS);
    for (int k = 0; k < nb; k++) bn[(size_t)k * cap + i] = bc[(size_t)k * cap + j];
  }
#undef SC_GATHER
  if (i == 0) a.ctr->n_sorted[a.pop] = n;
}

constexpr uint32_t kGapDirect = 64;    // empty-cell runs up to this length are written by one lane
constexpr uint32_t kGapChunk = 4096;   // longer runs are split into chunks of this many cells

struct CellStartArgs { const uint32_t* K; uint32_t* cs; uint4* gaps; uint32_t* gap_ctr; Counters* ctr; int pop;
                       uint32_t cells; };

// cs[c] = first sorted slot with key >= c.  Lane i (0 <= i <= n) owns the
// cells (K[i-1], K[i]] (all of them get the value i); long runs of empty cells
// are queued for k_cell_gaps so that no lane walks millions of cells.
SC_GLOBAL void k_cell_start(CellStartArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int n = a.ctr->n_active[a.pop];
  if (i > n) return;
  uint32_t lo = (i == 0) ? 0u : min(a.K[i - 1], a.cells) + 1u;
  uint32_t hi = (i == n) ? a.cells : min(a.K[i], a.cells);     // inclusive
  if (lo > hi) return;
  uint32_t len = hi - lo + 1u;
  uint32_t direct = len <= kGapDirect ? len : kGapDirect;
  for (uint32_t c = 0; c < direct; c++) a.cs[lo + c] = (uint32_t)i;
  if (len > kGapDirect) {   // the rest in chunks of kGapChunk cells, spread over blocks by k_cell_gaps
    const uint32_t rest = len - kGapDirect;
    const uint32_t chunks = (rest + kGapChunk - 1u) / kGapChunk;
    uint32_t q = atomicAdd(a.gap_ctr, chunks);
    for (uint32_t c = 0; c < chunks; c++) {
      uint32_t b = lo + kGapDirect + c * kGapChunk;
      a.gaps[q + c] = make_uint4(b, min(b + kGapChunk, hi + 1u), (uint32_t)i, 0u);
    }
  }
}

// After the sort of a batch that drops evicted instances: the live slots are
// the ones with keys up to the sentinel.
SC_GLOBAL void k_drop_evicted(CellStartArgs a) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int n = (int)a.cs[a.cells];
    a.ctr->n_active[a.pop] = n;
    a.ctr->n_sorted[a.pop] = n;
  }
}

SC_GLOBAL void k_cell_gaps(CellStartArgs a) {
  const uint32_t count = *a.gap_ctr;
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    uint4 gp = a.gaps[e];
    for (uint32_t c = gp.x + threadIdx.x; c < gp.y; c += kBlock) a.cs[c] = gp.z;
  }
}

// ============================================================================
// Integrator (dynamics.cuh:19-282)
// ============================================================================

struct MoveArgs { float4* dbg; float4* pnoise; float4* anoise; const float4* pos; const int4* hdr;
                  Counters* ctr; int pop; float noise; float rot_noise; GridP g; 
  int se; float se_translation, se_rotation;   // Stokes-Einstein noise (SC_NOISE_STOKES_EINSTEIN)
};

// move (debugVector.xyz = 0, dynamics.cuh:42-44) + brownianMovementAndRotation
// (dynamics.cuh:176-207) with Philox counters (id, step, array, block).
SC_GLOBAL void k_move(MoveArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[a.pop]) return;
  int4 h = a.hdr[i];
  if (!(h.z & SC_FLAG_ACTIVE)) return;
  float4 dbg = a.dbg[i];
  dbg.x = 0.0f; dbg.y = 0.0f; dbg.z = 0.0f;
  a.dbg[i] = dbg;
  const GridP g = grid_of(a.g, i);
  const uint32_t step = (uint32_t)local_step(g, a.ctr->step);
  scm::u32x4 c0 = {(uint32_t)h.x, step, (uint32_t)a.pop, 0u};
  scm::u32x4 c1 = {(uint32_t)h.x, step, (uint32_t)a.pop, 1u};
  scm::u32x4 r0 = scm::philox4x32_10(c0, g.seed_lo, g.seed_hi);
  scm::u32x4 r1 = scm::philox4x32_10(c1, g.seed_lo, g.seed_hi);
  float radius = a.pos[i].w;
  v4 dir = shoemake(scm::u01_(r0.x), scm::u01_(r0.y), scm::u01_(r0.z));
  float nrmv = scm::box_muller_(scm::u01_(r0.w), scm::u01_(r1.x));
  v4 rr = shoemake(scm::u01_(r1.y), scm::u01_(r1.z), scm::u01_(r1.w));
  float mag, frac;
  if (a.se) {
    // Stokes-Einstein scales of a sphere of this radius (sc_params)
    mag = scm::fabs_(nrmv) * scm::sqrt_(a.se_translation / radius);
    frac = scm::sqrt_(a.se_rotation / (radius * radius * radius));
  } else {
    mag = (float)((double)(scm::fabs_(nrmv) * a.noise) / scm::dmax_((double)radius / 2.5, 1.0));
    frac = (float)((double)a.rot_noise / ((double)radius / 2.5));
  }
  v3 pn = mul(rotate_up(dir), mag);
  float4 pno = a.pnoise[i];
  a.pnoise[i] = f4w(pn, pno.w);
  v4 an = slerp(q4(a.anoise[i]), rr, frac);
  a.anoise[i] = f4(an);
}

struct ApplyNoiseArgs { float4* vel; float4* avel; float4* pnoise; float4* anoise; const int4* hdr; Counters* ctr; int pop; };

// dynamics.cuh:217-242
SC_GLOBAL void k_apply_noise(ApplyNoiseArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[a.pop]) return;
  if (!(a.hdr[i].z & SC_FLAG_ACTIVE)) return;
  float4 v = a.vel[i], pn = a.pnoise[i];
  v.x += pn.x; v.y += pn.y; v.z += pn.z;
  a.vel[i] = v;
  a.pnoise[i] = make_float4(0.0f, 0.0f, 0.0f, pn.w);
  a.avel[i] = f4(qmul(q4(a.anoise[i]), q4(a.avel[i])));
  a.anoise[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
}

struct ApplyVelArgs { float4* pos; float4* vel; float4* rot; float4* avel; const int4* hdr; Counters* ctr; int pop;
                      GridP g; float vdecay; float adecay; };

__device__ __forceinline__ float wrapL(float x, float L) {
  float y = x - L * scm::floor_(x / L);
  if (y >= L) y = y - L;
  if (y < 0.0f) y = 0.0f;
  return y;
}

// dynamics.cuh:244-282 with the step fraction 0.45 of main.cu:553,560
SC_GLOBAL void k_apply_velocities(ApplyVelArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[a.pop]) return;
  if (!(a.hdr[i].z & SC_FLAG_ACTIVE)) return;
  const float frac = 0.45f;
  float4 p = a.pos[i], v = a.vel[i];
  const GridP g = grid_of(a.g, i);
  const float L = g.L;
  v3 np = add(xyz(p), mul(xyz(v), frac));
  if (g.boundary == SC_BOUNDARY_PERIODIC) { p.x = wrapL(np.x, L); p.y = wrapL(np.y, L); p.z = wrapL(np.z, L); }
  else { p.x = clampf(np.x, 0.0f, L); p.y = clampf(np.y, 0.0f, L); p.z = clampf(np.z, 0.0f, L); }
  a.pos[i] = p;
  v.x *= a.vdecay; v.y *= a.vdecay; v.z *= a.vdecay;
  a.vel[i] = v;
  v4 av = q4(a.avel[i]);
  v4 r = qmul(slerp(qid(), av, frac), q4(a.rot[i]));
  r = div4(r, len4(r));
  a.rot[i] = f4(r);
  a.avel[i] = f4(slerp(av, qid(), 1.0f - a.adecay));
}

// ============================================================================
// relax (dynamics.cuh:405-822)
// ============================================================================

struct RelaxArgs {
  const float4* pos; const int4* hdr; const int4* st; const int4* bonds; int cap;
  const float4* vel_c; const float4* avel_c; const float4* rot_c; const float4* dbg_c;
  float4* vel_n; float4* avel_n; float4* rot_n; float4* dbg_n;
  const uint32_t* cs; const TilePlan* plans; PairList list; const int* work; const int* work_cnt;
  const int* work_first;   // non-null: k_relax_pass takes each block's front group, k_relax_coll the rest
  const int* work_bond;    // per block: the range of its worklist's bonded slots, first << 16 | count (k_bond_terms)
  uint2* rbp; float4* bterm; float4* bang;   // bonded particles' partner entries and bond terms (Pop)
  uint32_t* flw;           // per slot pack_flags(flags, hydrophobic, id) for the step (k_relax_prepare)
  float4* pv_c; float4* pv_n;   // {position, velocity} records: the current ones, the pass's output
  Counters* ctr; GridP g; Tables t;
};

// gather width of the collision-only relax kernel (GridP::coll_gather: 4 or 8)
#define SC_COLL_KERNEL(g) ((g).coll_gather == 4 ? k_relax_coll<4> : k_relax_coll<8>)

// Before the first relax pass of a step (in the per-pass set-up, or in the
// fused relaxation's, which runs first): both buffers of rot and debugVector
// get the current values (only bonded particles change them during the
// passes, so the others never need to write them again), the packed flags
// the list builders read, and the current {position, velocity} records.
SC_GLOBAL void k_relax_prepare(RelaxArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[0]) return;
  a.rot_n[i] = a.rot_c[i];
  a.dbg_n[i] = a.dbg_c[i];
  const int4 h = a.hdr[i];
  a.flw[i] = pack_flags(h.z, h.w, h.x);
  a.pv_c[2 * (size_t)i] = a.pos[i];
  a.pv_c[2 * (size_t)i + 1] = a.vel_c[i];
}

// A pass's new velocity of slot i, in the velocity array and in its record
// (with the position, so that the record is one full 32-byte store).
__device__ __forceinline__ void put_vel(const RelaxArgs& a, int i, const float4 P, const float4 v) {
  a.vel_n[i] = v;
  a.pv_n[2 * (size_t)i] = P;
  a.pv_n[2 * (size_t)i + 1] = v;
}

// ---- LDS staging of a tile's neighbour columns --------------------------------
// A tile is 256 consecutive sorted slots.  Its particles occupy the column
// interval [colA, colB] (column = key >> bits); for each dx in {-1, 0, 1} the
// neighbour columns of all of them lie in [colA + dx*2^bits - 1, colB +
// dx*2^bits + 1], i.e. one contiguous key range and hence one contiguous slot
// range of the (sorted) target array.  At most three ranges, merged when they
// overlap, are copied to LDS.  Plans depend only on the sorted keys, so they
// are computed once per step (k_tile_plans).  A neighbour range outside the
// staged ones (periodic wrap-around, particles that moved since the sort) or
// a tile whose ranges exceed the LDS capacity reads global memory instead,
// with identical results.

struct PlanView {
  uint32_t nr, total;
  uint32_t s0, e0, s1, e1, s2, e2;
  int b0, b1, b2;
  // a view that stages nothing: every range is read from global memory
  __device__ __forceinline__ static PlanView none() {
    PlanView v;
    v.nr = 0; v.total = 0;
    v.s0 = v.e0 = v.s1 = v.e1 = v.s2 = v.e2 = 0u;
    v.b0 = v.b1 = v.b2 = 0;
    return v;
  }
  __device__ __forceinline__ void load(const TilePlan& p, uint32_t cap) {
    nr = p.nr; total = p.total;
    s0 = p.s[0]; e0 = p.e[0]; s1 = p.s[1]; e1 = p.e[1]; s2 = p.s[2]; e2 = p.e[2];
    if (total > cap) { nr = 0; total = 0; }
    b0 = -(int)s0;
    b1 = (int)(e0 - s0) - (int)s1;
    b2 = (int)(e0 - s0) + (int)(e1 - s1) - (int)s2;
  }
  __device__ __forceinline__ bool find(uint32_t lo, uint32_t hi, int& base) const {
    if (nr > 0 && lo >= s0 && hi <= e0) { base = b0; return true; }
    if (nr > 1 && lo >= s1 && hi <= e1) { base = b1; return true; }
    if (nr > 2 && lo >= s2 && hi <= e2) { base = b2; return true; }
    return false;
  }
  __device__ __forceinline__ uint32_t slot_of(uint32_t idx) const {
    if (nr > 2 && (int)idx >= b2 + (int)s2) return (uint32_t)((int)idx - b2);
    if (nr > 1 && (int)idx >= b1 + (int)s1) return (uint32_t)((int)idx - b1);
    return (uint32_t)((int)idx - b0);
  }
};

struct PlanArgs { const uint32_t* K; const uint32_t* cs_self; const uint32_t* cs_tgt; const uint32_t* cs_tgt2;
                  TilePlan* plans; Counters* ctr; int pop; GridP g; int tile; };

SC_GLOBAL void k_tile_plans(PlanArgs a) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int n = a.ctr->n_active[a.pop];
  const int t0 = t * a.tile;
  if (t0 >= n) return;
  const int t1 = min(t0 + a.tile, n);
  TilePlan sp;
  sp.nr = 0;
  sp.total = 0;
  for (int r = 0; r < 3; r++) { sp.s[r] = 0u; sp.e[r] = 0u; }
  const int nvalid = (int)a.cs_self[a.g.sentinel];
  const int last = min(t1, nvalid) - 1;
  // a tile that spans two instances of a batch stages nothing
  const bool one_cell = t0 <= last && (!a.g.table || a.g.inst[t0] == a.g.inst[last]);
  if (one_cell) {
    const GridP g = grid_of(a.g, t0);
    const uint32_t base = g.key_base;              // keys below are the instance's own
    const uint32_t zmask = (1u << g.bits) - 1u;
    const uint32_t KA = a.K[t0] - base, KB = a.K[last] - base;
    const int64_t colA = (int64_t)(KA >> g.bits), colB = (int64_t)(KB >> g.bits);
    // Column c0 (gy - 1 of the first column) is reached only from the first
    // column's particles, whose cells are >= zA: cells below zA - 1 are not
    // needed.  Likewise column c1 above zB + 1.
    const uint32_t zA = KA & zmask, zB = KB & zmask;
    const uint32_t zlo = zA > 0u ? zA - 1u : 0u, zhi = zB < zmask ? zB + 1u : zmask;
    const int64_t ncols = (int64_t)1 << (2 * g.bits);
    for (int dx = -1; dx <= 1; dx++) {
      int64_t c0 = colA + (int64_t)dx * ((int64_t)1 << g.bits) - 1;
      int64_t c1 = colB + (int64_t)dx * ((int64_t)1 << g.bits) + 1;
      const bool trim0 = c0 >= 0, trim1 = c1 <= ncols - 1;
      if (c0 < 0) c0 = 0;
      if (c1 > ncols - 1) c1 = ncols - 1;
      if (c0 > c1) continue;
      const uint32_t k0 = base + (((uint32_t)c0 << g.bits) | (trim0 ? zlo : 0u));
      const uint32_t k1 = base + (((uint32_t)c1 << g.bits) | (trim1 ? zhi : zmask));
      uint32_t s = a.cs_tgt[k0], e = a.cs_tgt[k1 + 1u];
      if (a.cs_tgt2) { s += a.cs_tgt2[k0]; e += a.cs_tgt2[k1 + 1u]; }   // merged metabolic + plain array
      if (e <= s) continue;
      if (sp.nr > 0 && s <= sp.e[sp.nr - 1]) {
        if (e > sp.e[sp.nr - 1]) sp.e[sp.nr - 1] = e;
      } else {
        sp.s[sp.nr] = s; sp.e[sp.nr] = e; sp.nr++;
      }
    }
    for (uint32_t r = 0; r < sp.nr; r++) sp.total += sp.e[r] - sp.s[r];
  }
  a.plans[t] = sp;
}

// Visits the neighbour slot ranges of pos in the reference order (columns of
// the 27-cell stencil by gx, gy; inside a column cells by gz and slots
// ascending).  The bounds of all nine columns are loaded before any range is
// walked.  For each range: lds(j, li) when the slots are staged (li = LDS
// index), else glb(j, sx, sy, sz) with the periodic image shift.
// cell starts of one sorted array, or of the merged metabolic + plain array
// (per cell the metabolic particles, then the plain ones: its cell starts are
// the sums of the two arrays' cell starts)
struct CsOne {
  const uint32_t* a;
  __device__ __forceinline__ uint32_t operator()(uint32_t k) const { return a[k]; }
};
struct CsSum {
  const uint32_t* a; const uint32_t* b;
  __device__ __forceinline__ uint32_t operator()(uint32_t k) const { return a[k] + b[k]; }
};

template <class CS, class FL, class FG>
__device__ __forceinline__ void visit_cells(const GridP& g, CS cs, v3 pos, const PlanView& pv, FL lds, FG glb) {
  const int cx = (int)grid_idx(g, pos.x), cy = (int)grid_idx(g, pos.y), cz = (int)grid_idx(g, pos.z);
  const bool periodic = g.boundary == SC_BOUNDARY_PERIODIC;
  const int nz = periodic ? g.n_per : g.n_cells;
  // z extent of the column ranges; periodic columns whose stencil wraps in z are split below
  const bool zwrap = periodic && (cz == 0 || cz == nz - 1);
  const int z0 = periodic ? (cz == 0 ? 0 : cz - 1) : max(cz - 1, 0);
  const int z1 = periodic ? (cz == nz - 1 ? nz - 1 : cz + 1) : min(cz + 1, g.n_cells - 1);
  uint32_t lo[9], hi[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const int dx = c / 3 - 1, dy = c % 3 - 1;
    int gx = cx + dx, gy = cy + dy;
    bool ok = true;
    if (periodic) {
      if (gx < 0) gx += g.n_per; else if (gx >= g.n_per) gx -= g.n_per;
      if (gy < 0) gy += g.n_per; else if (gy >= g.n_per) gy -= g.n_per;
    } else {
      ok = gx >= 0 && gx < g.n_cells && gy >= 0 && gy < g.n_cells;
    }
    const uint32_t kb = ok ? make_key(g, gx, gy, 0) : 0u;
    lo[c] = ok ? cs(kb | (uint32_t)z0) : 0u;
    hi[c] = ok ? cs((kb | (uint32_t)z1) + 1u) : 0u;
  }
  auto walk = [&](uint32_t l, uint32_t h, float sx, float sy, float sz) {
    int base;
    if (sx == 0.0f && sy == 0.0f && sz == 0.0f && pv.find(l, h, base)) {
      for (uint32_t j = l; j < h; j++) lds((int)j, base + (int)j);
    } else {
      for (uint32_t j = l; j < h; j++) glb((int)j, sx, sy, sz);
    }
  };
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const int dx = c / 3 - 1, dy = c % 3 - 1;
    float sx = 0.0f, sy = 0.0f;
    if (periodic) {
      const int gx = cx + dx, gy = cy + dy;
      if (gx < 0) sx = -g.L; else if (gx >= g.n_per) sx = g.L;
      if (gy < 0) sy = -g.L; else if (gy >= g.n_per) sy = g.L;
    }
    if (!zwrap) {
      walk(lo[c], hi[c], sx, sy, 0.0f);
    } else {
      // cells in stencil order dz = -1, 0, +1 with the wrapped one shifted by -L or +L
      int gx = cx + dx, gy = cy + dy;
      if (gx < 0) gx += g.n_per; else if (gx >= g.n_per) gx -= g.n_per;
      if (gy < 0) gy += g.n_per; else if (gy >= g.n_per) gy -= g.n_per;
      const uint32_t kb = make_key(g, gx, gy, 0);
      if (cz == 0) {
        const uint32_t kw = kb | (uint32_t)(nz - 1);
        walk(cs(kw), cs(kw + 1u), sx, sy, -g.L);
        walk(lo[c], hi[c], sx, sy, 0.0f);            // cells 0, 1
      } else {
        walk(lo[c], hi[c], sx, sy, 0.0f);            // cells nz-2, nz-1
        walk(cs(kb), cs(kb + 1u), sx, sy, g.L);
      }
    }
  }
}

template <class FL, class FG>
__device__ __forceinline__ void visit_neighbors(const GridP& g, const uint32_t* __restrict__ cs, v3 pos,
                                                const PlanView& pv, FL lds, FG glb) {
  visit_cells(g, CsOne{cs}, pos, pv, lds, glb);
}

// Visits neighbour slots of the stencil around pos: f(j, sx, sy, sz), where
// (sx, sy, sz) is the periodic image shift of the neighbour's cell, in the
// reference's order (dx, dy, dz ascending, slots ascending within a cell).
// The three cells of a column are one slot range (a periodic column whose
// stencil wraps in z is split in two), and all nine columns' ranges load
// together before the walk, so a particle in a sparse region waits for one
// round trip for its ranges instead of one per cell.
template <class F>
__device__ __forceinline__ void for_neighbors(const GridP& g, const uint32_t* __restrict__ cs, v3 pos, F f) {
  visit_cells(g, CsOne{cs}, pos, PlanView::none(), [](int, int) {}, f);
}

// ---- wave-wide helpers (every lane of the wave calls them together) ---------
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}
__device__ __forceinline__ uint32_t wave_min_u(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
  return v;
}

// What a stencil walk reads of a neighbour from global memory: its position
// (and radius or flag word in w) and one word of flags.
struct Cand { float4 p; uint32_t w; };
// lane k's value, to every lane (k per lane; every lane of the wave executes it)
__device__ __forceinline__ Cand gather_lane(const Cand& c, int k) {
  Cand r;
  r.p.x = __shfl(c.p.x, k);
  r.p.y = __shfl(c.p.y, k);
  r.p.z = __shfl(c.p.z, k);
  r.p.w = __shfl(c.p.w, k);
  r.w = (uint32_t)__shfl((int)c.w, k);
  return r;
}
// lane k's value, to every lane (k uniform)
__device__ __forceinline__ Cand bcast(const Cand& c, int k) {
  Cand r;
  r.p.x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.p.x), k));
  r.p.y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.p.y), k));
  r.p.z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.p.z), k));
  r.p.w = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(c.p.w), k));
  r.w = (uint32_t)__builtin_amdgcn_readlane((int)c.w, k);
  return r;
}

constexpr uint32_t kCoopSpan = 4096;   // widest union of the lanes' ranges a wave walks together
constexpr int kCoopGroups = 3;         // column groups a wave walks together per stencil column
constexpr int kWalkBatch = 4;          // neighbours a lane walking on its own loads at once

// visit_cells for a whole wave (every lane calls it; live: the lane walks).
// Per stencil column, in the same order: a lane whose range is staged in LDS
// reads it there (lds(j, li)); lanes with the same centre column and image
// shift (in dense regions nearly all lanes; up to kCoopGroups groups, the first walking lane
// not yet in a group naming the next one) walk the union of their ranges
// together -- each chunk of 64 neighbours loads with one instruction, lane l
// the l-th, and every lane takes them in turn from the loading lane's
// registers (use(j, c, sx, sy, sz) for those in its own range) -- instead of
// each lane issuing a load per neighbour; the rest (a periodic column split
// at the box's z faces, a union too wide, more groups) walk their own
// ranges, loading kWalkBatch neighbours (fetch(j)) at a time.  Each lane
// meets its neighbours in its own walk's order.  sync() runs on every lane
// together after each chunk walked together and after each column.
//
// With a cheap test (test(j, c, sx, sy): false where use would do nothing
// with the neighbour), a chunk walked together first runs only the test on
// every neighbour, marking the ones that pass in a 64-bit mask per lane;
// then every lane takes its marked neighbours in order, reading each from
// the loading lane's registers (gather_lane), so that use -- the costly
// part -- runs on all lanes with work at once instead of on the few lanes
// that pass a given neighbour.  Each lane still meets its neighbours in its
// own walk's order.
struct NoSync { __device__ __forceinline__ void operator()() const {} };
struct NoTest {};
template <class CS, class FL, class FF, class FU, class FS = NoSync, class FT = NoTest>
__device__ __forceinline__ void visit_cells_wave(const GridP& g, CS cs, v3 pos, bool live, const PlanView& pv,
                                                 FL lds, FF fetch, FU use, FS sync = NoSync(), FT test = FT()) {
  if (__ballot(live) == 0) return;
  const int lane = (int)(threadIdx.x & 63u);
  const int cx = (int)grid_idx(g, pos.x), cy = (int)grid_idx(g, pos.y), cz = (int)grid_idx(g, pos.z);
  const bool periodic = g.boundary == SC_BOUNDARY_PERIODIC;
  const int nz = periodic ? g.n_per : g.n_cells;
  const bool zwrap = periodic && (cz == 0 || cz == nz - 1);
  const int z0 = periodic ? (cz == 0 ? 0 : cz - 1) : max(cz - 1, 0);
  const int z1 = periodic ? (cz == nz - 1 ? nz - 1 : cz + 1) : min(cz + 1, g.n_cells - 1);
  auto walk = [&](uint32_t l, uint32_t h, float sx, float sy, float sz) {
    for (uint32_t j0 = l; j0 < h; j0 += kWalkBatch) {
      Cand c[kWalkBatch];
#pragma unroll
      for (int q = 0; q < kWalkBatch; q++) c[q] = fetch(min(j0 + (uint32_t)q, h - 1u));
#pragma unroll
      for (int q = 0; q < kWalkBatch; q++)
        if (j0 + (uint32_t)q < h) use((int)(j0 + (uint32_t)q), c[q], sx, sy, sz);
    }
  };
#pragma unroll 1
  for (int c = 0; c < 9; c++) {
    const int dx = c / 3 - 1, dy = c % 3 - 1;
    int gx = cx + dx, gy = cy + dy;
    bool ok = live;
    float sx = 0.0f, sy = 0.0f;
    if (periodic) {
      if (gx < 0) { gx += g.n_per; sx = -g.L; } else if (gx >= g.n_per) { gx -= g.n_per; sx = g.L; }
      if (gy < 0) { gy += g.n_per; sy = -g.L; } else if (gy >= g.n_per) { gy -= g.n_per; sy = g.L; }
    } else {
      ok = ok && gx >= 0 && gx < g.n_cells && gy >= 0 && gy < g.n_cells;
    }
    const uint32_t kb = ok ? make_key(g, gx, gy, 0) : 0u;
    const uint32_t lo = ok ? cs(kb | (uint32_t)z0) : 0u;
    const uint32_t hi = ok ? cs((kb | (uint32_t)z1) + 1u) : 0u;
    int base = 0;
    const bool staged = ok && !zwrap && sx == 0.0f && sy == 0.0f && pv.find(lo, hi, base);
    bool open = ok && !zwrap && !staged && lo < hi;   // may still join a group
    bool done = false;
#pragma unroll 1
    for (int grp = 0; grp < kCoopGroups; grp++) {
      const uint64_t om = __ballot(open);
      if (om == 0) break;
      const int leader = __ffsll((unsigned long long)om) - 1;
      // the same column through the same image (lanes of two instances of a
      // batch can share a column index with different box sizes)
      const bool member = open && cx == __builtin_amdgcn_readlane(cx, leader) &&
                          cy == __builtin_amdgcn_readlane(cy, leader) &&
                          __float_as_int(sx) == __builtin_amdgcn_readlane(__float_as_int(sx), leader) &&
                          __float_as_int(sy) == __builtin_amdgcn_readlane(__float_as_int(sy), leader);
      open = open && !member;
      const uint32_t ulo = wave_min_u(member ? lo : 0xffffffffu);
      const uint32_t uhi = wave_max_u(member ? hi : 0u);
      if (uhi - ulo > kCoopSpan) continue;    // its members walk on their own
      done = done || member;
#pragma unroll 1
      for (uint32_t b0 = ulo; b0 < uhi; b0 += 64u) {
        const Cand mine = fetch(min(b0 + (uint32_t)lane, uhi - 1u));
        const int cnt = (int)min(64u, uhi - b0);
        if constexpr (std::is_same<FT, NoTest>::value) {
#pragma unroll 1
          for (int k = 0; k < cnt; k++) {
            const Cand cv = bcast(mine, k);
            const uint32_t j = b0 + (uint32_t)k;
            if (member && j >= lo && j < hi) use((int)j, cv, sx, sy, 0.0f);
          }
        } else {
          // the chunk positions inside this lane's own range, and the
          // group's image shift (the same for every member)
          const int ka = member ? (int)(lo > b0 ? min(lo - b0, 64u) : 0u) : 64;
          const int kb = member ? (int)(hi > b0 ? min(hi - b0, 64u) : 0u) : 0;
          const float gsx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sx), leader));
          const float gsy = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(sy), leader));
          const int cntu = __builtin_amdgcn_readfirstlane(cnt);
          // the lane's own chunk positions as a mask (ka <= k < kb), applied
          // once after the loop, which then only tests and sets bits
          const uint64_t own = (kb > ka ? (kb - ka == 64 ? ~0ull : ((1ull << (kb - ka)) - 1ull)) << ka : 0ull);
          uint64_t pend = 0ull;
#pragma unroll 1
          for (int k = 0; k < cntu; k++) {
            const Cand cv = bcast(mine, k);
            const uint64_t bit = 1ull << k;   // the same on every lane
            pend |= test((int)(b0 + (uint32_t)k), cv, gsx, gsy) ? bit : 0ull;
          }
          pend &= own;
#pragma unroll 1
          while (__any(pend != 0ull)) {
            const int k = pend ? __ffsll((unsigned long long)pend) - 1 : 0;
            const Cand cv = gather_lane(mine, k);
            if (pend) {
              use((int)(b0 + (uint32_t)k), cv, sx, sy, 0.0f);
              pend &= pend - 1ull;
            }
          }
        }
        sync();
      }
    }
    if (ok && !done) {   // this lane's range on its own
      if (staged) {
        for (uint32_t j = lo; j < hi; j++) lds((int)j, base + (int)j);
      } else if (!zwrap) {
        walk(lo, hi, sx, sy, 0.0f);
      } else {
        // cells in stencil order dz = -1, 0, +1 with the wrapped one shifted by -L or +L
        if (cz == 0) {
          const uint32_t kw = kb | (uint32_t)(nz - 1);
          walk(cs(kw), cs(kw + 1u), sx, sy, -g.L);
          walk(lo, hi, sx, sy, 0.0f);            // cells 0, 1
        } else {
          walk(lo, hi, sx, sy, 0.0f);            // cells nz-2, nz-1
          walk(cs(kb), cs(kb + 1u), sx, sy, g.L);
        }
      }
    }
    sync();
  }
}

// ============================================================================
// coordinateNoise (dynamics.cuh:288-402)
// ============================================================================

struct CoordArgs {
  const float4* pos; const int4* hdr; const int4* st; const int4* bonds; int cap;
  const float4* pn_c; const float4* an_c; float4* pn_n; float4* an_n;
  const uint32_t* cs; const int* work; const int* work_cnt; Counters* ctr; GridP g; float frac;
  uint2* bnb;   // bond partners per slot (Pop::bnb)
  const uint32_t* flw;   // per slot pack_flags (written by the sort)
  int build;    // first pass of the step: walk the stencil and record the partners met
};

// Positions, flags and bonds do not change between the sort and the velocity
// update, so the neighbours that are bond partners -- the only ones a pass
// reads -- are the same in every pass of a step.  The first pass walks the
// stencil and records them in walk order ({slot, the bonds they are the
// partner of}; kNoBondNbr ends the list, kBondNbrWalk in the first entry:
// more than kBondNbrs, the passes walk); the others visit only those.
constexpr uint32_t kNoBondNbr = 0xffffffffu, kBondNbrWalk = 1u << 31;

SC_GLOBAL void k_coordinate(CoordArgs a) {
  // every lane stays to the end (the first pass walks the stencil as a wave)
  const int i = pass_slot(a.work, a.work_cnt, a.ctr->n_active[0], a.g.xcd_run);
  const int i0 = i >= 0 ? i : 0;
  int4 h = a.hdr[i0];
  int nb = a.st[i0].z;
  float4 pn4 = a.pn_c[i0], an4 = a.an_c[i0];
  const bool copy = i >= 0 && (!(h.z & SC_FLAG_ACTIVE) || nb < 1);
  if (copy) { a.pn_n[i] = pn4; a.an_n[i] = an4; }
  const bool live = i >= 0 && !copy;
  nb = min(max(nb, 0), SC_MAX_BONDS);
  int btype[SC_MAX_BONDS], bpart[SC_MAX_BONDS];
  uint32_t bloom = 0u;   // bit (id mod 32) of each partner's id
#pragma unroll
  for (int k = 0; k < SC_MAX_BONDS; k++) {
    if (live && k < nb) {
      int4 b = a.bonds[(size_t)k * a.cap + i0]; btype[k] = b.x; bpart[k] = b.z;
      bloom |= 1u << (b.z & 31);
    } else { btype[k] = 0; bpart[k] = 0; }
  }
  float4 P = a.pos[i0];
  v3 ppos = xyz(P);
  float pr = P.w;
  v3 ppn = xyz(pn4);
  v4 pan = q4(an4);
  v3 dPos = m3(0.0f, 0.0f, 0.0f);
  v4 dAng = qid();
  float cw = 0.0f;
  const float frac = a.frac;
  // partner j (radius tr) of the bonds in mask
  auto partner = [&](int j, float tr, uint32_t mask) __attribute__((always_inline)) {
    bool loaded = false;
    v3 tpn = m3(0.0f, 0.0f, 0.0f);
    v4 tan_ = qid();
#pragma unroll
    for (int k = 0; k < SC_MAX_BONDS; k++) {
      if (!((mask >> k) & 1u)) continue;
      if (btype[k] < 10 && frac > 0.25f) continue;
      if (!loaded) { tpn = xyz(a.pn_c[j]); tan_ = q4(a.an_c[j]); loaded = true; }
      float pmag = len(ppn);
      float w = 1.0f;
      dPos = add(dPos, mul(div(mul(mul(sub(tpn, ppn), 0.9f), tr), pr + tr), w));
      dPos = mul(dPos, scm::fmax_(len(dPos), pmag) / len(dPos));
      dAng = slerp(dAng, qmul(qinv(pan), tan_), 0.9f * tr / (pr + tr) * w);
      cw += w;
    }
  };
  // the bond partners among the neighbours in mask order: active, the low 16
  // bits of the id match a partner's (after the cheap bit test of bloom),
  // then the whole id
  auto partner_mask = [&](int j, uint32_t fl) -> uint32_t {
    if (!(fl & 1u) || !((bloom >> ((fl >> 16) & 31u)) & 1u)) return 0u;
    uint32_t mask = 0u;
#pragma unroll
    for (int k = 0; k < SC_MAX_BONDS; k++)
      if (k < nb && (uint32_t)(bpart[k] & 0xffff) == (fl >> 16)) mask |= 1u << k;
    if (!mask) return 0u;
    const int tid = a.hdr[j].x;
    mask = 0u;
#pragma unroll
    for (int k = 0; k < SC_MAX_BONDS; k++)
      if (k < nb && bpart[k] == tid) mask |= 1u << k;
    return mask;
  };
  uint2 e0 = make_uint2(kNoBondNbr, kBondNbrWalk);
  if (live && !a.build) e0 = a.bnb[i];
  // the first pass: the wave walks the stencil together and records each
  // lane's partners in walk order; up to kBondNbrs of them, the terms are then
  // added from that record like in the later passes
  int ne = 0;
  if (a.build) {
    visit_cells_wave(grid_of(a.g, i0), CsOne{a.cs}, ppos, live, PlanView::none(), [](int, int) {},
        [&](uint32_t j) { Cand c; c.p = a.pos[j]; c.w = a.flw[j]; return c; },
        [&](int j, const Cand& c, float sx, float sy, float sz) {
          if (j == i) return;
          const uint32_t mask = partner_mask(j, c.w);
          if (!mask || outside(sub(xyz(shift4(c.p, sx, sy, sz)), ppos), a.g.cutoff)) return;
          if (ne < kBondNbrs) a.bnb[(size_t)ne * a.cap + i] = make_uint2((uint32_t)j, mask);
          ne++;
        });
    if (live) {
      if (ne > kBondNbrs) a.bnb[i] = make_uint2(kNoBondNbr, kBondNbrWalk);
      else if (ne < kBondNbrs) a.bnb[(size_t)ne * a.cap + i] = make_uint2(kNoBondNbr, 0u);
      e0 = ne > kBondNbrs ? make_uint2(kNoBondNbr, kBondNbrWalk) : (ne > 0 ? a.bnb[i] : make_uint2(kNoBondNbr, 0u));
    }
  }
  if (!live) return;
  if (e0.y & kBondNbrWalk) {
    // more partner entries than recorded: this lane walks the stencil itself
    for_neighbors(grid_of(a.g, i), a.cs, ppos, [&](int j, float sx, float sy, float sz) {
      if (j == i) return;
      float4 T = shift4(a.pos[j], sx, sy, sz);
      if (outside(sub(xyz(T), ppos), a.g.cutoff)) return;
      const uint32_t mask = partner_mask(j, a.flw[j]);
      if (mask) partner(j, T.w, mask);
    });
  } else {
    uint2 e = e0;
    for (int k = 0; k < kBondNbrs && e.x != kNoBondNbr; ) {
      partner((int)e.x, a.pos[e.x].w, e.y);
      if (++k < kBondNbrs) e = a.bnb[(size_t)k * a.cap + i];
    }
  }
  ppn = add(ppn, div(dPos, scm::fmax_(cw, 1.0f)));
  pan = qmul(slerp(qid(), dAng, 1.0f / scm::fmax_(cw, 1.0f)), pan);
  a.pn_n[i] = f4w(ppn, pn4.w);
  a.an_n[i] = f4(pan);
}

constexpr uint32_t kRelaxCap = 1536;

// periodic image shift <-> 5-bit code (kNoShift: no shift)
__device__ __forceinline__ uint32_t shift_code(float sx, float sy, float sz) {
  int ix = sx < 0.0f ? 0 : (sx > 0.0f ? 2 : 1);
  int iy = sy < 0.0f ? 0 : (sy > 0.0f ? 2 : 1);
  int iz = sz < 0.0f ? 0 : (sz > 0.0f ? 2 : 1);
  return (uint32_t)(ix * 9 + iy * 3 + iz);
}
__device__ __forceinline__ float code_shift(uint32_t code, int axis, float L) {
  int v = axis == 0 ? (int)(code / 9) : (axis == 1 ? (int)((code / 3) % 3) : (int)(code % 3));
  return v == 0 ? -L : (v == 2 ? L : 0.0f);
}
constexpr uint32_t kNoShift = 13;

// ---- pair lists ------------------------------------------------------------------
// Positions, radii, flags and hydrophobicity do not change during the relax
// passes of a step (only velocities, angular velocities and the rotations of
// bonded particles do), nor during the diffusion passes, and every pair term
// of those passes is guarded by a test on positions and radii alone.  So the
// first pass of each kind walks the 27-cell stencil once, in the reference
// order, and keeps per particle the neighbours whose guard holds; every pass
// then walks that list, in the same order, and so adds the same terms in the
// same order as a walk of the whole stencil.
// An entry is {neighbour slot, image code | kind bits}; entries past the
// kPairSlots slots go to a run of the list's overflow pool (ListWriter).
constexpr uint32_t kEntryA = 1u << 8;   // relax: polar alignment term; metabolic relax: the neighbour is a lipid
constexpr uint32_t kEntryB = 1u << 9;   // relax: collision term
// relax list of a bonded particle: bit k of the 4 bits at kEntryBondShift is
// set when the neighbour is the partner of its bond k (no collision term then)
constexpr int kEntryBondShift = 10;
constexpr uint32_t kCodeMask = 31u;
// n[i] is a per-step word for every slot of the array, so that a pass reads
// 4 bytes instead of the header, the bond state and the count: the entry
// count (kMetaCount: the list did not fit, the passes walk the stencil),
// whether the pass copies the particle (inactive) or skips it (bonded: the
// relax pass's bonded half takes it; not a lipid: no back-reaction), and for
// the relax lists whether the particle is polar and whether an entry has the
// alignment term.
constexpr int kMetaCount = 0xffff, kMetaPolar = 1 << 16, kMetaSkip = 1 << 17, kMetaCopy = 1 << 18,
              kHasAlign = 1 << 19;
// set in the relax, metabolic relax and back-reaction meta words of a
// particle whose relaxation the fused kernel runs (the per-pass kernels skip it)
constexpr int kMetaFused = kMetaFusedBit;
constexpr int kMetaBonded = kMetaSkip;
// bonded: the whole pass runs in k_bond_terms (a list that did not fit, or
// more partner terms than bonds: the same partner seen through two images)
constexpr int kMetaBondHeavy = 1 << 21;
// bonded: 4 bits at this shift, the bonds whose terms add a velocity term
// (types >= 10 and chain bonds)
constexpr int kMetaContribShift = 22;

// Writes one particle's list.  First walk: the first kPairSlots entries into
// the slots, the rest only counted.  If some lane of the wave had more,
// reserve_wave() takes one region of the block's pool shard for the whole
// wave, kPoolStride entries per row, row r holding entry kPairSlots + r of
// every lane side by side (lane l at column l), so that a pass reading entry
// k of 64 particles reads a few contiguous lines instead of 64; a second
// walk of the same stencil (same order) writes the entries past the slots
// there.  A particle whose entries do not fit (shard full, or kMetaCount or
// more) gets the count kMetaCount: its passes walk the stencil themselves.
struct ListWriter {
  const PairList& L;
  int cap, i;
  int np = 0, idx = 0, phase = 0;
  uint32_t off = 0;
  uint32_t rows0 = 0, off0 = 0;   // pool rows reserved before the first walk, and this lane's column there
  bool walk = false;
  __device__ ListWriter(const PairList& l, int c, int slot) : L(l), cap(c), i(slot) {}
  // Before the first walk, by every lane of the wave together: a region for
  // the wave's estimated overflow (est: this lane's, from the last step's
  // count of its slot -- neighbouring slots have similar lists), so that a
  // lane whose list fits in it is written by the first walk alone.
  __device__ __forceinline__ void prereserve_wave(bool live, uint32_t est) {
    uint32_t rows = live ? est : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) rows = max(rows, (uint32_t)__shfl_xor((int)rows, o));
    if (rows == 0u) return;
    const uint32_t shard = blockIdx.x % kPoolShards;
    uint32_t at = 0u;
    if ((threadIdx.x & 63u) == 0u) at = atomicAdd(&L.pool_ctr[shard], rows * kPoolStride);
    at = (uint32_t)__shfl((int)at, 0);
    if ((uint64_t)at + (uint64_t)rows * kPoolStride > L.shard_cap) return;   // the count decides afterwards
    rows0 = rows;
    off0 = shard * L.shard_cap + at + (threadIdx.x & 63u);
  }
  // the estimate from the slot's meta word of the last step (count bits)
  __device__ __forceinline__ static uint32_t estimate(int prev_meta) {
    const uint32_t c = (uint32_t)(prev_meta & kMetaCount);
    if (c <= (uint32_t)kPairSlots || c >= (uint32_t)kMetaCount) return 0u;
    const uint32_t x = c - (uint32_t)kPairSlots;
    return x + (x >> 3) + 2u;
  }
  __device__ __forceinline__ void push(uint32_t j, uint32_t w) {
    if (phase == 0) {
      if (np < kPairSlots) L.e[(size_t)np * cap + i] = make_uint2(j, w);
      else if ((uint32_t)(np - kPairSlots) < rows0) L.pool[off0 + (uint32_t)(np - kPairSlots) * kPoolStride] = make_uint2(j, w);
      np++;
    } else {
      if (idx >= kPairSlots) L.pool[off + (uint32_t)(idx - kPairSlots) * kPoolStride] = make_uint2(j, w);
      idx++;
    }
  }
  // Called by every lane of the wave at the same point (live false: a lane
  // without a list).  True for a lane whose entries past the slots got their
  // column of the wave's pool region: it walks again to write them.
  __device__ __forceinline__ bool reserve_wave(bool live) {
    const bool over = live && np > kPairSlots;
    if (over && np >= kMetaCount) walk = true;
    // written by the first walk into the pre-reserved region
    const bool fits0 = over && np < kMetaCount && (uint32_t)(np - kPairSlots) <= rows0;
    if (fits0) { off = off0; L.pool_off[i] = off0; }
    const uint32_t extra = over && np < kMetaCount && !fits0 ? (uint32_t)(np - kPairSlots) : 0u;
    uint32_t rows = extra;                                   // the wave's longest overflow
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) rows = max(rows, (uint32_t)__shfl_xor((int)rows, o));
    if (rows == 0u) return false;
    const uint32_t shard = blockIdx.x % kPoolShards;
    uint32_t at = 0u;
    if ((threadIdx.x & 63u) == 0u) at = atomicAdd(&L.pool_ctr[shard], rows * kPoolStride);   // also records the demand
    at = (uint32_t)__shfl((int)at, 0);
    if (extra == 0u) return false;
    if ((uint64_t)at + (uint64_t)rows * kPoolStride > L.shard_cap) { walk = true; return false; }
    off = shard * L.shard_cap + at + (threadIdx.x & 63u);
    L.pool_off[i] = off;
    phase = 1;
    return true;
  }
  __device__ __forceinline__ int count() const { return walk ? kMetaCount : np; }
};


// ---- worklists -------------------------------------------------------------------
// A particle without list entries gets the same update in every pass, and
// that update reaches a fixed point: velocity v -> v + 0 after one pass,
// angular velocity q -> qmul(identity, q) after two (every product in it is
// 1 * x or 0 * x, so the result depends only on which components are zero
// and their signs; all such patterns settle by the second application).
// Once both Jacobi buffers hold the fixed point, further passes change
// nothing: relax passes from the fourth on, metabolic relax and diffusion
// passes from the third on visit only the particles listed here.  The order
// of the list does not matter, as each result depends on the current
// buffers only.
constexpr int kRelaxFullPasses = 3, kMetFullPasses = 2, kDiffFullPasses = 2;

// Worklists are per block: block b lists, in ascending order, the slots of
// its own range [256 b, 256 b + 256) that need passes, at work[256 b ...],
// and their number at cnt[b]; the pass kernel runs the same grid, so a block
// only keeps as many lanes busy as it has listed slots.
// Plain relax worklists put the unbonded particles of the general kernel
// (alignment terms, stencil walkers) first, then the bonded ones (those of
// the general kernel, then those of the collision-only kernel), then the
// rest, so that the waves running the long code paths are the ones that need
// them, the general kernel takes a front of each list and the bond terms a
// contiguous middle.
enum WorkKind {
  kWorkRelaxAll = 0,   // every slot the relax kernels write: inactive (copied), active unbonded, bonded
  kWorkRelax = 1,      // active unbonded with relax entries, or lipids with back-reaction entries
  kWorkEntries = 2,    // active with entries in n0
  kWorkBonded = 3      // active with bonds
};
struct WorkArgs { const int4* hdr; const int4* st; const int* n0; const int* n1; int* work; int* cnt; Counters* ctr;
                  int pop; int kind; int* first_cnt = nullptr; int* bond_cnt = nullptr;
                  int lean_bonded = 0;   // relax: bonded collision-only particles go to the collision-only kernel
};

// A relax worklist's bonded range (per tile): the first slot and the count,
// packed as first << 16 | count (a tile has at most 256 slots).
__device__ __forceinline__ int bond_range_lo(int wb) { return wb >> 16; }
__device__ __forceinline__ int bond_range_n(int wb) { return wb & 0xffff; }

SC_GLOBAL void k_worklist(WorkArgs a) {
  __shared__ int s_bond[kBlock / 64], s_first[kBlock / 64], s_rest[kBlock / 64], s_lean[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  bool need = false, first = false, bond = false, lean = false;
  if (i < a.ctr->n_active[a.pop]) {
    const int4 h = a.hdr[i];
    const bool act = (h.z & SC_FLAG_ACTIVE) != 0;
    if (a.kind == kWorkEntries) {
      need = act && (a.n0[i] & kMetaCount) != 0 && !(a.n0[i] & kMetaFused);
    } else if (a.kind == kWorkBonded) {
      need = act && a.st[i].z > 0;
    } else {
      const int meta = a.n0[i];
      const bool free_ = !(meta & (kMetaCopy | kMetaBonded | kMetaFused));
      // bonded particles take every pass (their partner terms change in each)
      bond = (meta & kMetaBonded) && !(meta & kMetaFused);
      need = bond || (a.kind == kWorkRelaxAll ? !(meta & kMetaFused)
                                              : free_ && ((meta & kMetaCount) != 0 ||
                                                          (h.y == SC_TYPE_LIPID && (a.n1[i] & kMetaCount) != 0)));
      // the general pass kernel: alignment entries, or a list that did not
      // fit (the pass walks the stencil), and bonded particles unless
      // lean_bonded lets those whose list is collision-only (and whose
      // bond terms k_bond_terms computes) take the collision-only kernel
      const bool general = (meta & kHasAlign) != 0 || (meta & kMetaCount) == kMetaCount;
      lean = bond && a.lean_bonded && !general && !(meta & kMetaBondHeavy);
      first = need && !bond && general;
    }
  }
  // each tile's worklist: [unbonded general][bonded general][bonded lean][the rest]
  // -- the general kernel takes a front, the bond terms a contiguous middle
  const bool bgen = bond && !lean;
  const uint64_t mb = __ballot(bgen), ml = __ballot(lean), mf = __ballot(first), mr = __ballot(need && !first && !bond);
  const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
  if (lane == 0) {
    s_bond[wave] = __popcll(mb); s_lean[wave] = __popcll(ml); s_first[wave] = __popcll(mf); s_rest[wave] = __popcll(mr);
  }
  __syncthreads();
  int bb = 0, tb = 0, bl = 0, tl = 0, bf = 0, tf = 0, br = 0, tr = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; w++) {
    const int cb = s_bond[w], cl = s_lean[w], cf = s_first[w], cr = s_rest[w];
    bb += w < wave ? cb : 0;
    bl += w < wave ? cl : 0;
    bf += w < wave ? cf : 0;
    br += w < wave ? cr : 0;
    tb += cb;
    tl += cl;
    tf += cf;
    tr += cr;
  }
  const uint64_t below = (1ull << lane) - 1ull;
  if (need)
    a.work[blockIdx.x * kBlock + (first  ? bf + __popcll(mf & below)
                                  : bgen ? tf + bb + __popcll(mb & below)
                                  : lean ? tf + tb + bl + __popcll(ml & below)
                                         : tf + tb + tl + br + __popcll(mr & below))] = i;
  if (threadIdx.x == 0) {
    a.cnt[blockIdx.x] = tf + tb + tl + tr;
    if (a.first_cnt) a.first_cnt[blockIdx.x] = tf + tb;
    if (a.bond_cnt) a.bond_cnt[blockIdx.x] = (tf << 16) | (tb + tl);
  }
}


// neighbours gathered before any pair arithmetic, so that a lane waits for
// one round trip for all of them instead of one per neighbour
constexpr int kGather = 4;

// Compiler-only barrier: loads issued above it stay above it (they are not
// sunk into the branches that use them), so they are in flight together.
// It emits no instruction and no wait.
__device__ __forceinline__ void loads_issued() { asm volatile("" ::: "memory"); }

// Entry k of slot i's list: the slots, then the slot's pool run at off.
__device__ __forceinline__ const uint2* entry_at(const PairList& L, int cap, int i, int k, uint32_t off) {
  return k < kPairSlots ? L.e + (size_t)k * cap + i : L.pool + off + (uint32_t)(k - kPairSlots) * kPoolStride;
}

// Visits entries 0 .. np-1 of slot i's list in order, kGather at a time.
// head holds entries 0 .. kGather-1 (read from the slots before the count
// was known).  fetch(c, live, f) issues the gathers of the chunk's c-th
// entry f (live false: past the count; it reads an allocated element whose
// value is not used), and use(k, c, w) consumes entry k, w being its .y
// word.  The gathers of a chunk are in flight together, and the next
// chunk's entries load while they are.
template <int G = kGather, class Fetch, class Use>
__device__ __forceinline__ void for_entries(const PairList& L, int cap, int i, int np, uint32_t off,
                                            const uint2 (&head)[G], Fetch fetch, Use use) {
  uint2 f[G];
#pragma unroll
  for (int c = 0; c < G; c++) f[c] = head[c];
  for (int k0 = 0; k0 < np; k0 += G) {
#pragma unroll
    for (int c = 0; c < G; c++) fetch(c, k0 + c < np, f[c]);
    uint32_t w[G];
#pragma unroll
    for (int c = 0; c < G; c++) w[c] = f[c].y;
    const int k1 = k0 + G;
    if (k1 < np) {
#pragma unroll
      for (int c = 0; c < G; c++) f[c] = *entry_at(L, cap, i, k1 + c < np ? k1 + c : 0, off);
    }
    loads_issued();
#pragma unroll
    for (int c = 0; c < G; c++)
      if (k0 + c < np) use(k0 + c, c, w[c]);
  }
}

__device__ __forceinline__ float4 shift_img(float4 t, uint32_t code, float L) {
  if (code == kNoShift) return t;   // most pairs: no image shift (shift4 of three zeros returns t as it is)
  return shift4(t, code_shift(code, 0, L), code_shift(code, 1, L), code_shift(code, 2, L));
}

// True when |d| <= lim cannot hold: |d|^2 more than 1 % past lim^2 (far
// above the rounding of either side), so a list builder skips the square
// root of its exact test for most candidates.  False for a NaN, and for a
// negative lim, where the exact test decides.
__device__ __forceinline__ bool beyond(v3 d, float lim) { return normsq(d) > lim * lim * 1.0201f; }

// The guards of the two pair terms of an unbonded particle
// (dynamics.cuh:405-822): hydrophobic alignment of two polar particles
// closer than 1.4 (r_p + r_t) (compared in double), and the collision
// response when closer than r_p + r_t.
__device__ __forceinline__ uint32_t relax_guards(v3 ppos, float pr, bool polar, float4 T, int thyd) {
  const v3 delta = sub(xyz(T), ppos);
  const float dist = norm(delta);
  const float tr = T.w;
  uint32_t k = 0u;
  if (polar && thyd == SC_POLAR && (double)dist <= (double)(pr + tr) + 0.8 * (double)(pr + tr) * 0.5) k |= kEntryA;
  if (dist < pr + tr) k |= kEntryB;
  return k;
}

// First relax pass of a step: the pair list of every active plain particle,
// from the LDS-staged stencil (kBonds: the cell can hold bonds, and a bonded
// particle's entries mark its bond partners).
template <bool kBonds>
SC_GLOBAL void k_relax_list(RelaxArgs a) {
  __shared__ float4 s_pos[kRelaxCap];
  __shared__ uint32_t s_fl[kRelaxCap];     // pack_flags
  const int n = a.ctr->n_active[0];
  const int tile = xcd_tile(a.g.xcd_run);
  const int t0 = tile * kBlock;
  if (t0 >= n) return;
  const int t1 = min(t0 + kBlock, n);
  PlanView pv;
  pv.load(a.plans[tile], kRelaxCap);
  for (uint32_t idx = threadIdx.x; idx < pv.total; idx += kBlock) {
    const uint32_t j = pv.slot_of(idx);
    s_pos[idx] = a.pos[j];
    s_fl[idx] = a.flw[j];
  }
  __syncthreads();
  // every lane stays to the end (the pool reservation is per wave)
  const int i = t0 + threadIdx.x;
  const bool in = i < t1;
  const int j0 = in ? i : t0;
  const int4 H = a.hdr[j0];
  const int off_meta = !(H.z & SC_FLAG_ACTIVE) ? kMetaCopy : 0;
  const bool live = in && off_meta == 0;
  const float4 P = a.pos[j0];
  const v3 ppos = xyz(P);
  const float pr = P.w;
  const bool polar = H.w == SC_POLAR;
  // a bonded particle's list also marks its bond partners (its pass, in the
  // bonded half, reads their ids' terms from the marks)
  const int nbd = kBonds && live ? min(a.st[j0].z, SC_MAX_BONDS) : 0;
  int bpart[SC_MAX_BONDS];
  uint32_t cm = 0u;     // the bonds whose terms add a velocity term
  uint32_t bloom = 0u;  // bit (id mod 32) of each partner's id: most neighbours fail it at once
#pragma unroll
  for (int k = 0; k < SC_MAX_BONDS; k++) {
    bpart[k] = 0;
    if (k < nbd) {
      const int4 q = a.bonds[(size_t)k * a.cap + j0];
      bpart[k] = q.z;
      bloom |= 1u << (q.z & 31);
      if (q.x >= 10 || q.x == SC_INTERACTION_TYPE_CHAIN) cm |= 1u << k;
    }
  }
  int npart = 0, nterm = 0, walk = 0;   // partner entries, their velocity terms; the walk in progress
  ListWriter lw(a.list, a.cap, i);
  lw.prereserve_wave(live, ListWriter::estimate(a.list.n[j0]));
  bool align = false;
  // no guard holds beyond the lane's widest contact distance, r_p plus the
  // largest radius the sort saw (times 1.4 for a polar particle); a zero
  // largest radius (no sort yet) tests every neighbour exactly
  const float rmax = __uint_as_float(a.ctr->rmax_bits[0]);
  const float lb = rmax > 0.0f ? (polar ? 1.4f * (pr + rmax) : pr + rmax) : 3.0e38f;
  auto cand = [&](float4 T, uint32_t fl, int j, uint32_t code) {
    // no guard holds beyond its contact distance (r_p + r_t, or 1.4 (r_p +
    // r_t) for two polar particles), and a neighbour whose id is no
    // partner's adds no bond mark: most neighbours leave on these cheap
    // tests (the lane's widest distance first), before the cutoff test, the
    // guards' square root and double compare, and the id comparisons
    const bool maybe_partner = kBonds && ((bloom >> ((fl >> 16) & 31u)) & 1u);
    if (beyond(sub(xyz(T), ppos), lb) && !maybe_partner) return;
    const int thyd = (int)((fl >> 8) & 0xffu);
    const float rs = pr + T.w;
    const bool near = !beyond(sub(xyz(T), ppos), (polar && thyd == SC_POLAR) ? 1.4f * rs : rs);
    if (!near && !maybe_partner) return;
    if (outside(sub(xyz(T), ppos), a.g.cutoff) || !(fl & 1u)) return;
    uint32_t k = near ? relax_guards(ppos, pr, polar, T, thyd) : 0u;
    if (maybe_partner) {
      uint32_t m = 0u;
#pragma unroll
      for (int b = 0; b < SC_MAX_BONDS; b++)
        if (b < nbd && (uint32_t)(bpart[b] & 0xffff) == (fl >> 16)) m |= 1u << b;
      if (m) {   // the low bits match: the whole id decides
        const int tid = a.hdr[j].x;
        m = 0u;
#pragma unroll
        for (int b = 0; b < SC_MAX_BONDS; b++)
          if (b < nbd && bpart[b] == tid) m |= 1u << b;
      }
      if (m) {
        k = (k & ~kEntryB) | (m << kEntryBondShift);
        if (walk == 0) {     // the first walk records the partner entries, in list order
          if (npart < kBondNbrs) a.rbp[(size_t)npart * a.cap + i] = make_uint2((uint32_t)j, k | code);
          npart++;
          nterm += __popc(m & cm);
        }
      }
    }
    align = align || (k & kEntryA) != 0u;
    if (k) lw.push((uint32_t)j, k | code);
  };
  // the second walk only fills the pool column; the reservation at its
  // start is made by every lane of the wave together
  for (int phase = 0; phase < 2; phase++) {
    bool go = live;
    if (phase == 1) {
      go = lw.reserve_wave(live);
      if (!__any(go)) break;
    }
    walk = phase;
    visit_cells_wave(grid_of(a.g, j0), CsOne{a.cs}, ppos, go, pv,
        [&](int j, int li) { if (j != i) cand(s_pos[li], s_fl[li], j, kNoShift); },
        [&](uint32_t j) { Cand c; c.p = a.pos[j]; c.w = a.flw[j]; return c; },
        [&](int j, const Cand& c, float sx, float sy, float sz) {
          if (j == i) return;
          cand(shift4(c.p, sx, sy, sz), c.w, j, shift_code(sx, sy, sz));
        },
        NoSync(),
        // cand's first test, as a filter without branches: no guard holds
        // beyond lb (every radius is at most the largest the sort saw), and
        // beyond() keeps 1 % of margin, so the difference taken as
        // p - (ppos - shift) -- one instruction per axis in the loop, the
        // bracket fixed per column -- instead of cand's (p + shift) - ppos,
        // which differs by a few ulps, drops no neighbour that a guard keeps
        [&](int j, const Cand& c, float sx, float sy) {
          const bool maybe_partner = kBonds & (((bloom >> ((c.w >> 16) & 31u)) & 1u) != 0u);
          const v3 d = m3(c.p.x - (ppos.x - sx), c.p.y - (ppos.y - sy), c.p.z - ppos.z);
          return maybe_partner | !beyond(d, lb);   // the slot itself passes; use() skips it
        });
  }
  int bmeta = 0;
  if (kBonds && nbd > 0) {
    const bool heavy = lw.count() == kMetaCount || npart > kBondNbrs || nterm > SC_MAX_BONDS;
    if (!heavy && npart < kBondNbrs) a.rbp[(size_t)npart * a.cap + i] = make_uint2(kNoBondNbr, 0u);
    bmeta = kMetaBonded | (heavy ? kMetaBondHeavy : 0) | (int)(cm << kMetaContribShift);
  }
  if (in)
    a.list.n[i] = live ? (lw.count() | (polar ? kMetaPolar : 0) | (align ? kHasAlign : 0) | bmeta) : off_meta;
}
template __global__ void k_relax_list<false>(RelaxArgs);
template __global__ void k_relax_list<true>(RelaxArgs);
#define SC_RELAX_LIST(g) ((g).bonds ? k_relax_list<true> : k_relax_list<false>)

// Neighbour values a pass reads that change between passes: from the
// current global buffers.
struct GlobalNeighbours {
  const float4* vel;    // plain velocities (relax: the plain array's; metabolic relax: lipids')
  const float4* avel;
  const float4* mvel;   // metabolic velocities
  __device__ __forceinline__ float4 v(uint32_t j) const { return vel[j]; }
  __device__ __forceinline__ float4 av(uint32_t j) const { return avel[j]; }
  __device__ __forceinline__ float4 mv(uint32_t j) const { return mvel[j]; }
  // a lipid's (plain) or a metabolic particle's: one load from the selected array
  // (the pointers are copied out first: a select between the members'
  // addresses would keep the struct in memory)
  __device__ __forceinline__ float4 vsel(bool lipid, uint32_t j) const {
    const float4* const pv = vel;
    const float4* const pm = mvel;
    return (lipid ? pv : pm)[j];
  }
};

// Sums of one relax pass of an unbonded particle (dynamics.cuh:405-822 with
// an empty bond list), its own values, and the pair terms whose guards hold
// (k: kEntryA hydrophobic alignment, kEntryB collision; T: the neighbour's
// position and radius, image-shifted; trot: its rotation, for kEntryA).
struct RelaxSum {
  v3 ppos, pvel; float pr; v4 pav;
  v3 up, nextUp;
  v3 dVel, phobicPos;
  float phobicSum, cw;
  v4 dang;   // the bond terms' angular term (identity without bonds)
  __device__ __forceinline__ RelaxSum(const float4 P, const float4 vel4, const float4 avel4)
      : ppos(xyz(P)), pvel(xyz(vel4)), pr(P.w), pav(q4(avel4)), up(m3(0.0f, 0.0f, 0.0f)),
        nextUp(m3(0.0f, 0.0f, 0.0f)), dVel(m3(0.0f, 0.0f, 0.0f)), phobicPos(m3(0.0f, 0.0f, 0.0f)),
        phobicSum(0.0f), cw(0.0f), dang(qid()) {}
  // the own up vector: needed only by the alignment term and the
  // realignment that follows it (kHasAlign implies a polar particle)
  __device__ __forceinline__ void orient(const float4 rot) { up = rotate_up(q4(rot)); nextUp = rotate(up, pav); }
  // kAlign false: the particle has no alignment entries (collision terms only)
  template <bool kAlign = true>
  __device__ __forceinline__ void term(const float4 T, const v3 tvel, const float4 trot, uint32_t k) {
    const v3 tpos = xyz(T);
    const v3 delta = sub(tpos, ppos);
    const float tr = T.w;
    const v3 nd = nrm(delta);
    if (kAlign && (k & kEntryA)) {
      const v3 tup = rotate_up(q4(trot));
      const v3 nextTup = rotate(tup, pav);
      v3 pPDV = add(add(mul(neg(nextUp), pr), ppos), pvel);
      v3 tPDV = add(add(mul(neg(nextTup), tr), tpos), tvel);
      v3 dd = sub(tPDV, pPDV);
      float ap = (float)(scm::dmin_(scm::dmax_((double)(len(dd) / (pr + tr)), 1e-6), 2.0) / 2.0);
      float w = 1.0f;
      cw += w;
      dVel = add(dVel, mul(mul(mul(dd, 0.05f), ap), w));
      float pap = scm::pow_(scm::fmin_(1.0f / scm::fmax_(len(dd) / (pr + tr), 1e-6f), 10.0f) / 10.0f, 0.2f);
      phobicSum += pap;
      phobicPos = add(phobicPos, mul(tPDV, pap));
    }
    if (k & kEntryB) {
      float cd = pr + tr;
      float w = 1.0f;
      cw += w;
      dVel = add(dVel, mul(div(mul(mul(sub(add(sub(tpos, mul(nd, cd)), tvel), add(ppos, pvel)), 0.9f), tr), pr + tr), w));
    }
  }
  template <bool kAlign = true>
  __device__ __forceinline__ void finish(float vel_w, float4& vel_out, float4& avel_out) const {
    v3 nv = add(pvel, div(dVel, scm::fmax_(cw, 1.0f)));
    v4 nav = qmul(slerp(qid(), dang, 1.0f / scm::fmax_(cw, 1.0f)), pav);
    if (kAlign && phobicSum > 0.0f) {     // without alignment entries phobicSum stays 0
      v3 dir = nrm(sub(div(phobicPos, phobicSum), add(ppos, nv)));
      nav = slerp(nav, quat_from_to(neg(up), dir), 0.2f);
    }
    vel_out = f4w(nv, vel_w);
    avel_out = f4(nav);
  }
};

// One relax pass of particle i (dynamics.cuh:405-822: hydrophobic alignment
// and collisions, and for a bonded particle its bond terms, which
// k_bond_terms computed for this pass and which join the sums at their
// partner's entries, in the list order) over its pair list: from its meta
// word, position, velocity, angular velocity and first list entries (already
// loaded), the new velocity and angular velocity.
template <class NB>
__device__ __forceinline__ void relax_free_update(const RelaxArgs& a, int i, int meta, const float4 P, const float4 vel4,
                                                  const float4 avel4, const uint2 (&e)[kGather], float Lb, NB nb,
                                                  float4& vel_out, float4& avel_out) {
  const int np = meta & kMetaCount;
  const bool align = (meta & kHasAlign) != 0;
  const uint32_t cm = (meta & kMetaBonded) ? ((uint32_t)meta >> kMetaContribShift) & 0xfu : 0u;
  RelaxSum S(P, vel4, avel4);
  if (meta & kMetaBonded) S.dang = q4(a.bang[i]);
  int nt = 0;
  if (np != kMetaCount) {
    // rotations only for a particle with alignment entries (its own and its
    // neighbours'); entries past the count read the particle itself
    const float4 R = align ? a.rot_c[i] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    const uint32_t off = np > kPairSlots ? a.list.pool_off[i] : 0u;
    float4 T[kGather], V[kGather], RJ[kGather];
    for_entries(a.list, a.cap, i, np, off, e,
        [&](int c, bool live, uint2 f) {
          const size_t j = live ? f.x : (uint32_t)i;
          T[c] = a.pv_c[2 * j];        // position and velocity from one record
          V[c] = a.pv_c[2 * j + 1];
          RJ[c] = align ? a.rot_c[j] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        },
        [&](int k, int c, uint32_t w) {
          if (k == 0 && align) S.orient(R);
          S.term(shift_img(T[c], w & kCodeMask, Lb), xyz(V[c]), RJ[c], w);
          // a partner entry: its bonds' velocity terms, in bond order
          for (uint32_t m = (w >> kEntryBondShift) & cm; m; m &= m - 1u) {
            S.dVel = add(S.dVel, xyz(a.bterm[(size_t)nt * a.cap + i]));
            S.cw += 1.0f;
            nt++;
          }
        });
  } else {
    const bool polar = (meta & kMetaPolar) != 0;
    if (align) S.orient(a.rot_c[i]);
    const v3 ppos = xyz(P);
    visit_neighbors(grid_of(a.g, i), a.cs, ppos, PlanView::none(), [&](int, int) {},
        [&](int j, float sx, float sy, float sz) {
          if (j == i) return;
          const float4 T = shift4(a.pos[j], sx, sy, sz);
          const int4 h = a.hdr[j];
          if (outside(sub(xyz(T), ppos), a.g.cutoff) || !(h.z & SC_FLAG_ACTIVE)) return;
          const uint32_t k = relax_guards(ppos, P.w, polar, T, h.w & 0xff);
          if (k) S.term(T, xyz(nb.v((uint32_t)j)), a.rot_c[j], k);
        });
  }
  S.finish(vel4.w, vel_out, avel_out);
}

// One relax pass of slot i (a particle without bonds) over its pair list.
__device__ __forceinline__ void relax_free_one(const RelaxArgs& a, int i) {
  // independent loads first: one memory round trip before the branches
  const int meta = a.list.n[i];
  const float4 vel4 = a.vel_c[i], avel4 = a.avel_c[i];
  const float4 P = a.pos[i];
  uint2 e[kGather];
#pragma unroll
  for (int k = 0; k < kGather; k++) e[k] = a.list.e[(size_t)k * a.cap + i];   // slots exist: read before the count
  const float Lb = box_of(a.g, i);   // periodic image shifts of the list entries
  loads_issued();
  if (meta & kMetaCopy) { put_vel(a, i, P, vel4); a.avel_n[i] = avel4; return; }
  if (meta & (kMetaBondHeavy | kMetaFused)) return;  // k_bond_terms; the fused relaxation
  float4 nv, nav;
  relax_free_update(a, i, meta, P, vel4, avel4, e, Lb, GlobalNeighbours{a.vel_c, a.avel_c, nullptr}, nv, nav);
  put_vel(a, i, P, nv);
  a.avel_n[i] = nav;
}

// One relax pass of slot i, an unbonded particle whose list holds only
// collision entries (no polar alignment term, no stencil walk) or an
// inactive slot: the same sums as relax_free_update, without the alignment
// and realignment code (which such a particle never reaches), so the kernel
// needs far fewer registers and runs more waves per SIMD.
// G: collision entries gathered per round trip (4: fewer registers, more
// waves per SIMD; 8: more loads in flight per wave; results do not depend on it)
template <int G>
__device__ __forceinline__ void relax_coll_one(const RelaxArgs& a, int i) {
  const int meta = a.list.n[i];
  const float4 vel4 = a.vel_c[i], avel4 = a.avel_c[i];
  const float4 P = a.pos[i];
  uint2 e[G];
#pragma unroll
  for (int k = 0; k < G; k++) e[k] = a.list.e[(size_t)k * a.cap + i];
  const float Lb = box_of(a.g, i);
  loads_issued();
  if (meta & kMetaCopy) { put_vel(a, i, P, vel4); a.avel_n[i] = avel4; return; }
  if (meta & (kMetaBondHeavy | kMetaFused)) return;
  const int np = meta & kMetaCount;
  // a bonded particle (the worklist sends only collision-only lists here):
  // its bond terms, from k_bond_terms, join the sums at its partners' entries
  const uint32_t cm = (meta & kMetaBonded) ? ((uint32_t)meta >> kMetaContribShift) & 0xfu : 0u;
  RelaxSum S(P, vel4, avel4);
  if (meta & kMetaBonded) S.dang = q4(a.bang[i]);
  int nt = 0;
  const uint32_t off = np > kPairSlots ? a.list.pool_off[i] : 0u;
  float4 T[G], V[G];
  for_entries<G>(a.list, a.cap, i, np, off, e,
      [&](int c, bool live, uint2 f) {
        const size_t j = live ? f.x : (uint32_t)i;
        T[c] = a.pv_c[2 * j];
        V[c] = a.pv_c[2 * j + 1];
      },
      [&](int, int c, uint32_t w) {
        S.term<false>(shift_img(T[c], w & kCodeMask, Lb), xyz(V[c]), make_float4(0.0f, 0.0f, 0.0f, 1.0f), w);
        for (uint32_t m = (w >> kEntryBondShift) & cm; m; m &= m - 1u) {
          S.dVel = add(S.dVel, xyz(a.bterm[(size_t)nt * a.cap + i]));
          S.cw += 1.0f;
          nt++;
        }
      });
  float4 nv, nav;
  S.finish<false>(vel4.w, nv, nav);
  put_vel(a, i, P, nv);
  a.avel_n[i] = nav;
}

// slot of a k_relax_coll thread: the listed slots of its tile past the front group
__device__ __forceinline__ int pass_slot_rest(int tile, const int* work, const int* cnt, const int* first) {
  const int c = cnt[tile], f = first[tile];
  const int t = f + (int)threadIdx.x;
  const int w = work[tile * kBlock + (t < kBlock ? t : kBlock - 1)];
  return t < c ? w : -1;
}

template <int G>
SC_GLOBAL void k_relax_coll(RelaxArgs a) {
  const int i = pass_slot_rest(xcd_tile(a.g.xcd_run), a.work, a.work_cnt, a.work_first);
  if (i >= 0) relax_coll_one<G>(a, i);
}
template __global__ void k_relax_coll<4>(RelaxArgs);
template __global__ void k_relax_coll<8>(RelaxArgs);

// interactions.cuh:191-197
__device__ __forceinline__ v4 target_rel_ori(bool ord, v4 trot, v4 q) {
  return ord ? qmul(trot, qinv(q)) : qmul(trot, q);
}

// Terms of bond k of a bonded particle with its partner (dynamics.cuh:405-822,
// the bond branches): bond types >= 10 and chain bonds add the velocity
// term X (weight 1) and turn dAng; type 0 turns the particle's rotation and
// moves its debug vector.  Returns whether X was set.  Only bond terms read
// or write prot, dbg and dAng, and they read nothing the other terms write.
struct BondCtx {
  v3 ppos, pvel; v4 pav; v3 up, nextUp; float pr; int ptype, pid;
};
__device__ __forceinline__ bool bond_term(const Tables& t, const BondCtx& c, int bt, int bg, const int4 TH,
                                          const v4 trot, const v4 tav, const v3 tpos, const v3 tvel, float tr,
                                          v4& prot, v4& dbg, v4& dAng, v3& X) {
  const v3 ppos = c.ppos, pvel = c.pvel;
  const v4 pav = c.pav;
  const v3 nextUp = c.nextUp;
  const float pr = c.pr;
  const int ptype = c.ptype, pid = c.pid;
  const float idist = pr + tr;
  if (bt >= 10) {
    sc_interaction_info pii = info_at(t, bt);
    v4 ro = ro_of(pii);
    v3 rp = rp_of(pii);
    bool ord_pt = participant_order(ptype, pid, TH.y, TH.x);
    bool ord_tp = participant_order(TH.y, TH.x, ptype, pid);
    v4 rel = qmul(qinv(qmul(pav, prot)), qmul(tav, target_rel_ori(ord_pt, trot, ro)));
    v4 oth = qmul(qinv(qmul(tav, trot)), qmul(pav, target_rel_ori(ord_tp, prot, ro)));
    dAng = slerp(dAng, qmul(slerp(rel, oth, 0.5f), rel), 0.9f * tr / (pr + tr));
    v3 rrp = ord_pt ? rotate(neg(rp), qmul(qinv(ro), qmul(tav, trot))) : rotate(rp, qmul(tav, trot));
    rrp = add(rrp, tpos);
    v3 orp = ord_tp ? rotate(neg(rp), qmul(qinv(ro), qmul(pav, prot))) : rotate(rp, qmul(pav, prot));
    orp = add(orp, ppos);
    float w = 1.0f;
    v3 pDV = mul(div(mul(mul(sub(add(rrp, tvel), add(ppos, pvel)), 0.9f), tr), pr + tr), w);
    v3 tDV = mul(div(mul(mul(sub(add(orp, pvel), add(tpos, tvel)), 0.9f), pr), pr + tr), w);
    X = mul(div(mul(mul(sub(pDV, div(add(pDV, tDV), 2.0f)), 0.9f), pr), pr + tr), w);
    return true;
  }
  if (bt == 0) {
    bool ord_pt = participant_order(ptype, pid, TH.y, TH.x);
    v4 trod = qmul(quat_axis_angle(up3(), -SCV_PI / 6), quat_axis_angle(m3(1.0f, 0.0f, 0.0f), SCV_PI / 6));
    prot = slerp(prot, target_rel_ori(ord_pt, trot, trod), 0.2f);
    v4 trpr = quat_axis_angle(m3(0.0f, 0.0f, 1.0f), -SCV_PI / 6);
    v3 rr = ord_pt ? neg(rotate_up(qmul(trot, qmul(qinv(trod), trpr)))) : rotate_up(qmul(trot, trpr));
    rr = mul(rr, idist);
    rr = add(rr, tpos);
    const double k02 = 1.0 - (double)0.2f;
    dbg.x = (float)((double)dbg.x + (double)(rr.x - ppos.x) * k02);
    dbg.y = (float)((double)dbg.y + (double)(rr.y - ppos.y) * k02);
    dbg.z = (float)((double)dbg.z + (double)(rr.z - ppos.z) * k02);
    return false;
  }
  if (bt == SC_INTERACTION_TYPE_CHAIN) {
    const v3 tup = rotate_up(trot);
    const v3 nextTup = rotate(tup, pav);
    const v3 nnd = nrm(sub(add(tpos, tvel), add(ppos, pvel)));
    const float sg = (bg == SC_GROUP_FORWARD) ? -1.0f : 1.0f;
    dAng = slerp(dAng, quat_from_to(nextUp, smul(sg, nnd)), 0.9f * tr / (pr + tr));
    const float flex = SCV_PI / 8;
    v3 relUp = smul((bg == SC_GROUP_FORWARD) ? 1.0f : -1.0f, nextTup);
    v3 ax = safe_cross(relUp, neg(nnd));
    float ca = angle_vv(relUp, neg(nnd));
    v3 relaxed = rotate(relUp, quat_axis_angle(ax, scm::fmin_(scm::fmax_(ca, -flex), flex)));
    relaxed = mul(relaxed, idist);
    relaxed = add(relaxed, tpos);
    v3 oUp = smul(sg, nextUp);
    v3 oax = safe_cross(oUp, nnd);
    float oa = angle_vv(oUp, nnd);
    v3 orel = rotate(oUp, quat_axis_angle(oax, scm::fmin_(scm::fmax_(oa, -flex), flex)));
    orel = mul(orel, idist);
    orel = add(orel, ppos);
    float w = 1.0f;
    v3 pDV = mul(div(mul(mul(sub(add(relaxed, tvel), add(ppos, pvel)), 0.9f), tr), pr + tr), w);
    v3 tDV = mul(div(mul(mul(sub(add(orel, pvel), add(tpos, tvel)), 0.9f), pr), pr + tr), w);
    X = mul(div(mul(mul(sub(pDV, div(add(pDV, tDV), 2.0f)), 0.9f), pr), pr + tr), w);
    const double k09 = 1.0 - (double)0.9f;
    dbg.x = (float)((double)dbg.x + (double)(relaxed.x - ppos.x) * k09);
    dbg.y = (float)((double)dbg.y + (double)(relaxed.y - ppos.y) * k09);
    dbg.z = (float)((double)dbg.z + (double)(relaxed.z - ppos.z) * k09);
    return true;
  }
  return false;
}

// The bond state of bonded slot i that bond_term reads: its constants,
// bond types, groups and partner ids.
struct BondedSelf {
  BondCtx c; int nb; int btype[SC_MAX_BONDS], bgroup[SC_MAX_BONDS], bpart[SC_MAX_BONDS];
  v4 prot, dbg; float vel_w; int phyd;
};
__device__ __forceinline__ BondedSelf bonded_self(const RelaxArgs& a, int i) {
  BondedSelf b;
  const int4 H = a.hdr[i];
  const float4 vel4 = a.vel_c[i], avel4 = a.avel_c[i], rot4 = a.rot_c[i], dbg4 = a.dbg_c[i];
  b.nb = min(a.st[i].z, SC_MAX_BONDS);
#pragma unroll
  for (int k = 0; k < SC_MAX_BONDS; k++) {
    if (k < b.nb) { int4 q = a.bonds[(size_t)k * a.cap + i]; b.btype[k] = q.x; b.bgroup[k] = q.y; b.bpart[k] = q.z; }
    else { b.btype[k] = 0; b.bgroup[k] = 0; b.bpart[k] = 0; }
  }
  const float4 P = a.pos[i];
  b.c.ppos = xyz(P);
  b.c.pr = P.w;
  b.c.ptype = H.y; b.c.pid = H.x; b.phyd = H.w;
  b.c.pvel = xyz(vel4);
  b.c.pav = q4(avel4);
  b.prot = q4(rot4);
  b.dbg = q4(dbg4);
  b.vel_w = vel4.w;
  b.c.up = rotate_up(b.prot);
  b.c.nextUp = rotate(b.c.up, b.c.pav);
  return b;
}

// relax for particles with bonds, the whole body over the step's list (or
// the stencil when the list did not fit): the particles whose bond terms
// cannot be precomputed (k_bond_terms).
__device__ __forceinline__ void relax_bonded_one(const RelaxArgs& a, int i) {
  BondedSelf B = bonded_self(a, i);
  const v3 ppos = B.c.ppos, pvel = B.c.pvel;
  const float pr = B.c.pr;
  const v4 pav = B.c.pav;
  const v3 up = B.c.up, nextUp = B.c.nextUp;
  const int nb = B.nb, phyd = B.phyd;
  v3 dVel = m3(0.0f, 0.0f, 0.0f);
  v4 dAng = qid();
  v3 phobicPos = m3(0.0f, 0.0f, 0.0f);
  float phobicSum = 0.0f;
  float cw = 0.0f;

  // the terms of neighbour j (T: its position, image-shifted, and radius; tvel its velocity):
  // kb the guards of the hydrophobic (kEntryA) and collision (kEntryB) terms,
  // mask the bonds whose partner it is (which rule out the collision term)
  auto pair = [&](int j, const float4 T, const v3 tvel, uint32_t kb, uint32_t mask) __attribute__((always_inline)) {
    const v3 tpos = xyz(T);
    const v3 delta = sub(tpos, ppos);
    const float tr = T.w;
    const v3 nd = nrm(delta);
    if (kb & kEntryA) {
      const v3 tup = rotate_up(q4(a.rot_c[j]));
      const v3 nextTup = rotate(tup, pav);
      v3 pPDV = add(add(mul(neg(nextUp), pr), ppos), pvel);
      v3 tPDV = add(add(mul(neg(nextTup), tr), tpos), tvel);
      v3 dd = sub(tPDV, pPDV);
      float ap = (float)(scm::dmin_(scm::dmax_((double)(len(dd) / (pr + tr)), 1e-6), 2.0) / 2.0);
      float w = 1.0f;
      cw += w;
      dVel = add(dVel, mul(mul(mul(dd, 0.05f), ap), w));
      float pap = scm::pow_(scm::fmin_(1.0f / scm::fmax_(len(dd) / (pr + tr), 1e-6f), 10.0f) / 10.0f, 0.2f);
      phobicSum += pap;
      phobicPos = add(phobicPos, mul(tPDV, pap));
    }
    if (!mask) {
      if (kb & kEntryB) {
        float cd = pr + tr;
        float w = 1.0f;
        cw += w;
        dVel = add(dVel, mul(div(mul(mul(sub(add(sub(tpos, mul(nd, cd)), tvel), add(ppos, pvel)), 0.9f), tr), pr + tr), w));
      }
      return;
    }
    const int4 TH = a.hdr[j];
    const v4 trot = q4(a.rot_c[j]), tav = q4(a.avel_c[j]);
    for (uint32_t m = mask; m; m &= m - 1u) {     // the bonds of this partner in order
      const int k = __ffs(m) - 1;
      int bt = B.btype[0], bg = B.bgroup[0];
#pragma unroll
      for (int l = 1; l < SC_MAX_BONDS; l++)
        if (l == k) { bt = B.btype[l]; bg = B.bgroup[l]; }
      v3 X;
      if (bond_term(a.t, B.c, bt, bg, TH, trot, tav, tpos, tvel, tr, B.prot, B.dbg, dAng, X)) {
        cw += 1.0f;
        dVel = add(dVel, X);
      }
    }
  };

  const int meta = a.list.n[i];
  const int np = meta & kMetaCount;
  if (np != kMetaCount) {
    // the step's list: the neighbours with a term, in stencil order
    const uint32_t off = np > kPairSlots ? a.list.pool_off[i] : 0u;
    const float Lb = box_of(a.g, i);
    uint2 e[kGather];
#pragma unroll
    for (int k = 0; k < kGather; k++) e[k] = a.list.e[(size_t)k * a.cap + i];
    float4 T[kGather], V[kGather];
    uint32_t J[kGather];
    for_entries(a.list, a.cap, i, np, off, e,
        [&](int c, bool live, uint2 f) {
          J[c] = live ? f.x : (uint32_t)i;
          T[c] = a.pos[J[c]];
          V[c] = a.vel_c[J[c]];
        },
        [&](int, int c, uint32_t w) {
          pair((int)J[c], shift_img(T[c], w & kCodeMask, Lb), xyz(V[c]), w & (kEntryA | kEntryB),
               (w >> kEntryBondShift) & 0xfu);
        });
  } else {
    // the list did not fit: the stencil, cell by cell in for_neighbors' order
    const GridP g = grid_of(a.g, i);
    const int cx = (int)grid_idx(g, ppos.x), cy = (int)grid_idx(g, ppos.y), cz = (int)grid_idx(g, ppos.z);
#pragma unroll 1
    for (int c = 0; c < 27; c++) {
      uint32_t lo, hi;
      float sx, sy, sz;
      if (!stencil_range(g, a.cs, cx, cy, cz, c, lo, hi, sx, sy, sz)) continue;
#pragma unroll 1
      for (uint32_t j = lo; j < hi; j++) {
        if ((int)j == i) continue;
        const float4 T = shift4(a.pos[j], sx, sy, sz);
        if (outside(sub(xyz(T), ppos), a.g.cutoff)) continue;
        const int4 TH = a.hdr[j];
        if (!(TH.z & SC_FLAG_ACTIVE)) continue;
        uint32_t kb = relax_guards(ppos, pr, phyd == SC_POLAR, T, TH.w);
        uint32_t mask = 0u;
#pragma unroll
        for (int k = 0; k < SC_MAX_BONDS; k++)
          if (k < nb && B.bpart[k] == TH.x) mask |= 1u << k;
        if (mask) kb &= ~kEntryB;
        if (kb | mask) pair((int)j, T, xyz(a.vel_c[j]), kb, mask);
      }
    }
  }

  v3 nv = add(pvel, div(dVel, scm::fmax_(cw, 1.0f)));
  v4 nav = qmul(slerp(qid(), dAng, 1.0f / scm::fmax_(cw, 1.0f)), pav);
  if (phobicSum > 0.0f) {
    v3 dir = nrm(sub(div(phobicPos, phobicSum), add(ppos, nv)));
    nav = slerp(nav, quat_from_to(neg(up), dir), 0.2f);
  }
  put_vel(a, i, a.pos[i], f4w(nv, B.vel_w));
  a.avel_n[i] = f4(nav);
  a.rot_n[i] = f4(B.prot);
  a.dbg_n[i] = f4(B.dbg);
}

// A relax pass over every tile's worklist (split: its front group -- bonded
// particles, alignment entries, walkers --, k_relax_coll takes the rest).
// The bond terms' code lives in a kernel of its own (k_bond_terms): it needs
// far more registers, and its spills would size every wave's private memory
// here.
// W: waves per SIMD the register budget is sized for (GridP::pass_waves):
// 3 keeps every value of the list walk in registers; 4 spills a few values
// that live across the walk (stored before it, reloaded after it), for more
// waves in flight to cover the gathers' latency.  Results do not depend on it.
template <int W>
SC_GLOBAL __attribute__((amdgpu_waves_per_eu(W))) void k_relax_pass(RelaxArgs a) {
  const int i = pass_slot_at(xcd_tile(a.g.xcd_run), a.work, a.work_first ? a.work_first : a.work_cnt,
                             a.ctr->n_active[0]);
  if (i >= 0) relax_free_one(a, i);
}
template __global__ void k_relax_pass<3>(RelaxArgs);
template __global__ void k_relax_pass<4>(RelaxArgs);

// Before each relax pass of a cell that can hold bonds: for each tile's
// bonded group, the terms of every bond with its partner in list order --
// the velocity terms into bterm, the angular term into bang, the new
// rotation and debug vector -- which the pass adds at the partners' entries;
// a particle marked heavy gets its whole pass here (bond_terms_one: true).
__device__ __forceinline__ bool bond_terms_one(const RelaxArgs& a, int i) {
  if (a.list.n[i] & kMetaBondHeavy) { relax_bonded_one(a, i); return true; }
  BondedSelf B = bonded_self(a, i);
  v4 dAng = qid();
  const float Lb = box_of(a.g, i);
  int nt = 0;
  for (int q = 0; q < kBondNbrs; q++) {
    const uint2 e = a.rbp[(size_t)q * a.cap + i];
    if (e.x == kNoBondNbr) break;
    const int j = (int)e.x;
    const float4 T = shift_img(a.pos[j], e.y & kCodeMask, Lb);
    const v3 tvel = xyz(a.vel_c[j]);
    const int4 TH = a.hdr[j];
    const v4 trot = q4(a.rot_c[j]), tav = q4(a.avel_c[j]);
    // the bonds of this partner in order (one copy of the bond code: the
    // bond's type and group by selects, not a register array index)
    for (uint32_t mask = (e.y >> kEntryBondShift) & 0xfu; mask; mask &= mask - 1u) {
      const int k = __ffs(mask) - 1;
      int bt = B.btype[0], bg = B.bgroup[0];
#pragma unroll
      for (int l = 1; l < SC_MAX_BONDS; l++)
        if (l == k) { bt = B.btype[l]; bg = B.bgroup[l]; }
      v3 X;
      if (bond_term(a.t, B.c, bt, bg, TH, trot, tav, xyz(T), tvel, T.w, B.prot, B.dbg, dAng, X)) {
        a.bterm[(size_t)nt * a.cap + i] = make_float4(X.x, X.y, X.z, 0.0f);
        nt++;
      }
    }
  }
  a.bang[i] = f4(dAng);
  a.rot_n[i] = f4(B.prot);
  a.dbg_n[i] = f4(B.dbg);
  return false;
}
SC_GLOBAL __attribute__((amdgpu_waves_per_eu(2))) void k_bond_terms(RelaxArgs a) {
  const int tile = xcd_tile(a.g.xcd_run);
  const int wb = a.work_bond[tile];
  if ((int)threadIdx.x >= bond_range_n(wb)) return;
  bond_terms_one(a, a.work[tile * kBlock + bond_range_lo(wb) + (int)threadIdx.x]);
}

// A relax pass whose threads of each tile's bonded group first compute
// their own bond terms (bond_terms_one, read back by the same thread's pass):
// one launch per pass instead of two, for cells whose passes are
// launch-bound (GridP::fuse_bonds); more registers per wave than either.
SC_GLOBAL __attribute__((amdgpu_waves_per_eu(2))) void k_relax_pass_bt(RelaxArgs a) {
  const int tile = xcd_tile(a.g.xcd_run);
  const int i = pass_slot_at(tile, a.work, a.work_first ? a.work_first : a.work_cnt, a.ctr->n_active[0]);
  if (i < 0) return;
  const int wb = a.work_bond[tile];
  if ((uint32_t)((int)threadIdx.x - bond_range_lo(wb)) < (uint32_t)bond_range_n(wb) && bond_terms_one(a, i)) return;
  relax_free_one(a, i);
}
// bonded collision-only particles take the collision-only kernel: with two
// relax kernels whose bond terms come from k_bond_terms (not fuse_bonds,
// where the pass kernel computes them and so must take every bonded particle)
#define SC_LEAN_BONDED(g) ((g).relax_split && !((g).bonds && (g).fuse_bonds) ? 1 : 0)
#define SC_PASS_KERNEL(g) ((g).bonds && (g).fuse_bonds ? k_relax_pass_bt : (g).pass_waves == 4 ? k_relax_pass<4> : k_relax_pass<3>)

// ============================================================================
// Metabolic relax + lipid back-reaction (metabolites.cuh:174-350)
// ============================================================================

struct RelaxMetArgs {
  const float4* mpos; const int4* mhdr; const float4* mvel_c; float4* mvel_n;
  const float4* ppos; const int4* phdr; const float4* pvel;
  const uint32_t* mcs; const uint32_t* pcs; const float4* mix_pos; const uint32_t* mix_meta;
  const TilePlan* plans_mix; PairList list; const int* work; const int* work_cnt; int cap;
  Counters* ctr; GridP g;
};

__device__ __forceinline__ bool is_lipid(int4 h) { return (h.z & SC_FLAG_ACTIVE) && h.y == SC_TYPE_LIPID; }

// ---- merged metabolic + plain array ---------------------------------------------
// The metabolic relax visits, per stencil cell, the metabolic particles and
// then the lipids (metabolites.cuh:174-250).  The merged array holds both
// sorted arrays in exactly that order: metabolic slot i goes to i + (plain
// slots in cells below its cell) = i + pcs[key], plain slot j to j +
// (metabolic slots in cells up to its own) = j + mcs[key + 1].  Built at the
// first metabolic relax pass (positions do not change during the passes).
constexpr uint32_t kMixPlain = 0x80000000u, kMixEligible = 0x40000000u, kMixSlot = 0x3fffffffu;

struct MixArgs {
  const float4* pos; const int4* hdr; const uint32_t* K; const uint32_t* cs_self; const uint32_t* cs_other;
  float4* mix_pos; uint32_t* mix_meta; GridP g; int metabolic;
};

SC_GLOBAL void k_mix(MixArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= (int)a.cs_self[a.g.sentinel]) return;   // slots with a cell at the last sort
  const uint32_t k = a.K[i];
  const int4 h = a.hdr[i];
  uint32_t dst, meta;
  if (a.metabolic) {
    dst = (uint32_t)i + a.cs_other[k];
    meta = (uint32_t)i | ((h.z & SC_FLAG_ACTIVE) ? kMixEligible : 0u);
  } else {
    dst = (uint32_t)i + a.cs_other[k + 1u];
    meta = (uint32_t)i | kMixPlain | (is_lipid(h) ? kMixEligible : 0u);
  }
  a.mix_pos[dst] = a.pos[i];
  a.mix_meta[dst] = meta;
}

constexpr uint32_t kMixCap = 2304;

// collision distance of metabolites.cuh:195-200 (double arithmetic, as the reference)
__device__ __forceinline__ float met_cd(float pr, float tr) {
  return (float)((double)(pr + tr) + 0.8 * (double)pr);
}

// The stencil of metabolic particle i at pp3 over the merged array in the
// reference order, after the eligibility and box tests: visit(T, lipid, j, code)
// with j the slot in the particle's own array.
template <class F>
__device__ __forceinline__ void met_stencil(const RelaxMetArgs& a, int i, v3 pp3, const PlanView& pv,
                                            const float4* s_pos, const uint32_t* s_meta, F visit) {
  auto cand = [&](float4 T, uint32_t meta, uint32_t code) {
    if (!(meta & kMixEligible)) return;
    const bool lipid = (meta & kMixPlain) != 0u;
    const int j = (int)(meta & kMixSlot);
    if (!lipid && j == i) return;
    if (outside(sub(xyz(T), pp3), a.g.cutoff)) return;
    visit(T, lipid, j, code);
  };
  visit_cells(grid_of(a.g, i), CsSum{a.mcs, a.pcs}, pp3, pv,
      [&](int, int li) { cand(s_pos[li], s_meta[li], kNoShift); },
      [&](int jc, float sx, float sy, float sz) {
        cand(shift4(a.mix_pos[jc], sx, sy, sz), a.mix_meta[jc], shift_code(sx, sy, sz));
      });
}

// First metabolic relax pass of a step: the colliding neighbours of every
// active metabolic particle (metabolic weight 1, lipids weight 2).
SC_GLOBAL void k_relax_metabolic_list(RelaxMetArgs a) {
  __shared__ float4 s_pos[kMixCap];
  __shared__ uint32_t s_meta[kMixCap];
  const int n = a.ctr->n_active[1];
  const int tile = xcd_tile(a.g.xcd_run);
  const int t0 = tile * kBlock;
  if (t0 >= n) return;
  const int t1 = min(t0 + kBlock, n);
  PlanView pv;
  pv.load(a.plans_mix[tile], kMixCap);
  for (uint32_t idx = threadIdx.x; idx < pv.total; idx += kBlock) {
    const uint32_t j = pv.slot_of(idx);
    s_pos[idx] = a.mix_pos[j];
    s_meta[idx] = a.mix_meta[j];
  }
  __syncthreads();
  // every lane stays to the end (the pool reservation is per wave)
  const int i = t0 + threadIdx.x;
  const bool in = i < t1;
  const int j0 = in ? i : t0;
  const bool live = in && (a.mhdr[j0].z & SC_FLAG_ACTIVE);
  const float4 P = a.mpos[j0];
  const v3 pp3 = xyz(P);
  ListWriter lw(a.list, a.cap, i);
  lw.prereserve_wave(live, ListWriter::estimate(a.list.n[j0]));
  // the second walk only fills the pool column; the reservation at its
  // start is made by every lane of the wave together
  for (int phase = 0; phase < 2; phase++) {
    if (phase == 1 && !lw.reserve_wave(live)) break;
    if (live)
      met_stencil(a, i, pp3, pv, s_pos, s_meta, [&](float4 T, bool lipid, int j, uint32_t code) {
        const v3 d = sub(xyz(T), pp3);
        const float cd = met_cd(P.w, T.w);
        if (!beyond(d, cd) && norm(d) <= cd) lw.push((uint32_t)j, (lipid ? kEntryA : 0u) | code);
      });
  }
  if (in) a.list.n[i] = live ? lw.count() : kMetaCopy;
}

// One metabolic relax pass of metabolic particle i over its pair list
// (metabolites.cuh:174-250): neighbours are metabolic particles (weight 1)
// and lipids (weight 2, kEntryA).
// Sums of one metabolic relax or back-reaction pass (metabolites.cuh:174-350):
// collisions within met_cd with a neighbour's position T (image-shifted)
// and velocity, weighted w.
struct MetSum {
  v3 pp3, pv; float pr;
  v3 dVel; float cw;
  __device__ __forceinline__ MetSum(const float4 P, const float4 vel4)
      : pp3(xyz(P)), pv(xyz(vel4)), pr(P.w), dVel(m3(0.0f, 0.0f, 0.0f)), cw(0.0f) {}
  __device__ __forceinline__ void collide(const float4 T, const v3 tvel, float w) {
    const v3 nd = nrm(sub(xyz(T), pp3));
    const float cd = met_cd(pr, T.w);
    cw += w;
    dVel = add(dVel, mul(mul(sub(add(sub(xyz(T), mul(nd, cd)), tvel), add(pp3, pv)), 0.9f), w));
  }
  __device__ __forceinline__ float4 finish(float vel_w) const {
    return f4w(add(pv, div(dVel, scm::fmax_(cw, 1.0f))), vel_w);
  }
};

template <class NB>
__device__ __forceinline__ float4 relax_metabolic_update(const RelaxMetArgs& a, int i, int meta, const float4 P,
                                                         const float4 vel4, const uint2 (&e)[kGather], float Lb, NB nb) {
  const int np = meta & kMetaCount;
  const v3 pp3 = xyz(P);
  const float pr = P.w;
  MetSum S(P, vel4);
  auto collide = [&](float4 T, v3 tvel, float w) { S.collide(T, tvel, w); };
  if (np != kMetaCount) {
    const uint32_t off = np > kPairSlots ? a.list.pool_off[i] : 0u;
    float4 T[kGather], V[kGather];
    for_entries(a.list, a.cap, i, np, off, e,
        [&](int c, bool live, uint2 f) {
          const bool lipid = live && (f.y & kEntryA) != 0u;
          const uint32_t j = live ? f.x : (uint32_t)i;
          T[c] = (lipid ? a.ppos : a.mpos)[j];
          V[c] = nb.vsel(lipid, j);
        },
        [&](int, int c, uint32_t w) {
          collide(shift_img(T[c], w & kCodeMask, Lb), xyz(V[c]), (w & kEntryA) ? 2.0f : 1.0f);
        });
  } else {
    met_stencil(a, i, pp3, PlanView::none(), nullptr, nullptr,
        [&](float4 T, bool lipid, int j, uint32_t) {
          if (norm(sub(xyz(T), pp3)) > met_cd(pr, T.w)) return;
          if (lipid) collide(T, xyz(nb.v((uint32_t)j)), 2.0f);
          else collide(T, xyz(nb.mv((uint32_t)j)), 1.0f);
        });
  }
  return S.finish(vel4.w);
}

// One metabolic relax pass over the pair list.
SC_GLOBAL void k_relax_metabolic(RelaxMetArgs a) {
  const int i = pass_slot(a.work, a.work_cnt, a.ctr->n_active[1], a.g.xcd_run);
  if (i < 0) return;
  const float4 vel4 = a.mvel_c[i];
  const int meta = a.list.n[i];
  const float4 P = a.mpos[i];
  uint2 e[kGather];
#pragma unroll
  for (int k = 0; k < kGather; k++) e[k] = a.list.e[(size_t)k * a.cap + i];
  const float Lb = box_of(a.g, i);   // periodic image shifts of the list entries
  loads_issued();
  if (meta & kMetaCopy) { a.mvel_n[i] = vel4; return; }
  if (meta & kMetaFused) return;                    // the fused relaxation
  a.mvel_n[i] = relax_metabolic_update(a, i, meta, P, vel4, e, Lb, GlobalNeighbours{a.pvel, nullptr, a.mvel_c});
}

struct RelaxPartArgs {
  const float4* ppos; const int4* phdr; float4* pvel;
  float4* ppv;   // the plain particles' current {position, velocity} records (the velocity is updated in place)
  const float4* mpos; const int4* mhdr; const float4* mvel;
  const uint32_t* mcs; const TilePlan* plans_pm; PairList list; int cap; Counters* ctr; GridP g;
};

constexpr uint32_t kPartCap = 1600;

// First lipid back-reaction pass of a step (metabolites.cuh:266-350): the
// colliding metabolic neighbours of every active lipid.
SC_GLOBAL void k_relax_partners_list(RelaxPartArgs a) {
  __shared__ float4 s_mpos[kPartCap];
  __shared__ uint8_t s_mact[kPartCap];
  const int n = a.ctr->n_active[0];
  const int tile = xcd_tile(a.g.xcd_run);
  const int t0 = tile * kBlock;
  if (t0 >= n) return;
  const int t1 = min(t0 + kBlock, n);
  PlanView pv;
  pv.load(a.plans_pm[tile], kPartCap);
  for (uint32_t idx = threadIdx.x; idx < pv.total; idx += kBlock) {
    const uint32_t j = pv.slot_of(idx);
    s_mpos[idx] = a.mpos[j];
    s_mact[idx] = (uint8_t)(a.mhdr[j].z & SC_FLAG_ACTIVE);
  }
  __syncthreads();
  // every lane stays to the end (the pool reservation is per wave)
  const int i = t0 + threadIdx.x;
  const bool in = i < t1;
  const int j0 = in ? i : t0;
  const bool live = in && is_lipid(a.phdr[j0]);
  const float4 P = a.ppos[j0];
  const v3 pp3 = xyz(P);
  ListWriter lw(a.list, a.cap, i);
  lw.prereserve_wave(live, ListWriter::estimate(a.list.n[j0]));
  auto cand = [&](float4 T, int j, uint32_t code) {
    const v3 delta = sub(xyz(T), pp3);
    if (outside(delta, a.g.cutoff)) return;
    const float cd = met_cd(P.w, T.w);
    if (!beyond(delta, cd) && norm(delta) <= cd) lw.push((uint32_t)j, code);
  };
  // the second walk only fills the pool column; the reservation at its
  // start is made by every lane of the wave together
  for (int phase = 0; phase < 2; phase++) {
    if (phase == 1 && !lw.reserve_wave(live)) break;
    if (live)
      visit_neighbors(grid_of(a.g, i), a.mcs, pp3, pv,
          [&](int j, int li) { if (s_mact[li]) cand(s_mpos[li], j, kNoShift); },
          [&](int j, float sx, float sy, float sz) {
            if (a.mhdr[j].z & SC_FLAG_ACTIVE) cand(shift4(a.mpos[j], sx, sy, sz), j, shift_code(sx, sy, sz));
          });
  }
  if (in) a.list.n[i] = live ? lw.count() : kMetaSkip;
}

// One lipid back-reaction pass of lipid i over its pair list
// (metabolites.cuh:266-350): metabolic neighbours, weight 2.
template <class NB>
__device__ __forceinline__ float4 relax_partners_update(const RelaxPartArgs& a, int i, int meta, const float4 P,
                                                        const float4 vel4, const uint2 (&e)[kGather], float Lb, NB nb) {
  const int np = meta & kMetaCount;
  const v3 pp3 = xyz(P);
  const float pr = P.w;
  MetSum S(P, vel4);
  auto collide = [&](float4 T, float4 V) { S.collide(T, xyz(V), 2.0f); };
  if (np != kMetaCount) {
    const uint32_t off = np > kPairSlots ? a.list.pool_off[i] : 0u;
    float4 T[kGather], V[kGather];
    for_entries(a.list, a.cap, i, np, off, e,
        [&](int c, bool live, uint2 f) {
          const uint32_t j = live ? f.x : 0u;   // unused: metabolic slot 0 (allocated)
          T[c] = a.mpos[j];
          V[c] = nb.mv(j);
        },
        [&](int, int c, uint32_t w) { collide(shift_img(T[c], w & kCodeMask, Lb), V[c]); });
  } else {
    visit_neighbors(grid_of(a.g, i), a.mcs, pp3, PlanView::none(), [&](int, int) {},
        [&](int j, float sx, float sy, float sz) {
          if (!(a.mhdr[j].z & SC_FLAG_ACTIVE)) return;
          const float4 T = shift4(a.mpos[j], sx, sy, sz);
          const v3 delta = sub(xyz(T), pp3);
          if (outside(delta, a.g.cutoff) || norm(delta) > met_cd(pr, T.w)) return;
          collide(T, nb.mv((uint32_t)j));
        });
  }
  return S.finish(vel4.w);
}

// One lipid back-reaction pass over the pair list; reads only metabolic
// neighbours, so it updates the plain velocities in place.
SC_GLOBAL void k_relax_partners(RelaxPartArgs a) {
  const int i = xcd_tile(a.g.xcd_run) * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[0]) return;
  const int meta = a.list.n[i];
  const float4 P = a.ppos[i];
  const float4 vel4 = a.pvel[i];
  uint2 e[kGather];
#pragma unroll
  for (int k = 0; k < kGather; k++) e[k] = a.list.e[(size_t)k * a.cap + i];
  const float Lb = box_of(a.g, i);   // periodic image shifts of the list entries
  loads_issued();
  if (meta & (kMetaSkip | kMetaFused)) return;
  const float4 v = relax_partners_update(a, i, meta, P, vel4, e, Lb, GlobalNeighbours{nullptr, nullptr, a.mvel});
  a.pvel[i] = v;
  a.ppv[2 * (size_t)i + 1] = v;
}

// ============================================================================
// Fused relaxation.  Every relaxation pass of a particle reads only the
// particles of its lists (and, in the back-reaction, its own velocity), so
// the relaxation of a step splits along the connected components of the
// graph whose edges are the entries of the relax, metabolic relax and
// back-reaction lists.  A component without bonded particles (whose passes
// read bond partners outside the lists) that fits a block runs all passes
// in one launch, its velocities held in LDS; every member computes exactly
// the per-pass kernels' update (same device functions, same operand order),
// so the results are identical.  The rest (and everything, in a step where
// some list walks the stencil) goes through the per-pass kernels, whose
// worklists leave the fused particles out.
// ============================================================================

__device__ __forceinline__ int uf_load(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Parents always point to a smaller node, so a find ends.
__device__ __forceinline__ int uf_find(const int* parent, int x) {
  int p = uf_load(parent + x);
  while (p != x) { x = p; p = uf_load(parent + x); }
  return x;
}
// Hooks the larger root under the smaller; a failed swap means the larger
// root was hooked meanwhile, and the larger of the two roots then strictly
// decreases, so the loop ends.  The final root of a component is its
// smallest node.
__device__ __forceinline__ void uf_unite(int* parent, int a, int b) {
  int ra = uf_find(parent, a), rb = uf_find(parent, b);
  while (ra != rb) {
    if (ra < rb) { const int t = ra; ra = rb; rb = t; }
    const int old = atomicCAS(parent + ra, ra, rb);
    if (old == ra) return;
    ra = uf_find(parent, old);
    rb = uf_find(parent, rb);
  }
}

struct FuseArgs {
  Fuse f; PairList prel, ppart, mrel; int pcap, mcap; const int4* pst; Counters* ctr;
};

__device__ __forceinline__ bool fuse_node(const FuseArgs& a, int x, bool& plain, int& i) {
  plain = x < a.pcap;
  i = plain ? x : x - a.pcap;
  return x < a.f.nodes && i < a.ctr->n_active[plain ? 0 : 1];
}

SC_GLOBAL void k_fuse_init(FuseArgs a) {
  const int x = blockIdx.x * kBlock + threadIdx.x;
  if (x == 0) *a.f.any_walk = 0;
  if (x >= a.f.nodes) return;
  a.f.parent[x] = x;
  a.f.size[x] = 0;
  a.f.flags[x] = 0;
  a.f.offset[x] = -1;
  a.f.fill[x] = 0;
  a.f.loc[x] = -1;
}

// The list entries of node x (relax and back-reaction lists of a plain
// particle, the metabolic relax list of a metabolic one) join its set.  A
// node that is not fused (bonded: its passes read bond partners outside the
// lists; a list past its slots: a dense region, where components are large
// anyway) joins nothing: every node it reads through the relax relation,
// which is symmetric, lists it in turn and so joins its set; the nodes of
// its other lists (metabolic relax and back-reaction guards differ by
// direction) are marked directly.  Either way they end up with the
// per-pass kernels.
SC_GLOBAL void k_fuse_link(FuseArgs a) {
  const int x = blockIdx.x * kBlock + threadIdx.x;
  bool plain;
  int i;
  if (!fuse_node(a, x, plain, i)) return;
  const int m0 = plain ? a.prel.n[i] : a.mrel.n[i];
  const int m1 = plain ? a.ppart.n[i] : kMetaSkip;
  const int n0 = (m0 & (kMetaCopy | kMetaBonded)) ? 0 : m0 & kMetaCount;
  const int n1 = (m1 & kMetaSkip) ? 0 : m1 & kMetaCount;
  if (n0 == kMetaCount || n1 == kMetaCount) { *a.f.any_walk = 1; return; }
  const bool bad = (m0 & kMetaBonded) || n0 > kPairSlots || n1 > kPairSlots;
  if (bad) a.f.flags[x] = 1;
  // kind 0: relax list (plain neighbours); 1: back-reaction list (metabolic);
  // 2: metabolic relax list (lipids, kEntryA, or metabolic)
  auto visit = [&](const PairList& L, int cap, int np, int kind) {
    const uint32_t off = np > kPairSlots ? L.pool_off[i] : 0u;
    for (int k = 0; k < np; k++) {
      const uint2 f = *entry_at(L, cap, i, k, off);
      const bool to_plain = kind == 0 || (kind == 2 && (f.y & kEntryA) != 0u);
      const int y = to_plain ? (int)f.x : a.pcap + (int)f.x;
      if (!bad) uf_unite(a.f.parent, x, y);
      else if (kind != 0) a.f.flags[y] = 1;
    }
  };
  if (plain) {
    visit(a.prel, a.pcap, n0, 0);
    visit(a.ppart, a.pcap, n1, 1);
  } else {
    visit(a.mrel, a.mcap, n0, 2);
  }
}

SC_GLOBAL void k_fuse_roots(FuseArgs a) {
  const int x = blockIdx.x * kBlock + threadIdx.x;
  bool plain;
  int i;
  if (!fuse_node(a, x, plain, i)) return;
  const int r = uf_find(a.f.parent, x);
  a.f.parent[x] = r;      // read again only after this kernel
  atomicAdd(a.f.size + r, 1);
  if (a.f.flags[x] & 1) atomicOr(a.f.flags + r, 2);
  if (plain && (a.prel.n[i] & kHasAlign)) atomicOr(a.f.flags + r, 4);
}

// One block per kFuseTile nodes: the fusable components rooted there, in
// node order, get member slots in the tile's kFuseWaves waves, a component
// never straddling two waves (first fit in order; what does not fit stays
// with the per-pass kernels).  Components with alignment terms (software
// pow, acos and sin in every pass) come first and get waves of their own,
// so that the other waves do not wait for them.
SC_GLOBAL void k_fuse_pack(FuseArgs a) {
  __shared__ int s_want[2 * kFuseTile], s_off[kFuseTile], s_fill[kFuseWaves];   // want, then root flags
  const int t = (int)threadIdx.x;
  const int x = blockIdx.x * kFuseTile + t;
  bool plain;
  int i;
  if (t < kFuseTile) {
    int want = 0;
    int fl = 0;
    if (fuse_node(a, x, plain, i) && a.f.parent[x] == x && !*a.f.any_walk) {
      const int sz = a.f.size[x];
      fl = a.f.flags[x];
      if (sz <= kFuseWave && !(fl & 2)) want = sz;
    }
    s_want[t] = want;
    s_want[t + kFuseTile] = fl;
  }
  __syncthreads();
  if (t == 0) {
    int w = 0, fill = 0;
    for (int k = 0; k < kFuseTile; k++) s_off[k] = -1;
    for (int heavy = 1; heavy >= 0; heavy--) {
      if (heavy == 0 && fill > 0) { if (w < kFuseWaves) s_fill[w] = fill; w++; fill = 0; }
      for (int k = 0; k < kFuseTile; k++) {
        const int want = s_want[k];
        if (want == 0 || ((s_want[k + kFuseTile] & 4) != 0) != (heavy == 1)) continue;
        if (fill + want > kFuseWave) { if (w < kFuseWaves) s_fill[w] = fill; w++; fill = 0; }
        if (w >= kFuseWaves) continue;
        s_off[k] = w * kFuseWave + fill;
        fill += want;
      }
    }
    if (w < kFuseWaves) s_fill[w] = fill;
    for (int k = w + 1; k < kFuseWaves; k++) s_fill[k] = 0;
  }
  __syncthreads();
  if (t < kFuseTile && s_want[t] > 0) a.f.offset[x] = s_off[t];
  if (t < kFuseWaves) a.f.count[blockIdx.x * kFuseWaves + t] = s_fill[t];
}

// Members into their blocks; fused particles marked in their meta words.
SC_GLOBAL void k_fuse_place(FuseArgs a) {
  const int x = blockIdx.x * kBlock + threadIdx.x;
  bool plain;
  int i;
  if (!fuse_node(a, x, plain, i)) return;
  const int r = a.f.parent[x];
  const int o = a.f.offset[r];
  if (o < 0) return;
  const int idx = o + atomicAdd(a.f.fill + r, 1);
  a.f.members[(size_t)(r / kFuseTile) * kFuseCap + idx] = x;
  a.f.loc[x] = idx % kFuseWave;
  if (plain) {
    a.prel.n[i] |= kMetaFused;
    a.ppart.n[i] |= kMetaFused;
  } else {
    a.mrel.n[i] |= kMetaFused;
  }
}

struct FusedArgs {
  RelaxArgs ra; RelaxMetArgs ma; RelaxPartArgs pa; Fuse f;
  const float4* pvel_in; const float4* pavel_in; const float4* mvel_in;     // current at relax pass 0
  float4* pvel_out; float4* pavel_out; float4* mvel_out;                    // current after the relaxation
  int passes;
};

// All relaxation passes of the components packed into one wave (main.cu:
// 468-565: R relax passes, a metabolic relax and a back-reaction pass after
// every even one), one lane per member, velocities in LDS.  The first
// kGather entries of each list have their constant neighbour data
// (image-shifted position, rotation, lane) gathered once; further entries
// are read in the pass, as the per-pass kernels read them.  A member without
// list entries reaches its fixed point as in the per-pass kernels (relax
// after kRelaxFullPasses passes, metabolic relax after kMetFullPasses; the
// back-reaction of a lipid without entries is then v + 0 = v) and is skipped
// from there on.
// (three waves per SIMD: capped registers, a few spills, +7 % on the 2^20-particle cell)
__global__ __launch_bounds__(kFuseWave) __attribute__((amdgpu_waves_per_eu(3))) void k_relax_fused(FusedArgs a) {
  __shared__ float4 s_pv[2][kFuseWave], s_pa[2][kFuseWave], s_mv[2][kFuseWave];
  const int t = (int)threadIdx.x;
  const int n = a.f.count[blockIdx.x];
  if (n == 0) return;
  const int tile = (int)blockIdx.x / kFuseWaves, wave = (int)blockIdx.x % kFuseWaves;
  const int x = t < n ? a.f.members[(size_t)tile * kFuseCap + wave * kFuseWave + t] : -1;
  const bool plain = x >= 0 && x < a.f.pcap, met = x >= a.f.pcap;
  const int i = plain ? x : x - a.f.pcap;
  const int* loc_p = a.f.loc;
  const int* loc_m = a.f.loc + a.f.pcap;
  int meta0 = 0, meta1 = kMetaSkip;
  float4 P = make_float4(0.0f, 0.0f, 0.0f, 0.0f), R = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  uint2 e0[kGather], e1[kGather];
  float Lb = 0.0f;
  uint32_t off0 = 0, off1 = 0;
  if (plain) {
    meta0 = a.ra.list.n[i];
    meta1 = a.pa.list.n[i];
    P = a.ra.pos[i];
    R = a.ra.rot_c[i];
#pragma unroll
    for (int k = 0; k < kGather; k++) {
      e0[k] = a.ra.list.e[(size_t)k * a.ra.cap + i];
      e1[k] = a.pa.list.e[(size_t)k * a.pa.cap + i];
    }
    Lb = box_of(a.ra.g, i);
    s_pv[0][t] = a.pvel_in[i];
    s_pa[0][t] = a.pavel_in[i];
  } else if (met) {
    meta0 = a.ma.list.n[i];
    P = a.ma.mpos[i];
#pragma unroll
    for (int k = 0; k < kGather; k++) e0[k] = a.ma.list.e[(size_t)k * a.ma.cap + i];
    Lb = box_of(a.ma.g, i);
    s_mv[0][t] = a.mvel_in[i];
  }
  const bool copy = (meta0 & kMetaCopy) != 0;
  const int n0 = copy ? 0 : meta0 & kMetaCount;
  const int n1 = (meta1 & kMetaSkip) ? 0 : meta1 & kMetaCount;
  const bool busy = n0 > 0 || n1 > 0;
  const bool align = (meta0 & kHasAlign) != 0;
  if (n0 > kPairSlots) off0 = plain ? a.ra.list.pool_off[i] : a.ma.list.pool_off[i];
  if (n1 > kPairSlots) off1 = a.pa.list.pool_off[i];
  // entry k of a list, its neighbour's constant data
  auto relax_nb = [&](int k, uint32_t& w, float4& T, float4& RJ, int& L) {
    const uint2 f = k < kGather ? e0[k] : *entry_at(a.ra.list, a.ra.cap, i, k, off0);
    w = f.y;
    T = shift_img(a.ra.pos[f.x], f.y & kCodeMask, Lb);
    RJ = align ? a.ra.rot_c[f.x] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    L = max(loc_p[f.x], 0);
  };
  auto met_nb = [&](int k, uint32_t& w, float4& T, int& L) {
    const uint2 f = k < kGather ? e0[k] : *entry_at(a.ma.list, a.ma.cap, i, k, off0);
    const bool lipid = (f.y & kEntryA) != 0u;
    w = f.y;
    T = shift_img((lipid ? a.ma.ppos : a.ma.mpos)[f.x], f.y & kCodeMask, Lb);
    L = max((lipid ? loc_p : loc_m)[f.x], 0);
  };
  auto part_nb = [&](int k, float4& T, int& L) {
    const uint2 f = k < kGather ? e1[k] : *entry_at(a.pa.list, a.pa.cap, i, k, off1);
    T = shift_img(a.pa.mpos[f.x], f.y & kCodeMask, Lb);
    L = max(loc_m[f.x], 0);
  };
  float4 T0[kGather], R0[kGather], T1[kGather];
  int L0[kGather], L1[kGather];
  uint32_t W0[kGather];
#pragma unroll
  for (int k = 0; k < kGather; k++) {
    T0[k] = T1[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    R0[k] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    L0[k] = L1[k] = 0;
    W0[k] = 0u;
    if (k < n0) {
      if (plain) relax_nb(k, W0[k], T0[k], R0[k], L0[k]);
      else met_nb(k, W0[k], T0[k], L0[k]);
    }
    if (plain && k < n1) part_nb(k, T1[k], L1[k]);
  }
  __syncthreads();
  int cp = 0, cm = 0;
  for (int j = 0; j < a.passes; j++) {
    // relax: both buffers of a member without entries hold its fixed point
    // after the full passes
    if (plain && (busy || j < kRelaxFullPasses)) {
      float4 nv = s_pv[cp][t], na = s_pa[cp][t];
      if (!copy) {
        RelaxSum S(P, nv, na);
        if (n0 > 0 && align) S.orient(R);
#pragma unroll
        for (int k = 0; k < kGather; k++)
          if (k < n0) S.term(T0[k], xyz(s_pv[cp][L0[k]]), R0[k], W0[k]);
        for (int k = kGather; k < n0; k++) {
          uint32_t w;
          float4 T, RJ;
          int L;
          relax_nb(k, w, T, RJ, L);
          S.term(T, xyz(s_pv[cp][L]), RJ, w);
        }
        S.finish(nv.w, nv, na);
      }
      s_pv[cp ^ 1][t] = nv;
      s_pa[cp ^ 1][t] = na;
    }
    __syncthreads();
    cp ^= 1;
    if (j % 2 == 0) {
      if (met && (busy || j / 2 < kMetFullPasses)) {
        float4 nv = s_mv[cm][t];
        if (!copy) {
          MetSum S(P, nv);
#pragma unroll
          for (int k = 0; k < kGather; k++) {
            if (k >= n0) break;
            const bool lipid = (W0[k] & kEntryA) != 0u;
            S.collide(T0[k], xyz(lipid ? s_pv[cp][L0[k]] : s_mv[cm][L0[k]]), lipid ? 2.0f : 1.0f);
          }
          for (int k = kGather; k < n0; k++) {
            uint32_t w;
            float4 T;
            int L;
            met_nb(k, w, T, L);
            const bool lipid = (w & kEntryA) != 0u;
            S.collide(T, xyz(lipid ? s_pv[cp][L] : s_mv[cm][L]), lipid ? 2.0f : 1.0f);
          }
          nv = S.finish(nv.w);
        }
        s_mv[cm ^ 1][t] = nv;
      }
      __syncthreads();
      cm ^= 1;
      // in place: each lipid reads only its own velocity and metabolic ones
      if (plain && !(meta1 & kMetaSkip) && (busy || j < kRelaxFullPasses)) {
        const float4 pv = s_pv[cp][t];
        MetSum S(P, pv);
#pragma unroll
        for (int k = 0; k < kGather; k++)
          if (k < n1) S.collide(T1[k], xyz(s_mv[cm][L1[k]]), 2.0f);
        for (int k = kGather; k < n1; k++) {
          float4 T;
          int L;
          part_nb(k, T, L);
          S.collide(T, xyz(s_mv[cm][L]), 2.0f);
        }
        s_pv[cp][t] = S.finish(pv.w);
      }
      __syncthreads();
    }
  }
  if (plain) {
    a.pvel_out[i] = s_pv[cp][t];
    a.pavel_out[i] = s_pa[cp][t];
  } else if (met) {
    a.mvel_out[i] = s_mv[cm][t];
  }
}

// ============================================================================
// diffuseMetabolites (metabolites.cuh:352-431)
// ============================================================================

struct DiffArgs {
  const float4* pos; const int4* hdr;
  const float4* m0c; const float4* m1c; const float2* m2c;
  float4* m0n; float4* m1n; float2* m2n;
  const uint32_t* cs; const TilePlan* plans_mm; PairList list; const int* work; const int* work_cnt; int cap;
  Counters* ctr; GridP g;
};

constexpr uint32_t kDiffCap = 1280;

// First diffusion pass of a step: the active metabolic neighbours of every
// active metabolic particle within the diffusion distance.
SC_GLOBAL void k_diffuse_list(DiffArgs a) {
  __shared__ float4 s_pos[kDiffCap];     // w: active flag (radius is not used)
  const int n = a.ctr->n_active[1];
  const int tile = xcd_tile(a.g.xcd_run);
  const int t0 = tile * kBlock;
  if (t0 >= n) return;
  const int t1 = min(t0 + kBlock, n);
  PlanView pv;
  pv.load(a.plans_mm[tile], kDiffCap);
  for (uint32_t idx = threadIdx.x; idx < pv.total; idx += kBlock) {
    const uint32_t j = pv.slot_of(idx);
    float4 p = a.pos[j];
    p.w = __int_as_float(a.hdr[j].z & SC_FLAG_ACTIVE);
    s_pos[idx] = p;
  }
  __syncthreads();
  // every lane stays to the end (the pool reservation is per wave)
  const int i = t0 + threadIdx.x;
  const bool in = i < t1;
  const int j0 = in ? i : t0;
  const bool live = in && (a.hdr[j0].z & SC_FLAG_ACTIVE);
  const v3 pp3 = xyz(a.pos[j0]);
  const float dcut = a.g.diff_cutoff;
  ListWriter lw(a.list, a.cap, i);
  lw.prereserve_wave(live, ListWriter::estimate(a.list.n[j0]));
  auto cand = [&](float4 T, int j, uint32_t code) {
    const v3 delta = sub(xyz(T), pp3);
    if (outside(delta, dcut) || beyond(delta, dcut)) return;
    if (norm(delta) <= dcut) lw.push((uint32_t)j, code);
  };
  // the second walk only fills the pool column; the reservation at its
  // start is made by every lane of the wave together
  for (int phase = 0; phase < 2; phase++) {
    if (phase == 1 && !lw.reserve_wave(live)) break;
    if (live)
      visit_neighbors(grid_of(a.g, i), a.cs, pp3, pv,
          [&](int j, int li) {
            if (j == i) return;
            const float4 T = s_pos[li];
            if (__float_as_int(T.w)) cand(T, j, kNoShift);
          },
          [&](int j, float sx, float sy, float sz) {
            if (j == i || !(a.hdr[j].z & SC_FLAG_ACTIVE)) return;
            cand(shift4(a.pos[j], sx, sy, sz), j, shift_code(sx, sy, sz));
          });
  }
  if (in) a.list.n[i] = live ? lw.count() : kMetaCopy;
}

// One diffusion pass over the pair list.
SC_GLOBAL void k_diffuse(DiffArgs a) {
  const int i = pass_slot(a.work, a.work_cnt, a.ctr->n_active[1], a.g.xcd_run);
  if (i < 0) return;
  const float4 A = a.m0c[i], B = a.m1c[i];
  const float2 C = a.m2c[i];
  const int meta = a.list.n[i];
  uint2 e[kGather];
#pragma unroll
  for (int k = 0; k < kGather; k++) e[k] = a.list.e[(size_t)k * a.cap + i];
  const float Lb = box_of(a.g, i);   // periodic image shifts of the list entries
  loads_issued();
  if (meta & kMetaCopy) { a.m0n[i] = A; a.m1n[i] = B; a.m2n[i] = C; return; }
  const int np = meta & kMetaCount;
  float m[SC_NUM_METABOLITES] = {A.x, A.y, A.z, A.w, B.x, B.y, B.z, B.w, C.x, C.y};
  float dm[SC_NUM_METABOLITES];
#pragma unroll
  for (int k = 0; k < SC_NUM_METABOLITES; k++) dm[k] = 0.0f;
  auto exchange = [&](float4 TA, float4 TB, float2 TC) {
    const float t[SC_NUM_METABOLITES] = {TA.x, TA.y, TA.z, TA.w, TB.x, TB.y, TB.z, TB.w, TC.x, TC.y};
#pragma unroll
    for (int k = 0; k < SC_NUM_METABOLITES; k++) {
      float diff = t[k] - m[k];
      // (double)|diff| > 0.001 (metabolites.cuh:416) for a float |diff| holds exactly
      // when |diff| exceeds the largest float below 0.001 (0x3A83126E)
      if (scm::fabs_(diff) > __uint_as_float(0x3A83126Eu)) dm[k] += diff * 0.1f;
    }
  };
  if (np != kMetaCount) {
    const uint32_t off = np > kPairSlots ? a.list.pool_off[i] : 0u;
    float4 TA[kGather], TB[kGather];
    float2 TC[kGather];
    for_entries(a.list, a.cap, i, np, off, e,
        [&](int c, bool live, uint2 f) {
          const uint32_t j = live ? f.x : (uint32_t)i;
          TA[c] = a.m0c[j];
          TB[c] = a.m1c[j];
          TC[c] = a.m2c[j];
        },
        [&](int, int c, uint32_t) { exchange(TA[c], TB[c], TC[c]); });
  } else {
    const v3 pp3 = xyz(a.pos[i]);
    const float dcut = a.g.diff_cutoff;
    visit_neighbors(grid_of(a.g, i), a.cs, pp3, PlanView::none(), [&](int, int) {},
        [&](int j, float sx, float sy, float sz) {
          if (j == i || !(a.hdr[j].z & SC_FLAG_ACTIVE)) return;
          const v3 delta = sub(xyz(shift4(a.pos[j], sx, sy, sz)), pp3);
          if (outside(delta, dcut) || norm(delta) > dcut) return;
          exchange(a.m0c[j], a.m1c[j], a.m2c[j]);
        });
  }
#pragma unroll
  for (int k = 0; k < SC_NUM_METABOLITES; k++) m[k] += dm[k];
  a.m0n[i] = make_float4(m[0], m[1], m[2], m[3]);
  a.m1n[i] = make_float4(m[4], m[5], m[6], m[7]);
  a.m2n[i] = make_float2(m[8], m[9]);
}

// ============================================================================
// removeInterferingMetabolicParticles (metabolites.cuh:18-88), in place
// ============================================================================

struct RemoveArgs { const float4* mpos; int4* mhdr; const float4* ppos; const int4* phdr; const uint32_t* pcs;
                    Counters* ctr; GridP g; };

SC_GLOBAL void k_remove_interfering(RemoveArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[1]) return;
  int4 h = a.mhdr[i];
  if (!(h.z & SC_FLAG_ACTIVE)) return;
  if (a.g.table && !grid_of(a.g, i).fresh) return;     // batch: init of the instances that have not run it
  const float4 P = a.mpos[i];
  const v3 pp = xyz(P);
  int flags = h.z;
  for_neighbors(grid_of(a.g, i), a.pcs, pp, [&](int j, float sx, float sy, float sz) {
    int4 th = a.phdr[j];
    if (!(th.z & SC_FLAG_ACTIVE) || th.y != SC_TYPE_LIPID) return;
    float4 T = shift4(a.ppos[j], sx, sy, sz);
    v3 delta = sub(xyz(T), pp);
    if (outside(delta, a.g.cutoff)) return;
    float dist = norm(delta);
    float idist = (float)((double)(P.w + T.w) + 0.001);
    if (dist <= idist) flags = flags & ~SC_FLAG_ACTIVE;
  });
  h.z = flags;
  a.mhdr[i] = h;
}

// ============================================================================
// metabolite reactions (no reference counterpart), in place: per active
// metabolic particle and reaction row r in order, the fraction
// min(rate * n, 1) of m[from] moves to m[to]; n = 1, or the number of active
// plain particles of the catalyst type within catalyst_distance.  One walk
// over the plain stencil counts the catalysts of every row.
// ============================================================================

struct ReactArgs { float4* m0; float4* m1; float2* m2; const float4* mpos; const int4* mhdr;
                   const float4* ppos; const int4* phdr; const uint32_t* pcs; Counters* ctr; GridP g; Tables t; };

SC_GLOBAL void k_react(ReactArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[1]) return;
  if (!(a.mhdr[i].z & SC_FLAG_ACTIVE)) return;
  const int nrx = a.t.n_rx;
  int cnt[SC_MAX_REACTIONS];
#pragma unroll
  for (int r = 0; r < SC_MAX_REACTIONS; r++) cnt[r] = 0;
  if (a.t.rx_catalysed) {
    const v3 pp = xyz(a.mpos[i]);
    float reach = 0.0f;     // the largest catalyst distance: a neighbour outside its cube counts for no row
#pragma unroll
    for (int r = 0; r < SC_MAX_REACTIONS; r++)
      if (r < nrx && a.t.rx[r].catalyst_type >= 0) reach = scm::fmax_(reach, a.t.rx[r].catalyst_distance);
    for_neighbors(grid_of(a.g, i), a.pcs, pp, [&](int j, float sx, float sy, float sz) {
      const v3 delta = sub(xyz(shift4(a.ppos[j], sx, sy, sz)), pp);
      if (outside(delta, reach)) return;
      const int4 th = a.phdr[j];
      if (!(th.z & SC_FLAG_ACTIVE)) return;
      const float dist = norm(delta);
#pragma unroll
      for (int r = 0; r < SC_MAX_REACTIONS; r++) {
        if (r < nrx) {
          const sc_reaction rx = a.t.rx[r];
          if (rx.catalyst_type == th.y && !outside(delta, rx.catalyst_distance) && dist <= rx.catalyst_distance)
            cnt[r]++;
        }
      }
    });
  }
  const float4 q0 = a.m0[i], q1 = a.m1[i];
  const float2 q2 = a.m2[i];
  float m[SC_NUM_METABOLITES] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w, q2.x, q2.y};
#pragma unroll
  for (int r = 0; r < SC_MAX_REACTIONS; r++) {
    if (r >= nrx) continue;
    const sc_reaction rx = a.t.rx[r];
    const float n = rx.catalyst_type >= 0 ? (float)cnt[r] : 1.0f;
    const float f = scm::fmin_(rx.rate * n, 1.0f);
    float from = 0.0f;     // m[rx.from] by selects: a register array indexed at run time would spill
#pragma unroll
    for (int k = 0; k < SC_NUM_METABOLITES; k++) from = k == rx.from ? m[k] : from;
    const float flux = from * f;
#pragma unroll
    for (int k = 0; k < SC_NUM_METABOLITES; k++) {
      if (k == rx.from) m[k] = m[k] - flux;
      if (k == rx.to) m[k] = m[k] + flux;
    }
  }
  a.m0[i] = make_float4(m[0], m[1], m[2], m[3]);
  a.m1[i] = make_float4(m[4], m[5], m[6], m[7]);
  a.m2[i] = make_float2(m[8], m[9]);
}

// ============================================================================
// complexify (interactions.cuh:216-346) and transitionInteractions (:349-555)
// ============================================================================

struct CplxArgs {
  const float4* pos; const float4* rot; const int4* hdr;
  const int4* st_c; int4* st_n; const int4* bonds_c; int4* bonds_n; int cap;
  const uint32_t* cs; Counters* ctr; GridP g; Tables t; int step_override; int use_override;
  int fresh_only;       // init complexify of a batch: only instances that have not run their init
};

__device__ __forceinline__ bool state_match(const sc_interaction_info& pii, bool ord, int ps, int ts) {
  int8_t fs = (int8_t)pii.first_partner_state, ss = (int8_t)pii.second_partner_state;
  int8_t p8 = (int8_t)ps, t8 = (int8_t)ts;
  return ord ? ((fs < 0 || fs == p8) && (ss < 0 || ss == t8)) : ((ss < 0 || ss == p8) && (fs < 0 || fs == t8));
}

__device__ __forceinline__ bool group_in(const int4* bonds, int cap, int j, int nb, int group) {
  for (int k = 0; k < nb && k < SC_MAX_BONDS; k++)
    if (bonds[(size_t)k * cap + j].y == group) return true;
  return false;
}

SC_GLOBAL void k_complexify(CplxArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[0]) return;
  const int4 H = a.hdr[i];
  int4 S = a.st_c[i];
  int nb = min(max(S.z, 0), SC_MAX_BONDS);
  int4 b[SC_MAX_BONDS];
#pragma unroll
  for (int k = 0; k < SC_MAX_BONDS; k++) b[k] = (k < nb) ? a.bonds_c[(size_t)k * a.cap + i] : make_int4(0, 0, 0, 0);
  if ((H.z & SC_FLAG_ACTIVE) && (!a.fresh_only || grid_of(a.g, i).fresh)) {
    int cs = 0, ce = 0;
    if (H.y >= 0 && 2 * H.y + 1 < a.t.n_pidx) { cs = a.t.pidx[2 * H.y]; ce = a.t.pidx[2 * H.y + 1]; }
    // no walk when every row of the type bonds by site contact or only by a
    // transition: the row loop below would skip each of them
    bool any = false;
    for (int k = cs; k < ce; k++) {
      const sc_interaction_info& r = a.t.partner[k];
      any = any || (!(r.bond_probability > 0.0f) && !r.only_via_transition);
    }
    if (any) {
      const float4 P = a.pos[i];
      const v3 pp = xyz(P);
      const v4 prot = q4(a.rot[i]);
      const v3 up = rotate_up(prot);
      for_neighbors(grid_of(a.g, i), a.cs, pp, [&](int j, float sx, float sy, float sz) {
        if (j == i) return;
        float4 T = shift4(a.pos[j], sx, sy, sz);
        v3 delta = sub(xyz(T), pp);
        if (outside(delta, a.g.cutoff)) return;
        int4 TH = a.hdr[j];
        if (!(TH.z & SC_FLAG_ACTIVE)) return;
        int4 TS = a.st_c[j];
        const v4 trot = q4(a.rot[j]);
        const v3 tup = rotate_up(trot);
        for (int k = cs; k < ce; k++) {
          if (nb >= SC_MAX_BONDS) continue;
          const sc_interaction_info pii = a.t.partner[k];
          if (!((H.y == pii.first_partner_type && TH.y == pii.second_partner_type) ||
                (H.y == pii.second_partner_type && TH.y == pii.first_partner_type)) ||
              pii.only_via_transition)
            continue;
          if (pii.bond_probability > 0.0f) continue;      // site-contact rows bond in k_site_choose/commit
          if (pii.group >= 0) {
            bool act = false;
            for (int l = 0; l < nb; l++) if (b[l].y == pii.group) { act = true; break; }
            if (!act) act = group_in(a.bonds_c, a.cap, j, min(max(TS.z, 0), SC_MAX_BONDS), pii.group);
            if (act) continue;
          }
          bool ord = participant_order(H.y, H.x, TH.y, TH.x);
          v4 ro = ro_of(pii);
          v3 rp = rp_of(pii);
          float oa = ord ? dot(rotate(up, ro), tup) : dot(rotate(tup, ro), up);
          float rpa = ord ? dot(rotate(rp, prot), delta) / normsq(delta) : dot(rotate(rp, trot), neg(delta)) / normsq(delta);
          if (oa > pii.orientation_alignment_threshold &&
              scm::dabs_((double)rpa - 1.0) < (1.0 - (double)pii.position_alignment_threshold) &&
              state_match(pii, ord, S.x, TS.x)) {
            // append (bounded by the check above)
#pragma unroll
            for (int l = 0; l < SC_MAX_BONDS; l++)
              if (l == nb) b[l] = make_int4(pii.id, pii.group, TH.x, 0);
            nb++;
          }
        }
      });
    }
  }
  S.z = nb;
  a.st_n[i] = S;
#pragma unroll
  for (int k = 0; k < SC_MAX_BONDS; k++)
    if (k < nb) a.bonds_n[(size_t)k * a.cap + i] = b[k];
}

SC_GLOBAL void k_transition(CplxArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[0]) return;
  const int4 H = a.hdr[i];
  int4 S = a.st_c[i];
  int nb = min(max(S.z, 0), SC_MAX_BONDS);
  int4 b[SC_MAX_BONDS];
#pragma unroll
  for (int k = 0; k < SC_MAX_BONDS; k++) b[k] = (k < nb) ? a.bonds_c[(size_t)k * a.cap + i] : make_int4(0, 0, 0, 0);
  const int step = local_step(grid_of(a.g, i), a.ctr->step);
  // the walk changes nothing unless a bond's row transitions or breaks (and
  // only such a bond's row can be changed by it)
  bool can = false;
  for (int k = 0; k < nb; k++)
    if (b[k].x >= 10) {
      const sc_interaction_info pii = info_at(a.t, b[k].x);
      can = can || pii.transition_to >= 0 || pii.break_after_transition;
    }
  if ((H.z & SC_FLAG_ACTIVE) && can) {
    const float4 P = a.pos[i];
    const v3 pp = xyz(P);
    const v4 prot = q4(a.rot[i]);
    const v3 up = rotate_up(prot);
    int src = -1;
    sc_interaction_info tgt = info_at(a.t, -1);
    for_neighbors(grid_of(a.g, i), a.cs, pp, [&](int j, float sx, float sy, float sz) {
      if (j == i) return;
      float4 T = shift4(a.pos[j], sx, sy, sz);
      v3 delta = sub(xyz(T), pp);
      if (outside(delta, a.g.cutoff)) return;
      int4 TH = a.hdr[j];
      if (!(TH.z & SC_FLAG_ACTIVE)) return;
      bool partner = false;
      for (int k = 0; k < nb; k++) partner = partner || (b[k].x >= 10 && b[k].z == TH.x);
      if (!partner) return;
      int4 TS = a.st_c[j];
      const v4 trot = q4(a.rot[j]);
      const v3 tup = rotate_up(trot);
      for (int k = 0; k < nb; k++) {
        int4 bk = b[0];
#pragma unroll
        for (int l = 1; l < SC_MAX_BONDS; l++) if (l == k) bk = b[l];
        if (bk.x < 10) continue;
        if (bk.z != TH.x) continue;
        const sc_interaction_info pii = info_at(a.t, bk.x);
        if (!(pii.transition_to >= 0 || pii.break_after_transition)) continue;
        bool ord = participant_order(H.y, H.x, TH.y, TH.x);
        v4 ro = ro_of(pii);
        v3 rp = rp_of(pii);
        float oa = ord ? dot(rotate(up, ro), tup) : dot(rotate(tup, ro), up);
        float rpa = ord ? dot(rotate(rp, prot), delta) / normsq(delta) : dot(rotate(rp, trot), neg(delta)) / normsq(delta);
        bool aligned = pii.break_after_transition || !pii.wait_for_alignment ||
                       (oa > pii.orientation_alignment_threshold &&
                        scm::dabs_((double)rpa - 1.0) < (1.0 - (double)pii.position_alignment_threshold));
        bool stok = !pii.wait_for_state || state_match(pii, ord, S.x, TS.x);
        if (!(aligned && stok)) continue;
        if (pii.transition_to >= 0) {
          const sc_interaction_info tpii = info_at(a.t, pii.transition_to);
          if (tpii.group >= 0 && tpii.group != pii.group) {
            bool act = false;
            for (int l = 0; l < nb; l++) if (b[l].y == tpii.group) { act = true; break; }
            if (!act) act = group_in(a.bonds_c, a.cap, j, min(max(TS.z, 0), SC_MAX_BONDS), tpii.group);
            if (act) continue;
          }
          int8_t tfs = (int8_t)tpii.first_partner_state, tss = (int8_t)tpii.second_partner_state;
          if (!tpii.set_state || !((H.y == tpii.first_partner_type && tfs >= 0) || (H.y == tpii.second_partner_type && tss >= 0))) {
#pragma unroll
            for (int l = 0; l < SC_MAX_BONDS; l++) if (l == k) { b[l].x = tpii.id; b[l].y = tpii.group; }
          } else {
            src = k;
            tgt = tpii;
          }
        } else if (pii.break_after_transition) {
#pragma unroll
          for (int l = 0; l < SC_MAX_BONDS - 1; l++) if (l >= k && l < nb - 1) b[l] = b[l + 1];
          nb--;
          if (src > k) src--;
          k--;
        }
      }
    });
    if (src >= 0) {
#pragma unroll
      for (int l = 0; l < SC_MAX_BONDS; l++) if (l == src) { b[l].x = tgt.id; b[l].y = tgt.group; }
      if (tgt.set_state) {
        int8_t tfs = (int8_t)tgt.first_partner_state, tss = (int8_t)tgt.second_partner_state;
        if (H.y == tgt.first_partner_type && tfs >= 0) { S.x = tfs; S.y = step; }
        if (H.y == tgt.second_partner_type && tss >= 0) { S.x = tss; S.y = step; }
      }
      if (tgt.polymerize >= 0) {
        const sc_interaction_info poly = info_at(a.t, tgt.polymerize);
        if (poly.first_partner_type == H.y || poly.second_partner_type == H.y) S.w = tgt.polymerize;
      }
    }
  }
  S.z = nb;
  a.st_n[i] = S;
#pragma unroll
  for (int k = 0; k < SC_MAX_BONDS; k++)
    if (k < nb) a.bonds_n[(size_t)k * a.cap + i] = b[k];
}

// ============================================================================
// propagateState (interactions.cuh:743-822): a particle takes the state of a
// bonded partner (bond type >= 10 whose interaction propagates state) that
// changed state at a later step; partners in stencil order, each compared
// with the state taken so far.  Bonds do not change.
// ============================================================================
SC_GLOBAL void k_propagate_state(CplxArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[0]) return;
  const int4 H = a.hdr[i];
  int4 S = a.st_c[i];
  const int nb = min(max(S.z, 0), SC_MAX_BONDS);
  if ((H.z & SC_FLAG_ACTIVE) && nb > 0) {
    int4 b[SC_MAX_BONDS];
#pragma unroll
    for (int k = 0; k < SC_MAX_BONDS; k++) b[k] = (k < nb) ? a.bonds_c[(size_t)k * a.cap + i] : make_int4(0, 0, 0, 0);
    const v3 pp = xyz(a.pos[i]);
    for_neighbors(grid_of(a.g, i), a.cs, pp, [&](int j, float sx, float sy, float sz) {
      if (j == i) return;
      const int4 TH = a.hdr[j];
      if (!(TH.z & SC_FLAG_ACTIVE)) return;
      if (outside(sub(xyz(shift4(a.pos[j], sx, sy, sz)), pp), a.g.cutoff)) return;
      const int4 TS = a.st_c[j];
#pragma unroll
      for (int k = 0; k < SC_MAX_BONDS; k++) {
        if (k >= nb || b[k].x < 10 || b[k].z != TH.x) continue;
        if (!info_at(a.t, b[k].x).propagate_state) continue;
        if (TS.y > S.y) { S.y = TS.y; S.x = TS.x; }
      }
    });
  }
  a.st_n[i] = S;
}

// ============================================================================
// polymerize (interactions.cuh:558-740): controllers processed serially in
// slot order, new monomers appended at slots n, n+1, ... with ids
// nextParticleId, nextParticleId+1, ...
// ============================================================================

struct PolyArgs {
  float4* pos; float4* vel; float4* rot; float4* avel; float4* pnoise; float4* anoise; float4* dbg;
  int4* hdr; int4* st; int4* bonds; int cap;
  int* list; const uint32_t* cs; Counters* ctr; GridP g; Tables t; int capacity;
  int* inst; InstCtr* ictr;   // batch: instance of each slot, per-instance ids and capacities
};

SC_GLOBAL void k_poly_collect(PolyArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[0]) return;
  if (!(a.hdr[i].z & SC_FLAG_ACTIVE)) return;
  if (a.st[i].w < 0) return;
  int slot = atomicAdd(&a.ctr->n_poly, 1);
  a.list[slot] = i;
}

__device__ void remove_group_bond(int4* bonds, int cap, int j, int4& st, int group) {
  int nb = min(max(st.z, 0), SC_MAX_BONDS);
  for (int k = 0; k < nb; k++) {
    if (bonds[(size_t)k * cap + j].y == group) {
      for (int l = k; l < nb - 1; l++) bonds[(size_t)l * cap + j] = bonds[(size_t)(l + 1) * cap + j];
      st.z = nb - 1;
      return;
    }
  }
}
__device__ void append_bond(int4* bonds, int cap, int j, int4& st, int type, int group, int partner) {
  int nb = min(max(st.z, 0), SC_MAX_BONDS);
  if (nb >= SC_MAX_BONDS) return;
  bonds[(size_t)nb * cap + j] = make_int4(type, group, partner, 0);
  st.z = nb + 1;
}

__global__ __launch_bounds__(64) void k_poly_serial(PolyArgs a) {
  if (threadIdx.x != 0) return;
  const int count = a.ctr->n_poly;
  if (count == 0) return;
  int* L = a.list;
  for (int x = 1; x < count; x++) {          // insertion sort: slot order
    int v = L[x], y = x - 1;
    while (y >= 0 && L[y] > v) { L[y + 1] = L[y]; y--; }
    L[y + 1] = v;
  }
  int nact = a.ctr->n_active[0];
  int last = nact - 1;
  int nextId = a.ctr->next_particle_id;
  for (int c = 0; c < count; c++) {
    const int i = L[c];
    int4 H = a.hdr[i];
    int4 S = a.st[i];
    if (!(H.z & SC_FLAG_ACTIVE) || S.w < 0) continue;
    // a batch counts ids and capacity per instance (the new slot is the next free one)
    InstCtr* ic = a.ictr ? &a.ictr[a.inst[i]] : nullptr;
    const int step = local_step(grid_of(a.g, i), a.ctr->step);
    const int sched = S.w;
    const sc_interaction_info poly = info_at(a.t, sched);
    int nb = min(max(S.z, 0), SC_MAX_BONDS);
    int activeIdx = -1;
    for (int k = 0; k < nb; k++) if (a.bonds[(size_t)k * a.cap + i].y == poly.group) { activeIdx = k; break; }
    int prev = -1;
    const float4 P = a.pos[i];
    if (activeIdx >= 0) {
      const int target = a.bonds[(size_t)activeIdx * a.cap + i].z;
      for_cells(grid_of(a.g, i), xyz(P), [&](uint32_t key, float, float, float) {
        const uint32_t cs0 = a.cs[key], cs1 = a.cs[key + 1u];
        for (uint32_t j = cs0; j < cs1; j++) {
          if ((int)j == i) continue;
          int4 th = a.hdr[j];
          if (!(th.z & SC_FLAG_ACTIVE)) continue;
          if (th.x == target) prev = (int)j;
        }
      });
    }
    const int ptype = H.y;
    const int newType = (poly.first_partner_type == ptype) ? poly.second_partner_type : poly.first_partner_type;
    const int newState = (int8_t)((poly.first_partner_type == ptype) ? poly.second_partner_state : poly.first_partner_state);
    if (newType >= 0) {
      bool room;
      int ni;
      if (ic) {
        room = ic->n_plain < ic->cap_plain;
        ni = room ? ++last : -1;
      } else {
        ni = ++last;
        room = ni < a.capacity;
      }
      if (room) {
        const int nid = ic ? ic->next_particle_id++ : nextId++;
        if (ic) { ic->n_plain++; a.inst[ni] = a.inst[i]; }
        float radius = 0.0f;
        int hyd = 0;
        if (newType < a.t.n_types) { radius = a.t.types[newType].radius; hyd = (int8_t)a.t.types[newType].hydrophobic; }
        v3 rp = rp_of(poly);
        v4 ro = ro_of(poly);
        v4 prot = q4(a.rot[i]);
        v3 np = add(xyz(P), rotate(ptype < newType ? rp : neg(rp), prot));
        v4 nr = qmul(prot, ptype < newType ? ro : qinv(ro));
        a.pos[ni] = f4w(np, radius);
        a.rot[ni] = f4(nr);
        a.vel[ni] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.pnoise[ni] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.avel[ni] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        a.anoise[ni] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        a.dbg[ni] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.hdr[ni] = make_int4(nid, newType, SC_FLAG_ACTIVE, hyd);
        int4 NS = make_int4(0, -1, 0, -1);
        if (newState >= 0) { NS.x = newState; NS.y = step; }
        append_bond(a.bonds, a.cap, ni, NS, sched, poly.group, H.x);
        if (prev >= 0) {
          int4 PS = a.st[prev];
          remove_group_bond(a.bonds, a.cap, prev, PS, poly.group);
          append_bond(a.bonds, a.cap, prev, PS, SC_INTERACTION_TYPE_CHAIN, SC_GROUP_FORWARD, nid);
          a.st[prev] = PS;
          append_bond(a.bonds, a.cap, ni, NS, SC_INTERACTION_TYPE_CHAIN, SC_GROUP_BACKWARD, a.hdr[prev].x);
          int4 bb = a.bonds[(size_t)activeIdx * a.cap + i];
          bb.x = sched;
          bb.z = nid;
          a.bonds[(size_t)activeIdx * a.cap + i] = bb;
        } else {
          append_bond(a.bonds, a.cap, i, S, sched, poly.group, nid);
        }
        a.st[ni] = NS;
        nact++;
      }
    } else {
      if (prev >= 0) {
        int4 PS = a.st[prev];
        remove_group_bond(a.bonds, a.cap, prev, PS, poly.group);
        a.st[prev] = PS;
      }
      if (activeIdx >= 0) {
        for (int l = activeIdx; l < nb - 1; l++) a.bonds[(size_t)l * a.cap + i] = a.bonds[(size_t)(l + 1) * a.cap + i];
        S.z = nb - 1;
      }
    }
    S.w = -1;
    a.st[i] = S;
  }
  a.ctr->n_active[0] = nact;
  a.ctr->next_particle_id = nextId;
  a.ctr->n_poly = 0;
}

SC_GLOBAL void k_end_step(Counters* ctr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) ctr->step += 1;
}

// ============================================================================
// Frame reduction (dynamics.cuh:824-844, types.cuh:238-294)
// ============================================================================

struct ReduceArgs { const float4* pos; const float4* rot; const float4* dbg; const int4* hdr; const int4* st;
                    const float4* m0; uint32_t* out; Counters* ctr; int pop; int cap; int words; };

__device__ __forceinline__ uint32_t pack2(float a, float b) {
  return (uint32_t)scm::f2h_(a) | ((uint32_t)scm::f2h_(b) << 16);
}

SC_GLOBAL void k_reduce(ReduceArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.cap) return;
  uint32_t* r = a.out + (size_t)i * a.words;
  if (i >= a.ctr->n_active[a.pop]) {
    for (int k = 0; k < a.words; k++) r[k] = 0u;
    return;
  }
  int4 h = a.hdr[i];
  int state = a.st[i].x;
  float4 p = a.pos[i], q = a.rot[i], d = a.dbg[i];
  r[0] = (uint32_t)h.x;
  r[1] = (uint32_t)h.y;
  r[2] = (uint32_t)(h.z | ((state & 0x0000000F) << 16));
  r[3] = pack2(p.w, p.x);
  r[4] = pack2(p.y, p.z);
  r[5] = pack2(q.x, q.y);
  r[6] = pack2(q.z, q.w);
  r[7] = pack2(d.x, d.y);
  r[8] = pack2(d.z, d.w);
  if (a.words == 11) {
    float4 m = a.m0[i];
    r[9] = pack2(m.x, m.y);
    r[10] = pack2(m.z, m.w);
  }
}

// ============================================================================
// Parity export: neighbour pairs in enumeration order
// ============================================================================

struct NbrArgs { const float4* pos; const int4* hdr; const uint32_t* cs; Counters* ctr; GridP g; int pop;
                 uint32_t* counts; const uint32_t* offsets; int32_t* out; };

SC_GLOBAL void k_neighbors(NbrArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int n = a.ctr->n_active[a.pop];
  if (i >= n) return;
  if (!(a.hdr[i].z & SC_FLAG_ACTIVE)) { if (a.counts) a.counts[i] = 0u; return; }
  const v3 pp = xyz(a.pos[i]);
  uint32_t c = 0;
  uint32_t o = a.offsets ? a.offsets[i] : 0u;
  for_neighbors(grid_of(a.g, i), a.cs, pp, [&](int j, float sx, float sy, float sz) {
    if (j == i) return;
    float4 T = shift4(a.pos[j], sx, sy, sz);
    if (outside(sub(xyz(T), pp), a.g.cutoff)) return;
    if (!(a.hdr[j].z & SC_FLAG_ACTIVE)) return;
    if (a.out) { a.out[2 * (size_t)(o + c)] = i; a.out[2 * (size_t)(o + c) + 1] = j; }
    c++;
  });
  if (a.counts) a.counts[i] = c;
}

// Copies the fields of slots[0, m) of one array into compact buffers
// (out[f][t] = field f of slot slots[t]; bonds at [k * m + t]).
// Rows [0, m) of the gathered fields (bonds k of row t at [k * stride + t]);
// with m_dev set, m = min(*m_dev, stride), read on the device.
struct GatherArgs { Pop p; const int* slots; int m; void* out[F_COUNT]; const int* m_dev; int stride; };

SC_GLOBAL void k_gather(GatherArgs a) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int m = a.m_dev ? min(*a.m_dev, a.stride) : a.m;
  if (t >= m) return;
  const int j = a.slots[t];
  const Pop& P = a.p;
#define SC_PICK(F, T) ((T*)a.out[F])[t] = P.c<T>(F)[j]
  SC_PICK(F_POS, float4); SC_PICK(F_VEL, float4); SC_PICK(F_ROT, float4); SC_PICK(F_AVEL, float4);
  SC_PICK(F_PNOISE, float4); SC_PICK(F_ANOISE, float4); SC_PICK(F_DBG, float4);
  SC_PICK(F_HDR, int4); SC_PICK(F_ST, int4); SC_PICK(F_INST, int);
  if (P.metabolic) {
    SC_PICK(F_MET0, float4); SC_PICK(F_MET1, float4); SC_PICK(F_MET2, float2);
  } else {
    const int4* b = P.c<int4>(F_BONDS);
    for (int k = 0; k < SC_MAX_BONDS; k++) ((int4*)a.out[F_BONDS])[(size_t)k * a.stride + t] = b[(size_t)k * P.cap + j];
  }
#undef SC_PICK
}

SC_GLOBAL void k_fill_u32(uint32_t* dst, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  size_t stride = (size_t)gridDim.x * kBlock;
  for (; i < n; i += stride) dst[i] = v;
}

// ============================================================================
// Scalar primitives evaluated on the device (host/device agreement checks)
// ============================================================================

__device__ __host__ float eval_primitive(int fn, float x) {
  switch (fn) {
    case 0: return scm::sin_(x);
    case 1: return scm::cos_(x);
    case 2: return scm::acos_(x);
    case 3: return scm::log_(x);
    case 4: return scm::log2_(x);
    case 5: return scm::exp2_(x);
    case 6: return scm::pow_(x, 0.2f);
    case 7: return scm::sqrt_(x);
    case 8: return scm::u2f((uint32_t)scm::f2h_(x));
    case 9: {
      uint32_t b = scm::f2u(x);
      return scm::box_muller_(scm::u01_(b), scm::u01_(b * 2654435761u));
    }
    case 10: {
      scm::u32x4 c = {scm::f2u(x), 1u, 2u, 3u};
      scm::u32x4 r = scm::philox4x32_10(c, 42u, 0u);
      return scm::u2f(r.x ^ r.y ^ r.z ^ r.w);
    }
    case 11: {
      uint32_t b = scm::f2u(x);
      v4 q = shoemake(scm::u01_(b), scm::u01_(b * 747796405u), scm::u01_(b * 2891336453u));
      v3 u = rotate_up(q);
      return u.x + u.y * 3.0f + u.z * 7.0f;
    }
    default: return 0.0f;
  }
}

SC_GLOBAL void k_eval_primitive(int fn, const float* x, float* y, size_t n) {
  size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) y[i] = eval_primitive(fn, x[i]);
}

// The grid constants a kernel over array p gets: in a batch, with p's slot
// instances (the table is shared).
static GridP pop_grid(const GridP& g, const Pop& p) {
  GridP r = g;
  if (g.table) r.inst = p.c<int>(F_INST);
  return r;
}

void launch_eval_primitive(int fn, const float* x, float* y, size_t n, hipStream_t s) {
  if (n == 0) return;
  SC_LAUNCH(k_eval_primitive, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, fn, x, y, n);
}

// ============================================================================
// launch wrappers
// ============================================================================

static inline unsigned blocks_for(size_t n) { return n == 0 ? 1u : (unsigned)((n + kBlock - 1) / kBlock); }

// ---- the slots of one instance of a batch, in slot order ------------------------
// Three passes over the first n slots: per 256-slot tile the number of the
// instance's slots, an exclusive scan of those counts by one block (the
// total goes to *total), and the slots themselves at their tile's offset.
SC_GLOBAL void k_inst_count(const int* inst, int n, int instance, int* cnt) {
  __shared__ int s_w[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const uint64_t m = __ballot(i < n && inst[i] == instance);
  if ((threadIdx.x & 63u) == 0) s_w[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int w = 0; w < kBlock / 64; w++) c += s_w[w];
    cnt[blockIdx.x] = c;
  }
}

constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_inst_scan(int* cnt, int tiles, int* total) {
  __shared__ int s_sum[kScanThreads];
  const int t = (int)threadIdx.x;
  const int per = (tiles + kScanThreads - 1) / kScanThreads;
  const int b0 = min(t * per, tiles), b1 = min(b0 + per, tiles);
  int own = 0;
  for (int b = b0; b < b1; b++) own += cnt[b];
  s_sum[t] = own;
  __syncthreads();
  for (int o = 1; o < kScanThreads; o <<= 1) {      // inclusive scan of the chunk sums
    const int v = t >= o ? s_sum[t - o] : 0;
    __syncthreads();
    s_sum[t] += v;
    __syncthreads();
  }
  int run = s_sum[t] - own;                         // exclusive
  for (int b = b0; b < b1; b++) {
    const int c = cnt[b];
    cnt[b] = run;
    run += c;
  }
  if (t == kScanThreads - 1) *total = s_sum[t];
}

SC_GLOBAL void k_inst_write(const int* inst, int n, int instance, const int* off, int* out) {
  __shared__ int s_w[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const bool mine = i < n && inst[i] == instance;
  const uint64_t m = __ballot(mine);
  const int lane = (int)(threadIdx.x & 63u), wave = (int)(threadIdx.x >> 6);
  if (lane == 0) s_w[wave] = __popcll(m);
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wave; w++) before += s_w[w];
  if (mine) out[off[blockIdx.x] + before + __popcll(m & ((1ull << lane) - 1ull))] = i;
}

void launch_instance_slots(Pop& p, int n, int instance, int* tile_cnt, int* out, int* total, hipStream_t s) {
  const unsigned tiles = blocks_for((size_t)std::max(n, 1));
  const int* inst = p.c<int>(F_INST);
  SC_LAUNCH(k_inst_count, dim3(tiles), dim3(kBlock), 0, s, inst, n, instance, tile_cnt);
  SC_LAUNCH(k_inst_scan, dim3(1), dim3(kScanThreads), 0, s, tile_cnt, (int)tiles, total);
  SC_LAUNCH(k_inst_write, dim3(tiles), dim3(kBlock), 0, s, inst, n, instance, (const int*)tile_cnt, out);
}

void launch_gather(Pop& p, const int* slots, int m, void* const* out, hipStream_t s) {
  GatherArgs a;
  a.p = p; a.slots = slots; a.m = m; a.m_dev = nullptr; a.stride = m;
  for (int f = 0; f < F_COUNT; f++) a.out[f] = out[f];
  SC_LAUNCH(k_gather, dim3(blocks_for((size_t)m)), dim3(kBlock), 0, s, a);
}

void launch_gather_dev(Pop& p, const int* slots, const int* m_dev, int rows, void* const* out, hipStream_t s) {
  GatherArgs a;
  a.p = p; a.slots = slots; a.m = 0; a.m_dev = m_dev; a.stride = rows;
  for (int f = 0; f < F_COUNT; f++) a.out[f] = out[f];
  SC_LAUNCH(k_gather, dim3(blocks_for((size_t)std::max(rows, 1))), dim3(kBlock), 0, s, a);
}

void launch_fill_u32(uint32_t* dst, uint32_t value, size_t n, hipStream_t s) {
  if (n == 0) return;
  unsigned b = blocks_for(n);
  if (b > 4096) b = 4096;
  SC_LAUNCH(k_fill_u32, dim3(b), dim3(kBlock), 0, s, dst, value, n);
}

// Standalone binning (sc_bin): the active count of the caller's array into
// the counters, and, after the sort, the inactive marker of the sorted keys
// back to the reference's 0xFFFFFFFF (grid.cuh:43; the marker sorts last
// either way).
SC_GLOBAL void k_bin_init(Counters* ctr, int n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctr->n_active[0] = n; ctr->n_active[1] = 0; ctr->n_sorted[0] = 0; ctr->n_sorted[1] = 0;
  }
}
SC_GLOBAL void k_bin_marker(uint32_t* keys, int n, uint32_t sentinel) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n && keys[i] == sentinel) keys[i] = 0xffffffffu;
}
void launch_bin_init(Counters* ctr, int n, hipStream_t s) { SC_LAUNCH(k_bin_init, dim3(1), dim3(64), 0, s, ctr, n); }
void launch_bin_marker(uint32_t* keys, int n, uint32_t sentinel, hipStream_t s) {
  if (n > 0) SC_LAUNCH(k_bin_marker, dim3(blocks_for((size_t)n)), dim3(kBlock), 0, s, keys, n, sentinel);
}

// Keys per thread of the onesweep tiles: GridP::sort_items when the
// look-back words hold that many tiles, else the default.
static int sort_items_of(const Pop& p, const GridP& g) {
  const int items = g.sort_items;
  if (items != 4 && items != 8 && items != 16) return kSortItems;
  const size_t tiles = ((size_t)p.launch + (size_t)kBlock * items - 1) / ((size_t)kBlock * items);
  return tiles <= (size_t)p.max_tiles ? items : kSortItems;
}

void launch_keys(Pop& p, int pop, const GridP& g, Counters* ctr, hipStream_t s) {
  PrepArgs2 z;
  z.hist = p.hist; z.status = p.status; z.status_stride = (size_t)p.max_tiles * kRadix;
  z.tile_ctr = p.tile_ctr; z.ctr = ctr; z.pop = pop; z.passes = g.passes;
  z.tile_keys = kBlock * sort_items_of(p, g);
  unsigned zb = blocks_for((size_t)p.max_tiles * kRadix);
  if (zb < 4) zb = 4;
  if (zb > 2048) zb = 2048;
  SC_LAUNCH(k_sort_zero, dim3(zb), dim3(kBlock), 0, s, z);
  KeysArgs a;
  a.pos = p.c<float4>(F_POS); a.hdr = p.c<int4>(F_HDR);
  a.keys = p.keys_in; a.hist = p.hist; a.gap_ctr = p.gap_ctr;
  a.ctr = ctr; a.g = pop_grid(g, p); a.pop = pop;
  const unsigned tiles = blocks_for(p.launch);
  a.rounds = (int)std::min<unsigned>(kKeysRounds, std::max<unsigned>(1u, tiles / kKeysMinBlocks));
  SC_LAUNCH(k_keys, dim3((tiles + a.rounds - 1) / a.rounds), dim3(kBlock), 0, s, a);
}

void launch_sort_passes(Pop& p, int pop, const GridP& g, Counters* ctr, hipStream_t s) {
  const uint32_t* src_k = p.keys_in;
  const uint32_t* src_v = nullptr;
  for (int ps = 0; ps < g.passes; ps++) {
    bool last = (ps == g.passes - 1);
    uint32_t* dst_k = last ? p.keys_sorted : (src_k == p.keys_in ? p.keys_tmp : p.keys_in);
    uint32_t* dst_v = last ? p.vals_sorted : (p.vals_tmp + (size_t)(ps & 1) * p.cap);
    SweepArgs a;
    a.keys_in = src_k; a.vals_in = src_v; a.keys_out = dst_k; a.vals_out = dst_v;
    a.hist = p.hist + ps * kRadix;
    a.status = p.status + (size_t)ps * p.max_tiles * kRadix;
    a.tile_ctr = p.tile_ctr + ps;
    a.ctr = ctr; a.pop = pop; a.shift = ps * kRadixBits;
    const int items = sort_items_of(p, g);
    const size_t tile = (size_t)kBlock * items;
    const dim3 grid((unsigned)(((size_t)p.launch + tile - 1) / tile));
    if (items == 4) SC_LAUNCH(k_onesweep<4>, grid, dim3(kBlock), 0, s, a);
    else if (items == 8) SC_LAUNCH(k_onesweep<8>, grid, dim3(kBlock), 0, s, a);
    else SC_LAUNCH(k_onesweep<16>, grid, dim3(kBlock), 0, s, a);
    src_k = dst_k;
    src_v = dst_v;
  }
}

void launch_cell_starts(Pop& p, int pop, const GridP& g, Counters* ctr, hipStream_t s) {
  CellStartArgs c;
  c.K = p.keys_sorted; c.cs = p.cs; c.gaps = p.gaps; c.gap_ctr = p.gap_ctr; c.ctr = ctr; c.pop = pop;
  c.cells = g.drop_evicted ? g.sentinel + 1u : g.sentinel;
  SC_LAUNCH(k_cell_start, dim3(blocks_for((size_t)p.launch + 1)), dim3(kBlock), 0, s, c);
  SC_LAUNCH(k_cell_gaps, dim3(1024), dim3(kBlock), 0, s, c);
  if (g.drop_evicted) SC_LAUNCH(k_drop_evicted, dim3(1), dim3(64), 0, s, c);
}

void launch_permute(Pop& p, int pop, const GridP& g, Counters* ctr, hipStream_t s) {
  PermArgs a;
  a.p = p; a.ctr = ctr; a.g = g; a.pop = pop;
  SC_LAUNCH(k_permute, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
  launch_cell_starts(p, pop, g, ctr, s);
  const Field fs[] = {F_POS, F_VEL, F_ROT, F_AVEL, F_PNOISE, F_ANOISE, F_DBG, F_HDR, F_ST, F_INST};
  for (Field f : fs) p.flip(f);
  if (p.metabolic) { p.flip(F_MET0); p.flip(F_MET1); p.flip(F_MET2); }
  else p.flip(F_BONDS);
}

void launch_tile_plans(Pop& p, Pop& m, const GridP& g, Counters* ctr, hipStream_t s) {
  // the metabolic tiles' second plan covers the merged metabolic + plain array
  struct Job { Pop* tile; int pop; Pop* tgt; Pop* tgt2; TilePlan* out; };
  Job jobs[4] = {{&p, 0, &p, nullptr, p.plans_self}, {&p, 0, &m, nullptr, p.plans_other},
                 {&m, 1, &m, nullptr, m.plans_self}, {&m, 1, &m, &p, m.plans_other}};
  for (const Job& jb : jobs) {
    PlanArgs a;
    a.K = jb.tile->keys_sorted; a.cs_self = jb.tile->cs; a.cs_tgt = jb.tgt->cs;
    a.cs_tgt2 = jb.tgt2 ? jb.tgt2->cs : nullptr; a.plans = jb.out;
    a.ctr = ctr; a.pop = jb.pop; a.g = pop_grid(g, *jb.tile); a.tile = kBlock;
    size_t tiles = ((size_t)jb.tile->launch + kBlock - 1) / kBlock;
    SC_LAUNCH(k_tile_plans, dim3(blocks_for(tiles)), dim3(kBlock), 0, s, a);
  }
}

void launch_move(Pop& p, int pop, const GridP& g, const Phys& ph, Counters* ctr, hipStream_t s) {
  MoveArgs a;
  a.dbg = p.c<float4>(F_DBG); a.pnoise = p.c<float4>(F_PNOISE); a.anoise = p.c<float4>(F_ANOISE);
  a.pos = p.c<float4>(F_POS); a.hdr = p.c<int4>(F_HDR); a.ctr = ctr; a.pop = pop;
  a.noise = ph.movement_noise[pop]; a.rot_noise = ph.rotation_noise; a.g = pop_grid(g, p);
  a.se = ph.noise_model == SC_NOISE_STOKES_EINSTEIN; a.se_translation = ph.se_translation; a.se_rotation = ph.se_rotation;
  SC_LAUNCH(k_move, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
}

// Unbonded and inactive particles only copy their noise, so after the first
// pass both buffers hold the same values: later passes visit the bonded
// particles only (listed in work[0], which the relax passes rebuild later).
void launch_coordinate(Pop& p, const GridP& g, Counters* ctr, int pass, float frac, hipStream_t s) {
  CoordArgs a;
  a.pos = p.c<float4>(F_POS); a.hdr = p.c<int4>(F_HDR); a.st = p.c<int4>(F_ST); a.bonds = p.c<int4>(F_BONDS);
  a.cap = p.cap; a.pn_c = p.c<float4>(F_PNOISE); a.an_c = p.c<float4>(F_ANOISE);
  a.pn_n = p.n<float4>(F_PNOISE); a.an_n = p.n<float4>(F_ANOISE);
  a.cs = p.cs; a.ctr = ctr; a.g = pop_grid(g, p); a.frac = frac;
  a.work = pass >= 1 ? p.work[0] : nullptr; a.work_cnt = p.work_cnt[0];
  a.bnb = p.bnb; a.flw = p.flw; a.build = pass == 0;
  if (pass == 0) {
    WorkArgs w{a.hdr, a.st, nullptr, nullptr, p.work[0], p.work_cnt[0], ctr, 0, kWorkBonded};
    SC_LAUNCH(k_worklist, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, w);
  }
  SC_LAUNCH(k_coordinate, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
  p.flip(F_PNOISE); p.flip(F_ANOISE);
}

void launch_apply_noise(Pop& p, int pop, Counters* ctr, hipStream_t s) {
  ApplyNoiseArgs a;
  a.vel = p.c<float4>(F_VEL); a.avel = p.c<float4>(F_AVEL); a.pnoise = p.c<float4>(F_PNOISE);
  a.anoise = p.c<float4>(F_ANOISE); a.hdr = p.c<int4>(F_HDR); a.ctr = ctr; a.pop = pop;
  SC_LAUNCH(k_apply_noise, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
}

static RelaxArgs relax_args(Pop& p, const GridP& g, const Tables& t, Counters* ctr) {
  RelaxArgs a;
  a.pos = p.c<float4>(F_POS); a.hdr = p.c<int4>(F_HDR); a.st = p.c<int4>(F_ST); a.bonds = p.c<int4>(F_BONDS);
  a.cap = p.cap;
  a.vel_c = p.c<float4>(F_VEL); a.avel_c = p.c<float4>(F_AVEL); a.rot_c = p.c<float4>(F_ROT); a.dbg_c = p.c<float4>(F_DBG);
  a.vel_n = p.n<float4>(F_VEL); a.avel_n = p.n<float4>(F_AVEL); a.rot_n = p.n<float4>(F_ROT); a.dbg_n = p.n<float4>(F_DBG);
  a.cs = p.cs; a.plans = p.plans_self; a.list = p.lists[0]; a.ctr = ctr; a.g = pop_grid(g, p); a.t = t;
  a.work = p.work[0]; a.work_cnt = p.work_cnt[0]; a.work_first = nullptr; a.work_bond = p.work_bond[0];
  a.flw = p.flw; a.rbp = p.rbp; a.bterm = p.bterm; a.bang = p.bang;
  a.pv_c = p.pv[p.cur[F_VEL]]; a.pv_n = p.pv[p.cur[F_VEL] ^ 1];
  return a;
}

void launch_relax(Pop& p, const GridP& g, const Tables& t, Counters* ctr, int pass, hipStream_t s,
                  bool lists_built, Fork* fork) {
  RelaxArgs a = relax_args(p, g, t, ctr);
  // work[1]: every slot (the first passes); work[0]: the later passes
  const bool later = pass >= kRelaxFullPasses;
  a.work = p.work[later ? 0 : 1]; a.work_cnt = p.work_cnt[later ? 0 : 1]; a.work_bond = p.work_bond[later ? 0 : 1];
  if (pass == 0 && !lists_built) {
    launch_fill_u32(p.lists[0].pool_ctr, 0u, kPoolShards, s);
    SC_LAUNCH(k_relax_prepare, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
    SC_LAUNCH(SC_RELAX_LIST(g), dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
    WorkArgs w{a.hdr, a.st, p.lists[0].n, nullptr, p.work[1], p.work_cnt[1], ctr, 0, kWorkRelaxAll, p.work_first[1],
               p.work_bond[1], SC_LEAN_BONDED(g)};
    SC_LAUNCH(k_worklist, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, w);
  } else if (pass == 1) {
    // the back-reaction lists exist from the first metabolic pass on
    WorkArgs w{a.hdr, a.st, p.lists[0].n, p.lists[1].n, p.work[0], p.work_cnt[0], ctr, 0, kWorkRelax, p.work_first[0],
               p.work_bond[0], SC_LEAN_BONDED(g)};
    SC_LAUNCH(k_worklist, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, w);
  }
  // the bonded particles' bond terms of this pass, which the pass adds (in a
  // launch of their own, or in the pass kernel itself: fuse_bonds)
  if (g.bonds && !g.fuse_bonds) SC_LAUNCH(k_bond_terms, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
  if (!g.relax_split) {
    // one kernel: every tile's whole worklist
    const dim3 g1(blocks_for(p.launch));
    if (g_pass_timer && g_pass_timer->n < g_pass_timer->cap) {
      PassTimer& pt = *g_pass_timer;
      ++g_launches;
      hipExtLaunchKernelGGL(SC_PASS_KERNEL(g), g1, dim3(kBlock), 0, s, pt.ev[2 * pt.n], pt.ev[2 * pt.n + 1], 0, a);
      pt.n++;
    } else {
      SC_LAUNCH(SC_PASS_KERNEL(g), g1, dim3(kBlock), 0, s, a);
    }
  } else {
    // the general kernel (bonded, alignment particles, walkers), then the
    // collision-only kernel over the rest of every tile's worklist
    a.work_first = p.work_first[later ? 0 : 1];
    const dim3 g1(blocks_for(p.launch)), g2(blocks_for(p.launch));
    if (g_pass_timer && g_pass_timer->n < g_pass_timer->cap) {
      // one pair of events spans the pass: the first dispatch's start, the second's end
      PassTimer& pt = *g_pass_timer;
      g_launches += 2;
      hipExtLaunchKernelGGL(SC_PASS_KERNEL(g), g1, dim3(kBlock), 0, s, pt.ev[2 * pt.n], nullptr, 0, a);
      hipExtLaunchKernelGGL(SC_COLL_KERNEL(g), g2, dim3(kBlock), 0, s, nullptr, pt.ev[2 * pt.n + 1], 0, a);
      pt.n++;
    } else if (fork) {
      // the two kernels touch disjoint particles and read only what earlier
      // launches wrote: the general one beside the collision-only one
      hipStream_t s2 = fork->begin();
      SC_LAUNCH(SC_PASS_KERNEL(g), g1, dim3(kBlock), 0, s2, a);
      fork->end();
      SC_LAUNCH(SC_COLL_KERNEL(g), g2, dim3(kBlock), 0, s, a);
      fork->join();
    } else {
      SC_LAUNCH(SC_PASS_KERNEL(g), g1, dim3(kBlock), 0, s, a);
      SC_LAUNCH(SC_COLL_KERNEL(g), g2, dim3(kBlock), 0, s, a);
    }
  }
  p.flip(F_VEL); p.flip(F_AVEL); p.flip(F_ROT); p.flip(F_DBG);
}

static FusedArgs fused_args(Pop& p, Pop& m, const GridP& g, const Tables& t, Counters* ctr, const Fuse& f,
                            int relax_steps);

// One later pass of a kernel without swapping buffers (timing): the current
// state stays as it is.  With events t0/t1 the launch records its start and
// end in them (hipExtLaunchKernelGGL: the dispatch's own timestamps).
#define PASS_LAUNCH(kernel, grid, block, shmem, stream, ...)                                  \
  do {                                                                                     \
    ++g_launches;                                                                          \
    if (t0) {                                                                              \
      hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, t0, t1, 0, __VA_ARGS__);  \
    } else {                                                                               \
      hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                \
    }                                                                                      \
  } while (0)

// Kernel 3: the fused relaxation of the last step's components, from the
// current buffers into the next ones.
void launch_pass_only(int kernel, Pop& p, Pop& m, const GridP& g, const Tables& t, Counters* ctr, const Fuse& f,
                      int relax_steps, hipStream_t s, hipEvent_t t0, hipEvent_t t1) {
  if (kernel == 3) {
    FusedArgs u = fused_args(p, m, g, t, ctr, f, relax_steps);
    u.pvel_out = p.n<float4>(F_VEL); u.pavel_out = p.n<float4>(F_AVEL); u.mvel_out = m.n<float4>(F_VEL);
    PASS_LAUNCH(k_relax_fused, dim3((unsigned)f.tiles * kFuseWaves), dim3(kFuseWave), 0, s, u);
  } else if (kernel == 0) {
    RelaxArgs a = relax_args(p, g, t, ctr);
    a.work_first = g.relax_split ? p.work_first[0] : nullptr;
    // both kernels of a pass; the events span them (start of the first, end of the second)
    const dim3 g1(blocks_for(p.launch)), g2(blocks_for(p.launch));
    if (g.bonds && !g.fuse_bonds) SC_LAUNCH(k_bond_terms, g1, dim3(kBlock), 0, s, a);   // outside the timed span
    if (!g.relax_split) {
      PASS_LAUNCH(SC_PASS_KERNEL(g), g1, dim3(kBlock), 0, s, a);
      return;
    }
    g_launches += 2;
    if (t0) {
      hipExtLaunchKernelGGL(SC_PASS_KERNEL(g), g1, dim3(kBlock), 0, s, t0, nullptr, 0, a);
      hipExtLaunchKernelGGL(SC_COLL_KERNEL(g), g2, dim3(kBlock), 0, s, nullptr, t1, 0, a);
    } else {
      hipLaunchKernelGGL(SC_PASS_KERNEL(g), g1, dim3(kBlock), 0, s, a);
      hipLaunchKernelGGL(SC_COLL_KERNEL(g), g2, dim3(kBlock), 0, s, a);
    }
  } else if (kernel == 1) {
    RelaxMetArgs a;
    a.mpos = m.c<float4>(F_POS); a.mhdr = m.c<int4>(F_HDR); a.mvel_c = m.c<float4>(F_VEL); a.mvel_n = m.n<float4>(F_VEL);
    a.ppos = p.c<float4>(F_POS); a.phdr = p.c<int4>(F_HDR); a.pvel = p.c<float4>(F_VEL);
    a.mcs = m.cs; a.pcs = p.cs; a.mix_pos = m.mix_pos; a.mix_meta = m.mix_meta;
    a.plans_mix = m.plans_other; a.list = m.lists[0]; a.cap = m.cap; a.ctr = ctr; a.g = pop_grid(g, m);
    a.work = m.work[0]; a.work_cnt = m.work_cnt[0];
    PASS_LAUNCH(k_relax_metabolic, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, a);
  } else {
    DiffArgs a;
    a.pos = m.c<float4>(F_POS); a.hdr = m.c<int4>(F_HDR);
    a.m0c = m.c<float4>(F_MET0); a.m1c = m.c<float4>(F_MET1); a.m2c = m.c<float2>(F_MET2);
    a.m0n = m.n<float4>(F_MET0); a.m1n = m.n<float4>(F_MET1); a.m2n = m.n<float2>(F_MET2);
    a.cs = m.cs; a.plans_mm = m.plans_self; a.list = m.lists[1]; a.cap = m.cap; a.ctr = ctr; a.g = pop_grid(g, m);
    a.work = m.work[1]; a.work_cnt = m.work_cnt[1];
    PASS_LAUNCH(k_diffuse, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, a);
  }
}
#undef PASS_LAUNCH

static RelaxMetArgs relax_met_args(Pop& m, Pop& p, const GridP& g, Counters* ctr) {
  RelaxMetArgs a;
  a.mpos = m.c<float4>(F_POS); a.mhdr = m.c<int4>(F_HDR); a.mvel_c = m.c<float4>(F_VEL); a.mvel_n = m.n<float4>(F_VEL);
  a.ppos = p.c<float4>(F_POS); a.phdr = p.c<int4>(F_HDR); a.pvel = p.c<float4>(F_VEL);
  a.mcs = m.cs; a.pcs = p.cs; a.mix_pos = m.mix_pos; a.mix_meta = m.mix_meta;
  a.plans_mix = m.plans_other; a.list = m.lists[0]; a.cap = m.cap; a.ctr = ctr; a.g = pop_grid(g, m);
  a.work = nullptr; a.work_cnt = m.work_cnt[0];
  return a;
}

static void launch_metabolic_lists(Pop& m, Pop& p, const GridP& g, Counters* ctr, const RelaxMetArgs& a, hipStream_t s) {
  MixArgs xm{a.mpos, a.mhdr, m.keys_sorted, m.cs, p.cs, m.mix_pos, m.mix_meta, g, 1};
  SC_LAUNCH(k_mix, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, xm);
  MixArgs xp{a.ppos, a.phdr, p.keys_sorted, p.cs, m.cs, m.mix_pos, m.mix_meta, g, 0};
  SC_LAUNCH(k_mix, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, xp);
  launch_fill_u32(m.lists[0].pool_ctr, 0u, kPoolShards, s);
  SC_LAUNCH(k_relax_metabolic_list, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, a);
}

void launch_relax_metabolic(Pop& m, Pop& p, const GridP& g, Counters* ctr, int pass, hipStream_t s, bool lists_built) {
  RelaxMetArgs a = relax_met_args(m, p, g, ctr);
  a.work = pass >= kMetFullPasses ? m.work[0] : nullptr;
  if (pass == 0 && !lists_built) {
    launch_metabolic_lists(m, p, g, ctr, a, s);
    WorkArgs w{m.c<int4>(F_HDR), nullptr, m.lists[0].n, nullptr, m.work[0], m.work_cnt[0], ctr, 1, kWorkEntries};
    SC_LAUNCH(k_worklist, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, w);
  }
  SC_LAUNCH(k_relax_metabolic, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, a);
  m.flip(F_VEL);
}

static RelaxPartArgs relax_part_args(Pop& p, Pop& m, const GridP& g, Counters* ctr) {
  RelaxPartArgs a;
  a.ppos = p.c<float4>(F_POS); a.phdr = p.c<int4>(F_HDR); a.pvel = p.c<float4>(F_VEL);
  a.ppv = p.pv[p.cur[F_VEL]];
  a.mpos = m.c<float4>(F_POS); a.mhdr = m.c<int4>(F_HDR); a.mvel = m.c<float4>(F_VEL);
  a.mcs = m.cs; a.plans_pm = p.plans_other; a.list = p.lists[1]; a.cap = p.cap; a.ctr = ctr; a.g = pop_grid(g, p);
  return a;
}

void launch_relax_partners(Pop& p, Pop& m, const GridP& g, Counters* ctr, bool first_pass, hipStream_t s,
                           bool lists_built) {
  RelaxPartArgs a = relax_part_args(p, m, g, ctr);
  if (first_pass && !lists_built) {
    launch_fill_u32(p.lists[1].pool_ctr, 0u, kPoolShards, s);
    SC_LAUNCH(k_relax_partners_list, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
  }
  SC_LAUNCH(k_relax_partners, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
}

static FusedArgs fused_args(Pop& p, Pop& m, const GridP& g, const Tables& t, Counters* ctr, const Fuse& f,
                            int relax_steps) {
  FusedArgs u;
  u.ra = relax_args(p, g, t, ctr);
  u.ma = relax_met_args(m, p, g, ctr);
  u.pa = relax_part_args(p, m, g, ctr);
  u.f = f;
  u.pvel_in = p.c<float4>(F_VEL); u.pavel_in = p.c<float4>(F_AVEL); u.mvel_in = m.c<float4>(F_VEL);
  // the buffers current after the relaxation: plain velocities flip every
  // relax pass, metabolic ones every metabolic relax pass
  const bool pflip = relax_steps % 2 != 0, mflip = ((relax_steps + 1) / 2) % 2 != 0;
  u.pvel_out = pflip ? p.n<float4>(F_VEL) : p.c<float4>(F_VEL);
  u.pavel_out = pflip ? p.n<float4>(F_AVEL) : p.c<float4>(F_AVEL);
  u.mvel_out = mflip ? m.n<float4>(F_VEL) : m.c<float4>(F_VEL);
  u.passes = relax_steps;
  return u;
}

void launch_relax_fused(Pop& p, Pop& m, const GridP& g, const Tables& t, Counters* ctr, const Fuse& f,
                        int relax_steps, hipStream_t s) {
  // every list of the relaxation (positions do not change during it)
  RelaxArgs ra = relax_args(p, g, t, ctr);
  launch_fill_u32(p.lists[0].pool_ctr, 0u, kPoolShards, s);
  SC_LAUNCH(k_relax_prepare, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, ra);
  SC_LAUNCH(SC_RELAX_LIST(g), dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, ra);
  RelaxMetArgs ma = relax_met_args(m, p, g, ctr);
  launch_metabolic_lists(m, p, g, ctr, ma, s);
  RelaxPartArgs pa = relax_part_args(p, m, g, ctr);
  launch_fill_u32(p.lists[1].pool_ctr, 0u, kPoolShards, s);
  SC_LAUNCH(k_relax_partners_list, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, pa);
  // components, packing, marks
  FuseArgs fa{f, p.lists[0], p.lists[1], m.lists[0], p.cap, m.cap, p.c<int4>(F_ST), ctr};
  const unsigned nb = blocks_for((size_t)f.nodes);
  SC_LAUNCH(k_fuse_init, dim3(nb), dim3(kBlock), 0, s, fa);
  SC_LAUNCH(k_fuse_link, dim3(nb), dim3(kBlock), 0, s, fa);
  SC_LAUNCH(k_fuse_roots, dim3(nb), dim3(kBlock), 0, s, fa);
  SC_LAUNCH(k_fuse_pack, dim3((unsigned)f.tiles), dim3(kBlock), 0, s, fa);
  SC_LAUNCH(k_fuse_place, dim3(nb), dim3(kBlock), 0, s, fa);
  // worklists of the per-pass kernels without the fused particles
  WorkArgs wr{ra.hdr, ra.st, p.lists[0].n, nullptr, p.work[1], p.work_cnt[1], ctr, 0, kWorkRelaxAll, p.work_first[1],
              p.work_bond[1], SC_LEAN_BONDED(g)};
  SC_LAUNCH(k_worklist, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, wr);
  WorkArgs wm{ma.mhdr, nullptr, m.lists[0].n, nullptr, m.work[0], m.work_cnt[0], ctr, 1, kWorkEntries};
  SC_LAUNCH(k_worklist, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, wm);
  // every pass of the fused components
  FusedArgs u = fused_args(p, m, g, t, ctr, f, relax_steps);
  SC_LAUNCH(k_relax_fused, dim3((unsigned)f.tiles * kFuseWaves), dim3(kFuseWave), 0, s, u);
}

void launch_apply_velocities(Pop& p, int pop, const GridP& g, const Phys& ph, Counters* ctr, hipStream_t s) {
  ApplyVelArgs a;
  a.pos = p.c<float4>(F_POS); a.vel = p.c<float4>(F_VEL); a.rot = p.c<float4>(F_ROT); a.avel = p.c<float4>(F_AVEL);
  a.hdr = p.c<int4>(F_HDR); a.ctr = ctr; a.pop = pop; a.g = pop_grid(g, p);
  a.vdecay = ph.velocity_decay; a.adecay = ph.angular_velocity_decay;
  SC_LAUNCH(k_apply_velocities, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
}

static CplxArgs cplx_args(Pop& p, const GridP& g, const Tables& t, Counters* ctr) {
  CplxArgs a;
  a.pos = p.c<float4>(F_POS); a.rot = p.c<float4>(F_ROT); a.hdr = p.c<int4>(F_HDR);
  a.st_c = p.c<int4>(F_ST); a.st_n = p.n<int4>(F_ST); a.bonds_c = p.c<int4>(F_BONDS); a.bonds_n = p.n<int4>(F_BONDS);
  a.cap = p.cap; a.cs = p.cs; a.ctr = ctr; a.g = pop_grid(g, p); a.t = t; a.step_override = 0; a.use_override = 0;
  a.fresh_only = 0;
  return a;
}

void launch_complexify(Pop& p, const GridP& g, const Tables& t, Counters* ctr, int step_override, hipStream_t s) {
  CplxArgs a = cplx_args(p, g, t, ctr);
  a.step_override = step_override;
  a.fresh_only = (g.table && step_override == -1) ? 1 : 0;
  SC_LAUNCH(k_complexify, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
  p.flip(F_ST); p.flip(F_BONDS);
}

// ---- site bonds (extension of interactions.cuh:216-346) --------------------
// Rows with bond_probability > 0 bond by contact: the second partner's
// centre within contact_distance of the first partner's site (its position
// plus the row's relative position in its frame; first = participant
// order), a Philox draw shared by both partners (counter: the two ids, the
// row, the step; key: the seed with its high word offset by kSiteKey), both
// free in the row's group and with room for a bond, and a mutual choice:
// k_site_choose picks per particle and group the candidate with the lowest
// (partner id, row id); k_site_commit adds the bond to a particle whose
// chosen partner chose it back, so both sides always agree.  Runs after
// k_complexify, on its result.
constexpr int kSiteChoices = 4;
constexpr uint32_t kSiteKey = 0x9E3779B9u;

struct SiteArgs {
  const float4* pos; const float4* rot; const int4* hdr; int4* st; int4* bonds; int cap;
  const uint32_t* cs; Counters* ctr; GridP g; Tables t; int2* site;
  int init;        // the init sequence's complexify (step -1)
  int fresh_only;  // batch init: only instances that have not run their init
};

__device__ __forceinline__ v3 min_image(const GridP& g, v3 d) {
  if (g.boundary != SC_BOUNDARY_PERIODIC) return d;
  const float L = g.L, h = 0.5f * g.L;
  const float x = d.x > h ? d.x - L : (d.x < -h ? d.x + L : d.x);
  const float y = d.y > h ? d.y - L : (d.y < -h ? d.y + L : d.y);
  const float z = d.z > h ? d.z - L : (d.z < -h ? d.z + L : d.z);
  return m3(x, y, z);
}

// A particle's choice over its candidates in walk order: per group the
// lowest (partner id, row id); a group met first once the choices are full
// (kSiteChoices, or the particle's free bond slots) is left out.
struct SiteChoice {
  int bgroup[kSiteChoices], brow[kSiteChoices], bslot[kSiteChoices];
  uint64_t bkey[kSiteChoices];
  int nc = 0, room;
  __device__ explicit SiteChoice(int r) : room(r) {}
  __device__ __forceinline__ void offer(int group, uint64_t key, int row, int slot) {
    int q = 0;
    while (q < nc && bgroup[q] != group) q++;
    if (q < nc) {
      if (key < bkey[q]) { bkey[q] = key; brow[q] = row; bslot[q] = slot; }
    } else if (nc < kSiteChoices && nc < room) {
      bgroup[nc] = group; bkey[nc] = key; brow[nc] = row; bslot[nc] = slot;
      nc++;
    }
  }
};

constexpr int kSiteChunk = 8;    // neighbours whose positions load together in the site walk
constexpr int kSiteQueue = 24;   // neighbours a lane keeps for the candidate tests (LDS column)

// Each particle's choices over its stencil, in for_neighbors' order.  The
// lanes of a wave walk the 27 stencil cells in step (chunk by chunk, as far
// as the wave's fullest cell), the position tests append the few
// neighbours that pass to the lane's queue in LDS, and whenever a lane's
// queue could not take another chunk every lane of the wave runs the
// candidate tests over its queue, in order: the costly tests then run with
// most lanes busy instead of one lane at a time per chunk.
SC_GLOBAL void k_site_choose(SiteArgs a) {
  __shared__ int s_q[kSiteQueue][kBlock];
  __shared__ uint8_t s_c[kSiteQueue][kBlock];   // the stencil cell of each queued neighbour
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int n = a.ctr->n_active[0];
  // every lane stays to the end (wave-wide chunk counts and queue flushes)
  bool live = i < n;
  if (live) {
#pragma unroll
    for (int q = 0; q < kSiteChoices; q++) a.site[(size_t)i * kSiteChoices + q] = make_int2(-1, -1);
  }
  const int i0 = live ? i : 0;
  const int4 H = a.hdr[i0];
  live = live && (H.z & SC_FLAG_ACTIVE);
  const GridP g = grid_of(a.g, i0);
  live = live && !(a.fresh_only && !g.fresh);
  int cs = 0, ce = 0;
  if (H.y >= 0 && 2 * H.y + 1 < a.t.n_pidx) { cs = a.t.pidx[2 * H.y]; ce = a.t.pidx[2 * H.y + 1]; }
  bool any = false;
  for (int k = cs; k < ce; k++) any = any || a.t.partner[k].bond_probability > 0.0f;
  live = live && any;
  const int4 S = a.st[i0];
  const int nb = min(max(S.z, 0), SC_MAX_BONDS);
  SiteChoice ch(SC_MAX_BONDS - nb);
  const float4 P = a.pos[i0];
  const v3 pp = xyz(P);
  const v4 prot = q4(a.rot[i0]);
  const uint32_t draw_step = a.init ? 0xffffffffu : (uint32_t)local_step(g, a.ctr->step);
  const int cx = (int)grid_idx(g, pp.x), cy = (int)grid_idx(g, pp.y), cz = (int)grid_idx(g, pp.z);
  // d = |D - R rp| >= | |D| - |R rp| | for the minimum-image centre offset D,
  // and |R rp| = |rp| to rounding for the unit rotations a step leaves (every
  // active rotation is normalised by the velocity update before this pass):
  // a neighbour with |D| outside [min |rp| - c, max |rp| + c] (widened by a
  // margin far above rounding) fails every row's contact test.  The init call
  // sees the setup's rotations as given and tests every neighbour.
  float w2lo = 0.0f, w2hi = 3.0e38f;
  if (!a.init) {
    float lo = 3.0e38f, hi = 0.0f;
    for (int k = cs; k < ce; k++) {
      const sc_interaction_info pii = a.t.partner[k];
      if (!(pii.bond_probability > 0.0f) || pii.only_via_transition) continue;
      const float r = norm(rp_of(pii));
      lo = scm::fmin_(lo, r - pii.contact_distance);
      hi = scm::fmax_(hi, r + pii.contact_distance);
    }
    const float m = 0.05f + 0.01f * hi;
    w2lo = lo - m > 0.0f ? (lo - m) * (lo - m) : 0.0f;
    w2hi = (hi + m) * (hi + m);
  }
  // The position test of the walk: the annulus, on the offset through the
  // cell's image, which is the minimum-image offset (to rounding, far below
  // the annulus margin) for a neighbour within the cutoff once the box is
  // wider than 2.5 cutoffs; the cutoff test itself then runs with the
  // candidate tests.  A narrower box tests both as the reference does.
  const bool wide = g.L > 2.5f * a.g.cutoff;
  // a neighbour in stencil cell c that passed the position test
  auto cand = [&](int j, int c) {
    const float4 Tj = a.pos[j];
    if (wide) {
      float sx, sy, sz;
      stencil_shift(g, cx, cy, cz, c, sx, sy, sz);
      if (outside(sub(xyz(shift4(Tj, sx, sy, sz)), pp), a.g.cutoff)) return;
    }
    const int4 TH = a.hdr[j];
    if (!(TH.z & SC_FLAG_ACTIVE)) return;
    const int4 TS = a.st[j];
    const int tnb = min(max(TS.z, 0), SC_MAX_BONDS);
    if (tnb >= SC_MAX_BONDS) return;
    for (int k = cs; k < ce; k++) {
      const sc_interaction_info pii = a.t.partner[k];
      if (!(pii.bond_probability > 0.0f) || pii.only_via_transition) continue;
      // a row listed again (equal partner types) repeats the same offer
      if (k > cs && a.t.partner[k - 1].id == pii.id) continue;
      if (!((H.y == pii.first_partner_type && TH.y == pii.second_partner_type) ||
            (H.y == pii.second_partner_type && TH.y == pii.first_partner_type)))
        continue;
      const bool ord = participant_order(H.y, H.x, TH.y, TH.x);
      if (!state_match(pii, ord, S.x, TS.x)) continue;
      // from the stored positions: the same operands whichever partner asks
      const v3 fpos = ord ? pp : xyz(Tj);
      const v3 spos = ord ? xyz(Tj) : pp;
      const v4 frot = ord ? prot : q4(a.rot[j]);
      const float d = norm(sub(min_image(g, sub(spos, fpos)), rotate(rp_of(pii), frot)));
      if (!(d <= pii.contact_distance)) continue;
      if (group_in(a.bonds, a.cap, i, nb, pii.group) || group_in(a.bonds, a.cap, j, tnb, pii.group)) continue;
      const uint32_t lo = (uint32_t)min(H.x, TH.x), hi = (uint32_t)max(H.x, TH.x);
      scm::u32x4 c = {lo, hi, (uint32_t)pii.id, draw_step};
      if (!(scm::u01_(scm::philox4x32_10(c, g.seed_lo, g.seed_hi + kSiteKey).x) < pii.bond_probability)) continue;
      ch.offer(pii.group, ((uint64_t)(uint32_t)TH.x << 32) | (uint32_t)pii.id, pii.id, j);
    }
  };
  int qn = 0;
  const int t = (int)threadIdx.x;
#pragma unroll 1
  for (int c = 0; c < 27; c++) {
    uint32_t lo = 0u, hi = 0u;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    if (!live || !stencil_range(g, a.cs, cx, cy, cz, c, lo, hi, sx, sy, sz)) { lo = hi = 0u; sx = sy = sz = 0.0f; }
    const uint32_t last = hi > lo ? hi - 1u : 0u;
    const int chunks = wave_max((int)((hi - lo + kSiteChunk - 1) / kSiteChunk));
#pragma unroll 1
    for (int b = 0; b < chunks; b++) {
      const uint32_t b0 = lo + (uint32_t)b * kSiteChunk;
      float4 T[kSiteChunk];
#pragma unroll
      for (int q = 0; q < kSiteChunk; q++) T[q] = a.pos[min(b0 + (uint32_t)q, last)];
      uint32_t pass = 0u;
      if (wide) {
#pragma unroll
        for (int q = 0; q < kSiteChunk; q++) {
          const uint32_t j = b0 + (uint32_t)q;
          const float D2 = normsq(sub(xyz(shift4(T[q], sx, sy, sz)), pp));
          if (j < hi && (int)j != i && !(D2 < w2lo || D2 > w2hi)) pass |= 1u << q;
        }
      } else {
#pragma unroll
        for (int q = 0; q < kSiteChunk; q++) {
          const uint32_t j = b0 + (uint32_t)q;
          if (j >= hi || (int)j == i || outside(sub(xyz(shift4(T[q], sx, sy, sz)), pp), a.g.cutoff)) continue;
          const float D2 = normsq(min_image(g, sub(xyz(T[q]), pp)));
          if (!(D2 < w2lo || D2 > w2hi)) pass |= 1u << q;
        }
      }
      for (; pass; pass &= pass - 1u) {
        s_q[qn][t] = (int)(b0 + (uint32_t)(__ffs(pass) - 1));
        s_c[qn++][t] = (uint8_t)c;
      }
      if (__any(qn > kSiteQueue - kSiteChunk)) {    // some queue could not take another chunk
        for (int r = 0; r < qn; r++) cand(s_q[r][t], s_c[r][t]);
        qn = 0;
      }
    }
  }
  for (int r = 0; r < qn; r++) cand(s_q[r][t], s_c[r][t]);
  if (live)
    for (int q = 0; q < ch.nc; q++) a.site[(size_t)i * kSiteChoices + q] = make_int2(ch.brow[q], ch.bslot[q]);
}

SC_GLOBAL void k_site_commit(SiteArgs a) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ctr->n_active[0]) return;
  int4 S = a.st[i];
  int nb = min(max(S.z, 0), SC_MAX_BONDS);
  const int nb0 = nb;
  for (int q = 0; q < kSiteChoices; q++) {
    const int2 c = a.site[(size_t)i * kSiteChoices + q];
    if (c.x < 0) break;
    bool mutual = false;
    for (int r = 0; r < kSiteChoices; r++) {
      const int2 d = a.site[(size_t)c.y * kSiteChoices + r];
      mutual = mutual || (d.x == c.x && d.y == i);
    }
    if (!mutual || nb >= SC_MAX_BONDS) continue;
    a.bonds[(size_t)nb * a.cap + i] = make_int4(c.x, info_at(a.t, c.x).group, a.hdr[c.y].x, 0);
    nb++;
  }
  if (nb != nb0) { S.z = nb; a.st[i] = S; }
}

void launch_site_bonds(Pop& p, const GridP& g, const Tables& t, Counters* ctr, int step_override, hipStream_t s) {
  SiteArgs a;
  a.pos = p.c<float4>(F_POS); a.rot = p.c<float4>(F_ROT); a.hdr = p.c<int4>(F_HDR);
  a.st = p.c<int4>(F_ST); a.bonds = p.c<int4>(F_BONDS); a.cap = p.cap;
  a.cs = p.cs; a.ctr = ctr; a.g = pop_grid(g, p); a.t = t; a.site = p.site;
  a.init = step_override == -1 ? 1 : 0;
  a.fresh_only = (g.table && step_override == -1) ? 1 : 0;
  SC_LAUNCH(k_site_choose, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
  SC_LAUNCH(k_site_commit, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
}

void launch_transition(Pop& p, const GridP& g, const Tables& t, Counters* ctr, hipStream_t s) {
  CplxArgs a = cplx_args(p, g, t, ctr);
  SC_LAUNCH(k_transition, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
  p.flip(F_ST); p.flip(F_BONDS);
}

void launch_propagate_state(Pop& p, const GridP& g, const Tables& t, Counters* ctr, hipStream_t s) {
  CplxArgs a = cplx_args(p, g, t, ctr);
  SC_LAUNCH(k_propagate_state, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
  p.flip(F_ST);
}

void launch_polymerize(Pop& p, const GridP& g, const Tables& t, const Phys& ph, Counters* ctr, InstCtr* ictr,
                       hipStream_t s) {
  PolyArgs a;
  a.pos = p.c<float4>(F_POS); a.vel = p.c<float4>(F_VEL); a.rot = p.c<float4>(F_ROT); a.avel = p.c<float4>(F_AVEL);
  a.pnoise = p.c<float4>(F_PNOISE); a.anoise = p.c<float4>(F_ANOISE); a.dbg = p.c<float4>(F_DBG);
  a.hdr = p.c<int4>(F_HDR); a.st = p.c<int4>(F_ST); a.bonds = p.c<int4>(F_BONDS); a.cap = p.cap;
  a.list = (int*)p.vals_tmp;   // free between sorts
  a.cs = p.cs; a.ctr = ctr; a.g = pop_grid(g, p); a.t = t; a.capacity = ph.capacity_plain;
  a.inst = p.c<int>(F_INST); a.ictr = ictr;
  SC_LAUNCH(k_poly_collect, dim3(blocks_for(p.launch)), dim3(kBlock), 0, s, a);
  SC_LAUNCH(k_poly_serial, dim3(1), dim3(64), 0, s, a);
}

void launch_diffuse(Pop& m, const GridP& g, Counters* ctr, int pass, hipStream_t s) {
  DiffArgs a;
  a.pos = m.c<float4>(F_POS); a.hdr = m.c<int4>(F_HDR);
  a.m0c = m.c<float4>(F_MET0); a.m1c = m.c<float4>(F_MET1); a.m2c = m.c<float2>(F_MET2);
  a.m0n = m.n<float4>(F_MET0); a.m1n = m.n<float4>(F_MET1); a.m2n = m.n<float2>(F_MET2);
  a.cs = m.cs; a.plans_mm = m.plans_self; a.list = m.lists[1]; a.cap = m.cap; a.ctr = ctr; a.g = pop_grid(g, m);
  a.work = pass >= kDiffFullPasses ? m.work[1] : nullptr; a.work_cnt = m.work_cnt[1];
  if (pass == 0) {
    launch_fill_u32(m.lists[1].pool_ctr, 0u, kPoolShards, s);
    SC_LAUNCH(k_diffuse_list, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, a);
    WorkArgs w{a.hdr, nullptr, m.lists[1].n, nullptr, m.work[1], m.work_cnt[1], ctr, 1, kWorkEntries};
    SC_LAUNCH(k_worklist, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, w);
  }
  SC_LAUNCH(k_diffuse, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, a);
  m.flip(F_MET0); m.flip(F_MET1); m.flip(F_MET2);
}

void launch_remove_interfering(Pop& m, Pop& p, const GridP& g, Counters* ctr, hipStream_t s) {
  RemoveArgs a;
  a.mpos = m.c<float4>(F_POS); a.mhdr = m.c<int4>(F_HDR); a.ppos = p.c<float4>(F_POS); a.phdr = p.c<int4>(F_HDR);
  a.pcs = p.cs; a.ctr = ctr; a.g = pop_grid(g, m);
  SC_LAUNCH(k_remove_interfering, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, a);
}

void launch_react(Pop& m, Pop& p, const GridP& g, const Tables& t, Counters* ctr, hipStream_t s) {
  ReactArgs a;
  a.m0 = m.c<float4>(F_MET0); a.m1 = m.c<float4>(F_MET1); a.m2 = m.c<float2>(F_MET2);
  a.mpos = m.c<float4>(F_POS); a.mhdr = m.c<int4>(F_HDR); a.ppos = p.c<float4>(F_POS); a.phdr = p.c<int4>(F_HDR);
  a.pcs = p.cs; a.ctr = ctr; a.g = pop_grid(g, m); a.t = t;
  SC_LAUNCH(k_react, dim3(blocks_for(m.launch)), dim3(kBlock), 0, s, a);
}

void launch_end_step(Counters* ctr, hipStream_t s) {
  SC_LAUNCH(k_end_step, dim3(1), dim3(64), 0, s, ctr);
}

void launch_reduce(Pop& p, int pop, Counters* ctr, void* out, hipStream_t s) {
  ReduceArgs a;
  a.pos = p.c<float4>(F_POS); a.rot = p.c<float4>(F_ROT); a.dbg = p.c<float4>(F_DBG); a.hdr = p.c<int4>(F_HDR);
  a.st = p.c<int4>(F_ST); a.m0 = p.metabolic ? p.c<float4>(F_MET0) : nullptr;
  a.out = (uint32_t*)out; a.ctr = ctr; a.pop = pop; a.cap = p.cap; a.words = p.metabolic ? 11 : 9;
  SC_LAUNCH(k_reduce, dim3(blocks_for(p.cap)), dim3(kBlock), 0, s, a);
}

void launch_neighbor_count(Pop& p, int pop, const GridP& g, Counters* ctr, uint32_t* counts, hipStream_t s) {
  NbrArgs a;
  a.pos = p.c<float4>(F_POS); a.hdr = p.c<int4>(F_HDR); a.cs = p.cs; a.ctr = ctr; a.g = pop_grid(g, p); a.pop = pop;
  a.counts = counts; a.offsets = nullptr; a.out = nullptr;
  SC_LAUNCH(k_neighbors, dim3(blocks_for(p.cap)), dim3(kBlock), 0, s, a);
}

void launch_neighbor_fill(Pop& p, int pop, const GridP& g, Counters* ctr, const uint32_t* offsets, int32_t* out,
                          hipStream_t s) {
  NbrArgs a;
  a.pos = p.c<float4>(F_POS); a.hdr = p.c<int4>(F_HDR); a.cs = p.cs; a.ctr = ctr; a.g = pop_grid(g, p); a.pop = pop;
  a.counts = nullptr; a.offsets = offsets; a.out = out;
  SC_LAUNCH(k_neighbors, dim3(blocks_for(p.cap)), dim3(kBlock), 0, s, a);
}

}  // namespace scd
