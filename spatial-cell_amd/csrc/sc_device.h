// Device data layout and kernel launch interface of the gfx950 pipeline.
//
// State lives in HBM as structure-of-arrays, one 16-byte lane per field so
// that every per-particle access is a single coalesced dwordx4:
//   pos    float4 {x, y, z, radius}     vel   float4 {x, y, z, 0}
//   rot    float4 quaternion            avel  float4 quaternion (angular velocity)
//   pnoise float4 {x, y, z, 0}          anoise float4 quaternion
//   dbg    float4 debug vector          hdr   int4 {id, type, flags, hydrophobic}
//   st     int4 {state, lastStateChangeStep, nBonds, scheduledPolymerization}
//   bonds  int4 {type, group, partnerId, 0}, bond k of slot i at [k * cap + i]
//   met0/met1 float4, met2 float2       the 10 metabolite concentrations
// Every field has two buffers; a Jacobi pass reads one and writes the other
// (the reference double-buffers whole 256/304-byte structs, memory.cuh:139-160).
// Cells: the reference's power-of-two grid and packed keys (grid.cuh:13-25).
// Cell ranges are stored as a CSR table cs[k] = first sorted slot whose key is
// >= k, for every key k in [0, cells] (cells = 2^(3*bits)), so the slots of
// any contiguous key range [k0, k1] are [cs[k0], cs[k1 + 1]) -- the reference's
// {start, end} pairs (grid.cuh:46-73) are cs[k], cs[k + 1].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/spatial_cell.h"

namespace scd {

constexpr int kBlock = 256;
constexpr int kSortItems = 16;                    // keys per thread of a onesweep tile (default)
constexpr int kSortTile = kBlock * kSortItems;   // keys per onesweep tile (default)
constexpr int kSortItemsMin = 4;                 // smallest tile the sort can take (GridP::sort_items)
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kMaxPasses = 4;

enum Field { F_POS, F_VEL, F_ROT, F_AVEL, F_PNOISE, F_ANOISE, F_DBG, F_HDR, F_ST, F_BONDS,
             F_MET0, F_MET1, F_MET2, F_INST, F_COUNT };   // F_INST: instance of each slot (batches)

// LDS staging plan of one 256-slot tile: up to three contiguous slot ranges
// of the target array (see k_tile_plans).
struct TilePlan {
  uint32_t nr, total;
  uint32_t s[3], e[3];
};

// Per-particle neighbour lists kept for the passes of one step: entry k of
// particle i at e[k * cap + i] = {neighbour slot, image code | kind bits};
// n[i] entries, n[i] > kPairSlots when they did not fit.
constexpr int kPairSlots = 16;
// Entries past the slots go to a pool split into shards (one per group of
// blocks, to spread the allocation atomics): a particle with more entries
// reserves one run in its block's shard; pool_off[i] is where it starts.
constexpr int kPoolShards = 64;
constexpr uint32_t kPoolStride = 64;   // pool rows: one column per lane of the wave that reserved the region
constexpr int kBondNbrs = 4;            // bond partner entries per slot (Pop::bnb)
constexpr int kPoolFactor = 8;       // initial pool entries per slot of the array
struct PairList {
  uint2* e; int* n;
  uint2* pool; uint32_t* pool_off; uint32_t* pool_ctr;   // pool_ctr[kPoolShards]: entries asked per shard
  uint32_t shard_cap;                                     // entries per shard
};

struct Pop {
  void* buf[F_COUNT][2];   // [field][buffer]
  int cur[F_COUNT];        // which buffer holds the current values
  int cap;
  int launch;              // slots the kernel grids cover: >= the active slots, <= cap
  bool metabolic;
  uint32_t* keys_in;       // unsorted keys
  uint32_t* keys_tmp;      // ping-pong for intermediate passes
  uint32_t* keys_sorted;   // output of the last pass (kept for clearing ranges)
  uint32_t* vals_tmp;
  uint32_t* vals_sorted;   // permutation: pre-sort slot of each sorted slot
  uint32_t* cs;            // CSR cell starts, cells + 1 entries
  uint4* gaps;             // long runs of empty cells: {first cell, end cell, slot, 0}
  uint32_t* gap_ctr;
  uint32_t* hist;          // [kMaxPasses][kRadix]
  uint32_t* status;        // [kMaxPasses][max tiles][kRadix]
  uint32_t* tile_ctr;      // [kMaxPasses]
  uint32_t* flw;           // per slot: active flag and hydrophobicity packed for the list builders
  int2* site;              // plain only: per slot the site-bond choices of the step, 4 x {row id, partner slot}
  uint2* bnb;              // plain only: per slot the bond partners met by the step's stencil walk,
                           // kBondNbrs x {slot, bond mask}, entry k at [k * cap + slot]
  uint2* rbp;              // plain only: per bonded slot the partner entries of its relax list, in list
                           // order, kBondNbrs x {slot, entry word}, entry k at [k * cap + slot]
  float4* bterm;           // plain only: per bonded slot the velocity terms of its bonds in list order
                           // (SC_MAX_BONDS x float4, entry k at [k * cap + slot]), rewritten every pass
  float4* bang;            // plain only: per bonded slot the angular term of its bonds, every pass
  float4* pv[2];           // plain only: per slot its position and velocity side by side (2 x float4,
                           // one 32-byte record: a relax pass gathers a listed neighbour's pair from one
                           // cache line); during the relax passes pv[cur[F_VEL]] holds the current
                           // velocities, written wherever F_VEL's are
  TilePlan* plans_self;    // tiles of this array over this array
  TilePlan* plans_other;   // plain: tiles over the metabolic array; metabolic: tiles over the merged array
  float4* mix_pos;         // metabolic only: merged metabolic + plain array (positions, slot | flags)
  uint32_t* mix_meta;
  PairList lists[2];       // plain: relax, lipid back-reaction; metabolic: relax, diffusion
  int* work[2];            // slots passes visit (per 256-slot block); plain: later coordination and relax
                           // passes, first relax passes; metabolic: relax, diffusion
  int* work_cnt[2];        // listed slots per block
  int* work_bond[2];       // relax worklists: per block, the range of its bonded slots (first << 16 | count; k_bond_terms)
  int* work_first[2];      // relax worklists: per block, the listed slots at its front that take the
                           // general pass kernel (alignment entries, stencil walkers, bonded particles
                           // but the collision-only ones when those take the other kernel); the rest
                           // take the collision-only kernel
  int max_tiles;

  template <class T> __host__ __device__ T* c(Field f) const { return (T*)buf[f][cur[f]]; }
  template <class T> __host__ __device__ T* n(Field f) const { return (T*)buf[f][cur[f] ^ 1]; }
  void flip(Field f) { cur[f] ^= 1; }
};

// Fused relaxation: the connected components of a step's coupling graph
// (relax, metabolic relax and back-reaction lists) that fit a block run all
// passes of the relaxation in one kernel.  Nodes: plain slots [0, pcap), then
// metabolic slots [pcap, pcap + mcap).
constexpr int kFuseTile = 128;   // nodes whose components one tile's waves take
constexpr int kFuseWave = 64;    // members per fused block (one wave, one lane each)
constexpr int kFuseWaves = 4;    // fused blocks per tile
constexpr int kMetaFusedBit = 1 << 20;   // in a list's meta word: the particle's relaxation ran fused
constexpr int kFuseCap = kFuseWave * kFuseWaves;   // member slots per tile
struct Fuse {
  int* parent;    // union-find parent (after the roots pass: the root)
  int* size;      // per root: members
  int* flags;     // per node: 1 = needs the per-pass kernels; per root: 2 = its component does,
                  // 4 = it has a particle with alignment terms
  int* offset;    // per root: first member slot in its tile, or -1 (per-pass kernels)
  int* fill;      // per root: members placed so far
  int* members;   // per tile: kFuseCap node ids, kFuseWave per fused block
  int* count;     // per fused block (tile * kFuseWaves + wave): members
  int* loc;       // per node: lane of its fused block, or -1
  int* any_walk;  // some list walks the stencil: nothing is fused this step
  int pcap, nodes, tiles;
};

struct Counters {
  int n_active[2];
  int next_particle_id;
  int step;
  int n_sorted[2];      // slots covered by the last sort (cell starts built from them)
  int n_poly;           // scheduled polymerisations found this step
  uint32_t rmax_bits[2];   // largest radius the last sort saw (float bits; a negative one as +inf; 0: none)
};

// Constants of one cell (instance).  A batch keeps one per instance in a
// device table; its particles' kernels look theirs up through the slot's
// instance (grid_of in the kernels).
// Tiles per XCD run in the neighbour kernels (GridP::xcd_run).
constexpr int kXcdRunDefault = 0;
constexpr int kCollGatherDefault = 8;
constexpr int kPassWavesDefault = 3;
constexpr int kSortItemsDefault = kSortItems;

struct GridP {
  float cell;           // gridCellSize
  float L;              // simSize
  float cutoff;         // maxInteractionDistance
  float diff_cutoff;    // maxDiffusionDistance
  int n_cells;          // power-of-two cells per axis (config.cuh:66-74)
  int bits;             // bits per axis
  int boundary;         // SC_BOUNDARY_*
  int n_per;            // periodic cells per axis (L / cell)
  uint32_t sentinel;    // sort key of an inactive slot = number of cells (of all instances)
  int key_bits;         // bits of the largest key
  int passes;           // radix passes
  uint32_t key_base;    // first cell key of this instance (aligned to 1 << 3 * bits); 0 for a single cell
  uint32_t seed_lo, seed_hi;   // Philox key of this instance
  int step0;            // global step at which this instance's step 0 ran
  int fresh;            // batch: the instance has not run its init sequence yet
  int evicted;          // batch: the instance's slots are dropped at the next sort
  int drop_evicted;     // (uniform) evicted slots sort past the sentinel and leave the array
  int xcd_run;          // neighbour kernels: tiles per run handed to one XCD (0: dispatch order)
  int coll_gather;      // collision-only relax kernel: neighbours gathered per round trip (4 or 8)
  int relax_split;      // 1: relax passes as two kernels (general, collision-only); 0: one kernel
  int pass_waves;       // waves per SIMD the general relax pass kernel is compiled for (3 or 4)
  int sort_items;       // onesweep keys per thread (4, 8 or 16: tiles of 1024, 2048 or 4096 keys)
  int fuse_bonds;       // 1: each relax pass computes its bonded particles' bond terms itself (no k_bond_terms)
  int bonds;            // (uniform) a bond can exist: interaction rows, or a particle created with bonds;
                        // each relax pass then has its bond terms (k_bond_terms or k_relax_pass_bt)
  const GridP* table;   // batch: per-instance table (device); null for a single cell
  const int* inst;      // batch: instance of each slot of the array a kernel runs over
};

// Per-instance counters of a batch (device): ids and slots of the plain array.
struct InstCtr {
  int next_particle_id;
  int n_plain;          // plain slots of the instance
  int cap_plain;        // its capacity (numParticles)
  int n_metabolic;
};

struct Phys {
  float movement_noise[2];   // plain, metabolic
  float rotation_noise;
  int noise_model;           // SC_NOISE_*
  float se_translation, se_rotation;   // Stokes-Einstein coefficients (nm^3)
  float velocity_decay;
  float angular_velocity_decay;
  uint32_t seed_lo, seed_hi;
  int coord_steps;
  int capacity_plain;
};

struct Tables {
  const sc_interaction_info* flat; int n_flat;
  const sc_interaction_info* partner; int n_partner;
  const int32_t* pidx; int n_pidx;
  const sc_type_info* types; int n_types;
  const sc_reaction* rx; int n_rx;     // metabolite reactions, at most SC_MAX_REACTIONS
  int rx_catalysed;                     // some reaction counts catalysts
};

// ---- launch wrappers (sc_kernels.hip) ----------------------------------------
void launch_keys(Pop& p, int pop, const GridP& g, Counters* ctr, hipStream_t s);
void launch_sort_passes(Pop& p, int pop, const GridP& g, Counters* ctr, hipStream_t s);
void launch_permute(Pop& p, int pop, const GridP& g, Counters* ctr, hipStream_t s);
// the CSR cell starts of p.keys_sorted (part of launch_permute)
void launch_cell_starts(Pop& p, int pop, const GridP& g, Counters* ctr, hipStream_t s);
// standalone binning: set the counters' active count; sorted inactive markers -> 0xFFFFFFFF
void launch_bin_init(Counters* ctr, int n, hipStream_t s);
void launch_bin_marker(uint32_t* keys, int n, uint32_t sentinel, hipStream_t s);
void launch_tile_plans(Pop& p, Pop& m, const GridP& g, Counters* ctr, hipStream_t s);
void launch_move(Pop& p, int pop, const GridP& g, const Phys& ph, Counters* ctr, hipStream_t s);
void launch_coordinate(Pop& p, const GridP& g, Counters* ctr, int pass, float frac, hipStream_t s);
void launch_apply_noise(Pop& p, int pop, Counters* ctr, hipStream_t s);
// Kernel launches made by this thread so far (one graph node each when captured).
extern thread_local uint64_t g_launches;

// While set, the relax passes of the steps this thread enqueues launch with
// hipExtLaunchKernelGGL and record each dispatch's start and end in the next
// pair of ev (up to cap pairs; n counts them).
struct PassTimer {
  hipEvent_t* ev;
  int cap;
  int n;
};
extern thread_local PassTimer* g_pass_timer;

// Work beside the main stream's within a step: begin() returns the stream
// for it (forked from the main stream's current point), end() returns the
// main stream to that point, join() makes the main stream's next work wait
// for both.
struct Fork {
  virtual hipStream_t begin() = 0;
  virtual void end() = 0;
  virtual void join() = 0;
};

// lists_built: launch_relax_fused built the pass-0 lists and worklists;
// fork (two relax kernels only): the general kernel runs beside the
// collision-only one
void launch_relax(Pop& p, const GridP& g, const Tables& t, Counters* ctr, int pass, hipStream_t s,
                  bool lists_built = false, Fork* fork = nullptr);
// Before relax pass 0 of a step: every list of the relaxation, the coupling
// graph's components, the worklists without the fused particles, and the
// fused kernel for all relax_steps passes of the fused components.
void launch_relax_fused(Pop& p, Pop& m, const GridP& g, const Tables& t, Counters* ctr, const Fuse& f,
                        int relax_steps, hipStream_t s);
void launch_relax_metabolic(Pop& m, Pop& p, const GridP& g, Counters* ctr, int pass, hipStream_t s,
                            bool lists_built = false);
void launch_relax_partners(Pop& p, Pop& m, const GridP& g, Counters* ctr, bool first_pass, hipStream_t s, bool lists_built = false);
void launch_apply_velocities(Pop& p, int pop, const GridP& g, const Phys& ph, Counters* ctr, hipStream_t s);
void launch_complexify(Pop& p, const GridP& g, const Tables& t, Counters* ctr, int step_override, hipStream_t s);
// site bonds of rows with bond_probability > 0, after launch_complexify (same step_override)
void launch_site_bonds(Pop& p, const GridP& g, const Tables& t, Counters* ctr, int step_override, hipStream_t s);
void launch_transition(Pop& p, const GridP& g, const Tables& t, Counters* ctr, hipStream_t s);
void launch_propagate_state(Pop& p, const GridP& g, const Tables& t, Counters* ctr, hipStream_t s);
void launch_polymerize(Pop& p, const GridP& g, const Tables& t, const Phys& ph, Counters* ctr, InstCtr* ictr,
                       hipStream_t s);
void launch_diffuse(Pop& m, const GridP& g, Counters* ctr, int pass, hipStream_t s);
void launch_react(Pop& m, Pop& p, const GridP& g, const Tables& t, Counters* ctr, hipStream_t s);
void launch_gather(Pop& p, const int* slots, int m, void* const* out, hipStream_t s);
// the same with the row count on the device (at most `rows`, the field stride)
void launch_gather_dev(Pop& p, const int* slots, const int* m_dev, int rows, void* const* out, hipStream_t s);
// the slots among the first n whose instance is `instance`, in slot order, to
// out; their number to *total (tile_cnt: one int per 256 slots)
void launch_instance_slots(Pop& p, int n, int instance, int* tile_cnt, int* out, int* total, hipStream_t s);
void launch_pass_only(int kernel, Pop& p, Pop& m, const GridP& g, const Tables& t, Counters* ctr, const Fuse& f,
                      int relax_steps, hipStream_t s,
                      hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
void launch_remove_interfering(Pop& m, Pop& p, const GridP& g, Counters* ctr, hipStream_t s);
void launch_end_step(Counters* ctr, hipStream_t s);
void launch_reduce(Pop& p, int pop, Counters* ctr, void* out, hipStream_t s);
void launch_neighbor_count(Pop& p, int pop, const GridP& g, Counters* ctr, uint32_t* counts, hipStream_t s);
void launch_neighbor_fill(Pop& p, int pop, const GridP& g, Counters* ctr, const uint32_t* offsets,
                          int32_t* out, hipStream_t s);
void launch_fill_u32(uint32_t* dst, uint32_t value, size_t n, hipStream_t s);
void launch_eval_primitive(int fn, const float* x, float* y, size_t n, hipStream_t s);
__host__ __device__ float eval_primitive(int fn, float x);

}  // namespace scd
